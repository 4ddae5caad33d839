// Host-side geometry for libbem_hip: RWG construction and the per-triangle
// records the assembly kernels read.  Compiled with -ffp-contract=off so the
// tier-deciding quantities (centroids, side lengths) round exactly as the
// reference's x86-64 build does.
//
//  edge table ........ src/ShapeMesh.cpp:856-926 (creation order, orientation),
//                      neighbour search src/ShapeMesh.cpp:712-728
//  half-RWG fields ... src/RWG.cpp:7-39
//  centroid, length .. src/Triangle.cpp:43-47, src/ShapeMesh.cpp:940-947
//  push order ........ src/Impedance.cpp:397-406 (pos then neg, RWG order)
#include "geometry.h"

#include <cmath>
#include <cstring>

namespace bemgeo {

static inline void sub3(const double *a, const double *b, double *c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}
static inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm2(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

int build_edges(int nv, int nt, const int32_t *tri, std::vector<int32_t> &edges) {
  // triangles at each vertex, ascending
  std::vector<int> off(nv + 1, 0), fill(nv, 0), tap(3 * (size_t)nt);
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < 3; k++) {
      int v = tri[3 * i + k];
      if (v < 0 || v >= nv) return -1;
      off[v + 1]++;
    }
  for (int v = 0; v < nv; v++) off[v + 1] += off[v];
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < 3; k++) {
      int v = tri[3 * i + k];
      tap[off[v] + fill[v]++] = i;
    }
  auto neighbour = [&](int i, int va, int vb) {
    for (int a = off[va]; a < off[va + 1]; a++) {
      int ta = tap[a];
      if (ta == i) continue;
      for (int b = off[vb]; b < off[vb + 1]; b++)
        if (tap[b] != i && tap[b] == ta) return ta;
    }
    return -1;
  };
  std::vector<char> touched(nt, 0);
  edges.clear();
  const int sides[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  for (int i = 0; i < nt; i++) {
    for (int s = 0; s < 3; s++) {
      int va = tri[3 * i + sides[s][0]], vb = tri[3 * i + sides[s][1]];
      int nb = neighbour(i, va, vb);
      if (nb != -1 && !touched[nb]) {
        edges.push_back(va);
        edges.push_back(vb);
        edges.push_back(i);
        edges.push_back(nb);
      }
    }
    touched[i] = 1;
  }
  return (int)(edges.size() / 4);
}

// Fill the frame fields of a TriGeom from the half-RWG (triangle tv, edge ea-eb).
static void frame_from_half(TriGeom &g, const double *X, const int32_t *tv, int ea, int eb, double elen) {
  int k;
  if (tv[0] != ea && tv[0] != eb) k = 0;
  else if (tv[1] != ea && tv[1] != eb) k = 1;
  else k = 2;
  const double *P0 = X + 3 * (size_t)tv[k], *P1 = X + 3 * (size_t)tv[(k + 1) % 3], *P2 = X + 3 * (size_t)tv[(k + 2) % 3];
  for (int c = 0; c < 3; c++) { g.p0[c] = P0[c]; g.p1[c] = P1[c]; g.p2[c] = P2[c]; }
  sub3(g.p1, g.p0, g.v01);
  sub3(g.p2, g.p1, g.v12);
  sub3(g.p0, g.p2, g.v20);
  sub3(g.p2, g.p0, g.v02);
  g.normv01i = 1. / std::sqrt(norm2(g.v01));
  g.normv12i = 1. / std::sqrt(norm2(g.v12));
  g.normv20i = 1. / std::sqrt(norm2(g.v20));
  double n[3];
  cross3(g.v01, g.v02, n);
  double A2 = std::sqrt(norm2(n));
  double s = 1. / A2;
  for (int c = 0; c < 3; c++) g.n[c] = s * n[c];
  g.Ai = 2.0 / A2;
  double m[3];
  cross3(g.v01, g.n, m);
  double mi = 1. / std::sqrt(norm2(m));
  for (int c = 0; c < 3; c++) g.m01[c] = m[c] * mi;
  cross3(g.v12, g.n, m);
  mi = 1. / std::sqrt(norm2(m));
  for (int c = 0; c < 3; c++) g.m12[c] = m[c] * mi;
  cross3(g.v20, g.n, m);
  mi = 1. / std::sqrt(norm2(m));
  for (int c = 0; c < 3; c++) g.m20[c] = m[c] * mi;
  g.elen = elen;
}

static int opposite_vertex(const int32_t *tv, int ea, int eb) {
  if (tv[0] != ea && tv[0] != eb) return tv[0];
  if (tv[1] != ea && tv[1] != eb) return tv[1];
  return tv[2];
}

void build_tree(const std::vector<const BodyMesh *> &bodies, const double *shifts, Tree &out) {
  out.tg.clear();
  out.rw.clear();
  out.close.clear();
  out.tri_off.clear();
  out.rwg_off.clear();
  int toff = 0, eoff = 0, voff = 0;
  for (size_t b = 0; b < bodies.size(); b++) {
    const BodyMesh &B = *bodies[b];
    // translated coordinates, rounded as the reference's body transform
    std::vector<double> X(3 * (size_t)B.nv);
    for (int v = 0; v < B.nv; v++)
      for (int c = 0; c < 3; c++) X[3 * v + c] = B.xyz[3 * v + c] + shifts[3 * b + c];
    out.tri_off.push_back(toff);
    out.rwg_off.push_back(eoff);
    size_t t0 = out.tg.size();
    out.tg.resize(t0 + B.nt);
    std::vector<int> nh(B.nt, 0);
    for (int t = 0; t < B.nt; t++) {
      TriGeom &g = out.tg[t0 + t];
      std::memset(&g, 0, sizeof g);
      const int32_t *tv = &B.tri[3 * (size_t)t];
      // centroid ((P1 + P2) + P3) * 1/3 and longest side
      double cx = 0, cy = 0, cz = 0, em = 0;
      for (int k = 0; k < 3; k++) {
        cx += X[3 * tv[k]];
        cy += X[3 * tv[k] + 1];
        cz += X[3 * tv[k] + 2];
      }
      g.centre[0] = cx * 0.33333333333333333333333333333333;
      g.centre[1] = cy * 0.33333333333333333333333333333333;
      g.centre[2] = cz * 0.33333333333333333333333333333333;
      for (int k = 0; k < 3; k++) {
        double d[3];
        sub3(&X[3 * tv[(k + 1) % 3]], &X[3 * tv[k]], d);
        double l = std::sqrt(norm2(d));
        if (l > em) em = l;
      }
      g.emax = em;
      g.body = (int32_t)b;
      for (int k = 0; k < 3; k++) g.vid[k] = voff + tv[k];
    }
    for (int e = 0; e < B.ne; e++) {
      const int32_t *E = &B.edges[4 * (size_t)e];
      double d[3];
      sub3(&X[3 * E[1]], &X[3 * E[0]], d);
      double len = std::sqrt(norm2(d));
      RwgRec r;
      r.len = len;
      for (int h = 0; h < 2; h++) {
        int t = E[2 + h];
        TriGeom &g = out.tg[t0 + t];
        int slot = nh[t];
        if (slot < 3) {
          if (slot == 0) frame_from_half(g, X.data(), &B.tri[3 * (size_t)t], E[0], E[1], len);
          int pv = opposite_vertex(&B.tri[3 * (size_t)t], E[0], E[1]);
          for (int c = 0; c < 3; c++) g.hp0[slot][c] = X[3 * pv + c];
          nh[t]++;
        }
        if (h == 0) { r.tp = toff + t; r.sp = slot; } else { r.tn = toff + t; r.sn = slot; }
      }
      out.rw.push_back(r);
    }
    for (int t = 0; t < B.nt; t++) out.tg[t0 + t].nh = nh[t] < 3 ? nh[t] : 3;
    for (const auto &p : B.close) out.close.push_back({toff + p.first, toff + p.second});
    toff += B.nt;
    eoff += B.ne;
    voff += B.nv;
  }
  out.T = toff;
  out.E = eoff;
}

int flux_table(const BodyMesh &B, const double *shift, int rwg_off, std::vector<FluxTri> &out) {
  // translated coordinates, rounded as in build_tree
  std::vector<double> X(3 * (size_t)B.nv);
  for (int v = 0; v < B.nv; v++)
    for (int c = 0; c < 3; c++) X[3 * v + c] = B.xyz[3 * v + c] + shift[c];
  out.assign(B.nt, FluxTri());
  for (int t = 0; t < B.nt; t++) {
    FluxTri &f = out[t];
    std::memset(&f, 0, sizeof f);
    for (int k = 0; k < 3; k++) {
      for (int c = 0; c < 3; c++) f.P[k][c] = X[3 * B.tri[3 * (size_t)t + k] + c];
      f.e[k] = -1;
    }
  }
  for (int e = 0; e < B.ne; e++) {
    const int32_t *E = &B.edges[4 * (size_t)e];
    double d[3];
    sub3(&X[3 * E[1]], &X[3 * E[0]], d);
    const double len = std::sqrt(norm2(d));
    for (int h = 0; h < 2; h++) {
      const int t = E[2 + h];
      const int32_t *tv = &B.tri[3 * (size_t)t];
      for (int k = 0; k < 3; k++) {
        const int a = tv[k], b = tv[(k + 1) % 3];
        if ((a == E[0] && b == E[1]) || (a == E[1] && b == E[0])) {
          out[t].e[k] = rwg_off + e;
          out[t].c[k] = h == 0 ? len : -len;   // + on the RWG's positive (creating) triangle
        }
      }
    }
  }
  for (const FluxTri &f : out)
    if (f.e[0] < 0 || f.e[1] < 0 || f.e[2] < 0) return -1;
  return 0;
}

void build_close_pairs(BodyMesh &B) {
  // pairs (a, b) sharing at least one vertex (a == b included),
  // src/Triangle.h:23-30, grouped by test triangle a
  std::vector<std::vector<int>> at(B.nv);
  for (int t = 0; t < B.nt; t++)
    for (int k = 0; k < 3; k++) at[B.tri[3 * (size_t)t + k]].push_back(t);
  B.close.clear();
  std::vector<int> mark(B.nt, -1);
  for (int a = 0; a < B.nt; a++) {
    std::vector<int> nb;
    for (int k = 0; k < 3; k++)
      for (int t : at[B.tri[3 * (size_t)a + k]])
        if (mark[t] != a) { mark[t] = a; nb.push_back(t); }
    std::sort(nb.begin(), nb.end());
    for (int b : nb) B.close.push_back({a, b});
  }
}

}  // namespace bemgeo
