// zts_ref.cpp -- driver around the reference's own assembly code, compiled
// together with the reference sources (see oracle/ref/Makefile).  It is test
// infrastructure: it generates the golden impedance fixtures and validates the
// C restatement in oracle/bem_oracle.c.  It never ships and never runs on the
// product path.
//
// Input (text, stdin):
//   omega
//   nbody
//   per body: eps_re eps_im mu_re mu_im nv nt
//             nv lines "x y z", nt lines "a b c" (0-based vertex ids)
//   with argv[2] == "--source", after the bodies:
//     region px py pz jx jy jz    (dipole in the exterior for region -1, else
//                                  inside body `region`)
//     2E lines "re im"            (a solution vector X, [J, M] per RWG)
// Output (binary, file named by argv[1]; the reference's mesher prints to stdout):
//   per body: int64 ne, then ne x 4 int64 {p1, p2, t1, t2} of the reference's
//             ShapeMesh edge table (body-local ids), in Edge::index order
//   then complex128 column-major matrices:
//     Zts(top, top, exterior g)   (2E x 2E, E = total RWGs, body 1 first)
//     Zts(body_b, body_b, g_b)    for each body b
//   or, with --tree (edge tables as above, then per body):
//     int64 nseg: segments the reference's ShapeMesh keeps (m->segment)
//     int64 n, then n int64: Edge::index of each RWG in the body's subtree of
//       the reference's RWGTree(meshDiel, ...) after unshuffle (tree order)
//     int64 nleaf, then nleaf int64: RWG counts of the leaves in that order
//   or, with --source:
//     B (2E complex128): Bdense of each body's tree in body order, the
//       leaf level of BH (src/Source.cpp:199-215)
//     float64 flux(top, body_b, X) for each body b (src/Flux.cpp:100)
//
// Reference entry points used: ShapeMesh::ShapeMesh (src/ShapeMesh.cpp:821),
// RWG::RWG (src/RWG.cpp:5), RWGTree protected ctor (src/RWGTree.cpp:590),
// GreensFunction (src/GreensFunction.cpp:11), Zts (src/Impedance.cpp:388),
// DipolePointSource (src/Source.cpp:14), Bdense (src/Source.cpp:167),
// flux (src/Flux.cpp:100), RWGTree(meshDiel, edgesAtPoint, minSize,
// minRadius) (src/RWGTree.cpp:602) with the arguments parseShapes gives it
// (src/HeatTransfer.cpp:98-105).
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>
#include "ShapeMesh.h"
#include "RWGTree.h"
#include "Impedance.h"
#include "Options.h"
#include "ReferenceCountingPointer.h"
#include "Source.h"
#include "Flux.h"
#include "Progress.h"
#include <iostream>
#include <sstream>

struct FlatTree : public RWGTree {
  FlatTree(int n, PointEdgeVectorMap &m, Dielectric *d, GreensFunction *g) : RWGTree(n, m, d, g) {}
};

static FILE *out;

static void write_matrix(ComplexDenseMatrix *Z) {
  int m = Z->m, n = Z->n;
  std::vector<double> col(2 * (size_t)m);
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < m; i++) {
      Complex z = (*Z)(i, j);
      col[2 * i] = real(z);
      col[2 * i + 1] = imag(z);
    }
    fwrite(col.data(), sizeof(double), col.size(), out);
  }
}

static void leaves(RWGTree *t, std::vector<int64_t> &sizes) {
  if (!t->subs) {
    sizes.push_back(t->rwgs);
    return;
  }
  for (int i = 0; i < t->subs; i++) leaves(t->sub[i], sizes);
}

int main(int argc, char **argv) {
  const std::string mode = argc == 3 ? argv[2] : "";
  if (argc != 2 && !(argc == 3 && (mode == "--source" || mode == "--tree"))) return 2;
  const bool source_mode = mode == "--source";
  const bool tree_mode = mode == "--tree";
  out = fopen(argv[1], "wb");
  if (!out) return 2;
  initReferenceCountingPointerLock();
  double omega;
  int nbody;
  if (scanf("%lf %d", &omega, &nbody) != 2) return 2;
  Options options;
  PointEdgeVectorMap eap;
  std::vector<ShapeMesh *> meshes;
  std::vector<std::vector<Triangle *> > body_tris(nbody);
  std::vector<std::vector<RWG *> > rwgs(nbody);
  std::vector<Complex> eps(nbody), mu(nbody);
  for (int b = 0; b < nbody; b++) {
    double er, ei, mr, mi;
    int nv, nt;
    if (scanf("%lf %lf %lf %lf %d %d", &er, &ei, &mr, &mi, &nv, &nt) != 6) return 2;
    eps[b] = Complex(er, ei);
    mu[b] = Complex(mr, mi);
    Shape *shape = new Shape();
    for (int i = 0; i < nv; i++) {
      double x, y, z;
      if (scanf("%lf %lf %lf", &x, &y, &z) != 3) return 2;
      shape->coord.push_back(Vector3(x, y, z));
    }
    for (int i = 0; i < nt; i++) {
      int a, c, d;
      if (scanf("%d %d %d", &a, &c, &d) != 3) return 2;
      char buf[96];
      snprintf(buf, sizeof buf, "%d %d %d -1 ", a, c, d);
      shape->faceIndex.push_back(FaceIndex(buf));
    }
    ShapeMesh *m = new ShapeMesh(*shape, options.sharpThresh, options.concaveThresh,
                                 options.relativeCostAngleConcave, NULL, NULL);
    meshes.push_back(m);
    Array<Point> &pts = *m->refinedPoint[0];
    Array<Edge> &edges = *m->refinedEdge[0];
    Array<Triangle> &tris = *m->refinedTriangle[0];
    for (int i = 0; i < tris.n; i++) body_tris[b].push_back(tris + i);
    int64_t ne = edges.n;
    fwrite(&ne, sizeof ne, 1, out);
    for (int i = 0; i < edges.n; i++) {
      Edge *e = edges + i;
      int64_t rec[4] = {(int64_t)(e->p1 - (pts + 0)), (int64_t)(e->p2 - (pts + 0)),
                        (int64_t)(e->t1 - (tris + 0)), (int64_t)(e->t2 - (tris + 0))};
      fwrite(rec, sizeof(int64_t), 4, out);
      rwgs[b].push_back(new RWG(e));
    }
  }
  if (tree_mode) {
    // the edges-at-point map of parseShapes (getEdgesAtPoints,
    // src/ShapeMesh.cpp:1962-1973): every edge of every kept segment, in
    // segment and edge-index order
    ShapeMeshDielectricPairList meshDiel;
    for (int b = 0; b < nbody; b++) {
      for (SegmentSet::const_iterator i = meshes[b]->segment.begin(); i != meshes[b]->segment.end(); ++i)
        for (EdgeSet::const_iterator k = (*i)->edge.begin(); k != (*i)->edge.end(); ++k) {
          eap[(*k)->p1].push_back(*k);
          eap[(*k)->p2].push_back(*k);
        }
      meshDiel.push_back(ShapeMeshDielectricPair(meshes[b], (Dielectric *)NULL));
    }
    RWGTree top(meshDiel, eap, options.minSegmentSize, 0.0);
    for (int b = 0; b < nbody; b++) {
      RWGTree *t = nbody > 1 ? top.sub[b] : &top;
      int64_t nseg = meshes[b]->segment.size(), n = t->rwgs;
      fwrite(&nseg, sizeof nseg, 1, out);
      fwrite(&n, sizeof n, 1, out);
      for (int i = 0; i < t->rwgs; i++) {
        int64_t k = t->rwg[i]->e->index;
        fwrite(&k, sizeof k, 1, out);
      }
      std::vector<int64_t> sizes;
      leaves(t, sizes);
      int64_t nl = sizes.size();
      fwrite(&nl, sizeof nl, 1, out);
      fwrite(sizes.data(), sizeof(int64_t), sizes.size(), out);
    }
    fclose(out);
    return 0;
  }
  size_t total = 0;
  for (int b = 0; b < nbody; b++) total += rwgs[b].size();
  GreensFunction gext(Complex(1., 0.), Complex(1., 0.), Complex(omega, 0.), true);
  std::vector<GreensFunction *> gb(nbody);
  std::vector<FlatTree *> sub(nbody);
  FlatTree top((int)total, eap, NULL, &gext);
  for (int b = 0; b < nbody; b++) {
    gb[b] = new GreensFunction(eps[b], mu[b], Complex(omega, 0.), false);
    sub[b] = new FlatTree((int)rwgs[b].size(), eap, NULL, gb[b]);
    for (size_t i = 0; i < rwgs[b].size(); i++) {
      sub[b]->rwg[sub[b]->rwgs++] = rwgs[b][i];
      top.rwg[top.rwgs++] = rwgs[b][i];
    }
  }
  if (source_mode) {
    int region;
    double p[3], j[3];
    if (scanf("%d %lf %lf %lf %lf %lf %lf", &region, p, p + 1, p + 2, j, j + 1, j + 2) != 7) return 2;
    if (region < -1 || region >= nbody) return 2;
    ComplexDenseMatrix X(2 * (int)total, 1, 0);
    for (size_t i = 0; i < 2 * total; i++) {
      double re, im;
      if (scanf("%lf %lf", &re, &im) != 2) return 2;
      X(i, 0) = Complex(re, im);
    }
    Point pp(p[0], p[1], p[2]);
    Vector3 jj(j[0], j[1], j[2]);
    DipolePointSource dip(region < 0 ? &gext : gb[region], pp, jj);
    SourceSet sources;
    sources.insert(&dip);
    Progress progress;
    int off = 0;
    for (int b = 0; b < nbody; b++) {
      for (int i = 0; i < sub[b]->rwgs; i++) sub[b]->index.push_back(off + i);
      off += sub[b]->rwgs;
      sub[b]->triangle.insert(body_tris[b].begin(), body_tris[b].end());
      ComplexDenseMatrix *B = Bdense(sub[b], &gext, sources, progress, options);
      write_matrix(B);
      delete B;
    }
    // flux() prints a line per triangle on stdout: keep it out of the result
    std::ostringstream sink;
    std::streambuf *keep = std::cout.rdbuf(sink.rdbuf());
    for (int b = 0; b < nbody; b++) {
      double S = flux(&top, sub[b], X);
      fwrite(&S, sizeof S, 1, out);
    }
    std::cout.rdbuf(keep);
    fclose(out);
    return 0;
  }
  ComplexDenseMatrix *W = Zts(&top, &top, &gext, options);
  write_matrix(W);
  delete W;
  for (int b = 0; b < nbody; b++) {
    ComplexDenseMatrix *G = Zts(sub[b], sub[b], gb[b], options);
    write_matrix(G);
    delete G;
  }
  fclose(out);
  return 0;
}
