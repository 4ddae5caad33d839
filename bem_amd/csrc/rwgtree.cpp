// The reference's order of a body's unknowns: the RWGTree that its drivers
// build before assembling (parseShapes, src/HeatTransfer.cpp:75-105), as a
// host computation.  The tree splits a surface recursively along planes whose
// orientation a Powell search picks (the plane that crosses the least edge
// length), and lists the RWGs leaf by leaf.  Every step below follows the
// reference's arithmetic and container orders, because the plane search
// compares floating-point values: a different association or set order would
// send it to another plane and permute the unknowns.  Compiled without
// floating-point contraction, as the reference's x86-64 build rounds.
//
//  sharp edges, first segment ..... src/ShapeMesh.cpp:10-24, :930-970,
//                                   :1917-1927 (see keep_first_segment)
//  edges at points ................ src/ShapeMesh.cpp:1962-1973
//  segment centroid/normal/radius . src/RWGTree.cpp:21-65
//  plane cut and segments ......... src/RWGTree.cpp:67-436
//  cut length ..................... src/RWGTree.cpp:438-479
//  plane search ................... src/RWGTree.cpp:481-559
//  tree and leaf order ............ src/RWGTree.cpp:602-722, :748-761
//  line search, Powell ............ src/util/MathUtils.cpp:27-92, :193-349
//                                   (Numerical Recipes in C, 2nd ed.,
//                                   sections 10.1, 10.2 and 10.5)
//  rotation ....................... src/util/Rotation.cpp:6-11, :67-74
//  triangle normal and area ....... src/Triangle.cpp:5-28
#include "geometry.h"

#include <cfloat>
#include <cmath>
#include <memory>
#include <set>
#include <stdexcept>
#include <vector>

namespace bemgeo {
namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator-(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(double c, const V3 &a) { return {c * a.x, c * a.y, c * a.z}; }
inline V3 cross(const V3 &a, const V3 &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm2(const V3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline void normalize(V3 &a) {
  const double r = 1. / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  a.x *= r;
  a.y *= r;
  a.z *= r;
}

const double kOneThird = 0.33333333333333333333333333333333;
const double kHalfPi = 1.5707963267948966192313216916398;
const double kSharp = 2.0;  // Options::sharpThresh, src/Options.cpp

// Point sets of the reference order points by coordinates (x, then y, then
// z), so two vertices at the same place count as one.
struct PointLess {
  const V3 *P;
  bool operator()(int a, int b) const {
    const V3 &p = P[a], &q = P[b];
    if (p.x == q.x) {
      if (p.y == q.y) return p.z < q.z;
      return p.y < q.y;
    }
    return p.x < q.x;
  }
};
typedef std::set<int, PointLess> PointSet;

struct Segment {
  std::set<int> edge, boundary, tri;
};
// Segments are ordered by their lowest edge index; the reference's
// comparator calls any two segments equal when either has no edge.
struct SegmentLess {
  bool operator()(const Segment *a, const Segment *b) const {
    return !a->edge.empty() && !b->edge.empty() && *a->edge.begin() < *b->edge.begin();
  }
};

struct Surface {
  int nv = 0, nt = 0, ne = 0;
  std::vector<V3> P;
  std::vector<int32_t> tv;        // triangle vertices, 3 per triangle
  std::vector<int32_t> E;         // {p1, p2, t1, t2} per edge
  std::vector<int> side;          // edges e12, e23, e31 of each triangle
  std::vector<double> len;
  std::vector<std::vector<int>> eap;  // edges at each vertex

  int other(int e, int p) const { return E[4 * e] == p ? E[4 * e + 1] : E[4 * e]; }
  // (p2 - p1) x (p3 - p1) from the triangle's own vertices
  V3 raw_normal(int t) const {
    const V3 &a = P[tv[3 * t]], &b = P[tv[3 * t + 1]], &c = P[tv[3 * t + 2]];
    return cross(b - a, c - a);
  }
  V3 normal(int t) const {
    V3 n = raw_normal(t);
    normalize(n);
    return n;
  }
  double area(int t) const { return 0.5 * std::sqrt(norm2(raw_normal(t))); }
  // angle between the normals of the edge's two triangles
  double sharpness(int e) const {
    const V3 n1 = normal(E[4 * e + 2]), n2 = normal(E[4 * e + 3]);
    V3 z = cross(n1, n2);
    const double nz = std::sqrt(norm2(z));
    if (nz / std::sqrt(norm2(n1)) < 1e-6) return 0.;
    z = (1. / nz) * z;
    return std::atan2(dot(n2, cross(z, n1)), dot(n2, n1));
  }
};

// Flood fill over triangles from t, stopping at boundary edges: each edge
// goes to the first segment that reaches it (depth first, sides in the
// order e12, e23, e31).
void traverse(const Surface &S, Segment &s, int t, std::vector<char> &tri_done, std::vector<char> &edge_done,
              const std::vector<char> &is_boundary, const std::vector<char> *allowed) {
  std::vector<int> stack(1, t);
  while (!stack.empty()) {
    t = stack.back();
    stack.pop_back();
    s.tri.insert(t);
    tri_done[t] = 1;
    for (int k = 0; k < 3; k++) {
      const int e = S.side[3 * t + k];
      if (is_boundary[e]) s.boundary.insert(e);
      if (edge_done[e]) continue;
      s.edge.insert(e);
      edge_done[e] = 1;
      if (is_boundary[e]) continue;
      const int t2 = S.E[4 * e + 2] == t ? S.E[4 * e + 3] : S.E[4 * e + 2];
      if ((!allowed || (*allowed)[t2]) && !tri_done[t2]) stack.push_back(t2);
    }
  }
}

struct Node {
  std::vector<int> rwg;  // edge indices, ascending
  std::set<int> boundary, tri;
  V3 centroid{0, 0, 0}, n{0, 0, 0};
  double radius = 0, area = 0;
  std::vector<std::unique_ptr<Node>> sub;
};

// Area-weighted centroid and mean normal of the segment's triangles, and the
// largest distance of its edges' endpoints from that centroid.
void segment_info(const Surface &S, const Segment &s, Node &out) {
  V3 c{0., 0., 0.}, n{0., 0., 0.};
  double total = 0.;
  for (int t : s.tri) {
    const double a = S.area(t);
    V3 p{0., 0., 0.};
    p = p + S.P[S.tv[3 * t]];
    p = p + S.P[S.tv[3 * t + 1]];
    p = p + S.P[S.tv[3 * t + 2]];
    p = (kOneThird * a) * p;
    c = c + p;
    n = n + a * S.normal(t);
    total += a;
  }
  c = (1. / total) * c;
  n = (1. / total) * n;
  if (norm2(n) < 1e-7)
    n = {0., 0., 1.};
  else
    normalize(n);
  double r2max = 0.;
  for (int e : s.edge)
    for (int k = 0; k < 2; k++) {
      const double r2 = norm2(S.P[S.E[4 * e + k]] - c);
      if (r2 > r2max) r2max = r2;
    }
  out.centroid = c;
  out.n = n;
  out.radius = std::sqrt(r2max);
  out.area = total;
}

// z component of the rotation by angle psi about the axis at polar angle phi,
// azimuth theta, applied to v.
struct Plane {
  double x, y, z, s, c, t;
  Plane(double phi, double theta, double psi) {
    const double sp = std::sin(phi), cp = std::cos(phi), st = std::sin(theta), ct = std::cos(theta);
    x = sp * ct;
    y = sp * st;
    z = cp;
    s = std::sin(psi);
    c = std::cos(psi);
    t = 1. - c;
  }
  double height(const V3 &v) const {
    return (t * x * z + s * y) * v.x + (t * y * z - s * x) * v.y + (t * z * z + c) * v.z;
  }
};

inline bool crosses(double z1, double z2, double small) {
  return (z1 <= small && z2 >= -small) || (z1 >= -small && z2 <= small);
}

inline double sign_of(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Length of the node's edges that the plane crosses, along the plane, as a
// function of the step alpha along the search line p + alpha xi.
struct CutLength {
  const Surface &S;
  const Node &node;
  double p[3], xi[3];
  double operator()(double alpha) const {
    const Plane R(p[0] + alpha * xi[0], p[1] + alpha * xi[1], p[2] + alpha * xi[2]);
    double length = 0.;
    for (int e : node.rwg) {
      const double z1 = R.height(S.P[S.E[4 * e]] - node.centroid);
      const double z2 = R.height(S.P[S.E[4 * e + 1]] - node.centroid);
      if (crosses(z1, z2, 1e-6 * S.len[e])) length += std::sqrt(S.len[e] * S.len[e] - (z1 - z2) * (z1 - z2));
    }
    return length;
  }
};

// Bracket a minimum along the line from a = 0, b = 1 (golden-section growth
// with parabolic extrapolation).
void bracket(const CutLength &f, double &ax, double &bx, double &cx, double &fa, double &fb, double &fc) {
  const double gold = 1.618034, glimit = 100.0, tiny = 1.0e-20;
  fa = f(ax);
  fb = f(bx);
  if (fb > fa) {
    std::swap(ax, bx);
    std::swap(fa, fb);
  }
  cx = bx + gold * (bx - ax);
  fc = f(cx);
  while (fb > fc) {
    const double r = (bx - ax) * (fb - fc);
    const double q = (bx - cx) * (fb - fa);
    double u = bx - ((bx - cx) * q - (bx - ax) * r) / (2.0 * sign_of(std::max(std::fabs(q - r), tiny), q - r));
    const double ulim = bx + glimit * (cx - bx);
    double fu;
    if ((bx - u) * (u - cx) > 0.0) {
      fu = f(u);
      if (fu < fc) {
        ax = bx;
        bx = u;
        fa = fb;
        fb = fu;
        return;
      } else if (fu > fb) {
        cx = u;
        fc = fu;
        return;
      }
      u = cx + gold * (cx - bx);
      fu = f(u);
    } else if ((cx - u) * (u - ulim) > 0.0) {
      fu = f(u);
      if (fu < fc) {
        bx = cx;
        cx = u;
        u = cx + gold * (cx - bx);
        fb = fc;
        fc = fu;
        fu = f(u);
      }
    } else if ((u - ulim) * (ulim - cx) >= 0.0) {
      u = ulim;
      fu = f(u);
    } else {
      u = cx + gold * (cx - bx);
      fu = f(u);
    }
    ax = bx;
    bx = cx;
    cx = u;
    fa = fb;
    fb = fc;
    fc = fu;
  }
}

// Brent's minimisation inside the bracket (parabolic steps, golden-section
// fallback); returns the minimum, its abscissa in xmin.
double brent_min(const CutLength &f, double ax, double bx, double cx, double tol, double &xmin) {
  const double cgold = 0.3819660, zeps = 1.0e-10;
  double a = ax < cx ? ax : cx, b = ax > cx ? ax : cx;
  double d = 0., e = 0., x, w, v, fx, fw, fv;
  x = w = v = bx;
  fw = fv = fx = f(x);
  for (int iter = 1; iter <= 100; iter++) {
    const double xm = 0.5 * (a + b);
    const double tol1 = tol * std::fabs(x) + zeps, tol2 = 2.0 * tol1;
    if (std::fabs(x - xm) <= (tol2 - 0.5 * (b - a))) {
      xmin = x;
      return fx;
    }
    if (std::fabs(e) > tol1) {
      const double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double p = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) p = -p;
      q = std::fabs(q);
      const double etemp = e;
      e = d;
      if (std::fabs(p) >= std::fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
        e = x >= xm ? a - x : b - x;
        d = cgold * e;
      } else {
        d = p / q;
        const double u = x + d;
        if (u - a < tol2 || b - u < tol2) d = sign_of(tol1, xm - x);
      }
    } else {
      e = x >= xm ? a - x : b - x;
      d = cgold * e;
    }
    const double u = std::fabs(d) >= tol1 ? x + d : x + sign_of(tol1, d);
    const double fu = f(u);
    if (fu <= fx) {
      if (u >= x)
        a = x;
      else
        b = x;
      v = w;
      w = x;
      x = u;
      fv = fw;
      fw = fx;
      fx = fu;
    } else {
      if (u < x)
        a = u;
      else
        b = u;
      if (fu <= fw || w == x) {
        v = w;
        w = u;
        fv = fw;
        fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u;
        fv = fu;
      }
    }
  }
  throw std::runtime_error("plane search: line minimisation did not converge in 100 steps");
}

// Minimise along direction xi from p: p moves to the minimum, xi becomes the
// step taken.  The function's line (f.p, f.xi) is set to (p, xi).
double line_min(CutLength &f, double *p, double *xi, double tol) {
  for (int j = 0; j < 3; j++) {
    f.p[j] = p[j];
    f.xi[j] = xi[j];
  }
  double ax = 0.0, xx = 1.0, bx, fa, fx, fb, xmin;
  bracket(f, ax, xx, bx, fa, fx, fb);
  const double fret = brent_min(f, ax, xx, bx, tol, xmin);
  for (int j = 0; j < 3; j++) {
    xi[j] *= xmin;
    p[j] += xi[j];
  }
  return fret;
}

// Powell's direction-set minimisation of the cut length over the three plane
// angles.  dir[j][i] is component j of direction i; the directions carry
// over between calls, as the reference passes the same array twice.
double powell(CutLength &f, double *p, double dir[3][3], double ftol, double lintol) {
  double pt[3], ptt[3], xit[3];
  for (int j = 0; j < 3; j++) {
    f.p[j] = p[j];
    f.xi[j] = dir[j][0];
  }
  double fret = f(0.);
  for (int j = 0; j < 3; j++) pt[j] = p[j];
  for (int iter = 1;; ++iter) {
    const double fp = fret;
    int ibig = 0;
    double del = 0.0;
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) xit[j] = dir[j][i];
      const double fptt = fret;
      fret = line_min(f, p, xit, lintol);
      if (std::fabs(fptt - fret) > del) {
        del = std::fabs(fptt - fret);
        ibig = i;
      }
    }
    if (2.0 * std::fabs(fp - fret) <= ftol * (std::fabs(fp) + std::fabs(fret))) return fret;
    if (iter == 200) throw std::runtime_error("plane search: Powell did not converge in 200 iterations");
    for (int j = 0; j < 3; j++) {
      ptt[j] = 2.0 * p[j] - pt[j];
      xit[j] = p[j] - pt[j];
      pt[j] = p[j];
      f.p[j] = ptt[j];
    }
    const double fptt = f(0.);
    if (fptt < fp) {
      const double a = fp - fret - del, b = fp - fptt;
      const double t = 2.0 * (fp - 2.0 * fret + fptt) * (a * a) - del * (b * b);
      if (t < 0.0) {
        fret = line_min(f, p, xit, lintol);
        for (int j = 0; j < 3; j++) {
          dir[j][ibig] = dir[j][2];
          dir[j][2] = xit[j];
        }
      }
    }
  }
}

class TreeBuilder {
 public:
  TreeBuilder(const Surface &S, int min_size, double min_radius)
      : S(S), min_size(min_size), min_radius(min_radius), less{S.P.data()} {}

  // Pick the cutting plane: start from the plane that holds the mean
  // normal, minimise the cut length from there and from the same plane
  // turned a quarter, keep the better.
  void partition(Node &node) {
    const V3 &n = node.n;
    const V3 xy{n.x, n.y, 0.};
    const double nxy = std::sqrt(norm2(xy));
    const double psi = (n.z == 0. && nxy == 0.) ? 0. : std::atan2(n.z, nxy);
    V3 v = cross(n, xy);
    const double nv = std::sqrt(norm2(v));
    double phi, theta;
    if (nv < 1e-9) {
      phi = kHalfPi;
      theta = 0.;
    } else {
      v = (1. / nv) * v;
      phi = std::acos(v.z);
      theta = (v.y == 0. && v.x == 0.) ? 0. : std::atan2(v.y, v.x);
    }
    double best1[3] = {phi, theta, psi}, best2[3];
    double dir[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    CutLength f{S, node, {0, 0, 0}, {0, 0, 0}};
    const double min1 = powell(f, best1, dir, 1e-3, 1e-3);
    best2[0] = best1[0];
    best2[1] = best1[1];
    best2[2] = best1[2] + kHalfPi;
    const double min2 = powell(f, best2, dir, 1e-3, 1e-3);
    split(node, min1 < min2 ? best1 : best2);
  }

 private:
  const Surface &S;
  const int min_size;
  const double min_radius;
  const PointLess less;

  bool on_boundary(const Node &node, int p) const {
    for (int e : S.eap[p])
      if (node.boundary.count(e)) return true;
    return false;
  }

  // Cut the node along the plane: walk the crossed edges into new boundary
  // paths, flood-fill the node's triangles into segments between the old and
  // new boundaries, and make each segment a child when all are large enough.
  void split(Node &node, const double *angles) {
    const Plane R(angles[0], angles[1], angles[2]);
    std::vector<double> height(S.nv, 0.);
    PointSet cut(less), on_b(less), off_b(less);
    for (int e : node.rwg) {
      const int p1 = S.E[4 * e], p2 = S.E[4 * e + 1];
      const double z1 = R.height(S.P[p1] - node.centroid), z2 = R.height(S.P[p2] - node.centroid);
      height[p1] = z1;
      height[p2] = z2;
      if (!crosses(z1, z2, 1e-6 * S.len[e])) continue;
      (on_boundary(node, p1) ? on_b : off_b).insert(p1);
      (on_boundary(node, p2) ? on_b : off_b).insert(p2);
      cut.insert(p1);
      cut.insert(p2);
    }
    // walks start at boundary points, then at the others
    std::vector<int> starts(on_b.begin(), on_b.end());
    starts.insert(starts.end(), off_b.begin(), off_b.end());

    std::set<int> fresh;       // new boundary edges
    PointSet fresh_pts(less);  // points of the new boundary
    bool reached = false;
    for (int p : starts) {
      bool skip = false;
      for (int e : S.eap[p])
        if (fresh_pts.count(S.other(e, p))) {
          skip = true;
          break;
        }
      if (skip) continue;
      V3 heading{0., 0., 0.};
      int prev = -1;
      for (;;) {
        // next edge: to a crossed point, near the plane, straight ahead
        double best = DBL_MAX;
        int be = -1, bq = -1;
        const bool pb = on_b.count(p) != 0;
        for (int e : S.eap[p]) {
          if (e == prev) continue;
          const int q = S.other(e, p);
          if (pb && on_b.count(q)) continue;
          if (!cut.count(q)) continue;
          double ahead = 0.;
          if (prev >= 0) ahead = norm2((S.P[p] + S.len[e] * heading) - S.P[q]);
          const double merit = (ahead + height[q] * height[q]) / (S.len[e] * S.len[e]);
          if (merit < best) {
            best = merit;
            be = e;
            bq = q;
          }
        }
        if (bq < 0) break;
        fresh.insert(be);
        fresh_pts.insert(p);
        on_b.insert(p);
        if (on_b.count(bq)) {
          fresh_pts.insert(bq);
          reached = true;
          break;
        }
        heading = (1. / S.len[be]) * (S.P[bq] - S.P[p]);
        prev = be;
        p = bq;
      }
    }

    if (reached) {
      // drop new boundary edges with an end that meets no other boundary edge
      for (;;) {
        std::vector<int> loose;
        for (int e : fresh) {
          for (int k = 0; k < 2; k++) {
            bool joined = false;
            for (int e1 : S.eap[S.E[4 * e + k]])
              if (e1 != e && (fresh.count(e1) || node.boundary.count(e1))) {
                joined = true;
                break;
              }
            if (!joined) {
              loose.push_back(e);
              break;
            }
          }
        }
        if (loose.empty()) break;
        for (int e : loose) fresh.erase(e);
      }
    }
    fresh.insert(node.boundary.begin(), node.boundary.end());

    std::vector<char> is_b(S.ne, 0), allowed(S.nt, 0), tri_done(S.nt, 0), edge_done(S.ne, 0);
    for (int e : fresh) is_b[e] = 1;
    for (int t : node.tri) allowed[t] = 1;
    std::vector<std::unique_ptr<Segment>> pool;
    std::set<Segment *, SegmentLess> segments;
    for (int t : node.tri) {
      if (tri_done[t]) continue;
      pool.emplace_back(new Segment);
      traverse(S, *pool.back(), t, tri_done, edge_done, is_b, &allowed);
      segments.insert(pool.back().get());
    }

    bool split = false;
    if (segments.size() > 1)
      for (Segment *s : segments) {
        split = true;
        if ((int)s->edge.size() < min_size) {
          split = false;
          break;
        }
        Node probe;
        segment_info(S, *s, probe);
        if (probe.radius < min_radius) {
          split = false;
          break;
        }
      }
    if (!split) return;
    for (Segment *s : segments) {
      std::unique_ptr<Node> child(new Node);
      for (int e : node.rwg)
        if (s->edge.count(e)) child->rwg.push_back(e);
      child->boundary = s->boundary;
      child->tri = s->tri;
      segment_info(S, *s, *child);
      partition(*child);
      node.sub.push_back(std::move(child));
    }
  }
};

void leaves(const Node &n, int32_t *order, int &k, std::vector<int32_t> &sizes) {
  if (n.sub.empty()) {
    for (int e : n.rwg) order[k++] = e;
    sizes.push_back((int32_t)n.rwg.size());
    return;
  }
  for (const auto &c : n.sub) leaves(*c, order, k, sizes);
}

}  // namespace

int rwg_tree_order(int nv, const double *xyz, int nt, const int32_t *tri, int min_size, double min_radius,
                   int32_t *order, int &n, std::vector<int32_t> &leaf_sizes) {
  Surface S;
  S.nv = nv;
  S.nt = nt;
  S.P.resize(nv);
  for (int i = 0; i < nv; i++) S.P[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  S.tv.assign(tri, tri + 3 * (size_t)nt);
  S.ne = build_edges(nv, nt, tri, S.E);
  if (S.ne < 0) return -1;
  if (nt == 0 || 2 * S.ne != 3 * nt) return -2;
  // each triangle's side opposite its third vertex
  S.side.assign(3 * (size_t)nt, -1);
  for (int e = 0; e < S.ne; e++) {
    const int a = S.E[4 * e], b = S.E[4 * e + 1];
    for (int k = 2; k < 4; k++) {
      const int t = S.E[4 * e + k];
      const int32_t *v = &S.tv[3 * t];
      const int slot = (v[0] != a && v[0] != b) ? 1 : (v[1] != a && v[1] != b) ? 2 : 0;
      S.side[3 * t + slot] = e;
    }
  }
  for (int s : S.side)
    if (s < 0) return -2;
  S.len.resize(S.ne);
  for (int e = 0; e < S.ne; e++) S.len[e] = std::sqrt(norm2(S.P[S.E[4 * e + 1]] - S.P[S.E[4 * e]]));

  // keep_first_segment: the reference's mesher splits the surface at edges
  // sharper than 2 rad and keeps only the part that holds triangle 0 (it
  // inserts each segment before filling it, and its segment order calls an
  // empty segment equal to any other, so later segments are never added).
  std::vector<char> sharp(S.ne, 0), tri_done(nt, 0), edge_done(S.ne, 0);
  for (int e = 0; e < S.ne; e++) sharp[e] = S.sharpness(e) > kSharp;
  Segment first;
  traverse(S, first, 0, tri_done, edge_done, sharp, nullptr);
  S.eap.assign(nv, std::vector<int>());
  for (int e : first.edge) {
    S.eap[S.E[4 * e]].push_back(e);
    S.eap[S.E[4 * e + 1]].push_back(e);
  }

  Node root;
  root.rwg.assign(first.edge.begin(), first.edge.end());
  root.boundary = first.boundary;
  root.tri = first.tri;
  segment_info(S, first, root);
  try {
    TreeBuilder(S, min_size, min_radius).partition(root);
  } catch (const std::exception &) {
    return -3;
  }
  int k = 0;
  leaf_sizes.clear();
  leaves(root, order, k, leaf_sizes);
  n = k;
  return 0;
}

}  // namespace bemgeo
