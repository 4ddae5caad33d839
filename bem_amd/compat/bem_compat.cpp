// Reference-shaped C++ API over libbem_hip (see bem_compat.h).
#include "bem_compat.h"
#include "x3d.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>

namespace bem {
bem_ctx *context() {
  static bem_ctx *ctx = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    bem_cfg cfg = {0, 0, 0, 0};
    const char *dev = std::getenv("BEM_DEVICE");
    if (dev) cfg.device = std::atoi(dev);
    int rc = bem_ctx_create(&ctx, &cfg);
    if (rc) throw Error(rc, "bem_ctx_create failed: no HIP device");
  });
  return ctx;
}
void check(int rc) {
  if (rc) throw Error(rc, bem_last_error(context()));
}
}  // namespace bem

// src/Options.cpp:23-26 (only the far thresholds act on this path; the
// library applies the same constants)
// The reference's drivers call this first (src/transfer.cpp:5) so that its
// reference-counted matrix storage may be copied; the storage here is the
// library's device memory, and the device context is created on first use.
void initReferenceCountingPointerLock() {}

Options::Options() : farThresh6(12.0), farThresh10(8.0), farThresh15(4.0), farThresh21(2.0) {}

GreensFunction::GreensFunction(bool ext) : bExternal(ext) {}
GreensFunction::GreensFunction(Complex e, Complex m, Complex w, bool ext) : bExternal(ext) { init(e, m, w); }

// src/GreensFunction.cpp:17-30
void GreensFunction::init(Complex e, Complex m, Complex w) {
  eps = e;
  mu = m;
  omega = w;
  ksquared = eps * mu * omega * omega;
  k = std::sqrt(ksquared);
  Complex I(0., 1.);
  ik = I * k;
  k2 = -0.5 * ksquared;
  normk = std::sqrt(std::norm(k));
  Zeps = 1. / (I * omega * eps);
  Zmu = 1. / (I * omega * mu);
}

RWGTree::RWGTree(const std::vector<double> &xyz, const std::vector<int> &tri, Dielectric *d, const Real s[3])
    : rwgs(0), subs(0), g(new GreensFunction(false)), diel(d), reference_order(false), body(nullptr),
      owns_g_(true) {
  sub[0] = sub[1] = nullptr;
  for (int c = 0; c < 3; c++) shift[c] = s ? s[c] : 0.0;
  const int nv = (int)(xyz.size() / 3), nt = (int)(tri.size() / 3);
  bem::check(bem_body_upload(bem::context(), xyz.data(), nv, tri.data(), nt, &body));
  int ne = 0;
  bem::check(bem_body_edges(body, nullptr, &ne));
  rwgs = ne;
  // the tree is built on the translated body, as the reference meshes it
  std::vector<double> moved(xyz);
  for (size_t i = 0; i < moved.size(); i++) moved[i] += shift[i % 3];
  std::vector<int32_t> ord(ne), lv(ne);
  int n = 0, nl = 0;
  if (bem_rwg_tree(moved.data(), nv, tri.data(), nt, 32, 0.0, ord.data(), &n, lv.data(), &nl) == 0) {
    order.assign(ord.begin(), ord.begin() + n);
    leaves.assign(lv.begin(), lv.begin() + nl);
    reference_order = n == ne;
  }
  std::vector<char> held(ne, 0);
  for (int e : order) held[e] = 1;
  for (int e = 0; e < ne; e++)
    if (!held[e]) order.push_back(e);
}

RWGTree::RWGTree(RWGTree *s, RWGTree *t)
    : rwgs(s->rwgs + t->rwgs), subs(2), g(nullptr), diel(nullptr),
      reference_order(s->reference_order && t->reference_order), body(nullptr), owns_g_(false) {
  sub[0] = s;
  sub[1] = t;
  shift[0] = shift[1] = shift[2] = 0.0;
  order = s->order;
  for (int e : t->order) order.push_back(e + s->rwgs);
  leaves = s->leaves;
  leaves.insert(leaves.end(), t->leaves.begin(), t->leaves.end());
}

RWGTree::~RWGTree() {
  if (body) bem_body_free(body);
  if (owns_g_) delete g;
}

static void collect(RWGTree *t, std::vector<bem_body *> &b, std::vector<double> &sh) {
  if (t->subs == 2) {
    collect(t->sub[0], b, sh);
    collect(t->sub[1], b, sh);
    return;
  }
  b.push_back(t->body);
  sh.insert(sh.end(), t->shift, t->shift + 3);
}

// Rows and columns from the library's mesh order into the tree's: unknown
// 2i + a of the tree (a = 0 for J, 1 for M) is unknown 2 order[i] + a of the
// mesh.  Rows are gathered one column at a time, columns are moved along the
// permutation's cycles, so one column of scratch suffices.
static void to_tree_order(ComplexDenseMatrix &Z, const std::vector<int> &order) {
  const int n = Z.n;
  std::vector<int> src(n);
  bool identity = true;
  for (int i = 0; i < n / 2; i++) {
    src[2 * i] = 2 * order[i];
    src[2 * i + 1] = 2 * order[i] + 1;
    identity = identity && order[i] == i;
  }
  if (identity) return;
  std::vector<Complex> col(n);
  for (int j = 0; j < n; j++) {
    Complex *c = &Z(0, j);
    for (int i = 0; i < n; i++) col[i] = c[src[i]];
    std::copy(col.begin(), col.end(), c);
  }
  std::vector<char> done(n, 0);
  for (int j0 = 0; j0 < n; j0++) {
    if (done[j0] || src[j0] == j0) continue;
    std::copy(&Z(0, j0), &Z(0, j0) + n, col.begin());
    int j = j0;
    for (;;) {
      done[j] = 1;
      const int from = src[j];
      if (from == j0) {
        std::copy(col.begin(), col.end(), &Z(0, j));
        break;
      }
      std::copy(&Z(0, from), &Z(0, from) + n, &Z(0, j));
      j = from;
    }
  }
}

// Zts (src/Impedance.cpp:388-413), rows and columns in the tree's order.  The
// hot path calls it with t == s (the self-interaction of a tree); for t != s
// with both inside one two-body tree, the (t, s) block of the union's matrix
// is returned.
ComplexDenseMatrix *Zts(RWGTree *t, RWGTree *s, GreensFunction *g, Options &) {
  if (t == s) {
    std::vector<bem_body *> b;
    std::vector<double> sh;
    collect(t, b, sh);
    ComplexDenseMatrix *Z = new ComplexDenseMatrix(2 * t->rwgs, 2 * t->rwgs);
    double e[2] = {g->eps.real(), g->eps.imag()}, m[2] = {g->mu.real(), g->mu.imag()};
    try {
      bem::check(bem_assemble(bem::context(), (int)b.size(), b.data(), sh.data(), e, m, g->omega.real(), Z->p.data(),
                              Z->lda, 0));
    } catch (...) {
      delete Z;
      throw;
    }
    to_tree_order(*Z, t->order);
    return Z;
  }
  if (t->subs || s->subs) throw bem::Error(BEM_E_UNSUPPORTED, "Zts(t, s): t and s must be single bodies");
  RWGTree top(t, s);
  Options o;
  ComplexDenseMatrix *U = Zts(&top, &top, g, o);
  ComplexDenseMatrix *Z = new ComplexDenseMatrix(2 * t->rwgs, 2 * s->rwgs);
  for (int j = 0; j < Z->n; j++)
    for (int i = 0; i < Z->m; i++) (*Z)(i, j) = (*U)(i, 2 * t->rwgs + j);
  top.sub[0] = top.sub[1] = nullptr;
  delete U;
  return Z;
}

// Zdense (src/Impedance.cpp:415-445): Zts with gExt, plus t->g when t and s
// share a medium.
void Zdense(RWGTree *t, RWGTree *s, ComplexDenseMatrix *dense, GreensFunction *gExt, Options &o) {
  ComplexDenseMatrix *Z = Zts(t, s, gExt, o);
  if (t->g && t->g == s->g) {
    ComplexDenseMatrix *Zi = Zts(t, s, t->g, o);
    for (size_t k = 0; k < Z->p.size(); k++) Z->p[k] += Zi->p[k];
    delete Zi;
  }
  *dense = *Z;
  delete Z;
}

// CalcHeatTransfer (src/HeatTransfer.cpp:8-59): media from the bodies'
// Dielectrics (temperatures set first, as the reference does), the gap from
// body 1's translation.
Real CalcHeatTransfer(Real a, Real eV, Real T1, Real T2, RWGTree *top, RWGTree *s, RWGTree *t, Options &) {
  if (!top || !s || !t || !s->body || !t->body) throw bem::Error(BEM_E_INVALID, "CalcHeatTransfer: bad trees");
  s->diel->setTemperature(T1);
  t->diel->setTemperature(T2);
  Complex eps1 = s->diel->eps(eV), mu1 = s->diel->mu(eV);
  Complex eps2 = t->diel->eps(eV), mu2 = t->diel->mu(eV);
  bem_request q = {};
  q.eV = eV;
  q.a = a;
  q.T1 = T1;
  q.T2 = T2;
  for (int c = 0; c < 3; c++) q.gap[c] = s->shift[c] - t->shift[c];
  q.eps1[0] = eps1.real(); q.eps1[1] = eps1.imag();
  q.mu1[0] = mu1.real(); q.mu1[1] = mu1.imag();
  q.eps2[0] = eps2.real(); q.eps2[1] = eps2.imag();
  q.mu2[0] = mu2.real(); q.mu2[1] = mu2.imag();
  q.precision = BEM_PREC_FP64;
  bem_result r;
  bem::check(bem_spectral_power(bem::context(), s->body, t->body, &q, &r));
  Real omega = r.omega;
  s->g->init(eps1, mu1, Complex(omega, 0.0));
  t->g->init(eps2, mu2, Complex(omega, 0.0));
  return r.P;
}

// src/ShapeParser.cpp:17-63
void parseShapes(RWGTree *&top, RWGTree *&s, RWGTree *&t, const char *body1x3d, const char *body2x3d,
                 const Vector3 &gap, Options &) {
  X3DShape b1, b2;
  read_x3d(body1x3d, b1);
  read_x3d(body2x3d, b2);
  for (const X3DShape *b : {&b1, &b2})
    if (!b->dielectric) throw bem::Error(BEM_E_INVALID, "parseShapes: a body file has no <Dielectric> element");
  auto diel = [](const X3DShape &b) {
    Params p;
    p.params = b.params;
    return Dielectrics::getDielectric(b.material, p);
  };
  const Real shift[3] = {gap.x, gap.y, gap.z};
  s = new RWGTree(b1.coord, b1.face, diel(b1), shift);
  t = new RWGTree(b2.coord, b2.face, diel(b2));
  top = new RWGTree(s, t);
}

extern "C" int bemc_tree_zts(const double *xyz1, int nv1, const int *tri1, int nt1, const double *shift1,
                             const double *xyz2, int nv2, const int *tri2, int nt2, double omega, double *Z,
                             int *order, int *reference_order) {
  try {
    ConstantDielectric vac(1.0, 1.0);
    RWGTree s(std::vector<double>(xyz1, xyz1 + 3 * (size_t)nv1), std::vector<int>(tri1, tri1 + 3 * (size_t)nt1), &vac,
              shift1);
    RWGTree t(std::vector<double>(xyz2, xyz2 + 3 * (size_t)nv2), std::vector<int>(tri2, tri2 + 3 * (size_t)nt2), &vac);
    RWGTree top(&s, &t);
    GreensFunction g(Complex(1.0, 0.0), Complex(1.0, 0.0), Complex(omega, 0.0), true);
    Options o;
    ComplexDenseMatrix *W = Zts(&top, &top, &g, o);
    std::copy(W->p.begin(), W->p.end(), reinterpret_cast<Complex *>(Z));
    delete W;
    std::copy(top.order.begin(), top.order.end(), order);
    *reference_order = top.reference_order ? 1 : 0;
    top.sub[0] = top.sub[1] = nullptr;
    return 0;
  } catch (const bem::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return e.code;
  }
}

// src/FrequencySweep.cpp:4-17
FrequencySweep::FrequencySweep(char *filename) {
  FILE *file = std::fopen(filename, "r");
  if (file) {
    char line[128];
    while (std::fgets(line, sizeof line, file)) f.push_back(std::atof(line));
    std::fclose(file);
  } else {
    std::perror(filename);
  }
}
