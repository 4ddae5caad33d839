// Reference-shaped C++ API of the hot path, implemented over libbem_hip's C ABI.
//
// A driver written against the reference's headers for this path
//   src/Impedance.h:20-21      Zdense / Zts
//   src/HeatTransfer.h:33      CalcHeatTransfer
//   src/FrequencySweep.h:9-13  FrequencySweep
//   src/GreensFunction.h:7-28  GreensFunction (init, eps, mu, omega, k, Zeps, Zmu)
//   src/util/Dielectric.h:7-13 Dielectric (eps(eV), mu(eV), setTemperature)
//   src/Options.h              Options (farThresh6/10/15/21)
//   src/util/DenseMatrix.h     ComplexDenseMatrix (column-major, operator()(i,j), lda)
//   src/ShapeParser.h:47       parseShapes (X3D files, x3d.h)
//   src/util/Dielectrics.h     Dielectrics::getDielectric (material library)
// compiles against this header and links libbem_hip.so instead.  RWGTree is
// built from vertex/triangle arrays or by parseShapes; a two-body tree is
// RWGTree(s, t) with s, t the bodies, as the reference's top tree holds
// sub[0] = body 1 and sub[1] = body 2.
//
// Error behaviour follows the reference: no return codes on these calls.
// Where the reference would abort (dimension mismatch) or silently continue
// (LAPACK info ignored), this layer throws bem::Error carrying the C ABI's
// message, which an unchanged driver turns into an abort.
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bem_hip.h"

typedef double Real;
typedef std::complex<double> Complex;

namespace bem {
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
// Process-wide device context used by the reference-shaped calls.
bem_ctx *context();
// Throws Error with the context's last message when rc is not 0.
void check(int rc);
}  // namespace bem

class Options {
 public:
  Options();
  Real farThresh6, farThresh10, farThresh15, farThresh21;
};

class Dielectric {
 public:
  Dielectric() {}
  virtual ~Dielectric() {}
  virtual Complex eps(Real eV) = 0;
  virtual Complex mu(Real eV) = 0;
  virtual void setTemperature(Real T) {}
};

class ConstantDielectric : public Dielectric {
 public:
  ConstantDielectric(Complex e, Complex m) : e_(e), m_(m) {}
  Complex eps(Real) { return e_; }
  Complex mu(Real) { return m_; }

 private:
  Complex e_, m_;
};

// Material library (src/util/Dielectrics.cpp:15-63 and the models it
// registers): "SiN" (tabulated n, k, src/util/SiN.cpp:3352-3358), "SiC"
// (temperature-dependent Lorentz, src/util/SiC.cpp:4-19), "Al" (Lorentz,
// src/util/Al.cpp:4-12), "Au" (Drude, src/util/Au.cpp:4-12),
// "SiliconDoped" with carrier density N (src/util/SiliconDoped.cpp), and for
// any other name a constant medium from the complex parameters eps and mu
// (default 1).  One instance per (name, parameters), owned by the registry.
class SiN : public Dielectric {
 public:
  Complex eps(Real eV);
  Complex mu(Real eV);
};
class SiC : public Dielectric {
 public:
  SiC();
  void setTemperature(Real T);
  Complex eps(Real eV);
  Complex mu(Real eV);
  Real epsinf, gamma, wL, wT;
};
class Al : public Dielectric {
 public:
  Complex eps(Real eV);
  Complex mu(Real eV);
};
class Au : public Dielectric {
 public:
  Complex eps(Real eV);
  Complex mu(Real eV);
};
class SiliconDoped : public Dielectric {
 public:
  explicit SiliconDoped(Real N) : N(N) {}
  Complex eps(Real eV);
  Complex mu(Real eV);
  Real N;   // carriers per m^3: < 0 n-type, > 0 p-type
};
class Params {
 public:
  std::vector<std::pair<std::string, std::string>> params;
  const std::string *find(const std::string &k) const;
};
namespace Dielectrics {
Dielectric *getDielectric(const std::string &name, const Params &params = Params());
}
// "re+im" or "re-im" with an optional trailing i or j; "re" alone is real.
// Throws bem::Error(BEM_E_INVALID) on anything else.
Complex parseComplex(const std::string &s);

class GreensFunction {
 public:
  explicit GreensFunction(bool bExternal);
  GreensFunction(Complex eps, Complex mu, Complex omega, bool bExternal = false);
  void init(Complex eps, Complex mu, Complex omega);
  bool isExternal() { return bExternal; }
  bool bExternal;
  Complex Zeps, Zmu;
  Real normk;
  Complex ik, k2, k, ksquared, eps, mu, omega;
};

class ComplexDenseMatrix {
 public:
  ComplexDenseMatrix(int m, int n) : m(m), n(n), lda(m), p((size_t)m * n) {}
  Complex &operator()(int i, int j) { return p[(size_t)j * lda + i]; }
  const Complex &operator()(int i, int j) const { return p[(size_t)j * lda + i]; }
  int m, n, lda;
  std::vector<Complex> p;
};

class RWGTree {
 public:
  // One body: closed triangle mesh (xyz nv x 3, tri nt x 3), its medium and a
  // translation (the reference applies the gap to body 1's shape).  The
  // unknowns take the reference's tree order (bem_rwg_tree on the translated
  // mesh, minSegmentSize 32, minRadius 0 as parseShapes passes them).
  RWGTree(const std::vector<double> &xyz, const std::vector<int> &tri, Dielectric *diel,
          const Real shift[3] = nullptr);
  // Two-body top tree over bodies s (body 1) and t (body 2): s's unknowns,
  // then t's (src/RWGTree.cpp:748-761).
  RWGTree(RWGTree *s, RWGTree *t);
  ~RWGTree();
  int rwgs;
  RWGTree *sub[2];
  int subs;
  GreensFunction *g;
  Dielectric *diel;
  // The reference's rwg[] order: order[i] is the mesh edge index
  // (bem_body_edges order, body 1's edges first in a two-body tree) of the
  // RWG at tree position i; Zts rows and columns follow it.  leaves holds
  // the RWG counts of the tree's leaves.  reference_order is false where the
  // reference's tree would not hold every edge (a closed loop of edges
  // sharper than 2 rad: its mesher keeps only the part holding triangle 0)
  // or its plane search fails; the missing edges then follow the tree's in
  // mesh order, so that every unknown stays in the problem.
  std::vector<int> order, leaves;
  bool reference_order;
  // library handles
  bem_body *body;
  Real shift[3];

 private:
  bool owns_g_;
};

struct Vector3 {
  Real x, y, z;
  Vector3(Real x = 0, Real y = 0, Real z = 0) : x(x), y(y), z(z) {}
};

// src/ShapeParser.cpp:17-63: both bodies read from X3D files (x3d.h), body 1
// translated by gap, each with the medium of its <Dielectric> element (a
// file without one is an error), the two-body tree in top.  The trees come
// back through references (the reference's pointer arguments are by value).
// Meshes are used as read: the reference's transfer drivers refine no edge
// (refineEdgeFuncBody returns false, src/HeatTransfer.h:28-31); the trees
// carry the reference's unknown order.
void parseShapes(RWGTree *&top, RWGTree *&s, RWGTree *&t, const char *body1x3d, const char *body2x3d,
                 const Vector3 &gap, Options &options);

// src/util/ReferenceCountingPointer.cpp:10-14: the reference's drivers call
// it before anything else; nothing needs initialising here.
void initReferenceCountingPointerLock();

ComplexDenseMatrix *Zts(RWGTree *t, RWGTree *s, GreensFunction *g, Options &options);
void Zdense(RWGTree *t, RWGTree *s, ComplexDenseMatrix *dense, GreensFunction *gExt, Options &options);
Real CalcHeatTransfer(Real a, Real eV, Real T1, Real T2, RWGTree *top, RWGTree *s, RWGTree *t, Options &options);

// C entry point for callers outside C++ (bindings, tests); needs a device.
// Two bodies as parseShapes makes them (body 1 translated by shift1), the
// exterior Zts(top, top, g) at dimensionless omega (vacuum) written to Z
// (2E x 2E complex128 column-major, E = total RWGs) in the tree's order, and
// the tree's order (mesh RWG index per position, body 1's edges first) to
// order[0..E).  *reference_order is 1 where that order is the reference's.
// Returns 0 or a BEM_E_* code.
extern "C" int bemc_tree_zts(const double *xyz1, int nv1, const int *tri1, int nt1, const double *shift1,
                             const double *xyz2, int nv2, const int *tri2, int nt2, double omega, double *Z,
                             int *order, int *reference_order);

class FrequencySweep {
 public:
  explicit FrequencySweep(char *file);
  std::vector<Real> f;
};
