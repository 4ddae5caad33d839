/*
 * bem_oracle.c -- CPU restatement of the reference's RWG/PMCHWT impedance
 * assembly, used ONLY as test infrastructure: the parity checker for the HIP
 * library and the timed CPU baseline ("port").  Nothing in the product path
 * links, loads or calls this file.
 *
 * It follows the reference's algorithm function by function, in plain C99
 * (FP64 / complex128), with the same per-point accumulation order as the
 * reference integrands (no algebraic regrouping), so that it is an
 * independent check of the regrouped GPU kernels:
 *
 *   edges / orientation ........ src/ShapeMesh.cpp:856-947, src/ShapeMesh.cpp:712-728
 *   half-RWG precompute ........ src/RWG.cpp:5-39
 *   shared-vertex test ......... src/Triangle.h:23-30 (sharedpoints)
 *   Green's function ........... src/GreensFunction.cpp:17-69
 *   EHNonSmoothIntegral ........ src/Source.cpp:16-86
 *   LKEdgeIntegral ............. src/Impedance.cpp:13-73
 *   inner/outer integrands ..... src/Impedance.cpp:75-274
 *   LKTriangle (tiers, scatter)  src/Impedance.cpp:276-385
 *   Zts ........................ src/Impedance.cpp:388-413
 *   Options far thresholds ..... src/Options.cpp:23-26
 *   dipole right-hand side ..... src/Source.cpp:88-193 (EHSmoothIntegrand,
 *                                DipolePointSource::EHTriangle, Bdense)
 *   flux ....................... src/Flux.cpp:100-219
 *
 * Parity pins (see tests/): every function here is checked against the
 * reference's own Zts compiled from its sources (oracle/_ref, golden fixtures
 * in tests/golden/), and the analytic integrals additionally against
 * brute-force adaptive quadrature, the Green's function against closed forms.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../bem_amd/csrc/quadrature_tables.h"

typedef double complex cplx;

#define ONETHIRD 0.33333333333333333333333333333333
#define ONEOVER4PI 0.079577471545947667884441881686257
#define FAR_THRESH6 12.0
#define FAR_THRESH10 8.0
#define FAR_THRESH15 4.0
#define FAR_THRESH21 2.0

/* ------------------------------------------------------------------ vectors */
typedef struct { double x, y, z; } vec3;
typedef struct { cplx x, y, z; } cvec3;

static inline vec3 v3(double x, double y, double z) { vec3 r = {x, y, z}; return r; }
static inline vec3 vadd(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline vec3 vsub(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline vec3 vscale(double c, vec3 a) { return v3(c * a.x, c * a.y, c * a.z); }
/* cross product (the reference's Vector3 operator*) */
static inline vec3 vcross(vec3 a, vec3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vdot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* squared length (the reference's norm(Vector3)) */
static inline double vnorm2(vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline vec3 vnormalize(vec3 a) {
  double ni = 1. / sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return v3(a.x * ni, a.y * ni, a.z * ni);
}

static inline cvec3 cv3(cplx x, cplx y, cplx z) { cvec3 r = {x, y, z}; return r; }
static inline cvec3 cvzero(void) { return cv3(0, 0, 0); }
static inline cvec3 cvadd(cvec3 a, cvec3 b) { return cv3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline cvec3 cvscale(cplx c, cvec3 a) { return cv3(c * a.x, c * a.y, c * a.z); }
static inline cvec3 cvscale_r(double c, cvec3 a) { return cv3(c * a.x, c * a.y, c * a.z); }
static inline cvec3 cv_from(cplx c, vec3 a) { return cv3(c * a.x, c * a.y, c * a.z); }
static inline cvec3 cv_real(vec3 a) { return cv3(a.x, a.y, a.z); }
/* complex x real cross product */
static inline cvec3 cvcross_r(cvec3 a, vec3 b) {
  return cv3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* dot(real, complex), no conjugation */
static inline cplx cvdot_r(vec3 a, cvec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* ----------------------------------------------------------- mesh / edges */

/* First triangle j != i attached to both vertices (src/ShapeMesh.cpp:712-728).
 * tap: per-vertex triangle lists in ascending triangle order (CSR). */
static int neighbor_triangle(int i, const int *tap_off, const int *tap, int va, int vb) {
  for (int a = tap_off[va]; a < tap_off[va + 1]; a++) {
    int ta = tap[a];
    if (ta == i) continue;
    for (int b = tap_off[vb]; b < tap_off[vb + 1]; b++) {
      int tb = tap[b];
      if (tb != i && ta == tb) return ta;
    }
  }
  return -1;
}

/*
 * Edge table in the reference's creation order (src/ShapeMesh.cpp:856-926):
 * scan triangles in order; for sides (0,1),(1,2),(2,0) create the edge when
 * the neighbour across it exists and has not been scanned yet.  Edge p1->p2
 * follows the creating triangle's winding; t1 (the RWG's positive half) is
 * the creating triangle, t2 the neighbour.
 * edges_out: ne x 4 ints {p1, p2, t1, t2}.  Returns ne, or -1 when max_edges
 * is exceeded.
 */
int oracle_mesh_edges(int nv, int nt, const int *tri, int max_edges, int *edges_out) {
  int *cnt = calloc((size_t)nv + 1, sizeof(int));
  int *tap = malloc(sizeof(int) * 3 * (size_t)nt);
  int *fill = calloc((size_t)nv + 1, sizeof(int));
  unsigned char *touched = calloc((size_t)nt, 1);
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < 3; k++) cnt[tri[3 * i + k] + 1]++;
  for (int v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < 3; k++) {
      int v = tri[3 * i + k];
      tap[cnt[v] + fill[v]++] = i;
    }
  int ne = 0;
  for (int i = 0; i < nt; i++) {
    const int *f = tri + 3 * i;
    const int sides[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int s = 0; s < 3; s++) {
      int va = f[sides[s][0]], vb = f[sides[s][1]];
      int nb = neighbor_triangle(i, cnt, tap, va, vb);
      if (nb != -1 && !touched[nb]) {
        if (ne >= max_edges) { ne = -1; goto done; }
        edges_out[4 * ne + 0] = va;
        edges_out[4 * ne + 1] = vb;
        edges_out[4 * ne + 2] = i;
        edges_out[4 * ne + 3] = nb;
        ne++;
      }
    }
    touched[i] = 1;
  }
done:
  free(cnt); free(tap); free(fill); free(touched);
  return ne;
}

/* ---------------------------------------------------------------- half RWG */
typedef struct {
  int tri;          /* triangle id in the tree */
  int rwg;          /* RWG index in the tree */
  int positive;     /* 1 = Edge::t1 half */
  double elen;      /* RWG edge length */
  vec3 p0, p1, p2, v01, v12, v20, v02, n, m01, m12, m20;
  double normv01i, normv12i, normv20i, Ai;
} halfrwg;

/* src/RWG.cpp:7-39: p0 is the triangle vertex not on the edge, (p0,p1,p2)
 * follow the triangle's own vertex cycle. */
static void make_half(halfrwg *h, const double *xyz, const int *tv, int ea, int eb,
                      int tri, int rwg, int positive, double elen) {
  int k;
  if (tv[0] != ea && tv[0] != eb) k = 0;
  else if (tv[1] != ea && tv[1] != eb) k = 1;
  else k = 2;
  int i0 = tv[k], i1 = tv[(k + 1) % 3], i2 = tv[(k + 2) % 3];
  h->tri = tri; h->rwg = rwg; h->positive = positive; h->elen = elen;
  h->p0 = v3(xyz[3 * i0], xyz[3 * i0 + 1], xyz[3 * i0 + 2]);
  h->p1 = v3(xyz[3 * i1], xyz[3 * i1 + 1], xyz[3 * i1 + 2]);
  h->p2 = v3(xyz[3 * i2], xyz[3 * i2 + 1], xyz[3 * i2 + 2]);
  h->v01 = vsub(h->p1, h->p0);
  h->v12 = vsub(h->p2, h->p1);
  h->v20 = vsub(h->p0, h->p2);
  h->v02 = vsub(h->p2, h->p0);
  h->normv01i = 1. / sqrt(vnorm2(h->v01));
  h->normv12i = 1. / sqrt(vnorm2(h->v12));
  h->normv20i = 1. / sqrt(vnorm2(h->v20));
  vec3 n = vcross(h->v01, h->v02);
  double A2 = sqrt(vnorm2(n));
  h->n = vscale(1. / A2, n);
  h->Ai = 2.0 / A2;
  h->m01 = vnormalize(vcross(h->v01, h->n));
  h->m12 = vnormalize(vcross(h->v12, h->n));
  h->m20 = vnormalize(vcross(h->v20, h->n));
}

/* --------------------------------------------------------- Green function */
typedef struct {
  cplx eps, mu, ksq, k, ik, k2, Zeps, Zmu;
  double omega;
} green;

/* src/GreensFunction.cpp:17-30 */
static void green_init(green *g, cplx eps, cplx mu, double omega) {
  g->eps = eps; g->mu = mu; g->omega = omega;
  g->ksq = eps * mu * omega * omega;
  g->k = csqrt(g->ksq);
  g->ik = I * g->k;
  g->k2 = -0.5 * g->ksq;
  g->Zeps = 1. / (I * omega * eps);
  g->Zmu = 1. / (I * omega * mu);
}

/* src/GreensFunction.cpp:41-52 */
static inline void scalar_gradient(const green *g, vec3 rt, vec3 rs, cplx *gs, cvec3 *gg) {
  double x = rt.x - rs.x, y = rt.y - rs.y, z = rt.z - rs.z;
  double R2 = x * x + y * y + z * z;
  double R = sqrt(R2);
  cplx ikR = g->ik * R;
  *gs = cexp(ikR) / R;
  cplx q = *gs * (1. - ikR) / R2;
  *gg = cv3(q * x, q * y, q * z);
}

/* src/GreensFunction.cpp:54-69 */
static inline void scalar_gradient_smooth(const green *g, vec3 rt, vec3 rs, cplx *gs, cvec3 *gg) {
  double x = rt.x - rs.x, y = rt.y - rs.y, z = rt.z - rs.z;
  double R2 = x * x + y * y + z * z;
  double R = sqrt(R2);
  cplx ikR = g->ik * R;
  cplx e = cexp(ikR);
  *gs = (e - 1.0) / R - g->k2 * R;
  double R3 = R2 * R;
  cplx q = (e * (1.0 - ikR) + g->k2 * R2 - 1.0) / R3;
  *gg = cv3(q * x, q * y, q * z);
}

/* ---------------------------------------------------- analytic integrals */

/* src/Source.cpp:16-86 */
void oracle_eh_nonsmooth(const halfrwg *sh, vec3 rt, double *ISn3h, double *ISn1, double *IS1,
                         vec3 *ILn1, vec3 *IL1, vec3 *IL3) {
  double tol = 1e-6 * sh->elen;
  vec3 p0 = vsub(rt, sh->p0), p1 = vsub(rt, sh->p1), p2 = vsub(rt, sh->p2);
  double R02 = vnorm2(p0), R0 = sqrt(R02);
  double R12 = vnorm2(p1), R1 = sqrt(R12);
  double R22 = vnorm2(p2), R2 = sqrt(R22);
  vec3 a0 = vscale(1. / R0, p0), a1 = vscale(1. / R1, p1), a2 = vscale(1. / R2, p2);
  vec3 n = sh->n;
  double x = 1.0 + vdot(a0, vadd(a1, a2)) + vdot(a1, a2);
  double tp = vdot(a0, vcross(a1, a2));
  double y = sqrt(tp * tp);
  double h = vdot(p0, n);
  *ISn3h = (h < 0. ? -2. : 2.) * atan2(y, x);
  double num, den;
  double s010 = sh->normv01i * vdot(sh->v01, p0);
  double s011 = sh->normv01i * vdot(sh->v01, p1);
  num = (R1 - s011); den = (R0 - s010);
  double ILn1_01 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s121 = sh->normv12i * vdot(sh->v12, p1);
  double s122 = sh->normv12i * vdot(sh->v12, p2);
  num = (R2 - s122); den = (R1 - s121);
  double ILn1_12 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s202 = sh->normv20i * vdot(sh->v20, p2);
  double s200 = sh->normv20i * vdot(sh->v20, p0);
  num = (R0 - s200); den = (R2 - s202);
  double ILn1_20 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double t01 = vdot(sh->m01, p0), t12 = vdot(sh->m12, p1), t20 = vdot(sh->m20, p2);
  *ISn1 = -h * *ISn3h - (t01 * ILn1_01 + t12 * ILn1_12 + t20 * ILn1_20);
  double R02_01 = (R02 - s010 * s010);
  double IL1_01 = 0.5 * (R02_01 * ILn1_01 - (R1 * s011 - R0 * s010));
  double R02_12 = (R12 - s121 * s121);
  double IL1_12 = 0.5 * (R02_12 * ILn1_12 - (R2 * s122 - R1 * s121));
  double R02_20 = (R22 - s202 * s202);
  double IL1_20 = 0.5 * (R02_20 * ILn1_20 - (R0 * s200 - R2 * s202));
  double IL3_01 = 0.75 * R02_01 * IL1_01 - 0.25 * (R12 * R1 * s011 - R02 * R0 * s010);
  double IL3_12 = 0.75 * R02_12 * IL1_12 - 0.25 * (R22 * R2 * s122 - R12 * R1 * s121);
  double IL3_20 = 0.75 * R02_20 * IL1_20 - 0.25 * (R02 * R0 * s200 - R22 * R2 * s202);
  *IS1 = ONETHIRD * (h * h * *ISn1 - (t01 * IL1_01 + t12 * IL1_12 + t20 * IL1_20));
  *ILn1 = vadd(vadd(vscale(ILn1_01, sh->m01), vscale(ILn1_12, sh->m12)), vscale(ILn1_20, sh->m20));
  *IL1 = vadd(vadd(vscale(IL1_01, sh->m01), vscale(IL1_12, sh->m12)), vscale(IL1_20, sh->m20));
  *IL3 = vadd(vadd(vscale(IL3_01, sh->m01), vscale(IL3_12, sh->m12)), vscale(IL3_20, sh->m20));
}

/* src/Impedance.cpp:13-73 */
void oracle_lk_edge(const halfrwg *th, vec3 rs, double *ISn1, vec3 *IL1) {
  vec3 p0 = vsub(rs, th->p0), p1 = vsub(rs, th->p1), p2 = vsub(rs, th->p2);
  double R02 = vnorm2(p0), R0 = sqrt(R02);
  double R12 = vnorm2(p1), R1 = sqrt(R12);
  double R22 = vnorm2(p2), R2 = sqrt(R22);
  double tol = 1e-8 * th->elen;
  vec3 a0 = vscale(1. / R0, p0), a1 = vscale(1. / R1, p1), a2 = vscale(1. / R2, p2);
  vec3 n = th->n;
  double h = vdot(p0, n);
  double x = 1.0 + vdot(a0, vadd(a1, a2)) + vdot(a1, a2);
  double tp = vdot(a0, vcross(a1, a2));
  double y = sqrt(tp * tp);
  double ISn3h = (h < 0 ? -2. : 2.) * atan2(y, x);
  double t01 = vdot(th->m01, p0), t12 = vdot(th->m12, p1), t20 = vdot(th->m20, p2);
  double s010 = th->normv01i * vdot(th->v01, p0);
  double s011 = th->normv01i * vdot(th->v01, p1);
  double num, den;
  num = (R1 - s011); den = (R0 - s010);
  double ILn1_01 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s121 = th->normv12i * vdot(th->v12, p1);
  double s122 = th->normv12i * vdot(th->v12, p2);
  num = (R2 - s122); den = (R1 - s121);
  double ILn1_12 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s202 = th->normv20i * vdot(th->v20, p2);
  double s200 = th->normv20i * vdot(th->v20, p0);
  num = (R0 - s200); den = (R2 - s202);
  double ILn1_20 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  *ISn1 = -h * ISn3h - (t01 * ILn1_01 + t12 * ILn1_12 + t20 * ILn1_20);
  double R02_01 = (R02 - s010 * s010);
  double IL1_01 = 0.5 * (R02_01 * ILn1_01 - (R1 * s011 - R0 * s010));
  double R02_12 = (R12 - s121 * s121);
  double IL1_12 = 0.5 * (R02_12 * ILn1_12 - (R2 * s122 - R1 * s121));
  double R02_20 = (R22 - s202 * s202);
  double IL1_20 = 0.5 * (R02_20 * ILn1_20 - (R0 * s200 - R2 * s202));
  *IL1 = vadd(vadd(vscale(IL1_01, th->m01), vscale(IL1_12, th->m12)), vscale(IL1_20, th->m20));
}

/* ------------------------------------------------------------ quadrature */
static inline vec3 quad_point(vec3 v0, vec3 v1, vec3 v2, double a, double b) {
  /* v0 + a*v1 + b*v2, evaluated left to right as in TriangleQuadrature.h */
  vec3 r = vadd(v0, vscale(a, v1));
  return vadd(r, vscale(b, v2));
}

/* ------------------------------------------------------------ LKTriangle */
typedef struct {
  int n;
  const halfrwg *h[3];
} halflist;

/* Inner integrand accumulators (src/Impedance.cpp:75-139). */
typedef struct {
  cplx G;
  cvec3 Gphi[3];
  cvec3 DGxphi[3];
} inner_acc;

static void inner_integrate(const green *g, const halflist *s, vec3 rt, int off, int npts, int smooth,
                            inner_acc *acc) {
  const halfrwg *sh = s->h[0];
  acc->G = 0;
  for (int j = 0; j < 3; j++) { acc->Gphi[j] = cvzero(); acc->DGxphi[j] = cvzero(); }
  for (int q = 0; q < npts; q++) {
    const double *abw = bem_quad_abw[off + q];
    vec3 v = quad_point(sh->p0, sh->v01, sh->v02, abw[0], abw[1]);
    double w = abw[2];
    cplx gs; cvec3 gg;
    if (smooth) scalar_gradient_smooth(g, rt, v, &gs, &gg);
    else scalar_gradient(g, rt, v, &gs, &gg);
    gs *= w;
    gg = cvscale_r(w, gg);
    acc->G += gs;
    for (int i = 0; i < s->n; i++) {
      vec3 phi = vscale(0.5, vsub(v, s->h[i]->p0));
      acc->Gphi[i] = cvadd(acc->Gphi[i], cv_from(gs, phi));
      acc->DGxphi[i] = cvadd(acc->DGxphi[i], cvcross_r(gg, phi));
    }
  }
}

/* Far pair (src/Impedance.cpp:141-169, tiers :329-361). */
static void lk_far(const green *g, const halflist *t, const halflist *s, int off, int npts, cplx *L, cplx *K) {
  const halfrwg *th = t->h[0];
  inner_acc in;
  for (int q = 0; q < npts; q++) {
    const double *abw = bem_quad_abw[off + q];
    vec3 v = quad_point(th->p0, th->v01, th->v02, abw[0], abw[1]);
    double w = abw[2];
    inner_integrate(g, s, v, off, npts, 0, &in);
    int ij = 0;
    for (int i = 0; i < t->n; i++) {
      vec3 phi = vscale(0.5, vsub(v, t->h[i]->p0));
      for (int j = 0; j < s->n; j++) {
        L[ij] += w * (g->ksq * cvdot_r(phi, in.Gphi[j]) - in.G);
        K[ij] -= w * cvdot_r(phi, in.DGxphi[j]);
        ij++;
      }
    }
  }
}

/* Close pair: Gauss105 outer x (Gauss91 smooth inner + analytic singular
 * part) (src/Impedance.cpp:171-251). */
static void lk_close(const green *g, const halflist *t, const halflist *s, cplx *L, cplx *K) {
  const halfrwg *th = t->h[0];
  const halfrwg *sh = s->h[0];
  vec3 vts[9];
  int ij = 0;
  for (int i = 0; i < t->n; i++)
    for (int j = 0; j < s->n; j++) vts[ij++] = vsub(s->h[j]->p0, t->h[i]->p0);
  inner_acc in;
  for (int q = 0; q < 105; q++) {
    const double *abw = bem_quad_abw[BEM_QUAD_OFF_105 + q];
    vec3 v = quad_point(th->p0, th->v01, th->v02, abw[0], abw[1]);
    double w = abw[2];
    inner_integrate(g, s, v, BEM_QUAD_OFF_91, 91, 1, &in);
    double ISn3h, ISn1, IS1;
    vec3 ILn1, IL1, IL3;
    oracle_eh_nonsmooth(sh, v, &ISn3h, &ISn1, &IS1, &ILn1, &IL1, &IL3);
    cplx k2 = g->k2;
    double Ai2 = 0.5 * sh->Ai;
    /* IL13 = Ai2 * (IL1 + (ONETHIRD*k2) * IL3) */
    cplx c13 = ONETHIRD * k2;
    cvec3 IL13 = cvscale_r(Ai2, cvadd(cv_real(IL1), cv_from(c13, IL3)));
    cplx ISn11 = Ai2 * (ISn1 + k2 * IS1);
    vec3 n = sh->n;
    double h = vdot(vsub(v, sh->p0), n);
    if (fabs(h) < 1e-8 * sh->elen) {
      cvec3 K21[3];
      for (int j = 0; j < s->n; j++) K21[j] = cvadd(IL13, cv_from(ISn11, vsub(v, s->h[j]->p0)));
      cplx L2 = 2.0 * ISn11 + in.G;
      for (int i = 0, ij2 = 0; i < t->n; i++) {
        vec3 phi = vscale(0.5, vsub(v, t->h[i]->p0));
        for (int j = 0; j < s->n; j++, ij2++)
          L[ij2] += w * (g->ksq * cvdot_r(phi, cvadd(in.Gphi[j], K21[j])) - L2);
        /* coplanar pairs accumulate no K (src/Impedance.cpp:223-225) */
      }
    } else {
      vec3 ISn3hn = vscale(Ai2 * ISn3h, n);
      cvec3 K21[3], K41[3];
      vec3 rho = vsub(v, vscale(h, n));
      cplx ak2 = Ai2 * k2;
      cvec3 K31 = cv_from(ak2, vsub(IL1, vscale(h * ISn1, n)));
      for (int j = 0; j < s->n; j++) {
        K21[j] = cvadd(IL13, cv_from(ISn11, vsub(rho, s->h[j]->p0)));
        K41[j] = cvcross_r(K31, vsub(v, s->h[j]->p0));
      }
      cplx L2 = 2.0 * ISn11 + in.G;
      for (int i = 0, ij2 = 0; i < t->n; i++) {
        vec3 phi = vscale(0.5, vsub(v, t->h[i]->p0));
        for (int j = 0; j < s->n; j++, ij2++) {
          L[ij2] += w * (g->ksq * cvdot_r(phi, cvadd(in.Gphi[j], K21[j])) - L2);
          cvec3 kk = cvadd(cvadd(in.DGxphi[j], cv_real(vcross(vts[ij2], ISn3hn))), K41[j]);
          K[ij2] -= w * cvdot_r(phi, kk);
        }
      }
    }
  }
}

/* Neighbour edge correction (src/Impedance.cpp:253-274, :301-324). */
static void lk_neighbor_edge(const halflist *t, vec3 e0, vec3 ev, vec3 *Kout) {
  const halfrwg *th = t->h[0];
  for (int i = 0; i < 3; i++) Kout[i] = v3(0, 0, 0);
  for (int q = 0; q < 15; q++) {
    double tq = bem_line15_tw[q][0], w = bem_line15_tw[q][1];
    vec3 v = vadd(e0, vscale(tq, ev));
    double ISn1; vec3 IL1;
    oracle_lk_edge(th, v, &ISn1, &IL1);
    vec3 n = th->n;
    double h = vdot(vsub(v, th->p0), n);
    vec3 rho = vsub(v, vscale(h, n));
    for (int i = 0; i < t->n; i++) {
      vec3 rho_p0 = vsub(rho, t->h[i]->p0);
      Kout[i] = vadd(Kout[i], vscale(w, vadd(IL1, vscale(ISn1, rho_p0))));
    }
  }
}

static int shared_points(const int *a, const int *b) {
  int c = 0;
  for (int i = 0; i < 3; i++)
    if (a[i] == b[0] || a[i] == b[1] || a[i] == b[2]) c++;
  return c;
}

/* tier index 0..5 = close, g6, g10, g15, g21, g36 */
int oracle_pair_class(const double *xyz, const int *tva, const int *tvb, const double *emax_a,
                      const double *emax_b, int same_body) {
  if (same_body) {
    if (tva == tvb) return 0;
    if (shared_points(tva, tvb) > 0) return 0;
  }
  vec3 pa = v3(0, 0, 0), pb = v3(0, 0, 0);
  for (int k = 0; k < 3; k++) {
    pa = vadd(pa, v3(xyz[3 * tva[k]], xyz[3 * tva[k] + 1], xyz[3 * tva[k] + 2]));
    pb = vadd(pb, v3(xyz[3 * tvb[k]], xyz[3 * tvb[k] + 1], xyz[3 * tvb[k] + 2]));
  }
  pa = vscale(ONETHIRD, pa);
  pb = vscale(ONETHIRD, pb);
  double d = vnorm2(vsub(pa, pb));
  double em = *emax_a > *emax_b ? *emax_a : *emax_b;
  double emax2 = em * em;
  if (d * FAR_THRESH6 > emax2) return 1;
  if (d * FAR_THRESH10 > emax2) return 2;
  if (d * FAR_THRESH15 > emax2) return 3;
  if (d * FAR_THRESH21 > emax2) return 4;
  return 5;
}

/* ------------------------------------------------------------------ tree */
typedef struct {
  int nv, nt, ne;
  const double *xyz;  /* nv x 3 */
  const int *tri;     /* nt x 3 (global vertex ids) */
  const int *edges;   /* ne x 4 {p1, p2, t1, t2} (global ids) */
  halfrwg *halves;    /* 2*ne */
  halflist *tlist;    /* nt */
  double *temax;      /* nt: longest side */
} tree;

static void tree_build(tree *T, int nv, const double *xyz, int nt, const int *tri, int ne, const int *edges) {
  T->nv = nv; T->nt = nt; T->ne = ne; T->xyz = xyz; T->tri = tri; T->edges = edges;
  T->halves = malloc(sizeof(halfrwg) * 2 * (size_t)ne);
  T->tlist = calloc((size_t)nt, sizeof(halflist));
  T->temax = malloc(sizeof(double) * (size_t)nt);
  for (int e = 0; e < ne; e++) {
    const int *E = edges + 4 * e;
    vec3 a = v3(xyz[3 * E[0]], xyz[3 * E[0] + 1], xyz[3 * E[0] + 2]);
    vec3 b = v3(xyz[3 * E[1]], xyz[3 * E[1] + 1], xyz[3 * E[1] + 2]);
    double len = sqrt(vnorm2(vsub(b, a)));
    make_half(&T->halves[2 * e], xyz, tri + 3 * E[2], E[0], E[1], E[2], e, 1, len);
    make_half(&T->halves[2 * e + 1], xyz, tri + 3 * E[3], E[0], E[1], E[3], e, 0, len);
    /* push order of Zts (src/Impedance.cpp:397-401): pos then neg, RWG order */
    halflist *lp = &T->tlist[E[2]];
    if (lp->n < 3) lp->h[lp->n++] = &T->halves[2 * e];
    halflist *ln = &T->tlist[E[3]];
    if (ln->n < 3) ln->h[ln->n++] = &T->halves[2 * e + 1];
  }
  for (int t = 0; t < nt; t++) {
    const int *f = tri + 3 * t;
    double m = 0;
    for (int k = 0; k < 3; k++) {
      int a = f[k], b = f[(k + 1) % 3];
      vec3 va = v3(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]);
      vec3 vb = v3(xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]);
      double l = sqrt(vnorm2(vsub(vb, va)));
      if (l > m) m = l;
    }
    T->temax[t] = m;
  }
}

static void tree_free(tree *T) {
  free(T->halves); free(T->tlist); free(T->temax);
}

/* One triangle pair -> 3x3 L, K (src/Impedance.cpp:276-363). */
static void lk_triangle(const green *g, const tree *T, int ta, int tb, int same_body, cplx *L, cplx *K) {
  for (int i = 0; i < 9; i++) { L[i] = 0; K[i] = 0; }
  const halflist *t = &T->tlist[ta], *s = &T->tlist[tb];
  if (t->n == 0 || s->n == 0) return;
  int cls = oracle_pair_class(T->xyz, T->tri + 3 * ta, T->tri + 3 * tb, &T->temax[ta], &T->temax[tb], same_body);
  if (cls == 0) {
    lk_close(g, t, s, L, K);
    if (ta != tb) {
      const halfrwg *sh = s->h[0];
      vec3 K01[3], K12[3], K20[3];
      lk_neighbor_edge(t, sh->p0, sh->v01, K01);
      lk_neighbor_edge(t, sh->p1, sh->v12, K12);
      lk_neighbor_edge(t, sh->p2, sh->v20, K20);
      double c = 0.125 * t->h[0]->Ai * s->h[0]->Ai;
      int ij = 0;
      for (int i = 0; i < t->n; i++)
        for (int j = 0; j < s->n; j++) {
          vec3 v = vsub(t->h[i]->p0, s->h[j]->p0);
          double add = vdot(vcross(v, sh->m01), K01[i]) / sh->normv01i +
                       vdot(vcross(v, sh->m12), K12[i]) / sh->normv12i +
                       vdot(vcross(v, sh->m20), K20[i]) / sh->normv20i;
          K[ij] += c * add;
          ij++;
        }
    }
    return;
  }
  static const int offs[6] = {0, BEM_QUAD_OFF_6, BEM_QUAD_OFF_10, BEM_QUAD_OFF_15, BEM_QUAD_OFF_21, BEM_QUAD_OFF_36};
  static const int npts[6] = {0, 6, 10, 15, 21, 36};
  lk_far(g, t, s, offs[cls], npts[cls], L, K);
}

/*
 * Self-interaction impedance of one tree (Zts(t, t, g), src/Impedance.cpp:388-413):
 *   Z is 2ne x 2ne complex128, column-major with leading dimension ldz, rows
 *   and columns interleaved [J_0, M_0, J_1, M_1, ...].  Z is overwritten.
 * body_of_tri: body id per triangle (pairs in different bodies are never close).
 * Only test triangles in [ta_lo, ta_hi) are evaluated (bounded timing samples);
 * pass 0, nt for the full matrix.
 * Returns 0.
 */
int oracle_zts(int nv, const double *xyz, int nt, const int *tri, const int *body_of_tri, int ne,
               const int *edges, double eps_re, double eps_im, double mu_re, double mu_im, double omega,
               int ta_lo, int ta_hi, double *Z, int64_t ldz) {
  tree T;
  tree_build(&T, nv, xyz, nt, tri, ne, edges);
  green g;
  green_init(&g, eps_re + I * eps_im, mu_re + I * mu_im, omega);
  cplx *Zc = (cplx *)Z;
  int64_t M = 2 * (int64_t)ne;
  for (int64_t j = 0; j < M; j++) memset(Zc + j * ldz, 0, sizeof(cplx) * (size_t)M);
  const int chunk = 64;
  cplx *tiles = malloc(sizeof(cplx) * 18 * (size_t)chunk * (size_t)nt);
  for (int c0 = ta_lo; c0 < ta_hi; c0 += chunk) {
    int c1 = c0 + chunk < ta_hi ? c0 + chunk : ta_hi;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int ta = c0; ta < c1; ta++)
      for (int tb = 0; tb < nt; tb++) {
        cplx *tile = tiles + 18 * ((size_t)(ta - c0) * nt + tb);
        lk_triangle(&g, &T, ta, tb, body_of_tri[ta] == body_of_tri[tb], tile, tile + 9);
      }
    /* deterministic scatter (src/Impedance.cpp:365-384) */
    for (int ta = c0; ta < c1; ta++)
      for (int tb = 0; tb < nt; tb++) {
        const cplx *L = tiles + 18 * ((size_t)(ta - c0) * nt + tb), *K = L + 9;
        const halflist *t = &T.tlist[ta], *s = &T.tlist[tb];
        int ij = 0;
        for (int i = 0; i < t->n; i++) {
          const halfrwg *th = t->h[i];
          int64_t ii = 2 * (int64_t)th->rwg;
          for (int j = 0; j < s->n; j++) {
            const halfrwg *sh = s->h[j];
            int64_t jj = 2 * (int64_t)sh->rwg;
            double c = ONEOVER4PI * th->elen * sh->elen;
            if (th->positive != sh->positive) c = -c;
            cplx Lij = c * L[ij], Kij = c * K[ij];
            Zc[ii + jj * ldz] += g.Zeps * Lij;
            Zc[ii + (jj + 1) * ldz] += Kij;
            Zc[ii + 1 + jj * ldz] -= Kij;
            Zc[ii + 1 + (jj + 1) * ldz] += g.Zmu * Lij;
            ij++;
          }
        }
      }
  }
  free(tiles);
  tree_free(&T);
  return 0;
}

/* Pair class histogram of a tree (0 close, 1..5 far tiers 6/10/15/21/36). */
int oracle_pair_histogram(int nv, const double *xyz, int nt, const int *tri, const int *body_of_tri,
                          int64_t *hist6) {
  double *temax = malloc(sizeof(double) * (size_t)nt);
  for (int t = 0; t < nt; t++) {
    const int *f = tri + 3 * t;
    double m = 0;
    for (int k = 0; k < 3; k++) {
      int a = f[k], b = f[(k + 1) % 3];
      vec3 va = v3(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]);
      vec3 vb = v3(xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]);
      double l = sqrt(vnorm2(vsub(vb, va)));
      if (l > m) m = l;
    }
    temax[t] = m;
  }
  for (int k = 0; k < 6; k++) hist6[k] = 0;
  for (int a = 0; a < nt; a++)
    for (int b = 0; b < nt; b++)
      hist6[oracle_pair_class(xyz, tri + 3 * a, tri + 3 * b, &temax[a], &temax[b],
                              body_of_tri[a] == body_of_tri[b])]++;
  free(temax);
  return 0;
}

/* ------------------------------------------------------ dipole and flux */
static inline cvec3 cvsub(cvec3 a, cvec3 b) { return cv3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline cvec3 cvcross(cvec3 a, cvec3 b) {
  return cv3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline vec3 cvreal(cvec3 a) { return v3(creal(a.x), creal(a.y), creal(a.z)); }
static inline cvec3 cvconj(cvec3 a) { return cv3(conj(a.x), conj(a.y), conj(a.z)); }

/*
 * Right-hand side of a point dipole (moment j, real, at p) in the medium
 * (eps, mu) -- DipolePointSource::EHTriangle (src/Source.cpp:121-165) summed
 * over the triangles of a tree as Bdense does (src/Source.cpp:167-193): per
 * triangle, the smooth part of the field by the 21-point rule on the
 * triangle's first half-RWG, the singular part by EHNonSmoothIntegral, then
 * B(2i) += E_i (J rows) and B(2i+1) += H_i (M rows) for each of its halves,
 * with the sign of the exterior rule (:150-153) when external is set.  Only
 * triangles with tri_on[t] != 0 take part (those bordering the dipole's
 * region).  B: 2ne complex128, overwritten.
 */
int oracle_dipole_rhs(int nv, const double *xyz, int nt, const int *tri, int ne, const int *edges, double eps_re,
                      double eps_im, double mu_re, double mu_im, double omega, int external, const double *p,
                      const double *jv, const int *tri_on, double *B) {
  tree T;
  tree_build(&T, nv, xyz, nt, tri, ne, edges);
  green g;
  green_init(&g, eps_re + I * eps_im, mu_re + I * mu_im, omega);
  cplx *Bc = (cplx *)B;
  memset(Bc, 0, sizeof(cplx) * 2 * (size_t)ne);
  const vec3 pp = v3(p[0], p[1], p[2]), j = v3(jv[0], jv[1], jv[2]);
  for (int t = 0; t < nt; t++) {
    const halflist *tl = &T.tlist[t];
    if (!tri_on[t] || tl->n == 0) continue;
    const halfrwg *th = tl->h[0];
    /* EHSmoothIntegrand (src/Source.cpp:88-119) over the 21-point rule */
    cvec3 DG = cvzero(), Gphi[3], DGxphi[3];
    for (int i = 0; i < 3; i++) Gphi[i] = DGxphi[i] = cvzero();
    for (int q = 0; q < 21; q++) {
      const double *abw = bem_quad_abw[BEM_QUAD_OFF_21 + q];
      vec3 v = quad_point(th->p0, th->v01, th->v02, abw[0], abw[1]);
      cplx gs;
      cvec3 gg;
      scalar_gradient_smooth(&g, pp, v, &gs, &gg);
      gs *= abw[2];
      gg = cvscale_r(abw[2], gg);
      DG = cvadd(DG, gg);
      for (int i = 0; i < tl->n; i++) {
        vec3 phi = vscale(0.5, vsub(v, tl->h[i]->p0));
        Gphi[i] = cvadd(Gphi[i], cv_from(gs, phi));
        DGxphi[i] = cvadd(DGxphi[i], cvcross_r(gg, phi));
      }
    }
    double ISn3h, ISn1, IS1;
    vec3 ILn1, IL1, IL3;
    oracle_eh_nonsmooth(th, pp, &ISn3h, &ISn1, &IS1, &ILn1, &IL1, &IL3);
    const cplx k2 = g.k2;
    const double Ai = th->Ai, Ai2 = 0.5 * Ai;
    cvec3 IL13 = cvscale_r(Ai2, cvadd(cv_real(IL1), cv_from(ONETHIRD * k2, IL3)));
    cplx ISn11 = Ai2 * (ISn1 + k2 * IS1);
    vec3 n = th->n;
    double h = vdot(vsub(pp, th->p0), n);
    vec3 rho = vsub(pp, vscale(h, n));
    cvec3 K31 = cvscale_r(Ai2, cvadd(cv_real(vadd(ILn1, vscale(ISn3h, n))),
                                     cv_from(k2, vsub(IL1, vscale(h * ISn1, n)))));
    cvec3 E1 = cvadd(DG, cvscale_r(Ai, cvadd(cv_from(ISn3h - k2 * h * ISn1, n),
                                             cvadd(cv_real(ILn1), cv_from(k2, IL1)))));
    for (int i = 0; i < tl->n; i++) {
      const halfrwg *hi = tl->h[i];
      int64_t ii = 2 * (int64_t)hi->rwg;
      double c = ONEOVER4PI * hi->elen;
      if ((external != 0) != (hi->positive != 0)) c = -c;
      cvec3 a = cvadd(cvadd(Gphi[i], IL13), cv_from(ISn11, vsub(rho, hi->p0)));
      cplx Ei = c * g.Zeps * cvdot_r(j, cvsub(E1, cvscale(g.ksq, a)));
      cvec3 hv = cvadd(DGxphi[i], cvcross_r(K31, vsub(pp, hi->p0)));
      cplx Hi = c * cvdot_r(j, cvscale_r(-1.0, hv));
      Bc[ii] += Ei;
      Bc[ii + 1] += Hi;
    }
  }
  tree_free(&T);
  return 0;
}

/*
 * Power flux through the triangles [t_lo, t_hi) of a tree from the solution X
 * (2ne complex128, [J, M] interleaved) -- flux(top, t, X)
 * (src/Flux.cpp:100-219): per triangle the surface currents at its vertices
 * from its three RWGs (signed by whether the triangle is the RWG's positive
 * half), M scaled by 1/omega and J conjugated, then
 * dS = n . Re(j1 x m1 + j2 x m2 + j3 x m3 + (j1+j2+j3) x (m1+m2+m3)) / (48 A2),
 * A2 twice the triangle's area.  Returns the sum of dS in triangle order.
 */
double oracle_flux(int nv, const double *xyz, int nt, const int *tri, int ne, const int *edges, int t_lo, int t_hi,
                   double omega, const double *X) {
  (void)nv;
  const cplx *x = (const cplx *)X;
  int *tedge = malloc(sizeof(int) * 3 * (size_t)nt);
  for (int i = 0; i < 3 * nt; i++) tedge[i] = -1;
  for (int e = 0; e < ne; e++) {
    const int *E = edges + 4 * e;
    for (int side = 2; side <= 3; side++) {
      const int t = E[side], *f = tri + 3 * t;
      for (int k = 0; k < 3; k++) {   /* e12, e23, e31 */
        int a = f[k], b = f[(k + 1) % 3];
        if ((a == E[0] && b == E[1]) || (a == E[1] && b == E[0])) tedge[3 * t + k] = e;
      }
    }
  }
  double S = 0.;
  for (int t = t_lo; t < t_hi; t++) {
    const int *f = tri + 3 * t;
    vec3 P1 = v3(xyz[3 * f[0]], xyz[3 * f[0] + 1], xyz[3 * f[0] + 2]);
    vec3 P2 = v3(xyz[3 * f[1]], xyz[3 * f[1] + 1], xyz[3 * f[1] + 2]);
    vec3 P3 = v3(xyz[3 * f[2]], xyz[3 * f[2] + 1], xyz[3 * f[2] + 2]);
    vec3 n = vcross(vsub(P2, P1), vsub(P3, P1));   /* normalUnnormalized, src/Triangle.cpp:5-12 */
    double A2 = sqrt(vnorm2(n));
    n = vscale(1. / A2, n);
    vec3 v12 = vsub(P2, P1), v23 = vsub(P3, P2), v31 = vsub(P1, P3);
    cvec3 j1 = cvzero(), j2 = cvzero(), j3 = cvzero(), m1 = cvzero(), m2 = cvzero(), m3 = cvzero();
    int e12 = tedge[3 * t], e23 = tedge[3 * t + 1], e31 = tedge[3 * t + 2];
    if (e12 < 0 || e23 < 0 || e31 < 0) continue;   /* open edge: no RWG */
    const int *E;
    double l;
    E = edges + 4 * e12;
    l = sqrt(vnorm2(vsub(v3(xyz[3 * E[1]], xyz[3 * E[1] + 1], xyz[3 * E[1] + 2]),
                         v3(xyz[3 * E[0]], xyz[3 * E[0] + 1], xyz[3 * E[0] + 2]))));
    double c12 = t == E[2] ? l : -l;
    cplx j12 = c12 * x[2 * e12], m12 = c12 * x[2 * e12 + 1];
    j1 = cvadd(j1, cv_from(j12, v31));
    j2 = cvsub(j2, cv_from(j12, v23));
    m1 = cvadd(m1, cv_from(m12, v31));
    m2 = cvsub(m2, cv_from(m12, v23));
    E = edges + 4 * e23;
    l = sqrt(vnorm2(vsub(v3(xyz[3 * E[1]], xyz[3 * E[1] + 1], xyz[3 * E[1] + 2]),
                         v3(xyz[3 * E[0]], xyz[3 * E[0] + 1], xyz[3 * E[0] + 2]))));
    double c23 = t == E[2] ? l : -l;
    cplx j23 = c23 * x[2 * e23], m23 = c23 * x[2 * e23 + 1];
    j2 = cvadd(j2, cv_from(j23, v12));
    j3 = cvsub(j3, cv_from(j23, v31));
    m2 = cvadd(m2, cv_from(m23, v12));
    m3 = cvsub(m3, cv_from(m23, v31));
    E = edges + 4 * e31;
    l = sqrt(vnorm2(vsub(v3(xyz[3 * E[1]], xyz[3 * E[1] + 1], xyz[3 * E[1] + 2]),
                         v3(xyz[3 * E[0]], xyz[3 * E[0] + 1], xyz[3 * E[0] + 2]))));
    double c31 = t == E[2] ? l : -l;
    cplx j31 = c31 * x[2 * e31], m31 = c31 * x[2 * e31 + 1];
    j3 = cvadd(j3, cv_from(j31, v23));
    j1 = cvsub(j1, cv_from(j31, v12));
    m3 = cvadd(m3, cv_from(m31, v23));
    m1 = cvsub(m1, cv_from(m31, v12));
    cplx omegai = 1.0 / (cplx)omega;
    m1 = cvscale(omegai, m1);
    m2 = cvscale(omegai, m2);
    m3 = cvscale(omegai, m3);
    j1 = cvconj(j1);
    j2 = cvconj(j2);
    j3 = cvconj(j3);
    vec3 s = vadd(vadd(vadd(cvreal(cvcross(j1, m1)), cvreal(cvcross(j2, m2))), cvreal(cvcross(j3, m3))),
                  cvreal(cvcross(cvadd(cvadd(j1, j2), j3), cvadd(cvadd(m1, m2), m3))));
    S += vdot(n, s) / (48. * A2);
  }
  free(tedge);
  return S;
}

/* Test hooks for the analytic integrals: half-RWG built from a triangle
 * (vertex triple p[9]) and an edge given by two of its vertices. */
int oracle_eh_nonsmooth_tri(const double *p9, int ea, int eb, const double *rt, double *out12) {
  int tv[3] = {0, 1, 2};
  halfrwg h;
  double elen = sqrt(vnorm2(vsub(v3(p9[3 * eb], p9[3 * eb + 1], p9[3 * eb + 2]),
                                 v3(p9[3 * ea], p9[3 * ea + 1], p9[3 * ea + 2]))));
  make_half(&h, p9, tv, ea, eb, 0, 0, 1, elen);
  double ISn3h, ISn1, IS1; vec3 ILn1, IL1, IL3;
  oracle_eh_nonsmooth(&h, v3(rt[0], rt[1], rt[2]), &ISn3h, &ISn1, &IS1, &ILn1, &IL1, &IL3);
  double o[12] = {ISn3h, ISn1, IS1, ILn1.x, ILn1.y, ILn1.z, IL1.x, IL1.y, IL1.z, IL3.x, IL3.y, IL3.z};
  memcpy(out12, o, sizeof o);
  return 0;
}

/* Green's function values for tests: out = {g, grad g (3)} as complex pairs. */
int oracle_green(double eps_re, double eps_im, double mu_re, double mu_im, double omega, const double *rt,
                 const double *rs, int smooth, double *out8) {
  green g;
  green_init(&g, eps_re + I * eps_im, mu_re + I * mu_im, omega);
  cplx gs; cvec3 gg;
  vec3 a = v3(rt[0], rt[1], rt[2]), b = v3(rs[0], rs[1], rs[2]);
  if (smooth) scalar_gradient_smooth(&g, a, b, &gs, &gg);
  else scalar_gradient(&g, a, b, &gs, &gg);
  cplx o[4] = {gs, gg.x, gg.y, gg.z};
  memcpy(out8, o, sizeof o);
  return 0;
}

/* Threads of the following assembly calls (the CPU baseline times the
   reference-faithful single-threaded assembly beside the all-cores one). */
void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
