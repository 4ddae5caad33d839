// FP64 RWG/PMCHWT impedance assembly on gfx950.
//
// The reference evaluates one triangle pair at a time into a 3x3 tile of
// L (single-layer) and K (double-layer) integrals and scatters the tile into
// Z (src/Impedance.cpp:276-385).  Here the pair space of a tree is split by
// class and each class gets its own launch shape:
//   far_g6_kernel   thread per pair over the dense pair grid (16x16 pairs per
//                   256-thread block, triangle data staged in LDS); computes
//                   every 6-point pair, queues 36-point pairs, skips close ones
//   pair_block_kernel<36,36>   one 64-lane block per queued 36-point pair
//   pair_block_kernel<105,91>  one 128-lane block per close pair (self or
//                   shared vertex): smooth Gauss105 x Gauss91 quadrature plus
//                   the analytic singular part and the neighbour edge term
// Tiles land in a structure-of-arrays buffer (36 real planes); gather_kernel
// then forms the 2x2 [J,M] blocks of Z for every RWG pair in a fixed order
// (deterministic: no floating-point atomics anywhere).
//
// Algebra.  The reference accumulates, per outer point and source half j,
//   Gphi_j = sum w g (r'-p0_j)/2,  DGxphi_j = sum w grad g x (r'-p0_j)/2
// (src/Impedance.cpp:93-104).  Both are affine in p0_j, so the inner loop
// keeps only the four j-independent moments
//   G0 = sum w g, G1 = sum w g rb, D0 = sum w grad g, D1 = sum w grad g x rb
// with rb = r' - o (o = the source frame vertex, keeping magnitudes local),
// and Gphi_j = (G1 - G0 pb_j)/2, DGxphi_j = (D1 - D0 x pb_j)/2, pb_j = p0_j - o.
// The outer step (src/Impedance.cpp:154-167, :194-250) is then applied per
// (i, j) exactly as written there.
#include "bem_common.h"

#define BEM_QUAD_NO_HOST_TABLES
#include "quadrature_tables.h"

__constant__ double c_quad[BEM_QUAD_TOTAL_POINTS][3] = BEM_QUAD_ABW_INIT;
__constant__ double c_line15[15][2] = BEM_LINE15_INIT;

namespace {

struct C2 {
  double re, im;
};

__device__ __forceinline__ C2 cmk(double r, double i) { C2 c; c.re = r; c.im = i; return c; }
__device__ __forceinline__ C2 cmul(C2 a, C2 b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Pair classification with the reference's exact arithmetic (no contraction):
// same body and a shared vertex -> close; else d*12 > emax^2 -> 6-point rule.
__device__ int classify_pair(const TriGeom &A, const TriGeom &B, bool same_tri) {
#pragma clang fp contract(off)
  if (A.body == B.body) {
    if (same_tri) return BEM_CLASS_CLOSE;
    for (int i = 0; i < 3; i++)
      if (A.vid[i] == B.vid[0] || A.vid[i] == B.vid[1] || A.vid[i] == B.vid[2]) return BEM_CLASS_CLOSE;
  }
  double vx = A.centre[0] - B.centre[0];
  double vy = A.centre[1] - B.centre[1];
  double vz = A.centre[2] - B.centre[2];
  double d = vx * vx + vy * vy + vz * vz;
  double em = A.emax > B.emax ? A.emax : B.emax;
  double emax2 = em * em;
  if (d * 12.0 > emax2) return BEM_CLASS_G6;
  return BEM_CLASS_G36;
}

// Quadrature point v0 + a v01 + b v02 (left to right, as TriangleQuadrature.h).
__device__ __forceinline__ void quad_point(const TriGeom &T, int q, double *v) {
#pragma clang fp contract(off)
  double a = c_quad[q][0], b = c_quad[q][1];
  for (int c = 0; c < 3; c++) {
    double t = T.p0[c] + a * T.v01[c];
    v[c] = t + b * T.v02[c];
  }
}

// exp(a) (cos b + i sin b) for the Helmholtz kernel exp(ikR), a = -Im(k) R,
// b = Re(k) R, written from the series: Cody-Waite reduction (Cody & Waite,
// Software Manual for the Elementary Functions, 1980) of b by pi/2 and of a
// by ln 2 with two-part constants (32 significant bits in the leading part,
// so n times it is exact for |n| < 2^20), then Taylor polynomials on the
// reduced ranges (|r| <= pi/4: sin to r^17, cos to r^16; |t| <= ln2/2: exp to
// t^13; truncation below 1e-17 relative).  A lossless medium (a == 0 for
// every lane of the launch) skips the exponential.
__device__ __forceinline__ void cexp_i(double a, double b, double &er, double &ei) {
  const double n = rint(b * 0.6366197723675814);
  const double r = fma(-n, 6.077100506506192e-11, fma(-n, 1.5707963267341256, b));
  const double r2 = r * r;
  double q = 2.8114572543455206e-15;
  q = fma(q, r2, -7.647163731819816e-13);
  q = fma(q, r2, 1.6059043836821613e-10);
  q = fma(q, r2, -2.505210838544172e-08);
  q = fma(q, r2, 2.7557319223985893e-06);
  q = fma(q, r2, -0.0001984126984126984);
  q = fma(q, r2, 0.008333333333333333);
  q = fma(q, r2, -0.16666666666666666);
  const double sr = fma(r * r2, q, r);
  double c = 4.779477332387385e-14;
  c = fma(c, r2, -1.1470745597729725e-11);
  c = fma(c, r2, 2.08767569878681e-09);
  c = fma(c, r2, -2.755731922398589e-07);
  c = fma(c, r2, 2.48015873015873e-05);
  c = fma(c, r2, -0.001388888888888889);
  c = fma(c, r2, 0.041666666666666664);
  c = fma(c, r2, -0.5);
  const double cr = fma(c, r2, 1.0);
  const int k = (int)n & 3;
  double sb = (k & 1) ? cr : sr, cb = (k & 1) ? sr : cr;
  if (k == 1 || k == 2) cb = -cb;
  if (k >= 2) sb = -sb;
  double ex = 1.0;
  if (a != 0.0) {
    const double m = rint(a * 1.4426950408889634);
    const double t = fma(-m, 1.9082149292705877e-10, fma(-m, 0.6931471803691238, a));
    double p = 1.6059043836821613e-10;
    p = fma(p, t, 2.08767569878681e-09);
    p = fma(p, t, 2.505210838544172e-08);
    p = fma(p, t, 2.755731922398589e-07);
    p = fma(p, t, 2.7557319223985893e-06);
    p = fma(p, t, 2.48015873015873e-05);
    p = fma(p, t, 0.0001984126984126984);
    p = fma(p, t, 0.001388888888888889);
    p = fma(p, t, 0.008333333333333333);
    p = fma(p, t, 0.041666666666666664);
    p = fma(p, t, 0.16666666666666666);
    p = fma(p, t, 0.5);
    p = fma(p, t, 1.0);
    ex = ldexp(fma(p, t, 1.0), (int)m);
  }
  er = ex * cb;
  ei = ex * sb;
}

// R = sqrt(x) and 1/R for a normal x > 0: the v_rsq_f64 estimate refined by
// one coupled Newton step for (sqrt x, 1/(2 sqrt x)) (Markstein, IA-64 and
// Elementary Functions, 2000), a second step on the reciprocal and a Newton
// correction of the root.
__device__ __forceinline__ void root_and_inverse(double x, double &R, double &invR) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  R = fma(d, h, g);
  invR = 2.0 * h;
}

// Inner moments (see file comment).
struct Inner {
  C2 G0;
  double G1r[3], G1i[3];
  double D0r[3], D0i[3];
  double D1r[3], D1i[3];
};

__device__ __forceinline__ void inner_zero(Inner &m) {
  m.G0 = cmk(0, 0);
  for (int c = 0; c < 3; c++) { m.G1r[c] = m.G1i[c] = m.D0r[c] = m.D0i[c] = m.D1r[c] = m.D1i[c] = 0.0; }
}

// An Inner in INNER_PLANES planes of stride doubles (structure of arrays, so
// neighbouring outer points are neighbouring addresses).
#define INNER_PLANES 20
__device__ __forceinline__ void inner_store(const Inner &m, double *__restrict__ mom, size_t stride, size_t i) {
  mom[i] = m.G0.re;
  mom[stride + i] = m.G0.im;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    mom[(2 + c) * stride + i] = m.G1r[c];
    mom[(5 + c) * stride + i] = m.G1i[c];
    mom[(8 + c) * stride + i] = m.D0r[c];
    mom[(11 + c) * stride + i] = m.D0i[c];
    mom[(14 + c) * stride + i] = m.D1r[c];
    mom[(17 + c) * stride + i] = m.D1i[c];
  }
}
__device__ __forceinline__ void inner_load(Inner &m, const double *__restrict__ mom, size_t stride, size_t i) {
  m.G0 = cmk(mom[i], mom[stride + i]);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    m.G1r[c] = mom[(2 + c) * stride + i];
    m.G1i[c] = mom[(5 + c) * stride + i];
    m.D0r[c] = mom[(8 + c) * stride + i];
    m.D0i[c] = mom[(11 + c) * stride + i];
    m.D1r[c] = mom[(14 + c) * stride + i];
    m.D1i[c] = mom[(17 + c) * stride + i];
  }
}

// One inner point: Green's function (regular src/GreensFunction.cpp:41-52 or
// smooth :54-69) at d = r - r', weighted moment update.
template <bool SMOOTH>
__device__ __forceinline__ void inner_point(const Medium &md, const double *r, const double *rp, const double *ob,
                                            double w, Inner &m) {
  double d[3] = {r[0] - rp[0], r[1] - rp[1], r[2] - rp[2]};
  double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double R, invR;
  root_and_inverse(R2, R, invR);
  double ar = md.ik_re * R, ai = md.ik_im * R;  // ikR
  double er, ei;
  cexp_i(ar, ai, er, ei);                       // exp(ikR)
  C2 g, q;
  if (!SMOOTH) {
    g = cmk(er * invR, ei * invR);
    // q = g (1 - ikR) / R^2
    C2 t = cmul(g, cmk(1.0 - ar, -ai));
    double invR2 = invR * invR;
    q = cmk(t.re * invR2, t.im * invR2);
  } else {
    // g = (e - 1)/R - k2 R ; q = (e (1 - ikR) + k2 R^2 - 1) / R^3
    g = cmk((er - 1.0) * invR - md.k2_re * R, ei * invR - md.k2_im * R);
    C2 t = cmul(cmk(er, ei), cmk(1.0 - ar, -ai));
    double invR3 = invR * invR * invR;
    q = cmk((t.re + md.k2_re * R2 - 1.0) * invR3, (t.im + md.k2_im * R2) * invR3);
  }
  double wgr = w * g.re, wgi = w * g.im;
  double wqr = w * q.re, wqi = w * q.im;
  double rb[3] = {rp[0] - ob[0], rp[1] - ob[1], rp[2] - ob[2]};
  double cr[3];
  cross3(d, rb, cr);
  m.G0.re += wgr;
  m.G0.im += wgi;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    m.G1r[c] += wgr * rb[c];
    m.G1i[c] += wgi * rb[c];
    m.D0r[c] += wqr * d[c];
    m.D0i[c] += wqi * d[c];
    m.D1r[c] += wqr * cr[c];
    m.D1i[c] += wqi * cr[c];
  }
}

// Far outer step (src/Impedance.cpp:154-167): L_ij += w (k^2 phi_i.Gphi_j - G),
// K_ij -= w phi_i.DGxphi_j, phi_i = (r - p0_i)/2.
__device__ __forceinline__ void outer_far(const Medium &md, const double *r, double w, const Inner &m,
                                          const double (*p0t)[3], int nt, const double (*pbs)[3], int ns,
                                          double *Lr, double *Li, double *Kr, double *Ki) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double phi[3] = {0.5 * (r[0] - p0t[i][0]), 0.5 * (r[1] - p0t[i][1]), 0.5 * (r[2] - p0t[i][2])};
    C2 a = cmk(dot3(phi, m.G1r), dot3(phi, m.G1i));
    C2 c = cmk(dot3(phi, m.D1r), dot3(phi, m.D1i));
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (i >= nt || j >= ns) continue;
      double b = dot3(phi, pbs[j]);
      double e[3];
      cross3(pbs[j], phi, e);
      C2 gp = cmk(0.5 * (a.re - b * m.G0.re), 0.5 * (a.im - b * m.G0.im));
      C2 kg = cmul(cmk(md.ksq_re, md.ksq_im), gp);
      int ij = 3 * i + j;
      Lr[ij] += w * (kg.re - m.G0.re);
      Li[ij] += w * (kg.im - m.G0.im);
      double dr = 0.5 * (c.re - dot3(m.D0r, e));
      double di = 0.5 * (c.im - dot3(m.D0i, e));
      Kr[ij] -= w * dr;
      Ki[ij] -= w * di;
    }
  }
}

// EHNonSmoothIntegral (src/Source.cpp:16-86) on the frame half of triangle S.
// ILn1 (optional) receives sum L_e m_e, which the dipole source needs.
__device__ __forceinline__ void eh_nonsmooth(const TriGeom &S, const double *rt, double &ISn3h, double &ISn1, double &IS1,
                             double *IL1, double *IL3, double *ILn1 = nullptr) {
  double tol = 1e-6 * S.elen;
  double p0[3], p1[3], p2[3];
  for (int c = 0; c < 3; c++) { p0[c] = rt[c] - S.p0[c]; p1[c] = rt[c] - S.p1[c]; p2[c] = rt[c] - S.p2[c]; }
  double R02 = dot3(p0, p0), R0 = sqrt(R02);
  double R12 = dot3(p1, p1), R1 = sqrt(R12);
  double R22 = dot3(p2, p2), R2 = sqrt(R22);
  double i0 = 1. / R0, i1 = 1. / R1, i2 = 1. / R2;
  double a0[3] = {i0 * p0[0], i0 * p0[1], i0 * p0[2]};
  double a1[3] = {i1 * p1[0], i1 * p1[1], i1 * p1[2]};
  double a2[3] = {i2 * p2[0], i2 * p2[1], i2 * p2[2]};
  double a12[3] = {a1[0] + a2[0], a1[1] + a2[1], a1[2] + a2[2]};
  double x = 1.0 + dot3(a0, a12) + dot3(a1, a2);
  double cx[3];
  cross3(a1, a2, cx);
  double tp = dot3(a0, cx);
  double y = sqrt(tp * tp);
  double h = dot3(p0, S.n);
  ISn3h = (h < 0. ? -2. : 2.) * atan2(y, x);
  double num, den;
  double s010 = S.normv01i * dot3(S.v01, p0);
  double s011 = S.normv01i * dot3(S.v01, p1);
  num = R1 - s011; den = R0 - s010;
  double L01 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s121 = S.normv12i * dot3(S.v12, p1);
  double s122 = S.normv12i * dot3(S.v12, p2);
  num = R2 - s122; den = R1 - s121;
  double L12 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s202 = S.normv20i * dot3(S.v20, p2);
  double s200 = S.normv20i * dot3(S.v20, p0);
  num = R0 - s200; den = R2 - s202;
  double L20 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double t01 = dot3(S.m01, p0), t12 = dot3(S.m12, p1), t20 = dot3(S.m20, p2);
  ISn1 = -h * ISn3h - (t01 * L01 + t12 * L12 + t20 * L20);
  double R02_01 = R02 - s010 * s010;
  double IL1_01 = 0.5 * (R02_01 * L01 - (R1 * s011 - R0 * s010));
  double R02_12 = R12 - s121 * s121;
  double IL1_12 = 0.5 * (R02_12 * L12 - (R2 * s122 - R1 * s121));
  double R02_20 = R22 - s202 * s202;
  double IL1_20 = 0.5 * (R02_20 * L20 - (R0 * s200 - R2 * s202));
  double IL3_01 = 0.75 * R02_01 * IL1_01 - 0.25 * (R12 * R1 * s011 - R02 * R0 * s010);
  double IL3_12 = 0.75 * R02_12 * IL1_12 - 0.25 * (R22 * R2 * s122 - R12 * R1 * s121);
  double IL3_20 = 0.75 * R02_20 * IL1_20 - 0.25 * (R02 * R0 * s200 - R22 * R2 * s202);
  IS1 = BEM_ONETHIRD * (h * h * ISn1 - (t01 * IL1_01 + t12 * IL1_12 + t20 * IL1_20));
  for (int c = 0; c < 3; c++) {
    IL1[c] = IL1_01 * S.m01[c] + IL1_12 * S.m12[c] + IL1_20 * S.m20[c];
    IL3[c] = IL3_01 * S.m01[c] + IL3_12 * S.m12[c] + IL3_20 * S.m20[c];
  }
  if (ILn1)
    for (int c = 0; c < 3; c++) ILn1[c] = L01 * S.m01[c] + L12 * S.m12[c] + L20 * S.m20[c];
}

// LKEdgeIntegral (src/Impedance.cpp:13-73) on the frame half of triangle T.
__device__ __forceinline__ void lk_edge(const TriGeom &T, const double *rs, double &ISn1, double *IL1) {
  double p0[3], p1[3], p2[3];
  for (int c = 0; c < 3; c++) { p0[c] = rs[c] - T.p0[c]; p1[c] = rs[c] - T.p1[c]; p2[c] = rs[c] - T.p2[c]; }
  double R02 = dot3(p0, p0), R0 = sqrt(R02);
  double R12 = dot3(p1, p1), R1 = sqrt(R12);
  double R22 = dot3(p2, p2), R2 = sqrt(R22);
  double tol = 1e-8 * T.elen;
  double i0 = 1. / R0, i1 = 1. / R1, i2 = 1. / R2;
  double a0[3] = {i0 * p0[0], i0 * p0[1], i0 * p0[2]};
  double a1[3] = {i1 * p1[0], i1 * p1[1], i1 * p1[2]};
  double a2[3] = {i2 * p2[0], i2 * p2[1], i2 * p2[2]};
  double h = dot3(p0, T.n);
  double a12[3] = {a1[0] + a2[0], a1[1] + a2[1], a1[2] + a2[2]};
  double x = 1.0 + dot3(a0, a12) + dot3(a1, a2);
  double cx[3];
  cross3(a1, a2, cx);
  double tp = dot3(a0, cx);
  double y = sqrt(tp * tp);
  double ISn3h = (h < 0 ? -2. : 2.) * atan2(y, x);
  double t01 = dot3(T.m01, p0), t12 = dot3(T.m12, p1), t20 = dot3(T.m20, p2);
  double s010 = T.normv01i * dot3(T.v01, p0);
  double s011 = T.normv01i * dot3(T.v01, p1);
  double num, den;
  num = R1 - s011; den = R0 - s010;
  double L01 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s121 = T.normv12i * dot3(T.v12, p1);
  double s122 = T.normv12i * dot3(T.v12, p2);
  num = R2 - s122; den = R1 - s121;
  double L12 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  double s202 = T.normv20i * dot3(T.v20, p2);
  double s200 = T.normv20i * dot3(T.v20, p0);
  num = R0 - s200; den = R2 - s202;
  double L20 = fabs(num) < tol || fabs(den) < tol ? 0. : log(num / den);
  ISn1 = -h * ISn3h - (t01 * L01 + t12 * L12 + t20 * L20);
  double R02_01 = R02 - s010 * s010;
  double IL1_01 = 0.5 * (R02_01 * L01 - (R1 * s011 - R0 * s010));
  double R02_12 = R12 - s121 * s121;
  double IL1_12 = 0.5 * (R02_12 * L12 - (R2 * s122 - R1 * s121));
  double R02_20 = R22 - s202 * s202;
  double IL1_20 = 0.5 * (R02_20 * L20 - (R0 * s200 - R2 * s202));
  for (int c = 0; c < 3; c++) IL1[c] = IL1_01 * T.m01[c] + IL1_12 * T.m12[c] + IL1_20 * T.m20[c];
}

// Close-pair outer step (src/Impedance.cpp:194-250) for one outer point v.
__device__ __forceinline__ void outer_close(const Medium &md, const TriGeom &S, const double *v, double w, const Inner &m,
                            const double (*p0t)[3], int nt, const double (*p0s)[3], const double (*pbs)[3], int ns,
                            double *Lr, double *Li, double *Kr, double *Ki) {
  double ISn3h, ISn1, IS1, IL1[3], IL3[3];
  eh_nonsmooth(S, v, ISn3h, ISn1, IS1, IL1, IL3);
  double Ai2 = 0.5 * S.Ai;
  // IL13 = Ai2 (IL1 + (k2/3) IL3), ISn11 = Ai2 (ISn1 + k2 IS1)
  C2 c13 = cmk(BEM_ONETHIRD * md.k2_re, BEM_ONETHIRD * md.k2_im);
  double IL13r[3], IL13i[3];
  for (int c = 0; c < 3; c++) {
    IL13r[c] = Ai2 * (IL1[c] + c13.re * IL3[c]);
    IL13i[c] = Ai2 * (c13.im * IL3[c]);
  }
  C2 ISn11 = cmk(Ai2 * (ISn1 + md.k2_re * IS1), Ai2 * (md.k2_im * IS1));
  double vp[3] = {v[0] - S.p0[0], v[1] - S.p0[1], v[2] - S.p0[2]};
  double h = dot3(vp, S.n);
  C2 L2 = cmk(2.0 * ISn11.re + m.G0.re, 2.0 * ISn11.im + m.G0.im);
  C2 ksq = cmk(md.ksq_re, md.ksq_im);
  bool coplanar = fabs(h) < 1e-8 * S.elen;
  double rho[3] = {v[0] - h * S.n[0], v[1] - h * S.n[1], v[2] - h * S.n[2]};
  // K31 = (Ai2 k2) u, u = IL1 - h ISn1 n ; ISn3hn = Ai2 ISn3h n
  C2 ak2 = cmk(Ai2 * md.k2_re, Ai2 * md.k2_im);
  double hI = h * ISn1;
  double u[3] = {IL1[0] - hI * S.n[0], IL1[1] - hI * S.n[1], IL1[2] - hI * S.n[2]};
  double s3 = Ai2 * ISn3h;
  double ISn3hn[3] = {s3 * S.n[0], s3 * S.n[1], s3 * S.n[2]};
  const double *base = coplanar ? v : rho;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double phi[3] = {0.5 * (v[0] - p0t[i][0]), 0.5 * (v[1] - p0t[i][1]), 0.5 * (v[2] - p0t[i][2])};
    C2 a = cmk(dot3(phi, m.G1r), dot3(phi, m.G1i));
    C2 c = cmk(dot3(phi, m.D1r), dot3(phi, m.D1i));
    C2 dIL13 = cmk(dot3(phi, IL13r), dot3(phi, IL13i));
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (i >= nt || j >= ns) continue;
      double b = dot3(phi, pbs[j]);
      C2 gp = cmk(0.5 * (a.re - b * m.G0.re), 0.5 * (a.im - b * m.G0.im));
      double bj[3] = {base[0] - p0s[j][0], base[1] - p0s[j][1], base[2] - p0s[j][2]};
      double pb = dot3(phi, bj);
      C2 k21 = cmk(dIL13.re + ISn11.re * pb, dIL13.im + ISn11.im * pb);
      C2 kg = cmul(ksq, cmk(gp.re + k21.re, gp.im + k21.im));
      int ij = 3 * i + j;
      Lr[ij] += w * (kg.re - L2.re);
      Li[ij] += w * (kg.im - L2.im);
      if (!coplanar) {
        double e[3];
        cross3(pbs[j], phi, e);
        double dr = 0.5 * (c.re - dot3(m.D0r, e));
        double di = 0.5 * (c.im - dot3(m.D0i, e));
        // vts_ij x ISn3hn, vts_ij = p0s_j - p0t_i
        double vts[3] = {p0s[j][0] - p0t[i][0], p0s[j][1] - p0t[i][1], p0s[j][2] - p0t[i][2]};
        double vx[3];
        cross3(vts, ISn3hn, vx);
        double t1 = dot3(phi, vx);
        // K41_j = (Ai2 k2) (u x (v - p0s_j))
        double vj[3] = {v[0] - p0s[j][0], v[1] - p0s[j][1], v[2] - p0s[j][2]};
        double ux[3];
        cross3(u, vj, ux);
        double t2 = dot3(phi, ux);
        Kr[ij] -= w * (dr + t1 + ak2.re * t2);
        Ki[ij] -= w * (di + ak2.im * t2);
      }
    }
  }
}

__device__ __forceinline__ void store_tile(double *tiles, size_t npairs, size_t idx, const double *Lr, const double *Li,
                                           const double *Kr, const double *Ki) {
#pragma unroll
  for (int q = 0; q < 9; q++) {
    tiles[(size_t)(2 * q) * npairs + idx] = Lr[q];
    tiles[(size_t)(2 * q + 1) * npairs + idx] = Li[q];
    tiles[(size_t)(18 + 2 * q) * npairs + idx] = Kr[q];
    tiles[(size_t)(18 + 2 * q + 1) * npairs + idx] = Ki[q];
  }
}

}  // namespace

// ------------------------------------------------------------ dipole source

// Per triangle of a source's region, the two dot products of
// DipolePointSource::EHTriangle (src/Source.cpp:121-165) for each half on
// the triangle, before the per-RWG factor c = +-len/(4 pi) (and Zeps for E):
//   dE_i = j . (E1 - k^2 (Gphi_i + IL13 + ISn11 (rho - p0_i)))
//   dH_i = j . -(DGxphi_i + K31 x (p - p0_i))
// The smooth part is the 21-point rule on the frame half with the smooth
// Green's function and its gradient (src/GreensFunction.cpp:54-69, R = |p -
// r'|), the singular part EHNonSmoothIntegral at p.  tb: nsrc x T x 3 pairs
// (dE, dH); blockIdx.y = source.
__global__ void __launch_bounds__(256) dipole_tri_kernel(const TriGeom *__restrict__ tg, int T,
                                                         const DipoleSrc *__restrict__ src,
                                                         double2 *__restrict__ tb) {
  const DipoleSrc &S = src[blockIdx.y];
  const int t = S.t_lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.t_hi) return;
  const TriGeom &G = tg[t];
  const Medium md = S.md;
  const double *p = S.p, *jv = S.j;
  const int nh = G.nh;
  double DGr[3] = {0, 0, 0}, DGi[3] = {0, 0, 0};
  double Gpr[3][3] = {}, Gpi[3][3] = {}, Xr[3][3] = {}, Xi[3][3] = {};
  for (int q = 0; q < 21; q++) {
    double v[3];
    quad_point(G, BEM_QUAD_OFF_21 + q, v);
    const double w = c_quad[BEM_QUAD_OFF_21 + q][2];
    double d[3] = {p[0] - v[0], p[1] - v[1], p[2] - v[2]};
    double R2 = dot3(d, d), R, invR;
    root_and_inverse(R2, R, invR);
    double ar = md.ik_re * R, ai = md.ik_im * R;
    double er, ei;
    cexp_i(ar, ai, er, ei);
    // gs = (e - 1)/R - k2 R ; q = (e (1 - ikR) + k2 R^2 - 1)/R^3 ; grad = q d
    C2 gs = cmk((er - 1.0) * invR - md.k2_re * R, ei * invR - md.k2_im * R);
    C2 tq = cmul(cmk(er, ei), cmk(1.0 - ar, -ai));
    double invR3 = invR * invR * invR;
    C2 qq = cmk((tq.re + md.k2_re * R2 - 1.0) * invR3, (tq.im + md.k2_im * R2) * invR3);
    gs.re *= w;
    gs.im *= w;
    double ggr[3], ggi[3];
    for (int c = 0; c < 3; c++) {
      ggr[c] = w * (qq.re * d[c]);
      ggi[c] = w * (qq.im * d[c]);
      DGr[c] += ggr[c];
      DGi[c] += ggi[c];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (i >= nh) break;
      double phi[3] = {0.5 * (v[0] - G.hp0[i][0]), 0.5 * (v[1] - G.hp0[i][1]), 0.5 * (v[2] - G.hp0[i][2])};
      double xr[3], xi[3];
      cross3(ggr, phi, xr);
      cross3(ggi, phi, xi);
      for (int c = 0; c < 3; c++) {
        Gpr[i][c] += gs.re * phi[c];
        Gpi[i][c] += gs.im * phi[c];
        Xr[i][c] += xr[c];
        Xi[i][c] += xi[c];
      }
    }
  }
  double ISn3h, ISn1, IS1, IL1[3], IL3[3], ILn1[3];
  eh_nonsmooth(G, p, ISn3h, ISn1, IS1, IL1, IL3, ILn1);
  const C2 k2 = cmk(md.k2_re, md.k2_im), ksq = cmk(md.ksq_re, md.ksq_im);
  const double Ai = G.Ai, Ai2 = 0.5 * Ai;
  const C2 c13 = cmk(BEM_ONETHIRD * k2.re, BEM_ONETHIRD * k2.im);
  const C2 ISn11 = cmk(Ai2 * (ISn1 + k2.re * IS1), Ai2 * (k2.im * IS1));
  const double h = (p[0] - G.p0[0]) * G.n[0] + (p[1] - G.p0[1]) * G.n[1] + (p[2] - G.p0[2]) * G.n[2];
  const C2 kh = cmk(k2.re * h * ISn1, k2.im * h * ISn1);   // (k2 h) ISn1
  double IL13r[3], IL13i[3], K31r[3], K31i[3], E1r[3], E1i[3], rho[3];
  for (int c = 0; c < 3; c++) {
    IL13r[c] = Ai2 * (IL1[c] + c13.re * IL3[c]);
    IL13i[c] = Ai2 * (c13.im * IL3[c]);
    rho[c] = p[c] - h * G.n[c];
    const double u = IL1[c] - h * ISn1 * G.n[c];
    K31r[c] = Ai2 * ((ILn1[c] + ISn3h * G.n[c]) + k2.re * u);
    K31i[c] = Ai2 * (k2.im * u);
    E1r[c] = DGr[c] + Ai * ((ISn3h - kh.re) * G.n[c] + (ILn1[c] + k2.re * IL1[c]));
    E1i[c] = DGi[c] + Ai * ((-kh.im) * G.n[c] + k2.im * IL1[c]);
  }
  double2 *o = tb + ((size_t)blockIdx.y * T + t) * 6;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (i >= nh) break;
    double dEr = 0, dEi = 0, dHr = 0, dHi = 0;
    double pq[3] = {p[0] - G.hp0[i][0], p[1] - G.hp0[i][1], p[2] - G.hp0[i][2]};
    double kr[3], ki[3];
    cross3(K31r, pq, kr);
    cross3(K31i, pq, ki);
    for (int c = 0; c < 3; c++) {
      const double rc = rho[c] - G.hp0[i][c];
      const double ar = (Gpr[i][c] + IL13r[c]) + (ISn11.re * rc);
      const double ai = (Gpi[i][c] + IL13i[c]) + (ISn11.im * rc);
      const double er = E1r[c] - (ksq.re * ar - ksq.im * ai);
      const double eim = E1i[c] - (ksq.re * ai + ksq.im * ar);
      dEr += jv[c] * er;
      dEi += jv[c] * eim;
      dHr -= jv[c] * (Xr[i][c] + kr[c]);
      dHi -= jv[c] * (Xi[i][c] + ki[c]);
    }
    o[2 * i] = make_double2(dEr, dEi);
    o[2 * i + 1] = make_double2(dHr, dHi);
  }
}

// B(2e) = sum over the RWG's halves in the region of (c Zeps) dE,
// B(2e+1) = sum of c dH, c = len/(4 pi), negated when "external" differs
// from "positive half" (src/Source.cpp:150-161).  B: column blockIdx.y of a
// 2E x nsrc matrix (leading dimension ldb).
__global__ void __launch_bounds__(256) dipole_gather_kernel(const RwgRec *__restrict__ rw, int E, int T,
                                                            const DipoleSrc *__restrict__ src,
                                                            const double2 *__restrict__ tb,
                                                            double2 *__restrict__ B, int64_t ldb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const DipoleSrc &S = src[blockIdx.y];
  const RwgRec r = rw[e];
  const double2 *base = tb + (size_t)blockIdx.y * T * 6;
  const double c0 = BEM_ONEOVER4PI * r.len;
  double2 bE = make_double2(0.0, 0.0), bH = make_double2(0.0, 0.0);
  const int th[2] = {r.tp, r.tn}, sl[2] = {r.sp, r.sn};
  for (int hh = 0; hh < 2; hh++) {
    const int t = th[hh];
    if (t < S.t_lo || t >= S.t_hi || sl[hh] >= 3) continue;
    const double c = (S.external != 0) != (hh == 0) ? -c0 : c0;
    const double2 dE = base[(size_t)t * 6 + 2 * sl[hh]], dH = base[(size_t)t * 6 + 2 * sl[hh] + 1];
    const double czr = c * S.md.Zeps_re, czi = c * S.md.Zeps_im;
    bE.x += czr * dE.x - czi * dE.y;
    bE.y += czr * dE.y + czi * dE.x;
    bH.x += c * dH.x;
    bH.y += c * dH.y;
  }
  double2 *col = B + (int64_t)blockIdx.y * ldb;
  col[2 * (int64_t)e] = bE;
  col[2 * (int64_t)e + 1] = bH;
}

// ------------------------------------------------------------------- flux

#define FLUX_BLOCK 256
// Per triangle, dS of flux(top, t, X) (src/Flux.cpp:100-219):
//   j_k, m_k   vertex currents from the three RWGs, m / omega, j conjugated
//   dS = n . Re(j1 x m1 + j2 x m2 + j3 x m3 + (j1+j2+j3) x (m1+m2+m3)) / (48 A2)
// then a block sum in a fixed tree order.  part: one sum per block, for each
// column (blockIdx.y) of X (2E x ncol, leading dimension ldx).
__global__ void __launch_bounds__(FLUX_BLOCK) flux_kernel(const FluxTri *__restrict__ ft, int nt, double omega,
                                                          const double2 *__restrict__ X, int64_t ldx,
                                                          double *__restrict__ part) {
  __shared__ double red[FLUX_BLOCK];
  const int t = blockIdx.x * FLUX_BLOCK + threadIdx.x;
  double dS = 0.0;
  if (t < nt) {
    const FluxTri &F = ft[t];
    const double2 *x = X + (int64_t)blockIdx.y * ldx;
    double v12[3], v23[3], v31[3], n[3];
    for (int c = 0; c < 3; c++) {
      v12[c] = F.P[1][c] - F.P[0][c];
      v23[c] = F.P[2][c] - F.P[1][c];
      v31[c] = F.P[0][c] - F.P[2][c];
    }
    double v13[3] = {F.P[2][0] - F.P[0][0], F.P[2][1] - F.P[0][1], F.P[2][2] - F.P[0][2]};
    cross3(v12, v13, n);
    const double A2 = sqrt(dot3(n, n));
    for (int c = 0; c < 3; c++) n[c] *= 1.0 / A2;
    // currents at the three vertices: jr/ji, mr/mi [vertex][component]
    double jr[3][3] = {}, ji[3][3] = {}, mr[3][3] = {}, mi[3][3] = {};
    const double *ve[3][2] = {{v31, v23}, {v12, v31}, {v23, v12}};   // (+ at vertex k, - at vertex k+1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double2 xj = x[2 * (int64_t)F.e[k]], xm = x[2 * (int64_t)F.e[k] + 1];
      const double cj_r = F.c[k] * xj.x, cj_i = F.c[k] * xj.y, cm_r = F.c[k] * xm.x, cm_i = F.c[k] * xm.y;
      const int a = k, b = (k + 1) % 3;
      for (int c = 0; c < 3; c++) {
        jr[a][c] += cj_r * ve[k][0][c];
        ji[a][c] += cj_i * ve[k][0][c];
        mr[a][c] += cm_r * ve[k][0][c];
        mi[a][c] += cm_i * ve[k][0][c];
        jr[b][c] -= cj_r * ve[k][1][c];
        ji[b][c] -= cj_i * ve[k][1][c];
        mr[b][c] -= cm_r * ve[k][1][c];
        mi[b][c] -= cm_i * ve[k][1][c];
      }
    }
    const double oi = 1.0 / omega;
    double s[3] = {0, 0, 0}, Jr[3] = {0, 0, 0}, Ji[3] = {0, 0, 0}, Mr[3] = {0, 0, 0}, Mi[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
      for (int c = 0; c < 3; c++) {
        mr[k][c] *= oi;
        mi[k][c] *= oi;
        ji[k][c] = -ji[k][c];
        Jr[c] += jr[k][c];
        Ji[c] += ji[k][c];
        Mr[c] += mr[k][c];
        Mi[c] += mi[k][c];
      }
      // Re(j x m) = jr x mr - ji x mi
      double a1[3], a2[3];
      cross3(jr[k], mr[k], a1);
      cross3(ji[k], mi[k], a2);
      for (int c = 0; c < 3; c++) s[c] += a1[c] - a2[c];
    }
    double a1[3], a2[3];
    cross3(Jr, Mr, a1);
    cross3(Ji, Mi, a2);
    for (int c = 0; c < 3; c++) s[c] += a1[c] - a2[c];
    dS = dot3(n, s) / (48.0 * A2);
  }
  red[threadIdx.x] = dS;
  __syncthreads();
  for (int w = FLUX_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// flux[col] = sum of the column's block sums in block order.
__global__ void flux_finish_kernel(const double *__restrict__ part, int nblk, int ncol, double *__restrict__ out) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  double S = 0.0;
  for (int b = 0; b < nblk; b++) S += part[(size_t)col * nblk + b];
  out[col] = S;
}

// ---------------------------------------------------------------------------
// Dense pair grid: every 6-point pair computed here; 36-point pairs queued.
// Block = 16 x 16 pairs.  tri_lo/tri_hi bound the test rows (row-stripe
// launches) and the source columns cover [col_lo, col_hi).
#define FAR_TB 16
__global__ void __launch_bounds__(256) far_g6_kernel(const TriGeom *__restrict__ tg, int T, Medium md,
                                                     int row_lo, int row_hi, int col_lo, int col_hi,
                                                     double *__restrict__ tiles, size_t npairs,
                                                     int2 *__restrict__ q36, unsigned int *__restrict__ q36_count,
                                                     unsigned int q36_cap) {
  __shared__ double s_out[FAR_TB][6][4];   // outer points + weight
  __shared__ double s_in[FAR_TB][6][4];    // inner points + weight
  __shared__ double s_tp0[FAR_TB][3][3];   // test half p0
  __shared__ double s_spb[FAR_TB][3][3];   // source half p0 - frame p0
  __shared__ double s_sob[FAR_TB][3];      // source frame p0
  __shared__ int s_tnh[FAR_TB], s_snh[FAR_TB];

  int tx = threadIdx.x % FAR_TB, ty = threadIdx.x / FAR_TB;
  int a0 = row_lo + blockIdx.y * FAR_TB;
  int b0 = col_lo + blockIdx.x * FAR_TB;
  // stage: thread (ty, tx) with tx < 6 loads quadrature point tx of test row ty / source col ty
  {
    int t = threadIdx.x;
    if (t < FAR_TB * 6) {
      int k = t / 6, q = t % 6;
      int a = a0 + k;
      if (a < row_hi) {
        double v[3];
        quad_point(tg[a], BEM_QUAD_OFF_6 + q, v);
        s_out[k][q][0] = v[0]; s_out[k][q][1] = v[1]; s_out[k][q][2] = v[2];
        s_out[k][q][3] = c_quad[BEM_QUAD_OFF_6 + q][2];
      }
    } else if (t < 2 * FAR_TB * 6) {
      int u = t - FAR_TB * 6;
      int k = u / 6, q = u % 6;
      int b = b0 + k;
      if (b < col_hi) {
        double v[3];
        quad_point(tg[b], BEM_QUAD_OFF_6 + q, v);
        s_in[k][q][0] = v[0]; s_in[k][q][1] = v[1]; s_in[k][q][2] = v[2];
        s_in[k][q][3] = c_quad[BEM_QUAD_OFF_6 + q][2];
      }
    } else if (t < 2 * FAR_TB * 6 + FAR_TB) {
      int k = t - 2 * FAR_TB * 6;
      int a = a0 + k;
      if (a < row_hi) {
        const TriGeom &A = tg[a];
        s_tnh[k] = A.nh;
        for (int i = 0; i < 3; i++)
          for (int c = 0; c < 3; c++) s_tp0[k][i][c] = A.hp0[i][c];
      }
    } else if (t < 2 * FAR_TB * 6 + 2 * FAR_TB) {
      int k = t - 2 * FAR_TB * 6 - FAR_TB;
      int b = b0 + k;
      if (b < col_hi) {
        const TriGeom &B = tg[b];
        s_snh[k] = B.nh;
        for (int c = 0; c < 3; c++) s_sob[k][c] = B.p0[c];
        for (int j = 0; j < 3; j++)
          for (int c = 0; c < 3; c++) s_spb[k][j][c] = B.hp0[j][c] - B.p0[c];
      }
    }
  }
  __syncthreads();
  int a = a0 + ty, b = b0 + tx;
  if (a >= row_hi || b >= col_hi) return;
  int cls = classify_pair(tg[a], tg[b], a == b);
  if (cls == BEM_CLASS_CLOSE) return;
  if (cls != BEM_CLASS_G6) {
    unsigned int slot = atomicAdd(q36_count, 1u);
    if (slot < q36_cap) q36[slot] = make_int2(a, b);
    return;
  }
  double Lr[9], Li[9], Kr[9], Ki[9];
#pragma unroll
  for (int q = 0; q < 9; q++) Lr[q] = Li[q] = Kr[q] = Ki[q] = 0.0;
  int nt = s_tnh[ty], ns = s_snh[tx];
  const double *ob = s_sob[tx];
  for (int o = 0; o < 6; o++) {
    const double *r = s_out[ty][o];
    Inner m;
    inner_zero(m);
#pragma unroll 2
    for (int q = 0; q < 6; q++) inner_point<false>(md, r, s_in[tx][q], ob, s_in[tx][q][3], m);
    outer_far(md, r, s_out[ty][o][3], m, s_tp0[ty], nt, s_spb[tx], ns, Lr, Li, Kr, Ki);
  }
  store_tile(tiles, npairs, (size_t)(a - row_lo) * (size_t)(col_hi - col_lo) + (size_t)(b - col_lo), Lr, Li, Kr, Ki);
}

// Block reduction of 36 accumulators; result valid in thread 0.
template <int NT>
__device__ __forceinline__ void wave_reduce9(double (&v)[9]) {
#pragma unroll
  for (int q = 0; q < 9; q++) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_down(x, off, 64);
    v[q] = x;
  }
}

template <int NT>
__device__ __forceinline__ void block_reduce36(double (&Lr)[9], double (&Li)[9], double (&Kr)[9], double (&Ki)[9],
                                               double *red /* [NT/64][36] */) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  wave_reduce9<NT>(Lr);
  wave_reduce9<NT>(Li);
  wave_reduce9<NT>(Kr);
  wave_reduce9<NT>(Ki);
  if (NT > 64) {
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 9; q++) {
        red[wv * 36 + q] = Lr[q];
        red[wv * 36 + 9 + q] = Li[q];
        red[wv * 36 + 18 + q] = Kr[q];
        red[wv * 36 + 27 + q] = Ki[q];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < 9; q++) {
        double a = red[q], b = red[9 + q], c = red[18 + q], d = red[27 + q];
        for (int w2 = 1; w2 < NT / 64; w2++) {
          a += red[w2 * 36 + q];
          b += red[w2 * 36 + 9 + q];
          c += red[w2 * 36 + 18 + q];
          d += red[w2 * 36 + 27 + q];
        }
        Lr[q] = a; Li[q] = b; Kr[q] = c; Ki[q] = d;
      }
    }
  }
}

// Smooth inner sums of the close pairs, one thread per (pair, outer point)
// of the listed pairs [0, np): the 91-point inner rule on the source triangle
// at one 105-point outer point of the test triangle, summed in the order of
// pair_block_kernel's own loop, stored to mom for the pair's outer part.
// A flat grid keeps every lane busy, and without the analytic part's state
// the kernel holds more waves per SIMD.
__global__ void __launch_bounds__(256) close_inner_kernel(const TriGeom *__restrict__ tg, Medium md,
                                                          const int2 *__restrict__ pairs, unsigned int np,
                                                          int row_lo, int row_hi, int col_lo, int col_hi,
                                                          double *__restrict__ mom, size_t mstride) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)np * 105) return;
  const unsigned int p = (unsigned int)(t / 105);
  const int o = (int)(t % 105);
  const int2 pr = pairs[p];
  const int a = pr.x, b = pr.y;
  if (a < row_lo || a >= row_hi || b < col_lo || b >= col_hi) return;   // outside the range: not used
  const TriGeom &A = tg[a];
  const TriGeom &B = tg[b];
  double r[3];
  quad_point(A, BEM_QUAD_OFF_105 + o, r);
  Inner m;
  inner_zero(m);
  for (int q = 0; q < 91; q++) {
    double v[4];
    quad_point(B, BEM_QUAD_OFF_91 + q, v);
    v[3] = c_quad[BEM_QUAD_OFF_91 + q][2];
    inner_point<true>(md, r, v, B.p0, v[3], m);
  }
  inner_store(m, mom, mstride, t);
}

// One block per listed pair.  CLOSE: Gauss105 outer x Gauss91 smooth inner +
// analytic terms + neighbour edge correction; otherwise 36 x 36 regular.
// PRE: the inner sums of every (pair, outer point) were formed beforehand by
// close_inner_kernel and are read from mom (INNER_PLANES planes of mstride
// doubles, index pair * NOUT + outer point) instead of being summed here.
template <int NOUT, int NIN, bool CLOSE, int NT, bool PRE = false>
__global__ void __launch_bounds__(NT) pair_block_kernel(const TriGeom *__restrict__ tg, int T, Medium md,
                                                        const int2 *__restrict__ pairs,
                                                        const unsigned int *__restrict__ count_dev,
                                                        unsigned int count_host, int row_lo, int row_hi,
                                                        int col_lo, int col_hi, double *__restrict__ tiles,
                                                        size_t npairs, const double *__restrict__ mom = nullptr,
                                                        size_t mstride = 0) {
  __shared__ double s_in[NIN][4];
  __shared__ double s_red[(NT / 64) * 36];
  __shared__ double s_tp0[3][3], s_sp0[3][3], s_spb[3][3];
  const int offo = CLOSE ? BEM_QUAD_OFF_105 : BEM_QUAD_OFF_36;
  const int offi = CLOSE ? BEM_QUAD_OFF_91 : BEM_QUAD_OFF_36;
  unsigned int count = count_host;
  if (count_dev && *count_dev < count_host) count = *count_dev;   // queue may have overflowed: never read past it
  for (unsigned int p = blockIdx.x; p < count; p += gridDim.x) {
    int2 pr = pairs[p];
    int a = pr.x, b = pr.y;
    if (a < row_lo || a >= row_hi || b < col_lo || b >= col_hi) continue;   // uniform per block
    const TriGeom &A = tg[a];
    const TriGeom &B = tg[b];
    __syncthreads();
    if (!PRE) {
      for (int t = threadIdx.x; t < NIN; t += NT) {
        double v[3];
        quad_point(B, offi + t, v);
        s_in[t][0] = v[0]; s_in[t][1] = v[1]; s_in[t][2] = v[2];
        s_in[t][3] = c_quad[offi + t][2];
      }
    }
    if (threadIdx.x < 9) {
      int i = threadIdx.x / 3, c = threadIdx.x % 3;
      s_tp0[i][c] = A.hp0[i][c];
      s_sp0[i][c] = B.hp0[i][c];
      s_spb[i][c] = B.hp0[i][c] - B.p0[c];
    }
    __syncthreads();
    double Lr[9], Li[9], Kr[9], Ki[9];
#pragma unroll
    for (int q = 0; q < 9; q++) Lr[q] = Li[q] = Kr[q] = Ki[q] = 0.0;
    int nt = A.nh, ns = B.nh;
    for (int o = threadIdx.x; o < NOUT; o += NT) {
      double r[3];
      quad_point(A, offo + o, r);
      double w = c_quad[offo + o][2];
      Inner m;
      if constexpr (PRE) {
        inner_load(m, mom, mstride, (size_t)p * NOUT + o);
      } else {
        inner_zero(m);
        for (int q = 0; q < NIN; q++) inner_point<CLOSE>(md, r, s_in[q], B.p0, s_in[q][3], m);
      }
      if (CLOSE)
        outer_close(md, B, r, w, m, s_tp0, nt, s_sp0, s_spb, ns, Lr, Li, Kr, Ki);
      else
        outer_far(md, r, w, m, s_tp0, nt, s_spb, ns, Lr, Li, Kr, Ki);
    }
    if (CLOSE && a != b) {
      // neighbour edge correction (src/Impedance.cpp:253-274, :301-324):
      // 15-point line rule on each of the 3 source edges, LKEdgeIntegral on
      // the test frame; each point's contribution to K_ij is linear, so it is
      // added to this thread's accumulator before the block reduction.
      for (int e = threadIdx.x; e < 45; e += NT) {
        int edge = e / 15, q = e % 15;
        const double *e0 = edge == 0 ? B.p0 : (edge == 1 ? B.p1 : B.p2);
        const double *ev = edge == 0 ? B.v01 : (edge == 1 ? B.v12 : B.v20);
        const double *me = edge == 0 ? B.m01 : (edge == 1 ? B.m12 : B.m20);
        double nvi = edge == 0 ? B.normv01i : (edge == 1 ? B.normv12i : B.normv20i);
        double tq = c_line15[q][0], wq = c_line15[q][1];
        double v[3];
        for (int c = 0; c < 3; c++) v[c] = e0[c] + tq * ev[c];
        double ISn1, IL1[3];
        lk_edge(A, v, ISn1, IL1);
        double vp[3] = {v[0] - A.p0[0], v[1] - A.p0[1], v[2] - A.p0[2]};
        double h = dot3(vp, A.n);
        double rho[3] = {v[0] - h * A.n[0], v[1] - h * A.n[1], v[2] - h * A.n[2]};
        double cc = 0.125 * A.Ai * B.Ai;
#pragma unroll
        for (int i = 0; i < 3; i++) {
          double kv[3];
          for (int c = 0; c < 3; c++) kv[c] = wq * (IL1[c] + ISn1 * (rho[c] - s_tp0[i][c]));
#pragma unroll
          for (int j = 0; j < 3; j++) {
            if (i >= nt || j >= ns) continue;
            double vv[3] = {s_tp0[i][0] - s_sp0[j][0], s_tp0[i][1] - s_sp0[j][1], s_tp0[i][2] - s_sp0[j][2]};
            double vm[3];
            cross3(vv, me, vm);
            Kr[3 * i + j] += cc * (dot3(vm, kv) / nvi);
          }
        }
      }
    }
    block_reduce36<NT>(Lr, Li, Kr, Ki, s_red);
    if (threadIdx.x == 0)
      store_tile(tiles, npairs, (size_t)(a - row_lo) * (size_t)(col_hi - col_lo) + (size_t)(b - col_lo), Lr, Li, Kr,
                 Ki);
  }
}

// ---------------------------------------------------------------------------
// Gather (src/Impedance.cpp:365-384): for RWG pair (i, j) sum the four half
// pairs with c = l_i l_j / 4pi and the half sign, then
//   Z(2i,2j) = Zeps L, Z(2i,2j+1) = K, Z(2i+1,2j) = -K, Z(2i+1,2j+1) = Zmu L.
// With ADD the values are added to Z (W = Z_ext + Z_int on the diagonal blocks);
// Zcopy (optional) receives the same values (the separate G_b matrices).
// Rows are RWGs [r_lo, r_hi) of the tree; the tile buffer holds test rows
// [row_lo, ..) x source columns [col_lo, col_lo + tld) of triangles.
#define GATHER_GI 32   // test RWGs (rows) per block
#define GATHER_GJ 16   // source RWGs (columns) per block
__global__ void __launch_bounds__(256) gather_kernel(const double *__restrict__ tiles, size_t npairs, int tld,
                                                     int row_lo, int col_lo, const RwgRec *__restrict__ rw, int r_lo,
                                                     int r_hi, int c_lo, int c_hi, Medium md,
                                                     double2 *__restrict__ Z, int64_t ldz, int zr0, int zc0, int add,
                                                     double2 *__restrict__ Zcopy, int64_t ldc, int cr0, int cc0) {
  // The block's 2GI x 2GJ output goes through LDS: lanes first run along
  // source RWGs (neighbouring tile columns, so the tile planes are read
  // nearly contiguously), then along output rows (coalesced column writes).
  __shared__ double2 sz[2 * GATHER_GJ][2 * GATHER_GI + 1];
  const int i0 = r_lo + blockIdx.x * GATHER_GI, j0 = c_lo + blockIdx.y * GATHER_GJ;
  const int t = threadIdx.x, jl = t % GATHER_GJ;
  const int j = j0 + jl;
  if (j < c_hi) {
    const RwgRec rj = rw[j];
    const int tb[2] = {rj.tp, rj.tn}, sb[2] = {rj.sp, rj.sn};
    for (int il = t / GATHER_GJ; il < GATHER_GI; il += 256 / GATHER_GJ) {
      const int i = i0 + il;
      if (i >= r_hi) break;
      const RwgRec ri = rw[i];
      const int ta[2] = {ri.tp, ri.tn}, sa[2] = {ri.sp, ri.sn};
      double Lr = 0, Li = 0, Kr = 0, Ki = 0;
      for (int hi = 0; hi < 2; hi++)
        for (int hj = 0; hj < 2; hj++) {
          double sg = (hi == hj) ? 1.0 : -1.0;
          size_t idx = (size_t)(ta[hi] - row_lo) * (size_t)tld + (size_t)(tb[hj] - col_lo);
          int q = 3 * sa[hi] + sb[hj];
          Lr += sg * tiles[(size_t)(2 * q) * npairs + idx];
          Li += sg * tiles[(size_t)(2 * q + 1) * npairs + idx];
          Kr += sg * tiles[(size_t)(18 + 2 * q) * npairs + idx];
          Ki += sg * tiles[(size_t)(18 + 2 * q + 1) * npairs + idx];
        }
      double c = BEM_ONEOVER4PI * ri.len * rj.len;
      Lr *= c; Li *= c; Kr *= c; Ki *= c;
      sz[2 * jl][2 * il] = make_double2(md.Zeps_re * Lr - md.Zeps_im * Li, md.Zeps_re * Li + md.Zeps_im * Lr);
      sz[2 * jl][2 * il + 1] = make_double2(-Kr, -Ki);
      sz[2 * jl + 1][2 * il] = make_double2(Kr, Ki);
      sz[2 * jl + 1][2 * il + 1] = make_double2(md.Zmu_re * Lr - md.Zmu_im * Li, md.Zmu_re * Li + md.Zmu_im * Lr);
    }
  }
  __syncthreads();
  const int rows = 2 * min(GATHER_GI, r_hi - i0), cols = 2 * min(GATHER_GJ, c_hi - j0);
  const int r = t % (2 * GATHER_GI);
  if (r >= rows) return;
  const int64_t zi = zr0 + 2 * (int64_t)(i0 - r_lo) + r, zc = zc0 + 2 * (int64_t)(j0 - c_lo);
  const int64_t yi = cr0 + 2 * (int64_t)(i0 - r_lo) + r, yc = cc0 + 2 * (int64_t)(j0 - c_lo);
  for (int cc = t / (2 * GATHER_GI); cc < cols; cc += 256 / (2 * GATHER_GI)) {
    double2 v = sz[cc][r];
    double2 *p = Z + (zc + cc) * ldz + zi;
    if (add) {
      double2 o = *p;
      *p = make_double2(o.x + v.x, o.y + v.y);
    } else {
      *p = v;
    }
    if (Zcopy) Zcopy[(yc + cc) * ldc + yi] = v;
  }
}

// ------------------------------------------------------------------ launchers
extern "C" {

int bem_launch_far_g6(const TriGeom *tg, int T, const Medium *md, int row_lo, int row_hi, int col_lo, int col_hi,
                      double *tiles, size_t npairs, int2 *q36, unsigned int *q36_count, unsigned int q36_cap,
                      hipStream_t st) {
  dim3 grid((col_hi - col_lo + FAR_TB - 1) / FAR_TB, (row_hi - row_lo + FAR_TB - 1) / FAR_TB);
  if (grid.x == 0 || grid.y == 0) return 0;
  hipLaunchKernelGGL(far_g6_kernel, grid, dim3(256), 0, st, tg, T, *md, row_lo, row_hi, col_lo, col_hi, tiles,
                     npairs, q36, q36_count, q36_cap);
  return (int)hipGetLastError();
}

int bem_launch_far_g36(const TriGeom *tg, int T, const Medium *md, const int2 *pairs, const unsigned int *count_dev,
                       unsigned int max_count, int row_lo, int row_hi, int col_lo, int col_hi, double *tiles,
                       size_t npairs, hipStream_t st) {
  unsigned int blocks = max_count < 4096u ? max_count : 4096u;
  if (blocks == 0) return 0;
  hipLaunchKernelGGL((pair_block_kernel<36, 36, false, 64>), dim3(blocks), dim3(64), 0, st, tg, T, *md, pairs,
                     count_dev, max_count, row_lo, row_hi, col_lo, col_hi, tiles, npairs,
                     (const double *)nullptr, (size_t)0);
  return (int)hipGetLastError();
}

// Close pairs in chunks of at most mom_pairs pairs: the inner sums of a chunk
// (close_inner_kernel) into mom (INNER_PLANES x 105 x mom_pairs doubles), then
// the outer part per pair (pair_block_kernel reading them).  mom == nullptr:
// one pass of pair_block_kernel doing both.
size_t bem_close_mom_doubles(unsigned int mom_pairs) { return (size_t)INNER_PLANES * 105 * mom_pairs; }

int bem_launch_close(const TriGeom *tg, int T, const Medium *md, const int2 *pairs, unsigned int count, int row_lo,
                     int row_hi, int col_lo, int col_hi, double *tiles, size_t npairs, double *mom,
                     unsigned int mom_pairs, hipStream_t st) {
  if (count == 0) return 0;
  if (!mom || mom_pairs == 0) {
    unsigned int blocks = count < 65535u ? count : 65535u;
    hipLaunchKernelGGL((pair_block_kernel<105, 91, true, 128>), dim3(blocks), dim3(128), 0, st, tg, T, *md, pairs,
                       (const unsigned int *)nullptr, count, row_lo, row_hi, col_lo, col_hi, tiles, npairs,
                       (const double *)nullptr, (size_t)0);
    return (int)hipGetLastError();
  }
  const size_t mstride = (size_t)105 * mom_pairs;
  for (unsigned int p0 = 0; p0 < count; p0 += mom_pairs) {
    const unsigned int np = count - p0 < mom_pairs ? count - p0 : mom_pairs;
    const unsigned int ib = (unsigned int)(((size_t)np * 105 + 255) / 256);
    hipLaunchKernelGGL(close_inner_kernel, dim3(ib), dim3(256), 0, st, tg, *md, pairs + p0, np, row_lo, row_hi,
                       col_lo, col_hi, mom, mstride);
    const unsigned int blocks = np < 65535u ? np : 65535u;
    hipLaunchKernelGGL((pair_block_kernel<105, 91, true, 128, true>), dim3(blocks), dim3(128), 0, st, tg, T, *md,
                       pairs + p0, (const unsigned int *)nullptr, np, row_lo, row_hi, col_lo, col_hi, tiles, npairs,
                       (const double *)mom, mstride);
    int e = (int)hipGetLastError();
    if (e) return e;
  }
  return 0;
}

// flag <- 1 when the 36-point pair queue overflowed its capacity.
__global__ void q36_check_kernel(const unsigned int *count, unsigned int cap, int *flag) {
  if (threadIdx.x == 0 && *count > cap) *flag = 1;
}

int bem_launch_q36_check(const unsigned int *count, unsigned int cap, int *flag, hipStream_t st) {
  hipLaunchKernelGGL(q36_check_kernel, dim3(1), dim3(64), 0, st, count, cap, flag);
  return (int)hipGetLastError();
}

// Dipole right-hand sides: src (nsrc device records), tb (nsrc x T x 6
// double2 workspace), B (2E x nsrc, leading dimension ldb).  tmax: the
// largest t_hi - t_lo over the sources.
int bem_launch_dipole_rhs(const TriGeom *tg, int T, const RwgRec *rw, int E, const void *src, int nsrc, int tmax,
                          double2 *tb, double2 *B, int64_t ldb, hipStream_t st) {
  if (nsrc <= 0) return 0;
  if (tmax > 0) {
    dim3 g1((tmax + 255) / 256, nsrc);
    hipLaunchKernelGGL(dipole_tri_kernel, g1, dim3(256), 0, st, tg, T, (const DipoleSrc *)src, tb);
  }
  if (E > 0) {
    dim3 g2((E + 255) / 256, nsrc);
    hipLaunchKernelGGL(dipole_gather_kernel, g2, dim3(256), 0, st, rw, E, T, (const DipoleSrc *)src, tb, B, ldb);
  }
  return (int)hipGetLastError();
}

// Flux of ncol solution columns through a surface of nt triangles: part
// holds ncol x ceil(nt / 256) block sums, out ncol values (device).
size_t bem_flux_partials(int nt, int ncol) { return (size_t)((nt + FLUX_BLOCK - 1) / FLUX_BLOCK) * ncol; }

int bem_launch_flux(const void *ft, int nt, double omega, const double2 *X, int64_t ldx, int ncol, double *part,
                    double *out, hipStream_t st) {
  if (ncol <= 0) return 0;
  const int nblk = (nt + FLUX_BLOCK - 1) / FLUX_BLOCK;
  if (nblk > 0)
    hipLaunchKernelGGL(flux_kernel, dim3(nblk, ncol), dim3(FLUX_BLOCK), 0, st, (const FluxTri *)ft, nt, omega, X,
                       ldx, part);
  hipLaunchKernelGGL(flux_finish_kernel, dim3((ncol + 63) / 64), dim3(64), 0, st, part, nblk, ncol, out);
  return (int)hipGetLastError();
}

int bem_launch_gather(const double *tiles, size_t npairs, int tld, int row_lo, int col_lo, const RwgRec *rw, int r_lo,
                      int r_hi, int c_lo, int c_hi, const Medium *md, double2 *Z, int64_t ldz, int zr0, int zc0,
                      int add, double2 *Zcopy, int64_t ldc, int cr0, int cc0, hipStream_t st) {
  if (r_hi <= r_lo || c_hi <= c_lo) return 0;
  dim3 grid((r_hi - r_lo + GATHER_GI - 1) / GATHER_GI, (c_hi - c_lo + GATHER_GJ - 1) / GATHER_GJ);
  hipLaunchKernelGGL(gather_kernel, grid, dim3(256), 0, st, tiles, npairs, tld, row_lo, col_lo, rw, r_lo, r_hi, c_lo,
                     c_hi, *md, Z, ldz, zr0, zc0, add, Zcopy, ldc, cr0, cc0);
  return (int)hipGetLastError();
}

}  // extern "C"
