// Kernels of the low-rank coupling between the two bodies, hand-written for
// gfx950: the cross blocks B = Z_ext(1,2) and C = Z_ext(2,1) of two disjoint
// bodies are far-field couplings of numerically low rank, stored as factor
// pairs U V (the reference's Rk, src/util/Rk.h) instead of dense blocks.
//
//  chol_inv64     Cholesky G = R^H R of a 64 x 64 Gram block and R^-1, for
//                 the CholQR2 orthonormalisation of the randomized range
//                 finder's 64-column blocks
//  transpose_op   dst = op(src)^T with optional conjugation and J/M signs
//  aca_*          adaptive cross approximation with full pivoting on a dense
//                 device block (src/util/DenseMatrix.h:213-279): one fused
//                 pass per step (rank-1 update of the residual + argmax), and
//                 a one-workgroup pivot step (cross vectors, norms, stop rule)
//  jacobi_round   one round of a parallel one-sided Jacobi SVD (complex
//                 rotations of column pairs, round-robin order), the SVD
//                 behind Rk::truncate (src/util/Rk.cpp:4-53)
#include "bem_common.h"

#include <algorithm>

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// ------------------------------------------------------------- CholQR
// G (w x w, leading dimension ldg, Hermitian positive definite, w <= 64) =
// R^H R with R upper triangular (right-looking, one column per step, the
// block in LDS); Rinv (64 x 64, leading dimension 64) <- R^-1 with the
// identity beyond w.  A pivot that is not positive -- G singular to working
// precision, so the block's columns are dependent -- sets info[0] and is
// replaced by 1 so the kernel still ends.
__global__ void __launch_bounds__(256) chol_inv64_kernel(const double2 *__restrict__ G, int64_t ldg, int w,
                                                        double2 *__restrict__ Rinv, int *__restrict__ info) {
  __shared__ double2 s[64][65];   // s[i][k]: R(i, k) for k >= i once step i is done
  __shared__ double2 x[64][65];   // x[i][k]: R^-1(i, k)
  const int t = threadIdx.x;
  for (int e = t; e < 64 * 64; e += 256) {
    const int i = e & 63, k = e >> 6;
    s[i][k] = (i < w && k < w) ? G[(int64_t)k * ldg + i] : make_double2(i == k ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int j = 0; j < 64; j++) {
    if (t == 0) {
      double d = s[j][j].x;
      if (!(d > 0.0)) {
        info[0] = 1;
        d = 1.0;
      }
      s[j][j] = make_double2(sqrt(d), 0.0);
    }
    __syncthreads();
    const double rjj = s[j][j].x;
    for (int k = j + 1 + t; k < 64; k += 256) s[j][k] = make_double2(s[j][k].x / rjj, s[j][k].y / rjj);
    __syncthreads();
    const int nt = 63 - j;
    for (int e = t; e < nt * nt; e += 256) {
      const int i = j + 1 + e % nt, k = j + 1 + e / nt;
      if (k >= i) {
        const double2 a = cconj(s[j][i]), b = s[j][k];
        s[i][k].x -= a.x * b.x - a.y * b.y;
        s[i][k].y -= a.x * b.y + a.y * b.x;
      }
    }
    __syncthreads();
  }
  // column k of R^-1 by back substitution (thread k)
  if (t < 64) {
    const int k = t;
    for (int i = 63; i > k; i--) x[i][k] = make_double2(0.0, 0.0);
    x[k][k] = make_double2(1.0 / s[k][k].x, 0.0);
    for (int i = k - 1; i >= 0; i--) {
      double2 acc = make_double2(0.0, 0.0);
      for (int l = i + 1; l <= k; l++) {
        const double2 p = cmul(s[i][l], x[l][k]);
        acc.x += p.x;
        acc.y += p.y;
      }
      const double inv = 1.0 / s[i][i].x;
      x[i][k] = make_double2(-acc.x * inv, -acc.y * inv);
    }
  }
  __syncthreads();
  for (int e = t; e < 64 * 64; e += 256) {
    const int i = e & 63, k = e >> 6;
    Rinv[(int64_t)k * 64 + i] = x[i][k];
  }
}

// --------------------------------------------------------- transposes
// dst (n x m) <- op(src (m x n))^T, op = conj when conj != 0; with
// sign_rows, row i of dst (an interleaved [J, M] unknown) is negated when i
// is odd.  32 x 32 tiles through LDS.
__global__ void __launch_bounds__(256) transpose_op_kernel(const double2 *__restrict__ src, int64_t lds, int m, int n,
                                                          double2 *__restrict__ dst, int64_t ldd, int conj,
                                                          int sign_rows) {
  __shared__ double2 tl[32][33];
  const int b0 = blockIdx.x * 32, c0 = blockIdx.y * 32;   // rows, columns of src
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int r = b0 + tx, c = c0 + j;
    if (r < m && c < n) tl[j][tx] = src[(int64_t)c * lds + r];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = c0 + tx, c = b0 + j;   // dst(r, c) = op(src(c, r))
    if (r < n && c < m) {
      double2 v = tl[tx][j];
      if (conj) v.y = -v.y;
      if (sign_rows && (r & 1)) v = make_double2(-v.x, -v.y);
      dst[(int64_t)c * ldd + r] = v;
    }
  }
}

// dst (m x n) <- s_i conj(src(i, j)), s_i = -1 on odd rows when sign_rows.
__global__ void conj_sign_kernel(const double2 *__restrict__ src, int64_t lds, int m, int n, double2 *__restrict__ dst,
                                 int64_t ldd, int sign_rows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= m) return;
  double2 v = src[(int64_t)c * lds + r];
  v.y = -v.y;
  if (sign_rows && (r & 1)) v = make_double2(-v.x, -v.y);
  dst[(int64_t)c * ldd + r] = v;
}

// ------------------------------------------------------------------ ACA
// State of one ACA run on the device (src/util/DenseMatrix.h:213-279).
struct AcaState {
  double epsabstot2;   // running sum of ||a||^2 ||b||^2
  int nu;              // steps taken
  int rank;            // pivots kept (r.size() in the reference)
  int done;            // the loop has ended
  int step_max;        // min(m, n)
  int full;            // ended because the caller's kmax columns were used up
};

// Argmax key of |M(i, j)|^2 with the reference's scan order: rows outer,
// columns inner, strict '>' -- among equal values the smallest i, then the
// smallest j, wins.
struct AcaMax {
  double v;
  int i, j;
};
__device__ __forceinline__ bool aca_better(const AcaMax &a, const AcaMax &b) {
  if (a.v != b.v) return a.v > b.v;
  if (a.i != b.i) return a.i < b.i;
  return a.j < b.j;
}

// One fused pass over the residual M (m x n, column-major): the rank-1
// update M -= a b^T of the previous step (a = U(:, nu-1), b = V(nu-1, :),
// none before the first step), then |M(i, j)|^2 = re^2 + im^2 and a
// per-workgroup argmax.  Contraction is off so the products and norms round
// as the reference's separate multiplies and adds.
__global__ void __launch_bounds__(256) aca_update_max_kernel(double2 *__restrict__ M, int64_t ldm, int m, int n,
                                                            const double2 *__restrict__ U, int64_t ldu,
                                                            const double2 *__restrict__ V, int64_t ldv,
                                                            const AcaState *__restrict__ st,
                                                            AcaMax *__restrict__ part) {
#pragma clang fp contract(off)
  if (st->done) return;
  const int k = st->nu;
  const bool upd = k > 0;
  const double2 *a = U + (int64_t)(k - 1) * ldu, *b = V + (k - 1);
  AcaMax best{-1.0, 0x7fffffff, 0x7fffffff};
  const int rows_per = 256;
  const int r = blockIdx.x * rows_per + threadIdx.x;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    if (r >= m) break;
    double2 v = M[(int64_t)c * ldm + r];
    if (upd) {
      const double2 ai = a[r], bj = b[(int64_t)c * ldv];
      const double pr = ai.x * bj.x - ai.y * bj.y, pi = ai.x * bj.y + ai.y * bj.x;
      v.x = v.x - pr;
      v.y = v.y - pi;
      M[(int64_t)c * ldm + r] = v;
    }
    const double nv = v.x * v.x + v.y * v.y;
    AcaMax cand{nv, r, c};
    if (aca_better(cand, best)) best = cand;
  }
  __shared__ AcaMax sm[256];
  sm[threadIdx.x] = best;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w && aca_better(sm[threadIdx.x + w], sm[threadIdx.x])) sm[threadIdx.x] = sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sm[0];
}

// The pivot step (one workgroup): argmax over the partials, then -- unless
// the maximum is zero (the loop's break) -- a = M(:, js) into U(:, k),
// b = M(is, :) / delta into V(k, :), their squared norms, and the stop rule
// epsrel2 = |a|^2 |b|^2 / sum <= eps2 or nu = min(m, n); the pivot is kept
// (rows[rank], cols[rank]) only while the loop continues, as the reference
// pushes it.
__global__ void __launch_bounds__(256) aca_pivot_kernel(const double2 *__restrict__ M, int64_t ldm, int m, int n,
                                                       const AcaMax *__restrict__ part, int np, double eps2, int kmax,
                                                       double2 *__restrict__ U, int64_t ldu, double2 *__restrict__ V,
                                                       int64_t ldv, int *__restrict__ rows, int *__restrict__ cols,
                                                       AcaState *__restrict__ st) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ AcaMax sm[256];
  __shared__ double sa[256], sb[256];
  const int t = threadIdx.x;
  AcaMax best{-1.0, 0x7fffffff, 0x7fffffff};
  for (int p = t; p < np; p += 256)
    if (aca_better(part[p], best)) best = part[p];
  sm[t] = best;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w && aca_better(sm[t + w], sm[t])) sm[t] = sm[t + w];
    __syncthreads();
  }
  const AcaMax piv = sm[0];
  const int k = st->nu;
  if (piv.v == 0.0 || k >= kmax) {   // norm(delta) == 0: break; k == kmax: no room left
    __syncthreads();
    if (t == 0) {
      st->done = 1;
      st->full = piv.v != 0.0;
    }
    return;
  }
  const int is = piv.i, js = piv.j;
  const double2 delta = M[(int64_t)js * ldm + is];
  double na = 0.0, nb = 0.0;
  for (int i = t; i < m; i += 256) {
    const double2 ai = M[(int64_t)js * ldm + i];
    U[(int64_t)k * ldu + i] = ai;
    na += ai.x * ai.x + ai.y * ai.y;
  }
  // b_j = M(is, j) / delta (complex division as (x conj(d)) / |d|^2)
  const double dd = delta.x * delta.x + delta.y * delta.y;
  for (int j = t; j < n; j += 256) {
    const double2 x = M[(int64_t)j * ldm + is];
    const double2 bj = make_double2((x.x * delta.x + x.y * delta.y) / dd, (x.y * delta.x - x.x * delta.y) / dd);
    V[(int64_t)j * ldv + k] = bj;
    nb += bj.x * bj.x + bj.y * bj.y;
  }
  sa[t] = na;
  sb[t] = nb;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) {
      sa[t] += sa[t + w];
      sb[t] += sb[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double epsabs2 = sa[0] * sb[0];
    const double tot = st->epsabstot2 + epsabs2;
    st->epsabstot2 = tot;
    const double epsrel2 = epsabs2 / tot;
    const int nu = k + 1;
    st->nu = nu;
    const bool cont = nu < st->step_max && epsrel2 > eps2;
    if (cont) {
      rows[st->rank] = is;
      cols[st->rank] = js;
      st->rank = st->rank + 1;
    } else {
      st->done = 1;
    }
  }
}

// --------------------------------------------------------------- Jacobi
// One round of the one-sided Jacobi SVD of A (m x k, k even): workgroup p
// takes the column pair of round `round` in the round-robin order (column 0
// fixed, the others rotating), forms alpha = |a_p|^2, beta = |a_q|^2,
// gamma = a_p^H a_q, and when |gamma| > tol sqrt(alpha beta) applies the
// complex rotation that makes the pair orthogonal,
//   [a_p, a_q] <- [a_p, a_q] [[c, s e^{i phi}], [-s e^{-i phi}, c]],
// gamma = |gamma| e^{i phi}, t = s / c the smaller root of
// t^2 + 2 zeta t - 1 = 0, zeta = (beta - alpha) / (2 |gamma|), to A and to
// the accumulated right factor W (kw x k).  rotated[0] is set when any pair
// rotated (the sweep has not converged).
__global__ void __launch_bounds__(256) jacobi_round_kernel(double2 *__restrict__ A, int64_t lda, int m, int k,
                                                          double2 *__restrict__ W, int64_t ldw, int kw, int round,
                                                          double tol, int *__restrict__ rotated) {
  const int pair = blockIdx.x, t = threadIdx.x;
  auto player = [&](int i) { return i == 0 ? 0 : 1 + (i - 1 + round) % (k - 1); };
  int p = player(pair), q = player(k - 1 - pair);
  if (p > q) {
    const int x = p;
    p = q;
    q = x;
  }
  double2 *ap = A + (int64_t)p * lda, *aq = A + (int64_t)q * lda;
  double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
  for (int i = t; i < m; i += 256) {
    const double2 x = ap[i], y = aq[i];
    al += x.x * x.x + x.y * x.y;
    be += y.x * y.x + y.y * y.y;
    gr += x.x * y.x + x.y * y.y;   // conj(x) y
    gi += x.x * y.y - x.y * y.x;
  }
  __shared__ double red[4][256];
  red[0][t] = al;
  red[1][t] = be;
  red[2][t] = gr;
  red[3][t] = gi;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w)
      for (int u = 0; u < 4; u++) red[u][t] += red[u][t + w];
    __syncthreads();
  }
  al = red[0][0];
  be = red[1][0];
  gr = red[2][0];
  gi = red[3][0];
  const double g = sqrt(gr * gr + gi * gi);
  if (!(g > tol * sqrt(al * be)) || g == 0.0) return;
  if (t == 0) rotated[0] = 1;
  const double zeta = (be - al) / (2.0 * g);
  const double tt = 1.0 / (zeta + (zeta >= 0.0 ? 1.0 : -1.0) * sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
  const double er = gr / g, ei = gi / g;   // e^{i phi}
  // a_p' = c a_p - s e^{-i phi} a_q,  a_q' = s e^{i phi} a_p + c a_q
  auto rot = [&](double2 *x, double2 *y, int len) {
    for (int i = t; i < len; i += 256) {
      const double2 u = x[i], v = y[i];
      const double2 emv = make_double2(er * v.x + ei * v.y, er * v.y - ei * v.x);   // e^{-i phi} v
      const double2 epu = make_double2(er * u.x - ei * u.y, er * u.y + ei * u.x);   // e^{i phi} u
      x[i] = make_double2(c * u.x - s * emv.x, c * u.y - s * emv.y);
      y[i] = make_double2(s * epu.x + c * v.x, s * epu.y + c * v.y);
    }
  };
  rot(ap, aq, m);
  if (W) rot(W + (int64_t)p * ldw, W + (int64_t)q * ldw, kw);
}

// sigma[j] <- |A(:, j)| (one workgroup per column, fixed-order sum).
__global__ void __launch_bounds__(256) col_norms_kernel(const double2 *__restrict__ A, int64_t lda, int m,
                                                       double *__restrict__ sigma) {
  __shared__ double red[256];
  const int j = blockIdx.x, t = threadIdx.x;
  double v = 0.0;
  for (int i = t; i < m; i += 256) {
    const double2 x = A[(int64_t)j * lda + i];
    v += x.x * x.x + x.y * x.y;
  }
  red[t] = v;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) sigma[j] = sqrt(red[0]);
}

// dst(:, c) <- scale[c] * src(:, idx[c]) for c < nc (a column gather with
// per-column scale, scale NULL for 1).
__global__ void gather_cols_kernel(const double2 *__restrict__ src, int64_t lds, int m, const int *__restrict__ idx,
                                   const double *__restrict__ scale, double2 *__restrict__ dst, int64_t ldd) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= m) return;
  double2 v = src[(int64_t)idx[c] * lds + r];
  if (scale) {
    const double f = scale[c];
    v = make_double2(v.x * f, v.y * f);
  }
  dst[(int64_t)c * ldd + r] = v;
}

}  // namespace

extern "C" {

int bem_chol_inv64(const double2 *G, int64_t ldg, int w, double2 *Rinv, int *info, hipStream_t st) {
  if (w < 1 || w > 64) return -1;
  hipLaunchKernelGGL(chol_inv64_kernel, dim3(1), dim3(256), 0, st, G, ldg, w, Rinv, info);
  return (int)hipGetLastError();
}

int bem_transpose_op(const double2 *src, int64_t lds, int m, int n, double2 *dst, int64_t ldd, int conj, int sign_rows,
                     hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  if ((n + 31) / 32 > 65535) return -1;
  hipLaunchKernelGGL(transpose_op_kernel, dim3((m + 31) / 32, (n + 31) / 32), dim3(256), 0, st, src, lds, m, n, dst,
                     ldd, conj, sign_rows);
  return (int)hipGetLastError();
}

int bem_conj_sign(const double2 *src, int64_t lds, int m, int n, double2 *dst, int64_t ldd, int sign_rows,
                  hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  if (n > 65535) return -1;
  hipLaunchKernelGGL(conj_sign_kernel, dim3((m + 255) / 256, n), dim3(256), 0, st, src, lds, m, n, dst, ldd, sign_rows);
  return (int)hipGetLastError();
}

size_t bem_aca_state_bytes(void) { return sizeof(AcaState); }

// Launch grid of the fused ACA pass: 256-row strips by up to 1024 column
// groups; partials hold one AcaMax per workgroup.
static dim3 aca_grid(int m, int n) {
  const int gx = (m + 255) / 256;
  int gy = n < 1024 ? n : 1024;
  if (gy < 1) gy = 1;
  return dim3(gx, gy);
}
size_t bem_aca_partial_bytes(int m, int n) {
  dim3 g = aca_grid(m, n);
  return sizeof(AcaMax) * (size_t)g.x * g.y;
}

// Initial state: nothing kept, step limit min(m, n).
int bem_aca_init(void *state, int m, int n, hipStream_t st) {
  AcaState s{0.0, 0, 0, 0, m < n ? m : n, 0};
  return (int)hipMemcpyAsync(state, &s, sizeof s, hipMemcpyHostToDevice, st);
}

// Reads the state back (after the stream has drained): steps taken, pivots
// kept, whether the loop ended and whether it ran out of columns.
int bem_aca_read(const void *state, int *nu, int *rank, int *done, int *full) {
  AcaState s;
  hipError_t e = hipMemcpy(&s, state, sizeof s, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return (int)e;
  *nu = s.nu;
  *rank = s.rank;
  *done = s.done;
  *full = s.full;
  return 0;
}

// `steps` ACA steps on M (m x n, overwritten by the residual), U (m x kmax,
// the columns a), V (kmax x n, the rows b), rows/cols (kmax pivots each);
// state and partial from bem_aca_state_bytes / bem_aca_partial_bytes.  The
// step index lives on the device, so batches of steps are issued without
// reading it back; a step after the loop ended does nothing.
int bem_aca_steps(double2 *M, int64_t ldm, int m, int n, double eps2, int kmax, double2 *U, int64_t ldu, double2 *V,
                  int64_t ldv, int *rows, int *cols, void *state, void *partial, int steps, hipStream_t st) {
  if (m <= 0 || n <= 0 || kmax <= 0) return -1;
  AcaState *s = (AcaState *)state;
  AcaMax *part = (AcaMax *)partial;
  const dim3 g = aca_grid(m, n);
  const int np = (int)(g.x * g.y);
  for (int q = 0; q < steps; q++) {
    hipLaunchKernelGGL(aca_update_max_kernel, g, dim3(256), 0, st, M, ldm, m, n, (const double2 *)U, ldu,
                       (const double2 *)V, ldv, (const AcaState *)s, part);
    hipLaunchKernelGGL(aca_pivot_kernel, dim3(1), dim3(256), 0, st, (const double2 *)M, ldm, m, n,
                       (const AcaMax *)part, np, eps2, kmax, U, ldu, V, ldv, rows, cols, s);
  }
  return (int)hipGetLastError();
}

// ---- one-sided Jacobi SVD
// One sweep (k - 1 rounds, k even) over the columns of A (m x k), rotations
// accumulated into W (kw x k, may be NULL); rotated (device int) is set when
// any pair rotated.
int bem_jacobi_sweep(double2 *A, int64_t lda, int m, int k, double2 *W, int64_t ldw, int kw, double tol, int *rotated,
                     hipStream_t st) {
  if (k < 2 || (k & 1)) return -1;
  for (int r = 0; r < k - 1; r++)
    hipLaunchKernelGGL(jacobi_round_kernel, dim3(k / 2), dim3(256), 0, st, A, lda, m, k, W, ldw, kw, r, tol, rotated);
  return (int)hipGetLastError();
}

int bem_col_norms(const double2 *A, int64_t lda, int m, int k, double *sigma, hipStream_t st) {
  if (k <= 0) return 0;
  hipLaunchKernelGGL(col_norms_kernel, dim3(k), dim3(256), 0, st, A, lda, m, sigma);
  return (int)hipGetLastError();
}

int bem_gather_cols(const double2 *src, int64_t lds, int m, const int *idx, const double *scale, int nc, double2 *dst,
                    int64_t ldd, hipStream_t st) {
  if (m <= 0 || nc <= 0) return 0;
  if (nc > 65535) return -1;
  hipLaunchKernelGGL(gather_cols_kernel, dim3((m + 255) / 256, nc), dim3(256), 0, st, src, lds, m, idx, scale, dst,
                     ldd);
  return (int)hipGetLastError();
}

}  // extern "C"
