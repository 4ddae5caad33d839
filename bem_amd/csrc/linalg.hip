// Dense complex128 linear algebra for the PMCHWT solve, hand-written for gfx950.
//
//  zgemm_kernel     C = alpha A B + beta C (column-major, complex128), 64x64
//                   C tile per 256-thread block, 2x2 waves of 32x32, FP64 MFMA
//                   v_mfma_f64_16x16x4 with the complex product split into four
//                   real MFMAs (re*re, -im*im, re*im, im*re) into two
//                   accumulators; A/B tiles staged through LDS as re/im planes.
//                   TRACE variant: instead of storing C, the epilogue forms
//                   sum conj(C_ij) * E_ij against a second matrix E and writes
//                   one partial per block (the HeatTransfer trace
//                   tr(B2^H B1) = sum conj(B2) .* B1, src/HeatTransfer.cpp:53-55).
//  tslu_*           tournament pivot selection for a panel (CALU)
//  getrf_nopiv_small / trsm_right_upper   the panel's factors once its
//                   pivot rows are in place
//  laswp_kernel     row interchanges on a column range
//  trsm_small       X = L^-1 X (unit lower) / U^-1 X (upper) with the nb x nb
//                   triangle staged in LDS, one thread per right-hand side
//  symm_kernel      G <- G + G^H (src/util/DenseMatrix.cpp:477-486)
#include "bem_common.h"

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace {

// ------------------------------------------------------------------- ZGEMM
constexpr int GBM = 64, GBN = 64, GBK = 16, GPAD = 1;

template <bool TRACE>
__global__ void __launch_bounds__(256) zgemm_kernel(int m, int n, int k, double2 alpha, const double2 *__restrict__ A,
                                                    int64_t lda, const double2 *__restrict__ B, int64_t ldb,
                                                    double2 beta, double2 *__restrict__ C, int64_t ldc,
                                                    const double2 *__restrict__ E, int64_t lde,
                                                    double2 *__restrict__ partial) {
  __shared__ double sAr[GBK][GBM + GPAD], sAi[GBK][GBM + GPAD];
  __shared__ double sBr[GBK][GBN + GPAD], sBi[GBK][GBN + GPAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;

  dbl4 accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      accr[a][b] = (dbl4){0, 0, 0, 0};
      acci[a][b] = (dbl4){0, 0, 0, 0};
    }

  // register prefetch of the next tiles
  double2 ra[4], rb[4];
  const int arow = tid & 63, acg = tid >> 6;      // A: row, column group of 4
  const int bkr = tid & 15, bcol = tid >> 4;      // B: k-row, column (+16q)
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int col = k0 + acg * 4 + c, row = m0 + arow;
      ra[c] = (row < m && col < k) ? A[(int64_t)col * lda + row] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int col = n0 + bcol + 16 * q, row = k0 + bkr;
      rb[q] = (row < k && col < n) ? B[(int64_t)col * ldb + row] : make_double2(0.0, 0.0);
    }
  };
  load_tiles(0);
  for (int k0 = 0; k0 < k; k0 += GBK) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++) {
      sAr[acg * 4 + c][arow] = ra[c].x;
      sAi[acg * 4 + c][arow] = ra[c].y;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      sBr[bkr][bcol + 16 * q] = rb[q].x;
      sBi[bkr][bcol + 16 * q] = rb[q].y;
    }
    __syncthreads();
    if (k0 + GBK < k) load_tiles(k0 + GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        ar[x] = sAr[kk + lk][wr * 32 + x * 16 + li];
        ai[x] = sAi[kk + lk][wr * 32 + x * 16 + li];
        nai[x] = -ai[x];
        br[x] = sBr[kk + lk][wc * 32 + x * 16 + li];
        bi[x] = sBi[kk + lk][wc * 32 + x * 16 + li];
      }
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
          accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], accr[x][y], 0, 0, 0);
          accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[x], bi[y], accr[x][y], 0, 0, 0);
          acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], acci[x][y], 0, 0, 0);
          acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], acci[x][y], 0, 0, 0);
        }
    }
  }
  // epilogue: for v_mfma_f64_16x16x4f64, D[i][j] of a 16x16 block sits in
  // lane j + 16*(i%4), element i/4 (measured on gfx950; the f32 variant differs)
  if (!TRACE) {
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = m0 + wr * 32 + x * 16 + 4 * r + lk;
          int col = n0 + wc * 32 + y * 16 + li;
          if (row < m && col < n) {
            double cr = accr[x][y][r], ci = acci[x][y][r];
            double vr = alpha.x * cr - alpha.y * ci, vi = alpha.x * ci + alpha.y * cr;
            double2 *p = C + (int64_t)col * ldc + row;
            if (beta.x != 0.0 || beta.y != 0.0) {
              double2 o = *p;
              vr += beta.x * o.x - beta.y * o.y;
              vi += beta.x * o.y + beta.y * o.x;
            }
            *p = make_double2(vr, vi);
          }
        }
  } else {
    __shared__ double sred[2][4];
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = m0 + wr * 32 + x * 16 + 4 * r + lk;
          int col = n0 + wc * 32 + y * 16 + li;
          if (row < m && col < n) {
            double cr = accr[x][y][r], ci = acci[x][y][r];
            double vr = alpha.x * cr - alpha.y * ci, vi = alpha.x * ci + alpha.y * cr;
            double2 e = E[(int64_t)col * lde + row];
            // conj(v) * e
            sr += vr * e.x + vi * e.y;
            si += vr * e.y - vi * e.x;
          }
        }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      sr += __shfl_down(sr, off, 64);
      si += __shfl_down(si, off, 64);
    }
    if (lane == 0) { sred[0][wave] = sr; sred[1][wave] = si; }
    __syncthreads();
    if (tid == 0) {
      double tr = ((sred[0][0] + sred[0][1]) + sred[0][2]) + sred[0][3];
      double ti = ((sred[1][0] + sred[1][1]) + sred[1][2]) + sred[1][3];
      partial[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(tr, ti);
    }
  }
}

// Fixed-order sum of the trace partials (deterministic).
__global__ void sum_partials_kernel(const double2 *__restrict__ partial, int n, double2 *__restrict__ out) {
  __shared__ double sr[256], si[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    a += partial[i].x;
    b += partial[i].y;
  }
  sr[threadIdx.x] = a;
  si[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(sr[0], si[0]);
}

// ------------------------------------------------ tournament-pivoted panel
// Communication-avoiding partial pivoting (CALU, Grigori-Demmel-Xiang 2011):
// the panel's candidate pivot rows are chosen by a reduction tree of
// partial-pivoting eliminations (izamax criterion |re|+|im|) instead of one
// global argmax per column, so the panel needs no grid-wide barrier.
//
// gepp_select: rows held one per thread (a[NB]); runs Gaussian elimination with
// partial pivoting on the valid rows and writes the chosen rows' ids in pivot
// order to out[0..w).  Rows are eliminated in place (not moved).
struct GeppShared {
  double *val;
  int *idx;
  double2 *prow;
  int *pick;
};

// Column step J of the elimination; recursion keeps J a compile-time index
// so the row stays in registers.
template <int NB, int NT, int J>
__device__ __forceinline__ void gepp_step(double2 (&a)[NB], bool valid, bool &chosen, int rowid, int w, int *out,
                                          const GeppShared &sh) {
  if constexpr (J < NB) {
    if (J < w) {
      const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
      double best = (valid && !chosen) ? fabs(a[J].x) + fabs(a[J].y) : -1.0;
      int bidx = (valid && !chosen) ? tid : 0x7fffffff;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_down(best, off, 64);
        int oi = __shfl_down(bidx, off, 64);
        if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
      }
      if (lane == 0) { sh.val[wave] = best; sh.idx[wave] = bidx; }
      __syncthreads();
      if (tid == 0) {
        double bv = sh.val[0];
        int bi = sh.idx[0];
        for (int x = 1; x < NT / 64; x++)
          if (sh.val[x] > bv || (sh.val[x] == bv && sh.idx[x] < bi)) { bv = sh.val[x]; bi = sh.idx[x]; }
        *sh.pick = (bv >= 0.0) ? bi : -1;   // -1: no valid row left
      }
      __syncthreads();
      const int p = *sh.pick;
      if (p < 0) {
        if (tid == 0) out[J] = -1;
      } else {
        if (tid == p) {
          chosen = true;
          out[J] = rowid;
#pragma unroll
          for (int c = 0; c < NB; c++) sh.prow[c] = a[c];
        }
        __syncthreads();
        const double2 u = sh.prow[J];
        const double den = u.x * u.x + u.y * u.y;
        if (valid && !chosen && den > 0.0) {
          const double2 ui = make_double2(u.x / den, -u.y / den);
          double2 x = a[J];
          double2 l = make_double2(x.x * ui.x - x.y * ui.y, x.x * ui.y + x.y * ui.x);
          a[J] = l;
#pragma unroll
          for (int c = J + 1; c < NB; c++) {
            double2 uc = sh.prow[c];
            a[c].x -= l.x * uc.x - l.y * uc.y;
            a[c].y -= l.x * uc.y + l.y * uc.x;
          }
        }
      }
      __syncthreads();
      gepp_step<NB, NT, J + 1>(a, valid, chosen, rowid, w, out, sh);
    }
  }
}

template <int NB, int NT>
__device__ __forceinline__ void gepp_select(double2 (&a)[NB], bool valid, int rowid, int w, int *out, double *s_val, int *s_idx,
                            double2 *s_prow, int *s_pick) {
  bool chosen = false;
  GeppShared sh{s_val, s_idx, s_prow, s_pick};
  gepp_step<NB, NT, 0>(a, valid, chosen, rowid, w, out, sh);
}

// Leaves of the tournament: block b takes panel rows [b*256, b*256+256) of
// the m-row panel starting at matrix row r0 and emits w candidate row ids.
template <int NB>
__global__ void __launch_bounds__(256) tslu_leaf_kernel(const double2 *__restrict__ A, int64_t lda, int m, int w,
                                                        int r0, int c0, int *__restrict__ cand) {
  __shared__ double s_val[4];
  __shared__ int s_idx[4], s_pick;
  __shared__ double2 s_prow[NB];
  int row = blockIdx.x * 256 + threadIdx.x;
  bool valid = row < m;
  double2 a[NB];
#pragma unroll
  for (int c = 0; c < NB; c++)
    a[c] = (valid && c < w) ? A[(int64_t)(c0 + c) * lda + r0 + row] : make_double2(0.0, 0.0);
  gepp_select<NB, 256>(a, valid, r0 + row, w, cand + (size_t)blockIdx.x * NB, s_val, s_idx, s_prow, &s_pick);
}

// One tournament round: block b merges candidate sets 2b and 2b+1 (ids into
// the matrix, -1 = empty) by GEPP on their original rows.
template <int NB>
__global__ void __launch_bounds__(2 * NB) tslu_merge_kernel(const double2 *__restrict__ A, int64_t lda, int w, int c0,
                                                            const int *__restrict__ cin, int nsets,
                                                            int *__restrict__ cout) {
  __shared__ double s_val[(2 * NB + 63) / 64];
  __shared__ int s_idx[(2 * NB + 63) / 64], s_pick;
  __shared__ double2 s_prow[NB];
  int set = 2 * blockIdx.x + threadIdx.x / NB;
  int slot = threadIdx.x % NB;
  int id = (set < nsets && slot < w) ? cin[(size_t)set * NB + slot] : -1;
  bool valid = id >= 0;
  double2 a[NB];
#pragma unroll
  for (int c = 0; c < NB; c++)
    a[c] = (valid && c < w) ? A[(int64_t)(c0 + c) * lda + id] : make_double2(0.0, 0.0);
  gepp_select<NB, 2 * NB>(a, valid, id, w, cout + (size_t)blockIdx.x * NB, s_val, s_idx, s_prow, &s_pick);
}

// Turn the winner list into LAPACK-style interchanges ipiv[k0+i] (row k0+i
// swapped with row ipiv[k0+i], applied in order).  Single thread: w is small.
__global__ void winners_to_ipiv_kernel(const int *__restrict__ win, int w, int k0, int *__restrict__ ipiv,
                                       int *__restrict__ info) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int pos[128];
  for (int i = 0; i < w; i++) pos[i] = win[i];
  for (int i = 0; i < w; i++) {
    int target = k0 + i;
    int src = pos[i];
    if (src < 0) {   // rank-deficient panel: keep the row, flag it
      src = target;
      if (info && *info == 0) *info = target + 1;
    }
    ipiv[target] = src;
    // the row at 'target' moves to 'src'
    for (int t = i + 1; t < w; t++)
      if (pos[t] == target) pos[t] = src;
  }
}

// LU without pivoting of the w x w diagonal block (rows already in pivot
// order), in place: unit L below the diagonal, U on and above.
__global__ void __launch_bounds__(256) getrf_nopiv_small_kernel(double2 *__restrict__ A, int64_t lda, int w,
                                                               int *__restrict__ info, int k0) {
  extern __shared__ double2 sA[];   // w x w column-major
  for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) sA[idx] = A[(int64_t)(idx / w) * lda + idx % w];
  __syncthreads();
  for (int j = 0; j < w; j++) {
    double2 u = sA[j * w + j];
    double den = u.x * u.x + u.y * u.y;
    if (den == 0.0) {
      if (threadIdx.x == 0 && info && *info == 0) *info = k0 + j + 1;
      __syncthreads();
      continue;
    }
    double2 ui = make_double2(u.x / den, -u.y / den);
    __syncthreads();
    for (int i = j + 1 + (int)threadIdx.x; i < w; i += blockDim.x) {
      double2 x = sA[j * w + i];
      sA[j * w + i] = make_double2(x.x * ui.x - x.y * ui.y, x.x * ui.y + x.y * ui.x);
    }
    __syncthreads();
    int rem = w - j - 1;
    for (int idx = threadIdx.x; idx < rem * rem; idx += blockDim.x) {
      int i = j + 1 + idx % rem, c = j + 1 + idx / rem;
      double2 l = sA[j * w + i], uc = sA[c * w + j];
      double2 x = sA[c * w + i];
      x.x -= l.x * uc.x - l.y * uc.y;
      x.y -= l.x * uc.y + l.y * uc.x;
      sA[c * w + i] = x;
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) A[(int64_t)(idx / w) * lda + idx % w] = sA[idx];
}

// L21 <- L21 U11^-1 for the rows below the diagonal block (one thread per row).
__global__ void __launch_bounds__(256) trsm_right_upper_kernel(const double2 *__restrict__ U, int64_t ldu, int w,
                                                               double2 *__restrict__ X, int64_t ldx, int nrows) {
  extern __shared__ double2 sU[];
  for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) sU[idx] = U[(int64_t)(idx / w) * ldu + idx % w];
  __syncthreads();
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double2 x[64];
  for (int j = 0; j < w; j++) x[j] = X[(int64_t)j * ldx + r];
  for (int j = 0; j < w; j++) {
    double2 s = x[j];
    for (int i = 0; i < j; i++) {
      double2 u = sU[j * w + i];
      s.x -= x[i].x * u.x - x[i].y * u.y;
      s.y -= x[i].x * u.y + x[i].y * u.x;
    }
    double2 d = sU[j * w + j];
    double den = d.x * d.x + d.y * d.y;
    x[j] = den > 0.0 ? make_double2((s.x * d.x + s.y * d.y) / den, (s.y * d.x - s.x * d.y) / den) : s;
  }
  for (int j = 0; j < w; j++) X[(int64_t)j * ldx + r] = x[j];
}

// Row interchanges: for p in [k0, k1): swap rows p and ipiv[p] in every column
// of [c0, c1) except [skip0, skip1).
__global__ void laswp_kernel(double2 *__restrict__ A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1,
                             const int *__restrict__ ipiv) {
  int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1 || (c >= skip0 && c < skip1)) return;
  double2 *col = A + (int64_t)c * lda;
  for (int p = k0; p < k1; p++) {
    int q = ipiv[p];
    if (q != p) {
      double2 t = col[p];
      col[p] = col[q];
      col[q] = t;
    }
  }
}

// X (nb x ncols) <- L^-1 X with L unit lower (UPPER=false) or U^-1 X with U
// upper non-unit (UPPER=true); the nb x nb triangle comes from T (ld ldt).
template <bool UPPER>
__global__ void __launch_bounds__(256) trsm_small_kernel(const double2 *__restrict__ T, int64_t ldt, int nb,
                                                         double2 *__restrict__ X, int64_t ldx, int ncols) {
  extern __shared__ double2 sT[];  // nb x nb, column-major
  for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
    int r = idx % nb, c = idx / nb;
    sT[idx] = T[(int64_t)c * ldt + r];
  }
  __syncthreads();
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  double2 *x = X + (int64_t)col * ldx;
  if (!UPPER) {
    for (int j = 0; j < nb; j++) {
      double2 xj = x[j];
      for (int i = j + 1; i < nb; i++) {
        double2 l = sT[j * nb + i];
        double2 xi = x[i];
        xi.x -= l.x * xj.x - l.y * xj.y;
        xi.y -= l.x * xj.y + l.y * xj.x;
        x[i] = xi;
      }
    }
  } else {
    for (int j = nb - 1; j >= 0; j--) {
      double2 d = sT[j * nb + j];
      double den = d.x * d.x + d.y * d.y;
      double2 xj = x[j];
      double2 v = make_double2((xj.x * d.x + xj.y * d.y) / den, (xj.y * d.x - xj.x * d.y) / den);
      x[j] = v;
      for (int i = 0; i < j; i++) {
        double2 u = sT[j * nb + i];
        double2 xi = x[i];
        xi.x -= u.x * v.x - u.y * v.y;
        xi.y -= u.x * v.y + u.y * v.x;
        x[i] = xi;
      }
    }
  }
}

__global__ void symm_kernel(double2 *__restrict__ G, int64_t ld, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= n || i < j) return;
  if (i == j) {
    double2 a = G[(int64_t)j * ld + i];
    G[(int64_t)j * ld + i] = make_double2(a.x + a.x, 0.0);
    return;
  }
  double2 a = G[(int64_t)j * ld + i];   // (i, j), lower
  double2 b = G[(int64_t)i * ld + j];   // (j, i), upper
  double2 s = make_double2(a.x + b.x, a.y - b.y);   // a + conj(b)
  G[(int64_t)j * ld + i] = s;
  G[(int64_t)i * ld + j] = make_double2(s.x, -s.y);
}

__global__ void set_identity_kernel(double2 *__restrict__ X, int64_t ld, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= n) return;
  X[(int64_t)j * ld + i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
}

}  // namespace

// ------------------------------------------------------------------ launchers
extern "C" {

int bem_zgemm_raw(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
              double2 beta, double2 *C, int64_t ldc, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid((m + GBM - 1) / GBM, (n + GBN - 1) / GBN);
  hipLaunchKernelGGL(zgemm_kernel<false>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                     (const double2 *)nullptr, (int64_t)0, (double2 *)nullptr);
  return (int)hipGetLastError();
}

// partial must hold ceil(m/64)*ceil(n/64) entries; out receives sum conj(alpha A B) .* E.
int bem_zgemm_trace(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                    const double2 *E, int64_t lde, double2 *partial, double2 *out, hipStream_t st) {
  dim3 grid((m + GBM - 1) / GBM, (n + GBN - 1) / GBN);
  hipLaunchKernelGGL(zgemm_kernel<true>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb,
                     make_double2(0.0, 0.0), (double2 *)nullptr, (int64_t)0, E, lde, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, (int)(grid.x * grid.y), out);
  return (int)hipGetLastError();
}

size_t bem_trace_partials(int m, int n) { return (size_t)((m + GBM - 1) / GBM) * (size_t)((n + GBN - 1) / GBN); }

// Tournament pivot selection for the panel A(r0:r0+m, c0:c0+w), w <= 64.
// cand must hold ceil(m/256)*64 ints, cand2 half of that (ping-pong);
// the winners end up in ipiv[k0..k0+w) as interchanges (k0 = r0).
int bem_calu_pivots(const double2 *A, int64_t lda, int m, int w, int r0, int c0, int *cand, int *cand2, int *ipiv,
                    int *info, hipStream_t st) {
  if (w <= 0) return 0;
  if (w > 64) return -1;
  int nsets = (m + 255) / 256;
  hipLaunchKernelGGL(tslu_leaf_kernel<64>, dim3(nsets), dim3(256), 0, st, A, lda, m, w, r0, c0, cand);
  int *cin = cand, *cout = cand2;
  while (nsets > 1) {
    int nout = (nsets + 1) / 2;
    hipLaunchKernelGGL(tslu_merge_kernel<64>, dim3(nout), dim3(128), 0, st, A, lda, w, c0, cin, nsets, cout);
    int *t = cin; cin = cout; cout = t;
    nsets = nout;
  }
  hipLaunchKernelGGL(winners_to_ipiv_kernel, dim3(1), dim3(64), 0, st, cin, w, r0, ipiv, info);
  return (int)hipGetLastError();
}

int bem_getrf_nopiv_small(double2 *A, int64_t lda, int w, int *info, int k0, hipStream_t st) {
  if (w <= 0) return 0;
  hipLaunchKernelGGL(getrf_nopiv_small_kernel, dim3(1), dim3(256), sizeof(double2) * w * w, st, A, lda, w, info, k0);
  return (int)hipGetLastError();
}

int bem_trsm_right_upper(const double2 *U, int64_t ldu, int w, double2 *X, int64_t ldx, int nrows, hipStream_t st) {
  if (nrows <= 0 || w <= 0) return 0;
  if (w > 64) return -1;
  hipLaunchKernelGGL(trsm_right_upper_kernel, dim3((nrows + 255) / 256), dim3(256), sizeof(double2) * w * w, st, U,
                     ldu, w, X, ldx, nrows);
  return (int)hipGetLastError();
}

int bem_laswp(double2 *A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1, const int *ipiv,
              hipStream_t st) {
  if (c1 <= c0 || k1 <= k0) return 0;
  hipLaunchKernelGGL(laswp_kernel, dim3((c1 - c0 + 127) / 128), dim3(128), 0, st, A, lda, c0, c1, skip0, skip1, k0, k1,
                     ipiv);
  return (int)hipGetLastError();
}

int bem_trsm_small(int upper, const double2 *T, int64_t ldt, int nb, double2 *X, int64_t ldx, int ncols,
                   hipStream_t st) {
  if (ncols <= 0 || nb <= 0) return 0;
  size_t sh = sizeof(double2) * (size_t)nb * (size_t)nb;
  if (upper)
    hipLaunchKernelGGL(trsm_small_kernel<true>, dim3((ncols + 255) / 256), dim3(256), sh, st, T, ldt, nb, X, ldx, ncols);
  else
    hipLaunchKernelGGL(trsm_small_kernel<false>, dim3((ncols + 255) / 256), dim3(256), sh, st, T, ldt, nb, X, ldx,
                       ncols);
  return (int)hipGetLastError();
}

int bem_symm(double2 *G, int64_t ld, int n, hipStream_t st) {
  hipLaunchKernelGGL(symm_kernel, dim3((n + 255) / 256, n), dim3(256), 0, st, G, ld, n);
  return (int)hipGetLastError();
}

int bem_set_identity(double2 *X, int64_t ld, int n, hipStream_t st) {
  hipLaunchKernelGGL(set_identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, st, X, ld, n);
  return (int)hipGetLastError();
}

}  // extern "C"
