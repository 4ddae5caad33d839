// Dense complex128 linear algebra for the PMCHWT solve, hand-written for gfx950.
//
//  zgemm_kernel     C = alpha A B + beta C (column-major, complex128) on FP64
//                   MFMA v_mfma_f64_16x16x4; A/B tiles staged through LDS as
//                   re/im planes.  Default: 32 x 64 C tile per 256-thread
//                   block and the complex product as three real MFMAs
//                   (Ar Br, Ai Bi, (Ar + Ai)(Br + Bi)) into three
//                   accumulators; the four-product form stays available.
//  zgemm_htrace     the HeatTransfer trace tr(B2^H B1) = sum conj(B2) .* B1
//                   (src/HeatTransfer.cpp:53-55) as tr(sym G2 X^H sym G1 X):
//                   only the upper half of the Hermitian X^H (sym G1 X) is
//                   formed and contracted against sym G2 in the epilogue, one
//                   partial per block
//  tslu_*           tournament pivot selection for a panel (CALU)
//  diag_block       a 64x64 diagonal block once its pivot rows are in place:
//                   LU without pivoting and the inverses of both triangles,
//                   which turn every triangular solve into ZGEMMs
//  diag_lu_piv      a 64x64 diagonal block with partial pivoting inside it
//                   (the default LU), plus both inverses
//  laswp_kernel     row interchanges on a column range
//  symm_kernel      G <- G + G^H (src/util/DenseMatrix.cpp:477-486)
#include "bem_common.h"

#include <algorithm>
#include <type_traits>

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

namespace {

// complex element types: double2 (complex128) and float2 (complex64)
template <typename T> struct Cplx;
template <> struct Cplx<double2> {
  using R = double;
  __device__ static double2 mk(double a, double b) { return make_double2(a, b); }
};
template <> struct Cplx<float2> {
  using R = float;
  __device__ static float2 mk(float a, float b) { return make_float2(a, b); }
};

// ------------------------------------------------------------------- ZGEMM
// Block tile BM x BN, WM x WN waves, each wave FM x FN blocks of 16 x 16;
// k staged through LDS in steps of GBK as separate re/im planes, the next
// step's tiles prefetched into registers while the MFMAs run.
// G3M: the complex product takes three real MFMAs per k step instead of
// four (T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi); re = T1 - T2,
// im = T3 - T1 - T2), with error bounded by u (|Ar| + |Ai|)(|Br| + |Bi|)
// per term as for the four-product form.
constexpr int GPAD = 1;

// The k loop of the C tile at (m0, n0) of one block: acc receives the real
// products (two accumulators, three with G3M).  TA: the A operand is the
// conjugate transpose of the k x m column-major array at A, read with threads
// along k (as the B tile) and conjugated on its way into LDS.
template <int BM, int BN, int WM, int WN, int GBK, bool G3M, bool TA>
__device__ __forceinline__ void zgemm_tile(int m, int n, int k, const double2 *__restrict__ A, int64_t lda,
                                           const double2 *__restrict__ B, int64_t ldb, int m0, int n0,
                                           dbl4 (&acc)[G3M ? 3 : 2][BM / WM / 16][BN / WN / 16]) {
  constexpr int NT = 64 * WM * WN;
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int AL = BM * GBK / NT;   // A elements loaded per thread per k step
  constexpr int BL = BN * GBK / NT;
  constexpr int NACC = G3M ? 3 : 2;
  static_assert(AL >= 1 && BL >= 1 && NT % BM == 0 && NT % GBK == 0, "tile shape");
  __shared__ double sAr[GBK][BM + GPAD], sAi[GBK][BM + GPAD];
  __shared__ double sBr[GBK][BN + GPAD], sBi[GBK][BN + GPAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;

#pragma unroll
  for (int q = 0; q < NACC; q++)
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
      for (int b = 0; b < FN; b++) acc[q][a][b] = (dbl4){0, 0, 0, 0};

  double2 ra[AL], rb[BL];
  // A tile (BM x GBK): thread t reads row t % BM of columns (t / BM) + (NT/BM) q
  // (TA: k index t % GBK of rows t / GBK + (NT/GBK) q)
  // B tile (GBK x BN): thread t reads k-row t % GBK of columns t / GBK + (NT/GBK) q
  const int arow = TA ? tid / GBK : tid % BM, acol0 = TA ? tid % GBK : tid / BM;
  constexpr int ASTEP = TA ? NT / GBK : NT / BM;
  const int bkr = tid % GBK, bcol0 = tid / GBK;
  constexpr int BSTEP = NT / GBK;
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int q = 0; q < AL; q++) {
      if constexpr (TA) {
        int row = m0 + arow + ASTEP * q, kk = k0 + acol0;
        ra[q] = (row < m && kk < k) ? A[(int64_t)row * lda + kk] : make_double2(0.0, 0.0);
      } else {
        int col = k0 + acol0 + ASTEP * q, row = m0 + arow;
        ra[q] = (row < m && col < k) ? A[(int64_t)col * lda + row] : make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int q = 0; q < BL; q++) {
      int col = n0 + bcol0 + BSTEP * q, row = k0 + bkr;
      rb[q] = (row < k && col < n) ? B[(int64_t)col * ldb + row] : make_double2(0.0, 0.0);
    }
  };
  load_tiles(0);
  for (int k0 = 0; k0 < k; k0 += GBK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < AL; q++) {
      if constexpr (TA) {
        sAr[acol0][arow + ASTEP * q] = ra[q].x;
        sAi[acol0][arow + ASTEP * q] = -ra[q].y;
      } else {
        sAr[acol0 + ASTEP * q][arow] = ra[q].x;
        sAi[acol0 + ASTEP * q][arow] = ra[q].y;
      }
    }
#pragma unroll
    for (int q = 0; q < BL; q++) {
      sBr[bkr][bcol0 + BSTEP * q] = rb[q].x;
      sBi[bkr][bcol0 + BSTEP * q] = rb[q].y;
    }
    __syncthreads();
    if (k0 + GBK < k) load_tiles(k0 + GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      double ar[FM], ai[FM], br[FN], bi[FN];
#pragma unroll
      for (int x = 0; x < FM; x++) {
        ar[x] = sAr[kk + lk][wr * (BM / WM) + x * 16 + li];
        ai[x] = sAi[kk + lk][wr * (BM / WM) + x * 16 + li];
      }
#pragma unroll
      for (int y = 0; y < FN; y++) {
        br[y] = sBr[kk + lk][wc * (BN / WN) + y * 16 + li];
        bi[y] = sBi[kk + lk][wc * (BN / WN) + y * 16 + li];
      }
      if constexpr (G3M) {
        double as[FM], bs[FN];
#pragma unroll
        for (int x = 0; x < FM; x++) as[x] = ar[x] + ai[x];
#pragma unroll
        for (int y = 0; y < FN; y++) bs[y] = br[y] + bi[y];
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) {
            acc[0][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[y], ar[x], acc[0][x][y], 0, 0, 0);
            acc[1][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[y], ai[x], acc[1][x][y], 0, 0, 0);
            acc[2][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bs[y], as[x], acc[2][x][y], 0, 0, 0);
          }
      } else {
        double nai[FM];
#pragma unroll
        for (int x = 0; x < FM; x++) nai[x] = -ai[x];
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) {
            acc[0][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[y], ar[x], acc[0][x][y], 0, 0, 0);
            acc[1][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[y], ar[x], acc[1][x][y], 0, 0, 0);
          }
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) {
            acc[0][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[y], nai[x], acc[0][x][y], 0, 0, 0);
            acc[1][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[y], ai[x], acc[1][x][y], 0, 0, 0);
          }
      }
    }
  }
}

template <int BM, int BN, int WM, int WN, int GBK, int OCC, bool G3M>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(OCC)))
zgemm_kernel(int m, int n, int k, double2 alpha, const double2 *__restrict__ A, int64_t lda,
             const double2 *__restrict__ B, int64_t ldb, double2 beta, double2 *__restrict__ C, int64_t ldc) {
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int NACC = G3M ? 3 : 2;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  dbl4 acc[NACC][FM][FN];
  zgemm_tile<BM, BN, WM, WN, GBK, G3M, false>(m, n, k, A, lda, B, ldb, m0, n0, acc);
  // the (re, im) of output element r of block (x, y)
  auto result = [&](int x, int y, int r, double &cr, double &ci) {
    if constexpr (G3M) {
      double t1 = acc[0][x][y][r], t2 = acc[1][x][y][r];
      cr = t1 - t2;
      ci = acc[2][x][y][r] - t1 - t2;
    } else {
      cr = acc[0][x][y][r];
      ci = acc[1][x][y][r];
    }
  };
  // epilogue: for v_mfma_f64_16x16x4f64, D[i][j] of a 16x16 block sits in
  // lane j + 16*(i%4), element i/4 (measured on gfx950; the f32 variant
  // differs).  The products above run with the B tile as the instruction's
  // first operand, so D is the transposed block: C[i][j] sits in lane
  // i + 16*(j%4), element j/4, and lanes run down a column of C (coalesced
  // column-major loads and stores).  With beta != 0 the old C values of one
  // row of 16 x 16 blocks are loaded first, from clamped addresses and
  // without branches, so those loads are in flight together.
  const bool rd = beta.x != 0.0 || beta.y != 0.0;
#pragma unroll
  for (int x = 0; x < FM; x++) {
    const int row = m0 + wr * (BM / WM) + x * 16 + li;
    double2 old[FN][4];
    if (rd) {
#pragma unroll
      for (int y = 0; y < FN; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int col = n0 + wc * (BN / WN) + y * 16 + 4 * r + lk;
          old[y][r] = C[(row < m && col < n) ? (int64_t)col * ldc + row : 0];
        }
    }
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int col = n0 + wc * (BN / WN) + y * 16 + 4 * r + lk;
        if (row < m && col < n) {
          double cr, ci;
          result(x, y, r, cr, ci);
          double vr = alpha.x * cr - alpha.y * ci, vi = alpha.x * ci + alpha.y * cr;
          if (rd) {
            double2 o = old[y][r];
            vr += beta.x * o.x - beta.y * o.y;
            vi += beta.x * o.y + beta.y * o.x;
          }
          C[(int64_t)col * ldc + row] = make_double2(vr, vi);
        }
      }
  }
}

// Hermitian trace: sum over i, j of conj(P_ij) E_ij for P = A^H B (A: k x n,
// B: k x n, column-major) when P and E are both Hermitian.  The (i, j) and
// (j, i) terms are then complex conjugates of each other, so only blocks that
// meet the upper triangle are computed: an entry above the diagonal adds twice
// its real part, a diagonal entry adds itself, an entry below adds nothing
// (the imaginary part of the sum is that of the diagonal terms).  Blocks below
// the diagonal write a zero partial.  Half the flops of the full product.
template <int BM, int BN, int WM, int WN, int GBK, int OCC, bool G3M>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(OCC)))
zgemm_htrace_kernel(int n, int k, const double2 *__restrict__ A, int64_t lda, const double2 *__restrict__ B,
                    int64_t ldb, const double2 *__restrict__ E, int64_t lde, double2 *__restrict__ partial) {
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int NACC = G3M ? 3 : 2;
  constexpr int NW = WM * WN;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  if (m0 > n0 + BN - 1) {   // every row of the block below every column
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(0.0, 0.0);
    return;
  }
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  dbl4 acc[NACC][FM][FN];
  zgemm_tile<BM, BN, WM, WN, GBK, G3M, true>(n, n, k, A, lda, B, ldb, m0, n0, acc);
  __shared__ double sred[2][NW];
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int x = 0; x < FM; x++)
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // element layout as in zgemm_kernel's epilogue
        int row = m0 + wr * (BM / WM) + x * 16 + li;
        int col = n0 + wc * (BN / WN) + y * 16 + 4 * r + lk;
        if (row <= col && col < n) {
          double cr, ci;
          if constexpr (G3M) {
            double t1 = acc[0][x][y][r], t2 = acc[1][x][y][r];
            cr = t1 - t2;
            ci = acc[2][x][y][r] - t1 - t2;
          } else {
            cr = acc[0][x][y][r];
            ci = acc[1][x][y][r];
          }
          double2 e = E[(int64_t)col * lde + row];
          double re = cr * e.x + ci * e.y;   // conj(c) * e
          if (row < col) {
            sr += 2.0 * re;
          } else {
            sr += re;
            si += cr * e.y - ci * e.x;
          }
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
    double tr = 0.0, ti = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) { tr += sred[0][w]; ti += sred[1][w]; }
    partial[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(tr, ti);
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
// the panel's pivot rows are chosen by a reduction tree of partial-pivoting
// eliminations (izamax criterion |re|+|im|) over blocks of rows, instead of
// one global argmax per column, so a panel needs no grid-wide barrier.
//
// Row storage: the NB (<= 32) columns of a row are split over a lane pair
// (2r, 2r+1), NB/2 complex values each, so a 1024-thread block holds 512 rows
// in registers.  A column step costs two block barriers; the LDS used by
// step J is double-buffered on J's parity so no closing barrier is needed.
#define GEPP_NB 32
#define GEPP_NT 1024
#define GEPP_ROWS (GEPP_NT / 2)
#define GEPP_HALF (GEPP_NB / 2)
#define GEPP_MERGE (GEPP_ROWS / GEPP_NB)

struct GeppLds {
  double val[2][GEPP_NT / 64];
  int idx[2][GEPP_NT / 64];
  double2 prow[2][GEPP_NB];
};

template <int J>
__device__ __forceinline__ void gepp2_step(double2 (&a)[GEPP_HALF], bool valid, bool &chosen, int rowid, int w,
                                           int *out, GeppLds &L) {
  if constexpr (J < GEPP_NB) {
    if (J < w) {
      constexpr int H = J / GEPP_HALF, CJ = J % GEPP_HALF, P = J & 1;
      const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, half = tid & 1;
      const bool owner = (half == H);
      double best = (valid && !chosen && owner) ? fabs(a[CJ].x) + fabs(a[CJ].y) : -1.0;
      int bidx = (valid && !chosen && owner) ? tid : 0x7fffffff;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(best, off, 64);
        int oi = __shfl_xor(bidx, off, 64);
        if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
      }
      if (lane == 0) { L.val[P][wave] = best; L.idx[P][wave] = bidx; }
      __syncthreads();
      double bv = L.val[P][0];
      int bi = L.idx[P][0];
#pragma unroll
      for (int x = 1; x < GEPP_NT / 64; x++) {
        double v = L.val[P][x];
        int ii = L.idx[P][x];
        if (v > bv || (v == bv && ii < bi)) { bv = v; bi = ii; }
      }
      const bool none = bv < 0.0;
      const int prow_t = none ? -1 : (bi & ~1);     // first lane of the pivot row's pair
      const bool is_piv = (tid & ~1) == prow_t;
      if (is_piv) {
        chosen = true;
#pragma unroll
        for (int c = 0; c < GEPP_HALF; c++)
          if (GEPP_HALF * half + c >= J) L.prow[P][GEPP_HALF * half + c] = a[c];
        if (half == 0) out[J] = rowid;
      }
      if (none && tid == 0) out[J] = -1;
      __syncthreads();
      if (!none) {
        const double2 u = L.prow[P][J];
        const double den = u.x * u.x + u.y * u.y;
        double2 l = make_double2(0.0, 0.0);
        const bool act = valid && !chosen && den > 0.0;
        if (act && owner) {
          double2 x = a[CJ];
          double ur = u.x / den, ui = -u.y / den;
          l = make_double2(x.x * ur - x.y * ui, x.x * ui + x.y * ur);
          a[CJ] = l;
        }
        // the partner lane receives the multiplier
        double lx = __shfl_xor(l.x, 1, 64), ly = __shfl_xor(l.y, 1, 64);
        if (!owner) l = make_double2(lx, ly);
        if (act) {
#pragma unroll
          for (int c = 0; c < GEPP_HALF; c++) {
            const int g = GEPP_HALF * half + c;   // column of this value
            if (g > J) {
              double2 uc = L.prow[P][g];
              a[c].x -= l.x * uc.x - l.y * uc.y;
              a[c].y -= l.x * uc.y + l.y * uc.x;
            }
          }
        }
      }
      gepp2_step<J + 1>(a, valid, chosen, rowid, w, out, L);
    }
  }
}

// Loads row `id` (matrix row, or -1 for none) columns [c0, c0+w) into the lane
// pair's halves.
__device__ __forceinline__ void gepp2_load(double2 (&a)[GEPP_HALF], const double2 *__restrict__ A, int64_t lda,
                                           int id, int w, int c0) {
  const int half = threadIdx.x & 1;
#pragma unroll
  for (int c = 0; c < GEPP_HALF; c++) {
    int g = GEPP_HALF * half + c;
    a[c] = (id >= 0 && g < w) ? A[(int64_t)(c0 + g) * lda + id] : make_double2(0.0, 0.0);
  }
}

// Leaves: block b takes panel rows [b*512, b*512+512) of the m-row panel that
// starts at matrix row r0 and emits w candidate row ids (GEPP_NB slots per block).
__global__ void __launch_bounds__(GEPP_NT) tslu_leaf_kernel(const double2 *__restrict__ A, int64_t lda, int m, int w,
                                                            int r0, int c0, int *__restrict__ cand) {
  __shared__ GeppLds L;
  int r = blockIdx.x * GEPP_ROWS + (threadIdx.x >> 1);
  bool valid = r < m;
  int id = valid ? r0 + r : -1;
  double2 a[GEPP_HALF];
  gepp2_load(a, A, lda, id, w, c0);
  bool chosen = false;
  gepp2_step<0>(a, valid, chosen, id, w, cand + (size_t)blockIdx.x * GEPP_NB, L);
}

// One tournament round: block b merges candidate sets 16b..16b+15 (GEPP_NB
// slots each, ids into the matrix, -1 = empty) by GEPP on their original rows.
__global__ void __launch_bounds__(GEPP_NT) tslu_merge_kernel(const double2 *__restrict__ A, int64_t lda, int w, int c0,
                                                             const int *__restrict__ cin, int nsets,
                                                             int *__restrict__ cout) {
  __shared__ GeppLds L;
  int r = threadIdx.x >> 1;            // 0..511
  int set = GEPP_MERGE * blockIdx.x + r / GEPP_NB, slot = r % GEPP_NB;
  int id = (set < nsets && slot < w) ? cin[(size_t)set * GEPP_NB + slot] : -1;
  double2 a[GEPP_HALF];
  gepp2_load(a, A, lda, id, w, c0);
  bool chosen = false;
  gepp2_step<0>(a, id >= 0, chosen, id, w, cout + (size_t)blockIdx.x * GEPP_NB, L);
}

// Winner list -> LAPACK-style interchanges ipiv[k0+i] (row k0+i swapped with
// row ipiv[k0+i], applied in order).  One wave: lane t tracks where winner t
// currently sits while the swaps are replayed in order.
__global__ void __launch_bounds__(64) winners_to_ipiv_kernel(const int *__restrict__ win, int w, int k0,
                                                             int *__restrict__ ipiv, int *__restrict__ info) {
  const int t = threadIdx.x;
  int pos = t < w ? win[t] : -1;
  for (int i = 0; i < w; i++) {
    int src = __shfl(pos, i, 64);
    bool bad = src < 0;
    if (bad) src = k0 + i;   // rank-deficient panel: keep the row, flag it
    if (t == 0) {
      ipiv[k0 + i] = src;
      if (bad && info && *info == 0) *info = k0 + i + 1;
    }
    // the row that sat at k0+i moves to src
    if (t > i && pos == k0 + i) pos = src;
  }
}

// The w x w diagonal block (rows already in pivot order) in one 256-thread
// block: LU without pivoting in place in LDS (unit L below, U on and above),
// written back to A; then the inverses of both triangles, each by a
// right-looking sweep of rank-1 updates in a second LDS buffer, written to
// inv + blk*2*64*64 as two 64 x 64 column-major matrices (Linv, Uinv),
// identity-padded when w < 64.  With factor=0 only the inverses are formed
// from an existing packed LU (the batched form used before triangular solves).
// Thread t works on row t & 63 and on every fourth column from t >> 6.
__global__ void __launch_bounds__(256) diag_block_kernel(double2 *__restrict__ A, int64_t lda, int n, int blk0,
                                                         int bw, int factor, double2 *__restrict__ inv,
                                                         int *__restrict__ info) {
  __shared__ double2 sA[64 * 64];   // column-major, padded to 64 x 64
  __shared__ double2 sX[64 * 64];   // inverse being formed, column-major
  const int blk = blk0 + blockIdx.x, k0 = blk * bw;
  const int w = min(bw, n - k0);
  const int tid = threadIdx.x, ri = tid & 63, cg = tid >> 6;
  double2 *Ak = A + (int64_t)k0 * lda + k0;
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    int r = idx & 63, c = idx >> 6;
    sA[idx] = (r < w && c < w) ? Ak[(int64_t)c * lda + r] : make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  if (factor) {
    for (int j = 0; j < w; j++) {
      double2 u = sA[j * 64 + j];
      double den = u.x * u.x + u.y * u.y;
      if (den == 0.0) {   // uniform: den comes from LDS
        if (tid == 0 && info && *info == 0) *info = k0 + j + 1;
        continue;
      }
      if (cg == 0 && ri > j) {
        double2 ui = make_double2(u.x / den, -u.y / den);
        double2 x = sA[j * 64 + ri];
        sA[j * 64 + ri] = make_double2(x.x * ui.x - x.y * ui.y, x.x * ui.y + x.y * ui.x);
      }
      __syncthreads();
      if (ri > j) {
        const double2 l = sA[j * 64 + ri];
        for (int c = j + 1 + cg; c < 64; c += 4) {
          double2 uc = sA[c * 64 + j];
          double2 x = sA[c * 64 + ri];
          x.x -= l.x * uc.x - l.y * uc.y;
          x.y -= l.x * uc.y + l.y * uc.x;
          sA[c * 64 + ri] = x;
        }
      }
      __syncthreads();
    }
    for (int idx = tid; idx < w * w; idx += 256) {
      int r = idx % w, c = idx / w;
      Ak[(int64_t)c * lda + r] = sA[c * 64 + r];
    }
  }
  // Linv: X = I; for j: X[i][c] -= L[i][j] X[j][c] for rows i > j, columns c <= j.
  for (int idx = tid; idx < 64 * 64; idx += 256) sX[idx] = make_double2((idx & 63) == (idx >> 6) ? 1.0 : 0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < 63; j++) {
    if (ri > j) {
      const double2 l = sA[j * 64 + ri];
      for (int c = cg; c <= j; c += 4) {
        double2 xj = sX[c * 64 + j];
        double2 x = sX[c * 64 + ri];
        x.x -= l.x * xj.x - l.y * xj.y;
        x.y -= l.x * xj.y + l.y * xj.x;
        sX[c * 64 + ri] = x;
      }
    }
    __syncthreads();
  }
  double2 *Li = inv + (size_t)blk * 2 * 64 * 64, *Ui = Li + 64 * 64;
  for (int idx = tid; idx < 64 * 64; idx += 256) Li[idx] = sX[idx];
  __syncthreads();
  // Uinv: X = I; for k = 63..0: X[k][:] /= U[k][k]; X[i][c] -= U[i][k] X[k][c]
  // for rows i < k, columns c >= k.
  for (int idx = tid; idx < 64 * 64; idx += 256) sX[idx] = make_double2((idx & 63) == (idx >> 6) ? 1.0 : 0.0, 0.0);
  __syncthreads();
  for (int k = 63; k >= 0; k--) {
    if (tid >= k && tid < 64) {   // row k, column tid >= k
      double2 d = sA[k * 64 + k];
      double den = d.x * d.x + d.y * d.y;
      double2 x = sX[tid * 64 + k];
      sX[tid * 64 + k] = make_double2((x.x * d.x + x.y * d.y) / den, (x.y * d.x - x.x * d.y) / den);
    }
    __syncthreads();
    if (ri < k) {
      const double2 u = sA[k * 64 + ri];
      for (int c = k + cg; c < 64; c += 4) {
        double2 xk = sX[c * 64 + k];
        double2 x = sX[c * 64 + ri];
        x.x -= u.x * xk.x - u.y * xk.y;
        x.y -= u.x * xk.y + u.y * xk.x;
        sX[c * 64 + ri] = x;
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < 64 * 64; idx += 256) Ui[idx] = sX[idx];
}

// Block-local partial pivoting (the default LU).  The w x w diagonal block at
// (k0, k0) is factored with partial pivoting over its own rows (izamax
// criterion |re|+|im|), the reference's block LU over tree blocks pivots the
// same way (src/util/HMatrix.h:277-316).  One 256-thread block:
//  1. LU with pivoting in registers: thread (row r, wave g) holds row r of
//     the columns c = g + 4t, t < 16.  At column j the wave holding it finds
//     the pivot among the rows not yet chosen (ties to the lowest row
//     position), forms the multipliers and publishes them with the pivot's
//     lane through LDS (double-buffered by the parity of j, one barrier per
//     column); every wave then reads the pivot row's entries of its columns
//     with readlane and updates its rows.  Rows are not moved: each thread
//     tracks its row's position under the LAPACK sequence of interchanges,
//     which goes to ipiv[k0..k0+w) as absolute row indices, and the packed
//     factors are stored by position (LDS, then A).
//  2. Inverses of both triangles by recursive doubling over 1, 2, 4, .. 32
//     blocks: Linv21 = -Linv22 L21 Linv11, Uinv12 = -Uinv11 U12 Uinv22, each
//     level two small products, three barriers per level.
//  Output: inv (64 x 64 Linv then 64 x 64 Uinv, column-major, identity
//  padded when w < 64).
// Lane l's copy of x (l wave-uniform): readlane on each 32-bit half.
__device__ __forceinline__ double lane_read(double x, int l) {
  long long b = __double_as_longlong(x);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_read(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// State of one thread of the register-resident block LU: row ri's entries in
// the columns wave + 4t (a[t]), the row's position under the interchanges so
// far, and whether it has been chosen as a pivot.  The column loop and the
// update loop are unrolled by template recursion, so every index into a[] is
// a constant and the row stays in registers.
template <typename T>
struct LuRow {
  T a[16];
  int pos;
  bool done;
};

// rank-1 update of this thread's column indices [TT, END) except SKIP, those
// with c > J, by the multiplier l and the pivot row held by lane p of the
// same wave
template <int TT, int END, int J, int SKIP, typename T>
__device__ __forceinline__ void lu_update(LuRow<T> &r, int wave, int p, T l) {
  if constexpr (TT < END) {
    if (TT != SKIP && wave + 4 * TT > J) {
      // rows already chosen as pivots (and the pivot row) carry l = 0, so
      // their U entries come out unchanged without a select
      // x - l u as four fused multiply-adds
      const T uc = Cplx<T>::mk(lane_read(r.a[TT].x, p), lane_read(r.a[TT].y, p));
      T x = r.a[TT];
      x.x = std::fma(-l.x, uc.x, x.x);
      x.x = std::fma(l.y, uc.y, x.x);
      x.y = std::fma(-l.x, uc.y, x.y);
      x.y = std::fma(-l.y, uc.x, x.y);
      r.a[TT] = x;
    }
    lu_update<TT + 1, END, J, SKIP>(r, wave, p, l);
  }
}

// One step of an inclusive max-scan across the wave through DPP: lanes whose
// source lies outside the row (or whose row is masked off) take ident.
template <int CTRL, int ROWM>
__device__ __forceinline__ double dpp_max(double v, double ident) {
  const long long b = __double_as_longlong(v), ib = __double_as_longlong(ident);
  const int lo = __builtin_amdgcn_update_dpp((int)(ib & 0xffffffffll), (int)(b & 0xffffffffll), CTRL, ROWM, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ib >> 32), (int)(b >> 32), CTRL, ROWM, 0xf, false);
  return fmax(v, __longlong_as_double(((long long)hi << 32) | (unsigned int)lo));
}
template <int CTRL, int ROWM>
__device__ __forceinline__ float dpp_max(float v, float ident) {
  return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(v), CTRL, ROWM,
                                                             0xf, false)));
}
// Maximum of v over the 64 lanes (all active), wave-uniform: row_shr 1, 2,
// 4, 8 scan each row of 16 lanes, row_bcast 15 and 31 carry the row maxima
// up, and lane 63 ends with the whole wave's.
template <typename R>
__device__ __forceinline__ R wave_max(R v, R ident) {
  v = dpp_max<0x111, 0xf>(v, ident);
  v = dpp_max<0x112, 0xf>(v, ident);
  v = dpp_max<0x114, 0xf>(v, ident);
  v = dpp_max<0x118, 0xf>(v, ident);
  v = dpp_max<0x142, 0xa>(v, ident);
  v = dpp_max<0x143, 0xc>(v, ident);
  return lane_read(v, 63);
}

// Column J's pivot by the wave holding it (column J already updated): max
// |re|+|im| over the rows not yet chosen, ties to the lowest position; the
// multipliers replace the column below the pivot and go to lbuf with the
// pivot's lane to pbuf (buffers of J's parity); the interchange to ipiv.
// The maximum comes from a DPP scan, the lanes holding it from a ballot; a
// tie between several lanes is settled by a scalar loop over their positions.
template <int J, typename T>
__device__ __forceinline__ void lu_pivot(LuRow<T> &r, int ri, int k0, int *__restrict__ ipiv, int *__restrict__ info,
                                         T (*lbuf)[64], int *pbuf) {
  using R = typename Cplx<T>::R;
  constexpr int t = J >> 2, b = J & 1;
  const R v = r.done ? (R)-1.0 : std::fabs(r.a[t].x) + std::fabs(r.a[t].y);
  const R vmax = wave_max(v, (R)-1.0);
  int who, pp;
  if (vmax > (R)0.0) {
    const unsigned long long m = __ballot(v == vmax);
    who = __builtin_ctzll(m);
    pp = __builtin_amdgcn_readlane(r.pos, who);
    for (unsigned long long mm = m & (m - 1); mm; mm &= mm - 1) {
      const int l = __builtin_ctzll(mm);
      const int q = __builtin_amdgcn_readlane(r.pos, l);
      if (q < pp) {
        pp = q;
        who = l;
      }
    }
  } else {   // zero column: keep the row at position J, flag it
    who = __builtin_ctzll(__ballot(r.pos == J));
    pp = J;
    if (ri == 0 && info && *info == 0) *info = k0 + J + 1;
  }
  who = __builtin_amdgcn_readfirstlane(who);
  const T u = Cplx<T>::mk(lane_read(r.a[t].x, who), lane_read(r.a[t].y, who));
  const R den = u.x * u.x + u.y * u.y;
  T l = Cplx<T>::mk(0.0, 0.0);
  if (!r.done && ri != who && den > (R)0.0) {
    T x = r.a[t];
    const R rden = (R)1.0 / den;
    const R ur = u.x * rden, ui = -u.y * rden;
    l = Cplx<T>::mk(x.x * ur - x.y * ui, x.x * ui + x.y * ur);
  }
  if (!r.done && ri != who) r.a[t] = l;   // L below the pivot, in place
  lbuf[b][ri] = l;
  if (ri == 0) {
    pbuf[b] = who;
    ipiv[k0 + J] = k0 + pp;
  }
}

// Step J of the block LU (then J + 1, ..): one barrier publishes column J's
// pivot; every wave updates its columns.  The wave holding column J + 1
// updates that column first and finds its pivot before its other columns, so
// the next pivot search overlaps the other waves' updates.
template <int J, typename T>
__device__ __forceinline__ void lu_column(LuRow<T> &r, int w, int wave, int ri, int k0, int *__restrict__ ipiv,
                                          int *__restrict__ info, T (*lbuf)[64], int *pbuf) {
  if constexpr (J < 64) {
    constexpr int b = J & 1;
    if (J < w) {
      if (J == 0 && wave == 0) lu_pivot<0>(r, ri, k0, ipiv, info, lbuf, pbuf);
      __syncthreads();
      const int p = __builtin_amdgcn_readfirstlane(pbuf[b]);
      const T l = lbuf[b][ri];
      const int ppos = __builtin_amdgcn_readlane(r.pos, p);
      if (r.pos == J) r.pos = ppos;   // interchange of positions J and ppos
      if (ri == p) {
        r.pos = J;
        r.done = true;
      }
      if constexpr (J + 1 < 64) {
        if (J + 1 < w && wave == ((J + 1) & 3)) {
          constexpr int tn = (J + 1) >> 2;
          lu_update<tn, tn + 1, J, -1>(r, wave, p, l);   // column J + 1 (index tn in this wave)
          lu_pivot<J + 1>(r, ri, k0, ipiv, info, lbuf, pbuf);
          lu_update<0, 16, J, tn>(r, wave, p, l);
        } else {
          lu_update<0, 16, J, -1>(r, wave, p, l);
        }
      }
    }
    lu_column<J + 1>(r, w, wave, ri, k0, ipiv, info, lbuf, pbuf);
  }
}

// One level of the triangular inverses of a 64 x 64 packed LU (sA, by
// position) into sX: pairs of B-blocks at offsets o (multiples of 2B) with
// their diagonal B-blocks' inverses already in sX.
//  L: T = L21 Linv11 (rows o+B.., cols o..), Linv21 = -Linv22 T
//  U: T = U12 Uinv22 (rows o.., cols o+B..), Uinv12 = -Uinv11 T
// The L factor part of sA (strictly lower) and its U part are read-only until
// phase 1 overwrites the L21/U12 blocks with T.  B is a template argument so
// the output indexing is shifts and masks and the inner sums unroll.
template <int B, typename T>
__device__ __forceinline__ void inv_level(T *sA, T *sX, int tid) {
  constexpr int PER = 32 * B;   // outputs per matrix: (32/B) pairs x B*B
  constexpr int SLOTS = (2 * PER + 255) / 256;
  T t[SLOTS];
  auto index = [&](int q, bool &upper, int &o, int &i, int &jj) {
    upper = q >= PER;
    const int qq = upper ? q - PER : q;
    const int pair = qq / (B * B), e = qq % (B * B);
    o = pair * 2 * B;
    i = e % B;
    jj = e / B;
  };
  // phase 1: T into registers (slot s holds output q = tid + 256 s)
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    const int q = tid + 256 * sl;
    T acc = Cplx<T>::mk(0.0, 0.0);
    if (q < 2 * PER) {
      bool upper;
      int o, i, jj;
      index(q, upper, o, i, jj);
      if (!upper) {
        // T[i][jj] = sum_{k >= jj} L21[i][k] Linv11[k][jj], Linv11 unit lower
#pragma unroll
        for (int k = 0; k < B; k++) {
          if (k >= jj) {
            const T l = sA[(o + k) * 64 + o + B + i];
            const T x = (k == jj) ? Cplx<T>::mk(1.0, 0.0) : sX[(o + jj) * 64 + o + k];
            acc.x = std::fma(l.x, x.x, acc.x);
            acc.x = std::fma(-l.y, x.y, acc.x);
            acc.y = std::fma(l.x, x.y, acc.y);
            acc.y = std::fma(l.y, x.x, acc.y);
          }
        }
      } else {
        // T[i][jj] = sum_{k <= jj} U12[i][k] Uinv22[k][jj]
#pragma unroll
        for (int k = 0; k < B; k++) {
          if (k <= jj) {
            const T u = sA[(o + B + k) * 64 + o + i];
            const T x = sX[(o + B + jj) * 64 + o + B + k];
            acc.x = std::fma(u.x, x.x, acc.x);
            acc.x = std::fma(-u.y, x.y, acc.x);
            acc.y = std::fma(u.x, x.y, acc.y);
            acc.y = std::fma(u.y, x.x, acc.y);
          }
        }
      }
    }
    t[sl] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    const int q = tid + 256 * sl;
    if (q < 2 * PER) {
      bool upper;
      int o, i, jj;
      index(q, upper, o, i, jj);
      if (!upper) sA[(o + jj) * 64 + o + B + i] = t[sl];
      else sA[(o + B + jj) * 64 + o + i] = t[sl];
    }
  }
  __syncthreads();
  // phase 2: Linv21 = -Linv22 T, Uinv12 = -Uinv11 T
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl++) {
    const int q = tid + 256 * sl;
    if (q < 2 * PER) {
      bool upper;
      int o, i, jj;
      index(q, upper, o, i, jj);
      T acc = Cplx<T>::mk(0.0, 0.0);
      if (!upper) {
        // sum_{k <= i} Linv22[i][k] T[k][jj], Linv22 unit lower
#pragma unroll
        for (int k = 0; k < B; k++) {
          if (k <= i) {
            const T x = (k == i) ? Cplx<T>::mk(1.0, 0.0) : sX[(o + B + k) * 64 + o + B + i];
            const T tt = sA[(o + jj) * 64 + o + B + k];
            acc.x = std::fma(x.x, tt.x, acc.x);
            acc.x = std::fma(-x.y, tt.y, acc.x);
            acc.y = std::fma(x.x, tt.y, acc.y);
            acc.y = std::fma(x.y, tt.x, acc.y);
          }
        }
        sX[(o + jj) * 64 + o + B + i] = Cplx<T>::mk(-acc.x, -acc.y);
      } else {
        // sum_{k >= i} Uinv11[i][k] T[k][jj]
#pragma unroll
        for (int k = 0; k < B; k++) {
          if (k >= i) {
            const T x = sX[(o + k) * 64 + o + i];
            const T tt = sA[(o + B + jj) * 64 + o + k];
            acc.x = std::fma(x.x, tt.x, acc.x);
            acc.x = std::fma(-x.y, tt.y, acc.x);
            acc.y = std::fma(x.x, tt.y, acc.y);
            acc.y = std::fma(x.y, tt.x, acc.y);
          }
        }
        sX[(o + B + jj) * 64 + o + i] = Cplx<T>::mk(-acc.x, -acc.y);
      }
    }
  }
  __syncthreads();
}

// The same level for B = 16 and 32 (complex128) on FP64 MFMA: each product
// of the level (T = L21 Linv11, T = U12 Uinv22, then -Linv22 T and
// -Uinv11 T) is split into 16 x 16 output tiles; a wave takes one tile per
// round and runs v_mfma_f64_16x16x4f64 over k in steps of 4, four real
// products per complex step.  The triangles' structure (zeros, the unit
// diagonal of Linv) is applied while loading the fragments, so every lane
// stays active.  Fragment layout as in zgemm_kernel: the right factor's
// fragment is the instruction's first operand (k = lane / 16, column =
// lane % 16), the left factor's the second (row = lane % 16, k = lane / 16),
// and element e of lane l holds row l % 16, column 4e + l / 16.
template <int B>
__device__ __forceinline__ void inv_level_mfma(double2 *sA, double2 *sX, int tid) {
  constexpr int NT = B / 16;                 // tiles per block side
  constexpr int PAIRS = 32 / B;
  constexpr int TILES = 2 * PAIRS * NT * NT; // per phase, both matrices
  constexpr int ROUNDS = TILES / 4;
  static_assert(TILES % 4 == 0, "whole rounds of four waves");
  const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  auto tile_of = [&](int tt, bool &upper, int &o, int &ti, int &tj) {
    upper = tt >= TILES / 2;
    const int q = upper ? tt - TILES / 2 : tt;
    const int pair = q / (NT * NT), e = q % (NT * NT);
    o = pair * 2 * B;
    ti = e % NT;
    tj = e / NT;
  };
  auto mma = [&](dbl4 &cr, dbl4 &ci, double2 a, double2 b) {   // c += a b (a: left factor, b: right)
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(b.x, a.x, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(b.y, -a.y, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(b.y, a.x, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(b.x, a.y, ci, 0, 0, 0);
  };
  // phase 1: T tiles into registers
  dbl4 tr[ROUNDS], tim[ROUNDS];
#pragma unroll
  for (int rd = 0; rd < ROUNDS; rd++) {
    bool upper;
    int o, ti, tj;
    tile_of(wave + 4 * rd, upper, o, ti, tj);
    dbl4 cr = (dbl4){0, 0, 0, 0}, ci = (dbl4){0, 0, 0, 0};
    const int i = ti * 16 + li, jj = tj * 16 + li;   // left-factor row, right-factor column of this lane
#pragma unroll
    for (int k0 = 0; k0 < B; k0 += 4) {
      const int k = k0 + lk;
      double2 a, b;
      if (!upper) {   // L21[i][k] Linv11[k][jj] (unit lower)
        a = sA[(o + k) * 64 + o + B + i];
        b = k == jj ? one : (k > jj ? sX[(o + jj) * 64 + o + k] : zero);
      } else {        // U12[i][k] Uinv22[k][jj] (upper)
        a = sA[(o + B + k) * 64 + o + i];
        b = k <= jj ? sX[(o + B + jj) * 64 + o + B + k] : zero;
      }
      mma(cr, ci, a, b);
    }
    tr[rd] = cr;
    tim[rd] = ci;
  }
  __syncthreads();
#pragma unroll
  for (int rd = 0; rd < ROUNDS; rd++) {
    bool upper;
    int o, ti, tj;
    tile_of(wave + 4 * rd, upper, o, ti, tj);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int i = ti * 16 + li, jj = tj * 16 + 4 * e + lk;
      const double2 v = make_double2(tr[rd][e], tim[rd][e]);
      if (!upper) sA[(o + jj) * 64 + o + B + i] = v;
      else sA[(o + B + jj) * 64 + o + i] = v;
    }
  }
  __syncthreads();
  // phase 2: Linv21 = -Linv22 T, Uinv12 = -Uinv11 T (each tile reads only
  // blocks no tile of this phase writes)
#pragma unroll
  for (int rd = 0; rd < ROUNDS; rd++) {
    bool upper;
    int o, ti, tj;
    tile_of(wave + 4 * rd, upper, o, ti, tj);
    dbl4 cr = (dbl4){0, 0, 0, 0}, ci = (dbl4){0, 0, 0, 0};
    const int i = ti * 16 + li, jj = tj * 16 + li;
#pragma unroll
    for (int k0 = 0; k0 < B; k0 += 4) {
      const int k = k0 + lk;
      double2 a, b;
      if (!upper) {   // Linv22[i][k] (unit lower) T[k][jj]
        a = k == i ? one : (k < i ? sX[(o + B + k) * 64 + o + B + i] : zero);
        b = sA[(o + jj) * 64 + o + B + k];
      } else {        // Uinv11[i][k] (upper) T[k][jj]
        a = k >= i ? sX[(o + k) * 64 + o + i] : zero;
        b = sA[(o + B + jj) * 64 + o + k];
      }
      mma(cr, ci, a, b);
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int i2 = ti * 16 + li, j2 = tj * 16 + 4 * e + lk;
      const double2 v = make_double2(-cr[e], -ci[e]);
      if (!upper) sX[(o + j2) * 64 + o + B + i2] = v;
      else sX[(o + B + j2) * 64 + o + i2] = v;
    }
  }
  __syncthreads();
}

// Levels 1 .. 32 of the triangular inverses: scalar up to B = 8, MFMA for
// B = 16 and 32 in complex128.
template <typename T>
__device__ __forceinline__ void inv_levels(T *sA, T *sX, int tid) {
  inv_level<1>(sA, sX, tid);
  inv_level<2>(sA, sX, tid);
  inv_level<4>(sA, sX, tid);
  inv_level<8>(sA, sX, tid);
  if constexpr (std::is_same<T, double2>::value) {
    inv_level_mfma<16>(sA, sX, tid);
    inv_level_mfma<32>(sA, sX, tid);
  } else {
    inv_level<16>(sA, sX, tid);
    inv_level<32>(sA, sX, tid);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) diag_lu_piv_kernel(T *__restrict__ A, int64_t lda, int n, int k0,
                                                          T *__restrict__ inv, int *__restrict__ ipiv,
                                                          int *__restrict__ info) {
  using R = typename Cplx<T>::R;
  __shared__ T sA[64 * 64];   // column-major: packed LU by position, later workspace
  __shared__ T sX[64 * 64];   // packed inverses (Linv strictly below, Uinv on/above)
  __shared__ T lbuf[2][64];   // multipliers of column j, by row (parity of j)
  __shared__ int pbuf[2];     // pivot row of column j (parity of j)
  const int w = min(64, n - k0);
  const int tid = threadIdx.x, ri = tid & 63, wave = tid >> 6;
  T *Ak = A + (int64_t)k0 * lda + k0;
  LuRow<T> row;   // row ri, columns wave + 4t (identity padded beyond w)
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const int c = wave + 4 * t;
    row.a[t] = (ri < w && c < w) ? Ak[(int64_t)c * lda + ri] : Cplx<T>::mk(ri == c ? 1.0 : 0.0, 0.0);
  }
  row.pos = ri;
  row.done = false;
  lu_column<0>(row, w, wave, ri, k0, ipiv, info, lbuf, pbuf);
  // packed factors by position
#pragma unroll
  for (int t = 0; t < 16; t++) sA[(wave + 4 * t) * 64 + row.pos] = row.a[t];
  __syncthreads();
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    int r = idx & 63, c = idx >> 6;
    if (r < w && c < w) Ak[(int64_t)c * lda + r] = sA[idx];
  }
  // level 0: diagonal of the inverses (Linv unit, Uinv 1/u)
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    int r = idx & 63, c = idx >> 6;
    T v = Cplx<T>::mk(0.0, 0.0);
    if (r == c) {
      T d = sA[idx];
      R den = d.x * d.x + d.y * d.y;
      v = Cplx<T>::mk(d.x / den, -d.y / den);
    }
    sX[idx] = v;
  }
  __syncthreads();
  inv_levels(sA, sX, tid);
  T *Li = inv, *Ui = inv + 64 * 64;
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    int r = idx & 63, c = idx >> 6;
    T v = sX[idx];
    Li[idx] = r > c ? v : Cplx<T>::mk(r == c ? 1.0 : 0.0, 0.0);
    Ui[idx] = r <= c ? v : Cplx<T>::mk(0.0, 0.0);
  }
}

// max |re|+|im| over an m x n block (upper=1: only entries with row <= col,
// upper=2: only row > col),
// folded into *out as the bit pattern of a non-negative double (atomic max).
template <typename T>
__global__ void absmax_kernel(const T *__restrict__ A, int64_t lda, int m, int n, int upper,
                              unsigned long long *__restrict__ out) {
  // 256 threads; the block's maximum goes out in one atomic (a non-negative
  // double orders as its bit pattern)
  __shared__ double wmax[4];
  double v = 0.0;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    const int rmax = upper == 1 ? min(m, c + 1) : m, rmin = upper == 2 ? c + 1 : 0;
#pragma unroll 4
    for (int r = rmin + blockIdx.x * blockDim.x + threadIdx.x; r < rmax; r += gridDim.x * blockDim.x) {
      T x = A[(int64_t)c * lda + r];
      v = fmax(v, std::fabs((double)x.x) + std::fabs((double)x.y));
    }
  }
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    atomicMax(out, (unsigned long long)__double_as_longlong(v));
  }
}

// Column interchanges for X <- X P^T with P = P_k0 .. P_{k1-1} the row
// interchanges of an LU (P_i swaps i and ipiv[i]): columns i and ipiv[i] of X
// (rows [0, m)) are swapped for i = k1-1 down to k0.  One thread per row.
template <typename T>
__global__ void col_swap_kernel(T *__restrict__ X, int64_t ldx, int m, int k0, int k1,
                                const int *__restrict__ ipiv) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int i = k1 - 1; i >= k0; i--) {
    int q = ipiv[i];
    if (q != i) {
      T t = X[(int64_t)i * ldx + r];
      X[(int64_t)i * ldx + r] = X[(int64_t)q * ldx + r];
      X[(int64_t)q * ldx + r] = t;
    }
  }
}

// Interchanges confined to 64-blocks (block-local pivoting: every ipiv[p]
// lies in p's block) as one permutation per block: perm[k0 + i] = the index
// whose row (or column) ends at position k0 + i after the block's
// interchanges, applied in increasing order (rows) or decreasing order
// (REV: the column interchanges X P^T).  One 64-thread block per 64-block.
template <bool REV>
__global__ void __launch_bounds__(64) block_perm_kernel(const int *__restrict__ ipiv, int n, int *__restrict__ perm) {
  __shared__ int idx[64];
  const int k0 = blockIdx.x * 64, w = min(64, n - k0), t = threadIdx.x;
  idx[t] = k0 + t;
  __syncthreads();
  if (t == 0) {
    for (int s = 0; s < w; s++) {
      const int p = REV ? w - 1 - s : s, q = ipiv[k0 + p] - k0;
      if (q < 0 || q >= w) continue;   // not block-local: callers never pass such ipiv
      const int a = idx[p];
      idx[p] = idx[q];
      idx[q] = a;
    }
  }
  __syncthreads();
  if (t < w) perm[k0 + t] = idx[t];
}

// X (n x r) rows permuted within each 64-block: row k0 + i <- row perm[k0 + i].
// A block stages 64 rows x 4 columns in LDS (coalesced both ways).
template <typename T>
__global__ void __launch_bounds__(256) row_permute_kernel(T *__restrict__ X, int64_t ldx, int n, int r,
                                                          const int *__restrict__ perm) {
  __shared__ T s[4][64];
  const int k0 = blockIdx.x * 64, i = threadIdx.x & 63, cc = threadIdx.x >> 6;
  const int c = blockIdx.y * 4 + cc, w = min(64, n - k0);
  const bool in = i < w && c < r;
  if (in) s[cc][i] = X[(int64_t)c * ldx + k0 + i];
  __syncthreads();
  if (in) X[(int64_t)c * ldx + k0 + i] = s[cc][perm[k0 + i] - k0];
}

// X (m x n) columns permuted within each 64-block: column k0 + j <- column
// perm[k0 + j].  A block stages 32 rows x 64 columns in LDS.
template <typename T>
__global__ void __launch_bounds__(256) col_permute_kernel(T *__restrict__ X, int64_t ldx, int m, int n,
                                                          const int *__restrict__ perm) {
  __shared__ T s[64][33];
  const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 64, w = min(64, n - k0);
  const int row = threadIdx.x & 31, j0 = threadIdx.x >> 5;
  const bool rin = r0 + row < m;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int j = j0 + 8 * q;
    if (rin && j < w) s[j][row] = X[(int64_t)(k0 + j) * ldx + r0 + row];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int j = j0 + 8 * q;
    if (rin && j < w) X[(int64_t)(k0 + j) * ldx + r0 + row] = s[perm[k0 + j] - k0][row];
  }
}

// Row interchanges: for p in [k0, k1): swap rows p and ipiv[p] in every column
// of [c0, c1) except [skip0, skip1).
template <typename T>
__global__ void laswp_kernel(T *__restrict__ A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1,
                             const int *__restrict__ ipiv) {
  int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1 || (c >= skip0 && c < skip1)) return;
  T *col = A + (int64_t)c * lda;
  for (int p = k0; p < k1; p++) {
    int q = ipiv[p];
    if (q != p) {
      T t = col[p];
      col[p] = col[q];
      col[q] = t;
    }
  }
}

// G <- G + G^H in place, by 32 x 32 tile pairs (lower tile (bi, bj), bi >= bj,
// and its mirror) staged through LDS so both sides are read and written
// coalesced.  Each block owns its pair; diagonal tiles are their own mirror.
__global__ void __launch_bounds__(256) symm_kernel(double2 *__restrict__ G, int64_t ld, int n) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi < bj) return;
  __shared__ double2 lo[32][33], up[32][33];   // lo[c][r] = G(r0+r, c0+c), up[c][r] = G(c0+r, r0+c)
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int r0 = bi * 32, c0 = bj * 32;
  for (int c = ty; c < 32; c += 8) {
    if (r0 + tx < n && c0 + c < n) lo[c][tx] = G[(int64_t)(c0 + c) * ld + r0 + tx];
    if (c0 + tx < n && r0 + c < n) up[c][tx] = G[(int64_t)(r0 + c) * ld + c0 + tx];
  }
  __syncthreads();
  // lower side and diagonal: G(i, j) = G(i, j) + conj(G(j, i)), i = r0 + tx, j = c0 + c
  for (int c = ty; c < 32; c += 8) {
    const int i = r0 + tx, j = c0 + c;
    if (i >= n || j >= n || i < j) continue;
    double2 a = lo[c][tx];
    if (i == j) {
      G[(int64_t)j * ld + i] = make_double2(a.x + a.x, 0.0);
    } else {
      double2 b = up[tx][c];
      G[(int64_t)j * ld + i] = make_double2(a.x + b.x, a.y - b.y);
    }
  }
  // upper side: G(j, i) = conj of the new G(i, j), j = c0 + tx, i = r0 + c
  for (int c = ty; c < 32; c += 8) {
    const int i = r0 + c, j = c0 + tx;
    if (i >= n || j >= n || i <= j) continue;
    double2 a = lo[tx][c], b = up[c][tx];
    G[(int64_t)i * ld + j] = make_double2(a.x + b.x, b.y - a.y);
  }
}


// ------------------------------------------------------------------- CGEMM
// complex64 C = alpha A B + beta C on FP32 MFMA v_mfma_f32_16x16x4f32, the
// same staging as zgemm_kernel: BM x BN block tile, WM x WN waves of
// (BM/WM) x (BN/WN), k staged through LDS in steps of GBK as re/im planes
// with the next step prefetched into registers.  Four real products per
// complex k step into two FP32 accumulators, or with G3M three (Ar Br,
// Ai Bi, (Ar + Ai)(Br + Bi)) into three, as zgemm_kernel.
//
// The k loop of one block's C tile, as zgemm_tile (TA: A read as the
// conjugate transpose of a k x m array).
template <int BM, int BN, int WM, int WN, int GBK, bool G3M, bool TA>
__device__ __forceinline__ void cgemm_tile(int m, int n, int k, const float2 *__restrict__ A, int64_t lda,
                                           const float2 *__restrict__ B, int64_t ldb, int m0, int n0,
                                           flt4 (&acc)[G3M ? 3 : 2][BM / WM / 16][BN / WN / 16]) {
  constexpr int NT = 64 * WM * WN;
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int AL = BM * GBK / NT;
  constexpr int BL = BN * GBK / NT;
  static_assert(AL >= 1 && BL >= 1 && NT % BM == 0 && NT % GBK == 0, "tile shape");
  __shared__ float sAr[GBK][BM + 1], sAi[GBK][BM + 1];
  __shared__ float sBr[GBK][BN + 1], sBi[GBK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  constexpr int NACC = G3M ? 3 : 2;
#pragma unroll
  for (int q = 0; q < NACC; q++)
#pragma unroll
    for (int a = 0; a < FM; a++)
#pragma unroll
      for (int b = 0; b < FN; b++) acc[q][a][b] = (flt4){0.f, 0.f, 0.f, 0.f};
  float2 ra[AL], rb[BL];
  const int arow = TA ? tid / GBK : tid % BM, acol0 = TA ? tid % GBK : tid / BM;
  constexpr int ASTEP = TA ? NT / GBK : NT / BM;
  const int bkr = tid % GBK, bcol0 = tid / GBK;
  constexpr int BSTEP = NT / GBK;
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int q = 0; q < AL; q++) {
      if constexpr (TA) {
        int row = m0 + arow + ASTEP * q, kk = k0 + acol0;
        ra[q] = (row < m && kk < k) ? A[(int64_t)row * lda + kk] : make_float2(0.f, 0.f);
      } else {
        int col = k0 + acol0 + ASTEP * q, row = m0 + arow;
        ra[q] = (row < m && col < k) ? A[(int64_t)col * lda + row] : make_float2(0.f, 0.f);
      }
    }
#pragma unroll
    for (int q = 0; q < BL; q++) {
      int col = n0 + bcol0 + BSTEP * q, row = k0 + bkr;
      rb[q] = (row < k && col < n) ? B[(int64_t)col * ldb + row] : make_float2(0.f, 0.f);
    }
  };
  load_tiles(0);
  for (int k0 = 0; k0 < k; k0 += GBK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < AL; q++) {
      if constexpr (TA) {
        sAr[acol0][arow + ASTEP * q] = ra[q].x;
        sAi[acol0][arow + ASTEP * q] = -ra[q].y;
      } else {
        sAr[acol0 + ASTEP * q][arow] = ra[q].x;
        sAi[acol0 + ASTEP * q][arow] = ra[q].y;
      }
    }
#pragma unroll
    for (int q = 0; q < BL; q++) {
      sBr[bkr][bcol0 + BSTEP * q] = rb[q].x;
      sBi[bkr][bcol0 + BSTEP * q] = rb[q].y;
    }
    __syncthreads();
    if (k0 + GBK < k) load_tiles(k0 + GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      float ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
      for (int x = 0; x < FM; x++) {
        ar[x] = sAr[kk + lk][wr * (BM / WM) + x * 16 + li];
        ai[x] = sAi[kk + lk][wr * (BM / WM) + x * 16 + li];
        nai[x] = -ai[x];
      }
#pragma unroll
      for (int y = 0; y < FN; y++) {
        br[y] = sBr[kk + lk][wc * (BN / WN) + y * 16 + li];
        bi[y] = sBi[kk + lk][wc * (BN / WN) + y * 16 + li];
      }
      if constexpr (G3M) {
        float as[FM], bs[FN];
#pragma unroll
        for (int x = 0; x < FM; x++) as[x] = ar[x] + ai[x];
#pragma unroll
        for (int y = 0; y < FN; y++) bs[y] = br[y] + bi[y];
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) {
            acc[0][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(br[y], ar[x], acc[0][x][y], 0, 0, 0);
            acc[1][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi[y], ai[x], acc[1][x][y], 0, 0, 0);
            acc[NACC - 1][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(bs[y], as[x], acc[NACC - 1][x][y], 0, 0, 0);
          }
      } else {
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) {
            acc[0][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(br[y], ar[x], acc[0][x][y], 0, 0, 0);
            acc[1][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi[y], ar[x], acc[1][x][y], 0, 0, 0);
          }
#pragma unroll
        for (int x = 0; x < FM; x++)
#pragma unroll
          for (int y = 0; y < FN; y++) {
            acc[0][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi[y], nai[x], acc[0][x][y], 0, 0, 0);
            acc[1][x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(br[y], ai[x], acc[1][x][y], 0, 0, 0);
          }
      }
    }
  }
}

template <int BM, int BN, int WM, int WN, int GBK, bool G3M>
__global__ void __launch_bounds__(64 * WM * WN)
cgemm_kernel(int m, int n, int k, float2 alpha, const float2 *__restrict__ A, int64_t lda,
             const float2 *__restrict__ B, int64_t ldb, float2 beta, float2 *__restrict__ C, int64_t ldc) {
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int NACC = G3M ? 3 : 2;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  flt4 acc[NACC][FM][FN];
  cgemm_tile<BM, BN, WM, WN, GBK, G3M, false>(m, n, k, A, lda, B, ldb, m0, n0, acc);
  // v_mfma_f32_16x16x4f32 output: D[i][j] of a 16 x 16 block in lane
  // j + 16 (i / 4), element i % 4; with the B tile as the first operand D is
  // the transposed block, so C[i][j] sits in lane i + 16 (j / 4), element
  // j % 4 and lanes run down a column of C (coalesced).  Old C values are
  // loaded a row of blocks at a time, as in zgemm_kernel.
  const bool rd = beta.x != 0.f || beta.y != 0.f;
#pragma unroll
  for (int x = 0; x < FM; x++) {
    const int row = m0 + wr * (BM / WM) + x * 16 + li;
    float2 old[FN][4];
    if (rd) {
#pragma unroll
      for (int y = 0; y < FN; y++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int col = n0 + wc * (BN / WN) + y * 16 + 4 * lk + r;
          old[y][r] = C[(row < m && col < n) ? (int64_t)col * ldc + row : 0];
        }
    }
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int col = n0 + wc * (BN / WN) + y * 16 + 4 * lk + r;
        if (row < m && col < n) {
          float cr, ci;
          if constexpr (G3M) {
            float t1 = acc[0][x][y][r], t2 = acc[1][x][y][r];
            cr = t1 - t2;
            ci = acc[NACC - 1][x][y][r] - t1 - t2;
          } else {
            cr = acc[0][x][y][r];
            ci = acc[1][x][y][r];
          }
          float vr = alpha.x * cr - alpha.y * ci, vi = alpha.x * ci + alpha.y * cr;
          if (rd) {
            float2 o = old[y][r];
            vr += beta.x * o.x - beta.y * o.y;
            vi += beta.x * o.y + beta.y * o.x;
          }
          C[(int64_t)col * ldc + row] = make_float2(vr, vi);
        }
      }
  }
}

// Hermitian trace in complex64, as zgemm_htrace_kernel (partials in FP64).
template <int BM, int BN, int WM, int WN, int GBK, bool G3M>
__global__ void __launch_bounds__(64 * WM * WN)
cgemm_htrace_kernel(int n, int k, const float2 *__restrict__ A, int64_t lda, const float2 *__restrict__ B,
                    int64_t ldb, const float2 *__restrict__ E, int64_t lde, double2 *__restrict__ partial) {
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int NACC = G3M ? 3 : 2;
  constexpr int NW = WM * WN;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  if (m0 > n0 + BN - 1) {
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(0.0, 0.0);
    return;
  }
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  flt4 acc[NACC][FM][FN];
  cgemm_tile<BM, BN, WM, WN, GBK, G3M, true>(n, n, k, A, lda, B, ldb, m0, n0, acc);
  __shared__ double sred[2][NW];
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int x = 0; x < FM; x++)
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // element layout as in cgemm_kernel's epilogue
        int row = m0 + wr * (BM / WM) + x * 16 + li;
        int col = n0 + wc * (BN / WN) + y * 16 + 4 * lk + r;
        if (row <= col && col < n) {
          double cr, ci;
          if constexpr (G3M) {
            double t1 = acc[0][x][y][r], t2 = acc[1][x][y][r];
            cr = t1 - t2;
            ci = (double)acc[2][x][y][r] - t1 - t2;
          } else {
            cr = acc[0][x][y][r];
            ci = acc[1][x][y][r];
          }
          float2 e = E[(int64_t)col * lde + row];
          double re = cr * (double)e.x + ci * (double)e.y;
          if (row < col) {
            sr += 2.0 * re;
          } else {
            sr += re;
            si += cr * (double)e.y - ci * (double)e.x;
          }
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
    double tr = 0.0, ti = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) { tr += sred[0][w]; ti += sred[1][w]; }
    partial[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(tr, ti);
  }
}

// p[i] <- a value in [-1, 1) hashed from (seed, i) (splitmix64 finaliser):
// operands with realistic bit patterns for timing runs (T double or float).
template <typename T>
__global__ void fill_hash_kernel(T *__restrict__ p, size_t n, unsigned long long seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[i] = (T)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
  }
}

// C (n x m) <- S B^T S for B (m x n), S = diag(+1, -1, +1, ...) over the
// interleaved [J, M] unknowns: the exterior block Z(2,1) from Z(1,2).  Two
// disjoint bodies couple through far pairs only, whose quadrature is the same
// rule on both triangles, so Z_ext is reciprocal to rounding there: the J-J
// and M-M couplings are symmetric and the J-M couplings change sign under
// transposition.  32 x 32 tiles through LDS, both sides coalesced.
__global__ void __launch_bounds__(256) signed_transpose_kernel(const double2 *__restrict__ B, int64_t ldb, int m, int n,
                                                               double2 *__restrict__ C, int64_t ldc) {
  __shared__ double2 t[32][33];
  const int b0 = blockIdx.x * 32, c0 = blockIdx.y * 32;   // rows of B, columns of B
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    int r = b0 + tx, c = c0 + j;
    if (r < m && c < n) t[j][tx] = B[(int64_t)c * ldb + r];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    int r = c0 + tx, c = b0 + j;   // C(r, c) = s_r s_c B(c, r)
    if (r < n && c < m) {
      double2 v = t[tx][j];
      if ((r ^ c) & 1) v = make_double2(-v.x, -v.y);
      C[(int64_t)c * ldc + r] = v;
    }
  }
}

// complex128 <-> complex64 copies of an m x n block (one thread per row,
// blockIdx.y per column); z2c rounds to nearest.
__global__ void z2c_kernel(const double2 *__restrict__ src, int64_t lds, float2 *__restrict__ dst, int64_t ldd, int m,
                           int n) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    double2 v = src[(int64_t)c * lds + r];
    dst[(int64_t)c * ldd + r] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ void c2z_kernel(const float2 *__restrict__ src, int64_t lds, double2 *__restrict__ dst, int64_t ldd, int m,
                           int n) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    float2 v = src[(int64_t)c * lds + r];
    dst[(int64_t)c * ldd + r] = make_double2(v.x, v.y);
  }
}

// dst (complex128) <- dst + src (complex64)
__global__ void c2z_add_kernel(const float2 *__restrict__ src, int64_t lds, double2 *__restrict__ dst, int64_t ldd,
                               int m, int n) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    float2 v = src[(int64_t)c * lds + r];
    double2 o = dst[(int64_t)c * ldd + r];
    dst[(int64_t)c * ldd + r] = make_double2(o.x + (double)v.x, o.y + (double)v.y);
  }
}

// dst (complex64) <- dst + src (complex128) rounded: the refinement update.
__global__ void add_z_to_c_kernel(const double2 *__restrict__ src, int64_t lds, float2 *__restrict__ dst, int64_t ldd,
                                  int m, int n) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    double2 v = src[(int64_t)c * lds + r];
    float2 o = dst[(int64_t)c * ldd + r];
    dst[(int64_t)c * ldd + r] = make_float2((float)(o.x + v.x), (float)(o.y + v.y));
  }
}

// Sum of |re|^2 + |im|^2 over an m x n complex128 block: each block of 256
// threads writes one partial (grid (gx, gy), partial[by * gx + bx]);
// sumsq_finish adds them in index order, so the norm does not depend on
// scheduling (no floating-point atomics).
__global__ void __launch_bounds__(256) sumsq_kernel(const double2 *__restrict__ A, int64_t lda, int m, int n,
                                                    double *__restrict__ partial) {
  __shared__ double ws[4];
  double v = 0.0;
  for (int c = blockIdx.y; c < n; c += gridDim.y)
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
      const double2 x = A[(int64_t)c * lda + r];
      v = fma(x.x, x.x, fma(x.y, x.y, v));
    }
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// out[0] <- (add ? out[0] : 0) + sum of partial[0..np) in a fixed order.
__global__ void __launch_bounds__(256) sumsq_finish_kernel(const double *__restrict__ partial, int np, int add,
                                                           double *__restrict__ out) {
  __shared__ double s[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) v += partial[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (add ? out[0] : 0.0) + s[0];
}

// --------------------------------------------------- few-column kernels
// Substitutions and products with at most FEW_MAX_COLS columns: the error
// probe's.
constexpr int FEW_MAX_COLS = 8;

template <typename T>
__device__ __forceinline__ T cfma(T a, T b, T c) {   // c + a b
  c.x = fma(a.x, b.x, fma(-a.y, b.y, c.x));
  c.y = fma(a.x, b.y, fma(a.y, b.x, c.y));
  return c;
}

// ------------------------------------------- few-column substitution step
// Step k of a substitution with a packed LU for r <= 8 right-hand sides, as
// one launch with no waiting between workgroups: every workgroup forms
// x_k = inv_k b_k from the rows of block k in Bw (only read in this launch),
// workgroup 0 writes x_k to Xo, and workgroup g >= 1 subtracts the coupling
// A(rows of its chunk, block k) x_k from its 64 rows of Bw -- the chunks
// below block k (LOWER: forward with the unit-lower L) or above it (upper:
// backward with U).  Thread (i, q) takes row i and the quarter q of the inner
// index for all columns; its slice of the inverse and of the coupling block
// are loaded before anything waits, and the quarters meet in LDS.
template <typename T, bool LOWER>
__global__ void __launch_bounds__(256) trsm_step_kernel(const T *__restrict__ A, int64_t lda, int n,
                                                        const T *__restrict__ inv, T *__restrict__ Bw, int64_t ldb,
                                                        T *__restrict__ Xo, int64_t ldx, int r, int k) {
  using R = typename Cplx<T>::R;
  __shared__ T xs[64][FEW_MAX_COLS];
  __shared__ T part[4][FEW_MAX_COLS][64];
  const int t = threadIdx.x, i = t & 63, q = t >> 6, l0 = q * 16;
  const int g = blockIdx.x;
  const int r0 = k * 64, w = min(64, n - r0);
  const int row0 = LOWER ? (k + g) * 64 : (g - 1) * 64;
  const bool live = g > 0 && row0 + i < n;
  const T *M = inv + (size_t)k * 2 * 64 * 64 + (LOWER ? 0 : 64 * 64);
  T m[16], a[16];
#pragma unroll
  for (int l = 0; l < 16; l++) {
    m[l] = M[(l0 + l) * 64 + i];
    a[l] = (live && l0 + l < w) ? A[(int64_t)(r0 + l0 + l) * lda + row0 + i] : Cplx<T>::mk((R)0, (R)0);
  }
  for (int e = t; e < 64 * r; e += 256) {
    const int l = e & 63, c = e >> 6;
    xs[l][c] = l < w ? Bw[(int64_t)c * ldb + r0 + l] : Cplx<T>::mk((R)0, (R)0);
  }
  __syncthreads();
  T acc[FEW_MAX_COLS];
#pragma unroll
  for (int c = 0; c < FEW_MAX_COLS; c++) {
    acc[c] = Cplx<T>::mk((R)0, (R)0);
    if (c < r) {
#pragma unroll
      for (int l = 0; l < 16; l++) acc[c] = cfma(m[l], xs[l0 + l][c], acc[c]);
      part[q][c][i] = acc[c];
    }
  }
  __syncthreads();
  T xk[2];
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int c = q + 4 * u;
    if (c < r) {
      const T s0 = part[0][c][i], s1 = part[1][c][i], s2 = part[2][c][i], s3 = part[3][c][i];
      xk[u] = Cplx<T>::mk((s0.x + s1.x) + (s2.x + s3.x), (s0.y + s1.y) + (s2.y + s3.y));
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int c = q + 4 * u;
    if (c < r) xs[i][c] = i < w ? xk[u] : Cplx<T>::mk((R)0, (R)0);
  }
  if (g == 0) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int c = q + 4 * u;
      if (c < r && i < w) Xo[(int64_t)c * ldx + r0 + i] = xk[u];
    }
    return;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < FEW_MAX_COLS; c++) {
    acc[c] = Cplx<T>::mk((R)0, (R)0);
    if (c < r) {
#pragma unroll
      for (int l = 0; l < 16; l++) acc[c] = cfma(a[l], xs[l0 + l][c], acc[c]);
      part[q][c][i] = acc[c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int c = q + 4 * u;
    if (c < r && live) {
      const T s0 = part[0][c][i], s1 = part[1][c][i], s2 = part[2][c][i], s3 = part[3][c][i];
      const T b = Bw[(int64_t)c * ldb + row0 + i];
      Bw[(int64_t)c * ldb + row0 + i] =
          Cplx<T>::mk(b.x - ((s0.x + s1.x) + (s2.x + s3.x)), b.y - ((s0.y + s1.y) + (s2.y + s3.y)));
    }
  }
}

// ------------------------------------------------------------ skinny ZGEMM
// C (m x r, r <= 8) <- alpha A (m x K) B (K x r) + beta C in complex128: the
// probe's products, bound by reading A once.  Workgroup (bx, by) takes rows
// [64 bx, 64 bx + 64) and the by-th of nsplit ranges of K; thread (i, q) row i
// and every fourth k of the range (coalesced column reads of A), the B rows
// staged in LDS.  The quarters meet in LDS and each workgroup writes its
// partial; skinny_finish adds the partials in split order (deterministic).
constexpr int SKINNY_KCHUNK = 64;

__global__ void __launch_bounds__(256) zgemm_skinny_kernel(int m, int K, int r, const double2 *__restrict__ A,
                                                           int64_t lda, const double2 *__restrict__ B, int64_t ldb,
                                                           int kper, double2 *__restrict__ part) {
  __shared__ double2 bs[SKINNY_KCHUNK][FEW_MAX_COLS];
  __shared__ double2 red[4][FEW_MAX_COLS][64];
  const int t = threadIdx.x, i = t & 63, q = t >> 6;
  const int row = blockIdx.x * 64 + i;
  const bool live = row < m;
  const int kb0 = blockIdx.y * kper, kb1 = min(K, kb0 + kper);
  double2 acc[FEW_MAX_COLS];
#pragma unroll
  for (int c = 0; c < FEW_MAX_COLS; c++) acc[c] = make_double2(0.0, 0.0);
  for (int k0 = kb0; k0 < kb1; k0 += SKINNY_KCHUNK) {
    const int kw = min(SKINNY_KCHUNK, kb1 - k0);
    __syncthreads();
    for (int e = t; e < SKINNY_KCHUNK * r; e += 256) {
      const int l = e % SKINNY_KCHUNK, c = e / SKINNY_KCHUNK;
      bs[l][c] = l < kw ? B[(int64_t)c * ldb + k0 + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    double2 a[SKINNY_KCHUNK / 4];
#pragma unroll
    for (int u = 0; u < SKINNY_KCHUNK / 4; u++) {
      const int l = q + 4 * u;
      a[u] = (live && l < kw) ? A[(int64_t)(k0 + l) * lda + row] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < SKINNY_KCHUNK / 4; u++)
#pragma unroll
      for (int c = 0; c < FEW_MAX_COLS; c++)
        if (c < r) acc[c] = cfma(a[u], bs[q + 4 * u][c], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < FEW_MAX_COLS; c++)
    if (c < r) red[q][c][i] = acc[c];
  __syncthreads();
  for (int c = q; c < r; c += 4) {
    const double2 s0 = red[0][c][i], s1 = red[1][c][i], s2 = red[2][c][i], s3 = red[3][c][i];
    if (live)
      part[((int64_t)blockIdx.y * r + c) * m + row] =
          make_double2((s0.x + s1.x) + (s2.x + s3.x), (s0.y + s1.y) + (s2.y + s3.y));
  }
}

__global__ void zgemm_skinny_finish_kernel(int m, int r, int nsplit, const double2 *__restrict__ part,
                                           double2 alpha, double2 beta, double2 *__restrict__ C, int64_t ldc) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (row >= m) return;
  double2 s = make_double2(0.0, 0.0);
  for (int p = 0; p < nsplit; p++) {
    const double2 v = part[((int64_t)p * r + c) * m + row];
    s.x += v.x;
    s.y += v.y;
  }
  double2 o = make_double2(0.0, 0.0);
  if (beta.x != 0.0 || beta.y != 0.0) o = C[(int64_t)c * ldc + row];
  double2 v = make_double2(beta.x * o.x - beta.y * o.y, beta.x * o.y + beta.y * o.x);
  v = cfma(alpha, s, v);
  C[(int64_t)c * ldc + row] = v;
}

// out[j] <- sum_i conj(G(i, j)) D(i, j) for the n columns of two m-row
// complex128 blocks: one 256-thread block per column, summed in a fixed
// order.
__global__ void __launch_bounds__(256) zdotc_cols_kernel(const double2 *__restrict__ G, int64_t ldg,
                                                         const double2 *__restrict__ D, int64_t ldd, int m,
                                                         double2 *__restrict__ out) {
  __shared__ double sr[256], si[256];
  const int j = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) {
    const double2 g = G[(int64_t)j * ldg + i], d = D[(int64_t)j * ldd + i];
    re = fma(g.x, d.x, fma(g.y, d.y, re));
    im = fma(g.x, d.y, fma(-g.y, d.x, im));
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = make_double2(sr[0], si[0]);
}

// dst (complex128) <- dst + alpha src over an m x n block.
__global__ void zaxpy_kernel(double2 alpha, const double2 *__restrict__ src, int64_t lds, double2 *__restrict__ dst,
                             int64_t ldd, int m, int n) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    const double2 v = src[(int64_t)c * lds + r];
    double2 o = dst[(int64_t)c * ldd + r];
    o.x = fma(alpha.x, v.x, fma(-alpha.y, v.y, o.x));
    o.y = fma(alpha.x, v.y, fma(alpha.y, v.x, o.y));
    dst[(int64_t)c * ldd + r] = o;
  }
}

// p[0..n) <- 0 (32-bit words): clears counters without a memset node.
__global__ void fill_zero_kernel(unsigned int *__restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// C (m x n, complex128) <- I - C on the square part (the residual of an
// identity right-hand side).
__global__ void identity_minus_kernel(double2 *__restrict__ Cm, int64_t ldc, int n) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (r >= n) return;
  double2 v = Cm[(int64_t)c * ldc + r];
  Cm[(int64_t)c * ldc + r] = make_double2((r == c ? 1.0 : 0.0) - v.x, -v.y);
}

template <int BM, int BN, int WM, int WN, int GBK, bool G3M = false>
void launch_cgemm(int m, int n, int k, float2 alpha, const float2 *A, int64_t lda, const float2 *B, int64_t ldb,
                  float2 beta, float2 *C, int64_t ldc, hipStream_t st) {
  dim3 grid((m + BM - 1) / BM, (n + BN - 1) / BN);
  hipLaunchKernelGGL((cgemm_kernel<BM, BN, WM, WN, GBK, G3M>), grid, dim3(64 * WM * WN), 0, st, m, n, k, alpha, A,
                     lda, B, ldb, beta, C, ldc);
}

template <int BM, int BN, int WM, int WN, int GBK, int OCC, bool G3M>
void launch_zgemm(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                  double2 beta, double2 *C, int64_t ldc, hipStream_t st) {
  dim3 grid((m + BM - 1) / BM, (n + BN - 1) / BN);
  hipLaunchKernelGGL((zgemm_kernel<BM, BN, WM, WN, GBK, OCC, G3M>), grid, dim3(64 * WM * WN), 0, st, m, n, k, alpha, A,
                     lda, B, ldb, beta, C, ldc);
}


// C (m x n) <- alpha A^H B + beta C with A (k x m) and B (k x n) column-major:
// the projections and small Gram products of the low-rank coupling, whose
// output is small and whose k is long.  K is split over gridDim.z: with one
// split the epilogue applies alpha and beta to C; otherwise split z writes its
// partial tile to part + z m n (leading dimension m) and ta_finish adds the
// splits in z order (deterministic, no atomics).
template <int BM, int BN, int WM, int WN, int GBK>
__global__ void __launch_bounds__(64 * WM * WN)
zgemm_ta_kernel(int m, int n, int k, int kper, const double2 *__restrict__ A, int64_t lda,
                const double2 *__restrict__ B, int64_t ldb, double2 alpha, double2 beta, double2 *__restrict__ C,
                int64_t ldc, double2 *__restrict__ part) {
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kz0 = blockIdx.z * kper, kl = min(kper, k - kz0);
  dbl4 acc[3][FM][FN];
  zgemm_tile<BM, BN, WM, WN, GBK, true, true>(m, n, kl, A + kz0, lda, B + kz0, ldb, m0, n0, acc);
  const bool split = gridDim.z > 1;
  double2 *out = split ? part + (int64_t)blockIdx.z * m * n : C;
  const int64_t ldo = split ? m : ldc;
  const bool rd = !split && (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int x = 0; x < FM; x++)
#pragma unroll
    for (int y = 0; y < FN; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // element layout as in zgemm_kernel's epilogue
        const int row = m0 + wr * (BM / WM) + x * 16 + li;
        const int col = n0 + wc * (BN / WN) + y * 16 + 4 * r + lk;
        if (row < m && col < n) {
          const double t1 = acc[0][x][y][r], t2 = acc[1][x][y][r];
          const double cr = t1 - t2, ci = acc[2][x][y][r] - t1 - t2;
          double2 v;
          if (split) {
            v = make_double2(cr, ci);
          } else {
            v = make_double2(alpha.x * cr - alpha.y * ci, alpha.x * ci + alpha.y * cr);
            if (rd) {
              const double2 o = C[(int64_t)col * ldc + row];
              v.x += beta.x * o.x - beta.y * o.y;
              v.y += beta.x * o.y + beta.y * o.x;
            }
          }
          out[(int64_t)col * ldo + row] = v;
        }
      }
}

__global__ void zgemm_ta_finish_kernel(int m, int n, int nz, const double2 *__restrict__ part, double2 alpha,
                                       double2 beta, double2 *__restrict__ C, int64_t ldc) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y;
  if (row >= m) return;
  double2 s = make_double2(0.0, 0.0);
  for (int z = 0; z < nz; z++) {
    const double2 v = part[(int64_t)z * m * n + (int64_t)col * m + row];
    s.x += v.x;
    s.y += v.y;
  }
  double2 v = make_double2(alpha.x * s.x - alpha.y * s.y, alpha.x * s.y + alpha.y * s.x);
  if (beta.x != 0.0 || beta.y != 0.0) {
    const double2 o = C[(int64_t)col * ldc + row];
    v.x += beta.x * o.x - beta.y * o.y;
    v.y += beta.x * o.y + beta.y * o.x;
  }
  C[(int64_t)col * ldc + row] = v;
}

}  // namespace

// ------------------------------------------------------------------ launchers
// X (n x r, r <= 8) <- L^-1 X (lower = 1) or U^-1 X (lower = 0) with a
// packed LU and its diagonal-block inverses: one trsm_step_kernel launch per
// 64-row block, the solution gathered in Y (n x r, leading dimension ldy) and
// copied back to X.
template <typename T>
static int trsm_few(const T *A, int64_t lda, int n, const T *inv, T *X, int64_t ldx, int r, int lower, T *Y,
                    int64_t ldy, hipStream_t st) {
  if (n <= 0 || r <= 0) return 0;
  if (r > FEW_MAX_COLS) return -1;
  const int nb = (n + 63) / 64;
  for (int s = 0; s < nb; s++) {
    const int k = lower ? s : nb - 1 - s;
    const int grid = 1 + (lower ? nb - 1 - k : k);
    if (lower)
      hipLaunchKernelGGL((trsm_step_kernel<T, true>), dim3(grid), dim3(256), 0, st, A, lda, n, inv, X, ldx, Y, ldy, r,
                         k);
    else
      hipLaunchKernelGGL((trsm_step_kernel<T, false>), dim3(grid), dim3(256), 0, st, A, lda, n, inv, X, ldx, Y, ldy,
                         r, k);
  }
  if (hipError_t e = hipGetLastError()) return (int)e;
  return (int)hipMemcpy2DAsync(X, sizeof(T) * ldx, Y, sizeof(T) * ldy, sizeof(T) * n, r, hipMemcpyDeviceToDevice, st);
}

// Block-local interchanges ipiv[0..n) applied to the rows of X (n x r) or,
// cols = 1, to the columns of X (m x n) as X P^T, through the permutation of
// each 64-block (perm: n ints of workspace).
template <typename T>
static int permute_local(T *X, int64_t ldx, int m, int n, int r, int cols, const int *ipiv, int *perm,
                         hipStream_t st) {
  if (n <= 0) return 0;
  const int nb = (n + 63) / 64;
  if (cols) hipLaunchKernelGGL(block_perm_kernel<true>, dim3(nb), dim3(64), 0, st, ipiv, n, perm);
  else hipLaunchKernelGGL(block_perm_kernel<false>, dim3(nb), dim3(64), 0, st, ipiv, n, perm);
  if (cols) {
    if (m <= 0) return 0;
    if (nb > 65535) return -1;
    hipLaunchKernelGGL(col_permute_kernel<T>, dim3((m + 31) / 32, nb), dim3(256), 0, st, X, ldx, m, n,
                       (const int *)perm);
  } else {
    if (r <= 0) return 0;
    if ((r + 3) / 4 > 65535) return -1;
    hipLaunchKernelGGL(row_permute_kernel<T>, dim3(nb, (r + 3) / 4), dim3(256), 0, st, X, ldx, n, r,
                       (const int *)perm);
  }
  return (int)hipGetLastError();
}

extern "C" {

// Tile variants (complex128): 32 / 64 are the four-product 32 x 32 and 64 x 64
// tiles, 256 the four-product 128 x 128 tile on eight waves; the
// three-product ones are 1032 (32 x 32), 2032 (32 x 32, k step 32), 1033
// (64 x 32), 1034 (32 x 64), 1064 (64 x 64), 2064 (64 x 64, k step 32), 1065
// (64 x 64, one wave per SIMD), 2034 (32 x 64, k step 32), and on eight waves
// 1066 (64 x 64), 1128 (64 x 128), 1129 (128 x 64).  Default rule (measured on gfx950 at
// lda = 3840): 1032 when either output dimension is at most 128, else 1034;
// the context's autotuner may replace it per shape class.
static const int ZTILES[] = {32, 64, 256, 1032, 2032, 1033, 1034, 1064, 2064, 1065, 2034, 1066, 1128, 1129};

int bem_zgemm_default_tile(int m, int n) { return (m <= 128 || n <= 128) ? 1032 : 1034; }

int bem_zgemm_tile_valid(int tile) {
  for (int t : ZTILES)
    if (t == tile) return 1;
  return 0;
}

int bem_zgemm_tiled(int tile, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B,
                    int64_t ldb, double2 beta, double2 *C, int64_t ldc, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  if (!tile) tile = bem_zgemm_default_tile(m, n);
  switch (tile) {
#define ZG(code, BM, BN, WM, WN, KB, OCC, G3)                                                 \
  case code:                                                                                 \
    launch_zgemm<BM, BN, WM, WN, KB, OCC, G3>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); \
    break;
    ZG(32, 32, 32, 2, 2, 16, 1, false)
    ZG(1032, 32, 32, 2, 2, 16, 1, true)
    ZG(2032, 32, 32, 2, 2, 32, 1, true)
    ZG(1033, 64, 32, 2, 2, 16, 2, true)
    ZG(1034, 32, 64, 2, 2, 16, 2, true)
    ZG(1064, 64, 64, 2, 2, 16, 2, true)
    ZG(2064, 64, 64, 2, 2, 32, 2, true)
    ZG(1065, 64, 64, 2, 2, 16, 1, true)
    ZG(2034, 32, 64, 2, 2, 32, 2, true)
    ZG(1066, 64, 64, 2, 4, 16, 2, true)
    ZG(1128, 64, 128, 2, 4, 16, 2, true)
    ZG(1129, 128, 64, 4, 2, 16, 2, true)
    ZG(256, 128, 128, 2, 4, 16, 2, false)
    ZG(64, 64, 64, 2, 2, 16, 1, false)
#undef ZG
    default: return -1;
  }
  return (int)hipGetLastError();
}

int bem_zgemm_raw(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                  double2 beta, double2 *C, int64_t ldc, hipStream_t st) {
  return bem_zgemm_tiled(0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st);
}

// Hermitian trace (zgemm_htrace_kernel): out <- sum conj(A^H B) .* E for an
// n x n product with k-long columns, partial holding bem_htrace_partials(n).
size_t bem_htrace_partials(int n) { return (size_t)((n + 31) / 32) * (size_t)((n + 63) / 64); }

int bem_zgemm_htrace(int n, int k, const double2 *A, int64_t lda, const double2 *B, int64_t ldb, const double2 *E,
                     int64_t lde, double2 *partial, double2 *out, hipStream_t st) {
  dim3 grid((n + 31) / 32, (n + 63) / 64);
  hipLaunchKernelGGL((zgemm_htrace_kernel<32, 64, 2, 2, 16, 2, true>), grid, dim3(256), 0, st, n, k, A, lda, B, ldb,
                     E, lde, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, (int)bem_htrace_partials(n), out);
  return (int)hipGetLastError();
}

// Tournament pivot selection for the panel A(r0:r0+m, c0:c0+w), w <= 32.
// cand must hold ceil(m/512)*32 ints, cand2 ceil(ceil(m/512)/16)*32 (ping-pong);
// the winners end up in ipiv[k0..k0+w) as interchanges (k0 = r0).
int bem_calu_pivots(const double2 *A, int64_t lda, int m, int w, int r0, int c0, int *cand, int *cand2, int *ipiv,
                    int *info, hipStream_t st) {
  if (w <= 0) return 0;
  if (w > GEPP_NB) return -1;
  int nsets = (m + GEPP_ROWS - 1) / GEPP_ROWS;
  hipLaunchKernelGGL(tslu_leaf_kernel, dim3(nsets), dim3(GEPP_NT), 0, st, A, lda, m, w, r0, c0, cand);
  int *cin = cand, *cout = cand2;
  while (nsets > 1) {
    int nout = (nsets + GEPP_MERGE - 1) / GEPP_MERGE;
    hipLaunchKernelGGL(tslu_merge_kernel, dim3(nout), dim3(GEPP_NT), 0, st, A, lda, w, c0, cin, nsets, cout);
    int *t = cin; cin = cout; cout = t;
    nsets = nout;
  }
  hipLaunchKernelGGL(winners_to_ipiv_kernel, dim3(1), dim3(64), 0, st, cin, w, r0, ipiv, info);
  return (int)hipGetLastError();
}

// Diagonal blocks [blk0, blk0+nblk) (bw <= 64 wide) of an n x n matrix:
// factor=1 factors each block without pivoting and inverts its triangles;
// factor=0 only forms the inverses of an existing packed LU.  Inverses go to
// inv + blk*2*64*64 (64 x 64, identity-padded).
int bem_diag_block(double2 *A, int64_t lda, int n, int blk0, int nblk, int bw, int factor, double2 *inv, int *info,
                   hipStream_t st) {
  if (nblk <= 0) return 0;
  if (bw <= 0 || bw > 64) return -1;
  hipLaunchKernelGGL(diag_block_kernel, dim3(nblk), dim3(256), 0, st, A, lda, n, blk0, bw, factor, inv, info);
  return (int)hipGetLastError();
}

// Block-local pivoted LU of the diagonal block at (k0, k0) plus the inverses
// of its triangles (to inv, 2 x 64 x 64).
int bem_diag_lu_piv(double2 *A, int64_t lda, int n, int k0, double2 *inv, int *ipiv, int *info, hipStream_t st) {
  hipLaunchKernelGGL(diag_lu_piv_kernel<double2>, dim3(1), dim3(256), 0, st, A, lda, n, k0, inv, ipiv, info);
  return (int)hipGetLastError();
}

int bem_absmax(const double2 *A, int64_t lda, int m, int n, int upper, unsigned long long *out, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid((m + 255) / 256 < 8 ? (m + 255) / 256 : 8, n < 512 ? n : 512);
  hipLaunchKernelGGL(absmax_kernel<double2>, grid, dim3(256), 0, st, A, lda, m, n, upper, out);
  return (int)hipGetLastError();
}

int bem_col_swap(double2 *X, int64_t ldx, int m, int k0, int k1, const int *ipiv, hipStream_t st) {
  if (m <= 0 || k1 <= k0) return 0;
  hipLaunchKernelGGL(col_swap_kernel<double2>, dim3((m + 255) / 256), dim3(256), 0, st, X, ldx, m, k0, k1, ipiv);
  return (int)hipGetLastError();
}

int bem_laswp(double2 *A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1, const int *ipiv,
              hipStream_t st) {
  if (c1 <= c0 || k1 <= k0) return 0;
  hipLaunchKernelGGL(laswp_kernel<double2>, dim3((c1 - c0 + 127) / 128), dim3(128), 0, st, A, lda, c0, c1, skip0, skip1,
                     k0, k1, ipiv);
  return (int)hipGetLastError();
}

int bem_permute_rows_local(double2 *X, int64_t ldx, int n, int r, const int *ipiv, int *perm, hipStream_t st) {
  return permute_local(X, ldx, 0, n, r, 0, ipiv, perm, st);
}
int bem_permute_cols_local(double2 *X, int64_t ldx, int m, int n, const int *ipiv, int *perm, hipStream_t st) {
  return permute_local(X, ldx, m, n, 0, 1, ipiv, perm, st);
}
int bem_permute_rows_local_c(float2 *X, int64_t ldx, int n, int r, const int *ipiv, int *perm, hipStream_t st) {
  return permute_local(X, ldx, 0, n, r, 0, ipiv, perm, st);
}
int bem_permute_cols_local_c(float2 *X, int64_t ldx, int m, int n, const int *ipiv, int *perm, hipStream_t st) {
  return permute_local(X, ldx, m, n, 0, 1, ipiv, perm, st);
}

int bem_symm(double2 *G, int64_t ld, int n, hipStream_t st) {
  if (n <= 0) return 0;
  const int nt = (n + 31) / 32;
  if (nt > 65535) return -1;
  hipLaunchKernelGGL(symm_kernel, dim3(nt, nt), dim3(256), 0, st, G, ld, n);
  return (int)hipGetLastError();
}


// ------------------------------------------------------- complex64 (mixed)
// C <- alpha A B + beta C in complex64.  Tile variants: 1 = 64 x 64, 2 = 32 x
// 32, 3 = 64 x 128, 4 = 128 x 64, 5 = 128 x 128 (2 x 2 waves each) with four
// real products per complex multiply-add; 11..15 the same tiles with three;
// 16 = 64 x 64 and 17 = 128 x 128 on eight waves (2 x 4), three products;
// 21 = 64 x 64, 23 = 64 x 128 and 26 = 64 x 64 on eight waves with k steps of
// 32 (three products): half the barriers per flop.
// Default rule: 32 x 32 when either dimension is at most 128, else 64 x 64.
int bem_cgemm_default_tile(int m, int n) { return (m <= 128 || n <= 128) ? 2 : 1; }
int bem_cgemm_tile_valid(int tile) {
  return (tile >= 1 && tile <= 5) || (tile >= 11 && tile <= 17) || tile == 21 || tile == 23 || tile == 26;
}
int bem_cgemm_wide(int m, int n) { return !(m <= 128 || n <= 128); }

int bem_cgemm_tiled(int tile, int m, int n, int k, float2 alpha, const float2 *A, int64_t lda, const float2 *B,
                    int64_t ldb, float2 beta, float2 *C, int64_t ldc, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  if (!tile) tile = bem_cgemm_default_tile(m, n);
  switch (tile) {
    case 1: launch_cgemm<64, 64, 2, 2, 16>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 2: launch_cgemm<32, 32, 2, 2, 16>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 3: launch_cgemm<64, 128, 2, 2, 16>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 4: launch_cgemm<128, 64, 2, 2, 16>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 5: launch_cgemm<128, 128, 2, 2, 16>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 11: launch_cgemm<64, 64, 2, 2, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 12: launch_cgemm<32, 32, 2, 2, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 13: launch_cgemm<64, 128, 2, 2, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 14: launch_cgemm<128, 64, 2, 2, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 15: launch_cgemm<128, 128, 2, 2, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 16: launch_cgemm<64, 64, 2, 4, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 17: launch_cgemm<128, 128, 2, 4, 16, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 21: launch_cgemm<64, 64, 2, 2, 32, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 23: launch_cgemm<64, 128, 2, 2, 32, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    case 26: launch_cgemm<64, 64, 2, 4, 32, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st); break;
    default: return -1;
  }
  return (int)hipGetLastError();
}

int bem_cgemm_raw(int m, int n, int k, float2 alpha, const float2 *A, int64_t lda, const float2 *B, int64_t ldb,
                  float2 beta, float2 *C, int64_t ldc, hipStream_t st) {
  return bem_cgemm_tiled(0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, st);
}

size_t bem_chtrace_partials(int n) { return (size_t)((n + 63) / 64) * (size_t)((n + 63) / 64); }

// out (complex128) <- sum conj(A^H B) .* E over an n x n Hermitian product in
// complex64 (cgemm_htrace_kernel), the partial sums and the final sum in FP64.
int bem_cgemm_htrace(int n, int k, const float2 *A, int64_t lda, const float2 *B, int64_t ldb, const float2 *E,
                     int64_t lde, double2 *partial, double2 *out, hipStream_t st) {
  dim3 grid((n + 63) / 64, (n + 63) / 64);
  hipLaunchKernelGGL((cgemm_htrace_kernel<64, 64, 2, 2, 16, false>), grid, dim3(256), 0, st, n, k, A, lda, B, ldb, E,
                     lde, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, partial, (int)bem_chtrace_partials(n), out);
  return (int)hipGetLastError();
}

int bem_diag_lu_piv_c(float2 *A, int64_t lda, int n, int k0, float2 *inv, int *ipiv, int *info, hipStream_t st) {
  hipLaunchKernelGGL(diag_lu_piv_kernel<float2>, dim3(1), dim3(256), 0, st, A, lda, n, k0, inv, ipiv, info);
  return (int)hipGetLastError();
}

int bem_absmax_c(const float2 *A, int64_t lda, int m, int n, int upper, unsigned long long *out, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  dim3 grid((m + 255) / 256 < 8 ? (m + 255) / 256 : 8, n < 512 ? n : 512);
  hipLaunchKernelGGL(absmax_kernel<float2>, grid, dim3(256), 0, st, A, lda, m, n, upper, out);
  return (int)hipGetLastError();
}

int bem_col_swap_c(float2 *X, int64_t ldx, int m, int k0, int k1, const int *ipiv, hipStream_t st) {
  if (m <= 0 || k1 <= k0) return 0;
  hipLaunchKernelGGL(col_swap_kernel<float2>, dim3((m + 255) / 256), dim3(256), 0, st, X, ldx, m, k0, k1, ipiv);
  return (int)hipGetLastError();
}

int bem_laswp_c(float2 *A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1, const int *ipiv,
                hipStream_t st) {
  if (c1 <= c0 || k1 <= k0) return 0;
  hipLaunchKernelGGL(laswp_kernel<float2>, dim3((c1 - c0 + 127) / 128), dim3(128), 0, st, A, lda, c0, c1, skip0, skip1,
                     k0, k1, ipiv);
  return (int)hipGetLastError();
}

static dim3 copy_grid(int m, int n) { return dim3((m + 255) / 256, n < 65535 ? n : 65535); }

int bem_z2c(const double2 *src, int64_t lds, float2 *dst, int64_t ldd, int m, int n, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(z2c_kernel, copy_grid(m, n), dim3(256), 0, st, src, lds, dst, ldd, m, n);
  return (int)hipGetLastError();
}

int bem_c2z(const float2 *src, int64_t lds, double2 *dst, int64_t ldd, int m, int n, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(c2z_kernel, copy_grid(m, n), dim3(256), 0, st, src, lds, dst, ldd, m, n);
  return (int)hipGetLastError();
}

int bem_c2z_add(const float2 *src, int64_t lds, double2 *dst, int64_t ldd, int m, int n, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(c2z_add_kernel, copy_grid(m, n), dim3(256), 0, st, src, lds, dst, ldd, m, n);
  return (int)hipGetLastError();
}

int bem_add_z_to_c(const double2 *src, int64_t lds, float2 *dst, int64_t ldd, int m, int n, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(add_z_to_c_kernel, copy_grid(m, n), dim3(256), 0, st, src, lds, dst, ldd, m, n);
  return (int)hipGetLastError();
}

// bytes of p filled with hashed doubles (as_float = 0) or floats (1) in [-1, 1)
int bem_fill_hash(void *p, size_t bytes, int as_float, unsigned long long seed, hipStream_t st) {
  if (bytes == 0) return 0;
  if (as_float)
    hipLaunchKernelGGL(fill_hash_kernel<float>, dim3(2048), dim3(256), 0, st, (float *)p, bytes / sizeof(float), seed);
  else
    hipLaunchKernelGGL(fill_hash_kernel<double>, dim3(2048), dim3(256), 0, st, (double *)p, bytes / sizeof(double),
                       seed);
  return (int)hipGetLastError();
}

int bem_signed_transpose(const double2 *B, int64_t ldb, int m, int n, double2 *C, int64_t ldc, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  if ((n + 31) / 32 > 65535) return -1;
  hipLaunchKernelGGL(signed_transpose_kernel, dim3((m + 31) / 32, (n + 31) / 32), dim3(256), 0, st, B, ldb, m, n, C,
                     ldc);
  return (int)hipGetLastError();
}

// Partials bem_sumsq needs for an m x n block (its fixed launch grid).
static dim3 sumsq_grid(int m, int n) {
  const int gx = (m + 255) / 256 < 8 ? (m + 255) / 256 : 8;
  return dim3(gx > 0 ? gx : 1, n < 256 ? (n > 0 ? n : 1) : 256);
}
size_t bem_sumsq_partials(int m, int n) {
  dim3 g = sumsq_grid(m, n);
  return (size_t)g.x * g.y;
}

// *out <- (add ? *out : 0) + ||A||_F^2 of an m x n complex128 block
// (partial: bem_sumsq_partials(m, n) doubles); deterministic.
int bem_sumsq(const double2 *A, int64_t lda, int m, int n, double *partial, int add, double *out, hipStream_t st) {
  dim3 g = sumsq_grid(m, n);
  if (m > 0 && n > 0) {
    hipLaunchKernelGGL(sumsq_kernel, g, dim3(256), 0, st, A, lda, m, n, partial);
    hipLaunchKernelGGL(sumsq_finish_kernel, dim3(1), dim3(256), 0, st, (const double *)partial, (int)(g.x * g.y), add,
                       out);
  } else {
    hipLaunchKernelGGL(sumsq_finish_kernel, dim3(1), dim3(256), 0, st, (const double *)partial, 0, add, out);
  }
  return (int)hipGetLastError();
}

int bem_trsm_few(const double2 *A, int64_t lda, int n, const double2 *inv, double2 *X, int64_t ldx, int r, int lower,
                 double2 *Y, int64_t ldy, hipStream_t st) {
  return trsm_few(A, lda, n, inv, X, ldx, r, lower, Y, ldy, st);
}
int bem_trsm_few_c(const float2 *A, int64_t lda, int n, const float2 *inv, float2 *X, int64_t ldx, int r, int lower,
                   float2 *Y, int64_t ldy, hipStream_t st) {
  return trsm_few(A, lda, n, inv, X, ldx, r, lower, Y, ldy, st);
}

// Partials zgemm_skinny needs for C (m x r) with inner dimension K.
static int skinny_split(int m, int K) {
  const int mb = (m + 63) / 64, kc = (K + SKINNY_KCHUNK - 1) / SKINNY_KCHUNK;
  int ns = (1024 + mb - 1) / mb;
  return ns < 1 ? 1 : (ns > kc ? (kc > 0 ? kc : 1) : ns);
}
size_t bem_zgemm_skinny_partials(int m, int r, int K) { return (size_t)skinny_split(m, K) * r * (size_t)m; }

// C (m x r, r <= 8) <- alpha A B + beta C (A m x K, B K x r) in complex128
// with part holding bem_zgemm_skinny_partials(m, r, K) values.
int bem_zgemm_skinny(int m, int r, int K, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                     double2 beta, double2 *C, int64_t ldc, double2 *part, hipStream_t st) {
  if (m <= 0 || r <= 0) return 0;
  if (r > FEW_MAX_COLS) return -1;
  const int ns = skinny_split(m, K);
  const int kc = (K + SKINNY_KCHUNK - 1) / SKINNY_KCHUNK;
  const int kper = ((kc + ns - 1) / ns) * SKINNY_KCHUNK;
  const int nsplit = K > 0 ? (K + kper - 1) / kper : 1;
  if (K > 0)
    hipLaunchKernelGGL(zgemm_skinny_kernel, dim3((m + 63) / 64, nsplit), dim3(256), 0, st, m, K, r, A, lda, B, ldb,
                       kper, part);
  hipLaunchKernelGGL(zgemm_skinny_finish_kernel, dim3((m + 255) / 256, r), dim3(256), 0, st, m, r, K > 0 ? nsplit : 0,
                     (const double2 *)part, alpha, beta, C, ldc);
  return (int)hipGetLastError();
}

int bem_zdotc_cols(const double2 *G, int64_t ldg, const double2 *D, int64_t ldd, int m, int n, double2 *out,
                   hipStream_t st) {
  if (n <= 0) return 0;
  if (n > 65535) return -1;
  hipLaunchKernelGGL(zdotc_cols_kernel, dim3(n), dim3(256), 0, st, G, ldg, D, ldd, m, out);
  return (int)hipGetLastError();
}

int bem_zaxpy(double2 alpha, const double2 *src, int64_t lds, double2 *dst, int64_t ldd, int m, int n, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  hipLaunchKernelGGL(zaxpy_kernel, copy_grid(m, n), dim3(256), 0, st, alpha, src, lds, dst, ldd, m, n);
  return (int)hipGetLastError();
}

// bytes (a multiple of 4) of p <- 0 by a kernel.
int bem_fill_zero(void *p, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 4;
  if (n == 0) return 0;
  const unsigned int blocks = (unsigned int)std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, st, (unsigned int *)p, n);
  return (int)hipGetLastError();
}

int bem_identity_minus(double2 *Cm, int64_t ldc, int n, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(identity_minus_kernel, dim3((n + 255) / 256, n), dim3(256), 0, st, Cm, ldc, n);
  return (int)hipGetLastError();
}

// Tile of bem_zgemm_ta: 64 x 64 on eight waves when both output dimensions
// reach 64 (the projections onto a range basis, rank x 64), else 32 x 32.
static int ta_tile(int m, int n) { return (m >= 64 && n >= 64) ? 64 : 32; }

// K splits of bem_zgemm_ta for an m x n output over k: enough workgroups to
// fill the device (about 512), each split at least 64 long.
static int ta_splits(int m, int n, int k) {
  const int t = ta_tile(m, n);
  const int blocks = ((m + t - 1) / t) * ((n + t - 1) / t);
  int nz = (512 + blocks - 1) / blocks;
  const int kmax = (k + 63) / 64;
  if (nz > kmax) nz = kmax;
  return nz < 1 ? 1 : nz;
}
size_t bem_zgemm_ta_partials(int m, int n, int k) {
  const int nz = ta_splits(m, n, k);
  return nz > 1 ? (size_t)nz * m * n : 0;
}

// C (m x n) <- alpha A^H B + beta C, A k x m and B k x n (column-major),
// part holding bem_zgemm_ta_partials(m, n, k) values.
int bem_zgemm_ta(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                 double2 beta, double2 *C, int64_t ldc, double2 *part, hipStream_t st) {
  if (m <= 0 || n <= 0) return 0;
  if (k <= 0) return -1;
  int nz = ta_splits(m, n, k);
  int kper = (k + nz - 1) / nz;
  kper = (kper + 15) / 16 * 16;
  nz = (k + kper - 1) / kper;
  const int t = ta_tile(m, n);
  if ((n + t - 1) / t > 65535 || nz > 65535) return -1;
  dim3 grid((m + t - 1) / t, (n + t - 1) / t, nz);
  if (t == 64)
    hipLaunchKernelGGL((zgemm_ta_kernel<64, 64, 2, 4, 16>), grid, dim3(512), 0, st, m, n, k, kper, A, lda, B, ldb,
                       alpha, beta, C, ldc, part);
  else
    hipLaunchKernelGGL((zgemm_ta_kernel<32, 32, 2, 2, 16>), grid, dim3(256), 0, st, m, n, k, kper, A, lda, B, ldb,
                       alpha, beta, C, ldc, part);
  if (nz > 1)
    hipLaunchKernelGGL(zgemm_ta_finish_kernel, dim3((m + 255) / 256, n), dim3(256), 0, st, m, n, nz,
                       (const double2 *)part, alpha, beta, C, ldc);
  return (int)hipGetLastError();
}

}  // extern "C"

