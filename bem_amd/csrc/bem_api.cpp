// libbem_hip host runtime: contexts, bodies, assembly orchestration and the
// PMCHWT solve behind the C ABI of include/bem_hip.h.
//
// CalcHeatTransfer (src/HeatTransfer.cpp:8-59) inverts W = Z_ext + blkdiag(G1,G2)
// and forms S = 1/4 tr(B2^H B1) with B1 = sym(G1) X, B2 = X sym(G2),
// X = W^-1[0:m1, m1:M].  Only X is needed, so the solve here is the block
// elimination of W = [[A, B], [C, D]]:
//   Y = A^-1 B,  Sc = D - C Y,  X = -Y Sc^-1
// with tournament-pivoted LU of A and of Sc, followed by B1 = sym(G1) X and a
// GEMM whose epilogue contracts X sym(G2) against B1 (no B2 stored).
//
// Concurrency.  A context owns `lanes` HIP streams, each with its own named
// workspaces.  A sweep keeps up to `lanes` independent solves in flight: the
// latency-bound LU panels of one solve overlap the GEMMs of another.  Every
// solve's kernels stay on its own stream, so results do not depend on what
// else is in flight.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <list>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bem_hip.h"
#include "bem_common.h"
#include "geometry.h"

extern "C" {
int bem_launch_far_g6(const TriGeom *tg, int T, const Medium *md, int row_lo, int row_hi, int col_lo, int col_hi,
                      double *tiles, size_t npairs, int2 *q36, unsigned int *q36_count, unsigned int q36_cap,
                      hipStream_t st);
int bem_launch_far_g36(const TriGeom *tg, int T, const Medium *md, const int2 *pairs, const unsigned int *count_dev,
                       unsigned int max_count, int row_lo, int row_hi, double *tiles, size_t npairs, hipStream_t st);
int bem_launch_close(const TriGeom *tg, int T, const Medium *md, const int2 *pairs, unsigned int count, int row_lo,
                     int row_hi, double *tiles, size_t npairs, hipStream_t st);
int bem_launch_gather(const double *tiles, size_t npairs, int T, int row_lo, const RwgRec *rw, int r_lo, int r_hi,
                      int c_lo, int c_hi, const Medium *md, double2 *Z, int64_t ldz, int zr0, int zc0, int add,
                      double2 *Zcopy, int64_t ldc, int cr0, int cc0, hipStream_t st);
int bem_zgemm_raw(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                  double2 beta, double2 *C, int64_t ldc, hipStream_t st);
int bem_zgemm_trace(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                    const double2 *E, int64_t lde, double2 *partial, double2 *out, hipStream_t st);
size_t bem_trace_partials(int m, int n);
int bem_calu_pivots(const double2 *A, int64_t lda, int m, int w, int r0, int c0, int *cand, int *cand2, int *ipiv,
                    int *info, hipStream_t st);
int bem_laswp(double2 *A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1, const int *ipiv,
              hipStream_t st);
int bem_symm(double2 *G, int64_t ld, int n, hipStream_t st);
int bem_diag_block(double2 *A, int64_t lda, int n, int blk0, int nblk, int bw, int factor, double2 *inv, int *info,
                   hipStream_t st);
int bem_diag_lu_piv(double2 *A, int64_t lda, int n, int k0, double2 *inv, int *ipiv, int *info, hipStream_t st);
int bem_absmax(const double2 *A, int64_t lda, int m, int n, int upper, unsigned long long *out, hipStream_t st);
int bem_col_swap(double2 *X, int64_t ldx, int m, int k0, int k1, const int *ipiv, hipStream_t st);
}

// src/util/SI.cpp:3-9
static const double SI_HBAR = 1.05457148e-34;
static const double SI_KB = 1.3806503e-23;
static const double SI_EV = 1.60217646e-19;
static const double SI_C = 2.99792458e8;

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
};

struct EventRec {
  hipEvent_t a, b;
  double flops;
};

struct DevTree {
  int T = 0, E = 0;
  TriGeom *tg = nullptr;
  RwgRec *rw = nullptr;
  int2 *close = nullptr;
  unsigned int nclose = 0;
};

// Device geometry of one (body 1, body 2, gap): the two-body tree and the
// two one-body trees, uploaded once and reused by every frequency.
struct PairGeom {
  const void *b1 = nullptr, *b2 = nullptr;
  double gap[3] = {0, 0, 0};
  DevTree top, s1, s2;
  std::vector<void *> allocs;
  size_t bytes = 0;
};

// One stream and its workspaces; holds at most one solve in flight.
struct Lane {
  int id = 0;
  hipStream_t st = nullptr;
  std::string tag;
  std::vector<EventRec> ev;
  size_t ev_used = 0;
  hipEvent_t ph[4] = {nullptr, nullptr, nullptr, nullptr};
  double2 *h_trace = nullptr;   // pinned host slots of the results
  int *h_info = nullptr;
  unsigned long long *h_growth = nullptr;   // max|U|, max|A| of LU(A), then of LU(S)
  bool busy = false;
  bem_result *res = nullptr;
  bem_request req;
  std::chrono::steady_clock::time_point t0;
};

}  // namespace

struct bem_ctx {
  int device = 0;
  size_t budget = 0;
  size_t bytes = 0;
  std::string err;
  std::recursive_mutex mu;
  std::unordered_map<std::string, DevBuf> ws;   // named, grow-only workspaces
  std::unordered_map<void *, size_t> user_allocs;
  std::vector<Lane> lanes;
  std::list<PairGeom> pairs;                    // most recently used first
  bool profile_gemm = true;
  // statistics of the last API call
  double t_last[9] = {0};
  double gemm_ms = 0, gemm_n = 0, gemm_flops = 0;
};

struct bem_body {
  bem_ctx *ctx;
  bemgeo::BodyMesh mesh;
};

static int fail(bem_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define DEV(c, expr)                                                                           \
  do {                                                                                         \
    hipError_t _e = (hipError_t)(expr);                                                        \
    if (_e != hipSuccess)                                                                      \
      return fail(c, BEM_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

#define CHK(expr)          \
  do {                     \
    int _r = (expr);       \
    if (_r != 0) return _r; \
  } while (0)

// Named workspace of a lane (grow-only).  Growing frees the old buffer after
// the lane's stream has drained.
static int ws_get(bem_ctx *c, Lane &L, const char *name, size_t bytes, void **out) {
  DevBuf &b = c->ws[L.tag + name];
  if (b.cap < bytes) {
    if (b.p) {
      DEV(c, hipStreamSynchronize(L.st));
      DEV(c, hipFree(b.p));
      c->bytes -= b.cap;
      b.p = nullptr;
      b.cap = 0;
    }
    if (c->budget && c->bytes + bytes > c->budget)
      return fail(c, BEM_E_NOMEM, "workspace '%s%s' (%zu bytes) exceeds the HBM budget", L.tag.c_str(), name, bytes);
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess)
      return fail(c, BEM_E_NOMEM, "hipMalloc(%zu) for '%s%s': %s", bytes, L.tag.c_str(), name, hipGetErrorString(e));
    b.cap = bytes;
    c->bytes += bytes;
  }
  *out = b.p;
  return 0;
}

static int ev_begin(bem_ctx *c, Lane &L, double flops, EventRec **rec) {
  if (L.ev_used == L.ev.size()) {
    EventRec r;
    DEV(c, hipEventCreate(&r.a));
    DEV(c, hipEventCreate(&r.b));
    L.ev.push_back(r);
  }
  EventRec &r = L.ev[L.ev_used++];
  r.flops = flops;
  DEV(c, hipEventRecord(r.a, L.st));
  *rec = &r;
  return 0;
}

// Accumulate the lane's GEMM event times into the context statistics
// (the lane's stream must have drained).
static void ev_collect(bem_ctx *c, Lane &L) {
  for (size_t i = 0; i < L.ev_used; i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, L.ev[i].a, L.ev[i].b) == hipSuccess) {
      c->gemm_ms += ms;
      c->gemm_n += 1;
      c->gemm_flops += L.ev[i].flops;
    }
  }
  L.ev_used = 0;
}

static void stats_reset(bem_ctx *c) { c->gemm_ms = c->gemm_n = c->gemm_flops = 0; }

// GEMM on the lane's stream (bracketed by events when profiling).
static int gemm(bem_ctx *c, Lane &L, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda,
                const double2 *B, int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  EventRec *r = nullptr;
  if (c->profile_gemm) CHK(ev_begin(c, L, 8.0 * m * (double)n * k, &r));
  int e = bem_zgemm_raw(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "zgemm launch: %s", hipGetErrorString((hipError_t)e));
  if (r) DEV(c, hipEventRecord(r->b, L.st));
  return 0;
}

// --------------------------------------------------------------- LU / solves
// Panel width of the blocked LU; the diagonal-block inverses are 64 x 64.
static const int LU_NB = 64;

static size_t inv_bytes(int n) { return sizeof(double2) * 2 * 64 * 64 * (size_t)((n + LU_NB - 1) / LU_NB); }

// dst block (rows x cols, ld) <- tmp (rows x cols, ld rows)
static int copy_back(bem_ctx *c, Lane &L, double2 *dst, int64_t ld, const double2 *tmp, int rows, int cols) {
  DEV(c, hipMemcpy2DAsync(dst, sizeof(double2) * ld, tmp, sizeof(double2) * rows, sizeof(double2) * rows, cols,
                          hipMemcpyDeviceToDevice, L.st));
  return 0;
}

// Blocked right-looking LU with tournament-pivoted panels.  Each 64-column
// block is factored as two 32-column sub-panels (tournament pivots, diagonal
// 32 x 32 factor + inverses, L21 by GEMM, then the second sub-panel updated);
// inv receives the inverses of the 64 x 64 unit-lower and upper diagonal
// blocks (inv_bytes(n)), and every triangular solve runs as a GEMM.
static const int CALU_W = 32;

static int lu_factor(bem_ctx *c, Lane &L, double2 *A, int64_t lda, int n, int *ipiv, int *info, double2 *inv) {
  const int nb = LU_NB, sb = CALU_W;
  int sets = (n + 511) / 512;
  int *cand, *cand2;
  double2 *tmp, *inv32;
  CHK(ws_get(c, L, "calu_cand", sizeof(int) * sb * (size_t)sets, (void **)&cand));
  CHK(ws_get(c, L, "calu_cand2", sizeof(int) * sb * (size_t)((sets + 15) / 16 + 1), (void **)&cand2));
  CHK(ws_get(c, L, "lu_tmp", sizeof(double2) * 64 * (size_t)std::max(n, 1), (void **)&tmp));
  CHK(ws_get(c, L, "inv32", sizeof(double2) * 2 * 64 * 64 * (size_t)((n + sb - 1) / sb), (void **)&inv32));
  const double2 m1 = make_double2(-1.0, 0.0), one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    for (int s0 = k0; s0 < k0 + jb; s0 += sb) {
      int sw = std::min(sb, k0 + jb - s0);
      int e = bem_calu_pivots(A, lda, n - s0, sw, s0, s0, cand, cand2, ipiv, info, L.st);
      if (e) return fail(c, BEM_E_DEVICE, "calu pivots: %d", e);
      e = bem_laswp(A, lda, 0, n, 0, 0, s0, s0 + sw, ipiv, L.st);
      if (e) return fail(c, BEM_E_DEVICE, "laswp: %d", e);
      const int b32 = s0 / sb;
      e = bem_diag_block(A, lda, n, b32, 1, sb, 1, inv32, info, L.st);
      if (e) return fail(c, BEM_E_DEVICE, "diag block: %d", e);
      const double2 *Li = inv32 + (size_t)b32 * 2 * 64 * 64, *Ui = Li + 64 * 64;
      int rest = n - s0 - sw;
      double2 *Ass = A + (int64_t)s0 * lda + s0;
      if (rest > 0) {
        CHK(gemm(c, L, rest, sw, sw, one, Ass + sw, lda, Ui, 64, zero, tmp, rest));   // L21 = A21 U11^-1
        CHK(copy_back(c, L, Ass + sw, lda, tmp, rest, sw));
      }
      int cols2 = k0 + jb - (s0 + sw);   // columns of this 64-block right of the sub-panel
      if (cols2 > 0) {
        double2 *A12 = A + (int64_t)(s0 + sw) * lda + s0;
        CHK(gemm(c, L, sw, cols2, sw, one, Li, 64, A12, lda, zero, tmp, sw));      // U12 = L11^-1 A12
        CHK(copy_back(c, L, A12, lda, tmp, sw, cols2));
        if (rest > 0) CHK(gemm(c, L, rest, cols2, sw, m1, Ass + sw, lda, A12, lda, one, A12 + sw, lda));
      }
    }
    const int blk = k0 / nb;
    int e = bem_diag_block(A, lda, n, blk, 1, nb, 0, inv, nullptr, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "diag block inverses: %d", e);
    const double2 *Linv = inv + (size_t)blk * 2 * 64 * 64;
    int rest = n - k0 - jb;
    if (rest > 0) {
      double2 *Akk = A + (int64_t)k0 * lda + k0;
      double2 *A21 = Akk + jb, *A12 = A + (int64_t)(k0 + jb) * lda + k0;
      CHK(gemm(c, L, jb, rest, jb, one, Linv, 64, A12, lda, zero, tmp, jb));     // U12 = L11^-1 A12
      CHK(copy_back(c, L, A12, lda, tmp, jb, rest));
      CHK(gemm(c, L, rest, rest, jb, m1, A21, lda, A12, lda, one, A12 + jb, lda));
    }
  }
  return 0;
}

// Blocked right-looking LU with partial pivoting inside each 64 x 64
// diagonal block (one kernel factors the block in LDS and inverts its
// triangles); the row interchanges are then applied to the other columns and
// the panel's triangular solves and the trailing update run as GEMMs.
// growth (optional, device) receives max|A| before in growth[1] and
// max|U| after in growth[0].
static int lu_factor_block(bem_ctx *c, Lane &L, double2 *A, int64_t lda, int n, int *ipiv, int *info, double2 *inv,
                           unsigned long long *growth) {
  const int nb = LU_NB;
  double2 *tmp;
  CHK(ws_get(c, L, "lu_tmp", sizeof(double2) * 64 * (size_t)std::max(n, 1), (void **)&tmp));
  const double2 m1 = make_double2(-1.0, 0.0), one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  int e;
  if (growth && (e = bem_absmax(A, lda, n, n, 0, growth + 1, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    const int blk = k0 / nb;
    double2 *Linv = inv + (size_t)blk * 2 * 64 * 64, *Uinv = Linv + 64 * 64;
    e = bem_diag_lu_piv(A, lda, n, k0, Linv, ipiv, info, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "diag block LU: %d", e);
    e = bem_laswp(A, lda, 0, n, k0, k0 + jb, k0, k0 + jb, ipiv, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "laswp: %d", e);
    int rest = n - k0 - jb;
    if (rest > 0) {
      double2 *Akk = A + (int64_t)k0 * lda + k0;
      double2 *A21 = Akk + jb, *A12 = A + (int64_t)(k0 + jb) * lda + k0;
      CHK(gemm(c, L, rest, jb, jb, one, A21, lda, Uinv, 64, zero, tmp, rest));   // L21 = A21 U11^-1
      CHK(copy_back(c, L, A21, lda, tmp, rest, jb));
      CHK(gemm(c, L, jb, rest, jb, one, Linv, 64, A12, lda, zero, tmp, jb));     // U12 = L11^-1 A12
      CHK(copy_back(c, L, A12, lda, tmp, jb, rest));
      CHK(gemm(c, L, rest, rest, jb, m1, A21, lda, A12, lda, one, A12 + jb, lda));
    }
  }
  if (growth && (e = bem_absmax(A, lda, n, n, 1, growth, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  return 0;
}

// X (n x r) <- A^-1 X from the packed LU, its interchanges and the diagonal
// block inverses.
static int lu_solve(bem_ctx *c, Lane &L, const double2 *A, int64_t lda, int n, const int *ipiv, const double2 *inv,
                    double2 *X, int64_t ldx, int r) {
  const int nb = LU_NB;
  const double2 m1 = make_double2(-1.0, 0.0), one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  double2 *tmp;
  CHK(ws_get(c, L, "lu_tmp", sizeof(double2) * 64 * (size_t)std::max(std::max(n, r), 1), (void **)&tmp));
  int e = bem_laswp(X, ldx, 0, r, 0, 0, 0, n, ipiv, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "laswp rhs: %d", e);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    const double2 *Linv = inv + (size_t)(k0 / nb) * 2 * 64 * 64;
    CHK(gemm(c, L, jb, r, jb, one, Linv, 64, X + k0, ldx, zero, tmp, jb));
    CHK(copy_back(c, L, X + k0, ldx, tmp, jb, r));
    int rest = n - k0 - jb;
    if (rest > 0)
      CHK(gemm(c, L, rest, r, jb, m1, A + (int64_t)k0 * lda + k0 + jb, lda, X + k0, ldx, one, X + k0 + jb, ldx));
  }
  int last = ((n - 1) / nb) * nb;
  for (int k0 = last; k0 >= 0; k0 -= nb) {
    int jb = std::min(nb, n - k0);
    const double2 *Uinv = inv + (size_t)(k0 / nb) * 2 * 64 * 64 + 64 * 64;
    CHK(gemm(c, L, jb, r, jb, one, Uinv, 64, X + k0, ldx, zero, tmp, jb));
    CHK(copy_back(c, L, X + k0, ldx, tmp, jb, r));
    if (k0 > 0) CHK(gemm(c, L, k0, r, jb, m1, A + (int64_t)k0 * lda, lda, X + k0, ldx, one, X, ldx));
  }
  return 0;
}

// X (r x n) <- X A^-1 from the packed LU of A (A = P L U): X U^-1 over column
// blocks left to right, then L^-1 right to left, then the column
// interchanges X P^T; the diagonal blocks go through their inverses, the
// coupling terms through GEMMs.
static int lu_solve_right(bem_ctx *c, Lane &L, const double2 *A, int64_t lda, int n, const int *ipiv,
                          const double2 *inv, double2 *X, int64_t ldx, int r) {
  const int nb = LU_NB;
  const double2 m1 = make_double2(-1.0, 0.0), one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  double2 *tmp;
  CHK(ws_get(c, L, "lu_tmp", sizeof(double2) * 64 * (size_t)std::max(std::max(n, r), 1), (void **)&tmp));
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    const double2 *Uinv = inv + (size_t)(k0 / nb) * 2 * 64 * 64 + 64 * 64;
    double2 *Xk = X + (int64_t)k0 * ldx;
    CHK(gemm(c, L, r, jb, jb, one, Xk, ldx, Uinv, 64, zero, tmp, r));
    CHK(copy_back(c, L, Xk, ldx, tmp, r, jb));
    int rest = n - k0 - jb;
    if (rest > 0)
      CHK(gemm(c, L, r, rest, jb, m1, Xk, ldx, A + (int64_t)(k0 + jb) * lda + k0, lda, one, Xk + (int64_t)jb * ldx, ldx));
  }
  int last = ((n - 1) / nb) * nb;
  for (int k0 = last; k0 >= 0; k0 -= nb) {
    int jb = std::min(nb, n - k0);
    const double2 *Linv = inv + (size_t)(k0 / nb) * 2 * 64 * 64;
    double2 *Xk = X + (int64_t)k0 * ldx;
    CHK(gemm(c, L, r, jb, jb, one, Xk, ldx, Linv, 64, zero, tmp, r));
    CHK(copy_back(c, L, Xk, ldx, tmp, r, jb));
    if (k0 > 0) CHK(gemm(c, L, r, k0, jb, m1, Xk, ldx, A + k0, lda, one, X, ldx));
  }
  int e = bem_col_swap(X, ldx, r, 0, n, ipiv, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "column interchanges: %d", e);
  return 0;
}

// ------------------------------------------------------------------ assembly
static Medium make_medium(double er, double ei, double mr, double mi, double omega) {
  // src/GreensFunction.cpp:17-30 (complex arithmetic written out)
  Medium md;
  double emr = er * mr - ei * mi, emi = er * mi + ei * mr;     // eps mu
  double kr = emr * omega * omega, ki = emi * omega * omega;   // (eps mu omega) omega
  md.ksq_re = kr;
  md.ksq_im = ki;
  // principal square root
  double mod = std::hypot(kr, ki);
  double sr = std::sqrt(0.5 * (mod + kr));
  double si = std::sqrt(0.5 * (mod - kr));
  if (ki < 0) si = -si;
  md.ik_re = -si;
  md.ik_im = sr;
  md.k2_re = -0.5 * kr;
  md.k2_im = -0.5 * ki;
  // 1/(i omega eps) = 1/(-omega ei + i omega er)
  double ar = -omega * ei, ai = omega * er, d = ar * ar + ai * ai;
  md.Zeps_re = ar / d;
  md.Zeps_im = -ai / d;
  ar = -omega * mi; ai = omega * mr; d = ar * ar + ai * ai;
  md.Zmu_re = ar / d;
  md.Zmu_im = -ai / d;
  return md;
}

// Persistent upload of a tree (owned by a PairGeom, or by a one-off holder).
static int upload_tree(bem_ctx *c, const bemgeo::Tree &t, DevTree &d, std::vector<void *> &allocs, size_t &bytes) {
  std::vector<int2> cl(t.close.size());
  for (size_t i = 0; i < cl.size(); i++) cl[i] = make_int2(t.close[i].first, t.close[i].second);
  size_t b1 = sizeof(TriGeom) * t.tg.size(), b2 = sizeof(RwgRec) * t.rw.size(), b3 = sizeof(int2) * (cl.size() + 1);
  void *p1, *p2, *p3;
  DEV(c, hipMalloc(&p1, b1));
  allocs.push_back(p1);
  DEV(c, hipMalloc(&p2, b2));
  allocs.push_back(p2);
  DEV(c, hipMalloc(&p3, b3));
  allocs.push_back(p3);
  bytes += b1 + b2 + b3;
  c->bytes += b1 + b2 + b3;
  DEV(c, hipMemcpy(p1, t.tg.data(), b1, hipMemcpyHostToDevice));
  DEV(c, hipMemcpy(p2, t.rw.data(), b2, hipMemcpyHostToDevice));
  if (!cl.empty()) DEV(c, hipMemcpy(p3, cl.data(), sizeof(int2) * cl.size(), hipMemcpyHostToDevice));
  d.tg = (TriGeom *)p1;
  d.rw = (RwgRec *)p2;
  d.close = (int2 *)p3;
  d.T = t.T;
  d.E = t.E;
  d.nclose = (unsigned int)cl.size();
  return 0;
}

static void free_allocs(bem_ctx *c, std::vector<void *> &allocs, size_t bytes) {
  (void)hipDeviceSynchronize();
  for (void *p : allocs) (void)hipFree(p);
  allocs.clear();
  c->bytes -= bytes;
}

// Body-local view of a tree: triangles/RWGs of body b re-based to 0.
static void body_subtree(const bemgeo::Tree &t, int b, bemgeo::Tree &out) {
  int t0 = t.tri_off[b], t1 = (b + 1 < (int)t.tri_off.size()) ? t.tri_off[b + 1] : t.T;
  int e0 = t.rwg_off[b], e1 = (b + 1 < (int)t.rwg_off.size()) ? t.rwg_off[b + 1] : t.E;
  out.T = t1 - t0;
  out.E = e1 - e0;
  out.tg.assign(t.tg.begin() + t0, t.tg.begin() + t1);
  out.rw.assign(t.rw.begin() + e0, t.rw.begin() + e1);
  for (auto &r : out.rw) { r.tp -= t0; r.tn -= t0; }
  out.close.clear();
  for (auto &p : t.close)
    if (p.first >= t0 && p.first < t1) out.close.push_back({p.first - t0, p.second - t0});
  out.tri_off = {0};
  out.rwg_off = {0};
}

static const size_t PAIR_CACHE = 8;

static int pair_geom(bem_ctx *c, bem_body *b1, bem_body *b2, const double *gap, PairGeom **out) {
  for (auto it = c->pairs.begin(); it != c->pairs.end(); ++it)
    if (it->b1 == b1 && it->b2 == b2 && it->gap[0] == gap[0] && it->gap[1] == gap[1] && it->gap[2] == gap[2]) {
      c->pairs.splice(c->pairs.begin(), c->pairs, it);
      *out = &c->pairs.front();
      return 0;
    }
  if (c->pairs.size() >= PAIR_CACHE) {
    PairGeom &old = c->pairs.back();
    free_allocs(c, old.allocs, old.bytes);
    c->pairs.pop_back();
  }
  c->pairs.emplace_front();
  PairGeom &g = c->pairs.front();
  g.b1 = b1;
  g.b2 = b2;
  for (int k = 0; k < 3; k++) g.gap[k] = gap[k];
  std::vector<const bemgeo::BodyMesh *> bm = {&b1->mesh, &b2->mesh};
  double shifts[6] = {gap[0], gap[1], gap[2], 0, 0, 0};
  bemgeo::Tree top, s1, s2;
  bemgeo::build_tree(bm, shifts, top);
  body_subtree(top, 0, s1);
  body_subtree(top, 1, s2);
  int e = upload_tree(c, top, g.top, g.allocs, g.bytes);
  if (!e) e = upload_tree(c, s1, g.s1, g.allocs, g.bytes);
  if (!e) e = upload_tree(c, s2, g.s2, g.allocs, g.bytes);
  if (e) {
    free_allocs(c, g.allocs, g.bytes);
    c->pairs.pop_front();
    return e;
  }
  *out = &g;
  return 0;
}

// Triangle-pair tiles of one tree in medium md, on the lane's stream.
static int assemble_tiles(bem_ctx *c, Lane &L, const DevTree &d, const Medium &md, double *tiles) {
  const int T = d.T;
  size_t npairs = (size_t)T * (size_t)T;
  unsigned int *q36cnt;
  int2 *q36;
  unsigned int cap = (unsigned int)std::min<size_t>(npairs, (size_t)1 << 24);
  CHK(ws_get(c, L, "q36_count", sizeof(unsigned int), (void **)&q36cnt));
  CHK(ws_get(c, L, "q36", sizeof(int2) * (size_t)cap, (void **)&q36));
  DEV(c, hipMemsetAsync(q36cnt, 0, sizeof(unsigned int), L.st));
  int e = bem_launch_far_g6(d.tg, T, &md, 0, T, 0, T, tiles, npairs, q36, q36cnt, cap, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "far_g6 launch: %s", hipGetErrorString((hipError_t)e));
  e = bem_launch_far_g36(d.tg, T, &md, q36, q36cnt, cap, 0, T, tiles, npairs, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "far_g36 launch: %s", hipGetErrorString((hipError_t)e));
  e = bem_launch_close(d.tg, T, &md, d.close, d.nclose, 0, T, tiles, npairs, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "close launch: %s", hipGetErrorString((hipError_t)e));
  return 0;
}

static double theta(double omega, double T) {
  // src/HeatTransfer.cpp:3-6 (T = 0 gives exp(inf) and exactly 0)
  return SI_HBAR * omega / (std::exp(SI_HBAR * omega / (SI_KB * T)) - 1);
}

// ------------------------------------------------------------- one solve
// Issues every kernel of one FP64 spectral-power solve on the lane's stream
// and returns without waiting; lane_finish() collects the result.
static int lane_issue(bem_ctx *c, Lane &L, PairGeom &g, const bem_request *q, bem_result *res) {
  L.t0 = std::chrono::steady_clock::now();
  L.req = *q;
  L.res = res;
  double omegaSI = q->eV * SI_EV / SI_HBAR;
  double omega = omegaSI * q->a / SI_C;
  const int m1 = 2 * g.s1.E, m2 = 2 * g.s2.E, M = m1 + m2;
  Medium mext = make_medium(1.0, 0.0, 1.0, 0.0, omega);
  Medium m1d = make_medium(q->eps1[0], q->eps1[1], q->mu1[0], q->mu1[1], omega);
  Medium m2d = make_medium(q->eps2[0], q->eps2[1], q->mu2[0], q->mu2[1], omega);

  double2 *W, *G1, *G2, *B1, *partial, *invA, *invS;
  double *tiles;
  int *ipA, *info;
  unsigned long long *gro;
  const bool tournament = (q->flags & BEM_SOLVE_TOURNAMENT) != 0;
  size_t npairs_top = (size_t)g.top.T * g.top.T;
  CHK(ws_get(c, L, "W", sizeof(double2) * (size_t)M * M, (void **)&W));
  CHK(ws_get(c, L, "G1", sizeof(double2) * (size_t)m1 * m1, (void **)&G1));
  CHK(ws_get(c, L, "G2", sizeof(double2) * (size_t)m2 * m2, (void **)&G2));
  CHK(ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * npairs_top, (void **)&tiles));
  CHK(ws_get(c, L, "ipiv", sizeof(int) * (size_t)(M + 64), (void **)&ipA));
  int *ipS = ipA + m1;
  CHK(ws_get(c, L, "info", sizeof(int) * 4, (void **)&info));
  CHK(ws_get(c, L, "growth", sizeof(unsigned long long) * 4, (void **)&gro));
  CHK(ws_get(c, L, "invA", inv_bytes(m1), (void **)&invA));
  CHK(ws_get(c, L, "invS", inv_bytes(m2), (void **)&invS));
  CHK(ws_get(c, L, "B1", sizeof(double2) * (size_t)m1 * m2, (void **)&B1));
  size_t np = bem_trace_partials(m1, m2);
  CHK(ws_get(c, L, "trace_partial", sizeof(double2) * (np + 1), (void **)&partial));
  double2 *trace_out = partial + np;
  DEV(c, hipMemsetAsync(info, 0, sizeof(int) * 4, L.st));
  DEV(c, hipMemsetAsync(gro, 0, sizeof(unsigned long long) * 4, L.st));

  DEV(c, hipEventRecord(L.ph[0], L.st));
  // ---- assembly: W = Z_ext(top), then + Z_int(b) on the diagonal blocks
  CHK(assemble_tiles(c, L, g.top, mext, tiles));
  int e = bem_launch_gather(tiles, npairs_top, g.top.T, 0, g.top.rw, 0, g.top.E, 0, g.top.E, &mext, W, M, 0, 0, 0,
                            nullptr, 0, 0, 0, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "gather: %d", e);
  CHK(assemble_tiles(c, L, g.s1, m1d, tiles));
  e = bem_launch_gather(tiles, (size_t)g.s1.T * g.s1.T, g.s1.T, 0, g.s1.rw, 0, g.s1.E, 0, g.s1.E, &m1d, W, M, 0, 0, 1,
                        G1, m1, 0, 0, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "gather: %d", e);
  CHK(assemble_tiles(c, L, g.s2, m2d, tiles));
  e = bem_launch_gather(tiles, (size_t)g.s2.T * g.s2.T, g.s2.T, 0, g.s2.rw, 0, g.s2.E, 0, g.s2.E, &m2d, W, M, m1, m1,
                        1, G2, m2, 0, 0, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "gather: %d", e);
  DEV(c, hipEventRecord(L.ph[1], L.st));

  // ---- solve: Y = A^-1 B, Sc = D - C Y, X = -Y Sc^-1
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  (void)zero;
  double2 *A = W, *Bm = W + (int64_t)m1 * M, *Cm = W + m1, *D = W + (int64_t)m1 * M + m1;
  if (tournament) CHK(lu_factor(c, L, A, M, m1, ipA, info, invA));
  else CHK(lu_factor_block(c, L, A, M, m1, ipA, info, invA, gro));
  CHK(lu_solve(c, L, A, M, m1, ipA, invA, Bm, M, m2));   // B <- Y
  CHK(gemm(c, L, m2, m2, m1, mone, Cm, M, Bm, M, one, D, M));
  if (tournament) CHK(lu_factor(c, L, D, M, m2, ipS, info + 1, invS));
  else CHK(lu_factor_block(c, L, D, M, m2, ipS, info + 1, invS, gro + 2));
  // B <- Y Sc^-1 = -X; the trace below is even in X, so the sign is dropped
  CHK(lu_solve_right(c, L, D, M, m2, ipS, invS, Bm, M, m1));
  double2 *X = Bm;
  DEV(c, hipEventRecord(L.ph[2], L.st));

  // ---- products and trace: S = 1/4 sum conj(X sym(G2)) .* (sym(G1) X)
  e = bem_symm(G1, m1, m1, L.st);
  if (!e) e = bem_symm(G2, m2, m2, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "symm: %d", e);
  CHK(gemm(c, L, m1, m2, m1, one, G1, m1, X, M, zero, B1, m1));
  {
    EventRec *r = nullptr;
    if (c->profile_gemm) CHK(ev_begin(c, L, 8.0 * m1 * (double)m2 * m2, &r));
    e = bem_zgemm_trace(m1, m2, m2, one, X, M, G2, m2, B1, m1, partial, trace_out, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "trace gemm: %d", e);
    if (r) DEV(c, hipEventRecord(r->b, L.st));
  }
  DEV(c, hipEventRecord(L.ph[3], L.st));
  DEV(c, hipMemcpyAsync(L.h_trace, trace_out, sizeof(double2), hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_info, info, sizeof(int) * 4, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_growth, gro, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost, L.st));
  L.busy = true;
  return 0;
}

static int lane_finish(bem_ctx *c, Lane &L) {
  if (!L.busy) return 0;
  L.busy = false;
  DEV(c, hipStreamSynchronize(L.st));
  bem_result *res = L.res;
  const bem_request *q = &L.req;
  float ta = 0, tf = 0, tp = 0;
  (void)hipEventElapsedTime(&ta, L.ph[0], L.ph[1]);
  (void)hipEventElapsedTime(&tf, L.ph[1], L.ph[2]);
  (void)hipEventElapsedTime(&tp, L.ph[2], L.ph[3]);
  ev_collect(c, L);
  double omega = q->eV * SI_EV / SI_HBAR * q->a / SI_C;
  double2 tr = *L.h_trace;
  res->omega = omega;
  res->ReS = 0.25 * tr.x;
  res->ImS = 0.25 * tr.y;
  res->theta1 = theta(omega, q->T1);
  res->theta2 = theta(omega, q->T2);
  res->P = res->ReS * (res->theta1 - res->theta2);
  res->rel_err_est = 0;
  res->iters = 0;
  res->ms_assemble = ta;
  res->ms_factor = tf;
  res->ms_products = tp;
  res->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - L.t0).count();
  res->status = 0;
  res->pivoting = (q->flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
  res->growth = 0.0;
  if (!res->pivoting) {
    double g[4];
    std::memcpy(g, L.h_growth, sizeof g);
    double ga = g[1] > 0 ? g[0] / g[1] : 0.0, gs = g[3] > 0 ? g[2] / g[3] : 0.0;
    res->growth = std::max(ga, gs);
    if (!std::isfinite(res->growth)) res->growth = INFINITY;
  }
  c->t_last[0] = ta;
  c->t_last[4] = tf;
  c->t_last[5] = tp;
  if (L.h_info[0] || L.h_info[1]) {
    res->status = BEM_E_SINGULAR;
    return fail(c, BEM_E_SINGULAR, "exactly singular pivot (A: %d, Schur: %d)", L.h_info[0], L.h_info[1]);
  }
  return 0;
}

static int check_request(bem_ctx *c, const bem_request *q) {
  if (!(q->eV > 0) || !(q->a > 0)) return fail(c, BEM_E_INVALID, "eV and a must be positive");
  if (q->precision != BEM_PREC_FP64) return fail(c, BEM_E_UNSUPPORTED, "precision mode %d not available", q->precision);
  return 0;
}

// ------------------------------------------------------------------- C ABI
extern "C" {

int bem_ctx_create(bem_ctx **out, const bem_cfg *cfg) {
  if (!out) return BEM_E_INVALID;
  bem_ctx *c = new bem_ctx();
  int nl = 2;
  if (cfg) {
    c->device = cfg->device;
    c->budget = cfg->hbm_budget;
    if (cfg->lanes > 0) nl = std::min(cfg->lanes, 8);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || hipSetDevice(c->device) != hipSuccess) {
    delete c;
    return BEM_E_DEVICE;
  }
  c->lanes.resize(nl);
  for (int i = 0; i < nl; i++) {
    Lane &L = c->lanes[i];
    L.id = i;
    L.tag = "L" + std::to_string(i) + ":";
    bool ok = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) == hipSuccess;
    for (int k = 0; k < 4 && ok; k++) ok = hipEventCreate(&L.ph[k]) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_trace, sizeof(double2), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_info, sizeof(int) * 4, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_growth, sizeof(unsigned long long) * 4, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
      bem_ctx_destroy(c);
      return BEM_E_DEVICE;
    }
  }
  *out = c;
  return 0;
}

void bem_ctx_destroy(bem_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  for (auto &kv : c->ws)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto &kv : c->user_allocs) (void)hipFree(kv.first);
  for (auto &g : c->pairs)
    for (void *p : g.allocs) (void)hipFree(p);
  for (auto &L : c->lanes) {
    for (auto &r : L.ev) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (int k = 0; k < 4; k++)
      if (L.ph[k]) (void)hipEventDestroy(L.ph[k]);
    if (L.h_trace) (void)hipHostFree(L.h_trace);
    if (L.h_info) (void)hipHostFree(L.h_info);
    if (L.h_growth) (void)hipHostFree(L.h_growth);
    if (L.st) (void)hipStreamDestroy(L.st);
  }
  delete c;
}

const char *bem_last_error(bem_ctx *c) { return c ? c->err.c_str() : "null context"; }
size_t bem_ctx_device_bytes(bem_ctx *c) { return c ? c->bytes : 0; }

int bem_body_upload(bem_ctx *c, const double *xyz, int nv, const int32_t *tri, int nt, bem_body **out) {
  if (!c || !xyz || !tri || !out || nv < 3 || nt < 1) return fail(c, BEM_E_INVALID, "bad body arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  bem_body *b = new bem_body();
  b->ctx = c;
  auto &M = b->mesh;
  M.nv = nv;
  M.nt = nt;
  M.xyz.assign(xyz, xyz + 3 * (size_t)nv);
  M.tri.assign(tri, tri + 3 * (size_t)nt);
  int ne = bemgeo::build_edges(nv, nt, tri, M.edges);
  if (ne < 0) {
    delete b;
    return fail(c, BEM_E_INVALID, "triangle refers to a vertex outside [0, nv)");
  }
  if (2 * ne != 3 * nt) {
    delete b;
    return fail(c, BEM_E_MESH, "mesh is not closed: %d edges for %d triangles", ne, nt);
  }
  M.ne = ne;
  bemgeo::build_close_pairs(M);
  *out = b;
  return 0;
}

int bem_body_edges(bem_body *b, int32_t *edges, int *ne) {
  if (!b || !ne) return BEM_E_INVALID;
  *ne = b->mesh.ne;
  if (edges) std::memcpy(edges, b->mesh.edges.data(), sizeof(int32_t) * b->mesh.edges.size());
  return 0;
}

void bem_body_free(bem_body *b) {
  if (!b) return;
  bem_ctx *c = b->ctx;
  {
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    for (auto it = c->pairs.begin(); it != c->pairs.end();) {
      if (it->b1 == b || it->b2 == b) {
        free_allocs(c, it->allocs, it->bytes);
        it = c->pairs.erase(it);
      } else {
        ++it;
      }
    }
  }
  delete b;
}

int bem_pair_histogram(bem_ctx *c, int nb, bem_body *const *bodies, const double *shifts, int64_t *hist6) {
  if (!c || nb < 1 || nb > 2 || !bodies || !shifts || !hist6) return fail(c, BEM_E_INVALID, "bad arguments");
  std::vector<const bemgeo::BodyMesh *> bm;
  for (int i = 0; i < nb; i++) bm.push_back(&bodies[i]->mesh);
  bemgeo::Tree t;
  bemgeo::build_tree(bm, shifts, t);
  for (int k = 0; k < 6; k++) hist6[k] = 0;
  // host restatement of the device classifier (same arithmetic, no contraction)
  for (int a = 0; a < t.T; a++)
    for (int b = 0; b < t.T; b++) {
      const TriGeom &A = t.tg[a], &B = t.tg[b];
      bool close = false;
      if (A.body == B.body) {
        if (a == b) close = true;
        for (int i = 0; i < 3 && !close; i++)
          if (A.vid[i] == B.vid[0] || A.vid[i] == B.vid[1] || A.vid[i] == B.vid[2]) close = true;
      }
      if (close) { hist6[0]++; continue; }
      double vx = A.centre[0] - B.centre[0], vy = A.centre[1] - B.centre[1], vz = A.centre[2] - B.centre[2];
      volatile double d = vx * vx;
      d = d + vy * vy;
      d = d + vz * vz;
      double em = A.emax > B.emax ? A.emax : B.emax;
      double e2 = em * em;
      if (d * 12.0 > e2) hist6[1]++;
      else hist6[5]++;
    }
  return 0;
}

int bem_assemble(bem_ctx *c, int nb, bem_body *const *bodies, const double *shifts, const double *eps, const double *mu,
                 double omega, void *Z, int64_t ldz, int flags) {
  if (!c || nb < 1 || nb > 2 || !bodies || !shifts || !eps || !mu || !Z) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_finish(c, L));
  std::vector<const bemgeo::BodyMesh *> bm;
  for (int i = 0; i < nb; i++) bm.push_back(&bodies[i]->mesh);
  bemgeo::Tree t;
  bemgeo::build_tree(bm, shifts, t);
  int M = 2 * t.E;
  if (ldz < M) return fail(c, BEM_E_INVALID, "ldz %lld < %d", (long long)ldz, M);
  DevTree d;
  std::vector<void *> allocs;
  size_t tb = 0;
  int e = upload_tree(c, t, d, allocs, tb);
  Medium md = make_medium(eps[0], eps[1], mu[0], mu[1], omega);
  double *tiles = nullptr;
  size_t npairs = (size_t)t.T * t.T;
  double2 *Zd = nullptr;
  bool dev = flags & BEM_ASM_DEVICE_PTR;
  int64_t ld = dev ? ldz : M;
  if (!e) e = ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * npairs, (void **)&tiles);
  if (!e) e = assemble_tiles(c, L, d, md, tiles);
  if (!e) {
    if (dev) Zd = (double2 *)Z;
    else e = ws_get(c, L, "asm_Z", sizeof(double2) * (size_t)M * M, (void **)&Zd);
  }
  if (!e && bem_launch_gather(tiles, npairs, t.T, 0, d.rw, 0, t.E, 0, t.E, &md, Zd, ld, 0, 0, 0, nullptr, 0, 0, 0, L.st))
    e = fail(c, BEM_E_DEVICE, "gather launch failed");
  if (!e && !dev && hipMemcpy2DAsync(Z, sizeof(double2) * ldz, Zd, sizeof(double2) * M, sizeof(double2) * M, M,
                                     hipMemcpyDeviceToHost, L.st) != hipSuccess)
    e = fail(c, BEM_E_DEVICE, "copy of Z failed");
  if (!e && hipStreamSynchronize(L.st) != hipSuccess) e = fail(c, BEM_E_DEVICE, "assembly failed on the device");
  free_allocs(c, allocs, tb);
  return e;
}

int bem_spectral_power(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *q, bem_result *res) {
  if (!c || !b1 || !b2 || !q || !res) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  std::memset(res, 0, sizeof *res);
  CHK(check_request(c, q));
  stats_reset(c);
  PairGeom *g;
  CHK(pair_geom(c, b1, b2, q->gap, &g));
  Lane &L = c->lanes[0];
  CHK(lane_finish(c, L));
  int e = lane_issue(c, L, *g, q, res);
  int f = lane_finish(c, L);
  if (e) return e;
  if (!f && !(res->growth <= BEM_GROWTH_LIMIT) && !(q->flags & BEM_SOLVE_TOURNAMENT)) {
    bem_request q2 = *q;
    q2.flags |= BEM_SOLVE_TOURNAMENT;
    e = lane_issue(c, L, *g, &q2, res);
    f = lane_finish(c, L);
    if (e) return e;
  }
  c->t_last[6] = c->gemm_ms;
  c->t_last[7] = c->gemm_n;
  c->t_last[8] = c->gemm_flops;
  return f;
}

int bem_sweep(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *reqs, int n, bem_result *res) {
  if (!c || !b1 || !b2 || !reqs || !res || n < 0) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  stats_reset(c);
  int first_err = 0;
  const int nl = (int)c->lanes.size();
  for (int i = 0; i < n; i++) {
    std::memset(&res[i], 0, sizeof res[i]);
    Lane &L = c->lanes[i % nl];
    int f = lane_finish(c, L);
    if (f && !first_err) first_err = f;
    int e = check_request(c, &reqs[i]);
    PairGeom *g = nullptr;
    if (!e) e = pair_geom(c, b1, b2, reqs[i].gap, &g);
    if (!e) e = lane_issue(c, L, *g, &reqs[i], &res[i]);
    if (e) {
      res[i].status = e;
      if (!first_err) first_err = e;
      (void)hipStreamSynchronize(L.st);
      L.busy = false;
    }
  }
  for (auto &L : c->lanes) {
    int f = lane_finish(c, L);
    if (f && !first_err) first_err = f;
  }
  // requests whose block-local pivoting grew too much are solved again with
  // tournament pivoting
  for (int i = 0; i < n; i++) {
    if (res[i].status || res[i].pivoting || res[i].growth <= BEM_GROWTH_LIMIT) continue;
    bem_request q2 = reqs[i];
    q2.flags |= BEM_SOLVE_TOURNAMENT;
    Lane &L = c->lanes[0];
    PairGeom *g = nullptr;
    int e = pair_geom(c, b1, b2, q2.gap, &g);
    if (!e) e = lane_issue(c, L, *g, &q2, &res[i]);
    int f = e ? e : lane_finish(c, L);
    if (f) {
      res[i].status = f;
      if (!first_err) first_err = f;
    }
  }
  c->t_last[6] = c->gemm_ms;
  c->t_last[7] = c->gemm_n;
  c->t_last[8] = c->gemm_flops;
  return first_err;
}

int bem_zgetrf(bem_ctx *c, void *A, int64_t lda, int n, int32_t *ipiv) {
  if (!c || !A || !ipiv || n < 0 || lda < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_finish(c, L));
  stats_reset(c);
  int *info;
  double2 *inv;
  CHK(ws_get(c, L, "info_getrf", sizeof(int), (void **)&info));
  CHK(ws_get(c, L, "inv_getrf", inv_bytes(n), (void **)&inv));
  DEV(c, hipMemsetAsync(info, 0, sizeof(int), L.st));
  CHK(lu_factor(c, L, (double2 *)A, lda, n, ipiv, info, inv));
  int h = 0;
  DEV(c, hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, L.st));
  DEV(c, hipStreamSynchronize(L.st));
  ev_collect(c, L);
  if (h) return fail(c, BEM_E_SINGULAR, "zero pivot at column %d", h - 1);
  return 0;
}

int bem_zgetrs(bem_ctx *c, const void *A, int64_t lda, int n, const int32_t *ipiv, void *B, int64_t ldb, int nrhs) {
  if (!c || !A || !ipiv || !B || n < 0 || lda < n || ldb < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_finish(c, L));
  stats_reset(c);
  double2 *inv;
  CHK(ws_get(c, L, "inv_getrs", inv_bytes(n), (void **)&inv));
  int e = bem_diag_block((double2 *)A, lda, n, 0, (n + LU_NB - 1) / LU_NB, LU_NB, 0, inv, nullptr, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "diag block inverses: %d", e);
  CHK(lu_solve(c, L, (const double2 *)A, lda, n, ipiv, inv, (double2 *)B, ldb, nrhs));
  DEV(c, hipStreamSynchronize(L.st));
  ev_collect(c, L);
  return 0;
}

int bem_zgemm(bem_ctx *c, int m, int n, int k, const double *alpha, const void *A, int64_t lda, const void *B,
              int64_t ldb, const double *beta, void *C, int64_t ldc) {
  if (!c || !alpha || !beta || m < 0 || n < 0 || k < 0) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_finish(c, L));
  stats_reset(c);
  CHK(gemm(c, L, m, n, k, make_double2(alpha[0], alpha[1]), (const double2 *)A, lda, (const double2 *)B, ldb,
           make_double2(beta[0], beta[1]), (double2 *)C, ldc));
  DEV(c, hipStreamSynchronize(L.st));
  ev_collect(c, L);
  c->t_last[6] = c->gemm_ms;
  c->t_last[7] = c->gemm_n;
  c->t_last[8] = c->gemm_flops;
  return 0;
}

int bem_device_alloc(bem_ctx *c, size_t bytes, void **ptr) {
  if (!c || !ptr) return BEM_E_INVALID;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  if (c->budget && c->bytes + bytes > c->budget) return fail(c, BEM_E_NOMEM, "allocation exceeds the HBM budget");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(c, BEM_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  c->user_allocs[*ptr] = bytes;
  c->bytes += bytes;
  return 0;
}

int bem_device_free(bem_ctx *c, void *ptr) {
  if (!c) return BEM_E_INVALID;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  auto it = c->user_allocs.find(ptr);
  if (it == c->user_allocs.end()) return fail(c, BEM_E_INVALID, "pointer not allocated by this context");
  DEV(c, hipDeviceSynchronize());
  DEV(c, hipFree(ptr));
  c->bytes -= it->second;
  c->user_allocs.erase(it);
  return 0;
}

int bem_memcpy_h2d(bem_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return BEM_E_INVALID;
  DEV(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

int bem_memcpy_d2h(bem_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return BEM_E_INVALID;
  DEV(c, hipDeviceSynchronize());
  DEV(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int bem_last_timings(bem_ctx *c, double *out9) {
  if (!c || !out9) return BEM_E_INVALID;
  for (int i = 0; i < 9; i++) out9[i] = c->t_last[i];
  return 0;
}

}  // extern "C"
