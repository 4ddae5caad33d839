// libbem_hip host runtime: contexts, bodies, assembly orchestration and the
// PMCHWT solve behind the C ABI of include/bem_hip.h.
//
// CalcHeatTransfer (src/HeatTransfer.cpp:8-59) inverts W = Z_ext + blkdiag(G1,G2)
// and forms S = 1/4 tr(B2^H B1) with B1 = sym(G1) X, B2 = X sym(G2),
// X = W^-1[0:m1, m1:M].  Only X is needed, so the solve here is the block
// elimination of W = [[A, B], [C, D]]:
//   Y = A^-1 B,  Sc = D - C Y,  X = -Y Sc^-1
// with blocked LU of A and of Sc (partial pivoting inside each 64 x 64
// diagonal block, tournament pivoting when the growth factor says so),
// followed by B1 = sym(G1) X and the trace: since sym(G) = G + G^H is
// Hermitian, 4 S = tr(sym(G2) X^H B1) with X^H B1 Hermitian, so one GEMM forms
// the upper half of X^H B1 and contracts it against sym(G2) in its epilogue
// (no B2 stored).
//
// Concurrency.  A context owns `lanes` HIP streams, each with its own named
// workspaces.  A sweep keeps up to `lanes` independent solves in flight: the
// latency-bound LU panels of one solve overlap the GEMMs of another.  Every
// solve's kernels stay on its own stream, so results do not depend on what
// else is in flight.
//
// Serving.  bem_submit queues a request and returns a ticket; a scheduler
// thread of the context admits queued requests (shortest estimated run time
// first, with aging) onto free lanes whenever the request's device footprint (lane workspaces plus any
// self blocks not yet cached) fits the HBM budget, and retires lanes whose
// stream signalled completion through a host callback.  Requests arriving
// while others run are therefore admitted as lanes and memory free up, and
// their kernels overlap on the device with those already in flight.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/bem_hip.h"
#include "bem_common.h"
#include "geometry.h"

extern "C" {
int bem_launch_far_g6(const TriGeom *tg, int T, const Medium *md, int row_lo, int row_hi, int col_lo, int col_hi,
                      double *tiles, size_t npairs, int2 *q36, unsigned int *q36_count, unsigned int q36_cap,
                      hipStream_t st);
int bem_launch_far_g36(const TriGeom *tg, int T, const Medium *md, const int2 *pairs, const unsigned int *count_dev,
                       unsigned int max_count, int row_lo, int row_hi, int col_lo, int col_hi, double *tiles,
                       size_t npairs, hipStream_t st);
int bem_launch_close(const TriGeom *tg, int T, const Medium *md, const int2 *pairs, unsigned int count, int row_lo,
                     int row_hi, int col_lo, int col_hi, double *tiles, size_t npairs, double *mom,
                     unsigned int mom_pairs, hipStream_t st);
size_t bem_close_mom_doubles(unsigned int mom_pairs);
int bem_launch_gather(const double *tiles, size_t npairs, int tld, int row_lo, int col_lo, const RwgRec *rw, int r_lo,
                      int r_hi, int c_lo, int c_hi, const Medium *md, double2 *Z, int64_t ldz, int zr0, int zc0,
                      int add, double2 *Zcopy, int64_t ldc, int cr0, int cc0, hipStream_t st);
int bem_launch_q36_check(const unsigned int *count, unsigned int cap, int *flag, hipStream_t st);
int bem_zgemm_tiled(int tile, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B,
                    int64_t ldb, double2 beta, double2 *C, int64_t ldc, hipStream_t st);
int bem_zgemm_tile_valid(int tile);
int bem_cgemm_tiled(int tile, int m, int n, int k, float2 alpha, const float2 *A, int64_t lda, const float2 *B,
                    int64_t ldb, float2 beta, float2 *C, int64_t ldc, hipStream_t st);
int bem_cgemm_tile_valid(int tile);
int bem_zgemm_raw(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                  double2 beta, double2 *C, int64_t ldc, hipStream_t st);
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
int bem_permute_rows_local(double2 *X, int64_t ldx, int n, int r, const int *ipiv, int *perm, hipStream_t st);
int bem_permute_cols_local(double2 *X, int64_t ldx, int m, int n, const int *ipiv, int *perm, hipStream_t st);
int bem_permute_rows_local_c(float2 *X, int64_t ldx, int n, int r, const int *ipiv, int *perm, hipStream_t st);
int bem_permute_cols_local_c(float2 *X, int64_t ldx, int m, int n, const int *ipiv, int *perm, hipStream_t st);
int bem_cgemm_wide(int m, int n);
int bem_cgemm_raw(int m, int n, int k, float2 alpha, const float2 *A, int64_t lda, const float2 *B, int64_t ldb,
                  float2 beta, float2 *C, int64_t ldc, hipStream_t st);
size_t bem_htrace_partials(int n);
size_t bem_chtrace_partials(int n);
int bem_zgemm_htrace(int n, int k, const double2 *A, int64_t lda, const double2 *B, int64_t ldb, const double2 *E,
                     int64_t lde, double2 *partial, double2 *out, hipStream_t st);
int bem_cgemm_htrace(int n, int k, const float2 *A, int64_t lda, const float2 *B, int64_t ldb, const float2 *E,
                     int64_t lde, double2 *partial, double2 *out, hipStream_t st);
int bem_diag_lu_piv_c(float2 *A, int64_t lda, int n, int k0, float2 *inv, int *ipiv, int *info, hipStream_t st);
int bem_absmax_c(const float2 *A, int64_t lda, int m, int n, int upper, unsigned long long *out, hipStream_t st);
int bem_col_swap_c(float2 *X, int64_t ldx, int m, int k0, int k1, const int *ipiv, hipStream_t st);
int bem_laswp_c(float2 *A, int64_t lda, int c0, int c1, int skip0, int skip1, int k0, int k1, const int *ipiv,
                hipStream_t st);
int bem_z2c(const double2 *src, int64_t lds, float2 *dst, int64_t ldd, int m, int n, hipStream_t st);
int bem_c2z(const float2 *src, int64_t lds, double2 *dst, int64_t ldd, int m, int n, hipStream_t st);
int bem_add_z_to_c(const double2 *src, int64_t lds, float2 *dst, int64_t ldd, int m, int n, hipStream_t st);
int bem_identity_minus(double2 *Cm, int64_t ldc, int n, hipStream_t st);
int bem_signed_transpose(const double2 *B, int64_t ldb, int m, int n, double2 *C, int64_t ldc, hipStream_t st);
int bem_fill_hash(void *p, size_t bytes, int as_float, unsigned long long seed, hipStream_t st);
int bem_c2z_add(const float2 *src, int64_t lds, double2 *dst, int64_t ldd, int m, int n, hipStream_t st);
size_t bem_sumsq_partials(int m, int n);
int bem_sumsq(const double2 *A, int64_t lda, int m, int n, double *partial, int add, double *out, hipStream_t st);
int bem_zaxpy(double2 alpha, const double2 *src, int64_t lds, double2 *dst, int64_t ldd, int m, int n, hipStream_t st);
int bem_zdotc_cols(const double2 *G, int64_t ldg, const double2 *D, int64_t ldd, int m, int n, double2 *out,
                   hipStream_t st);
int bem_trsm_few(const double2 *A, int64_t lda, int n, const double2 *inv, double2 *X, int64_t ldx, int r, int lower,
                 double2 *Y, int64_t ldy, hipStream_t st);
int bem_trsm_few_c(const float2 *A, int64_t lda, int n, const float2 *inv, float2 *X, int64_t ldx, int r, int lower,
                   float2 *Y, int64_t ldy, hipStream_t st);
size_t bem_zgemm_skinny_partials(int m, int r, int K);
int bem_fill_zero(void *p, size_t bytes, hipStream_t st);
int bem_zgemm_skinny(int m, int r, int K, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                     double2 beta, double2 *C, int64_t ldc, double2 *part, hipStream_t st);
size_t bem_zgemm_ta_partials(int m, int n, int k);
int bem_zgemm_ta(int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B, int64_t ldb,
                 double2 beta, double2 *C, int64_t ldc, double2 *part, hipStream_t st);
int bem_chol_inv64(const double2 *G, int64_t ldg, int w, double2 *Rinv, int *info, hipStream_t st);
int bem_transpose_op(const double2 *src, int64_t lds, int m, int n, double2 *dst, int64_t ldd, int conj, int sign_rows,
                     hipStream_t st);
int bem_conj_sign(const double2 *src, int64_t lds, int m, int n, double2 *dst, int64_t ldd, int sign_rows,
                  hipStream_t st);
size_t bem_aca_state_bytes(void);
size_t bem_aca_partial_bytes(int m, int n);
int bem_aca_init(void *state, int m, int n, hipStream_t st);
int bem_aca_read(const void *state, int *nu, int *rank, int *done, int *full);
int bem_aca_steps(double2 *M, int64_t ldm, int m, int n, double eps2, int kmax, double2 *U, int64_t ldu, double2 *V,
                  int64_t ldv, int *rows, int *cols, void *state, void *partial, int steps, hipStream_t st);
int bem_jacobi_sweep(double2 *A, int64_t lda, int m, int k, double2 *W, int64_t ldw, int kw, double tol, int *rotated,
                     hipStream_t st);
int bem_col_norms(const double2 *A, int64_t lda, int m, int k, double *sigma, hipStream_t st);
int bem_gather_cols(const double2 *src, int64_t lds, int m, const int *idx, const double *scale, int nc, double2 *dst,
                    int64_t ldd, hipStream_t st);
int bem_launch_dipole_rhs(const TriGeom *tg, int T, const RwgRec *rw, int E, const void *src, int nsrc, int tmax,
                          double2 *tb, double2 *B, int64_t ldb, hipStream_t st);
size_t bem_flux_partials(int nt, int ncol);
int bem_launch_flux(const void *ft, int nt, double omega, const double2 *X, int64_t ldx, int ncol, double *part,
                    double *out, hipStream_t st);
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

// One timed launch: its kernel class (BEM_KS_*) and algorithmic work.
struct EventRec {
  hipEvent_t a, b;
  int kind;
  double flops, bytes;
};

struct DevTree {
  int T = 0, E = 0;
  TriGeom *tg = nullptr;
  RwgRec *rw = nullptr;
  int2 *close = nullptr;
  unsigned int nclose = 0;
};

// Device geometry of one (body 1, body 2, gap): the two-body tree, whose
// cross blocks every frequency of that gap assembles.
struct PairGeom {
  const void *b1 = nullptr, *b2 = nullptr;
  double gap[3] = {0, 0, 0};
  DevTree top;
  int T1 = 0, E1 = 0;   // triangles / RWGs of body 1 (the first ones of top)
  std::vector<std::pair<void *, size_t>> allocs;
  size_t bytes = 0;
  int pins = 0;         // lanes whose solve in flight reads this tree
  int lr_blocks = 0;    // range-finder blocks the last low-rank solve on this pair needed (issue hint)
};

// Host-visible status of a self entry, written by its producing stream.
struct SelfStatus {
  int info[4];                    // [0] zero pivot, [1] 36-point queue overflow
  unsigned long long growth[3];   // max|U|, max|A|, max|L| (strictly lower) of the LU (role 0)
};

// Gap-independent blocks of one body at one frequency and precision (src/HeatTransfer.cpp:
// 26-29 assembles them inside Z_ext and Z_int on every call): the self block
// Z_ext(b, b) + Z_int(b) -- packed LU with interchanges and diagonal-block
// inverses when the body sits in the first role (A), raw in the second (D) --
// and sym(Z_int(b)).  Shared by every request on any lane; `ready` is
// recorded on the producing stream, consumers wait on it.
struct SelfEntry {
  const void *body = nullptr;
  double key[5] = {0, 0, 0, 0, 0};   // eps re/im, mu re/im, omega
  int role = 0, pivot = 0, prec = BEM_PREC_FP64;
  int n = 0;
  double2 *Z = nullptr, *Gs = nullptr, *inv = nullptr;
  // the unfactored FP64 self block (residuals of the error probe and of
  // refinement): a copy taken before the LU for an FP64 entry in role 0, Z
  // itself otherwise
  double2 *Araw = nullptr;
  // BEM_PREC_MIXED: Z stays the raw FP64 block (refinement residuals) and
  // Z32 holds its complex64 copy (packed LU with inv32 in role 0)
  float2 *Z32 = nullptr, *Gs32 = nullptr, *inv32 = nullptr;
  int *ipiv = nullptr, *dinfo = nullptr;
  unsigned long long *dgrowth = nullptr;
  SelfStatus *h = nullptr;   // pinned
  hipEvent_t ready = nullptr;
  std::vector<std::pair<void *, size_t>> allocs;   // device blocks and their sizes
  size_t bytes = 0;
  int pins = 0;
  uint64_t stamp = 0;
};

// Device arena of the self-block cache in a context without an HBM budget.
// Entries come and go with every frequency of a sweep, and hipFree
// synchronises the whole device: once the cache reached its cap, each
// eviction made the next request wait for every lane to drain (a 4-lane sweep
// of 1280-triangle sphere pairs fell from 24 to 14.5 solves/s).  The arena
// takes device memory in segments of at least ARENA_SEG bytes and keeps it
// until the context ends or a body's release leaves a segment empty; blocks
// are carved first fit from per-segment free
// lists (ARENA_ALIGN granules, neighbours coalesced on release), so an
// eviction is host bookkeeping only.  A context with a budget allocates and
// frees each block instead, so that its footprint is exactly what it holds.
static const size_t ARENA_SEG = (size_t)2 << 30;
static const size_t ARENA_ALIGN = 256;
struct ArenaSeg {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free;   // offset -> length, disjoint, coalesced
};

// Error probe of a solve, written on the device (see bem_hip.h, Error probe):
// max|d_1| and max|x_1| as the bit patterns of non-negative doubles, and the
// squared Frobenius norms of the probe residual r and of w.
struct ProbeOut {
  unsigned long long dmax, xmax;
  double rr, ww;
  double2 h[BEM_PROBE_COLS];   // g_j^H d_1j, g_j = sym G1 X sym G2 w_j (first-order change of Re S)
};

// Stages of the request in flight on a lane.
enum { ST_IDLE = 0, ST_SOLVE = 1, ST_REFINE = 2, ST_COMPRESS = 3 };

// One stream and its workspaces; holds at most one solve in flight.
struct Lane {
  int id = 0;
  hipStream_t st = nullptr;
  hipStream_t aux = nullptr;            // look-ahead panel factorisations
  hipEvent_t la_a = nullptr, la_b = nullptr;
  std::string tag;
  std::vector<EventRec> ev;
  size_t ev_used = 0;
  hipEvent_t ph[4] = {nullptr, nullptr, nullptr, nullptr};
  double2 *h_trace = nullptr;   // pinned host slots of the results
  int *h_info = nullptr;        // [0] zero pivot in the Schur LU, [1] queue overflow
  unsigned long long *h_growth = nullptr;   // max|U|, max|A|, max|L| of LU(S)
  bool busy = false;
  bem_result *res = nullptr;
  bem_request req;
  SelfEntry *e1 = nullptr, *e2 = nullptr;   // pinned while in flight
  PairGeom *pg = nullptr;                   // pinned while in flight
  std::chrono::steady_clock::time_point t0;
  size_t ws_bytes = 0;            // bytes of this lane's named workspaces
  int prec = BEM_PREC_FP64;       // precision of the solve in flight
  bool refine = false;
  int stage = ST_IDLE;            // what the lane's stream is running for the request
  int iters = 0;                  // refinement steps issued for the request
  int restarts = 0;               // re-solves of the request so far
  bool probe_async = false;       // the stage's probe runs on aux (probe_join pending)
  ProbeOut *h_probe = nullptr;    // pinned copy of the last stage's probe
  double *h_rfull = nullptr;      // pinned: squared norm of the last refinement residual
  // range finder in flight (ST_COMPRESS): the block being compressed, its
  // factors' workspaces, the blocks issued and those whose estimates were read
  struct {
    const double2 *B = nullptr;
    double2 *Q = nullptr, *Vh = nullptr;
    int m1 = 0, m2 = 0, rmax = 0, nbmax = 0, issued = 0, checked = 0;
    double tol = 0.0;
    std::chrono::steady_clock::time_point t0;
  } lr;
  int rank = 0;                   // rank of the coupling block of a low-rank solve (0: dense)
  double ms_compress = 0.0;       // its compression time (host clock)
  double *h_lr = nullptr;         // pinned: norms read back by the low-rank compression
  int *h_lri = nullptr;           // pinned: its status words
  // served request in flight (ticket 0: none)
  bem_ctx *ctx = nullptr;
  uint64_t ticket = 0;
  bem_body *sb1 = nullptr, *sb2 = nullptr;
  bem_result sres;
  std::chrono::steady_clock::time_point t_submit, t_issue;
  volatile int done_flag = 0;     // set by the stream's host callback
};

// A queued served request.
struct Pending {
  uint64_t ticket;
  bem_body *b1, *b2;
  bem_request req;
  std::chrono::steady_clock::time_point t_submit;
  int restarts = 0;
};

}  // namespace

struct bem_ctx {
  int device = 0;
  void *scratch = nullptr;          // device address of the allocation-free entry point's array
  int *h_noalloc = nullptr;         // pinned: its zero-pivot flag and growth maxima
  size_t budget = 0;
  size_t bytes = 0;
  std::string err;
  std::recursive_mutex mu;
  std::unordered_map<std::string, DevBuf> ws;   // named, grow-only workspaces
  std::unordered_map<void *, size_t> user_allocs;
  std::vector<Lane> lanes;
  std::list<PairGeom> pairs;                    // most recently used first
  std::list<SelfEntry> selfs;
  size_t self_bytes = 0, self_cap = 0;
  std::vector<ArenaSeg> arena;              // self-entry memory without a budget
  size_t arena_bytes = 0;
  std::vector<std::pair<void *, size_t>> ws_spare;   // workspaces outgrown by their lane (no budget)
  size_t ws_spare_bytes = 0;
  std::vector<SelfStatus *> spare_status;   // pinned slots and events of evicted entries
  std::vector<hipEvent_t> spare_events;
  uint64_t clock = 0;
  bool profile = true;   // bracket timed kernel classes with events
  bool lookahead = false;   // factor the next LU diagonal block on the lane's aux stream
  // GEMM tile per shape class (gemm_class), 0 = the built-in rule
  int ztile[3] = {0, 0, 0}, ctile[3] = {0, 0, 0};
  // statistics of the last API call
  double t_last[9] = {0};
  double gemm_ms = 0, gemm_n = 0, gemm_flops = 0;
  double kst[BEM_KS_COUNT][4] = {};   // per class: ms, launches, flops, bytes
  int self_hits = 0, self_misses = 0;
  std::unordered_set<bem_body *> bodies;
  // serving (queue, finished results and the scheduler; guarded by qmu)
  std::mutex qmu;
  std::condition_variable qcv, done_cv;
  std::deque<Pending> queue;
  std::unordered_map<uint64_t, bem_result> done;
  std::unordered_set<uint64_t> live;   // queued or running
  uint64_t next_ticket = 1;
  int running = 0;
  bool stop = false, kick = false;
  std::thread sched;
};

struct bem_body {
  bem_ctx *ctx;
  bemgeo::BodyMesh mesh;
  uint64_t fp = 0;              // FNV-1a of the vertex and triangle arrays
  DevTree dt;                   // the body alone, untranslated (self blocks)
  bool up = false;
  std::vector<std::pair<void *, size_t>> allocs;
  size_t bytes = 0;
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
static const size_t WS_SPARE_MAX = (size_t)1 << 30, WS_SPARE_TOTAL = (size_t)8 << 30;

static int ws_get(bem_ctx *c, Lane &L, const char *name, size_t bytes, void **out) {
  DevBuf &b = c->ws[L.tag + name];
  if (b.cap < bytes) {
    // a workspace that has to grow again takes headroom without a budget
    // (three times its old size, at most 1 GiB more): requests served
    // shortest first grow in size, and a regrowth that frees the old buffer
    // synchronises the device (hipFree)
    const size_t want = bytes;
    if (b.p && !c->budget) bytes = std::max(bytes, std::min(3 * b.cap, b.cap + ((size_t)1 << 30)));
    if (b.p) {
      DEV(c, hipStreamSynchronize(L.st));
      // idle now that the lane has drained: without a budget a buffer under
      // WS_SPARE_MAX is kept for any lane's next workspace instead of
      // hipFree'd (it stays in c->bytes), up to WS_SPARE_TOTAL in all
      if (!c->budget && b.cap < WS_SPARE_MAX && c->ws_spare_bytes + b.cap <= WS_SPARE_TOTAL) {
        c->ws_spare.emplace_back(b.p, b.cap);
        c->ws_spare_bytes += b.cap;
      } else {
        DEV(c, hipFree(b.p));
        c->bytes -= b.cap;
      }
      L.ws_bytes -= b.cap;
      b.p = nullptr;
      b.cap = 0;
    }
    if (c->budget && c->bytes + bytes > c->budget)
      return fail(c, BEM_E_NOMEM, "workspace '%s%s' (%zu bytes) exceeds the HBM budget", L.tag.c_str(), name, bytes);
    if (!c->budget) {   // the smallest outgrown workspace that holds the request
      size_t best = c->ws_spare.size();
      for (size_t i = 0; i < c->ws_spare.size(); i++)
        if (c->ws_spare[i].second >= want &&
            (best == c->ws_spare.size() || c->ws_spare[i].second < c->ws_spare[best].second))
          best = i;
      if (best < c->ws_spare.size()) {
        b.p = c->ws_spare[best].first;
        b.cap = c->ws_spare[best].second;
        c->ws_spare_bytes -= b.cap;
        c->ws_spare.erase(c->ws_spare.begin() + (long)best);
        L.ws_bytes += b.cap;
        *out = b.p;
        return 0;
      }
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess && bytes > want) {   // no room for the headroom: exactly what is asked
      (void)hipGetLastError();
      bytes = want;
      e = hipMalloc(&b.p, bytes);
    }
    if (e != hipSuccess)
      return fail(c, BEM_E_NOMEM, "hipMalloc(%zu) for '%s%s': %s", bytes, L.tag.c_str(), name, hipGetErrorString(e));
    b.cap = bytes;
    c->bytes += bytes;
    L.ws_bytes += bytes;
  }
  *out = b.p;
  return 0;
}

// Frees every workspace of an idle lane (its stream has drained).
static void ws_trim(bem_ctx *c, Lane &L) {
  for (auto &kv : c->ws)
    if (kv.second.p && kv.first.compare(0, L.tag.size(), L.tag) == 0) {
      (void)hipFree(kv.second.p);
      c->bytes -= kv.second.cap;
      kv.second.p = nullptr;
      kv.second.cap = 0;
    }
  L.ws_bytes = 0;
}

static int ev_begin(bem_ctx *c, Lane &L, int kind, double flops, double bytes, EventRec **rec) {
  *rec = nullptr;
  if (!c->profile) return 0;
  if (L.ev_used == L.ev.size()) {
    EventRec r;
    DEV(c, hipEventCreate(&r.a));
    DEV(c, hipEventCreate(&r.b));
    L.ev.push_back(r);
  }
  EventRec &r = L.ev[L.ev_used++];
  r.kind = kind;
  r.flops = flops;
  r.bytes = bytes;
  DEV(c, hipEventRecord(r.a, L.st));
  *rec = &r;
  return 0;
}

static int ev_end(bem_ctx *c, Lane &L, EventRec *r) {
  if (r) DEV(c, hipEventRecord(r->b, L.st));
  return 0;
}

// Accumulate the lane's GEMM event times into the context statistics
// (the lane's stream must have drained).
static void ev_collect(bem_ctx *c, Lane &L) {
  for (size_t i = 0; i < L.ev_used; i++) {
    float ms = 0;
    const EventRec &r = L.ev[i];
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      double *k = c->kst[r.kind];
      k[0] += ms;
      k[1] += 1;
      k[2] += r.flops;
      k[3] += r.bytes;
      if (r.kind == BEM_KS_ZGEMM || r.kind == BEM_KS_ZGEMM_NARROW || r.kind == BEM_KS_ZTRACE) {
        c->gemm_ms += ms;
        c->gemm_n += 1;
        c->gemm_flops += r.flops;
      }
    }
  }
  L.ev_used = 0;
}

static void stats_reset(bem_ctx *c) {
  c->gemm_ms = c->gemm_n = c->gemm_flops = 0;
  std::memset(c->kst, 0, sizeof c->kst);
}

// Shape class of a GEMM for the tile tables: 0 narrow (an output dimension
// of at most 128: panel products), 1 rank update (k <= 128: LU trailing
// updates and blocked triangular solves), 2 general.
static int gemm_class(int m, int n, int k) { return (m <= 128 || n <= 128) ? 0 : (k <= 128 ? 1 : 2); }

// GEMM on the lane's stream (bracketed by events when profiling).
static int gemm(bem_ctx *c, Lane &L, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda,
                const double2 *B, int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  EventRec *r;
  const int cls = gemm_class(m, n, k);
  const int kind = cls == 0 ? BEM_KS_ZGEMM_NARROW : BEM_KS_ZGEMM;
  const bool rd_c = beta.x != 0.0 || beta.y != 0.0;
  CHK(ev_begin(c, L, kind, 8.0 * m * (double)n * k,
               16.0 * ((double)m * k + (double)k * n + (rd_c ? 2.0 : 1.0) * (double)m * n), &r));
  int e = bem_zgemm_tiled(c->ztile[cls], m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "zgemm launch: %s", hipGetErrorString((hipError_t)e));
  return ev_end(c, L, r);
}

// --------------------------------------------------------------- LU / solves
// Panel width of the blocked LU; the diagonal-block inverses are 64 x 64.
static const int LU_NB = 64;

template <typename T = double2>
static size_t inv_bytes(int n) { return sizeof(T) * 2 * 64 * 64 * (size_t)((n + LU_NB - 1) / LU_NB); }

// Super-block of the two-level LU and triangular solves: inside a super-block
// of LU_NB2 rows/columns the 64-wide blocks are processed with rank-64
// updates confined to the super-block; what lies beyond it gets one
// rank-LU_NB2 update per super-block, a GEMM with four times the arithmetic
// intensity of a rank-64 update.
#ifndef BEM_LU_NB2
#define BEM_LU_NB2 256
#endif
static const int LU_NB2 = BEM_LU_NB2;
static_assert(LU_NB2 % LU_NB == 0, "super-block of whole 64-wide blocks");

// dst block (rows x cols, ld) <- tmp (rows x cols, ld rows)
template <typename T>
static int copy_back(bem_ctx *c, Lane &L, T *dst, int64_t ld, const T *tmp, int rows, int cols) {
  DEV(c, hipMemcpy2DAsync(dst, sizeof(T) * ld, tmp, sizeof(T) * rows, sizeof(T) * rows, cols,
                          hipMemcpyDeviceToDevice, L.st));
  return 0;
}

// complex64 GEMM on the lane's stream, timed as BEM_KS_CGEMM.
static int gemm(bem_ctx *c, Lane &L, int m, int n, int k, float2 alpha, const float2 *A, int64_t lda, const float2 *B,
                int64_t ldb, float2 beta, float2 *C, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  EventRec *r;
  const bool rd_c = beta.x != 0.f || beta.y != 0.f;
  CHK(ev_begin(c, L, BEM_KS_CGEMM, 8.0 * m * (double)n * k,
               8.0 * ((double)m * k + (double)k * n + (rd_c ? 2.0 : 1.0) * (double)m * n), &r));
  int e = bem_cgemm_tiled(c->ctile[gemm_class(m, n, k)], m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "cgemm launch: %s", hipGetErrorString((hipError_t)e));
  return ev_end(c, L, r);
}

// In-place product with a 64 x 64 factor M on the lane's stream:
//   side 'R': X (m x n, n <= 64) <- X M;  side 'L': X (m x n, m <= 64) <- M X.
// Safe in place because one thread block covers the whole aliased dimension
// (32 x 64 tiles for 'R', 64 x 32 for 'L' in complex128, 64 x 64 in
// complex64): a block reads every entry it overwrites during its k loop and
// stores only in its epilogue.
static int gemm_inplace(bem_ctx *c, Lane &L, char side, int m, int n, const double2 *M, double2 *X, int64_t ldx) {
  if (m <= 0 || n <= 0) return 0;
  const int k = side == 'R' ? n : m;
  EventRec *r;
  CHK(ev_begin(c, L, BEM_KS_ZGEMM_NARROW, 8.0 * m * (double)n * k, 16.0 * (64.0 * 64 + 2.0 * m * n), &r));
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  int e = side == 'R' ? bem_zgemm_tiled(1034, m, n, k, one, X, ldx, M, 64, zero, X, ldx, L.st)
                      : bem_zgemm_tiled(1033, m, n, k, one, M, 64, X, ldx, zero, X, ldx, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "zgemm launch: %s", hipGetErrorString((hipError_t)e));
  return ev_end(c, L, r);
}

static int gemm_inplace(bem_ctx *c, Lane &L, char side, int m, int n, const float2 *M, float2 *X, int64_t ldx) {
  if (m <= 0 || n <= 0) return 0;
  const int k = side == 'R' ? n : m;
  EventRec *r;
  CHK(ev_begin(c, L, BEM_KS_CGEMM, 8.0 * m * (double)n * k, 8.0 * (64.0 * 64 + 2.0 * m * n), &r));
  const float2 one = make_float2(1, 0), zero = make_float2(0, 0);
  int e = side == 'R' ? bem_cgemm_tiled(1, m, n, k, one, X, ldx, M, 64, zero, X, ldx, L.st)
                      : bem_cgemm_tiled(1, m, n, k, one, M, 64, X, ldx, zero, X, ldx, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "cgemm launch: %s", hipGetErrorString((hipError_t)e));
  return ev_end(c, L, r);
}

// precision-generic launches of the LU building blocks
static int diag_lu_piv(double2 *A, int64_t lda, int n, int k0, double2 *inv, int *ipiv, int *info, hipStream_t st) {
  return bem_diag_lu_piv(A, lda, n, k0, inv, ipiv, info, st);
}
static int diag_lu_piv(float2 *A, int64_t lda, int n, int k0, float2 *inv, int *ipiv, int *info, hipStream_t st) {
  return bem_diag_lu_piv_c(A, lda, n, k0, inv, ipiv, info, st);
}
static int laswp(double2 *A, int64_t lda, int c0, int c1, int s0, int s1, int k0, int k1, const int *ipiv,
                 hipStream_t st) {
  return bem_laswp(A, lda, c0, c1, s0, s1, k0, k1, ipiv, st);
}
static int laswp(float2 *A, int64_t lda, int c0, int c1, int s0, int s1, int k0, int k1, const int *ipiv,
                 hipStream_t st) {
  return bem_laswp_c(A, lda, c0, c1, s0, s1, k0, k1, ipiv, st);
}
static int permute_rows_local(double2 *X, int64_t ldx, int n, int r, const int *ipiv, int *perm, hipStream_t st) {
  return bem_permute_rows_local(X, ldx, n, r, ipiv, perm, st);
}
static int permute_rows_local(float2 *X, int64_t ldx, int n, int r, const int *ipiv, int *perm, hipStream_t st) {
  return bem_permute_rows_local_c(X, ldx, n, r, ipiv, perm, st);
}
static int permute_cols_local(double2 *X, int64_t ldx, int m, int n, const int *ipiv, int *perm, hipStream_t st) {
  return bem_permute_cols_local(X, ldx, m, n, ipiv, perm, st);
}
static int permute_cols_local(float2 *X, int64_t ldx, int m, int n, const int *ipiv, int *perm, hipStream_t st) {
  return bem_permute_cols_local_c(X, ldx, m, n, ipiv, perm, st);
}
static int col_swap(double2 *X, int64_t ldx, int m, int k0, int k1, const int *ipiv, hipStream_t st) {
  return bem_col_swap(X, ldx, m, k0, k1, ipiv, st);
}
static int col_swap(float2 *X, int64_t ldx, int m, int k0, int k1, const int *ipiv, hipStream_t st) {
  return bem_col_swap_c(X, ldx, m, k0, k1, ipiv, st);
}
static int absmax(const double2 *A, int64_t lda, int m, int n, int upper, unsigned long long *out, hipStream_t st) {
  return bem_absmax(A, lda, m, n, upper, out, st);
}
static int absmax(const float2 *A, int64_t lda, int m, int n, int upper, unsigned long long *out, hipStream_t st) {
  return bem_absmax_c(A, lda, m, n, upper, out, st);
}
template <typename T> static T cnum(double re, double im);
template <> double2 cnum<double2>(double re, double im) { return make_double2(re, im); }
template <> float2 cnum<float2>(double re, double im) { return make_float2((float)re, (float)im); }

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

// Two-level blocked LU (no look-ahead): super-blocks of LU_NB2 columns are
// factored 64 columns at a time with the rank-64 updates confined to the
// super-block; the columns right of it then get their U rows by forward
// substitution with the super-block's unit-lower triangle and one
// rank-LU_NB2 trailing update.  Same pivots and factors as the one-level
// form (block-local partial pivoting); growth as there.
template <typename T>
static int lu_factor_block2(bem_ctx *c, Lane &L, T *A, int64_t lda, int n, int *ipiv, int *info, T *inv,
                            unsigned long long *growth) {
  const int nb = LU_NB;
  const T m1 = cnum<T>(-1.0, 0.0), one = cnum<T>(1.0, 0.0);
  int e;
  if (growth && (e = absmax(A, lda, n, n, 0, growth + 1, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  for (int K0 = 0; K0 < n; K0 += LU_NB2) {
    const int K1 = std::min(n, K0 + LU_NB2);
    for (int k0 = K0; k0 < K1; k0 += nb) {
      const int jb = std::min(nb, n - k0);
      const int blk = k0 / nb;
      T *Linv = inv + (size_t)blk * 2 * 64 * 64, *Uinv = Linv + 64 * 64;
      if ((e = diag_lu_piv(A, lda, n, k0, Linv, ipiv, info, L.st)))
        return fail(c, BEM_E_DEVICE, "diag block LU: %d", e);
      if ((e = laswp(A, lda, 0, n, k0, k0 + jb, k0, k0 + jb, ipiv, L.st)))
        return fail(c, BEM_E_DEVICE, "laswp: %d", e);
      const int rest = n - k0 - jb;
      if (rest <= 0) continue;
      T *A21 = A + (int64_t)k0 * lda + k0 + jb;
      CHK(gemm_inplace(c, L, 'R', rest, jb, Uinv, A21, lda));             // L21 = A21 U11^-1
      const int pc = K1 - k0 - jb;                                         // panel columns right of the block
      if (pc > 0) {
        T *A12 = A + (int64_t)(k0 + jb) * lda + k0;
        CHK(gemm_inplace(c, L, 'L', jb, pc, Linv, A12, lda));             // U12 inside the super-block
        CHK(gemm(c, L, rest, pc, jb, m1, A21, lda, A12, lda, one, A12 + jb, lda));
      }
    }
    const int rc = n - K1;
    if (rc <= 0) continue;
    // U rows of the super-block for columns [K1, n): forward substitution
    for (int k0 = K0; k0 < K1; k0 += nb) {
      const int jb = std::min(nb, n - k0);
      const T *Linv = inv + (size_t)(k0 / nb) * 2 * 64 * 64;
      T *Ukr = A + (int64_t)K1 * lda + k0;
      CHK(gemm_inplace(c, L, 'L', jb, rc, Linv, Ukr, lda));
      const int inner = K1 - k0 - jb;
      if (inner > 0)
        CHK(gemm(c, L, inner, rc, jb, m1, A + (int64_t)k0 * lda + k0 + jb, lda, Ukr, lda, one, Ukr + jb, lda));
    }
    // rank-(K1 - K0) trailing update
    CHK(gemm(c, L, n - K1, rc, K1 - K0, m1, A + (int64_t)K0 * lda + K1, lda, A + (int64_t)K1 * lda + K0, lda, one,
             A + (int64_t)K1 * lda + K1, lda));
  }
  if (growth && (e = absmax(A, lda, n, n, 1, growth, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  if (growth && (e = absmax(A, lda, n, n, 2, growth + 2, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  return 0;
}

// Blocked right-looking LU with partial pivoting inside each 64 x 64
// diagonal block (one kernel factors the block, rows in registers, and inverts
// its triangles); the row interchanges are then applied to the other columns and
// the panel's triangular solves and the trailing update run as GEMMs.
// Look-ahead: the trailing update first covers the next 64 columns, then the
// next diagonal block is factored on the lane's auxiliary stream while the
// rest of the trailing update (disjoint columns) runs on the main stream, so
// the latency-bound block factorisation hides behind the large GEMM.
// growth (optional, device) receives max|A| before in growth[1], max|U|
// after in growth[0] and max|L| (strictly lower part) in growth[2].
template <typename T>
static int lu_factor_block(bem_ctx *c, Lane &L, T *A, int64_t lda, int n, int *ipiv, int *info, T *inv,
                           unsigned long long *growth) {
  if (!c->lookahead) return lu_factor_block2(c, L, A, lda, n, ipiv, info, inv, growth);
  const int nb = LU_NB;
  const T m1 = cnum<T>(-1.0, 0.0), one = cnum<T>(1.0, 0.0);
  int e;
  if (growth && (e = absmax(A, lda, n, n, 0, growth + 1, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  if (n > 0 && (e = diag_lu_piv(A, lda, n, 0, inv, ipiv, info, L.st)))
    return fail(c, BEM_E_DEVICE, "diag block LU: %d", e);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    const int blk = k0 / nb;
    T *Linv = inv + (size_t)blk * 2 * 64 * 64, *Uinv = Linv + 64 * 64;
    e = laswp(A, lda, 0, n, k0, k0 + jb, k0, k0 + jb, ipiv, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "laswp: %d", e);
    int rest = n - k0 - jb;
    if (rest > 0) {
      T *Akk = A + (int64_t)k0 * lda + k0;
      T *A21 = Akk + jb, *A12 = A + (int64_t)(k0 + jb) * lda + k0;
      CHK(gemm_inplace(c, L, 'R', rest, jb, Uinv, A21, lda));   // L21 = A21 U11^-1
      CHK(gemm_inplace(c, L, 'L', jb, rest, Linv, A12, lda));   // U12 = L11^-1 A12
      const int nxt = std::min(nb, rest);
      // next block column first, then its diagonal block on the aux stream
      CHK(gemm(c, L, rest, nxt, jb, m1, A21, lda, A12, lda, one, A12 + jb, lda));
      DEV(c, hipEventRecord(L.la_a, L.st));
      DEV(c, hipStreamWaitEvent(L.aux, L.la_a, 0));
      T *Linv2 = inv + (size_t)(blk + 1) * 2 * 64 * 64;
      if ((e = diag_lu_piv(A, lda, n, k0 + jb, Linv2, ipiv, info, L.aux)))
        return fail(c, BEM_E_DEVICE, "diag block LU: %d", e);
      DEV(c, hipEventRecord(L.la_b, L.aux));
      if (rest > nxt)
        CHK(gemm(c, L, rest, rest - nxt, jb, m1, A21, lda, A12 + (int64_t)nxt * lda, lda, one,
                 A12 + (int64_t)nxt * lda + jb, lda));
      DEV(c, hipStreamWaitEvent(L.st, L.la_b, 0));
    }
  }
  if (growth && (e = absmax(A, lda, n, n, 1, growth, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  if (growth && (e = absmax(A, lda, n, n, 2, growth + 2, L.st))) return fail(c, BEM_E_DEVICE, "absmax: %d", e);
  return 0;
}

// X (n x r) <- A^-1 X from the packed LU, its interchanges and the diagonal
// block inverses: forward substitution with the unit-lower L over super-blocks
// top to bottom, then backward substitution with U bottom to top.
template <typename T>
static int lu_solve(bem_ctx *c, Lane &L, const T *A, int64_t lda, int n, const int *ipiv, const T *inv,
                    T *X, int64_t ldx, int r, bool local = false) {
  const int nb = LU_NB;
  const T m1 = cnum<T>(-1.0, 0.0), one = cnum<T>(1.0, 0.0);
  int e;
  if (local) {   // block-local interchanges: one permutation per 64-row block, applied in parallel
    int *perm;
    CHK(ws_get(c, L, "perm", sizeof(int) * (size_t)(n + 64), (void **)&perm));
    e = permute_rows_local(X, ldx, n, r, ipiv, perm, L.st);
  } else {
    e = laswp(X, ldx, 0, r, 0, 0, 0, n, ipiv, L.st);
  }
  if (e) return fail(c, BEM_E_DEVICE, "laswp rhs: %d", e);
  for (int K0 = 0; K0 < n; K0 += LU_NB2) {
    const int K1 = std::min(n, K0 + LU_NB2);
    for (int k0 = K0; k0 < K1; k0 += nb) {
      int jb = std::min(nb, n - k0);
      const T *Linv = inv + (size_t)(k0 / nb) * 2 * 64 * 64;
      CHK(gemm_inplace(c, L, 'L', jb, r, Linv, X + k0, ldx));
      int inner = K1 - k0 - jb;
      if (inner > 0)
        CHK(gemm(c, L, inner, r, jb, m1, A + (int64_t)k0 * lda + k0 + jb, lda, X + k0, ldx, one, X + k0 + jb, ldx));
    }
    if (n > K1)
      CHK(gemm(c, L, n - K1, r, K1 - K0, m1, A + (int64_t)K0 * lda + K1, lda, X + K0, ldx, one, X + K1, ldx));
  }
  const int nsb = (n + LU_NB2 - 1) / LU_NB2;
  for (int sb = nsb - 1; sb >= 0; sb--) {
    const int K0 = sb * LU_NB2, K1 = std::min(n, K0 + LU_NB2);
    for (int k0 = ((K1 - 1) / nb) * nb; k0 >= K0; k0 -= nb) {
      int jb = std::min(nb, n - k0);
      const T *Uinv = inv + (size_t)(k0 / nb) * 2 * 64 * 64 + 64 * 64;
      CHK(gemm_inplace(c, L, 'L', jb, r, Uinv, X + k0, ldx));
      if (k0 > K0)
        CHK(gemm(c, L, k0 - K0, r, jb, m1, A + (int64_t)k0 * lda + K0, lda, X + k0, ldx, one, X + K0, ldx));
    }
    if (K0 > 0) CHK(gemm(c, L, K0, r, K1 - K0, m1, A + (int64_t)K0 * lda, lda, X + K0, ldx, one, X, ldx));
  }
  return 0;
}

// X (r x n) <- X A^-1 from the packed LU of A (A = P L U): X U^-1 over column
// super-blocks left to right, then L^-1 right to left, then the column
// interchanges X P^T; the diagonal blocks go through their inverses, the
// coupling terms through GEMMs (rank-64 inside a super-block, rank-LU_NB2
// beyond it).
template <typename T>
static int lu_solve_right(bem_ctx *c, Lane &L, const T *A, int64_t lda, int n, const int *ipiv,
                          const T *inv, T *X, int64_t ldx, int r, bool local = false) {
  const int nb = LU_NB;
  const T m1 = cnum<T>(-1.0, 0.0), one = cnum<T>(1.0, 0.0);
  for (int K0 = 0; K0 < n; K0 += LU_NB2) {
    const int K1 = std::min(n, K0 + LU_NB2);
    for (int k0 = K0; k0 < K1; k0 += nb) {
      int jb = std::min(nb, n - k0);
      const T *Uinv = inv + (size_t)(k0 / nb) * 2 * 64 * 64 + 64 * 64;
      T *Xk = X + (int64_t)k0 * ldx;
      CHK(gemm_inplace(c, L, 'R', r, jb, Uinv, Xk, ldx));
      int inner = K1 - k0 - jb;
      if (inner > 0)
        CHK(gemm(c, L, r, inner, jb, m1, Xk, ldx, A + (int64_t)(k0 + jb) * lda + k0, lda, one, Xk + (int64_t)jb * ldx,
                 ldx));
    }
    if (n > K1)
      CHK(gemm(c, L, r, n - K1, K1 - K0, m1, X + (int64_t)K0 * ldx, ldx, A + (int64_t)K1 * lda + K0, lda, one,
               X + (int64_t)K1 * ldx, ldx));
  }
  const int nsb = (n + LU_NB2 - 1) / LU_NB2;
  for (int sb = nsb - 1; sb >= 0; sb--) {
    const int K0 = sb * LU_NB2, K1 = std::min(n, K0 + LU_NB2);
    for (int k0 = ((K1 - 1) / nb) * nb; k0 >= K0; k0 -= nb) {
      int jb = std::min(nb, n - k0);
      const T *Linv = inv + (size_t)(k0 / nb) * 2 * 64 * 64;
      T *Xk = X + (int64_t)k0 * ldx;
      CHK(gemm_inplace(c, L, 'R', r, jb, Linv, Xk, ldx));
      if (k0 > K0)
        CHK(gemm(c, L, r, k0 - K0, jb, m1, Xk, ldx, A + k0 + (int64_t)K0 * lda, lda, one, X + (int64_t)K0 * ldx, ldx));
    }
    if (K0 > 0)
      CHK(gemm(c, L, r, K0, K1 - K0, m1, X + (int64_t)K0 * ldx, ldx, A + K0, lda, one, X, ldx));
  }
  int e;
  if (local) {   // block-local interchanges: one permutation per 64-column block
    int *perm;
    CHK(ws_get(c, L, "perm", sizeof(int) * (size_t)(n + 64), (void **)&perm));
    e = permute_cols_local(X, ldx, r, n, ipiv, perm, L.st);
  } else {
    e = col_swap(X, ldx, r, 0, n, ipiv, L.st);
  }
  if (e) return fail(c, BEM_E_DEVICE, "column interchanges: %d", e);
  return 0;
}

// Few-column substitutions (a launch per 64-row block) and the split-K
// skinny product serve up to FEW_COLS columns.
static const int FEW_COLS = 8;
static int trsm_few(const double2 *A, int n, const double2 *inv, double2 *X, int64_t ldx, int r, int lower,
                    double2 *Y, hipStream_t st) {
  return bem_trsm_few(A, n, n, inv, X, ldx, r, lower, Y, n, st);
}
static int trsm_few(const float2 *A, int n, const float2 *inv, float2 *X, int64_t ldx, int r, int lower, float2 *Y,
                    hipStream_t st) {
  return bem_trsm_few_c(A, n, n, inv, X, ldx, r, lower, Y, n, st);
}

// X (n x r) <- A^-1 X for few right-hand sides (r <= FEW_COLS): the
// interchanges, then the forward and backward substitutions with one launch
// per 64-row block each (A packed n x n, leading dimension n), instead of the
// blocked GEMMs; wider systems take lu_solve.
template <typename T>
static int lu_solve_few(bem_ctx *c, Lane &L, const T *A, int n, const int *ipiv, const T *inv, T *X, int64_t ldx,
                        int r, bool local) {
  if (r > FEW_COLS) return lu_solve(c, L, A, n, n, ipiv, inv, X, ldx, r, local);
  int e;
  T *Y;
  CHK(ws_get(c, L, "few_y", sizeof(T) * (size_t)n * FEW_COLS, (void **)&Y));
  if (local) {
    int *perm;
    CHK(ws_get(c, L, "perm", sizeof(int) * (size_t)(n + 64), (void **)&perm));
    e = permute_rows_local(X, ldx, n, r, ipiv, perm, L.st);
  } else {
    e = laswp(X, ldx, 0, r, 0, 0, 0, n, ipiv, L.st);
  }
  if (!e) e = trsm_few(A, n, inv, X, ldx, r, 1, Y, L.st);
  if (!e) e = trsm_few(A, n, inv, X, ldx, r, 0, Y, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "few-column substitution launch: %d", e);
  return 0;
}

// C (m x n) <- alpha A B + beta C for n <= FEW_COLS columns (the probe's
// products) by the split-K skinny kernel; wider products take gemm().
static int gemm_few(bem_ctx *c, Lane &L, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda,
                    const double2 *B, int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
  if (n > FEW_COLS) return gemm(c, L, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (m <= 0 || n <= 0) return 0;
  double2 *part;
  CHK(ws_get(c, L, "skinny_part", sizeof(double2) * bem_zgemm_skinny_partials(m, FEW_COLS, k), (void **)&part));
  EventRec *r;
  const bool rd_c = beta.x != 0.0 || beta.y != 0.0;
  CHK(ev_begin(c, L, BEM_KS_ZGEMM_NARROW, 8.0 * m * (double)n * k,
               16.0 * ((double)m * k + (double)k * n + (rd_c ? 2.0 : 1.0) * (double)m * n), &r));
  int e = bem_zgemm_skinny(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, part, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "skinny zgemm launch: %d", e);
  return ev_end(c, L, r);
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
static int dev_take(bem_ctx *c, size_t bytes, void **p, const char *what);
static void dev_give(bem_ctx *c, void *p, size_t bytes);

static int upload_tree(bem_ctx *c, const bemgeo::Tree &t, DevTree &d, std::vector<std::pair<void *, size_t>> &allocs,
                       size_t &bytes) {
  std::vector<int2> cl(t.close.size());
  for (size_t i = 0; i < cl.size(); i++) cl[i] = make_int2(t.close[i].first, t.close[i].second);
  size_t b1 = sizeof(TriGeom) * t.tg.size(), b2 = sizeof(RwgRec) * t.rw.size(), b3 = sizeof(int2) * (cl.size() + 1);
  void *p1, *p2, *p3;
  CHK(dev_take(c, b1, &p1, "triangle geometry"));
  allocs.emplace_back(p1, b1);
  CHK(dev_take(c, b2, &p2, "RWG records"));
  allocs.emplace_back(p2, b2);
  CHK(dev_take(c, b3, &p3, "close-pair list"));
  allocs.emplace_back(p3, b3);
  bytes += b1 + b2 + b3;
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

// Releases geometry blocks nothing in flight reads (an unpinned pair tree,
// or a body whose lanes have been retired): back to the arena without a
// budget, hipFree'd (after the device drains) with one.
static void free_allocs(bem_ctx *c, std::vector<std::pair<void *, size_t>> &allocs, size_t &bytes) {
  if (c->budget && !allocs.empty()) (void)hipDeviceSynchronize();
  for (auto &a : allocs) dev_give(c, a.first, a.second);
  allocs.clear();
  bytes = 0;
}

static const size_t PAIR_CACHE = 16;   // more than the lanes (at most 8), so one is always unpinned

static int pair_geom(bem_ctx *c, bem_body *b1, bem_body *b2, const double *gap, PairGeom **out) {
  for (auto it = c->pairs.begin(); it != c->pairs.end(); ++it)
    if (it->b1 == b1 && it->b2 == b2 && it->gap[0] == gap[0] && it->gap[1] == gap[1] && it->gap[2] == gap[2]) {
      c->pairs.splice(c->pairs.begin(), c->pairs, it);
      *out = &c->pairs.front();
      return 0;
    }
  if (c->pairs.size() >= PAIR_CACHE) {
    // least recently used tree that no solve in flight reads
    for (auto it = c->pairs.end(); it != c->pairs.begin();) {
      --it;
      if (it->pins == 0) {
        free_allocs(c, it->allocs, it->bytes);
        c->pairs.erase(it);
        break;
      }
    }
  }
  c->pairs.emplace_front();
  PairGeom &g = c->pairs.front();
  g.b1 = b1;
  g.b2 = b2;
  for (int k = 0; k < 3; k++) g.gap[k] = gap[k];
  std::vector<const bemgeo::BodyMesh *> bm = {&b1->mesh, &b2->mesh};
  double shifts[6] = {gap[0], gap[1], gap[2], 0, 0, 0};
  bemgeo::Tree top;
  bemgeo::build_tree(bm, shifts, top);
  g.T1 = top.tri_off.size() > 1 ? top.tri_off[1] : top.T;
  g.E1 = top.rwg_off.size() > 1 ? top.rwg_off[1] : top.E;
  top.close.clear();   // cross blocks never hold close pairs
  int e = upload_tree(c, top, g.top, g.allocs, g.bytes);
  if (e) {
    free_allocs(c, g.allocs, g.bytes);
    c->pairs.pop_front();
    return e;
  }
  *out = &g;
  return 0;
}

// The one-body tree of b (untranslated), uploaded on first use.
static int body_tree(bem_ctx *c, bem_body *b, const DevTree **out) {
  if (!b->up) {
    std::vector<const bemgeo::BodyMesh *> bm = {&b->mesh};
    double shift[3] = {0, 0, 0};
    bemgeo::Tree t;
    bemgeo::build_tree(bm, shift, t);
    CHK(upload_tree(c, t, b->dt, b->allocs, b->bytes));
    b->up = true;
  }
  *out = &b->dt;
  return 0;
}

// Algorithmic work per triangle pair of the reference's assembly, counted
// operation by operation on its restatement (additions, multiplications,
// divisions, and sqrt / exp / log / atan2 / sincos as one each; the
// classification included, the scatter into Z not): 8638 for a 6-point far
// pair, and 1939966 for a close pair on average over the 1280-triangle
// sphere's self block (1.89e6 self, 1.94e6 neighbours).  Each pair writes one
// 36-double tile (288 bytes).
static const double F_FAR_PAIR = 8638.0;
static const double F_CLOSE_PAIR = 1939966.0;
// close pairs whose inner sums share one pass of the moment workspace
// (20 doubles per (pair, outer point): 69 MB at 4096 pairs)
static const size_t CLOSE_CHUNK = 4096;
static const double TILE_BYTES = 8.0 * BEM_TILE_PLANES;

// Triangle-pair tiles of test rows [r0, r1) x source columns [c0, c1) of a
// tree in medium md, on the lane's stream (tile leading dimension c1 - c0).
// The range holds every close pair of the tree (a whole one-body tree) or
// none (a cross-body range of a two-body tree).
// A 36-point queue overflow sets *flag (device int).
static int assemble_tiles(bem_ctx *c, Lane &L, const DevTree &d, const Medium &md, double *tiles, int r0, int r1,
                          int c0, int c1, int *flag) {
  const int T = d.T;
  size_t npairs = (size_t)(r1 - r0) * (size_t)(c1 - c0);
  const bool full = r0 == 0 && r1 == T && c0 == 0 && c1 == T;
  const double nclose = full ? (double)d.nclose : 0.0;
  unsigned int *q36cnt;
  int2 *q36;
  unsigned int cap = (unsigned int)std::min<size_t>(std::max<size_t>(npairs, 1), (size_t)1 << 24);
  CHK(ws_get(c, L, "q36_count", sizeof(unsigned int), (void **)&q36cnt));
  CHK(ws_get(c, L, "q36", sizeof(int2) * (size_t)cap, (void **)&q36));
  DEV(c, hipMemsetAsync(q36cnt, 0, sizeof(unsigned int), L.st));
  EventRec *r;
  const double nfar = (double)npairs - nclose;
  CHK(ev_begin(c, L, BEM_KS_FAR, nfar * F_FAR_PAIR, nfar * TILE_BYTES, &r));
  int e = bem_launch_far_g6(d.tg, T, &md, r0, r1, c0, c1, tiles, npairs, q36, q36cnt, cap, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "far_g6 launch: %s", hipGetErrorString((hipError_t)e));
  CHK(ev_end(c, L, r));
  e = bem_launch_far_g36(d.tg, T, &md, q36, q36cnt, cap, r0, r1, c0, c1, tiles, npairs, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "far_g36 launch: %s", hipGetErrorString((hipError_t)e));
  if (nclose > 0) {
    // inner sums of at most CLOSE_CHUNK pairs at a time in a lane workspace
    const unsigned int chunk = (unsigned int)std::min<size_t>((size_t)d.nclose, CLOSE_CHUNK);
    double *mom;
    CHK(ws_get(c, L, "close_mom", sizeof(double) * bem_close_mom_doubles(chunk), (void **)&mom));
    CHK(ev_begin(c, L, BEM_KS_CLOSE, nclose * F_CLOSE_PAIR, nclose * TILE_BYTES, &r));
    e = bem_launch_close(d.tg, T, &md, d.close, d.nclose, r0, r1, c0, c1, tiles, npairs, mom, chunk, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "close launch: %s", hipGetErrorString((hipError_t)e));
    CHK(ev_end(c, L, r));
  }
  if (flag && (e = bem_launch_q36_check(q36cnt, cap, flag, L.st)))
    return fail(c, BEM_E_DEVICE, "queue check launch: %s", hipGetErrorString((hipError_t)e));
  return 0;
}

// Gather of the RWG blocks (see bem_launch_gather), timed as BEM_KS_GATHER:
// every tile of the range is read once and each RWG pair's 2 x 2 block is
// written (read too when added, and written twice with a copy).
static int gather(bem_ctx *c, Lane &L, const double *tiles, size_t npairs, int tld, int row_lo, int col_lo,
                  const RwgRec *rw, int r_lo, int r_hi, int c_lo, int c_hi, const Medium *md, double2 *Z, int64_t ldz,
                  int zr0, int zc0, int add, double2 *Zcopy, int64_t ldc, int cr0, int cc0) {
  EventRec *r;
  double nrwg = (double)(r_hi - r_lo) * (double)(c_hi - c_lo);
  CHK(ev_begin(c, L, BEM_KS_GATHER, 0.0, npairs * TILE_BYTES + nrwg * 64.0 * (1 + (add ? 1 : 0) + (Zcopy ? 1 : 0)),
               &r));
  int e = bem_launch_gather(tiles, npairs, tld, row_lo, col_lo, rw, r_lo, r_hi, c_lo, c_hi, md, Z, ldz, zr0, zc0, add,
                            Zcopy, ldc, cr0, cc0, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "gather launch: %s", hipGetErrorString((hipError_t)e));
  return ev_end(c, L, r);
}

static double theta(double omega, double T) {
  // src/HeatTransfer.cpp:3-6 (T = 0 gives exp(inf) and exactly 0)
  return SI_HBAR * omega / (std::exp(SI_HBAR * omega / (SI_KB * T)) - 1);
}

// ------------------------------------------------------------- self blocks
// Bytes of a self entry of an n x n block in role 0 (factored) or 1 (raw).
static size_t self_entry_bytes(int n, int role, int prec) {
  size_t nn = sizeof(double2) * (size_t)n * n;
  size_t b = 2 * nn + 256;
  if (role == 0) b += sizeof(int) * (size_t)(n + 64);
  if (prec == BEM_PREC_MIXED) b += nn + (role == 0 ? inv_bytes<float2>(n) : 0);
  else if (role == 0) b += inv_bytes(n) + nn;   // inverses and the unfactored copy
  return b;
}

static size_t arena_round(size_t bytes) { return (std::max<size_t>(bytes, 1) + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

// Best fit over the segments' free lists (the smallest free block that
// holds the request, so small blocks fill leftovers and large holes stay
// whole); nullptr when no block is large enough.
static void *arena_take(bem_ctx *c, size_t bytes) {
  bytes = arena_round(bytes);
  ArenaSeg *best = nullptr;
  std::map<size_t, size_t>::iterator bit;
  for (auto &sg : c->arena)
    for (auto it = sg.free.begin(); it != sg.free.end(); ++it)
      if (it->second >= bytes && (!best || it->second < bit->second)) {
        best = &sg;
        bit = it;
      }
  if (!best) return nullptr;
  const size_t off = bit->first, len = bit->second;
  best->free.erase(bit);
  if (len > bytes) best->free.emplace(off + bytes, len - bytes);
  return best->base + off;
}

// Returns a block to its segment, merged with free neighbours.
static void arena_give(bem_ctx *c, void *p, size_t bytes) {
  bytes = arena_round(bytes);
  char *q = (char *)p;
  for (auto &sg : c->arena) {
    if (q < sg.base || q >= sg.base + sg.size) continue;
    size_t off = (size_t)(q - sg.base);
    auto nx = sg.free.lower_bound(off);
    if (nx != sg.free.end() && nx->first == off + bytes) {
      bytes += nx->second;
      nx = sg.free.erase(nx);
    }
    if (nx != sg.free.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += bytes;
        return;
      }
    }
    sg.free.emplace(off, bytes);
    return;
  }
}

// Segments with no block in use go back to the runtime.  Called when a body
// is freed (which already waits for the lanes and frees the body's device
// memory), so that blocks of bodies that are gone do not stay held.
static void arena_release_empty(bem_ctx *c) {
  for (auto it = c->arena.begin(); it != c->arena.end();) {
    const bool empty = it->free.size() == 1 && it->free.begin()->first == 0 && it->free.begin()->second == it->size;
    if (empty) {
      (void)hipFree(it->base);
      c->bytes -= it->size;
      c->arena_bytes -= it->size;
      it = c->arena.erase(it);
    } else {
      ++it;
    }
  }
}

// A new segment that holds at least `bytes` (ARENA_SEG, or exactly the
// request when the device cannot give a whole segment).
static int arena_grow(bem_ctx *c, size_t bytes) {
  size_t sz = std::max(ARENA_SEG, arena_round(bytes));
  void *p = nullptr;
  hipError_t r = hipMalloc(&p, sz);
  if (r != hipSuccess && sz > arena_round(bytes)) {
    (void)hipGetLastError();
    sz = arena_round(bytes);
    r = hipMalloc(&p, sz);
  }
  if (r != hipSuccess) return fail(c, BEM_E_NOMEM, "hipMalloc(%zu) for the self-block arena: %s", sz, hipGetErrorString(r));
  ArenaSeg sg;
  sg.base = (char *)p;
  sg.size = sz;
  sg.free.emplace(0, sz);
  c->arena.push_back(std::move(sg));
  c->arena_bytes += sz;
  c->bytes += sz;
  return 0;
}

// Releases an entry's memory: to the arena without a budget, to the runtime
// with one.  Its pinned status slot and event are kept for the next entry
// (pinned-memory frees synchronise the device like hipFree).  Only unpinned
// entries come here, and a lane unpins its entries after its stream has
// drained, so no kernel still reads the blocks.
static void self_free(bem_ctx *c, SelfEntry &e) {
  if (e.ready) (void)hipEventSynchronize(e.ready);
  for (auto &a : e.allocs) dev_give(c, a.first, a.second);
  e.allocs.clear();
  if (e.h) c->spare_status.push_back(e.h);
  e.h = nullptr;
  if (e.ready) c->spare_events.push_back(e.ready);
  e.ready = nullptr;
  c->self_bytes -= e.bytes;
  e.bytes = 0;
}

// The least recently used unpinned entry, or end().
static std::list<SelfEntry>::iterator self_lru(bem_ctx *c) {
  auto victim = c->selfs.end();
  for (auto it = c->selfs.begin(); it != c->selfs.end(); ++it)
    if (it->pins == 0 && (victim == c->selfs.end() || it->stamp < victim->stamp)) victim = it;
  return victim;
}

// Drop least recently used unpinned entries until `need` more bytes fit.
static void self_evict(bem_ctx *c, size_t need) {
  while (c->self_bytes + need > c->self_cap) {
    auto victim = self_lru(c);
    if (victim == c->selfs.end()) return;   // everything in use: the cache runs over its cap
    self_free(c, *victim);
    c->selfs.erase(victim);
  }
}

// Device memory for geometry and self entries.  Without a budget: a free
// block of the arena, else a new segment while the arena is under the
// cache's cap, else room made by evicting cached entries (every entry
// pinned: a segment beyond the cap).  With a budget: hipMalloc, counted
// against it.
static int dev_take(bem_ctx *c, size_t bytes, void **p, const char *what) {
  if (c->budget) {
    if (c->bytes + bytes > c->budget)
      return fail(c, BEM_E_NOMEM, "%s (%zu bytes) exceeds the HBM budget", what, bytes);
    hipError_t r = hipMalloc(p, bytes ? bytes : 1);
    if (r != hipSuccess) return fail(c, BEM_E_NOMEM, "hipMalloc(%zu) for %s: %s", bytes, what, hipGetErrorString(r));
    c->bytes += bytes;
    return 0;
  }
  *p = arena_take(c, bytes);
  while (!*p) {
    if (c->arena_bytes + std::max(ARENA_SEG, arena_round(bytes)) <= c->self_cap) {
      CHK(arena_grow(c, bytes));
    } else {
      auto victim = self_lru(c);
      if (victim == c->selfs.end()) {
        CHK(arena_grow(c, bytes));
      } else {
        self_free(c, *victim);
        c->selfs.erase(victim);
      }
    }
    *p = arena_take(c, bytes);
  }
  return 0;
}

static void dev_give(bem_ctx *c, void *p, size_t bytes) {
  if (c->budget) {
    (void)hipFree(p);
    c->bytes -= bytes;
  } else {
    arena_give(c, p, bytes);
  }
}

static int self_alloc(bem_ctx *c, SelfEntry &e, void **p, size_t bytes) {
  CHK(dev_take(c, bytes, p, "self-block cache entry"));
  e.allocs.emplace_back(*p, bytes);
  e.bytes += bytes;
  c->self_bytes += bytes;
  return 0;
}

// Two bodies with identical vertex and triangle arrays have identical self
// blocks (they are assembled untranslated).
static bool same_mesh(const bem_body *a, const bem_body *b) {
  return a->fp == b->fp && a->mesh.xyz == b->mesh.xyz && a->mesh.tri == b->mesh.tri;
}

// Self entry of body b in medium (eps, mu) at omega for role 0 (factored) or
// 1 (raw), pinned for the lane; produced on the lane's stream when missing.
// twin (optional): a raw (role 1) entry pinned by the caller for a body with
// the same mesh, medium, omega and precision; a missing entry is then copied
// from it instead of assembled.
static int self_get(bem_ctx *c, Lane &L, bem_body *b, const double *eps, const double *mu, double omega, int role,
                    int pivot, int prec, const SelfEntry *twin, SelfEntry **out) {
  const double key[5] = {eps[0], eps[1], mu[0], mu[1], omega};
  for (auto &e : c->selfs)
    if (e.body == b && e.role == role && e.pivot == pivot && e.prec == prec &&
        std::memcmp(e.key, key, sizeof key) == 0) {
      e.pins++;
      e.stamp = ++c->clock;
      DEV(c, hipStreamWaitEvent(L.st, e.ready, 0));
      c->self_hits++;
      *out = &e;
      return 0;
    }
  c->self_misses++;
  const DevTree *d;
  CHK(body_tree(c, b, &d));
  const int T = d->T, n = 2 * d->E;
  size_t nn = sizeof(double2) * (size_t)n * n;
  size_t need = self_entry_bytes(n, role, prec);
  self_evict(c, need);
  c->selfs.emplace_front();
  SelfEntry &e = c->selfs.front();
  e.body = b;
  std::memcpy(e.key, key, sizeof key);
  e.role = role;
  e.pivot = pivot;
  e.prec = prec;
  e.n = n;
  e.pins = 1;
  e.stamp = ++c->clock;
  auto undo = [&](int code) {
    self_free(c, e);
    c->selfs.pop_front();
    return code;
  };
  int r = self_alloc(c, e, (void **)&e.Z, nn);
  if (!r) r = self_alloc(c, e, (void **)&e.Gs, nn);
  if (!r) r = self_alloc(c, e, (void **)&e.dinfo, sizeof(int) * 4 + sizeof(unsigned long long) * 3);
  const bool mixed = prec == BEM_PREC_MIXED;
  if (!r && role == 0 && !mixed) r = self_alloc(c, e, (void **)&e.inv, inv_bytes(n));
  if (!r && role == 0 && !mixed) r = self_alloc(c, e, (void **)&e.Araw, nn);
  if (!e.Araw) e.Araw = e.Z;
  if (!r && role == 0) r = self_alloc(c, e, (void **)&e.ipiv, sizeof(int) * (size_t)(n + 64));
  if (!r && mixed) r = self_alloc(c, e, (void **)&e.Z32, nn / 2);
  if (!r && mixed) r = self_alloc(c, e, (void **)&e.Gs32, nn / 2);
  if (!r && mixed && role == 0) r = self_alloc(c, e, (void **)&e.inv32, inv_bytes<float2>(n));
  if (!r && !c->spare_status.empty()) {
    e.h = c->spare_status.back();
    c->spare_status.pop_back();
  } else if (!r && hipHostMalloc((void **)&e.h, sizeof(SelfStatus), hipHostMallocDefault) != hipSuccess) {
    r = fail(c, BEM_E_NOMEM, "pinned status slot");
  }
  if (!r && !c->spare_events.empty()) {
    e.ready = c->spare_events.back();
    c->spare_events.pop_back();
  } else if (!r && hipEventCreateWithFlags(&e.ready, hipEventDisableTiming) != hipSuccess) {
    r = fail(c, BEM_E_DEVICE, "event creation");
  }
  if (r) return undo(r);
  e.dgrowth = (unsigned long long *)(e.dinfo + 4);
  // Z_ext(b, b) (vacuum) then + Z_int(b), whose copy is G
  Medium mext = make_medium(1.0, 0.0, 1.0, 0.0, omega);
  Medium mint = make_medium(eps[0], eps[1], mu[0], mu[1], omega);
  double *tiles;
  size_t np = (size_t)T * T;
  r = ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * np, (void **)&tiles);
  if (!r && hipMemsetAsync(e.dinfo, 0, sizeof(int) * 4 + sizeof(unsigned long long) * 3, L.st) != hipSuccess)
    r = fail(c, BEM_E_DEVICE, "memset");
  if (twin) {
    // same geometry, medium and omega: the blocks are bitwise those of twin
    if (!r && hipStreamWaitEvent(L.st, twin->ready, 0) != hipSuccess) r = fail(c, BEM_E_DEVICE, "event wait");
    if (!r && (hipMemcpyAsync(e.Z, twin->Z, nn, hipMemcpyDeviceToDevice, L.st) != hipSuccess ||
               hipMemcpyAsync(e.Gs, twin->Gs, nn, hipMemcpyDeviceToDevice, L.st) != hipSuccess ||
               (e.Z32 && hipMemcpyAsync(e.Z32, twin->Z32, nn / 2, hipMemcpyDeviceToDevice, L.st) != hipSuccess) ||
               (e.Gs32 && hipMemcpyAsync(e.Gs32, twin->Gs32, nn / 2, hipMemcpyDeviceToDevice, L.st) != hipSuccess)))
      r = fail(c, BEM_E_DEVICE, "self block copy");
  } else {
    if (!r) r = assemble_tiles(c, L, *d, mext, tiles, 0, T, 0, T, e.dinfo + 1);
    if (!r) r = gather(c, L, tiles, np, T, 0, 0, d->rw, 0, d->E, 0, d->E, &mext, e.Z, n, 0, 0, 0, nullptr, 0, 0, 0);
    if (!r) r = assemble_tiles(c, L, *d, mint, tiles, 0, T, 0, T, e.dinfo + 1);
    if (!r) r = gather(c, L, tiles, np, T, 0, 0, d->rw, 0, d->E, 0, d->E, &mint, e.Z, n, 0, 0, 1, e.Gs, n, 0, 0);
    if (!r && bem_symm(e.Gs, n, n, L.st)) r = fail(c, BEM_E_DEVICE, "symm launch");
  }
  if (!r && mixed) {
    if (!twin && (bem_z2c(e.Z, n, e.Z32, n, n, n, L.st) || bem_z2c(e.Gs, n, e.Gs32, n, n, n, L.st)))
      r = fail(c, BEM_E_DEVICE, "complex64 copy launch");
    if (!r && role == 0) r = lu_factor_block(c, L, e.Z32, n, n, e.ipiv, e.dinfo, e.inv32, e.dgrowth);
  } else if (!r && role == 0) {
    if (hipMemcpyAsync(e.Araw, e.Z, nn, hipMemcpyDeviceToDevice, L.st) != hipSuccess)
      r = fail(c, BEM_E_DEVICE, "self block copy");
    else if (pivot) r = lu_factor(c, L, e.Z, n, n, e.ipiv, e.dinfo, e.inv);
    else r = lu_factor_block(c, L, e.Z, n, n, e.ipiv, e.dinfo, e.inv, e.dgrowth);
  }
  if (!r && hipMemcpyAsync(e.h, e.dinfo, sizeof(SelfStatus), hipMemcpyDeviceToHost, L.st) != hipSuccess)
    r = fail(c, BEM_E_DEVICE, "status copy");
  if (!r && hipEventRecord(e.ready, L.st) != hipSuccess) r = fail(c, BEM_E_DEVICE, "event record");
  if (r) {
    (void)hipStreamSynchronize(L.st);
    return undo(r);
  }
  *out = &e;
  return 0;
}

static void self_release(bem_ctx *c, Lane &L) {
  for (SelfEntry **p : {&L.e1, &L.e2})
    if (*p) {
      (*p)->pins--;
      *p = nullptr;
    }
  if (L.pg) {
    L.pg->pins--;
    L.pg = nullptr;
  }
  (void)c;
}

// Drops every cached entry and tree of body b (after the device drained).
static void self_drop_body(bem_ctx *c, const void *b) {
  for (auto it = c->selfs.begin(); it != c->selfs.end();) {
    if (it->body == b) {
      self_free(c, *it);
      it = c->selfs.erase(it);
    } else {
      ++it;
    }
  }
}

// ------------------------------------------------------------- one solve
// Cross blocks of Z_ext for the pair's gap: B = Z_ext(1,2) (m1 x m2) and
// C = Z_ext(2,1) (m2 x m1), complex128.  Braw (optional) receives a second
// copy of B for the residuals of the error probe and of refinement, since the
// FP64 elimination overwrites B.
static int assemble_cross(bem_ctx *c, Lane &L, PairGeom &g, const Medium &mext, double2 *Bm, double2 *Cm,
                          double2 *Braw, int *flag) {
  const int T1 = g.T1, T2 = g.top.T - g.T1, Eb1 = g.E1, E = g.top.E;
  const int m1 = 2 * Eb1, m2 = 2 * (E - Eb1);
  double *tiles;
  size_t npc = (size_t)T1 * T2;
  CHK(ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * std::max<size_t>(npc, 1), (void **)&tiles));
  CHK(assemble_tiles(c, L, g.top, mext, tiles, 0, T1, T1, T1 + T2, flag));
  CHK(gather(c, L, tiles, npc, T2, 0, T1, g.top.rw, 0, Eb1, Eb1, E, &mext, Bm, m1, 0, 0, 0, Braw, m1, 0, 0));
  // C = Z_ext(2,1) = S B^T S (reciprocity of the far-pair rules, S = +1 on J
  // and -1 on M unknowns) instead of a second assembly
  int e = bem_signed_transpose(Bm, m1, m1, m2, Cm, m2, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "transpose launch: %d", e);
  return 0;
}

// Factors of a square block for the triangular solves: packed LU, its
// interchanges, the diagonal-block inverses, and whether every interchange
// stays inside its 64-block (block-local pivoting).
template <typename T>
struct LuRef {
  const T *lu;
  int n;
  const int *ipiv;
  const T *inv;
  bool local;
};

// FP64 <-> working-precision copies of an m x n block (leading dimension m):
// a plain copy in complex128 (none when src and dst coincide), rounding to
// complex64 and widening back in the mixed path.
static int to_work(bem_ctx *c, Lane &L, const double2 *src, double2 *dst, int m, int n) {
  if (src != dst) DEV(c, hipMemcpyAsync(dst, src, sizeof(double2) * (size_t)m * n, hipMemcpyDeviceToDevice, L.st));
  return 0;
}
static int to_work(bem_ctx *c, Lane &L, const double2 *src, float2 *dst, int m, int n) {
  if (bem_z2c(src, m, dst, m, m, n, L.st)) return fail(c, BEM_E_DEVICE, "complex64 copy launch");
  return 0;
}
static int from_work(bem_ctx *c, Lane &L, const double2 *src, double2 *dst, int m, int n) {
  return to_work(c, L, src, dst, m, n);
}
static int from_work(bem_ctx *c, Lane &L, const float2 *src, double2 *dst, int m, int n) {
  if (bem_c2z(src, m, dst, m, m, n, L.st)) return fail(c, BEM_E_DEVICE, "complex128 copy launch");
  return 0;
}
// dst (FP64) += src (working precision)
static int add_work(bem_ctx *c, Lane &L, const double2 *src, double2 *dst, int m, int n) {
  if (bem_zaxpy(make_double2(1, 0), src, m, dst, m, m, n, L.st)) return fail(c, BEM_E_DEVICE, "update launch");
  return 0;
}
static int add_work(bem_ctx *c, Lane &L, const float2 *src, double2 *dst, int m, int n) {
  if (bem_c2z_add(src, m, dst, m, m, n, L.st)) return fail(c, BEM_E_DEVICE, "update launch");
  return 0;
}

// Seed of the probe's columns (fixed: every solve of a request, alone or in a
// mix, draws the same ones).
static const unsigned long long PROBE_SEED = 0x5EED0B5Eull;

// Error probe of a solution held as Xn = -X (m1 x m2, FP64) and, when X2 is
// given, Z2 = Sc^-1 (m2 x m2, FP64; otherwise Sc^-1 w comes from the factors
// of Sc).  w: BEM_PROBE_COLS seeded columns with complex64 values (exact in
// both precisions); x = [X w; Z2 w]; r = [0; w] - W x in FP64 against the
// unfactored blocks; d_1 = [W^-1 r]_1 through the factors of A in the
// working precision T and the solve's X.  The top rows of W^-1 [0; w] are
// X_exact w, so d_1 = dX w with dX = X_exact - X, up to the accuracy of the
// correction.
// Re S = 1/4 tr(sym G2 X^H sym G1 X) changes to first order by 1/2 Re t,
// t = tr(F dX), F = sym G2 X^H sym G1; with E[w w^H] = s^2 I,
// E[w^H F dX w] = s^2 t, and w^H F = g^H for g = sym G1 X sym G2 w, so
// h_j = g_j^H d_1j are the terms of a trace estimate of t.
// out: max|d_1|, max|x_1|, ||r||^2, ||w||^2, h_j.
template <typename T>
static int probe(bem_ctx *c, Lane &L, const LuRef<T> &A, const LuRef<T> &S, const double2 *Araw,
                 const double2 *Braw, const double2 *Cm, const double2 *Draw, const double2 *Gs1, const double2 *Gs2,
                 const double2 *Xn, const double2 *X2, ProbeOut *out) {
  const int m1 = A.n, m2 = S.n, k = BEM_PROBE_COLS;
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  float2 *w32;
  double2 *w, *x1, *x2, *r1, *r2, *t1, *t2;
  T *a, *b;
  double *part;
  CHK(ws_get(c, L, "pr_w32", sizeof(float2) * (size_t)m2 * k, (void **)&w32));
  CHK(ws_get(c, L, "pr_w", sizeof(double2) * (size_t)m2 * k, (void **)&w));
  CHK(ws_get(c, L, "pr_x1", sizeof(double2) * (size_t)m1 * k, (void **)&x1));
  CHK(ws_get(c, L, "pr_x2", sizeof(double2) * (size_t)m2 * k, (void **)&x2));
  CHK(ws_get(c, L, "pr_r1", sizeof(double2) * (size_t)m1 * k, (void **)&r1));
  CHK(ws_get(c, L, "pr_r2", sizeof(double2) * (size_t)m2 * k, (void **)&r2));
  CHK(ws_get(c, L, "pr_t1", sizeof(double2) * (size_t)m1 * k, (void **)&t1));
  CHK(ws_get(c, L, "pr_t2", sizeof(double2) * (size_t)m2 * k, (void **)&t2));
  CHK(ws_get(c, L, "pr_a", sizeof(T) * (size_t)m1 * k, (void **)&a));
  CHK(ws_get(c, L, "pr_b", sizeof(T) * (size_t)m2 * k, (void **)&b));
  CHK(ws_get(c, L, "pr_part", sizeof(double) * std::max(bem_sumsq_partials(m1, k), bem_sumsq_partials(m2, k)),
             (void **)&part));
  {   // workspaces at their largest, so none grows between the probe's launches
    T *fy;
    CHK(ws_get(c, L, "few_y", sizeof(T) * (size_t)std::max(m1, m2) * FEW_COLS, (void **)&fy));
    double2 *sp;
    size_t np = 0;
    for (int mm : {m1, m2})
      for (int kk : {m1, m2}) np = std::max(np, bem_zgemm_skinny_partials(mm, FEW_COLS, kk));
    CHK(ws_get(c, L, "skinny_part", sizeof(double2) * np, (void **)&sp));
  }
  const size_t wb2 = sizeof(double2) * (size_t)m2 * k;
  if (bem_fill_hash(w32, sizeof(float2) * (size_t)m2 * k, 1, PROBE_SEED, L.st) ||
      bem_c2z(w32, m2, w, m2, m2, k, L.st))
    return fail(c, BEM_E_DEVICE, "probe fill launch");
  DEV(c, hipMemsetAsync(out, 0, 2 * sizeof(unsigned long long), L.st));
  // x = [X w; Z2 w]
  if (X2) {
    CHK(gemm_few(c, L, m2, k, m2, one, X2, m2, w, m2, zero, x2, m2));
  } else {
    CHK(to_work(c, L, w, b, m2, k));
    CHK(lu_solve_few(c, L, S.lu, m2, S.ipiv, S.inv, b, m2, k, S.local));
    CHK(from_work(c, L, b, x2, m2, k));
  }
  CHK(gemm_few(c, L, m1, k, m2, mone, Xn, m1, w, m2, zero, x1, m1));
  // r1 = -(A x1 + B x2), r2 = w - C x1 - D x2
  CHK(gemm_few(c, L, m1, k, m1, mone, Araw, m1, x1, m1, zero, r1, m1));
  CHK(gemm_few(c, L, m1, k, m2, mone, Braw, m1, x2, m2, one, r1, m1));
  DEV(c, hipMemcpyAsync(r2, w, wb2, hipMemcpyDeviceToDevice, L.st));
  CHK(gemm_few(c, L, m2, k, m1, mone, Cm, m2, x1, m1, one, r2, m2));
  CHK(gemm_few(c, L, m2, k, m2, mone, Draw, m2, x2, m2, one, r2, m2));
  if (bem_sumsq(r1, m1, m1, k, part, 0, &out->rr, L.st) || bem_sumsq(r2, m2, m2, k, part, 1, &out->rr, L.st) ||
      bem_sumsq(w, m2, m2, k, part, 0, &out->ww, L.st))
    return fail(c, BEM_E_DEVICE, "probe norm launch");
  // d_1 = [W^-1 r]_1 = u + X (r2 - C u) with u = A^-1 r1: the top block row
  // of W^-1 is [A^-1 - X C A^-1, X], and the X of the solve stands in for the
  // exact one (the difference times r2 - C u is of second order)
  CHK(to_work(c, L, r1, a, m1, k));
  CHK(lu_solve_few(c, L, A.lu, m1, A.ipiv, A.inv, a, m1, k, A.local));
  CHK(from_work(c, L, a, t1, m1, k));                                // u
  DEV(c, hipMemcpyAsync(t2, r2, wb2, hipMemcpyDeviceToDevice, L.st));
  CHK(gemm_few(c, L, m2, k, m1, mone, Cm, m2, t1, m1, one, t2, m2));    // r2 - C u
  CHK(gemm_few(c, L, m1, k, m2, mone, Xn, m1, t2, m2, one, t1, m1));    // d1 = u - Xn (r2 - C u)
  int e = bem_absmax(t1, m1, m1, k, 0, &out->dmax, L.st);
  if (!e) e = bem_absmax(x1, m1, m1, k, 0, &out->xmax, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "probe absmax launch: %d", e);
  // h_j = g_j^H d_1j with g = sym G1 X sym G2 w = -sym G1 Xn sym G2 w
  CHK(gemm_few(c, L, m2, k, m2, one, Gs2, m2, w, m2, zero, r2, m2));
  CHK(gemm_few(c, L, m1, k, m2, one, Xn, m1, r2, m2, zero, r1, m1));
  CHK(gemm_few(c, L, m1, k, m1, mone, Gs1, m1, r1, m1, zero, x1, m1));
  if ((e = bem_zdotc_cols(x1, m1, t1, m1, m1, k, out->h, L.st))) return fail(c, BEM_E_DEVICE, "probe dot launch: %d", e);
  return 0;
}

// Starts the probe of a solve stage.  When look-ahead gave the lane a second
// stream (the lanes fit the hardware queues twice), the probe runs there,
// overlapping the products and trace, which read the same X; probe_join()
// then makes the lane's stream wait for it.  Otherwise it runs in line.  Every
// workspace the probe uses is sized before its first launch, so none is
// reallocated while its kernels run on the other stream.
template <typename T>
static int probe_start(bem_ctx *c, Lane &L, const LuRef<T> &A, const LuRef<T> &S, const double2 *Araw,
                       const double2 *Braw, const double2 *Cm, const double2 *Draw, const double2 *Gs1,
                       const double2 *Gs2, const double2 *Xn, ProbeOut *out) {
  if (!c->lookahead) return probe<T>(c, L, A, S, Araw, Braw, Cm, Draw, Gs1, Gs2, Xn, nullptr, out);
  DEV(c, hipEventRecord(L.la_a, L.st));
  DEV(c, hipStreamWaitEvent(L.aux, L.la_a, 0));
  std::swap(L.st, L.aux);
  int e = probe<T>(c, L, A, S, Araw, Braw, Cm, Draw, Gs1, Gs2, Xn, nullptr, out);
  std::swap(L.st, L.aux);
  if (e) {
    (void)hipStreamSynchronize(L.aux);
    return e;
  }
  DEV(c, hipEventRecord(L.la_b, L.aux));
  L.probe_async = true;
  return 0;
}

static int probe_join(bem_ctx *c, Lane &L) {
  if (!L.probe_async) return 0;
  L.probe_async = false;
  DEV(c, hipStreamWaitEvent(L.st, L.la_b, 0));
  return 0;
}

// One step of iterative refinement of [X; Z2] = W^-1 [0; I]: the full
// residual R = [0; I] - W [X; Z2] in FP64 against the unfactored blocks (its
// squared Frobenius norm to *rfull), the correction [dX; dZ2] = W^-1 R through
// the factors in the working precision T, with Y = A^-1 B kept from the
// elimination:  u = A^-1 T1 (T1 = -R1),  dZ2 = Sc^-1 (R2 + C u),
// -dX = u + Y dZ2;  then the FP64 products and trace of the corrected X and
// the probe.  Xn holds -X and X2 holds Z2 (FP64); `first` forms Z2 from the
// factors of Sc before the residual.
template <typename T>
static int refine_step(bem_ctx *c, Lane &L, const SelfEntry &E1, const SelfEntry &E2, const LuRef<T> &A,
                       const LuRef<T> &S, const T *Ct, const T *Yt, const double2 *Braw, const double2 *Cm,
                       double2 *Xn, double2 *X2, bool first, double2 *trace_out, ProbeOut *pout, double *rfull) {
  const int m1 = A.n, m2 = S.n;
  const double2 d1 = make_double2(1, 0), dm1 = make_double2(-1, 0), d0 = make_double2(0, 0);
  const T one = cnum<T>(1, 0);
  double2 *R1, *R2, *B1, *partial;
  T *R1t, *R2t;
  double *part;
  CHK(ws_get(c, L, "R1_64", sizeof(double2) * (size_t)m1 * m2, (void **)&R1));
  CHK(ws_get(c, L, "R2_64", sizeof(double2) * (size_t)m2 * m2, (void **)&R2));
  if (std::is_same<T, double2>::value) {
    R1t = (T *)R1;
    R2t = (T *)R2;
  } else {
    CHK(ws_get(c, L, "R1_w", sizeof(T) * (size_t)m1 * m2, (void **)&R1t));
    CHK(ws_get(c, L, "R2_w", sizeof(T) * (size_t)m2 * m2, (void **)&R2t));
  }
  CHK(ws_get(c, L, "B1", sizeof(double2) * (size_t)m1 * m2, (void **)&B1));
  CHK(ws_get(c, L, "trace_partial", sizeof(double2) * (bem_htrace_partials(m2) + 1), (void **)&partial));
  CHK(ws_get(c, L, "rf_part", sizeof(double) * std::max(bem_sumsq_partials(m1, m2), bem_sumsq_partials(m2, m2)),
             (void **)&part));
  int e;
  if (first) {   // Z2 = Sc^-1: the identity through the factors
    DEV(c, hipMemsetAsync(R2, 0, sizeof(double2) * (size_t)m2 * m2, L.st));
    if ((e = bem_identity_minus(R2, m2, m2, L.st))) return fail(c, BEM_E_DEVICE, "identity launch: %d", e);
    CHK(to_work(c, L, R2, R2t, m2, m2));
    CHK(lu_solve(c, L, S.lu, m2, m2, S.ipiv, S.inv, R2t, m2, m2, S.local));
    CHK(from_work(c, L, R2t, X2, m2, m2));
  }
  // T1 = B Z2 - A Xn (= -R1);  R2 = I - (D Z2 - C Xn)
  CHK(gemm(c, L, m1, m2, m2, d1, Braw, m1, X2, m2, d0, R1, m1));
  CHK(gemm(c, L, m1, m2, m1, dm1, E1.Araw, m1, Xn, m1, d1, R1, m1));
  CHK(gemm(c, L, m2, m2, m2, d1, E2.Araw, m2, X2, m2, d0, R2, m2));
  CHK(gemm(c, L, m2, m2, m1, dm1, Cm, m2, Xn, m1, d1, R2, m2));
  if ((e = bem_identity_minus(R2, m2, m2, L.st))) return fail(c, BEM_E_DEVICE, "identity launch: %d", e);
  if (bem_sumsq(R1, m1, m1, m2, part, 0, rfull, L.st) || bem_sumsq(R2, m2, m2, m2, part, 1, rfull, L.st))
    return fail(c, BEM_E_DEVICE, "residual norm launch");
  // correction through the factors
  CHK(to_work(c, L, R1, R1t, m1, m2));
  CHK(lu_solve(c, L, A.lu, m1, m1, A.ipiv, A.inv, R1t, m1, m2, A.local));    // u
  CHK(to_work(c, L, R2, R2t, m2, m2));
  CHK(gemm(c, L, m2, m2, m1, one, Ct, m2, R1t, m1, one, R2t, m2));           // R2 + C u
  CHK(lu_solve(c, L, S.lu, m2, m2, S.ipiv, S.inv, R2t, m2, m2, S.local));    // dZ2
  CHK(gemm(c, L, m1, m2, m2, one, Yt, m1, R2t, m2, one, R1t, m1));           // u + Y dZ2 = -dX
  CHK(add_work(c, L, R1t, Xn, m1, m2));
  CHK(add_work(c, L, R2t, X2, m2, m2));
  // FP64 products and trace of the corrected X
  CHK(gemm(c, L, m1, m2, m1, d1, E1.Gs, m1, Xn, m1, d0, B1, m1));
  EventRec *r;
  CHK(ev_begin(c, L, BEM_KS_ZTRACE, 4.0 * m1 * (double)m2 * m2, 16.0 * (2.0 * m1 * m2 + 0.5 * m2 * m2), &r));
  if ((e = bem_zgemm_htrace(m2, m1, Xn, m1, B1, m1, E2.Gs, m2, partial, trace_out, L.st)))
    return fail(c, BEM_E_DEVICE, "trace gemm: %d", e);
  CHK(ev_end(c, L, r));
  return probe<T>(c, L, A, S, E1.Araw, Braw, Cm, E2.Araw, E1.Gs, E2.Gs, Xn, X2, pout);
}

// Mixed-precision elimination (BEM_PREC_MIXED) after the FP64 assembly of B,
// C: Y = A^-1 B, Sc = D - C Y, X = -Y Sc^-1 and the products in complex64 on
// FP32 MFMA from the cached complex64 LU of A (the trace sums in FP64), then
// the probe of the complex64 X (widened to FP64 in Xn64) against the FP64
// blocks.  With refinement requested, Y is kept (Y32) for the correction
// steps.
static int solve_mixed(bem_ctx *c, Lane &L, const SelfEntry &E1, const SelfEntry &E2, double2 *Bm, double2 *Cm,
                       int *info, unsigned long long *gro, double2 *trace_out, ProbeOut *pout, bool refine) {
  const int m1 = E1.n, m2 = E2.n;
  const float2 one = cnum<float2>(1, 0), mone = cnum<float2>(-1, 0), zero = cnum<float2>(0, 0);
  float2 *B32, *C32, *S32, *B1, *invS, *Y32 = nullptr;
  double2 *Xn;
  int *ipS;
  double2 *partial;
  CHK(ws_get(c, L, "B32", sizeof(float2) * (size_t)m1 * m2, (void **)&B32));
  CHK(ws_get(c, L, "C32", sizeof(float2) * (size_t)m2 * m1, (void **)&C32));
  CHK(ws_get(c, L, "S32", sizeof(float2) * (size_t)m2 * m2, (void **)&S32));
  CHK(ws_get(c, L, "B1_32", sizeof(float2) * (size_t)m1 * m2, (void **)&B1));
  CHK(ws_get(c, L, "invS32", inv_bytes<float2>(m2), (void **)&invS));
  CHK(ws_get(c, L, "ipiv", sizeof(int) * (size_t)(m2 + 64), (void **)&ipS));
  CHK(ws_get(c, L, "Xn64", sizeof(double2) * (size_t)m1 * m2, (void **)&Xn));
  size_t np = std::max(bem_chtrace_partials(m2), bem_htrace_partials(m2));
  CHK(ws_get(c, L, "trace_partial", sizeof(double2) * (np + 1), (void **)&partial));
  if (bem_z2c(Bm, m1, B32, m1, m1, m2, L.st) || bem_z2c(Cm, m2, C32, m2, m2, m1, L.st))
    return fail(c, BEM_E_DEVICE, "complex64 copy launch");
  DEV(c, hipMemcpyAsync(S32, E2.Z32, sizeof(float2) * (size_t)m2 * m2, hipMemcpyDeviceToDevice, L.st));
  CHK(lu_solve(c, L, E1.Z32, m1, m1, E1.ipiv, E1.inv32, B32, m1, m2, true));   // B32 <- Y
  CHK(gemm(c, L, m2, m2, m1, mone, C32, m2, B32, m1, one, S32, m2));     // Sc = D - C Y
  CHK(lu_factor_block(c, L, S32, m2, m2, ipS, info, invS, gro));
  if (refine) {
    CHK(ws_get(c, L, "Y32", sizeof(float2) * (size_t)m1 * m2, (void **)&Y32));
    DEV(c, hipMemcpyAsync(Y32, B32, sizeof(float2) * (size_t)m1 * m2, hipMemcpyDeviceToDevice, L.st));
  }
  CHK(lu_solve_right(c, L, S32, m2, m2, ipS, invS, B32, m1, m1, true));  // B32 <- Y Sc^-1 = -X
  if (bem_c2z(B32, m1, Xn, m1, m1, m2, L.st)) return fail(c, BEM_E_DEVICE, "complex128 copy launch");
  const LuRef<float2> A{E1.Z32, m1, E1.ipiv, E1.inv32, true}, S{S32, m2, ipS, invS, true};
  CHK(probe_start<float2>(c, L, A, S, E1.Araw, Bm, Cm, E2.Araw, E1.Gs, E2.Gs, Xn, pout));
  // products: 4 S = tr(sym G2 X^H sym G1 X), even in X, over the Hermitian
  // X^H (sym G1 X): half of it is formed
  CHK(gemm(c, L, m1, m2, m1, one, E1.Gs32, m1, B32, m1, zero, B1, m1));
  EventRec *r;
  CHK(ev_begin(c, L, BEM_KS_CGEMM, 4.0 * m1 * (double)m2 * m2, 8.0 * (2.0 * m1 * m2 + 0.5 * m2 * m2), &r));
  int e = bem_cgemm_htrace(m2, m1, B32, m1, B1, m1, E2.Gs32, m2, partial, trace_out, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "complex64 trace gemm: %d", e);
  return ev_end(c, L, r);
}

// ------------------------------------------------- low-rank coupling solve
// Columns per block of the randomized range finder (one CholQR2 block), and
// the most blocks one compression takes (ranks up to 32768).
static const int LR_BLK = 64;
static const int LR_MAX_BLOCKS = 512;
// Seed of the range finder's test columns (fixed, so a request gives the same
// factors alone and in a mix).
static const unsigned long long LR_SEED = 0x10C0FFEEull;
// Blocks whose products Y = B W_k are formed by one GEMM.
static const int LR_YGROUP = 8;

// C (m x n) <- alpha A^H B + beta C (A k x m, B k x n) on the lane's stream.
static int gemm_ta(bem_ctx *c, Lane &L, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda,
                   const double2 *B, int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  double2 *part = nullptr;
  const size_t np = bem_zgemm_ta_partials(m, n, k);
  if (np) CHK(ws_get(c, L, "lr_tapart", sizeof(double2) * np, (void **)&part));
  EventRec *r;
  CHK(ev_begin(c, L, BEM_KS_ZGEMM_NARROW, 8.0 * m * (double)n * k,
               16.0 * ((double)m * k + (double)k * n + (double)m * n), &r));
  int e = bem_zgemm_ta(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, part, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "zgemm (A^H B) launch: %d", e);
  return ev_end(c, L, r);
}

// Blocked randomized range finder of B (m1 x m2), in chunks of blocks on the
// lane's stream: Q (m1 x r, orthonormal columns, r a multiple of LR_BLK,
// leading dimension m1) and then Vh = Q^H B (r x m2) with B ~ Q Vh.  Block k
// draws LR_BLK seeded test columns W_k (entries with real and imaginary parts
// uniform in [-1, 1), E|w|^2 = 2/3), Y = B W_k, projects out the blocks
// before it twice (Y -= Q (Q^H Y)), and orthonormalises Y by CholQR2 (G =
// Y^H Y = R^H R, Y <- Y R^-1, twice).  E||Y_k||_F^2 = LR_BLK (2/3)
// ||B - Q_<k Q_<k^H B||_F^2, so r ends with the first block whose estimate is
// at most tol^2 ||B||_F^2 (that block still joins Q, which only lowers the
// error); a block whose Gram is singular ends it before that block when its
// estimate is under the tolerance, else the block is not compressible here.
// lr_rand_issue() queues `nblocks` more blocks and the copies of their
// estimates; lr_rand_eval() reads them once the stream has drained.  Block k
// depends only on the blocks before it, so r, Q and Vh do not depend on the
// chunking (the first chunk is what the pair needed last time).
static const int LR_CHUNK = 4;
static int lr_rand_issue(bem_ctx *c, Lane &L, int nblocks) {
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  auto &S = L.lr;
  const int m1 = S.m1, m2 = S.m2;
  double2 *W, *Cb, *G, *Rinv, *tap;
  double *est, *part;
  int *binfo;
  CHK(ws_get(c, L, "lr_W", sizeof(double2) * (size_t)m2 * LR_BLK * LR_YGROUP, (void **)&W));
  CHK(ws_get(c, L, "lr_Cb", sizeof(double2) * (size_t)S.nbmax * LR_BLK * LR_BLK, (void **)&Cb));
  CHK(ws_get(c, L, "lr_G", sizeof(double2) * LR_BLK * LR_BLK, (void **)&G));
  CHK(ws_get(c, L, "lr_Rinv", sizeof(double2) * 64 * 64, (void **)&Rinv));
  CHK(ws_get(c, L, "lr_est", sizeof(double) * (LR_MAX_BLOCKS + 1), (void **)&est));
  CHK(ws_get(c, L, "lr_binfo", sizeof(int) * (LR_MAX_BLOCKS + 1), (void **)&binfo));
  const size_t np = std::max(bem_sumsq_partials(m1, m2), bem_sumsq_partials(m1, LR_BLK));
  CHK(ws_get(c, L, "lr_ssq", sizeof(double) * np, (void **)&part));
  // the split-K partials of every product below, at their largest, so the
  // buffer does not grow between launches
  size_t ntap = bem_zgemm_ta_partials(LR_BLK, LR_BLK, m1);
  for (int rk = LR_BLK; rk <= S.rmax; rk += LR_BLK) ntap = std::max(ntap, bem_zgemm_ta_partials(rk, LR_BLK, m1));
  CHK(ws_get(c, L, "lr_tapart", sizeof(double2) * std::max<size_t>(ntap, 1), (void **)&tap));
  if (S.issued == 0) {
    if (bem_sumsq(S.B, m1, m1, m2, part, 0, est, L.st)) return fail(c, BEM_E_DEVICE, "norm launch");   // ||B||^2
    DEV(c, hipMemsetAsync(binfo, 0, sizeof(int) * (LR_MAX_BLOCKS + 1), L.st));
  }
  const int k1 = std::min(S.nbmax, S.issued + nblocks);
  for (int k = S.issued; k < k1; k++) {
    double2 *Y = S.Q + (int64_t)k * LR_BLK * m1;
    const int rk = k * LR_BLK;
    if ((k - S.issued) % LR_YGROUP == 0) {
      // Y = B W for up to LR_YGROUP blocks in one GEMM.  With the
      // three-product tiles every output element sums over m2 in the same
      // order whatever the product's width, so block k's columns are the
      // numbers its own product gives and the grouping changes only the rate
      const int g = std::min(LR_YGROUP, k1 - k);
      for (int j = 0; j < g; j++)
        if (bem_fill_hash(W + (size_t)j * m2 * LR_BLK, sizeof(double2) * (size_t)m2 * LR_BLK, 0,
                          LR_SEED + 0x9E37ull * (k + j), L.st))
          return fail(c, BEM_E_DEVICE, "test column fill launch");
      CHK(gemm(c, L, m1, LR_BLK * g, m2, one, S.B, m1, W, m2, zero, Y, m1));
    }
    if (rk > 0)
      for (int pass = 0; pass < 2; pass++) {
        CHK(gemm_ta(c, L, rk, LR_BLK, m1, one, S.Q, m1, Y, m1, zero, Cb, rk));
        CHK(gemm(c, L, m1, LR_BLK, rk, mone, S.Q, m1, Cb, rk, one, Y, m1));
      }
    if (bem_sumsq(Y, m1, m1, LR_BLK, part, 0, est + 1 + k, L.st)) return fail(c, BEM_E_DEVICE, "norm launch");
    for (int pass = 0; pass < 2; pass++) {
      CHK(gemm_ta(c, L, LR_BLK, LR_BLK, m1, one, Y, m1, Y, m1, zero, G, LR_BLK));
      if (bem_chol_inv64(G, LR_BLK, LR_BLK, Rinv, binfo + k, L.st)) return fail(c, BEM_E_DEVICE, "cholesky launch");
      CHK(gemm_inplace(c, L, 'R', m1, LR_BLK, Rinv, Y, m1));
    }
  }
  S.issued = k1;
  DEV(c, hipMemcpyAsync(L.h_lr, est, sizeof(double) * (S.issued + 1), hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_lri, binfo, sizeof(int) * S.issued, hipMemcpyDeviceToHost, L.st));
  return 0;
}

// Reads the estimates of the blocks issued since the last call (the lane's
// stream has drained): *done with *rank = r, 0 when B is zero or -1 when it
// does not compress within rmax; otherwise more blocks are needed.
static void lr_rand_eval(Lane &L, bool *done, int *rank) {
  auto &S = L.lr;
  *done = false;
  *rank = -1;
  const double bn2 = L.h_lr[0];
  if (!(bn2 > 0)) {
    *done = true;
    *rank = bn2 == 0 ? 0 : -1;
    return;
  }
  for (int k = S.checked; k < S.issued; k++) {
    const bool last = L.h_lr[1 + k] / (LR_BLK * (2.0 / 3.0)) <= S.tol * S.tol * bn2;
    if (L.h_lri[k]) {   // dependent columns: the range is exhausted if the estimate is under the tolerance
      *done = true;
      *rank = last && k > 0 ? k * LR_BLK : -1;
      return;
    }
    if (last) {
      *done = true;
      *rank = (k + 1) * LR_BLK;
      return;
    }
  }
  S.checked = S.issued;
  if (S.issued >= S.nbmax) *done = true;   // rmax columns without reaching the tolerance
}

// Adaptive cross approximation with full pivoting on the device (see
// bem_aca): batches of 16 steps, the state read back after each batch.
// rows/cols: device arrays of kmax ints.  *full: kmax columns were used up
// before the loop ended.
static int aca_run(bem_ctx *c, Lane &L, double2 *M, int64_t ldm, int m, int n, double eps2, int kmax, double2 *U,
                   int64_t ldu, double2 *V, int64_t ldv, int *rows, int *cols, int *rank, int *full) {
  void *state, *part;
  CHK(ws_get(c, L, "aca_state", bem_aca_state_bytes(), &state));
  CHK(ws_get(c, L, "aca_part", bem_aca_partial_bytes(m, n), &part));
  if (bem_aca_init(state, m, n, L.st)) return fail(c, BEM_E_DEVICE, "aca state upload");
  const int smax = std::min(std::min(m, n), kmax) + 1;
  int nu = 0, done = 0;
  *rank = 0;
  *full = 0;
  for (int s0 = 0; s0 < smax && !done; s0 += 16) {
    if (bem_aca_steps(M, ldm, m, n, eps2, kmax, U, ldu, V, ldv, rows, cols, state, part, std::min(16, smax - s0),
                      L.st))
      return fail(c, BEM_E_DEVICE, "aca launch");
    DEV(c, hipStreamSynchronize(L.st));
    if (bem_aca_read(state, &nu, rank, &done, full)) return fail(c, BEM_E_DEVICE, "aca state read");
  }
  return 0;
}

// Rk::truncate (see bem_rk_truncate) of U (m x k, ldu) V (k x n, ldv) on the
// lane's stream.  Jacobi sweeps run until a sweep rotates no pair (at most
// 60).  sigma_host (optional) receives the k singular values, decreasing.
static int rk_truncate(bem_ctx *c, Lane &L, double2 *U, int64_t ldu, double2 *V, int64_t ldv, int m, int n, int k,
                       double eps, int *knew, double *sigma_host) {
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  *knew = 0;
  if (k <= 0) return 0;
  const int ke = k + (k & 1);   // even column count for the round-robin pairs (a zero column pads)
  double2 *Vt, *W, *U1, *QW;
  double *sig;
  int *rot, *idx;
  CHK(ws_get(c, L, "rk_Vt", sizeof(double2) * (size_t)n * ke, (void **)&Vt));
  CHK(ws_get(c, L, "rk_W", sizeof(double2) * (size_t)ke * ke, (void **)&W));
  CHK(ws_get(c, L, "rk_U1", sizeof(double2) * (size_t)m * ke, (void **)&U1));
  CHK(ws_get(c, L, "rk_QW", sizeof(double2) * (size_t)n * ke, (void **)&QW));
  CHK(ws_get(c, L, "rk_sig", sizeof(double) * 2 * (size_t)ke, (void **)&sig));
  CHK(ws_get(c, L, "rk_rot", sizeof(int) * 4, (void **)&rot));
  CHK(ws_get(c, L, "rk_idx", sizeof(int) * (size_t)ke, (void **)&idx));
  std::vector<double2> eye((size_t)ke * ke, zero);
  for (int i = 0; i < ke; i++) eye[(size_t)i * ke + i] = one;
  auto sweeps = [&](double2 *A, int rows_a) -> int {
    for (int sw = 0; sw < 60; sw++) {
      DEV(c, hipMemsetAsync(rot, 0, sizeof(int), L.st));
      if (bem_jacobi_sweep(A, rows_a, rows_a, ke, W, ke, ke, 1e-15, rot, L.st))
        return fail(c, BEM_E_DEVICE, "jacobi launch");
      DEV(c, hipMemcpyAsync(L.h_lri, rot, sizeof(int), hipMemcpyDeviceToHost, L.st));
      DEV(c, hipStreamSynchronize(L.st));
      if (!L.h_lri[0]) return 0;
    }
    return 0;
  };
  // pass 1: Vt = V^H (n x k), Vt W1 = Y1 with orthogonal columns; Q1 = Y1 / |Y1|
  DEV(c, hipMemsetAsync(Vt, 0, sizeof(double2) * (size_t)n * ke, L.st));
  if (bem_transpose_op(V, ldv, k, n, Vt, n, 1, 0, L.st)) return fail(c, BEM_E_DEVICE, "transpose launch");
  DEV(c, hipMemcpyAsync(W, eye.data(), sizeof(double2) * eye.size(), hipMemcpyHostToDevice, L.st));
  CHK(sweeps(Vt, n));
  if (bem_col_norms(Vt, n, n, ke, sig, L.st)) return fail(c, BEM_E_DEVICE, "norm launch");
  // U1 = U W1 Sigma1 (the zero pad column of U contributes nothing)
  DEV(c, hipMemsetAsync(U1, 0, sizeof(double2) * (size_t)m * ke, L.st));
  CHK(gemm(c, L, m, ke, k, one, U, ldu, W, ke, zero, U1, m));
  std::vector<double> s1(ke);
  DEV(c, hipMemcpyAsync(s1.data(), sig, sizeof(double) * ke, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipStreamSynchronize(L.st));
  std::vector<int> id(ke);
  std::vector<double> inv1(ke);
  for (int j = 0; j < ke; j++) {
    id[j] = j;
    inv1[j] = s1[j] > 0 ? 1.0 / s1[j] : 0.0;
  }
  DEV(c, hipMemcpyAsync(idx, id.data(), sizeof(int) * ke, hipMemcpyHostToDevice, L.st));
  DEV(c, hipMemcpyAsync(sig + ke, inv1.data(), sizeof(double) * ke, hipMemcpyHostToDevice, L.st));
  // U1 <- U1 Sigma1 (column scale), Vt <- Q1 (normalised columns), in place
  if (bem_gather_cols(U1, m, m, idx, sig, ke, U1, m, L.st) ||
      bem_gather_cols(Vt, n, n, idx, sig + ke, ke, Vt, n, L.st))
    return fail(c, BEM_E_DEVICE, "column scale launch");
  // pass 2: U1 W2 = Y2 orthogonal columns, sigma = |Y2_j|; U V = Y2 (Q1 W2)^H
  DEV(c, hipMemcpyAsync(W, eye.data(), sizeof(double2) * eye.size(), hipMemcpyHostToDevice, L.st));
  CHK(sweeps(U1, m));
  if (bem_col_norms(U1, m, m, ke, sig, L.st)) return fail(c, BEM_E_DEVICE, "norm launch");
  CHK(gemm(c, L, n, ke, ke, one, Vt, n, W, ke, zero, QW, n));
  std::vector<double> s2(ke);
  DEV(c, hipMemcpyAsync(s2.data(), sig, sizeof(double) * ke, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipStreamSynchronize(L.st));
  std::stable_sort(id.begin(), id.end(), [&](int a, int b) { return s2[a] > s2[b]; });
  const double smax = s2[id[0]];
  int kn = 0;
  for (int j = 0; j < k; j++)
    if (s2[id[j]] >= eps * smax && s2[id[j]] > 0) kn++;
  if (sigma_host)
    for (int j = 0; j < k; j++) sigma_host[j] = s2[id[j]];
  DEV(c, hipMemcpyAsync(idx, id.data(), sizeof(int) * ke, hipMemcpyHostToDevice, L.st));
  // U <- Y2(:, kept) (= U_svd S), V <- (Q1 W2)(:, kept)^H
  if (kn > 0) {
    if (bem_gather_cols(U1, m, m, idx, nullptr, kn, U, ldu, L.st) ||
        bem_gather_cols(QW, n, n, idx, nullptr, kn, Vt, n, L.st) ||
        bem_transpose_op(Vt, n, n, kn, V, ldv, 1, 0, L.st))
      return fail(c, BEM_E_DEVICE, "truncation gather launch");
  }
  DEV(c, hipStreamSynchronize(L.st));
  *knew = kn;
  return 0;
}

// Compression of B by the reference's route: ACA with full pivoting on a
// copy of B, then Rk::truncate.  ACA stops on the size of its last increment
// and the truncation on singular values relative to the largest, neither a
// bound on ||B - U Vh||_F, so both run a decade tighter than tol:
// eps2 = (tol / 10)^2 and eps = tol / 10.  U (m1 x rmax) and Vh (rmax x m2,
// leading dimension ldv) receive the factors; *rank as lr_rand_eval's.
static int lr_compress_aca(bem_ctx *c, Lane &L, const double2 *B, int m1, int m2, double tol, int rmax, double2 *U,
                           double2 *Vh, int ldv, int *rank) {
  *rank = -1;
  double2 *M;
  int *rc;
  CHK(ws_get(c, L, "lr_M", sizeof(double2) * (size_t)m1 * m2, (void **)&M));
  CHK(ws_get(c, L, "lr_rc", sizeof(int) * 2 * (size_t)rmax, (void **)&rc));
  DEV(c, hipMemcpyAsync(M, B, sizeof(double2) * (size_t)m1 * m2, hipMemcpyDeviceToDevice, L.st));
  int r = 0, full = 0;
  const double t = 0.1 * tol;
  CHK(aca_run(c, L, M, m1, m1, m2, t * t, rmax, U, m1, Vh, ldv, rc, rc + rmax, &r, &full));
  if (full) return 0;
  if (r == 0) {
    *rank = 0;
    return 0;
  }
  int kn = 0;
  CHK(rk_truncate(c, L, U, m1, Vh, ldv, m1, m2, r, t, &kn, nullptr));
  *rank = kn;
  return 0;
}

// Error probe of a low-rank solve, as probe() for the dense one: w seeded
// (m2 x k); x1 = X w = -P (Z w); x2 = Sc~^-1 w = D^-1 (w + U2 K Z w) (the
// Woodbury form of the compressed Schur complement); r = [0; w] - W x with
// the dense A, B, C, D; d_1 = u + X (r2 - C u), u = A^-1 r1; h_j = g_j^H d_1j,
// g = sym G1 X sym G2 w.  Since W holds the dense B and C, the residual and
// the estimate include the truncation of B.
static int probe_lowrank(bem_ctx *c, Lane &L, const LuRef<double2> &A, const LuRef<double2> &D, const double2 *Araw,
                         const double2 *Draw, const double2 *Braw, const double2 *Cm, const double2 *Gs1,
                         const double2 *Gs2, const double2 *P, const double2 *Z, const double2 *K, const double2 *U2,
                         int r, ProbeOut *out) {
  const int m1 = A.n, m2 = D.n, k = BEM_PROBE_COLS;
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  float2 *w32;
  double2 *w, *x1, *x2, *r1, *r2, *t1, *t2, *zr, *yr;
  double *part;
  CHK(ws_get(c, L, "pr_w32", sizeof(float2) * (size_t)m2 * k, (void **)&w32));
  CHK(ws_get(c, L, "pr_w", sizeof(double2) * (size_t)m2 * k, (void **)&w));
  CHK(ws_get(c, L, "pr_x1", sizeof(double2) * (size_t)m1 * k, (void **)&x1));
  CHK(ws_get(c, L, "pr_x2", sizeof(double2) * (size_t)m2 * k, (void **)&x2));
  CHK(ws_get(c, L, "pr_r1", sizeof(double2) * (size_t)m1 * k, (void **)&r1));
  CHK(ws_get(c, L, "pr_r2", sizeof(double2) * (size_t)m2 * k, (void **)&r2));
  CHK(ws_get(c, L, "pr_t1", sizeof(double2) * (size_t)m1 * k, (void **)&t1));
  CHK(ws_get(c, L, "pr_t2", sizeof(double2) * (size_t)m2 * k, (void **)&t2));
  CHK(ws_get(c, L, "pr_zr", sizeof(double2) * (size_t)r * k, (void **)&zr));
  CHK(ws_get(c, L, "pr_yr", sizeof(double2) * (size_t)r * k, (void **)&yr));
  CHK(ws_get(c, L, "pr_part", sizeof(double) * std::max(bem_sumsq_partials(m1, k), bem_sumsq_partials(m2, k)),
             (void **)&part));
  const size_t wb2 = sizeof(double2) * (size_t)m2 * k;
  if (bem_fill_hash(w32, sizeof(float2) * (size_t)m2 * k, 1, PROBE_SEED, L.st) ||
      bem_c2z(w32, m2, w, m2, m2, k, L.st))
    return fail(c, BEM_E_DEVICE, "probe fill launch");
  DEV(c, hipMemsetAsync(out, 0, 2 * sizeof(unsigned long long), L.st));
  // x1 = -P (Z w);  x2 = D^-1 (w + U2 K (Z w))
  CHK(gemm_few(c, L, r, k, m2, one, Z, r, w, m2, zero, zr, r));
  CHK(gemm_few(c, L, m1, k, r, mone, P, m1, zr, r, zero, x1, m1));
  CHK(gemm_few(c, L, r, k, r, one, K, r, zr, r, zero, yr, r));
  DEV(c, hipMemcpyAsync(x2, w, wb2, hipMemcpyDeviceToDevice, L.st));
  CHK(gemm_few(c, L, m2, k, r, one, U2, m2, yr, r, one, x2, m2));
  CHK(lu_solve_few(c, L, D.lu, m2, D.ipiv, D.inv, x2, m2, k, D.local));
  // r1 = -(A x1 + B x2), r2 = w - C x1 - D x2
  CHK(gemm_few(c, L, m1, k, m1, mone, Araw, m1, x1, m1, zero, r1, m1));
  CHK(gemm_few(c, L, m1, k, m2, mone, Braw, m1, x2, m2, one, r1, m1));
  DEV(c, hipMemcpyAsync(r2, w, wb2, hipMemcpyDeviceToDevice, L.st));
  CHK(gemm_few(c, L, m2, k, m1, mone, Cm, m2, x1, m1, one, r2, m2));
  CHK(gemm_few(c, L, m2, k, m2, mone, Draw, m2, x2, m2, one, r2, m2));
  if (bem_sumsq(r1, m1, m1, k, part, 0, &out->rr, L.st) || bem_sumsq(r2, m2, m2, k, part, 1, &out->rr, L.st) ||
      bem_sumsq(w, m2, m2, k, part, 0, &out->ww, L.st))
    return fail(c, BEM_E_DEVICE, "probe norm launch");
  // d_1 = u + X (r2 - C u) = u - P (Z (r2 - C u)), u = A^-1 r1
  DEV(c, hipMemcpyAsync(t1, r1, sizeof(double2) * (size_t)m1 * k, hipMemcpyDeviceToDevice, L.st));
  CHK(lu_solve_few(c, L, A.lu, m1, A.ipiv, A.inv, t1, m1, k, A.local));   // u
  DEV(c, hipMemcpyAsync(t2, r2, wb2, hipMemcpyDeviceToDevice, L.st));
  CHK(gemm_few(c, L, m2, k, m1, mone, Cm, m2, t1, m1, one, t2, m2));
  CHK(gemm_few(c, L, r, k, m2, one, Z, r, t2, m2, zero, zr, r));
  CHK(gemm_few(c, L, m1, k, r, mone, P, m1, zr, r, one, t1, m1));
  int e = bem_absmax(t1, m1, m1, k, 0, &out->dmax, L.st);
  if (!e) e = bem_absmax(x1, m1, m1, k, 0, &out->xmax, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "probe absmax launch: %d", e);
  // g = sym G1 X sym G2 w = -sym G1 P Z sym G2 w
  CHK(gemm_few(c, L, m2, k, m2, one, Gs2, m2, w, m2, zero, r2, m2));
  CHK(gemm_few(c, L, r, k, m2, one, Z, r, r2, m2, zero, zr, r));
  CHK(gemm_few(c, L, m1, k, r, one, P, m1, zr, r, zero, r1, m1));
  CHK(gemm_few(c, L, m1, k, m1, mone, Gs1, m1, r1, m1, zero, x1, m1));
  if ((e = bem_zdotc_cols(x1, m1, t1, m1, m1, k, out->h, L.st))) return fail(c, BEM_E_DEVICE, "probe dot launch: %d", e);
  return 0;
}

// Elimination on the compressed coupling B ~ U Vh (U m1 x r, Vh r x m2,
// leading dimension ldv), C = S B^T S ~ U2 V2^H with U2 = S Vh^T and
// V2 = S conj(U), against LU(A) (E1) and LU(D) (E2):
//   P = A^-1 U, K = V2^H P, Rh = Vh D^-1, F = Rh U2, T = I - F K,
//   Z = T^-1 Rh, X = -P Z,
//   4 S = tr(M1 M2), M1 = P^H sym G1 P, M2 = Z sym G2 Z^H (both Hermitian, so
//   the trace is the Hermitian-trace GEMM of Zt^H (sym G2 Zt) against M1,
//   Zt = Z^H),
// then the probe against the dense B and C.
static int solve_lowrank(bem_ctx *c, Lane &L, const SelfEntry &E1, const SelfEntry &E2, const double2 *Bm,
                         const double2 *Cm, const double2 *U, const double2 *Vh, int ldv, int r, int *info,
                         unsigned long long *gro, double2 *trace_out, ProbeOut *pout) {
  const int m1 = E1.n, m2 = E2.n;
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  double2 *P, *Qc, *K, *Rh, *U2, *F, *T, *invT, *Zt, *H1, *M1, *W2, *partial;
  int *ipT;
  CHK(ws_get(c, L, "lr_P", sizeof(double2) * (size_t)m1 * r, (void **)&P));
  CHK(ws_get(c, L, "lr_Qc", sizeof(double2) * (size_t)m1 * r, (void **)&Qc));
  CHK(ws_get(c, L, "lr_K", sizeof(double2) * (size_t)r * r, (void **)&K));
  CHK(ws_get(c, L, "lr_Rh", sizeof(double2) * (size_t)r * m2, (void **)&Rh));
  CHK(ws_get(c, L, "lr_U2", sizeof(double2) * (size_t)m2 * r, (void **)&U2));
  CHK(ws_get(c, L, "lr_F", sizeof(double2) * (size_t)r * r, (void **)&F));
  CHK(ws_get(c, L, "lr_T", sizeof(double2) * (size_t)r * r, (void **)&T));
  CHK(ws_get(c, L, "lr_invT", inv_bytes(r), (void **)&invT));
  CHK(ws_get(c, L, "lr_ipT", sizeof(int) * (size_t)(r + 64), (void **)&ipT));
  CHK(ws_get(c, L, "lr_Zt", sizeof(double2) * (size_t)m2 * r, (void **)&Zt));
  CHK(ws_get(c, L, "lr_H1", sizeof(double2) * (size_t)m1 * r, (void **)&H1));
  CHK(ws_get(c, L, "lr_M1", sizeof(double2) * (size_t)r * r, (void **)&M1));
  CHK(ws_get(c, L, "lr_W2", sizeof(double2) * (size_t)m2 * r, (void **)&W2));
  CHK(ws_get(c, L, "trace_partial", sizeof(double2) * (bem_htrace_partials(r) + 1), (void **)&partial));
  // P = A^-1 U;  K = (S conj U)^H P
  DEV(c, hipMemcpyAsync(P, U, sizeof(double2) * (size_t)m1 * r, hipMemcpyDeviceToDevice, L.st));
  CHK(lu_solve(c, L, E1.Z, m1, m1, E1.ipiv, E1.inv, P, m1, r, !E1.pivot));
  if (bem_conj_sign(U, m1, m1, r, Qc, m1, 1, L.st)) return fail(c, BEM_E_DEVICE, "conjugate launch");
  CHK(gemm_ta(c, L, r, r, m1, one, Qc, m1, P, m1, zero, K, r));
  // Rh = Vh D^-1;  U2 = S Vh^T;  F = Rh U2;  T = I - F K
  DEV(c, hipMemcpy2DAsync(Rh, sizeof(double2) * r, Vh, sizeof(double2) * ldv, sizeof(double2) * r, m2,
                          hipMemcpyDeviceToDevice, L.st));
  CHK(lu_solve_right(c, L, E2.Z, m2, m2, E2.ipiv, E2.inv, Rh, r, r, !E2.pivot));
  if (bem_transpose_op(Vh, ldv, r, m2, U2, m2, 0, 1, L.st)) return fail(c, BEM_E_DEVICE, "transpose launch");
  CHK(gemm(c, L, r, r, m2, one, Rh, r, U2, m2, zero, F, r));
  CHK(gemm(c, L, r, r, r, one, F, r, K, r, zero, T, r));
  if (bem_identity_minus(T, r, r, L.st)) return fail(c, BEM_E_DEVICE, "identity launch");
  // Z = T^-1 Rh; T = I - F K is the identity less the coupling loop, and its
  // block-local LU reports its growth like the Schur complement's
  CHK(lu_factor_block(c, L, T, r, r, ipT, info, invT, gro));
  CHK(lu_solve(c, L, T, r, r, ipT, invT, Rh, r, m2, true));
  double2 *Z = Rh;
  DEV(c, hipEventRecord(L.ph[2], L.st));
  // 4 S = tr(M1 M2)
  if (bem_transpose_op(Z, r, r, m2, Zt, m2, 1, 0, L.st)) return fail(c, BEM_E_DEVICE, "transpose launch");
  CHK(gemm(c, L, m1, r, m1, one, E1.Gs, m1, P, m1, zero, H1, m1));
  CHK(gemm_ta(c, L, r, r, m1, one, P, m1, H1, m1, zero, M1, r));
  CHK(gemm(c, L, m2, r, m2, one, E2.Gs, m2, Zt, m2, zero, W2, m2));
  EventRec *er;
  CHK(ev_begin(c, L, BEM_KS_ZTRACE, 4.0 * m2 * (double)r * r, 16.0 * (2.0 * m2 * r + 0.5 * r * r), &er));
  int e = bem_zgemm_htrace(r, m2, Zt, m2, W2, m2, M1, r, partial, trace_out, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "trace gemm: %d", e);
  CHK(ev_end(c, L, er));
  const LuRef<double2> A{E1.Z, m1, E1.ipiv, E1.inv, !E1.pivot}, D{E2.Z, m2, E2.ipiv, E2.inv, !E2.pivot};
  return probe_lowrank(c, L, A, D, E1.Araw, E2.Araw, Bm, Cm, E1.Gs, E2.Gs, P, Z, K, U2, r, pout);
}

// The low-rank coupling could not be used: the dense FP64 solve on the same
// blocks needs its own copy of B for the probe, since it factors B in place.
static int lr_dense_fallback(bem_ctx *c, Lane &L) {
  const int m1 = L.e1->n, m2 = L.e2->n;
  double2 *Bm, *Braw;
  CHK(ws_get(c, L, "B", sizeof(double2) * (size_t)m1 * m2, (void **)&Bm));
  CHK(ws_get(c, L, "Braw", sizeof(double2) * (size_t)m1 * m2, (void **)&Braw));
  DEV(c, hipMemcpyAsync(Braw, Bm, sizeof(double2) * (size_t)m1 * m2, hipMemcpyDeviceToDevice, L.st));
  return 0;
}

// Issues the elimination, products, trace and probe of the lane's request
// once its blocks are assembled (lane_issue) -- dense, complex64, or, with
// r > 0, on the compressed coupling (Vh = Q^H B first unless vh_ready) --
// and the copies of the results; lane_step() collects them.
static int issue_solve(bem_ctx *c, Lane &L, int r, bool vh_ready) {
  const SelfEntry &E1 = *L.e1, &E2 = *L.e2;
  const int m1 = E1.n, m2 = E2.n;
  const int pivot = (L.req.flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
  const int prec = L.prec;
  double2 *Bm, *Cm, *Braw, *trace_out;
  int *info;
  unsigned long long *gro;
  ProbeOut *pout;
  CHK(ws_get(c, L, "B", sizeof(double2) * (size_t)m1 * m2, (void **)&Bm));
  CHK(ws_get(c, L, "C", sizeof(double2) * (size_t)m2 * m1, (void **)&Cm));
  if (prec == BEM_PREC_FP64 && r == 0) CHK(ws_get(c, L, "Braw", sizeof(double2) * (size_t)m1 * m2, (void **)&Braw));
  else Braw = Bm;   // the mixed and low-rank eliminations leave B unfactored
  CHK(ws_get(c, L, "info", sizeof(int) * 4, (void **)&info));
  CHK(ws_get(c, L, "growth", sizeof(unsigned long long) * 3, (void **)&gro));
  CHK(ws_get(c, L, "trace_out", sizeof(double2) * 2, (void **)&trace_out));
  CHK(ws_get(c, L, "probe_out", sizeof(ProbeOut), (void **)&pout));
  L.stage = ST_SOLVE;
  if (r > 0) {
    L.rank = r;
    if (!vh_ready)
      CHK(gemm_ta(c, L, r, m2, m1, make_double2(1, 0), L.lr.Q, m1, Bm, m1, make_double2(0, 0), L.lr.Vh, L.lr.rmax));
    CHK(solve_lowrank(c, L, E1, E2, Bm, Cm, L.lr.Q, L.lr.Vh, L.lr.rmax, r, info, gro, trace_out, pout));
  } else if (prec == BEM_PREC_MIXED) {
    CHK(solve_mixed(c, L, E1, E2, Bm, Cm, info, gro, trace_out, pout, L.refine));
    DEV(c, hipEventRecord(L.ph[2], L.st));
  } else {
    double2 *Sc, *B1, *partial, *invS;
    int *ipS;
    CHK(ws_get(c, L, "Sc", sizeof(double2) * (size_t)m2 * m2, (void **)&Sc));
    CHK(ws_get(c, L, "B1", sizeof(double2) * (size_t)m1 * m2, (void **)&B1));
    CHK(ws_get(c, L, "ipiv", sizeof(int) * (size_t)(m2 + 64), (void **)&ipS));
    CHK(ws_get(c, L, "invS", inv_bytes(m2), (void **)&invS));
    size_t np = bem_htrace_partials(m2);
    CHK(ws_get(c, L, "trace_partial", sizeof(double2) * (np + 1), (void **)&partial));
    DEV(c, hipMemcpyAsync(Sc, E2.Araw, sizeof(double2) * (size_t)m2 * m2, hipMemcpyDeviceToDevice, L.st));
    // ---- solve: Y = A^-1 B, Sc = D - C Y, X = -Y Sc^-1
    const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
    CHK(lu_solve(c, L, E1.Z, m1, m1, E1.ipiv, E1.inv, Bm, m1, m2, !E1.pivot));   // B <- Y
    CHK(gemm(c, L, m2, m2, m1, mone, Cm, m2, Bm, m1, one, Sc, m2));
    if (pivot) CHK(lu_factor(c, L, Sc, m2, m2, ipS, info, invS));
    else CHK(lu_factor_block(c, L, Sc, m2, m2, ipS, info, invS, gro));
    if (L.refine) {   // Y for the refinement's corrections
      double2 *Y;
      CHK(ws_get(c, L, "Y64", sizeof(double2) * (size_t)m1 * m2, (void **)&Y));
      DEV(c, hipMemcpyAsync(Y, Bm, sizeof(double2) * (size_t)m1 * m2, hipMemcpyDeviceToDevice, L.st));
    }
    // B <- Y Sc^-1 = -X; the trace below is even in X, so the sign is dropped
    CHK(lu_solve_right(c, L, Sc, m2, m2, ipS, invS, Bm, m1, m1, !pivot));
    double2 *X = Bm;
    DEV(c, hipEventRecord(L.ph[2], L.st));
    const LuRef<double2> A{E1.Z, m1, E1.ipiv, E1.inv, !E1.pivot}, S{Sc, m2, ipS, invS, !pivot};
    CHK(probe_start<double2>(c, L, A, S, E1.Araw, Braw, Cm, E2.Araw, E1.Gs, E2.Gs, Bm, pout));
    // ---- products and trace: S = 1/4 sum conj(X sym(G2)) .* (sym(G1) X)
    // = 1/4 tr(sym(G2) P) with P = X^H sym(G1) X Hermitian (sym(G) is G + G^H),
    // so only the upper half of P is formed
    CHK(gemm(c, L, m1, m2, m1, one, E1.Gs, m1, X, m1, zero, B1, m1));
    EventRec *rec;
    CHK(ev_begin(c, L, BEM_KS_ZTRACE, 4.0 * m1 * (double)m2 * m2, 16.0 * (2.0 * m1 * m2 + 0.5 * m2 * m2), &rec));
    int e = bem_zgemm_htrace(m2, m1, X, m1, B1, m1, E2.Gs, m2, partial, trace_out, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "trace gemm: %d", e);
    CHK(ev_end(c, L, rec));
  }
  CHK(probe_join(c, L));
  DEV(c, hipEventRecord(L.ph[3], L.st));
  DEV(c, hipMemcpyAsync(L.h_trace, trace_out, sizeof(double2) * 2, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_info, info, sizeof(int) * 4, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_growth, gro, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_probe, pout, sizeof(ProbeOut), hipMemcpyDeviceToHost, L.st));
  L.busy = true;
  return 0;
}

// Issues every kernel of the solve stage of one spectral-power request on the
// lane's stream and returns without waiting; lane_step() collects the result
// and decides what follows.
//   A = Z_ext(1,1) + Z_int(1) (cached, factored), D = Z_ext(2,2) + Z_int(2)
//   (cached), B = Z_ext(1,2), C = Z_ext(2,1) (assembled for this gap);
//   Y = A^-1 B, Sc = D - C Y, X = -Y Sc^-1, S = 1/4 tr((X sym G2)^H sym G1 X),
//   then the error probe of X.
// BEM_PREC_MIXED runs the elimination in complex64 (solve_mixed); a mixed
// request with tournament pivoting runs the FP64 path.
static int lane_issue(bem_ctx *c, Lane &L, PairGeom &g, bem_body *b1, bem_body *b2, const bem_request *q,
                      bem_result *res) {
  const bem_request qq = *q;
  if (L.restarts == 0) L.t0 = std::chrono::steady_clock::now();
  L.req = qq;
  L.res = res;
  L.sb1 = b1;
  L.sb2 = b2;
  L.pg = &g;
  g.pins++;
  L.stage = ST_SOLVE;
  L.iters = 0;
  double omegaSI = qq.eV * SI_EV / SI_HBAR;
  double omega = omegaSI * qq.a / SI_C;
  const int pivot = (qq.flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
  const int prec = pivot ? BEM_PREC_FP64 : qq.precision;
  L.prec = prec;
  // the low-rank coupling runs in FP64 with block-local pivoting and no
  // refinement (the probe checks it against the dense blocks instead)
  const bool lowrank = (qq.flags & BEM_SOLVE_LOWRANK) && prec == BEM_PREC_FP64 && !pivot;
  L.refine = !lowrank && (qq.flags & BEM_SOLVE_REFINE) != 0;
  L.rank = 0;
  L.ms_compress = 0.0;
  Medium mext = make_medium(1.0, 0.0, 1.0, 0.0, omega);
  DEV(c, hipEventRecord(L.ph[0], L.st));
  const bool twins = same_mesh(b1, b2) && std::memcmp(qq.eps1, qq.eps2, sizeof qq.eps1) == 0 &&
                     std::memcmp(qq.mu1, qq.mu2, sizeof qq.mu1) == 0;
  if (lowrank) {
    // both self blocks factored (LU(D) serves the Woodbury solves); twin
    // bodies share one entry, since A and D are then the same matrix
    CHK(self_get(c, L, b2, qq.eps2, qq.mu2, omega, 0, 0, prec, nullptr, &L.e2));
    if (twins) {
      L.e1 = L.e2;
      L.e2->pins++;
    } else {
      CHK(self_get(c, L, b1, qq.eps1, qq.mu1, omega, 0, 0, prec, nullptr, &L.e1));
    }
  } else {
    // body 2 (raw) first: when both bodies share one mesh and medium, body 1's
    // factored entry is then a copy of it instead of a second assembly
    CHK(self_get(c, L, b2, qq.eps2, qq.mu2, omega, 1, 0, prec, nullptr, &L.e2));
    CHK(self_get(c, L, b1, qq.eps1, qq.mu1, omega, 0, pivot, prec, twins ? L.e2 : nullptr, &L.e1));
  }
  const SelfEntry &E1 = *L.e1, &E2 = *L.e2;
  const int m1 = E1.n, m2 = E2.n;
  if (m1 != 2 * g.E1 || m2 != 2 * (g.top.E - g.E1)) return fail(c, BEM_E_INVALID, "body sizes do not match the pair tree");

  double2 *Bm, *Cm, *Braw, *trace_out;
  int *info;
  unsigned long long *gro;
  ProbeOut *pout;
  CHK(ws_get(c, L, "B", sizeof(double2) * (size_t)m1 * m2, (void **)&Bm));
  CHK(ws_get(c, L, "C", sizeof(double2) * (size_t)m2 * m1, (void **)&Cm));
  if (prec == BEM_PREC_FP64 && !lowrank) CHK(ws_get(c, L, "Braw", sizeof(double2) * (size_t)m1 * m2, (void **)&Braw));
  else Braw = Bm;   // the mixed and low-rank eliminations leave B unfactored
  CHK(ws_get(c, L, "info", sizeof(int) * 4, (void **)&info));
  CHK(ws_get(c, L, "growth", sizeof(unsigned long long) * 3, (void **)&gro));
  CHK(ws_get(c, L, "trace_out", sizeof(double2) * 2, (void **)&trace_out));
  CHK(ws_get(c, L, "probe_out", sizeof(ProbeOut), (void **)&pout));
  DEV(c, hipMemsetAsync(info, 0, sizeof(int) * 4, L.st));
  DEV(c, hipMemsetAsync(gro, 0, sizeof(unsigned long long) * 3, L.st));
  CHK(assemble_cross(c, L, g, mext, Bm, Cm, prec == BEM_PREC_FP64 && !lowrank ? Braw : nullptr, info + 1));
  DEV(c, hipEventRecord(L.ph[1], L.st));

  if (lowrank) {
    const double tol = qq.rank_tol > 0 ? qq.rank_tol : BEM_LOWRANK_TOL;
    const int rmax = ((int)(BEM_LOWRANK_MAX_FRAC * std::min(m1, m2)) / LR_BLK) * LR_BLK;
    if (rmax >= LR_BLK) {
      auto &S = L.lr;
      S = {};
      S.B = Bm;
      S.m1 = m1;
      S.m2 = m2;
      S.rmax = rmax;
      S.nbmax = std::min(rmax / LR_BLK, LR_MAX_BLOCKS);
      S.tol = tol;
      S.t0 = std::chrono::steady_clock::now();
      CHK(ws_get(c, L, "lr_Q", sizeof(double2) * (size_t)m1 * rmax, (void **)&S.Q));
      CHK(ws_get(c, L, "lr_Vh", sizeof(double2) * (size_t)rmax * m2, (void **)&S.Vh));
      if (qq.flags & BEM_LOWRANK_ACA) {   // the reference's route, run to its end here
        int r = -1;
        CHK(lr_compress_aca(c, L, Bm, m1, m2, tol, rmax, S.Q, S.Vh, rmax, &r));
        L.ms_compress = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - S.t0).count();
        if (r > 0) return issue_solve(c, L, r, true);
      } else {
        // first chunk: what the pair needed last time; lane_step() reads the
        // estimates and issues more blocks or the solve
        CHK(lr_rand_issue(c, L, std::max(g.lr_blocks, LR_CHUNK)));
        L.stage = ST_COMPRESS;
        L.busy = true;
        return 0;
      }
    }
    CHK(lr_dense_fallback(c, L));
  }
  return issue_solve(c, L, 0, false);
}

// Issues one refinement step (refine_step) of the lane's request, whose solve
// stage left its factors and X in the lane's workspaces.
static int lane_refine(bem_ctx *c, Lane &L) {
  const SelfEntry &E1 = *L.e1, &E2 = *L.e2;
  const int m1 = E1.n, m2 = E2.n;
  const bool first = L.iters == 0;
  L.iters++;
  L.stage = ST_REFINE;
  double2 *Bm, *Cm, *X2, *trace_out;
  double *rfull;
  ProbeOut *pout;
  int *ipS;
  CHK(ws_get(c, L, "B", sizeof(double2) * (size_t)m1 * m2, (void **)&Bm));
  CHK(ws_get(c, L, "C", sizeof(double2) * (size_t)m2 * m1, (void **)&Cm));
  CHK(ws_get(c, L, "X2_64", sizeof(double2) * (size_t)m2 * m2, (void **)&X2));
  CHK(ws_get(c, L, "trace_out", sizeof(double2) * 2, (void **)&trace_out));
  CHK(ws_get(c, L, "probe_out", sizeof(ProbeOut), (void **)&pout));
  CHK(ws_get(c, L, "rfull", sizeof(double), (void **)&rfull));
  CHK(ws_get(c, L, "ipiv", sizeof(int) * (size_t)(m2 + 64), (void **)&ipS));
  if (L.prec == BEM_PREC_MIXED) {
    float2 *S32, *invS, *C32, *Y32;
    double2 *Xn;
    CHK(ws_get(c, L, "S32", sizeof(float2) * (size_t)m2 * m2, (void **)&S32));
    CHK(ws_get(c, L, "invS32", inv_bytes<float2>(m2), (void **)&invS));
    CHK(ws_get(c, L, "C32", sizeof(float2) * (size_t)m2 * m1, (void **)&C32));
    CHK(ws_get(c, L, "Y32", sizeof(float2) * (size_t)m1 * m2, (void **)&Y32));
    CHK(ws_get(c, L, "Xn64", sizeof(double2) * (size_t)m1 * m2, (void **)&Xn));
    const LuRef<float2> A{E1.Z32, m1, E1.ipiv, E1.inv32, true}, S{S32, m2, ipS, invS, true};
    CHK(refine_step<float2>(c, L, E1, E2, A, S, C32, Y32, Bm, Cm, Xn, X2, first, trace_out, pout, rfull));
  } else {
    const int pivot = (L.req.flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
    double2 *Sc, *invS, *Y, *Braw;
    CHK(ws_get(c, L, "Sc", sizeof(double2) * (size_t)m2 * m2, (void **)&Sc));
    CHK(ws_get(c, L, "invS", inv_bytes(m2), (void **)&invS));
    CHK(ws_get(c, L, "Y64", sizeof(double2) * (size_t)m1 * m2, (void **)&Y));
    CHK(ws_get(c, L, "Braw", sizeof(double2) * (size_t)m1 * m2, (void **)&Braw));
    const LuRef<double2> A{E1.Z, m1, E1.ipiv, E1.inv, !E1.pivot}, S{Sc, m2, ipS, invS, !pivot};
    CHK(refine_step<double2>(c, L, E1, E2, A, S, Cm, Y, Braw, Cm, Bm, X2, first, trace_out, pout, rfull));
  }
  DEV(c, hipMemcpyAsync(L.h_trace, trace_out, sizeof(double2) * 2, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_probe, pout, sizeof(ProbeOut), hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(L.h_rfull, rfull, sizeof(double), hipMemcpyDeviceToHost, L.st));
  L.busy = true;
  return 0;
}

// Growth of the block-local LU: max(1, max|L|) max|U| / max|A| (|re|+|im|
// norms), the factor by which |L||U| can exceed |A| in the backward error.
// Partial pivoting over whole columns keeps |L| <= 1; pivots confined to a
// diagonal block can leave large multipliers below it without any growth of
// U, which the max|L| term catches.
static double growth_ratio(const unsigned long long *g) {
  double u, a, l;
  std::memcpy(&u, &g[0], sizeof u);
  std::memcpy(&a, &g[1], sizeof a);
  std::memcpy(&l, &g[2], sizeof l);
  double r = a > 0 ? std::max(1.0, l) * u / a : 0.0;
  return std::isfinite(r) ? r : INFINITY;
}

static double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

// Waits for the lane's stage to finish and reads it into *L.res: the solve
// stage fills every field, a refinement step updates the trace, the probe and
// the residual.  Returns 0 or the stage's error (also in res->status).
static int lane_collect(bem_ctx *c, Lane &L) {
  L.busy = false;
  hipError_t se = hipStreamSynchronize(L.st);
  const SelfEntry *E1 = L.e1, *E2 = L.e2;
  bem_result *res = L.res;
  const bem_request *q = &L.req;
  if (se != hipSuccess) {
    res->status = BEM_E_DEVICE;
    return fail(c, BEM_E_DEVICE, "solve failed on the device: %s", hipGetErrorString(se));
  }
  ev_collect(c, L);
  const double omega = q->eV * SI_EV / SI_HBAR * q->a / SI_C;
  const double2 tr = L.h_trace[0];
  res->omega = omega;
  res->ReS = 0.25 * tr.x;
  res->ImS = 0.25 * tr.y;
  res->theta1 = theta(omega, q->T1);
  res->theta2 = theta(omega, q->T2);
  res->P = res->ReS * (res->theta1 - res->theta2);
  res->precision = L.prec;
  res->iters = L.iters;
  res->restarts = L.restarts;
  res->pivoting = (q->flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
  const ProbeOut &p = *L.h_probe;
  const double xm = bits_to_double(p.xmax), dm = bits_to_double(p.dmax);
  res->probe_err = xm > 0 ? dm / xm : INFINITY;
  if (!std::isfinite(res->probe_err)) res->probe_err = INFINITY;
  res->residual = p.ww > 0 ? std::sqrt(p.rr / p.ww) : INFINITY;
  // first-order change of Re S: 1/2 Re t, t ~ (1 / (k s^2)) sum_j h_j with
  // s^2 = ||w||^2 / (m2 k) the sample's mean |w_i|^2; the stated error sums
  // |h_j| (an estimate that errs on the high side of |t|)
  {
    const int m2 = E2 ? E2->n : 1;
    double habs = 0.0, hre = 0.0;
    for (int j = 0; j < BEM_PROBE_COLS; j++) {
      habs += std::hypot(p.h[j].x, p.h[j].y);
      hre += p.h[j].x;
    }
    const double scale = p.ww > 0 ? (double)m2 / p.ww : INFINITY;   // 1 / (k s^2)
    const double aS = std::fabs(res->ReS);
    res->rel_err_est = aS > 0 ? 0.5 * habs * scale / aS : INFINITY;
    res->err_first_order = aS > 0 ? 0.5 * hre * scale / aS : INFINITY;
    if (!std::isfinite(res->rel_err_est)) res->rel_err_est = INFINITY;
  }
  res->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - L.t0).count();
  if (L.stage == ST_REFINE) {
    res->full_residual = std::sqrt(*L.h_rfull / (double)(E2 ? E2->n : 1));
    res->status = 0;
    return 0;
  }
  float ta = 0, tf = 0, tp = 0;
  (void)hipEventElapsedTime(&ta, L.ph[0], L.ph[1]);
  (void)hipEventElapsedTime(&tf, L.ph[1], L.ph[2]);
  (void)hipEventElapsedTime(&tp, L.ph[2], L.ph[3]);
  res->ms_assemble = ta;
  res->ms_factor = tf;
  res->ms_products = tp;
  res->full_residual = -1.0;
  res->status = 0;
  res->growth = 0.0;
  res->rank = L.rank;
  res->ms_compress = L.ms_compress;
  int sing_a = E1 ? E1->h->info[0] : 0, ovf = L.h_info[1] || (E1 && E1->h->info[1]) || (E2 && E2->h->info[1]);
  if (!res->pivoting && E1) res->growth = std::max(growth_ratio(E1->h->growth), growth_ratio(L.h_growth));
  if (E2 && E2 != E1 && E2->role == 0) {   // a factored D (low-rank coupling)
    sing_a = sing_a || E2->h->info[0];
    res->growth = std::max(res->growth, growth_ratio(E2->h->growth));
  }
  c->t_last[0] = ta;
  c->t_last[4] = tf;
  c->t_last[5] = tp;
  if (ovf) {
    res->status = BEM_E_DEVICE;
    return fail(c, BEM_E_DEVICE, "36-point pair queue overflowed");
  }
  if (sing_a || L.h_info[0]) {
    res->status = BEM_E_SINGULAR;
    return fail(c, BEM_E_SINGULAR, "exactly singular pivot (A: %d, Schur: %d)", sing_a, L.h_info[0]);
  }
  if (!std::isfinite(res->ReS) || !std::isfinite(res->ImS)) {
    res->status = BEM_E_SINGULAR;
    return fail(c, BEM_E_SINGULAR, "non-finite spectral power: the system is singular or not finite "
                                   "(coincident bodies?)");
  }
  return 0;
}

// What follows a finished stage (lane_step).
enum { STEP_DONE = 0, STEP_MORE = 1, STEP_RESTART = 2 };

// Collects the lane's finished stage and decides what follows:
//  - Recovery: a block-local solve that met an exactly zero pivot, or (FP64)
//    whose probe estimates an error of Re S above BEM_PROBE_LIMIT (a tenth
//    of it when the growth exceeds BEM_GROWTH_LIMIT), or (complex64) whose
//    growth exceeds BEM_GROWTH_LIMIT, is solved again with tournament
//    pivoting and refinement (STEP_RESTART: the lane is released and L.req
//    holds the request with its new flags);
//  - with refinement requested, a result whose relative residual is above
//    BEM_REFINE_TOL gets another refinement step on the lane (STEP_MORE),
//    at most BEM_REFINE_MAX; a mixed request that reaches the cap is solved
//    again on the FP64 path (STEP_RESTART);
//  - otherwise the result is final and the lane is released (STEP_DONE, *err
//    the request's status).
static int lane_step(bem_ctx *c, Lane &L, int *err) {
  if (L.stage == ST_COMPRESS) {
    // range-finder blocks done: more blocks, the low-rank solve, or the
    // dense solve when the block does not compress within rmax
    L.busy = false;
    hipError_t se = hipStreamSynchronize(L.st);
    int e = 0;
    if (se != hipSuccess) {
      e = fail(c, BEM_E_DEVICE, "compression failed on the device: %s", hipGetErrorString(se));
    } else {
      bool done;
      int rk;
      lr_rand_eval(L, &done, &rk);
      if (!done) {
        e = lr_rand_issue(c, L, LR_CHUNK);
        if (!e) {
          L.busy = true;
          return STEP_MORE;
        }
      } else {
        L.ms_compress = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - L.lr.t0).count();
        if (rk > 0) {
          L.pg->lr_blocks = rk / LR_BLK;
          e = issue_solve(c, L, rk, false);
        } else {
          e = lr_dense_fallback(c, L);
          if (!e) e = issue_solve(c, L, 0, false);
        }
        if (!e) return STEP_MORE;
      }
    }
    (void)hipStreamSynchronize(L.st);
    L.busy = false;
    L.res->status = e;
    *err = e;
    self_release(c, L);
    L.stage = ST_IDLE;
    return STEP_DONE;
  }
  int r = lane_collect(c, L);
  bem_result *res = L.res;
  *err = r;
  const bool tourn = (L.req.flags & BEM_SOLVE_TOURNAMENT) != 0;
  const bool recover = !(L.req.flags & BEM_SOLVE_NORECOVER);
  // growth above the limit makes the probe's limit ten times tighter rather
  // than forcing a re-solve: the probe measures the result itself
  const double accept = L.req.err_tol > 0 ? L.req.err_tol : BEM_PROBE_LIMIT;
  const double plimit = res->growth <= BEM_GROWTH_LIMIT ? accept : 0.1 * accept;
  bool restart = false;
  if (recover && L.stage == ST_SOLVE && L.rank > 0 && L.restarts < 2 &&
      (r == BEM_E_SINGULAR || (!r && !(res->rel_err_est <= plimit)))) {
    // a low-rank result the probe rejects (or a singular factor): the dense solve
    L.req.flags &= ~BEM_SOLVE_LOWRANK;
    restart = true;
  } else if (recover && L.stage == ST_SOLVE && !tourn && L.restarts < 2 &&
      (r == BEM_E_SINGULAR ||
       (!r && (L.prec == BEM_PREC_FP64 ? !(res->rel_err_est <= plimit) : !(res->growth <= BEM_GROWTH_LIMIT))))) {
    L.req.flags |= BEM_SOLVE_TOURNAMENT | BEM_SOLVE_REFINE;
    L.req.precision = BEM_PREC_FP64;
    restart = true;
  } else if (!r && L.refine && !(res->residual <= BEM_REFINE_TOL)) {
    if (L.iters < BEM_REFINE_MAX) {
      int e = lane_refine(c, L);
      if (!e) return STEP_MORE;
      (void)hipStreamSynchronize(L.st);
      L.busy = false;
      res->status = e;
      *err = e;
    } else if (recover && L.prec == BEM_PREC_MIXED && L.restarts < 2) {
      L.req.precision = BEM_PREC_FP64;
      restart = true;
    }
  }
  self_release(c, L);
  L.stage = ST_IDLE;
  if (restart) {
    L.restarts++;
    *err = 0;
    return STEP_RESTART;
  }
  return STEP_DONE;
}

// Abandons a partially issued stage: waits for what was issued and releases
// the lane.
static void lane_abort(bem_ctx *c, Lane &L) {
  (void)hipStreamSynchronize(L.aux);
  L.probe_async = false;
  (void)hipStreamSynchronize(L.st);
  L.busy = false;
  L.stage = ST_IDLE;
  self_release(c, L);
}

// Drives a synchronous caller's lane (no ticket) until its request is final:
// collects each stage, issues refinement steps, and solves restarted requests
// again on the same lane.  Returns the request's status.
static int lane_drive(bem_ctx *c, Lane &L) {
  for (;;) {
    if (!L.busy) {
      self_release(c, L);
      return 0;
    }
    int err = 0;
    int st = lane_step(c, L, &err);
    if (st == STEP_MORE) continue;
    if (st == STEP_DONE) return err;
    const bem_request q = L.req;
    PairGeom *g = nullptr;
    int e = pair_geom(c, L.sb1, L.sb2, q.gap, &g);
    if (!e) e = lane_issue(c, L, *g, L.sb1, L.sb2, &q, L.res);
    if (e) {
      lane_abort(c, L);
      L.res->status = e;
      return e;
    }
  }
}

static int check_request(bem_ctx *c, const bem_request *q) {
  if (!(q->eV > 0) || !(q->a > 0)) return fail(c, BEM_E_INVALID, "eV and a must be positive");
  if (q->precision != BEM_PREC_FP64 && q->precision != BEM_PREC_MIXED)
    return fail(c, BEM_E_UNSUPPORTED, "precision mode %d not available", q->precision);
  return 0;
}


// ------------------------------------------------------------------ serving
static bool self_cached(bem_ctx *c, const void *b, const double *eps, const double *mu, double omega, int role,
                        int pivot, int prec) {
  const double key[5] = {eps[0], eps[1], mu[0], mu[1], omega};
  for (auto &e : c->selfs)
    if (e.body == b && e.role == role && e.pivot == pivot && e.prec == prec &&
        std::memcmp(e.key, key, sizeof key) == 0)
      return true;
  return false;
}

// Device bytes a request would newly allocate on lane L: the lane's
// workspaces beyond what it holds, plus the self entries not yet cached.
static size_t footprint(bem_ctx *c, const Lane &L, const Pending &p) {
  const bemgeo::BodyMesh &M1 = p.b1->mesh, &M2 = p.b2->mesh;
  const size_t m1 = 2 * (size_t)M1.ne, m2 = 2 * (size_t)M2.ne, T1 = M1.nt, T2 = M2.nt;
  const bem_request &q = p.req;
  const double omega = q.eV * SI_EV / SI_HBAR * q.a / SI_C;
  const int pivot = (q.flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
  const int prec = pivot ? BEM_PREC_FP64 : q.precision;
  const bool h1 = self_cached(c, p.b1, q.eps1, q.mu1, omega, 0, pivot, prec);
  const bool h2 = self_cached(c, p.b2, q.eps2, q.mu2, omega, 1, 0, prec);
  size_t selfb = (h1 ? 0 : self_entry_bytes((int)m1, 0, prec)) + (h2 ? 0 : self_entry_bytes((int)m2, 1, prec));
  size_t pairs = std::max(T1 * T2, std::max(h1 ? 0 : T1 * T1, h2 ? 0 : T2 * T2));
  // lane workspaces in complex128 units: B, C, the unfactored B, B1 and Sc
  // (FP64); B, C, -X in FP64 plus complex64 B, C, B1 and Sc (mixed); Y,
  // Sc^-1 and the residual blocks with refinement (complex64 copies too when
  // mixed); the probe's columns are negligible
  const bool refine = (q.flags & BEM_SOLVE_REFINE) != 0;
  const double mm = (double)m1 * m2, m22 = (double)m2 * m2;
  const double w16 = prec == BEM_PREC_MIXED ? 4.5 * mm + 0.5 * m22 + (refine ? 3.0 * mm + 2.5 * m22 : 0.0)
                                            : 4.0 * mm + m22 + (refine ? 2.0 * mm + 2.0 * m22 : 0.0);
  size_t ws = (size_t)(sizeof(double2) * w16) + 8 * (size_t)BEM_TILE_PLANES * pairs +
              sizeof(int2) * std::min<size_t>(pairs, (size_t)1 << 24) + inv_bytes((int)m2) +
              sizeof(double2) * 64 * std::max(m1, m2) + sizeof(double2) * (bem_htrace_partials((int)m2) + 1) +
              sizeof(double) * bem_close_mom_doubles((unsigned int)CLOSE_CHUNK);
  return (ws > L.ws_bytes ? ws - L.ws_bytes : 0) + selfb;
}

// Frees cached self entries (least recently used, unpinned) and the
// workspaces of idle lanes other than `keep` until `need` more bytes fit the
// budget or nothing more can go.
static void make_room(bem_ctx *c, size_t need, const Lane *keep) {
  while (c->budget && c->bytes + need > c->budget) {
    auto victim = c->selfs.end();
    for (auto it = c->selfs.begin(); it != c->selfs.end(); ++it)
      if (it->pins == 0 && (victim == c->selfs.end() || it->stamp < victim->stamp)) victim = it;
    if (victim != c->selfs.end()) {
      self_free(c, *victim);
      c->selfs.erase(victim);
      continue;
    }
    Lane *idle = nullptr;
    for (auto &l : c->lanes)
      if (&l != keep && !l.busy && l.ws_bytes > 0) idle = &l;
    if (!idle) return;
    ws_trim(c, *idle);
  }
}

static double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

static void lane_done_cb(void *arg);

// Finishes what the lane has in flight.  A synchronous caller's lane is
// driven to its final result (lane_drive).  A served request's stage is
// collected: a refinement step that follows is issued on the lane (with
// `block`, waited for in turn; without, its completion callback is enqueued
// and the call returns), a restarted request goes back to the head of the
// queue, and a final result goes to the table of finished tickets; its errors
// are reported through the ticket, not as this call's return value.
static int lane_retire(bem_ctx *c, Lane &L, bool block = true) {
  if (!L.ticket) return lane_drive(c, L);
  for (;;) {
    const uint64_t t = L.ticket;
    int err = 0;
    const int st = L.busy ? lane_step(c, L, &err) : STEP_DONE;
    if (st == STEP_MORE) {
      if (block) continue;
      __atomic_store_n(&L.done_flag, 0, __ATOMIC_RELEASE);
      if (hipLaunchHostFunc(L.st, lane_done_cb, &L) == hipSuccess) return 0;
      continue;   // no callback: wait for the step here
    }
    L.ticket = 0;
    bem_result res = L.sres;
    if (err && !res.status) res.status = err;
    auto now = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(c->qmu);
    c->running--;
    if (st == STEP_RESTART) {
      Pending p{t, L.sb1, L.sb2, L.req, L.t_submit, L.restarts};
      c->queue.push_front(p);
      c->kick = true;
      c->qcv.notify_one();
      return 0;
    }
    res.ms_queue = ms_between(L.t_submit, L.t_issue);
    res.ms_total = ms_between(L.t_submit, now);
    c->done[t] = res;
    c->live.erase(t);
    c->done_cv.notify_all();
    return 0;
  }
}

// Host callback enqueued behind a served solve: marks the lane done and wakes
// the scheduler (no HIP calls here).
static void lane_done_cb(void *arg) {
  Lane *L = (Lane *)arg;
  __atomic_store_n(&L->done_flag, 1, __ATOMIC_RELEASE);
  bem_ctx *c = L->ctx;
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    c->kick = true;
  }
  c->qcv.notify_one();
}

static void finish_ticket(bem_ctx *c, uint64_t t, int status) {
  std::lock_guard<std::mutex> lk(c->qmu);
  bem_result r;
  std::memset(&r, 0, sizeof r);
  r.status = status;
  c->done[t] = r;
  c->live.erase(t);
  c->done_cv.notify_all();
}

// Estimated device milliseconds of a request: the dense solve's flops at a
// nominal 50 TFLOP/s (the factorisation of A is skipped when it is cached).
static double est_ms(bem_ctx *c, const Pending &p) {
  const double m1 = 2.0 * p.b1->mesh.ne, m2 = 2.0 * p.b2->mesh.ne;
  const bem_request &q = p.req;
  const double omega = q.eV * SI_EV / SI_HBAR * q.a / SI_C;
  const int pivot = (q.flags & BEM_SOLVE_TOURNAMENT) ? 1 : 0;
  const int prec = pivot ? BEM_PREC_FP64 : q.precision;
  double f = 8.0 * (2 * m1 * m1 * m2 + 3 * m1 * m2 * m2) + 8.0 / 3.0 * m2 * m2 * m2;
  if (!self_cached(c, p.b1, q.eps1, q.mu1, omega, 0, pivot, prec)) f += 8.0 / 3.0 * m1 * m1 * m1;
  if (prec == BEM_PREC_MIXED) f *= (q.flags & BEM_SOLVE_REFINE) ? 2.0 : 0.5;
  return f / 50e12 * 1e3;
}

// Index of the queued request to admit next: the smallest estimated run time
// minus time already waited (shortest job first with aging, so a long
// request is admitted once it has waited about its own run time longer than
// the shorter ones behind it).  Caller holds c->mu and c->qmu.
static size_t pick_next(bem_ctx *c) {
  auto now = std::chrono::steady_clock::now();
  size_t best = 0;
  double key = 0;
  for (size_t i = 0; i < c->queue.size(); i++) {
    const Pending &p = c->queue[i];
    double k = est_ms(c, p) - ms_between(p.t_submit, now);
    if (i == 0 || k < key) {
      best = i;
      key = k;
    }
  }
  return best;
}

// Admits queued requests while a lane is free and the chosen request's
// footprint fits the budget (caller holds c->mu).
static void admit(bem_ctx *c) {
  for (;;) {
    Pending p;
    size_t idx;
    {
      std::lock_guard<std::mutex> lk(c->qmu);
      if (c->queue.empty()) return;
      idx = pick_next(c);
      p = c->queue[idx];
    }
    Lane *L = nullptr;
    for (auto &l : c->lanes)
      if (!l.busy) {
        L = &l;
        break;
      }
    if (!L) return;
    int e = check_request(c, &p.req);
    if (!e && c->budget) {
      size_t extra = footprint(c, *L, p);
      if (c->bytes + extra > c->budget) make_room(c, extra, L);
      if (c->bytes + extra > c->budget) {
        for (auto &l : c->lanes)
          if (l.busy) return;   // wait until a running request frees its lane
        e = fail(c, BEM_E_NOMEM, "request needs %zu more device bytes; budget %zu, held %zu", extra, c->budget,
                 c->bytes);
      }
    }
    {
      std::lock_guard<std::mutex> lk(c->qmu);
      c->queue.erase(c->queue.begin() + idx);
      if (!e) c->running++;
    }
    if (e) {
      finish_ticket(c, p.ticket, e);
      continue;
    }
    PairGeom *g = nullptr;
    e = pair_geom(c, p.b1, p.b2, p.req.gap, &g);
    if (!e) {
      L->ticket = p.ticket;
      L->sb1 = p.b1;
      L->sb2 = p.b2;
      L->t_submit = p.t_submit;
      L->t_issue = std::chrono::steady_clock::now();
      std::memset(&L->sres, 0, sizeof L->sres);
      __atomic_store_n(&L->done_flag, 0, __ATOMIC_RELEASE);
      L->restarts = p.restarts;
      e = lane_issue(c, *L, *g, p.b1, p.b2, &p.req, &L->sres);
      if (!e && hipLaunchHostFunc(L->st, lane_done_cb, L) != hipSuccess)
        e = fail(c, BEM_E_DEVICE, "host callback could not be enqueued");
      if (e) {
        lane_abort(c, *L);
        L->ticket = 0;
      }
    }
    if (e) {
      {
        std::lock_guard<std::mutex> lk(c->qmu);
        c->running--;
      }
      finish_ticket(c, p.ticket, e);
    }
  }
}

static void sched_main(bem_ctx *c) {
  (void)hipSetDevice(c->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(c->qmu);
      c->qcv.wait_for(lk, std::chrono::milliseconds(5), [c] { return c->stop || c->kick; });
      if (c->stop) return;
      c->kick = false;
    }
    std::lock_guard<std::recursive_mutex> g(c->mu);
    for (auto &L : c->lanes)
      if (L.ticket && __atomic_load_n(&L.done_flag, __ATOMIC_ACQUIRE)) (void)lane_retire(c, L, false);
    admit(c);
  }
}


// ---------------------------------------------------------------- autotuner
// Representative shape (m, n, k) of each GEMM class.
static const int TUNE_SHAPE[3][3] = {{4096, 64, 64}, {4096, 4096, 64}, {4096, 4096, 256}};
static const int ZCANDS[] = {1032, 2032, 1033, 1034, 1064, 2064, 1065, 2034, 1066, 1128, 1129, 32, 64, 256};
static const int CCANDS[] = {1, 2, 3, 4, 5, 11, 12, 13, 14, 15, 16, 17, 21, 23, 26};

// Mean milliseconds of `reps` launches of one tile variant on lane 0.
static int tune_time(bem_ctx *c, Lane &L, bool c64, int tile, int cls, void *A, void *B, void *Cm, int reps,
                     double *ms) {
  const int m = TUNE_SHAPE[cls][0], n = TUNE_SHAPE[cls][1], k = TUNE_SHAPE[cls][2];
  hipEvent_t e0, e1;
  DEV(c, hipEventCreate(&e0));
  DEV(c, hipEventCreate(&e1));
  int err = 0;
  for (int rep = 0; rep <= reps && !err; rep++) {
    if (rep == 1) err = (int)hipEventRecord(e0, L.st);
    if (!err && c64)
      err = bem_cgemm_tiled(tile, m, n, k, make_float2(1, 0), (const float2 *)A, m, (const float2 *)B, k,
                            make_float2(1, 0), (float2 *)Cm, m, L.st);
    else if (!err)
      err = bem_zgemm_tiled(tile, m, n, k, make_double2(1, 0), (const double2 *)A, m, (const double2 *)B, k,
                            make_double2(1, 0), (double2 *)Cm, m, L.st);
  }
  float t = 0;
  if (!err) err = (int)hipEventRecord(e1, L.st);
  if (!err) err = (int)hipEventSynchronize(e1);
  if (!err) err = (int)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err) return fail(c, BEM_E_DEVICE, "autotune launch of tile %d failed (%d)", tile, err);
  *ms = t / reps;
  return 0;
}

static bool tune_load(const char *path, const std::string &arch, int *tiles6) {
  FILE *f = path ? std::fopen(path, "r") : nullptr;
  if (!f) return false;
  char buf[4096];
  size_t n = std::fread(buf, 1, sizeof buf - 1, f);
  std::fclose(f);
  buf[n] = 0;
  std::string t(buf);
  size_t pa = t.find("\"arch\": \""), pz = t.find("\"zgemm\": ["), pc = t.find("\"cgemm\": [");
  if (pa == std::string::npos || pz == std::string::npos || pc == std::string::npos) return false;
  size_t qa = t.find('"', pa + 9);
  if (qa == std::string::npos || t.substr(pa + 9, qa - pa - 9) != arch) return false;
  int v[6];
  if (std::sscanf(t.c_str() + pz + 10, "%d , %d , %d", &v[0], &v[1], &v[2]) != 3) return false;
  if (std::sscanf(t.c_str() + pc + 10, "%d , %d , %d", &v[3], &v[4], &v[5]) != 3) return false;
  for (int i = 0; i < 3; i++)
    if (!bem_zgemm_tile_valid(v[i]) || !bem_cgemm_tile_valid(v[3 + i])) return false;
  std::memcpy(tiles6, v, sizeof v);
  return true;
}

static void tune_store(const char *path, const std::string &arch, const int *tiles6, const double (*ms)[16],
                       const int *nz, const int *nc) {
  FILE *f = path ? std::fopen(path, "w") : nullptr;
  if (!f) return;
  std::fprintf(f, "{\"arch\": \"%s\", \"zgemm\": [%d, %d, %d], \"cgemm\": [%d, %d, %d],\n", arch.c_str(), tiles6[0],
               tiles6[1], tiles6[2], tiles6[3], tiles6[4], tiles6[5]);
  std::fprintf(f, " \"classes\": [\"narrow 4096x64x64\", \"rank update 4096x4096x64\", \"general 4096x4096x256\"],\n");
  for (int c64 = 0; c64 < 2; c64++) {
    std::fprintf(f, " \"%s_ms\": [", c64 ? "cgemm" : "zgemm");
    for (int cls = 0; cls < 3; cls++) {
      std::fprintf(f, "%s{", cls ? ", " : "");
      const int nn = c64 ? *nc : *nz;
      for (int i = 0; i < nn; i++)
        std::fprintf(f, "%s\"%d\": %.4f", i ? ", " : "", c64 ? CCANDS[i] : ZCANDS[i], ms[3 * c64 + cls][i]);
      std::fprintf(f, "}");
    }
    std::fprintf(f, "]%s\n", c64 ? "" : ",");
  }
  std::fprintf(f, "}\n");
  std::fclose(f);
}

// ------------------------------------------------- allocation-free LU entry
// Workspace of bem_zgetrf_local_noalloc in a fixed device array: the
// diagonal-block inverses of an n <= BEM_NOALLOC_MAX_N factorisation, then its
// zero-pivot flag (at +0) and growth maxima (at +64).  Each GPU holds its own
// copy until the process ends; host code reaches it only through the address
// that hipGetSymbolAddress returns with that device current (the array's name
// is a host-side stand-in), looked up once per device by bem_ctx_create.
static const size_t NOALLOC_INV_BYTES = sizeof(double2) * 2 * 64 * 64 * (size_t)((BEM_NOALLOC_MAX_N + 63) / 64);
static const size_t NOALLOC_BYTES = NOALLOC_INV_BYTES + 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];   // per device ordinal

static int scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= 64) return BEM_E_INVALID;
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BEM_E_DEVICE;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

// ------------------------------------------------- dipole sources and flux
// Right-hand sides of ns dipoles on tree d (B: 2E x ns, leading dimension
// ldb, device): toff[b]..toff[b+1] are body b's triangles, eps/mu (re, im)
// per body; the exterior is vacuum.
static int dipole_rhs(bem_ctx *c, Lane &L, const DevTree &d, const std::vector<int> &toff, const double *eps,
                      const double *mu, double omega, int ns, const bem_dipole *src, double2 *B, int64_t ldb) {
  const int nb = (int)toff.size() - 1;
  std::vector<DipoleSrc> h(ns);
  int tmax = 0;
  for (int s = 0; s < ns; s++) {
    const bem_dipole &q = src[s];
    if (q.region < -1 || q.region >= nb) return fail(c, BEM_E_INVALID, "dipole %d: region %d", s, (int)q.region);
    DipoleSrc &o = h[s];
    std::memset(&o, 0, sizeof o);
    for (int k = 0; k < 3; k++) {
      o.p[k] = q.p[k];
      o.j[k] = q.j[k];
    }
    if (q.region < 0) {
      o.md = make_medium(1.0, 0.0, 1.0, 0.0, omega);
      o.t_lo = 0;
      o.t_hi = d.T;
      o.external = 1;
    } else {
      const int b = q.region;
      o.md = make_medium(eps[2 * b], eps[2 * b + 1], mu[2 * b], mu[2 * b + 1], omega);
      o.t_lo = toff[b];
      o.t_hi = toff[b + 1];
      o.external = 0;
    }
    tmax = std::max(tmax, o.t_hi - o.t_lo);
  }
  DipoleSrc *dsrc;
  double2 *tb;
  CHK(ws_get(c, L, "dip_src", sizeof(DipoleSrc) * (size_t)std::max(ns, 1), (void **)&dsrc));
  CHK(ws_get(c, L, "dip_tb", sizeof(double2) * 6 * (size_t)std::max(ns, 1) * std::max(d.T, 1), (void **)&tb));
  DEV(c, hipMemcpyAsync(dsrc, h.data(), sizeof(DipoleSrc) * ns, hipMemcpyHostToDevice, L.st));
  int e = bem_launch_dipole_rhs(d.tg, d.T, d.rw, d.E, dsrc, ns, tmax, tb, B, ldb, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "dipole launch: %s", hipGetErrorString((hipError_t)e));
  // the host records must outlive the copy
  DEV(c, hipStreamSynchronize(L.st));
  return 0;
}

// Flux of ncol columns of X (device) through the surface ft: out (host).
static int flux_columns(bem_ctx *c, Lane &L, const std::vector<FluxTri> &ft, double omega,
                        const double2 *X, int64_t ldx, int ncol, double *out) {
  const int nt = (int)ft.size();
  FluxTri *dft;
  double *part, *dout;
  CHK(ws_get(c, L, "flux_tri", sizeof(FluxTri) * (size_t)std::max(nt, 1), (void **)&dft));
  CHK(ws_get(c, L, "flux_part", sizeof(double) * std::max<size_t>(bem_flux_partials(nt, ncol), 1), (void **)&part));
  CHK(ws_get(c, L, "flux_out", sizeof(double) * (size_t)std::max(ncol, 1), (void **)&dout));
  if (nt) DEV(c, hipMemcpyAsync(dft, ft.data(), sizeof(FluxTri) * nt, hipMemcpyHostToDevice, L.st));
  int e = bem_launch_flux(dft, nt, omega, X, ldx, ncol, part, dout, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "flux launch: %s", hipGetErrorString((hipError_t)e));
  DEV(c, hipMemcpyAsync(out, dout, sizeof(double) * ncol, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipStreamSynchronize(L.st));
  return 0;
}

// ------------------------------------------------------------------- C ABI
extern "C" {

int bem_ctx_create(bem_ctx **out, const bem_cfg *cfg) {
  if (!out) return BEM_E_INVALID;
  bem_ctx *c = new bem_ctx();
  // look-ahead needs a second stream per lane; it pays only while every
  // stream gets a hardware queue of its own (GPU_MAX_HW_QUEUES, default 4)
  int hwq = 4;
  if (const char *v = std::getenv("GPU_MAX_HW_QUEUES")) hwq = std::max(1, std::atoi(v));
  int nl = std::min(4, hwq);   // default: one lane per hardware queue
  if (cfg) {
    c->device = cfg->device;
    c->budget = cfg->hbm_budget;
    if (cfg->lanes > 0) nl = std::min(cfg->lanes, 8);
  }
  const int la = cfg ? cfg->lookahead : 0;
  c->lookahead = la > 0 || (la == 0 && 2 * nl <= hwq);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || hipSetDevice(c->device) != hipSuccess) {
    delete c;
    return BEM_E_DEVICE;
  }
  if (scratch_address(c->device, &c->scratch) ||
      hipHostMalloc((void **)&c->h_noalloc, 128, hipHostMallocDefault) != hipSuccess) {
    delete c;
    return BEM_E_DEVICE;
  }
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = 0;
  c->self_cap = (cfg && cfg->cache_bytes) ? cfg->cache_bytes : (size_t)(0.4 * (double)tot);
  c->lanes.resize(nl);
  for (int i = 0; i < nl; i++) {
    Lane &L = c->lanes[i];
    L.id = i;
    L.ctx = c;
    L.tag = "L" + std::to_string(i) + ":";
    bool ok = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&L.aux, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&L.la_a, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&L.la_b, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 4 && ok; k++) ok = hipEventCreate(&L.ph[k]) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_trace, sizeof(double2) * 2, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_info, sizeof(int) * 4, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_growth, sizeof(unsigned long long) * 4, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_probe, sizeof(ProbeOut), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_rfull, sizeof(double), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_lr, sizeof(double) * (LR_MAX_BLOCKS + 1), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&L.h_lri, sizeof(int) * (LR_MAX_BLOCKS + 1), hipHostMallocDefault) == hipSuccess;
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
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    c->stop = true;
  }
  c->qcv.notify_all();
  if (c->sched.joinable()) c->sched.join();
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  for (auto &kv : c->ws)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto &w : c->ws_spare) (void)hipFree(w.first);
  for (auto &kv : c->user_allocs) (void)hipFree(kv.first);
  for (auto &g : c->pairs) free_allocs(c, g.allocs, g.bytes);
  for (auto &e : c->selfs) self_free(c, e);
  for (SelfStatus *h : c->spare_status) (void)hipHostFree(h);
  for (hipEvent_t ev : c->spare_events) (void)hipEventDestroy(ev);
  for (bem_body *b : c->bodies) {
    free_allocs(c, b->allocs, b->bytes);
    b->up = false;
    b->ctx = nullptr;   // bem_body_free then only releases the host object
  }
  for (auto &sg : c->arena) (void)hipFree(sg.base);
  if (c->h_noalloc) (void)hipHostFree(c->h_noalloc);
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
    if (L.h_probe) (void)hipHostFree(L.h_probe);
    if (L.h_rfull) (void)hipHostFree(L.h_rfull);
    if (L.h_lr) (void)hipHostFree(L.h_lr);
    if (L.h_lri) (void)hipHostFree(L.h_lri);
    if (L.st) (void)hipStreamDestroy(L.st);
    if (L.aux) (void)hipStreamDestroy(L.aux);
    if (L.la_a) (void)hipEventDestroy(L.la_a);
    if (L.la_b) (void)hipEventDestroy(L.la_b);
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
  // a triangle of zero area (or a non-finite vertex) has no RWG basis: its
  // integrals divide by the area
  for (int t = 0; t < nt; t++) {
    const double *p0 = xyz + 3 * (size_t)tri[3 * t], *p1 = xyz + 3 * (size_t)tri[3 * t + 1],
                 *p2 = xyz + 3 * (size_t)tri[3 * t + 2];
    const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double v[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
    const double a2 = cx * cx + cy * cy + cz * cz;
    if (!(a2 > 0.0) || !std::isfinite(a2)) {
      delete b;
      return fail(c, BEM_E_MESH, "triangle %d has zero area or a non-finite vertex", t);
    }
  }
  M.ne = ne;
  bemgeo::build_close_pairs(M);
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](const void *p, size_t nbytes) {
    const unsigned char *q = (const unsigned char *)p;
    for (size_t i = 0; i < nbytes; i++) h = (h ^ q[i]) * 1099511628211ull;
  };
  mix(M.xyz.data(), sizeof(double) * M.xyz.size());
  mix(M.tri.data(), sizeof(int32_t) * M.tri.size());
  b->fp = h;
  c->bodies.insert(b);
  *out = b;
  return 0;
}

int bem_body_edges(bem_body *b, int32_t *edges, int *ne) {
  if (!b || !ne) return BEM_E_INVALID;
  *ne = b->mesh.ne;
  if (edges) std::memcpy(edges, b->mesh.edges.data(), sizeof(int32_t) * b->mesh.edges.size());
  return 0;
}

int bem_rwg_tree(const double *xyz, int nv, const int32_t *tri, int nt, int min_size, double min_radius,
                 int32_t *order, int *n, int32_t *leaves, int *nleaves) {
  if (!xyz || !tri || nv <= 0 || nt <= 0 || !order || !n || !leaves || !nleaves || min_size < 0)
    return BEM_E_INVALID;
  std::vector<int32_t> sizes;
  const int rc = bemgeo::rwg_tree_order(nv, xyz, nt, tri, min_size, min_radius, order, *n, sizes);
  if (rc == -1) return BEM_E_INVALID;
  if (rc == -2) return BEM_E_MESH;
  if (rc) return BEM_E_UNSUPPORTED;
  std::memcpy(leaves, sizes.data(), sizeof(int32_t) * sizes.size());
  *nleaves = (int)sizes.size();
  return 0;
}

void bem_body_free(bem_body *b) {
  if (!b) return;
  bem_ctx *c = b->ctx;
  if (c) {
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    for (auto &L : c->lanes) (void)lane_retire(c, L);
    for (auto it = c->pairs.begin(); it != c->pairs.end();) {
      if (it->b1 == b || it->b2 == b) {
        free_allocs(c, it->allocs, it->bytes);
        it = c->pairs.erase(it);
      } else {
        ++it;
      }
    }
    self_drop_body(c, b);
    if (!b->allocs.empty()) free_allocs(c, b->allocs, b->bytes);
    arena_release_empty(c);
    c->bodies.erase(b);
    std::vector<uint64_t> dropped;
    {
      std::lock_guard<std::mutex> lq(c->qmu);
      for (auto it = c->queue.begin(); it != c->queue.end();) {
        if (it->b1 == b || it->b2 == b) {
          dropped.push_back(it->ticket);
          it = c->queue.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (uint64_t t : dropped) finish_ticket(c, t, BEM_E_INVALID);
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
  CHK(lane_retire(c, L));
  std::vector<const bemgeo::BodyMesh *> bm;
  for (int i = 0; i < nb; i++) bm.push_back(&bodies[i]->mesh);
  bemgeo::Tree t;
  bemgeo::build_tree(bm, shifts, t);
  int M = 2 * t.E;
  if (ldz < M) return fail(c, BEM_E_INVALID, "ldz %lld < %d", (long long)ldz, M);
  DevTree d;
  std::vector<std::pair<void *, size_t>> allocs;
  size_t tb = 0;
  int e = upload_tree(c, t, d, allocs, tb);
  Medium md = make_medium(eps[0], eps[1], mu[0], mu[1], omega);
  double *tiles = nullptr;
  size_t npairs = (size_t)t.T * t.T;
  double2 *Zd = nullptr;
  bool dev = flags & BEM_ASM_DEVICE_PTR;
  int64_t ld = dev ? ldz : M;
  if (!e) e = ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * npairs, (void **)&tiles);
  int *flag = nullptr;
  if (!e) e = ws_get(c, L, "info", sizeof(int) * 4, (void **)&flag);
  if (!e && hipMemsetAsync(flag, 0, sizeof(int) * 4, L.st) != hipSuccess) e = fail(c, BEM_E_DEVICE, "memset");
  if (!e) e = assemble_tiles(c, L, d, md, tiles, 0, t.T, 0, t.T, flag + 1);
  if (!e) {
    if (dev) Zd = (double2 *)Z;
    else e = ws_get(c, L, "asm_Z", sizeof(double2) * (size_t)M * M, (void **)&Zd);
  }
  if (!e && bem_launch_gather(tiles, npairs, t.T, 0, 0, d.rw, 0, t.E, 0, t.E, &md, Zd, ld, 0, 0, 0, nullptr, 0, 0, 0,
                             L.st))
    e = fail(c, BEM_E_DEVICE, "gather launch failed");
  if (!e && !dev && hipMemcpy2DAsync(Z, sizeof(double2) * ldz, Zd, sizeof(double2) * M, sizeof(double2) * M, M,
                                     hipMemcpyDeviceToHost, L.st) != hipSuccess)
    e = fail(c, BEM_E_DEVICE, "copy of Z failed");
  int hflag[4] = {0, 0, 0, 0};
  if (!e && hipMemcpyAsync(hflag, flag, sizeof hflag, hipMemcpyDeviceToHost, L.st) != hipSuccess)
    e = fail(c, BEM_E_DEVICE, "flag copy");
  // drained before the tree's blocks go back, also after a failed launch
  const hipError_t se = hipStreamSynchronize(L.st);
  if (!e && se != hipSuccess) e = fail(c, BEM_E_DEVICE, "assembly failed on the device");
  if (!e && hflag[1]) e = fail(c, BEM_E_DEVICE, "36-point pair queue overflowed");
  free_allocs(c, allocs, tb);
  return e;
}

int bem_dipole_rhs(bem_ctx *c, int nb, bem_body *const *bodies, const double *shifts, const double *eps,
                   const double *mu, double omega, int ns, const bem_dipole *src, void *B, int64_t ldb, int flags) {
  if (!c || nb < 1 || nb > 2 || !bodies || !shifts || !eps || !mu || ns < 0 || (ns && (!src || !B)))
    return fail(c, BEM_E_INVALID, "bad arguments");
  if (!(omega > 0)) return fail(c, BEM_E_INVALID, "omega must be positive");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  std::vector<const bemgeo::BodyMesh *> bm;
  for (int i = 0; i < nb; i++) bm.push_back(&bodies[i]->mesh);
  bemgeo::Tree t;
  bemgeo::build_tree(bm, shifts, t);
  const int M = 2 * t.E;
  if (ldb < M) return fail(c, BEM_E_INVALID, "ldb %lld < %d", (long long)ldb, M);
  if (ns == 0) return 0;
  t.close.clear();
  DevTree d;
  std::vector<std::pair<void *, size_t>> allocs;
  size_t tb = 0;
  int e = upload_tree(c, t, d, allocs, tb);
  std::vector<int> toff(t.tri_off);
  toff.push_back(t.T);
  const bool dev = flags & BEM_ASM_DEVICE_PTR;
  double2 *Bd = dev ? (double2 *)B : nullptr;
  const int64_t ld = dev ? ldb : M;
  if (!e && !dev) e = ws_get(c, L, "dip_B", sizeof(double2) * (size_t)M * ns, (void **)&Bd);
  if (!e) e = dipole_rhs(c, L, d, toff, eps, mu, omega, ns, src, Bd, ld);
  if (!e && !dev && hipMemcpy2DAsync(B, sizeof(double2) * ldb, Bd, sizeof(double2) * M, sizeof(double2) * M, ns,
                                     hipMemcpyDeviceToHost, L.st) != hipSuccess)
    e = fail(c, BEM_E_DEVICE, "copy of B failed");
  // drained before the tree's blocks go back, also after a failed launch
  const hipError_t se = hipStreamSynchronize(L.st);
  if (!e && se != hipSuccess) e = fail(c, BEM_E_DEVICE, "dipole sources failed on the device");
  free_allocs(c, allocs, tb);
  return e;
}

int bem_flux(bem_ctx *c, int nb, bem_body *const *bodies, const double *shifts, int body, double omega,
             const void *X, int64_t ldx, int ncol, int flags, double *flux) {
  if (!c || nb < 1 || nb > 2 || !bodies || !shifts || body < 0 || body >= nb || ncol < 0 || (ncol && (!X || !flux)))
    return fail(c, BEM_E_INVALID, "bad arguments");
  if (!(omega > 0)) return fail(c, BEM_E_INVALID, "omega must be positive");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  int off = 0, E = 0;
  for (int i = 0; i < nb; i++) {
    if (i < body) off += bodies[i]->mesh.ne;
    E += bodies[i]->mesh.ne;
  }
  const int M = 2 * E;
  if (ldx < M) return fail(c, BEM_E_INVALID, "ldx %lld < %d", (long long)ldx, M);
  if (ncol == 0) return 0;
  std::vector<FluxTri> ft;
  if (bemgeo::flux_table(bodies[body]->mesh, shifts + 3 * body, off, ft))
    return fail(c, BEM_E_INVALID, "body %d has a triangle edge without an RWG", body);
  const double2 *Xd = (const double2 *)X;
  int64_t ld = ldx;
  if (!(flags & BEM_ASM_DEVICE_PTR)) {
    double2 *tmp;
    CHK(ws_get(c, L, "flux_X", sizeof(double2) * (size_t)M * ncol, (void **)&tmp));
    DEV(c, hipMemcpy2DAsync(tmp, sizeof(double2) * M, X, sizeof(double2) * ldx, sizeof(double2) * M, ncol,
                            hipMemcpyHostToDevice, L.st));
    Xd = tmp;
    ld = M;
  }
  return flux_columns(c, L, ft, omega, Xd, ld, ncol, flux);
}

int bem_dipole_flux(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *q, int ns, const bem_dipole *src,
                    double *flux) {
  if (!c || !b1 || !b2 || !q || ns < 0 || (ns && (!src || !flux))) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  CHK(check_request(c, q));
  if (ns == 0) return 0;
  stats_reset(c);
  PairGeom *g;
  CHK(pair_geom(c, b1, b2, q->gap, &g));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  const double omega = q->eV * SI_EV / SI_HBAR * q->a / SI_C;
  const int T1 = g->T1, T = g->top.T, E1 = g->E1, E = g->top.E;
  const int m1 = 2 * E1, m2 = 2 * (E - E1), M = m1 + m2;
  const Medium mext = make_medium(1.0, 0.0, 1.0, 0.0, omega);
  double2 *W, *Bx, *inv;
  int *ipiv, *info;
  CHK(ws_get(c, L, "dflux_W", sizeof(double2) * (size_t)M * M, (void **)&W));
  CHK(ws_get(c, L, "dflux_B", sizeof(double2) * (size_t)M * ns, (void **)&Bx));
  CHK(ws_get(c, L, "dflux_ipiv", sizeof(int) * (size_t)(M + 64), (void **)&ipiv));
  CHK(ws_get(c, L, "dflux_inv", inv_bytes(M), (void **)&inv));
  CHK(ws_get(c, L, "info", sizeof(int) * 4, (void **)&info));
  DEV(c, hipMemsetAsync(info, 0, sizeof(int) * 4, L.st));
  // diagonal blocks: Z_ext(b, b) + Z_int(b) on the body's own tree
  bem_body *bs[2] = {b1, b2};
  const double *eps[2] = {q->eps1, q->eps2}, *mu[2] = {q->mu1, q->mu2};
  for (int b = 0; b < 2; b++) {
    const DevTree *dt;
    CHK(body_tree(c, bs[b], &dt));
    const int Tb = dt->T, o = b ? m1 : 0;
    const size_t np = (size_t)Tb * Tb;
    double *tiles;
    CHK(ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * std::max<size_t>(np, 1), (void **)&tiles));
    const Medium mint = make_medium(eps[b][0], eps[b][1], mu[b][0], mu[b][1], omega);
    CHK(assemble_tiles(c, L, *dt, mext, tiles, 0, Tb, 0, Tb, info + 1));
    CHK(gather(c, L, tiles, np, Tb, 0, 0, dt->rw, 0, dt->E, 0, dt->E, &mext, W, M, o, o, 0, nullptr, 0, 0, 0));
    CHK(assemble_tiles(c, L, *dt, mint, tiles, 0, Tb, 0, Tb, info + 1));
    CHK(gather(c, L, tiles, np, Tb, 0, 0, dt->rw, 0, dt->E, 0, dt->E, &mint, W, M, o, o, 1, nullptr, 0, 0, 0));
  }
  // cross blocks: Z_ext(1, 2) assembled, Z_ext(2, 1) = S Z_ext(1, 2)^T S
  {
    const size_t npc = (size_t)T1 * (T - T1);
    double *tiles;
    CHK(ws_get(c, L, "tiles", sizeof(double) * BEM_TILE_PLANES * std::max<size_t>(npc, 1), (void **)&tiles));
    CHK(assemble_tiles(c, L, g->top, mext, tiles, 0, T1, T1, T, info + 1));
    CHK(gather(c, L, tiles, npc, T - T1, 0, T1, g->top.rw, 0, E1, E1, E, &mext, W, M, 0, m1, 0, nullptr, 0, 0, 0));
    int e = bem_signed_transpose(W + (size_t)m1 * M, M, m1, m2, W + m1, M, L.st);
    if (e) return fail(c, BEM_E_DEVICE, "transpose launch: %d", e);
  }
  std::vector<int> toff = {0, T1, T};
  const double epsv[4] = {q->eps1[0], q->eps1[1], q->eps2[0], q->eps2[1]};
  const double muv[4] = {q->mu1[0], q->mu1[1], q->mu2[0], q->mu2[1]};
  CHK(dipole_rhs(c, L, g->top, toff, epsv, muv, omega, ns, src, Bx, M));
  CHK(lu_factor(c, L, W, M, M, ipiv, info, inv));
  int hinfo[4] = {0, 0, 0, 0};
  DEV(c, hipMemcpyAsync(hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipStreamSynchronize(L.st));
  if (hinfo[1]) return fail(c, BEM_E_DEVICE, "36-point pair queue overflowed");
  if (hinfo[0]) return fail(c, BEM_E_SINGULAR, "zero pivot at column %d", hinfo[0] - 1);
  CHK(lu_solve_few(c, L, (const double2 *)W, M, ipiv, (const double2 *)inv, Bx, M, ns, false));
  const double shift1[3] = {q->gap[0], q->gap[1], q->gap[2]}, shift2[3] = {0, 0, 0};
  std::vector<FluxTri> ft;
  std::vector<double> out(ns);
  for (int b = 0; b < 2; b++) {
    if (bemgeo::flux_table(bs[b]->mesh, b ? shift2 : shift1, b ? E1 : 0, ft))
      return fail(c, BEM_E_INVALID, "body %d has a triangle edge without an RWG", b + 1);
    CHK(flux_columns(c, L, ft, omega, Bx, M, ns, out.data()));
    for (int s = 0; s < ns; s++) flux[2 * s + b] = out[s];
  }
  ev_collect(c, L);
  return 0;
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
  CHK(lane_retire(c, L));
  L.restarts = 0;
  int e = lane_issue(c, L, *g, b1, b2, q, res);
  if (e) {
    lane_abort(c, L);
    res->status = e;
    return e;
  }
  int f = lane_drive(c, L);
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
  // request i runs on lane i % lanes; a lane's previous request is driven to
  // its final result (refinement steps, re-solves) before the lane takes the
  // next one, while the other lanes' kernels keep running
  for (int i = 0; i < n; i++) {
    std::memset(&res[i], 0, sizeof res[i]);
    Lane &L = c->lanes[i % nl];
    int f = lane_retire(c, L);
    if (f && !first_err) first_err = f;
    int e = check_request(c, &reqs[i]);
    PairGeom *g = nullptr;
    if (!e) e = pair_geom(c, b1, b2, reqs[i].gap, &g);
    L.restarts = 0;
    if (!e) e = lane_issue(c, L, *g, b1, b2, &reqs[i], &res[i]);
    if (e) {
      res[i].status = e;
      if (!first_err) first_err = e;
      lane_abort(c, L);
    }
  }
  for (auto &L : c->lanes) {
    int f = lane_retire(c, L);
    if (f && !first_err) first_err = f;
  }
  c->t_last[6] = c->gemm_ms;
  c->t_last[7] = c->gemm_n;
  c->t_last[8] = c->gemm_flops;
  return first_err;
}

int bem_submit(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *q, uint64_t *ticket) {
  if (!c || !b1 || !b2 || !q || !ticket) return BEM_E_INVALID;
  std::lock_guard<std::mutex> lk(c->qmu);
  if (c->stop) return BEM_E_INVALID;
  Pending p{c->next_ticket++, b1, b2, *q, std::chrono::steady_clock::now(), 0};
  c->queue.push_back(p);
  c->live.insert(p.ticket);
  *ticket = p.ticket;
  if (!c->sched.joinable()) c->sched = std::thread(sched_main, c);
  c->kick = true;
  c->qcv.notify_one();
  return 0;
}

int bem_poll(bem_ctx *c, uint64_t ticket, bem_result *res, int wait_ms) {
  if (!c || !res) return BEM_E_INVALID;
  std::unique_lock<std::mutex> lk(c->qmu);
  auto ready = [&] { return c->done.count(ticket) > 0; };
  if (!ready()) {
    if (!c->live.count(ticket)) return BEM_E_INVALID;
    if (wait_ms == 0) return BEM_PENDING;
    if (wait_ms < 0) c->done_cv.wait(lk, ready);
    else if (!c->done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms), ready)) return BEM_PENDING;
  }
  auto it = c->done.find(ticket);
  *res = it->second;
  c->done.erase(it);
  return res->status;
}

int bem_queue_depth(bem_ctx *c, int *queued, int *running) {
  if (!c) return BEM_E_INVALID;
  std::lock_guard<std::mutex> lk(c->qmu);
  if (queued) *queued = (int)c->queue.size();
  if (running) *running = c->running;
  return 0;
}

int bem_gemm_tiles(bem_ctx *c, const int *set6, int *get6) {
  if (!c) return BEM_E_INVALID;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (set6) {
    for (int i = 0; i < 3; i++)
      if ((set6[i] && !bem_zgemm_tile_valid(set6[i])) || (set6[3 + i] && !bem_cgemm_tile_valid(set6[3 + i])))
        return fail(c, BEM_E_INVALID, "unknown tile code");
    for (auto &L : c->lanes) CHK(lane_retire(c, L));
    for (int i = 0; i < 3; i++) {
      c->ztile[i] = set6[i];
      c->ctile[i] = set6[3 + i];
    }
  }
  if (get6)
    for (int i = 0; i < 3; i++) {
      get6[i] = c->ztile[i];
      get6[3 + i] = c->ctile[i];
    }
  return 0;
}

int bem_autotune(bem_ctx *c, const char *cache_path, int flags, int *tiles6) {
  if (!c) return BEM_E_INVALID;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  hipDeviceProp_t prop;
  DEV(c, hipGetDeviceProperties(&prop, c->device));
  const std::string arch = prop.gcnArchName;
  int best[6];
  if (!(flags & BEM_TUNE_FORCE) && tune_load(cache_path, arch, best)) {
    CHK(bem_gemm_tiles(c, best, tiles6));
    return 0;
  }
  Lane &L = c->lanes[0];
  for (auto &l : c->lanes) CHK(lane_retire(c, l));
  const size_t bytes = sizeof(double2) * (size_t)4096 * 4096;
  void *A = nullptr, *B = nullptr, *Cm = nullptr;
  int r = 0;
  if (hipMalloc(&A, bytes) != hipSuccess || hipMalloc(&B, bytes) != hipSuccess || hipMalloc(&Cm, bytes) != hipSuccess)
    r = fail(c, BEM_E_NOMEM, "autotune scratch (3 x %zu bytes)", bytes);
  // operands: hashed values in [-1, 1) of the timed precision, so the MFMAs
  // see realistic bit patterns (all-zero operands draw less power and time
  // faster than data)
  auto fill = [&](int as_float) {
    if (!r && (bem_fill_hash(A, bytes, as_float, 1, L.st) || bem_fill_hash(B, bytes, as_float, 2, L.st) ||
               bem_fill_hash(Cm, bytes, as_float, 3, L.st)))
      r = fail(c, BEM_E_DEVICE, "autotune scratch fill");
  };
  double ms[6][16];
  static_assert(sizeof ZCANDS / sizeof ZCANDS[0] <= 16 && sizeof CCANDS / sizeof CCANDS[0] <= 16, "timing table");
  const int nz = (int)(sizeof ZCANDS / sizeof ZCANDS[0]), nc = (int)(sizeof CCANDS / sizeof CCANDS[0]);
  for (int cls = 0; cls < 3 && !r; cls++) {
    best[cls] = 0;
    best[3 + cls] = 0;
    double bz = 1e30, bc = 1e30;
    fill(0);
    for (int i = 0; i < nz && !r; i++) {
      r = tune_time(c, L, false, ZCANDS[i], cls, A, B, Cm, 3, &ms[cls][i]);
      if (!r && ms[cls][i] < bz) { bz = ms[cls][i]; best[cls] = ZCANDS[i]; }
    }
    fill(1);
    for (int i = 0; i < nc && !r; i++) {
      r = tune_time(c, L, true, CCANDS[i], cls, A, B, Cm, 3, &ms[3 + cls][i]);
      if (!r && ms[3 + cls][i] < bc) { bc = ms[3 + cls][i]; best[3 + cls] = CCANDS[i]; }
    }
  }
  (void)hipStreamSynchronize(L.st);
  for (void *p : {A, B, Cm})
    if (p) (void)hipFree(p);
  if (r) return r;
  tune_store(cache_path, arch, best, ms, &nz, &nc);
  return bem_gemm_tiles(c, best, tiles6);
}

int bem_zgetrf(bem_ctx *c, void *A, int64_t lda, int n, int32_t *ipiv) {
  if (!c || !A || !ipiv || n < 0 || lda < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
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

int bem_zgetrf_local(bem_ctx *c, void *A, int64_t lda, int n, int32_t *ipiv, double *growth) {
  if (!c || !A || !ipiv || n < 0 || lda < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  stats_reset(c);
  int *info;
  unsigned long long *gro;
  double2 *inv;
  CHK(ws_get(c, L, "info_getrf", sizeof(int), (void **)&info));
  CHK(ws_get(c, L, "growth_getrf", sizeof(unsigned long long) * 3, (void **)&gro));
  CHK(ws_get(c, L, "inv_getrf", inv_bytes(n), (void **)&inv));
  DEV(c, hipMemsetAsync(info, 0, sizeof(int), L.st));
  DEV(c, hipMemsetAsync(gro, 0, sizeof(unsigned long long) * 3, L.st));
  CHK(lu_factor_block(c, L, (double2 *)A, lda, n, ipiv, info, inv, gro));
  int h = 0;
  unsigned long long g[3] = {0, 0, 0};
  DEV(c, hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, L.st));
  DEV(c, hipMemcpyAsync(g, gro, sizeof g, hipMemcpyDeviceToHost, L.st));
  DEV(c, hipStreamSynchronize(L.st));
  ev_collect(c, L);
  if (growth) *growth = growth_ratio(g);
  if (h) return fail(c, BEM_E_SINGULAR, "zero pivot at column %d", h - 1);
  return 0;
}

int bem_aca(bem_ctx *c, void *A, int64_t lda, int m, int n, double eps2, int kmax, void *U, int64_t ldu, void *V,
            int64_t ldv, int32_t *rows, int32_t *cols, int *rank) {
  if (!c || !A || !U || !V || !rows || !cols || !rank || m <= 0 || n <= 0 || kmax <= 0 || lda < m || ldu < m ||
      ldv < kmax)
    return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  int *rc;
  CHK(ws_get(c, L, "aca_rc", sizeof(int) * 2 * (size_t)kmax, (void **)&rc));
  int r = 0, full = 0;
  CHK(aca_run(c, L, (double2 *)A, lda, m, n, eps2, kmax, (double2 *)U, ldu, (double2 *)V, ldv, rc, rc + kmax, &r,
              &full));
  DEV(c, hipMemcpy(rows, rc, sizeof(int) * (size_t)kmax, hipMemcpyDeviceToHost));
  DEV(c, hipMemcpy(cols, rc + kmax, sizeof(int) * (size_t)kmax, hipMemcpyDeviceToHost));
  *rank = r;
  if (full) return fail(c, BEM_E_UNSUPPORTED, "aca: %d columns used up before the tolerance was met", kmax);
  return 0;
}

int bem_rk_truncate(bem_ctx *c, void *U, int64_t ldu, void *V, int64_t ldv, int m, int n, int k, double eps, int *knew,
                    double *sigma) {
  if (!c || !U || !V || !knew || m <= 0 || n <= 0 || k < 0 || ldu < m || ldv < k || !(eps >= 0))
    return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  return rk_truncate(c, L, (double2 *)U, ldu, (double2 *)V, ldv, m, n, k, eps, knew, sigma);
}

int bem_zgetrf_local_noalloc(bem_ctx *c, void *A, int64_t lda, int n, int32_t *ipiv, double *growth) {
  if (!c || !A || !ipiv || n < 0 || lda < n) return fail(c, BEM_E_INVALID, "bad arguments");
  if (n > BEM_NOALLOC_MAX_N)
    return fail(c, BEM_E_UNSUPPORTED, "n = %d exceeds BEM_NOALLOC_MAX_N = %d of the allocation-free factorisation", n,
                BEM_NOALLOC_MAX_N);
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane *L = nullptr;
  for (auto &l : c->lanes)
    if (!l.busy) {
      L = &l;
      break;
    }
  if (!L) return fail(c, BEM_E_UNSUPPORTED, "every lane has a request in flight; the allocation-free factorisation "
                                            "runs on an idle lane");
  unsigned char *base = (unsigned char *)c->scratch;
  double2 *inv = (double2 *)base;
  int *info = (int *)(base + NOALLOC_INV_BYTES);
  unsigned long long *gro = (unsigned long long *)(base + NOALLOC_INV_BYTES + 64);
  if (bem_fill_zero(base + NOALLOC_INV_BYTES, 256, L->st)) return fail(c, BEM_E_DEVICE, "scratch clear launch");
  // the same kernels in the same order as bem_zgetrf_local; no timing events
  // (nothing is created)
  const bool prof = c->profile;
  c->profile = false;
  int e = lu_factor_block(c, *L, (double2 *)A, lda, n, ipiv, info, inv, gro);
  c->profile = prof;
  if (e) {
    (void)hipStreamSynchronize(L->st);
    return e;
  }
  DEV(c, hipMemcpyAsync(c->h_noalloc, info, sizeof(int), hipMemcpyDeviceToHost, L->st));
  DEV(c, hipMemcpyAsync(c->h_noalloc + 16, gro, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost, L->st));
  DEV(c, hipStreamSynchronize(L->st));
  if (growth) *growth = growth_ratio((const unsigned long long *)(c->h_noalloc + 16));
  if (c->h_noalloc[0]) return fail(c, BEM_E_SINGULAR, "zero pivot at column %d", c->h_noalloc[0] - 1);
  return 0;
}

int bem_zgetrs(bem_ctx *c, const void *A, int64_t lda, int n, const int32_t *ipiv, void *B, int64_t ldb, int nrhs) {
  if (!c || !A || !ipiv || !B || n < 0 || lda < n || ldb < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  stats_reset(c);
  double2 *inv;
  CHK(ws_get(c, L, "inv_getrs", inv_bytes(n), (void **)&inv));
  int e = bem_diag_block((double2 *)A, lda, n, 0, (n + LU_NB - 1) / LU_NB, LU_NB, 0, inv, nullptr, L.st);
  if (e) return fail(c, BEM_E_DEVICE, "diag block inverses: %d", e);
  if (nrhs <= FEW_COLS && lda == n)   // few right-hand sides: a launch per 64-row block
    CHK(lu_solve_few(c, L, (const double2 *)A, n, ipiv, inv, (double2 *)B, ldb, nrhs, false));
  else
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
  CHK(lane_retire(c, L));
  stats_reset(c);
  // at most FEW_COLS columns: the split-K skinny kernel (the probe's products)
  CHK(gemm_few(c, L, m, n, k, make_double2(alpha[0], alpha[1]), (const double2 *)A, lda, (const double2 *)B, ldb,
               make_double2(beta[0], beta[1]), (double2 *)C, ldc));
  DEV(c, hipStreamSynchronize(L.st));
  ev_collect(c, L);
  c->t_last[6] = c->gemm_ms;
  c->t_last[7] = c->gemm_n;
  c->t_last[8] = c->gemm_flops;
  return 0;
}

int bem_cgemm(bem_ctx *c, int m, int n, int k, const float *alpha, const void *A, int64_t lda, const void *B,
              int64_t ldb, const float *beta, void *C, int64_t ldc) {
  if (!c || !alpha || !beta || m < 0 || n < 0 || k < 0) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  Lane &L = c->lanes[0];
  CHK(lane_retire(c, L));
  stats_reset(c);
  CHK(gemm(c, L, m, n, k, make_float2(alpha[0], alpha[1]), (const float2 *)A, lda, (const float2 *)B, ldb,
           make_float2(beta[0], beta[1]), (float2 *)C, ldc));
  DEV(c, hipStreamSynchronize(L.st));
  ev_collect(c, L);
  return 0;
}

int bem_synchronize(bem_ctx *c) {
  if (!c) return BEM_E_INVALID;
  DEV(c, hipSetDevice(c->device));
  DEV(c, hipDeviceSynchronize());
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

int bem_kernel_stats(bem_ctx *c, int kind, double *out4) {
  if (!c || !out4 || kind < 0 || kind >= BEM_KS_COUNT) return BEM_E_INVALID;
  for (int i = 0; i < 4; i++) out4[i] = c->kst[kind][i];
  return 0;
}

int bem_last_timings(bem_ctx *c, double *out9) {
  if (!c || !out9) return BEM_E_INVALID;
  for (int i = 0; i < 9; i++) out9[i] = c->t_last[i];
  return 0;
}

}  // extern "C"
