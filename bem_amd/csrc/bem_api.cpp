// libbem_hip host runtime: contexts, bodies, assembly orchestration and the
// PMCHWT solve behind the C ABI of include/bem_hip.h.
//
// CalcHeatTransfer (src/HeatTransfer.cpp:8-59) inverts W = Z_ext + blkdiag(G1,G2)
// and forms S = 1/4 tr(B2^H B1) with B1 = sym(G1) X, B2 = X sym(G2),
// X = W^-1[0:m1, m1:M].  Only X is needed, so the solve here is the block
// elimination of W = [[A, B], [C, D]]:
//   Y = A^-1 B,  Sc = D - C Y,  X = -Y Sc^-1
// with tournament-pivoted LU of A and of Sc, followed by B1 = sym(G1) X and a
// GEMM whose epilogue contracts X sym(G2) against conj/B1 (no B2 stored).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
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
int bem_diag_block(double2 *A, int64_t lda, int n, int blk0, int nblk, int factor, double2 *inv, int *info,
                   hipStream_t st);
int bem_set_identity(double2 *X, int64_t ld, int n, hipStream_t st);
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
  int kind;       // 0 gemm
  double flops;
};

}  // namespace

struct bem_ctx {
  int device = 0;
  size_t budget = 0;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::recursive_mutex mu;
  std::unordered_map<std::string, DevBuf> ws;   // named, grow-only workspaces
  std::unordered_map<void *, size_t> user_allocs;
  // timing
  bool profile_gemm = true;
  std::vector<EventRec> ev_pool;
  size_t ev_used = 0;
  double t_last[9] = {0};
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

static int ws_get(bem_ctx *c, const char *name, size_t bytes, void **out) {
  DevBuf &b = c->ws[name];
  if (b.cap < bytes) {
    if (b.p) {
      DEV(c, hipStreamSynchronize(c->stream));
      DEV(c, hipFree(b.p));
      c->bytes -= b.cap;
      b.p = nullptr;
      b.cap = 0;
    }
    if (c->budget && c->bytes + bytes > c->budget)
      return fail(c, BEM_E_NOMEM, "workspace '%s' (%zu bytes) exceeds the HBM budget", name, bytes);
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(c, BEM_E_NOMEM, "hipMalloc(%zu) for '%s': %s", bytes, name, hipGetErrorString(e));
    b.cap = bytes;
    c->bytes += bytes;
  }
  *out = b.p;
  return 0;
}

static void ev_reset(bem_ctx *c) { c->ev_used = 0; }
static int ev_begin(bem_ctx *c, int kind, double flops, EventRec **rec) {
  if (c->ev_used == c->ev_pool.size()) {
    EventRec r;
    DEV(c, hipEventCreate(&r.a));
    DEV(c, hipEventCreate(&r.b));
    c->ev_pool.push_back(r);
  }
  EventRec &r = c->ev_pool[c->ev_used++];
  r.kind = kind;
  r.flops = flops;
  DEV(c, hipEventRecord(r.a, c->stream));
  *rec = &r;
  return 0;
}
static int ev_end(bem_ctx *c, EventRec *r) {
  DEV(c, hipEventRecord(r->b, c->stream));
  return 0;
}

// GEMM through the context (timed when profiling).
static int gemm(bem_ctx *c, int m, int n, int k, double2 alpha, const double2 *A, int64_t lda, const double2 *B,
                int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  EventRec *r = nullptr;
  if (c->profile_gemm) CHK(ev_begin(c, 0, 8.0 * m * (double)n * k, &r));
  int e = bem_zgemm_raw(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "zgemm launch: %s", hipGetErrorString((hipError_t)e));
  if (r) CHK(ev_end(c, r));
  return 0;
}

static void ev_collect(bem_ctx *c, double *gemm_ms, double *gemm_n, double *gemm_flops) {
  *gemm_ms = 0;
  *gemm_n = 0;
  *gemm_flops = 0;
  for (size_t i = 0; i < c->ev_used; i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev_pool[i].a, c->ev_pool[i].b) == hipSuccess && c->ev_pool[i].kind == 0) {
      *gemm_ms += ms;
      *gemm_n += 1;
      *gemm_flops += c->ev_pool[i].flops;
    }
  }
}

// --------------------------------------------------------------- LU / solves
// Panel width of the blocked LU; the diagonal-block inverses are 64 x 64.
static const int LU_NB = 64;

static size_t inv_bytes(int n) { return sizeof(double2) * 2 * 64 * 64 * (size_t)((n + LU_NB - 1) / LU_NB); }

// X(r x c block at dst, ld) <- tmp (r x c, ld r)
static int copy_back(bem_ctx *c, double2 *dst, int64_t ld, const double2 *tmp, int rows, int cols) {
  DEV(c, hipMemcpy2DAsync(dst, sizeof(double2) * ld, tmp, sizeof(double2) * rows, sizeof(double2) * rows, cols,
                          hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// Blocked right-looking LU with tournament-pivoted panels.  inv receives the
// inverses of the unit-lower and upper diagonal blocks (inv_bytes(n)); the
// panel's two triangular solves run as GEMMs against them.
static int lu_factor(bem_ctx *c, double2 *A, int64_t lda, int n, int *ipiv, int *info, double2 *inv) {
  const int nb = LU_NB;
  int sets = (n + 255) / 256;
  int *cand, *cand2;
  double2 *tmp;
  CHK(ws_get(c, "calu_cand", sizeof(int) * 64 * (size_t)sets, (void **)&cand));
  CHK(ws_get(c, "calu_cand2", sizeof(int) * 64 * (size_t)((sets + 3) / 4 + 1), (void **)&cand2));
  CHK(ws_get(c, "lu_tmp", sizeof(double2) * 64 * (size_t)std::max(n, 1), (void **)&tmp));
  const double2 m1 = make_double2(-1.0, 0.0), one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    int e = bem_calu_pivots(A, lda, n - k0, jb, k0, k0, cand, cand2, ipiv, info, c->stream);
    if (e) return fail(c, BEM_E_DEVICE, "calu pivots: %d", e);
    e = bem_laswp(A, lda, 0, n, 0, 0, k0, k0 + jb, ipiv, c->stream);
    if (e) return fail(c, BEM_E_DEVICE, "laswp: %d", e);
    double2 *Akk = A + (int64_t)k0 * lda + k0;
    const int blk = k0 / nb;
    e = bem_diag_block(A, lda, n, blk, 1, 1, inv, info, c->stream);
    if (e) return fail(c, BEM_E_DEVICE, "diag block: %d", e);
    const double2 *Linv = inv + (size_t)blk * 2 * 64 * 64, *Uinv = Linv + 64 * 64;
    int rest = n - k0 - jb;
    if (rest > 0) {
      double2 *A21 = Akk + jb, *A12 = A + (int64_t)(k0 + jb) * lda + k0;
      CHK(gemm(c, rest, jb, jb, one, A21, lda, Uinv, 64, zero, tmp, rest));   // L21 = A21 U11^-1
      CHK(copy_back(c, A21, lda, tmp, rest, jb));
      CHK(gemm(c, jb, rest, jb, one, Linv, 64, A12, lda, zero, tmp, jb));     // U12 = L11^-1 A12
      CHK(copy_back(c, A12, lda, tmp, jb, rest));
      CHK(gemm(c, rest, rest, jb, m1, A21, lda, A12, lda, one, A12 + jb, lda));
    }
  }
  return 0;
}

// X (n x r) <- A^-1 X from the packed LU, its interchanges and the diagonal
// block inverses.
static int lu_solve(bem_ctx *c, const double2 *A, int64_t lda, int n, const int *ipiv, const double2 *inv, double2 *X,
                    int64_t ldx, int r) {
  const int nb = LU_NB;
  const double2 m1 = make_double2(-1.0, 0.0), one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  double2 *tmp;
  CHK(ws_get(c, "lu_tmp", sizeof(double2) * 64 * (size_t)std::max(std::max(n, r), 1), (void **)&tmp));
  int e = bem_laswp(X, ldx, 0, r, 0, 0, 0, n, ipiv, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "laswp rhs: %d", e);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int jb = std::min(nb, n - k0);
    const double2 *Linv = inv + (size_t)(k0 / nb) * 2 * 64 * 64;
    CHK(gemm(c, jb, r, jb, one, Linv, 64, X + k0, ldx, zero, tmp, jb));
    CHK(copy_back(c, X + k0, ldx, tmp, jb, r));
    int rest = n - k0 - jb;
    if (rest > 0) CHK(gemm(c, rest, r, jb, m1, A + (int64_t)k0 * lda + k0 + jb, lda, X + k0, ldx, one, X + k0 + jb, ldx));
  }
  int last = ((n - 1) / nb) * nb;
  for (int k0 = last; k0 >= 0; k0 -= nb) {
    int jb = std::min(nb, n - k0);
    const double2 *Uinv = inv + (size_t)(k0 / nb) * 2 * 64 * 64 + 64 * 64;
    CHK(gemm(c, jb, r, jb, one, Uinv, 64, X + k0, ldx, zero, tmp, jb));
    CHK(copy_back(c, X + k0, ldx, tmp, jb, r));
    if (k0 > 0) CHK(gemm(c, k0, r, jb, m1, A + (int64_t)k0 * lda, lda, X + k0, ldx, one, X, ldx));
  }
  return 0;
}

// ------------------------------------------------------------------ assembly
static Medium make_medium(double er, double ei, double mr, double mi, double omega) {
  // src/GreensFunction.cpp:17-30 (complex arithmetic written out)
  Medium md;
  // eps*mu
  double emr = er * mr - ei * mi, emi = er * mi + ei * mr;
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

struct DevTree {
  int T = 0, E = 0;
  TriGeom *tg = nullptr;
  RwgRec *rw = nullptr;
  int2 *close = nullptr;
  unsigned int nclose = 0;
};

static int upload_tree(bem_ctx *c, const bemgeo::Tree &t, const char *tag, DevTree &d) {
  std::string n1 = std::string(tag) + ":tg", n2 = std::string(tag) + ":rw", n3 = std::string(tag) + ":close";
  CHK(ws_get(c, n1.c_str(), sizeof(TriGeom) * t.tg.size(), (void **)&d.tg));
  CHK(ws_get(c, n2.c_str(), sizeof(RwgRec) * t.rw.size(), (void **)&d.rw));
  CHK(ws_get(c, n3.c_str(), sizeof(int2) * (t.close.size() + 1), (void **)&d.close));
  std::vector<int2> cl(t.close.size());
  for (size_t i = 0; i < cl.size(); i++) cl[i] = make_int2(t.close[i].first, t.close[i].second);
  DEV(c, hipMemcpyAsync(d.tg, t.tg.data(), sizeof(TriGeom) * t.tg.size(), hipMemcpyHostToDevice, c->stream));
  DEV(c, hipMemcpyAsync(d.rw, t.rw.data(), sizeof(RwgRec) * t.rw.size(), hipMemcpyHostToDevice, c->stream));
  if (!cl.empty())
    DEV(c, hipMemcpyAsync(d.close, cl.data(), sizeof(int2) * cl.size(), hipMemcpyHostToDevice, c->stream));
  // the host vectors die at return: make the copies complete first
  DEV(c, hipStreamSynchronize(c->stream));
  d.T = t.T;
  d.E = t.E;
  d.nclose = (unsigned int)cl.size();
  return 0;
}

// Triangle-pair tiles of one tree (triangles [t_lo, t_lo+T) of tg) in medium md.
static int assemble_tiles(bem_ctx *c, const TriGeom *tg, int T, const int2 *close, unsigned int nclose,
                          const Medium &md, double *tiles) {
  size_t npairs = (size_t)T * (size_t)T;
  unsigned int *q36cnt;
  int2 *q36;
  unsigned int cap = (unsigned int)std::min<size_t>(npairs, (size_t)1 << 24);
  CHK(ws_get(c, "q36_count", sizeof(unsigned int), (void **)&q36cnt));
  CHK(ws_get(c, "q36", sizeof(int2) * (size_t)cap, (void **)&q36));
  DEV(c, hipMemsetAsync(q36cnt, 0, sizeof(unsigned int), c->stream));
  int e = bem_launch_far_g6(tg, T, &md, 0, T, 0, T, tiles, npairs, q36, q36cnt, cap, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "far_g6 launch: %s", hipGetErrorString((hipError_t)e));
  e = bem_launch_far_g36(tg, T, &md, q36, q36cnt, cap, 0, T, tiles, npairs, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "far_g36 launch: %s", hipGetErrorString((hipError_t)e));
  e = bem_launch_close(tg, T, &md, close, nclose, 0, T, tiles, npairs, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "close launch: %s", hipGetErrorString((hipError_t)e));
  return 0;
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

// ------------------------------------------------------------------- C ABI
extern "C" {

int bem_ctx_create(bem_ctx **out, const bem_cfg *cfg) {
  if (!out) return BEM_E_INVALID;
  bem_ctx *c = new bem_ctx();
  if (cfg) {
    c->device = cfg->device;
    c->budget = cfg->hbm_budget;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device) {
    delete c;
    return BEM_E_DEVICE;
  }
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return BEM_E_DEVICE;
  }
  *out = c;
  return 0;
}

void bem_ctx_destroy(bem_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &kv : c->ws)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto &kv : c->user_allocs) (void)hipFree(kv.first);
  for (auto &r : c->ev_pool) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  (void)hipStreamDestroy(c->stream);
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

void bem_body_free(bem_body *b) { delete b; }

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
  std::vector<const bemgeo::BodyMesh *> bm;
  for (int i = 0; i < nb; i++) bm.push_back(&bodies[i]->mesh);
  bemgeo::Tree t;
  bemgeo::build_tree(bm, shifts, t);
  int M = 2 * t.E;
  if (ldz < M) return fail(c, BEM_E_INVALID, "ldz %lld < %d", (long long)ldz, M);
  DevTree d;
  CHK(upload_tree(c, t, "asm", d));
  Medium md = make_medium(eps[0], eps[1], mu[0], mu[1], omega);
  double *tiles;
  size_t npairs = (size_t)t.T * t.T;
  CHK(ws_get(c, "tiles", sizeof(double) * BEM_TILE_PLANES * npairs, (void **)&tiles));
  CHK(assemble_tiles(c, d.tg, d.T, d.close, d.nclose, md, tiles));
  double2 *Zd;
  bool dev = flags & BEM_ASM_DEVICE_PTR;
  int64_t ld = dev ? ldz : M;
  if (dev) Zd = (double2 *)Z;
  else CHK(ws_get(c, "asm_Z", sizeof(double2) * (size_t)M * M, (void **)&Zd));
  int e = bem_launch_gather(tiles, npairs, t.T, 0, d.rw, 0, t.E, 0, t.E, &md, Zd, ld, 0, 0, 0, nullptr, 0, 0, 0, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "gather launch: %d", e);
  if (!dev) {
    DEV(c, hipMemcpy2DAsync(Z, sizeof(double2) * ldz, Zd, sizeof(double2) * M, sizeof(double2) * M, M,
                            hipMemcpyDeviceToHost, c->stream));
  }
  DEV(c, hipStreamSynchronize(c->stream));
  return 0;
}

static double theta(double omega, double T) {
  // src/HeatTransfer.cpp:3-6 (T = 0 gives exp(inf) and exactly 0)
  return SI_HBAR * omega / (std::exp(SI_HBAR * omega / (SI_KB * T)) - 1);
}

static int spectral_power_fp64(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *q, bem_result *res) {
  hipEvent_t e0, e1, e2, e3;
  DEV(c, hipEventCreate(&e0));
  DEV(c, hipEventCreate(&e1));
  DEV(c, hipEventCreate(&e2));
  DEV(c, hipEventCreate(&e3));
  auto t_host0 = std::chrono::steady_clock::now();
  double omegaSI = q->eV * SI_EV / SI_HBAR;
  double omega = omegaSI * q->a / SI_C;
  std::vector<const bemgeo::BodyMesh *> bm = {&b1->mesh, &b2->mesh};
  double shifts[6] = {q->gap[0], q->gap[1], q->gap[2], 0, 0, 0};
  bemgeo::Tree top, s1, s2;
  bemgeo::build_tree(bm, shifts, top);
  body_subtree(top, 0, s1);
  body_subtree(top, 1, s2);
  const int E1 = s1.E, E2 = s2.E;
  const int m1 = 2 * E1, m2 = 2 * E2, M = m1 + m2;
  DevTree dt, d1, d2;
  CHK(upload_tree(c, top, "top", dt));
  CHK(upload_tree(c, s1, "s1", d1));
  CHK(upload_tree(c, s2, "s2", d2));
  Medium mext = make_medium(1.0, 0.0, 1.0, 0.0, omega);
  Medium m1d = make_medium(q->eps1[0], q->eps1[1], q->mu1[0], q->mu1[1], omega);
  Medium m2d = make_medium(q->eps2[0], q->eps2[1], q->mu2[0], q->mu2[1], omega);

  double2 *W, *G1, *G2, *Sinv, *X, *B1, *partial, *trace_out;
  double *tiles;
  int *ipA, *ipS, *info;
  size_t npairs_top = (size_t)top.T * top.T;
  CHK(ws_get(c, "W", sizeof(double2) * (size_t)M * M, (void **)&W));
  CHK(ws_get(c, "G1", sizeof(double2) * (size_t)m1 * m1, (void **)&G1));
  CHK(ws_get(c, "G2", sizeof(double2) * (size_t)m2 * m2, (void **)&G2));
  CHK(ws_get(c, "tiles", sizeof(double) * BEM_TILE_PLANES * npairs_top, (void **)&tiles));
  CHK(ws_get(c, "ipiv", sizeof(int) * (size_t)(M + 64), (void **)&ipA));
  ipS = ipA + m1;
  CHK(ws_get(c, "info", sizeof(int) * 4, (void **)&info));
  DEV(c, hipMemsetAsync(info, 0, sizeof(int) * 4, c->stream));

  ev_reset(c);
  DEV(c, hipEventRecord(e0, c->stream));
  // ---- assembly: W = Z_ext(top), then + Z_int(b) on the diagonal blocks
  CHK(assemble_tiles(c, dt.tg, dt.T, dt.close, dt.nclose, mext, tiles));
  int e = bem_launch_gather(tiles, npairs_top, dt.T, 0, dt.rw, 0, dt.E, 0, dt.E, &mext, W, M, 0, 0, 0, nullptr, 0, 0, 0,
                            c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "gather: %d", e);
  CHK(assemble_tiles(c, d1.tg, d1.T, d1.close, d1.nclose, m1d, tiles));
  e = bem_launch_gather(tiles, (size_t)d1.T * d1.T, d1.T, 0, d1.rw, 0, d1.E, 0, d1.E, &m1d, W, M, 0, 0, 1, G1, m1, 0, 0,
                        c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "gather: %d", e);
  CHK(assemble_tiles(c, d2.tg, d2.T, d2.close, d2.nclose, m2d, tiles));
  e = bem_launch_gather(tiles, (size_t)d2.T * d2.T, d2.T, 0, d2.rw, 0, d2.E, 0, d2.E, &m2d, W, M, m1, m1, 1, G2, m2, 0,
                        0, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "gather: %d", e);
  DEV(c, hipEventRecord(e1, c->stream));

  // ---- solve: Y = A^-1 B, Sc = D - C Y, X = -Y Sc^-1
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  double2 *A = W, *Bm = W + (int64_t)m1 * M, *Cm = W + m1, *D = W + (int64_t)m1 * M + m1;
  double2 *invA, *invS;
  CHK(ws_get(c, "invA", inv_bytes(m1), (void **)&invA));
  CHK(ws_get(c, "invS", inv_bytes(m2), (void **)&invS));
  CHK(lu_factor(c, A, M, m1, ipA, info, invA));
  CHK(lu_solve(c, A, M, m1, ipA, invA, Bm, M, m2));   // B <- Y
  CHK(gemm(c, m2, m2, m1, mone, Cm, M, Bm, M, one, D, M));
  CHK(lu_factor(c, D, M, m2, ipS, info + 1, invS));
  CHK(ws_get(c, "Sinv", sizeof(double2) * (size_t)m2 * m2, (void **)&Sinv));
  e = bem_set_identity(Sinv, m2, m2, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "identity: %d", e);
  CHK(lu_solve(c, D, M, m2, ipS, invS, Sinv, m2, m2));
  CHK(ws_get(c, "X", sizeof(double2) * (size_t)m1 * m2, (void **)&X));
  CHK(gemm(c, m1, m2, m2, mone, Bm, M, Sinv, m2, zero, X, m1));
  DEV(c, hipEventRecord(e2, c->stream));

  // ---- products and trace: S = 1/4 sum conj(X sym(G2)) .* (sym(G1) X)
  e = bem_symm(G1, m1, m1, c->stream);
  if (!e) e = bem_symm(G2, m2, m2, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "symm: %d", e);
  CHK(ws_get(c, "B1", sizeof(double2) * (size_t)m1 * m2, (void **)&B1));
  CHK(gemm(c, m1, m2, m1, one, G1, m1, X, m1, zero, B1, m1));
  CHK(ws_get(c, "trace_partial", sizeof(double2) * (bem_trace_partials(m1, m2) + 1), (void **)&partial));
  trace_out = partial + bem_trace_partials(m1, m2);
  {
    EventRec *r = nullptr;
    if (c->profile_gemm) CHK(ev_begin(c, 0, 8.0 * m1 * (double)m2 * m2, &r));
    e = bem_zgemm_trace(m1, m2, m2, one, X, m1, G2, m2, B1, m1, partial, trace_out, c->stream);
    if (e) return fail(c, BEM_E_DEVICE, "trace gemm: %d", e);
    if (r) CHK(ev_end(c, r));
  }
  DEV(c, hipEventRecord(e3, c->stream));
  double2 tr;
  int hinfo[4];
  DEV(c, hipMemcpyAsync(&tr, trace_out, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  DEV(c, hipMemcpyAsync(hinfo, info, sizeof(int) * 4, hipMemcpyDeviceToHost, c->stream));
  DEV(c, hipStreamSynchronize(c->stream));
  float ta = 0, tf = 0, tp = 0;
  (void)hipEventElapsedTime(&ta, e0, e1);
  (void)hipEventElapsedTime(&tf, e1, e2);
  (void)hipEventElapsedTime(&tp, e2, e3);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  (void)hipEventDestroy(e3);
  double gms, gn, gfl;
  ev_collect(c, &gms, &gn, &gfl);
  c->t_last[0] = ta;
  c->t_last[4] = tf;
  c->t_last[5] = tp;
  c->t_last[6] = gms;
  c->t_last[7] = gn;
  c->t_last[8] = gfl;

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
  res->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
  res->status = 0;
  if (hinfo[0] || hinfo[1]) {
    res->status = BEM_E_SINGULAR;
    return fail(c, BEM_E_SINGULAR, "exactly singular pivot (A: %d, Schur: %d)", hinfo[0], hinfo[1]);
  }
  return 0;
}

int bem_spectral_power(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *q, bem_result *res) {
  if (!c || !b1 || !b2 || !q || !res) return fail(c, BEM_E_INVALID, "bad arguments");
  if (!(q->eV > 0) || !(q->a > 0)) return fail(c, BEM_E_INVALID, "eV and a must be positive");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  std::memset(res, 0, sizeof *res);
  if (q->precision != BEM_PREC_FP64) return fail(c, BEM_E_UNSUPPORTED, "precision mode %d not available", q->precision);
  return spectral_power_fp64(c, b1, b2, q, res);
}

int bem_sweep(bem_ctx *c, bem_body *b1, bem_body *b2, const bem_request *reqs, int n, bem_result *res) {
  if (!c || !reqs || !res || n < 0) return fail(c, BEM_E_INVALID, "bad arguments");
  int first_err = 0;
  for (int i = 0; i < n; i++) {
    int e = bem_spectral_power(c, b1, b2, &reqs[i], &res[i]);
    res[i].status = e;
    if (e && !first_err) first_err = e;
  }
  return first_err;
}

int bem_zgetrf(bem_ctx *c, void *A, int64_t lda, int n, int32_t *ipiv) {
  if (!c || !A || !ipiv || n < 0 || lda < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  int *info;
  double2 *inv;
  CHK(ws_get(c, "info_getrf", sizeof(int), (void **)&info));
  CHK(ws_get(c, "inv_getrf", inv_bytes(n), (void **)&inv));
  DEV(c, hipMemsetAsync(info, 0, sizeof(int), c->stream));
  ev_reset(c);
  CHK(lu_factor(c, (double2 *)A, lda, n, ipiv, info, inv));
  int h = 0;
  DEV(c, hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DEV(c, hipStreamSynchronize(c->stream));
  if (h) return fail(c, BEM_E_SINGULAR, "zero pivot at column %d", h - 1);
  return 0;
}

int bem_zgetrs(bem_ctx *c, const void *A, int64_t lda, int n, const int32_t *ipiv, void *B, int64_t ldb, int nrhs) {
  if (!c || !A || !ipiv || !B || n < 0 || lda < n || ldb < n) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  ev_reset(c);
  double2 *inv;
  CHK(ws_get(c, "inv_getrs", inv_bytes(n), (void **)&inv));
  int e = bem_diag_block((double2 *)A, lda, n, 0, (n + LU_NB - 1) / LU_NB, 0, inv, nullptr, c->stream);
  if (e) return fail(c, BEM_E_DEVICE, "diag block inverses: %d", e);
  CHK(lu_solve(c, (const double2 *)A, lda, n, ipiv, inv, (double2 *)B, ldb, nrhs));
  DEV(c, hipStreamSynchronize(c->stream));
  return 0;
}

int bem_zgemm(bem_ctx *c, int m, int n, int k, const double *alpha, const void *A, int64_t lda, const void *B,
              int64_t ldb, const double *beta, void *C, int64_t ldc) {
  if (!c || !alpha || !beta || m < 0 || n < 0 || k < 0) return fail(c, BEM_E_INVALID, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  DEV(c, hipSetDevice(c->device));
  ev_reset(c);
  CHK(gemm(c, m, n, k, make_double2(alpha[0], alpha[1]), (const double2 *)A, lda, (const double2 *)B, ldb,
           make_double2(beta[0], beta[1]), (double2 *)C, ldc));
  DEV(c, hipStreamSynchronize(c->stream));
  double gms, gn, gfl;
  ev_collect(c, &gms, &gn, &gfl);
  c->t_last[6] = gms;
  c->t_last[7] = gn;
  c->t_last[8] = gfl;
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
  DEV(c, hipStreamSynchronize(c->stream));
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
  DEV(c, hipStreamSynchronize(c->stream));
  DEV(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int bem_last_timings(bem_ctx *c, double *out9) {
  if (!c || !out9) return BEM_E_INVALID;
  for (int i = 0; i < 9; i++) out9[i] = c->t_last[i];
  return 0;
}

}  // extern "C"
