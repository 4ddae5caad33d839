/*
 * bem_hip.h -- C ABI of libbem_hip.so, the MI355X (gfx950) implementation of
 * the frequency-sweep hot path of the boundary-element heat-transfer code:
 * RWG/PMCHWT impedance assembly, the dense complex solve and the trace that
 * yields spectral power.
 *
 * Plain C types only (no HIP or torch types): host pointers unless stated,
 * complex numbers as (re, im) double pairs, matrices column-major complex128
 * interleaved with an explicit leading dimension.  Every function returns 0
 * or a negative BEM_E_* code; bem_last_error() gives the message.  A context
 * owns its device buffers; calls on one context are serialised internally.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   bem_assemble        Zts(RWGTree*, RWGTree*, GreensFunction*, Options&)
 *                       src/Impedance.h:21, src/Impedance.cpp:388-413, and
 *                       Zdense src/Impedance.h:20 (external + shared medium)
 *   bem_spectral_power  Real CalcHeatTransfer(Real a, Real eV, Real T1, Real T2,
 *                       RWGTree *top, RWGTree *s, RWGTree *t, Options&)
 *                       src/HeatTransfer.h:33, src/HeatTransfer.cpp:8-59
 *   bem_sweep           the per-frequency loop a driver runs over
 *                       FrequencySweep::f (src/FrequencySweep.h:9-13) /
 *                       the job farm of src/transfer-mpi.cpp:114-179
 *   bem_body_upload     mesh -> RWG construction of ShapeMesh + RWG
 *                       (src/ShapeMesh.cpp:821-947, src/RWG.cpp:5-39), with
 *                       the reference's edge order and orientation
 *   bem_rwg_tree        RWGTree(meshDiel, edgesAtPoint, minSize, minRadius)
 *                       src/RWGTree.cpp:602-761: the reference's unknown order
 *   bem_zgetrf / bem_zgetrf_local / bem_zgetrs / bem_zgemm
 *                       the LAPACK/BLAS FFI the reference binds
 *                       (zgetrf_/zgetri_ src/util/DenseMatrix.cpp:38-77,
 *                       zgemm_ src/util/MatrixOperations.cpp:70-74), on
 *                       device pointers
 */
#ifndef BEM_HIP_H
#define BEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEM_OK 0
#define BEM_E_INVALID (-1)     /* bad argument */
#define BEM_E_DEVICE (-2)      /* HIP runtime error */
#define BEM_E_NOMEM (-3)       /* device allocation failed / over budget */
#define BEM_E_SINGULAR (-4)    /* exactly singular pivot met */
#define BEM_E_MESH (-5)        /* mesh is not a closed 2-manifold */
#define BEM_E_UNSUPPORTED (-6) /* size or mode not supported */
#define BEM_PENDING 1          /* bem_poll: the request is queued or running */

#define BEM_PREC_FP64 0        /* complex128 throughout */
#define BEM_PREC_MIXED 1       /* FP64 assembly; complex64 elimination and products on FP32 MFMA
                                  (FP64 trace sum); FP64 refinement with BEM_SOLVE_REFINE */

typedef struct bem_ctx bem_ctx;
typedef struct bem_body bem_body;

typedef struct {
  int device;        /* HIP device ordinal */
  int lu_block;      /* reserved (the LU panel width is fixed at 64), pass 0 */
  size_t hbm_budget; /* bytes a context may hold, 0 = no limit */
  int lanes;         /* solves kept in flight, one stream each (0 = one per hardware queue,
                        GPU_MAX_HW_QUEUES, at most 4; max 8) */
  size_t cache_bytes;/* device bytes for cached self blocks (0 = 40% of the device).  Without
                        an hbm_budget the cache keeps its memory in segments of 2 GiB or more
                        and reuses it for new blocks (no hipFree, which would synchronise
                        the device, at an eviction); segments left empty when a body is
                        freed go back to the runtime.  With a budget each block is
                        allocated and freed on its own */
  int lookahead;     /* LU look-ahead on a second stream per lane: 1 on, -1 off, 0 on when the
                        2 x lanes streams fit the hardware queues (GPU_MAX_HW_QUEUES, default 4) */
} bem_cfg;

typedef struct {
  double eV;         /* photon energy (eV) */
  double a;          /* metres per geometry unit (x3d scale) */
  double T1, T2;     /* body temperatures (K) */
  double gap[3];     /* translation applied to body 1 (geometry units) */
  double eps1[2], mu1[2];  /* body 1 medium (re, im) at eV */
  double eps2[2], mu2[2];  /* body 2 medium (re, im) at eV */
  int precision;     /* BEM_PREC_* */
  int flags;         /* BEM_SOLVE_* bits */
  double rank_tol;   /* BEM_SOLVE_LOWRANK: relative Frobenius tolerance of the compressed
                        coupling block (0: BEM_LOWRANK_TOL) */
  double err_tol;    /* accuracy the caller accepts: a result whose rel_err_est is at most
                        err_tol is final (0: BEM_PROBE_LIMIT); see Recovery */
} bem_request;

/* Pivoting of the dense solve.  Default: partial pivoting inside each 64 x 64
 * diagonal block (as the reference's block LU over tree blocks,
 * src/util/HMatrix.h:277-316), with the growth factor
 * max(1, max|L|) max|U| / max|A| measured and reported (|L| <= 1 under
 * partial pivoting over whole columns; block-local pivots can leave large
 * multipliers below the block).  BEM_SOLVE_TOURNAMENT requests tournament
 * (communication-avoiding) partial pivoting over whole columns.
 *
 * Error probe.  Every solve ends with a probe of its result: for
 * BEM_PROBE_COLS seeded random vectors w, x = [X w; Sc^-1 w] is checked
 * against the unfactored FP64 blocks -- r = [0; w] - W x in FP64 -- and one
 * correction d = W^-1 r through the solve's own factors gives d_1 ~ dX w,
 * dX the error of X.  probe_err = max|d_1| / max|x_1|; residual =
 * ||r||_F / ||w||_F estimates the relative residual
 * ||[0; I] - W [X; Sc^-1]||_F / ||I||_F.  The error of Re S is, to first
 * order, 1/2 Re tr(sym G2 X^H sym G1 dX); the probe estimates that trace from
 * the same columns (g_j^H d_1j with g_j = sym G1 X sym G2 w_j): rel_err_est
 * sums the terms' magnitudes, err_first_order keeps their signs.  Where W is
 * ill-conditioned, dX holds the rounding of the residual amplified by W^-1
 * in directions Re S hardly sees, so probe_err can exceed the error of Re S
 * by orders of magnitude while rel_err_est does not.
 *
 * Recovery.  A block-local FP64 solve is solved again with tournament
 * pivoting and BEM_SOLVE_REFINE when it met an exactly zero pivot or when
 * rel_err_est exceeds the request's err_tol (BEM_PROBE_LIMIT when 0) -- a
 * tenth of it when the growth exceeds BEM_GROWTH_LIMIT, since large growth
 * may hide an inaccurate factorisation; a complex64 solve is when its growth
 * exceeds the limit.  `restarts` counts such re-solves. */
#define BEM_SOLVE_TOURNAMENT 1
/* BEM_SOLVE_NORECOVER: the first result is final, with its growth, probe and
 * rel_err_est as measured -- no re-solve (Recovery) and no switch from the
 * low-rank to the dense solve; refinement steps still run when requested.
 * For callers that take the error statement and decide themselves. */
#define BEM_SOLVE_NORECOVER 16
#define BEM_GROWTH_LIMIT 16.0
#define BEM_PROBE_COLS 8
#define BEM_PROBE_LIMIT 1e-9
/* Iterative refinement: each step forms the full residual
 * R = [0; I] - W [X; Z2] of the current X and Z2 = Sc^-1 in FP64 against the
 * FP64-assembled blocks, the correction through the solve's factors
 * (complex64 for BEM_PREC_MIXED, complex128 otherwise), and the FP64
 * products and trace of the corrected X, followed by the probe.  Steps
 * repeat until the probe's relative residual is at most BEM_REFINE_TOL, at
 * most BEM_REFINE_MAX of them; a BEM_PREC_MIXED request that reaches the
 * cap is solved again on the FP64 path.  A mixed request that needs
 * tournament pivoting (growth above the limit) is solved on the FP64 path. */
#define BEM_SOLVE_REFINE 2
#define BEM_REFINE_TOL 1e-10
#define BEM_REFINE_MAX 3
/* Low-rank coupling (BEM_SOLVE_LOWRANK, FP64 with block-local pivoting).  The
 * cross block B = Z_ext(1,2) of the two disjoint bodies is compressed to
 * B ~ U V^H with ||B - U V^H||_F <= rank_tol ||B||_F (default
 * BEM_LOWRANK_TOL), and C = Z_ext(2,1) = S B^T S follows from it.  The
 * elimination of the reference's block LU over the two-body tree
 * (src/util/HMatrix.h:277-360: factor the diagonal blocks, substitute into
 * the coupling blocks, update and factor the Schur complement) then runs on
 * the factors: with LU(A) and LU(D) of the cached self blocks,
 *   P = A^-1 U, K = V2^H P, R^H = V^H D^-1, F = R^H U2, T = I - F K,
 *   X = W^-1[0:m1, m1:] = -P T^-1 R^H        (Woodbury for Sc^-1),
 *   4 S = tr(M1 M2), M1 = P^H sym G1 P, M2 = Z sym G2 Z^H, Z = T^-1 R^H,
 * with U2 = S conj(V), V2 = S conj(U); O(m^2 r) per gap instead of O(m^3).
 * Compression: a blocked randomized range finder (64 seeded columns per
 * block, block Gram-Schmidt twice and CholQR2 per block, stop when the
 * residual estimate of the new block is under the tolerance), or, with
 * BEM_LOWRANK_ACA, the reference's adaptive cross approximation with full
 * pivoting (bem_aca) followed by SVD truncation (bem_rk_truncate), both at
 * rank_tol / 10 since neither stopping rule bounds the residual.  The error
 * probe runs against the dense B and C, so rel_err_est includes the
 * truncation; a result above BEM_PROBE_LIMIT, a rank above
 * BEM_LOWRANK_MAX_FRAC min(m1, m2), or a recovery condition of the dense
 * path sends the request to the dense solve.  bem_result.rank reports the
 * rank used (0: dense). */
#define BEM_SOLVE_LOWRANK 4
#define BEM_LOWRANK_ACA 8
#define BEM_LOWRANK_TOL 1e-11
#define BEM_LOWRANK_MAX_FRAC 0.5

typedef struct {
  double P;          /* Re(S) (Theta(omega,T1) - Theta(omega,T2)) */
  double ReS, ImS;   /* S = 1/4 tr(B2^H B1) */
  double theta1, theta2;
  double omega;      /* dimensionless omega a / c */
  double rel_err_est;/* estimated relative error of Re S (see Error probe) */
  int iters;         /* refinement steps taken */
  int status;        /* 0 or BEM_E_* for this request */
  double ms_assemble, ms_factor, ms_products, ms_total;  /* device-event times */
  double growth;     /* max(1, max|L|) max|U| / max|A|, the larger of the two LU factorisations */
  int pivoting;      /* 0 block-local, 1 tournament */
  double ms_queue;   /* served requests: submit to issue (ms); ms_total is then submit to result */
  int precision;     /* BEM_PREC_* the result was computed in */
  int restarts;      /* re-solves with tournament pivoting and refinement (Recovery), or on the FP64
                        path after the refinement cap of a mixed request */
  double probe_err;  /* max|d_1| / max|x_1| of the probe of the returned X */
  double residual;   /* the probe's relative residual of the returned [X; Sc^-1] */
  double full_residual; /* ||R||_F / ||I||_F of the last refinement step's full residual
                           (of the X before that step's correction); -1 without refinement */
  double err_first_order; /* the probe's signed estimate of (Re S_exact - Re S) / |Re S| */
  int rank;          /* rank of the coupling block of a low-rank solve, 0 for the dense solve */
  int pad;
  double ms_compress;/* low-rank solves: compression of B (host clock, ms) */
} bem_result;

/* ---- context ---------------------------------------------------------- */
int bem_ctx_create(bem_ctx **ctx, const bem_cfg *cfg);
void bem_ctx_destroy(bem_ctx *ctx);
const char *bem_last_error(bem_ctx *ctx);
/* Bytes of device memory currently held by the context. */
size_t bem_ctx_device_bytes(bem_ctx *ctx);

/* ---- geometry --------------------------------------------------------- */
/* Upload a closed triangle mesh (xyz: nv x 3, tri: nt x 3 0-based).  Builds
 * the RWG table in the reference's order: edges in creation order while
 * scanning triangles, positive half = creating triangle. */
int bem_body_upload(bem_ctx *ctx, const double *xyz, int nv, const int32_t *tri, int nt, bem_body **out);
/* Number of RWGs (edges) of a body; copies the edge table {p1, p2, t1, t2}
 * (ne x 4) when edges != NULL. */
int bem_body_edges(bem_body *b, int32_t *edges, int *ne);
void bem_body_free(bem_body *b);
/* The order in which the reference lists a body's unknowns: the leaves of
 * the RWGTree its drivers build (RWGTree(meshDiel, edgesAtPoint,
 * minSegmentSize, minRadius), src/RWGTree.cpp:602-761, called from
 * parseShapes src/HeatTransfer.cpp:98-105; its partition src/RWGTree.cpp:
 * 21-559 and the mesher's segmentation src/ShapeMesh.cpp:930-970,
 * 1917-1927).  Host only: no context, no device.  xyz, tri as
 * bem_body_upload, with the body's translation already applied (the
 * reference translates body 1 before meshing); min_size is
 * Options::minSegmentSize (32), min_radius 0 as parseShapes passes.
 * order[i] (room for every edge) is the edge index (bem_body_edges order) of
 * tree position i < *n; leaves[k] (room for every edge) the RWG count of
 * leaf k, k < *nleaves.  *n is below the edge count where the surface has a
 * closed loop of edges sharper than 2 rad: the reference keeps only the part
 * that holds triangle 0.  BEM_E_MESH for an open surface, BEM_E_UNSUPPORTED
 * when the plane search does not converge (the reference aborts). */
int bem_rwg_tree(const double *xyz, int nv, const int32_t *tri, int nt, int min_size, double min_radius,
                 int32_t *order, int *n, int32_t *leaves, int *nleaves);

/* ---- assembly (Zts) --------------------------------------------------- */
/* Impedance matrix of the union of nb (1 or 2) bodies, body k translated by
 * shifts[3k..3k+2], in medium (eps, mu) at dimensionless omega: 2E x 2E,
 * E = total RWGs, rows/cols [J_0, M_0, J_1, M_1, ...], body 0 first.
 * Z is a host pointer (ldz >= 2E) unless flags & BEM_ASM_DEVICE_PTR. */
#define BEM_ASM_DEVICE_PTR 1
int bem_assemble(bem_ctx *ctx, int nb, bem_body *const *bodies, const double *shifts, const double *eps,
                 const double *mu, double omega, void *Z, int64_t ldz, int flags);
/* Pair-class histogram of the same union: {close, g6, g10, g15, g21, g36}. */
int bem_pair_histogram(bem_ctx *ctx, int nb, bem_body *const *bodies, const double *shifts, int64_t *hist6);

/* ---- dipole sources and flux ------------------------------------------ */
/* A point dipole: moment j (real) at p, in the exterior medium (vacuum) for
 * region -1 or inside body `region` (0-based), whose medium it then takes. */
typedef struct {
  double p[3];
  double j[3];
  int32_t region;
  int32_t pad;
} bem_dipole;

/* Right-hand sides of ns dipoles over the union of nb bodies, body k
 * translated by shifts[3k..3k+2]: DipolePointSource::EHTriangle summed over
 * the triangles of each body's tree as Bdense does (src/Source.cpp:121-193),
 * the leaf level of BH (src/Source.cpp:199-215).  B is 2E x ns (leading
 * dimension ldb >= 2E), rows [J_0, M_0, J_1, M_1, ...]; column s is zero on
 * the RWGs outside dipole s's region.  eps, mu: 2 nb doubles, (re, im) of
 * each body's medium.  B is a host pointer unless flags & BEM_ASM_DEVICE_PTR. */
int bem_dipole_rhs(bem_ctx *ctx, int nb, bem_body *const *bodies, const double *shifts, const double *eps,
                   const double *mu, double omega, int ns, const bem_dipole *src, void *B, int64_t ldb, int flags);
/* Power flux through the surface of body `body` of the same union for each
 * of ncol solution columns X (2E x ncol, leading dimension ldx, [J, M] per
 * RWG): flux(top, t, X) (src/Flux.cpp:100-219), summed over the body's
 * triangles.  flux: ncol doubles (host).  X is a host pointer unless
 * flags & BEM_ASM_DEVICE_PTR. */
int bem_flux(bem_ctx *ctx, int nb, bem_body *const *bodies, const double *shifts, int body, double omega,
             const void *X, int64_t ldx, int ncol, int flags, double *flux);
/* flux(ZLU, j, green, p, rwgTree, t, greenExt) (src/Flux.cpp:221-236) for ns
 * dipoles at once: their right-hand sides B, X = W^-1 B with W = Z_ext +
 * blkdiag(Z_int(1), Z_int(2)) of the pair at the request's omega, gap and
 * media (the matrix of bem_spectral_power), factored with tournament partial
 * pivoting, then the flux of each column through each body.
 * flux: ns x 2 doubles, flux[2s + b] through body b+1. */
int bem_dipole_flux(bem_ctx *ctx, bem_body *b1, bem_body *b2, const bem_request *req, int ns,
                    const bem_dipole *src, double *flux);

/* ---- spectral power (CalcHeatTransfer) -------------------------------- */
/* Gap-independent blocks -- Z_ext(b,b) + Z_int(b) (factored for body 1) and
 * sym(Z_int(b)) -- are cached on the device per (body, medium, omega) and
 * reused by later requests on any lane; only the cross blocks Z_ext(1,2),
 * Z_ext(2,1) are assembled per gap.  Cache capacity: cfg.cache_bytes. */
int bem_spectral_power(bem_ctx *ctx, bem_body *b1, bem_body *b2, const bem_request *req, bem_result *res);
/* n independent requests on the same body pair, results in request order.
 * Up to cfg.lanes requests run concurrently on separate streams; each result
 * is computed by the same kernels in the same order as when run alone. */
int bem_sweep(bem_ctx *ctx, bem_body *b1, bem_body *b2, const bem_request *reqs, int n, bem_result *res);

/* ---- serving: continuous batching ------------------------------------ */
/* Queues one request on the body pair and returns at once with a ticket.  A
 * scheduler thread of the context admits queued requests -- shortest
 * estimated run time first, less the time already waited -- onto free lanes
 * (cfg.lanes streams) whenever the request's device footprint
 * -- lane workspaces plus self blocks not yet cached -- fits cfg.hbm_budget
 * (evicting unused cached blocks and idle workspaces first), so requests that
 * arrive while others run join them on the device as lanes and memory free
 * up.  Each result is computed by the same kernels in the same order as when
 * the request runs alone.  There is no reference counterpart beyond the job
 * farm of src/transfer-mpi.cpp:114-179, whose servants this replaces. */
int bem_submit(bem_ctx *ctx, bem_body *b1, bem_body *b2, const bem_request *req, uint64_t *ticket);
/* Result of a ticket: returns BEM_PENDING while it is queued or running
 * (waiting up to wait_ms; wait_ms < 0 waits until done), else the request's
 * status (0 or BEM_E_*) with *res filled, after which the ticket is retired.
 * BEM_E_INVALID for an unknown ticket. */
int bem_poll(bem_ctx *ctx, uint64_t ticket, bem_result *res, int wait_ms);
/* Requests waiting in the queue and running on lanes. */
int bem_queue_depth(bem_ctx *ctx, int *queued, int *running);

/* ---- low-rank blocks (Rk) on device pointers --------------------------- */
/* Adaptive cross approximation with full pivoting of A (m x n, lda; overwritten
 * by the residual A - U V), as the reference's aca(DenseMatrix&, r, c, eps2)
 * (src/util/DenseMatrix.h:213-279): each step takes the entry of largest
 * |.|^2 of the residual (first in row-major scan order among equal ones),
 * a = its column, b = its row / the pivot, subtracts a b^T, and continues
 * while nu < min(m, n) and |a|^2 |b|^2 / sum_steps |a|^2 |b|^2 > eps2; the
 * step that stops the loop is not kept, as the reference pushes no pivot for
 * it.  Outputs: rank kept pivots rows[]/cols[] (host, kmax ints each), U
 * (device, m x kmax, ldu) and V (device, kmax x n, ldv) whose first rank
 * columns / rows are the kept steps' a and b, so A ~ U(:, :rank) V(:rank, :).
 * BEM_E_UNSUPPORTED when kmax columns are used up before the loop ends. */
int bem_aca(bem_ctx *ctx, void *A, int64_t lda, int m, int n, double eps2, int kmax, void *U, int64_t ldu, void *V,
            int64_t ldv, int32_t *rows, int32_t *cols, int *rank);
/* Rk::truncate (src/util/Rk.cpp:4-53) of the rank-k product U V (U m x k,
 * ldu; V k x n, ldv; device): its singular value decomposition -- here two
 * one-sided Jacobi passes, on V^H and then on U times the first pass's
 * factors, in place of the reference's QR, RQ and SVD of the core -- keeping
 * the singular values s_i >= eps s_max in decreasing order; U (m x knew) <-
 * U_svd S and V (knew x n) <- V_svd^H, as the reference's A and B.  sigma
 * (optional, host, k doubles) receives every singular value in decreasing
 * order. */
int bem_rk_truncate(bem_ctx *ctx, void *U, int64_t ldu, void *V, int64_t ldv, int m, int n, int k, double eps,
                    int *knew, double *sigma);

/* ---- dense complex128 kernels on device pointers ---------------------- */
/* A (n x n, lda) <- P L U in place with tournament partial pivoting; ipiv
 * (device, n ints, 0-based interchanges row i <-> ipiv[i], applied in
 * order).  Returns BEM_E_SINGULAR on a zero pivot. */
int bem_zgetrf(bem_ctx *ctx, void *A, int64_t lda, int n, int32_t *ipiv);
/* A (n x n, lda) <- P L U in place with partial pivoting inside each 64 x 64
 * diagonal block: the factorisation the solve uses by default, as the
 * reference's tree-block LU factors each diagonal block on its own
 * (src/util/HMatrix.h:276-307).  ipiv as bem_zgetrf, every interchange
 * within its block; growth (optional, host) receives max(1, max|L|) max|U| / max|A|.
 * Returns BEM_E_SINGULAR on a zero pivot. */
int bem_zgetrf_local(bem_ctx *ctx, void *A, int64_t lda, int n, int32_t *ipiv, double *growth);
/* Allocation-free variant of bem_zgetrf_local.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels in the same order as bem_zgetrf_local, so the
 * factors, interchanges, growth and status are the same; its workspace --
 * the diagonal-block inverses, the zero-pivot flag and the growth maxima --
 * sits in a fixed device array of the library (8 MiB + 256 bytes per GPU,
 * held until the process ends), and the call allocates nothing.
 * n <= BEM_NOALLOC_MAX_N, else BEM_E_UNSUPPORTED.  Every call on one GPU, from
 * any context, shares that array and finds the last call's contents there:
 * calls that overlap on one GPU race.  It runs on an idle lane of the
 * context (BEM_E_UNSUPPORTED when every lane has a request in flight). */
#define BEM_NOALLOC_MAX_N 4096
int bem_zgetrf_local_noalloc(bem_ctx *ctx, void *A, int64_t lda, int n, int32_t *ipiv, double *growth);
/* B (n x nrhs, ldb) <- A^-1 B from the factors of bem_zgetrf. */
int bem_zgetrs(bem_ctx *ctx, const void *A, int64_t lda, int n, const int32_t *ipiv, void *B, int64_t ldb, int nrhs);
/* C <- alpha A B + beta C (no transposes), alpha/beta as (re, im). */
int bem_zgemm(bem_ctx *ctx, int m, int n, int k, const double *alpha, const void *A, int64_t lda, const void *B,
              int64_t ldb, const double *beta, void *C, int64_t ldc);
/* ---- GEMM tiles and the autotuner ----------------------------------- */
/* Tile variant per GEMM shape class {narrow: an output dimension <= 128,
 * rank update: k <= 128, general} for complex128 (tiles[0..2]; codes 32, 64,
 * 256, 1032, 2032, 1033, 1034, 1064, 2064, 1065, 2034, 1066, 1128, 1129) and
 * complex64 (tiles[3..5]; codes 1..5 four real products per complex
 * multiply-add, 11..15 the same tiles with three, 16 / 17 the 64 x 64 /
 * 128 x 128 tiles on eight waves, 21 / 23 / 26 tiles 11 / 13 / 16 with k
 * steps of 32); 0 selects the built-in shape rule.
 * set6 / get6 may be NULL. */
int bem_gemm_tiles(bem_ctx *ctx, const int *set6, int *get6);
/* Times every tile variant on a representative shape of each class (4096 x
 * 64 x 64, 4096 x 4096 x 64, and 4096 x 4096 x 256, the rank-256 trailing
 * update that dominates the general class of the two-level LU and solves) and
 * installs the fastest per class.
 * With cache_path, a cache written for the same GPU architecture is reused
 * (unless flags & BEM_TUNE_FORCE) and a new measurement is written there as
 * JSON with every timing.  tiles6 (optional) receives the choice. */
#define BEM_TUNE_FORCE 1
int bem_autotune(bem_ctx *ctx, const char *cache_path, int flags, int *tiles6);

/* complex64 C <- alpha A B + beta C on FP32 MFMA (alpha/beta as (re, im)
 * floats): the kernel of the mixed-precision elimination.  No reference
 * counterpart (the reference binds only complex128 BLAS). */
int bem_cgemm(bem_ctx *ctx, int m, int n, int k, const float *alpha, const void *A, int64_t lda, const void *B,
              int64_t ldb, const float *beta, void *C, int64_t ldc);
/* Waits until every stream of the context's device is idle (no reference
 * counterpart: timing harnesses bracket a measured region with it). */
int bem_synchronize(bem_ctx *ctx);
/* Device memory helpers for callers without their own allocator. */
int bem_device_alloc(bem_ctx *ctx, size_t bytes, void **ptr);
int bem_device_free(bem_ctx *ctx, void *ptr);
int bem_memcpy_h2d(bem_ctx *ctx, void *dst, const void *src, size_t bytes);
int bem_memcpy_d2h(bem_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-event timings of the last call: out[0] assembly, out[4] factor and
 * solves, out[5] products and trace (ms, last solve of the call); out[6] sum
 * of the durations of every zgemm launch of the call (ms), out[7] their
 * number, out[8] their real flops (8 m n k each).  out[1..3] are reserved. */
int bem_last_timings(bem_ctx *ctx, double *out9);
/* Kernel classes timed with device events on their own stream during the
 * last bem_spectral_power / bem_sweep call (a class may hold several
 * launches of one kernel). */
#define BEM_KS_ZGEMM 0         /* zgemm, rank-update and general classes */
#define BEM_KS_ZGEMM_NARROW 1  /* zgemm, narrow class (an output side <= 128) */
#define BEM_KS_ZTRACE 2        /* Hermitian-trace GEMM (upper half of X^H B1) */
#define BEM_KS_FAR 3           /* far-pair assembly (6-point rule) */
#define BEM_KS_CLOSE 4         /* close-pair assembly (105 x 91 + analytic) */
#define BEM_KS_GATHER 5        /* tiles -> 2 x 2 RWG blocks of Z */
#define BEM_KS_CGEMM 6         /* complex64 GEMM (mixed precision) */
#define BEM_KS_COUNT 8
/* out4 = {summed kernel ms, launches, algorithmic flops, algorithmic bytes}
 * of class `kind` over the last call.  Algorithmic counts: 8 m n k real
 * flops per complex GEMM and 16 bytes per operand entry read or written once;
 * 8308 flops per 6-point far pair and 1.83e6 per close pair, 288 bytes of
 * tile per pair. */
int bem_kernel_stats(bem_ctx *ctx, int kind, double *out4);

#ifdef __cplusplus
}
#endif
#endif
