"""ctypes binding of libbem_hip.so (include/bem_hip.h).

Names follow the reference's interface where one exists:
  Context.zts(...)               Zts (src/Impedance.cpp:388-413)
  Context.calc_heat_transfer(...) CalcHeatTransfer (src/HeatTransfer.cpp:8-59)
  Context.sweep(...)             the per-frequency loop over FrequencySweep::f
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

BEM_OK = 0
BEM_E_INVALID = -1
BEM_E_DEVICE = -2
BEM_E_NOMEM = -3
BEM_E_SINGULAR = -4
BEM_E_MESH = -5
BEM_E_UNSUPPORTED = -6
BEM_PENDING = 1
BEM_PREC_FP64 = 0
BEM_PREC_MIXED = 1
BEM_ASM_DEVICE_PTR = 1
BEM_SOLVE_TOURNAMENT = 1
BEM_SOLVE_REFINE = 2
BEM_SOLVE_LOWRANK = 4
BEM_LOWRANK_ACA = 8
BEM_SOLVE_NORECOVER = 16
BEM_GROWTH_LIMIT = 16.0
BEM_PROBE_COLS = 8
BEM_PROBE_LIMIT = 1e-9
BEM_REFINE_TOL = 1e-10
BEM_REFINE_MAX = 3
BEM_NOALLOC_MAX_N = 4096

# src/util/SI.cpp:3-9
SI_HBAR = 1.05457148e-34
SI_KB = 1.3806503e-23
SI_EV = 1.60217646e-19
SI_C = 2.99792458e8


class BemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bem error {code}: {msg}")
        self.code = code


class _Cfg(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("lu_block", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("lanes", ctypes.c_int), ("cache_bytes", ctypes.c_size_t), ("lookahead", ctypes.c_int)]


class _Req(ctypes.Structure):
    _fields_ = [("eV", ctypes.c_double), ("a", ctypes.c_double), ("T1", ctypes.c_double), ("T2", ctypes.c_double),
                ("gap", ctypes.c_double * 3), ("eps1", ctypes.c_double * 2), ("mu1", ctypes.c_double * 2),
                ("eps2", ctypes.c_double * 2), ("mu2", ctypes.c_double * 2), ("precision", ctypes.c_int),
                ("flags", ctypes.c_int), ("rank_tol", ctypes.c_double), ("err_tol", ctypes.c_double)]


class _Res(ctypes.Structure):
    _fields_ = [("P", ctypes.c_double), ("ReS", ctypes.c_double), ("ImS", ctypes.c_double),
                ("theta1", ctypes.c_double), ("theta2", ctypes.c_double), ("omega", ctypes.c_double),
                ("rel_err_est", ctypes.c_double), ("iters", ctypes.c_int), ("status", ctypes.c_int),
                ("ms_assemble", ctypes.c_double), ("ms_factor", ctypes.c_double),
                ("ms_products", ctypes.c_double), ("ms_total", ctypes.c_double), ("growth", ctypes.c_double),
                ("pivoting", ctypes.c_int), ("ms_queue", ctypes.c_double), ("precision", ctypes.c_int),
                ("restarts", ctypes.c_int), ("probe_err", ctypes.c_double), ("residual", ctypes.c_double),
                ("full_residual", ctypes.c_double), ("err_first_order", ctypes.c_double),
                ("rank", ctypes.c_int), ("pad", ctypes.c_int), ("ms_compress", ctypes.c_double)]


class _Dipole(ctypes.Structure):
    _fields_ = [("p", ctypes.c_double * 3), ("j", ctypes.c_double * 3), ("region", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


def _dipoles(src):
    """[(region, p, j), ...] -> a bem_dipole array."""
    n = len(src)
    arr = (_Dipole * max(n, 1))()
    for i, (region, p, j) in enumerate(src):
        arr[i].region = int(region)
        for k in range(3):
            arr[i].p[k] = float(p[k])
            arr[i].j[k] = float(j[k])
    return arr


def lib_path():
    return os.path.join(HERE, "libbem_hip.so")


_lib = None

# every symbol include/bem_hip.h declares
EXPORTS = ["bem_ctx_create", "bem_ctx_destroy", "bem_last_error", "bem_ctx_device_bytes", "bem_body_upload",
           "bem_body_edges", "bem_body_free", "bem_assemble", "bem_pair_histogram", "bem_spectral_power",
           "bem_sweep", "bem_zgetrf", "bem_zgetrf_local", "bem_zgetrs", "bem_zgemm", "bem_device_alloc",
           "bem_device_free", "bem_memcpy_h2d", "bem_memcpy_d2h", "bem_last_timings", "bem_kernel_stats",
           "bem_submit", "bem_poll", "bem_queue_depth", "bem_cgemm", "bem_gemm_tiles", "bem_autotune",
           "bem_synchronize", "bem_zgetrf_local_noalloc", "bem_dipole_rhs", "bem_flux", "bem_dipole_flux",
           "bem_aca", "bem_rk_truncate", "bem_rwg_tree"]


def load_library():
    """Load libbem_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() or make -C bem_amd/csrc")
    L = ctypes.CDLL(p)
    vp, i32, i64, dbl, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
    P = ctypes.POINTER
    L.bem_ctx_create.argtypes = [P(vp), P(_Cfg)]
    L.bem_ctx_destroy.argtypes = [vp]
    L.bem_synchronize.argtypes = [vp]
    L.bem_ctx_destroy.restype = None
    L.bem_last_error.argtypes = [vp]
    L.bem_last_error.restype = ctypes.c_char_p
    L.bem_ctx_device_bytes.argtypes = [vp]
    L.bem_ctx_device_bytes.restype = sz
    L.bem_body_upload.argtypes = [vp, vp, i32, vp, i32, P(vp)]
    L.bem_body_edges.argtypes = [vp, vp, P(i32)]
    L.bem_body_free.argtypes = [vp]
    L.bem_body_free.restype = None
    L.bem_assemble.argtypes = [vp, i32, vp, vp, vp, vp, dbl, vp, i64, i32]
    L.bem_pair_histogram.argtypes = [vp, i32, vp, vp, vp]
    L.bem_dipole_rhs.argtypes = [vp, i32, vp, vp, vp, vp, dbl, i32, vp, vp, i64, i32]
    L.bem_flux.argtypes = [vp, i32, vp, vp, i32, dbl, vp, i64, i32, i32, vp]
    L.bem_dipole_flux.argtypes = [vp, vp, vp, P(_Req), i32, vp, vp]
    L.bem_spectral_power.argtypes = [vp, vp, vp, P(_Req), P(_Res)]
    L.bem_sweep.argtypes = [vp, vp, vp, P(_Req), i32, P(_Res)]
    L.bem_zgetrf.argtypes = [vp, vp, i64, i32, vp]
    L.bem_zgetrf_local.argtypes = [vp, vp, i64, i32, vp, vp]
    L.bem_zgetrf_local_noalloc.argtypes = [vp, vp, i64, i32, vp, vp]
    L.bem_zgetrs.argtypes = [vp, vp, i64, i32, vp, vp, i64, i32]
    L.bem_zgemm.argtypes = [vp, i32, i32, i32, vp, vp, i64, vp, i64, vp, vp, i64]
    L.bem_cgemm.argtypes = [vp, i32, i32, i32, vp, vp, i64, vp, i64, vp, vp, i64]
    L.bem_gemm_tiles.argtypes = [vp, vp, vp]
    L.bem_autotune.argtypes = [vp, ctypes.c_char_p, i32, vp]
    L.bem_device_alloc.argtypes = [vp, sz, P(vp)]
    L.bem_device_free.argtypes = [vp, vp]
    L.bem_memcpy_h2d.argtypes = [vp, vp, vp, sz]
    L.bem_memcpy_d2h.argtypes = [vp, vp, vp, sz]
    L.bem_last_timings.argtypes = [vp, vp]
    L.bem_kernel_stats.argtypes = [vp, i32, vp]
    L.bem_submit.argtypes = [vp, vp, vp, P(_Req), P(ctypes.c_uint64)]
    L.bem_poll.argtypes = [vp, ctypes.c_uint64, P(_Res), i32]
    L.bem_queue_depth.argtypes = [vp, P(i32), P(i32)]
    L.bem_aca.argtypes = [vp, vp, i64, i32, i32, dbl, i32, vp, i64, vp, i64, vp, vp, P(i32)]
    L.bem_rk_truncate.argtypes = [vp, vp, i64, vp, i64, i32, i32, i32, dbl, P(i32), vp]
    L.bem_rwg_tree.argtypes = [vp, i32, vp, i32, i32, dbl, vp, P(i32), vp, P(i32)]
    _lib = L
    return L


def rwg_tree(xyz, tri, shift=None, min_size=32, min_radius=0.0):
    """The reference's order of a body's unknowns (bem_rwg_tree; the RWGTree
    of src/RWGTree.cpp:602-761 as parseShapes builds it, with
    Options::minSegmentSize = 32).  The body is translated by shift first,
    as the reference translates body 1 before meshing.  Returns (order,
    leaves): the mesh edge index at each tree position and the RWG count of
    each leaf.  Host only: needs the library, no device."""
    L = load_library()
    X = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    if shift is not None:
        X = X + np.asarray(shift, dtype=np.float64)
    T = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    cap = 3 * len(T)
    order = np.zeros(cap, dtype=np.int32)
    leaves = np.zeros(cap, dtype=np.int32)
    n, nl = ctypes.c_int(0), ctypes.c_int(0)
    rc = L.bem_rwg_tree(X.ctypes.data, len(X), T.ctypes.data, len(T), int(min_size), float(min_radius),
                        order.ctypes.data, ctypes.byref(n), leaves.ctypes.data, ctypes.byref(nl))
    if rc:
        raise BemError(rc, {BEM_E_MESH: "rwg_tree: the surface is not closed",
                            BEM_E_UNSUPPORTED: "rwg_tree: the plane search did not converge"}.get(rc, "rwg_tree: bad argument"))
    return order[:n.value].copy(), leaves[:nl.value].copy()


def omega_tilde(eV, a):
    """Dimensionless frequency omega a / c (src/HeatTransfer.cpp:10-11)."""
    return eV * SI_EV / SI_HBAR * a / SI_C


def theta(omega, T):
    """src/HeatTransfer.cpp:3-6 with the dimensionless omega; T = 0 gives 0."""
    if T == 0.0:
        return 0.0
    return SI_HBAR * omega / (np.exp(SI_HBAR * omega / (SI_KB * T)) - 1)


def lorentz_eps(eV, eps_inf=6.7, wLO=0.120, wTO=0.0983, gamma=0.00059):
    """SiC-like Lorentz permittivity in eV used by the benchmark workloads."""
    w = eV
    return eps_inf * (w * w - wLO * wLO + 1j * gamma * w) / (w * w - wTO * wTO + 1j * gamma * w)


@dataclass
class Request:
    eV: float
    a: float = 1e-6
    T1: float = 300.0
    T2: float = 0.0
    gap: tuple = (0.0, 0.0, 0.0)
    eps1: complex = 1.0
    mu1: complex = 1.0
    eps2: complex = 1.0
    mu2: complex = 1.0
    precision: int = BEM_PREC_FP64
    flags: int = 0
    rank_tol: float = 0.0
    err_tol: float = 0.0

    def to_c(self):
        r = _Req()
        r.eV, r.a, r.T1, r.T2 = self.eV, self.a, self.T1, self.T2
        for i in range(3):
            r.gap[i] = float(self.gap[i])
        for name in ("eps1", "mu1", "eps2", "mu2"):
            z = complex(getattr(self, name))
            arr = getattr(r, name)
            arr[0], arr[1] = z.real, z.imag
        r.precision = self.precision
        r.flags = self.flags
        r.rank_tol = self.rank_tol
        r.err_tol = self.err_tol
        return r


def _res_dict(r: _Res):
    return {k: getattr(r, k) for k, _ in _Res._fields_}


class Body:
    def __init__(self, ctx, handle, mesh):
        self.ctx = ctx
        self.handle = handle
        self.mesh = mesh

    @property
    def ne(self):
        n = ctypes.c_int(0)
        self.ctx._check(self.ctx.lib.bem_body_edges(self.handle, None, ctypes.byref(n)))
        return n.value

    def edges(self):
        ne = self.ne
        out = np.zeros((ne, 4), dtype=np.int32)
        n = ctypes.c_int(0)
        self.ctx._check(self.ctx.lib.bem_body_edges(self.handle, out.ctypes.data, ctypes.byref(n)))
        return out

    def free(self):
        # after bem_ctx_destroy the body holds no device memory and
        # bem_body_free only releases the host object
        if self.handle:
            self.ctx.lib.bem_body_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP device context (bem_ctx)."""

    def __init__(self, device=0, hbm_budget=0, lanes=0, cache_bytes=0, lookahead=0):
        self.lib = load_library()
        cfg = _Cfg(device, 0, hbm_budget, lanes, cache_bytes, lookahead)
        h = ctypes.c_void_p()
        rc = self.lib.bem_ctx_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc != 0:
            raise BemError(rc, "bem_ctx_create failed (no HIP device?)")
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise BemError(rc, self.lib.bem_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.bem_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        """Wait until every stream of the device is idle (bem_synchronize)."""
        self._check(self.lib.bem_synchronize(self.h))

    def device_bytes(self):
        return self.lib.bem_ctx_device_bytes(self.h)

    def body(self, mesh) -> Body:
        xyz = np.ascontiguousarray(mesh.xyz, dtype=np.float64)
        tri = np.ascontiguousarray(mesh.tri, dtype=np.int32)
        h = ctypes.c_void_p()
        self._check(self.lib.bem_body_upload(self.h, xyz.ctypes.data, xyz.shape[0], tri.ctypes.data, tri.shape[0],
                                             ctypes.byref(h)))
        return Body(self, h, mesh)

    def zts(self, bodies, shifts, eps, mu, omega) -> np.ndarray:
        """Impedance matrix of the union of bodies (Zts), (2E, 2E) complex128 Fortran order."""
        nb = len(bodies)
        arr = (ctypes.c_void_p * nb)(*[b.handle for b in bodies])
        sh = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(nb, 3))
        E = sum(b.ne for b in bodies)
        M = 2 * E
        Z = np.zeros((M, M), dtype=np.complex128, order="F")
        e = np.array([complex(eps).real, complex(eps).imag])
        m = np.array([complex(mu).real, complex(mu).imag])
        self._check(self.lib.bem_assemble(self.h, nb, arr, sh.ctypes.data, e.ctypes.data, m.ctypes.data,
                                          float(omega), Z.ctypes.data, M, 0))
        return Z

    def pair_histogram(self, bodies, shifts):
        nb = len(bodies)
        arr = (ctypes.c_void_p * nb)(*[b.handle for b in bodies])
        sh = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(nb, 3))
        h = np.zeros(6, dtype=np.int64)
        self._check(self.lib.bem_pair_histogram(self.h, nb, arr, sh.ctypes.data, h.ctypes.data))
        return h

    def dipole_rhs(self, bodies, shifts, eps, mu, omega, src) -> np.ndarray:
        """Right-hand sides of point dipoles (Bdense, src/Source.cpp:167-193)
        over the union of bodies: (2E, ns) complex128.  src: [(region, p, j)]
        with region -1 (exterior vacuum) or a body index; eps, mu: per body."""
        nb = len(bodies)
        arr = (ctypes.c_void_p * nb)(*[b.handle for b in bodies])
        sh = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(nb, 3))
        e = np.array([[complex(x).real, complex(x).imag] for x in eps], dtype=np.float64).ravel()
        m = np.array([[complex(x).real, complex(x).imag] for x in mu], dtype=np.float64).ravel()
        M = 2 * sum(b.ne for b in bodies)
        B = np.zeros((M, max(len(src), 1)), dtype=np.complex128, order="F")
        self._check(self.lib.bem_dipole_rhs(self.h, nb, arr, sh.ctypes.data, e.ctypes.data, m.ctypes.data,
                                            float(omega), len(src), _dipoles(src), B.ctypes.data, M, 0))
        return B[:, :len(src)]

    def flux(self, bodies, shifts, body, omega, X) -> np.ndarray:
        """Flux through body `body`'s surface of each column of X (flux(top,
        t, X), src/Flux.cpp:100-219)."""
        nb = len(bodies)
        arr = (ctypes.c_void_p * nb)(*[b.handle for b in bodies])
        sh = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(nb, 3))
        X = np.asarray(X, dtype=np.complex128)
        if X.ndim == 1:
            X = X[:, None]
        X = np.asfortranarray(X)
        out = np.zeros(X.shape[1])
        self._check(self.lib.bem_flux(self.h, nb, arr, sh.ctypes.data, int(body), float(omega), X.ctypes.data,
                                      X.shape[0], X.shape[1], 0, out.ctypes.data))
        return out

    def dipole_flux(self, b1: Body, b2: Body, req: Request, src) -> np.ndarray:
        """flux(ZLU, j, green, p, ...) (src/Flux.cpp:221-236) for each dipole
        of src through both bodies: (ns, 2)."""
        r = req.to_c()
        out = np.zeros((max(len(src), 1), 2))
        self._check(self.lib.bem_dipole_flux(self.h, b1.handle, b2.handle, ctypes.byref(r), len(src), _dipoles(src),
                                             out.ctypes.data))
        return out[:len(src)]

    def calc_heat_transfer(self, b1: Body, b2: Body, req: Request) -> dict:
        """CalcHeatTransfer: spectral power for one (frequency, gap) request."""
        r = req.to_c()
        out = _Res()
        self._check(self.lib.bem_spectral_power(self.h, b1.handle, b2.handle, ctypes.byref(r), ctypes.byref(out)))
        return _res_dict(out)

    def sweep(self, b1: Body, b2: Body, reqs) -> list:
        n = len(reqs)
        arr = (_Req * n)(*[q.to_c() for q in reqs])
        out = (_Res * n)()
        rc = self.lib.bem_sweep(self.h, b1.handle, b2.handle, arr, n, out)
        res = [_res_dict(out[i]) for i in range(n)]
        if rc != 0:
            raise BemError(rc, self.lib.bem_last_error(self.h).decode())
        return res

    # ---- serving (continuous batching)
    def submit(self, b1: Body, b2: Body, req: Request) -> int:
        """Queue one request; returns its ticket (bem_submit)."""
        t = ctypes.c_uint64(0)
        self._check(self.lib.bem_submit(self.h, b1.handle, b2.handle, ctypes.byref(req.to_c()), ctypes.byref(t)))
        return t.value

    def poll(self, ticket: int, wait_ms: int = 0):
        """Result dict of a ticket, or None while it is pending (bem_poll).
        A failed request returns its dict with a nonzero 'status'."""
        r = _Res()
        rc = self.lib.bem_poll(self.h, ctypes.c_uint64(ticket), ctypes.byref(r), int(wait_ms))
        if rc == BEM_PENDING:
            return None
        if rc == BEM_E_INVALID and r.status == 0:
            raise BemError(rc, f"unknown ticket {ticket}")
        return _res_dict(r)

    def queue_depth(self):
        q, r = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.bem_queue_depth(self.h, ctypes.byref(q), ctypes.byref(r)))
        return q.value, r.value

    # ---- GEMM tiles
    def gemm_tiles(self, tiles=None):
        """Set (6 codes: complex128 narrow/rank/general, complex64 same) and
        return the tile table (bem_gemm_tiles)."""
        out = np.zeros(6, dtype=np.int32)
        st = None if tiles is None else np.ascontiguousarray(tiles, dtype=np.int32)
        self._check(self.lib.bem_gemm_tiles(self.h, None if st is None else st.ctypes.data, out.ctypes.data))
        return [int(x) for x in out]

    def autotune(self, cache_path=None, force=False):
        """Run the tile autotuner (bem_autotune); returns the chosen table."""
        out = np.zeros(6, dtype=np.int32)
        p = None if cache_path is None else os.fsencode(cache_path)
        self._check(self.lib.bem_autotune(self.h, p, 1 if force else 0, out.ctypes.data))
        return [int(x) for x in out]

    KERNEL_CLASSES = {"zgemm": 0, "zgemm_narrow": 1, "ztrace": 2, "far": 3, "close": 4, "gather": 5, "cgemm": 6}

    def kernel_stats(self):
        """Per kernel class of the last solve/sweep call: summed kernel ms,
        launches, algorithmic flops and bytes (bem_kernel_stats)."""
        out = {}
        buf = np.zeros(4)
        for name, k in self.KERNEL_CLASSES.items():
            self._check(self.lib.bem_kernel_stats(self.h, k, buf.ctypes.data))
            out[name] = dict(ms=buf[0], launches=int(buf[1]), flops=buf[2], bytes=buf[3])
        return out

    def last_timings(self):
        out = np.zeros(9)
        self._check(self.lib.bem_last_timings(self.h, out.ctypes.data))
        return dict(assemble_ms=out[0], factor_ms=out[4], products_ms=out[5], gemm_ms=out[6],
                    gemm_launches=int(out[7]), gemm_flops=out[8])

    # ---- dense kernels on host arrays (device round trip), for tests
    def zgemm_host(self, A, B, C=None, alpha=1.0, beta=0.0):
        A = np.asfortranarray(A, dtype=np.complex128)
        B = np.asfortranarray(B, dtype=np.complex128)
        m, k = A.shape
        k2, n = B.shape
        assert k == k2
        C = np.zeros((m, n), dtype=np.complex128, order="F") if C is None else np.asfortranarray(C, dtype=np.complex128)
        dA, dB, dC = self._dev(A), self._dev(B), self._dev(C)
        try:
            al = np.array([complex(alpha).real, complex(alpha).imag])
            be = np.array([complex(beta).real, complex(beta).imag])
            self._check(self.lib.bem_zgemm(self.h, m, n, k, al.ctypes.data, dA, m, dB, k, be.ctypes.data, dC, m))
            out = np.zeros_like(C)
            self._check(self.lib.bem_memcpy_d2h(self.h, out.ctypes.data, dC, out.nbytes))
            return out
        finally:
            for p in (dA, dB, dC):
                self.lib.bem_device_free(self.h, p)

    def cgemm_host(self, A, B, C=None, alpha=1.0, beta=0.0):
        """complex64 alpha A B + beta C on the device (bem_cgemm)."""
        A = np.asfortranarray(A, dtype=np.complex64)
        B = np.asfortranarray(B, dtype=np.complex64)
        m, k = A.shape
        n = B.shape[1]
        C = np.zeros((m, n), dtype=np.complex64, order="F") if C is None else np.asfortranarray(C, dtype=np.complex64)
        dA, dB, dC = self._dev(A), self._dev(B), self._dev(C)
        try:
            al = np.array([complex(alpha).real, complex(alpha).imag], dtype=np.float32)
            be = np.array([complex(beta).real, complex(beta).imag], dtype=np.float32)
            self._check(self.lib.bem_cgemm(self.h, m, n, k, al.ctypes.data, dA, m, dB, k, be.ctypes.data, dC, m))
            out = np.zeros_like(C)
            self._check(self.lib.bem_memcpy_d2h(self.h, out.ctypes.data, dC, out.nbytes))
            return out
        finally:
            for p in (dA, dB, dC):
                self.lib.bem_device_free(self.h, p)

    def zgetrf_host(self, A):
        """LU factors and interchanges of a square matrix (device round trip)."""
        A = np.asfortranarray(A, dtype=np.complex128)
        n = A.shape[0]
        dA = self._dev(A)
        dp = self._dev(np.zeros(n + 1, dtype=np.int32))
        try:
            rc = self.lib.bem_zgetrf(self.h, dA, n, n, dp)
            LU = np.zeros_like(A)
            piv = np.zeros(n + 1, dtype=np.int32)
            self._check(self.lib.bem_memcpy_d2h(self.h, LU.ctypes.data, dA, LU.nbytes))
            self._check(self.lib.bem_memcpy_d2h(self.h, piv.ctypes.data, dp, piv.nbytes))
            self._check(rc)
            return LU, piv[:n]
        finally:
            self.lib.bem_device_free(self.h, dA)
            self.lib.bem_device_free(self.h, dp)

    def zgetrf_local_host(self, A):
        """LU factors, interchanges and growth factor with pivoting inside each
        64 x 64 diagonal block (bem_zgetrf_local, device round trip); a
        singular block raises BemError(BEM_E_SINGULAR)."""
        A = np.asfortranarray(A, dtype=np.complex128)
        n = A.shape[0]
        dA = self._dev(A)
        dp = self._dev(np.zeros(n + 1, dtype=np.int32))
        g = ctypes.c_double(0.0)
        try:
            rc = self.lib.bem_zgetrf_local(self.h, dA, n, n, dp, ctypes.byref(g))
            LU = np.zeros_like(A)
            piv = np.zeros(n + 1, dtype=np.int32)
            self._check(self.lib.bem_memcpy_d2h(self.h, LU.ctypes.data, dA, LU.nbytes))
            self._check(self.lib.bem_memcpy_d2h(self.h, piv.ctypes.data, dp, piv.nbytes))
            self._check(rc)
            return LU, piv[:n], g.value
        finally:
            self.lib.bem_device_free(self.h, dA)
            self.lib.bem_device_free(self.h, dp)

    def aca_host(self, A, eps2, kmax=None):
        """Adaptive cross approximation with full pivoting of A (bem_aca,
        device round trip): (rows, cols, U, V, residual) with A ~ U V."""
        A = np.asfortranarray(A, dtype=np.complex128)
        m, n = A.shape
        kmax = min(m, n) + 1 if kmax is None else kmax
        dA = self._dev(A)
        dU = self._dev(np.zeros((m, kmax), dtype=np.complex128, order="F"))
        dV = self._dev(np.zeros((kmax, n), dtype=np.complex128, order="F"))
        rows = np.zeros(kmax, dtype=np.int32)
        cols = np.zeros(kmax, dtype=np.int32)
        rank = ctypes.c_int(0)
        try:
            self._check(self.lib.bem_aca(self.h, dA, m, m, n, float(eps2), kmax, dU, m, dV, kmax, rows.ctypes.data,
                                         cols.ctypes.data, ctypes.byref(rank)))
            r = rank.value
            U = np.zeros((m, kmax), dtype=np.complex128, order="F")
            V = np.zeros((kmax, n), dtype=np.complex128, order="F")
            R = np.zeros_like(A)
            self._check(self.lib.bem_memcpy_d2h(self.h, U.ctypes.data, dU, U.nbytes))
            self._check(self.lib.bem_memcpy_d2h(self.h, V.ctypes.data, dV, V.nbytes))
            self._check(self.lib.bem_memcpy_d2h(self.h, R.ctypes.data, dA, R.nbytes))
            return rows[:r].copy(), cols[:r].copy(), U[:, :r].copy(), V[:r, :].copy(), R
        finally:
            for p in (dA, dU, dV):
                self.lib.bem_device_free(self.h, p)

    def rk_truncate_host(self, U, V, eps):
        """Rk::truncate of U V (bem_rk_truncate, device round trip):
        (U', V', sigma) with U' V' the SVD truncation at eps s_max."""
        U = np.asfortranarray(U, dtype=np.complex128)
        V = np.asfortranarray(V, dtype=np.complex128)
        m, k = U.shape
        n = V.shape[1]
        dU, dV = self._dev(U), self._dev(V)
        kn = ctypes.c_int(0)
        sig = np.zeros(max(k, 1), dtype=np.float64)
        try:
            self._check(self.lib.bem_rk_truncate(self.h, dU, m, dV, k, m, n, k, float(eps), ctypes.byref(kn),
                                                 sig.ctypes.data))
            Uo = np.zeros_like(U)
            Vo = np.zeros_like(V)
            self._check(self.lib.bem_memcpy_d2h(self.h, Uo.ctypes.data, dU, Uo.nbytes))
            self._check(self.lib.bem_memcpy_d2h(self.h, Vo.ctypes.data, dV, Vo.nbytes))
            r = kn.value
            return Uo[:, :r].copy(), Vo[:r, :].copy(), sig[:k].copy()
        finally:
            for p in (dU, dV):
                self.lib.bem_device_free(self.h, p)

    def zgesv_host(self, A, B):
        A = np.asfortranarray(A, dtype=np.complex128)
        B = np.asfortranarray(B, dtype=np.complex128)
        n = A.shape[0]
        nrhs = B.shape[1]
        dA, dB = self._dev(A), self._dev(B)
        dp = self._dev(np.zeros(n + 1, dtype=np.int32))
        try:
            self._check(self.lib.bem_zgetrf(self.h, dA, n, n, dp))
            self._check(self.lib.bem_zgetrs(self.h, dA, n, n, dp, dB, n, nrhs))
            X = np.zeros_like(B)
            self._check(self.lib.bem_memcpy_d2h(self.h, X.ctypes.data, dB, X.nbytes))
            return X
        finally:
            for p in (dA, dB, dp):
                self.lib.bem_device_free(self.h, p)

    def _dev(self, arr):
        p = ctypes.c_void_p()
        self._check(self.lib.bem_device_alloc(self.h, arr.nbytes, ctypes.byref(p)))
        self._check(self.lib.bem_memcpy_h2d(self.h, p, arr.ctypes.data, arr.nbytes))
        return p

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
