"""Python side of the oracle: ctypes bindings of libbem_oracle.so (the C
restatement of the reference's impedance assembly) and a numpy restatement
of the reference's CalcHeatTransfer composition.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the product path.

  CalcHeatTransfer ....... src/HeatTransfer.cpp:8-59
  Theta .................. src/HeatTransfer.cpp:3-6
  SI constants ........... src/util/SI.cpp:3-9
  inverse (zgetrf+zgetri)  src/util/DenseMatrix.cpp:65-77
  symm ................... src/util/DenseMatrix.cpp:477-486
  multiply_Adagger_B ..... src/util/DenseMatrix.cpp:468-475
  trace .................. src/util/DenseMatrix.h:489-497
  dipole RHS (C) ......... src/Source.cpp:88-193
  flux (C) ............... src/Flux.cpp:100-219
  aca .................... src/util/DenseMatrix.h:213-279
  Rk::truncate ........... src/util/Rk.cpp:4-53
  block LU over the two-body tree with Rk coupling blocks (Woodbury form)
                           src/util/HMatrix.h:277-360
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbem_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "zts_ref")
REF_ACA = os.path.join(HERE, "_ref", "aca_ref")

# src/util/SI.cpp:3-9
SI_HBAR = 1.05457148e-34
SI_KB = 1.3806503e-23
SI_EV = 1.60217646e-19
SI_C = 2.99792458e8

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.oracle_mesh_edges.argtypes = [ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, i32p]
        L.oracle_mesh_edges.restype = ctypes.c_int
        L.oracle_zts.argtypes = [ctypes.c_int, f64p, ctypes.c_int, i32p, i32p, ctypes.c_int, i32p,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
        L.oracle_zts.restype = ctypes.c_int
        L.oracle_pair_histogram.argtypes = [ctypes.c_int, f64p, ctypes.c_int, i32p, i32p, i64p]
        L.oracle_eh_nonsmooth_tri.argtypes = [f64p, ctypes.c_int, ctypes.c_int, f64p, f64p]
        L.oracle_green.argtypes = [ctypes.c_double] * 5 + [f64p, f64p, ctypes.c_int, f64p]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_dipole_rhs.argtypes = [ctypes.c_int, f64p, ctypes.c_int, i32p, ctypes.c_int, i32p] + \
            [ctypes.c_double] * 5 + [ctypes.c_int, f64p, f64p, i32p, ctypes.c_void_p]
        L.oracle_dipole_rhs.restype = ctypes.c_int
        L.oracle_flux.argtypes = [ctypes.c_int, f64p, ctypes.c_int, i32p, ctypes.c_int, i32p, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        L.oracle_flux.restype = ctypes.c_double
        _lib = L
    return _lib


# ----------------------------------------------------------------- geometry
def mesh_edges(tri: np.ndarray, nv: int) -> np.ndarray:
    """Reference edge table (ne, 4) int32 {p1, p2, t1, t2} (src/ShapeMesh.cpp:856-926)."""
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    nt = tri.shape[0]
    out = np.zeros((3 * nt, 4), dtype=np.int32)
    ne = lib().oracle_mesh_edges(nv, nt, tri, 3 * nt, out)
    if ne < 0:
        raise ValueError("edge table overflow")
    return out[:ne].copy()


class Tree:
    """A set of bodies in one RWG index space (body 0's RWGs first), the
    analogue of the reference's RWGTree (src/RWGTree.cpp:748-761 order)."""

    def __init__(self, meshes):
        xyz, tri, body, edges = [], [], [], []
        voff = toff = 0
        self.ne_body = []
        for b, m in enumerate(meshes):
            e = mesh_edges(m.tri, m.nv)
            xyz.append(m.xyz)
            tri.append(m.tri + voff)
            body.append(np.full(m.nt, b, dtype=np.int32))
            ee = e.copy()
            ee[:, :2] += voff
            ee[:, 2:] += toff
            edges.append(ee)
            self.ne_body.append(e.shape[0])
            voff += m.nv
            toff += m.nt
        self.xyz = np.ascontiguousarray(np.concatenate(xyz), dtype=np.float64)
        self.tri = np.ascontiguousarray(np.concatenate(tri), dtype=np.int32)
        self.body = np.ascontiguousarray(np.concatenate(body), dtype=np.int32)
        self.edges = np.ascontiguousarray(np.concatenate(edges), dtype=np.int32)

    @property
    def ne(self):
        return self.edges.shape[0]

    def zts(self, eps, mu, omega, ta_range=None) -> np.ndarray:
        """Zts(t, t, g) (src/Impedance.cpp:388-413): (2ne, 2ne) complex128, Fortran order."""
        M = 2 * self.ne
        Z = np.zeros((M, M), dtype=np.complex128, order="F")
        lo, hi = (0, self.tri.shape[0]) if ta_range is None else ta_range
        eps, mu = complex(eps), complex(mu)
        lib().oracle_zts(self.xyz.shape[0], self.xyz, self.tri.shape[0], self.tri, self.body, self.ne,
                         self.edges, eps.real, eps.imag, mu.real, mu.imag, float(omega), lo, hi,
                         Z.ctypes.data, M)
        return Z

    def body_triangles(self, b):
        """Triangle range [lo, hi) of body b."""
        lo = int(np.searchsorted(self.body, b, side="left"))
        return lo, int(np.searchsorted(self.body, b, side="right"))

    def dipole_rhs(self, eps, mu, omega, region, p, j) -> np.ndarray:
        """Right-hand side of a point dipole j at p (src/Source.cpp:121-193):
        2ne complex128.  region -1: the dipole is in the exterior medium
        (eps, mu) and every body's triangles take part; region b: it is inside
        body b, whose medium (eps, mu) is, and only body b's triangles do."""
        eps, mu = complex(eps), complex(mu)
        on = np.ones(self.tri.shape[0], dtype=np.int32) if region < 0 else (self.body == region).astype(np.int32)
        B = np.zeros(2 * self.ne, dtype=np.complex128)
        lib().oracle_dipole_rhs(self.xyz.shape[0], self.xyz, self.tri.shape[0], self.tri, self.ne, self.edges,
                                eps.real, eps.imag, mu.real, mu.imag, float(omega), 1 if region < 0 else 0,
                                np.ascontiguousarray(p, dtype=np.float64), np.ascontiguousarray(j, dtype=np.float64),
                                on, B.ctypes.data)
        return B

    def flux(self, omega, X, body) -> float:
        """Power through the surface of body `body` from the solution X
        (2ne complex128, [J, M] per RWG) -- src/Flux.cpp:100-219."""
        X = np.ascontiguousarray(X, dtype=np.complex128)
        assert X.shape == (2 * self.ne,)
        lo, hi = self.body_triangles(body)
        return lib().oracle_flux(self.xyz.shape[0], self.xyz, self.tri.shape[0], self.tri, self.ne, self.edges,
                                 lo, hi, float(omega), X.ctypes.data)

    def pair_histogram(self):
        h = np.zeros(6, dtype=np.int64)
        lib().oracle_pair_histogram(self.xyz.shape[0], self.xyz, self.tri.shape[0], self.tri, self.body, h)
        return h


# ------------------------------------------------------------------ physics
def omega_tilde(eV: float, a: float) -> float:
    """Dimensionless frequency omega*a/c (src/HeatTransfer.cpp:10-11)."""
    omegaSI = eV * SI_EV / SI_HBAR
    return omegaSI * a / SI_C


def theta(omega: float, T: float) -> float:
    """Bose factor with the dimensionless omega (src/HeatTransfer.cpp:3-6)."""
    if T == 0.0:
        return 0.0
    return SI_HBAR * omega / (np.exp(SI_HBAR * omega / (SI_KB * T)) - 1)


def lorentz_eps(eV: float, eps_inf=6.7, wLO=0.120, wTO=0.0983, gamma=0.00059) -> complex:
    """Fixed eV Lorentz permittivity used by the benchmark configurations."""
    w = eV
    return eps_inf * (w * w - wLO * wLO + 1j * gamma * w) / (w * w - wTO * wTO + 1j * gamma * w)


def symm(G: np.ndarray) -> np.ndarray:
    """G + G^H (src/util/DenseMatrix.cpp:477-486)."""
    return G + G.conj().T


def heat_transfer_from_blocks(W, G1, G2):
    """Re(S) of CalcHeatTransfer from the three impedance matrices
    (src/HeatTransfer.cpp:30-55): inverse of W + blkdiag(G1, G2), then
    S = 1/4 tr(B2^H B1) with B1 = sym(G1) W^-1[0:m1, m1:], B2 = W^-1[0:m1, m1:] sym(G2)."""
    m1 = G1.shape[0]
    W = np.array(W, dtype=np.complex128, order="F", copy=True)
    W[:m1, :m1] += G1
    W[m1:, m1:] += G2
    Winv = np.linalg.inv(W)
    X = Winv[:m1, m1:]
    B1 = symm(G1) @ X
    B2 = X @ symm(G2)
    S = 0.25 * np.sum(np.conj(B2) * B1)
    return S


def spectral_power(mesh1, mesh2, eV, a, T1, T2, eps1, eps2, mu1=1.0, mu2=1.0, return_parts=False):
    """CalcHeatTransfer (src/HeatTransfer.cpp:8-59) with both bodies given in
    their final positions (the gap already applied to body 1)."""
    omega = omega_tilde(eV, a)
    top = Tree([mesh1, mesh2])
    b1 = Tree([mesh1])
    b2 = Tree([mesh2])
    W = top.zts(1.0, 1.0, omega)
    G1 = b1.zts(eps1, mu1, omega)
    G2 = b2.zts(eps2, mu2, omega)
    S = heat_transfer_from_blocks(W, G1, G2)
    th1, th2 = theta(omega, T1), theta(omega, T2)
    P = S.real * (th1 - th2)
    if return_parts:
        return dict(P=P, ReS=S.real, S=S, theta1=th1, theta2=th2, omega=omega)
    return P


# ------------------------------------------- reference driver (fixtures only)
def run_reference(meshes, eps_list, mu_list, omega, workdir):
    """Run the reference's own Zts (oracle/_ref/zts_ref) on the given bodies.

    Returns (edge tables per body, W_top, [G_b ...]).  Only usable where the
    reference was built (this is how the golden fixtures are generated)."""
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    lines = [repr(float(omega)), str(len(meshes))]
    for m, e, mu in zip(meshes, eps_list, mu_list):
        e, mu = complex(e), complex(mu)
        lines.append(f"{e.real!r} {e.imag!r} {mu.real!r} {mu.imag!r} {m.nv} {m.nt}")
        lines += [f"{x!r} {y!r} {z!r}" for x, y, z in m.xyz.tolist()]
        lines += [f"{a} {b} {c}" for a, b, c in m.tri.tolist()]
    out = os.path.join(workdir, "zts_ref.bin")
    subprocess.run([REF_DRIVER, out], input="\n".join(lines) + "\n", text=True, check=True,
                   stdout=subprocess.DEVNULL)
    raw = open(out, "rb").read()
    pos = 0
    edges = []
    for _ in meshes:
        ne = int(np.frombuffer(raw, dtype=np.int64, count=1, offset=pos)[0])
        pos += 8
        e = np.frombuffer(raw, dtype=np.int64, count=4 * ne, offset=pos).reshape(ne, 4)
        pos += 32 * ne
        edges.append(e.astype(np.int32))
    E = sum(e.shape[0] for e in edges)
    mats = []
    for n in [E] + [e.shape[0] for e in edges]:
        cnt = (2 * n) * (2 * n)
        z = np.frombuffer(raw, dtype=np.complex128, count=cnt, offset=pos).reshape(2 * n, 2 * n, order="F")
        pos += 16 * cnt
        mats.append(np.array(z, order="F"))
    return edges, mats[0], mats[1:]


def run_reference_tree(meshes, workdir):
    """Run the reference's RWGTree construction (oracle/_ref/zts_ref --tree)
    on the bodies as parseShapes meshes them (src/HeatTransfer.cpp:75-105).

    Returns, per body, (nseg, order, leaves): the number of segments its
    ShapeMesh keeps, the Edge::index of each RWG in the body's subtree in
    tree order, and the RWG counts of the tree's leaves in that order.
    Fixture generation only."""
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    lines = ["1.0", str(len(meshes))]
    for m in meshes:
        lines.append(f"1.0 0.0 1.0 0.0 {m.nv} {m.nt}")
        lines += [f"{x!r} {y!r} {z!r}" for x, y, z in m.xyz.tolist()]
        lines += [f"{a} {b} {c}" for a, b, c in m.tri.tolist()]
    out = os.path.join(workdir, "tree_ref.bin")
    subprocess.run([REF_DRIVER, out, "--tree"], input="\n".join(lines) + "\n", text=True, check=True,
                   stdout=subprocess.DEVNULL)
    raw = np.frombuffer(open(out, "rb").read(), dtype=np.int64)
    pos = 0
    for _ in meshes:
        pos += 1 + 4 * int(raw[pos])
    res = []
    for _ in meshes:
        nseg, n = int(raw[pos]), int(raw[pos + 1])
        order = raw[pos + 2:pos + 2 + n].astype(np.int32)
        pos += 2 + n
        nl = int(raw[pos])
        leaves = raw[pos + 1:pos + 1 + nl].astype(np.int32)
        pos += 1 + nl
        res.append((nseg, order, leaves))
    return res


def run_reference_source(meshes, eps_list, mu_list, omega, region, p, j, X, workdir):
    """Run the reference's Bdense and flux (oracle/_ref/zts_ref --source).

    Returns (B, [flux over body b ...]): B is the dipole right-hand side
    (2E complex128) of each body's tree in body order, the flux is
    flux(top, body b, X).  Fixture generation only."""
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    lines = [repr(float(omega)), str(len(meshes))]
    for m, e, mu in zip(meshes, eps_list, mu_list):
        e, mu = complex(e), complex(mu)
        lines.append(f"{e.real!r} {e.imag!r} {mu.real!r} {mu.imag!r} {m.nv} {m.nt}")
        lines += [f"{x!r} {y!r} {z!r}" for x, y, z in m.xyz.tolist()]
        lines += [f"{a} {b} {c}" for a, b, c in m.tri.tolist()]
    lines.append(f"{int(region)} " + " ".join(repr(float(v)) for v in list(p) + list(j)))
    lines += [f"{z.real!r} {z.imag!r}" for z in np.asarray(X, dtype=np.complex128).tolist()]
    out = os.path.join(workdir, "source_ref.bin")
    subprocess.run([REF_DRIVER, out, "--source"], input="\n".join(lines) + "\n", text=True, check=True,
                   stdout=subprocess.DEVNULL)
    raw = open(out, "rb").read()
    pos = 0
    E = 0
    for _ in meshes:
        ne = int(np.frombuffer(raw, dtype=np.int64, count=1, offset=pos)[0])
        pos += 8 + 32 * ne
        E += ne
    B = np.frombuffer(raw, dtype=np.complex128, count=2 * E, offset=pos).copy()
    pos += 16 * 2 * E
    S = np.frombuffer(raw, dtype=np.float64, count=len(meshes), offset=pos).copy()
    return B, [float(v) for v in S]


# ------------------------------------------------------ low-rank blocks
def aca(A: np.ndarray, eps2: float):
    """Adaptive cross approximation with full pivoting, step by step as
    src/util/DenseMatrix.h:213-279: the pivot is the entry of largest
    |M(i, j)|^2 of the residual M (first in row-major order among equal
    ones, the reference's i-outer / j-inner scan with '>'), a = M(:, js),
    b = M(is, :) / M(is, js), M -= a b^T, and the loop goes on while
    nu < min(m, n) and |a|^2 |b|^2 / sum |a|^2 |b|^2 > eps2; the step that
    stops it keeps no pivot; a zero maximum breaks the loop.
    Returns (rows, cols, U, V, M) with U V the kept steps' product and
    M = A - U V (the reference's residual is a local copy it drops)."""
    M = np.array(A, dtype=np.complex128, copy=True)
    m, n = M.shape
    k = min(m, n)
    rows, cols, us, vs = [], [], [], []
    tot = 0.0
    nu = 0
    while True:
        a2 = M.real * M.real + M.imag * M.imag
        flat = int(np.argmax(a2))            # first maximum in row-major order
        i, j = divmod(flat, n)
        delta = M[i, j]
        if a2[i, j] == 0.0:
            break
        a = M[:, j].copy()
        b = M[i, :] / delta
        e = float(np.sum(a.real ** 2 + a.imag ** 2)) * float(np.sum(b.real ** 2 + b.imag ** 2))
        tot += e
        nu += 1
        cont = nu < k and e / tot > eps2
        if not cont:
            break        # the reference also subtracts this step from its local copy, which it then drops
        M -= np.outer(a, b)
        rows.append(i)
        cols.append(j)
        us.append(a)
        vs.append(b)
    U = np.array(us, dtype=np.complex128).T.reshape(m, len(us))
    V = np.array(vs, dtype=np.complex128).reshape(len(vs), n)
    return np.array(rows, dtype=np.int32), np.array(cols, dtype=np.int32), U, V, M


def run_reference_aca(A: np.ndarray, eps2: float, workdir: str):
    """Pivots of the reference's own aca() (oracle/_ref/aca_ref) on A."""
    if not os.path.exists(REF_ACA):
        raise RuntimeError("reference ACA driver not built (make -C oracle/ref)")
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape
    payload = np.array([m, n], dtype=np.int32).tobytes() + np.array([eps2], dtype=np.float64).tobytes() + \
        A.tobytes(order="F")
    out = os.path.join(workdir, "aca_ref.bin")
    subprocess.run([REF_ACA, out], input=payload, check=True, stdout=subprocess.DEVNULL)
    raw = open(out, "rb").read()
    cnt = int(np.frombuffer(raw, dtype=np.int32, count=1)[0])
    r = np.frombuffer(raw, dtype=np.int32, count=cnt, offset=4).copy()
    c = np.frombuffer(raw, dtype=np.int32, count=cnt, offset=4 + 4 * cnt).copy()
    return r, c


def rk_truncate(U: np.ndarray, V: np.ndarray, eps: float):
    """Rk::truncate (src/util/Rk.cpp:4-53): QR of U, RQ of V, SVD of the core
    R_U R_V, keep the singular values s >= eps s_max; returns (U S, V^H of the
    kept part, every singular value in decreasing order)."""
    Qa, Ra = np.linalg.qr(U)
    Qb, Rb = np.linalg.qr(V.conj().T)          # V = Rb^H Qb^H (the RQ of V)
    core = Ra @ Rb.conj().T
    W, S, Zh = np.linalg.svd(core)
    keep = int(np.sum(S >= eps * S[0])) if S.size else 0
    Un = (Qa @ W[:, :keep]) * S[:keep]
    Vn = Zh[:keep, :] @ Qb.conj().T
    return Un, Vn, S


def heat_transfer_lowrank(A, D, U, Vh, G1, G2):
    """4 S of the block elimination with the coupling B = U Vh and
    C = S B^T S (J/M signs), in the Woodbury form the GPU path uses:
    P = A^-1 U, K = (S conj U)^H P, Rh = Vh D^-1, F = Rh (S Vh^T),
    Z = (I - F K)^-1 Rh, X = -P Z, 4 S = tr(P^H sym G1 P Z sym G2 Z^H).
    A and D are the self blocks including the interior impedances."""
    m1, m2 = A.shape[0], D.shape[0]
    s1 = np.where(np.arange(m1) % 2 == 0, 1.0, -1.0)
    s2 = np.where(np.arange(m2) % 2 == 0, 1.0, -1.0)
    P = np.linalg.solve(A, U)
    K = (s1[:, None] * U.conj()).conj().T @ P
    Rh = np.linalg.solve(D.T, Vh.T).T
    F = Rh @ (s2[:, None] * Vh.T)
    Z = np.linalg.solve(np.eye(U.shape[1]) - F @ K, Rh)
    M1 = P.conj().T @ symm(G1) @ P
    M2 = Z @ symm(G2) @ Z.conj().T
    return np.trace(M1 @ M2) / 4.0
