"""Dipole sources and flux on the GPU through the C ABI (bem_dipole_rhs,
bem_flux, bem_dipole_flux) against the reference's own Bdense and flux
(tests/golden/make_dipole_golden.py) and the oracle.

Tolerances:
  right-hand sides:  max|B_gpu - B_ref| / max|B_ref| <= 1e-10, the impedance
                     entries' bound: B is a column of the same integrals, and
                     the smooth Green's function's gradient
                     (e (1 - ikR) + k2 R^2 - 1) / R^3 cancels to O((kR)^3), so
                     triangles near the dipole carry ~1e-16 (kR)^-3 relative
                     rounding in either implementation
  flux of a given X: |f_gpu - f_ref| <= 1e-12 sum|dS| (a sum over triangles in
                     a different order), checked as <= 1e-11 |f_ref| here
  dipole flux:       |f_gpu - f_ref| / |f_ref| <= 1e-8 (a dense solve, as Re S)
"""
import sys

import numpy as np
import pytest

from conftest import golden_meshes, load_golden

pytestmark = pytest.mark.gpu

B_TOL = 1e-10
F_TOL = 1e-11
S_TOL = 1e-8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _case():
    g = load_golden("sphere1_cube1_gap0p05")
    d = load_golden("dipole_flux")
    return g, d, golden_meshes(g)


def _src(d):
    return [(int(d["region"][s]), d["p"][s], d["j"][s]) for s in range(len(d["region"]))]


def test_dipole_rhs_matches_reference(gpu_ctx):
    g, d, ms = _case()
    bodies = [gpu_ctx.body(m) for m in ms]
    eps = [complex(g["eps_b0"]), complex(g["eps_b1"])]
    mu = [complex(g["mu_b0"]), complex(g["mu_b1"])]
    B = gpu_ctx.dipole_rhs(bodies, [(0, 0, 0), (0, 0, 0)], eps, mu, float(g["omega"]), _src(d))
    m1 = 2 * bodies[0].ne
    for s in range(3):
        e = nerr(B[:, s], d["B"][s])
        say(f"dipole rhs region {d['region'][s]}: {e:.2e}")
        assert e <= B_TOL
    assert not B[m1:, 1].any() and not B[:m1, 2].any()


def test_flux_matches_reference(gpu_ctx):
    g, d, ms = _case()
    bodies = [gpu_ctx.body(m) for m in ms]
    om = float(g["omega"])
    X = np.stack([d["X_rand"], 2j * d["X_rand"]], axis=1)
    for b in range(2):
        f = gpu_ctx.flux(bodies, [(0, 0, 0), (0, 0, 0)], b, om, X)
        ref = d["flux_rand"][b]
        say(f"flux body {b}: {f[0]!r} vs {ref!r}")
        assert abs(f[0] - ref) <= F_TOL * abs(ref)
        assert abs(f[1] - 4 * ref) <= F_TOL * abs(4 * ref)


def test_dipole_flux_matches_reference(gpu_ctx):
    """The whole chain of flux(ZLU, j, green, p, ...): the pair's matrix
    assembled, factored and solved on the GPU for three dipoles at once."""
    from bem_amd import Request
    g, d, ms = _case()
    b1, b2 = gpu_ctx.body(ms[0]), gpu_ctx.body(ms[1])
    eV, a = float(g["eV"]), float(g["a"])
    req = Request(eV=eV, a=a, gap=(0, 0, 0), eps1=complex(g["eps_b0"]), eps2=complex(g["eps_b1"]))
    f = gpu_ctx.dipole_flux(b1, b2, req, _src(d))
    for s in range(3):
        for b in range(2):
            ref = d["flux_phys"][s][b]
            rel = abs(f[s, b] - ref) / abs(ref)
            say(f"dipole flux source {s} body {b}: {f[s, b]!r} vs {ref!r} rel {rel:.2e}")
            assert rel <= S_TOL


def test_dipole_paths_at_benchmark_size(gpu_ctx):
    """The 1280-triangle icosphere pair at centre gap 2.2 (M = 7680): right-hand
    sides of an exterior dipole in the gap and one inside each sphere, and
    the flux of a random X, against the oracle; the dipole flux of nine
    dipoles at once (the blocked solve) against each dipole alone (the
    few-column solve)."""
    from bem_amd import mesh, Request, lorentz_eps
    from oracle import oracle as O
    s = mesh.icosphere(3)
    s1 = s.translated((2.2, 0, 0))
    T = O.Tree([s1, s])
    eV = 0.118
    om = O.omega_tilde(eV, 1e-6)
    eps = lorentz_eps(eV)
    b1, b2 = gpu_ctx.body(s), gpu_ctx.body(s)
    src = [(-1, (1.1, 0.05, -0.02), (0.0, 0.0, 1.0)), (0, (2.3, 0.1, 0.2), (1.0, 0.0, 0.0)),
           (1, (-0.2, 0.1, 0.3), (0.3, 0.4, 0.5))]
    B = gpu_ctx.dipole_rhs([b1, b2], [(2.2, 0, 0), (0, 0, 0)], [eps, eps], [1.0, 1.0], om, src)
    for k, (r, p, j) in enumerate(src):
        ref = T.dipole_rhs(1.0 if r < 0 else eps, 1.0, om, r, p, j)
        e = nerr(B[:, k], ref)
        say(f"C2 dipole rhs region {r}: {e:.2e}")
        assert e <= B_TOL
    rng = np.random.default_rng(11)
    X = rng.normal(size=2 * T.ne) + 1j * rng.normal(size=2 * T.ne)
    for b in range(2):
        f = gpu_ctx.flux([b1, b2], [(2.2, 0, 0), (0, 0, 0)], b, om, X)[0]
        ref = T.flux(om, X, b)
        say(f"C2 flux body {b}: rel {abs(f - ref) / abs(ref):.2e}")
        assert abs(f - ref) <= F_TOL * abs(ref)
    req = Request(eV=eV, gap=(2.2, 0, 0), eps1=eps, eps2=eps)
    many = src * 3
    fm = gpu_ctx.dipole_flux(b1, b2, req, many)
    for k, sk in enumerate(src):
        f1 = gpu_ctx.dipole_flux(b1, b2, req, [sk])[0]
        for b in range(2):
            rel = abs(fm[k, b] - f1[b]) / abs(f1[b])
            say(f"C2 dipole flux {k} body {b}: {f1[b]:.6e}, batched rel {rel:.1e}")
            assert rel <= 1e-10
        assert np.array_equal(fm[k], fm[k + 3]) and np.array_equal(fm[k], fm[k + 6])


def test_dipole_errors(gpu_ctx):
    from bem_amd import Request
    from bem_amd.runtime import BemError
    g, d, ms = _case()
    bodies = [gpu_ctx.body(m) for m in ms]
    om = float(g["omega"])
    with pytest.raises(BemError):
        gpu_ctx.dipole_rhs(bodies, [(0, 0, 0), (0, 0, 0)], [1, 1], [1, 1], om, [(2, (0, 0, 0), (1, 0, 0))])
    assert gpu_ctx.dipole_rhs(bodies, [(0, 0, 0), (0, 0, 0)], [1, 1], [1, 1], om, []).shape[1] == 0
    req = Request(eV=float(g["eV"]), gap=(0, 0, 0), eps1=complex(g["eps_b0"]), eps2=complex(g["eps_b1"]))
    assert gpu_ctx.dipole_flux(bodies[0], bodies[1], req, []).shape == (0, 2)
