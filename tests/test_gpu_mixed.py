"""Mixed-precision path (BEM_PREC_MIXED) through the C ABI: complex64 GEMM on
FP32 MFMA, the complex64 block elimination, the error probe of every result
and FP64 iterative refinement to a relative residual of BEM_REFINE_TOL.

Tolerances:
  cgemm:                |C - C_ref| <= 4e-6 sqrt(k) max|C_ref| (complex64 inputs, FP32 accumulation)
  complex64 solve:      |ReS - ReS_oracle| / |ReS_oracle| <= 2e-5 (no refinement; FP32 rounding
                        amplified by the conditioning of W, ~1e-6 measured on these meshes)
  its error statement:  rel_err_est >= actual error / 10 (the probe estimates the error of X in
                        random directions; Re S is quadratic in X)
  refined:              <= 1e-8, the FP64 path's bound, with the probe's relative residual
                        <= BEM_REFINE_TOL (1e-10)
"""
import numpy as np
import pytest

from conftest import golden_meshes, load_golden

pytestmark = pytest.mark.gpu

MIX_TOL = 2e-5
REF_TOL = 1e-8


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("mnk", [(1, 1, 1), (16, 16, 4), (64, 64, 16), (65, 130, 17), (200, 64, 300), (257, 190, 64),
                                 (300, 301, 129)])
def test_cgemm(gpu_ctx, mnk):
    m, n, k = mnk
    rng = np.random.default_rng(m * 5 + n)
    A = (rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))).astype(np.complex64)
    B = (rng.normal(size=(k, n)) + 1j * rng.normal(size=(k, n))).astype(np.complex64)
    C = (rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))).astype(np.complex64)
    al, be = 0.7 - 0.2j, -1.0 + 0.5j
    got = gpu_ctx.cgemm_host(A, B, C, al, be).astype(np.complex128)
    ref = al * (A.astype(np.complex128) @ B.astype(np.complex128)) + be * C.astype(np.complex128)
    assert np.abs(got - ref).max() <= 4e-6 * np.abs(ref).max() * max(1, k) ** 0.5


def _sphere_pair_ref(eV, gap):
    from bem_amd import mesh, lorentz_eps
    from oracle import oracle as O
    s = mesh.icosphere(2)
    eps = lorentz_eps(eV)
    return s, eps, O.spectral_power(s.translated(gap), s, eV, 1e-6, 300.0, 0.0, eps, eps, return_parts=True)


@pytest.mark.parametrize("eV", [0.09, 0.105, 0.119, 0.135])
def test_mixed_spectral_power_vs_oracle(gpu_ctx, eV):
    from bem_amd import Request, BEM_PREC_MIXED, BEM_PREC_FP64
    from bem_amd.runtime import BEM_SOLVE_REFINE
    gap = (2.2, 0.0, 0.0)
    s, eps, ref = _sphere_pair_ref(eV, gap)
    b1, b2 = gpu_ctx.body(s), gpu_ctx.body(s)
    fast = gpu_ctx.calc_heat_transfer(b1, b2, Request(eV=eV, gap=gap, eps1=eps, eps2=eps, precision=BEM_PREC_MIXED))
    refd = gpu_ctx.calc_heat_transfer(b1, b2, Request(eV=eV, gap=gap, eps1=eps, eps2=eps, precision=BEM_PREC_MIXED,
                                                      flags=BEM_SOLVE_REFINE))
    e_fast = abs(fast["ReS"] - ref["ReS"]) / abs(ref["ReS"])
    e_ref = abs(refd["ReS"] - ref["ReS"]) / abs(ref["ReS"])
    say(f"eV={eV}: complex64 err {e_fast:.2e} (est {fast['rel_err_est']:.2e}, residual {fast['residual']:.2e}), "
        f"refined err {e_ref:.2e} after {refd['iters']} steps (est {refd['rel_err_est']:.2e}, "
        f"residual {refd['residual']:.2e}, full residual before the last step {refd['full_residual']:.2e})")
    assert fast["precision"] == BEM_PREC_MIXED and fast["iters"] == 0 and fast["full_residual"] == -1.0
    assert e_fast <= MIX_TOL
    assert e_fast / 10 <= fast["rel_err_est"] <= 1e-3
    assert refd["precision"] == BEM_PREC_MIXED and 1 <= refd["iters"] <= 3 and refd["restarts"] == 0
    assert refd["residual"] <= 1e-10 and refd["full_residual"] > refd["residual"]
    assert e_ref <= REF_TOL
    assert refd["rel_err_est"] <= 1e-8
    assert fast["P"] == fast["ReS"] * (fast["theta1"] - fast["theta2"])


def test_mixed_refined_matches_reference_fixture(gpu_ctx):
    """The reference's own Zts fixture (cube pair, 0.11 eV): the refined mixed
    result meets the FP64 bound against the oracle composition of the
    reference matrices."""
    from bem_amd import Request, BEM_PREC_MIXED
    from bem_amd.runtime import BEM_SOLVE_REFINE
    g = load_golden("c1_cubes_0p11")
    m1, m2 = golden_meshes(g)
    b1, b2 = gpu_ctx.body(m1), gpu_ctx.body(m2)
    got = gpu_ctx.calc_heat_transfer(b1, b2, Request(eV=float(g["eV"]), a=float(g["a"]), eps1=complex(g["eps_b0"]),
                                                     eps2=complex(g["eps_b1"]), precision=BEM_PREC_MIXED,
                                                     flags=BEM_SOLVE_REFINE))
    ref = float(g["ReS"])
    assert abs(got["ReS"] - ref) <= REF_TOL * abs(ref)


def test_mixed_tournament_request_runs_fp64(gpu_ctx):
    from bem_amd import Request, BEM_PREC_MIXED, BEM_PREC_FP64
    from bem_amd.runtime import BEM_SOLVE_TOURNAMENT
    eV, gap = 0.11, (2.2, 0.0, 0.0)
    s, eps, ref = _sphere_pair_ref(eV, gap)
    b1, b2 = gpu_ctx.body(s), gpu_ctx.body(s)
    r = gpu_ctx.calc_heat_transfer(b1, b2, Request(eV=eV, gap=gap, eps1=eps, eps2=eps, precision=BEM_PREC_MIXED,
                                                   flags=BEM_SOLVE_TOURNAMENT))
    assert r["precision"] == BEM_PREC_FP64 and r["pivoting"] == 1
    assert abs(r["ReS"] - ref["ReS"]) <= REF_TOL * abs(ref["ReS"])


def test_mixed_served_alone_equals_mix():
    """Mixed and FP64 requests served together: every result equals the same
    request solved alone (cached blocks are keyed by precision)."""
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps, BEM_PREC_MIXED, BEM_PREC_FP64
    from bem_amd.runtime import BEM_SOLVE_REFINE
    s = mesh.icosphere(2)
    reqs = []
    for f in (0.095, 0.11, 0.125):
        for prec, flags in ((BEM_PREC_FP64, 0), (BEM_PREC_MIXED, 0), (BEM_PREC_MIXED, BEM_SOLVE_REFINE)):
            reqs.append(Request(eV=f, gap=(2.3, 0, 0), eps1=lorentz_eps(f), eps2=lorentz_eps(f), precision=prec,
                                flags=flags))
    with bem_amd.Context(0, lanes=3, lookahead=-1) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        ts = [ctx.submit(b1, b2, q) for q in reqs]
        served = [ctx.poll(t, wait_ms=-1) for t in ts]
    with bem_amd.Context(0, lanes=1, cache_bytes=1, lookahead=-1) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        for q, r in zip(reqs, served):
            a = ctx.calc_heat_transfer(b1, b2, q)
            assert r["status"] == 0 and a["ReS"] == r["ReS"] and a["ImS"] == r["ImS"]
