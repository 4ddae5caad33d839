"""Serving through the C ABI (bem_submit / bem_poll): continuous batching of
requests on several lanes, admission by device footprint under an HBM budget.

Tolerance: each served result must be bit-identical to the same request
solved alone on a fresh context (the kernels and their order do not depend on
what else is in flight), and within 1e-8 relative of the oracle in Re S.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

S_TOL = 1e-8


def say(*a):
    print(*a, flush=True)


def _mix():
    """A request mix over three body pairs, several gaps and frequencies."""
    from bem_amd import mesh, Request, lorentz_eps
    cube, sph, box = mesh.cube(2), mesh.icosphere(2), mesh.cuboid(1.5, 1.0, 0.5, 3, 2, 1)
    pairs = [(sph, sph, (2.3, 0.0, 0.0)), (cube, cube, (1.2, 0.0, 0.0)), (sph, box, (0.0, 0.0, 1.7))]
    reqs = []
    for i, f in enumerate([0.09, 0.105, 0.118, 0.13]):
        for p, (m1, m2, gap) in enumerate(pairs):
            g = (gap[0] + 0.05 * (i % 2), gap[1], gap[2]) if p != 2 else (gap[0], gap[1], gap[2] + 0.1 * (i % 2))
            reqs.append((p, Request(eV=f, gap=g, eps1=lorentz_eps(f), eps2=lorentz_eps(f))))
    return pairs, reqs


def test_served_results_equal_alone():
    """Requests submitted in a burst run concurrently on 3 lanes and share
    cached self blocks; each result equals the request solved alone."""
    import bem_amd
    pairs, reqs = _mix()
    say(f"serving: {len(reqs)} requests over {len(pairs)} body pairs")
    with bem_amd.Context(0, lanes=3, lookahead=-1) as ctx:
        bodies = [(ctx.body(a), ctx.body(b)) for a, b, _ in pairs]
        tickets = [ctx.submit(*bodies[p], q) for p, q in reqs]
        served = [ctx.poll(t, wait_ms=-1) for t in tickets]
        assert ctx.queue_depth() == (0, 0)
    for r in served:
        assert r["status"] == 0, r
        assert r["ms_total"] >= r["ms_queue"] >= 0.0
    with bem_amd.Context(0, lanes=1, cache_bytes=1, lookahead=-1) as alone_ctx:
        bodies = [(alone_ctx.body(a), alone_ctx.body(b)) for a, b, _ in pairs]
        for (p, q), r in zip(reqs, served):
            a = alone_ctx.calc_heat_transfer(*bodies[p], q)
            assert a["ReS"] == r["ReS"] and a["ImS"] == r["ImS"] and a["P"] == r["P"], (p, q.eV)


def test_served_result_matches_oracle():
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps
    from oracle import oracle as O
    s, c = mesh.icosphere(2), mesh.cube(2)
    eV, gap = 0.112, (2.0, 0.1, 0.0)
    eps = lorentz_eps(eV)
    ref = O.spectral_power(s.translated(gap), c, eV, 1e-6, 300.0, 0.0, eps, eps, return_parts=True)
    with bem_amd.Context(0, lanes=2) as ctx:
        b1, b2 = ctx.body(s), ctx.body(c)
        t = ctx.submit(b1, b2, Request(eV=eV, gap=gap, eps1=eps, eps2=eps))
        r = ctx.poll(t, wait_ms=-1)
    assert r["status"] == 0
    assert abs(r["ReS"] - ref["ReS"]) <= S_TOL * abs(ref["ReS"])


def test_budget_admission_serialises_and_rejects():
    """Under a budget that holds one request's footprint but not two, the
    requests still all complete (admitted one at a time, evicting cached
    blocks and idle workspaces); a budget below one request's footprint fails
    that request with BEM_E_NOMEM instead of over-allocating."""
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps, BEM_E_NOMEM
    s = mesh.icosphere(2)   # 320 triangles, m = 960 per body
    reqs = [Request(eV=f, gap=(2.4, 0, 0), eps1=lorentz_eps(f), eps2=lorentz_eps(f)) for f in (0.09, 0.1, 0.11, 0.12)]
    # any budget makes the context allocate each block exactly (no arena
    # segments), so its device bytes are the requests' own footprint
    with bem_amd.Context(0, lanes=1, lookahead=-1, hbm_budget=1 << 50) as ref_ctx:
        b1, b2 = ref_ctx.body(s), ref_ctx.body(s)
        alone = [ref_ctx.calc_heat_transfer(b1, b2, q) for q in reqs]
        one = ref_ctx.device_bytes()
    say(f"one request footprint ~ {one / 1e6:.1f} MB")
    budget = int(one * 1.5)
    with bem_amd.Context(0, lanes=3, hbm_budget=budget, lookahead=-1) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        ts = [ctx.submit(b1, b2, q) for q in reqs]
        got = [ctx.poll(t, wait_ms=-1) for t in ts]
        assert ctx.device_bytes() <= budget
    for a, g in zip(alone, got):
        assert g["status"] == 0
        assert a["ReS"] == g["ReS"]
    with bem_amd.Context(0, lanes=2, hbm_budget=one // 4) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        t = ctx.submit(b1, b2, reqs[0])
        r = ctx.poll(t, wait_ms=-1)
        assert r["status"] == BEM_E_NOMEM


def test_freeing_a_body_fails_its_queued_requests():
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps, BEM_E_INVALID
    s = mesh.icosphere(2)
    with bem_amd.Context(0, lanes=1) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        ts = [ctx.submit(b1, b2, Request(eV=f, gap=(2.4, 0, 0), eps1=lorentz_eps(f), eps2=lorentz_eps(f)))
              for f in np.linspace(0.09, 0.13, 6)]
        b1.free()
        res = [ctx.poll(t, wait_ms=-1) for t in ts]
    st = [r["status"] for r in res]
    assert all(x in (0, BEM_E_INVALID) for x in st)
    assert st[-1] == BEM_E_INVALID   # queued behind others on one lane when the body went


def test_poll_pending_then_done():
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps
    s = mesh.icosphere(3)
    with bem_amd.Context(0, lanes=1) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        t = ctx.submit(b1, b2, Request(eV=0.1, gap=(2.2, 0, 0), eps1=lorentz_eps(0.1), eps2=lorentz_eps(0.1)))
        first = ctx.poll(t, wait_ms=0)
        deadline = time.time() + 60
        r = None
        while r is None and time.time() < deadline:
            r = ctx.poll(t, wait_ms=50)
        assert r is not None and r["status"] == 0
        assert first is None or first["status"] == 0


def test_cache_arena_reuses_evicted_memory():
    """A sweep over more frequencies than the self-block cache holds, in a
    context without an HBM budget: evicted blocks return to the cache's arena
    and new ones are carved from it, so the device footprint stops growing
    once the first requests have sized the lanes and the arena, and every
    result equals the same request solved on a fresh context bit for bit
    (blocks carved from reused memory hold exactly what fresh ones would)."""
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps
    s = mesh.icosphere(2)   # m = 960: about 76 MB of self entries per frequency
    freqs = np.linspace(0.09, 0.13, 24)
    reqs = [Request(eV=float(f), gap=(2.4, 0, 0), eps1=lorentz_eps(float(f)), eps2=lorentz_eps(float(f)))
            for f in freqs]
    got, held = [], []
    with bem_amd.Context(0, lanes=2, cache_bytes=200 << 20, lookahead=-1) as ctx:
        b1, b2 = ctx.body(s), ctx.body(s)
        for k in range(3):
            got += ctx.sweep(b1, b2, reqs[8 * k:8 * k + 8])
            held.append(ctx.device_bytes())
            say(f"after {8 * (k + 1)} requests: {held[-1] / 2**20:.0f} MiB held")
    assert held[1] == held[0] and held[2] == held[0]
    with bem_amd.Context(0, lanes=1, cache_bytes=1, lookahead=-1) as alone_ctx:
        a1, a2 = alone_ctx.body(s), alone_ctx.body(s)
        for q, g in zip(reqs, got):
            a = alone_ctx.calc_heat_transfer(a1, a2, q)
            assert g["status"] == 0 and a["status"] == 0
            assert g["ReS"] == a["ReS"], (q.eV, g["ReS"], a["ReS"])


def test_cache_arena_mixed_sizes():
    """Self entries of three sizes through a 40 MB cache, below one 2 GiB
    arena segment, on two lanes: the first block grows a segment past the
    cap, the cache evicts entries of one size, and the arena carves blocks of
    other sizes from their freed memory.  Each served result equals the
    request solved alone on a fresh context, bit for bit."""
    import bem_amd
    from bem_amd import mesh, Request, lorentz_eps
    bodies = [(mesh.icosphere(1), (2.4, 0, 0)), (mesh.icosphere(2), (2.4, 0, 0)), (mesh.cube(3), (1.3, 0, 0))]
    reqs = []
    for i, f in enumerate(np.linspace(0.09, 0.13, 12)):
        k = i % 3
        reqs.append((k, Request(eV=float(f), gap=bodies[k][1], eps1=lorentz_eps(float(f)), eps2=lorentz_eps(float(f)))))
    with bem_amd.Context(0, lanes=2, cache_bytes=40 << 20, lookahead=-1) as ctx:
        hs = [(ctx.body(m), ctx.body(m)) for m, _ in bodies]
        ts = [ctx.submit(hs[k][0], hs[k][1], q) for k, q in reqs]
        got = [ctx.poll(t, wait_ms=-1) for t in ts]
    with bem_amd.Context(0, lanes=1, cache_bytes=1, lookahead=-1) as alone_ctx:
        ha = [(alone_ctx.body(m), alone_ctx.body(m)) for m, _ in bodies]
        for (k, q), g in zip(reqs, got):
            a = alone_ctx.calc_heat_transfer(ha[k][0], ha[k][1], q)
            assert g["status"] == 0 and a["status"] == 0
            assert g["ReS"] == a["ReS"], (k, q.eV, g["ReS"], a["ReS"])
