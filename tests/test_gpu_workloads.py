"""The benchmark's batch, serving and precision workloads (bench.py --workload
c4, c5, c3) on the GPU through the C ABI, against oracle values.

Tolerances:
  FP64 results:        |ReS - ReS_oracle| / |ReS_oracle| <= 1e-8
  served results:      bit-identical to the same request solved alone
  mixed + refinement:  <= 1e-8 against the oracle, with the probe's relative
                       residual <= BEM_REFINE_TOL (1e-10)
"""
import json
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
S_TOL = 1e-8


def say(*a):
    print(*a, flush=True)


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_c4_instances_served_vs_oracle():
    """Seven instances of the seeded 256-instance batch (smallest M = 1728,
    largest M = 10368, every body kind) submitted together on four lanes:
    each within 1e-8 of the oracle (tests/golden/make_c4_fixtures.py)."""
    import bem_amd
    from bem_amd import Request, lorentz_eps
    path = os.path.join(GOLDEN, "c4_oracle.json")
    ref = json.load(open(path))
    inst = _bench().batch_instances(256)
    picks = sorted(int(k) for k in ref)
    assert len(picks) == 7 and {ref[str(i)]["kind"] for i in picks} == {0, 1, 2}
    with bem_amd.Context(0, lanes=4) as ctx:
        tickets = []
        for i in picks:
            kind, (m1, m2), gap, f = inst[i]
            assert abs(f - ref[str(i)]["eV"]) == 0 and list(gap) == ref[str(i)]["gap"]
            b1, b2 = ctx.body(m1), ctx.body(m2)
            tickets.append((i, b1, b2, ctx.submit(b1, b2, Request(eV=f, gap=gap, eps1=lorentz_eps(f),
                                                                  eps2=lorentz_eps(f)))))
        for i, _, _, t in tickets:
            r = ctx.poll(t, wait_ms=-1)
            d = ref[str(i)]
            rel = abs(r["ReS"] - d["ReS"]) / abs(d["ReS"])
            say(f"c4 instance {i}: kind {d['kind']} M {d['M']} rel {rel:.2e} est {r['rel_err_est']:.1e}")
            assert r["status"] == 0 and rel <= S_TOL


def test_c5_stream_replay():
    """The first 16 requests of the seed-0 Poisson stream (200 req/s: cube
    pairs, sphere pairs, the cone tip over the plate), submitted at their
    arrival offsets, with two cone-tip anchor points and a sphere-pair oracle
    point (C2 geometry) joining the stream: every served result equals the
    same request solved alone, and the three anchors meet the oracle."""
    import bem_amd
    from bem_amd import Request, lorentz_eps
    bench = _bench()
    pairs, stream = bench.serving_stream(64, 200.0)
    stream = stream[:16]
    tip = {d["eV"]: d for d in json.load(open(os.path.join(GOLDEN, "tip_anchor.json")))["points"]}
    c2 = json.load(open(os.path.join(GOLDEN, "c2_oracle.json")))
    kinds = [p[0] for p in pairs]
    k_tip, k_s3 = kinds.index("cone tip over plate"), kinds.index("sphere pair 2x1280")
    f_c2 = [float(f) for f in np.linspace(0.08, 0.14, 64) if f"{f:.6f}" == "0.118095"][0]
    anchors = [(k_tip, tuple(tip[0.118]["shift"]), 0.118, tip[0.118].get("ReS_refined", tip[0.118]["ReS"])),
               (k_tip, tuple(tip[0.105]["shift"]), 0.105, tip[0.105]["ReS"]),
               (k_s3, (2.2, 0.0, 0.0), f_c2, c2["0.118095"]["ReS"])]
    # the anchors join the stream after its 4th, 8th and 12th request
    reqs = []
    for j, (ta, k, gap, f) in enumerate(stream):
        reqs.append((ta, k, gap, f, None))
        if j in (3, 7, 11):
            k2, g2, f2, want = anchors[(j - 3) // 4]
            reqs.append((ta, k2, g2, f2, want))

    def request(gap, f):
        return Request(eV=f, gap=gap, eps1=lorentz_eps(f), eps2=lorentz_eps(f))

    with bem_amd.Context(0, lanes=4) as ctx:
        bodies = [(ctx.body(m1), ctx.body(m2)) for _, m1, m2, _ in pairs]
        t0 = time.perf_counter()
        tickets = []
        for ta, k, gap, f, _ in reqs:
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(ctx.submit(*bodies[k], request(gap, f)))
        served = [ctx.poll(t, wait_ms=-1) for t in tickets]
    say(f"c5 replay: {len(reqs)} requests served; p99 latency "
        f"{sorted(r['ms_total'] for r in served)[-1]:.0f} ms")
    with bem_amd.Context(0, lanes=1, cache_bytes=1, lookahead=-1) as alone_ctx:
        bodies = [(alone_ctx.body(m1), alone_ctx.body(m2)) for _, m1, m2, _ in pairs]
        for (ta, k, gap, f, want), r in zip(reqs, served):
            assert r["status"] == 0, (kinds[k], f)
            a = alone_ctx.calc_heat_transfer(*bodies[k], request(gap, f))
            assert a["ReS"] == r["ReS"] and a["ImS"] == r["ImS"], (kinds[k], gap, f)
            if want is not None:
                rel = abs(r["ReS"] - want) / abs(want)
                say(f"c5 anchor {kinds[k]} eV {f:.6f}: rel {rel:.2e}")
                assert rel <= S_TOL


def test_c3_mixed_refined_at_anchors():
    """Sphere over the plate (M = 33360) at the two oracle anchor points in
    the mixed precision with refinement: complex64 elimination on FP32 MFMA,
    FP64 refinement to a relative residual of 1e-10; Re S within 1e-8 of the
    oracle.  A point whose complex64 factorisation grows past the limit is
    re-solved in FP64, with the same bounds."""
    import bem_amd
    from bem_amd import Request, lorentz_eps, BEM_PREC_MIXED
    from bem_amd.runtime import BEM_SOLVE_REFINE, BEM_REFINE_TOL
    sys.path.insert(0, GOLDEN)
    from make_anchor import WORKLOADS
    _, body1, body2, _ = WORKLOADS["c3"]
    points = json.load(open(os.path.join(GOLDEN, "c3_anchor.json")))["points"]
    assert len(points) >= 2
    with bem_amd.Context(0, lanes=1) as ctx:
        b1, b2 = ctx.body(body1()), ctx.body(body2())
        for d in points:
            eps = lorentz_eps(d["eV"])
            r = ctx.calc_heat_transfer(b1, b2, Request(eV=d["eV"], a=d["a"], T1=d["T1"], T2=d["T2"],
                                                       gap=tuple(d["shift"]), eps1=eps, eps2=eps,
                                                       precision=BEM_PREC_MIXED, flags=BEM_SOLVE_REFINE))
            ref = d.get("ReS_refined", d["ReS"])
            rel = abs(r["ReS"] - ref) / abs(ref)
            say(f"c3 mixed+refine eV {d['eV']} gap {d['gap']}: rel {rel:.2e}, {r['iters']} steps, "
                f"residual {r['residual']:.1e}, precision {r['precision']}, restarts {r['restarts']}")
            assert r["status"] == 0 and rel <= S_TOL and r["residual"] <= BEM_REFINE_TOL
            if r["precision"] == BEM_PREC_MIXED:
                assert r["iters"] >= 1 and r["restarts"] == 0
            else:
                # growth over the limit in complex64 (0.08 eV, gap 0.05: 20.5):
                # the request is re-solved in FP64, once
                assert r["restarts"] == 1
