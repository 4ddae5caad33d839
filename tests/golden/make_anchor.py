"""Anchor values of full-size configurations from the oracle, one (gap,
frequency) point per run, computed on the CPU with the C restatement's
assembly and LAPACK and added to the points of <workload>_anchor.json:
  c3   level-4 icosphere over a 4 x 4 x 0.5 plate of 6000 triangles (M = 33360),
       sphere centre at z = 1 + gap;
  tip  the cone tip of the serving mix (apex radius 0.05, 3000 triangles) over
       a 4 x 4 x 0.5 plate of 1920 triangles (M = 14760), apex at z = gap.

Composition as CalcHeatTransfer (src/HeatTransfer.cpp:8-59): W = Zts(top) with
the interior blocks added, X = W^-1[0:m1, m1:M] (the only block the two
products use), B1 = sym(G1) X, B2 = X sym(G2), S = tr(B2^H B1) / 4.  X comes
from an LU solve of W against [0; I] rather than the full inverse (same
block, a third of the memory).  c3 needs about 45 GB of host memory and
20 minutes on 8 cores, tip a few minutes.

With --refine the point also carries the value after two steps of FP64
iterative refinement of [X; Z2] = W^-1 [0; I] against the unfactored W
(ReS_refined), LAPACK's reciprocal 1-norm condition estimate of W (rcond,
zgecon) and the relative change of Re S per step, so that the accuracy of
the plain LU value itself is known where W is ill-conditioned.

    python tests/golden/make_anchor.py <c3|tip> <eV> <gap> [--refine]
"""
import json
import os
import sys
import time

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O          # noqa: E402
from bem_amd import mesh, lorentz_eps   # noqa: E402


WORKLOADS = {
    "c3": ("icosphere level 4 (5120 tri, radius 1) over a 4 x 4 x 0.5 plate (6000 tri), "
           "sphere centre at z = 1 + gap, plate top at z = 0",
           lambda: mesh.icosphere(4), lambda: mesh.cuboid(4.0, 4.0, 0.5, 30, 30, 10, origin=(-2.0, -2.0, -0.5)),
           lambda gap: 1.0 + gap),
    "tip": ("cone tip (apex radius 0.05, 3000 tri) over a 4 x 4 x 0.5 plate (1920 tri), apex at z = gap",
            lambda: mesh.cone_tip(apex_radius=0.05, n_around=50, n_cap=6, n_along=19, n_base=6),
            lambda: mesh.cuboid(4.0, 4.0, 0.5, 20, 20, 2, origin=(-2.0, -2.0, -0.5)),
            lambda gap: gap),
}


def main():
    args = [x for x in sys.argv[1:] if not x.startswith("--")]
    refine = "--refine" in sys.argv
    name, eV, gap = args[0], float(args[1]), float(args[2])
    desc, body1, body2, shift = WORKLOADS[name]
    a, T1, T2 = 1e-6, 300.0, 0.0
    sph = body1().translated((0.0, 0.0, shift(gap)))
    plate = body2()
    eps = lorentz_eps(eV)
    omega = O.omega_tilde(eV, a)
    t0 = time.time()
    say = lambda *m: print(f"[{time.time() - t0:7.0f} s]", *m, flush=True)
    top, b1, b2 = O.Tree([sph, plate]), O.Tree([sph]), O.Tree([plate])
    m1, m2 = 2 * b1.ne, 2 * b2.ne
    say(f"M = {m1 + m2} (m1 {m1}, m2 {m2}); assembling W")
    W = top.zts(1.0, 1.0, omega)
    say("W done; G1")
    G = b1.zts(eps, 1.0, omega)
    W[:m1, :m1] += G
    S1 = G + G.conj().T
    del G
    say("G1 done; G2")
    G = b2.zts(eps, 1.0, omega)
    W[m1:, m1:] += G
    S2 = G + G.conj().T
    del G
    W0 = np.array(W, order="F", copy=True) if refine else None
    anorm = np.abs(W).sum(axis=0).max() if refine else 0.0
    say("G2 done; LU of W")
    lu, piv = sla.lu_factor(W, overwrite_a=True, check_finite=False)
    del W
    rcond = None
    if refine:
        rcond = float(sla.lapack.zgecon(lu, anorm, norm="1")[0])
        say(f"rcond (1-norm estimate) {rcond:.3e}")
    E = np.zeros((m1 + m2, m2), dtype=np.complex128, order="F")
    E[m1:, :] = np.eye(m2)
    say("solve")
    Z = sla.lu_solve((lu, piv), E, overwrite_b=False, check_finite=False)

    def trace_of(Z):
        X = np.asfortranarray(Z[:m1, :])
        B1 = S1 @ X
        B2 = X @ S2
        return 0.25 * np.vdot(B2, B1)

    say("products")
    S = trace_of(Z)
    th1, th2 = O.theta(omega, T1), O.theta(omega, T2)
    out = {"workload": desc, "eV": eV, "gap": gap, "shift": [0.0, 0.0, shift(gap)], "a": a, "T1": T1, "T2": T2, "eps": [eps.real, eps.imag], "M": m1 + m2,
           "ReS": S.real, "ImS": S.imag, "P": S.real * (th1 - th2)}
    if refine:
        steps, Sk = [], S
        for it in range(2):
            R = E - W0 @ Z
            rel_res = float(np.linalg.norm(R) / np.linalg.norm(E))
            Z = Z + sla.lu_solve((lu, piv), R, check_finite=False)
            Sn = trace_of(Z)
            steps.append({"residual_before": rel_res, "rel_change_ReS": abs(Sn.real - Sk.real) / abs(Sn.real)})
            say(f"refinement {it + 1}: ||R||/||E|| before {rel_res:.3e}, ReS {Sn.real:.15e}, "
                f"change {steps[-1]['rel_change_ReS']:.3e}")
            Sk = Sn
        out.update({"ReS_refined": Sk.real, "rcond": rcond, "refinement": steps,
                    "lu_vs_refined": abs(S.real - Sk.real) / abs(Sk.real)})
    out["seconds"] = time.time() - t0
    path = os.path.join(HERE, f"{name}_anchor.json")
    points = json.load(open(path))["points"] if os.path.exists(path) else []
    points = [p for p in points if (p["eV"], p["gap"]) != (eV, gap)] + [out]
    json.dump({"points": points}, open(path, "w"), indent=1)
    say("wrote", path, out["ReS"])


if __name__ == "__main__":
    main()
