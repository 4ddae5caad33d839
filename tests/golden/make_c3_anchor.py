"""Anchor value of the sphere-over-plate configuration (M = 33360) from the
oracle: one (gap, frequency) point, computed on the CPU with the C
restatement's assembly and LAPACK, added to the points of c3_anchor.json.

Composition as CalcHeatTransfer (src/HeatTransfer.cpp:8-59): W = Zts(top) with
the interior blocks added, X = W^-1[0:m1, m1:M] (the only block the two
products use), B1 = sym(G1) X, B2 = X sym(G2), S = tr(B2^H B1) / 4.  X comes
from an LU solve of W against [0; I] rather than the full inverse (same
block, a third of the memory).  Needs about 45 GB of host memory and an hour
or more on 8 cores.

    python tests/golden/make_c3_anchor.py [eV] [gap]
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


def main():
    eV = float(sys.argv[1]) if len(sys.argv) > 1 else 0.112
    gap = float(sys.argv[2]) if len(sys.argv) > 2 else 0.05
    a, T1, T2 = 1e-6, 300.0, 0.0
    sph = mesh.icosphere(4).translated((0.0, 0.0, 1.0 + gap))
    plate = mesh.cuboid(4.0, 4.0, 0.5, 30, 30, 10, origin=(-2.0, -2.0, -0.5))
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
    say("G2 done; LU of W")
    lu, piv = sla.lu_factor(W, overwrite_a=True, check_finite=False)
    del W
    E = np.zeros((m1 + m2, m2), dtype=np.complex128, order="F")
    E[m1:, :] = np.eye(m2)
    say("solve")
    Z = sla.lu_solve((lu, piv), E, overwrite_b=True, check_finite=False)
    del lu, E
    X = np.asfortranarray(Z[:m1, :])
    del Z
    say("products")
    B1 = S1 @ X
    del S1
    B2 = X @ S2
    del S2, X
    S = 0.25 * np.vdot(B2, B1)
    th1, th2 = O.theta(omega, T1), O.theta(omega, T2)
    out = {"workload": "icosphere level 4 (5120 tri, radius 1) over a 4 x 4 x 0.5 plate (6000 tri), "
                       "sphere centre at z = 1 + gap, plate top at z = 0",
           "eV": eV, "gap": gap, "a": a, "T1": T1, "T2": T2, "eps": [eps.real, eps.imag], "M": m1 + m2,
           "ReS": S.real, "ImS": S.imag, "P": S.real * (th1 - th2), "seconds": time.time() - t0}
    path = os.path.join(HERE, "c3_anchor.json")
    points = json.load(open(path))["points"] if os.path.exists(path) else []
    points = [p for p in points if (p["eV"], p["gap"]) != (eV, gap)] + [out]
    json.dump({"points": points}, open(path, "w"), indent=1)
    say("wrote", path, out["ReS"])


if __name__ == "__main__":
    main()
