"""Generate the golden impedance fixtures from the reference's own Zts.

Runs oracle/_ref/zts_ref (the reference's assembly compiled from its sources
by oracle/ref/Makefile) on small geometries and stores, per case:
  edges_b{0,1}   the reference's ShapeMesh edge tables {p1, p2, t1, t2}
  W              Zts(top, top, exterior medium)    (2E x 2E complex128)
  G1, G2         Zts(body, body, interior medium)  per body
  ReS            Re of 1/4 tr(B2^H B1) composed from W, G1, G2 with LAPACK
                 (numpy), as src/HeatTransfer.cpp:30-55 composes it
  xyz_b*, tri_b* the input meshes, omega, eps_b*, mu_b*
Usage (dev container only, needs /root/reference):
    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from bem_amd import mesh  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cases():
    c2 = mesh.cube(2)
    yield "c1_cubes_0p11", [c2.translated((1.1, 0, 0)), c2], 0.11, 1e-6, None
    yield "c1_cubes_0p09", [c2.translated((1.1, 0, 0)), c2], 0.09, 1e-6, None
    # icosphere + small cube at a 0.05 gap: cross-body pairs in every far tier,
    # two different interior media
    s1 = mesh.icosphere(1)
    xmax = s1.xyz[:, 0].max()
    cb = mesh.cube(1, edge=1.0, origin=(xmax + 0.05, -0.5, -0.5))
    yield "sphere1_cube1_gap0p05", [s1, cb], 0.10, 1e-6, [O.lorentz_eps(0.10), 3.0 + 0.4j]


def main():
    for name, meshes, eV, a, eps_list in cases():
        omega = O.omega_tilde(eV, a)
        if eps_list is None:
            eps_list = [O.lorentz_eps(eV)] * 2
        mu_list = [1.0, 1.0]
        with tempfile.TemporaryDirectory() as d:
            edges, W, Gs = O.run_reference(meshes, eps_list, mu_list, omega, d)
        S = O.heat_transfer_from_blocks(W, Gs[0], Gs[1])
        top = O.Tree(meshes)
        rec = dict(W=W, G1=Gs[0], G2=Gs[1], ReS=np.float64(S.real), ImS=np.float64(S.imag),
                   omega=np.float64(omega), eV=np.float64(eV), a=np.float64(a),
                   hist=top.pair_histogram())
        for b, m in enumerate(meshes):
            rec[f"xyz_b{b}"] = m.xyz
            rec[f"tri_b{b}"] = m.tri
            rec[f"edges_b{b}"] = edges[b]
            rec[f"eps_b{b}"] = np.complex128(eps_list[b])
            rec[f"mu_b{b}"] = np.complex128(mu_list[b])
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **rec)
        print(f"{name}: M={W.shape[0]} ReS={S.real:.16e} hist={rec['hist'].tolist()} -> {path}")


if __name__ == "__main__":
    main()
