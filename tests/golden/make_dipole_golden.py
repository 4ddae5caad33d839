"""Generate the dipole-source and flux fixture from the reference's own code.

Runs oracle/_ref/zts_ref --source (the reference's Bdense and flux compiled
from its sources by oracle/ref/Makefile) on the geometry of the
sphere1_cube1_gap0p05 fixture (icosphere + small cube, two interior media)
and stores:
  region, p, j     three dipoles: in the exterior, inside body 0, inside body 1
  B                (3, 2E) complex128: Bdense of each body's tree, body order
  X_rand           a seeded random solution vector (2E complex128)
  flux_rand        (2,) flux(top, body b, X_rand)
  flux_phys        (3, 2) flux(top, body b, X) with X = W_full^-1 B[s],
                   W_full = W + blkdiag(G1, G2) of the sphere1_cube1_gap0p05
                   fixture (the reference's Zts), solved with LAPACK (numpy)
Usage (dev container only, needs /root/reference):
    make -C oracle ref && python tests/golden/make_dipole_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from bem_amd.mesh import Mesh  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def dipoles(meshes):
    s1, cb = meshes
    cube_c = cb.xyz.mean(axis=0)
    region = np.array([-1, 0, 1], dtype=np.int32)
    p = np.array([[0.3, 1.7, 0.2],                 # outside both bodies
                  [0.1, -0.2, 0.15],               # inside the sphere
                  cube_c + [0.05, -0.1, 0.12]])    # inside the cube
    j = np.array([[0.3, -0.5, 0.8], [1.0, 0.0, 0.0], [0.2, 0.7, -0.4]])
    return region, p, j


def main():
    g = dict(np.load(os.path.join(OUT, "sphere1_cube1_gap0p05.npz"), allow_pickle=False))
    meshes = [Mesh(g[f"xyz_b{b}"], g[f"tri_b{b}"], f"b{b}") for b in range(2)]
    eps = [complex(g["eps_b0"]), complex(g["eps_b1"])]
    mu = [complex(g["mu_b0"]), complex(g["mu_b1"])]
    om = float(g["omega"])
    region, p, j = dipoles(meshes)
    M = g["W"].shape[0]
    m1 = g["G1"].shape[0]
    Wf = g["W"].copy()
    Wf[:m1, :m1] += g["G1"]
    Wf[m1:, m1:] += g["G2"]
    rng = np.random.default_rng(20261204)
    X_rand = rng.normal(size=M) + 1j * rng.normal(size=M)
    B, flux_phys = [], []
    flux_rand = None
    with tempfile.TemporaryDirectory() as d:
        for s in range(3):
            b, fr = O.run_reference_source(meshes, eps, mu, om, int(region[s]), p[s], j[s], X_rand, d)
            if flux_rand is None:
                flux_rand = np.array(fr)
            assert np.array_equal(flux_rand, fr)
            B.append(b)
            X = np.linalg.solve(Wf, b)
            _, fp = O.run_reference_source(meshes, eps, mu, om, int(region[s]), p[s], j[s], X, d)
            flux_phys.append(fp)
    rec = dict(region=region, p=p, j=j, B=np.array(B), X_rand=X_rand, flux_rand=flux_rand,
               flux_phys=np.array(flux_phys))
    path = os.path.join(OUT, "dipole_flux.npz")
    np.savez_compressed(path, **rec)
    print(f"dipole_flux: M={M} |B|max={np.abs(rec['B']).max(axis=1)} flux_rand={flux_rand} "
          f"flux_phys={rec['flux_phys'].tolist()} -> {path}")


if __name__ == "__main__":
    main()
