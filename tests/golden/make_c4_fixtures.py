"""Oracle spectral power of seven instances of the seeded 256-instance batch
(bench.batch_instances(256), seed 0): the smallest (a cuboid pair, M = 1728),
the two largest (cube pair M = 10368, cuboid pair M = 9924), and for each body
kind one instance of middling size, plus the largest icosphere pair.  The
oracle is the C restatement of the reference assembly (pinned to the
reference's own Zts by the golden fixtures) with the CalcHeatTransfer
composition in numpy/LAPACK (src/HeatTransfer.cpp:8-59).

    python tests/golden/make_c4_fixtures.py      (a few minutes on 8 cores)
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import bench                      # noqa: E402
from oracle import oracle as O    # noqa: E402

PICK = [59, 33, 92, 66, 80, 147]


def main():
    inst = bench.batch_instances(256)
    sph = [i for i, (k, (m1, m2), _, _) in enumerate(inst) if k == 2 and m1.nt == 1280 and m2.nt == 1280]
    picks = PICK + sph[:1]
    out = {}
    for i in picks:
        kind, (m1, m2), gap, f = inst[i]
        t0 = time.time()
        eps = O.lorentz_eps(f)
        r = O.spectral_power(m1.translated(gap), m2, f, 1e-6, 300.0, 0.0, eps, eps, return_parts=True)
        M = 2 * (O.mesh_edges(m1.tri, m1.nv).shape[0] + O.mesh_edges(m2.tri, m2.nv).shape[0])
        out[str(i)] = {"kind": kind, "M": M, "eV": f, "gap": list(gap), "ReS": float(r["ReS"]),
                       "P": float(r["P"]), "seconds": time.time() - t0}
        print(i, out[str(i)], flush=True)
    with open(os.path.join(HERE, "c4_oracle.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
