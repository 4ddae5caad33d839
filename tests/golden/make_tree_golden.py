"""Generate the RWGTree order fixture from the reference's own code.

Runs oracle/_ref/zts_ref --tree (the reference's ShapeMesh and
RWGTree(meshDiel, edgesAtPoint, minSegmentSize = 32, minRadius = 0), compiled
from its sources by oracle/ref/Makefile) on the bodies of the workloads, each
translated as the workload places it, and stores per case:
  <case>_order   int32: Edge::index of each RWG of the body in tree order
  <case>_leaves  int32: RWG counts of the tree's leaves, in that order
  <case>_nseg    segments the reference's mesher keeps for the body
CASES below rebuilds each body from bem_amd.mesh.
Usage (dev container only, needs /root/reference):
    make -C oracle ref && python tests/golden/make_tree_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from bem_amd import mesh  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def flat_octahedron(level, hz):
    """Octahedron of half-height hz subdivided level times (midpoints, not
    projected): for small hz its equator is a closed loop of edges sharper
    than 2 rad, so the reference's mesher keeps only the half holding
    triangle 0."""
    xyz = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, hz], [0, 0, -hz]]
    tri = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]]
    for _ in range(level):
        mid, nxt = {}, []

        def middle(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                xyz.append([0.5 * (xyz[a][c] + xyz[b][c]) for c in range(3)])
                mid[k] = len(xyz) - 1
            return mid[k]
        for a, b, c in tri:
            ab, bc, ca = middle(a, b), middle(b, c), middle(c, a)
            nxt += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        tri = nxt
    return mesh.Mesh(np.array(xyz, dtype=np.float64), np.array(tri, dtype=np.int32), "flat_octahedron")


# name: (body, translation)
CASES = {
    # C2: 1280-triangle spheres, body 1 at centre gap 2.2
    "c2_body1": (lambda: mesh.icosphere(3), (2.2, 0.0, 0.0)),
    "c2_body2": (lambda: mesh.icosphere(3), (0.0, 0.0, 0.0)),
    # C3: 5120-triangle sphere 0.05 over the 6000-triangle plate
    "c3_sphere": (lambda: mesh.icosphere(4), (0.0, 0.0, 1.05)),
    "c3_plate": (lambda: mesh.cuboid(4.0, 4.0, 0.5, 30, 30, 10, origin=(-2.0, -2.0, -0.5)), (0.0, 0.0, 0.0)),
    # C5: cone tip over a plate (its base rim is sharper than 2 rad)
    "tip_cone": (lambda: mesh.cone_tip(apex_radius=0.05, n_around=50, n_cap=6, n_along=19, n_base=6),
                 (0.0, 0.0, 0.05)),
    "tip_plate": (lambda: mesh.cuboid(4.0, 4.0, 0.5, 20, 20, 2, origin=(-2.0, -2.0, -0.5)), (0.0, 0.0, 0.0)),
    "cone_default": (lambda: mesh.cone_tip(), (0.0, 0.0, 0.0)),
    # C4 body kinds
    "c4_cube7": (lambda: mesh.cube(7), (1.33, 0.0, 0.0)),
    "c4_cuboid": (lambda: mesh.cuboid(1.0, 1.0, 0.5, 12, 12, 6), (0.0, 0.0, 0.0)),
    "ico2": (lambda: mesh.icosphere(2), (0.0, 0.0, 0.0)),
    "cube2": (lambda: mesh.cube(2), (0.0, 0.0, 0.0)),
    # a closed loop of sharp edges other than the cone's base rim
    "flat_octahedron": (lambda: flat_octahedron(3, 0.2), (0.0, 0.0, 0.0)),
}
# bodies whose reference tree holds only part of the edges
SHARP = {"tip_cone", "cone_default", "flat_octahedron"}


def body(name):
    make, shift = CASES[name]
    return make().translated(shift)


def main():
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name in CASES:
            (nseg, order, leaves), = O.run_reference_tree([body(name)], d)
            out[f"{name}_order"] = order
            out[f"{name}_leaves"] = leaves
            out[f"{name}_nseg"] = np.int32(nseg)
            print(f"{name}: {len(order)} RWGs in {len(leaves)} leaves, {nseg} segment(s)", flush=True)
    np.savez_compressed(os.path.join(OUT, "rwg_tree.npz"), **out)


if __name__ == "__main__":
    main()
