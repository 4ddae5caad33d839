"""The reference's unknown order (bem_rwg_tree), on the CPU.

bem_rwg_tree restates the RWGTree that the reference's drivers build
(src/RWGTree.cpp:602-761 through parseShapes, src/HeatTransfer.cpp:98-105):
the mesher's first segment (src/ShapeMesh.cpp:930-970, :1917-1927), then the
recursive split along Powell-chosen planes (src/RWGTree.cpp:21-559).  Its
order and leaf sizes must equal the reference's exactly:
  - against tests/golden/rwg_tree.npz, written by the reference's own
    RWGTree compiled from its sources (tests/golden/make_tree_golden.py), for
    every workload body at its workload position;
  - against that reference build run live on the reference's example X3D
    meshes, where the reference tree and the build are present.
"""
import ctypes
import glob
import os
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLDEN)
from make_tree_golden import CASES, SHARP, body  # noqa: E402

EXAMPLES = "/root/reference/examples"
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "zts_ref")


@pytest.fixture(scope="module")
def fixture():
    return np.load(os.path.join(GOLDEN, "rwg_tree.npz"))


@pytest.mark.parametrize("case", sorted(CASES))
def test_order_matches_reference(fixture, case):
    from bem_amd import rwg_tree
    m = body(case)
    order, leaves = rwg_tree(m.xyz, m.tri)
    assert np.array_equal(order, fixture[f"{case}_order"])
    assert np.array_equal(leaves, fixture[f"{case}_leaves"])


@pytest.mark.parametrize("case", sorted(CASES))
def test_order_invariants(case):
    """Distinct edge indices; leaves ascending inside and summing to the
    count; every edge covered unless the surface has a closed loop of edges
    sharper than 2 rad (the cone's base rim, the flattened octahedron's
    equator)."""
    from bem_amd import rwg_tree
    m = body(case)
    order, leaves = rwg_tree(m.xyz, m.tri)
    ne = 3 * m.nt // 2
    assert len(np.unique(order)) == len(order) and order.min() >= 0 and order.max() < ne
    assert leaves.sum() == len(order) and leaves.min() > 0
    for seg in np.split(order, np.cumsum(leaves)[:-1]):
        assert np.all(np.diff(seg) > 0)
    if case in SHARP:
        assert len(order) < ne
    else:
        assert len(order) == ne


def test_order_depends_on_position(fixture):
    """The sphere at the origin and at x = 2.2 split differently (the planes'
    comparisons round differently); each matches the reference at its own
    position."""
    from bem_amd import rwg_tree, mesh
    s = mesh.icosphere(3)
    o1, _ = rwg_tree(s.xyz, s.tri, shift=(2.2, 0.0, 0.0))
    o2, _ = rwg_tree(s.xyz, s.tri)
    assert np.array_equal(o1, fixture["c2_body1_order"])
    assert np.array_equal(o2, fixture["c2_body2_order"])
    assert not np.array_equal(o1, o2)


def test_open_surface_rejected():
    from bem_amd import rwg_tree, mesh
    from bem_amd.runtime import BemError, BEM_E_MESH
    s = mesh.icosphere(2)
    with pytest.raises(BemError) as e:
        rwg_tree(s.xyz, s.tri[:-1])
    assert e.value.code == BEM_E_MESH


def _examples():
    if not os.path.isdir(EXAMPLES):
        return []
    return sorted(os.path.basename(f) for f in glob.glob(os.path.join(EXAMPLES, "*.x3d")))


@pytest.mark.parametrize("name", _examples() or ["(reference examples absent)"])
def test_examples_match_reference_live(name):
    """The reference's example meshes (read by the X3D reader of the compat
    layer) through bem_rwg_tree and through the reference's own RWGTree."""
    if not os.path.exists(REF_DRIVER) or not os.path.isdir(EXAMPLES):
        pytest.skip("reference build or examples absent")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import test_x3d as X
    from bem_amd import rwg_tree
    from bem_amd.mesh import Mesh
    from oracle import oracle as O
    L = ctypes.CDLL(X.LIB)
    L.bemc_x3d_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    rc, xyz, tri, _ = X.read(L, os.path.join(EXAMPLES, name))
    if rc or len(xyz) < 4:
        pytest.skip(f"{name}: no single closed body")
    if 2 * (3 * len(tri) // 2) != 3 * len(tri) or name in ("plane.x3d", "tip.x3d"):
        pytest.skip(f"{name}: open surface (the reference's mesher needs every side shared)")
    m = Mesh(xyz, tri)
    order, leaves = rwg_tree(m.xyz, m.tri)
    with tempfile.TemporaryDirectory() as d:
        (_, ro, rl), = O.run_reference_tree([m], d)
    assert np.array_equal(order, ro) and np.array_equal(leaves, rl)
