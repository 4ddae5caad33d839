"""Host-side checks: mesh generators, the C-ABI library exports every symbol
declared in include/bem_hip.h (no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from bem_amd import mesh


def closed_manifold(m):
    e = np.sort(np.concatenate([m.tri[:, [0, 1]], m.tri[:, [1, 2]], m.tri[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    return np.all(cnt == 2)


@pytest.mark.parametrize("m,nv,nt", [(mesh.cube(2), 26, 48), (mesh.icosphere(3), 642, 1280),
                                    (mesh.icosphere(4), 2562, 5120),
                                    (mesh.cuboid(4, 4, 0.5, 30, 30, 10), None, 6000)])
def test_generators(m, nv, nt):
    assert m.nt == nt
    if nv is not None:
        assert m.nv == nv
    assert closed_manifold(m)
    # outward winding: signed volume > 0
    p = m.xyz[m.tri]
    vol = np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2])).sum() / 6
    assert vol > 0


def test_library_exports_declared_symbols():
    from bem_amd.runtime import EXPORTS, lib_path
    hdr = open(os.path.join(ROOT, "include", "bem_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*|size_t)\s*\*?\s*(bem_\w+)\s*\(", hdr, re.M))
    assert declared == set(EXPORTS)
    lib = ctypes.CDLL(lib_path())
    for name in declared:
        assert hasattr(lib, name), name


def test_oracle_library_loads():
    from oracle import oracle as O
    assert O.lib().oracle_num_threads() >= 1
