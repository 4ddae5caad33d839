import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_golden(name):
    import numpy as np
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def golden_meshes(g):
    from bem_amd.mesh import Mesh
    out = []
    b = 0
    while f"xyz_b{b}" in g:
        out.append(Mesh(g[f"xyz_b{b}"], g[f"tri_b{b}"], f"b{b}"))
        b += 1
    return out


@pytest.fixture(scope="session")
def gpu_ctx():
    from bem_amd import Context
    ctx = Context(0, lookahead=-1)
    yield ctx
    ctx.close()
