"""X3D shape reader and material library of the compat layer (host code, no
device): bem_amd/compat/x3d.cpp and materials.cpp through their C entry
points in libbem_compat.so.

Pinned by: scipy's rotations (an independent implementation) for the
Transform stack, with the reference's convention of src/util/Rotation.cpp
(its matrix is the transpose of the right-handed rotation, so a rotation by
phi about a unit axis is scipy's rotation by -phi); the reference's own
example files where /root/reference is present (counts, closed meshes,
<Dielectric>); closed forms and limits of each medium (static limits,
passivity Im eps > 0, table nodes)."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bem_amd", "libbem_compat.so")
EXAMPLES = "/root/reference/examples"
SIN_SRC = "/root/reference/src/util/SiN.cpp"


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libbem_compat.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.bemc_x3d_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    L.bemc_dielectric.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_double)]
    return L


def read(lib, path):
    nv, nt = ctypes.c_int(0), ctypes.c_int(0)
    mat = ctypes.create_string_buffer(64)
    rc = lib.bemc_x3d_read(str(path).encode(), None, ctypes.byref(nv), None, ctypes.byref(nt), mat, 64)
    if rc:
        return rc, None, None, None
    xyz = np.zeros((max(nv.value, 1), 3))
    tri = np.zeros((max(nt.value, 1), 3), dtype=np.int32)
    rc = lib.bemc_x3d_read(str(path).encode(), xyz.ctypes.data, ctypes.byref(nv), tri.ctypes.data,
                           ctypes.byref(nt), mat, 64)
    assert rc == 0
    return 0, xyz[:nv.value], tri[:nt.value], mat.value.decode()


def medium(lib, name, eV, T=300.0, params=None):
    out = (ctypes.c_double * 4)()
    rc = lib.bemc_dielectric(name.encode(), params.encode() if params else None, eV, T, out)
    assert rc == 0
    return complex(out[0], out[1]), complex(out[2], out[3])


def ref_rot(axis, phi):
    """The reference's rotation by phi about a unit axis."""
    return Rotation.from_rotvec(-phi * np.asarray(axis, float)).as_matrix()


DOC = """<?xml version="1.0" encoding="UTF-8"?>
<!DOCTYPE X3D PUBLIC "ISO//Web3D//DTD X3D 3.0//EN" "http://www.web3d.org/specifications/x3d-3.0.dtd">
<!-- a comment with <tags> inside -->
<X3D version='3.0'>
 <Scene>
  {open}
   <Shape>
    <Dielectric material="SiC"/>
    <IndexedFaceSet coordIndex="0 1 2 -1, 0 2 3 -1,
                                0 3 1 -1, 1 3 2 -1, ">
     <Coordinate point="{points}" />
    </IndexedFaceSet>
   </Shape>
  {close}
 </Scene>
</X3D>
"""
TET = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def fmt(p):
    return ", ".join(" ".join(repr(float(c)) for c in q) for q in p) + ", "


def test_plain_document(lib, tmp_path):
    f = tmp_path / "t.x3d"
    f.write_text(DOC.format(open="", close="", points=fmt(TET)))
    rc, xyz, tri, mat = read(lib, f)
    assert rc == 0 and mat == "SiC"
    assert np.array_equal(xyz, TET)
    assert tri.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 1], [1, 3, 2]]


@pytest.mark.parametrize("axis,phi", [((1, 0, 0), np.pi / 2), ((-1, 0, 0), 1.570796), ((0, 0.6, 0.8), 0.7)])
def test_rotation_convention(lib, tmp_path, axis, phi):
    f = tmp_path / "r.x3d"
    f.write_text(DOC.format(open=f'<Transform rotation="{axis[0]} {axis[1]} {axis[2]} {phi!r}">',
                            close="</Transform>", points=fmt(TET)))
    rc, xyz, _, _ = read(lib, f)
    assert rc == 0
    assert np.abs(xyz - TET @ ref_rot(axis, phi).T).max() < 1e-15


def test_full_transform(lib, tmp_path):
    """P' = T + C + R SR S SR^-1 (P - C) (the X3D order, src/util/Transform.cpp:53-70)."""
    C, T, S = np.array([0.3, -0.2, 0.5]), np.array([1.5, 2.0, -3.0]), np.array([2.0, 0.5, 1.25])
    ax_r, ph_r, ax_s, ph_s = (0.0, 0.0, 1.0), 0.4, (0.6, 0.0, 0.8), -1.1
    attrs = (f'center="{C[0]} {C[1]} {C[2]}" translation="{T[0]} {T[1]} {T[2]}" scale="{S[0]} {S[1]} {S[2]}" '
             f'rotation="{ax_r[0]} {ax_r[1]} {ax_r[2]} {ph_r}" scaleOrientation="{ax_s[0]} {ax_s[1]} {ax_s[2]} {ph_s}"')
    f = tmp_path / "f.x3d"
    f.write_text(DOC.format(open=f"<Transform {attrs}>", close="</Transform>", points=fmt(TET)))
    rc, xyz, _, _ = read(lib, f)
    assert rc == 0
    R, SR, SRi = ref_rot(ax_r, ph_r), ref_rot(ax_s, ph_s), ref_rot(ax_s, -ph_s)
    want = np.array([T + C + R @ (SR @ (S * (SRi @ (p - C)))) for p in TET])
    assert np.abs(xyz - want).max() < 1e-14


def test_nested_transforms_outermost_first(lib, tmp_path):
    """The reference maps a point through the stack from the outermost
    transform to the innermost (src/util/Transform.cpp:38-47)."""
    f = tmp_path / "n.x3d"
    f.write_text(DOC.format(open='<Transform translation="1 0 0"><Group><Transform rotation="0 0 1 1.5707963267948966">',
                            close="</Transform></Group></Transform>", points=fmt(TET)))
    rc, xyz, _, _ = read(lib, f)
    assert rc == 0
    Rz = ref_rot((0, 0, 1), np.pi / 2)
    want = (TET + [1.0, 0.0, 0.0]) @ Rz.T
    assert np.abs(xyz - want).max() < 1e-15
    # and the stack is popped: a sibling after the nested transforms is untouched
    g = tmp_path / "s.x3d"
    g.write_text(DOC.format(open='<Transform translation="5 5 5"></Transform>', close="", points=fmt(TET)))
    rc, xyz2, _, _ = read(lib, g)
    assert rc == 0 and np.array_equal(xyz2, TET)


def test_number_groups_as_the_reference_reads_them(lib, tmp_path):
    """A face group is read from its start onward (src/FaceIndex.cpp:4-13): a
    group missing its -1 takes its fourth number from the next group; groups
    with too few numbers (the trailing one) are skipped; tabs and line breaks
    in attribute values are spaces; entities are decoded."""
    doc = ('<X3D><Shape><Dielectric material="Const" eps="2.5+0.1"/>'
           '<IndexedFaceSet coordIndex="0 1 2, 3\t4 5 -1,\n 1 2 3 -1, 7">'
           '<Coordinate point="0 0 0, 1 0 0,0 1 0,\n0 0 1, 1 1 &#49;, 2 2"/></IndexedFaceSet></Shape></X3D>')
    f = tmp_path / "g.x3d"
    f.write_text(doc)
    rc, xyz, tri, mat = read(lib, f)
    assert rc == 0 and mat == "Const"
    assert tri.tolist() == [[0, 1, 2], [3, 4, 5], [1, 2, 3]]
    assert xyz.tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]]


@pytest.mark.parametrize("doc", ["<X3D><Shape></X3D>", "<X3D><Coordinate point='1 2 3'></X3D>",
                                 "<X3D><Transform rotation='1 0'/></X3D>", "<X3D a=1/>", "<X3D>&bogus;</X3D><"])
def test_malformed(lib, tmp_path, doc):
    f = tmp_path / "m.x3d"
    f.write_text(doc)
    rc, *_ = read(lib, f)
    assert rc == -1   # BEM_E_INVALID


def test_missing_file(lib, tmp_path):
    rc, *_ = read(lib, tmp_path / "none.x3d")
    assert rc == -1


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference examples not mounted")
def test_reference_examples():
    """Every example shape of the reference: indices within the points, every
    edge shared by two triangles (closed surfaces), and the media of the two
    farm bodies."""
    L = ctypes.CDLL(LIB)
    L.bemc_x3d_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    seen = 0
    for name in sorted(os.listdir(EXAMPLES)):
        if not name.endswith(".x3d"):
            continue
        rc, xyz, tri, mat = read(L, os.path.join(EXAMPLES, name))
        assert rc == 0, name
        text = open(os.path.join(EXAMPLES, name)).read()
        if "," not in re.search(r'point="([^"]*)"', text).group(1).strip().rstrip(","):
            # points without separating commas (a later Blender exporter): the
            # reference's reader takes the first point only, and so does this one
            assert len(xyz) == 1, name
            continue
        groups = re.search(r'coordIndex="([^"]*)"', text).group(1).split(",")
        assert len(tri) == sum(1 for g in groups if len(g.split()) >= 4), name
        assert tri.min() >= 0 and tri.max() < len(xyz), name
        e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        if name in ("body1.x3d", "body2.x3d", "cube1.x3d", "sphere.x3d"):
            assert (cnt == 2).all(), name
        if name in ("body1.x3d", "body2.x3d"):
            assert mat == "SiN"
        seen += 1
    assert seen >= 15


def test_sic_limits(lib):
    """SiC (src/util/SiC.cpp): static limit epsinf wL^2/wT^2, eps(wL) ~ 0
    (the longitudinal optical phonon), Im eps > 0 across the band."""
    hbar, eV = 1.05457148e-34, 1.60217646e-19
    e0, _ = medium(lib, "SiC", 1e-9)
    assert abs(e0 - 6.7 * (182.7e12 / 149.5e12) ** 2) < 1e-6
    eL, _ = medium(lib, "SiC", 182.7e12 * hbar / eV)
    assert abs(eL.real) < 1e-2 and abs(eL) < 0.1
    for f in np.linspace(0.08, 0.14, 13):
        assert medium(lib, "SiC", f)[0].imag > 0
    # temperature moves the resonance: at 600 K wT is lower
    hot, _ = medium(lib, "SiC", 1e-9, T=600.0)
    assert abs(hot - 6.7 * np.exp(5e-5 * 300) * ((182.7e12 - 5.463e9 * 300) / (149.5e12 - 4.106e9 * 300)) ** 2) < 1e-6


def test_metals_and_constants(lib):
    e, m = medium(lib, "Al", 1e-9)
    assert abs(e - (2.2281 + 1.30601)) < 1e-6 and m == 1
    for f in (0.05, 0.1, 1.0):
        ea, eu = medium(lib, "Al", f)[0], medium(lib, "Au", f)[0]
        assert ea.imag > 0 and eu.imag > 0 and eu.real < 0   # Drude metal below the plasma frequency
    assert abs(medium(lib, "Au", 1e4)[0] - 1.0) < 1e-3            # far above it
    assert medium(lib, "Const", 0.1, params="eps=2.5+0.1")[0] == complex(2.5, 0.1)
    assert medium(lib, "Const", 0.1, params="eps=2.5-0.1;mu=1.5")[0:2] == (complex(2.5, -0.1), complex(1.5, 0))
    assert medium(lib, "Vacuum", 0.1) == (1, 1)
    assert medium(lib, "SiliconDoped", 0.1) == (11.7, 1)
    # doped silicon: Drude term lowers Re eps and adds loss
    ed = medium(lib, "SiliconDoped", 0.1, params="N=-1e25")[0]
    assert ed.real < 11.7 and ed.imag > 0


def test_sin_table(lib):
    """SiN: (n + i k)^2 at the table's nodes, linear in n and k between."""
    import importlib.util
    rows = []
    for line in open(os.path.join(ROOT, "bem_amd", "compat", "sin_table.h")):
        m = re.match(r"\s*\{([^}]*)\},", line)
        if m:
            rows.append([float(v) for v in m.group(1).split(",")])
    rows = np.array(rows)
    assert rows.shape == (1114, 3) and (np.diff(rows[:, 0]) > 0).all()
    for i in (0, 100, 500, 1112):
        f, n, k = rows[i]
        assert abs(medium(lib, "SiN", f)[0] - complex(n, k) ** 2) < 1e-12 * abs(complex(n, k) ** 2)
        f2, n2, k2 = rows[i + 1]
        x = 0.3
        nn, kk = n + x * (n2 - n), k + x * (k2 - k)
        assert abs(medium(lib, "SiN", f + x * (f2 - f))[0] - complex(nn, kk) ** 2) < 1e-9
    assert abs(medium(lib, "SiN", rows[-1, 0])[0] - complex(rows[-1, 1], rows[-1, 2]) ** 2) < 1e-12
    if os.path.exists(SIN_SRC):   # the generated table holds the reference's rows
        text = open(SIN_SRC).read()
        num = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"
        for col, name in enumerate(("ftab", "ntab", "ktab")):
            body = re.search(r"SiN::" + name + r"\[SiN_TABLE_SIZE\]\s*=\s*\{(.*?)\}", text, re.S).group(1)
            assert np.array_equal(np.array([float(v) for v in re.findall(num, body)]), rows[:, col])
