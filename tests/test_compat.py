"""The reference-shaped C++ layer (bem_amd/compat): a transfer-style driver
written against RWGTree / Options / FrequencySweep / CalcHeatTransfer
(compare src/transfer.cpp:3-29), linked against libbem_hip.

Tolerance: Re S through the driver within 1e-8 relative of the oracle (the
driver prints P; P / (Theta(T1) - Theta(T2)) recovers Re S)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT = os.path.join(ROOT, "bem_amd", "compat")


def test_compat_library_exports_reference_api():
    """libbem_compat.so defines the reference's entry points for the path."""
    lib = os.path.join(ROOT, "bem_amd", "libbem_compat.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for sym in ["Zts(RWGTree*, RWGTree*, GreensFunction*, Options&)",
                "Zdense(RWGTree*, RWGTree*, ComplexDenseMatrix*, GreensFunction*, Options&)",
                "CalcHeatTransfer(double, double, double, double, RWGTree*, RWGTree*, RWGTree*, Options&)",
                "FrequencySweep::FrequencySweep(char*)",
                "parseShapes(RWGTree*&, RWGTree*&, RWGTree*&, char const*, char const*, Vector3 const&, Options&)",
                "Dielectrics::getDielectric(std::__cxx11::basic_string<char, std::char_traits<char>, "
                "std::allocator<char> > const&, Params const&)"]:
        assert sym in out, sym
    assert os.access(os.path.join(COMPAT, "transfer_sweep"), os.X_OK)
    assert os.access(os.path.join(COMPAT, "transfer_farm"), os.X_OK)
    assert os.access(os.path.join(COMPAT, "transfer"), os.X_OK)
    assert "initReferenceCountingPointerLock()" in out


REF_TRANSFER = "/root/reference/src/transfer.cpp"


@pytest.mark.skipif(not os.path.exists(REF_TRANSFER), reason="the reference tree is not mounted")
def test_reference_transfer_builds_unchanged(tmp_path):
    """The reference's own single-frequency driver (src/transfer.cpp), its
    text unchanged, compiles against the drop-in headers of
    bem_amd/compat/include (HeatTransfer.h, Impedance.h, FrequencySweep.h,
    ReferenceCountingPointer.h) and links against libbem_compat.so and
    libbem_hip.so.  The text goes to the compiler on stdin from a scratch
    directory, so the reference's own headers next to it are not picked up;
    the binary stays in that directory.  Without arguments it prints the
    reference's usage line and exits with its status (-2)."""
    exe = str(tmp_path / "transfer_ref")
    src = open(REF_TRANSFER, "rb").read()
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(COMPAT, "include"), "-x", "c++", "-",
           "-x", "none", "-L", os.path.join(ROOT, "bem_amd"), "-lbem_compat", "-lbem_hip",
           "-Wl,-rpath," + os.path.join(ROOT, "bem_amd"), "-o", exe]
    subprocess.run(cmd, input=src, cwd=str(tmp_path), check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 254 and r.stderr.startswith("Usage: transfer")


def test_transfer_usage():
    """transfer keeps the reference's command line (src/transfer.cpp:7-10)."""
    out = subprocess.run([os.path.join(COMPAT, "transfer")], capture_output=True, text=True)
    assert out.returncode == 2 and out.stderr.startswith("Usage: transfer <x3d file for body 1>")


def _write_mesh(path, m):
    with open(path, "w") as f:
        f.write(f"{m.nv} {m.nt}\n")
        for x in m.xyz:
            f.write(f"{x[0]:.17g} {x[1]:.17g} {x[2]:.17g}\n")
        for t in m.tri:
            f.write(f"{t[0]} {t[1]} {t[2]}\n")


@pytest.mark.gpu
def test_transfer_driver_matches_oracle(tmp_path):
    from bem_amd import mesh, lorentz_eps
    from oracle import oracle as O
    s = mesh.icosphere(1)
    c = mesh.cube(2)
    _write_mesh(tmp_path / "b1.mesh", s)
    _write_mesh(tmp_path / "b2.mesh", c)
    freqs = [0.095, 0.11, 0.125]
    (tmp_path / "freqs.txt").write_text("\n".join(str(f) for f in freqs) + "\n")
    gap = (0.0, 0.0, 2.1)
    print("running transfer_sweep", flush=True)
    out = subprocess.run([os.path.join(COMPAT, "transfer_sweep"), str(tmp_path / "b1.mesh"), str(tmp_path / "b2.mesh"),
                          str(tmp_path / "freqs.txt"), "1e-6", "300", "0", *[str(g) for g in gap]],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [l.split() for l in out.stdout.strip().splitlines()]
    assert len(lines) == len(freqs)
    for f, l in zip(freqs, lines):
        eV, P = float(l[4]), float(l[5])
        assert abs(eV - f) < 1e-15
        eps = lorentz_eps(f)
        ref = O.spectral_power(s.translated(gap), c, f, 1e-6, 300.0, 0.0, eps, eps, return_parts=True)
        ReS = P / (ref["theta1"] - ref["theta2"])
        assert abs(ReS - ref["ReS"]) <= 1e-8 * abs(ref["ReS"]), (f, ReS, ref["ReS"])


def _write_x3d(path, m, material, translation=(0.0, 0.0, 0.0)):
    """A shape file as Blender's exporter writes them: one Transform, one
    Shape with its <Dielectric>, comma-separated faces and points."""
    faces = ", ".join(f"{t[0]} {t[1]} {t[2]} -1" for t in m.tri)
    pts = ", ".join(" ".join(repr(float(v)) for v in x) for x in m.xyz)
    path.write_text(f"""<?xml version="1.0" encoding="UTF-8"?>
<X3D version="3.0"><Scene>
 <Transform translation="{translation[0]} {translation[1]} {translation[2]}" rotation="1 0 0 0">
  <Shape><Dielectric material="{material}"/>
   <IndexedFaceSet coordIndex="{faces}, "><Coordinate point="{pts}, "/></IndexedFaceSet>
  </Shape>
 </Transform>
</Scene></X3D>
""")


@pytest.mark.gpu
def test_transfer_farm_matches_oracle(tmp_path):
    """transfer_farm (the reference's farm inputs and ./output format,
    src/transfer-mpi.cpp) on X3D bodies with library media: every line
    within 1e-8 of the oracle, and a restart runs only the jobs missing from
    ./output."""
    import ctypes
    from bem_amd import mesh
    from oracle import oracle as O
    s, c = mesh.icosphere(1), mesh.cube(2)
    _write_x3d(tmp_path / "body1.x3d", s, "SiC")
    _write_x3d(tmp_path / "body2.x3d", c, "SiN")
    (tmp_path / "a.txt").write_text("1e-6\n")
    (tmp_path / "gaps.txt").write_text("0 0 2.1\n")
    (tmp_path / "freqs.txt").write_text("0.095\n0.11\n")
    (tmp_path / "temps.txt").write_text("300 0\n")
    farm = os.path.join(COMPAT, "transfer_farm")
    print("running transfer_farm", flush=True)
    out = subprocess.run([farm], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = (tmp_path / "output").read_text().splitlines()
    assert len(lines) == 2
    lib = ctypes.CDLL(os.path.join(ROOT, "bem_amd", "libbem_compat.so"))
    lib.bemc_dielectric.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double)]

    def eps(name, f, T):
        o = (ctypes.c_double * 4)()
        assert lib.bemc_dielectric(name.encode(), None, f, T, o) == 0
        return complex(o[0], o[1])

    got = {}
    for l in lines:
        a, f, T1, T2, g, P = (float(v) for v in l.split())
        assert (a, T1, T2, g) == (1e-6, 300.0, 0.0, 0.0)
        got[f] = P
        ref = O.spectral_power(s.translated((0.0, 0.0, 2.1)), c, f, 1e-6, 300.0, 0.0, eps("SiC", f, 300.0),
                               eps("SiN", f, 0.0))
        assert abs(P - ref) <= 1e-8 * abs(ref), (f, P, ref)
    assert sorted(got) == [0.095, 0.11]
    # restart: nothing left to run; then one line removed is run again
    out = subprocess.run([farm], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "0 jobs" in out.stderr, out.stderr
    assert (tmp_path / "output").read_text().splitlines() == lines
    (tmp_path / "output").write_text(lines[0] + "\n")
    out = subprocess.run([farm], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "1 jobs" in out.stderr, out.stderr
    again = (tmp_path / "output").read_text().splitlines()
    assert again[0] == lines[0] and again[1] == lines[1]
    # the single-frequency driver on the same files gives the same P
    out = subprocess.run([os.path.join(COMPAT, "transfer"), "body1.x3d", "body2.x3d", "0.11", "1e-6", "300", "0",
                          "0", "0", "2.1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    eV, P = out.stdout.split()
    assert float(eV) == 0.11 and float(P) == got[0.11]


def test_transfer_farm_refuses_ambiguous_restart(tmp_path):
    """./output keeps only a gap's x component: with distinct gaps of equal x in
    gaps.txt a restart could skip jobs that never ran, so transfer_farm
    refuses it (before any device work) and says why."""
    from bem_amd import mesh
    farm = os.path.join(COMPAT, "transfer_farm")
    if not os.path.exists(farm):
        pytest.skip("compat drivers not built")
    s, c = mesh.icosphere(1), mesh.cube(2)
    _write_x3d(tmp_path / "body1.x3d", s, "SiC")
    _write_x3d(tmp_path / "body2.x3d", c, "SiN")
    (tmp_path / "a.txt").write_text("1e-6\n")
    (tmp_path / "gaps.txt").write_text("0 0 2.1\n0 0.5 2.1\n")
    (tmp_path / "freqs.txt").write_text("0.095\n")
    (tmp_path / "temps.txt").write_text("300 0\n")
    (tmp_path / "output").write_text("1e-06 0.095 300 0 0 1.0e-24\n")
    out = subprocess.run([farm], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 2
    assert "same x" in out.stderr and "move ./output away" in out.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["spheres", "cone"])
def test_compat_zts_in_reference_tree_order(gpu_ctx, kind):
    """Compat Zts lists its unknowns in the reference's tree order: the
    library's mesh-order W with rows and columns taken in the order that
    bem_rwg_tree gives (itself equal to the reference's RWGTree,
    tests/test_rwgtree.py), J and M of each RWG kept together, bit for bit.
    The cone's base rim is sharper than 2 rad: the reference's tree would
    hold only part of its edges, so the order is flagged as not the
    reference's and the rest follow in mesh order."""
    import ctypes
    from bem_amd import mesh, rwg_tree
    if kind == "spheres":
        b1, b2, shift = mesh.icosphere(2), mesh.icosphere(2), (2.2, 0.0, 0.0)
    else:
        b1, b2, shift = mesh.cone_tip(n_around=12, n_along=6, n_cap=2, n_base=2), mesh.cube(3), (0.0, 0.0, 2.6)
    lib = ctypes.CDLL(os.path.join(ROOT, "bem_amd", "libbem_compat.so"))
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    lib.bemc_tree_zts.argtypes = [vp, i32, vp, i32, vp, vp, i32, vp, i32, ctypes.c_double, vp, vp,
                                  ctypes.POINTER(i32)]
    ne1, ne2 = 3 * b1.nt // 2, 3 * b2.nt // 2
    E = ne1 + ne2
    omega = 0.5
    Z = np.zeros((2 * E, 2 * E), dtype=np.complex128, order="F")
    order = np.zeros(E, dtype=np.int32)
    ref = i32(0)
    sh = np.array(shift, dtype=np.float64)
    print(f"compat tree Zts: {kind}, E = {E}", flush=True)
    rc = lib.bemc_tree_zts(b1.xyz.ctypes.data, b1.nv, b1.tri.ctypes.data, b1.nt, sh.ctypes.data,
                           b2.xyz.ctypes.data, b2.nv, b2.tri.ctypes.data, b2.nt, omega, Z.ctypes.data,
                           order.ctypes.data, ctypes.byref(ref))
    assert rc == 0
    o1, _ = rwg_tree(b1.xyz, b1.tri, shift=shift)
    o2, _ = rwg_tree(b2.xyz, b2.tri)
    if kind == "spheres":
        assert ref.value == 1
        assert np.array_equal(order, np.concatenate([o1, o2 + ne1]))
    else:
        assert ref.value == 0 and len(o1) < ne1
        rest = np.setdiff1d(np.arange(ne1), o1)
        assert np.array_equal(order, np.concatenate([o1, rest, o2 + ne1]))
    W = gpu_ctx.zts([gpu_ctx.body(b1), gpu_ctx.body(b2)], [shift, (0.0, 0.0, 0.0)], 1.0, 1.0, omega)
    idx = np.empty(2 * E, dtype=np.int64)
    idx[0::2], idx[1::2] = 2 * order, 2 * order + 1
    assert np.array_equal(Z, W[np.ix_(idx, idx)])
