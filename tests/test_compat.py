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
                "FrequencySweep::FrequencySweep(char*)"]:
        assert sym in out, sym
    assert os.access(os.path.join(COMPAT, "transfer_sweep"), os.X_OK)


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
