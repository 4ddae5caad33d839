"""Regenerate sin_table.h, the SiN optical constants of the material library.

The photon energies (eV) and the refractive index n + i k of silicon nitride
are the measured data of the reference's src/util/SiN.cpp (ftab :4-1119,
ntab :1120-2235, ktab :2236-3351; 1114 rows).  eps(eV) = (n + i k)^2 with n, k
linearly interpolated in that table, so the same rows are needed for the same
permittivity; they are carried over as data, in this project's own layout.

Run in a tree where the reference is mounted at /root/reference:
    python bem_amd/compat/gen_sin_table.py
"""
import os
import re

SRC = "/root/reference/src/util/SiN.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sin_table.h")
NUM = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"


def table(text, name):
    m = re.search(r"SiN::" + name + r"\[SiN_TABLE_SIZE\]\s*=\s*\{(.*?)\}", text, re.S)
    return [float(v) for v in re.findall(NUM, m.group(1))]


def main():
    text = open(SRC).read()
    f, n, k = (table(text, t) for t in ("ftab", "ntab", "ktab"))
    assert len(f) == len(n) == len(k) == 1114, (len(f), len(n), len(k))
    assert all(b > a for a, b in zip(f, f[1:])), "energies must increase"
    with open(OUT, "w") as o:
        o.write("// Generated by gen_sin_table.py: SiN photon energy (eV), n, k per row,\n"
                "// measured data of the reference's src/util/SiN.cpp.\n#pragma once\n\n")
        o.write("static const int kSiNRows = %d;\n" % len(f))
        o.write("static const double kSiN[%d][3] = {\n" % len(f))
        for a, b, c in zip(f, n, k):
            o.write("    {%r, %r, %r},\n" % (a, b, c))
        o.write("};\n")
    print("wrote", OUT, len(f), "rows")


if __name__ == "__main__":
    main()
