"""Regenerate quadrature_tables.h from the reference's rule tables.

The node/weight constants of the triangle and line rules are the numerical
data of src/TriangleQuadrature.h (triangle rules :31-540, line rule
QuadratureGauss15 :620-656).  Parity of the impedance entries with the
reference at 1e-10 depends on using exactly these printed constants (several
rules are not symmetric and are printed to ~13 digits), so the values are
carried over as data; the C layout below is this project's own.

Run in a tree where the reference is mounted at /root/reference:
    python bem_amd/csrc/gen_quadrature.py
"""
import os
import re
import sys

SRC = "/root/reference/src/TriangleQuadrature.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quadrature_tables.h")
TRI_RULES = [6, 10, 15, 21, 36, 91, 105]
NUM = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"


def parse_triangle_rules(text):
    rules = {}
    blocks = re.split(r"class TriangleQuadratureGauss(\d+)\b", text)
    for k in range(1, len(blocks), 2):
        npts = int(blocks[k])
        body = blocks[k + 1].split("class ", 1)[0]
        pts = []
        for line in body.splitlines():
            m = re.search(r"v = v0([^;]*);\s*integrand\.weighPoint\(\s*(" + NUM + r")\s*,v\)", line)
            if not m:
                continue
            expr, w = m.group(1), float(m.group(2))
            a = b = 0.0
            for term in re.findall(r"\+\s*(?:(" + NUM + r")\s*\*\s*)?(v1|v2)", expr):
                coef = float(term[0]) if term[0] else 1.0
                if term[1] == "v1":
                    a = coef
                else:
                    b = coef
            pts.append((a, b, w))
        if len(pts) != npts:
            raise SystemExit(f"rule {npts}: parsed {len(pts)} points")
        rules[npts] = pts
    return rules


def parse_line_rule15(text):
    body = text.split("class QuadratureGauss15", 1)[1]
    ts = [float(x) for x in re.findall(r"v0 \+\s*(" + NUM + r")\s*\* v1", body)]
    ws = [float(x) for x in re.findall(r"weighPoint\(\s*(" + NUM + r")\s*,v\)", body)]
    if len(ts) != 15 or len(ws) != 15:
        raise SystemExit("line rule 15 parse failed")
    return list(zip(ts, ws))


def main():
    text = open(SRC).read()
    rules = parse_triangle_rules(text)
    line15 = parse_line_rule15(text)
    out = []
    out.append("// Generated by gen_quadrature.py -- do not edit.\n")
    out.append("// Triangle rules: point = v0 + a*(v1-v0) + b*(v2-v0), weights sum to 1.\n")
    out.append("// Constants are the reference's rule data (src/TriangleQuadrature.h:20-540,\n")
    out.append("// line rule src/TriangleQuadrature.h:620-656, whose weights sum to 2).\n")
    out.append("#pragma once\n\n")
    offs = 0
    out.append("#define BEM_QUAD_TOTAL_POINTS %d\n" % sum(len(rules[n]) for n in TRI_RULES))
    for n in TRI_RULES:
        out.append("#define BEM_QUAD_OFF_%d %d\n" % (n, offs))
        offs += n
    out.append("\n#define BEM_QUAD_ABW_INIT { \\\n")
    for n in TRI_RULES:
        out.append("  /* %d-point rule */ \\\n" % n)
        for a, b, w in rules[n]:
            out.append("  {%r, %r, %r}, \\\n" % (a, b, w))
    out.append("}\n\n")
    out.append("#define BEM_LINE15_INIT { \\\n")
    for t, w in line15:
        out.append("  {%r, %r}, \\\n" % (t, w))
    out.append("}\n\n")
    out.append("#ifndef BEM_QUAD_NO_HOST_TABLES\n")
    out.append("static const double bem_quad_abw[BEM_QUAD_TOTAL_POINTS][3] = BEM_QUAD_ABW_INIT;\n")
    out.append("static const double bem_line15_tw[15][2] = BEM_LINE15_INIT;\n")
    out.append("#endif\n")
    with open(OUT, "w") as f:
        f.write("".join(out))
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
