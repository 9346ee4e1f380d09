"""Derive the Gauss-Radau spacing tables of the GR15 integrator with mpmath.

TEST INFRASTRUCTURE (oracle side). Writes ``oracle/radau_tables.h``, the
tables the CPU restatement in ``abie_oracle.c`` uses.

The reference types the tables as decimal literals
(``src/integrator_gauss_radau15.h:24-54``).  Here they are derived from their
definition instead of being copied:

* ``hs``: the 8 roots of ``P_7(2x-1) + P_8(2x-1)`` on [0, 1] (Gauss-Radau
  nodes with the left end point, ``hs[0] = 0``);
* ``rs[j][k] = 1 / (hs[j] - hs[k])`` for ``k < j``, 0 elsewhere;
* ``cs[j][k]`` = coefficient of ``x**k`` in ``prod_{i=1..j} (x - hs[i])``.

Each value is rounded once, from 60 significant digits, to the nearest double
and emitted as a C99 hexadecimal literal, so the header is exact.  The test
suite checks every entry bit-for-bit against the reference's compiled tables.
"""
import os
import sys

import mpmath as mp

mp.mp.dps = 60


def radau_nodes():
    x = mp.mpf
    poly = lambda t: mp.legendre(7, 2 * t - 1) + mp.legendre(8, 2 * t - 1)
    # bracket the 8 sign changes on a fine grid, then polish with findroot
    grid = [x(i) / 4000 for i in range(-1, 4001)]
    roots = []
    for a, b in zip(grid[:-1], grid[1:]):
        fa, fb = poly(a), poly(b)
        if fa == 0:
            roots.append(a)
        elif fa * fb < 0:
            roots.append(mp.findroot(poly, (a, b), solver="anderson"))
    roots = sorted(set(mp.nstr(r, 50) for r in roots), key=lambda s: mp.mpf(s))
    roots = [mp.mpf(r) for r in roots]
    # t = 0 is a root exactly
    roots[0] = mp.mpf(0)
    assert len(roots) == 8, roots
    return roots


def tables():
    hs = radau_nodes()
    rs = [[mp.mpf(0)] * 8 for _ in range(8)]
    for j in range(8):
        for k in range(j):
            rs[j][k] = 1 / (hs[j] - hs[k])
    cs = [[mp.mpf(0)] * 8 for _ in range(8)]
    # coefficients of prod_{i=1..j} (x - hs[i]), row j, power k
    coeffs = [mp.mpf(1)]
    cs[0][0] = mp.mpf(1)
    for j in range(1, 8):
        new = [mp.mpf(0)] * (len(coeffs) + 1)
        for k, c in enumerate(coeffs):
            new[k + 1] += c
            new[k] -= hs[j] * c
        coeffs = new
        for k, c in enumerate(coeffs):
            cs[j][k] = c
    return hs, rs, cs


def to_double(v):
    return float(mp.mpf(v))


def emit(path):
    hs, rs, cs = tables()
    out = ["/* Generated by oracle/gen_radau_tables.py from the Gauss-Radau",
           " * definition (see the script).  Do not edit. */",
           "#ifndef ORACLE_RADAU_TABLES_H", "#define ORACLE_RADAU_TABLES_H", ""]
    out.append("static const double ORC_HS[8] = {")
    out.append(",\n".join("    %s" % to_double(v).hex() for v in hs))
    out.append("};")
    for name, tab in (("ORC_RS", rs), ("ORC_CS", cs)):
        out.append("static const double %s[8][8] = {" % name)
        rows = []
        for row in tab:
            rows.append("    {" + ", ".join(to_double(v).hex() for v in row) + "}")
        out.append(",\n".join(rows))
        out.append("};")
    out += ["", "#endif", ""]
    text = "\n".join(out).replace("0x0.0p+0", "0.0")
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    emit(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "radau_tables.h"))
