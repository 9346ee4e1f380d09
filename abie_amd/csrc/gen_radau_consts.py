"""Emit radau_consts.h: the Gauss-Radau tables of the GR15 integrator.

The nodes are the roots of P_7(2x-1) + P_8(2x-1) on [0, 1] (Gauss-Radau with
the left end point); rs[j][k] = 1/(h_j - h_k); cs[j][k] is the coefficient of
x^k in prod_{i=1..j}(x - h_i).  Values are computed in 60-digit arithmetic,
rounded once to double and written as hexadecimal literals.  These are the
tables the reference types as decimals (src/integrator_gauss_radau15.h:24-54);
the test-suite checks them bit for bit.
"""
import os

import mpmath as mp

mp.mp.dps = 60


def nodes():
    f = lambda t: mp.legendre(7, 2 * t - 1) + mp.legendre(8, 2 * t - 1)
    # the polynomial has degree 8: its Taylor coefficients at 0 are exact
    coeffs = mp.taylor(f, 0, 8)[::-1]
    roots = mp.polyroots(coeffs, maxsteps=200, extraprec=200)
    found = sorted(mp.re(r) for r in roots)
    found[0] = mp.mpf(0)
    assert len(found) == 8 and all(0 <= r <= 1 for r in found), found
    return found


def main():
    h = nodes()
    rs = [[(1 / (h[j] - h[k]) if k < j else mp.mpf(0)) for k in range(8)] for j in range(8)]
    cs = [[mp.mpf(0)] * 8 for _ in range(8)]
    poly = [mp.mpf(1)]
    cs[0][0] = mp.mpf(1)
    for j in range(1, 8):
        nxt = [mp.mpf(0)] * (len(poly) + 1)
        for k, c in enumerate(poly):
            nxt[k + 1] += c
            nxt[k] -= h[j] * c
        poly = nxt
        for k, c in enumerate(poly):
            cs[j][k] = c
    lit = lambda v: "0.0" if v == 0 else float(v).hex()
    lines = ["// Generated by gen_radau_consts.py (Gauss-Radau definition); do not edit.",
             "#pragma once", ""]
    lines.append("#define ABIE_RADAU_H {" + ", ".join(lit(v) for v in h) + "}")
    for name, tab in (("RS", rs), ("CS", cs)):
        rows = ["{" + ", ".join(lit(v) for v in row) + "}" for row in tab]
        lines.append("#define ABIE_RADAU_%s {%s}" % (name, ", \\\n    ".join(rows)))
    lines.append("")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radau_consts.h")
    with open(out, "w") as fh:
        fh.write("\n".join(lines))


if __name__ == "__main__":
    main()
