"""Generates the Chebyshev expansions through which the device evaluates the
reference's Gauss-Laguerre sums for Ai(x) and Ai'(x) at x >= 1
(special_functions/airy.rs:41-116):

    python -m ptarmigan_amd.tables.gen_airy_cheb [out.f64]

For x >= 1 the reference writes Ai(x) = c s^(-1/6) e^(-s) F(t) and
Ai'(x) = -c' s^(1/6) e^(-s) G(t), s = 2 x^(3/2) / 3, t = 1/s, where F and G are
sums over the nodes x_k and weights w_k of a generalised Gauss-Laguerre rule,
F(t) = sum_k w_k (2 + x_k t)^(-1/6) and G(t) = sum_k w_k (2 + x_k t)^(1/6),
with 40, 16 or 4 nodes for Ai (x < 2, x < 10, x >= 10) and 32, 16, 8 or 4 for
Ai' (x < 2, x < 4, x < 11, x >= 11). Each of these sums is a smooth function of
t on its own interval of t, which is split into one or two segments; on each
segment the sum (evaluated from the shipped nodes and weights in 40-digit
arithmetic) is expanded in Chebyshev polynomials of degree 15 of
u = (t - c) / h. The expansions' truncation error is below 1e-17 relative
(checked here at 41 points per segment), so the device's Clenshaw sum gives
the reference's sum to its rounding, at 16 coefficients instead of 4-40 pairs
of logarithm and exponential.

Output: 10 segments of 19 doubles, [t_lo, c, 1/h, a_0 .. a_15], Ai's first
(x in [1, 2): 2 segments, [2, 10): 2, [10, inf): 1), then Ai''s ([1, 2): 2,
[2, 4): 1, [4, 11): 1, [11, inf): 1), each region's segments in increasing t.
The default output is generated/airy_gl_cheb.f64 beside this file; extract.py
packs it into ptg_tables.bin.
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEGREE = 15
REC = DEGREE + 4
# (node table, sign of the exponent 1/6, x range, segments), in the output's order
REGIONS = [
    ("GL_m1_6_NODES_40", "GL_m1_6_WEIGHTS_40", -1, 1.0, 2.0, 2),
    ("GL_m1_6_NODES_16", "GL_m1_6_WEIGHTS_16", -1, 2.0, 10.0, 2),
    ("GL_m1_6_NODES_4", "GL_m1_6_WEIGHTS_4", -1, 10.0, None, 1),
    ("GL_p1_6_NODES_32", "GL_p1_6_WEIGHTS_32", 1, 1.0, 2.0, 2),
    ("GL_p1_6_NODES_16", "GL_p1_6_WEIGHTS_16", 1, 2.0, 4.0, 1),
    ("GL_p1_6_NODES_8", "GL_p1_6_WEIGHTS_8", 1, 4.0, 11.0, 1),
    ("GL_p1_6_NODES_4", "GL_p1_6_WEIGHTS_4", 1, 11.0, None, 1),
]


def gl_sum(nodes, weights, sign, t):
    p = mp.mpf(sign) / 6
    return mp.fsum(mp.mpf(float(w)) * (2 + mp.mpf(float(x)) * t) ** p for x, w in zip(nodes, weights))


def t_of(x):
    return 3 / (2 * mp.mpf(x) ** mp.mpf(1.5))


def segments(tables):
    """[(t_lo, c, h, coefficients, worst relative truncation error)] in output order."""
    mp.mp.dps = 40
    out = []
    n = DEGREE + 1
    for kn, kw, sign, x_lo, x_hi, nseg in REGIONS:
        nodes = tables["special_functions/airy:" + kn]
        weights = tables["special_functions/airy:" + kw]
        t_hi = t_of(x_lo)
        t_lo = mp.mpf(0) if x_hi is None else t_of(x_hi)
        for j in range(nseg):
            a = t_lo + (t_hi - t_lo) * j / nseg
            b = t_lo + (t_hi - t_lo) * (j + 1) / nseg
            c, h = (a + b) / 2, (b - a) / 2
            us = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n) for k in range(n)]
            fs = [gl_sum(nodes, weights, sign, c + h * u) for u in us]
            coef = [2 * mp.fsum(fs[k] * mp.cos(mp.pi * m * (k + mp.mpf(1) / 2) / n) for k in range(n)) / n
                    for m in range(n)]
            coef[0] /= 2
            worst = mp.mpf(0)
            for u in np.linspace(-1.0, 1.0, 41):
                u = mp.mpf(u)
                val = mp.fsum(coef[m] * mp.chebyt(m, u) for m in range(n))
                worst = max(worst, abs(val / gl_sum(nodes, weights, sign, c + h * u) - 1))
            out.append((a, c, h, coef, worst))
    return out


def main():
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from ptarmigan_amd import tables
    out_path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "generated", "airy_gl_cheb.f64")
    segs = segments(tables.load())
    rows = []
    for a, c, h, coef, worst in segs:
        if worst > 1e-17:
            raise SystemExit(f"segment at t = {float(a):.6g}: truncation error {float(worst):.3g}")
        rows.append([float(a), float(c), float(1 / h)] + [float(v) for v in coef])
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    np.array(rows, dtype="<f8").tofile(out_path)
    print(f"wrote {out_path}: {len(rows)} segments x {REC} doubles; largest truncation error "
          f"{max(float(s[4]) for s in segs):.3g}")


if __name__ == "__main__":
    main()
