"""The Chebyshev expansions of the reference's Gauss-Laguerre Airy sums
(tables/gen_airy_cheb.py, special_functions/airy:CHEB_GL_SUMS in the table
blob), which the device evaluates for Ai and Ai' at x >= 1 in place of the
reference's 4-40 node sums (special_functions/airy.rs:41-116).

The table is held here, without a GPU, against the oracle's restatement of the
reference's sums: the expansion at the segment the device picks, times the
reference's prefactor, is Ai(x) and Ai'(x) of the oracle to a few units in the
last place over x in [1, 100] (dense, and on both sides of every node-count
switch and segment edge). A mistake in a segment's interval, its ordering or
the node set it was fitted to fails by orders of magnitude more.
"""
import numpy as np
import pytest


def _table():
    from ptarmigan_amd import tables
    return tables.load()["special_functions/airy:CHEB_GL_SUMS"]


def _clenshaw(rec, t):
    u = (t - rec[1]) * rec[2]
    b1 = b2 = 0.0
    for k in range(15, 0, -1):
        b1, b2 = 2.0 * u * b1 - b2 + rec[3 + k], b1
    return u * b1 - b2 + rec[3]


def _segment(C, x, t, first):
    """Segment for x: `first` lists each region's first segment and its upper x;
    a region's second segment, if any, starts at its t_lo."""
    for seg, x_hi in first:
        if x_hi is None or x < x_hi:
            if seg + 1 < len(C) and seg + 1 not in [s for s, _ in first] and t >= C[seg + 1][0]:
                return seg + 1
            return seg
    raise AssertionError


def _airy_from_table(C, x):
    # the reference's prefactors (airy.rs:51-53,108-110), as the oracle forms them,
    # so that only the sums differ
    s = 2.0 * x ** 1.5 / 3.0
    t = 1.0 / s
    ai_seg = _segment(C, x, t, [(0, 2.0), (2, 10.0), (4, None)])
    ap_seg = _segment(C, x, t, [(5, 2.0), (7, 4.0), (8, 11.0), (9, None)])
    e = np.exp(-s)
    ai = 0.262183997088323 * s ** (-1.0 / 6.0) * e * _clenshaw(C[ai_seg], t)
    aip = -0.2898380090915846 * s ** (1.0 / 6.0) * e * _clenshaw(C[ap_seg], t)
    return ai, aip


def test_table_layout():
    C = _table()
    assert C.shape == (10, 19)
    # segments of one region are contiguous in t and cover it; t = 1/s falls as x grows
    t_of = lambda x: 1.5 / x ** 1.5
    for lo_seg, x_lo, x_hi, n in [(0, 1, 2, 2), (2, 2, 10, 2), (4, 10, None, 1), (5, 1, 2, 2), (7, 2, 4, 1),
                                  (8, 4, 11, 1), (9, 11, None, 1)]:
        t_lo = 0.0 if x_hi is None else t_of(x_hi)
        assert C[lo_seg][0] == pytest.approx(t_lo, rel=1e-15, abs=1e-300)
        top = C[lo_seg + n - 1][1] + 1.0 / C[lo_seg + n - 1][2]
        assert top == pytest.approx(t_of(x_lo), rel=1e-14)


def test_expansions_reproduce_the_reference_sums(orc):
    C = _table()
    # up to x = 100 (s = 667): beyond ~115 e^(-s) leaves the normal range
    x = list(np.exp(np.linspace(0.0, np.log(100.0), 3000)))
    for edge in (2.0, 4.0, 10.0, 11.0):
        x += [np.nextafter(edge, 0.0), edge, np.nextafter(edge, 10 * edge)]
    for seg in (1, 3, 6):  # inner segment edges, in x
        x_edge = (1.5 / C[seg][0]) ** (2.0 / 3.0)
        x += [x_edge * (1 - 1e-12), x_edge, x_edge * (1 + 1e-12)]
    worst = [0.0, 0.0]
    for v in x:
        ref = orc.airy(v)[0]
        got = _airy_from_table(C, v)
        for c in (0, 1):
            worst[c] = max(worst[c], abs(got[c] / ref[c] - 1.0))
    assert worst[0] < 5e-15 and worst[1] < 5e-15, worst
