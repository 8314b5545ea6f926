"""Generates the tabulated rejection ceiling of LP nonlinear Compton emission at
high harmonics (csrc/dev_lma.h, namespace lp_ceil) with the library's own
sampling code, on the GPU:

    python -m ptarmigan_amd.tables.gen_lp_ceiling [out.f64]

For every node (a, eta, n) of the grid -- a = 3 ... 20 (33 nodes), eta = 1e-3
... 2 (44), n = 200 ... 40000 (110, rounded to integers), all log-spaced -- it
stores the largest of the reference's ceiling samples of the double-differential
rate at harmonic n (nonlinear_compton/lp/mod.rs:53-140, through ptg_eval
PTG_FN_LP_CEILING), and 0 where n is beyond twice the harmonic sum limit
(lp/mod.rs:411-416), where no harmonic is sampled. Then, for every cell whose
corners are tabulated, the largest sample at eight interior points (the
quarter points of each axis) against the trilinear interpolation of the
logarithms: the cell's margin is 1.02 x the largest ratio (at least 1), or 0
(not tabulated) above 1.25. The default output is generated/nlc_lp_ceiling.f64
beside this file; extract.py then packs it into ptg_tables.bin. Takes about a
minute on one MI355X.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

# the grid of csrc/dev_lma.h (lp_ceil::)
NA, NE, NN = 33, 44, 110
LN_A0, LN_A1 = math.log(3.0), math.log(20.0)
LN_E0, LN_E1 = math.log(1.0e-3), math.log(2.0)
LN_N0, LN_N1 = math.log(200.0), math.log(40000.0)


def grid():
    ln_a = LN_A0 + (LN_A1 - LN_A0) * np.arange(NA) / (NA - 1)
    ln_e = LN_E0 + (LN_E1 - LN_E0) * np.arange(NE) / (NE - 1)
    n = np.rint(np.exp(LN_N0 + (LN_N1 - LN_N0) * np.arange(NN) / (NN - 1))).astype(np.int64)
    return np.exp(ln_a), np.exp(ln_e), n


def sum_limit(a, eta):
    """Harmonics contributing to the LP rate (lp/mod.rs:411-416)."""
    m = 3.0 - 0.4 * eta ** 0.25
    return math.ceil(6.0 * (1.0 + a * a) + 2.0 * a ** m)


def main():
    from ptarmigan_amd import Engine, Fn
    out_path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "generated", "nlc_lp_ceiling.f64")
    a, eta, n = grid()
    rows, where = [], []
    for ia in range(NA):
        for ie in range(NE):
            top = 2 * sum_limit(a[ia], eta[ie])
            for k in range(NN):
                if n[k] <= top:
                    rows.append((float(n[k]), a[ia], eta[ie]))
                    where.append((ia, ie, k))
    rows = np.array(rows)
    table = np.zeros((NA, NE, NN))
    with Engine(0) as eng:
        def largest(r):
            return np.concatenate([eng.eval(Fn.LP_CEILING, r[i:i + 65536], 2)[:, 0] for i in range(0, len(r), 65536)])
        for (ia, ie, k), v in zip(where, largest(rows)):
            table[ia, ie, k] = v if np.isfinite(v) and v > 0.0 else 0.0
        margin = cell_margins(table, largest)
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    np.concatenate([table.ravel(), margin.ravel()]).astype("<f8").tofile(out_path)
    used = margin > 0.0
    print(f"wrote {out_path}: {len(rows)} of {table.size} nodes tabulated, {int((table > 0).sum())} positive; "
          f"{int(used.sum())} cells, margins {np.quantile(margin[used], [0.5, 0.99]).round(4)} (median, 99 %), "
          f"{int(((margin == 0.0) & corners_ok(table)).sum())} cells above 1.25 left to the direct ceiling")


def corners_ok(table):
    ok = table > 0.0
    c = np.ones((NA - 1, NE - 1, NN - 1), dtype=bool)
    for xa in (0, 1):
        for xe in (0, 1):
            for xn in (0, 1):
                c &= ok[xa:NA - 1 + xa, xe:NE - 1 + xe, xn:NN - 1 + xn]
    return c


def cell_margins(table, largest):
    """1.02 x the largest ratio of the sample to the interpolated value at the
    cells' eight quarter points (at least 1; 0 above 1.25 or where a corner is
    not tabulated)."""
    ok = corners_ok(table)
    ia, ie, k = np.nonzero(ok)
    lt = np.log(np.where(table > 0.0, table, 1.0))
    sa = (LN_A1 - LN_A0) / (NA - 1)
    se = (LN_E1 - LN_E0) / (NE - 1)
    sn = (LN_N1 - LN_N0) / (NN - 1)
    ratio = np.zeros(len(ia))
    for qa in (0.25, 0.75):
        for qe in (0.25, 0.75):
            for qn in (0.25, 0.75):
                a = np.exp(LN_A0 + (ia + qa) * sa)
                eta = np.exp(LN_E0 + (ie + qe) * se)
                n = np.rint(np.exp(LN_N0 + (k + qn) * sn))
                # where the device's interpolation puts the integer n
                fa, fe, fn = (np.log(a) - LN_A0) / sa - ia, (np.log(eta) - LN_E0) / se - ie, (np.log(n) - LN_N0) / sn - k
                acc = np.zeros(len(ia))
                for xa in (0, 1):
                    for xe in (0, 1):
                        for xn in (0, 1):
                            w = (fa if xa else 1 - fa) * (fe if xe else 1 - fe) * (fn if xn else 1 - fn)
                            acc += w * lt[ia + xa, ie + xe, k + xn]
                f = largest(np.stack([n, a, eta], 1))
                r = np.where(np.isfinite(f) & (f > 0.0), f / np.exp(acc), np.inf)
                ratio = np.maximum(ratio, r)
    mu = np.maximum(1.02 * ratio, 1.0)
    mu[~(mu <= 1.25)] = 0.0
    margin = np.zeros((NA - 1, NE - 1, NN - 1))
    margin[ia, ie, k] = mu
    return margin


if __name__ == "__main__":
    main()
