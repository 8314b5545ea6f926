"""Regenerates the mid-range CDF table of LP nonlinear Breit-Wheeler pair creation
(5652 nodes x 16 harmonics x (n, par, perp)), which the reference defines by its
generator (src/pair_creation/lp/mod.rs:1065-1163) but does not ship.

    python -m ptarmigan_amd.tables.gen_bw_tables [threads]

Builds gen_bw_tables.hip (host code; the rate functions are the library's own,
csrc/dev_bw.h), runs it over the reference's grid -- eta' = 2 / (1 + i/4) for
i < 157 and a = 10^(-7/5 + j/20) for j < 36, eta = (1 + a^2/2) eta' -- and writes
generated/nbw_lp_mid_cdf.f64. The last column of every row is the total rate,
which the reference does ship (mid_range/total.rs): the run fails unless the two
agree. Then run extract.py to rebuild ptg_tables.bin. Takes a few minutes on 8 cores.
"""
import math
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    from ptarmigan_amd import tables
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    t = tables.load()
    exe = os.path.join(tempfile.gettempdir(), "ptg_gen_bw_tables")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-pthread",
                    os.path.join(HERE, "gen_bw_tables.hip"), "-o", exe], check=True)
    mt = "pair_creation/lp/tables/mid_range/total:"
    ln_min_a = t[mt + "LN_MIN_A"][0]
    ln_max_ep = t[mt + "LN_MAX_ETA_PRIME"][0]
    density = t[mt + "ETA_PRIME_DENSITY"][0]
    n_rows, n_cols = 157, 36
    nodes = []
    for i in range(n_rows):
        ep = math.exp(ln_max_ep) / (1.0 + i / density)
        for j in range(n_cols):
            a = math.exp(ln_min_a) * 10.0 ** (j / 20.0)
            nodes.append((a, (1.0 + 0.5 * a * a) * ep))
    q = "quadrature/tables:"
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<Q", len(nodes)))
            for name in ("GAUSS_16_NODES", "GAUSS_16_WEIGHTS", "GAUSS_32_NODES", "GAUSS_32_WEIGHTS"):
                f.write(np.ascontiguousarray(t[q + name], dtype="<f8").tobytes())
            f.write(np.asarray(nodes, dtype="<f8").tobytes())
        subprocess.run([exe, fin, fout, str(threads)], check=True)
        table = np.fromfile(fout, dtype="<f8").reshape(n_rows * n_cols, 16, 3)
    total = t[mt + "TABLE"].reshape(n_rows * n_cols, 2)
    got = np.log(table[:, 15, 1:])
    ok = np.isfinite(total)
    err = np.abs(got[ok] - total[ok])
    print(f"ln(total) vs the shipped mid-range total table: max |diff| {err.max():.3e}, "
          f"median {np.median(err):.3e} over {ok.sum()} values")
    if not (err.max() < 1e-6):
        raise SystemExit("generated totals disagree with the shipped table")
    os.makedirs(os.path.join(HERE, "generated"), exist_ok=True)
    table.astype("<f8").tofile(os.path.join(HERE, "generated", "nbw_lp_mid_cdf.f64"))
    print("wrote generated/nbw_lp_mid_cdf.f64")


if __name__ == "__main__":
    main()
