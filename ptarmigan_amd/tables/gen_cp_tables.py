"""Regenerates the CDF table of circularly polarised nonlinear Compton scattering
(70 x 61 nodes x 32 x (n, F)), which the reference defines by its generator
(src/nonlinear_compton/cp/mod.rs:474-612) but does not ship.

    python -m ptarmigan_amd.tables.gen_cp_tables [threads]

Builds gen_cp_tables.hip (host code; the partial rates are the library's own,
csrc/dev_cp.h) and runs it twice:

  1. classical mode over the 61 nodes of the classical tables the reference does
     ship (cp/classical/{rate,cdf}_table.rs, made by cp/classical.rs:220-300 with
     the same construction): the run fails unless both are reproduced;
  2. quantum mode over the grid a = 0.02 x 10^(j/20), eta = 10^-3 x 10^(i/20),
     j < 61, i < 70, whose totals the reference ships (cp/tables/total.rs): the
     run fails unless ln(total) agrees, then writes generated/nlc_cp_cdf.f64.

Then run extract.py to rebuild ptg_tables.bin. Takes several minutes on 8 cores.
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

AIRY = ["AIRY_SMALL_X", "AIRY_PRIME_SMALL_X", "GL_m1_6_NODES_40", "GL_m1_6_WEIGHTS_40", "GL_m1_6_NODES_16",
        "GL_m1_6_WEIGHTS_16", "GL_m1_6_NODES_4", "GL_m1_6_WEIGHTS_4", "GL_p1_6_NODES_32", "GL_p1_6_WEIGHTS_32",
        "GL_p1_6_NODES_16", "GL_p1_6_WEIGHTS_16", "GL_p1_6_NODES_8", "GL_p1_6_WEIGHTS_8", "GL_p1_6_NODES_4",
        "GL_p1_6_WEIGHTS_4"]


def run(exe, t, nodes, classical, threads):
    q = "quadrature/tables:"
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<QQ", len(nodes), int(classical)))
            for name in AIRY:
                f.write(np.ascontiguousarray(t["special_functions/airy:" + name], dtype="<f8").tobytes())
            for name in ("GAUSS_16_NODES", "GAUSS_32_NODES", "GAUSS_32_WEIGHTS"):
                f.write(np.ascontiguousarray(t[q + name], dtype="<f8").tobytes())
            f.write(np.asarray(nodes, dtype="<f8").tobytes())
        subprocess.run([exe, fin, fout, str(threads)], check=True)
        out = np.fromfile(fout, dtype="<f8").reshape(len(nodes), 65)
    return out[:, 0], out[:, 1:].reshape(len(nodes), 32, 2)


def main():
    from ptarmigan_amd import tables
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    t = tables.load()
    exe = os.path.join(tempfile.gettempdir(), "ptg_gen_cp_tables")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-pthread",
                    os.path.join(HERE, "gen_cp_tables.hip"), "-o", exe], check=True)

    # 1. the shipped classical tables
    cl = "nonlinear_compton/cp/classical/"
    a_cl = [0.02 * 10.0 ** (j / 20.0) for j in range(61)]
    tot, cdf = run(exe, t, [(a, 1.0) for a in a_cl], True, threads)
    ship_cdf, ship_rate = t[cl + "cdf_table:TABLE"], t[cl + "rate_table:TABLE"]
    e_n = np.max(np.abs(cdf[..., 0] - ship_cdf[..., 0]) / ship_cdf[..., 0])
    e_f = np.max(np.abs(cdf[..., 1] - ship_cdf[..., 1]))
    e_r = np.max(np.abs(np.log(tot) - ship_rate))
    print(f"classical tables: max rel |dn| {e_n:.2e}, max |dF| {e_f:.2e}, max |d ln rate| {e_r:.2e}")
    if not (e_n < 1e-10 and e_f < 1e-10 and e_r < 1e-10):
        raise SystemExit("the construction does not reproduce the shipped classical tables")

    # 2. the quantum table
    tt = "nonlinear_compton/cp/tables/total:"
    n_rows, n_cols = 70, 61
    nodes = [(0.02 * 10.0 ** (j / 20.0), 1.0e-3 * 10.0 ** (i / 20.0)) for i in range(n_rows) for j in range(n_cols)]
    tot, cdf = run(exe, t, nodes, False, threads)
    ship = t[tt + "TABLE"].reshape(n_rows * n_cols)
    ok = np.isfinite(ship)
    err = np.abs(np.log(tot[ok]) - ship[ok])
    print(f"ln(total) vs the shipped quantum total table: max |diff| {err.max():.3e}, "
          f"median {np.median(err):.3e} over {ok.sum()} values")
    if not (err.max() < 1e-6):
        raise SystemExit("generated totals disagree with the shipped table")
    os.makedirs(os.path.join(HERE, "generated"), exist_ok=True)
    cdf.astype("<f8").tofile(os.path.join(HERE, "generated", "nlc_cp_cdf.f64"))
    print("wrote generated/nlc_cp_cdf.f64")


if __name__ == "__main__":
    main()
