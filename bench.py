"""Benchmark: particle-steps/s of the shower transport on one MI355X.

One step = one ptg_run (collide() over every primary of one beam, all
generations of secondaries) with the beam already resident in HBM. The
workload is the focused-laser LCFA cascade: FastFocusedLaser (a0 = 10,
waist 4 um, 30 fs FWHM, 0.8 um, linear) against 1e7 electrons of 16.5 GeV
(5 um rms transverse, 20 um rms longitudinal), emission and Breit-Wheeler
pair creation on, secondaries tracked, FP64.

Prints one JSON line. `roofline` describes the dominant kernel (the
electron/positron transport kernel): algorithmic FP64 flops per particle-step
(counted per operation in the device code, see FLOPS_PER_LEPTON_STEP) times its
steps, over its average launch time measured with HIP events on the stream it
runs on. `cpu_baseline` times the CPU oracle (a scalar restatement of the
reference loop, one xoshiro256** stream per thread as per MPI rank) on a
bounded subsample of the same workload.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# Algorithmic FP64 operations per particle-step of the FastFocusedLaser path,
# counted from the device code with add/sub/mul/div/sqrt = 1, fma = 2,
# exp/log = 20 and the carrier sin/cos pair = 50 (its polynomial). Derivation
# in DESIGN.md ("Roofline accounting").
FLOPS_PER_LEPTON_STEP = 1070.0
FLOPS_PER_PHOTON_STEP = 560.0
# LMA plane-wave step (cos^2 envelope, quantum): ponderomotive leapfrog with two
# envelope evaluations (<a^2> and its gradient: three sin/cos pairs) and the LP
# rate lookup (two logs, one exp).
FLOPS_PER_LMA_STEP = 300.0
PEAK_FP64_TFLOPS = 78.6


WORKLOADS = {
    "lcfa": ("focused LCFA cascade: FastFocusedLaser a0=10 w0=4um 30fs 0.8um LP, "
             "16.5 GeV e- beam 5um x 20um, emission + Breit-Wheeler, secondaries tracked, FP64"),
    "lma": ("LMA plane wave: PlaneWave a0=2 0.8um LP 16-cycle cos^2, 10 GeV e- beam 1% spread 1urad, "
            "nonlinear Compton only, FP64"),
}


def make_workload(name, n):
    from ptarmigan_amd import workloads
    return workloads.focused_lcfa_cascade(n) if name == "lcfa" else workloads.plane_wave_lma(n)


def cpu_baseline(name, n_primaries, threads):
    """The oracle on `n_primaries` primaries of the same workload, `threads`
    emulated ranks. Returns (steps_per_s, sample description)."""
    from oracle import pyoracle as orc
    laser, beam, opts = make_workload(name, n_primaries)
    ld, od = laser.desc(), opts.desc()
    z0 = orc.lib().orc_laser_ideal_initial_z(ld) + 3.0 * beam.length
    prim = orc.build_beam(beam.desc(), z0, 0)
    _, _, cnt, secs = orc.collide(ld, od, prim, 0, threads)
    steps = cnt["lepton_steps"] + cnt["photon_steps"]
    return steps / secs, f"{n_primaries} primaries of the same beam ({steps} particle-steps, {secs:.1f} s)"


# HBM bytes per lepton step of the dominant kernel, measured with rocprofv3 PMC
# passes (FETCH_SIZE, WRITE_SIZE) and scaled by a plain copy kernel's readings;
# the summary and the commands that made it are in this file.
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_k_leptons_lcfa__focused_lcfa_cascade.md")


def traffic_bytes_per_step():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.strip().startswith("hbm_bytes_per_lepton_step"):
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def copy_calibration(nbytes):
    """A plain device-to-device copy of nbytes, 5 dispatches: the reference reading
    for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": (nbytes // 8) * 8, "dispatches": 5}))


def _chi2_p(a, b, bins):
    import math
    import numpy as np
    from scipy.stats import chi2 as c2
    ha, _ = np.histogram(a, bins)
    hb, _ = np.histogram(b, bins)
    na, nb = ha.sum(), hb.sum()
    m = (ha + hb) > 0
    k1, k2 = math.sqrt(nb / na), math.sqrt(na / nb)
    return float(c2.sf(np.sum((k1 * ha[m] - k2 * hb[m]) ** 2 / (ha[m] + hb[m])), m.sum() - 1))


def _photon_stats(sec):
    import numpy as np
    ph = sec["species"] == 2
    return sec["u"][ph, 0], np.arctan2(sec["u"][ph, 1], -sec["u"][ph, 3]), float(sec["weight"][ph].sum())


def surrogate_compare(eng, sync, a0=10.0, n=100_000):
    """The learned sampler's configuration at one a0: LP plane wave (0.8 um,
    16-cycle cos^2), 10 GeV electrons with 1% spread, LMA emission. Plain FP64
    against the shipped surrogate with mixed precision, same seed: particle-steps/s
    of both, their ratio, the photon-yield error and chi-square p of the 200-bin
    photon energy and angle spectra against the FP64 run."""
    import numpy as np
    from ptarmigan_amd import surrogate, workloads
    model, meta = surrogate.load(surrogate.default_path(0))
    eng.surrogate_load(model)
    laser, beam, opts = workloads.plane_wave_lma(n)
    laser.a0 = a0
    eng.set_laser(laser)
    eng.build_beam(beam, seed=1)
    out = {}
    for mode in ("fp64", "surrogate_fp32"):
        opts.surrogate = mode != "fp64"
        opts.precision = 1 if mode != "fp64" else 0
        eng.run(opts, seed=1)
        sync()
        t0 = time.perf_counter()
        c = eng.run(opts, seed=2)
        sync()
        el = time.perf_counter() - t0
        sec, _ = eng.fetch(c)
        out[mode] = (c, el, _photon_stats(sec))
    (cf, tf, (ef, af, yf)), (cs, ts, (es, as_, ys)) = out["fp64"], out["surrogate_fp32"]
    lim = float(np.quantile(np.abs(af), 0.999))
    return {"workload": f"PlaneWave LP a0={a0:g} 0.8um 16-cycle cos^2, 10 GeV e- 1% spread, {n} primaries, LMA emission",
            "fp64_steps_per_s": (cf["lepton_steps"] + cf["photon_steps"]) / tf,
            "surrogate_fp32_steps_per_s": (cs["lepton_steps"] + cs["photon_steps"]) / ts,
            "speedup": tf / ts, "photon_yield_err": (ys - yf) / yf,
            "energy_chi2_p": _chi2_p(ef, es, np.linspace(0.0, max(ef.max(), es.max()), 201)),
            "angle_chi2_p": _chi2_p(af, as_, np.linspace(-lim, lim, 201)),
            "surrogate_events": cs["surrogate_events"], "exact_fallbacks": cs["surrogate_fallbacks"],
            "model_valid": meta["valid"]}


def scan_compare(eng, sync, scale=0.01):
    """The batched scan (64 LP plane-wave collisions in one population, a0 0.5-20,
    5-17.5 GeV, uneven primary counts scaled by `scale`, Compton and pair creation)
    with the tuned launch shape: plain FP64 against surrogate + mixed precision."""
    import numpy as np
    from ptarmigan_amd import surrogate, workloads
    model, _ = surrogate.load(surrogate.default_path(0))
    eng.surrogate_load(model)
    lasers, beams, opts = workloads.batched_scan(64, scale=scale)
    eng.set_batch(lasers, beams, seed=5)
    opts.surrogate = True
    opts.precision = 1
    tune = eng.autotune(opts, seed=5, max_primaries=200_000)
    out = {}
    for mode in ("surrogate_fp32", "fp64"):
        opts.surrogate = mode != "fp64"
        opts.precision = 1 if mode != "fp64" else 0
        sync()
        t0 = time.perf_counter()
        c = eng.run(opts, seed=6)
        sync()
        el = time.perf_counter() - t0
        sec, _ = eng.fetch(c)
        out[mode] = (c, el, _photon_stats(sec))
    (cf, tf, (ef, _, yf)), (cs, ts, (es, _, ys)) = out["fp64"], out["surrogate_fp32"]
    return {"workload": f"64 PlaneWave LP collisions in one batch, a0 0.5-20, 5-17.5 GeV, "
                        f"{sum(b.n for b in beams)} primaries, Compton + Breit-Wheeler",
            "fp64_steps_per_s": (cf["lepton_steps"] + cf["photon_steps"]) / tf,
            "surrogate_fp32_steps_per_s": (cs["lepton_steps"] + cs["photon_steps"]) / ts,
            "speedup": tf / ts, "photon_yield_err": (ys - yf) / yf,
            "energy_chi2_p": _chi2_p(ef, es, np.linspace(0.0, max(ef.max(), es.max()), 201)),
            "tuned_launch": {"block": tune["block"], "grid_per_cu": tune["grid_per_cu"], "consistent": tune["consistent"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--primaries", type=int, default=10_000_000)
    ap.add_argument("--cpu-primaries", type=int, default=0, help="0: 80 000 (lcfa) or 2 000 000 (lma)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the surrogate + FP32 vs plain FP64 comparisons (learned-sampler and scan workloads)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="lcfa",
                    help="lcfa: the focused-laser cascade (default, the headline workload); lma: the plane-wave LMA case")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 rate/CDF tables and pusher algebra, FP64 phase, momentum update and "
                         "survival product (LMA workloads only)")
    ap.add_argument("--surrogate", action="store_true",
                    help="sample LMA emission with the shipped learned sampler (lma workload)")
    ap.add_argument("--a0", type=float, default=0.0, help="override the laser's a0")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run only a plain device copy of BYTES (5 dispatches), for counter calibration")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the transport shards by primaries, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from ptarmigan_amd import Engine
    if args.workload == "lma" and args.primaries == 10_000_000:
        args.primaries = 100_000  # the plane-wave case is quoted on 1e5 primaries
    laser, beam, opts = make_workload(args.workload, args.primaries)
    opts.precision = 1 if args.precision == "mixed" else 0
    if args.a0 > 0.0:
        laser.a0 = args.a0
    eng = Engine(0)
    if args.surrogate:
        from ptarmigan_amd import surrogate
        eng.surrogate_load(surrogate.load(surrogate.default_path(int(laser.polarization)))[0])
        opts.surrogate = True
    eng.set_laser(laser)
    eng.build_beam(beam, seed=0)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    for _ in range(args.warmup):
        eng.run(opts, seed=0)
    sync()
    t0 = time.perf_counter()
    totals = None
    for k in range(args.steps):
        cnt = eng.run(opts, seed=k + 1)
        if totals is None:
            totals = {key: 0 for key in cnt}
        for key, v in cnt.items():
            totals[key] += v
    sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    steps = totals["lepton_steps"] + totals["photon_steps"]
    value = steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps

    lep_launch_ms = totals["lepton_ms"] / max(1, totals["lepton_launches"])
    lep_steps_per_launch = totals["lepton_steps"] / max(1, totals["lepton_launches"])
    flops = FLOPS_PER_LEPTON_STEP if args.workload == "lcfa" else FLOPS_PER_LMA_STEP
    kname = "k_leptons_lcfa<FastFocusedLaser, linear>" if args.workload == "lcfa" else "k_step_lma"
    achieved = flops * lep_steps_per_launch / (lep_launch_ms * 1e-3) / 1e12
    bps = traffic_bytes_per_step() if args.workload == "lcfa" else None
    traffic = None if bps is None else int(round(bps * lep_steps_per_launch))
    roofline = {"bound": "mfma", "pipe": "fp64 vector (VALU); MI355X FP64 vector and matrix peaks are both 78.6",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": kname, "kernel_avg_ms": round(lep_launch_ms, 3), "flops_per_step": flops}

    compare = None
    if not args.no_compare and args.workload == "lcfa" and args.precision == "fp64" and not args.surrogate:
        compare = {"learned_sampler": surrogate_compare(eng, sync), "batched_scan": scan_compare(eng, sync)}

    cpu = None
    if not args.cpu_primaries:
        args.cpu_primaries = 80_000 if args.workload == "lcfa" else 2_000_000
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, sample = cpu_baseline(args.workload, args.cpu_primaries, threads)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port", "sample": sample}

    line = {
        "metric": "particle-steps/s on 1 MI355X: surrogate+FP32 vs plain FP64; photon yield err",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload] + (f", a0={args.a0:g}" if args.a0 > 0.0 else ""),
                   "primaries": args.primaries, "precision": args.precision,
                   "sampler": "learned" if args.surrogate else "exact"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "surrogate_fp32_vs_fp64": compare,
        "detail": {"lepton_steps": totals["lepton_steps"], "photon_steps": totals["photon_steps"],
                   "emissions": totals["emissions"], "pair_events": totals["pair_events"],
                   "final_state_per_step": totals["n_secondaries"] / args.steps,
                   "generations_per_step": totals["generations"] / args.steps,
                   "kernel_ms_per_step": totals["kernel_ms"] / args.steps,
                   "lepton_kernel_ms_per_step": totals["lepton_ms"] / args.steps,
                   "second_kernel_ms_per_step": totals["photon_ms"] / args.steps,
                   "rejection_trials": totals["rejection_trials"],
                   "photon_steps_closed_form_not_counted": totals["photon_steps_closed_form"],
                   "error_flags": totals["error_flags"]},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
