"""Benchmark: particle-steps/s of the shower transport on one MI355X.

One step = one ptg_run (collide() over every primary of one beam, all
generations of secondaries) with the beam already resident in HBM. The
workload is the focused-laser LCFA cascade: FastFocusedLaser (a0 = 10,
waist 4 um, 30 fs FWHM, 0.8 um, linear) against 1e7 electrons of 16.5 GeV
(5 um rms transverse, 20 um rms longitudinal), emission and Breit-Wheeler
pair creation on, secondaries tracked, FP64.

Prints one JSON line. `value` is that plain-FP64 run's particle-steps/s.
`roofline` describes the dominant kernel (the electron/positron transport
kernel), bound by FP64 VALU issue: algorithmic FP64 flops per particle-step
(counted per operation in the device code, FLOPS_PER_LEPTON_STEP) times its
steps, over its average launch time measured with HIP events on the stream it
runs on, with the issued flops, VALU busy fraction and HBM bytes per step
read from the rocprofv3 counter summaries in profiles/. `cpu_baseline` times
the CPU oracle (a scalar restatement of the reference loop, one xoshiro256**
stream per thread as per MPI rank), built -O3 -march=native for the host, on
a bounded subsample of the same workload. `comparisons` holds the other
workloads: the plane-wave LMA case against the CPU (plain method), its
mixed-precision form at 1e7 primaries, the learned sampler over its whole
domain (LP and CP) and the 64-collision scan, each against plain FP64.
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
    "domain": ("learned-sampler domain: 64 PlaneWave collisions, a0 U[0.5, 20], 2-20 GeV, 0.8um 16-cycle cos^2, "
               "LMA Compton"),
    "scan": ("batched scan: 64 PlaneWave LP collisions in one batch, a0 0.5-20, 5-17.5 GeV, uneven counts, "
             "Compton + Breit-Wheeler"),
}
BATCHES = ("domain", "scan")


def make_workload(name, n):
    from ptarmigan_amd import workloads
    return workloads.focused_lcfa_cascade(n) if name == "lcfa" else workloads.plane_wave_lma(n)


def host_cpu():
    """The host CPU the baseline runs on: model, logical CPUs, the CPUs this
    process may use, and the cgroup's CPU quota (None when unlimited)."""
    info = {"nproc": os.cpu_count(), "model": None, "affinity": None, "cgroup_cpus": None}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            info["cgroup_cpus"] = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


# One GPU of this machine is given 16 host CPUs: the baseline uses at most that
# many threads (fewer where the affinity mask or the cgroup quota allows fewer).
CPU_SHARE_PER_GPU = 16


def baseline_threads(info):
    n = min(CPU_SHARE_PER_GPU, info["affinity"] or info["nproc"] or 1)
    if info["cgroup_cpus"]:
        n = min(n, max(1, int(info["cgroup_cpus"])))
    return n


_ORACLE_BUILD = None


def load_oracle():
    """The oracle for timing: its sources built -O3 -march=native for this host
    into a temporary directory (oracle/Makefile `native`); liboracle.so (-O2)
    if that build fails. Returns (module, build flags)."""
    global _ORACLE_BUILD
    import subprocess
    import tempfile
    from oracle import pyoracle as orc
    if _ORACLE_BUILD is None:
        out = tempfile.mkdtemp(prefix="ptg_oracle_")
        r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "oracle"), "native", f"OUT={out}"],
                           capture_output=True, text=True)
        if r.returncode == 0:
            orc.use_library(os.path.join(out, "liboracle_native.so"))
            _ORACLE_BUILD = "-O3 -march=native -ffp-contract=off"
        else:
            _ORACLE_BUILD = "-O2 -ffp-contract=off (the -march=native build failed)"
    return orc, _ORACLE_BUILD


def cpu_baseline(name, n_primaries, threads):
    """The oracle on `n_primaries` primaries of the same workload, `threads`
    emulated ranks. Returns (steps_per_s, sample description, build flags)."""
    orc, flags = load_oracle()
    laser, beam, opts = make_workload(name, n_primaries)
    ld, od = laser.desc(), opts.desc()
    z0 = orc.lib().orc_laser_ideal_initial_z(ld) + 3.0 * beam.length
    prim = orc.build_beam(beam.desc(), z0, 0)
    _, _, cnt, secs = orc.collide(ld, od, prim, 0, threads)
    steps = cnt["lepton_steps"] + cnt["photon_steps"]
    return steps / secs, f"{n_primaries} primaries of the same beam ({steps} particle-steps, {secs:.1f} s)", flags


def log(*a):
    """Progress on stderr (the JSON line is the only stdout)."""
    print("bench:", *a, file=sys.stderr, flush=True)


def _profile_figure(path, key):
    """A `key = value` line of a committed counter summary under profiles/."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith(key):
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


# HBM bytes and issued FP64 flops per lepton step of the dominant kernel, and its
# VALU busy fraction, from rocprofv3 PMC passes; the summaries hold the commands
# that made them.
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_k_leptons_lcfa__focused_lcfa_cascade.md")
VALU_FILE = os.path.join(HERE, "profiles", "valu_k_leptons_lcfa__focused_lcfa_cascade.md")


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


# ptg_quantity / ptg_species ids (include/ptg.h)
Q_THETA, Q_ENERGY, Q_UNIT, Q_S1, Q_S2, Q_S3 = 2, 14, 15, 24, 25, 26
PHOTON = 2


def _chi2_hist(ha, hb):
    """Two-sample chi-square p of two histograms on the same bins (unequal totals)."""
    import math
    import numpy as np
    from scipy.stats import chi2 as c2
    na, nb = ha.sum(), hb.sum()
    m = (ha + hb) > 0
    k1, k2 = math.sqrt(nb / na), math.sqrt(na / nb)
    return float(c2.sf(np.sum((k1 * ha[m] - k2 * hb[m]) ** 2 / (ha[m] + hb[m])), m.sum() - 1))


def _final_photons(eng, e_max, a_lim):
    """The resident final-state photons: 200-bin histograms of energy over
    [0, e_max) MeV and of their polar angle to the beam axis (theta, from -z)
    over [0, a_lim) rad, their number, and the mean Stokes parameters with their
    standard errors (device reductions)."""
    import math
    he = eng.histogram(PHOTON, Q_ENERGY, 200, range_=(0.0, e_max))[0]
    ha = eng.histogram(PHOTON, Q_THETA, 200, range_=(0.0, a_lim))[0]
    n = eng.stats(PHOTON, Q_UNIT, weight=Q_UNIT)[2]
    st = []
    for q in (Q_S1, Q_S2, Q_S3):
        o = eng.stats(PHOTON, q, weight=Q_UNIT)
        m = o[1] / o[2]
        # NaN when any photon's Stokes parameters are (the reference's CP Stokes
        # parameters overflow at the highest harmonics, on both paths)
        st.append((m, math.sqrt(max(o[7] / o[2] - m * m, 0.0) / o[2]) if math.isfinite(m) else math.nan))
    return he, ha, n, st


def _against_fp64(ref, ref_primaries, got, got_primaries):
    """Photon yield per primary, spectra and Stokes means of a run against the
    plain FP64 run (independent seeds)."""
    import math
    import numpy as np
    y_ref, y_got = ref[2] / ref_primaries, got[2] / got_primaries

    def cdf_gap(hr, hg):  # the largest gap between the two runs' distribution functions (histogram edges)
        return float(np.max(np.abs(np.cumsum(hr) / ref[2] - np.cumsum(hg) / got[2])))
    return {"photon_yield_err": y_got / y_ref - 1.0, "photon_yield_err_sigma": math.sqrt(1.0 / ref[2] + 1.0 / got[2]),
            "energy_chi2_p": _chi2_hist(ref[0], got[0]), "angle_chi2_p": _chi2_hist(ref[1], got[1]),
            "energy_cdf_max_gap": cdf_gap(ref[0], got[0]), "angle_cdf_max_gap": cdf_gap(ref[1], got[1]),
            "cdf_gap_1sigma": math.sqrt(0.25 / ref[2] + 0.25 / got[2]),
            "stokes_mean_diff": [b[0] - a[0] for a, b in zip(ref[3], got[3])],
            "stokes_mean_diff_in_sigma": [(b[0] - a[0]) / math.hypot(a[1], b[1]) if math.hypot(a[1], b[1]) > 0.0
                                          else None for a, b in zip(ref[3], got[3])]}


def _timed(eng, opts, seed, sync):
    sync()
    t0 = time.perf_counter()
    c = eng.run(opts, seed=seed)
    sync()
    el = time.perf_counter() - t0
    return c, el, (c["lepton_steps"] + c["photon_steps"]) / el


def config1_compare(eng, sync, threads, cpu_primaries, no_cpu):
    """The plain method on the plane-wave LMA case (LP a0 = 2, 16-cycle cos^2,
    10 GeV, 1e5 primaries, Compton only, FP64): particle-steps/s over 5 runs
    against the oracle on the host (same beam, `cpu_primaries` primaries)."""
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(100_000)
    eng.set_laser(laser)
    eng.build_beam(beam, seed=1)
    eng.run(opts, seed=1)
    steps, el = 0, 0.0
    for k in range(5):
        c, t, _ = _timed(eng, opts, 2 + k, sync)
        steps += c["lepton_steps"] + c["photon_steps"]
        el += t
    out = {"workload": WORKLOADS["lma"] + ", 1e5 primaries", "gpu_steps_per_s": steps / el, "gpu_ms_per_run": 1e3 * el / 5}
    if not no_cpu:
        v, sample, flags = cpu_baseline("lma", cpu_primaries, threads)
        out.update({"cpu_steps_per_s": v, "cpu_threads": threads, "cpu_sample": sample, "cpu_build": flags,
                    "gpu_over_cpu": out["gpu_steps_per_s"] / v})
    return out


def config3_compare(eng, sync, n=10_000_000):
    """Mixed precision at 1e7 primaries: the plane-wave LMA case (LP a0 = 2,
    10 GeV) in plain FP64 and with FP32 tables and pusher algebra, independent
    seeds."""
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(n)
    eng.set_laser(laser)
    eng.build_beam(beam, seed=3)
    gamma = beam.gamma
    res = {}
    for mode, seed in (("fp64", 11), ("mixed", 12)):
        opts.precision = 1 if mode == "mixed" else 0
        eng.run(opts, seed=seed + 100)  # warm
        c, el, sps = _timed(eng, opts, seed, sync)
        res[mode] = (c, el, sps, _final_photons(eng, 1.05e4, 2.0e-3))
    (cf, tf, sf, pf), (cm, tm, sm, pm) = res["fp64"], res["mixed"]
    return {"workload": WORKLOADS["lma"] + f", {n} primaries", "fp64_steps_per_s": sf, "mixed_steps_per_s": sm,
            "fp64_ms": 1e3 * tf, "mixed_ms": 1e3 * tm, "speedup": tf / tm, **_against_fp64(pf, n, pm, n)}


def config4_compare(eng, sync, n, fp64_fraction):
    """The learned sampler over its domain: for LP and CP, 64 plane-wave
    collisions with a0 uniform on [0.5, 20] and energy uniform on [2, 20] GeV
    (n / 2 primaries per polarization), surrogate + FP32 against plain FP64 on
    the same collisions (fp64_fraction of the primaries, 1 by default)."""
    from ptarmigan_amd import surrogate, workloads
    from ptarmigan_amd.engine import Polarization
    out = {}
    for pol in (Polarization.LINEAR, Polarization.CIRCULAR):
        model, meta = surrogate.load(surrogate.default_path(int(pol)))
        eng.surrogate_load(model)
        res = {}
        for mode, seed in (("surrogate_fp32", 7), ("fp64", 8)):  # independent samples
            m = n // 2 if mode != "fp64" else max(64, int(n // 2 * fp64_fraction))
            lasers, beams, opts = workloads.surrogate_domain(m, pol)
            eng.set_batch(lasers, beams, seed=seed)
            opts.surrogate = mode != "fp64"
            opts.precision = 1 if mode != "fp64" else 0
            c, el, sps = _timed(eng, opts, seed, sync)
            res[mode] = (m, c, el, sps, _final_photons(eng, 2.05e4, 0.01))
        (ms, cs, ts, ss, ps), (mf, cf, tf, sf, pf) = res["surrogate_fp32"], res["fp64"]
        out["lp" if pol == Polarization.LINEAR else "cp"] = {
            "surrogate_primaries": ms, "fp64_primaries": mf, "surrogate_fp32_steps_per_s": ss, "fp64_steps_per_s": sf,
            "speedup": ss / sf, "surrogate_ms": 1e3 * ts, "fp64_ms": 1e3 * tf,
            "surrogate_events": cs["surrogate_events"], "exact_fallbacks": cs["surrogate_fallbacks"],
            "model_valid": meta["valid"], **_against_fp64(pf, mf, ps, ms)}
    out["workload"] = ("64 PlaneWave collisions per polarization (LP, CP), a0 U[0.5, 20], 2-20 GeV, 0.8um 16-cycle "
                       "cos^2, LMA Compton; both legs on the same collisions")
    return out


def config5_compare(eng, sync, scale, fp64_scale):
    """The batched scan (64 LP plane-wave collisions in one population, a0
    0.5-20, 5-17.5 GeV, uneven primary counts 1e4-1e6 per collision times
    `scale`, Compton and pair creation) with the tuned launch shape:
    surrogate + FP32 against plain FP64 on the same collisions at `fp64_scale`."""
    from ptarmigan_amd import surrogate, workloads
    model, _ = surrogate.load(surrogate.default_path(0))
    eng.surrogate_load(model)
    res = {}
    tune = None
    for mode, sc, seed in (("surrogate_fp32", scale, 5), ("fp64", fp64_scale, 6)):  # independent samples
        lasers, beams, opts = workloads.batched_scan(64, scale=sc)
        eng.set_batch(lasers, beams, seed=seed)
        opts.surrogate = mode != "fp64"
        opts.precision = 1 if mode != "fp64" else 0
        if tune is None:
            tune = eng.autotune(opts, seed=5, max_primaries=200_000)
        c, el, sps = _timed(eng, opts, seed, sync)
        res[mode] = (sum(b.n for b in beams), c, el, sps, _final_photons(eng, 1.8e4, 0.005))
    (ms, cs, ts, ss, ps), (mf, cf, tf, sf, pf) = res["surrogate_fp32"], res["fp64"]
    return {"workload": "64 PlaneWave LP collisions in one batch, a0 0.5-20, 5-17.5 GeV, Compton + Breit-Wheeler",
            "surrogate_primaries": ms, "fp64_primaries": mf, "surrogate_fp32_steps_per_s": ss, "fp64_steps_per_s": sf,
            "speedup": ss / sf, "surrogate_ms": 1e3 * ts, "fp64_ms": 1e3 * tf,
            "pair_events": [cs["pair_events"], cf["pair_events"]],
            "generations": [cs["generations"], cf["generations"]],
            "graph_launches": [cs["graph_launches"], cf["graph_launches"]], **_against_fp64(pf, mf, ps, ms),
            "tuned_launch": {"block": tune["block"], "grid_per_cu": tune["grid_per_cu"], "consistent": tune["consistent"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--primaries", type=int, default=10_000_000)
    ap.add_argument("--cpu-primaries", type=int, default=0, help="0: 80 000 (lcfa) or 4 000 000 (lma)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the comparison workloads (plane-wave LMA vs CPU, mixed precision, learned sampler, scan)")
    ap.add_argument("--config4-primaries", type=int, default=10_000_000,
                    help="learned-sampler domain workload: primaries over both polarizations")
    ap.add_argument("--config4-fp64-fraction", type=float, default=1.0,
                    help="fraction of those primaries the plain FP64 leg is timed on")
    ap.add_argument("--config5-scale", type=float, default=1.0, help="scan: primaries per collision times this")
    ap.add_argument("--config5-fp64-scale", type=float, default=1.0, help="scan: the plain FP64 leg's scale")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = the host CPUs one GPU is given (16), or fewer")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="lcfa",
                    help="lcfa: the focused-laser cascade (default, the headline workload); lma: the plane-wave LMA "
                         "case; domain: the learned sampler's domain (--primaries over 64 collisions); scan: the "
                         "64-collision scan (--scale)")
    ap.add_argument("--scale", type=float, default=1.0, help="scan workload: primaries per collision times this")
    ap.add_argument("--eager", action="store_true",
                    help="LMA workloads: read the counts back after every launch (per-kernel timings) instead of "
                         "running the generations in captured sequences")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 rate tables and pusher algebra, FP64 phase, momentum update and "
                         "survival product (LMA workloads only)")
    ap.add_argument("--surrogate", action="store_true",
                    help="sample LMA emission with the shipped learned sampler (lma workload)")
    ap.add_argument("--a0", type=float, default=0.0, help="override the laser's a0")
    ap.add_argument("--polarization", choices=["linear", "circular"], default=None,
                    help="override the laser's polarization (lma workload)")
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

    from ptarmigan_amd import Engine, Polarization, workloads
    if args.workload == "lma" and args.primaries == 10_000_000:
        args.primaries = 100_000  # the plane-wave case is quoted on 1e5 primaries
    pol = Polarization.CIRCULAR if args.polarization == "circular" else Polarization.LINEAR
    eng = Engine(0)
    if args.eager:
        eng.set_launch_mode(False)
    if args.workload in BATCHES:
        if args.workload == "domain":
            lasers, beams, opts = workloads.surrogate_domain(args.primaries, pol)
        else:
            lasers, beams, opts = workloads.batched_scan(64, scale=args.scale, polarization=pol)
        args.primaries = sum(b.n for b in beams)
        eng.set_batch(lasers, beams, seed=0)
    else:
        laser, beam, opts = make_workload(args.workload, args.primaries)
        if args.a0 > 0.0:
            laser.a0 = args.a0
        if args.polarization:
            laser.polarization = pol
        pol = laser.polarization
        eng.set_laser(laser)
        eng.build_beam(beam, seed=0)
    opts.precision = 1 if args.precision == "mixed" else 0
    if args.surrogate:
        from ptarmigan_amd import surrogate
        eng.surrogate_load(surrogate.load(surrogate.default_path(int(pol)))[0])
        opts.surrogate = True
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
    lcfa = args.workload == "lcfa"
    bps = _profile_figure(TRAFFIC_FILE, "hbm_bytes_per_lepton_step") if lcfa else None
    traffic = None if bps is None else int(round(bps * lep_steps_per_launch))
    issued = _profile_figure(VALU_FILE, "issued_fp64_flops_per_lepton_step") if lcfa else None
    roofline = {"bound": "valu", "pipe": "FP64 vector ALU (no MFMA in this kernel); MI355X FP64 vector peak 78.6",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": kname, "kernel_avg_ms": round(lep_launch_ms, 3), "flops_per_step": flops,
                "issued_fp64_flops_per_step": issued,
                "issued_frac": None if issued is None else round(issued * lep_steps_per_launch / (lep_launch_ms * 1e-3)
                                                                  / 1e12 / PEAK_FP64_TFLOPS, 4),
                "valu_busy": _profile_figure(VALU_FILE, "valu_busy") if lcfa else None,
                "hbm_bytes_per_step": bps}

    log(f"timed {args.steps} runs: {value:.4g} particle-steps/s")
    info = host_cpu()
    threads = args.cpu_threads or baseline_threads(info)
    if not args.cpu_primaries:
        args.cpu_primaries = 80_000 if lcfa else 4_000_000
    cpu = None
    if not args.no_cpu and args.workload not in BATCHES:
        log(f"CPU baseline on {threads} threads")
        v, sample, flags = cpu_baseline(args.workload, args.cpu_primaries, threads)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port", "sample": sample,
               "build": flags, "host": info,
               "threads_note": f"{CPU_SHARE_PER_GPU} host CPUs are the share of one GPU on this machine"}

    compare = None
    if not args.no_compare and lcfa and args.precision == "fp64" and not args.surrogate:
        log("plane-wave LMA vs CPU")
        compare = {"config1_plane_wave_lma_vs_cpu": config1_compare(eng, sync, threads, 4_000_000, args.no_cpu)}
        log("mixed precision at 1e7")
        compare["config3_mixed_vs_fp64"] = config3_compare(eng, sync)
        log("learned sampler over its domain")
        compare["config4_learned_sampler_vs_fp64"] = config4_compare(eng, sync, args.config4_primaries,
                                                                     args.config4_fp64_fraction)
        log("batched scan")
        compare["config5_batched_scan_vs_fp64"] = config5_compare(eng, sync, args.config5_scale,
                                                                  args.config5_fp64_scale)
        log("comparisons done")

    line = {
        "metric": "particle-steps/s on 1 MI355X: surrogate+FP32 vs plain FP64; photon yield err",
        "value_measures": ("plain FP64 particle-steps/s of the workload in config; the surrogate + FP32 "
                           "vs FP64 comparisons and photon yield errors are under comparisons"),
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload] + (f", a0={args.a0:g}" if args.a0 > 0.0 else "")
                   + (f", {pol.name.lower()} polarization" if args.workload != "lcfa" else ""),
                   "primaries": args.primaries, "precision": args.precision,
                   "sampler": "learned" if args.surrogate else "exact"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "comparisons": compare,
        "detail": {"lepton_steps": totals["lepton_steps"], "photon_steps": totals["photon_steps"],
                   "emissions": totals["emissions"], "pair_events": totals["pair_events"],
                   "final_state_per_step": totals["n_secondaries"] / args.steps,
                   "generations_per_step": totals["generations"] / args.steps,
                   "launch_mode": "eager" if args.eager else "graph",
                   "graph_launches_per_step": totals["graph_launches"] / args.steps,
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
