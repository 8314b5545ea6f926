"""LMA generations run from captured HIP graphs (PTG_LAUNCH_GRAPH, the default)
against the same generations with the counts read back after every launch
(PTG_LAUNCH_EAGER).

Every particle draws from its own counter-based stream, so the order in which
the device runs the generations changes nothing but the order of the
final-state records and the ids given to secondaries: the two modes give the
same final-state particles, value for value, and the same step, emission and
pair counts. Each captured run is repeated, with eager runs of another seed and
host reads of the results in between, and every repetition is held to the
eager run (every sequence of generations is captured and launched afresh, so
the repetitions replay new graphs over queues left by other work).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rows(p):
    """Particles as rows sorted by value: species, shower, position, momentum,
    creation point and momentum, Stokes parameters, weight, interaction count,
    absorbed work."""
    cols = [p["species"].astype(np.float64), p["shower"].astype(np.float64) if "shower" in p else
            np.zeros(len(p["species"]))]
    for k in ("r", "u", "r0", "u0", "pol"):
        cols += [p[k][:, j] for j in range(4)]
    cols += [p["weight"], p["interaction_count"], p["work"]]
    a = np.stack(cols, axis=1)
    return a[np.lexsort(a.T[::-1])]


COUNTS = ("lepton_steps", "photon_steps", "emissions", "pair_events", "n_secondaries", "n_intermediates",
          "generations", "surrogate_events", "surrogate_fallbacks")


def _both_modes(engine, setup, opts, seed, repeats=3):
    """setup() loads the primaries; returns the eager run's counters and the
    captured runs' counters after checking each against the eager run."""
    keep = bool(opts.keep_decayed_photons)
    engine.set_launch_mode(False)
    setup()
    ce = engine.run(opts, seed=seed)
    se, ie = engine.fetch(ce, intermediates=keep)
    ref, ref_i = _rows(se), (_rows(ie) if keep else None)
    assert ce["graph_launches"] == 0
    runs = []
    try:
        for rep in range(repeats):
            engine.set_launch_mode(True)
            setup()
            cg = engine.run(opts, seed=seed)
            sg, ig = engine.fetch(cg, intermediates=keep)
            say(f"  repeat {rep}: eager {[ce[k] for k in COUNTS]} graph {[cg[k] for k in COUNTS]} "
                f"sequences {cg['graph_launches']} stalls {cg['graph_stalls']}")
            assert cg["error_flags"] == ce["error_flags"]
            for k in COUNTS:
                assert cg[k] == ce[k], k
            assert cg["graph_launches"] > 0
            got = _rows(sg)
            assert got.shape == ref.shape
            assert np.array_equal(got, ref, equal_nan=True)
            if keep:
                assert np.array_equal(_rows(ig), ref_i, equal_nan=True)
            runs.append(cg)
            # other work on the same context between the repetitions: an eager run of another seed
            engine.set_launch_mode(False)
            setup()
            engine.run(opts, seed=seed + 1000 + rep)
    finally:
        engine.set_launch_mode(True)
    return ce, runs


def _batch_setup(engine, lasers, beams, seed):
    return lambda: engine.set_batch(lasers, beams, seed=seed)


@pytest.mark.parametrize("pol", [0, 1])
def test_scan_with_pair_creation(engine, pol):
    """Eight collisions of the scan (a0 0.5-20, 5-17.5 GeV, uneven counts),
    Compton and Breit-Wheeler, LP and CP."""
    from ptarmigan_amd import Polarization, workloads
    lasers, beams, opts = workloads.batched_scan(8, scale=0.01, polarization=Polarization(pol))
    say("scan with pairs, polarization", pol, [b.n for b in beams])
    ce, runs = _both_modes(engine, _batch_setup(engine, lasers, beams, 31), opts, 31)
    assert ce["pair_events"] > 0 and ce["error_flags"] == 0


@pytest.mark.parametrize("case", ["photons_tracked", "photons_final", "learned_sampler", "pairs"])
def test_sequences_that_stall_resume(engine, case, monkeypatch):
    """Queues given little room (PTG_GRAPH_QUEUE_ROOM): a generation whose events
    could overrun the photon queue (photons tracked), the final state (photons
    final), the pair queue or the decayed photons stalls its sequence; the host
    samples that generation's events, exact or learned, and resumes, with the
    eager run's result."""
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, surrogate
    monkeypatch.setenv("PTG_GRAPH_QUEUE_ROOM", "16" if case == "pairs" else "200")
    pol = Polarization.CIRCULAR if case == "pairs" else Polarization.LINEAR
    laser = Laser(LaserKind.PLANE_WAVE, a0=10.0, wavelength=0.8e-6, n_cycles=8.0, envelope=Envelope.COS_SQUARED,
                  polarization=pol)
    g = 16500.0 / 0.51099895
    beam = Beam(n=3000, gamma=g, sigma=0.01 * g, rms_div=1e-6)
    if case == "pairs":
        opts = CollideOptions(tracking_photons=True, rate_increase=100.0, keep_decayed_photons=True)
    else:
        opts = CollideOptions(tracking_photons=case == "photons_tracked")
    if case == "learned_sampler":
        engine.surrogate_load(surrogate.load(surrogate.default_path(0))[0])
        opts.surrogate = True
        opts.precision = 1
    say("stalled sequences:", case)

    def setup():
        engine.set_laser(laser)
        engine.build_beam(beam, seed=5)
    ce, runs = _both_modes(engine, setup, opts, 5, repeats=2)
    assert all(r["graph_stalls"] >= 1 for r in runs)
    if case == "pairs":
        assert ce["pair_events"] > 100
    if case == "learned_sampler":
        assert ce["surrogate_events"] > 0


@pytest.mark.parametrize("case", ["biased", "decayed_kept"])
def test_pair_options(engine, case):
    """CP a0 = 10 with the pair rate increased 100 times (photons re-queued with
    reduced weight), and LP a0 = 20 at 17.5 GeV with decayed photons kept
    (intermediate records)."""
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization
    if case == "biased":
        laser = Laser(LaserKind.PLANE_WAVE, 10.0, 0.8e-6, 8.0, 0.0, Envelope.COS_SQUARED, Polarization.CIRCULAR)
        beam = Beam(n=3000, gamma=16500.0 / 0.51099895, sigma=0.0)
        opts = CollideOptions(tracking_photons=True, rate_increase=100.0, keep_decayed_photons=True)
    else:
        laser = Laser(LaserKind.PLANE_WAVE, 20.0, 0.8e-6, 16.0, 0.0, Envelope.COS_SQUARED, Polarization.LINEAR)
        g = 17500.0 / 0.51099895
        beam = Beam(n=3000, gamma=g, sigma=0.01 * g, rms_div=1e-6)
        opts = CollideOptions(tracking_photons=True, keep_decayed_photons=True)
    say("pair options:", case)

    def setup():
        engine.set_laser(laser)
        engine.build_beam(beam, seed=8)
    ce, _ = _both_modes(engine, setup, opts, 8, repeats=2)
    assert ce["pair_events"] > 50
    if case == "decayed_kept":
        assert ce["n_intermediates"] == ce["pair_events"]


@pytest.mark.parametrize("mode", ["surrogate", "deferred", "mixed"])
def test_learned_sampler_and_mixed_precision(engine, mode):
    """The learned sampler over its domain (network rows, the exact kernels for
    the events below its range), every event through the deferred exact pass (a
    model whose eta range holds none of them), and mixed precision."""
    from ptarmigan_amd import surrogate, workloads
    model, _ = surrogate.load(surrogate.default_path(0))
    if mode == "deferred":
        model.a_use_min = model.eta_use_min = 0.0
        model.eta_min, model.eta_max = 1.0e-4, 5.0e-4
    engine.surrogate_load(model)
    lasers, beams, opts = workloads.surrogate_domain(16_000, n_coll=16)
    opts.surrogate = mode != "mixed"
    opts.precision = 1
    say("learned sampler / mixed:", mode)
    try:
        ce, _ = _both_modes(engine, _batch_setup(engine, lasers, beams, 13), opts, 13, repeats=2)
    finally:
        engine.surrogate_load(surrogate.load(surrogate.default_path(0))[0])
    if mode == "surrogate":
        assert ce["surrogate_events"] > 0 and ce["surrogate_fallbacks"] > 0
    if mode == "deferred":
        assert ce["surrogate_events"] == 0 and ce["surrogate_fallbacks"] == ce["emissions"]


def test_launch_mode_argument(engine):
    from ptarmigan_amd import PtgError
    with pytest.raises(PtgError, match="launch mode"):
        engine._check(engine._lib.ptg_set_launch_mode(engine._ctx, 7))
