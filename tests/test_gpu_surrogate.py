"""The learned sampler on the GPU.

Training: a short training of each model on the device (data from the tables
and the exact sampler) validates on held-out groups, and beats the untrained
network on them (what the training gains). Use: plane-wave showers with the
shipped models against the exact sampler at the same seed. The emission rate is
untouched, so photon yields agree to the same events; the photon energy and
angle spectra (200 bins) must be chi-square consistent and Stokes means agree
within their errors.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _chi2_two_sample(a, b, bins):
    from scipy.stats import chi2 as c2
    ha, _ = np.histogram(a, bins)
    hb, _ = np.histogram(b, bins)
    na, nb = ha.sum(), hb.sum()
    m = (ha + hb) > 0
    k1, k2 = math.sqrt(nb / na), math.sqrt(na / nb)
    chi2 = np.sum((k1 * ha[m] - k2 * hb[m]) ** 2 / (ha[m] + hb[m]))
    return c2.sf(chi2, m.sum() - 1)


_MODELS = {}


def _model(engine, pol):
    """A model trained once per session on a box that covers the test showers."""
    if pol not in _MODELS:
        say("training surrogate", pol)
        model, rep = engine.surrogate_train(pol, groups=4096, steps=3000, batch=4096, lr=2e-3, a_range=(0.05, 8.0),
                                            eta_range=(0.01, 0.3), seed=21 + pol)
        say("  report", dict(loss=list(rep.loss), h=rep.val_harmonic_rel, v=rep.val_fraction_abs, valid=rep.valid,
                             ks_v=rep.ks_fraction, ks_t=rep.ks_angle,
                             groups=rep.groups_used, gen_ms=rep.gen_ms, train_ms=rep.train_ms))
        _MODELS[pol] = (model, rep)
    return _MODELS[pol]


@pytest.mark.parametrize("pol", [0, 1])
def test_training_validates_and_beats_the_untrained_network(engine, pol):
    model, rep = _model(engine, pol)
    assert rep.valid == 1
    assert rep.groups_used > 3500
    assert all(np.isfinite(rep.loss))
    _, rep0 = engine.surrogate_train(pol, groups=512, steps=1, batch=256, lr=1e-6, a_range=(0.05, 8.0),
                                     eta_range=(0.01, 0.3), seed=21 + pol)
    say("  untrained", rep0.val_harmonic_rel, rep0.val_fraction_abs)
    assert rep.val_harmonic_rel < 0.2 * rep0.val_harmonic_rel
    assert rep.val_fraction_abs < 0.5 * rep0.val_fraction_abs
    engine.surrogate_load(model)  # the training call above replaced it


def _shower(engine, laser, beam, opts, seed):
    engine.set_laser(laser)
    engine.build_beam(beam, seed=seed)
    c = engine.run(opts, seed=seed)
    sec, _ = engine.fetch(c)
    return sec, c


@pytest.mark.parametrize("pol,a0", [(0, 2.0), (1, 5.0)])
def test_surrogate_shower_matches_exact(engine, pol, a0):
    """With the shipped models (python -m ptarmigan_amd.surrogate)."""
    from ptarmigan_amd import surrogate, workloads
    from ptarmigan_amd.engine import Polarization
    model, meta = surrogate.load(surrogate.default_path(pol))
    assert meta["valid"] == 1
    engine.surrogate_load(model)
    laser, beam, opts = workloads.plane_wave_lma(100_000)
    laser.polarization = Polarization(pol)
    laser.a0 = a0
    say("surrogate shower", pol, a0)
    se, ce = _shower(engine, laser, beam, opts, 31)
    opts.surrogate = True
    ss, cs = _shower(engine, laser, beam, opts, 31)
    say("  exact    ", ce, "\n  surrogate", cs)
    assert cs["error_flags"] == 0
    assert cs["surrogate_events"] > 0.95 * cs["emissions"]
    assert cs["surrogate_events"] + cs["surrogate_fallbacks"] >= cs["emissions"]
    pe, ps = se["species"] == 2, ss["species"] == 2
    ye, ys = se["weight"][pe].sum(), ss["weight"][ps].sum()
    say(f"  yield exact {ye} surrogate {ys}")
    assert abs(ys - ye) / ye < 5e-3
    ee, es = se["u"][pe, 0], ss["u"][ps, 0]
    bins = np.linspace(0.0, max(ee.max(), es.max()), 201)
    p_e = _chi2_two_sample(ee, es, bins)
    say(f"  energy spectrum chi2 p = {p_e:.3g}; mean {ee.mean():.6g} vs {es.mean():.6g}")
    assert abs(es.mean() - ee.mean()) / ee.mean() < 0.01
    for comp in (1, 2):
        ae = np.arctan2(se["u"][pe, comp], -se["u"][pe, 3])
        as_ = np.arctan2(ss["u"][ps, comp], -ss["u"][ps, 3])
        lim = np.quantile(np.abs(ae), 0.999)
        p_a = _chi2_two_sample(ae, as_, np.linspace(-lim, lim, 201))
        core_e, core_s = ae[np.abs(ae) <= lim], as_[np.abs(as_) <= lim]
        # the wide-angle tail (soft photons at small s): counts beyond the exact 99.9% range
        te, ts = np.sum(np.abs(ae) > lim), np.sum(np.abs(as_) > lim)
        say(f"  angle {comp} chi2 p = {p_a:.3g}; core rms {core_e.std():.4g} vs {core_s.std():.4g}; tail {te} vs {ts}")
        # known limits of the shipped models (documented with the design): the core
        # width within 5%, and at most 0.2% of the photons moved into the soft,
        # wide-angle tail beyond the exact run's 99.9% range (up to 2-3x its count)
        assert abs(core_s.std() - core_e.std()) / core_e.std() < 0.05
        assert ts - te < 2e-3 * len(ae) + 5.0 * math.sqrt(te + ts + 1.0)
        assert p_a > 1e-4
    assert p_e > 1e-4
    assert np.all(np.isfinite(ss["pol"][ps])) and np.all(np.isfinite(ss["u"][ps]))
    for col in (1, 2, 3):
        x, y = se["pol"][pe, col], ss["pol"][ps, col]
        se_ = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-12
        assert abs(x.mean() - y.mean()) < 6 * se_, col


def test_surrogate_requires_a_valid_model_and_quantum_lma(engine):
    from ptarmigan_amd import workloads
    from ptarmigan_amd.engine import PtgError
    laser, beam, opts = workloads.plane_wave_lma(1000)
    opts.surrogate = True
    opts.classical = True
    engine.set_laser(laser)
    engine.build_beam(beam, seed=0)
    with pytest.raises(PtgError, match="surrogate"):
        engine.run(opts, seed=0)
    laser, beam, opts = workloads.focused_lcfa_cascade(1000)
    opts.surrogate = True
    engine.set_laser(laser)
    engine.build_beam(beam, seed=0)
    with pytest.raises(PtgError, match="surrogate"):
        engine.run(opts, seed=0)


def test_events_left_to_the_exact_sampler_are_exact(engine):
    """A model whose domain holds none of the shower's events leaves every emission
    to the exact kernels, through the deferred queue: the final state is the exact
    run's, particle for particle (every event draws from its own counters)."""
    from ptarmigan_amd import surrogate, workloads
    model, _ = surrogate.load(surrogate.default_path(0))
    model.a_min, model.a_max = 15.0, 20.0  # the a0 = 2 shower never reaches a = 15
    engine.surrogate_load(model)
    laser, beam, opts = workloads.plane_wave_lma(20_000)
    se, ce = _shower(engine, laser, beam, opts, 41)
    opts.surrogate = True
    ss, cs = _shower(engine, laser, beam, opts, 41)
    say("deferred: exact", ce["emissions"], "surrogate run", cs["emissions"], cs["surrogate_fallbacks"], cs["error_flags"])
    assert cs["surrogate_events"] == 0 and cs["surrogate_fallbacks"] == ce["emissions"]
    assert cs["error_flags"] == 0
    for sp in (0, 2):
        a = np.sort(se["u"][se["species"] == sp, 0])
        b = np.sort(ss["u"][ss["species"] == sp, 0])
        assert np.array_equal(a, b), sp
    engine.surrogate_load(surrogate.load(surrogate.default_path(0))[0])
