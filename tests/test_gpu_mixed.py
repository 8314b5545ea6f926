"""Mixed precision on the GPU against the FP64 GPU path (the FP64 path is the
oracle's GPU twin: its parity with the oracle is tested elsewhere).

Mixed precision keeps the phase, position and momentum updates, the mass shell
and the survival product in FP64 and runs the rate table lookup, the
plane-wave pusher algebra and the CDF table in FP32. The two runs use
independent seeds (beam and events), so they are independent samples: photon
yields per primary must agree within 5 standard errors, the photon energy and
angle spectra (200 bins, the reference's output binning) must be chi-square
consistent (p > 1e-3), and the Stokes means agree within 5 standard errors.
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


def _run(engine, laser, beam, opts, seed):
    engine.set_laser(laser)
    engine.build_beam(beam, seed=seed)
    c = engine.run(opts, seed=seed)
    sec, _ = engine.fetch(c)
    return sec, c


def _compare(sf, cf, sm, cm):
    say("  fp64 ", cf, "\n  mixed", cm)
    assert cm["error_flags"] == 0
    assert cm["lepton_steps"] == pytest.approx(cf["lepton_steps"], rel=1e-3)
    pf, pm = sf["species"] == 2, sm["species"] == 2
    yf, ym = sf["weight"][pf].sum(), sm["weight"][pm].sum()
    # equal primary counts and weights: the yields' ratio has a relative standard
    # error of about sqrt(1 / N_f + 1 / N_m) (photon counts nearly Poisson)
    sig = math.sqrt(1.0 / pf.sum() + 1.0 / pm.sum())
    say(f"  photon yield fp64 {yf:.6g} mixed {ym:.6g} rel {(ym - yf) / yf:.2e} (sigma {sig:.1e})")
    assert abs(ym - yf) / yf < 5.0 * sig
    ef, em = sf["u"][pf, 0], sm["u"][pm, 0]
    bins = np.linspace(0.0, max(ef.max(), em.max()), 201)
    assert _chi2_two_sample(ef, em, bins) > 1e-3
    for comp in (1, 2):
        af = np.arctan2(sf["u"][pf, comp], -sf["u"][pf, 3])
        am = np.arctan2(sm["u"][pm, comp], -sm["u"][pm, 3])
        lim = max(np.abs(af).max(), np.abs(am).max())
        assert _chi2_two_sample(af, am, np.linspace(-lim, lim, 201)) > 1e-3, comp
    for col in (1, 2, 3):
        x, y = sf["pol"][pf, col], sm["pol"][pm, col]
        se = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-12
        assert abs(x.mean() - y.mean()) < 5 * se, col
    lf, lm = sf["u"][sf["species"] == 0, 0], sm["u"][sm["species"] == 0, 0]
    assert abs(lm.mean() - lf.mean()) / lf.mean() < 1e-3


def test_mixed_precision_lp_plane_wave(engine):
    """LP plane wave, a0 = 2, 16-cycle cos^2, 10 GeV electrons with 1% spread
    (the plain-method configuration), 5e5 primaries."""
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(500_000)
    say("mixed precision LP")
    sf, cf = _run(engine, laser, beam, opts, 11)
    opts.precision = 1
    sm, cm = _run(engine, laser, beam, opts, 12)
    _compare(sf, cf, sm, cm)


def test_mixed_precision_cp_plane_wave(engine):
    """The same with circular polarisation and a0 = 5."""
    from ptarmigan_amd import workloads
    from ptarmigan_amd.engine import Polarization
    laser, beam, opts = workloads.plane_wave_lma(200_000)
    laser.polarization = Polarization.CIRCULAR
    laser.a0 = 5.0
    say("mixed precision CP")
    sf, cf = _run(engine, laser, beam, opts, 13)
    opts.precision = 1
    sm, cm = _run(engine, laser, beam, opts, 14)
    _compare(sf, cf, sm, cm)


def test_mixed_precision_unsupported_paths_fail_loudly(engine):
    """LCFA fields and classical emission have no mixed-precision kernels: the
    run is refused, not silently done in FP64."""
    from ptarmigan_amd import workloads
    from ptarmigan_amd.engine import PtgError
    laser, beam, opts = workloads.focused_lcfa_cascade(1000)
    opts.precision = 1
    engine.set_laser(laser)
    engine.build_beam(beam, seed=0)
    with pytest.raises(PtgError, match="mixed precision"):
        engine.run(opts, seed=0)
    laser, beam, opts = workloads.plane_wave_lma(1000)
    opts.precision = 1
    opts.classical = True
    engine.set_laser(laser)
    engine.build_beam(beam, seed=0)
    with pytest.raises(PtgError, match="mixed precision"):
        engine.run(opts, seed=0)
