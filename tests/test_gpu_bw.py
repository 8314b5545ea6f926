"""GPU parity of LP nonlinear Breit-Wheeler pair creation in the plane wave
against the CPU oracle.

Table lookups are compared value by value (total rates 1e-10 relative, harmonic
indices exactly but for ceil() at an integer); the sampler and the photon and
cascade showers statistically at matched sample size (two-sample KS /
chi-square at p > 1e-3, counts within 5 sigma).
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


def test_bw_rate_matches_oracle(engine, orc):
    from ptarmigan_amd import Fn
    say("bw rate")
    rng = np.random.default_rng(7)
    m = 4000
    a = np.exp(rng.uniform(np.log(0.03), np.log(24.0), m))
    eta = np.exp(rng.uniform(np.log(0.002), np.log(1.9), m))
    got = engine.eval(Fn.BW_RATE, np.stack([a, eta], 1), 3)
    for i in range(m):
        k, f = orc.bw_rate(a[i], eta[i])
        assert got[i, 2] == k, (a[i], eta[i])
        if k > 0:
            assert np.allclose(got[i, :2], f, rtol=1e-10, atol=0), (a[i], eta[i])


def test_bw_harmonic_matches_oracle(engine, orc):
    from ptarmigan_amd import Fn
    say("bw harmonic")
    rng = np.random.default_rng(8)
    m = 3000
    a = np.exp(rng.uniform(np.log(0.05), np.log(20.0), m))
    eta = np.exp(rng.uniform(np.log(0.05), np.log(1.5), m))
    sv1 = rng.uniform(-1, 1, m)
    frac = rng.uniform(size=m)
    keep = np.array([orc.bw_rate(a[i], eta[i])[0] > 0 for i in range(m)])
    a, eta, sv1, frac = a[keep], eta[keep], sv1[keep], frac[keep]
    got = engine.eval(Fn.BW_HARMONIC, np.stack([a, eta, sv1, frac], 1), 1)[:, 0]
    ref = np.array([orc.lib().orc_bw_harmonic(a[i], eta[i], sv1[i], frac[i]) for i in range(len(a))])
    assert len(a) > 1000
    assert (got == ref).mean() > 0.995
    assert np.all(np.abs(got - ref) <= 1)


@pytest.mark.parametrize("a,eta,sv1,sv2,fixed_n", [(1.0, 0.5, 0.0, 0.0, 8), (1.5, 0.4, 0.7, 0.2, 0), (6.0, 0.3, -0.5, 0.0, 0)])
def test_bw_sampler_statistics(engine, orc, a, eta, sv1, sv2, fixed_n):
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Fn
    # at a = 6, eta = 0.3 the harmonics run from 127 to ~660, about half from 200
    # up, where the wave shares each J(n, x, y) evaluation
    say("bw sample", a, eta, sv1, sv2, fixed_n)
    m = 20_000
    inp = np.stack([np.full(m, a), np.full(m, eta), np.full(m, sv1), np.full(m, sv2), np.arange(m, dtype=float),
                    np.full(m, float(fixed_n))], 1)
    got = engine.eval(Fn.BW_SAMPLE, inp, 4)
    ref = orc.bw_sample(a, eta, sv1, sv2, m, seed=5, fixed_n=fixed_n)
    assert np.all(got[:, 3] >= 0)
    if fixed_n:
        assert np.all(got[:, 0] == fixed_n)
    else:
        top = int(max(got[:, 0].max(), ref[:, 0].max()))
        lo = int(min(got[:, 0].min(), ref[:, 0].min()))
        assert _chi2_two_sample(got[:, 0], ref[:, 0], np.arange(lo - 0.5, top + 1.5)) > 1e-3
    for col in (1, 2):
        assert ks_2samp(got[:, col], ref[:, col]).pvalue > 1e-3, col


def _photon_beam(n, pol):
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, Species
    laser = Laser(LaserKind.PLANE_WAVE, 10.0, 0.8e-6, 16.0, envelope=Envelope.COS_SQUARED, polarization=Polarization.LINEAR)
    g = 16500.0 / 0.510999
    beam = Beam(n=n, gamma=g, sigma=0.01 * g, species=Species.PHOTON, rms_div=1e-6, pol=pol)
    return laser, beam


@pytest.mark.parametrize("pol_resolved", [False, True])
def test_photon_beam_pair_creation_matches_oracle(engine, orc, pol_resolved):
    """16.5 GeV photons through an a0 = 10 LP plane wave (pairs not tracked): pair
    yield, pair spectra, surviving photons and their Stokes parameters."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import CollideOptions
    laser, beam = _photon_beam(40_000, (1.0, 0.6, 0.0, 0.0))
    opts = CollideOptions(tracking_photons=True, tracking_secondaries=False, pol_resolved=pol_resolved)
    say("photon beam pairs", pol_resolved)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0, 4)
    sg, _, cg = engine.collide(prim, opts, seed=4)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 4, 16)
    say("  gpu", cg, "\n  cpu", cc)
    assert cg["error_flags"] == 0
    a, b = cg["pair_events"], cc["pair_events"]
    assert a > 500 and abs(a - b) < 5.0 * math.sqrt(a + b)
    assert cg["photon_steps"] == pytest.approx(cc["photon_steps"], rel=2e-3)
    for sp in (0, 1):
        eg, ec = sg["u"][sg["species"] == sp, 0], sc["u"][sc["species"] == sp, 0]
        assert len(eg) == a and len(ec) == b
        assert ks_2samp(eg, ec).pvalue > 1e-3
        pg = np.hypot(sg["u"][sg["species"] == sp, 1], sg["u"][sg["species"] == sp, 2])
        pc = np.hypot(sc["u"][sc["species"] == sp, 1], sc["u"][sc["species"] == sp, 2])
        assert ks_2samp(pg, pc).pvalue > 1e-3
    phg, phc = sg["species"] == 2, sc["species"] == 2
    assert abs(phg.sum() - phc.sum()) < 5.0 * math.sqrt(a + b)
    x, y = sg["pol"][phg, 1], sc["pol"][phc, 1]
    se = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-15
    assert abs(x.mean() - y.mean()) < 5 * se + 1e-12
    # electrons and positrons of one pair share position, weight and parent
    e = np.argsort(sg["parent_id"][sg["species"] == 0])
    p = np.argsort(sg["parent_id"][sg["species"] == 1])
    assert np.array_equal(sg["parent_id"][sg["species"] == 0][e], sg["parent_id"][sg["species"] == 1][p])


def test_plane_wave_cascade_matches_oracle(engine, orc):
    """16.5 GeV electrons, a0 = 10 LP plane wave, photons and secondaries tracked:
    the lepton and photon phases alternate until both are empty."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, Species
    laser = Laser(LaserKind.PLANE_WAVE, 10.0, 0.8e-6, 16.0, envelope=Envelope.COS_SQUARED, polarization=Polarization.LINEAR)
    g = 16500.0 / 0.510999
    beam = Beam(n=6000, gamma=g, sigma=0.01 * g, species=Species.ELECTRON, rms_div=1e-6)
    opts = CollideOptions(tracking_photons=True)
    say("plane-wave cascade")
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0, 6)
    sg, _, cg = engine.collide(prim, opts, seed=6)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 6, 16)
    say("  gpu", cg, "\n  cpu", cc)
    for key in ("emissions", "pair_events"):
        a, b = cg[key], cc[key]
        assert abs(a - b) < 5.0 * math.sqrt(a + b + 1.0), key
    for key in ("lepton_steps", "photon_steps"):
        assert cg[key] == pytest.approx(cc[key], rel=0.02), key
    for sp in (0, 2):
        eg, ec = sg["u"][sg["species"] == sp, 0], sc["u"][sc["species"] == sp, 0]
        assert ks_2samp(eg, ec).pvalue > 1e-3, sp
