"""GPU parity of circularly polarised nonlinear Compton scattering in the plane
wave (quantum and classical) against the CPU oracle.

Deterministic pieces value by value: J_{n-1,n,n+1} (1e-12 relative by recurrence,
1e-10 by the uniform expansion), total-rate lookups (1e-11 relative; 1e-8 on the
closed-form linear branch below a = 0.02, where two terms of order a^2/eta
cancel to order a^2 eta), partial rates and rejection ceilings (1e-10; 1e-8 from
n = 1000, where J comes from the uniform expansion), harmonic
indices (exactly but for ceil() at an integer). The sampler and the full shower
statistically at matched sample size (KS / chi-square at p > 1e-3).
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


def test_bessel_matches_oracle(engine, orc):
    from ptarmigan_amd import Fn
    say("bessel j_pm")
    rng = np.random.default_rng(11)
    n = np.concatenate([rng.integers(1, 1000, 3000), rng.integers(1000, 80000, 1000)]).astype(float)
    x = n * rng.uniform(0.05, 0.999, len(n))
    got = engine.eval(Fn.BESSEL_J, np.stack([n, x], 1), 3)
    for i in range(len(n)):
        ref = orc.bessel_j_pm(int(n[i]), x[i])
        tol = 1e-12 if n[i] < 1000 else 1e-10
        assert np.allclose(got[i], ref, rtol=tol, atol=1e-250), (n[i], x[i], got[i], ref)


@pytest.mark.parametrize("classical", [False, True])
def test_cp_rate_matches_oracle(engine, orc, classical):
    from ptarmigan_amd import Fn
    say("cp rate", classical)
    rng = np.random.default_rng(12)
    m = 4000
    a = np.exp(rng.uniform(np.log(0.005), np.log(25.0), m))
    eta = np.exp(rng.uniform(np.log(2e-4), np.log(2.5), m))
    got = engine.eval(Fn.CP_RATE, np.stack([a, eta, np.full(m, float(classical))], 1), 2)
    for i in range(m):
        ref = orc.cp_rate(a[i], eta[i], classical=classical)
        assert (got[i, 1] == 1.0) == (ref is not None), (a[i], eta[i])
        if ref is not None:
            linear = not classical and a[i] < 0.02 and eta[i] >= 1e-3
            assert got[i, 0] == pytest.approx(ref, rel=1e-8 if linear else 1e-11), (a[i], eta[i])


@pytest.mark.parametrize("classical", [False, True])
def test_cp_harmonic_matches_oracle(engine, orc, classical):
    from ptarmigan_amd import Fn
    say("cp harmonic", classical)
    rng = np.random.default_rng(13)
    m = 4000
    a = np.exp(rng.uniform(np.log(0.01), np.log(19.9), m))
    eta = np.exp(rng.uniform(np.log(5e-4), np.log(1.9), m))
    frac = rng.uniform(size=m)
    got = engine.eval(Fn.CP_HARMONIC, np.stack([a, eta, frac, np.full(m, float(classical))], 1), 1)[:, 0]
    L = orc.lib()
    ref = np.array([L.orc_cp_classical_harmonic(a[i], frac[i]) if classical else L.orc_cp_harmonic(a[i], eta[i], frac[i])
                    for i in range(m)])
    assert np.all(ref >= 1)
    assert (got == ref).mean() > 0.995
    assert np.all(np.abs(got - ref) <= 1)


def test_cp_partial_rates_match_oracle(engine, orc):
    """Partial rates and ceilings (quantum and classical) against the oracle, and
    the reference's partial-rate values (cp/mod.rs:309-346) at 1e-6 on the GPU."""
    from ptarmigan_amd import Fn
    say("cp partial rates")
    from test_oracle_cp import KATS
    rng = np.random.default_rng(14)
    m = 600
    a = np.exp(rng.uniform(np.log(0.05), np.log(20.0), m))
    eta = np.exp(rng.uniform(np.log(1e-3), np.log(1.5), m))
    n = np.maximum(1, (rng.uniform(size=m) * 10.0 * (1.0 + a ** 3))).astype(int).astype(float)
    L = orc.lib()
    for cl in (0.0, 1.0):
        got = engine.eval(Fn.CP_PARTIAL_RATE, np.stack([n, a, eta, np.full(m, cl)], 1), 2)
        for i in range(m):
            if cl:
                ref = (L.orc_cp_integrated_spectrum_low_eta(int(n[i]), a[i]), L.orc_cp_classical_ceiling(int(n[i]), a[i]))
            else:
                ref = (L.orc_cp_partial_rate(int(n[i]), a[i], eta[i]), L.orc_cp_ceiling(int(n[i]), a[i], eta[i]))
            # from n = 1000 the Bessel functions come from Debye's expansion: some 30 pow/exp/log
            # per value, where the device and host libm differ by an ulp or so each
            tol = 1e-10 if n[i] < 1000 else 1e-8
            assert np.allclose(got[i], ref, rtol=tol, atol=1e-300), (cl, n[i], a[i], eta[i], got[i], ref)
    k = np.array(KATS)
    got = engine.eval(Fn.CP_PARTIAL_RATE, np.stack([k[:, 0], k[:, 1], k[:, 2], np.zeros(len(k))], 1), 2)
    assert np.all(np.abs(got[:, 0] - k[:, 3]) / k[:, 3] < 1e-6)


@pytest.mark.parametrize("a,eta,fixed_n,classical", [(1.0, 0.2, 0, False), (3.0, 0.1, 0, False), (8.0, 0.05, 0, False),
                                                     (12.0, 0.05, 3000, False), (2.0, 0.05, 0, True), (6.0, 0.05, 0, True)])
def test_cp_sampler_statistics(engine, orc, a, eta, fixed_n, classical):
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Fn
    say("cp sample", a, eta, fixed_n, classical)
    m = 20_000
    inp = np.stack([np.full(m, a), np.full(m, eta), np.arange(m, dtype=float), np.full(m, float(fixed_n)),
                    np.full(m, float(classical))], 1)
    got = engine.eval(Fn.CP_SAMPLE, inp, 7)
    ref = orc.cp_sample(a, eta, m, classical=classical, seed=3, fixed_n=fixed_n)
    assert np.all(got[:, 6] >= 1)
    if fixed_n:
        assert np.all(got[:, 0] == fixed_n)
    else:
        top = int(max(got[:, 0].max(), ref[:, 0].max()))
        assert _chi2_two_sample(got[:, 0], ref[:, 0], np.unique(np.geomspace(1, top + 1, 60).astype(int)) - 0.5) > 1e-3
    for col in (1, 2, 3, 5):
        assert ks_2samp(got[:, col], ref[:, col]).pvalue > 1e-3, col


@pytest.mark.parametrize("classical", [False, True])
def test_cp_plane_wave_shower_matches_oracle(engine, orc, classical):
    """CP plane wave, a0 = 2, 16-cycle cos^2, 10 GeV electrons (quantum: recoil and
    absorption; classical: LL radiation reaction, no recoil): photon yield,
    photon energy and angle spectra, Stokes means, final electron energies."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Polarization, workloads
    laser, beam, opts = workloads.plane_wave_lma(30_000)
    laser.polarization = Polarization.CIRCULAR
    opts.classical = classical
    say("cp shower", classical)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0 + 3.0 * beam.length, 7)
    sg, _, cg = engine.collide(prim, opts, seed=7)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 7, 16)
    say("  gpu", cg, "\n  cpu", cc)
    assert cg["error_flags"] == 0
    assert cg["lepton_steps"] == pytest.approx(cc["lepton_steps"], rel=1e-3)
    a, b = cg["emissions"], cc["emissions"]
    assert a > 1000 and abs(a - b) < 5.0 * math.sqrt(a + b + 1.0)
    ph_g, ph_c = sg["species"] == 2, sc["species"] == 2
    for q in (lambda s, m: s["u"][m, 0], lambda s, m: np.arctan2(s["u"][m, 1], -s["u"][m, 3]),
              lambda s, m: np.arctan2(s["u"][m, 2], -s["u"][m, 3]), lambda s, m: s["payload"][m]):
        assert ks_2samp(q(sg, ph_g), q(sc, ph_c)).pvalue > 1e-3
    for col in (1, 2, 3):
        x, y = sg["pol"][ph_g, col], sc["pol"][ph_c, col]
        se = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-12
        assert abs(x.mean() - y.mean()) < 5 * se, col
    eg, ec = sg["u"][sg["species"] == 0, 0], sc["u"][sc["species"] == 0, 0]
    assert ks_2samp(eg, ec).pvalue > 1e-3
    wg, wc = sg["work"][sg["species"] == 0], sc["work"][sc["species"] == 0]
    assert ks_2samp(wg, wc).pvalue > 1e-3


# ---------------------------------------------------------------- CP nonlinear Breit-Wheeler
def test_cpbw_rates_match_oracle(engine, orc):
    """Totals (table, explicit sum below a = 0.3, zero where negligible) and partial
    rates against the oracle; the reference's partial-rate values at 1e-6."""
    from ptarmigan_amd import Fn
    from test_oracle_cp import CPBW_KATS
    say("cpbw rates")
    rng = np.random.default_rng(31)
    m = 2000
    a = np.exp(rng.uniform(np.log(0.01), np.log(20.0), m))
    eta = np.exp(rng.uniform(np.log(0.01), np.log(1.9), m))
    keep = ~((a < 0.3) & (2.0 * (0.25 + a * a) / eta > 2000))
    a, eta = a[keep], eta[keep]
    got = engine.eval(Fn.CPBW_RATE, np.stack([a, eta], 1), 3)
    for i in range(len(a)):
        k, f = orc.cpbw_total(a[i], eta[i])
        assert got[i, 2] == k, (a[i], eta[i])
        assert np.allclose(got[i, :2], f, rtol=1e-9 if k == 1 else 1e-11, atol=0), (a[i], eta[i], got[i], f)
    n = np.array([k[0] for k in CPBW_KATS], dtype=float)
    kat = np.array(CPBW_KATS)
    got = engine.eval(Fn.CPBW_PARTIAL_RATE, kat[:, :3], 2)
    assert np.all(np.abs(0.5 * got.sum(1) - kat[:, 3]) / kat[:, 3] < 1e-6)
    for i in range(len(n)):
        ref = orc.cpbw_partial_rate(int(n[i]), kat[i, 1], kat[i, 2])
        assert np.allclose(got[i], ref, rtol=1e-10 if n[i] < 1000 else 1e-8, atol=0), i


@pytest.mark.parametrize("a,eta,sv", [(1.0, 0.5, (0.0, 0.0, 0.0)), (3.0, 0.2, (0.5, -0.3, 0.7)), (0.2, 1.0, (0.0, 0.5, -0.5))])
def test_cpbw_sampler_matches_oracle(engine, orc, a, eta, sv):
    """Harmonic indices value by value on common fractions, then (n, s, theta)
    statistically."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Fn
    say("cpbw sample", a, eta, sv)
    frac = (np.arange(1000) + 0.5) / 1000
    got = engine.eval(Fn.CPBW_HARMONIC, np.stack([np.full(1000, a), np.full(1000, eta), np.full(1000, sv[2]), frac], 1), 2)
    ref = np.array([orc.lib().orc_cpbw_harmonic(a, eta, sv[2], f) for f in frac])
    assert (got[:, 0] == ref).mean() > 0.99 and np.all(np.abs(got[:, 0] - ref) <= 1)
    m = 6000
    inp = np.stack([np.full(m, a), np.full(m, eta), np.full(m, sv[0]), np.full(m, sv[1]), np.full(m, sv[2]),
                    np.arange(m, dtype=float), np.zeros(m)], 1)
    got = engine.eval(Fn.CPBW_SAMPLE, inp, 4)
    ref = orc.cpbw_sample(a, eta, *sv, m, seed=8)
    assert np.all(got[:, 3] >= 1)
    lo, hi = int(min(got[:, 0].min(), ref[:, 0].min())), int(max(got[:, 0].max(), ref[:, 0].max()))
    edges = np.unique(np.linspace(lo, hi + 1, 40).astype(int)) - 0.5
    assert _chi2_two_sample(got[:, 0], ref[:, 0], edges) > 1e-3
    for col in (1, 2):
        assert ks_2samp(got[:, col], ref[:, col]).pvalue > 1e-3, col


def test_cp_photon_beam_pair_creation_matches_oracle(engine, orc):
    """16.5 GeV photons, partly circularly polarised (S_3 = 0.5), through an a0 = 8
    CP plane wave (pairs not tracked): pair yield, pair spectra, surviving photons
    and their S_3."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, Species
    laser = Laser(LaserKind.PLANE_WAVE, 8.0, 0.8e-6, 16.0, envelope=Envelope.COS_SQUARED, polarization=Polarization.CIRCULAR)
    g = 16500.0 / 0.510999
    beam = Beam(n=20_000, gamma=g, sigma=0.01 * g, species=Species.PHOTON, rms_div=1e-6, pol=(1.0, 0.0, 0.0, 0.5))
    opts = CollideOptions(tracking_photons=True, tracking_secondaries=False, pol_resolved=True)
    say("cp photon beam pairs")
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0, 12)
    sg, _, cg = engine.collide(prim, opts, seed=12)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 12, 16)
    say("  gpu", cg, "\n  cpu", cc)
    assert cg["error_flags"] == 0
    a, b = cg["pair_events"], cc["pair_events"]
    assert a > 500 and abs(a - b) < 5.0 * math.sqrt(a + b)
    assert cg["photon_steps"] == pytest.approx(cc["photon_steps"], rel=5e-3)
    for sp in (0, 1):
        eg, ec = sg["u"][sg["species"] == sp, 0], sc["u"][sc["species"] == sp, 0]
        assert ks_2samp(eg, ec).pvalue > 1e-3
    phg, phc = sg["species"] == 2, sc["species"] == 2
    x, y = sg["pol"][phg, 3], sc["pol"][phc, 3]
    se = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-15
    assert abs(x.mean() - y.mean()) < 5 * se + 1e-12
