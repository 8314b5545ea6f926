"""GPU parity of the locally-monochromatic (plane-wave) path against the CPU oracle.

Deterministic pieces are compared value by value: the LP rate table lookup and
harmonic CDF inversion (1e-12 relative; harmonic indices exactly), J(n, x, y)
by the kernels' saddle point and quadrature and by the reference's recurrence
(1e-11 of the largest |J| in the batch),
field-only plane-wave trajectories (1e-12 relative). The sampler and the full
shower are compared statistically at matched sample size (two-sample KS /
chi-square at p > 1e-3): the RNG streams differ (xoshiro256** on the CPU,
Philox per particle on the GPU).
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C_LIGHT = 2.997925e8


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


def test_lp_rate_matches_oracle(engine, orc):
    from ptarmigan_amd import Fn
    say("lp rate")
    rng = np.random.default_rng(3)
    a = np.exp(rng.uniform(np.log(0.005), np.log(30.0), 3000))
    eta = np.exp(rng.uniform(np.log(1e-4), np.log(5.0), 3000))
    got = engine.eval(Fn.LP_RATE, np.stack([a, eta], 1), 2)
    for i in range(len(a)):
        ref = orc.lp_rate(a[i], eta[i])
        if ref is None:
            assert got[i, 1] == 0.0
        else:
            assert got[i, 1] == 1.0
            # 1e-9: below the table the closed form of lp/linear.rs cancels terms of
            # order 4/eta, so one ulp of log() between libraries grows to ~1e-11
            assert abs(got[i, 0] - ref) <= 1e-9 * abs(ref)


def test_lp_harmonic_matches_oracle(engine, orc):
    from ptarmigan_amd import Fn
    say("lp harmonic")
    rng = np.random.default_rng(4)
    m = 4000
    a = np.exp(rng.uniform(np.log(0.01), np.log(20.0), m))
    eta = np.exp(rng.uniform(np.log(1e-4), np.log(1.5), m))
    frac = rng.uniform(size=m)
    got = engine.eval(Fn.LP_HARMONIC, np.stack([a, eta, frac], 1), 1)[:, 0]
    ref = np.array([orc.lp_harmonic(a[i], eta[i], frac[i]) for i in range(m)])
    # ceil() of an interpolated real: an ulp-level difference may move a value
    # sitting on an integer by one
    same = got == ref
    assert same.mean() > 0.999
    assert np.all(np.abs(got - ref) <= 1)


@pytest.mark.parametrize("n", [3, 8, 12, 25, 100, 1000])
def test_double_bessel_matches_oracle(engine, orc, n):
    from ptarmigan_amd import Fn
    say("J(n, x, y)", n)
    rng = np.random.default_rng(n)
    m = 256
    x = n * math.sqrt(2.0) * rng.uniform(size=m)
    y = 0.5 * n * rng.uniform(size=m)
    for rec in (1.0, 0.0):
        inp = np.stack([np.full(m, n), x, y, np.full(m, rec)], 1)
        got = engine.eval(Fn.DOUBLE_BESSEL, inp, 5)
        ref = np.array([orc.double_bessel(n, x[i], y[i], recurrence=bool(rec)) for i in range(m)])
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-11 * scale, rec


@pytest.mark.parametrize("n", [3, 7, 8, 25, 1000, 2500, 10000])
def test_double_bessel_quadrature_matches_recurrence(engine, orc, n):
    """The kernels' trapezoid quadrature of J(n - 2 .. n + 2, x, y) against the
    reference's recurrence (double_bessel.rs:142-246), both on the device, over
    the whole domain |x| <= n sqrt 2, 0 <= y <= n / 2 (the saddle-point region
    included), and against the oracle's recurrence (built without FMA contraction,
    as the reference): within 1e-12 absolute (|J| <= 1). The quadrature's own
    rounding floor is ~1e-15 times the square root of its ~n nodes; the rest is
    the recurrences' rounding (the device's contracts a * b + c into FMAs)."""
    from ptarmigan_amd import Fn
    say("J quadrature", n)
    rng = np.random.default_rng(100 + n)
    m = 200 if n < 5000 else 40
    x = n * math.sqrt(2.0) * rng.uniform(size=m)
    y = 0.5 * n * rng.uniform(size=m)
    rec = engine.eval(Fn.DOUBLE_BESSEL, np.stack([np.full(m, n), x, y, np.full(m, 1.0)], 1), 5)
    quad = engine.eval(Fn.DOUBLE_BESSEL, np.stack([np.full(m, n), x, y, np.full(m, 2.0)], 1), 5)
    err = np.abs(quad - rec).max()
    # negative x (theta beyond pi / 2 in the rates): J(m, -x, y) = (-1)^m J(m, x, y)
    neg = engine.eval(Fn.DOUBLE_BESSEL, np.stack([np.full(m, n), -x, y, np.full(m, 2.0)], 1), 5)
    sign = (-1.0) ** (n - 2 + np.arange(5))
    err_neg = np.abs(neg - sign * rec).max()
    k = np.arange(0, m, 4)
    ora = np.array([orc.double_bessel(n, x[i], y[i], recurrence=True) for i in k])
    err_orc = np.abs(quad[k] - ora).max()
    # the quadrature shared by a wave (harmonics from 200 up in the LP event kernels):
    # the same nodes summed in 64 partial sums, every lane holding the same bits
    # (NaN otherwise); it differs from one lane's sum by rounding only
    wave = engine.eval(Fn.DOUBLE_BESSEL, np.stack([np.full(m, n), np.concatenate([x[: m // 2], -x[m // 2:]]), y,
                                                   np.full(m, 4.0)], 1), 5)
    lane = np.concatenate([quad[: m // 2], neg[m // 2:]])
    err_wave = np.abs(wave - lane).max()
    say("  max |dJ|", err, "negative x", err_neg, "oracle", err_orc, "wave-shared vs one lane", err_wave)
    assert err < 1e-12 and err_neg < 1e-12 and err_orc < 1e-12
    assert np.all(np.isfinite(wave)) and err_wave < 1e-13


def test_double_bessel_reference_kat_on_gpu(engine):
    """double_bessel.rs:417-597 known answers through the device recurrence (1e-9
    relative) and through the kernels' evaluation (1e-9 relative where |J| > 1e-6;
    the 1.9e-31 value lies below the quadrature's 1e-15 absolute floor)."""
    from ptarmigan_amd import Fn
    pts = [(10, 0.5 * 10 * math.sqrt(2), 0.5 * 5, 0.027510433345552824743),
           (100, 0.9375 * 100 * math.sqrt(2), 0.5 * 50, -0.024581413776829650737),
           (1000, 1000.0, 500.0, 5.6390291047140394025e-3),
           (1000, 1100.0, 250.0, 1.9079405288254989534e-31),
           (3000, 3500.0, 1500.0, 9.5390642676650740986e-3)]
    for n, x, y, t in pts:
        got = engine.eval(Fn.DOUBLE_BESSEL, [[n, x, y, 1.0]], 5)[0, 2]
        assert abs((got - t) / t) < 1e-9, (n, x, y)
        got = engine.eval(Fn.DOUBLE_BESSEL, [[n, x, y, 0.0]], 5)[0, 2]
        wave = engine.eval(Fn.DOUBLE_BESSEL, [[n, x, y, 4.0]], 5)[0, 2]
        assert abs(wave - t) < 1e-13 + 1e-9 * abs(t), (n, x, y)
        if abs(t) > 1e-6:
            assert abs((got - t) / t) < 1e-9, (n, x, y)
        else:
            assert abs(got) < 1e-13, (n, x, y)


@pytest.mark.parametrize("a,eta,fixed_n", [(2.0, 0.1, 0), (1.0, 0.1, 1), (2.0, 0.3, 8), (5.0, 0.05, 0), (10.0, 0.1, 300)])
def test_lp_sampler_statistics(engine, orc, a, eta, fixed_n):
    """(n, s, theta, S1, S2) from the device sampler against the oracle's
    (at n = 300 the wave shares each J(n, x, y) evaluation)."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Fn
    say("lp sample", a, eta, fixed_n)
    m = 20_000 if fixed_n < 200 else 5_000
    inp = np.stack([np.full(m, a), np.full(m, eta), np.arange(m, dtype=float), np.full(m, float(fixed_n))], 1)
    got = engine.eval(Fn.LP_SAMPLE, inp, 6)
    ref = orc.lp_sample(a, eta, m, seed=17, fixed_n=fixed_n)
    if fixed_n:
        assert np.all(got[:, 0] == fixed_n)
    else:
        top = int(max(got[:, 0].max(), ref[:, 0].max()))
        assert _chi2_two_sample(got[:, 0], ref[:, 0], np.arange(0.5, top + 1.5)) > 1e-3
    for col in (1, 2, 3, 4):
        assert ks_2samp(got[:, col], ref[:, col]).pvalue > 1e-3, col
    assert np.all(got[:, 3] ** 2 + got[:, 4] ** 2 <= 1.0 + 1e-12)


def _pw_lasers():
    from ptarmigan_amd import Envelope, Laser, LaserKind, Polarization
    return {
        "pw_lp_cos2": Laser(LaserKind.PLANE_WAVE, 2.0, 0.8e-6, 16.0, envelope=Envelope.COS_SQUARED,
                            polarization=Polarization.LINEAR),
        "pw_cp_flat": Laser(LaserKind.PLANE_WAVE, 100.0, 0.8e-6, 8.0, envelope=Envelope.FLATTOP,
                            polarization=Polarization.CIRCULAR),
        "pw_lp_gauss": Laser(LaserKind.PLANE_WAVE, 20.0, 0.8e-6, 8.0, envelope=Envelope.GAUSSIAN,
                             polarization=Polarization.LINEAR),
    }


@pytest.mark.parametrize("name", list(_pw_lasers()))
@pytest.mark.parametrize("rr", [False, True])
def test_plane_wave_trajectories(engine, orc, name, rr):
    """Field::push in the LMA plane wave (plane_wave.rs:196-249), Lorentz or
    Landau-Lifshitz, through the whole pulse."""
    from ptarmigan_amd import CollideOptions
    laser = _pw_lasers()[name]
    say("plane-wave push", name, rr)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    rng = np.random.default_rng(7)
    n = 256
    gam = rng.uniform(100.0, 2e4, n)
    u = np.stack([gam, rng.normal(0, 1.0, n), rng.normal(0, 1.0, n), -np.sqrt(gam**2 - 1.0)], 1)
    u[:, 0] = np.sqrt(1.0 + (u[:, 1:] ** 2).sum(1))
    r = np.stack([-z0 * np.ones(n), np.zeros(n), np.zeros(n), z0 * np.ones(n)], 1)
    sp = np.zeros(n, dtype=np.uint8)
    sp[::3] = 1
    opts = CollideOptions(classical=True, rr=rr)
    rg, ug, kg, wg = engine.push(sp, r, u, 200_000, opts, with_work=True)
    rc, uc, kc, wc = orc.push(laser.desc(), sp, r, u, 200_000, eqn=1 if rr else 0, with_work=True)
    assert np.array_equal(kg, kc)
    assert np.abs(ug - uc).max() / np.abs(uc).max() < 1e-12
    assert np.abs(rg - rc).max() / np.abs(rc).max() < 1e-12
    # work (units of m c^2) to the same 1e-12 of gamma as the momenta: without
    # radiation reaction it cancels to ~1e-9 over the pulse, pure rounding
    assert np.abs(wg - wc).max() <= 1e-12 * np.abs(uc[:, 0]).max()


def test_plane_wave_ll_depletion_on_gpu(engine):
    """plane_wave.rs:381-417 known answer, through the device pusher."""
    from ptarmigan_amd import CollideOptions, Envelope, Laser, LaserKind, Polarization
    n_cycles, a0 = 8.0, 20.0
    laser = Laser(LaserKind.PLANE_WAVE, a0, 0.8e-6, n_cycles, envelope=Envelope.GAUSSIAN,
                  polarization=Polarization.LINEAR)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    u = np.array([[math.sqrt(1.0 + 1e6), 0.0, 0.0, -1000.0]])
    r = np.array([[-z0, 0.0, 0.0, z0]])
    opts = CollideOptions(classical=True, rr=True, dt_multiplier=0.5)
    _, _, _, work = engine.push([0], r, u, 10_000_000, opts, with_work=True)
    alpha, tau_c = 7.29735257e-3, 1.28808867e-21
    omega_mc2 = C_LIGHT * tau_c * 2 * math.pi / 0.8e-6
    phase = math.pi**1.5 * n_cycles / math.sqrt(math.log(4.0))
    expected = alpha * a0**4 * omega_mc2 * (1.0 / 8.0) * phase / 3.0
    assert abs(work[0] - expected) / expected < 1e-3


@pytest.mark.parametrize("a0,n", [(2.0, 40_000), (12.0, 4_000)])
def test_lma_shower_statistics_match_oracle(engine, orc, a0, n):
    """LP plane wave, 16-cycle cos^2, 10 GeV electrons: photon yield, photon
    energy and angle spectra, harmonics, Stokes means and final electron energies
    agree with the oracle at matched sample size. At a0 = 2 the harmonics stay
    below 40 (the wave's lanes take a candidate each); at a0 = 12 about one
    emission in ten has a harmonic from 200 up (to ~2000), where the wave shares
    each J(n, x, y) evaluation."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(n)
    laser.a0 = a0
    say("lma shower statistics", a0, n)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0 + 3.0 * beam.length, 5)
    sg, _, cg = engine.collide(prim, opts, seed=5)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 5, 16)
    say("  gpu", cg, "\n  cpu", cc)
    assert cg["lepton_steps"] == pytest.approx(cc["lepton_steps"], rel=1e-3)
    a, b = cg["emissions"], cc["emissions"]
    assert abs(a - b) < 5.0 * math.sqrt(a + b + 1.0)
    ph_g, ph_c = sg["species"] == 2, sc["species"] == 2
    assert ph_g.sum() == a and ph_c.sum() == b
    for q in (lambda s, m: s["u"][m, 0], lambda s, m: np.arctan2(s["u"][m, 1], -s["u"][m, 3]),
              lambda s, m: np.arctan2(s["u"][m, 2], -s["u"][m, 3]), lambda s, m: s["payload"][m],
              lambda s, m: s["parent_chi"][m]):
        assert ks_2samp(q(sg, ph_g), q(sc, ph_c)).pvalue > 1e-3
    for col in (1, 2, 3):
        x, y = sg["pol"][ph_g, col], sc["pol"][ph_c, col]
        se = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-12
        assert abs(x.mean() - y.mean()) < 5 * se, col
    eg, ec = sg["u"][sg["species"] == 0, 0], sc["u"][sc["species"] == 0, 0]
    assert len(eg) == len(ec) == beam.n
    assert ks_2samp(eg, ec).pvalue > 1e-3


# ---------------------------------------------------------------- LP, classical limit (lp/classical.rs)
def test_lp_classical_rate_and_harmonic_match_oracle(engine, orc):
    from ptarmigan_amd import Fn
    say("lp classical rate, harmonic")
    rng = np.random.default_rng(21)
    m = 3000
    a = np.exp(rng.uniform(np.log(0.005), np.log(25.0), m))
    eta = np.exp(rng.uniform(np.log(1e-4), np.log(1.0), m))
    got = engine.eval(Fn.LP_CL_RATE, np.stack([a, eta], 1), 2)
    for i in range(m):
        ref = orc.lp_cl_rate(a[i], eta[i])
        assert (got[i, 1] == 1.0) == (ref is not None), a[i]
        if ref is not None:
            assert got[i, 0] == pytest.approx(ref, rel=1e-12), a[i]
    a = np.exp(rng.uniform(np.log(0.01), np.log(19.9), m))
    frac = rng.uniform(size=m)
    got = engine.eval(Fn.LP_CL_HARMONIC, np.stack([a, frac], 1), 1)[:, 0]
    ref = np.array([orc.lib().orc_lp_cl_harmonic(a[i], frac[i]) for i in range(m)])
    assert (got == ref).mean() > 0.999
    assert np.all(np.abs(got - ref) <= 1)


def test_lp_classical_ceiling_matches_oracle(engine, orc):
    """lp/classical.rs:43-110. The rate is (-g0^2 - a^2 (g0 g2 - g1^2)) / 2 pi in
    J(n - 2 .. n + 2); the device's J (quadrature) and the oracle's (recurrence)
    differ by up to ~5e-13 absolute, which moves the rate by up to about
    (1 + a^2) sqrt(rate) 1e-12: the bound is 1e-10 relative plus that."""
    from ptarmigan_amd import Fn
    say("lp classical ceiling")
    rng = np.random.default_rng(22)
    m = 400
    a = np.exp(rng.uniform(np.log(0.1), np.log(20.0), m))
    n = np.maximum(1, rng.uniform(size=m) * (6.0 * (1.0 + a * a) + 2.0 * a ** 3)).astype(int).astype(float)
    got = engine.eval(Fn.LP_CL_CEILING, np.stack([n, a], 1), 1)[:, 0]
    ref = np.array([orc.lib().orc_lp_cl_ceiling(int(n[i]), a[i]) for i in range(m)])
    bound = 1e-10 * np.abs(ref) + 1e-12 * (1.0 + a * a) * np.sqrt(np.abs(ref))
    say("  max |d| / bound", (np.abs(got - ref) / bound).max())
    assert np.all(np.abs(got - ref) <= bound)


@pytest.mark.parametrize("a,eta,fixed_n", [(1.0, 0.05, 0), (3.0, 0.05, 0), (8.0, 0.02, 0), (3.0, 0.05, 6), (8.0, 0.02, 300)])
def test_lp_classical_sampler_statistics(engine, orc, a, eta, fixed_n):
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Fn
    say("lp classical sample", a, eta, fixed_n)
    m = 20_000 if fixed_n < 200 else 5_000
    inp = np.stack([np.full(m, a), np.full(m, eta), np.arange(m, dtype=float), np.full(m, float(fixed_n))], 1)
    got = engine.eval(Fn.LP_CL_SAMPLE, inp, 6)
    ref = orc.lp_cl_sample(a, eta, m, seed=23, fixed_n=fixed_n)
    assert np.all(got[:, 5] >= 1)
    if fixed_n:
        assert np.all(got[:, 0] == fixed_n)
    else:
        top = int(max(got[:, 0].max(), ref[:, 0].max()))
        assert _chi2_two_sample(got[:, 0], ref[:, 0], np.unique(np.geomspace(1, top + 1, 60).astype(int)) - 0.5) > 1e-3
    for col in (1, 2):
        assert ks_2samp(got[:, col], ref[:, col]).pvalue > 1e-3, col
    # The reference evaluates J(n, x, y) at the negative x of theta near pi, where its
    # saddle-point branch returns NaN: those photons carry NaN Stokes parameters, in
    # the same proportion here as in the oracle; the rest are compared by KS.
    bad_g, bad_c = ~np.isfinite(got[:, 3]), ~np.isfinite(ref[:, 3])
    pg, pc = bad_g.mean(), bad_c.mean()
    assert abs(pg - pc) <= 5.0 * math.sqrt((pg + pc) / m) + 1e-12
    for col in (3, 4):
        assert ks_2samp(got[~bad_g, col], ref[~bad_c, col]).pvalue > 1e-3, col


def test_lp_classical_shower_matches_oracle(engine, orc):
    """LP plane wave, a0 = 2, 16-cycle cos^2, 10 GeV electrons, classical emission
    with Landau-Lifshitz radiation reaction and no recoil."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(30_000)
    opts.classical = True
    say("lp classical shower")
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0 + 3.0 * beam.length, 9)
    sg, _, cg = engine.collide(prim, opts, seed=9)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 9, 16)
    say("  gpu", cg, "\n  cpu", cc)
    assert cg["error_flags"] == 0
    a, b = cg["emissions"], cc["emissions"]
    assert a > 1000 and abs(a - b) < 5.0 * math.sqrt(a + b + 1.0)
    ph_g, ph_c = sg["species"] == 2, sc["species"] == 2
    for q in (lambda s, m: s["u"][m, 0], lambda s, m: np.arctan2(s["u"][m, 1], -s["u"][m, 3]),
              lambda s, m: np.arctan2(s["u"][m, 2], -s["u"][m, 3])):
        assert ks_2samp(q(sg, ph_g), q(sc, ph_c)).pvalue > 1e-3
    for col in (1, 2):
        x, y = sg["pol"][ph_g, col], sc["pol"][ph_c, col]
        x, y = x[np.isfinite(x)], y[np.isfinite(y)]  # NaN Stokes: see test_lp_classical_sampler_statistics
        se = math.sqrt(x.var() / len(x) + y.var() / len(y)) + 1e-12
        assert abs(x.mean() - y.mean()) < 5 * se, col
    eg, ec = sg["u"][sg["species"] == 0, 0], sc["u"][sc["species"] == 0, 0]
    assert ks_2samp(eg, ec).pvalue > 1e-3


# ---------------------------------------------------------------- FocusedLaser (focused_laser.rs)
def _fl_lasers():
    from ptarmigan_amd import Envelope, Laser, LaserKind, Polarization
    return {
        "fl_lp_cos2": Laser(LaserKind.FOCUSED_LASER, 2.0, 0.8e-6, 16.0, waist=4e-6, envelope=Envelope.COS_SQUARED,
                            polarization=Polarization.LINEAR),
        "fl_cp_flat": Laser(LaserKind.FOCUSED_LASER, 50.0, 0.8e-6, 8.0, waist=3e-6, envelope=Envelope.FLATTOP,
                            polarization=Polarization.CIRCULAR),
        "fl_lp_gauss": Laser(LaserKind.FOCUSED_LASER, 20.0, 0.8e-6, 8.0, waist=2e-6, envelope=Envelope.GAUSSIAN,
                             polarization=Polarization.LINEAR),
    }


@pytest.mark.parametrize("name", list(_fl_lasers()))
@pytest.mark.parametrize("rr", [False, True])
def test_focused_laser_trajectories(engine, orc, name, rr):
    """Field::push in the focused LMA field (focused_laser.rs:185-225) with the
    transverse gradient, Lorentz or Landau-Lifshitz, through the whole pulse."""
    from ptarmigan_amd import CollideOptions
    laser = _fl_lasers()[name]
    say("focused push", name, rr)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    rng = np.random.default_rng(17)
    n = 256
    gam = rng.uniform(100.0, 2e4, n)
    u = np.stack([gam, rng.normal(0, 1.0, n), rng.normal(0, 1.0, n), -np.sqrt(gam**2 - 1.0)], 1)
    u[:, 0] = np.sqrt(1.0 + (u[:, 1:] ** 2).sum(1))
    r = np.stack([-z0 * np.ones(n), rng.normal(0, 2e-6, n), rng.normal(0, 2e-6, n), z0 * np.ones(n)], 1)
    sp = np.zeros(n, dtype=np.uint8)
    sp[::3] = 1
    opts = CollideOptions(classical=True, rr=rr)
    rg, ug, kg, wg = engine.push(sp, r, u, 200_000, opts, with_work=True)
    rc, uc, kc, wc = orc.push(laser.desc(), sp, r, u, 200_000, eqn=1 if rr else 0, with_work=True)
    assert np.array_equal(kg, kc)
    assert np.abs(ug - uc).max() / np.abs(uc).max() < 1e-12
    assert np.abs(rg - rc).max() / np.abs(rc).max() < 1e-12
    assert np.abs(wg - wc).max() <= 1e-12 * np.abs(uc[:, 0]).max()


def test_focused_laser_deflection_on_gpu(engine):
    """focused_laser.rs:353-380 (deflection) through the device pusher, error < 0.5%."""
    from ptarmigan_amd import CollideOptions, Envelope, Laser, LaserKind, Polarization
    a0, w0, lam, gamma = 100.0, 4.0e-6, 0.8e-6, 1000.0
    n_cycles = C_LIGHT * 30.0e-15 / lam
    laser = Laser(LaserKind.FOCUSED_LASER, a0, lam, n_cycles, waist=w0, envelope=Envelope.GAUSSIAN,
                  polarization=Polarization.CIRCULAR)
    engine.set_laser(laser)
    omega = C_LIGHT * 2.0 * math.pi / lam
    t_start = -20.0 * lam / C_LIGHT
    b = np.arange(1, 8) * 1.0e-6
    u = np.tile([gamma, 0.0, 0.0, -math.sqrt(gamma * gamma - 1.0)], (7, 1))
    r = np.stack([np.zeros(7), b, np.zeros(7), np.zeros(7)], 1) + u * C_LIGHT * t_start / gamma
    opts = CollideOptions(classical=True, rr=False, dt_multiplier=0.25 * lam / C_LIGHT * omega)
    _, u2, _ = engine.push(np.zeros(7, dtype=np.uint8), r, u, 10_000_000, opts)
    theta = 1000.0 * np.arctan2(u2[:, 1], -u2[:, 3])
    i_phi = omega * 30.0e-15 * math.sqrt(math.pi / (16.0 * math.log(2.0)))
    theory = 1000.0 * 2.0 * a0 * a0 * i_phi * b * lam * np.exp(-2.0 * b * b / (w0 * w0)) / (2.0 * math.pi * gamma**2 * w0 * w0)
    assert np.all(np.abs(theta - theory) / theory < 5.0e-3)


@pytest.mark.parametrize("pol,classical", [("linear", False), ("circular", False), ("linear", True)])
def test_focused_laser_shower_matches_oracle(engine, orc, pol, classical):
    """FocusedLaser, a0 = 5, w0 = 3 um, 16-cycle cos^2, 10 GeV electrons with a 2 um
    Gaussian spot (photons final state): emission count, photon spectra and
    angles, electron energies, against the oracle."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, Species
    laser = Laser(LaserKind.FOCUSED_LASER, 5.0, 0.8e-6, 16.0, waist=3e-6, envelope=Envelope.COS_SQUARED,
                  polarization=Polarization.LINEAR if pol == "linear" else Polarization.CIRCULAR)
    g = 10000.0 / 0.510999
    beam = Beam(n=20_000, gamma=g, sigma=0.01 * g, species=Species.ELECTRON, radius=2e-6, rms_div=1e-6)
    opts = CollideOptions(tracking_photons=False, classical=classical)
    say("focused shower", pol, classical)
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0, 21)
    sg, _, cg = engine.collide(prim, opts, seed=21)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 21, 16)
    say("  gpu", cg, "\n  cpu", cc)
    assert cg["error_flags"] == 0
    assert cg["lepton_steps"] == pytest.approx(cc["lepton_steps"], rel=1e-3)
    a, b = cg["emissions"], cc["emissions"]
    assert a > 1000 and abs(a - b) < 5.0 * math.sqrt(a + b + 1.0)
    ph_g, ph_c = sg["species"] == 2, sc["species"] == 2
    for q in (lambda s, m: s["u"][m, 0], lambda s, m: np.arctan2(s["u"][m, 1], -s["u"][m, 3]),
              lambda s, m: np.arctan2(s["u"][m, 2], -s["u"][m, 3]), lambda s, m: s["payload"][m]):
        assert ks_2samp(q(sg, ph_g), q(sc, ph_c)).pvalue > 1e-3
    eg, ec = sg["u"][sg["species"] == 0, 0], sc["u"][sc["species"] == 0, 0]
    assert ks_2samp(eg, ec).pvalue > 1e-3
    xg, xc = sg["u"][sg["species"] == 0, 1], sc["u"][sc["species"] == 0, 1]
    assert ks_2samp(xg, xc).pvalue > 1e-3


def test_focused_laser_pairs_match_oracle(engine, orc):
    """CP FocusedLaser, a0 = 15, w0 = 3 um, photons and secondaries tracked: the
    lepton and photon phases of the cascade in the focused LMA field."""
    from scipy.stats import ks_2samp
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, Species
    laser = Laser(LaserKind.FOCUSED_LASER, 15.0, 0.8e-6, 16.0, waist=3e-6, envelope=Envelope.COS_SQUARED,
                  polarization=Polarization.CIRCULAR)
    g = 16500.0 / 0.510999
    beam = Beam(n=4000, gamma=g, sigma=0.01 * g, species=Species.ELECTRON, radius=1e-6, rms_div=1e-6)
    opts = CollideOptions(tracking_photons=True)
    say("focused cascade")
    engine.set_laser(laser)
    dt, z0 = engine.field_info()
    prim = orc.build_beam(beam.desc(), z0, 22)
    sg, _, cg = engine.collide(prim, opts, seed=22)
    sc, _, cc, _ = orc.collide(laser.desc(), opts.desc(), prim, 22, 16)
    say("  gpu", cg, "\n  cpu", cc)
    for key in ("emissions", "pair_events"):
        a, b = cg[key], cc[key]
        assert abs(a - b) < 5.0 * math.sqrt(a + b + 1.0), key
    assert cg["pair_events"] + cc["pair_events"] > 20
    for sp in (0, 2):
        eg, ec = sg["u"][sg["species"] == sp, 0], sc["u"][sc["species"] == sp, 0]
        assert ks_2samp(eg, ec).pvalue > 1e-3, sp
