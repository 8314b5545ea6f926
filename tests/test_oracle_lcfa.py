"""Pins the CPU oracle of the LCFA path (no GPU).

1. The reference's own known-answer tests, with the reference's tolerances
   (special_functions/airy.rs:140-250, pwmci.rs:308-387,
   lcfa/photon_emission/mod.rs:357-487, lcfa/pair_creation/mod.rs:229-392,
   field/fast_plane_wave.rs:246-350, field/fast_focused_laser.rs:271-292).
2. Golden vectors from mpmath/scipy (tests/golden/gen_golden.py).
3. Distributional checks: the inverse-CDF photon sampler against the CDF of the
   quadrature spectrum.
"""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C = 2.997925e8
RQM = -1.602177e-19 / 9.109383e-31


# ---------------------------------------------------------------- Airy / K (airy.rs:140-250)
@pytest.mark.parametrize("x,target", [(0.0, 0.3550280538878172), (2.0, 0.03492413042327438),
                                      (17.0, 7.05019729838861e-22), (20.0, 1.69167286867e-27)])
def test_airy_kat(orc, x, target):
    assert abs(orc.airy(x)[0][0] - target) / target < 1e-12


@pytest.mark.parametrize("x,target", [(0.5, -0.2249105326646839), (1.2, -0.1327853785572262),
                                      (3.0, -0.01191297670595132), (5.0, -0.0002474138908684625),
                                      (100.0, -2.635140361604410e-290)])
def test_airy_prime_kat(orc, x, target):
    assert abs(orc.airy(x)[0][1] - target) / abs(target) < 1e-12


@pytest.mark.parametrize("x,k13,k23", [(1.0, 0.43843063344153436171, 0.49447506210420826699),
                                       (2.0, 0.11654496129616524876, 0.12483892748812831057),
                                       (10.0, 0.000017874608271055334883, 0.000018161187569530204281),
                                       (20.0, 5.7568278247790870062e-10, 5.8038484271925806951e-10)])
def test_bessel_k_kat(orc, x, k13, k23):
    v = orc.airy(x)[0]
    assert abs(v[2] - k13) / k13 < 1e-12
    assert abs(v[3] - k23) / k23 < 1e-12


def test_airy_golden(orc):
    g = np.load(os.path.join(GOLDEN, "special.npz"))
    for x, ai, aip in zip(g["x"], g["ai"], g["aip"]):
        v = orc.airy(x)[0]
        assert abs(v[0] - ai) <= 5e-12 * abs(ai)
        assert abs(v[1] - aip) <= 5e-12 * abs(aip)
    for x, k13, k23 in zip(g["xk"], g["k13"], g["k23"]):
        v = orc.airy(x)[0]
        assert abs(v[2] - k13) <= 5e-12 * k13
        assert abs(v[3] - k23) <= 5e-12 * k23


# ---------------------------------------------------------------- PWMCI (pwmci.rs:308-387)
def test_pwmci_invert_x_sqd(orc):
    t = np.array([[i / 20.0, (i / 20.0) ** 2] for i in range(20)])
    assert abs(orc.pwmci(t, 0.73, invert=True) - math.sqrt(0.73)) < 1e-4


def test_pwmci_invert_tanh(orc):
    t = np.array([[5.0 * i / 40.0, math.tanh(5.0 * i / 40.0)] for i in range(40)])
    assert abs(orc.pwmci(t, 0.22, invert=True) - math.atanh(0.22)) < 1e-4


def test_pwmci_invert_shifted_tanh(orc):
    t = np.array([[5.0 * i / 20.0, 1.0 + math.tanh(5.0 * i / 20.0 - 2.0)] for i in range(20)])
    assert abs(orc.pwmci(t, 1.24, invert=True) - (2.0 - math.atanh(1.0 - 1.24))) < 1e-4


def test_pwmci_extrapolation(orc):
    t = np.array([[i / 20.0, math.tanh(i / 20.0)] for i in range(20)])
    assert abs(orc.pwmci(t, -1.0, extrapolate=True) + 1.0) < 1e-3
    th = math.tanh(1.0)
    target = (-1.0 - 0.9 + th + th * th) / (th * th - 1.0)
    assert abs(orc.pwmci(t, 0.9, invert=True, extrapolate=True) - target) < 0.1
    assert orc.pwmci(t, -1.0, extrapolate=False) is None


def test_pwmci_is_monotone_and_interpolating(orc):
    """Closed-form properties: passes through the nodes, monotone between them,
    and evaluate(invert(f)) == f."""
    x = np.linspace(0.0, 3.0, 16)
    t = np.stack([x, np.log1p(x) + 0.1 * x ** 2], 1)
    for xi, fi in t[1:-1]:
        assert abs(orc.pwmci(t, xi) - fi) < 1e-14 * max(1.0, abs(fi)) or xi == t[0, 0]
    xs = np.linspace(0.001, 2.999, 997)
    fs = np.array([orc.pwmci(t, v) for v in xs])
    assert np.all(np.diff(fs) > 0)
    for f in fs[::50]:
        assert abs(orc.pwmci(t, orc.pwmci(t, f, invert=True)) - f) < 1e-6 * abs(f)


# ---------------------------------------------------------------- LCFA emission
@pytest.mark.parametrize("chi,target", [(0.026, 2.07935e14), (3.5, 1.58485e16), (9.98, 3.45844e16),
                                        (12.4, 4.04647e16), (403.0, 4.46834e17)])
def test_lcfa_rate_kat(orc, chi, target):
    assert abs(orc.lib().orc_lcfa_rate(chi, 1000.0) - target) / target < 1e-3


def test_lcfa_rate_golden(orc):
    g = np.load(os.path.join(GOLDEN, "lcfa.npz"))
    for chi, w in zip(g["chi_w"], g["w_gamma1000"]):
        assert abs(orc.lib().orc_lcfa_rate(chi, 1000.0) - w) / w < 1e-3
    # rate scales as 1/gamma (lcfa/photon_emission/mod.rs:32)
    assert math.isclose(orc.lib().orc_lcfa_rate(0.7, 500.0), 2 * orc.lib().orc_lcfa_rate(0.7, 1000.0), rel_tol=1e-14)


@pytest.mark.parametrize("chi,target", [(0.0025, 0.98540791120703309744), (0.012, 0.93473708092320246516),
                                        (0.12, 0.61600532094295304700), (0.7, 0.23882354190470674907),
                                        (2.5, 0.081136810639749760802), (7.0, 0.027867522424020332230),
                                        (12.0, 0.015120883496044083748), (25.0, 0.0063062413153582144900),
                                        (70.0, 0.0017465062499607358868), (250.0, 0.00033811666381767191642)])
def test_gaunt_kat(orc, chi, target):
    assert abs(orc.lib().orc_gaunt(chi) - target) / target < 1e-3


@pytest.mark.parametrize("chi", [0.1, 1.0, 10.0])
def test_photon_energy_sampler_matches_spectrum(orc, chi):
    """lcfa::photon_emission::sample inverts the spectrum's CDF: the sampled
    u = omega/(gamma - omega) follows the CDF of the quadrature spectrum."""
    g = np.load(os.path.join(GOLDEN, "lcfa.npz"))
    k = list(g["cdf_chi"]).index(chi)
    gamma = 1e5
    rng = np.random.default_rng(11)
    n = 20000
    r = rng.uniform(size=(n, 3))
    om = np.array([orc.lcfa_sample(chi, gamma, *row)[0] for row in r])
    u = om / (gamma - om)
    emp = np.searchsorted(np.sort(u), g["cdf_u"]) / n
    assert np.max(np.abs(emp - g["cdf"][k])) < 4.0 / math.sqrt(n) + 5e-3


def test_stokes_vector_kat(orc):
    """lcfa/photon_emission/mod.rs:396-463: averaged over angles at fixed energy the
    Stokes vector is (1, (W11 - W22)/(W11 + W22), 0, 0). Azimuths are drawn in
    mirrored pairs (r3, 1 - r3), so the components odd in the out-of-plane angle
    cancel exactly instead of to the statistical noise of one random stream."""
    chi, gamma = 1.0, 1000.0
    w = np.array([1.0, 0.0, 0.0])
    lng = np.array([0.0, 0.0, 1.0])
    perp0 = np.array([1.0, 0.0, 0.0])
    rng = np.random.default_rng(0)
    sv = np.zeros(4)
    n = 40000
    for _ in range(n // 2):
        r2, r3 = rng.uniform(), rng.uniform()
        for rr3 in (r3, 1.0 - r3):
            om, ok, th, ph = orc.lcfa_sample(chi, gamma, 0.98, r2, rr3)
            perp = perp0 * math.cos(ph) + np.cross(lng, perp0) * math.sin(ph)
            k3 = om * (math.cos(th) * lng + math.sin(th) * perp)
            k = np.array([np.linalg.norm(k3), *k3])
            sv += orc.lcfa_stokes(k, chi, gamma, lng, w) / n
    assert abs(sv[2]) < 1e-3 and abs(sv[3]) < 1e-3
    from scipy import special, integrate
    u = om / (gamma - om)
    xi = 2 * u / (3 * chi)
    k23 = special.kv(2 / 3, xi)
    int53 = integrate.quad(lambda y: special.kv(5 / 3, y), xi, np.inf)[0]
    target = k23 / (u * u * k23 / (1 + u) + int53)
    assert abs(sv[1] - target) / target < 0.01


# ---------------------------------------------------------------- LCFA pair creation
@pytest.mark.parametrize("chi,target", [(0.042, 6.077538994929929904e-29), (0.105, 2.1082097875655204834e-12),
                                        (0.42, 0.00037796132366581330636), (1.05, 0.015977478443872017101),
                                        (4.2, 0.08917816786414408900), (12.0, 0.10884579479913803705),
                                        (42.0, 0.09266735324318656466)])
def test_aux_t_kat(orc, chi, target):
    t = orc.aux_t(chi)
    assert abs(0.5 * (t[0] + t[1]) - target) / target < 1e-2


def test_aux_t_golden(orc):
    g = np.load(os.path.join(GOLDEN, "lcfa.npz"))
    for chi, tp, tq in zip(g["chi_t"], g["t_par"], g["t_perp"]):
        t = orc.aux_t(chi)
        assert abs(t[0] - tp) / tp < 1e-2, chi
        assert abs(t[1] - tq) / tq < 1e-2, chi


def test_pair_spectrum_ceiling(orc):
    """lcfa/pair_creation/mod.rs:262-287: the ceiling bounds the spectrum."""
    rng = np.random.default_rng(0)
    L = orc.lib()
    for _ in range(100):
        chi = math.exp(rng.uniform(0.0, math.log(100.0)))
        sv1 = rng.uniform(-1, 1)
        mx = max(L.orc_pair_spectrum(0.5 * i / 10000.0, chi, sv1) for i in range(1, 10000, 7))
        assert L.orc_pair_spectrum_ceiling(chi, sv1) > mx


def test_pair_angular_spectrum_ceiling(orc):
    rng = np.random.default_rng(1)
    L = orc.lib()
    for _ in range(300):
        chi = math.exp(math.log(0.1) + (math.log(100.0) - math.log(1.0)) * rng.uniform())
        sv1 = rng.uniform(-1, 1)
        s = 0.5 * rng.uniform()
        mx = max(L.orc_pair_angular_spectrum(0.5 * i / 100.0, s, chi, sv1) for i in range(100))
        assert L.orc_pair_angular_spectrum_ceiling(s, chi, sv1) > mx or mx < 1e-200


def test_pure_states_remain_so(orc):
    """lcfa/pair_creation/mod.rs:379-392."""
    for sv in ([1.0, 1.0, 0.0, 0.0], [1.0, -1.0, 0.0, 0.0]):
        _, new = orc.pair_probability([100.0, 0.0, 0.0, -100.0], sv, 1.0, [1.0, 0.0, 0.0], 1e-6 / C)
        assert abs(sv[1] - new[1]) < 1e-12


def test_pair_rate_pol_resolved(orc):
    """T_par/T_perp equals the ratio of the integrated polarised spectra (mod.rs:262-291)."""
    L = orc.lib()
    n32 = np.polynomial.legendre.leggauss(32)
    for chi in (0.15, 0.75, 1.2, 8.7, 43.0):
        s = 0.5 * (1 + n32[0])
        tp = sum(w * L.orc_pair_spectrum(x, chi, 1.0) for x, w in zip(s, n32[1]))
        tq = sum(w * L.orc_pair_spectrum(x, chi, -1.0) for x, w in zip(s, n32[1]))
        t = orc.aux_t(chi)
        assert abs(t[0] / t[1] - tp / tq) / (t[0] / t[1]) < 1e-2


# ---------------------------------------------------------------- fields and pusher
def _laser(orc, kind, env, pol, a0, wl=0.8e-6, waist=4e-6, nc=10.0):
    return orc.laser_desc(kind=kind, envelope=env, polarization=pol, a0=a0, wavelength=wl, waist=waist, n_cycles=nc)


def test_fast_focused_on_axis(orc):
    """fast_focused_laser.rs:271-292: an on-axis electron leaves with no transverse momentum."""
    L = _laser(orc, 3, 2, 1, 100.0)
    u = np.array([[math.sqrt(1 + 1000.0 ** 2), 0, 0, -1000.0]])
    t_start = -20.0 * 0.8e-6 / C
    r = np.array([[0.0, 0.0, 0.0, 0.0]]) + u * C * t_start / u[0, 0]
    r2, u2, k = orc.push(L, [0], r, u, 10 ** 7)
    assert abs(u2[0, 1]) < 1e-3 and abs(u2[0, 2]) < 1e-3
    assert abs(u2[0, 0] ** 2 - (u2[0, 1:] ** 2).sum() - 1.0) < 1e-3


def test_fast_plane_wave_max_u_perp(orc):
    """fast_plane_wave.rs:245-288: in a CP plane wave max |u_perp| = a0."""
    L = _laser(orc, 2, 0, 1, 100.0, nc=8.0)
    dt_target = 0.005 * 0.8e-6 / C
    mult = dt_target / orc.lib().orc_laser_max_timestep(L)
    u = np.array([[math.sqrt(1 + 100.0 ** 2), 0, 0, -100.0]])
    r = u * C * (-0.5 * 8 * 0.8e-6 / C) / u[0, 0]
    upmax = 0.0
    for _ in range(1600):
        r, u, _ = orc.push(L, [0], r, u, 1, dt_multiplier=mult)
        upmax = max(upmax, math.hypot(u[0, 1], u[0, 2]))
    assert abs(upmax - 100.0) / 100.0 < 1e-3


def test_fast_plane_wave_ll_depletion(orc):
    """fast_plane_wave.rs:315-350: work done under Landau-Lifshitz equals the
    classical radiated energy (closed form)."""
    a0, nc = 20.0, 8.0
    L = _laser(orc, 2, 2, 0, a0, nc=nc)
    z0 = orc.lib().orc_laser_ideal_initial_z(L)
    r = np.array([[-z0, 0, 0, z0]])
    u = np.array([[math.sqrt(1 + 1000.0 ** 2), 0, 0, -1000.0]])
    _, _, _, work = orc.push(L, [0], r, u, 10 ** 8, dt_multiplier=0.1, eqn=1, with_work=True)
    phase = math.pi ** 1.5 * nc / math.sqrt(math.log(4.0))
    omega_mc2 = C * 1.28808867e-21 * 2 * math.pi / 0.8e-6
    expected = 7.29735257e-3 * a0 ** 4 * omega_mc2 * (1.0 / 8.0) * phase / 3.0
    assert abs(work[0] - expected) / expected < 1e-3


def test_xoshiro_jump_changes_stream(orc):
    a = orc.xoshiro(0x8658b90036b165eb, 4)
    b = orc.xoshiro(0x8658b90036b165eb, 4, jumps=1)
    assert not np.array_equal(a, b)
