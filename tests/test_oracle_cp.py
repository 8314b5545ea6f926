"""The oracle's restatement of Bessel J_n and of circularly polarised nonlinear
Compton scattering (oracle/cp.cpp), pinned by the reference's own tests
(special_functions/bessel.rs:206-341, nonlinear_compton/cp/mod.rs:236-470,
cp/classical.rs:184-216), by scipy's Bessel functions, and by the reference's
shipped tables: the classical CDF and rate tables, which the same construction
must reproduce, and the quantum total-rate table, against which the regenerated
quantum CDF table's totals are checked.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as orc


def _xoshiro_f64(seed):
    """rand_xoshiro 0.4 Xoshiro256** seeded by SplitMix64; rand 0.7 gen::<f64>."""
    M = (1 << 64) - 1
    z = seed
    s = []
    for _ in range(4):
        z = (z + 0x9E3779B97F4A7C15) & M
        x = z
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        s.append(x ^ (x >> 31))

    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M

    while True:
        r = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        yield (r >> 11) * (1.0 / 9007199254740992.0)


# ---------------------------------------------------------------- Bessel J
@pytest.mark.parametrize("n,x,target", [
    (70, 5.0, 5.484562820751933e-73), (31, 30.0, 0.1023416331626096), (24, 2.0, 1.548492695958880e-24),
    (8, 2.0, 0.000022179552287925904088), (1, 0.87, 0.39512125869620047961), (1, 0.2, 0.099500832639235995398),
    (0, 1.2, 0.6711327442643626735), (0, 12.0, 0.047689310796833536624), (0, 120.0, 0.0718234158291561276)])
def test_bessel_known_values(n, x, target):
    """bessel.rs:220-301, error < 1e-8."""
    assert abs(orc.lib().orc_bessel_j(n, x) - target) / target < 1e-8


@pytest.mark.parametrize("n,x,target,which", [(400, 350.0, 5.912510654628083e-10, 0), (400, 400.0, 0.06070867128509718, 0),
                                              (1000, 990.0, 0.012361942456230178547, 1)])
def test_bessel_recurrence_and_uniform_expansion(n, x, target, which):
    """bessel.rs:303-341: the recurrence at n = 400 and the uniform expansion at n = 1000, error < 1e-8."""
    out = np.zeros(2)
    orc.lib().orc_bessel_j_parts(n, x, orc.dptr(out))
    assert abs(out[which] - target) / target < 1e-8


def test_bessel_triple_matches_scipy():
    """(J_{n-1}, J_n, J_{n+1}) against scipy.special.jv on 0 < x < n, both methods
    (recurrence: 1e-8, the reference's own bound; uniform expansion: 1e-5). Values
    below 1e-200 are exempt: the recurrence's normalisation overflows there and
    returns 0, as the reference's does."""
    from scipy.special import jv
    rng = np.random.default_rng(3)
    for n in list(rng.integers(1, 999, 60)) + list(rng.integers(1000, 60000, 30)):
        n = int(n)
        x = n * rng.uniform(0.3, 0.999)
        got = orc.bessel_j_pm(n, x)
        ref = jv([n - 1, n, n + 1], x)
        tol = 1e-8 if n < 1000 else 1e-5
        assert np.allclose(got, ref, rtol=tol, atol=1e-200), (n, x, got, ref)


# ---------------------------------------------------------------- quantum CP rates
KATS = [
    (2, 0.5, 0.15, 2.748486539e-3), (10, 1.0, 0.2, 1.984654425e-4), (80, 2.0, 0.2, 3.751480198e-6),
    (160, 2.0, 0.2, 6.842944878e-9), (50, 3.0, 0.1, 5.090018978e-4), (200, 3.0, 0.1, 3.504645316e-6),
    (200, 4.0, 0.1, 5.564288841e-5), (500, 4.0, 0.1, 9.722534139e-7), (100, 5.0, 0.1, 6.745093014e-4),
    (500, 5.0, 0.1, 1.258283729e-5), (1000, 5.0, 0.1, 4.137051481e-7), (40, 7.0, 0.1, 3.198368332e-3),
    (160, 7.0, 0.1, 6.698029091e-4), (640, 7.0, 0.1, 5.063579159e-5), (2560, 7.0, 0.1, 2.322138448e-7),
    (100, 9.5, 0.1, 1.656736051e-3), (1000, 9.5, 0.1, 6.425026440e-5), (8000, 9.5, 0.1, 2.056455838e-8),
    (100, 9.5, 0.01, 1.981917068e-4), (1000, 9.5, 0.01, 1.624055198e-5), (100, 9.5, 0.0012, 2.424143218e-5),
    (1000, 9.5, 0.0012, 2.319090168e-6), (5000, 9.5, 0.0012, 3.366698038e-8)]


@pytest.mark.parametrize("n,a,eta,target", KATS)
def test_cp_partial_rate_known_values(n, a, eta, target):
    """cp/mod.rs:309-346 (partial_rate_accuracy), error < 1e-6."""
    assert abs(orc.lib().orc_cp_partial_rate(n, a, eta) - target) / target < 1e-6


@pytest.mark.parametrize("nmax,a,eta", [(10, 0.5, 0.2), (20, 1.0, 0.2), (280, 3.0, 0.12), (650, 4.0, 0.12), (20, 1.0, 0.75),
                                        (280, 3.0, 0.75), (1260, 5.0, 0.12)])
def test_cp_total_rate_is_sum_of_partial_rates(nmax, a, eta):
    """cp/mod.rs:348-379 (total_rate, the cheaper points): the table against the
    explicit sum of partial rates, error < 1e-3."""
    total = sum(orc.lib().orc_cp_partial_rate(n, a, eta) for n in range(1, nmax + 1))
    assert abs(total - orc.cp_rate(a, eta)) / orc.cp_rate(a, eta) < 1e-3


@pytest.mark.parametrize("nmax,a,eta", [(280, 3.0, 0.0005), (10010, 10.0, 0.0005)])
def test_cp_total_rate_low_eta(nmax, a, eta):
    """cp/mod.rs:381-397 (total_rate_low_eta): the classical table below the quantum
    one's eta range against eta x the sum of classical partial rates, error < 1e-3."""
    total = eta * sum(orc.lib().orc_cp_integrated_spectrum_low_eta(n, a) for n in range(1, nmax + 1))
    assert abs(total - orc.cp_rate(a, eta)) / orc.cp_rate(a, eta) < 1e-3


def test_cp_rate_limits():
    """Below a = 0.02 the rate is the closed-form linear Compton rate (cp/mod.rs:168-169),
    which must meet the table at its edge; below eta = 1e-3 it is the classical rate,
    which must meet the quantum table at its edge up to the leading quantum
    correction, of order chi = a eta."""
    for eta in (0.002, 0.05, 0.5, 1.5):
        below, above = orc.cp_rate(0.0199999, eta), orc.cp_rate(0.0200001, eta)
        assert abs(below / above - 1.0) < 2e-3, eta
    for a in (0.05, 0.5, 2.0, 10.0, 15.0):
        below, above = orc.cp_rate(a, 0.00099999), orc.cp_rate(a, 0.00100001)
        assert 0.0 < below / above - 1.0 < 2e-3 + 1.2 * a * 1e-3, a
    assert orc.cp_rate(25.0, 0.1) is None and orc.cp_rate(1.0, 2.5) is None


def test_cp_rate_ceiling():
    """cp/mod.rs:238-282 (rate_ceiling): the first 25 of its 100 seeded draws of
    (a, eta); the ceiling exceeds the maximum over a 10^4-point grid at every harmonic."""
    rng = _xoshiro_f64(0)
    L = orc.lib()
    for _ in range(25):
        a = math.exp(math.log(0.2) + (math.log(20.0) - math.log(0.2)) * next(rng))
        eta = math.exp(math.log(0.001) + (math.log(1.0) - math.log(0.001)) * next(rng))
        n_max = int(10.0 * (1.0 + a ** 3))
        if n_max > 200:
            harmonics = [int(math.exp(math.log(2.0) + 0.1 * i * (math.log(n_max) - math.log(2.0)))) for i in range(11)]
        elif n_max > 10:
            harmonics = [1, 2, 3] + [int(math.exp(math.log(5.0) + 0.25 * i * (math.log(n_max) - math.log(5.0)))) for i in range(5)]
        else:
            harmonics = list(range(1, n_max))
        for n in harmonics:
            true_max = L.orc_cp_grid_max(n, a, eta, 0)
            assert L.orc_cp_ceiling(n, a, eta) > true_max, (a, eta, n)


def test_cp_classical_rate_ceiling():
    """cp/classical.rs:184-216 (rate_ceiling), all 100 seeded draws of a."""
    rng = _xoshiro_f64(0)
    L = orc.lib()
    for _ in range(100):
        a = math.exp(math.log(0.2) + (math.log(20.0) - math.log(0.2)) * next(rng))
        n_max = int(10.0 * (1.0 + a ** 3))
        if n_max > 200:
            harmonics = [int(math.exp(math.log(2.0) + 0.1 * i * (math.log(n_max) - math.log(2.0)))) for i in range(11)]
        elif n_max > 10:
            harmonics = [1, 2, 3] + [int(math.exp(math.log(5.0) + 0.25 * i * (math.log(n_max) - math.log(5.0)))) for i in range(5)]
        else:
            harmonics = list(range(1, n_max))
        for n in harmonics:
            true_max = L.orc_cp_grid_max(n, a, 0.1, 1)
            ceiling = L.orc_cp_classical_ceiling(n, a)
            assert ceiling > true_max or true_max == 0.0, (a, n)


# ---------------------------------------------------------------- tables
@pytest.mark.parametrize("j", [0, 20, 33, 38, 40, 44, 50])
def test_cdf_construction_reproduces_shipped_classical_table(golden_tables, j):
    """The construction of cp/classical.rs:220-300 (the one cp/mod.rs:529-611 uses
    for the quantum table) reproduces the shipped classical rows to their printed
    precision: every-harmonic rows (n_max <= 32), mixed rows (n_max < 100) and
    log-spaced rows."""
    a = 0.02 * 10.0 ** (j / 20.0)
    total, row = orc.cp_cdf_row(a, 1.0, classical=True)
    ship = golden_tables["nonlinear_compton/cp/classical/cdf_table:TABLE"][j]
    assert np.allclose(row[:, 0], ship[:, 0], rtol=1e-11, atol=0)
    assert np.allclose(row[:, 1], ship[:, 1], rtol=0, atol=1e-11)
    assert abs(math.log(total) - golden_tables["nonlinear_compton/cp/classical/rate_table:TABLE"][j]) < 1e-11


def test_regenerated_quantum_cdf_totals_match_shipped_table(golden_tables):
    """The regenerated quantum CDF table ends at 1 and is monotone in every row;
    its rows' totals were checked against the shipped total table by the
    generator, and every row with n_max <= 32 is reproduced here by the oracle."""
    cdf = golden_tables["nonlinear_compton/cp/tables/cdf:TABLE"]
    assert cdf.shape == (4270, 32, 2)
    assert np.all(np.diff(cdf[:, :, 0], axis=1) >= 0.0)
    assert np.all(np.diff(cdf[:, :, 1], axis=1) >= -1e-15)
    assert np.allclose(cdf[:, 31, 1], 1.0, atol=1e-12)
    ship = golden_tables["nonlinear_compton/cp/tables/total:TABLE"]
    for i in (0, 23, 47, 69):
        for j in (0, 15, 30, 37):  # a <= 1.4: n_max <= 32
            a, eta = 0.02 * 10.0 ** (j / 20.0), 1.0e-3 * 10.0 ** (i / 20.0)
            total, row = orc.cp_cdf_row(a, eta)
            assert np.allclose(row, cdf[61 * i + j], rtol=1e-13, atol=1e-15), (i, j)
            assert abs(math.log(total) - ship[i, j]) < 1e-9, (i, j)


@pytest.mark.parametrize("i,j", [(20, 42), (46, 48), (60, 52)])
def test_regenerated_quantum_cdf_rows_with_many_harmonics(golden_tables, i, j):
    """Rows with 32 < n_max < 100 and n_max >= 100 against the oracle."""
    a, eta = 0.02 * 10.0 ** (j / 20.0), 1.0e-3 * 10.0 ** (i / 20.0)
    total, row = orc.cp_cdf_row(a, eta)
    cdf = golden_tables["nonlinear_compton/cp/tables/cdf:TABLE"][61 * i + j]
    assert np.allclose(row, cdf, rtol=1e-12, atol=1e-14)
    assert abs(math.log(total) - golden_tables["nonlinear_compton/cp/tables/total:TABLE"][i, j]) < 1e-9


@pytest.mark.parametrize("a,eta", [(0.946303, 0.105925), (2.37700, 0.105925), (4.74275, 0.105925)])
def test_cp_harmonic_index_sampling(a, eta):
    """cp/mod.rs:399-470 (harmonic_index_sampling): the table inversion against the
    direct cumulative sum of partial rates; the reference prints the error
    statistics, here the mean |error| must stay below one harmonic and the two
    methods agree exactly for most draws."""
    L = orc.lib()
    n_max = int(10.0 * (1.0 + a * a * a))
    rates = np.cumsum([L.orc_cp_partial_rate(n, a, eta) for n in range(1, n_max)])
    rng = np.random.default_rng(0)
    frac = rng.uniform(size=20000)
    n_interp = np.array([L.orc_cp_harmonic(a, eta, f) for f in frac])
    n_direct = 1 + np.searchsorted(rates, frac * rates[-1], side="right")
    err = n_direct - n_interp
    assert np.mean(np.abs(err)) < 1.0
    assert np.mean(err == 0) > 0.5


# ---------------------------------------------------------------- sampler
def _s_marginal(a, eta, n, classical, bins):
    """Expected probability per bin of s at fixed harmonic n from the spectrum."""
    L = orc.lib()
    sn = 2.0 * n * eta / (1.0 + a * a)
    scale = sn if classical else sn / (1.0 + sn)
    v = np.linspace(0.0, 1.0, 20001)[:-1] + 0.5 / 20000
    f = np.array([L.orc_cp_spectrum_low_eta(n, a, x) if classical else L.orc_cp_ddpr(n, a, eta, x) for x in v])
    h, _ = np.histogram(v * scale, bins, weights=f)
    return h / h.sum()


@pytest.mark.parametrize("a,eta,n,classical", [(1.0, 0.2, 3, False), (3.0, 0.1, 40, False), (3.0, 0.1, 40, True),
                                               (12.0, 0.05, 2000, False)])
def test_cp_sampler_s_marginal(a, eta, n, classical):
    """The rejection sampler's s at fixed n against the spectrum it samples (chi^2),
    and the Stokes vector is a physical polarisation state."""
    from scipy.stats import chisquare
    out = orc.cp_sample(a, eta, 40000, classical=classical, seed=1, fixed_n=n)
    sn = 2.0 * n * eta / (1.0 + a * a)
    top = sn if classical else sn / (1.0 + sn)
    bins = np.linspace(0.0, top, 41)
    h, _ = np.histogram(out[:, 1], bins)
    p = _s_marginal(a, eta, n, classical, bins)
    keep = p * len(out) > 5
    assert chisquare(h[keep], p[keep] / p[keep].sum() * h[keep].sum()).pvalue > 1e-3
    pol = np.linalg.norm(out[:, 3:6], axis=1)
    assert np.all(pol <= 1.0 + 1e-9)
    assert np.all((out[:, 2] >= 0.0) & (out[:, 2] < 2.0 * math.pi))


# ---------------------------------------------------------------- CP nonlinear Breit-Wheeler (pair_creation/cp/mod.rs)
CPBW_KATS = [
    (50, 1.0, 0.1, 1.4338498960396018931e-18), (100, 1.0, 0.1, 1.3654528056555865291e-18),
    (200, 1.0, 0.1, 2.0884399327604375975e-34), (4000, 5.0, 0.02, 9.2620552570985535880e-61),
    (5000, 5.0, 0.02, 5.9413657401296979089e-17), (6000, 5.0, 0.02, 1.1629168979497463847e-18),
    (8000, 5.0, 0.02, 1.6722921069034930599e-23), (3, 0.1, 1.0, 2.7926363804797338348e-7),
    (5, 0.1, 1.0, 1.4215822192894496185e-10), (20, 0.1, 1.0, 1.1028787567587238051e-37),
    (3000, 10.0, 0.1, 2.8532353822421244101e-48), (4000, 10.0, 0.1, 1.4356925580571873594e-5),
    (8000, 10.0, 0.1, 1.5847977567888444504e-7), (6, 1.0, 1.0, 0.0031666996194000280745),
    (20, 1.0, 1.0, 1.2280171339973043031e-5), (50, 1.0, 1.0, 7.7268893728561315057e-11),
    (300, 10.0, 1.0, 1.6230747656967905300e-7), (400, 10.0, 1.0, 0.0031791285538880908649),
    (2000, 10.0, 1.0, 5.9533991784012215406e-5), (640, 1.0, 0.01, 9.7351336009776642509e-115),
    (1000, 1.0, 0.01, 2.3257373741363993054e-156), (25, 0.1, 0.1, 5.7778053795802739886e-52),
    (50, 0.1, 0.1, 3.3444371706672986244e-90)]


@pytest.mark.parametrize("n,a,eta,target", CPBW_KATS)
def test_cpbw_partial_rate_known_values(n, a, eta, target):
    """pair_creation/cp/mod.rs:454-499 (partial_rates): the polarisation average, error < 1e-6."""
    r = orc.cpbw_partial_rate(n, a, eta)
    assert abs(0.5 * (r[0] + r[1]) - target) / target < 1e-6


@pytest.mark.parametrize("i,j", [(0, 6), (0, 9), (0, 13), (2, 4), (2, 9), (3, 3), (3, 8), (4, 2), (4, 7), (5, 4), (5, 10),
                                 (6, 6), (6, 9), (7, 8), (7, 12), (8, 13)])
def test_cpbw_table_matches_summation(i, j):
    """pair_creation/cp/mod.rs:501-523 (total_rates) on the points of its grid
    with at most 3000 harmonics: the table against the explicit sum, error < 1%."""
    a, eta = 0.305 * 10.0 ** (i / 5.0), 0.003 * 10.0 ** (j / 5.0)
    method, table = orc.cpbw_total(a, eta)
    if method == 0:
        pytest.skip("negligible rate")
    hi = orc.C.c_int(0)
    orc.lib().orc_cpbw_sum_limits(a, eta, orc.C.byref(hi))
    assert hi.value <= 3000
    total = orc.cpbw_by_summation(a, eta)
    assert method == 2
    assert np.all(np.abs(table - total) / total < 0.01)


def test_cpbw_probability_and_stokes():
    """Probability for S_3 = +1 / -1 from the two tabulated totals, and the Stokes
    update that keeps the photon ensemble unbiased (pair_creation/cp/mod.rs:286-322):
    for S_3 = +-1 the parameter stays +-1, and S_1, S_2 scale by (1 - P_avg) / (1 - P)."""
    L = orc.lib()
    a, eta, dphi = 3.0, 0.2, 0.1
    method, f = orc.cpbw_total(a, eta)
    assert method == 2
    for sv3, fk in ((1.0, f[0]), (-1.0, f[1])):
        sv = np.array([1.0, 0.0, 0.0, sv3])
        out = np.zeros(4)
        p = L.orc_cpbw_probability(a, eta, orc.dptr(sv), dphi, orc.dptr(out))
        assert p == pytest.approx(7.29735257e-3 * fk * dphi / eta, rel=1e-14)
        assert out[3] == pytest.approx(sv3, abs=1e-14)
    sv = np.array([1.0, 0.3, -0.4, 0.2])
    out = np.zeros(4)
    p = L.orc_cpbw_probability(a, eta, orc.dptr(sv), dphi, orc.dptr(out))
    p_avg = 7.29735257e-3 * 0.5 * (f[0] + f[1]) * dphi / eta
    assert out[1] == pytest.approx(0.3 * (1 - p_avg) / (1 - p), rel=1e-14)
    assert out[2] == pytest.approx(-0.4 * (1 - p_avg) / (1 - p), rel=1e-14)
    # decay-weighted average of S_3 over both outcomes equals the prior S_3
    p_plus, p_minus = 7.29735257e-3 * f[0] * dphi / eta, 7.29735257e-3 * f[1] * dphi / eta
    no_decay = (0.6 * (1 - p_plus) - 0.4 * (1 - p_minus)) / (0.6 * (1 - p_plus) + 0.4 * (1 - p_minus))
    assert out[3] == pytest.approx(no_decay, rel=1e-12)


@pytest.mark.parametrize("a,eta,sv3", [(1.0, 0.5, 0.0), (3.0, 0.2, 0.8), (0.2, 1.0, -0.5)])
def test_cpbw_harmonic_distribution(a, eta, sv3):
    """Harmonics drawn by explicit summation against the partial rates (chi^2); the
    rates below a = 0.3 are summed too (pair_creation/cp/mod.rs:271-281)."""
    from scipy.stats import chisquare
    L = orc.lib()
    hi = orc.C.c_int(0)
    lo = L.orc_cpbw_sum_limits(a, eta, orc.C.byref(hi))
    pr = np.array([orc.cpbw_partial_rate(n, a, eta) for n in range(lo, hi.value + 1)])
    w = 0.5 * ((1 + sv3) * pr[:, 0] + (1 - sv3) * pr[:, 1])
    frac = (np.arange(2000) + 0.5) / 2000
    n = np.array([L.orc_cpbw_harmonic(a, eta, sv3, f) for f in frac])
    assert np.all((n >= lo) & (n <= hi.value))
    h = np.bincount(n - lo, minlength=len(w))[:len(w)]
    p = w / w.sum()
    keep = p * len(n) > 5
    # the inversion is deterministic on evenly spaced fractions; the table total and
    # the sum differ by < 1%, which shifts the top harmonic's share only
    assert np.max(np.abs(h[keep] / len(n) - p[keep])) < 0.012


@pytest.mark.parametrize("n,a,eta,sv", [(30, 1.0, 0.5, (0.0, 0.0, 0.0)), (300, 3.0, 0.2, (0.5, -0.3, 0.7)),
                                        (1500, 5.0, 0.2, (0.0, 0.9, -0.2))])
def test_cpbw_sampler_marginals(n, a, eta, sv):
    """(s, theta) at fixed n against the double-differential rate's marginals (chi^2)."""
    from scipy.stats import chisquare
    L = orc.lib()
    out = orc.cpbw_sample(a, eta, *sv, 20000, seed=4, fixed_n=n)
    s_grid = np.linspace(0.0, 1.0, 801)[:-1] + 0.5 / 800
    th_grid = np.linspace(0.0, 2 * math.pi, 121)[:-1] + math.pi / 120
    sn = 2.0 * n * eta / (1.0 + a * a)
    s_min = 0.5 * (1.0 - math.sqrt(1.0 - 4.0 / sn))
    inside = (s_grid > s_min) & (s_grid < 1.0 - s_min)
    dens = np.array([[L.orc_cpbw_fully_resolved(n, a, eta, min(s, 1 - s), t, *sv) if ok else 0.0 for t in th_grid]
                     for s, ok in zip(s_grid, inside)])
    for col, grid, marg, bins in ((1, s_grid, dens.sum(1), np.linspace(0, 1, 41)),
                                  (2, th_grid, dens.sum(0), np.linspace(0, 2 * math.pi, 25))):
        h, _ = np.histogram(out[:, col], bins)
        p, _ = np.histogram(grid, bins, weights=marg)
        p = p / p.sum()
        keep = p * len(out) > 5
        assert chisquare(h[keep], p[keep] / p[keep].sum() * h[keep].sum()).pvalue > 1e-3, col
