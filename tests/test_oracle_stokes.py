"""Pins of the oracle's LMA photon Stokes parameters (nonlinear_compton/lp/mod.rs:622-671,
lp/classical.rs:294-330, cp/mod.rs:203-232, cp/classical.rs:155-175) by what the
mathematics fixes, not by retyping the formulas:

* Thomson limit (a -> 0, eta -> 0, first harmonic): in the zero-momentum frame the
  emission is the radiation of a dipole driven by the laser's polarization vector
  eps, so the photon's field is E = eps - (eps . n) n, fully polarized. Its Stokes
  parameters are taken on the light-front basis the reference defines
  (polarization.rs:128-137: e1 = x - k'_x (k' - omega' z) / (omega' (omega' - k'_z)),
  which a boost along the laser axis leaves unchanged) with the reference's own
  Jones convention (polarization.rs:166-182: e2 = e1 x n, u = 2 Re E1 E2*,
  v = 2 Im E1 E2*).
* The quantum forms tend to the separately written classical forms as eta -> 0 at
  finite a (many harmonics, J(n, x, y) away from small arguments).
* xi_0 of the LP Stokes block is the double-differential rate (lp/mod.rs:21-50,
  pinned by the reference's rate tests): xi_0 = 4 pi d2W / a^2 exactly.

Reference quirks these pins record (the oracle reproduces them, as does the device):
* quantum LP: S2 has the opposite sign to the dipole on the reference's basis and
  to the reference's classical LP sampler (cos theta > 0, where the classical
  sampler's double parity correction cancels);
* CP: the linear part is the dipole's with e1 and e2 exchanged, (Q, U) -> (-Q, U);
  the quantum S3 carries LASER_S3 = -1 (cp/mod.rs:9, 224), the classical S3 does not.
"""
import math

import numpy as np
import pytest


def _lf_basis(n):
    """The reference's light-front basis e1 for a photon along n (laser along +z)
    and e2 = e1 x n."""
    e1 = np.array([1.0, 0.0, 0.0]) - n[0] * (n - np.array([0.0, 0.0, 1.0])) / (1.0 - n[2])
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(e1, n)


def _stokes(E, e1, e2):
    E1, E2 = np.dot(E, e1), np.dot(E, e2)
    i = abs(E1) ** 2 + abs(E2) ** 2
    c = E1 * np.conj(E2)
    return np.array([(abs(E1) ** 2 - abs(E2) ** 2) / i, 2.0 * c.real / i, 2.0 * c.imag / i])


def _dipole(eps, cos_t, phi):
    """Stokes (S1, S2, S3) of dipole radiation driven along eps (complex), emitted at
    polar angle arccos(cos_t) from the laser axis and azimuth phi from its x axis."""
    st = math.sqrt(1.0 - cos_t * cos_t)
    n = np.array([st * math.cos(phi), st * math.sin(phi), cos_t])
    e1, e2 = _lf_basis(n)
    return _stokes(eps - np.dot(eps, n) * n, e1, e2)


X = np.array([1.0, 0.0, 0.0])
CIRC = np.array([1.0, 1.0j, 0.0]) / math.sqrt(2.0)


def test_lp_quantum_thomson_limit(orc):
    """n = 1, a = 1e-4, eta = 1e-7: S1 is the dipole's; S2 is minus the dipole's
    (quirk above). The ZMF polar angle follows from s as in nonlinear_compton::generate
    (mod.rs:79-82): cos = 1 - s (1 + a_rms^2 + 2 eta) / eta."""
    a, eta = 1e-4, 1e-7
    rng = np.random.default_rng(1)
    sn = 2.0 * eta / (1.0 + 0.5 * a * a)
    smax = sn / (1.0 + sn)
    for _ in range(300):
        s = rng.uniform(0.01, 0.99) * smax
        th = rng.uniform(0.0, 2.0 * math.pi)
        c = 1.0 - s * (1.0 + 0.5 * a * a + 2.0 * eta) / eta
        d = _dipole(X, c, th)
        r = orc.lp_stokes(1, a, eta, s, th)
        assert abs(r[1] - d[0]) < 1e-5 and abs(r[2] + d[1]) < 1e-5 and r[3] == 0.0, (s, th, r, d)


def test_lp_classical_thomson_limit(orc):
    """Classical LP (lp/classical.rs), n = 1, a = 1e-4, photon at v = s / s_n with
    cos = 1 - 2 v: the dipole's Stokes parameters on the reference's basis, for
    cos theta > 0."""
    a = 1e-4
    rng = np.random.default_rng(2)
    for _ in range(300):
        v = rng.uniform(0.01, 0.99)
        th = rng.uniform(-0.5 * math.pi, 0.5 * math.pi) % (2.0 * math.pi)
        d = _dipole(X, 1.0 - 2.0 * v, th)
        r = orc.lp_cl_stokes(1, a, v, th)
        assert abs(r[1] - d[0]) < 1e-5 and abs(r[2] - d[1]) < 1e-5, (v, th, r, d)


@pytest.mark.parametrize("a", [0.3, 1.0, 3.0, 8.0])
def test_lp_quantum_tends_to_classical(orc, a):
    """At eta = 1e-7 the quantum LP Stokes parameters equal the classical ones at
    the same harmonic, v = s / s_n and theta (cos theta > 0), up to the S2 sign."""
    eta = 1e-7
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 5, 8, 13):
        sn = 2.0 * n * eta / (1.0 + 0.5 * a * a)
        smax = sn / (1.0 + sn)
        for _ in range(30):
            s = rng.uniform(0.02, 0.98) * smax
            th = rng.uniform(-0.5 * math.pi, 0.5 * math.pi) % (2.0 * math.pi)
            q = orc.lp_stokes(n, a, eta, s, th)
            c = orc.lp_cl_stokes(n, a, s / sn, th)
            assert abs(q[1] - c[1]) < 1e-3 and abs(q[2] + c[2]) < 1e-3, (n, s, th, q, c)


def test_lp_xi0_is_the_double_differential_rate(orc):
    """xi_0 of the Stokes block equals 4 pi d2W(s, theta) / a^2 (lp/mod.rs:21-50),
    whose integrals the reference's partial-rate and table tests pin."""
    rng = np.random.default_rng(4)
    lib = orc.lib()
    for _ in range(400):
        n = int(rng.integers(1, 40))
        a = math.exp(rng.uniform(math.log(0.05), math.log(10.0)))
        eta = math.exp(rng.uniform(math.log(1e-3), math.log(0.5)))
        sn = 2.0 * n * eta / (1.0 + 0.5 * a * a)
        s = rng.uniform(0.01, 0.99) * sn / (1.0 + sn)
        th = rng.uniform(0.0, 0.5 * math.pi)
        xi0 = orc.lp_stokes(n, a, eta, s, th)[0]
        w = lib.orc_lp_ddpr(n, a, eta, s, th)
        assert xi0 == pytest.approx(4.0 * math.pi * w / (a * a), rel=1e-10, abs=1e-300), (n, a, eta, s, th)


def test_cp_quantum_thomson_limit(orc):
    """CP, n = 1, a = 1e-4, eta = 1e-7: the dipole of eps = (x + i y)/sqrt 2 with
    (S1, S2, S3) -> (-S1, S2, -S3), i.e. taken on (e2, e1); fully polarized."""
    a, eta = 1e-4, 1e-7
    rng = np.random.default_rng(5)
    sn = 2.0 * eta / (1.0 + a * a)
    smax = sn / (1.0 + sn)
    for _ in range(300):
        v = rng.uniform(0.01, 0.99)
        th = rng.uniform(0.0, 2.0 * math.pi)
        c = 1.0 - v * smax * (1.0 + a * a + 2.0 * eta) / eta
        d = _dipole(CIRC, c, th)
        r = orc.cp_stokes(1, a, eta, v, th, False)
        assert np.allclose(r[1:], [-d[0], d[1], -d[2]], rtol=0, atol=1e-6), (v, th, r, d)
        assert abs(np.dot(r[1:], r[1:]) - 1.0) < 1e-5


def test_cp_classical_thomson_limit(orc):
    """Classical CP, n = 1, a = 1e-4, cos = 1 - 2 v: (-S1, S2, +S3) of the dipole."""
    a = 1e-4
    rng = np.random.default_rng(6)
    for _ in range(300):
        v = rng.uniform(0.01, 0.99)
        th = rng.uniform(0.0, 2.0 * math.pi)
        d = _dipole(CIRC, 1.0 - 2.0 * v, th)
        r = orc.cp_stokes(1, a, 1e-7, v, th, True)
        assert np.allclose(r[1:], [-d[0], d[1], d[2]], rtol=0, atol=1e-6), (v, th, r, d)


@pytest.mark.parametrize("a", [0.3, 1.0, 3.0, 8.0])
def test_cp_quantum_tends_to_classical(orc, a):
    """At eta = 1e-7 the quantum CP Stokes parameters equal the classical ones with
    S3 reversed (LASER_S3)."""
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 8, 21):
        for _ in range(30):
            v = rng.uniform(0.02, 0.98)
            th = rng.uniform(0.0, 2.0 * math.pi)
            q = orc.cp_stokes(n, a, 1e-7, v, th, False)
            c = orc.cp_stokes(n, a, 1e-7, v, th, True)
            assert np.allclose(q[1:], c[1:] * np.array([1.0, 1.0, -1.0]), rtol=0, atol=1e-5), (n, v, th)


@pytest.mark.parametrize("kind", ["lp", "cp"])
def test_degree_of_polarization_at_most_one(orc, kind):
    """Away from the limits (finite a and eta, many harmonics) the photon is a
    physical state: S1^2 + S2^2 + S3^2 <= 1, wherever the rate is resolved (LP: xi_0
    above 1e-12; a harmonic far above the sum limit at small a has a rate ~ a^(2n),
    ~1e-30 and below, where the ratios are rounding noise and are never drawn)."""
    rng = np.random.default_rng(8)
    for _ in range(500):
        n = int(rng.integers(1, 60))
        a = math.exp(rng.uniform(math.log(0.05), math.log(15.0)))
        eta = math.exp(rng.uniform(math.log(1e-3), math.log(0.5)))
        th = rng.uniform(0.0, 2.0 * math.pi)
        if kind == "lp":
            sn = 2.0 * n * eta / (1.0 + 0.5 * a * a)
            r = orc.lp_stokes(n, a, eta, rng.uniform(0.01, 0.99) * sn / (1.0 + sn), th)
            if not r[0] > 1e-12:
                continue
        else:
            r = orc.cp_stokes(n, a, eta, rng.uniform(0.01, 0.99), th, False)
        assert np.dot(r[1:], r[1:]) <= 1.0 + 1e-9, (n, a, eta, r)
