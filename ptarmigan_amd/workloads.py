"""The benchmark workloads, stated with their sizes.

Each function returns (laser, beam, options) for one collision setup, built
with the reference's own parameter semantics (src/main.rs:339-880).
"""
import math

from .engine import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization, Species

ELECTRON_MASS_MEV = 0.510999
SPEED_OF_LIGHT = 2.997925e8


def focused_lcfa_cascade(n=10_000_000):
    """Focused Gaussian laser (waist 4 um, a0 = 10, 30 fs FWHM, 0.8 um, linear
    polarisation) against a 16.5 GeV electron beam (n primaries, 5 um rms
    transverse, 20 um rms longitudinal Gaussian); LCFA emission and Breit-Wheeler
    pair creation with all secondaries tracked; FP64 state and tables."""
    wavelength = 0.8e-6
    n_cycles = SPEED_OF_LIGHT * 30e-15 / wavelength  # laser:fwhm_duration (src/main.rs:455-460)
    laser = Laser(LaserKind.FAST_FOCUSED_LASER, a0=10.0, wavelength=wavelength, n_cycles=n_cycles, waist=4e-6,
                  envelope=Envelope.GAUSSIAN, polarization=Polarization.LINEAR)
    beam = Beam(n=n, gamma=16500.0 / ELECTRON_MASS_MEV, sigma=0.0, species=Species.ELECTRON, radius=5e-6, radial=0,
                length=20e-6)
    options = CollideOptions(tracking_photons=True, tracking_secondaries=True)
    return laser, beam, options


def plane_wave_lma(n=100_000):
    """Linearly polarised plane-wave pulse, 0.8 um, a0 = 2, 16-cycle cos^2 envelope,
    head-on with a 10 GeV electron beam (n primaries, 1% Gaussian energy spread,
    1 urad rms divergence); LMA nonlinear Compton only (pair creation off)."""
    gamma = 10000.0 / ELECTRON_MASS_MEV
    laser = Laser(LaserKind.PLANE_WAVE, a0=2.0, wavelength=0.8e-6, n_cycles=16.0, envelope=Envelope.COS_SQUARED,
                  polarization=Polarization.LINEAR)
    beam = Beam(n=n, gamma=gamma, sigma=0.01 * gamma, species=Species.ELECTRON, rms_div=1e-6)
    options = CollideOptions(tracking_photons=False)
    return laser, beam, options


def batched_scan(n_coll=64, scale=1.0, seed=2024, polarization=Polarization.LINEAR):
    """An a0 and energy scan as one batch: n_coll linearly polarised plane-wave
    collisions (0.8 um, 16-cycle cos^2), a0 log-spaced over [0.5, 20], electron
    energy linear over [5, 17.5] GeV (1% spread, 1 urad divergence), primaries per
    collision floor(10^U(4, 6)) from a fixed seed (uneven; their sum is about
    1.4e7 for 64 collisions), times `scale`; LMA Compton and pair creation on.
    Returns (lasers, beams, options)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    counts = np.floor(10.0 ** rng.uniform(4.0, 6.0, n_coll)).astype(np.int64)
    counts = np.maximum(1, np.floor(counts * scale)).astype(np.int64)
    a0s = np.exp(np.linspace(np.log(0.5), np.log(20.0), n_coll))
    energies = np.linspace(5000.0, 17500.0, n_coll)  # MeV
    lasers, beams = [], []
    for k in range(n_coll):
        lasers.append(Laser(LaserKind.PLANE_WAVE, a0=float(a0s[k]), wavelength=0.8e-6, n_cycles=16.0,
                            envelope=Envelope.COS_SQUARED, polarization=polarization))
        gamma = energies[k] / ELECTRON_MASS_MEV
        beams.append(Beam(n=int(counts[k]), gamma=gamma, sigma=0.01 * gamma, species=Species.ELECTRON, rms_div=1e-6))
    options = CollideOptions(tracking_photons=True)
    return lasers, beams, options


def surrogate_domain(n, polarization=Polarization.LINEAR, n_coll=64, seed=4096):
    """The learned sampler's domain as one batch: n_coll plane-wave collisions
    (0.8 um, 16-cycle cos^2, the given polarization) with a0 uniform on
    [0.5, 20] and electron energy uniform on [2, 20] GeV (1% spread, 1 urad
    divergence), both stratified (one draw in each of n_coll equal slices, the
    energies shuffled against the a0s), n primaries split evenly between them;
    LMA nonlinear Compton only. The same seed gives the same collisions at any
    n, so a smaller n samples the same distribution. Returns (lasers, beams, options)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    k = np.arange(n_coll)
    a0s = 0.5 + 19.5 * (k + rng.uniform(size=n_coll)) / n_coll
    energies = rng.permutation(2000.0 + 18000.0 * (k + rng.uniform(size=n_coll)) / n_coll)  # MeV
    per = [n // n_coll + (1 if k < n % n_coll else 0) for k in range(n_coll)]
    lasers, beams = [], []
    for k in range(n_coll):
        lasers.append(Laser(LaserKind.PLANE_WAVE, a0=float(a0s[k]), wavelength=0.8e-6, n_cycles=16.0,
                            envelope=Envelope.COS_SQUARED, polarization=polarization))
        gamma = energies[k] / ELECTRON_MASS_MEV
        beams.append(Beam(n=max(1, per[k]), gamma=gamma, sigma=0.01 * gamma, species=Species.ELECTRON, rms_div=1e-6))
    return lasers, beams, CollideOptions(tracking_photons=False)
