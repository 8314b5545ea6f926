"""ctypes bindings to the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module. The oracle is the checker, never the product path.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
TABLES = os.path.join(HERE, "..", "ptarmigan_amd", "tables", "ptg_tables.bin")

_lib = None


# The oracle shares the boundary's struct layouts (include/ptg.h); it does not
# share any of its code paths.
sys.path.insert(0, os.path.join(HERE, ".."))
from ptarmigan_amd._abi import BeamDesc, CollideOptions, LaserDesc, Particles  # noqa: E402


FIELDS = {"species": np.uint8, "r": (np.float64, 4), "u": (np.float64, 4), "r0": (np.float64, 4),
          "u0": (np.float64, 4), "pol": (np.float64, 4), "payload": np.float64, "parent_chi": np.float64,
          "interaction_count": np.float64, "work": np.float64, "weight": np.float64, "id": np.uint64,
          "parent_id": np.uint64}


def alloc_particles(n):
    """Dict of zeroed numpy arrays for n particles plus the ctypes struct viewing them."""
    arrs = {}
    for k, spec in FIELDS.items():
        if isinstance(spec, tuple):
            arrs[k] = np.zeros((n, spec[1]), dtype=spec[0])
        else:
            arrs[k] = np.zeros(n, dtype=spec)
    return arrs, particles_struct(arrs)


def particles_struct(arrs):
    kw = {}
    for name, ctype in Particles._fields_:
        a = arrs.get(name)
        if a is None:
            kw[name] = None
        else:
            assert a.flags["C_CONTIGUOUS"]
            kw[name] = a.ctypes.data_as(ctype)
    return Particles(**kw)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_library(path):
    """Loads the oracle from `path` (a build of the same sources, e.g. the
    -O3 -march=native timing build) instead of liboracle.so; call before lib()."""
    global LIB, _lib
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    LIB = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        d = C.c_double
        P = C.POINTER
        L.orc_init.argtypes = [C.c_char_p]
        L.orc_collide.restype = C.c_void_p
        L.orc_collide.argtypes = [P(LaserDesc), P(CollideOptions), P(Particles), C.c_size_t, C.c_uint64, C.c_int, P(d)]
        L.orc_run_counts.argtypes = [C.c_void_p, P(C.c_uint64)]
        L.orc_run_fetch.argtypes = [C.c_void_p, P(Particles), P(Particles)]
        L.orc_run_free.argtypes = [C.c_void_p]
        L.orc_push.argtypes = [P(LaserDesc), P(C.c_uint8), P(d), P(d), C.c_size_t, C.c_uint32, d, C.c_int, P(C.c_uint32),
                               P(d)]
        L.orc_build_beam.argtypes = [P(BeamDesc), d, C.c_uint64, P(Particles)]
        for f in ("orc_laser_max_timestep", "orc_laser_ideal_initial_z"):
            getattr(L, f).restype = d
            getattr(L, f).argtypes = [P(LaserDesc)]
        L.orc_laser_fields.argtypes = [P(LaserDesc), P(d), C.c_size_t, P(d)]
        L.orc_lcfa_rate.restype = d
        L.orc_lcfa_rate.argtypes = [d, d]
        L.orc_lcfa_sample.argtypes = [d, d, d, d, d, P(d)]
        L.orc_lcfa_stokes.argtypes = [P(d), d, d, P(d), P(d), P(d)]
        L.orc_gaunt.restype = d
        L.orc_gaunt.argtypes = [d]
        L.orc_aux_t.argtypes = [d, P(d)]
        for f, n in (("orc_pair_spectrum", 3), ("orc_pair_spectrum_ceiling", 2),
                     ("orc_pair_angular_spectrum", 4), ("orc_pair_angular_spectrum_ceiling", 3)):
            getattr(L, f).restype = d
            getattr(L, f).argtypes = [d] * n
        L.orc_pair_probability.restype = d
        L.orc_pair_probability.argtypes = [P(d), P(d), d, P(d), d, P(d)]
        L.orc_airy.restype = C.c_int
        L.orc_airy.argtypes = [d, P(d)]
        L.orc_pwmci.restype = C.c_int
        L.orc_pwmci.argtypes = [P(d), C.c_int, d, C.c_int, C.c_int, P(d)]
        L.orc_xoshiro_stream.argtypes = [C.c_uint64, C.c_int, C.c_size_t, P(C.c_uint64)]
        # LMA (linearly polarised plane wave)
        L.orc_lp_rate.restype = C.c_int
        L.orc_lp_rate.argtypes = [d, d, P(d)]
        L.orc_lp_harmonic.restype = C.c_int
        L.orc_lp_harmonic.argtypes = [d, d, d]
        L.orc_double_bessel.argtypes = [C.c_int, d, d, P(d)]
        L.orc_double_bessel_recurrence.argtypes = [C.c_int, d, d, P(d)]
        L.orc_lp_stokes.argtypes = [C.c_int, d, d, d, d, P(d)]
        L.orc_lp_cl_stokes.argtypes = [C.c_int, d, d, d, P(d)]
        L.orc_cp_stokes.argtypes = [C.c_int, d, d, d, d, C.c_int, P(d)]
        L.orc_lp_ddpr.restype = d
        L.orc_lp_ddpr.argtypes = [C.c_int, d, d, d, d]
        L.orc_lp_partial_rate.restype = d
        L.orc_lp_partial_rate.argtypes = [C.c_int, d, d, P(d)]
        L.orc_lp_sum_limit.restype = C.c_int
        L.orc_lp_sum_limit.argtypes = [d, d]
        L.orc_lp_linear_rate.restype = d
        L.orc_lp_linear_rate.argtypes = [d, d]
        L.orc_lp_theta_bound.restype = d
        L.orc_lp_theta_bound.argtypes = [C.c_int, d, d, d]
        L.orc_lp_ceiling.restype = d
        L.orc_lp_ceiling.argtypes = [C.c_int, d, d]
        L.orc_lp_sample.argtypes = [d, d, C.c_uint64, C.c_int, C.c_size_t, P(d)]
        # LP Breit-Wheeler
        L.orc_bw_partial_rate.argtypes = [C.c_int, d, d, C.c_int, P(d)]
        L.orc_bw_ddpr.restype = d
        L.orc_bw_ddpr.argtypes = [C.c_int, d, d, d, d, d, d]
        L.orc_bw_ceiling.restype = d
        L.orc_bw_ceiling.argtypes = [C.c_int, d, d, d, d]
        L.orc_bw_theta_bound.restype = d
        L.orc_bw_theta_bound.argtypes = [C.c_int, d, d, d]
        L.orc_bw_sum_limits.restype = C.c_int
        L.orc_bw_sum_limits.argtypes = [d, d, P(C.c_int)]
        L.orc_bw_rate.restype = C.c_int
        L.orc_bw_rate.argtypes = [d, d, P(d)]
        L.orc_bw_probability.restype = d
        L.orc_bw_probability.argtypes = [d, d, P(d), d, P(d)]
        L.orc_bw_by_summation.argtypes = [d, d, P(d)]
        L.orc_bw_harmonic.restype = C.c_int
        L.orc_bw_harmonic.argtypes = [d, d, d, d]
        L.orc_bw_sample.argtypes = [d, d, d, d, C.c_uint64, C.c_int, C.c_size_t, P(d)]
        i = C.c_int
        L.orc_bessel_j_pm.argtypes = [i, d, P(d)]
        L.orc_bessel_j.argtypes = [i, d]
        L.orc_bessel_j.restype = d
        L.orc_bessel_j_parts.argtypes = [i, d, P(d)]
        for name in ("orc_cp_ddpr",):
            getattr(L, name).argtypes = [i, d, d, d]
            getattr(L, name).restype = d
        for name in ("orc_cp_ceiling", "orc_cp_partial_rate", "orc_cp_spectrum_low_eta"):
            getattr(L, name).argtypes = [i, d, d]
            getattr(L, name).restype = d
        for name in ("orc_cp_integrated_spectrum_low_eta", "orc_cp_classical_ceiling"):
            getattr(L, name).argtypes = [i, d]
            getattr(L, name).restype = d
        L.orc_lp_cl_partial_rate.argtypes = [i, d, P(d)]
        L.orc_lp_cl_partial_rate.restype = d
        L.orc_lp_cl_ceiling.argtypes = [i, d]
        L.orc_lp_cl_ceiling.restype = d
        L.orc_lp_cl_ddpr.argtypes = [i, d, d, d]
        L.orc_lp_cl_ddpr.restype = d
        L.orc_lp_theta_bound_low_eta.argtypes = [i, d, d]
        L.orc_lp_theta_bound_low_eta.restype = d
        L.orc_lp_cl_rate.argtypes = [d, d, P(d)]
        L.orc_lp_cl_harmonic.argtypes = [d, d]
        L.orc_lp_cl_harmonic_batch.argtypes = [d, P(d), C.c_size_t, P(C.c_int)]
        L.orc_lp_cl_v_marginal.argtypes = [i, d, i, i, P(d)]
        L.orc_lp_cl_sample.argtypes = [d, d, C.c_uint64, i, C.c_size_t, P(d)]
        L.orc_cpbw_partial_rate.argtypes = [i, d, d, P(d)]
        L.orc_cpbw_fully_resolved.argtypes = [i, d, d, d, d, d, d, d]
        L.orc_cpbw_fully_resolved.restype = d
        L.orc_cpbw_sum_limits.argtypes = [d, d, P(i)]
        L.orc_cpbw_total.argtypes = [d, d, P(d)]
        L.orc_cpbw_by_summation.argtypes = [d, d, P(d)]
        L.orc_cpbw_probability.argtypes = [d, d, P(d), d, P(d)]
        L.orc_cpbw_probability.restype = d
        L.orc_cpbw_harmonic.argtypes = [d, d, d, d]
        L.orc_cpbw_sample.argtypes = [d, d, d, d, d, C.c_uint64, i, C.c_size_t, P(d)]
        L.orc_cp_grid_max.argtypes = [i, d, d, i]
        L.orc_cp_grid_max.restype = d
        L.orc_cp_rate.argtypes = [d, d, P(d)]
        L.orc_cp_classical_rate.argtypes = [d, d, P(d)]
        L.orc_cp_harmonic.argtypes = [d, d, d]
        L.orc_cp_classical_harmonic.argtypes = [d, d]
        L.orc_cp_cdf_row.argtypes = [d, d, i, P(d)]
        L.orc_cp_cdf_row.restype = d
        L.orc_cp_sample.argtypes = [d, d, i, C.c_uint64, i, C.c_size_t, P(d)]
        rc = L.orc_init(os.path.abspath(TABLES).encode())
        if rc != 0:
            raise RuntimeError(f"oracle: cannot load tables from {TABLES}")
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def laser_desc(kind=3, envelope=2, polarization=0, a0=10.0, wavelength=0.8e-6, waist=4e-6, n_cycles=10.0,
               pol_angle=0.0, chirp_b=0.0, finite_bandwidth=0):
    return LaserDesc(kind, envelope, polarization, finite_bandwidth, a0, wavelength, waist, n_cycles, pol_angle, chirp_b)


def collide_options(**kw):
    o = dict(dt_multiplier=1.0, rate_increase=1.0, t_stop=np.inf, min_energy=0.0, max_angle=np.pi,
             multiplicity=-1, discard_bg_e=0, discard_bg_ph=0, rr=1, tracking_photons=1,
             tracking_secondaries=1, keep_decayed_photons=0, pol_resolved=0, rotate_stokes_pars=1,
             classical=0, gaunt_factor=0, precision=0, reserved=0)
    o.update(kw)
    return CollideOptions(**o)


def collide(laser, options, prim, seed=0, nthreads=1):
    """Run the oracle collide() over primaries `prim` (dict of arrays). Returns
    (secondaries dict, intermediates dict, counters dict, seconds)."""
    L = lib()
    n = len(prim["species"])
    ps = particles_struct(prim)
    secs = C.c_double(0.0)
    h = L.orc_collide(C.byref(laser), C.byref(options), C.byref(ps), n, seed, nthreads, C.byref(secs))
    c = np.zeros(6, dtype=np.uint64)
    L.orc_run_counts(h, c.ctypes.data_as(C.POINTER(C.c_uint64)))
    sec, sec_s = alloc_particles(int(c[4]))
    inter, inter_s = alloc_particles(int(c[5]))
    L.orc_run_fetch(h, C.byref(sec_s), C.byref(inter_s))
    L.orc_run_free(h)
    cnt = dict(lepton_steps=int(c[0]), photon_steps=int(c[1]), emissions=int(c[2]), pair_events=int(c[3]))
    return sec, inter, cnt, secs.value


def build_beam(beam, initial_z, seed=0):
    arrs, s = alloc_particles(int(beam.n))
    lib().orc_build_beam(C.byref(beam), initial_z, seed, C.byref(s))
    return arrs


def push(laser, species, r, u, n_steps, dt_multiplier=1.0, eqn=0, with_work=False):
    """Field::push repeated up to n_steps times per particle, stopping at the field edge."""
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 4).copy()
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 4).copy()
    sp = np.ascontiguousarray(species, dtype=np.uint8)
    steps = np.zeros(len(sp), dtype=np.uint32)
    work = np.zeros(len(sp))
    lib().orc_push(C.byref(laser), sp.ctypes.data_as(C.POINTER(C.c_uint8)), dptr(r), dptr(u), len(sp), n_steps,
                   dt_multiplier, eqn, steps.ctypes.data_as(C.POINTER(C.c_uint32)), dptr(work))
    return (r, u, steps, work) if with_work else (r, u, steps)


def fields(laser, r):
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((len(r), 7))
    lib().orc_laser_fields(C.byref(laser), dptr(r), len(r), dptr(out))
    return out


def airy(x):
    out = np.zeros(4)
    ok = lib().orc_airy(x, dptr(out))
    return out, ok


def pwmci(table, v, invert=False, extrapolate=False):
    t = np.ascontiguousarray(table, dtype=np.float64)
    out = np.zeros(1)
    ok = lib().orc_pwmci(dptr(t), len(t), v, int(invert), int(extrapolate), dptr(out))
    return out[0] if ok else None


def lcfa_sample(chi, gamma, r1, r2, r3):
    out = np.zeros(4)
    lib().orc_lcfa_sample(chi, gamma, r1, r2, r3, dptr(out))
    return out


def aux_t(chi):
    out = np.zeros(2)
    lib().orc_aux_t(chi, dptr(out))
    return out


def lcfa_stokes(k, chi, gamma, v, w):
    out = np.zeros(4)
    k, v, w = (np.ascontiguousarray(x, dtype=np.float64) for x in (k, v, w))
    lib().orc_lcfa_stokes(dptr(k), chi, gamma, dptr(v), dptr(w), dptr(out))
    return out


def pair_probability(ell, sv, chi, a_perp, dt):
    out = np.zeros(4)
    ell, sv, a_perp = (np.ascontiguousarray(x, dtype=np.float64) for x in (ell, sv, a_perp))
    p = lib().orc_pair_probability(dptr(ell), dptr(sv), chi, dptr(a_perp), dt, dptr(out))
    return p, out


def xoshiro(seed, n, jumps=0):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_xoshiro_stream(seed, jumps, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


# ---- LMA probes
def lp_rate(a, eta):
    out = np.zeros(1)
    ok = lib().orc_lp_rate(a, eta, dptr(out))
    return out[0] if ok else None


def lp_harmonic(a, eta, frac):
    return lib().orc_lp_harmonic(a, eta, frac)


def double_bessel(n, x, y, recurrence=False):
    """(J(n-2..n+2, x, y)); evaluate() picks the saddle point where the reference does."""
    out = np.zeros(5)
    f = lib().orc_double_bessel_recurrence if recurrence else lib().orc_double_bessel
    f(n, x, y, dptr(out))
    return out


def lp_partial_rate(n, a, eta):
    mx = np.zeros(1)
    v = lib().orc_lp_partial_rate(n, a, eta, dptr(mx))
    return v, mx[0]


def lp_sample(a, eta, count, seed=0, fixed_n=0):
    """count samples of (n, s, theta, S1, S2)."""
    out = np.zeros((count, 5))
    lib().orc_lp_sample(a, eta, seed, fixed_n, count, dptr(out))
    return out


# ---- LP Breit-Wheeler probes
def bw_partial_rate(n, a, eta, method=0):
    """(par, perp) partial rate of harmonic n; method 0 = integrate(), 1 direct, 2 adaptive."""
    out = np.zeros(2)
    lib().orc_bw_partial_rate(n, a, eta, method, dptr(out))
    return out


def bw_rate(a, eta):
    """(status, [par, perp]): status 0 negligible, 1 mid-range table, 2 high-range table, -1 outside."""
    out = np.zeros(2)
    k = lib().orc_bw_rate(a, eta, dptr(out))
    return k, out


def bw_sum_limits(a, eta):
    hi = C.c_int()
    lo = lib().orc_bw_sum_limits(a, eta, C.byref(hi))
    return lo, hi.value


def bw_by_summation(a, eta):
    out = np.zeros(48)
    lib().orc_bw_by_summation(a, eta, dptr(out))
    return out.reshape(16, 3)


def bw_sample(a, eta, sv1, sv2, count, seed=0, fixed_n=0):
    out = np.zeros((count, 3))
    lib().orc_bw_sample(a, eta, sv1, sv2, seed, fixed_n, count, dptr(out))
    return out


# ---- Bessel J and CP nonlinear Compton (oracle/cp.cpp)
def bessel_j_pm(n, x):
    out = np.zeros(3)
    lib().orc_bessel_j_pm(n, x, dptr(out))
    return out


def cp_rate(a, eta, classical=False):
    out = C.c_double(0.0)
    f = lib().orc_cp_classical_rate if classical else lib().orc_cp_rate
    return out.value if f(a, eta, C.byref(out)) else None


def cp_cdf_row(a, eta, classical=False):
    out = np.zeros((32, 2))
    total = lib().orc_cp_cdf_row(a, eta, int(classical), dptr(out))
    return total, out


def cp_sample(a, eta, count, classical=False, seed=0, fixed_n=0):
    """(n, s, theta, S1, S2, S3, trials) per sample from one seeded stream."""
    out = np.zeros((count, 7))
    lib().orc_cp_sample(a, eta, int(classical), seed, fixed_n, count, dptr(out))
    return out


# ---- LP nonlinear Compton, classical (oracle/lma.cpp)
def lp_cl_rate(a, eta):
    out = C.c_double(0.0)
    return out.value if lib().orc_lp_cl_rate(a, eta, C.byref(out)) else None


def lp_cl_partial_rate(n, a):
    mx = C.c_double(0.0)
    r = lib().orc_lp_cl_partial_rate(n, a, C.byref(mx))
    return r, mx.value


def lp_cl_sample(a, eta, count, seed=0, fixed_n=0):
    """(n, s, theta, S1, S2) per sample from one seeded stream."""
    out = np.zeros((count, 5))
    lib().orc_lp_cl_sample(a, eta, seed, fixed_n, count, dptr(out))
    return out


# ---- CP nonlinear Breit-Wheeler (oracle/cp.cpp)
def cpbw_partial_rate(n, a, eta):
    out = np.zeros(2)
    lib().orc_cpbw_partial_rate(n, a, eta, dptr(out))
    return out


def cpbw_total(a, eta):
    """(method: 0 zero, 1 summation, 2 lookup, (S3 = +1, S3 = -1) totals)."""
    out = np.zeros(2)
    k = lib().orc_cpbw_total(a, eta, dptr(out))
    return k, out


def cpbw_by_summation(a, eta):
    out = np.zeros(2)
    lib().orc_cpbw_by_summation(a, eta, dptr(out))
    return out


def cpbw_sample(a, eta, sv1, sv2, sv3, count, seed=0, fixed_n=0):
    out = np.zeros((count, 3))
    lib().orc_cpbw_sample(a, eta, sv1, sv2, sv3, seed, fixed_n, count, dptr(out))
    return out


def lp_stokes(n, a, eta, s, theta):
    """(xi_0, S1, S2, S3) of the quantum LP photon at harmonic n, (s, theta)."""
    out = np.zeros(4)
    lib().orc_lp_stokes(int(n), a, eta, s, theta, dptr(out))
    return out


def lp_cl_stokes(n, a, v, theta):
    """(1, S1, S2, S3) of the classical LP photon at harmonic n, (v, theta)."""
    out = np.zeros(4)
    lib().orc_lp_cl_stokes(int(n), a, v, theta, dptr(out))
    return out


def cp_stokes(n, a, eta, v, theta, classical=False):
    """(1, S1, S2, S3) of the CP photon at harmonic n, fraction v of the largest s, azimuth theta."""
    out = np.zeros(4)
    lib().orc_cp_stokes(int(n), a, eta, v, theta, int(classical), dptr(out))
    return out
