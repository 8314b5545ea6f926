"""ctypes view of the C ABI in include/ptg.h (libptg.so).

The library is the product: there is no fallback. Loading fails with
ptarmigan_amd.LibraryMissing when libptg.so has not been built.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptg.so")
TABLES_PATH = os.path.join(HERE, "tables", "ptg_tables.bin")


class LibraryMissing(ImportError):
    pass


class LaserDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("envelope", C.c_int32), ("polarization", C.c_int32),
                ("finite_bandwidth", C.c_int32), ("a0", C.c_double), ("wavelength", C.c_double),
                ("waist", C.c_double), ("n_cycles", C.c_double), ("pol_angle", C.c_double),
                ("chirp_b", C.c_double)]


class CollideOptions(C.Structure):
    _fields_ = [("dt_multiplier", C.c_double), ("rate_increase", C.c_double), ("t_stop", C.c_double),
                ("min_energy", C.c_double), ("max_angle", C.c_double), ("multiplicity", C.c_int64),
                ("discard_bg_e", C.c_int32), ("discard_bg_ph", C.c_int32), ("rr", C.c_int32),
                ("tracking_photons", C.c_int32), ("tracking_secondaries", C.c_int32),
                ("keep_decayed_photons", C.c_int32), ("pol_resolved", C.c_int32),
                ("rotate_stokes_pars", C.c_int32), ("classical", C.c_int32), ("gaunt_factor", C.c_int32),
                ("precision", C.c_int32), ("surrogate", C.c_int32)]


class Particles(C.Structure):
    _fields_ = [("species", C.POINTER(C.c_uint8)), ("r", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)),
                ("r0", C.POINTER(C.c_double)), ("u0", C.POINTER(C.c_double)), ("pol", C.POINTER(C.c_double)),
                ("payload", C.POINTER(C.c_double)), ("parent_chi", C.POINTER(C.c_double)),
                ("interaction_count", C.POINTER(C.c_double)), ("work", C.POINTER(C.c_double)),
                ("weight", C.POINTER(C.c_double)), ("id", C.POINTER(C.c_uint64)),
                ("parent_id", C.POINTER(C.c_uint64))]


class Counters(C.Structure):
    _fields_ = [("lepton_steps", C.c_uint64), ("photon_steps", C.c_uint64), ("emissions", C.c_uint64),
                ("pair_events", C.c_uint64), ("n_secondaries", C.c_uint64), ("n_intermediates", C.c_uint64),
                ("generations", C.c_uint64), ("rejection_trials", C.c_uint64), ("error_flags", C.c_uint32),
                ("graph_stalls", C.c_uint32), ("kernel_ms", C.c_double), ("lepton_ms", C.c_double),
                ("photon_ms", C.c_double), ("lepton_launches", C.c_uint64), ("photon_launches", C.c_uint64),
                ("photon_steps_closed_form", C.c_uint64), ("surrogate_events", C.c_uint64),
                ("surrogate_fallbacks", C.c_uint64), ("graph_launches", C.c_uint64), ("ceiling_restarts", C.c_uint64)]


MLP_NPARAM = 38304


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid_per_cu", C.c_int32), ("n_candidates", C.c_int32),
                ("consistent", C.c_int32), ("cand_block", C.c_int32 * 16), ("cand_grid_per_cu", C.c_int32 * 16),
                ("cand_ms", C.c_double * 16)]


class SurrogateModel(C.Structure):
    _fields_ = [("polarization", C.c_int32), ("valid", C.c_int32), ("a_min", C.c_double), ("a_max", C.c_double),
                ("eta_min", C.c_double), ("eta_max", C.c_double), ("a_use_min", C.c_double),
                ("eta_use_min", C.c_double),
                ("params", C.c_float * MLP_NPARAM)]


class SurrogateTrainDesc(C.Structure):
    _fields_ = [("polarization", C.c_int32), ("groups", C.c_int32), ("steps", C.c_int32), ("batch", C.c_int32),
                ("lr", C.c_double), ("a_min", C.c_double), ("a_max", C.c_double), ("eta_min", C.c_double),
                ("eta_max", C.c_double), ("seed", C.c_uint64)]


class SurrogateReport(C.Structure):
    _fields_ = [("loss", C.c_double * 3), ("val_harmonic_rel", C.c_double), ("val_fraction_abs", C.c_double),
                ("ks_fraction", C.c_double), ("ks_angle", C.c_double), ("valid", C.c_int32), ("groups_used", C.c_int32), ("table_rows", C.c_uint64),
                ("sampler_rows", C.c_uint64), ("gen_ms", C.c_double), ("train_ms", C.c_double)]


class BeamDesc(C.Structure):
    _fields_ = [("n", C.c_uint64), ("species", C.c_int32), ("radial", C.c_int32), ("gamma", C.c_double),
                ("sigma", C.c_double), ("energy_chirp", C.c_double), ("radius", C.c_double),
                ("max_radius", C.c_double), ("length", C.c_double), ("rms_div", C.c_double),
                ("angle", C.c_double), ("collision_plane_angle", C.c_double), ("offset", C.c_double * 3),
                ("pol", C.c_double * 4), ("weight", C.c_double)]


class HistSpec(C.Structure):
    _fields_ = [("species", C.c_int32), ("quantity", C.c_int32), ("weight_fn", C.c_int32),
                ("log_scale", C.c_int32), ("nbins", C.c_uint32), ("fixed_range", C.c_int32),
                ("min", C.c_double), ("max", C.c_double), ("cut", C.c_int32), ("cut_quantity", C.c_int32),
                ("cut_min", C.c_double), ("cut_max", C.c_double), ("scale", C.c_double)]


class StatsSpec(C.Structure):
    _fields_ = [("species", C.c_int32), ("variable", C.c_int32), ("weight", C.c_int32), ("cut", C.c_int32),
                ("cut_quantity", C.c_int32), ("reserved", C.c_int32), ("cut_min", C.c_double),
                ("cut_max", C.c_double), ("shift", C.c_double)]


STATS_N = 9

# ptg_launch_mode
LAUNCH_EAGER, LAUNCH_GRAPH = 0, 1

# every extern "C" symbol include/ptg.h declares
EXPORTS = ("ptg_create", "ptg_destroy", "ptg_last_error", "ptg_set_laser", "ptg_field_info", "ptg_build_beam",
           "ptg_load_primaries", "ptg_run", "ptg_collide", "ptg_fetch", "ptg_push", "ptg_histogram", "ptg_eval",
           "ptg_surrogate_train", "ptg_surrogate_load", "ptg_mlp_probe", "ptg_set_batch", "ptg_autotune",
           "ptg_fetch_showers", "ptg_histogram_2d", "ptg_stats", "ptg_finish",
           "ptg_set_beam_spectrum", "ptg_histogram_noalloc", "ptg_laser_energy", "ptg_set_launch_mode")

_lib = None


def load():
    """Load libptg.so and declare its prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `python -m ptarmigan_amd.build` "
                             "(the HIP library is the only implementation)")
    L = C.CDLL(LIB_PATH)
    P, d, vp = C.POINTER, C.c_double, C.c_void_p
    L.ptg_create.argtypes = [P(vp), C.c_char_p, C.c_int]
    L.ptg_create.restype = C.c_int
    L.ptg_destroy.argtypes = [vp]
    L.ptg_destroy.restype = None
    L.ptg_last_error.argtypes = [vp]
    L.ptg_last_error.restype = C.c_char_p
    L.ptg_set_laser.argtypes = [vp, P(LaserDesc)]
    L.ptg_field_info.argtypes = [vp, P(d), P(d)]
    L.ptg_build_beam.argtypes = [vp, P(BeamDesc), d, C.c_uint64]
    L.ptg_load_primaries.argtypes = [vp, P(Particles), C.c_size_t]
    L.ptg_run.argtypes = [vp, P(CollideOptions), C.c_uint64, P(Counters)]
    L.ptg_collide.argtypes = [vp, P(Particles), C.c_size_t, P(CollideOptions), C.c_uint64, P(Counters)]
    L.ptg_fetch.argtypes = [vp, P(Particles), P(Particles)]
    L.ptg_push.argtypes = [vp, P(C.c_uint8), P(d), P(d), C.c_size_t, C.c_uint32, P(CollideOptions), P(C.c_uint32), P(d)]
    L.ptg_histogram.argtypes = [vp, P(HistSpec), P(d), P(d)]
    L.ptg_histogram_noalloc.argtypes = [vp, P(HistSpec), P(d), P(d)]
    L.ptg_laser_energy.argtypes = [P(LaserDesc), P(d), P(C.c_int32)]
    L.ptg_laser_energy.restype = C.c_int
    L.ptg_eval.argtypes = [vp, C.c_int, P(d), C.c_size_t, C.c_size_t, P(d), C.c_size_t]
    f = C.c_float
    L.ptg_surrogate_train.argtypes = [vp, P(SurrogateTrainDesc), P(SurrogateModel), P(SurrogateReport)]
    L.ptg_surrogate_load.argtypes = [vp, P(SurrogateModel)]
    L.ptg_mlp_probe.argtypes = [vp, P(f), P(f), P(f), C.c_size_t, f, P(f), P(f), P(f)]
    L.ptg_set_batch.argtypes = [vp, P(LaserDesc), P(BeamDesc), C.c_int, C.c_uint64]
    L.ptg_fetch_showers.argtypes = [vp, P(C.c_uint32)]
    L.ptg_histogram_2d.argtypes = [vp, P(HistSpec), P(HistSpec), P(d), P(d)]
    L.ptg_stats.argtypes = [vp, P(StatsSpec), P(d)]
    L.ptg_finish.argtypes = [vp, C.c_int, d, C.c_int, d, d, P(d)]
    L.ptg_set_beam_spectrum.argtypes = [vp, C.c_int, P(d), C.c_size_t, d, d]
    L.ptg_autotune.argtypes = [vp, P(CollideOptions), C.c_uint64, C.c_uint64, P(TuneResult)]
    L.ptg_set_launch_mode.argtypes = [vp, C.c_int]
    for f in ("ptg_set_laser", "ptg_field_info", "ptg_build_beam", "ptg_load_primaries", "ptg_run", "ptg_collide",
              "ptg_fetch", "ptg_push", "ptg_histogram", "ptg_eval"):
        getattr(L, f).restype = C.c_int
    _lib = L
    return L
