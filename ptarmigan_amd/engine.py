"""Host-side interface to the MI355X shower engine (libptg.so).

Mirrors the reference's interface for the per-particle loop:

* ``Laser`` / ``LaserKind`` -- the ``Field`` implementations and their constructors
  (src/field/mod.rs:117-192, src/main.rs:1042-1060);
* ``CollideOptions`` -- field-for-field ``CollideOptions`` (src/main.rs:72-105) with
  the driver's defaults (src/main.rs:339-396, 862-880);
* ``Engine.collide`` -- ``collide`` (src/main.rs:134-289) applied to every primary,
  as the chunked map at src/main.rs:1118-1135, returning the final-state
  particles (``Shower::secondaries``) and decayed photons (``intermediates``).

Everything runs in libptg.so on the GPU; this module only marshals arguments.
"""
import ctypes as C
import dataclasses
import enum
import math

import numpy as np

from . import _abi

ELECTRON_MASS_MEV = 0.510999
SPEED_OF_LIGHT = 2.997925e8


class PtgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ptg error {code}: {msg}")
        self.code = code


class Species(enum.IntEnum):
    ELECTRON = 0
    POSITRON = 1
    PHOTON = 2


class LaserKind(enum.IntEnum):
    PLANE_WAVE = 0
    FOCUSED_LASER = 1
    FAST_PLANE_WAVE = 2
    FAST_FOCUSED_LASER = 3


class Envelope(enum.IntEnum):
    COS_SQUARED = 0
    FLATTOP = 1
    GAUSSIAN = 2


class Polarization(enum.IntEnum):
    LINEAR = 0
    CIRCULAR = 1


@dataclasses.dataclass
class Laser:
    """Laser pulse; `kind` selects the Field implementation as src/main.rs:1042-1060
    does from (focusing, lcfa)."""
    kind: LaserKind
    a0: float
    wavelength: float
    n_cycles: float
    waist: float = 0.0
    envelope: Envelope = Envelope.COS_SQUARED
    polarization: Polarization = Polarization.LINEAR
    pol_angle: float = 0.0
    chirp_b: float = 0.0
    finite_bandwidth: bool = False

    @staticmethod
    def select(a0, wavelength, n_cycles, waist=None, lcfa=False, **kw):
        """The reference's choice of field (src/main.rs:1042-1060)."""
        focusing = waist is not None
        if focusing and not lcfa:
            kind = LaserKind.FOCUSED_LASER
        elif focusing:
            kind = LaserKind.FAST_FOCUSED_LASER
        elif not lcfa:
            kind = LaserKind.PLANE_WAVE
        else:
            kind = LaserKind.FAST_PLANE_WAVE
        return Laser(kind, a0, wavelength, n_cycles, waist or 0.0, **kw)

    def desc(self):
        return _abi.LaserDesc(int(self.kind), int(self.envelope), int(self.polarization), int(self.finite_bandwidth),
                              self.a0, self.wavelength, self.waist, self.n_cycles, self.pol_angle, self.chirp_b)


@dataclasses.dataclass
class CollideOptions:
    """src/main.rs:72-105; defaults are the driver's (src/main.rs:339-396, 862-880)."""
    dt_multiplier: float = 1.0
    rate_increase: float = 1.0
    t_stop: float = math.inf
    discard_bg_e: bool = False
    discard_bg_ph: bool = False
    min_energy: float = 0.0          # MeV
    max_angle: float = math.pi
    multiplicity: int = None
    rr: bool = True
    tracking_photons: bool = True
    tracking_secondaries: bool = True
    keep_decayed_photons: bool = False
    pol_resolved: bool = False
    rotate_stokes_pars: bool = True
    classical: bool = False
    gaunt_factor: bool = False
    precision: int = 0               # 0 = FP64, 1 = mixed
    surrogate: bool = False          # sample LMA emission with the loaded surrogate inside its domain

    def desc(self):
        return _abi.CollideOptions(
            self.dt_multiplier, self.rate_increase, self.t_stop, self.min_energy, self.max_angle,
            -1 if self.multiplicity is None else int(self.multiplicity), int(self.discard_bg_e),
            int(self.discard_bg_ph), int(self.rr), int(self.tracking_photons), int(self.tracking_secondaries),
            int(self.keep_decayed_photons), int(self.pol_resolved), int(self.rotate_stokes_pars),
            int(self.classical), int(self.gaunt_factor), int(self.precision), int(self.surrogate))


@dataclasses.dataclass
class Beam:
    """BeamBuilder parameters for a normally distributed energy spectrum
    (src/particle/builder.rs:173-219, src/main.rs:520-535)."""
    n: int
    gamma: float
    sigma: float = 0.0
    species: Species = Species.ELECTRON
    radius: float = 0.0
    radial: int = 0                  # 0 normal, 1 truncated normal, 2 uniform
    max_radius: float = 0.0
    length: float = 0.0
    rms_div: float = 0.0
    energy_chirp: float = 0.0
    angle: float = 0.0
    collision_plane_angle: float = 0.0
    offset: tuple = (0.0, 0.0, 0.0)
    pol: tuple = (1.0, 0.0, 0.0, 0.0)
    weight: float = 1.0

    def desc(self):
        return _abi.BeamDesc(self.n, int(self.species), self.radial, self.gamma, self.sigma, self.energy_chirp,
                             self.radius, self.max_radius, self.length, self.rms_div, self.angle,
                             self.collision_plane_angle, (C.c_double * 3)(*self.offset),
                             (C.c_double * 4)(*self.pol), self.weight)


PARTICLE_FIELDS = {"species": (np.uint8, None), "r": (np.float64, 4), "u": (np.float64, 4), "r0": (np.float64, 4),
                   "u0": (np.float64, 4), "pol": (np.float64, 4), "payload": (np.float64, None),
                   "parent_chi": (np.float64, None), "interaction_count": (np.float64, None),
                   "work": (np.float64, None), "weight": (np.float64, None), "id": (np.uint64, None),
                   "parent_id": (np.uint64, None)}


def empty_particles(n):
    return {k: np.zeros((n, w) if w else n, dtype=t) for k, (t, w) in PARTICLE_FIELDS.items()}


def particles_struct(arrs):
    kw = {}
    for name, ctype in _abi.Particles._fields_:
        a = arrs.get(name)
        if a is None:
            kw[name] = None
        else:
            if not a.flags["C_CONTIGUOUS"]:
                raise ValueError(f"{name} must be C-contiguous")
            kw[name] = a.ctypes.data_as(ctype)
    return _abi.Particles(**kw)


def laser_energy(laser: Laser):
    """Field::energy (src/field/mod.rs:191) of `laser`: (energy, unit), unit "J" for
    the focused fields and "J/m^2" for the plane waves. Host arithmetic; needs no
    device."""
    lib = _abi.load()
    e, per_area = C.c_double(), C.c_int32()
    rc = lib.ptg_laser_energy(C.byref(laser.desc()), C.byref(e), C.byref(per_area))
    if rc != 0:
        raise PtgError(rc, "ptg_laser_energy: invalid laser parameters")
    return e.value, ("J/m^2" if per_area.value else "J")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """One libptg context on one GPU (one per host thread)."""

    def __init__(self, device=0, tables=None):
        self._lib = _abi.load()
        self._ctx = C.c_void_p()
        rc = self._lib.ptg_create(C.byref(self._ctx), tables.encode() if tables else None, device)
        if rc != 0:
            msg = self._lib.ptg_last_error(self._ctx).decode() if self._ctx else "ptg_create failed"
            if self._ctx:
                self._lib.ptg_destroy(self._ctx)
                self._ctx = None
            raise PtgError(rc, msg)
        self.laser = None

    def close(self):
        if self._ctx:
            self._lib.ptg_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise PtgError(rc, self._lib.ptg_last_error(self._ctx).decode())

    # -- field
    def set_laser(self, laser: Laser):
        self._check(self._lib.ptg_set_laser(self._ctx, C.byref(laser.desc())))
        self.laser = laser

    def field_info(self):
        dt, z0 = C.c_double(), C.c_double()
        self._check(self._lib.ptg_field_info(self._ctx, C.byref(dt), C.byref(z0)))
        return dt.value, z0.value

    # -- primaries
    def build_beam(self, beam: Beam, seed=0, initial_z=None):
        """BeamBuilder::build with the driver's initial_z = ideal_initial_z + 3 sigma_z
        (src/main.rs:1062-1066)."""
        if initial_z is None:
            initial_z = self.field_info()[1] + 3.0 * beam.length
        self._check(self._lib.ptg_build_beam(self._ctx, C.byref(beam.desc()), initial_z, seed))

    def set_beam_spectrum(self, kind=0, table=None, brem=(0.0, 0.0)):
        """ptg_set_beam_spectrum: 0 normal, 1 tabulated CDF ((gamma, cdf) rows), 2 bremsstrahlung (min, max)."""
        t = np.ascontiguousarray(table if table is not None else np.zeros((0, 2)), dtype=np.float64).reshape(-1, 2)
        self._check(self._lib.ptg_set_beam_spectrum(self._ctx, int(kind), _dptr(t), t.shape[0], float(brem[0]),
                                                    float(brem[1])))

    def load_primaries(self, prim):
        n = len(prim["species"])
        self._prim_keep = prim
        self._check(self._lib.ptg_load_primaries(self._ctx, C.byref(particles_struct(prim)), n))

    # -- transport
    def run(self, options: CollideOptions, seed=0):
        cnt = _abi.Counters()
        self._check(self._lib.ptg_run(self._ctx, C.byref(options.desc()), seed, C.byref(cnt)))
        return {k: getattr(cnt, k) for k, _ in _abi.Counters._fields_}

    def fetch(self, counters, intermediates=False):
        sec = empty_particles(int(counters["n_secondaries"]))
        inter = empty_particles(int(counters["n_intermediates"])) if intermediates else None
        self._check(self._lib.ptg_fetch(self._ctx, C.byref(particles_struct(sec)),
                                        C.byref(particles_struct(inter)) if inter is not None else None))
        shower = np.zeros(len(sec["species"]), dtype=np.uint32)
        if len(shower):
            self._check(self._lib.ptg_fetch_showers(self._ctx, shower.ctypes.data_as(C.POINTER(C.c_uint32))))
        sec["shower"] = shower
        return sec, inter

    def collide(self, prim, options: CollideOptions, seed=0):
        """collide() over every primary. Returns (secondaries, intermediates, counters)."""
        self.load_primaries(prim)
        cnt = self.run(options, seed)
        sec, inter = self.fetch(cnt, intermediates=options.keep_decayed_photons)
        return sec, inter, cnt

    def push(self, species, r, u, n_steps, options: CollideOptions = None, with_work=False):
        """Field-only Field::push of each particle for up to n_steps steps (stops at the field edge)."""
        options = options or CollideOptions()
        sp = np.ascontiguousarray(species, dtype=np.uint8)
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 4).copy()
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 4).copy()
        taken = np.zeros(len(sp), dtype=np.uint32)
        work = np.zeros(len(sp))
        self._check(self._lib.ptg_push(self._ctx, sp.ctypes.data_as(C.POINTER(C.c_uint8)), _dptr(r), _dptr(u),
                                       len(sp), n_steps, C.byref(options.desc()),
                                       taken.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(work)))
        return (r, u, taken, work) if with_work else (r, u, taken)

    @staticmethod
    def _hist_spec(species, quantity, weight_fn, log_scale, nbins, range_=None, cut=None, scale=1.0):
        return _abi.HistSpec(int(species), int(quantity), int(weight_fn), int(log_scale), nbins,
                             0 if range_ is None else 1, range_[0] if range_ else 0.0, range_[1] if range_ else 0.0,
                             1 if cut else 0, int(cut[0]) if cut else 0, float(cut[1]) if cut else 0.0,
                             float(cut[2]) if cut else 0.0, float(scale))

    def histogram(self, species, quantity, nbins, weight_fn=0, log_scale=False, range_=None, cut=None, scale=1.0):
        """Histogram::generate_1d over the resident final-state particles; returns (counts, min, max, total).
        `cut` = (quantity, min, max) bins only particles with min <= quantity < max; `scale`
        converts the binned quantity to another unit."""
        spec = self._hist_spec(species, quantity, weight_fn, log_scale, nbins, range_, cut, scale)
        out = np.zeros(nbins)
        tot = C.c_double()
        self._check(self._lib.ptg_histogram(self._ctx, C.byref(spec), _dptr(out), C.byref(tot)))
        return out, spec.min, spec.max, tot.value

    def histogram2d(self, species, qx, qy, nbins_x, nbins_y, weight_fn=0, log_x=False, log_y=False, cut=None,
                    scale=(1.0, 1.0), ranges=(None, None)):
        """Histogram::generate_2d over the resident final-state particles; returns
        (flat counts of nbins_x * nbins_y, (xmin, xmax), (ymin, ymax), total)."""
        sx = self._hist_spec(species, qx, weight_fn, log_x, nbins_x, ranges[0], cut, scale[0])
        sy = self._hist_spec(species, qy, weight_fn, log_y, nbins_y, ranges[1], None, scale[1])
        out = np.zeros(nbins_x * nbins_y)
        tot = C.c_double()
        self._check(self._lib.ptg_histogram_2d(self._ctx, C.byref(sx), C.byref(sy), _dptr(out), C.byref(tot)))
        return out, (sx.min, sx.max), (sy.min, sy.max), tot.value

    def finish(self, stop_mode=0, t_stop=0.0, beam_basis=False, theta=0.0, phi=0.0):
        """ptg_finish: the driver's time fix (0 none, 1 common time, 2 given time in s)
        and beam-frame rotation of the resident final state; returns the common time."""
        t = C.c_double()
        self._check(self._lib.ptg_finish(self._ctx, int(stop_mode), float(t_stop), int(bool(beam_basis)),
                                         float(theta), float(phi), C.byref(t)))
        return t.value

    def stats(self, species, variable, weight=15, cut=None, shift=0.0):
        """ptg_stats: the reductions of one summary statistic over the resident final
        state (weight quantity 15 = unit). `cut` = (quantity, min, max), strict."""
        spec = _abi.StatsSpec(int(species), int(variable), int(weight), 1 if cut else 0, int(cut[0]) if cut else 0, 0,
                              float(cut[1]) if cut else 0.0, float(cut[2]) if cut else 0.0, float(shift))
        out = np.zeros(_abi.STATS_N)
        self._check(self._lib.ptg_stats(self._ctx, C.byref(spec), _dptr(out)))
        return out

    def eval(self, fn, x, out_w):
        """Batched evaluation of one device function (ptg_eval_fn ids) for parity tests."""
        x = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float64)))
        out = np.zeros((x.shape[0], out_w))
        self._check(self._lib.ptg_eval(self._ctx, int(fn), _dptr(x), x.shape[0], x.shape[1], _dptr(out), out_w))
        return out


    # -- batches of collisions and launch tuning
    def set_batch(self, lasers, beams, seed=0):
        """ptg_set_batch: lasers differing only in a0, one beam per laser."""
        if len(lasers) != len(beams):
            raise ValueError("one beam per laser")
        n = len(lasers)
        la = (_abi.LaserDesc * n)(*[l.desc() for l in lasers])
        be = (_abi.BeamDesc * n)(*[b.desc() for b in beams])
        self._check(self._lib.ptg_set_batch(self._ctx, la, be, n, seed))
        self.laser = lasers[0]

    def set_launch_mode(self, graph: bool):
        """LMA generations in captured sequences (graph=True, the default) or with
        the counts read back after every launch (graph=False, per-kernel timings)."""
        self._check(self._lib.ptg_set_launch_mode(self._ctx, _abi.LAUNCH_GRAPH if graph else _abi.LAUNCH_EAGER))

    def autotune(self, options: CollideOptions, seed=0, max_primaries=0):
        r = _abi.TuneResult()
        self._check(self._lib.ptg_autotune(self._ctx, C.byref(options.desc()), seed, max_primaries, C.byref(r)))
        k = r.n_candidates
        return {"block": r.block, "grid_per_cu": r.grid_per_cu, "consistent": bool(r.consistent),
                "candidates": [(r.cand_block[i], r.cand_grid_per_cu[i], r.cand_ms[i]) for i in range(k)]}

    # -- learned sampler
    def surrogate_train(self, polarization, groups=8192, steps=6000, batch=4096, lr=2e-3, a_range=(0.25, 20.0),
                        eta_range=(0.01, 0.3), seed=1):
        """Trains the surrogate on the device (ptg_surrogate_train) and loads it.
        Returns (model, report) as ctypes structures."""
        d = _abi.SurrogateTrainDesc(int(polarization), int(groups), int(steps), int(batch), float(lr),
                                    float(a_range[0]), float(a_range[1]), float(eta_range[0]), float(eta_range[1]),
                                    int(seed))
        model = _abi.SurrogateModel()
        rep = _abi.SurrogateReport()
        self._check(self._lib.ptg_surrogate_train(self._ctx, C.byref(d), C.byref(model), C.byref(rep)))
        return model, rep

    def surrogate_load(self, model):
        self._check(self._lib.ptg_surrogate_load(self._ctx, C.byref(model)))

    def mlp_probe(self, params, x, t, w_t=1.0):
        """Forward and backward of the surrogate network on the device: (outputs n x 32,
        gradient of the mean loss, loss sums)."""
        params = np.ascontiguousarray(params, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n = x.shape[0]
        out = np.zeros((n, 32), dtype=np.float32)
        grads = np.zeros(_abi.MLP_NPARAM, dtype=np.float32)
        loss = np.zeros(3, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self._lib.ptg_mlp_probe(self._ctx, params.ctypes.data_as(fp), x.ctypes.data_as(fp),
                                            t.ctypes.data_as(fp), n, float(w_t), out.ctypes.data_as(fp),
                                            grads.ctypes.data_as(fp), loss.ctypes.data_as(fp)))
        return out, grads, loss


class Fn(enum.IntEnum):
    LCFA_RATE = 0
    LCFA_SAMPLE = 1
    AUX_T = 2
    AIRY = 3
    FIELDS = 4
    GAUNT = 5
    PAIR_SPECTRUM = 6
    LCFA_STOKES = 7
    LP_RATE = 8
    LP_HARMONIC = 9
    DOUBLE_BESSEL = 10
    LP_SAMPLE = 11
    BW_RATE = 12
    BW_HARMONIC = 13
    BW_SAMPLE = 14
    BESSEL_J = 15
    CP_RATE = 16
    CP_HARMONIC = 17
    CP_PARTIAL_RATE = 18
    CP_SAMPLE = 19
    LP_CL_RATE = 20
    LP_CL_HARMONIC = 21
    LP_CL_CEILING = 22
    LP_CL_SAMPLE = 23
    CPBW_RATE = 24
    CPBW_PARTIAL_RATE = 25
    CPBW_HARMONIC = 26
    CPBW_SAMPLE = 27
    LP_CEILING = 28
