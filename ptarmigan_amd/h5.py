"""HDF5 particle dump (output:dump_all_particles: hdf5) and beam import
(beam:from_hdf5), over libptg_h5.so (csrc/h5io.cpp, the HDF5 C library).

The layout is the reference's (src/main.rs:1289-1715, documented in
docs/output.md): `build/`, `config/` (the run's parameters, scalars), and
`final-state/{photon,electron,positron}` (+ `intermediate-state/photon` when
decayed photons are kept) with one dataset per particle property: weight,
a0_at_creation (alias xi), parent_chi, n_pos or n_gamma, id, parent_id,
absorption, polarization (alias polarisation), position, position_at_creation
(positrons), momentum; `final-state/laser/` holds the pulse energy and the
energy absorbed from it. Four-vectors and Stokes vectors are datasets of a
4 x f64 array type, flags enumerations over u8, strings fixed length, each
dataset carries `unit` and `desc` attributes as the reference writes them.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libptg_h5.so")
_lib = None

K_F64, K_U64, K_I64, K_BOOL, K_ENUM = 0, 1, 2, 3, 4
ELECTRON_MASS_MEV = 0.510999
SPEED_OF_LIGHT = 2.997925e8
ELEMENTARY_CHARGE = 1.602177e-19
VERSION = "1.0.0"

# unit conversions from the default units (m, MeV, MeV/c) of src/output/units.rs:140-170
UNITS = {"um": 1.0e6, "micron": 1.0e6, "mm": 1.0e3, "m": 1.0, "J": 1.0e6 * ELEMENTARY_CHARGE, "GeV": 1.0e-3,
         "MeV": 1.0, "kg/m/s": 1.0e6 * ELEMENTARY_CHARGE / SPEED_OF_LIGHT, "GeV/c": 1.0e-3, "MeV/c": 1.0}


class H5Error(RuntimeError):
    pass


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise H5Error("HDF5 support is not built (no libhdf5 was found at build time)")
        L = C.CDLL(LIB)
        L.ptg_h5_last_error.restype = C.c_char_p
        L.ptg_h5_create.restype = C.c_int64
        L.ptg_h5_create.argtypes = [C.c_char_p]
        L.ptg_h5_open.restype = C.c_int64
        L.ptg_h5_open.argtypes = [C.c_char_p]
        L.ptg_h5_close.argtypes = [C.c_int64]
        L.ptg_h5_group.argtypes = [C.c_int64, C.c_char_p]
        L.ptg_h5_write_scalar.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.POINTER(C.c_char_p),
                                          C.POINTER(C.c_uint8), C.c_int, C.c_char_p, C.c_char_p]
        L.ptg_h5_write_string.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]
        L.ptg_h5_write_array.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.c_uint64, C.c_int, C.c_char_p,
                                         C.c_char_p]
        L.ptg_h5_alias.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_char_p]
        L.ptg_h5_exists.argtypes = [C.c_int64, C.c_char_p]
        L.ptg_h5_info.argtypes = [C.c_int64, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ptg_h5_read.argtypes = [C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.c_uint64, C.c_int]
        L.ptg_h5_read_string.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_int]
        L.ptg_h5_read_attribute.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def _b(s):
    return None if s is None else str(s).encode()


def _check(rc):
    if rc < 0:
        raise H5Error(lib().ptg_h5_last_error().decode())
    return rc


class File:
    """An HDF5 file opened for writing (mode "w", truncating) or reading ("r")."""

    def __init__(self, path, mode="r"):
        L = lib()
        self.h = _check(L.ptg_h5_create(_b(path)) if mode == "w" else L.ptg_h5_open(_b(path)))

    def close(self):
        if self.h is not None:
            _check(lib().ptg_h5_close(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- writing
    def group(self, path):
        _check(lib().ptg_h5_group(self.h, _b(path)))

    def f64(self, path, v, unit=None, desc=None):
        x = C.c_double(float(v))
        _check(lib().ptg_h5_write_scalar(self.h, _b(path), K_F64, C.byref(x), None, None, 0, _b(unit), _b(desc)))

    def u64(self, path, v, unit=None, desc=None):
        x = C.c_uint64(int(v))
        _check(lib().ptg_h5_write_scalar(self.h, _b(path), K_U64, C.byref(x), None, None, 0, _b(unit), _b(desc)))

    def flag(self, path, v, unit=None, desc=None):
        x = C.c_uint8(1 if v else 0)
        _check(lib().ptg_h5_write_scalar(self.h, _b(path), K_BOOL, C.byref(x), None, None, 0, _b(unit), _b(desc)))

    def enum(self, path, names, value, unit=None, desc=None):
        """A u8 enumeration with members `names` (values 0, 1, ...), holding `value`."""
        arr = (C.c_char_p * len(names))(*[_b(n) for n in names])
        vals = (C.c_uint8 * len(names))(*range(len(names)))
        x = C.c_uint8(int(value))
        _check(lib().ptg_h5_write_scalar(self.h, _b(path), K_ENUM, C.byref(x), arr, vals, len(names), _b(unit),
                                         _b(desc)))

    def string(self, path, v, unit=None, desc=None):
        _check(lib().ptg_h5_write_string(self.h, _b(path), _b(v), _b(unit), _b(desc)))

    def array(self, path, a, unit=None, desc=None):
        """1-D dataset of f64 or u64 (from the array's dtype); an n x 4 or n x 3
        array becomes n elements of a 4- or 3-element f64 array type, a shape
        (4,) or (3,) array one such element in a scalar-shaped dataset of one."""
        a = np.asarray(a)
        kind = K_U64 if a.dtype.kind == "u" else K_F64
        a = np.ascontiguousarray(a, dtype=np.uint64 if kind == K_U64 else np.float64)
        width = a.shape[1] if a.ndim == 2 else 1
        n = a.shape[0] if a.ndim >= 1 else 1
        _check(lib().ptg_h5_write_array(self.h, _b(path), kind, a.ctypes.data, n, width, _b(unit), _b(desc)))

    def vector(self, path, v, unit=None, desc=None):
        """One three- or four-vector (a single element of the array type)."""
        self.array(path, np.asarray(v, dtype=np.float64).reshape(1, -1), unit, desc)

    def alias(self, group, target, alias):
        _check(lib().ptg_h5_alias(self.h, _b(group), _b(target), _b(alias)))

    # -- reading
    def exists(self, path):
        return _check(lib().ptg_h5_exists(self.h, _b(path))) == 1

    def info(self, path):
        n, w, c = C.c_uint64(), C.c_int(), C.c_int()
        _check(lib().ptg_h5_info(self.h, _b(path), C.byref(n), C.byref(w), C.byref(c)))
        return n.value, w.value, c.value

    def read(self, path, dtype=np.float64):
        """The dataset as an array: (n,) or (n, width) (array elements)."""
        n, w, _ = self.info(path)
        kind = K_U64 if np.dtype(dtype).kind == "u" else K_F64
        out = np.zeros((n, w) if w > 1 else n, dtype=np.uint64 if kind == K_U64 else np.float64)
        _check(lib().ptg_h5_read(self.h, _b(path), kind, out.ctypes.data, n, w))
        return out

    def read_string(self, path):
        buf = C.create_string_buffer(4096)
        _check(lib().ptg_h5_read_string(self.h, _b(path), buf, len(buf)))
        return buf.value.decode()

    def attribute(self, path, name):
        buf = C.create_string_buffer(4096)
        _check(lib().ptg_h5_read_attribute(self.h, _b(path), _b(name), buf, len(buf)))
        return buf.value.decode()


# ---------------------------------------------------------------- the particle dump
def _species_groups(f, root, sec, species, units, with_creation):
    """One group of final-state (or intermediate) particles (src/main.rs:1473-1662)."""
    ul, ue, up = units["length"], units["energy"], units["momentum"]
    sl, se, sp = UNITS[ul], UNITS[ue], UNITS[up]
    m = sec["species"] == species
    name = {0: "electron", 1: "positron", 2: "photon", 3: "photon"}[species]
    g = f"{root}/{name}"
    f.group(g)
    real = {0: "real electrons", 1: "real positrons", 2: "real photons", 3: "real photons"}[species]
    macro = {0: "macroelectron", 1: "macropositron", 2: "macrophoton", 3: "macrophoton"}[species]
    f.array(f"{g}/weight", sec["weight"][m], "1", f"number of {real} each {macro} represents")
    if species in (1, 2, 3):
        where = "emission" if species != 1 else "creation"
        f.array(f"{g}/a0_at_creation", sec["payload"][m], "1",
                f"normalized amplitude (RMS under LMA) at point of {where}")
        f.alias(g, "a0_at_creation", "xi")
        f.array(f"{g}/parent_chi", sec["parent_chi"][m], "1",
                f"quantum parameter (RMS under LMA) of parent at point of {where}")
    who = name
    if species in (2, 3):
        f.array(f"{g}/n_pos", sec["interaction_count"][m], "1", "total probability of pair creation for the photon")
        f.array(f"{g}/id", sec["id"][m], None, "unique ID of the photon")
        f.array(f"{g}/parent_id", sec["parent_id"][m], None,
                "ID of the particle that created the photon (for primary particles, parent_id = id")
        f.array(f"{g}/polarization", sec["pol"][m], "1", "Stokes parameters of the photon: I, Q, U, V")
        f.alias(g, "polarization", "polarisation")
    else:
        f.array(f"{g}/n_gamma", sec["interaction_count"][m], "1", f"total number of photons emitted by the {who}")
        f.array(f"{g}/id", sec["id"][m], None, f"unique ID of the {who}")
        f.array(f"{g}/parent_id", sec["parent_id"][m], None,
                f"ID of the particle that created the {who} (for primary particles, parent_id = id)")
        f.array(f"{g}/absorption", sec["work"][m] * ELECTRON_MASS_MEV * se, ue, "energy absorbed from the laser")
    f.array(f"{g}/position", sec["r"][m] * sl, ul, f"four-position of the {who}")
    if species == 1 and with_creation:
        f.array(f"{g}/position_at_creation", sec["r0"][m] * sl, ul, f"four-position at which the {who} was created")
    f.array(f"{g}/momentum", sec["u"][m] * ELECTRON_MASS_MEV * sp, up, f"four-momentum of the {who}")


def write_particles(path, run, sec, inter=None):
    """Writes the particle dump of one run to `path`. `run` holds the run's
    parameters (the dict built by deck.hdf5_run_info); `sec` and `inter` are the
    fetched final-state and intermediate particles (engine.fetch)."""
    units = run["units"]
    ul, ue = units["length"], units["energy"]
    sl, se = UNITS[ul], UNITS[ue]
    with File(path, "w") as f:
        f.group("build")
        f.string("build/version", VERSION)
        f.string("build/branch", "main")
        f.string("build/commit-hash", run.get("commit", "unknown"))
        f.string("build/features", "hdf5-output")
        f.u64("config/mpi-tasks", 1)
        f.string("config/input-file", run["input"])
        f.string("config/unit/position", ul)
        f.string("config/unit/momentum", units["momentum"])
        c = run["control"]
        f.f64("config/control/dt_multiplier", c["dt_multiplier"])
        f.flag("config/control/radiation_reaction", c["radiation_reaction"])
        f.flag("config/control/classical", c["classical"])
        f.flag("config/control/pair_creation", c["pair_creation"])
        f.flag("config/control/pair_creation_is_pol_resolved", c["pol_resolved"])
        f.flag("config/control/lcfa", c["lcfa"])
        f.u64("config/control/rng_seed", c["rng_seed"])
        f.f64("config/control/increase_pair_rate_by", c["increase_pair_rate_by"])
        f.flag("config/control/bandwidth_correction", c["bandwidth_correction"])
        if c.get("stop_at_time") is not None:
            f.f64("config/control/stop_at_time", c["stop_at_time"], "s",
                  "tracking has continued until all particles have this time coordinate")
        if c.get("select_multiplicity") is not None:
            f.u64("config/control/select_multiplicity", c["select_multiplicity"])
        else:
            f.flag("config/control/select_multiplicity", False)
        la = run["laser"]
        f.f64("config/laser/a0", la["a0"], "1", "peak value of the laser normalized amplitude")
        f.f64("config/laser/wavelength", la["wavelength"] * sl, ul, "wavelength of the carrier")
        f.enum("config/laser/polarization", ["linear", "circular"], la["polarization"], None, "linear/circular")
        if la["polarization"] == 0:
            f.f64("config/laser/polarization_angle", la["polarization_angle"], "rad",
                  "angle between x-axis and electric field")
        f.flag("config/laser/focusing", la["focusing"], None, "true/false => pulse is modelled in 3d/1d")
        f.enum("config/laser/envelope", ["cos^2", "flattop", "gaussian"], la["envelope"], None,
               "pulse envelope of the laser vector potential")
        f.f64("config/laser/chirp_b", la["chirp_b"], "1", "parameter that appears in carrier phase = phi + b phi^2")
        if la["focusing"]:
            f.f64("config/laser/waist", la["waist"] * sl, ul,
                  "radius in the focal plane at which intensity is 1/e^2 of its peak value")
        if la["envelope"] == 2:
            f.f64("config/laser/fwhm_duration", la["n_cycles"] * la["wavelength"] / SPEED_OF_LIGHT, "s",
                  "full width at half maximum of the temporal intensity profile")
        else:
            f.f64("config/laser/n_cycles", la["n_cycles"], "1",
                  "number of wavelengths corresponding to the total pulse duration")
        b = run["beam"]
        f.u64("config/beam/n", b["n"], "1", "number of primary macroparticles")
        f.string("config/beam/species", b["species"])
        f.f64("config/beam/collision_angle", b["collision_angle"], "rad",
              "polar angle between beam momentum and laser wavevector, zero if counterpropagating")
        f.f64("config/beam/collision_plane_angle", b["collision_plane_angle"], "rad",
              "azimuthal angle between beam momentum and laser wavevector, zero if p in E, k plane")
        f.vector("config/beam/offset", np.asarray(b["offset"]) * sl, ul)
        if b.get("imported_from_file"):
            f.flag("config/beam/imported_from_file", True)
            f.f64("config/beam/distance_bt_ips", b["distance_bt_ips"] * sl, ul,
                  "distance between particle beam source and focal spot")
            f.f64("config/beam/min_energy", b["min_energy"] * se, ue,
                  "primaries with energies below this threshold are discarded")
            f.f64("config/beam/max_angle", b["max_angle"], "rad",
                  "primaries propagating at angles larger than this are discarded")
        else:
            f.flag("config/beam/imported_from_file", False)
            f.f64("config/beam/n_real", b["n"] * b["weight"], "1",
                  "total number of real particles represented by the primary macroparticles")
            q = {"electron": -1.0, "positron": 1.0, "photon": 0.0}[b["species"]]
            f.f64("config/beam/charge", b["n"] * b["weight"] * q * ELEMENTARY_CHARGE * (1.0 if q else 0.0) *
                  (-1.0 if q < 0 else 1.0) if q else 0.0, "C")
            f.f64("config/beam/gamma", b["gamma"], "1")
            f.f64("config/beam/sigma", b["sigma"], "1")
            f.flag("config/beam/bremsstrahlung_source", b["bremsstrahlung_source"])
            if b["bremsstrahlung_source"]:
                f.f64("config/beam/gamma_min", b["gamma_min"], "1")
            f.f64("config/beam/radius", b["radius"] * sl, ul)
            f.f64("config/beam/radius_max", b["radius_max"] * sl, ul,
                  "density distribution is cut off at this perpendicular distance from the beam axis")
            f.f64("config/beam/length", b["length"] * sl, ul)
            f.f64("config/beam/rms_divergence", b["rms_divergence"], "rad")
            f.vector("config/beam/polarization", b["polarization"], "1",
                     "Stokes parameters of the primary particles: I, Q, U, V")
            f.alias("config/beam", "polarization", "polarisation")
            f.flag("config/beam/transverse_distribution_is_normal", b["transverse_normal"])
            f.flag("config/beam/longitudinal_distribution_is_normal", True)
        o = run["output"]
        f.flag("config/output/laser_defines_positive_z", o["laser_defines_positive_z"])
        f.flag("config/output/beam_defines_positive_z", not o["laser_defines_positive_z"])
        f.flag("config/output/discard_background_e", o["discard_background_e"])
        f.flag("config/output/discard_background_ph", o["discard_background_ph"])
        f.f64("config/output/min_energy", o["min_energy"] * se, ue)
        f.group("final-state")
        for species in (2, 0, 1):
            _species_groups(f, "final-state", sec, species, units, True)
        if inter is not None:
            f.group("intermediate-state")
            _species_groups(f, "intermediate-state", {**inter, "species": np.full(len(inter["species"]), 3, np.uint8)},
                            3, units, False)
        e, e_unit = run["laser_energy"]
        if la["focusing"]:
            f.f64("final-state/laser/energy", e, e_unit, "total energy of the laser pulse")
        else:
            f.f64("final-state/laser/energy_flux", e, e_unit, "total energy of the laser pulse, per unit area")
        f.f64("final-state/laser/absorption", run["absorption_j"], "J", "energy absorbed from the laser")


# ---------------------------------------------------------------- the beam import
def _rotate_y(v, theta):
    """Rotation of the (x, z) components about the y axis (four_vector.rs rotate_around_y)."""
    c, s = math.cos(theta), math.sin(theta)
    out = v.copy()
    out[..., 1] = c * v[..., 1] + s * v[..., 3]
    out[..., 3] = -s * v[..., 1] + c * v[..., 3]
    return out


def load_beam(path, species, distance_bt_ips, initial_z, collision_angle=0.0, offset=(0.0, 0.0, 0.0),
              auto_timing=True, min_energy=0.0, max_angle=math.pi):
    """BeamLoader::build (src/particle/loader.rs:135-283): the `species`
    particles of a particle dump as primaries, realigned to this run's
    collision: rotated from the file's beam axis onto the new one, offset, and
    (auto_timing) moved back along their worldlines by the distance between the
    interaction points. Returns a dict of numpy arrays (species, r, u, pol,
    weight, id, parent_id, ...) for Engine.load_primaries; `min_energy` in MeV."""
    sp = {"electron": 0, "positron": 1, "photon": 2}[species]
    with File(path, "r") as f:
        x_unit = f.read_string("config/unit/position")
        p_unit = f.read_string("config/unit/momentum")
        if x_unit not in UNITS or p_unit not in UNITS:
            raise H5Error(f"{path}: unit '{x_unit}' or '{p_unit}' not recognised")
        axis = None
        if f.exists("config/output/beam_defines_positive_z") and f.read("config/output/beam_defines_positive_z")[0]:
            axis = np.array([0.0, 0.0, 1.0])
        if axis is None and f.exists("config/output/laser_defines_positive_z") and \
                f.read("config/output/laser_defines_positive_z")[0] and f.exists("config/beam/collision_angle"):
            th = float(f.read("config/beam/collision_angle")[0])
            axis = np.array([-math.sin(th), 0.0, -math.cos(th)])
        if axis is None and f.exists("beam_axis"):
            s = f.read_string("beam_axis")
            axes = {"+x": [1.0, 0.0, 0.0], "-x": [-1.0, 0.0, 0.0], "+z": [0.0, 0.0, 1.0], "-z": [0.0, 0.0, -1.0]}
            if s not in axes:
                raise H5Error(f"beam axis \"{s}\" not recognised")
            axis = np.array(axes[s])
        if axis is None:
            import warnings
            warnings.warn("no coordinate system specified, assuming beam propagates towards positive z...")
            axis = np.array([0.0, 0.0, 1.0])
        g = f"final-state/{species}"
        weight = f.read(f"{g}/weight")
        pos = f.read(f"{g}/position")
        mom = f.read(f"{g}/momentum")
        pol = f.read(f"{g}/polarization") if f.exists(f"{g}/polarization") else None
    if pol is None:
        import warnings
        warnings.warn("no polarization data found, continuing with unpolarized particles...")
    n = min(len(weight), len(pos), len(mom))
    weight, pos, mom = weight[:n], pos[:n], mom[:n]
    pol = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1)) if pol is None else pol[:n]
    # rotation angle from the file's beam axis to the new one (loader.rs:109-113)
    new_axis = np.array([-math.sin(collision_angle), 0.0, -math.cos(collision_angle)])
    cross_y = axis[2] * new_axis[0] - axis[0] * new_axis[2]
    theta = math.copysign(abs(math.acos(max(-1.0, min(1.0, float(new_axis @ axis))))), cross_y)
    r = pos / UNITS[x_unit]
    p = mom / UNITS[p_unit] / ELECTRON_MASS_MEV
    p = _rotate_y(p, theta)
    r = _rotate_y(r, theta)
    c, s = math.cos(collision_angle), math.sin(collision_angle)
    off = np.array([0.0, c * offset[0] + s * offset[2], offset[1], -s * offset[0] + c * offset[2]])
    r = r + off
    if auto_timing:
        beta = p / p[:, :1]
        denom = (beta[:, 0] - beta[:, 3])[:, None]
        base = np.array([-1.0, math.sin(collision_angle), 0.0, math.cos(collision_angle)])
        shift = distance_bt_ips * (base + (1.0 + math.cos(collision_angle)) * beta / denom)
        shift = shift - (2.0 * initial_z + r[:, :1] - r[:, 3:4]) * beta / denom
        r = r + shift
    # kept: energy above min_energy, direction within max_angle of the new beam axis
    e = p[:, 0] * ELECTRON_MASS_MEV
    nrm = np.linalg.norm(p[:, 1:], axis=1)
    cosang = (p[:, 1:] @ new_axis) / np.where(nrm > 0.0, nrm, 1.0)
    keep = (e > min_energy) & (cosang > math.cos(max_angle))
    idx = np.arange(n, dtype=np.uint64)
    k = int(keep.sum())
    out = {"species": np.full(k, sp, np.uint8), "r": np.ascontiguousarray(r[keep]),
           "u": np.ascontiguousarray(p[keep]), "r0": np.ascontiguousarray(r[keep]), "u0": np.ascontiguousarray(p[keep]),
           "pol": np.ascontiguousarray(pol[keep]), "payload": np.zeros(k), "parent_chi": np.zeros(k),
           "interaction_count": np.zeros(k), "work": np.zeros(k), "weight": np.ascontiguousarray(weight[keep]),
           "id": idx[keep].copy(), "parent_id": idx[keep].copy()}
    return out
