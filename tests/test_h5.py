"""HDF5 particle dump and beam import (ptarmigan_amd/h5.py over libptg_h5.so).

Layout and types against the reference's writer (src/main.rs:1289-1715,
hdf5-writer/src/data.rs, dataset.rs, datatype.rs, docs/output.md): the dataset
names, their element types (4 x f64 arrays for four-vectors and Stokes vectors,
u64 ids, u8 enumerations for flags), the unit and desc attributes, the aliases,
and the unit conversion; the importer (src/particle/loader.rs:135-283) by
round trip and by what its realignment must do: rotate the file's beam axis
onto the new one, and with auto timing put every particle at
ct - z = -2 initial_z (the light-front time at which the reference starts a
generated beam)."""
import math
import os

import numpy as np
import pytest

from ptarmigan_amd import h5

pytestmark = pytest.mark.skipif(not h5.available(), reason="HDF5 support not built")

H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_ENUM = 0, 1, 3, 8


def test_scalars_arrays_strings_round_trip(tmp_path):
    path = str(tmp_path / "kinds.h5")
    with h5.File(path, "w") as f:
        f.f64("a/x", 1.25, "m", "a length")
        f.u64("a/n", 2 ** 40 + 3)
        f.flag("a/yes", True)
        f.flag("a/no", False)
        f.enum("a/pol", ["linear", "circular"], 1)
        f.string("a/s", "cos^2 and µ")
        f.array("b/v", np.arange(5, dtype=np.float64) * 0.5, "1")
        f.array("b/id", np.array([1, 2 ** 63, 7], dtype=np.uint64))
        f.array("b/r", np.arange(12, dtype=np.float64).reshape(3, 4), "mm", "four-position")
        f.array("b/empty", np.zeros((0, 4)))
        f.alias("b", "r", "r_alias")
    with h5.File(path) as f:
        assert f.read("a/x")[0] == 1.25 and f.info("a/x") == (1, 1, H5T_FLOAT)
        assert f.attribute("a/x", "unit") == "m" and f.attribute("a/x", "desc") == "a length"
        assert f.read("a/n", np.uint64)[0] == 2 ** 40 + 3 and f.info("a/n")[2] == H5T_INTEGER
        assert f.read("a/yes")[0] == 1 and f.read("a/no")[0] == 0 and f.info("a/yes")[2] == H5T_ENUM
        assert f.read("a/pol")[0] == 1
        assert f.read_string("a/s") == "cos^2 and µ" and f.info("a/s")[2] == H5T_STRING
        assert np.array_equal(f.read("b/v"), np.arange(5) * 0.5)
        assert np.array_equal(f.read("b/id", np.uint64), np.array([1, 2 ** 63, 7], dtype=np.uint64))
        assert f.info("b/r") == (3, 4, H5T_FLOAT)
        assert np.array_equal(f.read("b/r"), np.arange(12.0).reshape(3, 4))
        assert np.array_equal(f.read("b/r_alias"), f.read("b/r"))
        assert f.info("b/empty") == (0, 4, H5T_FLOAT)
        assert f.exists("b/r") and not f.exists("b/nothing") and not f.exists("c/d")
        assert f.attribute("b/id", "unit") == ""


def _particles(rng, n_ph=50, n_el=30, n_po=20):
    n = n_ph + n_el + n_po
    sp = np.concatenate([np.full(n_ph, 2), np.full(n_el, 0), np.full(n_po, 1)]).astype(np.uint8)
    u = np.zeros((n, 4))
    u[:, 1:] = rng.normal(size=(n, 3)) * [0.01, 0.01, 1.0] + [0.0, 0.0, 2000.0]
    u[:, 0] = np.where(sp == 2, np.linalg.norm(u[:, 1:], axis=1), np.sqrt(1.0 + (u[:, 1:] ** 2).sum(1)))
    r = rng.normal(size=(n, 4)) * 1e-6
    return {"species": sp, "r": r, "u": u, "r0": r - 1e-7, "u0": u, "pol": rng.uniform(-1, 1, (n, 4)),
            "payload": rng.uniform(size=n), "parent_chi": rng.uniform(size=n), "interaction_count": rng.uniform(size=n),
            "work": rng.uniform(size=n), "weight": rng.uniform(1, 2, n), "id": np.arange(n, dtype=np.uint64) + 10,
            "parent_id": np.arange(n, dtype=np.uint64)}


def _run_info(units=("m", "MeV", "MeV/c"), beam_defines_z=True, focusing=True):
    return {"input": "control:\n  lcfa: true\n", "units": dict(zip(("length", "energy", "momentum"), units)),
            "control": {"dt_multiplier": 1.0, "radiation_reaction": True, "classical": False, "pair_creation": True,
                        "pol_resolved": False, "lcfa": True, "rng_seed": 7, "increase_pair_rate_by": 1.0,
                        "bandwidth_correction": False, "stop_at_time": None, "select_multiplicity": None},
            "laser": {"a0": 10.0, "wavelength": 0.8e-6, "polarization": 0, "polarization_angle": 0.0,
                      "focusing": focusing, "envelope": 2, "chirp_b": 0.0, "waist": 4e-6, "n_cycles": 11.24},
            "beam": {"n": 1000, "species": "electron", "collision_angle": 0.0, "collision_plane_angle": 0.0,
                     "offset": (0.0, 0.0, 0.0), "imported_from_file": False, "weight": 2.0, "gamma": 3.2e4,
                     "sigma": 0.0, "bremsstrahlung_source": False, "gamma_min": 0.0, "radius": 5e-6,
                     "radius_max": math.inf, "length": 2e-5, "rms_divergence": 0.0,
                     "polarization": (1.0, 0.0, 0.0, 0.0), "transverse_normal": True},
            "output": {"laser_defines_positive_z": not beam_defines_z, "discard_background_e": False,
                       "discard_background_ph": False, "min_energy": 0.0},
            "laser_energy": (1.2, "J"), "absorption_j": 1e-3}


REFERENCE_DATASETS = {
    "photon": ["weight", "a0_at_creation", "xi", "parent_chi", "n_pos", "id", "parent_id", "polarization",
               "polarisation", "position", "momentum"],
    "electron": ["weight", "n_gamma", "id", "parent_id", "absorption", "position", "momentum"],
    "positron": ["weight", "a0_at_creation", "xi", "parent_chi", "n_gamma", "id", "parent_id", "absorption",
                 "position", "position_at_creation", "momentum"],
}


def test_particle_dump_layout_and_units(tmp_path):
    """The reference's group and dataset names, types and unit attributes; HEP
    units: positions in mm, momenta in GeV/c, absorbed energy in GeV."""
    rng = np.random.default_rng(1)
    sec = _particles(rng)
    path = str(tmp_path / "particles.h5")
    h5.write_particles(path, _run_info(("mm", "GeV", "GeV/c")), sec)
    with h5.File(path) as f:
        for name in ("build/version", "build/branch", "build/commit-hash", "build/features", "config/input-file",
                     "config/unit/position", "config/unit/momentum", "config/mpi-tasks", "config/control/lcfa",
                     "config/control/rng_seed", "config/control/select_multiplicity", "config/laser/a0",
                     "config/laser/polarization", "config/laser/envelope", "config/laser/waist",
                     "config/laser/fwhm_duration", "config/beam/n", "config/beam/charge", "config/beam/polarisation",
                     "config/output/beam_defines_positive_z", "final-state/laser/energy",
                     "final-state/laser/absorption"):
            assert f.exists(name), name
        assert not f.exists("config/laser/n_cycles") and not f.exists("final-state/laser/energy_flux")
        assert f.read_string("config/unit/position") == "mm" and f.read_string("config/unit/momentum") == "GeV/c"
        assert f.info("config/laser/polarization")[2] == H5T_ENUM and f.info("config/control/lcfa")[2] == H5T_ENUM
        assert f.read("config/beam/charge")[0] == pytest.approx(-1000 * 2.0 * 1.602177e-19)
        assert f.read_string("config/input-file").startswith("control:")
        for species, names in REFERENCE_DATASETS.items():
            m = sec["species"] == {"photon": 2, "electron": 0, "positron": 1}[species]
            for name in names:
                assert f.exists(f"final-state/{species}/{name}"), (species, name)
            g = f"final-state/{species}"
            assert f.info(f"{g}/position") == (m.sum(), 4, H5T_FLOAT)
            assert f.info(f"{g}/id")[2] == H5T_INTEGER
            assert f.attribute(f"{g}/position", "unit") == "mm" and f.attribute(f"{g}/momentum", "unit") == "GeV/c"
            assert np.allclose(f.read(f"{g}/position"), sec["r"][m] * 1e3, rtol=1e-15)
            assert np.allclose(f.read(f"{g}/momentum"), sec["u"][m] * 0.510999e-3, rtol=1e-15)
            assert np.array_equal(f.read(f"{g}/id", np.uint64), sec["id"][m])
            assert np.array_equal(f.read(f"{g}/weight"), sec["weight"][m])
            if species != "photon":
                assert f.attribute(f"{g}/absorption", "unit") == "GeV"
                assert np.allclose(f.read(f"{g}/absorption"), sec["work"][m] * 0.510999e-3, rtol=1e-15)
        assert np.array_equal(f.read("final-state/photon/polarisation"), sec["pol"][sec["species"] == 2])


def _dump(tmp_path, sec, beam_defines_z=True, units=("m", "MeV", "MeV/c")):
    path = str(tmp_path / "stage1_particles.h5")
    h5.write_particles(path, _run_info(units, beam_defines_z), sec)
    return path


@pytest.mark.parametrize("units", [("m", "MeV", "MeV/c"), ("mm", "GeV", "GeV/c"), ("m", "J", "kg/m/s")])
def test_loader_round_trip(tmp_path, units):
    """Without realignment (file beam axis = new beam axis, no timing shift) the
    imported photons are the dumped ones, in file order, ids 0, 1, ..."""
    rng = np.random.default_rng(2)
    sec = _particles(rng)
    sec["u"][:, 3] *= -1.0  # propagating along -z, as the new beam must
    path = _dump(tmp_path, sec, beam_defines_z=False, units=units)
    prim = h5.load_beam(path, "photon", 0.0, 0.0, auto_timing=False)
    m = sec["species"] == 2
    assert len(prim["species"]) == m.sum() and np.all(prim["species"] == 2)
    assert np.allclose(prim["r"], sec["r"][m], rtol=1e-14, atol=1e-22)
    assert np.allclose(prim["u"], sec["u"][m], rtol=1e-14)
    assert np.array_equal(prim["pol"], sec["pol"][m])
    assert np.array_equal(prim["weight"], sec["weight"][m])
    assert np.array_equal(prim["id"], np.arange(m.sum(), dtype=np.uint64))


def test_loader_realigns_and_times_the_beam(tmp_path):
    """A dump written in the beam frame (beam along +z) imported head-on
    (collision angle 0, beam along -z): momenta rotated by pi about y, and with
    auto timing every particle at ct - z = -2 initial_z, its transverse
    momentum untouched in size."""
    rng = np.random.default_rng(3)
    sec = _particles(rng)
    path = _dump(tmp_path, sec, beam_defines_z=True)
    z0 = 7.5e-6
    with pytest.warns(UserWarning, match="no polarization data"):  # electrons carry none, as in the reference
        prim = h5.load_beam(path, "electron", 0.5, z0)
    m = sec["species"] == 0
    assert np.allclose(prim["u"][:, 3], -sec["u"][m, 3], rtol=1e-13)
    assert np.allclose(prim["u"][:, 1], -sec["u"][m, 1], rtol=1e-13)
    assert np.allclose(prim["r"][:, 0] - prim["r"][:, 3], -2.0 * z0, rtol=1e-9, atol=1e-15)


def test_loader_cuts_energy_and_angle(tmp_path):
    """min_energy (MeV) and max_angle about the new beam axis discard primaries,
    which keep their file-order ids."""
    rng = np.random.default_rng(4)
    sec = _particles(rng)
    sec["u"][:, 3] *= -1.0
    m = sec["species"] == 2
    e = sec["u"][m, 0] * 0.510999
    cut_e = float(np.median(e))
    path = _dump(tmp_path, sec, beam_defines_z=False)
    prim = h5.load_beam(path, "photon", 0.0, 0.0, auto_timing=False, min_energy=cut_e)
    keep = np.where(e > cut_e)[0]
    assert np.array_equal(prim["id"], keep.astype(np.uint64))
    ang = np.arccos(-sec["u"][m, 3] / np.linalg.norm(sec["u"][m, 1:], axis=1))
    cut_a = float(np.median(ang))
    prim = h5.load_beam(path, "photon", 0.0, 0.0, auto_timing=False, max_angle=cut_a)
    assert np.array_equal(prim["id"], np.where(ang < cut_a)[0].astype(np.uint64))


def test_deck_reads_from_hdf5_keys(tmp_path):
    from ptarmigan_amd.deck import Deck, DeckError, LoadedBeam, dump_mode
    d = Deck.from_string("""
control: {lcfa: true}
laser: {a0: 30.0, wavelength: 0.8 * micro, fwhm_duration: 30.0 * femto, waist: 4.0 * micro}
beam:
  species: photon
  from_hdf5: {file: stage1_particles.h5, distance_bt_ips: 0.5, max_angle: 0.1 * milli, min_energy: 1.0e-13}
output: {dump_all_particles: hdf5}
""")
    b = d.beam()
    assert isinstance(b, LoadedBeam)
    assert b.distance_bt_ips == 0.5 and b.max_angle == pytest.approx(1e-4) and b.auto_timing
    assert b.min_energy == pytest.approx(1e-13 * 1e-6 / 1.602177e-19)
    assert dump_mode(d) == "hdf5"
    bad = Deck.from_string("output: {dump_all_particles: csv}\n")
    with pytest.raises(DeckError, match="dump_all_particles"):
        dump_mode(bad)
    auto = Deck.from_string("""
control: {lcfa: true, increase_pair_rate_by: auto}
laser: {a0: 30.0, wavelength: 0.8 * micro, fwhm_duration: 30.0 * femto, waist: 4.0 * micro, polarization: linear}
beam: {species: photon, from_hdf5: {file: x.h5, distance_bt_ips: 0.5}}
""")
    with pytest.raises(DeckError, match="loading particles from file"):
        auto.collisions()
