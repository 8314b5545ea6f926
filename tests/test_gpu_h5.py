"""The two-stage workflow of the reference's examples/hdf5 (stage1.yml ->
stage2.yml) on the GPU: stage 1 dumps its final state to HDF5
(output:dump_all_particles), stage 2 imports the photons of that dump as its
primaries (beam:from_hdf5) and collides them with a second pulse."""
import math
import os
import sys

import numpy as np
import pytest

from ptarmigan_amd import h5

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not h5.available(), reason="HDF5 support not built")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


STAGE1 = """
control: {lcfa: true, radiation_reaction: true, pair_creation: false, rng_seed: 3}
laser: {a0: 10.0, wavelength: 0.8 * micro, fwhm_duration: 30.0 * femto, waist: 4.0 * micro, polarization: linear}
beam: {n: 4000, species: electron, gamma: 10000.0 / 0.510999, sigma: 10.0, radius: 1.0 * micro, length: 2.0 * micro}
output: {ident: stage1, dump_all_particles: hdf5, coordinate_system: beam}
"""

STAGE2 = """
control: {lcfa: true, radiation_reaction: true, pair_creation: true, increase_pair_rate_by: 1000.0, rng_seed: 5}
laser: {a0: 30.0, wavelength: 0.8 * micro, fwhm_duration: 30.0 * femto, waist: 4.0 * micro, polarization: linear}
beam:
  species: photon
  from_hdf5: {file: stage1_particles.h5, distance_bt_ips: 0.5, max_angle: 1.0 * milli, min_energy: 100.0 * MeV}
output: {ident: stage2, dump_all_particles: hdf5, coordinate_system: beam}
"""


def test_stage1_dump_then_stage2_import(tmp_path, engine):
    from ptarmigan_amd.deck import Deck, run_deck
    out = str(tmp_path)
    d1 = Deck.from_string(STAGE1, "stage1")
    d1.dir = out
    (a0, sec, cnt), = run_deck(d1, engine, out_dir=out)
    path = os.path.join(out, "stage1_particles.h5")
    assert cnt["particle_dump"] == path and os.path.exists(path)
    say("stage 1:", cnt["emissions"], "emissions")
    with h5.File(path) as f:
        for species, code in (("photon", 2), ("electron", 0), ("positron", 1)):
            m = sec["species"] == code
            assert f.info(f"final-state/{species}/momentum")[:2] == (m.sum(), 4)
            assert np.allclose(f.read(f"final-state/{species}/momentum"), sec["u"][m] * 0.510999, rtol=1e-15)
            assert np.array_equal(f.read(f"final-state/{species}/id", np.uint64), sec["id"][m])
        assert f.read("config/beam/n", np.uint64)[0] == 4000
        assert f.read("config/output/beam_defines_positive_z")[0] == 1
        assert f.read_string("config/input-file") == STAGE1
        assert f.read("final-state/laser/energy")[0] > 0.0
    # stage 2: the photons within 1 mrad of the beam axis and above 100 MeV
    m = sec["species"] == 2
    u = sec["u"][m]
    keep = (u[:, 0] * 0.510999 > 100.0) & (u[:, 3] / np.linalg.norm(u[:, 1:], axis=1) > math.cos(1e-3))
    d2 = Deck.from_string(STAGE2, "stage2")
    d2.dir = out
    (a0, sec2, cnt2), = run_deck(d2, engine, out_dir=out)
    say("stage 2:", int(keep.sum()), "primaries,", cnt2["pair_events"], "pair events, flags", cnt2["error_flags"])
    assert cnt2["error_flags"] == 0
    assert keep.sum() > 100
    # every imported photon is either still a photon, or created a pair; ids are file order
    ph2 = sec2["species"] == 2
    primaries = ph2 & (sec2["parent_id"] == sec2["id"])
    assert primaries.sum() <= keep.sum()
    assert np.all(np.isin(sec2["id"][primaries], np.nonzero(keep)[0]))
    assert cnt2["pair_events"] > 0
    with h5.File(os.path.join(out, "stage2_particles.h5")) as f:
        assert f.read("config/beam/imported_from_file")[0] == 1
        assert f.read("config/beam/distance_bt_ips")[0] == 0.5
        assert f.info("final-state/positron/position_at_creation")[0] == int((sec2["species"] == 1).sum())
