"""A batch of collisions in one particle population (ptg_set_batch) and the
launch-shape autotuner (ptg_autotune).

Every particle draws from its own counter-based stream and reads its own
collision's laser, so the first collision of a batch (built as ptg_build_beam
builds it) yields exactly the particles of that collision run alone, whatever
else shares the launches; the autotuner's candidates give the same showers.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scan(n_coll, counts, a0s):
    from ptarmigan_amd import workloads
    lasers, beams, opts = workloads.batched_scan(n_coll)
    for k in range(n_coll):
        lasers[k].a0 = a0s[k]
        beams[k].n = counts[k]
    return lasers, beams, opts


@pytest.mark.parametrize("pairs", [False, True])
def test_first_collision_of_a_batch_is_the_single_collision(engine, pairs):
    lasers, beams, opts = _scan(3, [3000, 2000, 1000], [4.0, 1.0, 8.0])
    opts.tracking_photons = pairs
    say("batch vs single, pairs", pairs)
    engine.set_batch(lasers, beams, seed=9)
    cb = engine.run(opts, seed=9)
    sb, _ = engine.fetch(cb)
    engine.set_laser(lasers[0])
    engine.build_beam(beams[0], seed=9)
    cs = engine.run(opts, seed=9)
    ss, _ = engine.fetch(cs)
    say("  batch", cb, "\n  single", cs)
    assert cb["error_flags"] == 0 and cs["error_flags"] == 0
    n0 = beams[0].n
    for sp in (0, 1, 2):
        a = np.sort(ss["u"][ss["species"] == sp, 0])
        # particles of collision 0: their primary (shower) index is below n0
        b = np.sort(sb["u"][(sb["species"] == sp) & (sb["shower"] < n0), 0])
        assert len(a) == len(b), sp
        assert np.array_equal(a, b), sp
    # the other collisions ran at their own a0: more photons per electron at the larger a0
    nph = [np.sum((sb["species"] == 2) & (sb["shower"] >= lo) & (sb["shower"] < hi))
           for lo, hi in ((0, 3000), (3000, 5000), (5000, 6000))]
    say("  photons per collision", nph)
    assert nph[2] / 1000 > nph[0] / 3000 > nph[1] / 2000


def test_pair_events_read_their_own_collisions_laser(engine):
    """Breit-Wheeler events of a later collision in a batch: the photon's field
    comes from its own collision's pulse (a0 = 20 here, behind an a0 = 0.5 first
    collision), so the positrons per primary match that collision run alone (an
    independent seed) within the Poisson error, and no event falls outside the
    tables, as one would at the first collision's a0."""
    lasers, beams, opts = _scan(2, [4000, 4000], [0.5, 20.0])
    beams[1].gamma = beams[0].gamma = 17500.0 / 0.510999
    engine.set_batch(lasers, beams, seed=21)
    cb = engine.run(opts, seed=21)
    sb, _ = engine.fetch(cb)
    engine.set_laser(lasers[1])
    engine.build_beam(beams[1], seed=22)
    cs = engine.run(opts, seed=22)
    ss, _ = engine.fetch(cs)
    nb = int(np.sum((sb["species"] == 1) & (sb["shower"] >= 4000)))
    n0 = int(np.sum((sb["species"] == 1) & (sb["shower"] < 4000)))
    ns = int(np.sum(ss["species"] == 1))
    say("positrons: batch collision 1", nb, "collision 0", n0, "single", ns, "flags", cb["error_flags"], cs["error_flags"])
    assert cb["error_flags"] == 0 and cs["error_flags"] == 0
    assert ns > 100 and n0 < ns / 20
    assert abs(nb - ns) < 5.0 * np.sqrt(nb + ns)


def test_autotuner_picks_a_shape_and_keeps_results(engine):
    lasers, beams, opts = _scan(4, [20000, 5000, 10000, 2000], [1.0, 2.0, 4.0, 8.0])
    engine.set_batch(lasers, beams, seed=3)
    r = engine.autotune(opts, seed=3, max_primaries=20000)
    say("autotune", r)
    assert r["consistent"]
    assert len(r["candidates"]) == 9
    assert (r["block"], r["grid_per_cu"]) in [(b, g) for b, g, _ in r["candidates"]]
    best = min(ms for _, _, ms in r["candidates"])
    assert dict(((b, g), ms) for b, g, ms in r["candidates"])[(r["block"], r["grid_per_cu"])] == best
    c1 = engine.run(opts, seed=3)
    engine.set_batch(lasers, beams, seed=3)  # a new batch keeps the tuned shape
    c2 = engine.run(opts, seed=3)
    assert c1["emissions"] == c2["emissions"] and c1["lepton_steps"] == c2["lepton_steps"]
