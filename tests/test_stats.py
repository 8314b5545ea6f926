"""Summary statistics (the `stats:` block).

CPU: the reference's own parser tests (src/output/stats.rs:579-723, `parsing`
and `parse_expr`) restated, and the stats.txt line format. GPU: every reduction
evaluated by ptg_stats over a shower against numpy over the fetched particles
(the definitions of stats.rs:265-481: total, fraction, mean, the shifted
variance, min and max of v*w, and the circular mean, variance and standard
deviation from the weighted resultant).
"""
import math

import numpy as np
import pytest

from ptarmigan_amd import output as O
from ptarmigan_amd.deck import DeckError

CTX = {"a": 2.0}


def test_reference_parsing_cases():
    s = O.Statistic("mean energy in (1.0 * MeV; auto)", CTX)
    assert s.op == "mean" and s.min == 1.0 and s.max == math.inf
    s = O.Statistic("variance angle_x`energy", CTX)
    assert s.op == "variance" and s.filter is None and s.weight == "energy"
    assert (s.min, s.max) == (-math.inf, math.inf)
    with pytest.raises(DeckError):
        O.Statistic("frac pelican in (0.0; 1.0)", CTX)
    s = O.Statistic("frac number for p_perp in (auto; 100.0 * MeV/c)", CTX)
    assert s.op == "fraction" and s.filter == "p_perp" and s.min == -math.inf and s.max == 100.0
    s = O.Statistic("total r_perp for p^- in (a / (1.0 + a) * MeV/c; 1000.0 * MeV/c)", CTX)
    assert s.op == "total" and s.filter == "p^-"
    assert abs(s.min - 2.0 / 3.0) < 1e-10 * 2.0 / 3.0 and abs(s.max - 1000.0) < 1e-10 * 1000.0
    s = O.Statistic("mean p_x for x in (-1.0 * micro; 1.0 * micro)", CTX)
    assert s.op == "mean" and s.var_unit == "MeV/c" and s.filter_unit == "m"
    assert (s.min, s.max) == (-1.0e-6, 1.0e-6)
    for spec, op in (("circmean angle_x`energy", "circmean"), ("circvar angle_x`energy", "circvar"),
                     ("cstd angle_x`energy", "circstd")):
        s = O.Statistic(spec, CTX)
        assert s.op == op and s.filter is None and s.weight == "energy" and (s.min, s.max) == (-math.inf, math.inf)
    with pytest.raises(DeckError):
        O.Statistic("circvar energy", CTX)


def test_reference_expression_cases():
    ctx = {"a": 1.0, "b": 5.0}
    for spec, formula, unit in (("test a*b", None, "1"), ("test`formula a*b", "a*b", "1"), ("test a*b mm", None, "mm"),
                                ("test`formula a*b mm", "a*b", "mm")):
        e = O.StatsExpression(spec, ctx)
        assert (e.name, e.formula, e.value, e.unit) == ("test", formula, 5.0, unit)
    assert str(O.StatsExpression("test`formula a*b mm", ctx)) == "expr: test (a*b) = 5.000000e0 [mm]"
    with pytest.raises(DeckError):
        O.StatsExpression("lonely", ctx)


def test_line_format():
    s = O.Statistic("frac number for p_perp in (auto; 100.0 * MeV/c)", CTX)
    s.name, s.value = "photon", 0.5
    assert str(s) == "photon: fraction number (p_perp [MeV/c] < 1.000000e2) = 5.000000e-1 [1]"
    s = O.Statistic("variance energy`weight in (1.0 * MeV; 2.0 * MeV)", CTX)
    s.name, s.value = "electron", 2.0
    assert str(s) == ("electron: variance energy (weight-weighted) (1.000000e0 < energy [(MeV)^2] < 2.000000e0)"
                      " = 2.000000e0 [(MeV)^2]")
    for bad in ("mean energy in 1.0", "mean energy for x (0; 1)", "mean energy in (0; 1; 2)", "average energy",
                "mean energy`spin", "mean energy`a`b"):
        with pytest.raises(DeckError):
            O.Statistic(bad, CTX)


@pytest.mark.gpu
def test_statistics_match_numpy(engine):
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(20_000)
    engine.set_laser(laser)
    engine.build_beam(beam, seed=8)
    cnt = engine.run(opts, seed=8)
    sec, _ = engine.fetch(cnt)
    ph = sec["species"] == 2
    u, w = sec["u"][ph], sec["weight"][ph]
    me = 0.510999
    energy = me * u[:, 0]
    ax = np.arctan2(u[:, 1], u[:, 3])
    pperp = me * np.hypot(u[:, 1], u[:, 2])
    print("stats over photons", int(ph.sum()), flush=True)
    cut = (pperp > -math.inf) & (pperp < 1.0)
    sel = (energy > 1.0) & (energy < math.inf)
    checks = [
        ("total energy`weight", np.sum(energy * w)),
        ("frac number for p_perp in (auto; 1.0 * MeV/c)", np.sum(w[cut]) / np.sum(w)),
        ("mean energy in (1.0 * MeV; auto)", energy[sel].mean()),
        ("variance energy", energy.var()),
        ("min energy", energy.min()),
        ("max energy`weight", (energy * w).max()),
        ("circmean angle_x`energy", math.atan2(np.sum(energy * np.sin(ax)), np.sum(energy * np.cos(ax)))),
        ("circvar angle_x", 1.0 - np.hypot(np.cos(ax).mean(), np.sin(ax).mean())),
        ("cstd angle_x", math.sqrt(-2.0 * math.log(np.hypot(np.cos(ax).mean(), np.sin(ax).mean())))),
    ]
    for spec, want in checks:
        st = O.Statistic(spec, {})
        got = st.evaluate(engine, 2, "photon")
        print(" ", spec, got, want, flush=True)
        # 1 - R for a narrow beam is ~1e-8, so summation order alone moves the
        # circular variance and deviation at 1e-8 relative
        rel = 1e-6 if spec.startswith(("circvar", "cstd")) else 1e-9
        assert got == pytest.approx(want, rel=rel, abs=1e-15), spec
    st = O.Statistic("mean energy for x in (1.0; 2.0)", {})  # nothing within the bounds
    assert math.isnan(st.evaluate(engine, 2, "photon"))
    assert O.Statistic("min energy for x in (1.0; 2.0)", {}).evaluate(engine, 2, "photon") == math.inf


@pytest.mark.gpu
def test_deck_writes_its_output(engine, tmp_path):
    """A deck with output and stats sections, run end to end: one set of files per a0."""
    from ptarmigan_amd.deck import Deck, run_deck
    deck = Deck.from_string("""
control:
  rng_seed: 3
laser:
  a0: {start: 1.0, stop: 1.5, step: 0.5}
  wavelength: 0.8 * micro
  polarization: circular
  n_cycles: 8.0
beam:
  n: 3000
  gamma: 10.0 * GeV / (me * c^2)
output:
  ident: auto
  file_format: fits
  photon: [energy, "angle_x:angle_y:(20;auto)"]
stats:
  photon: [total number, mean energy, "frac number for energy in (auto; 100.0 * MeV)"]
  electron: [mean gamma]
  expression: [gamma0`formula 10.0*GeV/(me*c^2)]
""", name="scan")
    res = run_deck(deck, engine, out_dir=str(tmp_path))
    print("deck output", [r[0] for r in res], flush=True)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == sorted(f"scan_a0_{a:.3f}_{f}" for a in (1.0, 1.5) for f in
                           ("photon_energy.fits", "photon_angle_x-angle_y.fits", "stats.txt"))
    for a0, part, cnt in res:
        lines = (tmp_path / f"scan_a0_{a0:.3f}_stats.txt").read_text().splitlines()
        assert lines[0].startswith("electron: mean gamma = ") and lines[-1].startswith("expr: gamma0 (10.0*GeV/(me*c^2))")
        total = float(lines[1].split("= ")[1].split(" ")[0])
        assert total == pytest.approx(part["weight"][part["species"] == 2].sum(), rel=1e-6)
        raw = (tmp_path / f"scan_a0_{a0:.3f}_photon_energy.fits").read_bytes()
        assert len(raw) % 2880 == 0 and raw.startswith(b"SIMPLE  =                    T")
