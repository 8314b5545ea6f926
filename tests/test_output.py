"""Distribution output: specifiers, bins, units, cuts and the file layouts.

CPU: the specifier grammar and bin choices of src/output/mod.rs:150-230 and
src/output/hgram.rs:74-89 (including the reference's own read_ospec case),
unit conversions fixed by src/output/units.rs, and the plain-text and FITS
layouts against hand-written known answers. GPU: files written from a shower
against a numpy restatement of Histogram::generate_1d/2d
(src/output/hgram.rs:169-446) over the fetched particles.
"""
import math
import os

import numpy as np
import pytest

from ptarmigan_amd import output as O
from ptarmigan_amd.deck import Deck, DeckError


def test_reference_read_ospec():
    d = O.Distribution("angle_x:angle_y:(100;auto)")
    assert d.names == ["angle_x", "angle_y"] and d.bspec == ("fixed_number", 100) and d.weight == "auto"


@pytest.mark.parametrize("spec,names,bspec,weight,suffix", [
    ("energy", ["energy"], ("auto", None), "weight", "_energy"),
    ("p^-:p_perp", ["p_minus", "p_perp"], ("auto", None), "weight", "_p_minus-p_perp"),
    ("x:px:(energy)", ["x", "px"], ("auto", None), "energy", "_x-px_energy-weight"),
    ("energy:(log;auto)", ["energy"], ("log", None), "auto", "_energy_log"),
    ("rho:(0.5;pol_x)", ["radius"], ("fixed_size", 0.5), "pol_x", "_radius_pol_x-weight"),
    ("energy:(auto;auto;angle in 0, 1.0 * milli)", ["energy"], ("auto", None), "auto", "_energy_angle-cut"),
    ("gamma:(nonsense;helicity)", ["gamma"], ("auto", None), "helicity", "_gamma_helicity-weight"),
])
def test_specifiers(spec, names, bspec, weight, suffix):
    d = O.Distribution(spec)
    assert (d.names, d.bspec, d.weight, d.suffix()) == (names, bspec, weight, suffix)


@pytest.mark.parametrize("spec", ["energy:(weird)", "nonsense", "x:y:z", "energy:(auto;auto;angle 0)",
                                  "energy:(auto;auto;angle in unknown_name, 1.0)", "energy:(auto;auto;spin in 0, 1)"])
def test_bad_specifiers(spec):
    with pytest.raises(DeckError):
        O.Distribution(spec)


def test_bin_specs_and_cuts():
    assert O.bin_spec("100") == ("fixed_number", 100)
    assert O.bin_spec("2.5e-3") == ("fixed_size", 2.5e-3)
    assert O.bin_spec("log") == ("log", None) and O.bin_spec("auto") == ("auto", None)
    assert O.bin_spec("garbage") == ("auto", None)
    # cut bounds are SI, held in MeV, MeV/c and m; `auto` leaves the side open; the
    # words after `in` are joined without spaces before evaluation
    c = O.Cut("energy in 1.0 * MeV, auto", {})
    assert c.min == pytest.approx(1.0, rel=1e-15) and c.max == math.inf and c.quantity == 14
    c = O.Cut("px in auto, 2.0 * MeV / c", {})
    assert c.min == -math.inf and c.max == pytest.approx(2.0, rel=1e-15)
    c = O.Cut("x in -1.0 * micro, k * micro", {"k": 3.0})
    assert (c.min, c.max) == (pytest.approx(-1e-6), pytest.approx(3e-6))


def test_unit_systems():
    assert O.UNIT_SYSTEMS["hep"]["length"] == (1.0e3, "mm")
    assert O.UNIT_SYSTEMS["si"]["energy"][0] == 1.0e6 * 1.602177e-19
    assert O.UNIT_SYSTEMS["si"]["momentum"] == (1.0e6 * 1.602177e-19 / 2.997925e8, "kg/m/s")
    assert O.unit_system(Deck.from_string("output:\n  units: HEP\n"))["energy"] == (1.0e-3, "GeV")
    with pytest.raises(DeckError):
        O.unit_system(Deck.from_string("output:\n  units: cgs\n"))
    assert O.file_format(Deck.from_string("output:\n  file_format: ascii\n")) == "plain_text"
    assert O.file_format(Deck.from_string("output:\n  file_format: fits\n")) == "fits"


def test_rust_exponent_format():
    assert O.rust_exp(1234.5, 9) == "1.234500000e3"
    assert O.rust_exp(-1.25e-4, 9, True) == "-1.250000000E-4"
    assert O.rust_exp(0.0, 9) == "0.000000000e0" and O.rust_exp(-0.0, 9) == "-0.000000000e0"
    assert O.rust_exp(1e-300, 3) == "1.000e-300"
    assert (O.rust_exp(math.inf, 9), O.rust_exp(-math.inf, 9), O.rust_exp(math.nan, 9)) == ("inf", "-inf", "NaN")


def test_plain_text_layout():
    h = O.Histogram(1, 3.5, 2.0, [0.0], [1.0], [1.0, 2.5], [2], [0.5], "hgram/density/energy", "1/MeV",
                    ["energy"], ["MeV"])
    assert h.plain_text() == ("energy\thgram/density/energy\nMeV\t1/MeV\n"
                              "2.500000000e-1\t1.000000000e0\n7.500000000e-1\t2.500000000e0\n")
    # 2-D: x runs fastest
    h = O.Histogram(2, 4.0, 4.0, [0.0, 10.0], [2.0, 14.0], [1, 2, 3, 4], [2, 2], [1.0, 2.0],
                    "hgram/density/x_y", "1/(m.m)", ["x", "y"], ["m", "m"])
    rows = [r.split("\t") for r in h.plain_text().splitlines()[2:]]
    assert [[float(v) for v in r] for r in rows] == [[0.5, 11, 1], [1.5, 11, 2], [0.5, 13, 3], [1.5, 13, 4]]


def test_fits_layout():
    cts = np.arange(6, dtype=float) - 2.0
    h = O.Histogram(2, 10.0, 7.0, [0.0, -1.0], [3.0, 1.0], cts, [3, 2], [1.0, 1.0], "hgram/density/x_y", "1/(m.m)",
                    ["x", "y"], ["m", "m"])
    b = h.fits()
    assert len(b) == 2 * 2880
    cards = [b[i:i + 80].decode() for i in range(0, 2880, 80)]
    assert cards[0] == "SIMPLE  =                    T / file conforms to FITS standard                 "
    assert cards[1].startswith("BITPIX  =                  -64 / ")
    assert cards[3].startswith("NAXIS1  =                    3 / ") and cards[4].startswith("NAXIS2  =                    2")
    assert "CRVAL2  =       -1.000000000E0" + " " * 50 in cards
    assert "DATAMIN =       -2.000000000E0" + " " * 50 in cards and "DATAMAX =        3.000000000E0" + " " * 50 in cards
    assert "CNAME1  = 'x'" + " " * 67 in cards
    assert any(c.startswith("END ") for c in cards) and all(len(c) == 80 for c in cards)
    assert np.array_equal(np.frombuffer(b[2880:2880 + 48], dtype=">f8"), cts)
    assert b[2880 + 48:] == b"\0" * (2880 - 48)


# ---------------------------------------------------------------- GPU: files against numpy

M_E = 0.510999
_E = 1.602177e-19
_C = 2.997925e8


def _accessor(sec, name):
    u = sec["u"]
    p0, p1, p2, p3 = (M_E * u[:, k] for k in range(4))
    return {"energy": p0, "angle_x": np.arctan2(p1, p3), "angle_y": np.arctan2(p2, p3),
            "angle": np.arctan2(np.hypot(p1, p2), p3), "px": p1, "S_1": sec["pol"][:, 1]}[name]


def _gen1d(v, w, accept, bspec, n, log):
    """Histogram::generate_1d with HeightSpec::Density (hgram.rs:209-323)."""
    if log:
        v = np.log(v)
    ok = np.isfinite(v) & accept
    lo, hi = (v[ok].min(), v[ok].max()) if ok.any() else (np.finfo(float).max, -np.finfo(float).max)
    # with nothing in range the reference's max - min is f64::MIN - f64::MAX, which
    # IEEE arithmetic takes to -inf without a fault; nothing is binned then
    with np.errstate(over="ignore"):
        span = hi - lo
    if lo == hi:
        nb, vol = 1, 1.0
    elif bspec[0] in ("auto", "log"):
        nb = int(math.ceil(2.0 * np.cbrt(float(n))))
        vol = span / nb
    elif bspec[0] == "fixed_number":
        nb = bspec[1]
        vol = span / nb
    else:
        nb, vol = (int(math.ceil(span / bspec[1])) if np.isfinite(span) else 0), bspec[1]
    b = np.floor((v[ok] - lo) / vol).astype(np.int64)
    ww = w[ok]
    if log:
        ww = ww * vol / (np.exp(lo + b * vol) * np.expm1(vol))
    m = b < nb
    cts = np.bincount(b[m], weights=ww[m], minlength=nb) / vol
    return lo, vol, cts


def _read_dat(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines[:2], np.array([[float(x) for x in ln.split("\t")] for ln in lines[2:]])


@pytest.mark.gpu
def test_distribution_files_match_numpy(engine, tmp_path):
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(20_000)
    opts.pol_resolved = True
    engine.set_laser(laser)
    engine.build_beam(beam, seed=5)
    cnt = engine.run(opts, seed=5)
    sec, _ = engine.fetch(cnt)
    deck = Deck.from_string("""
output:
  units: hep
  file_format: plain_text
  photon: [energy, "energy:(log;auto)", "px:(0.0005;energy)", "energy:(50;auto;angle in 0, 2.0 * milli)",
           "angle_x:angle_y:(24;auto)", "S_1:(20;auto)"]
""")
    print("distribution files", cnt["n_secondaries"], flush=True)
    files = O.write_distributions(engine, deck, sec, prefix_dir=str(tmp_path), ident="run")
    names = sorted(os.path.basename(f) for f in files)
    assert names == sorted(["run_photon_energy.dat", "run_photon_energy_log.dat", "run_photon_px_energy-weight.dat",
                            "run_photon_energy_angle-cut.dat", "run_photon_angle_x-angle_y.dat",
                            "run_photon_S_1.dat"])
    ph = sec["species"] == 2
    part = {k: v[ph] for k, v in sec.items()}
    n, w = int(ph.sum()), part["weight"]
    all_ = np.ones(n, bool)
    gev = 1.0e-3
    cases = [("run_photon_energy.dat", _accessor(part, "energy") * gev, w, all_, ("auto", None), False),
             ("run_photon_energy_log.dat", _accessor(part, "energy") * gev, w, all_, ("log", None), True),
             ("run_photon_px_energy-weight.dat", _accessor(part, "px") * gev, w * _accessor(part, "energy"), all_,
              ("fixed_size", 0.0005), False),
             ("run_photon_energy_angle-cut.dat", _accessor(part, "energy") * gev, w,
              (_accessor(part, "angle") >= 0.0) & (_accessor(part, "angle") < 2.0e-3), ("fixed_number", 50), False),
             ("run_photon_S_1.dat", _accessor(part, "S_1"), w, all_, ("fixed_number", 20), False)]
    for fname, v, ww, acc, bspec, log in cases:
        head, data = _read_dat(os.path.join(tmp_path, fname))
        lo, vol, cts = _gen1d(v, ww, acc, bspec, n, log)
        print(" ", fname, len(cts), flush=True)
        assert len(data) == len(cts)
        centres = lo + (0.5 + np.arange(len(cts))) * (vol if len(cts) > 1 else 0.0)
        # the files hold ten significant digits
        assert np.allclose(data[:, 0], centres, rtol=1e-9, atol=1e-9 * abs(lo))
        # binned on the device: a particle within rounding of a bin edge may sit one bin over
        scale = np.abs(ww).max() / vol * (vol / (np.exp(lo) * np.expm1(vol)) if log else 1.0) * 4.0
        assert np.abs(data[:, 1] - cts).sum() <= scale + 1e-9 * np.abs(cts).sum(), fname
    head, _ = _read_dat(os.path.join(tmp_path, "run_photon_energy.dat"))
    assert head == ["energy\thgram/density/energy", "GeV\t1/GeV"]
    head, data = _read_dat(os.path.join(tmp_path, "run_photon_angle_x-angle_y.dat"))
    assert head == ["angle_x\tangle_y\thgram/density/angle_x_angle_y", "rad\trad\t1/(rad.rad)"]
    ax, ay = _accessor(part, "angle_x"), _accessor(part, "angle_y")
    sx, sy = (ax.max() - ax.min()) / 24, (ay.max() - ay.min()) / 24
    bx = np.floor((ax - ax.min()) / sx).astype(np.int64)
    by = np.floor((ay - ay.min()) / sy).astype(np.int64)
    f = by * 24 + bx
    ref = np.bincount(f[f < 576], weights=w[f < 576], minlength=576) / (sx * sy)
    assert len(data) == 576
    assert np.abs(data[:, 2] - ref).sum() <= 4.0 * w.max() / (sx * sy) + 1e-9 * np.abs(ref).sum()


def _project(stokes, u, axis):
    """Polarization projected on an axis (StokesVector::project_onto,
    geometry/polarization.rs:159-211) for arrays of Stokes vectors and momenta."""
    si, sq, su, sv = (stokes[:, k] for k in range(4))
    n = u[:, 1:4] / np.linalg.norm(u[:, 1:4], axis=1)[:, None]
    mag = np.hypot(n[:, 0], n[:, 2])
    x = np.where(mag[:, None] == 0.0, np.array([1.0, 0.0, 0.0]),
                 np.stack([n[:, 2] / np.where(mag == 0, 1, mag), 0 * mag, -n[:, 0] / np.where(mag == 0, 1, mag)], 1))
    y = np.cross(x, n)
    xa, ya = x @ axis, y @ axis
    unpol = 0.5 * (xa ** 2 + ya ** 2)
    frac = np.hypot(np.hypot(sq, su), sv) / si
    with np.errstate(all="ignore"):
        q, uu, v = sq / (frac * si), su / (frac * si), sv / (frac * si)
        s2 = uu * uu + v * v
        ex = np.where(1.0 + q < 1e-12, np.sqrt(0.5 * (0.5 * s2 + 0.125 * s2 * s2)), np.sqrt(0.5 * (1.0 + q)))
        eyr = np.where(ex == 0.0, 1.0, 0.5 * uu / ex)
        eyi = np.where(ex == 0.0, 0.0, -0.5 * v / ex)
    pol = (ex * xa + eyr * ya) ** 2 + (eyi * ya) ** 2
    return np.where(frac > 0.0, frac * pol + (1.0 - frac) * unpol, unpol)


def test_projection_restatement_matches_reference_cases():
    """The reference's project_polarization cases (particle/mod.rs:366-420): a photon
    linearly polarized along e1 (S_1 = 1) or along e1 x k (S_1 = -1) projects as
    the squared component of that vector on the axis."""
    for theta, phi in [(math.pi, 0.0), (0.75 * math.pi, 0.0), (0.6 * math.pi, math.pi / 2), (0.8 * math.pi, 1.0),
                       (0.2 * math.pi, 4.0)]:
        k = np.array([1.0, math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
        h = math.hypot(k[1], k[3])
        e1 = np.array([k[3] / h, 0.0, -k[1] / h])
        e2 = np.cross(e1, k[1:])
        e2 /= np.linalg.norm(e2)
        for s1, eps in ((1.0, e1), (-1.0, e2)):
            for axis in (np.array([1.0, 0, 0]), np.array([0, 1.0, 0])):
                got = _project(np.array([[1.0, s1, 0.0, 0.0]]), k[None, :], axis)[0]
                assert abs(got - (eps @ axis) ** 2) < 1e-6
        # unpolarized light splits the transverse intensity evenly
        for axis in (np.array([1.0, 0, 0]), np.array([0, 1.0, 0])):
            got = _project(np.array([[1.0, 0.0, 0.0, 0.0]]), k[None, :], axis)[0]
            assert abs(got - 0.5 * ((e1 @ axis) ** 2 + (e2 @ axis) ** 2)) < 1e-12


@pytest.mark.gpu
def test_polarization_projection_on_device(engine):
    """The device's pol_x / pol_y (quantity and weight) against the restatement above."""
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(20_000)
    opts.pol_resolved = True
    engine.set_laser(laser)
    engine.build_beam(beam, seed=6)
    cnt = engine.run(opts, seed=6)
    sec, _ = engine.fetch(cnt)
    ph = sec["species"] == 2
    print("pol projection", int(ph.sum()), flush=True)
    for q, wfn, axis in ((27, 0, np.array([1.0, 0, 0])), (28, 0, np.array([0, 1.0, 0]))):
        v = _project(sec["pol"][ph], sec["u"][ph], axis)
        counts, lo, hi, total = engine.histogram(2, q, 40, wfn)
        assert (lo, hi) == (pytest.approx(v.min(), abs=1e-12), pytest.approx(v.max(), abs=1e-12))
        ref = np.bincount(np.minimum(np.floor((v - lo) / ((hi - lo) / 40)).astype(int), 40), minlength=41)[:40]
        assert np.abs(counts - ref).sum() <= 4
    for wfn, axis in ((3, np.array([1.0, 0, 0])), (4, np.array([0, 1.0, 0]))):
        _, _, _, total = engine.histogram(2, 14, 10, wfn)  # energy, weighted by the projection
        assert total == pytest.approx(np.sum(sec["weight"][ph] * _project(sec["pol"][ph], sec["u"][ph], axis)),
                                      rel=1e-10)


@pytest.mark.gpu
def test_intermediate_photons(engine, tmp_path):
    """Histograms and statistics of the intermediate (decayed) photons: species 3
    selects the buffer ptg_fetch returns as intermediates."""
    from ptarmigan_amd import workloads
    # no event biasing: a biased pair event removes the photon only with probability 1 / rate
    laser, beam, opts = workloads.focused_lcfa_cascade(400_000)
    opts.keep_decayed_photons = True
    engine.set_laser(laser)
    engine.build_beam(beam, seed=21)
    cnt = engine.run(opts, seed=21)
    sec, inter = engine.fetch(cnt, intermediates=True)
    print("intermediates", len(inter["species"]), cnt["pair_events"], flush=True)
    assert len(inter["species"]) > 10
    e = 0.510999 * inter["u"][:, 0]
    counts, lo, hi, total = engine.histogram(O.INTERMEDIATE, 14, 16)
    assert (lo, hi) == (pytest.approx(e.min(), rel=1e-15), pytest.approx(e.max(), rel=1e-15))
    assert total == pytest.approx(inter["weight"].sum(), rel=1e-12)
    ref = np.bincount(np.floor((e - lo) / ((hi - lo) / 16)).astype(int), weights=inter["weight"], minlength=17)[:16]
    assert np.abs(counts - ref).sum() <= 2.0 * inter["weight"].max()
    st = O.Statistic("mean energy", {})
    assert st.evaluate(engine, O.INTERMEDIATE, "intermediate") == pytest.approx(e.mean(), rel=1e-10)
    deck = Deck.from_string("output:\n  intermediate: [energy]\n")
    files = O.write_distributions(engine, deck, sec, str(tmp_path), "", intermediates=inter)
    assert [os.path.basename(f) for f in files] == ["intermediate_energy.dat"]
