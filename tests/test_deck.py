"""Input decks: expressions, the constants block and the driver's key semantics.

The parser cases restate the reference's own tests of its input reader
(src/input/mod.rs:310-410: config_parser and looper). The rest checks each
mapped parameter against what the driver computes from the same keys
(src/main.rs:339-880), with values fixed by closed forms: 2 pi hbar c / E for a
wavelength given as a photon energy, c t / lambda for a duration, charge / (e n)
for a weight, J -> MeV for min_energy.
"""
import math

import numpy as np
import pytest

from ptarmigan_amd import deck as D
from ptarmigan_amd.engine import Envelope, LaserKind, Polarization, Species

CONFIG = """---
control:
  dx: 0.001
  nx: 4000
  ne: sin(a * x)
  ib: a * b^3

extra:
  dx: 160
  r: [0.0, b, 1.0, 2.0 * a]
  s: [0.0, none]

constants:
  N: 2.0
  a: N * pi
  b: 17.0
  y: 1.0

deep:
  nested:
    section:
      key: 1.0
"""


def test_reference_config_parser_cases():
    c = D.Deck.from_string(CONFIG)
    assert c.number("control:dx") == 0.001
    assert c.count("control:nx") == 4000
    assert c.number("control:ib") == 2.0 * math.pi * 17.0 ** 3
    assert c.number("extra:dx") == 160.0
    ne = c.func("control:ne", "x")
    assert ne(0.6) == math.sin(2.0 * math.pi * 0.6)
    with pytest.raises(D.DeckError):
        c.func("control:ne", "y")
    assert c.numbers("extra:r") == [0.0, 17.0, 1.0, 4.0 * math.pi]
    with pytest.raises(D.DeckError):
        c.numbers("extra:s")
    assert c.number("deep:nested:section:key") == 1.0
    assert D.evaluate("1.0 / (1.0 + y)", c.context) == 0.5
    assert c.contains("extra") and c.contains("deep:nested") and not c.contains("control:absent")


def test_reference_looper_cases():
    c = D.Deck.from_string("---\nlaser:\n    a0: 10.0\n")
    assert c.number_loop("laser:a0") == [10.0]
    c = D.Deck.from_string("---\nlaser:\n    a0:\n        start: 1.0\n        stop: 10.0\n        step: 2.0\n")
    assert c.number_loop("laser:a0") == [1.0, 3.0, 5.0, 7.0, 9.0]
    # accumulation as the reference does it: 1.0 + 0.1 + ... passes 1.5 after five steps
    c = D.Deck.from_string("---\nx:\n  start: 1.0\n  stop: 1.5\n  step: 0.1\n")
    assert len(c.number_loop("x")) == 5


@pytest.mark.parametrize("expr,value", [
    ("1 / 2", 0.0),                  # integers stay integers
    ("1.0 / 2", 0.5),
    ("7 % 3", 1.0),
    ("-7 / 2", -3.0),                # truncation toward zero
    ("2^3^2", 512.0),                # right-associative power
    ("-2^2", -4.0),                  # power binds tighter than negation
    ("2 * 3 + 4 * 5", 26.0),
    ("(2 + 3) * 4", 20.0),
    ("1.0e-6 + 2e3", 2000.000001),
    ("atan2(1.0, 1.0) / degree", 45.0),
    ("round(2.5) + round(-2.5)", 0.0),  # half away from zero
    ("signum(-3.0) + abs(-2.0)", 1.0),
    ("step(0.5, 0.0, 1.0) + step(1.0, 0.0, 1.0)", 1.0),  # [min, max)
    ("gauss(1.0, 1.0, 2.0)", 1.0),
    ("cbrt(-27.0)", -3.0),
    ("ln(exp(2.5))", 2.5),
    ("floor(-1.5) + ceil(1.2)", 0.0),
])
def test_expressions(expr, value):
    assert D.evaluate(expr) == pytest.approx(value, rel=1e-15, abs=1e-15)


def test_reference_constants_and_functions():
    ctx = {}
    # the reference's rounded SI values; e and qe are the electron's (negative) charge
    assert D.evaluate("e", ctx) == -1.602177e-19 and D.evaluate("qe", ctx) == -1.602177e-19
    assert D.evaluate("MeV") == 1.602177e-13 and D.evaluate("GeV / keV") == pytest.approx(1e6, rel=1e-15)
    assert D.evaluate("me * c^2 / MeV") == pytest.approx(0.5109989, rel=2e-6)
    assert D.evaluate("mp / me") == pytest.approx(1836.15, rel=1e-5)
    # erf is the reference's tanh approximation, good to 4e-4 against the true erf
    for x in np.linspace(-3, 3, 61):
        assert abs(D.evaluate(f"erf({float(x)!r})") - math.erf(x)) < 4e-4
    # critical density of a 0.8 um laser: eps0 m omega^2 / e^2 = 1.742e27 m^-3
    assert D.evaluate("critical(2.0 * pi * c / (0.8 * micro))") == pytest.approx(1.742e27, rel=1e-3)
    assert math.isnan(D.evaluate("sqrt(-1.0)")) and D.evaluate("1.0 / 0") == math.inf
    for bad in ("unknown_name", "sin(", "2 +", "foo(1.0)", "1.0 2.0", "sqrt(1.0, 2.0)"):
        with pytest.raises(D.DeckError):
            D.evaluate(bad)


def test_constants_block_in_order():
    c = D.Deck.from_string("constants:\n  a0: 15.0\n  gamma0: 1000.0 / 0.511\n  max_angle: atan(a0 / gamma0)\n"
                           "  flag: true\n")
    assert c.context["gamma0"] == pytest.approx(1956.95, abs=5e-3)  # the documented values
    assert c.context["max_angle"] == pytest.approx(7.6649e-3, rel=1e-4)
    assert "flag" not in c.context  # entries that are not numbers or strings are passed over
    with pytest.raises(D.DeckError):
        D.Deck.from_string("constants:\n  b: 2.0 * a\n  a: 1.0\n")  # no forward definitions
    with pytest.raises(D.DeckError):
        D.Deck.from_string("")


DECK = """---
control:
  radiation_reaction: true
  pair_creation: true
  rng_seed: 7
  dt_multiplier: 0.5
  increase_pair_rate_by: 10.0
  stop_at_time: 100.0 * femto

laser:
  a0:
    start: 1.0
    stop: 2.0
    step: 0.5
  omega: 1.55 * eV
  polarization: linear @ 30.0 * degree
  n_cycles: 16.0

beam:
  n: 1000
  species: positron
  charge: 1.0 * nano
  gamma: energy / (me * c^2)
  sigma: 0.01 * energy / (me * c^2)
  radius: [2.0 * micro, normally_distributed, 6.0 * micro]
  length: 10.0 * micro
  collision_angle: 5.0 * degree
  collision_plane: vertical
  rms_divergence: 50.0 * micro
  offset: [1.0 * micro, 0.0, -2.0 * micro]
  energy_chirp: 0.2
  stokes_pars: [0.0, 0.0, 1.0]

output:
  min_energy: 100.0 * MeV
  max_angle: 10.0 * milli
  discard_background: true

constants:
  energy: 10.0 * GeV
"""


def test_driver_key_semantics():
    d = D.Deck.from_string(DECK)
    cols, seed, stop = d.collisions()
    assert seed == 7 and stop == ("given_time", pytest.approx(1e-13))
    assert [c[0].a0 for c in cols] == [1.0, 1.5, 2.0]
    laser, beam, opts = cols[0]
    assert laser.kind == LaserKind.PLANE_WAVE  # no waist, LCFA off
    assert laser.wavelength == pytest.approx(1239.84198e-9 / 1.55, rel=2e-6)  # 2 pi hbar c / E
    assert laser.polarization == Polarization.LINEAR and laser.pol_angle == pytest.approx(math.pi / 6)
    assert laser.envelope == Envelope.COS_SQUARED and laser.n_cycles == 16.0
    gamma = 10.0e9 * 1.602177e-19 / (9.109383e-31 * 2.997925e8 ** 2)
    assert beam.species == Species.POSITRON and beam.n == 1000
    assert beam.gamma == pytest.approx(gamma, rel=1e-14) and beam.sigma == pytest.approx(0.01 * gamma, rel=1e-14)
    assert beam.weight == pytest.approx(1e-9 / (1.602177e-19 * 1000), rel=1e-14)
    assert (beam.radius, beam.radial, beam.max_radius) == (pytest.approx(2e-6), 1, pytest.approx(6e-6))
    assert beam.length == pytest.approx(1e-5) and beam.rms_div == pytest.approx(5e-5)
    assert beam.angle == pytest.approx(math.radians(5.0)) and beam.collision_plane_angle == pytest.approx(math.pi / 2)
    assert beam.offset == (pytest.approx(1e-6), 0.0, pytest.approx(-2e-6))
    assert beam.energy_chirp == 0.2 and beam.pol == (1.0, 0.0, 0.0, 1.0)
    assert opts.dt_multiplier == 0.5 and opts.rate_increase == 10.0 and opts.t_stop == pytest.approx(1e-13)
    assert opts.min_energy == pytest.approx(100.0) and opts.max_angle == pytest.approx(0.01)
    assert opts.discard_bg_e and opts.discard_bg_ph and opts.tracking_photons and not opts.classical


def _with(text, **edits):
    d = D.Deck.from_string(text)
    for path, v in edits.items():
        node = d.data
        parts = path.split("__")
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        if v is None:
            node.pop(parts[-1], None)
        else:
            node[parts[-1]] = v
    return D.Deck(d.data)


def test_driver_defaults_and_alternatives():
    d = _with(DECK, laser__waist="4.0 * micro", laser__fwhm_duration="30.0 * femto", laser__polarization="circular",
              control__lcfa=True, beam__charge=None, beam__n_real=5.0e9, output__discard_background=None,
              output__discard_background_e=True, control__stop_at_time="auto", control__classical="true",
              control__pair_creation=None, beam__radius="1.0 * micro", beam__collision_plane="30.0 * degree")
    cols, _, stop = d.collisions()
    laser, beam, opts = cols[0]
    assert laser.kind == LaserKind.FAST_FOCUSED_LASER and laser.waist == pytest.approx(4e-6)
    assert laser.envelope == Envelope.GAUSSIAN  # the default of a focused laser
    assert laser.n_cycles == pytest.approx(2.997925e8 * 30e-15 / laser.wavelength, rel=1e-14)
    assert laser.polarization == Polarization.CIRCULAR and laser.chirp_b == 0.0
    assert beam.weight == pytest.approx(5.0e6) and (beam.radial, beam.radius) == (0, pytest.approx(1e-6))
    assert beam.collision_plane_angle == pytest.approx(math.pi / 6)
    assert opts.discard_bg_e and not opts.discard_bg_ph
    assert stop[0] == "same_time" and opts.t_stop == math.inf
    assert opts.classical and not opts.tracking_photons  # pair creation defaults to off when classical
    d = _with(DECK, laser__envelope="flattop", laser__n_cycles=0.5)
    with pytest.raises(D.DeckError, match="flattop"):
        d.collisions()
    d = _with(DECK, laser__envelope="sawtooth")  # reported, then the default
    with pytest.warns(UserWarning):
        assert d.collisions()[0][0][0].envelope == Envelope.COS_SQUARED


@pytest.mark.parametrize("edit,match", [
    ({"beam__n_real": 1.0}, "conflicting"),
    ({"beam__stokes_pars": [1.0, 1.0, 0.0]}, "S_1"),
    ({"beam__stokes_pars": [1.0, 0.0]}, "three Stokes"),
    ({"beam__offset": [1.0, 0.0]}, "three-vector"),
    ({"beam__energy_chirp": 1.5}, "chirp"),
    ({"beam__species": "muon"}, "species"),
    ({"beam__radius": ["1.0 * micro", "gaussian"]}, "radius"),
    ({"beam__n": 1.0e3}, "beam:n"),
    ({"laser__polarization": "elliptical"}, "polarization"),
    ({"control__classical": "maybe"}, "classical"),
    ({"control__classical": "gaunt_factor_corrected"}, "LCFA"),
    ({"control__increase_pair_rate_by": 0.5}, ">= 1.0"),
    ({"control__increase_pair_rate_by": "auto"}, "pass an Engine"),
    ({"control__select_multiplicity": 2}, "incompatible"),
    ({"control__stop_at_time": "soon"}, "stop_at_time"),
    ({"beam__from_hdf5": {"file": "x.h5"}}, "distance_bt_ips is required"),
    ({"beam__spectrum": {"function": "1.0", "file": "x.dat", "min": 1.0}}, "not both"),
    ({"beam__spectrum": {"function": "sqrt(gamma - 1e6)", "min": 1.0, "max": 2.0}}, "non-numerical"),
    ({"beam__gamma": None}, "beam:gamma is required"),
])
def test_driver_errors(edit, match):
    with pytest.raises(D.DeckError, match=match):
        _with(DECK, **edit).collisions()


@pytest.mark.gpu
def test_deck_runs_an_a0_loop_as_one_batch(engine):
    """Three a0 values of a plane wave run as one batch: every collision gets its
    own primaries and final state; photon yields grow with a0 and agree with
    separate runs of each collision within their statistics."""
    d = _with(DECK, beam__n=4000, beam__charge=None, control__increase_pair_rate_by=None,
              control__stop_at_time=None, output__min_energy=None, output__max_angle=None,
              output__discard_background=None, beam__species="electron", laser__polarization="linear")
    res = D.run_deck(d, engine)
    assert [r[0] for r in res] == [1.0, 1.5, 2.0]
    yields = []
    for a0, part, cnt in res:
        print("deck batch a0", a0, {k: cnt[k] for k in ("emissions", "error_flags")}, flush=True)
        assert cnt["error_flags"] == 0
        e = part["species"] == 0
        assert abs(part["weight"][e].sum() - 4000.0) < 1e-9  # one electron per primary
        yields.append(np.sum(part["species"] == 2))
    assert yields[0] < yields[1] < yields[2]
    cols, seed, _ = d.collisions()
    for (laser, beam, opts), y in zip(cols, yields):
        engine.set_laser(laser)
        engine.build_beam(beam, seed=seed + 1)
        c = engine.run(opts, seed=seed + 1)
        sec, _ = engine.fetch(c)
        y1 = np.sum(sec["species"] == 2)
        assert abs(y - y1) < 5.0 * math.sqrt(2.0 * max(y, y1)), (laser.a0, y, y1)


FOCUSED_DECK = """---
control:
  lcfa: true
  pair_creation: false
  rng_seed: 3
laser:
  a0: 10.0
  wavelength: 0.8 * micro
  waist: 4.0 * micro
  fwhm_duration: 30.0 * femto
  polarization: linear
beam:
  n: 4000
  species: electron
  charge: 1.0 * nano
  gamma: 16.5 * GeV / (me * c^2)
  radius: 2.0 * micro
  length: 10.0 * micro
"""


@pytest.mark.gpu
def test_deck_reports_laser_depletion(engine):
    """The driver's depletion fraction (src/main.rs:1136-1147, 1198-1215): the
    weighted absorbed energy of the final state, from the device's reduction,
    against the same sum over the fetched particles, over Field::energy."""
    from ptarmigan_amd import laser_energy
    d = D.Deck.from_string(FOCUSED_DECK)
    (a0, part, cnt), = D.run_deck(d, engine)
    laser = d.collisions()[0][0][0]
    absorbed_j = float(np.sum(part["weight"] * part["work"])) * 0.510999 * 1e6 * 1.602177e-19
    energy, unit = laser_energy(laser)
    print("depletion", cnt["laser_depletion"], absorbed_j, energy, unit, flush=True)
    assert unit == "J" and energy > 0.0
    assert cnt["laser_depletion"] == pytest.approx(absorbed_j / energy, rel=1e-9, abs=1e-300)
    assert absorbed_j != 0.0


def test_depletion_warning_threshold():
    """Above a 10 % depletion the driver warns (src/main.rs:1208-1215)."""
    import warnings
    from ptarmigan_amd import Envelope, Laser, LaserKind, Polarization, laser_energy
    laser = Laser(LaserKind.FAST_FOCUSED_LASER, 10.0, 0.8e-6, 11.24, 4e-6, Envelope.GAUSSIAN, Polarization.LINEAR)
    e, _ = laser_energy(laser)
    mev = e / (1e6 * 1.602177e-19)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert D.laser_depletion(laser, 0.05 * mev) == pytest.approx(0.05)
    with pytest.warns(UserWarning, match="depletion of 20.00%"):
        assert D.laser_depletion(laser, 0.2 * mev) == pytest.approx(0.2)


def test_command_line_reports_bad_decks(tmp_path, capsys):
    from ptarmigan_amd.__main__ import main
    assert main([]) == 2
    bad = tmp_path / "bad.yml"
    bad.write_text("laser:\n  a0: 1.0\n")
    assert main([str(bad)]) == 1
    assert "error:" in capsys.readouterr().err


class _RateStub:
    """Records the arguments of the rate evaluations and returns fixed rates."""

    def __init__(self, f_ph=3.0, f0=1.0, f1=2.0, t=(0.1, 0.3), lcfa_rate=5.0):
        self.calls, self.f_ph, self.f0, self.f1, self.t, self.lcfa_rate = [], f_ph, f0, f1, t, lcfa_rate

    def eval(self, fn, x, out_w):
        self.calls.append((fn, list(x[0])))
        out = {8: [self.f_ph, 1.0], 16: [self.f_ph, 1.0], 12: [self.f0, self.f1, 1.0], 24: [self.f0, self.f1, 2.0],
               0: [self.lcfa_rate], 2: list(self.t)}[fn]
        return np.array([out])


def _mdot(a, b):
    return a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]


@pytest.mark.parametrize("pol", ["linear", "circular"])
def test_auto_pair_rate_kinematics(pol):
    """increase_pair_rate_by: auto (src/main.rs:290-319): the (a, eta) handed to the
    device's rate functions, recomputed here with explicit four-vectors as the
    reference builds them, and the ratio of the two probabilities."""
    d = _with(DECK, control__increase_pair_rate_by="auto", laser__polarization=pol)
    stub = _RateStub()
    cols, _, _ = d.collisions(stub)
    laser, beam, opts = cols[1]
    g = beam.gamma
    k = 2.997925e8 * 1.28808867e-21 * 2.0 * math.pi / laser.wavelength * np.array([1.0, 0.0, 0.0, 1.0])
    ell = g * np.array([1.0, 0.0, 0.0, -1.0])
    u = np.array([math.sqrt(1.0 + g * g), 0.0, 0.0, -g])
    a_rms = laser.a0 / math.sqrt(2.0) if pol == "linear" else laser.a0
    q = u + a_rms ** 2 * k / (2.0 * _mdot(k, u))
    assert _mdot(q, q) == pytest.approx(1.0 + a_rms ** 2, rel=1e-6)  # q^2 = 1 + a^2 (cancellation at gamma^2)
    calls = stub.calls[2:4]  # the second a0
    a_arg = a_rms * math.sqrt(2.0) if pol == "linear" else a_rms
    assert calls[0][1][:2] == [pytest.approx(a_arg, rel=1e-14), pytest.approx(_mdot(k, q), rel=1e-12)]
    assert calls[1][1][:2] == [pytest.approx(a_arg, rel=1e-14), pytest.approx(_mdot(k, ell), rel=1e-12)]
    # P = alpha f dphi / eta, dphi = eta dt / (tau_c E): photon f / q0 over pair (f0 + f1) / (2 ell0)
    want = max((3.0 / q[0]) / (0.5 * 3.0 / ell[0]), 1.0)
    assert opts.rate_increase == pytest.approx(want, rel=1e-12)
    # a zero pair rate leaves the rate alone
    cols, _, _ = d.collisions(_RateStub(f0=0.0, f1=0.0))
    assert all(c[2].rate_increase == 1.0 for c in cols)


def test_auto_pair_rate_lcfa():
    d = _with(DECK, control__increase_pair_rate_by="auto", control__lcfa=True)
    stub = _RateStub(t=(0.2, 0.4), lcfa_rate=1.0e20)
    cols, _, _ = d.collisions(stub)
    laser, beam, opts = cols[0]
    chi = 2.0 * beam.gamma * laser.a0 * 1.26e-6 / (0.510999e6 * laser.wavelength)
    assert stub.calls[0] == (0, [pytest.approx(chi, rel=1e-14), beam.gamma]) and stub.calls[1][1] == [pytest.approx(chi)]
    pair = 7.29735257e-3 * chi * 0.3 / (1.28808867e-21 * beam.gamma)
    assert opts.rate_increase == pytest.approx(1.0e20 / pair, rel=1e-12)
    with pytest.raises(D.DeckError, match=">= 1.0"):  # the LCFA ratio is not clamped
        _with(DECK, control__increase_pair_rate_by="auto", control__lcfa=True).collisions(_RateStub(lcfa_rate=1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("edit", [{"laser__polarization": "linear"}, {"laser__polarization": "circular"},
                                  {"control__lcfa": True, "laser__a0": 20.0}])
def test_auto_pair_rate_on_device(engine, edit):
    """The rates of increase_pair_rate_by: auto, evaluated by the device's own rate
    functions at the deck's (a0, gamma): finite, at least 1, growing as a0 falls for
    the LMA (pair creation is the rarer process further below threshold)."""
    d = _with(DECK, control__increase_pair_rate_by="auto", beam__charge=None, **edit)
    cols, _, _ = d.collisions(engine)
    rates = [c[2].rate_increase for c in cols]
    print("auto pair rate", edit, rates, flush=True)
    assert all(math.isfinite(r) and r >= 1.0 for r in rates)
    if "control__lcfa" not in edit:
        assert rates[0] >= rates[-1]


def _rot(axis, t):
    """ThreeVector::rotate_around (geometry/three_vector.rs:79-93) as a matrix."""
    ax, ay, az = axis
    s, c = math.sin(t), math.cos(t)
    return np.array([[c + ax * ax * (1 - c), ax * ay * (1 - c) - az * s, ax * az * (1 - c) + ay * s],
                     [ay * ax * (1 - c) + az * s, c + ay * ay * (1 - c), ay * az * (1 - c) - ax * s],
                     [az * ax * (1 - c) - ay * s, az * ay * (1 - c) + ax * s, c + az * az * (1 - c)]])


def _beam_basis(theta, phi):
    """The rotation of Particle::to_beam_coordinate_basis (particle/mod.rs:256-300)."""
    cos_v = math.hypot(math.cos(theta), math.sin(theta) * math.cos(phi))
    if phi == 0.0 or theta == 0.0:
        v, h = 0.0, math.pi - theta
    else:
        v = math.copysign(math.acos(cos_v), math.sin(theta) * math.sin(phi))
        h = math.copysign(math.acos(-math.cos(theta) / cos_v), math.cos(phi) * math.sin(theta))
    return _rot((1.0, 0.0, 0.0), -v) @ _rot((0.0, 1.0, 0.0), h)


@pytest.mark.parametrize("theta,phi", [(0.0, 0.0), (0.3, 0.0), (0.3, 0.7), (1.2, -2.0), (0.05, math.pi / 2)])
def test_beam_basis_sends_the_centroid_to_plus_z(theta, phi):
    centroid = np.array([-math.cos(phi) * math.sin(theta), -math.sin(theta) * math.sin(phi), -math.cos(theta)])
    assert np.allclose(_beam_basis(theta, phi) @ centroid, [0.0, 0.0, 1.0], atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,theta,phi", [(1, 0.2, 0.6), (2, 0.0, 0.0)])
def test_finish_matches_the_driver(engine, mode, theta, phi):
    """ptg_finish against the driver's post-processing restated with numpy
    (src/main.rs:1150-1196): the drift to a common time, then the beam frame."""
    from ptarmigan_amd import workloads
    laser, beam, opts = workloads.plane_wave_lma(5000)
    beam.angle, beam.collision_plane_angle = theta, phi
    engine.set_laser(laser)
    engine.build_beam(beam, seed=12)
    cnt = engine.run(opts, seed=12)
    before, _ = engine.fetch(cnt)
    t_stop = 2.0e-12
    t = engine.finish(mode, t_stop, True, theta, phi)
    after, _ = engine.fetch(cnt)
    print("finish", mode, theta, phi, t, flush=True)
    r, u = before["r"].copy(), before["u"]
    ct = r[:, 0].max() if mode == 1 else t_stop * 2.997925e8
    assert t == pytest.approx(ct / 2.997925e8, rel=1e-15)
    r += u * ((ct - r[:, 0]) / u[:, 0])[:, None]
    R = _beam_basis(theta, phi)
    scale = np.abs(r[:, 1:]).max()
    assert np.allclose(after["r"][:, 0], r[:, 0], rtol=1e-13)
    assert np.allclose(after["r"][:, 1:], r[:, 1:] @ R.T, rtol=1e-12, atol=1e-12 * scale)
    assert np.allclose(after["u"][:, 1:], u[:, 1:] @ R.T, rtol=1e-12, atol=1e-12 * np.abs(u).max())
    assert np.array_equal(after["u"][:, 0], u[:, 0])
    assert np.allclose(after["u0"][:, 1:], before["u0"][:, 1:] @ R.T, rtol=1e-12, atol=1e-12 * np.abs(u).max())
    s0 = np.abs(before["r0"][:, 1:]).max()
    assert np.allclose(after["r0"][:, 1:], before["r0"][:, 1:] @ R.T, rtol=1e-12, atol=1e-12 * s0)


@pytest.mark.gpu
def test_finish_rotates_intermediates_without_drift(engine):
    """The intermediate (decayed) photons take the beam-frame rotation but not the
    drift to the common time: the driver drifts bunch.iter_mut() and rotates
    bunch.iter_all_mut() (src/main.rs:1150-1196, particle/shower.rs:64-69)."""
    from ptarmigan_amd import Beam, CollideOptions, Envelope, Laser, LaserKind, Polarization
    laser = Laser(LaserKind.FAST_PLANE_WAVE, 40.0, 0.8e-6, 8.0, 0.0, Envelope.COS_SQUARED, Polarization.LINEAR)
    theta, phi = 0.25, -0.8
    beam = Beam(n=20_000, gamma=16500.0 / 0.510999, angle=theta, collision_plane_angle=phi)
    opts = CollideOptions(tracking_photons=True, keep_decayed_photons=True)
    engine.set_laser(laser)
    engine.build_beam(beam, seed=21)
    cnt = engine.run(opts, seed=21)
    sec0, inter0 = engine.fetch(cnt, intermediates=True)
    assert len(inter0["species"]) > 50
    engine.finish(1, 0.0, True, theta, phi)
    sec1, inter1 = engine.fetch(cnt, intermediates=True)
    print("intermediates", len(inter0["species"]), flush=True)
    R = _beam_basis(theta, phi)
    for k in ("r", "u", "r0", "u0"):
        scale = np.abs(inter0[k][:, 1:]).max()
        assert np.array_equal(inter1[k][:, 0], inter0[k][:, 0]), k   # times and energies unchanged: no drift
        assert np.allclose(inter1[k][:, 1:], inter0[k][:, 1:] @ R.T, rtol=1e-12, atol=1e-12 * scale), k
    # the final state is drifted to the latest time as well
    ct = sec0["r"][:, 0].max()
    assert np.allclose(sec1["r"][:, 0], ct, rtol=1e-13)


def test_spectrum_tables(tmp_path):
    """GammaDistribution::custom rows (particle/dstr.rs:88-106) from a function and
    from a file, and the spectrum's mean gamma."""
    d = _with(DECK, beam__spectrum={"function": "gamma / 1000.0", "min": 1000.0, "max": 2000.0, "resolution": 4})
    cols, _, _ = d.collisions()
    kind, table, _ = d.spectrum
    vals = [1.0, 1.25, 1.5, 1.75]
    assert kind == 1 and np.allclose(table[:, 0], [1000, 1250, 1500, 1750, 2000])
    assert np.allclose(table[:, 1], np.cumsum([0.0] + vals) / sum(vals), rtol=1e-15)
    assert cols[0][1].gamma == pytest.approx(sum(g * v for g, v in zip([1000, 1250, 1500, 1750], vals)) / sum(vals))
    f = tmp_path / "spec.dat"
    f.write_text("1.0\n3.0\n0.0\n")
    (tmp_path / "deck.yml").write_text(DECK.replace("  n: 1000\n", "  n: 1000\n  spectrum:\n    file: spec.dat\n"
                                                                  "    min: 100.0\n    step: 10.0\n"))
    d = D.Deck.from_file(str(tmp_path / "deck.yml"))
    d.collisions()
    assert np.allclose(d.spectrum[1], [[100.0, 0.0], [110.0, 0.25], [120.0, 1.0], [130.0, 1.0]])
    d = _with(DECK, beam__species="photon", beam__bremsstrahlung_source=True, beam__gamma_min=100.0, beam__energy_chirp=0.0)
    cols, _, _ = d.collisions()
    assert d.spectrum == (2, None, (100.0, pytest.approx(cols[0][1].gamma)))


def _photon_primaries(engine, d, seed):
    from ptarmigan_amd import workloads
    laser, _, opts = workloads.plane_wave_lma(10)
    cols, _, _ = d.collisions()
    beam = cols[0][1]
    opts.tracking_photons = False  # photons are not transported: the final state is the primaries
    engine.set_laser(laser)
    engine.set_beam_spectrum(*d.spectrum)
    engine.build_beam(beam, seed=seed)
    cnt = engine.run(opts, seed=seed)
    sec, _ = engine.fetch(cnt)
    engine.set_beam_spectrum(0)
    return sec["u"][:, 0]


@pytest.mark.gpu
def test_beam_spectra_on_device(engine):
    """Tabulated and bremsstrahlung spectra sampled by the device beam builder
    against their distributions: a flat table (uniform gamma up to the
    reference's erf approximation, 4e-4) and the bremsstrahlung density
    4/(3x) - 4/3 + x on [x_min, 1]."""
    n = 200_000
    d = _with(DECK, beam__species="photon", beam__n=n, beam__charge=None, beam__energy_chirp=0.0,
              beam__spectrum={"function": "1.0", "min": 1000.0, "max": 2000.0, "resolution": 50})
    g = np.sort(_photon_primaries(engine, d, 3))
    print("tabulated spectrum", g[0], g[-1], flush=True)
    assert g[0] >= 1000.0 and g[-1] <= 2000.0
    ks = np.abs(np.arange(1, n + 1) / n - (g - 1000.0) / 1000.0).max()
    assert ks < 5e-3
    d = _with(DECK, beam__species="photon", beam__n=n, beam__charge=None, beam__energy_chirp=0.0,
              beam__bremsstrahlung_source=True, beam__gamma_min=200.0, beam__gamma=2000.0)
    g = _photon_primaries(engine, d, 4) / 2000.0
    x = np.linspace(0.1, 1.0, 200001)
    p = 4.0 / (3.0 * x) - 4.0 / 3.0 + x
    mean = np.trapezoid(x * p, x) / np.trapezoid(p, x)
    print("brem spectrum", g.min(), g.max(), g.mean(), mean, flush=True)
    assert g.min() >= 0.1 and g.max() <= 1.0
    assert abs(g.mean() - mean) < 5.0 * g.std() / math.sqrt(n)
