"""Distribution output of the final state: the reference's per-species
specifiers `dstr1[:dstr2][:(bins;weight[;cut])]` (src/output/mod.rs:150-343),
binned on the GPU by ptg_histogram / ptg_histogram_2d and written as plain text
(`.dat`) or FITS (`.fits`) files laid out as the reference writes them
(src/output/hgram.rs:446-565).

The binning itself (range, number of bins, which particles count) is the
device's; this module chooses the bins as Histogram::generate_1d/2d does
(src/output/hgram.rs:169-446), applies the height (density, with the per-bin
factor of log-scaled bins) and formats the files.
"""
import math
import re

import numpy as np

from .deck import DeckError, ELEMENTARY_CHARGE, SPEED_OF_LIGHT, evaluate
from .engine import PtgError, Species

# output name -> (ptg_quantity, kind) as src/output/functions.rs:10-107
_Q = {}
for _names, _id, _kind in [
        (("angle_x",), 0, "angle"), (("angle_y",), 1, "angle"), (("theta", "pi_minus_angle"), 2, "angle"),
        (("angle", "polar_angle"), 3, "angle"), (("px", "p_x"), 4, "momentum"), (("py", "p_y"), 5, "momentum"),
        (("pz", "p_z"), 6, "momentum"), (("p_perp",), 7, "momentum"), (("r_perp",), 8, "1"),
        (("r_x", "rx"), 9, "1"), (("r_y", "ry"), 10, "1"), (("p^-", "p-", "p_minus"), 11, "momentum"),
        (("p^+", "p+", "p_plus"), 12, "momentum"), (("gamma",), 13, "1"), (("energy",), 14, "energy"),
        (("unit",), 15, "1"), (("x",), 16, "length"), (("y",), 17, "length"), (("z",), 18, "length"),
        (("radius", "rho"), 19, "length"), (("birth_a",), 20, "1"), (("parent_chi",), 21, "1"),
        (("n_inter", "n_gamma", "n_pos"), 22, "1"), (("absorption", "energy_abs"), 23, "energy"),
        (("S_1", "S1"), 24, "1"), (("S_2", "S2"), 25, "1"), (("S_3", "S3"), 26, "1"),
        (("pol_x", "pol_sigma"), 27, "1"), (("pol_y", "pol_pi"), 28, "1"), (("weight", "number"), 29, "1"),
        (("helicity",), 30, "1")]:
    for _n in _names:
        _Q[_n] = (_id, _kind)
_ALIAS = {"p^-": "p_minus", "p-": "p_minus", "p^+": "p_plus", "p+": "p_plus", "rho": "radius"}
_WEIGHTS = {"weight": 0, "auto": 0, "energy": 1, "helicity": 2, "pol_x": 3, "pol_y": 4}

# unit systems (src/output/units.rs): scale from the default unit (m, MeV, MeV/c) and name
UNIT_SYSTEMS = {
    "auto": {"length": (1.0, "m"), "energy": (1.0, "MeV"), "momentum": (1.0, "MeV/c")},
    "hep": {"length": (1.0e3, "mm"), "energy": (1.0e-3, "GeV"), "momentum": (1.0e-3, "GeV/c")},
    "si": {"length": (1.0, "m"), "energy": (1.0e6 * ELEMENTARY_CHARGE, "J"),
           "momentum": (1.0e6 * ELEMENTARY_CHARGE / SPEED_OF_LIGHT, "kg/m/s")},
}
UNIT_SYSTEMS["HEP"], UNIT_SYSTEMS["SI"] = UNIT_SYSTEMS["hep"], UNIT_SYSTEMS["si"]


def _unit(kind, system):
    if kind == "1":
        return 1.0, "1"
    if kind == "angle":
        return 1.0, "rad"
    return system[kind]


_USIZE = re.compile(r"^\+?\d+$")
_FLOAT = re.compile(r"^[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?|inf|infinity|nan)$", re.IGNORECASE)


def bin_spec(s):
    """BinSpec from a string (src/output/hgram.rs:74-89): a count, a bin size, auto or log."""
    if _USIZE.match(s):
        return ("fixed_number", int(s))
    if _FLOAT.match(s):
        return ("fixed_size", float(s))
    if s == "log":
        return ("log", None)
    return ("auto", None)


class Cut:
    """A range cut `var in min, max` (src/output/mod.rs:60-140): bounds in SI, held in
    the default units; `auto` leaves a side open."""

    def __init__(self, s, context):
        if "," not in s:
            raise DeckError(f"{s!r} is not a filter")
        prefix, suffix = s.split(",", 1)
        words = prefix.split()
        if len(words) < 2 or words[0] not in _Q:  # the second word is read but not checked
            raise DeckError(f"{s!r} is not a filter")
        self.name = _ALIAS.get(words[0], words[0])
        self.quantity, kind = _Q[words[0]]
        lo_expr = "".join(words[2:])
        lo = -math.inf if lo_expr == "auto" else _parse_si(lo_expr, context, s)
        hi = math.inf if suffix.strip() == "auto" else _parse_si(suffix, context, s)
        scale = _unit(kind, UNIT_SYSTEMS["si"])[0]
        self.min, self.max = lo / scale, hi / scale


def _parse_si(expr, context, spec):
    try:
        return float(evaluate(expr, context))
    except DeckError:
        raise DeckError(f"{spec!r} is not a filter")


class Distribution:
    """One specifier of output:electron / photon / positron (DistributionFunction::load)."""

    def __init__(self, spec, context=None):
        ss = spec.split(":")
        bspec, weight, cut = ("auto", None), "weight", None
        if len(ss) >= 2 and ss[-1].startswith("(") and ss[-1].endswith(")"):
            last = ss.pop().lstrip("(").rstrip(")").split(";")
            if len(last) == 1:
                weight = last[0].strip()
            elif len(last) in (2, 3):
                bspec, weight = bin_spec(last[0].strip()), last[1].strip()
                if len(last) == 3:
                    cut = last[2]
        if not ss or any(s not in _Q for s in ss) or weight not in _WEIGHTS or len(ss) > 2:
            raise DeckError(f"{spec!r} is not a distribution function")
        self.spec = spec
        self.names = [_ALIAS.get(s, s) for s in ss]
        self.quantities = [_Q[s][0] for s in ss]
        self.kinds = [_Q[s][1] for s in ss]
        self.bspec = bspec
        self.weight = weight
        self.cut = Cut(cut, context or {}) if cut is not None else None

    def suffix(self):
        s = "_" + "-".join(self.names)
        if self.weight not in ("weight", "auto"):
            s += f"_{self.weight}-weight"
        if self.bspec[0] == "log":
            s += "_log"
        if self.cut is not None:
            s += f"_{self.cut.name}-cut"
        return s

    def histogram(self, engine, species, n_species, system):
        """The histogram of the resident final state (None when there are no particles
        of the species), as Histogram::generate_1d/2d."""
        if n_species == 0:
            return None
        units = [_unit(k, system) for k in self.kinds]
        log = self.bspec[0] == "log"
        cut = (self.cut.quantity, self.cut.min, self.cut.max) if self.cut else None
        wfn = _WEIGHTS[self.weight]
        dim = len(self.quantities)
        # the range of each axis, as the device finds it
        ranges = []
        for a in range(dim):
            try:
                _, lo, hi, _ = engine.histogram(species, self.quantities[a], 1, wfn, log, cut=cut, scale=units[a][0])
            except PtgError as err:
                if err.code != -5:
                    raise
                lo, hi = np.finfo(float).max, -np.finfo(float).max  # nothing accepted
            ranges.append((lo, hi))
        nbins, sizes = [], []
        for lo, hi in ranges:
            if lo == hi:
                nbins.append(1)
            elif self.bspec[0] in ("auto", "log"):
                nbins.append(int(math.ceil(2.0 * float(np.cbrt(float(n_species))))))
            elif self.bspec[0] == "fixed_number":
                nbins.append(self.bspec[1])
            else:
                with np.errstate(all="ignore"):
                    nb = np.ceil(np.float64(hi - lo) / self.bspec[1])
                nbins.append(int(nb) if np.isfinite(nb) and nb > 0 else 0)
            if lo == hi:
                sizes.append(0.0)
            elif self.bspec[0] == "fixed_size":
                sizes.append(self.bspec[1])
            else:
                with np.errstate(all="ignore"):
                    sizes.append(float(np.float64(hi - lo) / nbins[-1]) if nbins[-1] else math.inf)
        empty = any(lo > hi for lo, hi in ranges) or 0 in nbins
        if dim == 1:
            bin_vol = 1.0 if ranges[0][0] == ranges[0][1] else sizes[0]
        else:
            bin_vol = float(np.prod([s if s != 0.0 else 1.0 for s in sizes]))
        if empty:
            cts = np.zeros(int(np.prod(nbins)))
            _, _, _, total = engine.histogram(species, self.quantities[0], 1, wfn, False, range_=(0.0, 1.0))
        else:
            # bins of the chosen size: the device bins over [lo, lo + n * size)
            fixed = [(lo, lo + n * s) if self.bspec[0] == "fixed_size" and lo != hi else (lo, hi)
                     for (lo, hi), n, s in zip(ranges, nbins, sizes)]
            if dim == 1:
                cts, _, _, total = engine.histogram(species, self.quantities[0], nbins[0], wfn, log, range_=fixed[0],
                                                    cut=cut, scale=units[0][0])
            else:
                cts, _, _, total = engine.histogram2d(species, self.quantities[0], self.quantities[1], nbins[0],
                                                      nbins[1], wfn, log, log, cut=cut, scale=[u[0] for u in units],
                                                      ranges=fixed)
            cts = cts.astype(float)
            if log:  # the actual width of each log-scaled bin (hgram.rs:135-139, 286-292)
                for a in range(dim):
                    if sizes[a] == 0.0:
                        continue
                    b = np.arange(nbins[a])
                    f = sizes[a] / (np.exp(ranges[a][0] + b * sizes[a]) * np.expm1(sizes[a]))
                    cts = cts * f if dim == 1 else _scale_axis(cts, nbins, a, f)
        cts = cts / bin_vol
        names = self.names
        if dim == 1:
            name, bunit = f"hgram/density/{names[0]}", f"1/{units[0][1]}"
        else:
            name, bunit = f"hgram/density/{names[0]}_{names[1]}", f"1/({units[0][1]}.{units[1][1]})"
        return Histogram(dim, total, float(n_species), [r[0] for r in ranges], [r[1] for r in ranges], cts, nbins,
                         [0.0] if dim == 1 and nbins[0] <= 1 else sizes, name, bunit, names, [u[1] for u in units])


def _scale_axis(cts, nbins, axis, f):
    """Multiplies each flat bin by the factor of its bin on `axis` (flat = by * n_y + bx)."""
    by, bx = np.divmod(np.arange(len(cts)), nbins[1])
    b = bx if axis == 0 else by
    m = b < len(f)
    out = cts.copy()
    out[m] *= f[b[m]]
    return out


def rust_exp(x, prec, upper=False):
    """Rust's `{:.prec e}` / `{:.prec E}`: no '+' and no zero padding in the exponent."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    mant, exp = f"{x:.{prec}e}".split("e")
    return f"{mant}{'E' if upper else 'e'}{int(exp)}"


class Histogram:
    def __init__(self, dim, total, unweighted_total, min_, max_, cts, bins, bin_sz, name, bunit, axis, unit):
        self.dim, self.total, self.unweighted_total = dim, total, unweighted_total
        self.min, self.max, self.cts, self.bins, self.bin_sz = min_, max_, np.asarray(cts, float), bins, bin_sz
        self.name, self.bunit, self.axis, self.unit = name, bunit, axis, unit

    def plain_text(self):
        """The `.dat` layout (hgram.rs:530-565): axis names and the histogram name,
        units and the height unit, then one line per bin: centres and height."""
        lines = ["\t".join(self.axis) + "\t" + self.name, "\t".join(self.unit) + "\t" + self.bunit]
        index = [0] * self.dim
        for ct in self.cts:
            for j in range(self.dim - 1):
                if index[j] >= self.bins[j]:
                    index[j] -= self.bins[j]
                    index[j + 1] += 1
            coord = [self.min[j] + (0.5 + index[j]) * self.bin_sz[j] for j in range(self.dim)]
            lines.append("".join(rust_exp(c, 9) + "\t" for c in coord) + rust_exp(float(ct), 9))
            index[0] += 1
        return "\n".join(lines) + "\n"

    def fits(self, version="ptarmigan_amd"):
        """A FITS primary HDU (hgram.rs:448-527): 80-character cards, big-endian f64 data,
        each part padded to 2880 bytes."""
        cards = [f"SIMPLE  = {'T':>20} / {'file conforms to FITS standard':<47}",
                 f"BITPIX  = {-64:>20} / {'number of bits per data pixel':<47}",
                 f"NAXIS   = {self.dim:>20} / {'number of data axes':<47}"]
        for n in range(1, self.dim + 1):
            cards.append(f"NAXIS{n:<3}= {self.bins[n - 1]:>20} / {'number of pixels along this axis':<47}")
        cards.append(f"EXTEND  = {'T':>20} / {'dataset may contain extensions':<47}")

        def num(key, v):
            return f"{key:<8}= {rust_exp(v, 9, True):>20}{'':<50}"

        def text(key, v):
            return f"{key:<8}= '{v}'" + " " * max(0, 80 - 12 - len(v))
        for n in range(1, self.dim + 1):
            cards += [num(f"CRPIX{n}", 0.5), num(f"CRVAL{n}", self.min[n - 1]), num(f"CDELT{n}", self.bin_sz[n - 1]),
                      text(f"CNAME{n}", self.axis[n - 1]), text(f"CUNIT{n}", self.unit[n - 1])]
        cards += [text("BUNIT", self.bunit), num("TOTAL", self.total), num("UWTOTAL", self.unweighted_total),
                  text("OBJECT", self.name)]
        if len(self.cts):
            cards += [num("DATAMIN", float(self.cts.min())), num("DATAMAX", float(self.cts.max()))]
        cards.append(f"{'COMMENT   Generated by ' + version:<80}")
        cards.append(f"{'END':<80}")
        head = "".join(c[:80] for c in cards).encode("ascii")
        if len(head) > 2880:
            raise ValueError("FITS header longer than one block")
        head += b" " * (2880 - len(head))
        data = self.cts.astype(">f8").tobytes()
        if len(data) % 2880:
            data += b"\0" * (2880 - len(data) % 2880)
        return head + data

    def write(self, filename, file_format):
        if file_format == "fits":
            with open(filename + ".fits", "wb") as f:
                f.write(self.fits())
            return filename + ".fits"
        with open(filename + ".dat", "w") as f:
            f.write(self.plain_text())
        return filename + ".dat"


INTERMEDIATE = 3  # the decayed photons, kept when the deck asks for them


def file_format(deck):
    """output:file_format (src/main.rs:908-925): plain_text (also plain-text, ascii) or
    fits; anything else falls back to plain_text."""
    s = deck.string("output:file_format", "plain_text")
    return "fits" if s == "fits" else "plain_text"


def unit_system(deck):
    s = deck.string("output:units", "auto")
    if s not in UNIT_SYSTEMS:
        raise DeckError(f'unit system requested, "{s}", is not one of "auto", "hep", or "si"')
    return UNIT_SYSTEMS[s]


def distributions(deck):
    """{species: [Distribution]} from output:electron / photon / positron."""
    out = {}
    for name, sp in (("electron", Species.ELECTRON), ("photon", Species.PHOTON), ("positron", Species.POSITRON),
                     ("intermediate", INTERMEDIATE)):
        specs = deck.raw_list(f"output:{name}")
        out[sp] = [Distribution(s, deck.context) for s in specs]
    return out


def write_distributions(engine, deck, particles, prefix_dir="", ident="", a0=None, n_a0=1, intermediates=None):
    """Writes every requested distribution of the resident final state (and of the
    intermediate photons); returns the file names. `particles` and `intermediates`
    are the fetched particles (for the species counts); file names follow
    src/main.rs:1218-1248."""
    system, fmt = unit_system(deck), file_format(deck)
    cur = ident if n_a0 <= 1 else f"{ident}{'_' if ident else ''}a0_{a0:.3f}"
    written = []
    names = {Species.ELECTRON: "electron", Species.PHOTON: "photon", Species.POSITRON: "positron",
             INTERMEDIATE: "intermediate"}
    for sp, dists in distributions(deck).items():
        if sp == INTERMEDIATE:
            n = 0 if intermediates is None else len(intermediates["species"])
        else:
            n = int(np.sum(particles["species"] == int(sp)))
        prefix = f"{prefix_dir}{'/' if prefix_dir else ''}{cur}{'_' if cur else ''}{names[sp]}"
        for d in dists:
            h = d.histogram(engine, int(sp), n, system)
            if h is not None:
                written.append(h.write(prefix + d.suffix(), fmt))
    return written


# ---------------------------------------------------------------- summary statistics
_OPS = {"total": "total", "fraction": "fraction", "frac": "fraction", "mean": "mean", "variance": "variance",
        "var": "variance", "minimum": "minimum", "min": "minimum", "maximum": "maximum", "max": "maximum",
        "circmean": "circmean", "cmean": "circmean", "circvariance": "circvar", "cvariance": "circvar",
        "circvar": "circvar", "cvar": "circvar", "circstd": "circstd", "cstd": "circstd"}
_DEFAULT = UNIT_SYSTEMS["auto"]


class Statistic:
    """One `stats:` specifier (SummaryStatistic::load, src/output/stats.rs:85-263):
    `op var[`weight]`, `op var[`weight] in (min; max)` or
    `op var[`weight] for var2 in (min; max)`, bounds in SI and strict."""

    def __init__(self, spec, context=None):
        context = context or {}
        bad = DeckError(f"{spec!r} is not a summary statistic")
        start, end = spec.find("("), spec.rfind(")")
        rng = spec[start:end + 1] if start >= 0 and end >= 0 else None
        words = (spec.replace(rng, "") if rng else spec).split()
        fname = None
        if len(words) == 5:
            if words[3] not in _Q or words[2] != "for":
                raise DeckError(f"{spec!r}: the requested filter is not recognised")
            fname = words[3]
        lo, hi = -math.inf, math.inf
        if len(words) in (3, 5):
            if rng is None or words[-1] != "in":
                raise bad
            rs = ["".join(r.split()) for r in rng[1:-1].split(";")]
            if len(rs) != 2:
                raise bad
            try:
                lo = -math.inf if rs[0] == "auto" else float(evaluate(rs[0], context))
                hi = math.inf if rs[1] == "auto" else float(evaluate(rs[1], context))
            except DeckError:
                raise bad
        if len(words) not in (2, 3, 5):
            raise bad
        if words[0] not in _OPS:
            raise DeckError(f"{spec!r}: the requested operation is not recognised")
        self.op = _OPS[words[0]]
        vs = words[1].split("`")
        if len(vs) > 2:
            raise bad
        var, wname = vs[0], (vs[1] if len(vs) == 2 else "unit")
        if var not in _Q:
            raise DeckError(f"{spec!r}: the requested variable is not recognised")
        kind = _Q[var][1]
        if self.op.startswith("circ") and kind != "angle":
            raise DeckError(f"{spec!r}: circular stats can only be applied to angles")
        u = _unit(kind, _DEFAULT)[1]
        if self.op in ("fraction", "circvar"):
            u = "1"
        elif self.op == "variance":
            u = "1" if u == "1" else f"({u})^2"
        if wname not in _Q:
            raise bad
        self.variable, self.var_unit, self.var_quantity = var, u, _Q[var][0]
        self.weight, self.weight_quantity = wname, _Q[wname][0]
        # a finite bound without `for` cuts on the variable itself (keeping its unit)
        if fname is None and (math.isfinite(lo) or math.isfinite(hi)):
            self.filter, self.filter_unit, fkind = var, u, kind
        elif fname is not None:
            self.filter, self.filter_unit, fkind = fname, _unit(_Q[fname][1], _DEFAULT)[1], _Q[fname][1]
        else:
            self.filter, fkind = None, kind
        scale = _unit(fkind, UNIT_SYSTEMS["si"])[0]
        self.min, self.max = lo / scale, hi / scale
        self.value = math.nan
        self.name = ""

    def evaluate(self, engine, species, name):
        """The statistic over the resident final state (stats.rs:412-481): one device
        pass for the sums, a second for the variance, shifted by the mean."""
        cut = (_Q[self.filter][0], self.min, self.max) if self.filter else None
        out = engine.stats(int(species), self.var_quantity, self.weight_quantity, cut)
        f = np.float64
        with np.errstate(all="ignore"):
            if self.op == "total":
                v = out[1]
            elif self.op == "fraction":
                v = f(out[1]) / f(out[0])
            elif self.op == "mean":
                v = f(out[1]) / f(out[2])
            elif self.op == "minimum":
                v = out[3]
            elif self.op == "maximum":
                v = out[4]
            elif self.op == "variance":
                mean = f(out[1]) / f(out[2])
                o2 = engine.stats(int(species), self.var_quantity, self.weight_quantity, cut,
                                  shift=float(mean) if np.isfinite(mean) else 0.0)
                v = (f(o2[7]) - f(o2[8]) ** 2 / f(o2[2])) / f(o2[2])
            else:
                mx, my = f(out[5]) / f(out[2]), f(out[6]) / f(out[2])
                r = np.hypot(mx, my)
                v = {"circmean": np.arctan2(my, mx), "circvar": 1.0 - r, "circstd": np.sqrt(-2.0 * np.log(r))}[self.op]
        self.value, self.name = float(v), name
        return self.value

    def __str__(self):
        bounds = ""
        if self.filter:
            if math.isfinite(self.min) and math.isinf(self.max):
                bounds = f" ({self.filter} [{self.filter_unit}] > {rust_exp(self.min, 6)})"
            elif math.isinf(self.min) and math.isfinite(self.max):
                bounds = f" ({self.filter} [{self.filter_unit}] < {rust_exp(self.max, 6)})"
            else:
                bounds = f" ({rust_exp(self.min, 6)} < {self.filter} [{self.filter_unit}] < {rust_exp(self.max, 6)})"
        wstr = f" ({self.weight}-weighted)" if self.weight != "unit" else ""
        return (f"{self.name}: {self.op} {self.variable}{wstr}{bounds} = {rust_exp(self.value, 6)} "
                f"[{self.var_unit}]")


class StatsExpression:
    """`name[`formula] expression [unit]` (StatsExpression, stats.rs:519-575)."""

    def __init__(self, spec, context=None):
        v = spec.split()
        if len(v) < 2:
            raise DeckError(f"{spec!r} is not a stats expression")
        self.name = v[0].split("`")[0]
        self.formula = v[1] if "`" in v[0] else None
        try:
            self.value = float(evaluate(v[1], context or {}))
        except DeckError:
            raise DeckError(f"{spec!r}: cannot evaluate {v[1]!r}")
        self.unit = v[2] if len(v) > 2 else "1"

    def __str__(self):
        f = f" ({self.formula})" if self.formula is not None else ""
        return f"expr: {self.name}{f} = {rust_exp(self.value, 6)} [{self.unit}]"


def statistics(deck):
    """[(species name, Statistic)] in the order the reference writes them
    (electron, positron, photon; src/main.rs:1274-1285), and the expressions."""
    # stats:intermediate is parsed but, as in the reference, not written
    [Statistic(s, deck.context) for s in deck.raw_list("stats:intermediate")]
    stats = []
    for name in ("electron", "positron", "photon"):
        stats += [(name, Statistic(s, deck.context)) for s in deck.raw_list(f"stats:{name}")]
    exprs = [StatsExpression(s, deck.context) for s in deck.raw_list("stats:expression")]
    return stats, exprs


def write_stats(engine, deck, prefix_dir="", ident="", a0=None, n_a0=1):
    """Evaluates the deck's statistics over the resident final state and writes
    `[ident_]stats.txt`; returns the file name, or None when none are asked for."""
    stats, exprs = statistics(deck)
    if not stats and not exprs:
        return None
    cur = ident if n_a0 <= 1 else f"{ident}{'_' if ident else ''}a0_{a0:.3f}"
    species = {"electron": Species.ELECTRON, "positron": Species.POSITRON, "photon": Species.PHOTON}
    lines = []
    for name, st in stats:
        st.evaluate(engine, species[name], name)
        lines.append(str(st))
    lines += [str(e) for e in exprs]
    fname = f"{prefix_dir}{'/' if prefix_dir else ''}{cur}{'_' if cur else ''}stats.txt"
    with open(fname, "w") as f:
        f.write("\n".join(lines) + "\n")
    return fname
