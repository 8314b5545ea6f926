"""Pack Ptarmigan's numeric lookup tables into one binary blob.

The reference embeds its rate tables, inverse-CDF tables and quadrature rules as
Rust array literals (e.g. ``src/lcfa/photon_emission/tables.rs:9-111``,
``src/nonlinear_compton/lp/cdf_table.rs:1-4``). These numbers are data: the
device kernels and the oracle must use the same values, so this script reads
them as text from a checkout of the reference and writes

* ``ptg_tables.bin``  -- the packed float64 blob read by ``libptg`` and the oracle
* ``manifest.json``   -- name, source location, shape and SHA-256 of every entry

Blob layout (little endian)::

    char[8]  magic "PTGTBL01"
    u32      n_entries
    u32      reserved (0)
    n_entries x { char name[96]; u64 offset; u64 count; u32 ndim; u32 shape[4]; u32 pad }
    f64      data[...]            (offset/count are in units of f64)

Usage: ``python extract.py /path/to/reference [outdir]``.
"""
import hashlib
import json
import math
import os
import re
import struct
import sys

# (source file relative to src/, names to pull; None = every top-level array/CDF/Series)
SOURCES = [
    ("constants.rs", None),
    ("quadrature/tables.rs", ["GAUSS_16_NODES", "GAUSS_16_WEIGHTS", "GAUSS_32_NODES",
                              "GAUSS_32_WEIGHTS", "GL_NODES", "GL_WEIGHTS"]),
    ("special_functions/airy.rs", None),
    ("lcfa/photon_emission/tables.rs", None),
    ("lcfa/photon_emission/mod.rs", ["GAUNT_FACTOR_TABLE"]),
    ("lcfa/pair_creation/tables.rs", None),
    ("nonlinear_compton/lp/rate_table.rs", None),
    ("nonlinear_compton/lp/cdf_table.rs", None),
    ("nonlinear_compton/lp/classical/rate_table.rs", None),
    ("nonlinear_compton/lp/classical/cdf_table.rs", None),
    ("nonlinear_compton/cp/tables/total.rs", None),
    ("nonlinear_compton/cp/classical/rate_table.rs", None),
    ("nonlinear_compton/cp/classical/cdf_table.rs", None),
    ("pair_creation/lp/tables/total.rs", None),
    ("pair_creation/lp/tables/cdf.rs", None),
    ("pair_creation/lp/tables/mid_range/total.rs", None),
    ("pair_creation/cp/total.rs", None),
]

DECL = re.compile(r"^(?:pub )?(?:static|const) ([A-Za-z_0-9]+)\s*:\s*([^=]+?)\s*=\s*", re.M)
CONSTS = {"LN_10": math.log(10.0), "LN_2": math.log(2.0), "PI": math.pi,
          "NEG_INFINITY": -math.inf, "INFINITY": math.inf}
TOKEN = re.compile(r"NEG_INFINITY|INFINITY|[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?(?:_?f64|_?usize|_?i32)?"
                   r"(?:\s*/\s*[-+]?(?:\d+\.?\d*)(?:[eE][-+]?\d+)?(?:_?f64)?)?")


def _num(tok):
    tok = re.sub(r"_?(f64|usize|i32)", "", tok)
    if tok in CONSTS:
        return CONSTS[tok]
    if "/" in tok:
        a, b = tok.split("/")
        return float(a) / float(b)
    return float(tok)


def _shape(typ):
    """Shape of a Rust array type such as ``[[[f64; 2]; 16]; 4148]`` (outermost first)."""
    dims = [int(d) for d in re.findall(r";\s*(\d+)\s*\]", typ)]
    return list(reversed(dims))


def _initializer(text, start):
    """Return the initializer text starting at ``start`` up to the terminating ``;``."""
    depth = 0
    i = start
    while i < len(text):
        c = text[i]
        if c in "[{(":
            depth += 1
        elif c in "]})":
            depth -= 1
        elif c == ";" and depth == 0:
            return text[start:i]
        i += 1
    raise ValueError("unterminated initializer")


def extract_file(ref_root, rel, names):
    path = os.path.join(ref_root, "src", rel)
    with open(path) as f:
        text = f.read()
    out = []
    for m in DECL.finditer(text):
        name, typ = m.group(1), m.group(2).strip()
        if names is not None and name not in names:
            continue
        init = _initializer(text, m.end())
        line = text.count("\n", 0, m.start()) + 1
        if typ == "usize":
            continue
        if typ == "f64":
            expr = init.replace("std::f64::consts::", "").replace("consts::", "")
            toks = re.findall(r"\d+\.?\d*(?:[eE][-+]?\d+)?|[A-Z_][A-Z_0-9]*|[-+*/().]", expr)
            try:
                val = eval("".join(str(CONSTS.get(t, t)) for t in toks), {"__builtins__": {}})
            except Exception:
                continue  # depends on other declarations; restated by hand where used
            vals, shape = [float(val)], [1]
        else:
            body = re.sub(r"//[^\n]*", "", init)
            body = re.sub(r"\b(CDF|Series|coeff|power|table|a|n)\b\s*:?", " ", body)
            vals = [_num(t) for t in TOKEN.findall(body)]
            if typ.startswith("["):
                shape = _shape(typ)
                if typ.startswith("[CDF"):
                    shape = [shape[0], 64]
            elif typ == "CDF":
                shape = [64]
            elif typ.startswith("Series"):
                shape = [2, len(vals) // 2]
            else:
                continue
            want = 1
            for d in shape:
                want *= d
            if want != len(vals):
                raise ValueError(f"{rel}:{name}: expected {want} values for {typ}, got {len(vals)}")
        key = rel[:-3] + ":" + name
        out.append((key, f"src/{rel}:{line}", shape, vals))
    return out


# Tables the reference generates with code of its own but does not ship; made by
# the generator beside this file and kept under generated/ (name, file, shape, source).
GENERATED = [
    ("pair_creation/lp/tables/mid_range/cdf:TABLE", "nbw_lp_mid_cdf.f64", [5652, 16, 3],
     "regenerated by tables/gen_bw_tables.py (reference generator: src/pair_creation/lp/mod.rs:1065-1163)"),
    ("nonlinear_compton/cp/tables/cdf:TABLE", "nlc_cp_cdf.f64", [4270, 32, 2],
     "regenerated by tables/gen_cp_tables.py (reference generator: src/nonlinear_compton/cp/mod.rs:474-612)"),
    # not a reference table: the largest of the reference's per-emission ceiling samples
    # (src/nonlinear_compton/lp/mod.rs:53-140) over a grid, for the tabulated ceiling
    ("nonlinear_compton/lp:CEILING_MAX", "nlc_lp_ceiling.f64", [33 * 44 * 110 + 32 * 43 * 109],
     "generated by tables/gen_lp_ceiling.py with the library's sampling code (ceiling samples of "
     "src/nonlinear_compton/lp/mod.rs:53-140)"),
    # not a reference table: Chebyshev expansions of the reference's Gauss-Laguerre sums
    # for Ai and Ai' at x >= 1 (src/special_functions/airy.rs:41-116)
    ("special_functions/airy:CHEB_GL_SUMS", "airy_gl_cheb.f64", [10, 19],
     "generated by tables/gen_airy_cheb.py from the GL_m1_6_*/GL_p1_6_* nodes and weights of "
     "src/special_functions/airy.rs:285-421,451-593"),
]


def generated_entries(outdir):
    out = []
    for key, fname, shape, src in GENERATED:
        path = os.path.join(outdir, "generated", fname)
        if not os.path.exists(path):
            print(f"note: {path} not present; {key} left out", file=sys.stderr)
            continue
        with open(path, "rb") as f:
            raw = f.read()
        n = len(raw) // 8
        want = 1
        for d in shape:
            want *= d
        if n != want:
            raise ValueError(f"{path}: {n} values, expected {want}")
        out.append((key, src, shape, list(struct.unpack(f"<{n}d", raw))))
    return out


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    outdir = sys.argv[2] if len(sys.argv) > 2 else os.path.dirname(os.path.abspath(__file__))
    entries = []
    for rel, names in SOURCES:
        got = extract_file(ref_root, rel, names)
        if names is not None:
            missing = set(names) - {k.split(":")[1] for k, *_ in got}
            if missing:
                raise ValueError(f"{rel}: missing {sorted(missing)}")
        entries.extend(got)
    entries.extend(generated_entries(outdir))

    header = bytearray(b"PTGTBL01")
    header += struct.pack("<II", len(entries), 0)
    data = bytearray()
    manifest = {}
    offset = 0
    for key, src, shape, vals in entries:
        blob = struct.pack(f"<{len(vals)}d", *vals)
        shp = (shape + [0, 0, 0, 0])[:4]
        header += struct.pack("<96sQQI4II", key.encode(), offset, len(vals), len(shape), *shp, 0)
        manifest[key] = {"source": src, "shape": shape, "sha256": hashlib.sha256(blob).hexdigest()}
        data += blob
        offset += len(vals)
    with open(os.path.join(outdir, "ptg_tables.bin"), "wb") as f:
        f.write(header + data)
    with open(os.path.join(outdir, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(f"wrote {len(entries)} tables, {offset} doubles")


if __name__ == "__main__":
    main()
