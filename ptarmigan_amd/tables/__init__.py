"""Reader for the packed table blob ``ptg_tables.bin`` (layout documented in ``extract.py``)."""
import os
import struct

import numpy as np

BLOB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ptg_tables.bin")


def load(path=BLOB):
    """Return ``{name: ndarray}`` for every table in the blob."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != b"PTGTBL01":
        raise ValueError(f"{path}: not a PTGTBL01 table blob")
    n, _ = struct.unpack_from("<II", raw, 8)
    rec = struct.Struct("<96sQQI4II")
    base = 16 + n * rec.size
    out = {}
    for i in range(n):
        name, off, cnt, ndim, s0, s1, s2, s3, _ = rec.unpack_from(raw, 16 + i * rec.size)
        shape = (s0, s1, s2, s3)[:ndim]
        arr = np.frombuffer(raw, dtype="<f8", count=cnt, offset=base + 8 * off)
        out[name.rstrip(b"\0").decode()] = arr.reshape(shape)
    return out
