"""Binary catalogue files of the reference CLI (cc/catalog_reader.h:16-71):
a 64-byte header of eight float64 -- posmin[3], posmax[3], then two unused
values -- followed by records of four float64 (x, y, z, weight)."""

import numpy as np

HEADER_DOUBLES = 8
TOOBIG = 1e10


def read_header(path):
    """(posmin, posmax) from the header; raises ValueError on a short or
    implausible header (the reference asserts |value| < 1e10)."""
    h = np.fromfile(path, dtype="<f8", count=HEADER_DOUBLES)
    if h.size != HEADER_DOUBLES:
        raise ValueError("%s: short catalogue header" % path)
    if not np.all(np.abs(h[:7]) < TOOBIG):
        raise ValueError("%s: implausible catalogue header" % path)
    return h[0:3].copy(), h[3:6].copy()


def read_catalog(path):
    """(posmin, posmax, records [n, 4] float64: x, y, z, weight)."""
    posmin, posmax = read_header(path)
    data = np.fromfile(path, dtype="<f8", offset=8 * HEADER_DOUBLES)
    n = data.size // 4
    return posmin, posmax, np.ascontiguousarray(data[: 4 * n].reshape(n, 4))


def write_catalog(path, records, posmin, posmax, max_sep=0.0):
    """Writes records [n, 4] (x, y, z, weight) with the 64-byte header."""
    rec = np.ascontiguousarray(records, dtype="<f8")
    if rec.ndim != 2 or rec.shape[1] != 4:
        raise ValueError("records must be (n, 4)")
    head = np.zeros(HEADER_DOUBLES, "<f8")
    head[0:3] = posmin
    head[3:6] = posmax
    head[6] = max_sep
    with open(path, "wb") as f:
        head.tofile(f)
        rec.tofile(f)
