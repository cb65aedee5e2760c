"""Readers for LDData generating-vector and generating-matrix files
(QMCSoftware/LDData), the inputs the reference loads at _problems.py:8,17:

    np.loadtxt(".../dnet/<name>.txt", dtype=np.uint64, skiprows=7)      -> C
    np.loadtxt(".../lattice/<name>.txt", dtype=np.uint64, skiprows=6)   -> g

Only local files are read (there is no network fetch).  Each reader returns
the array in the layout the generators take -- g as (1, d) for lat_gen_*, C as
(1, d, mmax) MSB-first columns for dnb2_gen_* -- together with the comment
lines of the header.  Integers are parsed exactly, as Python integers (all
64 bits), and rejected if they do not fit uint64.

read_lattice(path, d=None, skiprows=6) -> (g, header)
read_dnet(path, d=None, skiprows=7)    -> (C, header)
    d: keep the first d rows (the reference's g_full[:d], C_full[:d]); a
    larger d raises, as the reference's assert does.
    skiprows: header lines to skip, as the reference passes them; None skips
    the leading lines that start with '#'.
"""
import numpy as np


def _rows(path, skiprows):
    header, rows = [], []
    with open(path) as f:
        lines = f.read().splitlines()
    if skiprows is None:
        skiprows = 0
        while skiprows < len(lines) and lines[skiprows].lstrip().startswith("#"):
            skiprows += 1
    header = lines[:skiprows]
    for ln in lines[skiprows:]:
        ln = ln.split("#", 1)[0].strip()
        if not ln:
            continue
        vals = [int(tok, 10) for tok in ln.split()]
        for v in vals:
            if v < 0 or v >= 1 << 64:
                raise ValueError("%s: value %d does not fit uint64" % (path, v))
        rows.append(vals)
    return header, rows


def _first(rows, d, path):
    if d is not None:
        if d > len(rows):
            raise AssertionError("%s holds %d dimensions, %d requested" % (path, len(rows), d))
        rows = rows[:d]
    return rows


def read_lattice(path, d=None, skiprows=6):
    header, rows = _rows(path, skiprows)
    if any(len(r) != 1 for r in rows):
        raise ValueError("%s: a lattice file holds one integer per line" % path)
    rows = _first(rows, d, path)
    g = np.array([r[0] for r in rows], dtype=np.uint64)[None, :]
    return g, header


def read_dnet(path, d=None, skiprows=7):
    header, rows = _rows(path, skiprows)
    widths = {len(r) for r in rows}
    if len(widths) != 1:
        raise ValueError("%s: rows of unequal length %s" % (path, sorted(widths)))
    rows = _first(rows, d, path)
    C = np.array(rows, dtype=np.uint64)[None, :, :]
    return C, header
