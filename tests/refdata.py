"""Readers for the reference's own output files kept as fixtures in tests/golden."""
import os

import numpy as np


def read_phi_sites(path, L, H):
    """phi_sites.dat written by Setup_phi_table (J1J2_6tv.f90:617-631):
    '(3I8,1X,E20.12)' rows ix iy layer value, one header line."""
    rows = np.loadtxt(path, skiprows=1)
    p = np.zeros(L * L * H)
    for ix, iy, iz, v in rows:
        p[int(ix) + L * int(iy) + L * L * int(iz)] = v
    return p


def read_table(path):
    """Numeric rows after the 'disorder configurations processed' header of a
    cf*.dat file (Write_files, J1J2_6tv.f90:3221-3569)."""
    lines = open(path).read().splitlines()
    start = None
    for i, ln in enumerate(lines):
        if "disorder configurations processed" in ln:
            start = i
    assert start is not None, path
    rows = []
    for ln in lines[start + 1:]:
        parts = ln.split()
        try:
            rows.append([float(x) for x in parts])
        except ValueError:
            continue
    return np.array([r for r in rows if r])


def find(dirpath, prefix):
    for f in sorted(os.listdir(dirpath)):
        if f.startswith(prefix):
            return os.path.join(dirpath, f)
    raise FileNotFoundError(prefix)


def read_cfhs(path):
    """cfhs (Write_files, J1J2_6tv.f90:3442-3483): rows 'T BIN [<energy>] freq'
    for bins 1 .. BINS-1 of every temperature.  Returns a list, in file order,
    of (T, bin numbers, lower-edge labels, frequencies)."""
    rows = []
    for ln in open(path):
        t = ln.split()
        if len(t) != 4:
            continue
        try:
            rows.append([float(x) for x in t])
        except ValueError:
            continue
    a = np.array(rows)
    out = []
    for T in dict.fromkeys(a[:, 0].tolist()):
        r = a[a[:, 0] == T]
        out.append((T, r[:, 1].astype(int), r[:, 2], r[:, 3]))
    return out


def read_cfpt(path):
    """cfpt (Write_files, J1J2_6tv.f90:3535-3551): per pair (attempts,
    accepts, acceptance), then replicas_visiting_both_extremes and its
    fraction."""
    lines = open(path).read().splitlines()
    att, acc, rate = [], [], []
    visit = frac = None
    for ln in lines[1:]:
        t = ln.split()
        if t and t[0] == "replicas_visiting_both_extremes":
            visit, frac = int(t[1]), float(t[2])
        elif len(t) == 6:
            att.append(int(t[3]))
            acc.append(int(t[4]))
            rate.append(float(t[5]))
    return np.array(att), np.array(acc), np.array(rate), visit, frac
