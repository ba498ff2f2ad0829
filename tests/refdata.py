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
