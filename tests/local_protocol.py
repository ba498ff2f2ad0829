"""The LOCAL sampler's temperature protocols as the reference runs them, for
the tests (the engine's host restates the same in Fortran):

  * Temperature_for_index (J1J2_6tv.f90:986-996): a HOT start walks the
    ladder TMIN + (k-1) DT downwards, other starts upwards;
  * CARRY: one start per config, KISS seeded with IRINIT + iconf - 1 after
    Setup_vacancies has drawn 3 L^3 numbers (occupancy and the two bond
    masks, J1J2_6tv.f90:300-322, 1116-1175), carried from temperature to
    temperature;
  * INDEP: Prepare_temperature_state (J1J2_6tv.f90:1097-1108) reseeds KISS
    with IRINIT + (iconf - 1) NTEMP + itemp before every temperature and
    starts afresh (no vacancy draws).
"""
import numpy as np

import oracle as O

START_CODE = {"HOT": 0, "FER": 1, "STR": 2}


def temperatures(tmin, tmax, dt, start):
    ntemp = 1 + int(round((tmax - tmin) / dt))
    ladder = tmin + dt * np.arange(ntemp)
    return ladder[::-1].copy() if start == "HOT" else ladder


def start_spins(L, H, start, protocol, irinit, iconf, ntemp, itemp, straxis=0):
    """Spins (x fastest) with which temperature itemp (1-based) of config iconf begins."""
    N = L * L * H
    if start != "HOT":
        return O.initial_spins(L, H, START_CODE[start], straxis, 0)
    if protocol == "INDEP":
        seed, skip = irinit + (iconf - 1) * ntemp + itemp, 0
    else:
        seed, skip = irinit + iconf - 1, 3 * N
    return np.where(O.kiss_stream(seed, skip + N)[skip:] < 0.5, 1, -1).astype(np.int8)


def local_nematic(L, H, s):
    """l_i = s_i (s_{i+x} - s_{i+y} + s_{i-x} - s_{i-y}) (J1J2_6tv.f90:1017-1021)."""
    a = s.reshape(H, L, L).astype(np.int64)            # z, y, x
    return (a * (np.roll(a, -1, 2) - np.roll(a, -1, 1) + np.roll(a, 1, 2) - np.roll(a, 1, 1))).reshape(-1)


def read_locnem(path):
    """Local_nematic_OP's file (J1J2_6tv.f90:2845-2884): header values and the
    per-site <s>, <l> columns (x fastest)."""
    head = {}
    for ln in open(path):
        if ln.startswith("#") and "=" in ln:
            k, v = ln[1:].split("=", 1)
            head[k.strip()] = [float(x) for x in v.split()]
    a = np.loadtxt(path, comments="#")
    return head, a[:, 5], a[:, 6]
