"""The engine's C-ABI library on the CPU: it loads, exports every function the
header declares, its host-side helpers equal the oracle and the reference's
fixtures, and invalid run options fail with a status code (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import j1j2_engine as J
import oracle as O
from refdata import read_phi_sites

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "j1j2_engine.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(j1j2_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(J.LIB_PATH)
    names = header_functions()
    assert len(names) >= 25
    assert sorted(names) == sorted(J.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("L,H,seed", [(8, 8, 3), (6, 4, 8), (12, 10, 11)])
def test_make_disorder_matches_oracle_and_reference(golden_dir, L, H, seed):
    mine = J.make_disorder(L, H, 3.0, seed, True, 2.0)
    assert np.max(np.abs(mine - O.corr_field(L, L, H, 3.0, seed, True, 2.0))) < 1e-12
    variant = {8: "hot_init_l8", 6: "hot_init_l6h4", 12: "str_init_l12"}[L]
    ref = read_phi_sites(os.path.join(golden_dir, "ref_" + variant, "phi_sites.dat"), L, H)
    assert np.max(np.abs(mine - ref)) < 2e-11
    raw = J.make_disorder(L, H, 3.0, seed, False, 2.0)
    assert np.max(np.abs(raw - O.corr_field(L, L, H, 3.0, seed, False, 2.0))) < 1e-12


@pytest.mark.parametrize("start,axis", [(0, 0), (1, 0), (2, 0), (2, 1)])
def test_initial_spins_match_oracle(start, axis):
    for seed in (1, 17, 123457):
        a = J.initial_spins(6, 4, start, axis, seed)
        b = O.initial_spins(6, 4, start, axis, seed)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ladder,n", [(J.LADDER_LINEAR, 16), (J.LADDER_SPLIT, 11), (J.LADDER_LOCAL, 4)])
def test_ladder_matches_oracle(ladder, n):
    p = J.make_params(8, nreplica=n, ladder=ladder, tmin=8.0, tmax=14.0, dt=2.0, tladdermid=9.5,
                      tladderfrac=0.7)
    assert np.array_equal(J.ladder(p), O.ladder(ladder, n, 8.0, 14.0, 2.0, 9.5, 0.7))


@pytest.mark.parametrize("field,value,msg", [
    ("L", 7, "even"), ("H", 3, "even"), ("sampler", 5, "SAMPLERMODE"), ("ptswapever", 0, "PTSWAPEVER"),
    ("ptmeasindex", 9, "PTMEASINDEX"), ("tmin", 12.0, "TMAX"), ("nreplica", 0, "NREPLICA"),
    ("ladder", 7, "PTLADDERMODE")])
def test_invalid_options_fail_with_status(field, value, msg):
    p = J.make_params(8, nreplica=4)
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = J.lib().j1j2_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert rc == -1 and not h.value
    assert msg in J.last_error()


def test_split_ladder_validation():
    p = J.make_params(8, nreplica=2, ladder=J.LADDER_SPLIT)
    with pytest.raises(J.EngineError, match="SPLIT"):
        J.ladder(p)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_reports_device_error():
    h = ctypes.c_void_p()
    rc = J.lib().j1j2_create(ctypes.byref(J.make_params(8, nreplica=4)), 0, ctypes.byref(h))
    assert rc in (-1, -3) and not h.value
    assert J.last_error()


def test_disorder_batch_equals_per_config_fields():
    """j1j2_make_disorder_batch: config k's field is j1j2_make_disorder's for
    seed seed0 + k, whatever the thread count (host-only, no device)."""
    for L, H, seed0, n in ((8, 8, 1, 5), (6, 4, 7, 3), (12, 10, 11, 2)):
        ref = np.stack([J.make_disorder(L, H, 3.0, seed0 + k, True, 2.0) for k in range(n)])
        for nt in (1, 3, 0):
            got = J.make_disorder_batch(L, H, 3.0, seed0, n, True, 2.0, nthreads=nt)
            assert np.array_equal(got, ref)
