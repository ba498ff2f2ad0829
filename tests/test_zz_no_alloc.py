"""j1j2_run_cycles_noalloc (no device allocation, no graph, no wait) against
j1j2_run_cycles on twin engines with the same inputs: spins, maps,
observables, every fetched result and the raw sums equal bit for bit.

The calls run in a row with shrinking inputs (a mixed batch of byte-layout,
global-memory and bit-layout lattices; then fewer, smaller lattices; then the
FP64 plain form), so each later call finds the previous call's contents in
the static scratch array the calls share on the GPU."""
import numpy as np
import pytest

import j1j2_engine as J
from gpu_cases import SEED, build_engine, say

pytestmark = pytest.mark.gpu


def snapshot(eng):
    out = []
    for c in range(eng.nconf):
        f = eng.fetch(c)
        out.append((eng.spins(c), eng.maps(c), eng.observables(c), f, eng.sums(c)))
    return out


def assert_same(a, b):
    for (s1, m1, o1, f1, u1), (s2, m2, o2, f2, u2) in zip(a, b):
        assert np.array_equal(s1, s2)
        assert np.array_equal(m1[0], m2[0]) and np.array_equal(m1[1], m2[1])
        assert np.array_equal(o1, o2)
        for k in f1:
            assert np.array_equal(np.asarray(f1[k]), np.asarray(f2[k])), k
        assert np.array_equal(u1[0], u2[0]) and u1[1] == u2[1]
        assert np.array_equal(u1[2], u2[2]) and np.array_equal(u1[3], u2[3])


# (label, groups as (L, H, nreplica, nconf), precision, starts): larger input first
CASES = [
    ("mixed: bytes L=16, global L=10, bits L=64xH16", [(16, 16, 4, 3), (10, 10, 2, 2), (64, 16, 2, 1)],
     J.FP32_BATCHED, ["hot", "str", "hot", "str", "hot", "str"]),
    ("smaller: bytes L=8, fused local nematic", [(8, 8, 2, 2)], J.FP32_BATCHED, "hot"),
    ("smaller: FP64 plain L=8", [(8, 8, 2, 1)], J.FP64_PLAIN, "str"),
]


def test_noalloc_matches_run_cycles():
    neq, nmess = 7, 19           # 19 = graphs of 16, 2 and 1 cycles on the other path
    for label, shape, prec, starts in CASES:
        say(f"case noalloc: {label}")
        groups, iconf = [], 1
        for L, H, R, C in shape:
            groups.append(J.make_params(L, H, nreplica=R, nconf=C, iconf0=iconf, precision=prec, seed=SEED,
                                        flags=J.FLAG_WRITELOCNEM))
            iconf += C
        ref, _ = build_engine(groups, starts)
        ref.run_cycles(J.PHASE_EQUIL, neq, begin=True)
        ref.run_cycles(J.PHASE_MEAS, nmess, begin=True)
        want = snapshot(ref)
        ref.close()
        eng, _ = build_engine(groups, starts)
        eng.run_cycles_noalloc(J.PHASE_EQUIL, neq, begin=True)
        eng.run_cycles_noalloc(J.PHASE_MEAS, nmess, begin=True)
        eng.synchronize()
        got = snapshot(eng)
        eng.close()
        assert_same(want, got)
        say(f"  bit-identical over {eng.nconf} configs")


def test_noalloc_rejects_too_many_lattices():
    """256 replicas x 65 configs of a 2x2x2 lattice: 16640 lattices, more than
    the static scratch holds; the call fails with J1J2_EINVAL and a message,
    and launches nothing."""
    say("case noalloc: too many lattices")
    g = J.make_params(2, 2, nreplica=256, nconf=65, seed=SEED)
    eng = J.Engine([g])
    with pytest.raises(J.EngineError) as err:
        eng.run_cycles_noalloc(J.PHASE_EQUIL, 1, begin=True)
    assert err.value.code == -1 and "at most 16384" in str(err.value)
    eng.close()
