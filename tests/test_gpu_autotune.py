"""The autotuner times its candidate launch plans on real cycles of the run
(threads per block 256 / 512 / 576 / 1024, per-size launches in sequence or on
parallel graph branches, one launch over all sizes).  A plan must not change
any result: an autotuned run is compared bit for bit with an untuned one and,
for one config, with the oracle."""
import os

import numpy as np
import pytest

import j1j2_engine as J
from gpu_cases import SEED, build_engine, compare_state, oracle_for, say

pytestmark = pytest.mark.gpu

CYCLES = 60          # the tuner needs at most 54 of them


def run(groups, starts, tune, phase_cycles, equil_cycles=None):
    os.environ["J1J2_AUTOTUNE"] = "1" if tune else "0"
    try:
        eng, inputs = build_engine(groups, starts)
    finally:
        os.environ.pop("J1J2_AUTOTUNE", None)
    eng.run(J.PHASE_EQUIL, phase_cycles if equil_cycles is None else equil_cycles)
    eng.run(J.PHASE_MEAS, phase_cycles)
    plan = eng.launch_plan()
    state = [(eng.spins(c), eng.maps(c), eng.observables(c), eng.fetch(c)["basefields"]) for c in range(eng.nconf)]
    return eng, inputs, plan, state


def same(a, b):
    for (s1, m1, o1, f1), (s2, m2, o2, f2) in zip(a, b):
        assert np.array_equal(s1, s2)
        assert np.array_equal(m1[0], m2[0]) and np.array_equal(m1[1], m2[1])
        assert np.array_equal(o1, o2)
        assert np.array_equal(f1, f2)


def test_autotuned_uneven_batch_is_bit_identical():
    say("case autotune uneven L in {8,16,24,32,48}")
    groups = [J.make_params(L, L, nreplica=R, nconf=2, iconf0=1 + 2 * k, seed=SEED, flags=J.FLAG_WRITELOCNEM)
              for k, (L, R) in enumerate([(8, 4), (16, 8), (24, 12), (32, 16), (48, 6)])]
    starts = ["hot", "str"] * 5
    # 17 equilibration cycles: fewer than the candidates take (2 x 3 x 4 per-size
    # and 2 x 3 x 5 whole-batch plans), so the tuner waits for the second call and
    # never runs cycles beyond a call's count
    eng_t, inputs, plan_t, st_t = run(groups, starts, True, CYCLES, equil_cycles=17)
    say(f"  tuned plan: {plan_t}")
    assert "autotuned" in plan_t
    eng_t.close()
    eng_d, _, plan_d, st_d = run(groups, starts, False, CYCLES, equil_cycles=17)
    assert "default" in plan_d
    eng_d.close()
    same(st_t, st_d)


def test_autotuned_single_size_matches_oracle():
    say("case autotune L=24 vs oracle")
    g = J.make_params(24, 24, nreplica=4, nconf=3, seed=SEED)
    eng, inputs, plan, _ = run([g], "hot", True, 30)
    say(f"  tuned plan: {plan}")
    assert "autotuned" in plan
    gg, iconf, p, s = inputs[2]
    ora = oracle_for(gg, iconf, p, s)
    ora.run(30, False)
    ora.run(30, True)
    compare_state(eng, 2, ora, J.ladder(gg))
    eng.close()


def test_autotuned_bit_layout_is_bit_identical():
    """Bit-layout candidates: whole lattices, 2 or 4 z-slabs per lattice."""
    say("case autotune bits L=64 H=16")
    groups = [J.make_params(64, 16, nreplica=4, nconf=2, iconf0=1, seed=SEED, flags=J.FLAG_WRITELOCNEM)]
    eng_t, inputs, plan_t, st_t = run(groups, ["hot", "str"], True, CYCLES)
    say(f"  tuned plan: {plan_t}")
    assert "autotuned" in plan_t
    eng_t.close()
    eng_d, _, plan_d, st_d = run(groups, ["hot", "str"], False, CYCLES)
    assert "default" in plan_d and " s" not in plan_d
    eng_d.close()
    same(st_t, st_d)


@pytest.mark.parametrize("splits", [
    [[(32, 16, 1)], [(32, 8, 2)]],                                   # 16 lattices of L = 32 either way
    [[(16, 8, 2), (32, 16, 1)], [(16, 16, 1), (32, 8, 2)]],          # two sizes, 16 lattices each
])
def test_cached_plan_follows_each_engines_work_list(splits):
    """The autotuner's choice is cached per process.  Engines over the same
    lattices split differently into configs x replicas have work lists of
    different shapes; each engine must launch its own list whether it tuned
    or took a cached choice.  Every engine (tuned, then re-created with the
    cache filled, for each split in turn) equals an untuned run bit for bit,
    and its last config equals the oracle."""
    for rnd in range(2):
        for split in splits:
            groups, iconf = [], 1
            for L, R, C in split:
                groups.append(J.make_params(L, L, nreplica=R, nconf=C, iconf0=iconf, seed=SEED))
                iconf += C
            say(f"case cached plan: split {split}, round {rnd}")
            eng_t, inputs, plan_t, st_t = run(groups, "hot", True, CYCLES)
            say(f"  plan: {plan_t}")
            assert "autotuned" in plan_t
            gg, ic, p, s = inputs[-1]
            ora = oracle_for(gg, ic, p, s)
            ora.run(CYCLES, False)
            ora.run(CYCLES, True)
            compare_state(eng_t, eng_t.nconf - 1, ora, J.ladder(gg))
            eng_t.close()
            eng_d, _, plan_d, st_d = run(groups, "hot", False, CYCLES)
            eng_d.close()
            same(st_t, st_d)
