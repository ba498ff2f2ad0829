"""Parity at the benchmark configurations' full sizes: one config's whole
trajectory bit for bit against the oracle, and size-independent properties
(maps are permutations, observables equal the oracle's on the fetched spins,
both FP32 paths agree) for every config."""
import os

import numpy as np
import pytest

import j1j2_engine as J
import oracle as O
from gpu_cases import SEED, build_engine, compare_state, oracle_for, say
from refdata import find, read_phi_sites

pytestmark = pytest.mark.gpu


def check_properties(eng, inputs, sample):
    for c in sample:
        g, iconf, p, s0 = inputs[c]
        spins = eng.spins(c)
        rat, tor = eng.maps(c)
        R = g.nreplica
        assert sorted(rat) == list(range(R)) and all(tor[rat[k]] == k for k in range(R))
        lat = O.OracleLattice(g.L, g.H, p, g.J1, g.deltaJ1, g.J2, g.J3)
        obs = eng.observables(c)
        N = g.L * g.L * g.H
        for r in range(R):
            o = lat.observables(spins[r])
            assert obs[0, r] == o[0] / N and obs[2, r] == o[1] / N and obs[3, r] == o[2] / N
            assert obs[1, r] == lat.energy_total(o) / N
            assert obs[4, r] == 2 * (o[3] - o[4]) / N


def test_batched_l32_16x64():
    """Batched PT, L=32, 16 replicas x 64 configs (the benchmark workload)."""
    say("case L=32 16x64")
    g = J.make_params(32, 32, nreplica=16, nconf=64, flags=J.FLAG_WRITELOCNEM, seed=SEED)
    eng, inputs = build_engine([g], "str")
    assert eng.sweep_path()[0] == 1
    eng.run(J.PHASE_EQUIL, 2)
    eng.run(J.PHASE_MEAS, 2)
    gg, iconf, p, s = inputs[37]
    ora = oracle_for(gg, iconf, p, s)
    ora.run(2, False)
    ora.run(2, True)
    compare_state(eng, 37, ora, J.ladder(gg))
    check_properties(eng, inputs, [0, 5, 63])


@pytest.mark.parametrize("precision", [J.FP64_PLAIN, J.FP32_BATCHED])
def test_c2_full_shape(golden_dir, precision):
    """PT, L = 16, 16 replicas over T = 8.0-11.5, one config (iconf 1, IRINIT 1),
    STR-X start: the plain FP64 form (global-memory colour passes, FP64 field in
    the reference's order) and the FP32 form, three equilibration and four
    measuring cycles bit for bit against the oracle.  The disorder is the
    reference's own for this run (its phi_sites.dat)."""
    say(f"case c2 full shape L=16 16x1 precision {precision}")
    g = J.make_params(16, 16, nreplica=16, nconf=1, iconf0=1, precision=precision, seed=SEED,
                      flags=J.FLAG_WRITELOCNEM)
    eng, inputs = build_engine([g], "str")
    gg, iconf, p, s = inputs[0]
    ref_p = read_phi_sites(find(os.path.join(golden_dir, "ref_pt_l16_c2"), "phi_sites"), 16, 16)
    assert np.max(np.abs(p - ref_p)) < 2e-11
    assert eng.sweep_path()[0] == (0 if precision == J.FP64_PLAIN else 1)
    eng.run(J.PHASE_EQUIL, 3)
    eng.run(J.PHASE_MEAS, 4)
    ora = oracle_for(gg, iconf, p, s)
    ora.run(3, False)
    ora.run(4, True)
    compare_state(eng, 0, ora, J.ladder(gg))
    eng.close()


C4 = [(16, 8), (24, 12), (32, 16), (48, 24)]


def c4_engine(tune, monkeypatch, threads=None):
    """The uneven batch at full size: 16 configs of each L in {16, 24, 32, 48}
    with 8, 12, 16, 24 replicas, HOT and STR starts alternating."""
    monkeypatch.setenv("J1J2_AUTOTUNE", "1" if tune else "0")
    if threads:
        monkeypatch.setenv("J1J2_RES_THREADS", str(threads))
    else:
        monkeypatch.delenv("J1J2_RES_THREADS", raising=False)
    groups = [J.make_params(L, L, nreplica=R, nconf=16, iconf0=1 + 16 * k, seed=SEED, flags=J.FLAG_WRITELOCNEM)
              for k, (L, R) in enumerate(C4)]
    return build_engine(groups, ["hot", "str"] * 32)


def test_uneven_full_batch_mixed_launch_matches_oracle(monkeypatch):
    """The whole uneven batch (960 lattices) in one 1024-thread mixed launch,
    the plan the autotuner keeps for it: one config of every size bit for bit
    against the oracle, the properties for all 64."""
    say("case uneven full batch, one 1024-thread mixed launch")
    eng, inputs = c4_engine(False, monkeypatch, threads=1024)
    assert "nt1024 x960" in eng.launch_plan(), eng.launch_plan()
    eng.run(J.PHASE_EQUIL, 1)
    eng.run(J.PHASE_MEAS, 1)
    for c in (3, 18, 37, 60):
        gg, iconf, p, s = inputs[c]
        ora = oracle_for(gg, iconf, p, s)
        ora.run(1, False)
        ora.run(1, True)
        compare_state(eng, c, ora, J.ladder(gg))
    check_properties(eng, inputs, range(64))
    eng.close()


def test_uneven_full_batch_autotuned_is_bit_identical(monkeypatch):
    """The autotuned full uneven batch against the default plan, every config."""
    say("case uneven full batch, autotuned vs default plan")
    res = []
    for tune in (True, False):
        eng, _ = c4_engine(tune, monkeypatch)
        eng.run(J.PHASE_EQUIL, 60)
        eng.run(J.PHASE_MEAS, 3)
        plan = eng.launch_plan()
        say(f"  plan: {plan}")
        assert ("autotuned" in plan) == tune
        res.append([(eng.spins(c), eng.maps(c), eng.fetch(c)) for c in range(eng.nconf)])
        eng.close()
    for (s1, m1, f1), (s2, m2, f2) in zip(*res):
        assert np.array_equal(s1, s2) and np.array_equal(m1[0], m2[0]) and np.array_equal(m1[1], m2[1])
        for k in f1:
            assert np.array_equal(np.asarray(f1[k]), np.asarray(f2[k])), k


def test_uneven_l16_to_48():
    say("case uneven L in {16,24,32,48}")
    groups = [J.make_params(L, L, nreplica=R, nconf=2, iconf0=1 + 2 * k, seed=SEED)
              for k, (L, R) in enumerate([(16, 8), (24, 12), (32, 16), (48, 24)])]
    eng, inputs = build_engine(groups, ["hot", "str"] * 4)
    assert eng.sweep_path()[0] == 1
    eng.run(J.PHASE_EQUIL, 1)
    eng.run(J.PHASE_MEAS, 1)
    for c in (1, 6):
        gg, iconf, p, s = inputs[c]
        ora = oracle_for(gg, iconf, p, s)
        ora.run(1, False)
        ora.run(1, True)
        compare_state(eng, c, ora, J.ladder(gg))
    check_properties(eng, inputs, range(8))


@pytest.mark.parametrize("flags,path,slabs,measure", [(0, 3, 1, False), (J.FLAG_WRITELOCNEM, 3, 4, True),
                                                     (J.FLAG_FORCE_GLOBAL, 2, 1, False)])
def test_large_l96(flags, path, slabs, measure, monkeypatch):
    """L = 96, 24 replicas x 8 configs: the bit-layout resident path (whole
    lattices, and four z-slabs per lattice on 768 blocks -- the benchmark's
    plan, also through a measuring cycle with the fused local-nematic sums)
    and the global-memory path."""
    say(f"case L=96 24x8 path {path} slabs {slabs} measure {measure}")
    monkeypatch.setenv("J1J2_BIT_SLABS", str(slabs))
    g = J.make_params(96, 96, nreplica=24, nconf=8, seed=SEED, flags=flags)
    eng, inputs = build_engine([g], "str")
    assert eng.sweep_path()[0] == path
    if path == 3:
        assert (f"s{slabs}" in eng.launch_plan()) == (slabs > 1)
    eng.run(J.PHASE_EQUIL, 1)
    if measure:
        eng.run(J.PHASE_MEAS, 1)
    check_properties(eng, inputs, [0, 7])
    gg, iconf, p, s = inputs[3]
    ora = oracle_for(gg, iconf, p, s)
    ora.run(1, False)
    if measure:
        ora.run(1, True)
    compare_state(eng, 3, ora, J.ladder(gg), measured=measure)


@pytest.mark.parametrize("precision", [J.FP32_BATCHED, J.FP64_PLAIN])
def test_maximum_lattice(precision):
    """The largest lattice the engine takes (L*L*H = 2^22: L = 128, H = 256, on
    the global-memory sweep), one equilibration and one measurement cycle, bit
    for bit against the oracle."""
    say(f"case maximum lattice 128x128x256 precision {precision}")
    g = J.make_params(128, 256, nreplica=2, nconf=1, iconf0=1, precision=precision, seed=SEED)
    eng, inputs = build_engine([g], "hot")
    assert eng.sweep_path()[0] == (2 if precision == J.FP32_BATCHED else 0)
    eng.run(J.PHASE_EQUIL, 1)
    eng.run(J.PHASE_MEAS, 1)
    gg, iconf, p, s = inputs[0]
    ora = oracle_for(gg, iconf, p, s)
    ora.run(1, False)
    ora.run(1, True)
    compare_state(eng, 0, ora, J.ladder(gg))
