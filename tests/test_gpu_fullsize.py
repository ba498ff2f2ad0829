"""Parity at the benchmark configurations' full sizes: one config's whole
trajectory bit for bit against the oracle, and size-independent properties
(maps are permutations, observables equal the oracle's on the fetched spins,
both FP32 paths agree) for every config."""
import numpy as np
import pytest

import j1j2_engine as J
import oracle as O
from gpu_cases import SEED, build_engine, compare_state, oracle_for, say

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


@pytest.mark.parametrize("flags,path,slabs", [(0, 3, 1), (0, 3, 4), (J.FLAG_FORCE_GLOBAL, 2, 1)])
def test_large_l96(flags, path, slabs, monkeypatch):
    """L = 96, 24 replicas x 8 configs: the bit-layout resident path (whole
    lattices, and four z-slabs per lattice on 768 blocks) and the
    global-memory path."""
    say(f"case L=96 24x8 path {path} slabs {slabs}")
    monkeypatch.setenv("J1J2_BIT_SLABS", str(slabs))
    g = J.make_params(96, 96, nreplica=24, nconf=8, seed=SEED, flags=flags)
    eng, inputs = build_engine([g], "str")
    assert eng.sweep_path()[0] == path
    if path == 3:
        assert (f"s{slabs}" in eng.launch_plan()) == (slabs > 1)
    eng.run(J.PHASE_EQUIL, 1)
    check_properties(eng, inputs, [0, 7])
    gg, iconf, p, s = inputs[3]
    ora = oracle_for(gg, iconf, p, s)
    ora.run(1, False)
    compare_state(eng, 3, ora, J.ladder(gg), measured=False)


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
