"""GPU parity: the engine (through its C-ABI) against the oracle on the same
seeded inputs -- spins, replica maps, swap counters, moment sums, histograms,
derived fields and local-nematic sums bit for bit."""
import numpy as np
import pytest

import j1j2_engine as J
import oracle as O
from gpu_cases import SEED, build_engine, compare_state, oracle_for, say

pytestmark = pytest.mark.gpu

LOCNEM = J.FLAG_WRITELOCNEM


def run_and_compare(groups, starts, neq, nmess, ora_configs=None, shadow=False, flips=False):
    eng, inputs = build_engine(groups, starts)
    say(f"engine path {J.PATH_NAMES[eng.sweep_path()[0]]}, {eng.nconf} configs")
    eng.run(J.PHASE_EQUIL, neq)
    eng.run(J.PHASE_MEAS, nmess)
    eng.synchronize()
    checked = 0
    for c, (g, iconf, p, s) in enumerate(inputs):
        if ora_configs is not None and c not in ora_configs:
            continue
        ora = oracle_for(g, iconf, p, s, shadow=shadow, count_flips=flips)
        ora.run(neq, False)
        ora.run(nmess, True)
        compare_state(eng, c, ora, J.ladder(g), measured=True)
        if shadow:
            assert eng.mismatch(c) == ora.shadow
        if flips:
            assert np.array_equal(eng.flips(c), ora.flips)
        checked += 1
    return eng, checked


@pytest.mark.parametrize("L,H,R,start", [(8, 8, 4, "hot"), (16, 8, 3, "str"), (8, 2, 2, "hot")])
def test_fp32_resident_matches_oracle(L, H, R, start):
    say(f"case fp32 resident L={L} H={H} R={R} {start}")
    g = J.make_params(L, H, nreplica=R, nconf=2, iconf0=3, precision=J.FP32_BATCHED, flags=LOCNEM,
                      seed=SEED, ptmeasindex=min(2, R))
    eng, n = run_and_compare([g], start, 3, 6)
    assert eng.sweep_path()[0] == 1 and n == 2


@pytest.mark.parametrize("nt", [256, 512, 576, 1024])
def test_resident_block_sizes_match_oracle(nt, monkeypatch):
    """Every block size of the byte layout on every size it runs: the
    fixed-geometry stepping with even and odd layer steps (some threads idle
    when a step's layers do not fill the block, e.g. L = 48 with 1024
    threads) and the general path (L = 48 with 256 threads), one launch per
    size and one mixed launch."""
    say(f"case resident block size {nt}, L in 8..48")
    monkeypatch.setenv("J1J2_RES_THREADS", str(nt))
    # H large enough that odd steps (7 layers of L = 24 at 256 threads, 9 of
    # L = 32 at 576, 5 of L = 40, 3 and 7 of L = 48) reach their second parity
    sizes = [(8, 4, 2), (16, 4, 2), (24, 16, 2), (32, 20, 2), (40, 12, 2), (48, 16, 2)]
    groups = [J.make_params(L, H, nreplica=R, nconf=1, iconf0=1 + k, flags=LOCNEM, seed=SEED, ptmeasindex=2)
              for k, (L, H, R) in enumerate(sizes)]
    eng, n = run_and_compare(groups, ["hot", "str"] * 3, 2, 3)            # one mixed launch
    assert eng.sweep_path()[0] == 1 and n == len(sizes)
    assert f"nt{nt}" in eng.launch_plan()
    eng.close()
    for k, g in enumerate(groups):                                          # one size per launch
        eng, n = run_and_compare([g], "str" if k % 2 else "hot", 2, 3)
        assert n == 1 and f"L{g.L} nt{nt}" in eng.launch_plan()
        eng.close()
    # cubic lattices (H = L) take kernels specialised on the height: one
    # launch per size, and one mixed launch of all sizes
    cubic = [J.make_params(L, L, nreplica=2, nconf=1, iconf0=7 + k, flags=LOCNEM, seed=SEED, ptmeasindex=2)
             for k, L in enumerate((8, 16, 24, 32, 40, 48))]
    for k, g in enumerate(cubic):
        eng, n = run_and_compare([g], "hot" if k % 2 else "str", 2, 3)
        assert n == 1 and f"L{g.L} nt{nt}" in eng.launch_plan()
        eng.close()
    eng, n = run_and_compare(cubic, ["str", "hot"] * 3, 2, 3)
    assert n == len(cubic) and f"L0 nt{nt}" in eng.launch_plan()
    eng.close()


@pytest.mark.parametrize("flags", [LOCNEM, LOCNEM | J.FLAG_WRITENUMFLIPS])
def test_bundles_with_partial_groups_match_oracle(flags, monkeypatch):
    """The 1024-thread launch over several sizes sweeps small lattices in
    bundles (8 per block at L = 16, 2 at L = 24 and 32).  Replica counts that
    are not multiples of the bundle size leave groups with no lattice of
    their own (they repeat the first lattice's sweeps and write nothing):
    every lattice, including the last of each config, and the local-nematic
    sums and flip counts shared or kept per block, against the oracle."""
    say(f"case bundles with partial groups, flags {flags}")
    monkeypatch.setenv("J1J2_RES_THREADS", "1024")
    # (L, replicas, configs); L = 40 (no bundle) keeps flip counts resident
    sizes = [(16, 11, 2), (24, 3, 1), (32, 5, 1), (40, 2, 1)]
    groups = []
    for k, (L, R, C) in enumerate(sizes):
        groups.append(J.make_params(L, L, nreplica=R, nconf=C, iconf0=1 + 4 * k, flags=flags, seed=SEED,
                                    ptmeasindex=2))
    eng, n = run_and_compare(groups, ["hot", "str"] * 3, 2, 4, flips=bool(flags & J.FLAG_WRITENUMFLIPS))
    assert eng.sweep_path()[0] == 1 and n == sum(C for _, _, C in sizes)
    assert "nt1024" in eng.launch_plan() and "bundled" in eng.launch_plan(), eng.launch_plan()


@pytest.mark.parametrize("L,H,tmin,tmax", [(16, 8, 0.02, 0.05), (16, 8, 2.0e4, 1.0e5), (64, 4, 0.02, 0.05),
                                          (64, 4, 2.0e4, 1.0e5)])
def test_fp32_extreme_temperatures_match_oracle(L, H, tmin, tmax):
    """Far outside the benchmark's temperatures the hardware-exp2 filter of
    the decision meets its edge cases: 2^y below the normal range (every
    uphill move rejected unless r = 0) and y near 0 (thresholds within 2^-16
    of 2^32, left open and decided by the exact rule); byte and bit layouts."""
    say(f"case fp32 extreme T L={L} H={H} T={tmin}..{tmax}")
    g = J.make_params(L, H, nreplica=3, nconf=2, iconf0=11, precision=J.FP32_BATCHED, flags=LOCNEM,
                      seed=SEED, ptmeasindex=2, tmin=tmin, tmax=tmax)
    eng, n = run_and_compare([g], "hot", 3, 4)
    assert eng.sweep_path()[0] in (1, 3) and n == 2


@pytest.mark.parametrize("L,H,R,start", [(56, 4, 3, "hot"), (64, 2, 2, "str"), (72, 4, 2, "hot"),
                                          (80, 2, 3, "hot"), (88, 4, 2, "str"), (96, 4, 3, "hot")])
def test_fp32_resident_bits_matches_oracle(L, H, R, start):
    """L > 48: the one-bit-per-site LDS layout (kernels_bits.hip)."""
    say(f"case fp32 resident bits L={L} H={H} R={R} {start}")
    g = J.make_params(L, H, nreplica=R, nconf=2, iconf0=5, precision=J.FP32_BATCHED, flags=LOCNEM,
                      seed=SEED, ptmeasindex=min(2, R))
    eng, n = run_and_compare([g], start, 3, 4)
    assert eng.sweep_path()[0] == 3 and n == 2


@pytest.mark.parametrize("L,H,R,slabs,start", [(56, 8, 3, 4, "hot"), (64, 4, 2, 2, "str"), (72, 8, 2, 2, "hot"),
                                                (96, 8, 2, 4, "str")])
def test_fp32_bits_slabs_match_oracle(L, H, R, slabs, start, monkeypatch):
    """Bit-layout lattices split into z-slabs (one block each, halo layers
    exchanged through versioned words): bit-identical to the oracle."""
    say(f"case fp32 bits slabs L={L} H={H} R={R} S={slabs} {start}")
    monkeypatch.setenv("J1J2_BIT_SLABS", str(slabs))
    g = J.make_params(L, H, nreplica=R, nconf=2, iconf0=7, precision=J.FP32_BATCHED, flags=LOCNEM,
                      seed=SEED, ptmeasindex=min(2, R))
    eng, n = run_and_compare([g], start, 3, 4)
    assert eng.sweep_path()[0] == 3 and n == 2
    assert f"s{slabs}" in eng.launch_plan()


def test_bits_slabs_shadow_matches_oracle(monkeypatch):
    """Slabs with the FP64 shadow decisions counted, one cycle per swap."""
    say("case bits slabs + shadow")
    monkeypatch.setenv("J1J2_BIT_SLABS", "2")
    g = J.make_params(64, 8, nreplica=2, nconf=2, iconf0=2, ptswapever=1,
                      flags=J.FLAG_SHADOW_FP64 | LOCNEM, seed=SEED)
    eng, n = run_and_compare([g], "hot", 2, 3, shadow=True)
    assert n == 2 and "s2" in eng.launch_plan()


def test_bits_shadow_and_mixed_with_bytes():
    """Bit and byte layouts in one engine (separate launches), with the FP64
    shadow decisions counted on both, against the oracle."""
    say("case bits + bytes + shadow")
    groups = [J.make_params(16, 4, nreplica=2, nconf=1, iconf0=1, flags=J.FLAG_SHADOW_FP64 | LOCNEM, seed=SEED),
              J.make_params(64, 4, nreplica=3, nconf=1, iconf0=2, flags=J.FLAG_SHADOW_FP64 | LOCNEM, seed=SEED)]
    eng, n = run_and_compare(groups, "hot", 2, 3, shadow=True)
    assert eng.sweep_path()[0] == 3 and n == 2


@pytest.mark.parametrize("L,H,R", [(6, 4, 3), (2, 4, 2), (10, 6, 4)])
def test_fp32_global_matches_oracle(L, H, R):
    say(f"case fp32 global L={L} H={H} R={R}")
    g = J.make_params(L, H, nreplica=R, nconf=2, precision=J.FP32_BATCHED, flags=LOCNEM, seed=SEED)
    eng, n = run_and_compare([g], "hot", 3, 5)
    assert eng.sweep_path()[0] == 2


@pytest.mark.parametrize("L,H,R", [(8, 8, 4), (6, 4, 3)])
def test_fp64_plain_matches_oracle(L, H, R):
    say(f"case fp64 plain L={L} H={H} R={R}")
    g = J.make_params(L, H, nreplica=R, nconf=2, precision=J.FP64_PLAIN, flags=LOCNEM, seed=SEED)
    eng, n = run_and_compare([g], "hot", 3, 5)
    assert eng.sweep_path()[0] == 0


def test_resident_and_global_paths_agree():
    say("case resident vs global, L=16")
    outs = []
    for flags in (0, J.FLAG_FORCE_GLOBAL):
        g = J.make_params(16, 16, nreplica=4, nconf=3, flags=flags, seed=SEED)
        eng, _ = build_engine([g], "hot")
        eng.run(J.PHASE_EQUIL, 4)
        outs.append([eng.spins(c) for c in range(3)] + [eng.observables(c) for c in range(3)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_graph_and_eager_schedules_agree():
    """Replayed graphs (16, 8, 4, 2 and 1 cycles: 31 = all five, 21 = 16 + 4
    + 1 continuing the phase) give the eager schedule's results."""
    say("case graph vs eager")
    outs = []
    for flags in (0, J.FLAG_NO_GRAPH):
        g = J.make_params(8, 8, nreplica=4, nconf=2, flags=flags | LOCNEM, seed=SEED)
        eng, _ = build_engine([g], "hot")
        eng.run(J.PHASE_EQUIL, 3)
        eng.run(J.PHASE_MEAS, 31)
        eng.run_cycles(J.PHASE_MEAS, 21)
        outs.append([eng.spins(c) for c in range(2)] + [eng.fetch(c)["basefields"] for c in range(2)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_uneven_batch_matches_oracle():
    say("case uneven batch")
    groups = [J.make_params(L, L, nreplica=R, nconf=2, iconf0=1 + 2 * k, flags=LOCNEM, seed=SEED)
              for k, (L, R) in enumerate([(8, 2), (16, 3), (24, 4)])]
    eng, n = run_and_compare(groups, ["hot", "str"] * 3, 2, 3)
    assert n == 6 and eng.sweep_path()[0] == 1


def test_uneven_batch_with_global_group_matches_oracle():
    say("case uneven batch with a non-resident group")
    groups = [J.make_params(L, L, nreplica=2, nconf=1, iconf0=1 + k, flags=LOCNEM, seed=SEED)
              for k, L in enumerate([8, 6])]
    eng, n = run_and_compare(groups, "hot", 2, 3)
    assert eng.sweep_path()[0] == 2


@pytest.mark.parametrize("precision", [J.FP32_BATCHED, J.FP64_PLAIN])
def test_local_sampler_matches_oracle(precision):
    say(f"case LOCAL precision {precision}")
    g = J.make_params(8, 8, nreplica=1, ptswapever=1, nconf=1, sampler=J.SAMPLER_LOCAL,
                      precision=precision, ladder=J.LADDER_LOCAL, tmin=10.0, tmax=10.0, dt=1.0,
                      neq=20, nmess=50, flags=LOCNEM, seed=SEED)
    run_and_compare([g], "hot", 20, 50)


def test_shadow_mismatch_counts_match_oracle():
    say("case shadow FP64 mismatch counting")
    for flags in (J.FLAG_SHADOW_FP64, J.FLAG_SHADOW_FP64 | J.FLAG_FORCE_GLOBAL):
        g = J.make_params(8, 8, nreplica=3, nconf=2, flags=flags, seed=SEED)
        run_and_compare([g], "hot", 2, 2, shadow=True)


@pytest.mark.parametrize("flags,path", [(0, 1), (J.FLAG_FORCE_GLOBAL, 2), (J.FLAG_SHADOW_FP64, 1)])
def test_flip_counts_match_oracle(flags, path):
    """WRITENUMFLIPS: per-slot, per-site flip counts (confnumflips) on the
    LDS-resident sweep (a byte per site beside the lattice, added to the
    slot's counters at the end of each launch), on the global-memory sweep,
    and with the FP64 shadow."""
    say(f"case WRITENUMFLIPS flags {flags}")
    g = J.make_params(8, 4, nreplica=3, nconf=2, flags=J.FLAG_WRITENUMFLIPS | flags, seed=SEED)
    eng, _ = run_and_compare([g], "hot", 2, 3, flips=True, shadow=bool(flags & J.FLAG_SHADOW_FP64))
    assert eng.sweep_path()[0] == path


@pytest.mark.parametrize("nt,swapever,neq,nmess", [(256, 5, 8, 12), (256, 29, 4, 8), (256, 127, 2, 3),
                                                     (256, 128, 1, 2), (512, 5, 8, 12)])
def test_flip_counts_l32_match_oracle(monkeypatch, nt, swapever, neq, nmess):
    """Flip counts of cubic L = 32 lattices: in registers through the colour
    passes (256-thread blocks, PTSWAPEVER <= 127), kept in a pending byte per
    site and added to the slot's counters when a byte reaches 128 (several
    times over the run at PTSWAPEVER = 29 and 127), or in the LDS beside the
    lattice (PTSWAPEVER = 128, 512 threads), with the local-nematic sums."""
    say(f"case WRITENUMFLIPS L=32 nt {nt} PTSWAPEVER {swapever}")
    monkeypatch.setenv("J1J2_RES_THREADS", str(nt))
    g = J.make_params(32, 32, nreplica=3, nconf=2, iconf0=3, seed=SEED, ptmeasindex=2, ptswapever=swapever,
                      flags=J.FLAG_WRITENUMFLIPS | LOCNEM)
    eng, n = run_and_compare([g], ["hot", "str"], neq, nmess, flips=True)
    assert n == 2 and eng.sweep_path()[0] == 1 and f"L32 nt{nt}" in eng.launch_plan(), eng.launch_plan()


@pytest.mark.parametrize("nt,swapever,path", [(256, 5, 1), (576, 31, 1), (1024, 5, 1), (1024, 32, 2)])
def test_flip_counts_l48_match_oracle(monkeypatch, nt, swapever, path):
    """Flip counts of cubic L = 48 lattices (lattice and a byte of counts
    beside it would not fit one CU's LDS): in bits 3..7 of the sites' own
    bytes on the LDS-resident sweep for PTSWAPEVER <= 31, cleared at the end
    of each launch; the global-memory sweep above that."""
    say(f"case WRITENUMFLIPS L=48 nt {nt} PTSWAPEVER {swapever}")
    monkeypatch.setenv("J1J2_RES_THREADS", str(nt))
    g = J.make_params(48, 48, nreplica=2, nconf=2, iconf0=4, seed=SEED, ptmeasindex=2, ptswapever=swapever,
                      flags=J.FLAG_WRITENUMFLIPS | LOCNEM)
    eng, n = run_and_compare([g], ["hot", "str"], 3, 4, flips=True)
    assert n == 2 and eng.sweep_path()[0] == path, eng.launch_plan()


def test_flip_counts_uneven_batch_match_oracle():
    """Flip counts over an uneven batch: L = 16, 24 and 40 with a byte of
    counts beside the lattice in LDS, cubic L = 48 with the counts in the
    sites' own bytes, per-size launches (one launch over all sizes would need
    the counts' LDS for L = 48 too), 20 cycles (counts accumulate over
    launches)."""
    say("case WRITENUMFLIPS uneven batch")
    groups = [J.make_params(L, L, nreplica=R, nconf=1, iconf0=1 + k, seed=SEED, flags=J.FLAG_WRITENUMFLIPS)
              for k, (L, R) in enumerate([(16, 3), (24, 2), (40, 2), (48, 2)])]
    eng, _ = run_and_compare(groups, ["hot", "str", "hot", "str"], 8, 12, flips=True)
    assert eng.sweep_path()[0] == 1 and "L0" not in eng.launch_plan(), eng.launch_plan()


@pytest.mark.parametrize("precision,flags,R,ptswapever,ptmeasindex,bins", [
    (J.FP32_BATCHED, 0, 1, 3, 1, 7),                  # one temperature: no swap pairs
    (J.FP32_BATCHED, 0, 5, 3, 5, 13),                 # PTMEASINDEX at the last slot, odd PTSWAPEVER
    (J.FP32_BATCHED, J.FLAG_FORCE_GLOBAL, 2, 2, 2, 1),  # one histogram bin, global-memory sweep
    (J.FP64_PLAIN, 0, 1, 2, 1, 100),
    (J.FP64_PLAIN, 0, 4, 1, 4, 3)])
def test_pt_edge_parameters_match_oracle(precision, flags, R, ptswapever, ptmeasindex, bins):
    """Run parameters at the edges of their ranges (J1J2_6tv.f90:841-925 accepts
    NREPLICA = 1, PTSWAPEVER = 1, PTMEASINDEX = NREPLICA, BINS = 1)."""
    say(f"case PT edges precision {precision} flags {flags} R={R} swapever {ptswapever} "
        f"measindex {ptmeasindex} bins {bins}")
    g = J.make_params(8, 4, nreplica=R, nconf=2, iconf0=5, precision=precision, ptswapever=ptswapever,
                      ptmeasindex=ptmeasindex, bins=bins, flags=flags | LOCNEM | J.FLAG_WRITENUMFLIPS,
                      seed=SEED)
    eng, n = run_and_compare([g], ["hot", "str"], 2, 4, flips=True)
    assert n == 2


def test_zero_cycles_report_initial_state(golden_dir):
    """NEQ = NMESS = 0: the engine reports the HOT starts' energies exactly as
    the reference's cfen table does (fixture from the reference build)."""
    import os
    from refdata import find, read_table
    say("case NMESS=0 vs reference cfen")
    g = J.make_params(8, 8, nreplica=4, nconf=3, iconf0=1, seed=SEED)
    eng, _ = build_engine([g], "hot")
    en = np.stack([eng.fetch(c)["basefields"][9] for c in range(3)])
    tab = read_table(find(os.path.join(golden_dir, "ref_hot_init_l8"), "cfen"))
    assert np.allclose(tab[:, 1], en.mean(0), rtol=0, atol=1e-9)


@pytest.mark.parametrize("precision", [J.FP32_BATCHED, J.FP64_PLAIN])
def test_frozen_local_run_matches_reference(golden_dir, precision):
    """LOCAL, FER start, phi = 0, T = 0.01, NEQ = 10, NMESS = 200: no move is
    ever taken (dE = 4 at every site), so the engine's energy histogram and
    energy equal the reference's own run of these parameters (fixture
    ref_local_frozen) exactly: every sample in bin 50 of the degenerate
    range e -+ 1e-6."""
    import os
    from refdata import find, read_table
    say(f"case frozen LOCAL run vs reference, precision {precision}")
    L = 8
    g = J.make_params(L, L, nreplica=1, ptswapever=1, nconf=1, sampler=J.SAMPLER_LOCAL, precision=precision,
                      ladder=J.LADDER_LOCAL, tmin=0.01, tmax=0.01, dt=1.0, neq=10, nmess=200, seed=SEED)
    eng = J.Engine([g])
    eng.upload_disorder(0, np.zeros(L * L * L))
    eng.set_spins(0, J.initial_spins(L, L, J.START_FER, 0, 0))
    eng.run(J.PHASE_EQUIL, 10)
    eng.run(J.PHASE_MEAS, 200)
    out = eng.fetch(0)
    ref_dir = os.path.join(golden_dir, "ref_local_frozen")
    hs = read_table(find(ref_dir, "cfhs"))
    assert np.array_equal(out["hist"][:99, 0], hs[:, 3])
    assert out["hist_min"][0] == -1.0 - 1.0e-6
    # bin energies as Write_files prints them (e12.6: half an ulp is 5e-6 at |e| = 1)
    assert np.allclose(out["hist_min"][0] + out["hist_bin"][0] * (hs[:, 1] - 1), hs[:, 2], rtol=0, atol=5e-6)
    en = read_table(find(ref_dir, "cfen"))
    assert out["basefields"][9, 0] == en[0, 1] == -1.0


def test_pt_ladder_continues_across_calls():
    """run_cycles without begin continues the swap alternation and the sums."""
    say("case chunked runs")
    g = J.make_params(8, 8, nreplica=4, nconf=1, seed=SEED)
    eng, inputs = build_engine([g], "hot")
    eng.run(J.PHASE_EQUIL, 2)
    eng.run_cycles(J.PHASE_MEAS, 2, begin=True)
    eng.run_cycles(J.PHASE_MEAS, 3, begin=False)
    gg, iconf, p, s = inputs[0]
    ora = oracle_for(gg, iconf, p, s)
    ora.run(2, False)
    ora.run(5, True)
    compare_state(eng, 0, ora, J.ladder(gg))


def test_slab_wait_failure_is_reported(monkeypatch):
    """The error path of the slab exchange, forced from the host: with
    J1J2_SLAB_SPIN_MAX < 0 every bounded wait gives up at once and sets the
    error flag, every wave still finishes, and the run returns J1J2_EDEVICE
    with its message instead of hanging; a later engine runs normally."""
    say("case slab wait failure forced from the host")
    g = J.make_params(64, 8, nreplica=2, nconf=2, iconf0=3, seed=SEED)
    monkeypatch.setenv("J1J2_BIT_SLABS", "2")
    monkeypatch.setenv("J1J2_SLAB_SPIN_MAX", "-1")
    eng, _ = build_engine([g], "hot")
    assert "s2" in eng.launch_plan()
    with pytest.raises(J.EngineError) as e:
        eng.run(J.PHASE_EQUIL, 1)
    assert e.value.code == -3 and "slab" in str(e.value)
    eng.close()
    monkeypatch.delenv("J1J2_SLAB_SPIN_MAX")
    run_and_compare([g], "hot", 1, 1)


@pytest.mark.parametrize("L,H,flags,precision,slabs", [
    (16, 16, 0, J.FP32_BATCHED, None),                  # byte layout, cubic (height-specialised)
    (24, 8, 0, J.FP32_BATCHED, None),                   # byte layout, general height
    (64, 8, 0, J.FP32_BATCHED, "1"),                    # bit layout, whole lattices
    (64, 8, 0, J.FP32_BATCHED, "2"),                    # bit layout, two z-slabs
    (10, 6, 0, J.FP32_BATCHED, None),                   # global-memory sweep (L % 8 != 0)
    (8, 8, 0, J.FP64_PLAIN, None)])
def test_anisotropic_couplings_match_oracle(L, H, flags, precision, slabs, monkeypatch):
    """J1h != J1v (deltaJ1), J2 and J3 away from +-1: a coupling taken for
    another (x for y, J2 for J3, a sign) on any sweep path breaks bit
    equality with the oracle."""
    say(f"case anisotropic couplings L={L} H={H} precision {precision} slabs {slabs}")
    if slabs is not None:
        monkeypatch.setenv("J1J2_BIT_SLABS", slabs)
    g = J.make_params(L, H, nreplica=3, nconf=2, iconf0=9, precision=precision, flags=LOCNEM | flags,
                      J1=1.1, deltaJ1=0.3, J2=-0.7, J3=1.3, seed=SEED, ptmeasindex=2)
    eng, n = run_and_compare([g], "hot", 2, 3)
    assert n == 2


@pytest.mark.parametrize("precision", [J.FP32_BATCHED, J.FP64_PLAIN])
def test_split_ladder_matches_oracle(precision):
    """PTLADDERMODE = SPLIT (J1J2_6tv.f90:928-962): unevenly spaced
    temperatures, so a swap that used the wrong pair's betas breaks bit
    equality with the oracle."""
    say(f"case SPLIT ladder precision {precision}")
    g = J.make_params(8, 8, nreplica=5, nconf=2, iconf0=4, precision=precision, ladder=J.LADDER_SPLIT,
                      tladdermid=9.0, tladderfrac=0.6, flags=LOCNEM, seed=SEED, ptmeasindex=3)
    d = np.diff(J.ladder(g))
    assert np.all(d > 0) and not np.allclose(d, d[0])
    eng, n = run_and_compare([g], "str", 2, 3)
    assert n == 2
