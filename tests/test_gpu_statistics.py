"""Statistical parity on the GPU.

1. Against the unmodified reference (its own runs, kept as fixtures in
   tests/golden/): the same disorder config and start, the same run lengths;
   64 independent engine chains (different Philox keys) give the mean and the
   single-run spread; z = (reference - mean) / sqrt(s^2 + s^2/64), the
   reference's single run being assigned the engine's single-run spread s.
2. FP32 batched form against the FP64 form at L = 32, 16 replicas x 64 disorder
   configs: paired per-config differences, z = mean / (std / sqrt(64)).
3. Both forms against exact enumeration of the 2x2x4 lattice (2^16 states):
   64 independent LOCAL chains at one temperature, z = (mean of the chains'
   time averages - exact) / (std / sqrt(64)).
4. Against the reference's own production runs at L = 20 and 40 (phi = 0,
   across the stripe transition): 100 chains per temperature on both sides,
   z = (reference - mean) / sqrt(err_ref^2 + err^2) with the standard errors
   over the chains.

Every z-score must satisfy |z| < 4 and at least 75% (resp. 85%) of them |z| < 2.
"""
import json
import os

import numpy as np
import pytest

import j1j2_engine as J
import oracle as O
from gpu_cases import config_inputs, say
from refdata import find, read_table

pytestmark = pytest.mark.gpu

NCHAIN = 64
# z-scores of every comparison are also written as JSON when J1J2_STAT_DIR names a folder
OUT = os.environ.get("J1J2_STAT_DIR")

# (name, BASEFIELDS index, reference file prefix, column)
QUANTITIES = [("e", 9, "cfen", 1), ("c", 11, "cfen", 3), ("|m|", 0, "cfma", 1), ("chi", 3, "cfma", 4),
              ("|str|", 13, "cfst", 1), ("str2", 16, "cfst", 2), ("nem", 22, "cfst", 4),
              ("|strx|", 14, "cfst", 11), ("|stry|", 15, "cfst", 16), ("nbar", 37, "cfnm", 1),
              ("nbar2", 39, "cfnm", 7), ("UBN", 41, "cfnm", 10)]


def zsummary(zs, label):
    zs = np.asarray(zs, np.float64)
    frac2 = float(np.mean(np.abs(zs) < 2.0))
    say(f"{label}: {zs.size} z-scores, max |z| = {np.max(np.abs(zs)):.2f}, within 2 sigma: {100 * frac2:.0f}%")
    if OUT:
        os.makedirs(OUT, exist_ok=True)
        with open(os.path.join(OUT, f"stat_{label}.json"), "w") as f:
            json.dump({"z": zs.tolist(), "frac_within_2sigma": frac2}, f)
    return frac2


def chains_vs_reference(g, p, spins, variant, golden_dir):
    g.nconf = NCHAIN
    eng = J.Engine([g])
    for c in range(NCHAIN):
        eng.upload_disorder(c, p)
        eng.set_spins(c, spins)
    eng.run(J.PHASE_EQUIL, g.neq)
    eng.run(J.PHASE_MEAS, g.nmess)
    bf = np.stack([eng.fetch(c)["basefields"] for c in range(NCHAIN)])      # chain, field, slot
    outs = [eng.fetch(c) for c in range(NCHAIN)]
    eng.close()
    zs, rows = [], []
    gd = os.path.join(golden_dir, "ref_" + variant)
    for name, f, pref, col in QUANTITIES:
        ref = read_table(find(gd, pref))[:, col]
        mu, s = bf[:, f, :].mean(0), bf[:, f, :].std(0, ddof=1)
        z = (ref - mu) / np.sqrt(s ** 2 * (1 + 1.0 / NCHAIN) + 1e-300)
        zs.extend(z.tolist())
        rows.append((name, ref, mu, z))
    return zs, rows, outs


def test_pt_matches_reference_l8(golden_dir):
    """Reference run ref_pt_l8_stat: L = 8, NREPLICA = 4, T = 8.0-11.5, STR-X start,
    NEQ = 2000, NMESS = 40000 cycles, PTSWAPEVER = 5, disorder seed IRINIT = 1."""
    say("case PT L=8 vs reference (64 chains)")
    L, R = 8, 4
    g = J.make_params(L, L, nreplica=R, neq=2000, nmess=40000, seed=777)
    p = J.make_disorder(L, L, 3.0, 1, True, 2.0)
    spins = np.tile(J.initial_spins(L, L, J.START_STR, 0, 0), R)
    zs, rows, outs = chains_vs_reference(g, p, spins, "pt_l8_stat", golden_dir)
    # swap acceptance per pair (cfpt: attempts, accepts)
    pt = open(find(os.path.join(golden_dir, "ref_pt_l8_stat"), "cfpt")).read().splitlines()[1:R]
    ref_acc = np.array([float(line.split()[5]) for line in pt])
    acc = np.stack([o["swap_acc"] / np.maximum(o["swap_att"], 1) for o in outs])
    att = np.stack([o["swap_att"] for o in outs])
    assert np.all(att == 21000)                       # (NEQ + NMESS) / 2 per pair, as the reference
    mu, s = acc.mean(0), acc.std(0, ddof=1)
    zs.extend(((ref_acc - mu) / np.sqrt(s ** 2 * (1 + 1.0 / NCHAIN))).tolist())
    for name, ref, mu_, z in rows:
        say(f"  {name:7s} ref {np.array2string(ref, precision=5)} engine {np.array2string(mu_, precision=5)} "
            f"z {np.array2string(z, precision=2)}")
    frac = zsummary(zs, "pt_l8_vs_reference")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.75


def test_local_matches_reference_l8(golden_dir):
    """Reference run ref_local_l8 (the reference's CPU-runnable case): LOCAL, L = 8, T = 10, HOT
    start, NEQ = NMESS = 1000 sweeps.  The start is the reference's: KISS seeded
    with IRINIT = 1, 3*L^3 draws consumed by Setup_vacancies (J1J2_6tv.f90:
    1116-1175), then Initialize_spins."""
    say("case LOCAL L=8 vs reference (64 chains)")
    L = 8
    N = L * L * L
    spins = J.initial_spins(L, L, J.START_HOT, 0, 1, skip=3 * N)
    assert np.array_equal(spins, np.where(O.kiss_stream(1, 4 * N)[3 * N:] < 0.5, 1, -1))
    g = J.make_params(L, L, nreplica=1, ptswapever=1, sampler=J.SAMPLER_LOCAL, ladder=J.LADDER_LOCAL,
                      tmin=10.0, tmax=10.0, dt=1.0, neq=1000, nmess=1000, seed=4242)
    p = J.make_disorder(L, L, 3.0, 1, True, 2.0)
    zs, rows, _ = chains_vs_reference(g, p, spins, "local_l8", golden_dir)
    for name, ref, mu_, z in rows:
        say(f"  {name:7s} ref {np.array2string(ref, precision=5)} engine {np.array2string(mu_, precision=5)} "
            f"z {np.array2string(z, precision=2)}")
    frac = zsummary(zs, "local_l8_vs_reference")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.75


def test_fp32_matches_fp64_l32():
    """FP32 batched form vs FP64 plain form at L = 32, 16 replicas x 64 configs."""
    say("case FP32 vs FP64 at L=32")
    L, R, C = 32, 16, 64
    res = {}
    for prec in (J.FP32_BATCHED, J.FP64_PLAIN):
        g = J.make_params(L, L, nreplica=R, nconf=C, precision=prec, neq=300, nmess=1500, seed=99)
        eng = J.Engine([g])
        for c in range(C):
            eng.upload_disorder(c, J.make_disorder(L, L, 3.0, 1 + c, True, 2.0))
            eng.set_spins(c, np.tile(J.initial_spins(L, L, J.START_STR, 0, 0), R))
        eng.run(J.PHASE_EQUIL, g.neq)
        eng.run(J.PHASE_MEAS, g.nmess)
        outs = [eng.fetch(c) for c in range(C)]
        res[prec] = (np.stack([o["basefields"] for o in outs]),
                     np.stack([o["swap_acc"] / o["swap_att"] for o in outs]))
        eng.close()
    zs = []
    for name, f, _, _ in QUANTITIES:
        d = res[J.FP32_BATCHED][0][:, f, :] - res[J.FP64_PLAIN][0][:, f, :]
        z = d.mean(0) / (d.std(0, ddof=1) / np.sqrt(C) + 1e-300)
        zs.extend(z.tolist())
        say(f"  {name:7s} max |z| over slots {np.max(np.abs(z)):.2f}")
    d = res[J.FP32_BATCHED][1] - res[J.FP64_PLAIN][1]
    zs.extend((d.mean(0) / (d.std(0, ddof=1) / np.sqrt(C))).tolist())
    frac = zsummary(zs, "fp32_vs_fp64_l32")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.85


@pytest.mark.parametrize("precision", [J.FP32_BATCHED, J.FP64_PLAIN])
def test_engine_matches_exact_enumeration_2x2x4(precision):
    """Equilibrium averages of the engine's chains on the 2x2x4 lattice against
    the exact Boltzmann averages (oracle enumeration of all 2^16 states)."""
    L, H, T = 2, 4, 4.0
    say(f"case exact enumeration 2x2x4, precision {precision}")
    rng = np.random.default_rng(11)
    p = rng.uniform(-1.0, 1.0, L * L * H)
    exact = O.OracleLattice(L, H, p).enumerate(1.0 / T)
    neq, nmess = 500, 4000
    g = J.make_params(L, H, nreplica=1, ptswapever=1, sampler=J.SAMPLER_LOCAL, ladder=J.LADDER_LOCAL,
                      tmin=T, tmax=T, dt=1.0, nconf=NCHAIN, neq=neq, nmess=nmess, precision=precision,
                      seed=2718)
    eng = J.Engine([g])
    for c in range(NCHAIN):
        eng.upload_disorder(c, p)
        eng.set_spins(c, J.initial_spins(L, H, J.START_HOT, 0, 1 + c))
    eng.run(J.PHASE_EQUIL, neq)
    eng.run(J.PHASE_MEAS, nmess)
    means = []
    for c in range(NCHAIN):
        m, e, sx, sy, nb = eng.series(c, 0)[0]
        means.append([e.mean(), (e * e).mean(), np.abs(m).mean(), (m * m).mean(), nb.mean(),
                      (sx * sx).mean(), (sy * sy).mean()])
    eng.close()
    means = np.array(means)
    ref = exact[[0, 1, 2, 3, 4, 5, 6]]
    z = (means.mean(0) - ref) / (means.std(0, ddof=1) / np.sqrt(NCHAIN) + 1e-300)
    for name, mu, r, zz in zip(["e", "e2", "|m|", "m2", "nbar", "sx2", "sy2"], means.mean(0), ref, z):
        say(f"  {name:5s} engine {mu:.6f} exact {r:.6f} z {zz:.2f}")
    zsummary(z, f"enumeration_2x2x4_{'fp32' if precision == J.FP32_BATCHED else 'fp64'}")
    assert np.all(np.abs(z) < 4.0)


@pytest.mark.parametrize("L,nsweep", [(20, 100000), (40, 50000)])
def test_local_matches_reference_production_clean(golden_dir, L, nsweep):
    """The reference's own production runs (data/raw/corr_random_field/p0/c3/
    J2-1_P0.00_RAN_STRX_s100_scanclean_no_field/L20 and L40, copied to
    tests/golden/ref_prod_clean with their run_meta.txt): LOCAL sampler, phi = 0,
    STR-X start, T = 4.00..4.40 in steps of 0.04 with a fresh start at each
    temperature (INDEP), 100 chains (configs), NEQ = NMESS = 100000 sweeps; cfst
    gives per temperature <|str|>, the stripe susceptibility and the stripe
    Binder cumulant, each with its standard error over the 100 chains.  The
    engine runs the same: one chain per temperature and config, 100 configs of
    the zero field (L = 20 on the global-memory sweep, L = 40 LDS-resident),
    nsweep + nsweep sweeps; z = (reference - mean) / sqrt(err_ref^2 + err^2)
    across the stripe transition (T_c ~ 4.16-4.2)."""
    say(f"case LOCAL L={L} phi=0 vs the reference's production run ({nsweep}+{nsweep} sweeps)")
    NT, C = 11, 100
    g = J.make_params(L, L, nreplica=NT, ptswapever=1, nconf=C, sampler=J.SAMPLER_LOCAL, ladder=J.LADDER_LOCAL,
                      tmin=4.0, tmax=4.4, dt=0.04, neq=nsweep, nmess=nsweep, seed=31337)
    eng = J.Engine([g])
    start = np.tile(J.initial_spins(L, L, J.START_STR, 0, 0), NT)
    for c in range(C):
        eng.upload_disorder(c, np.zeros(L * L * L))
        eng.set_spins(c, start)
    assert eng.sweep_path()[0] == (2 if L % 8 else 1)
    eng.run(J.PHASE_EQUIL, nsweep)
    eng.run(J.PHASE_MEAS, nsweep)
    bf = np.stack([eng.fetch(c)["basefields"] for c in range(C)])          # config, field, temperature
    eng.close()
    tab = read_table(find(os.path.join(golden_dir, "ref_prod_clean"), f"cfst{L:04d}{L:04d}"))
    assert np.allclose(tab[:, 0], J.ladder(g), atol=5e-6)
    zs = []
    # (name, BASEFIELDS index, cfst column, cfst column of its standard error)
    for name, f, col, ecol in [("|str|", 13, 1, 9), ("chi_str", 18, 6, 10), ("U_str", 20, 7, 8)]:
        ref, err_ref = tab[:, col], tab[:, ecol]
        mu, err = bf[:, f, :].mean(0), bf[:, f, :].std(0) / np.sqrt(C)
        z = (ref - mu) / np.sqrt(err_ref ** 2 + err ** 2)
        zs.extend(z.tolist())
        say(f"  {name:8s} ref {np.array2string(ref, precision=4)}\n           engine "
            f"{np.array2string(mu, precision=4)}\n           z {np.array2string(z, precision=2)}")
    frac = zsummary(zs, f"local_l{L}_clean_vs_reference_production")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.75


# (name, BASEFIELDS index, table, column, column of its standard error over configs)
PROD_QUANTITIES = [("e", 9, "cfen", 1, 2), ("c", 11, "cfen", 3, 4), ("chi", 3, "cfma", 4, 5),
                   ("U_m", 5, "cfma", 6, 7), ("|str|", 13, "cfst", 1, 9), ("chi_str", 18, "cfst", 6, 10),
                   ("U_str", 20, "cfst", 7, 8), ("nbar", 37, "cfnm", 1, 2), ("U_n", 41, "cfnm", 6, 7)]


LONG = pytest.mark.skipif(not os.environ.get("J1J2_LONG_TESTS"),
                          reason="long run (~90 s on one MI355X); set J1J2_LONG_TESTS=1")


@pytest.mark.parametrize("L,neq,nmess,start", [(20, 2000, 4000, "str"), (24, 2000, 4000, "str"),
                                               (20, 2000, 4000, "hot"), (24, 2000, 4000, "hot"),
                                               pytest.param(40, 100000, 100000, "str", marks=LONG)])
def test_pt_matches_reference_production(golden_dir, L, neq, nmess, start):
    """The reference's own production PT runs (data/raw/corr_random_field/p2/c3/
    J2-1_P0.00_RAN_STRX_s100_scanpt_prod_main/L20, L24, L40, copied to
    tests/golden/ref_prod_pt with their run_meta.txt): PT SERIAL, phi = 2,
    xi = 3, 20 replicas over T = 8.0..11.5, STR-X start, PTSWAPEVER = 5,
    NEQ = NMESS = 1e5 cycles, disorder averages over 1000 configs (L = 40: 100)
    with their standard errors.  The engine runs 64 other disorder configs
    (iconf 1001..1064: independent draws of the same distribution, seeds the
    reference's generator would give them); z = (reference - mean) /
    sqrt(err_ref^2 + err^2) for energy, specific heat, magnetic and stripe
    susceptibilities and Binder cumulants, <|str|>, nbar and the nematic
    Binder cumulant at every temperature.  (<|m|> is left out: the reference
    reports |time average of m|, which shrinks with the run length.)  L = 20
    and 24 equilibrate within 2000 + 4000 cycles; at L = 40 the swap
    acceptance is ~0.2% at the cold end and the stripe order still grows after
    3e4 cycles, so it runs the reference's own 1e5 + 1e5 (a long test).
    The HOT-start runs (.../J2-1_P0.00_RAN_HOT_s100_scanpt_prod_main, fixtures
    in tests/golden/ref_prod_pt_hot) are compared the same way, the engine
    starting each replica from the reference's own HOT start for its seed."""
    say(f"case PT L={L} phi=2 {start} vs the reference's production run (64 configs, {neq}+{nmess} cycles)")
    R, C = 20, 64
    g = J.make_params(L, L, nreplica=R, nconf=C, iconf0=1001, neq=neq, nmess=nmess, seed=2718)
    eng = J.Engine([g])
    for c in range(C):
        p, s = config_inputs(L, L, R, 1001 + c, start)
        eng.upload_disorder(c, p)
        eng.set_spins(c, s.reshape(-1))
    assert eng.sweep_path()[0] == (2 if L % 8 else 1)
    eng.run(J.PHASE_EQUIL, g.neq)
    eng.run(J.PHASE_MEAS, g.nmess)
    bf = np.stack([eng.fetch(c)["basefields"] for c in range(C)])          # config, field, slot
    eng.close()
    gd = os.path.join(golden_dir, "ref_prod_pt" if start == "str" else "ref_prod_pt_hot")
    zs = []
    for name, f, pref, col, ecol in PROD_QUANTITIES:
        tab = read_table(find(gd, f"{pref}{L:04d}{L:04d}"))
        assert np.allclose(tab[:, 0], J.ladder(g), atol=5e-6)
        ref, err_ref = tab[:, col], tab[:, ecol]
        mu, err = bf[:, f, :].mean(0), bf[:, f, :].std(0) / np.sqrt(C)
        z = (ref - mu) / np.sqrt(err_ref ** 2 + err ** 2)
        zs.extend(z.tolist())
        say(f"  {name:8s} max |z| {np.max(np.abs(z)):.2f}  T=8: ref {ref[0]:.5g} engine {mu[0]:.5g}")
    frac = zsummary(zs, f"pt_l{L}_{start}_vs_reference_production")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.75
