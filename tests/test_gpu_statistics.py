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
from refdata import find, read_cfhs, read_cfpt, read_table

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
    # every replica of the reference's run visited both ends of the ladder; so
    # does every replica of every engine chain of the same length
    visit = read_cfpt(find(os.path.join(golden_dir, "ref_pt_l8_stat"), "cfpt"))[3]
    assert [o["visitboth"] for o in outs] == [visit] * NCHAIN
    hz = histogram_vs_reference(outs, os.path.join(golden_dir, "ref_pt_l8_stat"))
    frac = zsummary(hz, "pt_l8_histogram_vs_reference")
    assert np.all(np.abs(hz) < 4.0) and frac >= 0.75


def histogram_vs_reference(outs, gd):
    """The engine chains' energy histograms against the reference's (cfhs).
    Each histogram has its own range (first measured energy +- (|e0| + 1)),
    so they are compared as distributions: the cumulative distribution at
    the reference's bin edges nearest its deciles, the engine's read off a
    monotone cubic through its own bin edges (interpolation error ~2e-4),
    z = (reference - chain mean) / (chain spread * sqrt(1 + 1/64)).  Also:
    the labels (lower bin edges) average to <e> - width/2, the rule of
    1 + int((e - min) / width) binning that the reference's file obeys (NINT
    binning would put them at <e> - width): (label mean + width/2 - <e>) /
    width within 0.2 for every chain and 0.05 over the chains."""
    from scipy.interpolate import PchipInterpolator
    ref = read_cfhs(find(gd, "cfhs"))
    zs = []
    for s, (T, bins, labels, freq) in enumerate(ref):
        w_r = (labels[-1] - labels[0]) / (len(labels) - 1)
        edges_r = labels[0] + w_r * np.arange(len(labels) + 1)
        F_r = np.concatenate([[0.0], np.cumsum(freq)])
        ks = sorted({int(np.argmin(np.abs(F_r - q))) for q in np.arange(1, 10) / 10.0})
        x = edges_r[ks]
        F, off = [], []
        for o in outs:
            h, lo, w = o["hist"][:, s], o["hist_min"][s], o["hist_bin"][s]
            edges = lo + w * np.arange(h.size + 1)
            F.append(PchipInterpolator(edges, np.concatenate([[0.0], np.cumsum(h)]), extrapolate=True)(x))
            off.append((float((h * edges[:-1]).sum()) + w / 2 - o["basefields"][9][s]) / w)
        off = np.array(off)
        assert np.all(np.abs(off) < 0.2) and abs(off.mean()) < 0.05, (s, off.mean(), off.min(), off.max())
        F = np.clip(np.array(F), 0.0, 1.0)
        mu, sd = F.mean(0), F.std(0, ddof=1)
        z = (F_r[ks] - mu) / np.sqrt(sd ** 2 * (1 + 1.0 / len(outs)) + 1e-300)
        say(f"  histogram T={T:.3f}: CDF ref {np.array2string(F_r[ks], precision=3)} z {np.array2string(z, precision=2)}")
        zs.extend(z.tolist())
    return zs


@pytest.mark.parametrize("precision", [J.FP32_BATCHED, J.FP64_PLAIN])
def test_frozen_pt_run_matches_reference(golden_dir, precision):
    """Reference run ref_pt_frozen: L = 8, NREPLICA = 5, T = 0.01 .. 0.02, FER
    start, phi = 0, NEQ = 11, NMESS = 7.  Nothing flips, every energy is -1 and
    every swap is accepted, so the run is deterministic: the engine must
    reproduce the histogram (all in bin 51 of [-3, 1]), the swap attempts and
    accepts per pair, the visited-both-ends count and the energies exactly."""
    say(f"case frozen PT run vs reference, precision {precision}")
    gd = os.path.join(golden_dir, "ref_pt_frozen")
    L, R = 8, 5
    g = J.make_params(L, L, nreplica=R, tmin=0.01, tmax=0.02, neq=11, nmess=7, precision=precision, seed=4)
    eng = J.Engine([g])
    eng.upload_disorder(0, np.zeros(L ** 3))
    s0 = np.tile(J.initial_spins(L, L, J.START_FER, 0, 0), R)
    eng.set_spins(0, s0)
    eng.run(J.PHASE_EQUIL)
    eng.run(J.PHASE_MEAS)
    out = eng.fetch(0)
    assert np.array_equal(eng.spins(0).reshape(-1), s0)
    eng.close()
    att, acc, rate, visit, frac = read_cfpt(find(gd, "cfpt"))
    assert np.array_equal(out["swap_att"], att) and np.array_equal(out["swap_acc"], acc)
    assert out["visitboth"] == visit
    for s, (T, bins, labels, freq) in enumerate(read_cfhs(find(gd, "cfhs"))):
        w = (labels[-1] - labels[0]) / (len(labels) - 1)
        assert abs(out["hist_min"][s] - labels[0]) < 1e-6 and abs(out["hist_bin"][s] - w) < 1e-7
        assert np.array_equal(out["hist"][:99, s], freq)
    assert np.array_equal(out["basefields"][9], read_table(find(gd, "cfen"))[:, 1])


@pytest.mark.parametrize("precision", [J.FP64_PLAIN, J.FP32_BATCHED])
def test_local_matches_reference_l8(golden_dir, precision):
    """Reference run ref_local_l8 (the reference's CPU-runnable case, an FP64
    run): LOCAL, L = 8, T = 10, HOT start, NEQ = NMESS = 1000 sweeps, in both
    forms.  The start is the reference's: KISS seeded with IRINIT = 1, 3*L^3
    draws consumed by Setup_vacancies (J1J2_6tv.f90:1116-1175), then
    Initialize_spins."""
    say(f"case LOCAL L=8 vs reference (64 chains), precision {precision}")
    L = 8
    N = L * L * L
    spins = J.initial_spins(L, L, J.START_HOT, 0, 1, skip=3 * N)
    assert np.array_equal(spins, np.where(O.kiss_stream(1, 4 * N)[3 * N:] < 0.5, 1, -1))
    g = J.make_params(L, L, nreplica=1, ptswapever=1, sampler=J.SAMPLER_LOCAL, ladder=J.LADDER_LOCAL,
                      tmin=10.0, tmax=10.0, dt=1.0, neq=1000, nmess=1000, seed=4242, precision=precision)
    p = J.make_disorder(L, L, 3.0, 1, True, 2.0)
    zs, rows, _ = chains_vs_reference(g, p, spins, "local_l8", golden_dir)
    for name, ref, mu_, z in rows:
        say(f"  {name:7s} ref {np.array2string(ref, precision=5)} engine {np.array2string(mu_, precision=5)} "
            f"z {np.array2string(z, precision=2)}")
    frac = zsummary(zs, f"local_l8_vs_reference_{'fp64' if precision == J.FP64_PLAIN else 'fp32'}")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.75


def paired_z(d):
    """Paired differences d (pairs along axis 0): the t statistic of their mean,
    expressed as the normal z-score of the same two-sided p-value (t with
    n - 1 degrees of freedom), so that few pairs (8 configs at L = 96) are
    judged as fairly as many."""
    from scipy.stats import norm, t as tdist
    n = d.shape[0]
    sd = d.std(0, ddof=1)
    tt = np.where(sd > 0, d.mean(0) / (sd / np.sqrt(n) + 1e-300), 0.0)
    p = 2.0 * tdist.sf(np.abs(tt), n - 1)
    return np.sign(tt) * norm.isf(np.maximum(p, 1e-300) / 2.0)


def fp32_vs_fp64(label, groups, neq, nmess, same_disorder=None, sampler=J.SAMPLER_PT, temps=(8.0, 11.5)):
    """Both forms on the same configs, Philox keys and starts: groups are
    (L, nreplica, nconf, start); the FP32 run also counts, per config, the
    proposals whose FP32 decision differs from the FP64 rule's.  Returns the
    paired z-scores (per group, quantity and slot, and per swap pair) and
    the per-config mismatch counts."""
    res = {}
    for prec in (J.FP32_BATCHED, J.FP64_PLAIN):
        params, iconf = [], 1
        for L, R, C, _ in groups:
            params.append(J.make_params(L, L, nreplica=R, nconf=C, iconf0=iconf, precision=prec, neq=neq, nmess=nmess,
                                        seed=99, sampler=sampler, tmin=temps[0], tmax=temps[1],
                                        ptswapever=5 if sampler == J.SAMPLER_PT else 1,
                                        ladder=J.LADDER_LINEAR if sampler == J.SAMPLER_PT else J.LADDER_LOCAL,
                                        flags=J.FLAG_SHADOW_FP64 if prec == J.FP32_BATCHED else 0))
            iconf += C
        eng = J.Engine(params)
        c = 0
        for (L, R, C, start), g in zip(groups, params):
            for k in range(C):
                if same_disorder is not None:
                    p, sp = same_disorder
                else:
                    p, sp = config_inputs(L, L, R, g.iconf0 + k, "hot" if start == "mixed" and k % 2 == 0 else
                                          ("str" if start == "mixed" else start))
                eng.upload_disorder(c, p)
                eng.set_spins(c, np.asarray(sp).reshape(-1))
                c += 1
        eng.run(J.PHASE_EQUIL, neq)
        eng.run(J.PHASE_MEAS, nmess)
        outs = [eng.fetch(c) for c in range(eng.nconf)]
        mis = [eng.mismatch(c) for c in range(eng.nconf)] if prec == J.FP32_BATCHED else None
        res[prec] = (outs, mis)
        eng.close()
    zs, c0 = [], 0
    for L, R, C, _ in groups:
        sl = slice(c0, c0 + C)
        b32 = np.stack([o["basefields"] for o in res[J.FP32_BATCHED][0][sl]])
        b64 = np.stack([o["basefields"] for o in res[J.FP64_PLAIN][0][sl]])
        for name, f, _, _ in QUANTITIES:
            d = b32[:, f, :] - b64[:, f, :]
            zs.extend(paired_z(d).tolist())
        if sampler == J.SAMPLER_PT and R > 1:
            a32 = np.stack([o["swap_acc"] / o["swap_att"] for o in res[J.FP32_BATCHED][0][sl]])
            a64 = np.stack([o["swap_acc"] / o["swap_att"] for o in res[J.FP64_PLAIN][0][sl]])
            zs.extend(paired_z(a32 - a64).tolist())
        c0 += C
    mis = res[J.FP32_BATCHED][1]
    prop = np.array([m[0] for m in mis], np.float64)
    bad = np.array([m[1] for m in mis], np.float64)
    say(f"  {label}: flip-decision mismatch {bad.sum():.0f} of {prop.sum():.3g} proposals "
        f"({bad.sum() / prop.sum():.3g}); per config {np.min(bad / prop):.3g} .. {np.max(bad / prop):.3g}")
    if OUT:
        os.makedirs(OUT, exist_ok=True)
        with open(os.path.join(OUT, f"mismatch_{label}.json"), "w") as f:
            json.dump({"proposals": prop.tolist(), "mismatches": bad.tolist()}, f)
    assert np.all(prop > 0) and bad.sum() / prop.sum() < 1e-6
    return zs


def judge(zs, label):
    frac = zsummary(zs, label)
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.85


def test_fp32_matches_fp64_c1_local_l8():
    """c1 shape: LOCAL, L = 8, T = 10, the reference's HOT start and disorder,
    64 chains (independent Philox keys), 1000 + 1000 sweeps."""
    say("case FP32 vs FP64, c1 shape (LOCAL L=8)")
    L, N = 8, 512
    same = (J.make_disorder(L, L, 3.0, 1, True, 2.0), J.initial_spins(L, L, J.START_HOT, 0, 1, skip=3 * N))
    judge(fp32_vs_fp64("c1_local_l8", [(8, 1, 64, None)], 1000, 1000, same_disorder=same,
                       sampler=J.SAMPLER_LOCAL, temps=(10.0, 10.0)), "fp32_vs_fp64_c1")


def test_fp32_matches_fp64_c2_l16():
    """c2 shape: PT, L = 16, 16 replicas, STR-X start, the disorder of IRINIT = 1,
    64 chains (independent Philox keys), NEQ = NMESS = 2000."""
    say("case FP32 vs FP64, c2 shape (L=16 16 replicas)")
    L, R = 16, 16
    same = (J.make_disorder(L, L, 3.0, 1, True, 2.0), np.tile(J.initial_spins(L, L, J.START_STR, 0, 0), R))
    judge(fp32_vs_fp64("c2_pt_l16", [(16, 16, 64, None)], 2000, 2000, same_disorder=same), "fp32_vs_fp64_c2")


def test_fp32_matches_fp64_l32():
    """c3: L = 32, 16 replicas x 64 disorder configs, STR-X start."""
    say("case FP32 vs FP64 at L=32")
    judge(fp32_vs_fp64("c3_l32", [(32, 16, 64, "str")], 300, 1500), "fp32_vs_fp64_l32")


def test_fp32_matches_fp64_c4_uneven():
    """c4: L in {16, 24, 32, 48} x 16 configs, 8/12/16/24 replicas, HOT and STR
    starts alternating; paired per size over its 16 configs."""
    say("case FP32 vs FP64, c4 uneven batch")
    judge(fp32_vs_fp64("c4_uneven", [(16, 8, 16, "mixed"), (24, 12, 16, "mixed"), (32, 16, 16, "mixed"),
                                     (48, 24, 16, "mixed")], 300, 1500), "fp32_vs_fp64_c4")


def test_fp32_matches_fp64_c5_l96():
    """c5: L = 96, 24 replicas x 8 configs, STR-X start; paired over the 8 configs."""
    say("case FP32 vs FP64, c5 L=96")
    judge(fp32_vs_fp64("c5_l96", [(96, 24, 8, "str")], 200, 600), "fp32_vs_fp64_c5")


def test_pt_matches_reference_c2(golden_dir):
    """Reference run ref_pt_l16_c2 (the c2 configuration, run by the reference
    on the CPU in FP64): L = 16, 16 replicas over T = 8.0-11.5, STR-X start,
    NEQ = NMESS = 2000 cycles, PTSWAPEVER = 5, disorder IRINIT = 1.  64 chains of
    the plain FP64 form on the same disorder: every quantity per slot, the swap
    acceptances, the visited-both-ends count and the energy histograms."""
    say("case PT L=16 (c2) FP64 plain vs reference (64 chains)")
    L, R = 16, 16
    gd = os.path.join(golden_dir, "ref_pt_l16_c2")
    g = J.make_params(L, L, nreplica=R, neq=2000, nmess=2000, seed=1618, precision=J.FP64_PLAIN)
    p = J.make_disorder(L, L, 3.0, 1, True, 2.0)
    spins = np.tile(J.initial_spins(L, L, J.START_STR, 0, 0), R)
    zs, rows, outs = chains_vs_reference(g, p, spins, "pt_l16_c2", golden_dir)
    att, acc, rate, visit, frac = read_cfpt(find(gd, "cfpt"))
    assert all(np.array_equal(o["swap_att"], att) for o in outs)
    a = np.stack([o["swap_acc"] / o["swap_att"] for o in outs])
    zs.extend(((rate - a.mean(0)) / np.sqrt(a.std(0, ddof=1) ** 2 * (1 + 1.0 / NCHAIN))).tolist())
    v = np.array([o["visitboth"] for o in outs], np.float64)
    if v.std() > 0:
        zs.append(float((visit - v.mean()) / (v.std(ddof=1) * np.sqrt(1 + 1.0 / NCHAIN))))
    else:
        assert visit == v[0]
    say(f"  visited both ends: reference {visit}, engine {v.mean():.2f} +- {v.std():.2f}")
    for name, ref, mu_, z in rows:
        say(f"  {name:7s} max |z| {np.max(np.abs(z)):.2f}")
    frac = zsummary(zs, "pt_l16_c2_vs_reference")
    assert np.all(np.abs(zs) < 4.0) and frac >= 0.75
    hz = histogram_vs_reference(outs, gd)
    frac = zsummary(hz, "pt_l16_c2_histogram_vs_reference")
    assert np.all(np.abs(hz) < 4.0) and frac >= 0.75


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
