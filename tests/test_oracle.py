"""Pins of the CPU oracle (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a transposed neighbour, a wrong seed offset) fails:
reference outputs kept as fixtures, published known-answer vectors, closed
forms, brute force and exact enumeration.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from refdata import find, read_phi_sites, read_table

PHILOX_KAT = [  # Random123 kat_vectors, philox4x32_10
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(v) for v in O.philox(ctr, key)) == out


@pytest.mark.parametrize("L,H", [(8, 4), (16, 6), (6, 4), (10, 2)])
def test_site_words_use_each_draw_once(L, H):
    """The sweep's random words: every site of a colour gets exactly one
    word; every 16-bit half of every call's four words goes to exactly one
    site (L % 4 == 0; the short last group of a layer leaves some unused
    otherwise), high parts from the calls whose second counter word is
    c | rep << 2, low parts from the calls with bit 30 added.  Mistakes such
    as two layers taking the same half, or the low part coming from the high
    part's call, fail here."""
    seed, iconf, rep, sweep = 12345, 7, 3, (5 << 32) + 9
    half = L // 2
    lay, ncol = half * half, L * L * H // 4
    gl = (lay + 3) // 4
    for c in range(4):
        seen = np.zeros(ncol, np.int64)
        his, los, drawn_hi, drawn_lo = [], [], [], []
        for m in range(H // 2 * gl):
            qi, r = O.call_words(L, seed, iconf, rep, sweep, c, m)
            seen[qi] += 1
            his += list(r >> 16)
            los += list(r & 0xFFFF)
            ctr = (m, c | rep << 2, sweep & 0xFFFFFFFF, sweep >> 32)
            hw = O.philox(ctr, (seed, iconf))
            lw = O.philox((ctr[0], ctr[1] | 0x40000000, ctr[2], ctr[3]), (seed, iconf))
            # layer 2zp takes the high halves, layer 2zp + 1 the low halves
            n_even = sum(1 for q in qi if (q // lay) % 2 == 0)
            assert all((q // lay) // 2 == m // gl for q in qi)
            assert list(r[:n_even] >> 16) == [int(w) >> 16 for w in hw[:n_even]]
            assert list(r[n_even:] >> 16) == [int(w) & 0xFFFF for w in hw[:len(qi) - n_even]]
            drawn_hi += [int(w) >> 16 for w in hw] + [int(w) & 0xFFFF for w in hw]
            drawn_lo += [int(w) >> 16 for w in lw] + [int(w) & 0xFFFF for w in lw]
        assert np.all(seen == 1)
        if L % 4 == 0:
            assert sorted(his) == sorted(drawn_hi) and sorted(los) == sorted(drawn_lo)
        # the low parts are not the high parts' halves
        assert sum(a == b for a, b in zip(los, his)) < 3


def test_exp64_accuracy():
    xs = np.concatenate([np.linspace(-700, 0, 20001), np.linspace(-1, 1, 2001), [0.0, -1e-300]])
    for x in xs:
        ref = math.exp(x)
        got = O.exp64(x)
        assert abs(got - ref) <= 2.3e-16 * ref, x
    assert O.exp64(-800.0) == 0.0


def test_accept_threshold_accuracy():
    ys = np.linspace(-64, 0, 100001).astype(np.float32)
    for y in ys[::7]:
        ref = 2.0 ** (32.0 + float(y))
        assert abs(O.accept_threshold_f32(float(y)) - ref) <= 1.2e-7 * ref
    # every dE <= 0 move is taken: (float)r <= 2^32 for every 32-bit r
    assert O.accept_threshold_f32(0.0) == 2.0 ** 32
    assert O.accept_threshold_f32(-0.0) == 2.0 ** 32
    assert all(O.accept_threshold_f32(y) >= 2.0 ** 32 for y in (1e-7, 0.3, 0.5, 0.51, 7.0, 95.0))
    assert O.accept_threshold_f32(200.0) == math.inf
    assert O.accept_threshold_f32(-200.0) == 0.0
    assert O.accept_threshold_f32(-5000.0) == 0.0


# ------------------------------------------------------------------ reference fixtures
@pytest.mark.parametrize("variant,L,H,seed", [("hot_init_l8", 8, 8, 3), ("hot_init_l6h4", 6, 4, 8),
                                               ("str_init_l12", 12, 10, 11)])
def test_corr_field_matches_reference(golden_dir, variant, L, H, seed):
    """phi_sites.dat of the reference (last config: seed IRINIT+NCONF-1) equals the
    oracle's make_corr_rand_3d restatement: pins KISS seeding, the polar Gaussian
    order over the padded grid, the filter, the normalisation and the erf map."""
    ref = read_phi_sites(os.path.join(golden_dir, "ref_" + variant, "phi_sites.dat"), L, H)
    mine = O.corr_field(L, L, H, 3.0, seed, True, 2.0)
    assert np.max(np.abs(mine - ref)) < 2e-11
    assert np.max(np.abs(ref)) < 2.0


@pytest.mark.parametrize("variant,L,H,nrep,nconf,irinit", [("hot_init_l8", 8, 8, 4, 3, 1),
                                                           ("hot_init_l6h4", 6, 4, 2, 2, 7)])
def test_hot_start_energies_match_reference(golden_dir, variant, L, H, nrep, nconf, irinit):
    """With NEQ = NMESS = 0 the reference reports the energy of each replica's
    HOT start (seed IRINIT+(iconf-1)*NREPLICA+rep, J1J2_6tv.f90:1238) on the
    config's field (seed IRINIT+iconf-1), averaged over configs (cfen, es24.16)."""
    N = L * L * H
    en = np.zeros((nconf, nrep))
    for ic in range(1, nconf + 1):
        lat = O.OracleLattice(L, H, O.corr_field(L, L, H, 3.0, irinit + ic - 1, True, 2.0))
        for rep in range(1, nrep + 1):
            s = O.initial_spins(L, H, 0, 0, irinit + (ic - 1) * nrep + rep)
            en[ic - 1, rep - 1] = lat.measurement_ref(s)[1]
            # the engine's integer/fixed-point energy agrees with the literal one
            assert abs(lat.energy_total(lat.observables(s)) / N - en[ic - 1, rep - 1]) < 1e-9
    tab = read_table(find(os.path.join(golden_dir, "ref_" + variant), "cfen"))
    mean = en.mean(0)
    std = np.sqrt(np.maximum((en ** 2).mean(0) - mean ** 2, 0)) / math.sqrt(nconf)
    assert np.allclose(tab[:, 1], mean, rtol=0, atol=1e-14)
    assert np.allclose(tab[:, 2], std, rtol=0, atol=1e-12)


def test_stripe_start_matches_reference_and_closed_form(golden_dir):
    L, H = 12, 10
    p = O.corr_field(L, L, H, 3.0, 11, True, 2.0)
    lat = O.OracleLattice(L, H, p)
    s = O.initial_spins(L, H, 2, 1, 0)           # STR, STRAXIS = Y: s = (-1)^x, l_i = -4
    e = lat.measurement_ref(s)[1]
    assert abs(e - (-3.0 + 4.0 * p.mean())) < 1e-13
    sx = O.initial_spins(L, H, 2, 0, 0)          # STRAXIS = X: s = (-1)^y, l_i = +4
    assert abs(lat.measurement_ref(sx)[1] - (-3.0 - 4.0 * p.mean())) < 1e-13
    tab = read_table(find(os.path.join(golden_dir, "ref_str_init_l12"), "cfen"))
    assert np.allclose(tab[:, 1], e, atol=1e-14)


def test_ferro_closed_form():
    L, H = 6, 4
    rng = np.random.default_rng(1)
    lat = O.OracleLattice(L, H, rng.uniform(-2, 2, L * L * H))
    s = O.initial_spins(L, H, 1, 0, 0)
    m, e, sx, sy = lat.measurement_ref(s)
    assert (m, sx, sy) == (1.0, 0.0, 0.0)
    assert abs(e + 1.0) < 1e-14
    assert lat.nbar_ref(s) == 0.0


# ------------------------------------------------------------------ local field and sweep rule
@pytest.mark.parametrize("L,H", [(4, 4), (2, 4), (6, 2)])
def test_local_field_is_half_energy_difference(L, H):
    """Brute force: E(s with site i flipped) - E(s) = 2 s_i h_i for every site,
    with E the literal Measurement energy; the FP32 field agrees to FP32."""
    rng = np.random.default_rng(L * 10 + H)
    N = L * L * H
    p = rng.uniform(-2, 2, N)
    lat = O.OracleLattice(L, H, p, J1=1.0, deltaJ1=0.3, J2=-0.7, J3=1.3)
    s = np.where(rng.random(N) < 0.5, 1, -1).astype(np.int8)
    E0 = lat.measurement_ref(s)[1] * N
    for i in range(N):
        t = s.copy()
        t[i] = -t[i]
        dE = lat.measurement_ref(t)[1] * N - E0
        h = lat.field_f64(s, i)
        assert abs(dE - 2 * s[i] * h) < 1e-10
        assert abs(lat.field_f32(s, i) - h) < 2e-5 * (1 + abs(h))
        o = lat.observables(t)
        assert abs(lat.energy_total(o) - lat.measurement_ref(t)[1] * N) < 1e-6


@pytest.mark.parametrize("L,H", [(2, 2), (2, 4), (4, 2), (6, 4), (8, 8)])
def test_colouring_is_proper(L, H):
    N = L * L * H
    col = np.array([O.colour(L, i) for i in range(N)])
    assert all(np.sum(col == c) == N // 4 for c in range(4))
    for i in range(N):
        x, y, z = i % L, (i // L) % L, i // (L * L)
        for dx, dy, dz in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (-1, 1, 0)]:
            j = (x + dx) % L + L * ((y + dy) % L) + L * L * ((z + dz) % H)
            if j != i:
                assert col[i] != col[j]
    idx = [O.lib().ora_colour_index(L, i) for i in range(N)]
    for c in range(4):
        assert sorted(idx[i] for i in range(N) if col[i] == c) == list(range(N // 4))


@pytest.mark.parametrize("precision", [0, 1])
def test_infinite_temperature_flips_every_site(precision):
    L, H = 4, 4
    rng = np.random.default_rng(5)
    lat = O.OracleLattice(L, H, rng.uniform(-2, 2, L * L * H))
    s = np.where(rng.random(L * L * H) < 0.5, 1, -1).astype(np.int8)
    s0 = s.copy()
    lat.sweep(s, precision, 1e-300, 7, 1, 0, 0)
    assert np.array_equal(s, -s0)


@pytest.mark.parametrize("precision", [0, 1])
def test_zero_temperature_never_goes_uphill(precision):
    L, H = 6, 4
    rng = np.random.default_rng(9)
    N = L * L * H
    lat = O.OracleLattice(L, H, rng.uniform(-2, 2, N))
    s = np.where(rng.random(N) < 0.5, 1, -1).astype(np.int8)
    e_prev = lat.measurement_ref(s)[1]
    for k in range(20):
        lat.sweep(s, precision, 1e6, 3, 1, 0, k)
        e = lat.measurement_ref(s)[1]
        assert e <= e_prev + 1e-12
        e_prev = e


def test_equal_temperatures_accept_every_swap():
    L, H, R = 4, 2, 4
    rng = np.random.default_rng(2)
    p = rng.uniform(-2, 2, L * L * H)
    s0 = np.where(rng.random(R * L * L * H) < 0.5, 1, -1).astype(np.int8)
    pt = O.OraclePT(L, H, p, [9.0] * R, s0, ptswapever=1)
    pt.begin_phase()
    pt.cycle(False)                     # pairs (0,1), (2,3)
    assert list(pt.rep_at_temp) == [1, 0, 3, 2]
    pt.cycle(False)                     # pair (1,2)
    assert list(pt.rep_at_temp) == [1, 3, 0, 2]
    assert list(pt.attempts) == [1, 1, 1] and list(pt.accepts) == [1, 1, 1]
    rat, tor = pt.rep_at_temp, pt.temp_of_rep
    assert all(tor[rat[s]] == s for s in range(R))


# ------------------------------------------------------------------ exact enumeration
@pytest.mark.parametrize("precision", [0, 1])
def test_metropolis_matches_exact_enumeration_2x2x4(precision):
    """Equilibrium of the four-colour Metropolis chain on the 2x2x4 lattice
    (2^16 states) against the exact Boltzmann averages, within 4 sigma from
    batch means."""
    L, H, T = 2, 4, 4.0
    rng = np.random.default_rng(11)
    p = rng.uniform(-1.0, 1.0, L * L * H)
    lat = O.OracleLattice(L, H, p)
    exact = lat.enumerate(1.0 / T)
    nmess = 200000
    pt = O.OraclePT(L, H, p, [T], O.initial_spins(L, H, 1, 0, 0), ptswapever=1, local=True,
                    precision=precision, series_cap=nmess, seed=99)
    pt.run(2000, False)
    pt.run(nmess, True)
    ser = pt.series[0]
    m, e, sx, sy, nb = ser
    for vals, ref in ((e, exact[0]), (e * e, exact[1]), (np.abs(m), exact[2]), (m * m, exact[3]),
                      (nb, exact[4]), (sx * sx, exact[5]), (sy * sy, exact[6])):
        blocks = vals.reshape(200, -1).mean(1)
        sigma = blocks.std(ddof=1) / math.sqrt(len(blocks))
        assert abs(vals.mean() - ref) < 4 * sigma + 1e-12, (vals.mean(), ref, sigma)


# ------------------------------------------------------------------ derived observables
def test_derived_closed_forms():
    """Constant series: susceptibilities and specific heat vanish, Binder
    cumulants take their deterministic values (J1J2_6tv.f90:1581-1626)."""
    n, m, e, sx, sy, nb = 10, 0.5, -2.0, 0.3, -0.4, 0.25
    s2 = sx * sx + sy * sy
    sums = np.array([m, abs(m), m * m, m ** 4, e, e * e, abs(sx), abs(sy), sx * sx, sx ** 4,
                     sy * sy, sy ** 4, s2, s2 * s2, math.sqrt(s2), abs(sy * sy - sx * sx), nb,
                     abs(nb), nb * nb, nb ** 4]) * n
    o = O.derive(sums, n, 0.1, 512, -3.0, -2.5)
    assert abs(o[0] - 0.5) < 1e-15 and abs(o[3]) < 1e-12 and abs(o[11]) < 1e-12
    assert abs(o[5] - 2.0 / 3.0) < 1e-15 and abs(o[20] - 1.0) < 1e-14
    assert abs(o[27] - 2.0 / 3.0) < 1e-15 and abs(o[41] - 2.0 / 3.0) < 1e-15
    assert abs(o[8] - math.log(0.5)) < 1e-15
    assert o[35] == -3.0 and o[36] == -2.5
    assert abs(o[34] - (0.09 - 0.16)) < 1e-15
    # fluctuating magnetisation: susc = (<m^2> - <|m|>^2) N beta
    sums2 = np.zeros(20)
    sums2[0], sums2[1], sums2[2], sums2[3] = 0.0, 2.0, 2.0, 2.0   # m = +-1 twice
    o2 = O.derive(sums2, 2, 0.5, 100, 0.0, 0.0)
    assert o2[3] == 0.0 and abs(o2[5] - (1 - 1 / 3)) < 1e-15


def _occupied_bins(hist):
    nz = np.nonzero(hist)[0] + 1                           # 1-based bin numbers
    return nz.min(), nz.max()


def test_local_histogram_rule_matches_reference(golden_dir):
    """Process_data's histogram (J1J2_6tv.f90:2296-2350) from the reference's
    own LOCAL outputs.  A dynamic run (cfhs of ref_local_l8): the range is the
    series' [min, max] widened by half its width on both sides and a sample
    goes to bin NINT((e - lo) / width), so the lowest energy lands in bin 25
    and the highest in bin 75 whatever the series, and no sample falls
    outside (frequencies sum to 1).  A frozen run (ref_local_frozen: FER
    start, phi = 0, T = 0.01, every energy -1): the range is e -+ 1e-6 and
    all samples land in bin 50.  A floor+1 binning, another widening or a
    dropped degenerate branch moves these bins."""
    # rows: T, bin, energy, frequency (Write_files prints bins 1..BINS-1)
    tab = read_table(find(os.path.join(golden_dir, "ref_local_l8"), "cfhs"))
    assert np.array_equal(tab[:, 1], np.arange(1, 100))
    assert _occupied_bins(tab[:, 3]) == (25, 75) and abs(tab[:, 3].sum() - 1.0) < 1e-5
    rng = np.random.default_rng(5)
    n = 1000
    series = [rng.uniform(-1, 1, n), rng.normal(-4.5, 0.4, n), rng.uniform(-1, 1, n),
              rng.uniform(-1, 1, n), rng.uniform(-4, 4, n)]
    _, h, hmin, hbin = O.process_local(*series, 0.1, 512, 100)
    assert _occupied_bins(h) == (25, 75) and abs(h.sum() - 1.0) < 1e-12
    e = series[1]
    assert hmin == e.min() - 0.5 * (e.max() - e.min()) and hbin == 2.0 * (e.max() - e.min()) / 100

    frozen = read_table(find(os.path.join(golden_dir, "ref_local_frozen"), "cfhs"))
    en = read_table(find(os.path.join(golden_dir, "ref_local_frozen"), "cfen"))
    assert en[0, 1] == -1.0                                # FER closed form: no move was taken
    assert np.array_equal(frozen[:, 1], np.arange(1, 100))
    assert _occupied_bins(frozen[:, 3]) == (50, 50) and frozen[49, 3] == 1.0
    const = [np.full(200, v) for v in (1.0, -1.0, 0.0, 0.0, 0.0)]
    _, h, hmin, hbin = O.process_local(*const, 100.0, 512, 100)
    assert _occupied_bins(h) == (50, 50) and h[49] == 1.0
    assert hmin == -1.0 - 1.0e-6 and hbin == ((-1.0 + 1.0e-6) - (-1.0 - 1.0e-6)) / 100


def test_ladders():
    t = O.ladder(0, 16, 8.0, 11.5)
    assert t[0] == 8.0 and abs(t[-1] - 11.5) < 1e-15 and np.allclose(np.diff(t), 3.5 / 15)
    ts = O.ladder(1, 11, 8.0, 14.0, tmid=9.5, tfrac=0.7)
    assert list(ts[:1]) == [8.0] and abs(ts[7] - 9.5) < 1e-15 and abs(ts[-1] - 14.0) < 1e-15
    assert np.allclose(np.diff(ts[:8]), 1.5 / 7) and np.allclose(np.diff(ts[7:]), 4.5 / 3)
    tl = O.ladder(2, 4, 8.0, 14.0, dt=2.0)
    assert list(tl) == [8.0, 10.0, 12.0, 14.0]


def test_pt_nematic_and_energy_bookkeeping():
    L, H, R = 4, 4, 3
    rng = np.random.default_rng(4)
    p = rng.uniform(-2, 2, L * L * H)
    s0 = np.where(rng.random(R * L * L * H) < 0.5, 1, -1).astype(np.int8)
    pt = O.OraclePT(L, H, p, [8.0, 9.0, 10.0], s0)
    pt.run(3, True)
    lat = O.OracleLattice(L, H, p)
    for r in range(R):
        s = pt.spins[r]
        o = lat.observables(s)
        assert np.array_equal(o, pt.obs[r])
        assert abs(lat.nbar_ref(s) - 2 * (o[3] - o[4]) / s.size) < 1e-15
        assert abs(pt.etot[r] - lat.measurement_ref(s)[1] * s.size) < 1e-6
