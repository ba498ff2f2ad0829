"""The PT energy histogram, the swap bookkeeping and the visited-both-ends
count of Pt_measurement_loop (J1J2_6tv.f90:1437-1442, 1508-1524, 1540) and
their output (cfhs, cfpt: J1J2_6tv.f90:3442-3483, 3535-3551), pinned to the
reference's own runs (tests/golden/ref_pt_l8_stat, ref_pt_frozen).

What the reference's files fix, independently of the oracle:
  * range: each slot's range is e0 -/+ (|e0| + 1) around its first measured
    energy e0, so half the printed range equals |centre| + 1, and the centre
    is a single sample, not <e>;
  * binning: bin = 1 + int((e - min) / width), so the printed label (the
    bin's lower edge) averages to <e> - width/2 (NINT binning would give
    <e> - width), and in the frozen run, where every energy is exactly -1 on
    the range [-3, 1] of width 0.04, every sample lands in bin 51 (NINT: 50);
  * swaps: the frozen run accepts every swap (all energies equal), so the
    attempts per pair (10, 8, 10, 8 over 11 + 7 cycles) fix the odd/even
    alternation and its restart with the odd pairs at the measurement phase,
    and 3 of the 5 replicas visit both ends (a replica counts once it has
    sat at slot 1 and at slot NREPLICA after some measuring cycle's swaps).
The oracle's own output is then held to the same facts.
"""
import os

import numpy as np

import oracle as O
from refdata import find, read_cfhs, read_cfpt, read_table

START_FER, START_STR = 1, 2


def hist_geometry(labels):
    """Lower edge of bin 1 and the bin width from the printed labels of bins 1 .. BINS-1."""
    width = (labels[-1] - labels[0]) / (len(labels) - 1)
    return labels[0], width


def test_reference_histogram_range_is_first_energy_rule(golden_dir):
    gd = os.path.join(golden_dir, "ref_pt_l8_stat")
    hs = read_cfhs(find(gd, "cfhs"))
    en = read_table(find(gd, "cfen"))
    N, nmess = 8 ** 3, 40000
    assert len(hs) == 4
    far = []
    for (T, bins, labels, freq), row in zip(hs, en):
        assert np.array_equal(bins, np.arange(1, 100))
        lo, w = hist_geometry(labels)
        # labels are lower edges of equal bins (e12.6: 6 significant digits)
        assert np.allclose(labels, lo + w * np.arange(99), atol=2e-5)
        centre = lo + 50 * w
        assert abs(50 * w - (abs(centre) + 1.0)) < 2e-4, (T, 50 * w, centre)
        # the centre is one sample of e: within 5 standard deviations of <e>
        # (sd from the printed specific heat c = N beta^2 var e) ...
        e_mean, c = row[1], row[3]
        sd = np.sqrt(c * T * T / N)
        assert abs(centre - e_mean) < 5 * sd
        # ... and not <e> itself: far beyond any error of the mean
        far.append(abs(centre - e_mean) / (sd * np.sqrt(100.0 / nmess)))
    assert max(far) > 10.0, far


def test_reference_histogram_labels_are_lower_edges(golden_dir):
    gd = os.path.join(golden_dir, "ref_pt_l8_stat")
    hs = read_cfhs(find(gd, "cfhs"))
    en = read_table(find(gd, "cfen"))
    for (T, bins, labels, freq), row in zip(hs, en):
        lo, w = hist_geometry(labels)
        assert 0.999 <= freq.sum() <= 1.0 + 1e-5      # bin BINS is not printed
        m = float((freq * labels).sum())
        assert abs(m + w / 2 - row[1]) < 0.02 * w, (T, (m + w / 2 - row[1]) / w)
        assert abs(m + w - row[1]) > 0.4 * w        # what NINT binning would print


def frozen_reference(golden_dir):
    gd = os.path.join(golden_dir, "ref_pt_frozen")
    return gd, read_cfhs(find(gd, "cfhs")), read_cfpt(find(gd, "cfpt")), read_table(find(gd, "cfen"))


def test_reference_frozen_run(golden_dir):
    """FER start, phi = 0, T = 0.01 .. 0.02 (5 slots): no spin ever flips, every
    energy is exactly -1, every swap is accepted."""
    gd, hs, (att, acc, rate, visit, frac), en = frozen_reference(golden_dir)
    e, width, lo = -1.0, 4.0 / 100, -3.0
    assert int((e - lo) / width) + 1 == 51 and int(np.rint((e - lo) / width)) == 50
    for T, bins, labels, freq in hs:
        l0, w = hist_geometry(labels)
        assert abs(l0 - lo) < 1e-6 and abs(w - width) < 1e-7
        assert freq[50] == 1.0 and freq.sum() == 1.0          # bin 51
    assert np.array_equal(att, [10, 8, 10, 8]) and np.array_equal(acc, att) and np.all(rate == 1.0)
    assert visit == 3 and abs(frac - 0.6) < 1e-12
    assert np.all(en[:, 1] == -1.0)


def frozen_oracle(precision):
    L, R = 8, 5
    temps = O.ladder(0, R, 0.01, 0.02)
    s0 = np.tile(O.initial_spins(L, L, START_FER, 0, 0), R)
    ora = O.OraclePT(L, L, np.zeros(L ** 3), temps, s0, precision=precision, seed=123, iconf=1)
    ora.run(11, False)
    ora.run(7, True)
    return ora, s0


def test_oracle_frozen_run_matches_reference(golden_dir):
    gd, hs, (att, acc, rate, visit, frac), en = frozen_reference(golden_dir)
    for precision in (0, 1):
        ora, s0 = frozen_oracle(precision)
        assert np.array_equal(ora.spins.reshape(-1), s0)      # nothing flipped
        assert np.array_equal(ora.attempts, att) and np.array_equal(ora.accepts, acc)
        assert ora.visitboth == visit
        h = ora.hist()                                        # slot, bin: fractions
        for s, (T, bins, labels, freq) in enumerate(hs):
            lo, w = hist_geometry(labels)
            assert abs(ora.hist_min[s] - lo) < 1e-6 and abs(ora.hist_bin[s] - w) < 1e-7
            assert np.array_equal(h[s, :99], freq)
        assert np.array_equal(ora.basefields()[9], en[:, 1])


def test_oracle_histogram_rules_on_a_dynamic_run():
    """L = 8, 4 replicas, T = 8 .. 11.5, STR-X start, the disorder of IRINIT = 1:
    the oracle's ranges centre on each slot's first measured energy with half
    width |e0| + 1, and its labels average to <e> - width/2 as the
    reference's do."""
    L, R = 8, 4
    p = O.corr_field(L, L, L, 3.0, 1, True, 2.0)
    temps = O.ladder(0, R, 8.0, 11.5)
    s0 = np.tile(O.initial_spins(L, L, START_STR, 0, 0), R)
    ora = O.OraclePT(L, L, p, temps, s0, precision=1, seed=77, iconf=1)
    ora.run(200, False)
    ora.begin_phase()
    ora.cycle(True)
    e0 = ora.final_en.copy()
    for _ in range(3999):
        ora.cycle(True)
    n = ora.s.nmeasure
    assert n == 4000 and np.all(ora.hist_counts.sum(1) == n)
    e_mean = ora.sums[:, 4] / n
    for s in range(R):
        lo, w = ora.hist_min[s], ora.hist_bin[s]
        assert lo + 50 * w == e0[s] or abs(lo + 50 * w - e0[s]) < 1e-12 * abs(e0[s])
        assert abs(50 * w - (abs(e0[s]) + 1.0)) < 1e-12 * (abs(e0[s]) + 1.0)
        labels = lo + w * np.arange(100)
        m = float((ora.hist_counts[s] * labels).sum() / n)
        assert abs(m + w / 2 - e_mean[s]) < 0.02 * w, (s, (m + w / 2 - e_mean[s]) / w)
