"""The sweep kernels' first-stage decision (CPU only).

The FP32 rule accepts a move with the site's 32-bit word r = v 2^16 + l iff
(float)r <= T(y), T the oracle's threshold (ora_accept_threshold_f32).  The
kernels decide from the 16-bit v alone when that settles it for every l:
with F = 2^16 + v 2^-7 and e the hardware exp2 of y,

    d = RN_f32(F - 2^9 e)            (one fma)
    accept every l   if d < 2^16 - 2^-7
    reject every l   if d > 2^16
    otherwise        draw l and apply the exact rule.

This restates that filter with exact rational arithmetic (the fma's single
rounding done exactly) and checks it against the exact rule over the whole
range of y the engine allows, with e anywhere within 2 ulp of 2^y: what the
filter accepts the rule accepts for every l, what it rejects the rule rejects
for every l.  A filter bound moved by one float step, or the fma's operands
swapped, fails here.  It also checks that the undecided band stays a few
values of v wide (the kernels' cost depends on it).
"""
import os
import re
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

T_ACC = 65536.0 - 2.0 ** -7          # 0x1.fffffcp15
T_REJ = 65536.0                      # 0x1p16


def rne_f32(q):
    """Round the rational q to the nearest float32 (ties to even), as a float."""
    if q == 0:
        return 0.0
    sign = -1 if q < 0 else 1
    q = abs(q)
    ex = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** ex > q:
        ex -= 1
    if q >= Fraction(2) ** (ex + 1):
        ex += 1
    ex = max(ex, -126)
    scale = Fraction(2) ** (23 - ex)
    n = q * scale
    fl = n.numerator // n.denominator
    rem = n - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * float(Fraction(fl) / scale)


def f32_steps(x, k):
    """The float32 k steps away from float32 x."""
    v = np.float32(x)
    direction = np.float32(np.inf) if k > 0 else np.float32(-np.inf)
    for _ in range(abs(k)):
        v = np.nextafter(v, direction)
    return float(v)


def filter_d(v, e):
    return rne_f32(Fraction((1 << 23) + v, 1 << 7) - 512 * Fraction(e))


def rule_accepts(r, T):
    return float(np.float32(r)) <= T


def last_true(pred, lo, hi):
    """Largest v in [lo, hi] with pred(v) (pred monotone decreasing), or lo - 1."""
    if not pred(lo):
        return lo - 1
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if pred(mid):
            lo = mid
        else:
            hi = mid - 1
    return lo


Y_GRID = sorted(set([float(np.float32(y)) for y in np.concatenate([
    -np.geomspace(1e-7, 40.0, 260),          # the whole accepted range of negative exponents
    np.linspace(-16.0, 0.0, 97),             # around integer exponents
    [0.0, 1e-6, 0.25, 1.0, 3.5, 20.0],      # downhill moves: always accepted
])]))


@pytest.mark.parametrize("k", [-2, 0, 2])
def test_filter_decisions_agree_with_exact_rule(k):
    widest = 0
    for y in Y_GRID:
        T = float(O.accept_threshold_f32(y))
        e = f32_steps(2.0 ** y, k)              # the hardware exp2, within 2 ulp
        v_acc = last_true(lambda v: filter_d(v, e) < T_ACC, 0, 65535)
        v_rej = last_true(lambda v: not filter_d(v, e) > T_REJ, 0, 65535) + 1
        assert v_acc < v_rej, (y, k)
        if v_acc >= 0:   # the filter accepts v <= v_acc: so must the rule for the largest word
            assert rule_accepts(v_acc * 65536 + 65535, T), (y, k, v_acc)
        if v_rej <= 65535:   # it rejects v >= v_rej: so must the rule for the smallest word
            assert not rule_accepts(v_rej * 65536, T), (y, k, v_rej)
        if y <= 0.0 and 2.0 ** y > 2.0 ** -15:
            widest = max(widest, v_rej - v_acc - 1)
        if y >= 0.25:    # clearly downhill: every v decided as an accept
            assert v_acc == 65535, (y, k)
    assert widest <= 4, widest


def test_filter_bounds_are_the_kernel_constants():
    """The kernels' decide16_filter (device_common.h) uses exactly this fma
    and these two float32 bounds."""
    src = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "j1j2-public_amd", "csrc",
                            "device_common.h")).read()
    body = src[src.index("Filt decide16_filter(float y, float F)"):]
    body = body[:body.index("return Filt")]
    assert re.search(r"d = __builtin_fmaf\(e, -0x1p9f, F\);", body)
    acc = re.search(r"acc = d < (0x[0-9a-fA-F.]+p[-+]?\d+)f;", body)
    rej = re.search(r"rej = d > (0x[0-9a-fA-F.]+p[-+]?\d+)f;", body)
    assert acc and rej
    assert float.fromhex(acc.group(1)) == T_ACC and float(np.float32(T_ACC)) == T_ACC
    assert float.fromhex(rej.group(1)) == T_REJ and float(np.float32(T_REJ)) == T_REJ


def _violations(acc_bound, rej_bound):
    bad = 0
    for y in Y_GRID:
        T = float(O.accept_threshold_f32(y))
        for k in (-2, 0, 2):
            e = f32_steps(2.0 ** y, k)
            v_acc = last_true(lambda v: filter_d(v, e) < acc_bound, 0, 65535)
            v_rej = last_true(lambda v: not filter_d(v, e) > rej_bound, 0, 65535) + 1
            bad += v_acc >= 0 and not rule_accepts(v_acc * 65536 + 65535, T)
            bad += v_rej <= 65535 and rule_accepts(v_rej * 65536, T)
    return bad


def test_planted_looser_bounds_fail():
    """The bounds are tight: one float step looser on either side lets the
    filter decide a word against the rule somewhere on the grid, which the
    check above would catch."""
    up = float(np.nextafter(np.float32(T_ACC), np.float32(np.inf)))
    down = float(np.nextafter(np.float32(T_REJ), np.float32(-np.inf)))
    assert _violations(T_ACC, T_REJ) == 0
    assert _violations(up, T_REJ) > 0
    assert _violations(T_ACC, down) > 0
