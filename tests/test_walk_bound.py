"""The tile walk's FP32 screen and its bounds, checked in exact arithmetic.

The walk (kdtree-rs_amd/csrc/kdt_walk.hip) scores tree point p for query q in
the tree's frame (centre c = the middle of the root box, k_screen_frame) as
    A = fma(m_{D-1}, p~_{D-1}, ... fma(m_0, p~_0, fl32(pn + qn)))
with p~ = fl32(fl64(p - c)) (the build's rows32c), q~ likewise, m = -2 q~, pn
and qn the FP32 roundings of the FP64 sums of the squares, and uses
    e1 = (a (Q + P_c) + t)(1 + 2^-50),  N = ((1 + a)(Q + P_c) + t)(1 + 2^-50),
    E2 = (D + 3) 2^-24 N^2 (1 + 2^-10) + (D + 4) 2^-149,
    a = 2^-24 + 2^-52,  t = 2 sqrt(D) 2^-150,  Q = |fl64(q - c)| (1 + 2^-44),
P_c the build's bound on |p - c| (k_rows32c: rounded up to FP32), e1 and E2
rounded up to FP32, s = 1 + 2 (D + 3) 2^-53 and
    U(A) = s (sqrt(max(A, 0) + E2) + e1)^2 (1 + 2^-50)     (FP64 arithmetic),
    T(b) = fl32_up(((sqrt(s b) + e1)^2 + E2)(1 + 2^-50))  (screen_t32).
Every FP32 operation is emulated exactly here (rationals, round to nearest
even, subnormals), and the checks are the facts the walk relies on:
  core       |A - |q~ - p~|^2| <= E2
  inclusion  A <= T(d), d the reference's FP64 reduced distance (the oracle's
             point_dist): a point with d <= b is appended and is a candidate
  upper      d <= U(A): k points seen bound the k-th distance by U(A_k)
  box        bd32 <= TB(d) for any box holding p, bd32 the walk's FP32 box
             distance over the build's outward-rounded centred FP32 box
             (k_boxes32c) and TB(b) = fl32_up(((1 + (D + 3) u)(sqrt(s b) + e1)^2
             + (D + 1) 2^-149)(1 + 2^-50)): a node holding a point with d <= b
             is never skipped
  update     after an insert the kernel forms T and TB in FP32 from one square
             root (kdt_walk.hip header), with the square root taken one unit
             in the last place low here (v_sqrt_f32 is within one): they are
             never below T(U(A)) and TB(U(A))
Inputs stress each: near-equal rows far from the centre (the row rounding is
the whole error), cancellation at large norms, subnormal-range values, the
1e18 limit, offset data.  Margins with a term taken out must break on them:
E2 without its N^2 term or with u N^2 in place of (D + 3) u N^2 (core), U
without E2 (upper), T without E2 (inclusion), TB without its rounding terms
(TB = s d: box).  (The row-rounding term e1 is
needed by the proof but stays below E2's N^2 term -- 2 e1 sqrt(d) <= 2 u N^2 --
so no input breaks a margin that drops e1 alone; likewise TB's e1 and (D + 3)
u terms are each covered by the outward rounding of the boxes or by the
other on these inputs, so the negative test drops both.)
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import point_dist


def round_f32(x: Fraction) -> Fraction:
    """x rounded to the nearest FP32 value, ties to even (subnormals kept)."""
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    qe = max(e, -126) - 23
    scaled = a / Fraction(2) ** qe
    n = scaled.numerator // scaled.denominator
    r = scaled - n
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return sign * n * Fraction(2) ** qe


def f32_up(x: float) -> float:
    f = np.float32(x)
    if float(f) < x:
        f = np.nextafter(f, np.float32(np.inf))
    return float(f)


def centred(x: np.ndarray, c: np.ndarray) -> np.ndarray:
    return (x - c).astype(np.float32)  # fl32(fl64(x - c))


def sumsq64(v32) -> float:
    s = 0.0
    for v in v32:
        s += float(v) * float(v)
    return s


def walk_A(q32, p32) -> Fraction:
    qn = Fraction(float(np.float32(sumsq64(q32))))
    pn = Fraction(float(np.float32(sumsq64(p32))))
    acc = round_f32(pn + qn)
    for x, y in zip(q32, p32):
        acc = round_f32(Fraction(float(np.float32(-2.0) * x)) * Fraction(float(y)) + acc)
    return acc


def frame(points: np.ndarray):
    lo, hi = points.min(axis=0), points.max(axis=0)
    c = 0.5 * lo + 0.5 * hi
    pc = 0.0
    for x in points:
        y = x - c
        pc = max(pc, f32_up(math.sqrt(float(np.dot(y, y))) * (1.0 + 2.0 ** -44)))
    return c, pc


def margins(D, q, c, pc, drop=None):
    """(E2, e1 rounded up, E2 rounded up) as the kernel forms them; `drop`
    weakens E2 for the negative test."""
    y = q - c
    Q = math.sqrt(float(sum(float(v) * float(v) for v in y))) * (1.0 + 2.0 ** -44)
    a = 2.0 ** -24 + 2.0 ** -52
    t = 2.0 * math.sqrt(D) * 2.0 ** -150
    e1 = (a * (Q + pc) + t) * (1.0 + 2.0 ** -50)
    N = ((1.0 + a) * (Q + pc) + t) * (1.0 + 2.0 ** -50)
    E2 = (D + 3.0) * 2.0 ** -24 * N * N * (1.0 + 2.0 ** -10) + (D + 4.0) * 2.0 ** -149
    if drop == "E2-N2":
        E2 = (D + 4.0) * 2.0 ** -149
    if drop == "E2-uN2":
        E2 = 2.0 ** -24 * N * N + (D + 4.0) * 2.0 ** -149
    return E2, f32_up(e1), f32_up(E2)


def shrink(D):
    return 1.0 + 2.0 * (D + 3.0) * 2.0 ** -53


def T_of(D, b, e1, e2):
    r = math.sqrt(b * shrink(D)) + e1
    return f32_up((r * r + e2) * (1.0 + 2.0 ** -50))


def U_of(D, A, e1, e2):
    r = math.sqrt(max(float(A), 0.0) + e2) + e1
    return shrink(D) * r * r * (1.0 + 2.0 ** -50)


def f32_down(x: float) -> float:
    f = np.float32(x)
    if float(f) > x:
        f = np.nextafter(f, np.float32(-np.inf))
    return float(f)


def box32(lo, hi, c):
    """k_boxes32c: fl64(lo - c) rounded down to FP32 and one step further, hi up."""
    lo32 = [float(np.nextafter(np.float32(f32_down(float(v))), np.float32(-np.inf))) for v in lo - c]
    hi32 = [float(np.nextafter(np.float32(f32_up(float(v))), np.float32(np.inf))) for v in hi - c]
    return lo32, hi32


def walk_box_dist(q32, lo32, hi32) -> Fraction:
    """The kernel's FP32 box distance: g_j = max(fma(0.5, m_j, lo32_j),
    -fma(0.5, m_j, hi32_j), 0) with m = -2 q~, then fma(g, g, acc)."""
    acc = Fraction(0)
    for x, lo, hi in zip(q32, lo32, hi32):
        m = Fraction(float(np.float32(-2.0) * x))
        g = max(round_f32(m / 2 + Fraction(lo)), -round_f32(m / 2 + Fraction(hi)), Fraction(0))
        acc = round_f32(g * g + acc)
    return acc


def TB_of(D, b, e1, factor=True):
    t = math.sqrt(b * shrink(D)) + e1
    f = 1.0 + (D + 3.0) * 2.0 ** -24 if factor else 1.0
    return f32_up((f * t * t + (D + 1.0) * 2.0 ** -149) * (1.0 + 2.0 ** -50))


def after_insert(D, A, e1, e2):
    """T and TB as the kernel forms them after an insert: FP32 operations
    (numpy float32, round to nearest), the square root one ulp low."""
    f = np.float32
    up1, up2 = f(1.0 + 2.0 ** -22), f(1.0 + 2.0 ** -21)
    x = (f(max(float(A), 0.0)) + f(e2)) * up1
    r = np.nextafter(np.sqrt(x), f(0.0)) * up2
    t = ((r + f(e1)) + f(e1)) * up2
    tt2 = t * t
    T = (tt2 + f(e2) + f(2.0 ** -147)) * up2
    TB = (tt2 * f(1.0 + (D + 3.0) * 2.0 ** -24) + f((D + 1.0) * 2.0 ** -149) + f(2.0 ** -147)) * up2
    return float(T), float(TB)


def cases():
    """(D, q, p, extra rows of the tree that set the frame)."""
    rng = np.random.default_rng(7)
    out = []
    for D in (1, 2, 3, 4, 8, 16):
        for _ in range(25):
            out.append((D, rng.standard_normal(D), rng.standard_normal(D), rng.standard_normal((4, D))))
        # near-equal rows far from the centre: the rounding of both rows is the error
        for far in (1e3, 1e6, 1e12):
            base = far + rng.standard_normal(D)
            for eps in (0.0, 1e-9, 1e-6, 1e-3):
                extra = np.stack([base - 2 * far, base])
                out.append((D, base, base + eps * rng.standard_normal(D), extra))
        # cancellation at large norms around the centre
        for scale in (1.0, 1e5, 1e17):
            base = rng.standard_normal(D) * scale
            for eps in (1e-7, 1e-4):
                out.append((D, base, base + eps * scale * rng.standard_normal(D), np.stack([-base, 3 * base])))
        out.append((D, rng.standard_normal(D) * 1e-40, rng.standard_normal(D) * 1e-39,
                    rng.standard_normal((2, D)) * 1e-39))
        out.append((D, np.full(D, 9e17), np.full(D, 9e17) * (1 + 1e-6), np.stack([np.full(D, -9e17)])))
        out.append((D, 10.0 * rng.uniform(size=D), 10.0 * rng.uniform(size=D) + 0.3 * rng.standard_normal(D),
                    10.0 * rng.uniform(size=(8, D))))
        out.append((D, 1e6 + rng.uniform(size=D), 1e6 + rng.uniform(size=D), 1e6 + rng.uniform(size=(6, D))))
    return out


def evaluate(D, q, p, extra, drop=None, with_frame=False):
    pts = np.vstack([p[None, :], extra])
    c, pc = frame(np.vstack([pts, q[None, :]]))  # the query inside the root box, as in most calls
    q32, p32 = centred(q, c), centred(p, c)
    A = walk_A(q32, p32)
    exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(q32, p32))
    E2, e1f, e2f = margins(D, q, c, pc, drop)
    d = point_dist(q, p)
    if with_frame:
        return A, exact, E2, e1f, e2f, d, c, q32
    return A, exact, E2, e1f, e2f, d


def boxes_around(rng, p, q, D):
    """Boxes holding p: from p and a corner nearby, and reaching towards q."""
    out = []
    for _ in range(2):
        r = p + rng.standard_normal(D) * (1 + np.abs(p)) * 1e-3
        out.append((np.minimum(p, r), np.maximum(p, r)))
    r = p + 0.999 * (q - p)
    out.append((np.minimum(p, r), np.maximum(p, r)))
    # p the box's corner nearest to q: the exact box distance is |q - p|^2
    ext = np.abs(rng.standard_normal(D)) * (1 + np.abs(p))
    away = np.where(q < p, p + ext, p - ext)
    out.append((np.minimum(p, away), np.maximum(p, away)))
    return out


def test_walk_margin_holds():
    rng = np.random.default_rng(11)
    for D, q, p, extra in cases():
        A, exact, E2, e1f, e2f, d, c, q32 = evaluate(D, q, p, extra, with_frame=True)
        assert abs(A - exact) <= Fraction(E2), (D, float(A), float(exact), E2)
        T = T_of(D, d, e1f, e2f)
        assert A <= Fraction(T), (D, float(A), d, T)
        U = U_of(D, A, e1f, e2f)
        assert d <= U, (D, d, U, float(A))
        Ti, TBi = after_insert(D, A, e1f, e2f)
        assert Ti >= T_of(D, U, e1f, e2f) and TBi >= TB_of(D, U, e1f), (D, float(A))
        for lo, hi in boxes_around(rng, p, q, D):
            lo32, hi32 = box32(lo, hi, c)
            bd32 = walk_box_dist(q32, lo32, hi32)
            assert bd32 <= Fraction(TB_of(D, d, e1f)), (D, float(bd32), d)


def test_weakened_box_threshold_fails():
    rng = np.random.default_rng(12)
    broken = 0
    for D, q, p, extra in cases():
        A, exact, E2, e1f, e2f, d, c, q32 = evaluate(D, q, p, extra, with_frame=True)
        for lo, hi in boxes_around(rng, p, q, D):
            lo32, hi32 = box32(lo, hi, c)
            bd32 = walk_box_dist(q32, lo32, hi32)
            broken += bd32 > Fraction(shrink(D) * d)
    assert broken > 0


def test_update_thresholds_cover_every_magnitude():
    """T' and TB' after an insert (FP32, sqrt one ulp low) against T(U(A)) and
    TB(U(A)) over A, e1 and E2 from the subnormal range to 10^36."""
    rng = np.random.default_rng(13)
    for D in (1, 2, 3, 4, 8, 16):
        for _ in range(400):
            A = float(np.float32(10.0 ** rng.uniform(-44, 36)))
            e1 = f32_up(10.0 ** rng.uniform(-44, 17))
            e2 = f32_up(10.0 ** rng.uniform(-44, 36))
            U = U_of(D, A, e1, e2)
            if not math.isfinite(U) or U > 1e37:
                continue
            Ti, TBi = after_insert(D, A, e1, e2)
            assert Ti >= T_of(D, U, e1, e2) and TBi >= TB_of(D, U, e1), (D, A, e1, e2)


@pytest.mark.parametrize("drop", ["E2-N2", "E2-uN2", "U-no-E2", "T-no-E2"])
def test_weakened_walk_margin_fails(drop):
    broken = 0
    for D, q, p, extra in cases():
        if drop.startswith("E2"):
            A, exact, E2, _, _, _ = evaluate(D, q, p, extra, drop)
            broken += abs(A - exact) > Fraction(E2)
        else:
            A, exact, E2, e1f, e2f, d = evaluate(D, q, p, extra)
            if drop == "U-no-E2":
                broken += d > U_of(D, A, e1f, 0.0)
            else:
                broken += A > Fraction(T_of(D, d, e1f, 0.0))
    assert broken > 0
