"""The traversal's FP32 leaf screen and its margin, checked in exact arithmetic.

The screen (kdt_query_kernels.h, screen_score32 / screen_t32) scores a tree
point p for query q as
    A = fma(m_{D-1}, p~_{D-1}, ... fma(m_0, p~_0, fl32(pn + qn)))
with q~, p~ the rows rounded to FP32, m = -2 q~, qn and pn the FP32 roundings
of the FP64 sums of squares (pn from the build's k_norms32), and rejects the
point when A > T(b) = fl32_up(((sqrt(s b) + e1')^2 + E2') (1 + 2^-50)) (FP64
arithmetic), e1' = fl32_up(e1), E2' = fl32_up(E2),
    E2 = (D + 3) u N^2 (1 + 2^-10) + (D + 4) 2^-149,  e1 = u (1 + 2u) N + sqrt(D) 2^-148,
    N = sqrt(qn64) (1 + 2^-45) + P,  s = 1 + 2 (D + 3) 2^-53,  u = 2^-24,
P the tree's bound on |p~| (k_norms32: the largest |p~| rounded up, times 1 + 1e-6).

Here every FP32 operation is emulated exactly (rationals, round to nearest
even at FP32 precision, subnormals included), and the checks are the two
facts the screen relies on:
  core       |A - |q~ - p~|^2| <= E2
  inclusion  A <= T(d), d the reference's FP64 reduced distance (the oracle's
             point_dist, src/kernel.rs recipes): a point the reference offers
             (d <= bound) is never rejected, since T grows with the bound.
Inputs stress the margin: near-equal rows with large norms (cancellation in
|p|^2 + |q|^2 - 2 q.p), duplicates, values near FP32 underflow, the 1e18
limit, mixed magnitudes.  A margin without the N^2 term, or with u N^2 in
place of (D + 3) u N^2, must break the core inequality on these inputs.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import point_dist

U = 2.0 ** -24


def round_f32(x: Fraction) -> Fraction:
    """x rounded to the nearest FP32 value, ties to even (subnormals kept; no
    overflow in these inputs)."""
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    qe = max(e, -126) - 23  # exponent of one unit in the last place
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


def sumsq64(v32) -> float:
    s = 0.0
    for x in v32:
        s += float(x) * float(x)  # exact products (FP32 inputs), FP64 sums in order
    return s


def screen_A(q32, p32) -> Fraction:
    qn = Fraction(float(np.float32(sumsq64(q32))))
    pn = Fraction(float(np.float32(sumsq64(p32))))
    acc = round_f32(pn + qn)
    for x, y in zip(q32, p32):
        acc = round_f32(Fraction(float(np.float32(-2.0) * x)) * Fraction(float(y)) + acc)
    return acc


def margin(D, q32, p32, n2_term=True, factor=None):
    qn64 = sumsq64(q32)
    P = f32_up(math.sqrt(sumsq64(p32))) * (1.0 + 1e-6)
    N = math.sqrt(qn64) * (1.0 + 2.0 ** -45) + P
    c = (D + 3.0) if factor is None else factor
    E2 = (c * U * N * N * (1.0 + 2.0 ** -10) if n2_term else 0.0) + (D + 4.0) * 2.0 ** -149
    e1 = U * (1.0 + 2.0 * U) * N + math.sqrt(D) * 2.0 ** -148
    return E2, f32_up(e1), f32_up(E2)


def threshold(D: int, b: float, e1: float, e2: float) -> float:
    """T(b) as the kernel forms it (FP64 operations, then rounded up to FP32)."""
    s = 1.0 + 2.0 * (D + 3.0) * 2.0 ** -53
    r = math.sqrt(b * s) + e1
    return f32_up((r * r + e2) * (1.0 + 2.0 ** -50))


def cases():
    rng = np.random.default_rng(2026)
    out = []
    for D in (1, 2, 3, 4, 8, 9, 16):
        for _ in range(30):
            out.append((D, rng.standard_normal(D), rng.standard_normal(D)))
        for scale in (1.0, 1e3, 1e9, 1e17):  # near-equal rows, large norms: cancellation
            base = rng.standard_normal(D) * scale
            for eps in (0.0, 1e-7, 1e-5, 1e-3):
                out.append((D, base, base + eps * scale * rng.standard_normal(D)))
        out.append((D, rng.standard_normal(D) * np.logspace(-30, 10, D), rng.standard_normal(D)))
        out.append((D, rng.standard_normal(D) * 1e-38, rng.standard_normal(D) * 1e-39))
        out.append((D, np.full(D, 9.9e17), np.full(D, 9.9e17) * (1 + 1e-7)))
        out.append((D, 10.0 * rng.uniform(size=D), 10.0 * rng.uniform(size=D) + 0.3 * rng.standard_normal(D)))
    return out


def test_fp32_screen_margin_holds():
    for D, q, p in cases():
        q32, p32 = q.astype(np.float32), p.astype(np.float32)
        A = screen_A(q32, p32)
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(q32, p32))
        E2, e1f, e2f = margin(D, q32, p32)
        assert abs(A - exact) <= Fraction(E2), (D, float(A), float(exact), E2)
        d = point_dist(q, p)  # the reference's FP64 reduced distance
        T = threshold(D, d, e1f, e2f)
        assert math.isinf(T) or A <= Fraction(T), (D, float(A), d, T)


@pytest.mark.parametrize("weakened", ["no-N2-term", "u-N2-only"])
def test_weakened_fp32_margin_fails(weakened):
    broken = 0
    for D, q, p in cases():
        q32, p32 = q.astype(np.float32), p.astype(np.float32)
        A = screen_A(q32, p32)
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(q32, p32))
        E2, _, _ = margin(D, q32, p32, n2_term=weakened != "no-N2-term",
                          factor=0.5 if weakened == "u-N2-only" else None)
        broken += abs(A - exact) > Fraction(E2)
    assert broken > 0
