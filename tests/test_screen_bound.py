"""The matrix-core screen's FP32 error bound, checked against exact arithmetic.

The screen (kdt_screen.hip) forms A = fl32(|q~|^2 + |p~|^2 - 2 q~.p~) from the
rows rounded to FP32 and relies on two consequences of its bound, with
N = |q~| + |p~|, u = 2^-24, g = (2D+4)u / (1 - (2D+4)u), s = 1 + 2(D+3)2^-53:

  (upper)     d <= U(A) = s (sqrt(max(A, 0) + E2) + e1)^2
  (inclusion) A <= T(d) = (sqrt(s d) + e1)^2 + E2
  E2 = (2g + 8u) N^2 + D 2^-125,   e1 = u (1 + 2u) N + sqrt(D) 2^-148

where d is the reference's FP64 reduced distance (the oracle's point_dist, the
Streamed / Folded summation of src/kernel.rs).  (upper) makes tau an upper bound
of the k-th distance; (inclusion) puts every top-k point among the candidates.
Here A is formed as the kernel forms it, with the q.p contraction accumulated in
FP32 in several orders (the matrix cores' internal order is not specified), on
inputs chosen to stress the bound: near-equal rows with large norms (cancellation
in |q|^2 + |p|^2 - 2 q.p), mixed magnitudes, values near FP32 underflow and near
the 1e18 limit of the screen.  The exact |q~ - p~|^2 (rationals) checks the core
inequality |A - |q~ - p~|^2| <= E2 itself, and a bound without the cancellation
term must fail on the same inputs, so the test can tell a wrong bound.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import point_dist

U = 2.0 ** -24


def f32(x):
    return np.float32(x)


def dot32(a, b, order):
    """FP32 contraction of two float32 vectors in a given accumulation order."""
    if order == "seq":
        acc = f32(0.0)
        for x, y in zip(a, b):
            acc = f32(acc + f32(x * y))
        return acc
    if order == "fma":
        acc = 0.0
        for x, y in zip(a, b):
            acc = float(f32(float(acc) + float(x) * float(y)))  # one rounding per step
        return f32(acc)
    if order == "pairs":  # two products per step, as a K = 2 matrix instruction could
        acc = f32(0.0)
        for i in range(0, len(a), 2):
            t = f32(f32(a[i] * b[i]) + f32(a[i + 1] * b[i + 1])) if i + 1 < len(a) else f32(a[i] * b[i])
            acc = f32(acc + t)
        return acc
    if order == "tree":
        v = [f32(x * y) for x, y in zip(a, b)]
        while len(v) > 1:
            v = [f32(v[i] + v[i + 1]) if i + 1 < len(v) else v[i] for i in range(0, len(v), 2)]
        return v[0]
    raise ValueError(order)


def screen_A(q32, p32, order):
    qn = f32(float(np.sum(q32.astype(np.float64) ** 2)))
    pn = f32(float(np.sum(p32.astype(np.float64) ** 2)))
    dot = dot32(q32, p32, order)
    return float(f32(-2.0 * float(dot) + float(f32(qn + pn))))  # fmaf: one rounding of the exact value


def bounds(D, nq, npn, cancellation=True):
    m2 = (2.0 * D + 4.0) * U
    g = m2 / (1.0 - m2)
    N = nq * (1.0 + 2.0 ** -50) + npn
    E2 = ((2.0 * g + 8.0 * U) * N * N if cancellation else 8.0 * U * max(nq, npn) ** 2) + D * 2.0 ** -125
    e1 = U * (1.0 + 2.0 * U) * N + math.sqrt(D) * 2.0 ** -148
    s = 1.0 + 2.0 * (D + 3.0) * 2.0 ** -53
    return E2 * (1 + 1e-12), e1 * (1 + 1e-12), s


def cases():
    rng = np.random.default_rng(2024)
    out = []
    for D in (16, 33, 64):
        for _ in range(40):  # ordinary Gaussian rows
            out.append((D, rng.standard_normal(D), rng.standard_normal(D)))
        for scale in (1.0, 1e3, 1e9, 1e17):  # near-equal rows, large norms: cancellation
            base = rng.standard_normal(D) * scale
            for eps in (1e-7, 1e-5, 1e-3):
                out.append((D, base, base + eps * scale * rng.standard_normal(D)))
        # mixed magnitudes, values near FP32 underflow, the 1e18 screen limit
        out.append((D, rng.standard_normal(D) * np.logspace(-30, 10, D), rng.standard_normal(D)))
        out.append((D, rng.standard_normal(D) * 1e-38, rng.standard_normal(D) * 1e-39))
        out.append((D, np.full(D, 9.9e17), np.full(D, 9.9e17) * (1 + 1e-7)))
    return out


@pytest.mark.parametrize("order", ["seq", "fma", "pairs", "tree"])
def test_screen_bound_holds(order):
    for D, q, p in cases():
        q32, p32 = q.astype(np.float32), p.astype(np.float32)
        A = screen_A(q32, p32, order)
        nq = float(np.linalg.norm(q32.astype(np.float64)))
        npn = float(np.linalg.norm(p32.astype(np.float64)))
        E2, e1, s = bounds(D, nq, npn)
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(q32, p32))
        assert abs(Fraction(A) - exact) <= Fraction(E2), (D, order)
        d = point_dist(q, p)  # the reference's FP64 reduced distance
        upper = s * (math.sqrt(max(A, 0.0) + E2) + e1) ** 2 * (1 + 1e-12)
        assert d <= upper, (D, order, d, upper)
        incl = ((math.sqrt(s * d) + e1) ** 2 + E2) * (1 + 1e-12)
        assert A <= incl, (D, order, A, incl)


def test_bound_without_cancellation_term_fails():
    """Dropping the 2 g N^2 term (the FP32 error of q.p under cancellation) must
    break the core inequality on near-equal large-norm rows."""
    broken = 0
    for D, q, p in cases():
        q32, p32 = q.astype(np.float32), p.astype(np.float32)
        A = screen_A(q32, p32, "seq")
        nq = float(np.linalg.norm(q32.astype(np.float64)))
        npn = float(np.linalg.norm(p32.astype(np.float64)))
        E2, _, _ = bounds(D, nq, npn, cancellation=False)
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(q32, p32))
        broken += abs(Fraction(A) - exact) > Fraction(E2)
    assert broken > 0
