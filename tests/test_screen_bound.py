"""The matrix-core screen's error bound, checked against exact arithmetic.

The screen (kdt_screen.hip) rounds the rows to FP32 (q~, p~) and their squared
norms to FP32 (qn, pn), splits each value x into bf16 halves xh = bf16(x),
xl = bf16(x - xh), and has the matrix cores sum in FP32 the 3D exact products
of -2 q~ and p~ (xh.yh + xl.yh + xh.yl per coordinate) and the four exact terms
qn_h + qn_l + pn_h + pn_l: that sum is A.  It relies on two consequences of
its bound, with N = |q~| + |p~|, u = 2^-24, s = 1 + 2(D+3)2^-53,
g = 3.1 2^-16 + 1.03 (3D+4)u / (1 - (3D+4)u):

  (upper)     d <= U(A) = s (sqrt(max(A, 0) + E2) + e1)^2
  (inclusion) A <= T(d) = (sqrt(s d) + e1)^2 + E2
  E2 = (2g + 8u) N^2 + D 2^-110 (1 + N)^2,   e1 = u (1 + 2u) N + sqrt(D) 2^-148

where d is the reference's FP64 reduced distance (the oracle's point_dist, the
Streamed / Folded summation of src/kernel.rs).  (upper) makes tau an upper bound
of the k-th distance; (inclusion) puts every top-k point among the candidates.
Here A is formed as the kernel forms it (bf16 rounding emulated bit for bit,
round to nearest even), with the 3D + 4 terms accumulated in FP32 in several
orders (the matrix cores' internal order is not specified), on inputs chosen to
stress the bound: near-equal rows with large norms (cancellation in
|q|^2 + |p|^2 - 2 q.p), rows whose split residuals all add up, mixed
magnitudes, values near FP32 underflow and near the 1e18 limit of the screen.
The exact |q~ - p~|^2 (rationals) checks the core inequality
|A - |q~ - p~|^2| <= E2 itself, and bounds without the cancellation term or
without the split term must each fail on these inputs, so the test can tell a
wrong bound.
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


def bf16(x) -> np.float32:
    """Round a float32 to bfloat16 (nearest, ties to even), as float32."""
    b = int(np.array([x], dtype=np.float32).view(np.uint32)[0])
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return np.array([r & 0xFFFFFFFF], dtype=np.uint32).view(np.float32)[0]


def split_terms(a, b):
    """The 3 D exact FP32 products of the bf16 split, per coordinate."""
    out = []
    for x, y in zip(a, b):
        xh, yh = bf16(x), bf16(y)
        xl, yl = bf16(f32(x - xh)), bf16(f32(y - yh))
        out += [f32(float(xh) * float(yh)), f32(float(xl) * float(yh)), f32(float(xh) * float(yl))]
    return out


def fsum32(v, order):
    """FP32 sum of float32 terms in a given order."""
    if order == "seq":
        acc = f32(0.0)
        for t in v:
            acc = f32(acc + t)
        return acc
    if order == "fma":  # one rounding per step of a wider sum
        acc = 0.0
        for t in v:
            acc = float(f32(acc + float(t)))
        return f32(acc)
    if order == "pairs":
        acc = f32(0.0)
        for i in range(0, len(v), 2):
            t = f32(v[i] + v[i + 1]) if i + 1 < len(v) else v[i]
            acc = f32(acc + t)
        return acc
    if order == "tree":
        while len(v) > 1:
            v = [f32(v[i] + v[i + 1]) if i + 1 < len(v) else v[i] for i in range(0, len(v), 2)]
        return v[0]
    raise ValueError(order)


def screen_A(q32, p32, order):
    qn = f32(float(np.sum(q32.astype(np.float64) ** 2)))
    pn = f32(float(np.sum(p32.astype(np.float64) ** 2)))
    norms = []
    for x in (qn, pn):  # the norm block: x_h * 1 + x_l * 1
        xh = bf16(x)
        norms += [xh, bf16(f32(x - xh))]
    prods = split_terms((np.float32(-2.0) * q32).astype(np.float32), p32)
    if order == "norms-first":
        return float(fsum32(norms + prods, "seq"))
    return float(fsum32(prods + norms, order))  # the kernel issues the norm block last


def bounds(D, nq, npn, cancellation=True, split=True):
    m3 = (3.0 * D + 4.0) * U
    g = (3.1 * 2.0 ** -16 if split else 0.0) + 1.03 * m3 / (1.0 - m3)
    N = nq * (1.0 + 2.0 ** -50) + npn
    E2 = ((2.0 * g + 8.0 * U) * N * N if cancellation else 8.0 * U * max(nq, npn) ** 2)
    E2 += D * 2.0 ** -110 * (1.0 + N) ** 2
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
        # the FP32 value whose split contraction x.x is worst (relative error
        # 1.99 2^-16, found by exhausting [1, 2)), in every coordinate, so that
        # the split errors add up; as duplicates (A ~ 0) and opposite rows
        x = float(np.array([0x3F808040], dtype=np.uint32).view(np.float32)[0])
        for scale in (1.0, 2.0 ** -40, 2.0 ** 50):
            out.append((D, np.full(D, x * scale), np.full(D, x * scale)))
            out.append((D, np.full(D, x * scale), np.full(D, -x * scale)))
    return out


@pytest.mark.parametrize("order", ["seq", "fma", "pairs", "tree", "norms-first"])
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


@pytest.mark.parametrize("weakened", ["no-cancellation-term", "no-split-term"])
def test_weakened_bound_fails(weakened):
    """Dropping the 2 g N^2 term (the error of q.p under cancellation), or only
    its bf16 split part, must break the core inequality on these inputs."""
    broken = 0
    for D, q, p in cases():
        q32, p32 = q.astype(np.float32), p.astype(np.float32)
        A = screen_A(q32, p32, "seq")
        nq = float(np.linalg.norm(q32.astype(np.float64)))
        npn = float(np.linalg.norm(p32.astype(np.float64)))
        E2, _, _ = bounds(D, nq, npn, cancellation=weakened != "no-cancellation-term",
                          split=weakened != "no-split-term")
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(q32, p32))
        broken += abs(Fraction(A) - exact) > Fraction(E2)
    assert broken > 0
