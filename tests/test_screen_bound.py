"""The matrix-core screen's error bound, checked against exact arithmetic.

The screen (kdt_screen.hip) moves every row into a frame fixed per tree, centre
c (the middle of the root box) and a power of two sigma with sigma |x_j - c_j|
< 2^8 over the tree, and rounds it to fp16: x^_j = fp16(fl32(sigma fl64(x_j -
c_j))), nearest even, fp16 subnormals flushed to zero.  The matrix cores sum in
FP32 the D exact products (-2 q^_j) p^_j and the norms qn = fl32(|q^|^2), pn =
fl32(|p^|^2), each as three bf16 pieces (p1 = bf16(pn), p2 = bf16(pn - p1),
p3 = bf16(pn - p1 - p2)): that sum is A.  With a = 2^-11 (1 + 2^-12),
b = 2^-14, P_c >= |p - c| over the tree, s = 1 + 2(D+3)2^-53:

  per coordinate   |sigma (x_j - c_j) - x^_j| <= a sigma |x_j - c_j| + b
  e1 = a sigma (|q - c| + P_c) + 2 b sqrt(D)
  Nh = sigma (|q - c| + P_c)(1 + a) + 2 b sqrt(D)
  E2 = ((1 + 2^-7) gamma + 1.13 2^-24) Nh^2 + 2^-118,
       gamma = (D + 5) 2^-23 / (1 - (D + 5) 2^-23)
  (core)       |A - |q^ - p^|^2| <= E2
  (upper)      d <= U(A) = s (sqrt(max(A, 0) + E2) + e1)^2 / sigma^2
  (inclusion)  A <= T(d) = (sigma sqrt(s d) + e1)^2 + E2

where d is the reference's FP64 reduced distance (the oracle's point_dist).
(upper) makes tau an upper bound of the k-th distance; (inclusion) puts every
top-k point among the candidates.  Here the rounding is emulated bit for bit
(numpy's float64 -> float32 -> float16 conversions round to nearest even) and
A is accumulated in FP32 in several orders, and once rounding every addition
toward zero (the bound allows 2^-23 per addition; the matrix cores' internal
order and rounding are not specified).  The inputs stress each term: rows far
from the centre but close to each other (the fp16 rounding of both rows is the
whole distance error), near-equal rows with large norms (cancellation in
|q|^2 + |p|^2 - 2 q.p), coordinates below the fp16 normal range (flushed),
mixed magnitudes, offset data (the centre removes the offset), huge values
(no range limit on the data: sigma scales them) and queries near the 2^11
limit.  Bounds without the fp16 term or without the accumulation term must
each fail on these inputs, so the test can tell a wrong bound (the flush term
is dominated by the fp16 term in every frame; a test shows why).
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import point_dist

A_REL = 2.0 ** -11 * (1 + 2.0 ** -12)
B_ABS = 2.0 ** -14 * (1 + 2.0 ** -60)


def bf16(x) -> np.float32:
    """Round a float32 to bfloat16 (nearest, ties to even), as float32."""
    b = int(np.array([x], dtype=np.float32).view(np.uint32)[0])
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return np.array([r & 0xFFFFFFFF], dtype=np.uint32).view(np.float32)[0]


def frame(tree_rows: np.ndarray):
    """Centre, sigma and P_c as the build makes them (k_screen_frame, k_rows_h)."""
    lo, hi = tree_rows.min(axis=0), tree_rows.max(axis=0)
    c = 0.5 * lo + 0.5 * hi
    m = float(np.max(np.maximum(np.abs(hi - c), np.abs(lo - c))))
    s = min(8 - (math.frexp(m)[1]), 1000) if m > 0 else 0  # frexp exponent = ilogb + 1
    sigma = math.ldexp(1.0, s)
    pc = max(math.sqrt(float(np.sum((r - c) ** 2))) for r in tree_rows) * (1 + 2.0 ** -44)
    return c, sigma, pc


def f16_row(x: np.ndarray, c: np.ndarray, sigma: float) -> np.ndarray:
    """screen_f16 of every coordinate: fp16(fl32(sigma fl64(x - c))), clamped
    to +-2^11, fp16 subnormals flushed; as float64 values."""
    v = np.clip(((x - c) * sigma).astype(np.float32), -2048.0, 2048.0)
    h = v.astype(np.float16).astype(np.float64)
    h[np.abs(h) < 2.0 ** -14] = 0.0
    return h


def pieces(h: np.ndarray):
    pn = np.float32(float(np.sum(h * h)))
    p1 = bf16(pn)
    r1 = np.float32(pn - p1)
    p2 = bf16(r1)
    p3 = bf16(np.float32(r1 - p2))
    return [p1, p2, p3]


def rtz32(x: Fraction) -> np.float32:
    """x rounded toward zero to float32."""
    f = np.float32(float(x))
    zero = np.float32(0.0)
    while abs(Fraction(float(f))) > abs(x):
        f = np.nextafter(f, zero)
    away = np.float32(np.inf) if x > 0 else np.float32(-np.inf)
    while True:
        g = np.nextafter(f, away)
        if abs(Fraction(float(g))) <= abs(x):
            f = g
        else:
            return f


def fsum32(v, order):
    """FP32 sum of float32 terms in a given order."""
    if order == "seq":
        acc = np.float32(0.0)
        for t in v:
            acc = np.float32(acc + t)
        return acc
    if order == "rtz":
        acc = np.float32(0.0)
        for t in v:
            acc = rtz32(Fraction(float(acc)) + Fraction(float(t)))
        return acc
    if order == "pairs":
        acc = np.float32(0.0)
        for i in range(0, len(v), 2):
            t = np.float32(v[i] + v[i + 1]) if i + 1 < len(v) else v[i]
            acc = np.float32(acc + t)
        return acc
    if order == "tree":
        while len(v) > 1:
            v = [np.float32(v[i] + v[i + 1]) if i + 1 < len(v) else v[i] for i in range(0, len(v), 2)]
        return v[0]
    raise ValueError(order)


def screen_A(hq, hp, order):
    prods = [np.float32((-2.0 * a) * b) for a, b in zip(hq, hp)]  # exact: 22-bit products
    assert all(float(t) == (-2.0 * a) * b for t, a, b in zip(prods, hq, hp))
    norms = pieces(hq) + pieces(hp)
    if order == "norms-first":
        return float(fsum32(norms + prods, "seq"))
    return float(fsum32(prods + norms, order))  # the kernel issues the norm block last


def bounds(D, qc, pc, sigma, fp16_term=True, flush_term=True, accumulation=True):
    a = A_REL if fp16_term else 2.0 ** -24
    b = B_ABS if flush_term else 0.0
    eb = b * 2.0 * math.sqrt(D) * (1 + 2.0 ** -50)
    e1 = (a * sigma * (qc + pc) + eb) * (1 + 1e-12)
    nh = (sigma * (qc + pc) * (1 + a) + eb) * (1 + 1e-12)
    m = (D + 5.0) * 2.0 ** -23
    gam = m / (1 - m) if accumulation else 0.0
    E2 = (((1 + 2.0 ** -7) * gam + 1.13 * 2.0 ** -24) * nh * nh + 2.0 ** -118) * (1 + 1e-12)
    s = 1.0 + 2.0 * (D + 3.0) * 2.0 ** -53
    return E2, e1, s


def cases():
    """(D, tree rows, q, p): p is a row of the tree; q any query."""
    rng = np.random.default_rng(2024)
    out = []
    for D in (16, 33, 64):
        for _ in range(30):  # ordinary Gaussian rows
            t = rng.standard_normal((4, D))
            out.append((D, t, rng.standard_normal(D), t[0]))
        for scale in (1.0, 1e3, 1e17, 1e150):
            # rows far from the centre, close to each other: the fp16 rounding of
            # both rows is the distance error (and not cancelled by the centre)
            base = rng.standard_normal(D) * scale
            t = np.stack([base, -base, base * 0.5, -base * 0.5])
            for eps in (1e-7, 1e-4, 1e-2):
                out.append((D, t, base + eps * scale * rng.standard_normal(D), base))
            # near-equal rows with large norms, data offset from the origin
            t = base + eps * scale * rng.standard_normal((4, D))
            out.append((D, t, t[1] + 1e-3 * eps * scale, t[0]))
        # coordinates below the fp16 normal range after scaling (flushed)
        t = rng.standard_normal((4, D)) * np.logspace(-12, 0, D)
        out.append((D, t, rng.standard_normal(D) * np.logspace(-12, 0, D), t[2]))
        tiny = np.zeros((4, D))
        tiny[0, 0] = 1.0  # sets sigma; every other coordinate flushes
        tiny[1:, 1:] = rng.standard_normal((3, D - 1)) * 2.0 ** -24
        out.append((D, tiny, np.full(D, 2.0 ** -25), tiny[1]))
        # mixed magnitudes, huge values, queries near the 2^11 limit
        t = rng.standard_normal((4, D)) * np.logspace(-30, 10, D)
        out.append((D, t, rng.standard_normal(D) * np.logspace(-30, 10, D), t[3]))
        t = rng.standard_normal((4, D)) * 1e150  # (squared distances near 1e302)
        out.append((D, t, rng.standard_normal(D) * 1e150, t[0]))
        t = rng.uniform(-1.0, 1.0, (4, D))
        out.append((D, t, np.full(D, 7.9), t[1]))  # sigma = 2^8: 7.9 sigma = 2022 < 2^11
        # duplicates and exact opposites
        t = rng.standard_normal((4, D))
        out.append((D, t, t[0].copy(), t[0]))
        out.append((D, t, -t[0], t[0]))
    return out


def check_case(D, tree_rows, q, p, order, **weak):
    c, sigma, pc = frame(tree_rows)
    ymax = float(np.max(np.abs(q - c))) * sigma
    if not ymax <= 2048.0:
        return None  # the select flags this query: no bound is claimed
    hq, hp = f16_row(q, c, sigma), f16_row(p, c, sigma)
    qc = math.sqrt(float(np.sum((q - c) ** 2))) * (1 + 2.0 ** -44)
    E2, e1, s = bounds(D, qc, pc, sigma, **weak)
    return hq, hp, c, sigma, qc, pc, E2, e1, s


@pytest.mark.parametrize("order", ["seq", "rtz", "pairs", "tree", "norms-first"])
def test_screen_bound_holds(order):
    checked = 0
    for D, tree_rows, q, p in cases():
        r = check_case(D, tree_rows, q, p, order)
        if r is None:
            continue
        hq, hp, c, sigma, qc, pc, E2, e1, s = r
        # per coordinate, exactly: |sigma (x - c) - x^| <= a sigma |x - c| + b
        for x, h in ((q, hq), (p, hp)):
            for xj, cj, hj in zip(x, c, h):
                y = (Fraction(float(xj)) - Fraction(float(cj))) * Fraction(sigma)
                assert abs(y - Fraction(float(hj))) <= Fraction(A_REL) * abs(y) + Fraction(B_ABS)
        A = screen_A(hq, hp, order)
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(hq, hp))
        assert abs(Fraction(A) - exact) <= Fraction(E2), (D, order)
        d = point_dist(q, p)  # the reference's FP64 reduced distance
        upper = s * ((math.sqrt(max(A, 0.0) + E2) + e1) / sigma) ** 2 * (1 + 1e-12)
        assert d <= upper, (D, order, d, upper)
        incl = ((sigma * math.sqrt(s * d) + e1) ** 2 + E2) * (1 + 1e-12)
        assert A <= incl, (D, order, A, incl)
        checked += 1
    assert checked >= 150


def test_far_queries_are_not_screened():
    # rows spanning [-1, 1] in every coordinate: M = 1, sigma = 2^7, so a query
    # coordinate at 16.1 maps beyond 2^11 and one at 15.9 does not
    rng = np.random.default_rng(7)
    t = rng.uniform(-1.0, 1.0, (4, 16))
    t[0], t[1] = -1.0, 1.0
    assert frame(t)[1] == 2.0 ** 7
    assert check_case(16, t, np.full(16, 16.1), t[2], "seq") is None
    assert check_case(16, t, np.full(16, 15.9), t[2], "seq") is not None


def test_flush_term_is_dominated():
    """The flushed-coordinate term 2 b sqrt(D) of e1 never exceeds the fp16 term
    (for data spanning more than 2^-990, where sigma is not capped at 2^1000):
    sigma makes the tree's largest |x_j - c_j| at least 2^7 / sigma, so
    a sigma P_c >= 2^-11 2^7 = 2^-4 > 2 b sqrt(64) = 2^-10.  (It stays in the
    bound, which is derived without this property; no input can make a bound
    without it fail.)"""
    for D, tree_rows, q, p in cases():
        c, sigma, pc = frame(tree_rows)
        assert sigma * pc >= 2.0 ** 7 or pc == 0.0
        assert A_REL * sigma * pc >= 2 * B_ABS * math.sqrt(D) or pc == 0.0


@pytest.mark.parametrize("weakened", ["fp16_term", "accumulation"])
def test_weakened_bound_fails(weakened):
    """A bound that treats the rows as exact FP32 values (no fp16 term), or that
    leaves out the FP32 accumulation, must break (core), (upper) or
    (inclusion) on these inputs."""
    broken = 0
    for D, tree_rows, q, p in cases():
        r = check_case(D, tree_rows, q, p, "seq", **{weakened: False})
        if r is None:
            continue
        hq, hp, c, sigma, qc, pc, E2, e1, s = r
        A = screen_A(hq, hp, "seq")
        exact = sum((Fraction(float(x)) - Fraction(float(y))) ** 2 for x, y in zip(hq, hp))
        d = point_dist(q, p)
        incl = ((sigma * math.sqrt(s * d) + e1) ** 2 + E2) * (1 + 1e-12)
        upper = s * ((math.sqrt(max(A, 0.0) + E2) + e1) / sigma) ** 2 * (1 + 1e-12)
        broken += abs(Fraction(A) - exact) > Fraction(E2) or A > incl or d > upper
    assert broken > 0


# --- the group tests' integer minima ---------------------------------------
#
# k_screen tests a group of four values against a lane's threshold thr by the
# signed-integer minimum of their bit patterns against t = max(bits(thr), 0)
# (an integer minimum needs no canonicalising instruction before it).  The
# group test may pass more than needed, never less: A < thr must imply
# bits(A) < t for every float A the accumulator can hold (any sign, zero,
# subnormal, infinite; matrix results are never NaN) and every threshold.


def _bits(x: np.ndarray) -> np.ndarray:
    return np.asarray(x, np.float32).view(np.int32).astype(np.int64)


def _special_floats() -> np.ndarray:
    f = np.float32
    base = [0.0, -0.0, np.inf, -np.inf, 1.0, -1.0, 2.0**-149, -(2.0**-149), 2.0**-126, -(2.0**-126),
            np.finfo(np.float32).max, -np.finfo(np.float32).max, 1e-30, -1e-30, 3.5, -3.5]
    rng = np.random.default_rng(77)
    mags = 2.0 ** rng.uniform(-149, 127, size=400)
    vals = np.concatenate([np.array(base, f), (mags * rng.choice([-1.0, 1.0], size=400)).astype(f)])
    # neighbours one ulp apart, so that near-equal pairs are covered
    with np.errstate(over="ignore"):  # the largest finite values step to +-inf, which are in the set anyway
        return np.unique(np.concatenate([vals, np.nextafter(vals, f(np.inf)), np.nextafter(vals, f(-np.inf))]))


def test_integer_group_minimum_never_misses():
    v = _special_floats()
    A, thr = np.meshgrid(v, v, indexing="ij")
    A, thr = A.ravel(), thr.ravel()
    t = np.maximum(_bits(thr), 0)
    under = A < thr
    assert under.any()
    assert np.all(_bits(A)[under] < t[under])
    # and the four-value form: min of the bit patterns under t whenever any value is under thr
    rng = np.random.default_rng(78)
    g = rng.choice(v, size=(20000, 4))
    th = rng.choice(v, size=20000)
    anyu = (g < th[:, None]).any(axis=1)
    assert np.all(_bits(g).min(axis=1)[anyu] < np.maximum(_bits(th), 0)[anyu])


def test_integer_group_minimum_needs_the_clamp():
    """Without the clamp at 0 (t = bits(thr)), a negative threshold misses
    values below it: the bit patterns of negative floats order in reverse."""
    v = _special_floats()
    A, thr = np.meshgrid(v, v, indexing="ij")
    A, thr = A.ravel(), thr.ravel()
    under = A < thr
    assert not np.all(_bits(A)[under] < _bits(thr)[under])


# The main pass lowers a lane's threshold at each merge to a bound of tau_A
# from its list's k-th entry ak (an upper bound of A_k), computed in FP32
# (screen_tau_a32 in kdt_screen.hip) from the query's e1 and E2 rounded up to
# float (formed in FP64 from a sum of squares in another order than the
# select's: within a relative 2^-40 of the select's terms).  Values at or above it are dropped, so it must lie strictly above the
# tau_A that the select computes in FP64 from A_k <= ak (cutoff +inf, which
# can only lower it).  Emulated with numpy's float32 (round to nearest even),
# and once with every FP32 result one ulp lower still (sqrtf is only
# specified within one ulp, and the compiler may contract the products).

_F32 = np.float32


def _round_up_f32(x: float) -> np.float32:
    f = _F32(x)
    if float(f) < x:
        f = np.nextafter(f, _F32(np.inf))
    return f


def _select_tau_a(ak: float, e1: float, E2: float, D: int, sigma: float) -> np.float32:
    """k_screen_select's tau_A (FP64, then rounded up to float) for A_k = ak, no cutoff."""
    sh = 1.0 + 2.0 * (D + 3.0) * 2.0 ** -53
    if not ak < math.inf:
        return _F32(np.inf)
    r = (math.sqrt(max(0.0, ak + E2)) + e1) / sigma
    tau = sh * r * r * (1 + 1e-12)
    if not tau < math.inf:
        return _F32(np.inf)
    r2 = sigma * math.sqrt(sh * tau) + e1
    t = (r2 * r2 + E2) * (1 + 1e-12)
    return _round_up_f32(t) if t < math.inf else _F32(np.inf)


def _merge_tau_a32(ak, e1f, E2f, low=False, margin=True, two_e1=True):
    """screen_tau_a32: each FP32 operation rounded to nearest, or (low) one ulp lower."""
    def op(x):
        x = _F32(x)
        return np.nextafter(x, _F32(-np.inf)) if low and np.isfinite(x) else x
    with np.errstate(over="ignore", invalid="ignore"):
        r = op(op(np.sqrt(max(op(_F32(ak) + E2f), _F32(0.0)))) + e1f)
        w = op(r + e1f) if two_e1 else r
        v = op(op(w * w) + E2f)
        return op(v * _F32(1.0 + 2.0 ** -18)) if margin else v


def _merge_cases():
    rng = np.random.default_rng(91)
    for D in (8, 12, 16, 31, 32, 48, 64):
        for _ in range(40):
            sigma = 2.0 ** int(rng.integers(-30, 31))
            spread = 2.0 ** rng.uniform(-12, 11) * math.sqrt(D)  # sigma (|q - c| + P_c)
            qc = spread * rng.uniform(0.0, 1.0) / sigma
            pc = spread / sigma - qc
            E2, e1, _ = bounds(D, qc, pc, sigma)
            aks = [-E2 * 2.0, -E2, -E2 * rng.uniform(), 0.0, E2 * rng.uniform(), e1 * e1,
                   spread * spread * 2.0 ** rng.uniform(-40, 1), spread * spread * 4.0, 1e30, 3e38, math.inf]
            aks += list(spread * spread * 2.0 ** rng.uniform(-30, 2, size=8))
            for ak in aks:
                yield D, sigma, e1, E2, float(_F32(ak))


def _merge_bound_holds(**variant) -> bool:
    for D, sigma, e1, E2, ak in _merge_cases():
        # the select forms e1 and E2 from its own sum of the squares of q - c:
        # up to a relative 2^-40 above the main pass's
        ta = _select_tau_a(ak, e1 * (1 + 2.0 ** -40), E2 * (1 + 2.0 ** -40), D, sigma)
        e1f, E2f = _round_up_f32(e1), _round_up_f32(E2)
        for low in (False, True):
            lim = _merge_tau_a32(ak, e1f, E2f, low=low, **variant)
            if not (lim > ta or (np.isinf(ta) and np.isinf(lim))):
                return False
    return True


def test_merge_bound_exceeds_select_threshold():
    assert _merge_bound_holds()


@pytest.mark.parametrize("weakened", ["margin", "two_e1"])
def test_weakened_merge_bound_fails(weakened):
    """Without the 1 + 2^-18 factor, or with one e1 instead of two, the FP32
    bound falls to or below the select's tau_A on these inputs."""
    assert not _merge_bound_holds(**{weakened: False})
