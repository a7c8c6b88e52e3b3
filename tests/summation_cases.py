"""Closed forms that fix the summation order of each row width's distance recipe.

Shared by the oracle tests (tests/test_oracle.py) and the device tests
(tests/test_gpu_parity.py).
"""

from __future__ import annotations

import numpy as np

# Closed forms that fix the summation order of each width's recipe (src/kernel.rs
# Scan plans: Packed d <= 8 with the d4 pair order and the d8 chunk order for
# points in full chunks of the leaf, Folded 9..16, Streamed >= 17).  Terms are 1
# and T = 2^-53, half an ulp of 1: 1 + T rounds (to even) back to 1, so a T
# added to a partial sum of 1 vanishes, while T + T = 2^-52 added to 1 survives.
# Each recipe therefore gives exactly 1 or exactly 1 + 2^-52, and each vector
# below is chosen so that the named plausible mistake gives the other value.
# L1 terms are |p_i - q_i| (q = 0, p_i = the term), so no root or square blurs them.
T53 = 2.0**-53
ONE, ONE_UP = 1.0, 1.0 + 2.0**-52


def terms(d, **at):
    v = np.zeros(d)
    for k, x in at.items():
        v[int(k[1:])] = x
    return v


# (name, d, terms, offset in leaf, leaf length, expected, mistake it catches)
ORDER_CASES = [
    ("d3-in-order", 3, terms(3, a0=1, a1=T53, a2=T53), 0, 8, ONE, "summing from the last axis"),
    ("d3-in-order-late-one", 3, terms(3, a0=T53, a1=T53, a2=1), 0, 8, ONE_UP, "summing from the last axis"),
    ("d4-full-pairs", 4, terms(4, a0=T53, a1=1, a2=T53), 0, 2, ONE_UP, "in order or (a0+a1)+(a2+a3)"),
    ("d4-full-pairs-2", 4, terms(4, a0=1, a1=T53, a3=T53), 2, 4, ONE_UP, "in order or (a0+a1)+(a2+a3)"),
    ("d4-tail-in-order", 4, terms(4, a0=T53, a1=1, a2=T53), 2, 3, ONE, "the pair order on a tail row"),
    ("d8-full-chunk", 8, terms(8, a0=1, a4=T53, a5=T53), 0, 8, ONE_UP, "in order or the hsum8 lane pairing"),
    ("d8-full-chunk-2", 8, terms(8, a2=1, a6=T53, a7=T53), 8, 16, ONE_UP, "in order or the hsum8 lane pairing"),
    ("d8-tail-in-order", 8, terms(8, a0=1, a4=T53, a5=T53), 8, 9, ONE, "the chunk order on a tail row"),
    ("d16-folded", 16, terms(16, a0=1, a4=T53, a12=T53), 0, 32, ONE_UP, "streamed chunks or in order"),
    ("d16-hsum-pairs", 16, terms(16, a0=1, a2=T53, a6=T53), 0, 32, ONE_UP, "adjacent-lane pairing in hsum8"),
    ("d10-tail-after-hsum", 10, terms(10, a0=1, a1=T53, a9=T53), 0, 32, ONE, "folding the tail into the lanes"),
    ("d20-streamed", 20, terms(20, a0=1, a8=T53, a12=T53), 0, 32, ONE_UP, "folding all chunks into 8 lanes"),
    ("d20-tail-in-order", 20, terms(20, a0=1, a16=T53, a17=T53), 0, 32, ONE, "pairing the tail"),
    ("d33-streamed", 33, terms(33, a0=1, a16=T53, a20=T53), 0, 32, ONE_UP, "one running sum over all axes"),
    ("d33-tail-after-chunks", 33, terms(33, a0=1, a31=T53, a32=T53), 0, 32, ONE, "pairing across the last chunk"),
]
