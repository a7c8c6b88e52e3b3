"""The reference's FP64 reduced-distance recipes for the row widths of the
batching and serving workloads, written out term by term (src/kernel.rs):

  d = 2, 3, 5-7   Packed, in order: (a0 + a1) + a2 ...
  d = 4           Packed: full pairs (a0 + a2) + (a1 + a3), leaf tails in order
  d = 8           Packed: full 8-row chunks ((a0+a1)+(a2+a3))+((a4+a5)+(a6+a7)),
                  leaf tails in order
  d = 16          Folded: w_j = a_j + a_(8+j), ((w0+w4)+(w2+w6))+((w1+w5)+(w3+w7))
  d = 64          Streamed: ((0 + h(c0)) + h(c1)) + ... over the 8-wide chunks,
                  h(c) = ((c0+c4)+(c2+c6))+((c1+c5)+(c3+c7))

`full` marks points in a complete chunk of their leaf (counted from the leaf's
first row, src/kernel.rs:447).  Works on numpy arrays and on torch tensors
(elementwise IEEE operations in this order, no fused multiply-add).  Returns
the Euclidean distance (sqrt of the reduced distance, src/metric.rs:76).
"""

from __future__ import annotations


def _hsum8(a, base):
    c = [a[..., base + j] for j in range(8)]
    return ((c[0] + c[4]) + (c[2] + c[6])) + ((c[1] + c[5]) + (c[3] + c[7]))


def reduced(diff, full=None):
    """Reduced L2 distance from coordinate differences `diff` (..., d)."""
    a = diff * diff
    d = a.shape[-1]
    if d <= 8 and d not in (4, 8):
        s = a[..., 0]
        for j in range(1, d):
            s = s + a[..., j]
        return s
    if d in (4, 8):
        seq = a[..., 0]
        for j in range(1, d):
            seq = seq + a[..., j]
        if d == 4:
            packed = (a[..., 0] + a[..., 2]) + (a[..., 1] + a[..., 3])
        else:
            packed = ((a[..., 0] + a[..., 1]) + (a[..., 2] + a[..., 3])) + \
                     ((a[..., 4] + a[..., 5]) + (a[..., 6] + a[..., 7]))
        if full is None:
            raise ValueError("d = 4 and 8 need the full-chunk mask")
        return _where(full, packed, seq)
    if d == 16:
        w = [a[..., j] + a[..., 8 + j] for j in range(8)]
        return ((w[0] + w[4]) + (w[2] + w[6])) + ((w[1] + w[5]) + (w[3] + w[7]))
    if d % 8 == 0 and d > 16:
        s = _hsum8(a, 0)  # 0 + h(c0) == h(c0)
        for c in range(1, d // 8):
            s = s + _hsum8(a, 8 * c)
        return s
    raise ValueError(f"no recipe written here for d = {d}")


def _where(mask, x, y):
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return torch.where(mask, x, y)
    except ImportError:  # pragma: no cover
        pass
    import numpy as np

    return np.where(mask, x, y)


def leaf_full_mask(nodes, n, chunk):
    """Per tree position: whether it lies in a complete `chunk`-row chunk of its leaf."""
    import numpy as np

    full = np.zeros(n, bool)
    for nd in nodes[nodes["is_leaf"] == 1]:
        start, end = int(nd["a"]), int(nd["b"])
        m = (end - start) // chunk * chunk
        full[start:start + m] = True
    return full
