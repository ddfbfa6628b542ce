"""Independent numpy restatements used to pin the oracle (tests only).

Pure-Python loops over tiny inputs; no torch ops, so a mistake shared between
the oracle and torch's grid_sample conventions would show up here.
"""
import itertools
import math

import numpy as np


def unnormalize(c, size, align_corners):
    if align_corners:
        return (c + 1.0) / 2.0 * (size - 1)
    return ((c + 1.0) * size - 1.0) / 2.0


def grid_sample(inp, coords, align_corners, padding="zeros", mode="bilinear"):
    """inp [C, *S] with S = (H, W) or (D, H, W); coords [N, nd] ordered (x, y[, z]).
    Returns [N, C] in float64."""
    inp = np.asarray(inp, dtype=np.float64)
    coords = np.asarray(coords, dtype=np.float64)
    C = inp.shape[0]
    S = inp.shape[1:]
    nd = len(S)
    sizes = S[::-1]                      # (W, H[, D]) -- matches coordinate order
    out = np.zeros((coords.shape[0], C))
    for n in range(coords.shape[0]):
        pos = []
        for a in range(nd):
            p = unnormalize(coords[n, a], sizes[a], align_corners)
            if padding == "border":
                p = min(max(p, 0.0), sizes[a] - 1.0)
            pos.append(p)
        if mode == "nearest":
            idx = [int(np.rint(p)) for p in pos]       # round half to even
            if all(0 <= idx[a] < sizes[a] for a in range(nd)):
                out[n] = inp[(slice(None),) + tuple(reversed(idx))]
            continue
        base = [math.floor(p) for p in pos]
        frac = [p - b for p, b in zip(pos, base)]
        for bits in itertools.product((0, 1), repeat=nd):
            w = 1.0
            idx = []
            for a in range(nd):
                w *= frac[a] if bits[a] else (1.0 - frac[a])
                idx.append(base[a] + bits[a])
            if all(0 <= idx[a] < sizes[a] for a in range(nd)):
                out[n] += w * inp[(slice(None),) + tuple(reversed(idx))]
    return out


def sawtooth(p, a0, scale):
    """((p - a0) mod scale) / (scale / 2) - 1, mod with the sign of the divisor."""
    r = np.mod(p - a0, scale)
    return np.clip(r / (scale / 2.0) - 1.0, -1.0, 1.0)


def triangle(p, a0, scale):
    d = p - a0
    r = np.mod(d, scale)
    k = np.floor(d / scale)
    v = r / (scale / 2.0) - 1.0
    return np.where(np.mod(k, 2) == 1, -v, v)


def dct_rows(p, size, dim, n_select):
    """Closed-form DCT dictionary rows: outer products of 1-D cosines
    cos(pi*k*n/p) (mean removed for k > 0), selected at array_split starts,
    L2-normalised."""
    one = []
    for k in range(p):
        v = np.cos(np.arange(size) * k * np.pi / p)
        if k > 0:
            v = v - v.mean()
        one.append(v)
    total = p ** dim
    starts = [int(s[0]) for s in np.array_split(np.arange(total), n_select)] if n_select < total \
        else list(range(total))
    rows = []
    for r in starts:
        digits = []
        q = r
        for _ in range(dim):
            digits.append(q % p)
            q //= p
        digits = digits[::-1]                       # most significant first
        v = one[digits[0]]
        for dgt in digits[1:]:
            v = np.multiply.outer(v, one[dgt])
        v = v.reshape(-1)
        nrm = np.linalg.norm(v)
        rows.append(v / (nrm if nrm else 1.0))
    return np.stack(rows), starts


def mlp(h, layers):
    """layers: list of (W [out, in], b or None), ReLU between layers."""
    h = np.asarray(h, dtype=np.float64)
    for i, (w, b) in enumerate(layers):
        h = h @ np.asarray(w, dtype=np.float64).T
        if b is not None:
            h = h + np.asarray(b, dtype=np.float64)
        if i != len(layers) - 1:
            h = np.maximum(h, 0.0)
    return h
