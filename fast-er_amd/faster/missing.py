"""Missing values and their level in the pattern histogram.

No counterpart in the reference: its levels are fixed at 3 per fuzzy and 2 per
exact variable (faster/estimation.py:70,116) and a missing value is compared
like any other. With ``fit(..., Missing=True)`` every variable gets one more
level, its last, for pairs where either record's value is missing (None, NaN,
pd.NA, NaT: ``pd.isna``). Missing values are left out of the comparisons, so
no level list holds such a pair; the pattern merge then adds the missing
digits of every listed pair from per-row bit masks (fh_pattern_add_missing),
and the pairs that no list holds are counted in closed form per missing set M
(the variables missing on either row):

    Counts[c_M] = P(M) - (listed pairs whose missing set is M)

with c_M the code whose digits are "missing" on M and 0 elsewhere, and P(M)
the number of record pairs whose row masks combine to M (a product over the
distinct row masks of both sides). Their lists are lazy (``MissingPairs``).
"""
from __future__ import annotations

import itertools

import numpy as np
import pandas as pd
import torch

from . import _lib, _ops


def missing_rows(values) -> np.ndarray:
    """Boolean mask of the missing values of a column."""
    return np.asarray(pd.isna(pd.Series(np.asarray(values, dtype=object).reshape(-1))), dtype=bool)


def numeric_values(values) -> np.ndarray:
    """A column as float64 with NaN for missing values."""
    s = pd.to_numeric(pd.Series(np.asarray(values, dtype=object).reshape(-1)), errors="raise")
    return s.to_numpy(dtype=np.float64, na_value=np.nan)


def row_masks(missing_cols, n) -> np.ndarray:
    """uint32 per row: bit v set when variable v is missing on the row."""
    m = np.zeros(n, dtype=np.uint32)
    for v, col in enumerate(missing_cols):
        m |= (np.asarray(col, dtype=bool).astype(np.uint32) << np.uint32(v))
    return m


def pair_totals(mask_a, mask_b, dedup) -> dict:
    """P(M): record pairs (i, j) (i > j for deduplication) whose masks
    combine (bitwise or) to M."""
    out = {}
    ua, ca = np.unique(mask_a, return_counts=True)
    if dedup:
        for x, cx in zip(ua.tolist(), ca.tolist()):
            out[x] = out.get(x, 0) + cx * (cx - 1) // 2
        for (x, cx), (y, cy) in itertools.combinations(list(zip(ua.tolist(), ca.tolist())), 2):
            out[x | y] = out.get(x | y, 0) + cx * cy
        return out
    ub, cb = np.unique(mask_b, return_counts=True)
    for x, cx in zip(ua.tolist(), ca.tolist()):
        for y, cy in zip(ub.tolist(), cb.tolist()):
            out[x | y] = out.get(x | y, 0) + cx * cy
    return out


def strides(levels):
    """Pattern strides, first variable slowest."""
    st = [1] * len(levels)
    for v in range(len(levels) - 2, -1, -1):
        st[v] = st[v + 1] * int(levels[v + 1])
    return st


def missing_code(M, base_levels, st):
    """Code whose digit is "missing" (= the variable's base level count) on
    the variables of bit set M and 0 elsewhere."""
    return sum(int(base_levels[v]) * st[v] for v in range(len(base_levels)) if (M >> v) & 1)


def missing_set_of_code(c, base_levels, st):
    M = 0
    for v in range(len(base_levels)):
        if (c // st[v]) % (int(base_levels[v]) + 1) == int(base_levels[v]):
            M |= 1 << v
    return M


class MissingPairs:
    """The pairs of a missing-only pattern (every compared digit 0 or
    missing), not materialised: ``numel()`` is exact; ``materialize()``
    builds the ascending id list on the device when it fits in memory."""

    def __init__(self, count, builder):
        self._count = int(count)
        self._builder = builder

    def numel(self):
        return self._count

    @property
    def size(self):
        return self._count

    def __len__(self):
        return self._count

    def materialize(self):
        return self._builder()

    def __repr__(self):
        return f"MissingPairs(count={self._count})"


def _pairs_with_mask(M, mask_a, mask_b, n_b, dedup):
    """Ascending ids of all record pairs whose masks combine to exactly M."""
    dev = _lib.device()
    parts = []
    if dedup:
        ua = np.unique(mask_a)
        rows = {x: torch.from_numpy(np.nonzero(mask_a == x)[0].astype(np.int64)).to(dev) for x in ua.tolist()}
        for x in ua.tolist():
            for y in ua.tolist():
                if x > y or (x | y) != M:
                    continue
                rx, ry = rows[x], rows[y]
                if x == y:
                    ii, jj = torch.tril_indices(len(rx), len(rx), -1, device=dev)
                    parts.append(rx[ii] * n_b + rx[jj])
                else:
                    a, b = rx[:, None], ry[None, :]
                    parts.append((torch.maximum(a, b) * n_b + torch.minimum(a, b)).reshape(-1))
    else:
        ua, ub = np.unique(mask_a), np.unique(mask_b)
        for x in ua.tolist():
            for y in ub.tolist():
                if (x | y) != M:
                    continue
                ra = torch.from_numpy(np.nonzero(mask_a == x)[0].astype(np.int64)).to(dev)
                rb = torch.from_numpy(np.nonzero(mask_b == y)[0].astype(np.int64)).to(dev)
                parts.append((ra[:, None] * n_b + rb[None, :]).reshape(-1))
    if not parts:
        return torch.empty(0, dtype=torch.int64, device=dev)
    out = torch.cat(parts)
    return _ops.sort_keys(out, int(out.max().item()) + 1) if out.numel() > 1 else out


def apply_missing(ids, offs, base_levels, mask_a, mask_b, miss_a_dev, miss_b_dev, id_div, max_id, total_pairs, dedup):
    """Adds the missing digits to the merged lists and counts the missing-only
    patterns. ``offs`` are the merge's offsets over the levels base + 1.
    Returns (ids, offsets, Counts, Indices)."""
    levels = [int(x) + 1 for x in base_levels]
    st = strides(levels)
    add = [int(base_levels[v]) * st[v] for v in range(len(levels))]
    ids2, offs2 = _ops.pattern_add_missing(ids, offs, miss_a_dev, miss_b_dev, id_div, max_id, add)
    n_pat = len(offs2) - 1
    sizes = np.diff(np.asarray(offs2, dtype=np.int64))
    totals = pair_totals(mask_a, mask_b, dedup)
    listed = {}
    for c in range(1, n_pat):
        if sizes[c]:
            M = missing_set_of_code(c, base_levels, st)
            listed[M] = listed.get(M, 0) + int(sizes[c])
    counts = sizes.copy()
    lazy = {}
    for M, P in totals.items():
        if M == 0:
            continue
        c = missing_code(M, base_levels, st)
        counts[c] = P - listed.get(M, 0)
        lazy[c] = M
    counts[0] = total_pairs - int(counts[1:].sum())
    indices = []
    for c in range(1, n_pat):
        if c in lazy:
            M = lazy[c]

            def build(M=M):
                allp = _pairs_with_mask(M, mask_a, mask_b, id_div, dedup)
                parts = [ids2[offs2[k]:offs2[k + 1]] for k in range(1, n_pat)
                         if sizes[k] and missing_set_of_code(k, base_levels, st) == M]
                if not parts or allp.numel() == 0:
                    return allp
                lst = torch.cat(parts)
                lst = _ops.sort_keys(lst.clone(), int(lst.max().item()) + 1)
                return allp[_ops.member_mask(allp, lst, False)]

            indices.append(MissingPairs(counts[c], build))
        else:
            indices.append(ids2[offs2[c]:offs2[c + 1]])
    return ids2, offs2, counts.astype(np.int64), indices
