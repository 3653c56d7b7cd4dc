"""Pattern histogram with exact variables kept as per-row value ids.

An exact-match variable over a column with few distinct values implies a huge
pair list (a 2M-row deduplication on birth year has about 2.6e10 equal pairs,
far beyond what the reference could build at faster/deduplication.py:674).
Such variables are not listed: the pattern merge looks their level up per pair
from the value ids (fh_pattern_merge_rowids), and the patterns whose fuzzy
digits are all 0 are counted in closed form from value-group sizes by
inclusion-exclusion over the exact variables. Their pair lists stay lazy
(``RowIdPairs``).
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

from . import _lib, _ops


def _joint_ids(cols):
    if len(cols) == 1:
        return cols[0]
    stacked = np.stack(cols, axis=1)
    _, inv = np.unique(stacked, axis=0, return_inverse=True)
    return np.asarray(inv).reshape(-1)


def equal_pair_counts(ids_a, ids_b, dedup):
    """N(S) for every subset S of the exact variables: the number of pairs
    whose values are equal on every variable of S (S = () gives all pairs;
    pairs are i > j for deduplication)."""
    k = len(ids_a)
    out = {}
    for r in range(k + 1):
        for S in itertools.combinations(range(k), r):
            if not S:
                na = len(ids_a[0]) if k else 0
                out[S] = na * (na - 1) // 2 if dedup else na * len(ids_b[0])
                continue
            if dedup:
                g = _joint_ids([ids_a[q] for q in S])
                c = np.bincount(g).astype(np.int64)
                out[S] = int(np.sum(c * (c - 1) // 2))
            else:
                na = len(ids_a[0])
                g = _joint_ids([np.concatenate([ids_a[q], ids_b[q]]) for q in S])
                ca = np.bincount(g[:na], minlength=g.max() + 1).astype(np.int64)
                cb = np.bincount(g[na:], minlength=g.max() + 1).astype(np.int64)
                out[S] = int(np.dot(ca, cb))
    return out


def exact_pattern_totals(ids_a, ids_b, dedup):
    """T(d) for every digit vector d in {0,1}^k (first exact variable most
    significant): the number of pairs equal on exactly the variables with
    digit 1, by inclusion-exclusion over N(S)."""
    k = len(ids_a)
    N = equal_pair_counts(ids_a, ids_b, dedup)
    tot = {}
    for d in itertools.product(range(2), repeat=k):
        S = tuple(q for q in range(k) if d[q])
        val = 0
        for T, n in N.items():
            if set(S) <= set(T):
                val += (-1) ** (len(T) - len(S)) * n
        code = 0
        for q in range(k):
            code = code * 2 + d[q]
        tot[code] = val
    return tot


class RowIdPairs:
    """A pattern's pair list defined by exact-variable equality and not
    materialised. ``numel()`` is exact; ``materialize()`` builds the list
    (ascending int64 ids on the device) when it fits in memory."""

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
        return f"RowIdPairs(count={self._count})"


def merge_with_rowids(keys, levels, rowid_dev, rowid_host, id_div, dedup, total_pairs, materialize_fn):
    """Pattern merge of listed fuzzy variables plus row-id exact variables.

    Returns (ids, offsets, indices, counts): ``indices`` has a tensor view for
    every pattern with a fuzzy digit and a ``RowIdPairs`` for the others."""
    k = len(rowid_dev)
    ids, offs, n_union = _ops.pattern_merge(keys, levels, rowid_dev, id_div)
    n_pat = len(offs) - 1
    sizes = np.diff(np.asarray(offs, dtype=np.int64))
    low = 2 ** k  # codes below this have every fuzzy digit 0
    totals = exact_pattern_totals([h[0] for h in rowid_host], [h[1] for h in rowid_host], dedup)
    counts = np.zeros(n_pat, dtype=np.int64)
    counts[low:] = sizes[low:]
    listed_by_exact = np.zeros(low, dtype=np.int64)
    for c in range(low, n_pat):
        listed_by_exact[c % low] += sizes[c]
    for d in range(1, low):
        counts[d] = totals[d] - listed_by_exact[d]
    counts[0] = total_pairs - counts[1:].sum()
    indices = []
    for c in range(1, n_pat):
        if c < low:
            indices.append(RowIdPairs(counts[c], (lambda d=c: materialize_fn(d, ids, offs))))
        else:
            indices.append(ids[offs[c]:offs[c + 1]])
    return ids, offs, indices, counts


ROWID_EXACT_MIN_PAIRS = 1 << 31
"""Exact variables whose equal-pair list would exceed this many pairs are
kept as row ids ('auto' mode)."""


def choose_exact_mode(mode, n_fuzzy, exact_pair_totals):
    if mode not in ("auto", "list", "rowid"):
        raise ValueError("Exact_Mode must be 'auto', 'list' or 'rowid'")
    if n_fuzzy == 0 or not exact_pair_totals:
        return "list"
    if mode == "auto":
        return "rowid" if max(exact_pair_totals) > ROWID_EXACT_MIN_PAIRS else "list"
    return mode


def materialize_exact_only(all_equal_sorted, ids, offs, d, low):
    """Pairs of exact-only pattern d for one exact variable: every equal pair
    minus those already listed under a fuzzy pattern with exact digit d."""
    parts = [ids[offs[c]:offs[c + 1]] for c in range(low, len(offs) - 1) if c % low == d]
    listed = torch.cat(parts) if parts else ids[:0]
    listed = _ops.sort_keys(listed.clone(), int(listed.max().item()) + 1 if listed.numel() else 1)
    if listed.numel() == 0:
        return all_equal_sorted
    return all_equal_sorted[_ops.member_mask(all_equal_sorted, listed, False)]
