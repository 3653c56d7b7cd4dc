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


ROWID_EXACT_MIN_PAIRS = (1 << 32) - 2
"""Exact variables whose equal-pair list would exceed this many pairs (the
longest list the pattern merge accepts) are kept as row ids in 'auto' mode."""


def choose_exact_mode(mode, n_fuzzy, exact_pair_totals):
    if mode not in ("auto", "list", "rowid"):
        raise ValueError("Exact_Mode must be 'auto', 'list' or 'rowid'")
    if n_fuzzy == 0 or not exact_pair_totals:
        return "list"
    if mode == "auto":
        return "rowid" if max(exact_pair_totals) > ROWID_EXACT_MIN_PAIRS else "list"
    return mode


def exact_digit_vars(d, k):
    """Exact variables with digit 1 in the exact part ``d`` of a pattern code
    (first exact variable most significant)."""
    return [q for q in range(k) if (d >> (k - 1 - q)) & 1]


def joint_ids(cols_a, cols_b=None):
    """Joint value ids of several id columns (equal ids on every column <=>
    equal joint id); with ``cols_b`` the ids are shared by both sides."""
    if cols_b is None:
        return np.asarray(_joint_ids(cols_a), dtype=np.int64)
    na = len(cols_a[0])
    j = np.asarray(_joint_ids([np.concatenate([a, b]) for a, b in zip(cols_a, cols_b)]), dtype=np.int64)
    return j[:na], j[na:]


def materialize_exact_only(d, k, ids, offs, rowid_dev, id_div, equal_pairs):
    """Pair ids (ascending, on the device) of the exact-only pattern whose
    exact digits are ``d`` and whose fuzzy digits are all 0.

    ``equal_pairs(S)`` returns the ascending level keys of every pair equal on
    all exact variables in S (the reference's exact lists,
    faster/comparison.py:500-600 and faster/deduplication.py:628-701, over the
    joint values of S). Pairs that also agree on an exact variable outside S
    belong to another pattern and are dropped; pairs listed under a pattern
    with a fuzzy agreement and exact digits ``d`` are removed last, which
    leaves exactly the pairs of the reference's setdiff cascade
    (faster/comparison.py:700-725) for this pattern."""
    low = 2 ** k
    S = exact_digit_vars(d, k)
    keys = equal_pairs(S)
    pair = keys >> 2
    if pair.numel():
        r1 = pair // int(id_div)
        r2 = pair % int(id_div)
        keep = torch.ones(pair.numel(), dtype=torch.bool, device=pair.device)
        for q in range(k):
            if q not in S:
                a, b = rowid_dev[q]
                keep &= a[r1] != b[r2]
        pair = pair[keep]
    parts = [ids[offs[c]:offs[c + 1]] for c in range(low, len(offs) - 1) if c % low == d]
    listed = torch.cat(parts) if parts else ids[:0]
    if listed.numel() == 0 or pair.numel() == 0:
        return pair
    listed = _ops.sort_keys(listed.clone(), int(listed.max().item()) + 1)
    return pair[_ops.member_mask(pair, listed, False)]
