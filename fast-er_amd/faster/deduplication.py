"""Deduplication: the ``Deduplication`` class and its utility functions.

Drop-in for faster/deduplication.py (fast-er 0.3.1). Pair ids are
``row * len(df) + col`` with ``row > col``. The fuzzy comparison covers each
unordered pair of distinct values once and scores it with the
lexicographically larger value as string 1, which is the orientation the
reference's lower-triangle kernel gets from ``np.unique`` order
(faster/deduplication.py:150-183); identical values score 1 as on the
reference's forced diagonal (faster/deduplication.py:185-190).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import pandas as pd
import torch

from . import _lib, _ops, patterns
from .comparison import _split
from .encode import (exact_column, group_exact_dedup, group_exact_dedup_device, group_packed, group_strings_device,
                     packed_column_device)
from .levels import level_of_score

_LEVELS_FUZZY = 3
_LEVELS_EXACT = 2


def fuzzy_keys_dedup(g, p=0.1, lower_thr=0.88, upper_thr=0.94, stats=None):
    """Sorted level keys ((row * n + col) * 4 + level, row > col) of one fuzzy
    variable from device string groups."""
    n = g.n_rows
    if n == 0:
        return torch.empty(0, dtype=torch.int64, device=_lib.device())
    hits, st = _ops.jw_levels(g.chars, g.offsets, g.n_unique, g.chars, g.offsets, g.n_unique,
                              _lib.FH_MODE_DEDUP, p, lower_thr, upper_thr, want_stats=stats is not None)
    if stats is not None:
        stats.append(st)
    diag = level_of_score(1.0, lower_thr, upper_thr)
    keys = _ops.expand_dedup(hits, g.n_unique, g.starts, g.rows, n, diag)
    return _ops.sort_keys(keys, _ops.max_pair_key(n * n))


def jaro_winkler_dedup_gpu(string, p=0.1, lower_thr=0.88, upper_thr=0.94, num_threads=256, max_chunk_size=2.0):
    """Jaro-Winkler levels between all pairs of values of one array
    (faster/deduplication.py:393-626).

    :return: two ascending int64 tensors of ids ``i * len(string) + j``, i > j.
    """
    _lib.load()
    g = group_strings_device(string)
    return _split(fuzzy_keys_dedup(g, p, lower_thr, upper_thr), 3)


def exact_dedup_gpu(string, num_threads=256):
    """Exact-equality pairs within one array (faster/deduplication.py:628-701).

    :return: a one-element list with the ascending int64 ids ``i * len(string) + j``, i > j.
    """
    _lib.load()
    g = group_exact_dedup_device(exact_column(string))
    return _split(_ops.exact_dedup(g.group, g.starts, g.rows, g.n_groups), 2)


@dataclass
class DedupInputs:
    """Device-resident inputs of ``Deduplication.fit_device``."""
    n: int
    fuzzy: list
    exact: list


class Deduplication():
    """Compares the values of selected variables within one DataFrame
    (faster/deduplication.py:703-826).

    :param df: DataFrame to deduplicate.
    :param Vars_Fuzzy: Variables compared with Jaro-Winkler.
    :param Vars_Exact: Variables compared for exact equality.
    """

    def __init__(self, df: pd.DataFrame, Vars_Fuzzy, Vars_Exact=[]):
        if any(var not in df.columns for var in Vars_Fuzzy) or any(var not in df.columns for var in Vars_Exact):
            raise Exception("The variable names in Vars_Fuzzy and Vars_Exact must match variable names in df.")
        self.df = df
        self.Vars_Fuzzy = Vars_Fuzzy
        self.Vars_Exact = Vars_Exact
        self.Indices = None
        self.Stats = None
        self._Fit_flag = False

    def fit(self, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, Num_Threads=256, Max_Chunk_Size=2.0, *,
            Exact_Mode="auto"):
        """Compares all pairs of observations; the result is stored in ``Indices``
        (faster/deduplication.py:740-808).

        ``Exact_Mode`` (keyword-only): 'list' materialises every exact variable's
        equal pairs as the reference does; 'rowid' keeps exact variables as
        per-row value ids, counts the patterns without a fuzzy agreement in
        closed form and leaves their lists lazy (``patterns.RowIdPairs``);
        'auto' picks 'rowid' when an exact variable has more than
        ``patterns.ROWID_EXACT_MIN_PAIRS`` equal pairs."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        _lib.load()
        self.fit_device(self.device_inputs(), p, Lower_Thr, Upper_Thr, Exact_Mode=Exact_Mode)

    def device_inputs(self):
        """The host part of ``fit``: every compared column packed and copied to
        the device."""
        _lib.load()
        fuzzy = [packed_column_device(self.df[v].to_numpy()) for v in self.Vars_Fuzzy]
        exact = [exact_column(self.df[v].to_numpy()) for v in self.Vars_Exact]
        return DedupInputs(len(self.df), fuzzy, exact)

    def fit_device(self, inputs, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, *, Exact_Mode="auto"):
        """The device part of ``fit``, from ``device_inputs()``."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        n = inputs.n
        stats = []
        keys = [fuzzy_keys_dedup(group_packed(c), p, Lower_Thr, Upper_Thr, stats) for c in inputs.fuzzy]
        exact = [group_exact_dedup_device(c) for c in inputs.exact]
        totals = [_ops.exact_dedup_total(g.group, g.starts, g.rows, g.n_groups) for g in exact]
        mode = patterns.choose_exact_mode(Exact_Mode, len(inputs.fuzzy), totals)
        self.Exact_Mode = mode
        lv_f = [_LEVELS_FUZZY] * len(inputs.fuzzy)
        if mode == "rowid":
            rowid_dev = [(g.group, g.group) for g in exact]
            host = [g.host() for g in exact]
            rowid_host = [(h.group, h.group) for h in host]

            def equal_pairs(S):
                if len(S) == 1:
                    g = exact[S[0]]
                    return _ops.exact_dedup(g.group, g.starts, g.rows, g.n_groups)
                g = group_exact_dedup(patterns.joint_ids([host[q].group for q in S]))
                grp, st, rows = g.to_device()
                return _ops.exact_dedup(grp, st, rows, g.n_groups)

            def mat(d, ids, offs):
                return patterns.materialize_exact_only(d, len(exact), ids, offs, rowid_dev, n, equal_pairs)

            ids, offs, self.Indices, self._Counts = patterns.merge_with_rowids(
                keys, lv_f, rowid_dev, rowid_host, n, True, n * (n + 1) // 2, mat)
            self._rowid_exact = (exact, totals)
        else:
            for g, t in zip(exact, totals):
                keys.append(_ops.exact_dedup(g.group, g.starts, g.rows, g.n_groups, capacity=t))
            ids, offs, _ = _ops.pattern_merge(keys, lv_f + [_LEVELS_EXACT] * len(exact))
            self.Indices = [ids[offs[k]:offs[k + 1]] for k in range(1, len(offs) - 1)]
            self._rowid_exact = None
        self._n = n
        self._pattern_ids = ids
        self._pattern_offsets = offs
        self._var_keys = keys
        self._var_levels = lv_f + [_LEVELS_EXACT] * len(exact)
        self.Stats = stats
        self._Fit_flag = True
        return self

    @property
    def Counts(self):
        """Pair count of every pattern; pattern 0 is ``n(n+1)/2`` minus the rest,
        so it includes the n self-pairs as the reference's does
        (faster/deduplication.py:810-826)."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        try:
            return self._Counts
        except AttributeError:
            n = len(self.df)
            counts = [int(x.numel()) for x in self.Indices]
            self._Counts = np.concatenate([[int(n * (n + 1) // 2) - np.sum(counts, dtype=np.int64)], counts]).astype(np.int64)
            return self._Counts

    @property
    def Levels(self):
        """Per-variable, per-level pair lists (fuzzy variables first; exact
        variables kept as row ids appear as ``patterns.RowIdPairs``)."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        try:
            return self._Levels
        except AttributeError:
            out = [_split(k, L) for k, L in zip(self._var_keys, self._var_levels)]
            if self._rowid_exact is not None:
                exact, totals = self._rowid_exact
                for g, t in zip(exact, totals):
                    out.append([patterns.RowIdPairs(t, lambda g=g: _split(
                        _ops.exact_dedup(g.group, g.starts, g.rows, g.n_groups), 2)[0])])
            self._Levels = out
            return self._Levels
