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

from dataclasses import dataclass, field

import numpy as np
import pandas as pd
import torch

from . import _lib, _ops, concurrent, missing, patterns
from .comparison import _split
from .encode import (DeviceExactDedupGroups, exact_column, group_exact_dedup, group_exact_dedup_device, group_packed,
                     group_packed_multi, group_strings_device, packed_column_device)
from .levels import level_of_score

_LEVELS_FUZZY = 3
_LEVELS_EXACT = 2
_LEVELS_NUMERIC = 3


def fuzzy_keys_dedup(g, p=0.1, lower_thr=0.88, upper_thr=0.94, stats=None):
    """Sorted level keys ((row * n + col) * 4 + level, row > col) of one fuzzy
    variable from device string groups."""
    n = g.n_rows
    if n == 0 or g.n_unique == 0:
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
    """Device-resident inputs of ``Deduplication.fit_device`` (see
    ``comparison.LinkInputs``; one table, so one set of rows and masks)."""
    n: int
    fuzzy: list
    exact: list
    numeric: list = field(default_factory=list)
    missing: bool = False
    fuzzy_rows: list = None
    exact_rows: list = None
    mask: np.ndarray = None
    mask_dev: torch.Tensor = None


class Deduplication():
    """Compares the values of selected variables within one DataFrame
    (faster/deduplication.py:703-826).

    :param df: DataFrame to deduplicate.
    :param Vars_Fuzzy: Variables compared with Jaro-Winkler.
    :param Vars_Exact: Variables compared for exact equality.
    :param Vars_Numeric: (keyword-only; no counterpart in the reference)
        numeric variables compared by distance into three levels.
    """

    def __init__(self, df: pd.DataFrame, Vars_Fuzzy, Vars_Exact=[], *, Vars_Numeric=()):
        if any(var not in df.columns for var in Vars_Fuzzy) or any(var not in df.columns for var in Vars_Exact) or \
                any(var not in df.columns for var in Vars_Numeric):
            raise Exception("The variable names in Vars_Fuzzy and Vars_Exact must match variable names in df.")
        self.df = df
        self.Vars_Fuzzy = Vars_Fuzzy
        self.Vars_Exact = Vars_Exact
        self.Vars_Numeric = list(Vars_Numeric)
        self.Indices = None
        self.Stats = None
        self.Var_Levels = None
        self._Fit_flag = False

    def fit(self, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, Num_Threads=256, Max_Chunk_Size=2.0, *,
            Exact_Mode="auto", Numeric_Dist=(0.0, 1.0), Missing=False):
        """Compares all pairs of observations; the result is stored in ``Indices``
        (faster/deduplication.py:740-808).

        Keyword-only options as ``Comparison.fit``: ``Exact_Mode`` ('list',
        'rowid' or 'auto'), ``Numeric_Dist = (d2, d1)`` for numeric variables
        and ``Missing`` (a last level per variable for pairs with a missing
        value)."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        _lib.load()
        self.fit_device(self.device_inputs(Missing), p, Lower_Thr, Upper_Thr, Exact_Mode=Exact_Mode,
                        Numeric_Dist=Numeric_Dist)

    def device_inputs(self, Missing=False):
        """The host part of ``fit``: every compared column packed and copied to
        the device."""
        _lib.load()
        dev = _lib.device()
        n = len(self.df)
        numeric = [torch.from_numpy(missing.numeric_values(self.df[v].to_numpy())).to(dev) for v in self.Vars_Numeric]
        if not Missing:
            fuzzy = [packed_column_device(self.df[v].to_numpy()) for v in self.Vars_Fuzzy]
            exact = [exact_column(self.df[v].to_numpy()) for v in self.Vars_Exact]
            return DedupInputs(n, fuzzy, exact, numeric)
        names = list(self.Vars_Fuzzy) + list(self.Vars_Exact) + self.Vars_Numeric
        if len(names) > 32:
            raise Exception("Missing levels support at most 32 compared variables.")
        miss = [missing.missing_rows(self.df[v].to_numpy()) for v in names]
        nf = len(self.Vars_Fuzzy)

        def rows_of(mask):
            return torch.from_numpy(np.nonzero(~mask)[0].astype(np.int64)).to(dev)

        fuzzy = [packed_column_device(self.df[v].to_numpy()[~miss[k]]) for k, v in enumerate(self.Vars_Fuzzy)]
        fuzzy_rows = [rows_of(miss[k]) for k in range(nf)]
        exact = [exact_column(self.df[v].to_numpy()[~miss[nf + q]]) for q, v in enumerate(self.Vars_Exact)]
        exact_rows = [rows_of(miss[nf + q]) for q in range(len(self.Vars_Exact))]
        mask = missing.row_masks(miss, n)
        return DedupInputs(n, fuzzy, exact, numeric, True, fuzzy_rows, exact_rows, mask,
                           torch.from_numpy(mask.view(np.int32)).to(dev))

    def fit_device(self, inputs, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, *, Exact_Mode="auto",
                   Numeric_Dist=(0.0, 1.0)):
        """The device part of ``fit``, from ``device_inputs()``."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        n = inputs.n
        stats = []
        keys = []
        grouped = group_packed_multi(list(inputs.fuzzy)) if inputs.fuzzy else []

        def fuzzy_task(k):
            def task():
                g = grouped[k]
                if inputs.missing:
                    g.rows = inputs.fuzzy_rows[k][g.rows]
                    g.n_rows = n
                st = []
                return fuzzy_keys_dedup(g, p, Lower_Thr, Upper_Thr, st), st
            return task

        for kk, st in concurrent.run([fuzzy_task(k) for k in range(len(inputs.fuzzy))]):
            keys.append(kk)
            stats.extend(st)
        exact = []
        for q, c in enumerate(inputs.exact):
            g = group_exact_dedup_device(c)
            if inputs.missing:
                rows = inputs.exact_rows[q]
                grp = torch.full((n,), -1, dtype=torch.int64, device=rows.device)
                grp[rows] = g.group
                g = DeviceExactDedupGroups(g.n_groups, grp, g.starts, rows[g.rows])
            exact.append(g)
        totals = [_ops.exact_dedup_total(g.group, g.starts, g.rows, g.n_groups) for g in exact]
        if inputs.numeric or inputs.missing:
            if Exact_Mode == "rowid":
                raise Exception("Exact_Mode='rowid' cannot be combined with numeric variables or missing levels.")
            mode = "list"
        else:
            mode = patterns.choose_exact_mode(Exact_Mode, len(inputs.fuzzy), totals)
        self.Exact_Mode = mode
        lv_f = [_LEVELS_FUZZY] * len(inputs.fuzzy)
        base_levels = lv_f + [_LEVELS_EXACT] * len(exact) + [_LEVELS_NUMERIC] * len(inputs.numeric)
        self._base_levels = base_levels
        self.Var_Levels = [x + 1 for x in base_levels] if inputs.missing else list(base_levels)
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
            d2, d1 = float(Numeric_Dist[0]), float(Numeric_Dist[1])
            for x in inputs.numeric:
                keys.append(_ops.numeric_dedup(x, d2, d1))
            # Without missing levels, the exact lists are the equal pairs of the
            # value ids: the merge looks them up by id for the other lists.
            listed_ids = None if inputs.missing else (
                [None] * len(inputs.fuzzy) + [(g.group, g.group) for g in exact] + [None] * len(inputs.numeric))
            ids, offs, _ = _ops.pattern_merge(keys, self.Var_Levels, listed_ids=listed_ids, id_div=n)
            if inputs.missing:
                ids, offs, self._Counts, self.Indices = missing.apply_missing(
                    ids, offs, base_levels, inputs.mask, inputs.mask, inputs.mask_dev, inputs.mask_dev, n,
                    max(0, n * n - 1), n * (n + 1) // 2, True)
            else:
                self.Indices = [ids[offs[k]:offs[k + 1]] for k in range(1, len(offs) - 1)]
            self._rowid_exact = None
        self._n = n
        self._pattern_ids = ids
        self._pattern_offsets = offs
        self._var_keys = keys
        self._var_levels = base_levels
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
