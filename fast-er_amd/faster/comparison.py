"""Linkage comparison: the ``Comparison`` class and its utility functions.

Drop-in for faster/comparison.py (fast-er 0.3.1): same names, signatures,
defaults, exceptions and outputs. Device work runs through libfaster_hip on
the current ROCm device; results are int64 torch tensors on that device
(the reference returns CuPy arrays).

``Num_Threads``/``num_threads`` and ``Max_Chunk_Size``/``max_chunk_size`` are
accepted for compatibility and have no effect: the comparison kernel has fixed
gfx950 launch shapes, keeps scores in registers and compacts levels in-kernel,
so it needs neither a dense score matrix nor chunking.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import pandas as pd
import torch

from . import _lib, _ops, patterns
from .encode import (exact_column, group_exact_link, group_exact_link_device, group_packed, group_strings_device,
                     pack_strings, packed_column_device)
from .levels import level_of_score

_LEVELS_FUZZY = 3
_LEVELS_EXACT = 2


def fuzzy_keys_link(ga, gb, p=0.1, lower_thr=0.88, upper_thr=0.94, stats=None):
    """Sorted level keys ((row_A * n_B + row_B) * 4 + level) of one fuzzy
    variable from device string groups of both sides."""
    n_a, n_b = ga.n_rows, gb.n_rows
    if n_a == 0 or n_b == 0:
        return torch.empty(0, dtype=torch.int64, device=_lib.device())
    hits, st = _ops.jw_levels(ga.chars, ga.offsets, ga.n_unique, gb.chars, gb.offsets, gb.n_unique,
                              _lib.FH_MODE_LINK, p, lower_thr, upper_thr, want_stats=stats is not None)
    if stats is not None:
        stats.append(st)
    keys = _ops.expand_link(hits, gb.n_unique, ga.starts, ga.rows, gb.starts, gb.rows, n_b)
    return _ops.sort_keys(keys, _ops.max_pair_key(n_a * n_b))


def exact_keys_link(group_a, starts_b, rows_b, n_groups, n_b):
    """Ascending level keys (level 1) of exact-equality pairs."""
    return _ops.exact_link(group_a, starts_b, rows_b, n_groups, n_b)


def _split(keys, n_levels):
    ids, offs = _ops.split_levels(keys, n_levels)
    return [ids[offs[l]:offs[l + 1]] for l in range(1, n_levels)]


def jaro_winkler_gpu(str1, str2, offset=0, p=0.1, lower_thr=0.88, upper_thr=0.94, num_threads=256):
    """Jaro-Winkler similarity between all pairs of strings of two arrays,
    thresholded into two levels (faster/comparison.py:252-344).

    :return: two ascending int64 tensors with ids ``i * len(str2) + j + offset``:
             pairs with ``lower_thr <= s < upper_thr`` and pairs with
             ``s >= upper_thr``.
    """
    _lib.load()
    dev = _lib.device()
    c1, o1 = pack_strings(str1)
    c2, o2 = pack_strings(str2)
    n1, n2 = len(o1) - 1, len(o2) - 1
    t = lambda a: torch.from_numpy(a if a.size else np.zeros(1, a.dtype)).to(dev)
    hits, _ = _ops.jw_levels(t(c1), t(o1), n1, t(c2), t(o2), n2, _lib.FH_MODE_LINK, p, lower_thr, upper_thr)
    hits = _ops.sort_keys(hits, _ops.max_pair_key(n1 * n2))
    return [x + int(offset) for x in _split(hits, 3)]


def jaro_winkler_unique_gpu(str_A, str_B, p=0.1, lower_thr=0.88, upper_thr=0.94, num_threads=256,
                            max_chunk_size=2.0):
    """Jaro-Winkler levels over unique values, expanded to the original rows
    (faster/comparison.py:346-498).

    :return: two ascending int64 tensors of ids ``i * len(str_B) + j``.
    """
    _lib.load()
    ga = group_strings_device(str_A)
    gb = group_strings_device(str_B)
    return _split(fuzzy_keys_link(ga, gb, p, lower_thr, upper_thr), 3)


def exact_gpu(str_A, str_B, num_threads=256):
    """Exact-equality pairs (faster/comparison.py:500-600).

    :return: a one-element list with the ascending int64 ids ``i * len(str_B) + j``.
    """
    _lib.load()
    n_a = len(np.asarray(str_A).reshape(-1))
    g = group_exact_link_device(exact_column(str_A, str_B), n_a)
    keys = exact_keys_link(g.group_a, g.starts_b, g.rows_b, g.n_groups, len(np.asarray(str_B).reshape(-1)))
    return _split(keys, 2)


@dataclass
class LinkInputs:
    """Device-resident inputs of ``Comparison.fit_device``: per fuzzy variable
    the packed columns of A and B, per exact variable the column of A rows
    followed by B rows."""
    n_a: int
    n_b: int
    fuzzy: list
    exact: list


class Comparison():
    """Compares the values of selected variables between two pandas DataFrames
    (faster/comparison.py:602-748).

    :param df_A: First DataFrame to compare.
    :param df_B: Second DataFrame to compare.
    :param Vars_Fuzzy_A: Variables of ``df_A`` compared with Jaro-Winkler.
    :param Vars_Fuzzy_B: Variables of ``df_B`` matching ``Vars_Fuzzy_A``, in the same order.
    :param Vars_Exact_A: Variables of ``df_A`` compared for exact equality.
    :param Vars_Exact_B: Variables of ``df_B`` matching ``Vars_Exact_A``, in the same order.
    """

    def __init__(self, df_A: pd.DataFrame, df_B: pd.DataFrame, Vars_Fuzzy_A, Vars_Fuzzy_B, Vars_Exact_A=[],
                 Vars_Exact_B=[]):
        if len(Vars_Fuzzy_A) != len(Vars_Fuzzy_B):
            raise Exception("The lengths of Vars_Fuzzy_A and Vars_Fuzzy_B differ.")
        if len(Vars_Exact_A) != len(Vars_Exact_B):
            raise Exception("The lengths of Vars_Exact_A and Vars_Exact_B differ.")
        if any(var not in df_A.columns for var in Vars_Fuzzy_A) or any(var not in df_B.columns for var in Vars_Fuzzy_B):
            raise Exception("The names in Vars_Fuzzy_A and Vars_Fuzzy_B must match variables names in df_A and df_B, respectively.")
        if any(var not in df_A.columns for var in Vars_Exact_A) or any(var not in df_B.columns for var in Vars_Exact_B):
            raise Exception("The names in Vars_Exact_A and Vars_Exact_B must match variables names in df_A and df_B, respectively.")
        self.df_A = df_A
        self.df_B = df_B
        self.Vars_Fuzzy_A = Vars_Fuzzy_A
        self.Vars_Fuzzy_B = Vars_Fuzzy_B
        self.Vars_Exact_A = Vars_Exact_A
        self.Vars_Exact_B = Vars_Exact_B
        self.Indices = None
        """Per-pattern pair lists (pattern 0 omitted), ids ``i * len(df_B) + j``,
        patterns in ``itertools.product`` order with the first variable slowest."""
        self.Stats = None
        """Work accounting of the comparison kernels, one dict per fuzzy variable."""
        self._Fit_flag = False
        self._var_keys = None

    def fit(self, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, Num_Threads=256, Max_Chunk_Size=2.0, *,
            Exact_Mode="auto"):
        """Compares all pairs of observations on the selected variables; the
        result is stored in ``Indices`` (faster/comparison.py:662-730).

        ``Exact_Mode`` (keyword-only): 'list' materialises every exact
        variable's equal pairs as the reference does; 'rowid' keeps exact
        variables as per-row value ids, counts the patterns without a fuzzy
        agreement in closed form and leaves their lists lazy; 'auto' picks
        'rowid' when an exact variable has more than
        ``patterns.ROWID_EXACT_MIN_PAIRS`` equal pairs."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        _lib.load()
        self.fit_device(self.device_inputs(), p, Lower_Thr, Upper_Thr, Exact_Mode=Exact_Mode)

    def device_inputs(self):
        """The host part of ``fit``: every compared column packed (strings as
        UTF-8 bytes + offsets, exact columns as int64 / float64 values of A
        then B) and copied to the device."""
        _lib.load()
        fuzzy = [(packed_column_device(self.df_A[va].to_numpy()), packed_column_device(self.df_B[vb].to_numpy()))
                 for va, vb in zip(self.Vars_Fuzzy_A, self.Vars_Fuzzy_B)]
        exact = [exact_column(self.df_A[va].to_numpy(), self.df_B[vb].to_numpy())
                 for va, vb in zip(self.Vars_Exact_A, self.Vars_Exact_B)]
        return LinkInputs(len(self.df_A), len(self.df_B), fuzzy, exact)

    def fit_device(self, inputs, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, *, Exact_Mode="auto"):
        """The device part of ``fit``, from ``device_inputs()``: grouping of
        every column by value, all-pairs comparison, expansion, exact pairs and
        the pattern merge."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        n_a, n_b = inputs.n_a, inputs.n_b
        stats = []
        keys = []
        for ca, cb in inputs.fuzzy:
            keys.append(fuzzy_keys_link(group_packed(ca), group_packed(cb), p, Lower_Thr, Upper_Thr, stats))
        exact = [group_exact_link_device(c, n_a) for c in inputs.exact]
        totals = [_ops.exact_link_total(g.group_a, g.starts_b, g.rows_b, g.n_groups, n_b) for g in exact]
        mode = patterns.choose_exact_mode(Exact_Mode, len(inputs.fuzzy), totals)
        self.Exact_Mode = mode
        self.Stats = stats
        lv_f = [_LEVELS_FUZZY] * len(inputs.fuzzy)
        self._var_levels = lv_f + [_LEVELS_EXACT] * len(exact)
        self._n_pairs = n_a * n_b
        if mode == "rowid":
            rowid_dev = [(g.group_a, g.group_b) for g in exact]
            host = [g.host() for g in exact]
            rowid_host = [(h.group_a, h.group_b) for h in host]

            def equal_pairs(S):
                if len(S) == 1:
                    g = exact[S[0]]
                    return exact_keys_link(g.group_a, g.starts_b, g.rows_b, g.n_groups, n_b)
                ja, jb = patterns.joint_ids([host[q].group_a for q in S], [host[q].group_b for q in S])
                g = group_exact_link(ja, jb)
                ga, sb, rb = g.to_device()
                return exact_keys_link(ga, sb, rb, g.n_groups, n_b)

            def mat(d, ids, offs):
                return patterns.materialize_exact_only(d, len(exact), ids, offs, rowid_dev, n_b, equal_pairs)

            ids, offs, self.Indices, self._Counts = patterns.merge_with_rowids(
                keys, lv_f, rowid_dev, rowid_host, n_b, False, n_a * n_b, mat)
            self._pattern_ids = ids
            self._pattern_offsets = offs
            self._var_keys = keys
            self._rowid_exact = (exact, totals)
            self._Fit_flag = True
        else:
            for g, t in zip(exact, totals):
                keys.append(_ops.exact_link(g.group_a, g.starts_b, g.rows_b, g.n_groups, n_b, capacity=t))
            self._rowid_exact = None
            self._finish(keys, n_a * n_b)
        return self

    def _finish(self, keys, n_pairs):
        ids, offs, n_union = _ops.pattern_merge(keys, self._var_levels)
        self._pattern_ids = ids
        self._pattern_offsets = offs
        self.Indices = [ids[offs[k]:offs[k + 1]] for k in range(1, len(offs) - 1)]
        self._var_keys = keys
        self._n_pairs = n_pairs
        self._Fit_flag = True

    @property
    def Counts(self):
        """Pair count of every pattern, the all-disagree pattern first
        (faster/comparison.py:732-748)."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        try:
            return self._Counts
        except AttributeError:
            counts = [int(x.numel()) for x in self.Indices]
            self._Counts = np.concatenate([[self._n_pairs - np.sum(counts, dtype=np.int64)], counts]).astype(np.int64)
            return self._Counts

    @property
    def Levels(self):
        """Per-variable, per-level pair lists: ``Levels[v][l - 1]`` holds the
        ascending ids at level ``l`` of variable ``v`` (fuzzy variables first;
        exact variables kept as row ids appear as ``patterns.RowIdPairs``)."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        try:
            return self._Levels
        except AttributeError:
            out = [_split(k, L) for k, L in zip(self._var_keys, self._var_levels)]
            if self._rowid_exact is not None:
                exact, totals = self._rowid_exact
                n_b = len(self.df_B)
                for g, t in zip(exact, totals):
                    out.append([patterns.RowIdPairs(t, lambda g=g: _split(
                        exact_keys_link(g.group_a, g.starts_b, g.rows_b, g.n_groups, n_b), 2)[0])])
            self._Levels = out
            return self._Levels

    def posteriors(self, Ksi):
        """Posterior match probability of every pair in ``Indices`` order
        (pattern-major), gathered on the device: ``Ksi[pattern(pair)]``."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        ksi = torch.as_tensor(np.asarray(Ksi, dtype=np.float64))
        return _ops.pair_posteriors(self._pattern_offsets, ksi, self._pattern_ids.numel())
