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

from dataclasses import dataclass, field

import numpy as np
import pandas as pd
import torch

from . import _lib, _ops, concurrent, missing, patterns
from .encode import (DeviceExactLinkGroups, exact_column, group_exact_link, group_exact_link_device, group_packed,
                     group_packed_multi, group_strings_device, pack_strings, packed_column_device)
from .levels import level_of_score

_LEVELS_FUZZY = 3
_LEVELS_EXACT = 2
_LEVELS_NUMERIC = 3


def fuzzy_keys_link(ga, gb, p=0.1, lower_thr=0.88, upper_thr=0.94, stats=None):
    """Sorted level keys ((row_A * n_B + row_B) * 4 + level) of one fuzzy
    variable from device string groups of both sides."""
    n_a, n_b = ga.n_rows, gb.n_rows
    if n_a == 0 or n_b == 0 or ga.n_unique == 0 or gb.n_unique == 0:
        return torch.empty(0, dtype=torch.int64, device=_lib.device())
    hits, st = _ops.jw_levels(ga.chars, ga.offsets, ga.n_unique, gb.chars, gb.offsets, gb.n_unique,
                              _lib.FH_MODE_LINK, p, lower_thr, upper_thr, want_stats=stats is not None)
    if stats is not None:
        stats.append(st)
    return _ops.expand_link_sorted(hits, ga.n_unique, gb.n_unique, ga.starts, ga.rows, n_a, gb.starts, gb.rows, n_b)


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
    the packed columns of A and B (the non-missing values only when
    ``missing``, with the original row of every packed value in ``rows_*``);
    per exact variable the column of A rows followed by B rows (non-missing
    values only, with their original rows); per numeric variable the float64
    values of A and B (NaN = missing); and with ``missing`` the per-row masks
    of missing variables (host and device)."""
    n_a: int
    n_b: int
    fuzzy: list
    exact: list
    numeric: list = field(default_factory=list)
    missing: bool = False
    fuzzy_rows: list = None
    exact_rows: list = None
    mask_a: np.ndarray = None
    mask_b: np.ndarray = None
    mask_a_dev: torch.Tensor = None
    mask_b_dev: torch.Tensor = None


def _remap_rows(groups, rows_dev):
    """Groups of a packed subset of rows, with rows given as original rows."""
    if rows_dev is None:
        return groups
    groups.rows = rows_dev[groups.rows]
    return groups


def _exact_groups_link(col, n_a_sub, rows_a, rows_b, n_a, n_b):
    """Exact grouping of A's and B's non-missing values mapped to the original
    rows: missing A rows get value id -1 (no partner), missing B rows -2."""
    g = group_exact_link_device(col, n_a_sub)
    if rows_a is None:
        return g
    dev = g.group_a.device
    ga = torch.full((n_a,), -1, dtype=torch.int64, device=dev)
    gb = torch.full((n_b,), -2, dtype=torch.int64, device=dev)
    ga[rows_a] = g.group_a
    gb[rows_b] = g.group_b
    return DeviceExactLinkGroups(g.n_groups, ga, gb, g.starts_b, rows_b[g.rows_b])


class Comparison():
    """Compares the values of selected variables between two pandas DataFrames
    (faster/comparison.py:602-748).

    :param df_A: First DataFrame to compare.
    :param df_B: Second DataFrame to compare.
    :param Vars_Fuzzy_A: Variables of ``df_A`` compared with Jaro-Winkler.
    :param Vars_Fuzzy_B: Variables of ``df_B`` matching ``Vars_Fuzzy_A``, in the same order.
    :param Vars_Exact_A: Variables of ``df_A`` compared for exact equality.
    :param Vars_Exact_B: Variables of ``df_B`` matching ``Vars_Exact_A``, in the same order.
    :param Vars_Numeric_A: (keyword-only; no counterpart in the reference)
        numeric variables of ``df_A`` compared by distance into three levels.
    :param Vars_Numeric_B: the matching variables of ``df_B``.
    """

    def __init__(self, df_A: pd.DataFrame, df_B: pd.DataFrame, Vars_Fuzzy_A, Vars_Fuzzy_B, Vars_Exact_A=[],
                 Vars_Exact_B=[], *, Vars_Numeric_A=(), Vars_Numeric_B=()):
        if len(Vars_Fuzzy_A) != len(Vars_Fuzzy_B):
            raise Exception("The lengths of Vars_Fuzzy_A and Vars_Fuzzy_B differ.")
        if len(Vars_Exact_A) != len(Vars_Exact_B):
            raise Exception("The lengths of Vars_Exact_A and Vars_Exact_B differ.")
        if len(Vars_Numeric_A) != len(Vars_Numeric_B):
            raise Exception("The lengths of Vars_Numeric_A and Vars_Numeric_B differ.")
        if any(var not in df_A.columns for var in Vars_Fuzzy_A) or any(var not in df_B.columns for var in Vars_Fuzzy_B):
            raise Exception("The names in Vars_Fuzzy_A and Vars_Fuzzy_B must match variables names in df_A and df_B, respectively.")
        if any(var not in df_A.columns for var in Vars_Exact_A) or any(var not in df_B.columns for var in Vars_Exact_B):
            raise Exception("The names in Vars_Exact_A and Vars_Exact_B must match variables names in df_A and df_B, respectively.")
        if any(var not in df_A.columns for var in Vars_Numeric_A) or \
                any(var not in df_B.columns for var in Vars_Numeric_B):
            raise Exception("The names in Vars_Numeric_A and Vars_Numeric_B must match variables names in df_A and df_B, respectively.")
        self.df_A = df_A
        self.df_B = df_B
        self.Vars_Fuzzy_A = Vars_Fuzzy_A
        self.Vars_Fuzzy_B = Vars_Fuzzy_B
        self.Vars_Exact_A = Vars_Exact_A
        self.Vars_Exact_B = Vars_Exact_B
        self.Vars_Numeric_A = list(Vars_Numeric_A)
        self.Vars_Numeric_B = list(Vars_Numeric_B)
        self.Indices = None
        """Per-pattern pair lists (pattern 0 omitted), ids ``i * len(df_B) + j``,
        patterns in ``itertools.product`` order with the first variable slowest."""
        self.Stats = None
        """Work accounting of the comparison kernels, one dict per fuzzy variable."""
        self.Var_Levels = None
        """Levels of every variable in pattern order (fuzzy, exact, numeric);
        for ``Estimation(..., Levels=...)``."""
        self._Fit_flag = False
        self._var_keys = None

    def fit(self, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, Num_Threads=256, Max_Chunk_Size=2.0, *,
            Exact_Mode="auto", Numeric_Dist=(0.0, 1.0), Missing=False):
        """Compares all pairs of observations on the selected variables; the
        result is stored in ``Indices`` (faster/comparison.py:662-730).

        Keyword-only options (none has a counterpart in the reference):

        * ``Exact_Mode``: 'list' materialises every exact variable's equal
          pairs as the reference does; 'rowid' keeps exact variables as
          per-row value ids, counts the patterns without a fuzzy agreement in
          closed form and leaves their lists lazy; 'auto' picks 'rowid' when
          an exact variable has more than ``patterns.ROWID_EXACT_MIN_PAIRS``
          equal pairs (list mode is used with numeric variables or
          ``Missing``).
        * ``Numeric_Dist = (d2, d1)``: numeric variables are at level 2 when
          ``max(x, y) <= min(x, y) + d2`` (float64), at level 1 when within
          ``d1`` but not ``d2``, else 0.
        * ``Missing``: every variable gets one more level, its last, for
          pairs where either value is missing (``pd.isna``); see
          ``faster.missing``."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        _lib.load()
        self.fit_device(self.device_inputs(Missing), p, Lower_Thr, Upper_Thr, Exact_Mode=Exact_Mode,
                        Numeric_Dist=Numeric_Dist)

    def device_inputs(self, Missing=False):
        """The host part of ``fit``: every compared column packed (strings as
        UTF-8 bytes + offsets, exact columns as int64 / float64 values of A
        then B, numeric columns as float64) and copied to the device."""
        _lib.load()
        dev = _lib.device()
        n_a, n_b = len(self.df_A), len(self.df_B)
        numeric = [(torch.from_numpy(missing.numeric_values(self.df_A[va].to_numpy())).to(dev),
                    torch.from_numpy(missing.numeric_values(self.df_B[vb].to_numpy())).to(dev))
                   for va, vb in zip(self.Vars_Numeric_A, self.Vars_Numeric_B)]
        if not Missing:
            fuzzy = [(packed_column_device(self.df_A[va].to_numpy()), packed_column_device(self.df_B[vb].to_numpy()))
                     for va, vb in zip(self.Vars_Fuzzy_A, self.Vars_Fuzzy_B)]
            exact = [exact_column(self.df_A[va].to_numpy(), self.df_B[vb].to_numpy())
                     for va, vb in zip(self.Vars_Exact_A, self.Vars_Exact_B)]
            return LinkInputs(n_a, n_b, fuzzy, exact, numeric)
        pairs = list(zip(self.Vars_Fuzzy_A, self.Vars_Fuzzy_B)) + list(zip(self.Vars_Exact_A, self.Vars_Exact_B)) + \
            list(zip(self.Vars_Numeric_A, self.Vars_Numeric_B))
        if len(pairs) > 32:
            raise Exception("Missing levels support at most 32 compared variables.")
        miss_a = [missing.missing_rows(self.df_A[va].to_numpy()) for va, _ in pairs]
        miss_b = [missing.missing_rows(self.df_B[vb].to_numpy()) for _, vb in pairs]
        nf, ne = len(self.Vars_Fuzzy_A), len(self.Vars_Exact_A)

        def rows_of(mask):
            return torch.from_numpy(np.nonzero(~mask)[0].astype(np.int64)).to(dev)

        fuzzy, fuzzy_rows, exact, exact_rows = [], [], [], []
        for v, (va, vb) in enumerate(pairs[:nf]):
            xa = self.df_A[va].to_numpy()[~miss_a[v]]
            xb = self.df_B[vb].to_numpy()[~miss_b[v]]
            fuzzy.append((packed_column_device(xa), packed_column_device(xb)))
            fuzzy_rows.append((rows_of(miss_a[v]), rows_of(miss_b[v])))
        for q, (va, vb) in enumerate(pairs[nf:nf + ne]):
            v = nf + q
            xa = self.df_A[va].to_numpy()[~miss_a[v]]
            xb = self.df_B[vb].to_numpy()[~miss_b[v]]
            exact.append(exact_column(xa, xb))
            exact_rows.append((len(xa), rows_of(miss_a[v]), rows_of(miss_b[v])))
        mask_a = missing.row_masks(miss_a, n_a)
        mask_b = missing.row_masks(miss_b, n_b)
        return LinkInputs(n_a, n_b, fuzzy, exact, numeric, True, fuzzy_rows, exact_rows, mask_a, mask_b,
                          torch.from_numpy(mask_a.view(np.int32)).to(dev),
                          torch.from_numpy(mask_b.view(np.int32)).to(dev))

    def fit_device(self, inputs, p=0.1, Lower_Thr=0.88, Upper_Thr=0.94, *, Exact_Mode="auto",
                   Numeric_Dist=(0.0, 1.0)):
        """The device part of ``fit``, from ``device_inputs()``: grouping of
        every column by value, all-pairs comparison, expansion, exact and
        numeric pairs, the pattern merge and the missing levels."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        n_a, n_b = inputs.n_a, inputs.n_b
        stats = []
        keys = []
        # Every fuzzy column of both sides grouped by value in one device sort.
        flat = group_packed_multi([c for pair in inputs.fuzzy for c in pair]) if inputs.fuzzy else []
        grouped = [(flat[2 * v], flat[2 * v + 1]) for v in range(len(inputs.fuzzy))]

        def fuzzy_task(v):
            def task():
                ga, gb = grouped[v]
                if inputs.missing:
                    ra, rb = inputs.fuzzy_rows[v]
                    ga, gb = _remap_rows(ga, ra), _remap_rows(gb, rb)
                    ga.n_rows, gb.n_rows = n_a, n_b
                st = []
                return fuzzy_keys_link(ga, gb, p, Lower_Thr, Upper_Thr, st), st
            return task

        if (inputs.numeric or inputs.missing) and Exact_Mode == "rowid":
            raise Exception("Exact_Mode='rowid' cannot be combined with numeric variables or missing levels.")

        def exact_task():
            # The exact variables (grouping, pair counts and, in list mode,
            # their lists) beside the fuzzy chains: they need nothing of them.
            if inputs.missing:
                ex = [_exact_groups_link(c, r[0], r[1], r[2], n_a, n_b) for c, r in zip(inputs.exact, inputs.exact_rows)]
            else:
                ex = [group_exact_link_device(c, n_a) for c in inputs.exact]
            tot = [_ops.exact_link_total(g.group_a, g.starts_b, g.rows_b, g.n_groups, n_b) for g in ex]
            if inputs.numeric or inputs.missing:
                md = "list" if ex else patterns.choose_exact_mode(Exact_Mode, len(inputs.fuzzy), tot)
            else:
                md = patterns.choose_exact_mode(Exact_Mode, len(inputs.fuzzy), tot)
            lists = [_ops.exact_link(g.group_a, g.starts_b, g.rows_b, g.n_groups, n_b, capacity=t)
                     for g, t in zip(ex, tot)] if md == "list" else []
            return None, ex, tot, md, lists

        results = concurrent.run([fuzzy_task(v) for v in range(len(inputs.fuzzy))] + [exact_task])
        for k, st in results[:-1]:
            keys.append(k)
            stats.extend(st)
        _, exact, totals, mode, exact_lists = results[-1]
        concurrent.record_on_current(exact, exact_lists)
        self.Exact_Mode = mode
        self.Stats = stats
        lv_f = [_LEVELS_FUZZY] * len(inputs.fuzzy)
        base_levels = lv_f + [_LEVELS_EXACT] * len(exact) + [_LEVELS_NUMERIC] * len(inputs.numeric)
        self._base_levels = base_levels
        self._var_levels = base_levels
        self.Var_Levels = [x + 1 for x in base_levels] if inputs.missing else list(base_levels)
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
            return self
        keys.extend(exact_lists)
        d2, d1 = float(Numeric_Dist[0]), float(Numeric_Dist[1])
        for xa, xb in inputs.numeric:
            keys.append(_ops.numeric_link(xa, xb, d2, d1))
        self._rowid_exact = None
        if inputs.missing:
            ids, offs, _ = _ops.pattern_merge(keys, self.Var_Levels)
            ids, offs, self._Counts, self.Indices = missing.apply_missing(
                ids, offs, base_levels, inputs.mask_a, inputs.mask_b, inputs.mask_a_dev, inputs.mask_b_dev, n_b,
                max(0, n_a * n_b - 1), n_a * n_b, False)
            self._pattern_ids = ids
            self._pattern_offsets = offs
            self._var_keys = keys
            self._Fit_flag = True
            return self
        # The exact lists are the equal pairs of the value ids: the merge looks
        # them up by id for the other lists' entries.
        listed_ids = [None] * len(inputs.fuzzy) + [(g.group_a, g.group_b) for g in exact] + [None] * len(
            inputs.numeric)
        self._finish(keys, n_a * n_b, listed_ids, n_b)
        return self

    def _finish(self, keys, n_pairs, listed_ids=None, id_div=1):
        ids, offs, n_union = _ops.pattern_merge(keys, self._var_levels, listed_ids=listed_ids, id_div=id_div)
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
        ascending ids at level ``l`` of variable ``v`` (fuzzy, then exact, then
        numeric variables; exact variables kept as row ids appear as
        ``patterns.RowIdPairs``; the missing level is not listed)."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        try:
            return self._Levels
        except AttributeError:
            out = [_split(k, L) for k, L in zip(self._var_keys, self._base_levels)]
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
        (pattern-major, listed pairs), gathered on the device:
        ``Ksi[pattern(pair)]``."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        ksi = torch.as_tensor(np.asarray(Ksi, dtype=np.float64))
        return _ops.pair_posteriors(self._pattern_offsets, ksi, self._pattern_ids.numel())
