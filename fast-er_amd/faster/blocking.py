"""Batched blocked linkage: every block in one persistent, graph-captured launch.

The reference's blocking recipe runs one ``Comparison`` per block and sums the
``Counts`` (docs/source/usage.rst:34-76). ``BlockedLinkage`` takes the two
tables whole with a block id per row and runs all blocks at once through
libfaster_hip's blocked plan (fh_blocked_*): work items of every block are
popped from one queue by a persistent kernel, hits accumulate pattern codes in
a device pair table, and the run is captured into a hipGraph and replayed.
Patterns whose fuzzy digits are all 0 are counted per block in closed form
from value-group counts (as in ``patterns``). The results are the per-block
pattern counts the recipe would produce and their sum.

``autotune`` times candidate tile widths, tiles per work item, waves per work
item and persistent grid sizes on the actual workload and keeps the fastest.
"""
from __future__ import annotations

import ctypes
import itertools
import time

import numpy as np
import pandas as pd
import torch

from . import _lib
from ._lib import check, stream
from .encode import exact_column, group_exact_link_device, pack_column

DEFAULT_TUNING = dict(tile_width=256, tiles_per_item=4, waves_per_item=8, wgs_per_cu=2, hash_slots=0)

_TUNE_CACHE = {}


def _packed_device(values):
    chars, offs = pack_column(values)
    dev = _lib.device()
    if chars.size == 0:
        chars = np.zeros(1, np.uint8)
    return torch.from_numpy(chars).to(dev), torch.from_numpy(offs).to(dev)


def _equal_pairs_per_block(block_a, block_b, cols_a, cols_b, n_blocks):
    """Per block b: the record pairs (i in A, j in B) of block b with
    cols_a[q][i] == cols_b[q][j] for every q, as exact int64 sums of
    (A rows) x (B rows) over the groups of equal (block, values)."""
    key_a = np.asarray(block_a, dtype=np.int64)
    key_b = np.asarray(block_b, dtype=np.int64)
    span = 1  # product of the value ranges: key = block * span + mixed-radix values
    for ca, cb in zip(cols_a, cols_b):
        m = int(max(ca.max(initial=-1), cb.max(initial=-1))) + 1
        if m <= 0:
            return np.zeros(n_blocks, dtype=np.int64)
        if span > ((1 << 62) // max(n_blocks, 1)) // m:
            return _equal_pairs_per_block_rows(block_a, block_b, cols_a, cols_b, n_blocks)
        key_a = key_a * m + ca
        key_b = key_b * m + cb
        span *= m
    ua, na = np.unique(key_a, return_counts=True)
    ub, nb = np.unique(key_b, return_counts=True)
    out = np.zeros(n_blocks, dtype=np.int64)
    if ua.size == 0 or ub.size == 0:
        return out
    pos = np.minimum(np.searchsorted(ub, ua), ub.size - 1)
    hit = ub[pos] == ua
    np.add.at(out, ua[hit] // span, na[hit].astype(np.int64) * nb[pos[hit]].astype(np.int64))
    return out


def _equal_pairs_per_block_rows(block_a, block_b, cols_a, cols_b, n_blocks):
    """_equal_pairs_per_block for value ranges whose mixed-radix key would
    not fit in int64: groups from np.unique over the rows of (block, values)."""
    both = np.concatenate([np.stack([np.asarray(block_a, dtype=np.int64)] + list(cols_a), 1),
                           np.stack([np.asarray(block_b, dtype=np.int64)] + list(cols_b), 1)])
    _, inv = np.unique(both, axis=0, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    g = int(inv.max()) + 1 if inv.size else 0
    n_a = len(block_a)
    ca = np.bincount(inv[:n_a], minlength=g).astype(np.int64)
    cb = np.bincount(inv[n_a:], minlength=g).astype(np.int64)
    blk = np.zeros(g, dtype=np.int64)
    blk[inv[:n_a]] = block_a
    blk[inv[n_a:]] = block_b
    out = np.zeros(n_blocks, dtype=np.int64)
    np.add.at(out, blk, ca * cb)
    return out


def exact_totals_per_block(ids_a, ids_b, block_a, block_b, n_blocks):
    """T_b(d): pairs of block b whose exact variables are equal exactly on the
    variables with digit 1 in d (first variable most significant), by
    inclusion-exclusion over per-block equal-pair counts."""
    k = len(ids_a)
    na_b = np.bincount(block_a, minlength=n_blocks).astype(np.int64)
    nb_b = np.bincount(block_b, minlength=n_blocks).astype(np.int64)
    N = {(): na_b * nb_b}
    for r in range(1, k + 1):
        for S in itertools.combinations(range(k), r):
            N[S] = _equal_pairs_per_block(block_a, block_b, [np.asarray(ids_a[q], dtype=np.int64) for q in S],
                                          [np.asarray(ids_b[q], dtype=np.int64) for q in S], n_blocks)
    out = {}
    for d in itertools.product(range(2), repeat=k):
        S = tuple(q for q in range(k) if d[q])
        val = np.zeros(n_blocks, dtype=np.int64)
        for T, n in N.items():
            if set(S) <= set(T):
                val += (-1) ** (len(T) - len(S)) * n
        code = 0
        for q in range(k):
            code = code * 2 + d[q]
        out[code] = val
    return out


class BlockedLinkage:
    """Pattern counts of a blocked linkage of ``df_A`` and ``df_B``.

    :param block_A: block id of every row of ``df_A`` (ints in [0, n_blocks)).
    :param block_B: block id of every row of ``df_B``.
    Other parameters as ``Comparison``.
    """

    def __init__(self, df_A: pd.DataFrame, df_B: pd.DataFrame, block_A, block_B, Vars_Fuzzy_A, Vars_Fuzzy_B,
                 Vars_Exact_A=[], Vars_Exact_B=[], p=0.1, Lower_Thr=0.88, Upper_Thr=0.94):
        if len(Vars_Fuzzy_A) != len(Vars_Fuzzy_B) or len(Vars_Exact_A) != len(Vars_Exact_B):
            raise Exception("The lengths of the variable lists of df_A and df_B differ.")
        if not Vars_Fuzzy_A:
            raise Exception("Blocked linkage needs at least one fuzzy variable.")
        _lib.load()
        self.n_a, self.n_b = len(df_A), len(df_B)
        self.block_a = np.asarray(block_A, dtype=np.int32)
        self.block_b = np.asarray(block_B, dtype=np.int32)
        self.n_blocks = int(max(self.block_a.max(initial=-1), self.block_b.max(initial=-1)) + 1)
        self.p, self.lo, self.hi = p, Lower_Thr, Upper_Thr
        dev = _lib.device()
        self._a = [_packed_device(df_A[v].to_numpy()) for v in Vars_Fuzzy_A]
        self._b = [_packed_device(df_B[v].to_numpy()) for v in Vars_Fuzzy_B]
        # Joint value ids of each exact column, grouped on the device; the
        # host copies feed the closed-form per-block counts.
        ex = [group_exact_link_device(exact_column(df_A[a].to_numpy(), df_B[b].to_numpy()), self.n_a)
              for a, b in zip(Vars_Exact_A, Vars_Exact_B)]
        self._ex_dev = [(g.group_a.to(torch.int64).contiguous(), g.group_b.to(torch.int64).contiguous()) for g in ex]
        self._ex_host = [(a.cpu().numpy(), b.cpu().numpy()) for a, b in self._ex_dev]
        self._blk_a = torch.from_numpy(self.block_a).to(dev)
        self._blk_b = torch.from_numpy(self.block_b).to(dev)
        self.n_fuzzy = len(Vars_Fuzzy_A)
        self.n_exact = len(ex)
        self.n_patterns = 3 ** self.n_fuzzy * 2 ** self.n_exact
        self._exact_totals = exact_totals_per_block([a for a, _ in self._ex_host], [b for _, b in self._ex_host],
                                                    self.block_a, self.block_b, self.n_blocks)
        # Record pairs per block (rows of A x rows of B), fixed for the object.
        self._block_pairs = (np.bincount(self.block_a, minlength=self.n_blocks).astype(np.int64)
                             * np.bincount(self.block_b, minlength=self.n_blocks).astype(np.int64))
        self._plan = None
        self.tuning = dict(DEFAULT_TUNING)
        self.last_ms = None
        self.Block_Counts = None

    # -- plan management ----------------------------------------------------
    def _make_plan(self, tuning):
        L = _lib.load()
        nv, ne = self.n_fuzzy, self.n_exact
        VP = ctypes.c_void_p
        ca = (VP * nv)(*[c.data_ptr() for c, _ in self._a])
        oa = (VP * nv)(*[o.data_ptr() for _, o in self._a])
        cb = (VP * nv)(*[c.data_ptr() for c, _ in self._b])
        ob = (VP * nv)(*[o.data_ptr() for _, o in self._b])
        ia = (VP * max(ne, 1))(*[a.data_ptr() for a, _ in self._ex_dev])
        ib = (VP * max(ne, 1))(*[b.data_ptr() for _, b in self._ex_dev])
        tu = _lib.BlockedTuning(**{k: int(v) for k, v in tuning.items()})
        plan = VP(0)
        check(L.fh_blocked_plan_create(nv, ca, oa, self.n_a, cb, ob, self.n_b, VP(self._blk_a.data_ptr()),
                                       VP(self._blk_b.data_ptr()), self.n_blocks, ne, ia, ib,
                                       ctypes.c_float(self.p), ctypes.c_float(self.lo), ctypes.c_float(self.hi),
                                       ctypes.byref(tu), ctypes.byref(plan), stream()), "fh_blocked_plan_create")
        return plan

    def _drop_plan(self):
        if self._plan is not None:
            _lib.load().fh_blocked_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self._drop_plan()
        except Exception:
            pass

    def plan(self, tuning=None):
        if tuning is not None:
            self.tuning = dict(DEFAULT_TUNING, **tuning)
            self._drop_plan()
        if self._plan is None:
            self._plan = self._make_plan(self.tuning)
        return self._plan

    def info(self):
        buf = (ctypes.c_double * 8)()
        check(_lib.load().fh_blocked_info(self.plan(), buf), "fh_blocked_info")
        keys = ["items_short", "items_long", "patterns", "hash_slots", "pairs_kernel", "alg_lane_ops", "grid",
                "lds_bytes"]
        return dict(zip(keys, list(buf)))

    # -- runs ---------------------------------------------------------------
    def run(self, use_graph=True):
        """Enqueues one run (graph replay after the first). Returns self."""
        check(_lib.load().fh_blocked_run(self.plan(), int(bool(use_graph)), stream()), "fh_blocked_run")
        return self

    def run_per_block(self):
        """Enqueues the plain per-block launch loop over the same plan (one
        launch set per block, no graph, no host synchronisation); ``collect``
        reads its counts. Returns self."""
        check(_lib.load().fh_blocked_run_per_block(self.plan(), stream()), "fh_blocked_run_per_block")
        return self

    def collect(self):
        """Waits for the last run; returns the per-block counts (n_blocks x
        n_patterns, int64) including the patterns without fuzzy agreement."""
        L = _lib.load()
        hist = np.zeros((self.n_blocks, self.n_patterns), dtype=np.int64)
        nu = ctypes.c_int64(0)
        ov = ctypes.c_int(0)
        ms = ctypes.c_double(0)
        check(L.fh_blocked_counts(self.plan(), hist.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nu),
                                  ctypes.byref(ov), ctypes.byref(ms)), "fh_blocked_counts")
        if ov.value:
            raise OverflowError("blocked linkage pair table overflowed")
        self.last_ms = ms.value
        self.n_union = nu.value
        low = 2 ** self.n_exact
        counts = hist.copy()
        listed = np.zeros((self.n_blocks, low), dtype=np.int64)
        for c in range(low, self.n_patterns):
            listed[:, c % low] += hist[:, c]
        for d in range(1, low):
            counts[:, d] = self._exact_totals[d] - listed[:, d]
        counts[:, 0] = self._block_pairs - counts[:, 1:].sum(axis=1)
        self.Block_Counts = counts
        return counts

    def fit(self, use_graph=True):
        """One run; per-block counts in ``Block_Counts``, their sum in ``Counts``."""
        for _ in range(2):
            try:
                self.run(use_graph)
                self.collect()
                return self
            except OverflowError:
                slots = int(self.info()["hash_slots"]) * 4
                self.plan(dict(self.tuning, hash_slots=slots))
        raise OverflowError("blocked linkage pair table overflowed twice")

    @property
    def Counts(self):
        if self.Block_Counts is None:
            raise Exception("The model must be fitted first.")
        return self.Block_Counts.sum(axis=0)

    # -- autotuning ---------------------------------------------------------
    def signature(self):
        la = np.bincount(self.block_a, minlength=self.n_blocks)
        return (self.n_a, self.n_b, self.n_blocks, self.n_fuzzy, self.n_exact, int(la.max()),
                int(np.percentile(la, 99)))

    def autotune(self, space=None, repeats=2, verbose=False):
        """Times every configuration of ``space`` (dict of lists) with graph
        replays and keeps the fastest; results are cached per workload
        signature. Returns (best tuning, {config: ms})."""
        sig = self.signature()
        if space is None and sig in _TUNE_CACHE:
            self.plan(_TUNE_CACHE[sig])
            return self.tuning, {}
        space = space or dict(tile_width=[128, 256], tiles_per_item=[2, 8], waves_per_item=[4, 8],
                              wgs_per_cu=[1, 2, 3])
        keys = list(space)
        results = {}
        best, best_ms = None, float("inf")
        for vals in itertools.product(*(space[k] for k in keys)):
            cfg = dict(self.tuning, **dict(zip(keys, vals)))
            try:
                self.plan(cfg)
                self.run(True)
                self.collect()
                ms = []
                for _ in range(repeats):
                    self.run(True)
                    self.collect()
                    ms.append(self.last_ms)
                t = float(np.median(ms))
            except (RuntimeError, OverflowError):
                continue
            results[tuple(vals)] = t
            if verbose:
                print(f"[autotune] {dict(zip(keys, vals))}: {t:.2f} ms", flush=True)
            if t < best_ms:
                best, best_ms = cfg, t
        self.plan(best)
        _TUNE_CACHE[sig] = best
        return best, results


def per_block_loop(df_A, df_B, block_A, block_B, Vars_Fuzzy, Vars_Exact, n_blocks=None, progress=None, **fit_kw):
    """The reference's blocking recipe (docs/source/usage.rst:41-76) on this
    build's plain classes: one Comparison.fit per block. Returns the per-block
    Counts (n_blocks x n_patterns) and the wall time in seconds."""
    from .comparison import Comparison

    block_A = np.asarray(block_A)
    block_B = np.asarray(block_B)
    n_blocks = n_blocks or int(max(block_A.max(), block_B.max()) + 1)
    n_pat = 3 ** len(Vars_Fuzzy) * 2 ** len(Vars_Exact)
    out = np.zeros((n_blocks, n_pat), dtype=np.int64)
    ia = np.argsort(block_A, kind="stable")
    ib = np.argsort(block_B, kind="stable")
    sa = np.searchsorted(block_A[ia], np.arange(n_blocks + 1))
    sb = np.searchsorted(block_B[ib], np.arange(n_blocks + 1))
    t0 = time.perf_counter()
    for b in range(n_blocks):
        ra, rb = ia[sa[b]:sa[b + 1]], ib[sb[b]:sb[b + 1]]
        if len(ra) == 0 or len(rb) == 0:
            continue
        A = df_A.iloc[ra].reset_index(drop=True)
        B = df_B.iloc[rb].reset_index(drop=True)
        comp = Comparison(A, B, Vars_Fuzzy, Vars_Fuzzy, Vars_Exact, Vars_Exact)
        comp.fit(Exact_Mode="list", **fit_kw)
        out[b] = comp.Counts
        if progress is not None and (b + 1) % 500 == 0:
            progress(b + 1)
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0
