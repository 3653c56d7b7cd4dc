"""Thin torch-tensor wrappers over the C ABI (include/faster_hip.h).

Every function takes and returns device tensors on the current ROCm device
and runs on torch's current stream. Pair-list keys are int64 tensors holding
the ABI's uint64 level keys (id * 4 + level, always below 2**63).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream


def _i64(x) -> ctypes.c_int64:
    return ctypes.c_int64(int(x))


def _empty_i64(n: int) -> torch.Tensor:
    return torch.empty(max(int(n), 1), dtype=torch.int64, device=_lib.device())


def _retry(run, first_capacity: int):
    """Runs ``run(capacity) -> (required, result)`` and, when the arena was too
    small, once more with the exact size."""
    cap = int(first_capacity)
    need, res = run(cap)
    if need > cap:
        need2, res = run(int(need))
        if need2 != need:
            raise RuntimeError(f"output size changed between runs ({need} vs {need2})")
    return need, res


def bits_for(max_value: int) -> int:
    return max(1, int(max_value).bit_length())


class _Rates:
    """Output sizes per input unit seen by recent calls, to size the next
    call's output buffer. Only calls of a comparable size (at least 1/64 of
    the new call's units) count, so a tiny call with an extreme rate (every
    pair a hit at thresholds of 0) does not size a large one."""

    def __init__(self):
        self.seen = []  # (units, rate), most recent last

    def predict(self, units):
        r = [rate for u, rate in self.seen if u * 64 >= units]
        return max(r) if r else 0.0

    def add(self, units, rate):
        self.seen = (self.seen + [(units, rate)])[-32:]


_hit_rate = _Rates()  # hits per value pair (sizes the comparison's hit arena)


def jw_levels(chars_a, off_a, n_a, chars_b, off_b, n_b, mode, p, lower_thr, upper_thr,
              capacity=None, want_stats=False):
    """All-pairs JW levels (fh_jw_levels). Returns (hit keys, stats dict or None).

    The hit arena is sized from the rows and from the largest hit rate of
    earlier calls (with 25% headroom), so an overflow, which re-runs the
    pass with the exact size, is rare."""
    L = _lib.load()
    pairs = float(n_a) * float(n_b)
    if capacity is None:
        capacity = max(1 << 20, 4 * (int(n_a) + int(n_b)), int(1.25 * _hit_rate.predict(pairs) * pairs))
    stats = _lib.JwStats()

    def run(cap):
        hits = _empty_i64(cap)
        n = ctypes.c_int64(0)
        check(L.fh_jw_levels(ptr(chars_a), ptr(off_a), _i64(n_a), ptr(chars_b), ptr(off_b), _i64(n_b),
                             int(mode), ctypes.c_float(p), ctypes.c_float(lower_thr),
                             ctypes.c_float(upper_thr), ptr(hits), _i64(cap), ctypes.byref(n),
                             ctypes.byref(stats) if want_stats else None, stream()), "fh_jw_levels")
        return n.value, hits

    need, hits = _retry(run, capacity)
    if pairs > 0:
        _hit_rate.add(pairs, need / pairs)
    return hits[:need], (stats.as_dict() if want_stats else None)


def expand_link(hits, n_ub, starts_a, rows_a, starts_b, rows_b, n_b):
    L = _lib.load()
    n_hits = hits.numel()
    if n_hits == 0:
        return _empty_i64(0)[:0]

    def run(cap):
        out = _empty_i64(cap)
        n = ctypes.c_int64(0)
        check(L.fh_expand_link(ptr(hits), _i64(n_hits), _i64(n_ub), ptr(starts_a), ptr(rows_a),
                               ptr(starts_b), ptr(rows_b), _i64(n_b), ptr(out), _i64(cap),
                               ctypes.byref(n), stream()), "fh_expand_link")
        return n.value, out

    need, out = _retry(run, max(2 * n_hits, 1024))
    return out[:need]


def expand_link_sorted(hits, n_ua, n_ub, starts_a, rows_a, n_a, starts_b, rows_b, n_b):
    """Record-pair level keys of the unique-value hits, ascending
    (fh_expand_link_sorted)."""
    L = _lib.load()
    n_hits = hits.numel()
    if n_hits == 0:
        return _empty_i64(0)[:0]

    def run(cap):
        out = _empty_i64(cap)
        n = ctypes.c_int64(0)
        check(L.fh_expand_link_sorted(ptr(hits), _i64(n_hits), _i64(n_ua), _i64(n_ub), ptr(starts_a), ptr(rows_a),
                                      _i64(n_a), ptr(starts_b), ptr(rows_b), _i64(n_b), ptr(out), _i64(cap),
                                      ctypes.byref(n), stream()), "fh_expand_link_sorted")
        return n.value, out

    need, out = _retry(run, max(2 * n_hits, 1024, int(1.25 * _expand_rate.predict(n_hits) * n_hits)))
    _expand_rate.add(n_hits, need / n_hits)
    return out[:need]


_expand_rate = _Rates()  # record pairs per hit (sizes the sorted expansion's output)


def expand_dedup(hits, n_u, starts, rows, n, diag_level):
    L = _lib.load()
    n_hits = hits.numel()

    def run(cap):
        out = _empty_i64(cap)
        m = ctypes.c_int64(0)
        check(L.fh_expand_dedup(ptr(hits) if n_hits else ptr(None), _i64(n_hits), _i64(n_u), ptr(starts),
                                ptr(rows), _i64(n), int(diag_level), ptr(out), _i64(cap),
                                ctypes.byref(m), stream()), "fh_expand_dedup")
        return m.value, out

    need, out = _retry(run, max(2 * n_hits, 1024))
    return out[:need]


def exact_link(group_a, starts_b, rows_b, n_groups, n_b, capacity=None):
    L = _lib.load()
    n_a = group_a.numel()

    def run(cap):
        out = _empty_i64(cap)
        m = ctypes.c_int64(0)
        check(L.fh_exact_link(ptr(group_a), _i64(n_a), ptr(starts_b), ptr(rows_b), _i64(n_groups),
                              _i64(n_b), ptr(out), _i64(cap), ctypes.byref(m), stream()),
              "fh_exact_link")
        return m.value, out

    need, out = _retry(run, capacity if capacity is not None else 0)
    return out[:need]


def exact_link_total(group_a, starts_b, rows_b, n_groups, n_b) -> int:
    """Number of exact-equality pairs (fh_exact_link's count pass only)."""
    L = _lib.load()
    m = ctypes.c_int64(0)
    check(L.fh_exact_link(ptr(group_a), _i64(group_a.numel()), ptr(starts_b), ptr(rows_b), _i64(n_groups), _i64(n_b),
                          ptr(None), _i64(0), ctypes.byref(m), stream()), "fh_exact_link")
    return m.value


def exact_dedup_total(group, starts, rows, n_groups) -> int:
    """Number of exact-equality pairs i > j (fh_exact_dedup's count pass only)."""
    L = _lib.load()
    m = ctypes.c_int64(0)
    check(L.fh_exact_dedup(ptr(group), _i64(group.numel()), ptr(starts), ptr(rows), _i64(n_groups), ptr(None), _i64(0),
                           ctypes.byref(m), stream()), "fh_exact_dedup")
    return m.value


def exact_dedup(group, starts, rows, n_groups, capacity=None):
    L = _lib.load()
    n = group.numel()

    def run(cap):
        out = _empty_i64(cap)
        m = ctypes.c_int64(0)
        check(L.fh_exact_dedup(ptr(group), _i64(n), ptr(starts), ptr(rows), _i64(n_groups), ptr(out),
                               _i64(cap), ctypes.byref(m), stream()), "fh_exact_dedup")
        return m.value, out

    need, out = _retry(run, capacity if capacity is not None else 0)
    return out[:need]


def group_strings(chars: torch.Tensor, offsets: torch.Tensor, n: int):
    """Device grouping of n packed strings (fh_group_strings). Returns
    (unique chars, unique offsets, starts, rows, n_unique)."""
    L = _lib.load()
    dev = _lib.device()
    total = max(int(chars.numel()), 1)
    uchars = torch.empty(total, dtype=torch.uint8, device=dev)
    uoffsets = _empty_i64(n + 1)
    starts = _empty_i64(n + 1)
    rows = _empty_i64(n)
    nu = ctypes.c_int64(0)
    ub = ctypes.c_int64(0)
    check(L.fh_group_strings(ptr(chars), ptr(offsets), _i64(n), ptr(uchars), ptr(uoffsets), ptr(starts), ptr(rows),
                             ctypes.byref(nu), ctypes.byref(ub), stream()), "fh_group_strings")
    k = nu.value
    return uchars[:max(ub.value, 1)], uoffsets[:k + 1], starts[:k + 1], rows[:n], k


def group_strings_multi(cols):
    """fh_group_strings_multi over [(chars, offsets, n), ...] (up to 16
    columns): one (unique chars, unique offsets, starts, rows, n_unique) per
    column, as group_strings gives it."""
    L = _lib.load()
    dev = _lib.device()
    k = len(cols)
    outs = []
    for chars, _, n in cols:
        outs.append((torch.empty(max(int(chars.numel()), 1), dtype=torch.uint8, device=dev), _empty_i64(n + 1),
                     _empty_i64(n + 1), _empty_i64(n)))
    P = ctypes.c_void_p * k
    I = ctypes.c_int64 * k
    nu, ub = I(), I()
    check(L.fh_group_strings_multi(k, P(*[ptr(c[0]) for c in cols]), P(*[ptr(c[1]) for c in cols]),
                                   I(*[int(c[2]) for c in cols]), P(*[ptr(o[0]) for o in outs]),
                                   P(*[ptr(o[1]) for o in outs]), P(*[ptr(o[2]) for o in outs]),
                                   P(*[ptr(o[3]) for o in outs]), nu, ub, stream()), "fh_group_strings_multi")
    res = []
    for (uchars, uoff, starts, rows), (_, _, n), u, b in zip(outs, cols, nu, ub):
        res.append((uchars[:max(b, 1)], uoff[:u + 1], starts[:u + 1], rows[:n], u))
    return res


FH_VALUES_I64 = 0
FH_VALUES_F64 = 1


def group_values(values: torch.Tensor, kind: int):
    """Device grouping of n numeric values (fh_group_values; ``values`` holds
    int64 values or float64 bit patterns). Returns (group, starts, rows,
    n_groups)."""
    L = _lib.load()
    n = values.numel()
    group = _empty_i64(n)
    starts = _empty_i64(n + 1)
    rows = _empty_i64(n)
    ng = ctypes.c_int64(0)
    v = values.contiguous()
    check(L.fh_group_values(ptr(v) if n else ptr(None), _i64(n), int(kind), ptr(group), ptr(starts), ptr(rows),
                            ctypes.byref(ng), stream()), "fh_group_values")
    k = ng.value
    return group[:n], starts[:k + 1], rows[:n], k


def groups_side_b(starts: torch.Tensor, rows: torch.Tensor, n_groups: int, n_a: int, n_b: int, want_group=False):
    """B rows per group of a grouping of [A rows; B rows] (fh_groups_side_b).
    Returns (starts_b, rows_b, group or None)."""
    L = _lib.load()
    starts_b = _empty_i64(n_groups + 1)
    rows_b = _empty_i64(n_b)
    group = _empty_i64(n_a + n_b) if want_group else None
    check(L.fh_groups_side_b(ptr(starts), ptr(rows), _i64(n_groups), _i64(n_a), _i64(n_b), ptr(starts_b),
                             ptr(rows_b), ptr(group), stream()), "fh_groups_side_b")
    return starts_b[:n_groups + 1], rows_b[:n_b], (group[:n_a + n_b] if want_group else None)


def numeric_link(values_a: torch.Tensor, values_b: torch.Tensor, d2: float, d1: float) -> torch.Tensor:
    """Sorted level keys ((i * n_b + j) * 4 + level) of a numeric-distance
    variable (fh_numeric_link; float64 values, NaN = missing)."""
    L = _lib.load()
    va = values_a.to(torch.float64).contiguous()
    vb = values_b.to(torch.float64).contiguous()
    n_a, n_b = va.numel(), vb.numel()

    def run(cap):
        out = _empty_i64(cap)
        m = ctypes.c_int64(0)
        check(L.fh_numeric_link(ptr(va) if n_a else ptr(None), _i64(n_a), ptr(vb) if n_b else ptr(None), _i64(n_b),
                                ctypes.c_double(d2), ctypes.c_double(d1), ptr(out), _i64(cap), ctypes.byref(m),
                                stream()), "fh_numeric_link")
        return m.value, out

    need, out = _retry(run, 0)
    return sort_keys(out[:need], max_pair_key(n_a * n_b))


def numeric_dedup(values: torch.Tensor, d2: float, d1: float) -> torch.Tensor:
    """Sorted level keys ((row * n + col) * 4 + level, row > col) of a
    numeric-distance variable within one column (fh_numeric_dedup)."""
    L = _lib.load()
    v = values.to(torch.float64).contiguous()
    n = v.numel()

    def run(cap):
        out = _empty_i64(cap)
        m = ctypes.c_int64(0)
        check(L.fh_numeric_dedup(ptr(v) if n else ptr(None), _i64(n), ctypes.c_double(d2), ctypes.c_double(d1),
                                 ptr(out), _i64(cap), ctypes.byref(m), stream()), "fh_numeric_dedup")
        return m.value, out

    need, out = _retry(run, 0)
    return sort_keys(out[:need], max_pair_key(n * n))


def pattern_add_missing(ids: torch.Tensor, offsets, miss_a: torch.Tensor, miss_b: torch.Tensor, id_div: int,
                        max_id: int, add):
    """Adds missing-level digits to a pattern-major pair list
    (fh_pattern_add_missing). Returns (ids, host offsets)."""
    L = _lib.load()
    n_patterns = len(offsets) - 1
    n = ids.numel()
    d_offs = torch.tensor(offsets, dtype=torch.int64, device=_lib.device())
    out = _empty_i64(n)
    offs = (ctypes.c_int64 * (n_patterns + 1))()
    addv = (ctypes.c_int64 * len(add))(*[int(x) for x in add])
    check(L.fh_pattern_add_missing(ptr(ids) if n else ptr(None), ptr(d_offs), int(n_patterns), _i64(n), ptr(miss_a),
                                   ptr(miss_b), _i64(id_div), _i64(max_id), len(add), addv, ptr(out), offs, stream()),
          "fh_pattern_add_missing")
    return out[:n], [offs[i] for i in range(n_patterns + 1)]


def sort_keys(keys: torch.Tensor, max_key: int) -> torch.Tensor:
    """Ascending radix sort of level keys below ``max_key``: returns the
    sorted keys in a new tensor and leaves ``keys`` overwritten."""
    L = _lib.load()
    n = keys.numel()
    if n <= 1:
        return keys
    out = torch.empty_like(keys)
    check(L.fh_sort_u64_into(ptr(keys), ptr(out), _i64(n), 0, bits_for(max_key), stream()), "fh_sort_u64_into")
    return out


def split_levels(keys: torch.Tensor, n_levels: int):
    """Sorted level keys -> (ids, host offsets [n_levels + 1])."""
    L = _lib.load()
    n = keys.numel()
    ids = _empty_i64(n)
    offs = (ctypes.c_int64 * (n_levels + 1))()
    check(L.fh_split_levels(ptr(keys) if n else ptr(None), _i64(n), int(n_levels), ptr(ids), offs,
                            stream()), "fh_split_levels")
    return ids[:n], [offs[i] for i in range(n_levels + 1)]


def member_mask(arr1: torch.Tensor, arr2: torch.Tensor, inside: bool) -> torch.Tensor:
    L = _lib.load()
    n1, n2 = arr1.numel(), arr2.numel()
    mask = torch.empty(max(n1, 1), dtype=torch.uint8, device=_lib.device())
    a1 = arr1.contiguous().to(torch.int64)
    a2 = arr2.contiguous().to(torch.int64)
    check(L.fh_sorted_member_mask(ptr(a1), _i64(n1), ptr(a2) if n2 else ptr(None), _i64(n2),
                                  int(bool(inside)), ptr(mask), stream()), "fh_sorted_member_mask")
    return mask[:n1].bool()


def pattern_merge(var_keys, var_levels, rowid_vars=None, id_div=1, listed_ids=None):
    """Per-variable sorted level keys -> (ids, host pattern offsets, n_union).

    ``rowid_vars``: optional list of (ids_a, ids_b) device int64 tensors of
    two-level exact variables that follow the listed ones in pattern order and
    are looked up per pair instead of being listed (fh_pattern_merge_rowids).
    ``listed_ids``: optional per listed variable (ids_a, ids_b) or None; a
    pair of per-row value ids states that the variable's list is exactly the
    pairs of equal ids, which the other lists then look up by id
    (fh_pattern_merge_listed)."""
    L = _lib.load()
    nv = len(var_keys)
    keys = [k.contiguous() for k in var_keys]
    ptrs = (ctypes.c_void_p * nv)(*[k.data_ptr() if k.numel() else 0 for k in keys])
    sizes = (ctypes.c_int64 * nv)(*[k.numel() for k in keys])
    levels = (ctypes.c_int32 * nv)(*[int(x) for x in var_levels])
    ne = len(rowid_vars) if rowid_vars else 0
    n_patterns = int(np.prod([int(x) for x in var_levels])) * (2 ** ne)
    cap = max(1, sum(k.numel() for k in keys))
    ids = _empty_i64(cap)
    offs = (ctypes.c_int64 * (n_patterns + 1))()
    n_union = ctypes.c_int64(0)
    if ne:
        pa = (ctypes.c_void_p * ne)(*[a.data_ptr() for a, _ in rowid_vars])
        pb = (ctypes.c_void_p * ne)(*[b.data_ptr() for _, b in rowid_vars])
        check(L.fh_pattern_merge_rowids(nv, ptrs, sizes, levels, ne, pa, pb, _i64(id_div), ptr(ids), _i64(cap),
                                        offs, ctypes.byref(n_union), stream()), "fh_pattern_merge_rowids")
    elif listed_ids is not None and any(x is not None for x in listed_ids):
        assert len(listed_ids) == nv
        la = [None if x is None else x[0].contiguous() for x in listed_ids]
        lb = [None if x is None else x[1].contiguous() for x in listed_ids]
        pa = (ctypes.c_void_p * nv)(*[0 if x is None else x.data_ptr() for x in la])
        pb = (ctypes.c_void_p * nv)(*[0 if x is None else x.data_ptr() for x in lb])
        check(L.fh_pattern_merge_listed(nv, ptrs, sizes, levels, pa, pb, _i64(id_div), ptr(ids), _i64(cap), offs,
                                        ctypes.byref(n_union), stream()), "fh_pattern_merge_listed")
    else:
        check(L.fh_pattern_merge(nv, ptrs, sizes, levels, ptr(ids), _i64(cap), offs, ctypes.byref(n_union),
                                 stream()), "fh_pattern_merge")
    if n_union.value > cap:
        raise RuntimeError("pattern merge: union larger than the sum of its inputs")
    return ids[:n_union.value], [offs[i] for i in range(n_patterns + 1)], n_union.value


def _levels_arr(levels):
    return (ctypes.c_int32 * len(levels))(*[int(x) for x in levels])


def em_fit(levels, counts: torch.Tensor, params0: torch.Tensor, tol: float, max_iter: int):
    """FP64 EM on the device. Returns (params, ksi, iterations, converged)."""
    L = _lib.load()
    n_patterns = int(np.prod([int(x) for x in levels]))
    params = params0.clone().to(device=_lib.device(), dtype=torch.float64).contiguous()
    ksi = torch.empty(n_patterns, dtype=torch.float64, device=_lib.device())
    it = ctypes.c_int(0)
    conv = ctypes.c_int(0)
    c = counts.to(device=_lib.device(), dtype=torch.int64).contiguous()
    check(L.fh_em_fit(len(levels), _levels_arr(levels), ptr(c), ptr(params), ctypes.c_double(tol),
                      int(max_iter), ptr(ksi), ctypes.byref(it), ctypes.byref(conv), stream()),
          "fh_em_fit")
    return params, ksi, it.value, bool(conv.value)


def match_probability(levels, params: torch.Tensor) -> torch.Tensor:
    L = _lib.load()
    n_patterns = int(np.prod([int(x) for x in levels]))
    p = params.to(device=_lib.device(), dtype=torch.float64).contiguous()
    ksi = torch.empty(n_patterns, dtype=torch.float64, device=_lib.device())
    check(L.fh_match_probability(len(levels), _levels_arr(levels), ptr(p), ptr(ksi), stream()),
          "fh_match_probability")
    return ksi


def pair_posteriors(pattern_offsets, ksi: torch.Tensor, n_pairs: int) -> torch.Tensor:
    L = _lib.load()
    offs = torch.tensor(pattern_offsets, dtype=torch.int64, device=_lib.device())
    k = ksi.to(device=_lib.device(), dtype=torch.float64).contiguous()
    post = torch.empty(max(int(n_pairs), 1), dtype=torch.float64, device=_lib.device())
    check(L.fh_pair_posteriors(ptr(offs), len(pattern_offsets) - 1, ptr(k), ptr(post), _i64(n_pairs),
                               stream()), "fh_pair_posteriors")
    return post[:n_pairs]


def max_pair_key(n_ids: int) -> int:
    """Largest level key for ids below n_ids."""
    return max(0, int(n_ids) - 1) * 4 + 3


def ceil_log2(x: int) -> int:
    return int(math.ceil(math.log2(max(2, x))))
