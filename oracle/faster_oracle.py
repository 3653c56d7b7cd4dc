"""Oracle for the linkage hot path -- TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference's algorithm (fast-er 0.3.1, ``faster/``):
per-pair Jaro-Winkler in C (``oracle/faster_oracle.c``, loaded with ctypes and
run over row chunks on a thread pool, one chunk per call so the GIL is
released), and NumPy for the host-side steps (unique/inverse, expansion, set
operations, the pattern merge, FP64 EM).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker; the product package never
does.

Each function names the reference lines it restates. Where the reference's
behaviour relies on a third-party primitive (``np.unique``, CuPy's stable
``argsort``/``sort``/``cumsum``), the same primitive is taken from NumPy: the
outputs are sorted sets, so any correct sort gives identical lists.

The numeric-distance and missing-level functions at the end of the module
(``numeric_link``, ``numeric_dedup``, ``comparison_fit_full``,
``deduplication_fit_full``) have no reference counterpart; they are brute
force over every record pair, and tests/test_oracle.py pins them (equality
at distance 0 equals ``exact``, a pair loop on integer data, symmetry, the
embedding of the reference-restating ``comparison_fit`` when nothing is
missing, closed-form missing totals).

Documented deviations from the reference:

* EM runs entirely in FP64. The reference accumulates the E-step log-sum in a
  float32 buffer (``faster/estimation.py:80``); this oracle (and the product)
  use float64 throughout.
* Strings are compared as UTF-8 bytes with byte lengths. The reference packs
  UTF-8 bytes but takes lengths from Python ``len`` (``comparison.py:283,288``),
  which only agrees for ASCII text.
"""
from __future__ import annotations

import ctypes
import itertools
import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_LOCK = threading.Lock()


def lib():
    """Load ``oracle/liboracle.so`` (built by ``oracle/Makefile``)."""
    global _LIB
    with _LIB_LOCK:
        if _LIB is not None:
            return _LIB
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess

            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(path)
        c_u8p = ctypes.POINTER(ctypes.c_ubyte)
        c_i64p = ctypes.POINTER(ctypes.c_int64)
        L.oracle_jw_score.restype = ctypes.c_float
        L.oracle_jw_score.argtypes = [c_u8p, ctypes.c_int, c_u8p, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_int]
        L.oracle_jw_formula.restype = ctypes.c_float
        L.oracle_jw_formula.argtypes = [ctypes.c_int] * 5 + [ctypes.c_float, ctypes.c_int]
        L.oracle_jw_levels_link.restype = None
        L.oracle_jw_levels_link.argtypes = [
            c_u8p, c_i64p, ctypes.c_int64, ctypes.c_int64, c_u8p, c_i64p, ctypes.c_int64,
            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
            c_i64p, ctypes.c_int64, c_i64p, ctypes.c_int64, c_i64p]
        L.oracle_jw_levels_dedup.restype = None
        L.oracle_jw_levels_dedup.argtypes = [
            c_u8p, c_i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
            c_i64p, ctypes.c_int64, c_i64p, ctypes.c_int64, c_i64p]
        L.oracle_binary_search_mask.restype = None
        L.oracle_binary_search_mask.argtypes = [c_i64p, ctypes.c_int64, c_i64p,
                                                ctypes.c_int64, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_uint8)]
        _LIB = L
        return L


def _u8p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))


def _i64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def n_threads():
    return os.cpu_count() or 1


# ---------------------------------------------------------------------------
# Strings
# ---------------------------------------------------------------------------

def _as_bytes(s):
    return s if isinstance(s, (bytes, bytearray)) else str(s).encode("utf-8")


def pack_strings(strings):
    """Concatenate strings as UTF-8 bytes with int64 offsets (n + 1 entries);
    the layout the reference builds at comparison.py:283-302."""
    bs = [_as_bytes(s) for s in strings]
    lens = np.fromiter((len(b) for b in bs), dtype=np.int64, count=len(bs))
    offsets = np.zeros(len(bs) + 1, dtype=np.int64)
    np.cumsum(lens, out=offsets[1:])
    chars = np.frombuffer(b"".join(bs), dtype=np.uint8).copy()
    if chars.size == 0:
        chars = np.zeros(1, dtype=np.uint8)
    return chars, offsets


def jw_score(s1, s2, p=0.1, use_fma=True):
    """Jaro-Winkler score of one pair (comparison.py:11-131)."""
    b1 = np.frombuffer(_as_bytes(s1) + b"\0", dtype=np.uint8)
    b2 = np.frombuffer(_as_bytes(s2) + b"\0", dtype=np.uint8)
    return float(lib().oracle_jw_score(_u8p(b1), len(b1) - 1, _u8p(b2), len(b2) - 1,
                                       ctypes.c_float(p), int(use_fma)))


def jw_formula(len1, len2, m, tt, prefix, p=0.1, use_fma=True):
    """The fp32 formula of comparison.py:99-123 for given (m, tt, prefix)."""
    return float(lib().oracle_jw_formula(int(len1), int(len2), int(m), int(tt),
                                         int(prefix), ctypes.c_float(p), int(use_fma)))


def _f32(x):
    return float(np.float32(x))


def jaro_winkler_all_pairs(str1, str2, offset=0, p=0.1, lower_thr=0.88, upper_thr=0.94,
                           use_fma=True, threads=None):
    """All pairs of ``str1`` x ``str2`` thresholded into two ascending id lists
    ``idx * len(str2) + idy + offset`` (comparison.py:252-344: kernel
    :133-186, thresholds compared in fp32 :321-329). Rows of ``str1`` are split
    over a thread pool; chunk outputs are concatenated in row order, so each
    list is ascending like the reference's argwhere output."""
    chars1, off1 = str1 if isinstance(str1, tuple) else pack_strings(str1)
    chars2, off2 = str2 if isinstance(str2, tuple) else pack_strings(str2)
    n1 = len(off1) - 1
    n2 = len(off2) - 1
    L = lib()
    lo, hi = _f32(lower_thr), _f32(upper_thr)
    threads = threads or n_threads()
    nchunks = max(1, min(n1, threads * 8))
    bounds = np.linspace(0, n1, nchunks + 1).astype(np.int64)

    def run(k):
        r0, r1 = int(bounds[k]), int(bounds[k + 1])
        cap = 1024
        while True:
            o1 = np.empty(cap, dtype=np.int64)
            o2 = np.empty(cap, dtype=np.int64)
            cnt = np.zeros(2, dtype=np.int64)
            L.oracle_jw_levels_link(_u8p(chars1), _i64p(off1), r0, r1, _u8p(chars2),
                                    _i64p(off2), n2, ctypes.c_float(p), ctypes.c_float(lo),
                                    ctypes.c_float(hi), int(use_fma), _i64p(o1), cap,
                                    _i64p(o2), cap, _i64p(cnt))
            if cnt[0] <= cap and cnt[1] <= cap:
                return o1[:cnt[0]].copy(), o2[:cnt[1]].copy()
            cap = int(max(cnt[0], cnt[1]))

    if nchunks == 1 or threads == 1:
        parts = [run(k) for k in range(nchunks)]
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            parts = list(ex.map(run, range(nchunks)))
    out1 = np.concatenate([q[0] for q in parts] + [np.zeros(0, np.int64)]) + offset
    out2 = np.concatenate([q[1] for q in parts] + [np.zeros(0, np.int64)]) + offset
    return [out1.astype(np.int64), out2.astype(np.int64)]


def _unique_groups(values):
    """np.unique with inverse and counts, plus the rows of each unique value in
    ascending row order (CuPy's stable argsort of the inverse, as at
    comparison.py:377-391)."""
    uniq, inv, cnt = np.unique(values, return_inverse=True, return_counts=True)
    inv = np.asarray(inv).reshape(-1)
    rows_sorted = np.argsort(inv, kind="stable").astype(np.int64)
    starts = np.zeros(len(uniq) + 1, dtype=np.int64)
    np.cumsum(cnt, out=starts[1:])
    return uniq, inv, cnt.astype(np.int64), rows_sorted, starts


def group_strings(values):
    """Unique values of a string column (np.unique over Python ``str``,
    comparison.py:377,395 / deduplication.py:420) as UTF-8 bytes, with the
    rows of each value in ascending order (comparison.py:380-391).
    Returns (list of bytes, starts int64[n_unique + 1], rows int64[n])."""
    arr = np.array([s if isinstance(s, str) else str(s) for s in values], dtype=object)
    if len(arr) == 0:
        return [], np.zeros(1, np.int64), np.zeros(0, np.int64)
    uniq, _, _, rows, starts = _unique_groups(arr)
    return [u.encode("utf-8") for u in uniq], starts, rows


def _expand_link(ua, ub, cntA, cntB, rowsA, startsA, rowsB, startsB, n_B):
    """Unique-pair -> original-pair expansion (the kernel at
    comparison.py:196-245): each unique pair (a, b) yields
    rowA * n_B + rowB for all rows of a and b, in a-row-major order."""
    if len(ua) == 0:
        return np.zeros(0, dtype=np.int64)
    ca = cntA[ua]
    cb = cntB[ub]
    tot = ca * cb
    rep = np.repeat(np.arange(len(ua)), tot)
    first = np.zeros(len(ua) + 1, dtype=np.int64)
    np.cumsum(tot, out=first[1:])
    k = np.arange(first[-1], dtype=np.int64) - first[rep]
    ia = rowsA[startsA[ua[rep]] + k // cb[rep]]
    ib = rowsB[startsB[ub[rep]] + k % cb[rep]]
    return ia * n_B + ib


def jaro_winkler_unique(str_A, str_B, p=0.1, lower_thr=0.88, upper_thr=0.94, use_fma=True,
                        threads=None):
    """Per-level sorted original-pair lists for one fuzzy variable
    (comparison.py:346-498): unique values per side, all-pairs JW over unique
    values, expansion to ``rowA * len(str_B) + rowB``, sort."""
    uA, _, cA, rA, sA = _unique_groups(np.asarray(str_A, dtype=object))
    uB, _, cB, rB, sB = _unique_groups(np.asarray(str_B, dtype=object))
    nUB = len(uB)
    lv = jaro_winkler_all_pairs(list(uA), list(uB), 0, p, lower_thr, upper_thr, use_fma,
                                threads)
    out = []
    for ids in lv:
        out.append(np.sort(_expand_link(ids // nUB, ids % nUB, cA, cB, rA, sA, rB, sB,
                                        len(str_B))))
    return out


def exact(str_A, str_B):
    """Exact-equality pairs (comparison.py:500-600): joint unique of the two
    unique sets, values present on both sides (count == 2), expansion, sort."""
    uA, _, cA, rA, sA = _unique_groups(np.asarray(str_A, dtype=object))
    uB, _, cB, rB, sB = _unique_groups(np.asarray(str_B, dtype=object))
    both = np.concatenate((uA, uB))
    ua_all, inv_all, cnt_all = np.unique(both, return_inverse=True, return_counts=True)
    inv_all = np.asarray(inv_all).reshape(-1)
    order = np.argsort(inv_all, kind="stable")
    offs = np.cumsum(cnt_all)
    eq = np.nonzero(cnt_all == 2)[0]
    if eq.size == 0:
        return [np.zeros(0, dtype=np.int64)]
    ia = order[offs[eq] - 2]
    ib = order[offs[eq] - 1] - len(uA)
    return [np.sort(_expand_link(ia, ib, cA, cB, rA, sA, rB, sB, len(str_B)))]


# ---------------------------------------------------------------------------
# Set operations and the pattern merge
# ---------------------------------------------------------------------------

def _mask(arr1, arr2, inside):
    a1 = np.ascontiguousarray(arr1, dtype=np.int64)
    a2 = np.ascontiguousarray(arr2, dtype=np.int64)
    out = np.zeros(len(a1), dtype=np.uint8)
    lib().oracle_binary_search_mask(_i64p(a1), len(a1), _i64p(a2), len(a2), int(inside),
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out.astype(bool)


def intersect(arr1, arr2):
    """search.py:57-87 (empty input -> empty uint64 array, :71-72)."""
    if len(arr1) == 0 or len(arr2) == 0:
        return np.empty(0, dtype=np.uint64)
    return np.asarray(arr1)[_mask(arr1, arr2, True)]


def setdiff(arr1, arr2):
    """search.py:89-119 (either side empty -> arr1 itself, :103-104)."""
    if len(arr1) == 0 or len(arr2) == 0:
        return arr1
    return np.asarray(arr1)[_mask(arr1, arr2, False)]


def reduce_(function, iterable, initial=None):
    """search.py:121-157 without the memory-pool calls."""
    it = iter(iterable)
    value = next(it) if initial is None else initial
    for element in it:
        value = function(value, element)
    return value


def merge_patterns(per_var):
    """The merge cascade of comparison.py:695-725 (identical at
    deduplication.py:773-803): per-variable level lists -> per-pattern lists
    in itertools.product order, first variable slowest, pattern 0 omitted."""
    indices = [list(x) for x in per_var]
    cur = indices[0]
    del indices[0]
    while len(indices) > 0:
        output = []
        for j in range(len(indices[0])):
            output.append(reduce_(setdiff, cur, indices[0][j]))
        while len(cur) > 0:
            output.append(reduce_(setdiff, indices[0], cur[0]))
            for j in range(len(indices[0])):
                output.append(intersect(cur[0], indices[0][j]))
            del cur[0]
        cur = output
        del indices[0]
    return cur


def counts_link(indices, n_A, n_B):
    """comparison.py:732-748."""
    c = [len(x) for x in indices]
    return np.concatenate([[n_A * n_B - np.sum(c, dtype=np.int64)], c]).astype(np.int64)


def counts_dedup(indices, n):
    """deduplication.py:810-826 (pattern 0 includes the n self-pairs)."""
    c = [len(x) for x in indices]
    return np.concatenate([[int(n * (n + 1) / 2) - np.sum(c, dtype=np.int64)], c]).astype(np.int64)


def comparison_fit(cols_fuzzy_A, cols_fuzzy_B, cols_exact_A=(), cols_exact_B=(), p=0.1,
                   lower_thr=0.88, upper_thr=0.94, use_fma=True, threads=None):
    """Comparison.fit + Counts (comparison.py:662-748) on column arrays.
    Returns (per-variable level lists, Indices, Counts)."""
    per_var = []
    for a, b in zip(cols_fuzzy_A, cols_fuzzy_B):
        per_var.append(jaro_winkler_unique(a, b, p, lower_thr, upper_thr, use_fma, threads))
    for a, b in zip(cols_exact_A, cols_exact_B):
        per_var.append(exact(a, b))
    indices = merge_patterns(per_var)
    n_A = len(cols_fuzzy_A[0]) if cols_fuzzy_A else len(cols_exact_A[0])
    n_B = len(cols_fuzzy_B[0]) if cols_fuzzy_B else len(cols_exact_B[0])
    return per_var, indices, counts_link(indices, n_A, n_B)


# ---------------------------------------------------------------------------
# Deduplication
# ---------------------------------------------------------------------------

def jaro_winkler_dedup(string, p=0.1, lower_thr=0.88, upper_thr=0.94, use_fma=True,
                       threads=None):
    """deduplication.py:393-626: lower triangle of unique x unique (diagonal
    scores 1), expansion (kernels :206-243 and :253-338), sort. Ids are
    ``row * n + col`` with row > col."""
    u, _, cnt, rows, starts = _unique_groups(np.asarray(string, dtype=object))
    nu = len(u)
    n = len(string)
    chars, off = pack_strings(list(u))
    L = lib()
    lo, hi = _f32(lower_thr), _f32(upper_thr)
    threads = threads or n_threads()
    # Balanced row chunks over a triangle: equal-area split.
    nchunks = max(1, min(nu, threads * 8))
    bounds = np.unique(np.round(np.sqrt(np.linspace(0, 1, nchunks + 1)) * nu).astype(np.int64))

    def run(k):
        r0, r1 = int(bounds[k]), int(bounds[k + 1])
        cap = 1024
        while True:
            o1 = np.empty(cap, dtype=np.int64)
            o2 = np.empty(cap, dtype=np.int64)
            c = np.zeros(2, dtype=np.int64)
            L.oracle_jw_levels_dedup(_u8p(chars), _i64p(off), nu, r0, r1, ctypes.c_float(p),
                                     ctypes.c_float(lo), ctypes.c_float(hi), int(use_fma),
                                     _i64p(o1), cap, _i64p(o2), cap, _i64p(c))
            if c[0] <= cap and c[1] <= cap:
                return o1[:c[0]].copy(), o2[:c[1]].copy()
            cap = int(max(c[0], c[1]))

    nb = len(bounds) - 1
    if nb <= 1 or threads == 1:
        parts = [run(k) for k in range(nb)]
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            parts = list(ex.map(run, range(nb)))
    out = []
    for lvl in range(2):
        ids = np.concatenate([q[lvl] for q in parts] + [np.zeros(0, np.int64)])
        out.append(np.sort(_expand_dedup(ids // nu, ids % nu, cnt, rows, starts, n)))
    return out


def _expand_dedup(ua, ub, cnt, rows, starts, n):
    """Kernels deduplication.py:206-243 (counts) and :253-338 (ids): for
    a != b every row pair, oriented max * n + min; for a == b the pairs i > j
    within the group (rows ascending)."""
    res = []
    for a, b in zip(ua.tolist(), ub.tolist()):
        ra = rows[starts[a]:starts[a + 1]]
        rb = rows[starts[b]:starts[b + 1]]
        if a != b:
            x = np.repeat(ra, len(rb))
            y = np.tile(rb, len(ra))
            res.append(np.maximum(x, y) * n + np.minimum(x, y))
        else:
            ii, jj = np.tril_indices(len(ra), -1)
            res.append(ra[ii] * n + ra[jj])
    if not res:
        return np.zeros(0, dtype=np.int64)
    return np.concatenate(res).astype(np.int64)


def exact_dedup(string):
    """deduplication.py:628-701: pairs i > j with equal values, id i * n + j."""
    u, _, cnt, rows, starts = _unique_groups(np.asarray(string, dtype=object))
    n = len(string)
    res = []
    for g in np.nonzero(cnt > 1)[0].tolist():
        r = rows[starts[g]:starts[g + 1]]
        ii, jj = np.tril_indices(len(r), -1)
        res.append(r[ii] * n + r[jj])
    if not res:
        return [np.zeros(0, dtype=np.int64)]
    return [np.sort(np.concatenate(res).astype(np.int64))]


def deduplication_fit(cols_fuzzy, cols_exact=(), p=0.1, lower_thr=0.88, upper_thr=0.94,
                      use_fma=True, threads=None):
    """Deduplication.fit + Counts (deduplication.py:740-826)."""
    per_var = []
    for a in cols_fuzzy:
        per_var.append(jaro_winkler_dedup(a, p, lower_thr, upper_thr, use_fma, threads))
    for a in cols_exact:
        per_var.append(exact_dedup(a))
    indices = merge_patterns(per_var)
    n = len(cols_fuzzy[0]) if cols_fuzzy else len(cols_exact[0])
    return per_var, indices, counts_dedup(indices, n)


# ---------------------------------------------------------------------------
# Fellegi-Sunter EM (FP64)
# ---------------------------------------------------------------------------

def gamma(levels):
    """estimation.py:56-70 generalised to any level counts per variable."""
    return np.array(list(itertools.product(*(range(int(L)) for L in levels))))


def init_params(levels):
    """estimation.py:112-120: Lambda = 0.1, then every variable's pi_0 drawn
    from NumPy's global legacy RNG (descending), then every pi_1 (ascending)."""
    pi0 = [-np.sort(-np.random.dirichlet(np.arange(1, L * 50 + 1, 50))) for L in levels]
    pi1 = [np.sort(np.random.dirichlet(np.arange(1, L * 50 + 1, 50))) for L in levels]
    return 0.1, [pi0, pi1]


def match_probability(G, lam, Pi):
    """estimation.py:72-96 in float64."""
    cond = np.zeros((2, len(G)), dtype=np.float64)
    for m in range(2):
        for k in range(G.shape[1]):
            cond[m, :] += np.log(Pi[m][k][G[:, k]])
        cond[m, :] = np.exp(cond[m, :])
    return (lam * cond[1, :]) / (lam * cond[1, :] + (1 - lam) * cond[0, :])


def em_fit(levels, counts, tol=1e-4, max_iter=5000, init=None):
    """estimation.py:98-158 in float64. Returns (Lambda, Pi, iterations,
    converged, Ksi)."""
    G = gamma(levels)
    counts = np.asarray(counts, dtype=np.int64)
    lam, Pi = init if init is not None else init_params(levels)
    Pi = [[np.array(x, dtype=np.float64) for x in Pi[0]], [np.array(x, dtype=np.float64) for x in Pi[1]]]
    lam = float(lam)
    convergence = False
    it = 1
    cf = counts.astype(np.float64)
    while not convergence and it <= max_iter:
        ksi = match_probability(G, lam, Pi)
        lam = np.dot(ksi, cf) / np.sum(cf)
        d1 = np.dot(ksi, cf)
        pi1 = [np.array([np.dot((G[:, k] == l) * cf, ksi) for l in range(L)]) / d1
               for k, L in enumerate(levels)]
        d0 = np.dot(1 - ksi, cf)
        pi0 = [np.array([np.dot((G[:, k] == l) * cf, 1 - ksi) for l in range(L)]) / d0
               for k, L in enumerate(levels)]
        new = [pi0, pi1]
        delta = np.max(np.abs(np.concatenate([np.concatenate(x) for x in new]) -
                              np.concatenate([np.concatenate(x) for x in Pi])))
        if delta < tol:
            convergence = True
        Pi = new
        it += 1
    ksi = match_probability(G, lam, Pi)
    return float(lam), Pi, it - 1, convergence, ksi


def log_likelihood(levels, counts, lam, Pi):
    """Observed-data log-likelihood of the mixture over the pattern histogram:
    sum_p N_p log(lam * P1(p) + (1 - lam) * P0(p))."""
    G = gamma(levels)
    c0 = np.ones(len(G))
    c1 = np.ones(len(G))
    for k in range(G.shape[1]):
        c0 = c0 * Pi[0][k][G[:, k]]
        c1 = c1 * Pi[1][k][G[:, k]]
    return float(np.dot(np.asarray(counts, dtype=np.float64), np.log(lam * c1 + (1 - lam) * c0)))


def linkage_pairs(indices, ksi, n_B, threshold=0.85):
    """Pair selection of linkage.py:48-62 (including its Indices[i - 1]
    lookup, so pattern 0 maps to the last list). Returns (rows_A, rows_B)."""
    above = np.ravel(np.argwhere(np.asarray(ksi) >= threshold))
    if np.sum([len(indices[i - 1]) for i in above]) == 0:
        raise Exception("No pair of observations has a conditional match probability exceeding the threshold.")
    ids = np.concatenate([np.asarray(indices[i - 1], dtype=np.int64) for i in above])
    return ids // n_B, ids % n_B


# ---------------------------------------------------------------------------
# Numeric-distance and missing levels (no counterpart in the reference: its
# comparisons are Jaro-Winkler and equality only, faster/comparison.py:346-600,
# with the levels fixed at faster/estimation.py:70,116). Brute force over
# every record pair, on dense level matrices; small inputs only.
# ---------------------------------------------------------------------------

def within(x, y, d):
    """max(x, y) <= min(x, y) + d in float64 (NaN: False)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    lo, hi = np.minimum(x, y), np.maximum(x, y)
    with np.errstate(invalid="ignore"):
        return hi <= lo + np.float64(d)


def numeric_level_matrix(xa, xb, d2, d1):
    """Level of every pair (i, j): 2 within d2, 1 within d1, else 0."""
    a = np.asarray(xa, dtype=np.float64)[:, None]
    b = np.asarray(xb, dtype=np.float64)[None, :]
    return np.where(within(a, b, d2), 2, np.where(within(a, b, d1), 1, 0)).astype(np.int8)


def numeric_link(xa, xb, d2, d1):
    """[level-1 ids, level-2 ids], ids i * n_B + j ascending."""
    M = numeric_level_matrix(xa, xb, d2, d1).ravel()
    return [np.nonzero(M == 1)[0].astype(np.int64), np.nonzero(M == 2)[0].astype(np.int64)]


def numeric_dedup(x, d2, d1):
    """[level-1 ids, level-2 ids], ids i * n + j with i > j, ascending."""
    n = len(x)
    M = numeric_level_matrix(x, x, d2, d1)
    M[np.triu_indices(n)] = 0
    M = M.ravel()
    return [np.nonzero(M == 1)[0].astype(np.int64), np.nonzero(M == 2)[0].astype(np.int64)]


def _isna(values):
    import pandas as pd

    return np.asarray(pd.isna(pd.Series(np.asarray(values, dtype=object).reshape(-1))), dtype=bool)


def _as_float(values):
    import pandas as pd

    return pd.to_numeric(pd.Series(np.asarray(values, dtype=object).reshape(-1))).to_numpy(np.float64, na_value=np.nan)


def _fill(D, ids, n_cols, rows_a, rows_b, level):
    """D[rows_a[i], rows_b[j]] = level for subset pair ids i * n_cols + j."""
    ids = np.asarray(ids, dtype=np.int64)
    D[rows_a[ids // n_cols], rows_b[ids % n_cols]] = level


def _digits_link(cols_fuzzy_A, cols_fuzzy_B, cols_exact_A, cols_exact_B, cols_num_A, cols_num_B, d2, d1, missing,
                 p, lower_thr, upper_thr):
    n_A = len((list(cols_fuzzy_A) + list(cols_exact_A) + list(cols_num_A))[0])
    n_B = len((list(cols_fuzzy_B) + list(cols_exact_B) + list(cols_num_B))[0])
    digits, base = [], []
    for a, b in zip(cols_fuzzy_A, cols_fuzzy_B):
        a, b = np.asarray(a, dtype=object), np.asarray(b, dtype=object)
        ma = _isna(a) if missing else np.zeros(n_A, bool)
        mb = _isna(b) if missing else np.zeros(n_B, bool)
        ra, rb = np.nonzero(~ma)[0], np.nonzero(~mb)[0]
        D = np.zeros((n_A, n_B), np.int8)
        if len(ra) and len(rb):
            l1, l2 = jaro_winkler_all_pairs(list(a[ra]), list(b[rb]), p=p, lower_thr=lower_thr, upper_thr=upper_thr)
            _fill(D, l1, len(rb), ra, rb, 1)
            _fill(D, l2, len(rb), ra, rb, 2)
        digits.append((D, ma, mb))
        base.append(3)
    for a, b in zip(cols_exact_A, cols_exact_B):
        a, b = np.asarray(a, dtype=object), np.asarray(b, dtype=object)
        ma = _isna(a) if missing else np.zeros(n_A, bool)
        mb = _isna(b) if missing else np.zeros(n_B, bool)
        ids = np.full(n_A + n_B, -1, dtype=np.int64)
        keep = np.concatenate([~ma, ~mb])
        both = np.concatenate([a, b])[keep]
        if len(both):
            _, inv = np.unique(both, return_inverse=True)
            ids[keep] = np.asarray(inv).reshape(-1)
        ia, ib = ids[:n_A], ids[n_A:]
        D = ((ia[:, None] == ib[None, :]) & (ia[:, None] >= 0)).astype(np.int8)
        digits.append((D, ma, mb))
        base.append(2)
    for a, b in zip(cols_num_A, cols_num_B):
        xa, xb = _as_float(a), _as_float(b)
        D = numeric_level_matrix(xa, xb, d2, d1)
        ma = np.isnan(xa) if missing else np.zeros(n_A, bool)
        mb = np.isnan(xb) if missing else np.zeros(n_B, bool)
        digits.append((D, ma, mb))
        base.append(3)
    return n_A, n_B, digits, base


def _patterns_from_digits(digits, base, missing, pair_mask):
    levels = [b + 1 for b in base] if missing else list(base)
    strides = [int(np.prod(levels[k + 1:])) for k in range(len(levels))]
    code = np.zeros(digits[0][0].shape, dtype=np.int64)
    per_var = []
    for (D, ma, mb), L, st in zip(digits, base, strides):
        D = D.astype(np.int64)
        if missing:
            D = np.where(ma[:, None] | mb[None, :], L, D)
        code += D * st
        per_var.append([np.nonzero(((D == l) & pair_mask).ravel())[0].astype(np.int64) for l in range(1, L)])
    code = np.where(pair_mask, code, -1).ravel()
    n_pat = int(np.prod(levels))
    order = np.argsort(code, kind="stable")  # ids ascending within a code
    bounds = np.searchsorted(code[order], np.arange(1, n_pat + 1))
    indices = [order[bounds[c - 1]:bounds[c]].astype(np.int64) for c in range(1, n_pat)]
    return levels, per_var, indices


def comparison_fit_full(cols_fuzzy_A, cols_fuzzy_B, cols_exact_A=(), cols_exact_B=(), cols_num_A=(), cols_num_B=(),
                        d2=0.0, d1=1.0, missing=False, p=0.1, lower_thr=0.88, upper_thr=0.94):
    """Linkage with fuzzy, exact and numeric variables (in that pattern order)
    and, with ``missing``, a last level per variable for pairs where either
    value is missing (pd.isna); missing values are compared with nothing.
    Brute force over every record pair. Returns (levels, per-variable level
    lists, Indices, Counts)."""
    n_A, n_B, digits, base = _digits_link(cols_fuzzy_A, cols_fuzzy_B, cols_exact_A, cols_exact_B, cols_num_A,
                                          cols_num_B, d2, d1, missing, p, lower_thr, upper_thr)
    levels, per_var, indices = _patterns_from_digits(digits, base, missing, np.ones((n_A, n_B), bool))
    counts = [len(x) for x in indices]
    return levels, per_var, indices, np.concatenate([[n_A * n_B - sum(counts)], counts]).astype(np.int64)


def deduplication_fit_full(cols_fuzzy, cols_exact=(), cols_num=(), d2=0.0, d1=1.0, missing=False, p=0.1,
                           lower_thr=0.88, upper_thr=0.94):
    """Deduplication with fuzzy, exact and numeric variables and optional
    missing levels (pairs i > j, ids i * n + j; Counts[0] includes the n
    self-pairs as deduplication.py:825 does). Fuzzy levels come from
    jaro_winkler_dedup over the non-missing values (orientation and forced
    diagonal as deduplication.py:150-190)."""
    n = len((list(cols_fuzzy) + list(cols_exact) + list(cols_num))[0])
    digits, base = [], []
    for a in cols_fuzzy:
        a = np.asarray(a, dtype=object)
        m = _isna(a) if missing else np.zeros(n, bool)
        r = np.nonzero(~m)[0]
        D = np.zeros((n, n), np.int8)
        if len(r) > 1:
            l1, l2 = jaro_winkler_dedup(a[r], p, lower_thr, upper_thr)
            for lvl, ids in ((1, l1), (2, l2)):
                ids = np.asarray(ids, dtype=np.int64)
                D[r[ids // len(r)], r[ids % len(r)]] = lvl
        digits.append((D, m, m))
        base.append(3)
    for a in cols_exact:
        a = np.asarray(a, dtype=object)
        m = _isna(a) if missing else np.zeros(n, bool)
        ids = np.full(n, -1, dtype=np.int64)
        if (~m).any():
            _, inv = np.unique(a[~m], return_inverse=True)
            ids[~m] = np.asarray(inv).reshape(-1)
        D = ((ids[:, None] == ids[None, :]) & (ids[:, None] >= 0)).astype(np.int8)
        digits.append((D, m, m))
        base.append(2)
    for a in cols_num:
        x = _as_float(a)
        D = numeric_level_matrix(x, x, d2, d1)
        m = np.isnan(x) if missing else np.zeros(n, bool)
        digits.append((D, m, m))
        base.append(3)
    lower = np.tril(np.ones((n, n), bool), -1)
    levels, per_var, indices = _patterns_from_digits(digits, base, missing, lower)
    counts = [len(x) for x in indices]
    return levels, per_var, indices, np.concatenate([[n * (n + 1) // 2 - sum(counts)], counts]).astype(np.int64)
