"""GPU parity: the HIP path (through the C ABI) against the oracle.

Integer/byte/index outputs must match bit for bit; EM parameters and
posteriors within 1e-9 relative."""
import json
import os

import numpy as np
import pandas as pd
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _np(t):
    return t.detach().cpu().numpy().astype(np.int64) if isinstance(t, torch.Tensor) else np.asarray(t, np.int64)


def _eq_lists(got, ref, what=""):
    assert len(got) == len(ref), what
    for k, (g, r) in enumerate(zip(got, ref)):
        g, r = _np(g), np.asarray(r, dtype=np.int64)
        assert g.shape == r.shape, f"{what}[{k}]: {g.shape} vs {r.shape}"
        np.testing.assert_array_equal(g, r, err_msg=f"{what}[{k}]")


def _words(rng, n, alpha="abcde", lo=0, hi=8, dup=None):
    out = ["".join(alpha[k] for k in rng.integers(0, len(alpha), rng.integers(lo, hi + 1))) for _ in range(n)]
    if dup:
        for k in range(0, n, dup):
            out[k] = out[(k * 7) % max(1, k)] if k else out[0]
    return out


def test_library_and_device():
    import faster._lib as L

    lib = L.load()
    assert lib.fh_version() >= 100
    assert torch.cuda.is_available()


def test_valu_probe():
    """fh_probe_valu (the VALU-peak check of bench.py --valu-probe) reports
    the lane-ops it launched, and its per-thread words equal the recurrence
    restated on the CPU (or3, add, select on 8 chains per thread)."""
    import ctypes

    import faster._lib as L

    lib = L.load()
    blocks, iters = 3, 5
    out = torch.zeros(blocks * 256, dtype=torch.int32, device="cuda")
    ops = ctypes.c_double(0.0)
    L.check(lib.fh_probe_valu(L.ptr(out), blocks, iters, ctypes.byref(ops), L.stream()), "fh_probe_valu")
    torch.cuda.synchronize()
    assert ops.value == 3 * 8 * 8 * iters * blocks * 256
    seed = 0x9E3779B9
    tid = np.arange(256, dtype=np.uint32)
    m = [np.uint32((seed * (k + 1)) & 0xFFFFFFFF) ^ tid for k in range(8)]
    na = [np.zeros(256, dtype=np.uint32) for _ in range(8)]
    w = np.uint32(seed >> 7)
    for _ in range(iters):
        for s in range(8):
            for u in range(8):
                nc = m[(s + u) % 8] | w | na[u]
                t = nc + np.uint32(0x01010101)
                na[u] = (nc & na[u]) | (~nc & t)
    want = np.zeros(256, dtype=np.uint32)
    for x in na:
        want ^= x
    got = out.cpu().numpy().view(np.uint32).reshape(blocks, 256)
    for b in range(blocks):
        np.testing.assert_array_equal(got[b], want)
    with pytest.raises(RuntimeError, match="bad arguments"):
        L.check(lib.fh_probe_valu(L.ptr(out), 0, iters, ctypes.byref(ops), L.stream()), "fh_probe_valu")


def test_kat_levels(oracle):
    from faster.comparison import jaro_winkler_gpu

    with open(os.path.join(HERE, "golden", "jw_kat.json")) as f:
        kat = json.load(f)
    A = [c["s1"] for c in kat["cases"]]
    B = [c["s2"] for c in kat["cases"]]
    for lo, hi in ((0.88, 0.94), (0.8, 0.9), (0.95, 0.99)):
        got = jaro_winkler_gpu(A, B, 0, 0.1, lo, hi)
        ref = oracle.jaro_winkler_all_pairs(A, B, 0, 0.1, lo, hi)
        _eq_lists(got, ref, f"kat {lo},{hi}")


@pytest.mark.parametrize("seed,alpha,lo_len,hi_len", [
    (0, "abcde", 0, 8), (1, "ab", 1, 16), (2, "abcdefghijklmnopqrstuvwxyz", 2, 20),
    (3, "abc", 14, 40), (4, "xyz", 30, 70), (5, "aAbB 0'-", 0, 24)])
def test_all_pairs_levels(oracle, seed, alpha, lo_len, hi_len):
    """Random strings over small alphabets (many matches), all length classes
    (empty, 16-bit, 32-bit, 64-bit masks and the generic > 64 path)."""
    from faster.comparison import jaro_winkler_gpu

    rng = np.random.default_rng(seed)
    A = _words(rng, 300, alpha, lo_len, hi_len)
    B = _words(rng, 340, alpha, lo_len, hi_len)
    for k in range(0, 300, 9):  # shared values and near copies
        B[k] = A[k]
    got = jaro_winkler_gpu(A, B, 17)
    ref = oracle.jaro_winkler_all_pairs(A, B, 17)
    _eq_lists(got, ref, f"seed {seed}")


@pytest.mark.parametrize("mode", ["link", "dedup"])
@pytest.mark.parametrize("lens,alpha,n", [((1, 2, 3), "abcdef", 600), ((4, 5), "abcdefgh", 3000),
                                          ((5, 6, 7, 8), "abcd", 2600),
                                          ((10, 11, 15, 16), "abc", 1800), ((31, 32, 33), "ab", 900),
                                          ((63, 64, 65), "ab", 500)])
def test_mask_kind_boundaries(oracle, mode, lens, alpha, n):
    """Values at the length boundaries of the mask kinds (5 strings per dword
    up to 5 bytes, 4 up to 7, 3 up to 10, 2 up to 15, one 32-bit mask up to
    32, one 64-bit mask up to 64, the byte-wise path beyond), over alphabets
    small enough that every class holds near copies and long shared
    prefixes, with classes of more than one string-1 wave and more than one
    string-2 tile (1,280 five-byte strings fill a tile of the five-string
    kind)."""
    from faster.comparison import jaro_winkler_unique_gpu
    from faster.deduplication import jaro_winkler_dedup_gpu

    rng = np.random.default_rng(sum(lens) * 7 + len(alpha))

    def draw(m):
        return ["".join(alpha[k] for k in rng.integers(0, len(alpha), int(rng.choice(lens)))) for _ in range(m)]

    def near(s):
        b = list(s)
        b[int(rng.integers(0, len(b)))] = alpha[int(rng.integers(0, len(alpha)))]
        return "".join(b)

    A = draw(n)
    B = draw(n) + [near(s) for s in A[: n // 3]]
    if mode == "link":
        _eq_lists(jaro_winkler_unique_gpu(A, B), oracle.jaro_winkler_unique(A, B), f"link {lens}")
    else:
        S = sorted(set(A + B[n:]))
        _eq_lists(jaro_winkler_dedup_gpu(S), oracle.jaro_winkler_dedup(S), f"dedup {lens}")


def test_bytes_on_one_side_only(oracle):
    """Bytes that occur on one side only share the alphabet's last code with
    the positions past a value's end; they must match nothing, in the
    comparison pass and in the survivor scoring (prefix form: both values'
    row codes), while the shared bytes still match."""
    from faster.comparison import jaro_winkler_unique_gpu

    rng = np.random.default_rng(23)
    A = _words(rng, 1500, "abcdxy", 1, 12)
    B = _words(rng, 1500, "abcduv", 1, 12)
    B += [s.replace("x", "u").replace("y", "v") for s in A[:500]]  # near copies across the two alphabets
    B += [s for s in A[500:700] if "x" not in s and "y" not in s]  # exact copies over shared bytes
    _eq_lists(jaro_winkler_unique_gpu(A, B), oracle.jaro_winkler_unique(A, B), "one-sided bytes")


@pytest.mark.parametrize("n_sym", [60, 120, 0])
def test_large_alphabets(oracle, n_sym):
    """Alphabets of more byte values than a 256-dword tile holds rows for, so
    the per-lane form's tiles narrow to 128, 64 and (n_sym = 0: ASCII 1-127,
    two-byte characters covering every lead and continuation byte, and one
    three-byte character per lead byte, about 237 byte values on both sides)
    32 dwords; linkage and deduplication, with near copies."""
    from faster import _lib, _ops
    from faster.comparison import jaro_winkler_unique_gpu
    from faster.deduplication import jaro_winkler_dedup_gpu

    rng = np.random.default_rng(n_sym)
    if n_sym:
        syms = [chr(c) for c in range(1, 128)] + [chr(c) for c in range(0xA0, 0x800)]
        syms = [syms[int(k)] for k in rng.choice(len(syms), n_sym, replace=False)]
    else:
        syms = ([chr(c) for c in range(1, 128)] + [chr(((2 + k % 30) << 6) | ((7 * k) % 64)) for k in range(128)]
                + [chr((j << 12) | (0x900 if j == 0 else 0x100)) for j in range(16)])

    def draw(m):
        return ["".join(syms[int(k)] for k in rng.integers(0, len(syms), int(rng.integers(1, 21)))) for _ in range(m)]

    def near(s):
        b = list(s)
        b[int(rng.integers(0, len(b)))] = syms[int(rng.integers(0, len(syms)))]
        return "".join(b)

    A = draw(500)
    B = draw(400) + [near(s) for s in A[:200]] + A[200:260]
    _eq_lists(jaro_winkler_unique_gpu(A, B), oracle.jaro_winkler_unique(A, B), f"link {n_sym} symbols")
    dev = _lib.device()
    ca, oa = _pack(A, dev)
    cb, ob = _pack(B, dev)
    _, st = _ops.jw_levels(ca, oa, len(A), cb, ob, len(B), _lib.FH_MODE_LINK, 0.1, 0.88, 0.94, want_stats=True)
    assert st["alphabet"] > (220 if n_sym == 0 else 55)  # tiles of 32 dwords / narrower than 256
    assert st["method"] == _lib._METHODS["tiles"]
    S = sorted(set(A + B[400:]))
    _eq_lists(jaro_winkler_dedup_gpu(S), oracle.jaro_winkler_dedup(S), f"dedup {n_sym} symbols")


@pytest.mark.parametrize("lo,hi", [(0.88, 0.94), (0.0, 0.5)])
def test_values_of_any_length(oracle, lo, hi):
    """Values over 1,024 bytes (scored with their flags in global scratch)
    beside short, long (65-1,024 bytes) and empty ones, in linkage and
    deduplication: near copies of the long values reach both levels, and
    thresholds at 0 make every pair a hit, so each pair with a long member is
    emitted exactly once."""
    from faster.comparison import jaro_winkler_gpu
    from faster.deduplication import jaro_winkler_dedup_gpu

    rng = np.random.default_rng(13)

    def near(s, k):
        b = list(s)
        for i in rng.integers(0, len(b), k):
            b[int(i)] = "abcd"[int(rng.integers(0, 4))]
        return "".join(b)

    huge = _words(rng, 6, "abcd", 1025, 2600)
    A = _words(rng, 30, "abcd", 0, 12) + huge + _words(rng, 3, "abcd", 65, 300) + [""]
    B = _words(rng, 30, "abcd", 0, 12) + [near(h, 3) for h in huge[:4]] + [huge[4] + "ab", huge[5][:-1]] + \
        _words(rng, 2, "abcd", 900, 1100) + ["", huge[0][:40]]
    _eq_lists(jaro_winkler_gpu(A, B, 5, 0.1, lo, hi), oracle.jaro_winkler_all_pairs(A, B, 5, 0.1, lo, hi),
              f"link any length {lo},{hi}")
    S = A[25:] + B[28:]
    _eq_lists(jaro_winkler_dedup_gpu(S, 0.1, lo, hi), oracle.jaro_winkler_dedup(S, 0.1, lo, hi),
              f"dedup any length {lo},{hi}")


@pytest.mark.parametrize("lo,hi", [(0.88, 0.94), (0.7, 0.8), (0.94, 0.88), (0.0, 0.5), (0.99, 1.0)])
def test_threshold_edge_cases(oracle, lo, hi):
    """Thresholds at and below 0 (every pair is a hit), inverted, and 1.0."""
    from faster.comparison import jaro_winkler_gpu

    rng = np.random.default_rng(11)
    A = _words(rng, 60, "abc", 0, 6)
    B = _words(rng, 50, "abc", 0, 6)
    _eq_lists(jaro_winkler_gpu(A, B, 0, 0.1, lo, hi), oracle.jaro_winkler_all_pairs(A, B, 0, 0.1, lo, hi),
              f"thr {lo},{hi}")


def test_unique_link(oracle):
    from faster.comparison import jaro_winkler_unique_gpu

    rng = np.random.default_rng(7)
    A = _words(rng, 900, "abcdef", 1, 10)
    B = _words(rng, 800, "abcdef", 1, 10)
    for k in range(0, 800, 5):
        B[k] = A[(k * 3) % 900]
    for k in range(0, 900, 4):
        A[k] = A[0] if k % 8 else A[k]
    _eq_lists(jaro_winkler_unique_gpu(A, B), oracle.jaro_winkler_unique(A, B), "unique")


@pytest.mark.parametrize("n_ua,missing", [(40, False), (40, True), (2900, False)])
def test_expand_link_sorted(n_ua, missing):
    """fh_expand_link_sorted = every record pair of the unique-value hits,
    ascending: with heavily repeated A values (the per-value partner lists
    copied per row, rows without a value skipped) and with few repeats (the
    expansion sorted directly), against NumPy."""
    from faster import _lib, _ops

    rng = np.random.default_rng(n_ua + missing)
    dev = _lib.device()
    n_a, n_b, n_ub = 3000, 2000, 300
    val_a = rng.integers(0, n_ua, n_a)
    if missing:
        val_a[rng.random(n_a) < 0.1] = -1
    val_b = rng.integers(0, n_ub, n_b)

    def groups(vals, n_u):
        rows = [np.nonzero(vals == u)[0] for u in range(n_u)]
        starts = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
        return starts, np.concatenate(rows).astype(np.int64), rows

    sa, ra, rows_a = groups(val_a, n_ua)
    sb, rb, rows_b = groups(val_b, n_ub)
    ids = rng.choice(n_ua * n_ub, size=min(900, n_ua * n_ub // 3), replace=False)
    lvl = rng.integers(1, 3, ids.size)
    hits = (ids * 4 + lvl).astype(np.int64)
    want = []
    for i, l in zip(ids.tolist(), lvl.tolist()):
        ua, ub = divmod(i, n_ub)
        for x in rows_a[ua]:
            want.append((int(x) * n_b + rows_b[ub]) * 4 + l)
    want = np.sort(np.concatenate(want)) if want else np.zeros(0, np.int64)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    got = _ops.expand_link_sorted(t(hits), n_ua, n_ub, t(sa), t(ra), n_a, t(sb), t(rb), n_b)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_exact_link(oracle):
    from faster.comparison import exact_gpu

    rng = np.random.default_rng(8)
    A = rng.integers(0, 50, 3000)
    B = rng.integers(20, 80, 2500)
    _eq_lists(exact_gpu(A, B), oracle.exact(A, B), "exact")
    Sa = np.array(_words(rng, 500, "ab", 0, 3), dtype=object)
    Sb = np.array(_words(rng, 400, "ab", 0, 3), dtype=object)
    _eq_lists(exact_gpu(Sa, Sb), oracle.exact(Sa, Sb), "exact str")


def test_exact_link_load_balance(oracle):
    """Expansion tiles that span one huge group (3,000 B rows per A row), and
    tiles whose owners are spread over thousands of empty rows (beyond the
    LDS-staged row window), against the oracle."""
    from faster.comparison import exact_gpu
    from faster.deduplication import exact_dedup_gpu

    rng = np.random.default_rng(21)
    A = np.concatenate([rng.integers(10 ** 6, 2 * 10 ** 6, 19000), np.full(1000, 7)])
    rng.shuffle(A)
    B = np.concatenate([np.full(3000, 7), rng.integers(0, 100, 2000), rng.integers(10 ** 6, 2 * 10 ** 6, 3000)])
    rng.shuffle(B)
    _eq_lists(exact_gpu(A, B), oracle.exact(A, B), "exact load balance")
    E = np.concatenate([rng.integers(10 ** 6, 2 * 10 ** 6, 30000), np.full(900, 3)])
    rng.shuffle(E)
    _eq_lists(exact_dedup_gpu(E), oracle.exact_dedup(E), "exact dedup load balance")


def test_sort_and_scan_primitives():
    import faster._ops as ops

    from faster import _lib

    L = _lib.load()
    rng = np.random.default_rng(9)
    for n in (0, 1, 5, 4095, 4096, 4097, 100_000, 1_500_000):
        x = rng.integers(0, 2 ** 46, n, dtype=np.int64)
        t = torch.from_numpy(x).cuda()
        np.testing.assert_array_equal(ops.sort_keys(t, 2 ** 46).cpu().numpy(), np.sort(x))
    # Both C entry points at odd and even pass counts (8, 16, 36 and 64 bits),
    # with few distinct digits (bits 0..2 only) and with every digit in use.
    for bits in (8, 16, 36, 64):
        for hi in (8, 2 ** bits if bits < 64 else None):
            x = rng.integers(0, hi, 50_000, dtype=np.uint64) if hi else rng.integers(0, 2 ** 63, 50_000).astype(
                np.uint64) * np.uint64(2) + np.uint64(1)
            want = np.sort(x)
            print(f"sort bits={bits} hi={hi}", flush=True)
            a = torch.from_numpy(x.view(np.int64)).cuda()
            ops.check(L.fh_sort_u64(ops.ptr(a), ops._i64(a.numel()), 0, bits, ops.stream()), "fh_sort_u64")
            np.testing.assert_array_equal(a.cpu().numpy().view(np.uint64), want)
            a = torch.from_numpy(x.view(np.int64)).cuda()
            out = torch.empty_like(a)
            ops.check(L.fh_sort_u64_into(ops.ptr(a), ops.ptr(out), ops._i64(a.numel()), 0, bits, ops.stream()),
                      "fh_sort_u64_into")
            np.testing.assert_array_equal(out.cpu().numpy().view(np.uint64), want)


def test_set_ops():
    from faster.search import intersect, reduce, setdiff

    rng = np.random.default_rng(10)
    a = np.unique(rng.integers(0, 10 ** 6, 50_000))
    b = np.unique(rng.integers(0, 10 ** 6, 70_000))
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    np.testing.assert_array_equal(_np(intersect(ta, tb)), np.intersect1d(a, b))
    np.testing.assert_array_equal(_np(setdiff(ta, tb)), np.setdiff1d(a, b))
    assert intersect(ta[:0], tb).dtype == torch.uint64
    assert setdiff(ta[:0], tb) is not None
    np.testing.assert_array_equal(_np(reduce(setdiff, [tb[:10], tb[10:20]], ta)),
                                  np.setdiff1d(np.setdiff1d(a, b[:10]), b[10:20]))


def _frames(n_a, n_b, seed):
    from faster.synthetic import linkage_tables

    return linkage_tables(n_a, n_b, seed=seed)


def test_comparison_fit_small(oracle):
    import faster

    dfa, dfb, _ = _frames(1500, 1700, 3)
    vf = ["first_name", "last_name", "street"]
    comp = faster.Comparison(dfa, dfb, vf, vf, ["house_number"], ["house_number"])
    comp.fit()
    per_var, indices, counts = oracle.comparison_fit([dfa[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf],
                                                     [dfa["house_number"].to_numpy()], [dfb["house_number"].to_numpy()])
    for v in range(4):
        _eq_lists(comp.Levels[v], per_var[v], f"var {v}")
    _eq_lists(comp.Indices, indices, "Indices")
    np.testing.assert_array_equal(comp.Counts, counts)
    with pytest.raises(Exception, match="already been fitted"):
        comp.fit()


def test_comparison_fit_many_variables(oracle):
    """Six fuzzy variables (729 patterns) in one merge that mixes both owner
    paths: the first variable owns 486 codes and takes the general path
    (flag, compact, radix sort by code); the others own at most 162 and take
    the one-pass ranked scatter."""
    import faster

    dfa, dfb, _ = _frames(700, 800, 5)
    base = ["first_name", "last_name", "street"]
    for df in (dfa, dfb):
        for v in base:
            df[v + "_r"] = df[v].str[::-1]
    vf = base + [v + "_r" for v in base]
    comp = faster.Comparison(dfa, dfb, vf, vf)
    comp.fit()
    per_var, indices, counts = oracle.comparison_fit([dfa[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf])
    print(f"many variables: patterns={len(counts)} union={int(np.sum(counts[1:]))}", flush=True)
    assert len(counts) == 3 ** 6
    for v in range(6):
        _eq_lists(comp.Levels[v], per_var[v], f"var {v}")
    _eq_lists(comp.Indices, indices, "Indices")
    np.testing.assert_array_equal(comp.Counts, counts)


@pytest.mark.parametrize("empty", ["general", "one_pass"])
def test_merge_with_empty_lists(oracle, empty):
    """Empty level lists in both owner paths: a fuzzy variable whose two sides
    share no character (no pair above the threshold) as the variable with
    486 codes (general path) or as a one-pass variable, and an exact
    variable with no equal pair."""
    import faster

    dfa, dfb, _ = _frames(300, 350, 12)
    base = ["first_name", "last_name", "street"]
    for df in (dfa, dfb):
        for v in base:
            df[v + "_r"] = df[v].str[::-1]
    dfa["none"] = ["x" * (1 + i % 5) for i in range(len(dfa))]
    dfb["none"] = ["y" * (1 + i % 5) for i in range(len(dfb))]
    dfa["hn"] = np.arange(len(dfa))
    dfb["hn"] = np.arange(len(dfb)) + 10_000
    vf = ["none"] + base + ["last_name_r", "street_r"] if empty == "general" else \
        base + ["first_name_r", "none", "street_r"]
    comp = faster.Comparison(dfa, dfb, vf, vf, ["hn"], ["hn"])
    comp.fit()
    per_var, indices, counts = oracle.comparison_fit([dfa[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf],
                                                     [dfa["hn"].to_numpy()], [dfb["hn"].to_numpy()])
    k = vf.index("none")
    assert all(len(x) == 0 for x in per_var[k]) and len(per_var[-1][0]) == 0
    _eq_lists(comp.Indices, indices, "Indices")
    np.testing.assert_array_equal(comp.Counts, counts)


def test_pattern_merge_capacity_and_bad_keys():
    """fh_pattern_merge through the C ABI: with a capacity below the union it
    returns FH_OK, reports the full union size and writes nothing past the
    capacity (for a one-pass and a general-path merge); a level key with
    level 0, or a level not below its variable's level count, is rejected
    with FH_ERR_ARG."""
    import ctypes

    from faster import _lib

    L = _lib.load()
    rng = np.random.default_rng(4)
    dev = torch.device("cuda")

    def lists(nv, levels):
        out = []
        for v in range(nv):
            ids = np.unique(rng.integers(0, 5000, 1500))
            lv = rng.integers(1, levels[v], len(ids))
            out.append(torch.as_tensor(ids * 4 + lv, dtype=torch.int64, device=dev))
        return out

    def merge(keys, levels, capacity, total):
        nv = len(keys)
        ptrs = (ctypes.c_void_p * nv)(*[k.data_ptr() for k in keys])
        sizes = (ctypes.c_int64 * nv)(*[k.numel() for k in keys])
        lv = (ctypes.c_int32 * nv)(*levels)
        ids = torch.full((total,), -7, dtype=torch.int64, device=dev)
        n_pat = int(np.prod(levels))
        offs = (ctypes.c_int64 * (n_pat + 1))()
        nu = ctypes.c_int64(-1)
        st = L.fh_pattern_merge(nv, ptrs, sizes, lv, _lib.ptr(ids), ctypes.c_int64(capacity), offs, ctypes.byref(nu),
                                _lib.stream())
        torch.cuda.synchronize()
        return st, nu.value, ids.cpu().numpy(), list(offs)

    for levels in ([3, 3, 2], [3] * 6):
        keys = lists(len(levels), levels)
        total = sum(k.numel() for k in keys)
        st, n_union, ids_full, offs_full = merge(keys, levels, total, total)
        assert st == 0 and n_union == offs_full[-1]
        union = np.unique(np.concatenate([k.cpu().numpy() >> 2 for k in keys]))
        assert n_union == len(union)
        cap = n_union // 2
        st, nu2, ids_small, _ = merge(keys, levels, cap, total)
        assert st == 0 and nu2 == n_union
        assert np.all(ids_small[cap:] == -7)
        bad = [k.clone() for k in keys]
        bad[1][5] = (bad[1][5] >> 2) * 4  # level 0
        assert merge(bad, levels, total, total)[0] == 2
        assert b"level" in L.fh_last_error()
        bad = [k.clone() for k in keys]
        bad[-1][3] = (bad[-1][3] >> 2) * 4 + levels[-1]  # level == level count
        assert merge(bad, levels, total, total)[0] == 2



@pytest.mark.parametrize("levels,sizes", [([3, 3, 3, 2], [3_000_000, 20_000, 100_000, 1_500_000]),
                                          ([3] * 6, [2_000_000, 15_000, 300_000, 60_000, 1_000_000, 5_000])])
def test_pattern_merge_lists_of_very_different_sizes(levels, sizes):
    """Lists of 5k to 3M keys over one id range, so that a tile of one list
    meets ranges of the others shorter than the tile, a few LDS chunks long,
    and far longer (searched in place); the second case also has variables
    owning more than 256 codes (the general owner path). Ids and pattern
    offsets equal the merge's definition computed with NumPy: the pattern
    code of a pair is the sum over variables of its level times the
    variable's stride (first variable slowest, level 0 where unlisted), and
    the output lists the union's ids by code, ascending within a code."""
    import ctypes

    from faster import _lib

    L = _lib.load()
    rng = np.random.default_rng(11)
    dev = torch.device("cuda")
    id_range = 100_000_000
    host = []
    for n, nl in zip(sizes, levels):
        ids = np.unique(rng.integers(0, id_range, int(n * 1.05)))
        ids = np.sort(rng.choice(ids, n, replace=False))
        host.append((ids, rng.integers(1, nl, n)))
    keys = [torch.as_tensor(ids * 4 + lv, dtype=torch.int64, device=dev) for ids, lv in host]
    nv = len(keys)
    total = sum(sizes)
    n_pat = int(np.prod(levels))
    ptrs = (ctypes.c_void_p * nv)(*[k.data_ptr() for k in keys])
    c_sizes = (ctypes.c_int64 * nv)(*sizes)
    lv = (ctypes.c_int32 * nv)(*levels)
    out = torch.full((total,), -7, dtype=torch.int64, device=dev)
    offs = (ctypes.c_int64 * (n_pat + 1))()
    nu = ctypes.c_int64(-1)
    st = L.fh_pattern_merge(nv, ptrs, c_sizes, lv, _lib.ptr(out), ctypes.c_int64(total), offs, ctypes.byref(nu),
                            _lib.stream())
    torch.cuda.synchronize()
    assert st == 0, L.fh_last_error()
    union = np.unique(np.concatenate([ids for ids, _ in host]))
    code = np.zeros(len(union), dtype=np.int64)
    stride = 1
    for v in range(nv - 1, -1, -1):
        ids, lvl = host[v]
        code[np.searchsorted(union, ids)] += lvl * stride
        stride *= levels[v]
    order = np.lexsort((union, code))
    assert nu.value == len(union)
    np.testing.assert_array_equal(out.cpu().numpy()[:len(union)], union[order])
    expect = np.concatenate([[0], np.cumsum(np.bincount(code, minlength=n_pat))])
    np.testing.assert_array_equal(np.asarray(list(offs)), expect)


def test_pattern_merge_listed_exact_ids_equal_plain_merge():
    """fh_pattern_merge_listed through the C ABI: exact variables whose lists
    are the equal pairs of per-row value ids (negative ids: rows without a
    value), given with those ids, merge to the same ids and offsets as
    fh_pattern_merge over the same lists, with the exact variables first,
    between and last in pattern order; a row-id variable with 3 levels is
    rejected."""
    import ctypes

    from faster import _lib

    L = _lib.load()
    rng = np.random.default_rng(17)
    dev = torch.device("cuda")
    n_a, n_b = 2000, 3000
    ga = rng.integers(-1, 40, n_a)
    gb = rng.integers(-1, 40, n_b)
    ii, jj = np.nonzero((ga[:, None] == gb[None, :]) & (ga[:, None] >= 0))
    exact_keys = (ii.astype(np.int64) * n_b + jj) * 4 + 1
    ga2 = rng.integers(0, 7, n_a)
    gb2 = rng.integers(0, 7, n_b)
    ii2, jj2 = np.nonzero(ga2[:, None] == gb2[None, :])
    exact2_keys = (ii2.astype(np.int64) * n_b + jj2) * 4 + 1

    def fuzzy(m):
        ids = np.unique(rng.integers(0, n_a * n_b, m))
        return ids * 4 + rng.integers(1, 3, len(ids))

    tens = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int64), device=dev)
    ids_t = {"e": (tens(ga), tens(gb)), "e2": (tens(ga2), tens(gb2))}
    lists = {"e": exact_keys, "e2": exact2_keys, "f1": fuzzy(400_000), "f2": fuzzy(30_000)}
    for order in (["e", "f1", "f2"], ["f1", "e", "f2", "e2"], ["f1", "f2", "e"]):
        keys = [tens(lists[x]) for x in order]
        levels = [2 if x.startswith("e") else 3 for x in order]
        nv = len(keys)
        total = sum(k.numel() for k in keys)
        n_pat = int(np.prod(levels))
        ptrs = (ctypes.c_void_p * nv)(*[k.data_ptr() for k in keys])
        sizes = (ctypes.c_int64 * nv)(*[k.numel() for k in keys])
        lv = (ctypes.c_int32 * nv)(*levels)
        outs = []
        for listed in (False, True):
            out = torch.full((total,), -7, dtype=torch.int64, device=dev)
            offs = (ctypes.c_int64 * (n_pat + 1))()
            nu = ctypes.c_int64(-1)
            if listed:
                pa = (ctypes.c_void_p * nv)(*[ids_t[x][0].data_ptr() if x in ids_t else 0 for x in order])
                pb = (ctypes.c_void_p * nv)(*[ids_t[x][1].data_ptr() if x in ids_t else 0 for x in order])
                st = L.fh_pattern_merge_listed(nv, ptrs, sizes, lv, pa, pb, ctypes.c_int64(n_b), _lib.ptr(out),
                                               ctypes.c_int64(total), offs, ctypes.byref(nu), _lib.stream())
            else:
                st = L.fh_pattern_merge(nv, ptrs, sizes, lv, _lib.ptr(out), ctypes.c_int64(total), offs,
                                        ctypes.byref(nu), _lib.stream())
            torch.cuda.synchronize()
            assert st == 0, L.fh_last_error()
            outs.append((nu.value, out.cpu().numpy()[:nu.value], list(offs)))
        assert outs[0][0] == outs[1][0]
        np.testing.assert_array_equal(outs[0][1], outs[1][1])
        assert outs[0][2] == outs[1][2]
        union = np.unique(np.concatenate([lists[x] >> 2 for x in order]))
        assert outs[1][0] == len(union)
    bad = (ctypes.c_int32 * 2)(3, 3)
    k2 = [tens(lists["e"]), tens(lists["f1"])]
    ptrs = (ctypes.c_void_p * 2)(*[k.data_ptr() for k in k2])
    sizes = (ctypes.c_int64 * 2)(*[k.numel() for k in k2])
    pa = (ctypes.c_void_p * 2)(ids_t["e"][0].data_ptr(), 0)
    pb = (ctypes.c_void_p * 2)(ids_t["e"][1].data_ptr(), 0)
    out = torch.empty(sum(k.numel() for k in k2), dtype=torch.int64, device=dev)
    offs = (ctypes.c_int64 * 10)()
    nu = ctypes.c_int64(-1)
    assert L.fh_pattern_merge_listed(2, ptrs, sizes, bad, pa, pb, ctypes.c_int64(n_b), _lib.ptr(out),
                                     ctypes.c_int64(out.numel()), offs, ctypes.byref(nu), _lib.stream()) != 0

def test_deduplication_fit(oracle):
    import faster
    from faster.synthetic import dedup_table

    df, _ = dedup_table(1200, seed=4)
    df["first_name"] = [s[:3] for s in df["first_name"]]  # many identical values
    vf = ["first_name", "last_name", "street"]
    d = faster.Deduplication(df, vf, ["birth_year"])
    d.fit()
    per_var, indices, counts = oracle.deduplication_fit([df[v].to_numpy() for v in vf], [df["birth_year"].to_numpy()])
    for v in range(4):
        _eq_lists(d.Levels[v], per_var[v], f"dedup var {v}")
    _eq_lists(d.Indices, indices, "dedup Indices")
    np.testing.assert_array_equal(d.Counts, counts)


def test_dedup_edge_values(oracle):
    from faster.deduplication import exact_dedup_gpu, jaro_winkler_dedup_gpu

    rng = np.random.default_rng(12)
    S = _words(rng, 400, "ab", 0, 5) + ["a", "a", "", "", "x" * 70, "x" * 69 + "y", "x" * 70]
    _eq_lists(jaro_winkler_dedup_gpu(S), oracle.jaro_winkler_dedup(S), "dedup edge")
    E = rng.integers(0, 5, 700)
    _eq_lists(exact_dedup_gpu(E), oracle.exact_dedup(E), "exact dedup")


def test_em_matches_oracle(oracle):
    import faster

    rng = np.random.default_rng(13)
    counts = np.concatenate([[10 ** 10], rng.integers(0, 10 ** 6, 53)]).astype(np.int64)
    for tol, mi in ((1e-4, 5000), (1e-8, 5000), (1e-6, 3)):
        np.random.seed(0)
        est = faster.Estimation(3, 1, counts)
        est.fit(Tolerance=tol, Max_Iter=mi)
        np.random.seed(0)
        lam, Pi, it, conv, ksi = oracle.em_fit([3, 3, 3, 2], counts, tol, mi)
        assert est.Iterations == it
        assert est.Lambda == pytest.approx(lam, rel=1e-9)
        for m in range(2):
            for k in range(4):
                np.testing.assert_allclose(est.Pi[m][k], Pi[m][k], rtol=1e-9)
        np.testing.assert_allclose(est.Ksi, ksi, rtol=1e-9, atol=1e-300)


def test_pipeline_linkage_and_posteriors(oracle):
    import faster

    dfa, dfb, _ = _frames(800, 900, 5)
    vf = ["first_name", "last_name", "street"]
    comp = faster.Comparison(dfa, dfb, vf, vf, ["house_number"], ["house_number"])
    comp.fit()
    np.random.seed(0)
    est = faster.Estimation(3, 1, comp.Counts)
    est.fit()
    link = faster.Linkage(dfa, dfb, comp.Indices, est.Ksi)
    out = link.transform(0.85)
    ra, rb = oracle.linkage_pairs([_np(x) for x in comp.Indices], est.Ksi, len(dfb), 0.85)
    np.testing.assert_array_equal(out["Index_A"].to_numpy(), ra)
    np.testing.assert_array_equal(out["Index_B"].to_numpy(), rb)
    post = comp.posteriors(est.Ksi).cpu().numpy()
    ref = np.concatenate([np.full(x.numel(), est.Ksi[k + 1]) for k, x in enumerate(comp.Indices)])
    np.testing.assert_array_equal(post, ref)


def test_linkage_pattern_zero_compat():
    """Pattern 0 above the threshold: the reference links Indices[-1] (the
    last pattern's pairs, faster/linkage.py:54); Compat=False skips it."""
    import faster

    dfa, dfb, _ = _frames(300, 320, 6)
    vf = ["first_name", "last_name"]
    comp = faster.Comparison(dfa, dfb, vf, vf)
    comp.fit()
    last = _np(comp.Indices[-1])
    assert last.size > 0
    ksi = np.zeros(len(comp.Indices) + 1)
    ksi[0] = 1.0
    out = faster.Linkage(dfa, dfb, comp.Indices, ksi).transform(0.85)
    np.testing.assert_array_equal(out["Index_A"].to_numpy(), last // len(dfb))
    np.testing.assert_array_equal(out["Index_B"].to_numpy(), last % len(dfb))
    with pytest.raises(Exception, match="No pair of observations"):
        faster.Linkage(dfa, dfb, comp.Indices, ksi).transform(0.85, Compat=False)
    ksi[1] = 1.0
    out = faster.Linkage(dfa, dfb, comp.Indices, ksi).transform(0.85, Compat=False)
    first = _np(comp.Indices[0])
    np.testing.assert_array_equal(out["Index_B"].to_numpy(), first % len(dfb))


@pytest.mark.parametrize("which", ["dedup", "link"])
def test_rowid_exact_mode(oracle, which):
    """Two exact variables kept as row ids: Counts equal the oracle's (listed)
    Counts, fuzzy-pattern lists are identical, every lazy exact-only list
    materialises to the oracle's list, and Linkage.transform links the pairs
    of an exact-only pattern."""
    import faster
    from faster.patterns import RowIdPairs
    from faster.synthetic import dedup_table, linkage_tables

    vf = ["first_name", "last_name"]
    if which == "dedup":
        df, _ = dedup_table(900, seed=21)
        df["birth_year"] = (df["birth_year"] % 7).astype(np.int16)
        df["flag"] = (np.arange(len(df)) % 3).astype(np.int64)
        obj = faster.Deduplication(df, vf, ["birth_year", "flag"])
        obj.fit(Exact_Mode="rowid")
        _, indices, counts = oracle.deduplication_fit([df[v].to_numpy() for v in vf],
                                                      [df["birth_year"].to_numpy(), df["flag"].to_numpy()])
    else:
        dfa, dfb, _ = linkage_tables(700, 800, seed=22)
        for d, k in ((dfa, 1), (dfb, 2)):
            d["house_number"] = d["house_number"] % 5
            d["flag"] = (np.arange(len(d)) * k % 3).astype(np.int64)
        ve = ["house_number", "flag"]
        obj = faster.Comparison(dfa, dfb, vf, vf, ve, ve)
        obj.fit(Exact_Mode="rowid")
        _, indices, counts = oracle.comparison_fit([dfa[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf],
                                                   [dfa[v].to_numpy() for v in ve], [dfb[v].to_numpy() for v in ve])
    assert obj.Exact_Mode == "rowid"
    np.testing.assert_array_equal(obj.Counts, counts)
    lazy = []
    for k, (g, r) in enumerate(zip(obj.Indices, indices)):
        if isinstance(g, RowIdPairs):
            lazy.append(k + 1)
            assert g.numel() == len(r)
            np.testing.assert_array_equal(_np(g.materialize()), np.asarray(r, dtype=np.int64),
                                          err_msg=f"exact-only pattern {k + 1}")
        else:
            np.testing.assert_array_equal(_np(g), np.asarray(r, dtype=np.int64), err_msg=f"pattern {k + 1}")
    assert lazy == [1, 2, 3]
    if which == "link":
        ksi = np.zeros(len(counts))
        ksi[3] = 1.0  # exact digits (1, 1), both names disagree
        out = faster.Linkage(dfa, dfb, obj.Indices, ksi).transform(0.85, Compat=False)
        got = np.sort(out["Index_A"].to_numpy() * len(dfb) + out["Index_B"].to_numpy())
        np.testing.assert_array_equal(got, np.asarray(indices[2], dtype=np.int64))


def test_blocked_linkage_matches_per_block_loop(oracle):
    """One persistent graph-captured launch over all blocks gives the per-block
    Counts of the per-block recipe, bit for bit, on every graph replay (with
    eager work between replays), without the graph, and as the plain
    per-block launch loop over the same plan."""
    from faster.blocking import BlockedLinkage, per_block_loop
    from faster.synthetic import blocked_tables

    dfa, dfb, ba, bb, sizes = blocked_tables(9000, n_blocks=40, seed=3, kmin=5, kmax=3000)
    vf = ["first_name", "last_name", "street"]
    bl = BlockedLinkage(dfa, dfb, ba, bb, vf, vf, ["house_number"], ["house_number"])
    ref, _ = per_block_loop(dfa, dfb, ba, bb, vf, ["house_number"])
    for k in range(3):
        bl.fit(use_graph=True)
        np.testing.assert_array_equal(bl.Block_Counts, ref, err_msg=f"replay {k}")
        _ = torch.arange(1000, device="cuda").sum().item()  # eager work between replays
    bl.fit(use_graph=False)
    np.testing.assert_array_equal(bl.Block_Counts, ref)
    # The plain per-block launch loop over the same plan, twice (its queue
    # heads and table are reset by every run).
    for k in range(2):
        bl.run_per_block()
        np.testing.assert_array_equal(bl.collect(), ref, err_msg=f"per-block launch loop, run {k}")
    # The largest block against the oracle directly.
    b = int(np.argmax(sizes))
    ra, rb = np.nonzero(ba == b)[0], np.nonzero(bb == b)[0]
    _, _, counts = oracle.comparison_fit([dfa[v].to_numpy()[ra] for v in vf], [dfb[v].to_numpy()[rb] for v in vf],
                                         [dfa["house_number"].to_numpy()[ra]], [dfb["house_number"].to_numpy()[rb]])
    np.testing.assert_array_equal(bl.Block_Counts[b], counts)
    # Autotuned configuration gives the same counts.
    best, res = bl.autotune(space=dict(tile_width=[128, 256], tiles_per_item=[1, 4], wgs_per_cu=[1, 2]))
    assert res
    bl.fit()
    np.testing.assert_array_equal(bl.Block_Counts, ref)


def test_blocked_linkage_large_alphabet(oracle):
    """Blocks whose values use about 80 byte values on both sides (mixed
    case, digits, punctuation): the plan narrows its tiles below the
    requested 256 dwords so the tile rows fit the LDS, and the batched launch,
    the per-block launch loop, the per-block recipe and the oracle agree."""
    from faster.blocking import BlockedLinkage, per_block_loop

    rng = np.random.default_rng(17)
    syms = [chr(c) for c in range(0x21, 0x7F)][:80]
    n_blocks, n = 6, 2400

    def word():
        return "".join(syms[int(k)] for k in rng.integers(0, len(syms), int(rng.integers(2, 14))))

    def near(w):
        b = list(w)
        b[int(rng.integers(0, len(b)))] = syms[int(rng.integers(0, len(syms)))]
        return "".join(b)

    cols_a = {v: [word() for _ in range(n)] for v in ("x", "y")}
    cols_b = {v: [near(w) if k % 3 == 0 else word() for k, w in enumerate(cols_a[v])] for v in ("x", "y")}
    dfa = pd.DataFrame({**cols_a, "e": rng.integers(0, 5, n)})
    dfb = pd.DataFrame({**cols_b, "e": rng.integers(0, 5, n)})
    ba = rng.integers(0, n_blocks, n).astype(np.int32)
    bb = ba.copy()
    bb[::7] = rng.integers(0, n_blocks, len(bb[::7]))
    vf = ["x", "y"]
    bl = BlockedLinkage(dfa, dfb, ba, bb, vf, vf, ["e"], ["e"])
    ref, _ = per_block_loop(dfa, dfb, ba, bb, vf, ["e"])
    bl.fit(use_graph=True)
    np.testing.assert_array_equal(bl.Block_Counts, ref)
    bl.run_per_block()
    np.testing.assert_array_equal(bl.collect(), ref)
    ra, rb = np.nonzero(ba == 0)[0], np.nonzero(bb == 0)[0]
    _, _, counts = oracle.comparison_fit([dfa[v].to_numpy()[ra] for v in vf], [dfb[v].to_numpy()[rb] for v in vf],
                                         [dfa["e"].to_numpy()[ra]], [dfb["e"].to_numpy()[rb]])
    np.testing.assert_array_equal(bl.Block_Counts[0], counts)
    assert bl.Block_Counts[:, 1:].sum() > 0


def test_blocked_linkage_with_long_values(oracle):
    """Blocks holding values over 64 bytes (scored outside the bit-parallel
    passes) and over 1,024 bytes (flags in scratch), with near copies across
    the sides of one block: the batched launch, the per-block launch loop and
    the per-block recipe give the same Counts, and one block equals the
    oracle."""
    from faster.blocking import BlockedLinkage, per_block_loop
    from faster.synthetic import blocked_tables

    dfa, dfb, ba, bb, sizes = blocked_tables(2400, n_blocks=10, seed=5, kmin=20, kmax=800)
    rng = np.random.default_rng(6)
    vf = ["first_name", "last_name", "street"]
    b0 = int(np.argmax(sizes))
    ra, rb = np.nonzero(ba == b0)[0], np.nonzero(bb == b0)[0]
    rb2 = np.nonzero(bb == (b0 + 1) % len(sizes))[0]
    for v, (lo, hi) in zip(vf, [(65, 200), (1025, 1500), (70, 1300)]):
        long_a = _words(rng, 6, "abcd", lo, hi)
        col_a, col_b = dfa[v].to_numpy().copy(), dfb[v].to_numpy().copy()
        for k, s in enumerate(long_a):
            col_a[ra[k]] = s
            t = list(s)
            t[int(rng.integers(0, len(t)))] = "x"
            col_b[rb[k]] = "".join(t) if k % 2 == 0 else s[:-1]
        col_b[rb2[0]] = long_a[0]  # a long value alone in another block
        dfa[v], dfb[v] = col_a, col_b
    bl = BlockedLinkage(dfa, dfb, ba, bb, vf, vf, ["house_number"], ["house_number"])
    ref, _ = per_block_loop(dfa, dfb, ba, bb, vf, ["house_number"])
    bl.fit(use_graph=True)
    np.testing.assert_array_equal(bl.Block_Counts, ref)
    bl.fit(use_graph=True)
    np.testing.assert_array_equal(bl.Block_Counts, ref, err_msg="second replay")
    bl.run_per_block()
    np.testing.assert_array_equal(bl.collect(), ref, err_msg="per-block launch loop")
    _, _, counts = oracle.comparison_fit([dfa[v].to_numpy()[ra] for v in vf], [dfb[v].to_numpy()[rb] for v in vf],
                                         [dfa["house_number"].to_numpy()[ra]], [dfb["house_number"].to_numpy()[rb]])
    np.testing.assert_array_equal(bl.Block_Counts[b0], counts)
    assert counts[1:].sum() > 0


@pytest.mark.parametrize("mode,order", [("link", "random"), ("dedup", "random"), ("link", "sorted"),
                                        ("dedup", "sorted")])
def test_comparison_methods_agree(mode, order):
    """The per-lane bit-parallel form and the prefix form return the same
    level keys on synthetic data with every length
    class (1..64 bytes, every mask kind) and many near-duplicates. "sorted"
    passes the values in byte order over a small alphabet, so consecutive
    values share long prefixes and the prefix form reuses states at every
    depth; "random" leaves them unordered."""
    from faster import _lib, _ops
    from faster.synthetic import random_strings

    rng = np.random.default_rng(31)
    A = random_strings(rng, 3000, 8.0, 0.6, 1, 64)
    B = random_strings(rng, 2500, 8.0, 0.6, 1, 64)
    for k in range(0, 2500, 3):
        s = A[(k * 7) % 3000]
        B[k] = s[:-1] + "x" if len(s) > 1 else s
    if order == "sorted":
        # Few symbols, so that many values of one length share prefixes.
        tr = str.maketrans("abcdefghijklmnopqrstuvwxyz", "abcabcabcabcabcabcabcabcab")
        A = sorted({s.translate(tr) for s in A}, key=lambda s: s.encode())
        B = sorted({s.translate(tr) for s in B} | set(A[::5]), key=lambda s: s.encode())
    dev = _lib.device()
    out = {}
    default = _lib.jw_method()
    for method in ("tiles", "prefix"):
        _lib.set_jw_method(method)
        try:
            if mode == "link":
                ca, oa = _pack(A, dev)
                cb, ob = _pack(B, dev)
                hits, st = _ops.jw_levels(ca, oa, len(A), cb, ob, len(B), _lib.FH_MODE_LINK, 0.1, 0.88, 0.94,
                                          want_stats=True)
            else:
                ca, oa = _pack(A + B, dev)
                hits, st = _ops.jw_levels(ca, oa, len(A) + len(B), ca, oa, len(A) + len(B), _lib.FH_MODE_DEDUP, 0.1,
                                          0.88, 0.94, want_stats=True)
        finally:
            _lib.set_jw_method(default)
        assert st["method"] == _lib._METHODS[method]
        out[method] = np.sort(hits.cpu().numpy())
    assert out["tiles"].size > 500
    np.testing.assert_array_equal(out["prefix"], out["tiles"])


def _pack(strings, dev):
    from faster.encode import pack_strings

    c, o = pack_strings(strings)
    return torch.from_numpy(c if c.size else np.zeros(1, np.uint8)).to(dev), torch.from_numpy(o).to(dev)
