"""Pins of the oracle (CPU): published values, exact rationals, brute force,
matching theory and EM invariants. No GPU needed."""
import itertools
import json
import os
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
F32_LO = float(np.float32(0.88))
F32_HI = float(np.float32(0.94))


def _kat():
    with open(os.path.join(HERE, "golden", "jw_kat.json")) as f:
        return json.load(f)


def _exact_jw(l1, l2, m, t_half_count_x2, prefix, p=Fraction(1, 10)):
    """Closed form of docs/source/comparison.rst:8-18 in exact rationals;
    the transposition count enters as tt/2."""
    if m == 0:
        return Fraction(0)
    t = Fraction(t_half_count_x2, 2)
    j = (Fraction(m, l1) + Fraction(m, l2) + (m - t) / m) / 3
    return j + p * prefix * (1 - j)


def test_kat_published_values(oracle):
    kat = _kat()
    for c in kat["cases"]:
        s = oracle.jw_score(c["s1"], c["s2"], kat["p"])
        assert s == pytest.approx(c["fp32"], abs=0), c
        if c["published"] is not None:
            assert abs(s - c["published"]) < 5e-4, c
        if c["m"]:
            ex = _exact_jw(len(c["s1"]), len(c["s2"]), c["m"], 2 * c["t"], c["prefix"])
            assert abs(s - float(ex)) < 1e-6, c


def test_formula_vs_exact_rational(oracle):
    """Every (len1, len2, m, tt, prefix) up to length 12: the fp32 level equals
    the exact-rational level except where the exact value sits within 1e-6 of
    a threshold (where fp32 rounding legitimately decides)."""
    lo, hi = Fraction(22, 25), Fraction(47, 50)
    mism = 0
    for l1 in range(1, 13):
        for l2 in range(1, 13):
            for m in range(1, min(l1, l2) + 1):
                for tt in range(0, m + 1):
                    for pre in range(0, min(4, l1, l2) + 1):
                        s = oracle.jw_formula(l1, l2, m, tt, pre)
                        ex = _exact_jw(l1, l2, m, tt, pre)
                        assert abs(s - float(ex)) < 1e-6
                        lv_f = 2 if s >= F32_HI else (1 if s >= F32_LO else 0)
                        lv_e = 2 if ex >= hi else (1 if ex >= lo else 0)
                        if lv_f != lv_e:
                            near = min(abs(ex - lo), abs(ex - hi))
                            assert near < Fraction(1, 10 ** 6)
                            mism += 1
    # The reference's float arithmetic differs from exact arithmetic on a few
    # tuples that land exactly on a threshold; there must be some.
    assert mism > 0


def test_fma_does_not_change_levels(oracle):
    """The fused and unfused final update give the same level for every
    tuple with lengths <= 32 at the default thresholds."""
    for l1 in range(1, 33):
        for l2 in range(1, 33):
            for m in range(1, min(l1, l2) + 1):
                for tt in range(0, m + 1):
                    for pre in range(0, min(4, l1, l2) + 1):
                        a = oracle.jw_formula(l1, l2, m, tt, pre, 0.1, True)
                        b = oracle.jw_formula(l1, l2, m, tt, pre, 0.1, False)
                        la = 2 if a >= F32_HI else (1 if a >= F32_LO else 0)
                        lb = 2 if b >= F32_HI else (1 if b >= F32_LO else 0)
                        assert la == lb


def _greedy_m(s1, s2):
    w = max(len(s1), len(s2)) // 2 - 1
    used = [False] * len(s2)
    m = 0
    for i, c in enumerate(s1):
        for j in range(max(0, i - w), min(len(s2), i + w + 1)):
            if s2[j] == c and not used[j]:
                used[j] = True
                m += 1
                break
    return m


def _max_matching(s1, s2):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import maximum_bipartite_matching

    w = max(len(s1), len(s2)) // 2 - 1
    rows, cols = [], []
    for i, c in enumerate(s1):
        for j in range(max(0, i - w), min(len(s2), i + w + 1)):
            if s2[j] == c:
                rows.append(i)
                cols.append(j)
    if not rows:
        return 0
    g = csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(len(s1), len(s2)))
    return int((maximum_bipartite_matching(g, perm_type="column") >= 0).sum())


def test_match_count_is_maximum_matching(oracle):
    """The reference's greedy first-unused match count equals a maximum
    bipartite matching of the window graph (so it is symmetric in the two
    strings), checked on random small-alphabet strings; the oracle's score
    uses that count."""
    rng = np.random.default_rng(5)
    for _ in range(3000):
        l1, l2 = rng.integers(1, 14, 2)
        s1 = "".join("abc"[k] for k in rng.integers(0, 3, l1))
        s2 = "".join("abc"[k] for k in rng.integers(0, 3, l2))
        m = _greedy_m(s1, s2)
        assert m == _max_matching(s1, s2) == _greedy_m(s2, s1)
        if m == 0:
            assert oracle.jw_score(s1, s2) == 0.0
        else:
            # score with t = 0 and the longest possible prefix bounds it.
            pre = 0
            for a, b in zip(s1[:4], s2[:4]):
                if a != b:
                    break
                pre += 1
            assert oracle.jw_score(s1, s2) <= oracle.jw_formula(len(s1), len(s2), m, 0, pre) + 1e-7


def _brute_link(A, B, oracle, lo=0.88, hi=0.94):
    l1, l2 = [], []
    for i, a in enumerate(A):
        for j, b in enumerate(B):
            s = np.float32(oracle.jw_score(a, b))
            if s >= np.float32(hi):
                l2.append(i * len(B) + j)
            elif s >= np.float32(lo):
                l1.append(i * len(B) + j)
    return [np.array(l1, dtype=np.int64), np.array(l2, dtype=np.int64)]


def _words(rng, n, alpha="abcde", lmax=7, dup=None):
    out = ["".join(alpha[k] for k in rng.integers(0, len(alpha), rng.integers(0, lmax + 1))) for _ in range(n)]
    if dup:
        for k in range(0, n, dup):
            out[k] = out[0]
    return out


def test_unique_path_equals_brute_force(oracle):
    rng = np.random.default_rng(1)
    A = _words(rng, 60, dup=5)
    B = _words(rng, 70, dup=7)
    got = oracle.jaro_winkler_unique(A, B)
    ref = _brute_link(A, B, oracle)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    allp = oracle.jaro_winkler_all_pairs(A, B)
    for g, r in zip(allp, ref):
        np.testing.assert_array_equal(g, r)


def test_exact_equals_brute_force(oracle):
    rng = np.random.default_rng(2)
    A = rng.integers(0, 12, 80)
    B = rng.integers(5, 20, 90)
    got = oracle.exact(A, B)[0]
    ref = np.array([i * len(B) + j for i in range(len(A)) for j in range(len(B)) if A[i] == B[j]], dtype=np.int64)
    np.testing.assert_array_equal(got, ref)


def test_merge_equals_brute_force_codes(oracle):
    """Pattern lists from the setdiff/intersect cascade equal the lists built
    directly from each pair's pattern code (first variable slowest)."""
    rng = np.random.default_rng(3)
    nA, nB = 40, 50
    A1, B1 = _words(rng, nA, dup=4), _words(rng, nB, dup=5)
    A2, B2 = _words(rng, nA, "ab", 5), _words(rng, nB, "ab", 5)
    A3, B3 = rng.integers(0, 4, nA), rng.integers(0, 4, nB)
    per_var, indices, counts = oracle.comparison_fit([A1, A2], [B1, B2], [A3], [B3])
    lv1 = np.zeros(nA * nB, int)
    lv2 = np.zeros(nA * nB, int)
    for lv, (l_1, l_2) in ((lv1, per_var[0]), (lv2, per_var[1])):
        lv[l_1] = 1
        lv[l_2] = 2
    lv3 = np.zeros(nA * nB, int)
    lv3[per_var[2][0]] = 1
    code = lv1 * 6 + lv2 * 2 + lv3
    assert len(indices) == 17
    for k in range(1, 18):
        np.testing.assert_array_equal(np.asarray(indices[k - 1], dtype=np.int64), np.nonzero(code == k)[0])
    assert counts[0] == np.sum(code == 0)
    assert counts.sum() == nA * nB


def test_dedup_equals_brute_force(oracle):
    """Deduplication lists: every pair i > j, string 1 the lexicographically
    larger value, identical values at level 2."""
    rng = np.random.default_rng(4)
    S = _words(rng, 70, dup=6)
    got = oracle.jaro_winkler_dedup(S)
    l1, l2 = [], []
    for i in range(len(S)):
        for j in range(i):
            a, b = (S[i], S[j]) if S[i] > S[j] else (S[j], S[i])
            s = np.float32(1.0) if S[i] == S[j] else np.float32(oracle.jw_score(a, b))
            if s >= np.float32(0.94):
                l2.append(i * len(S) + j)
            elif s >= np.float32(0.88):
                l1.append(i * len(S) + j)
    np.testing.assert_array_equal(got[0], np.array(l1, dtype=np.int64))
    np.testing.assert_array_equal(got[1], np.array(l2, dtype=np.int64))
    E = rng.integers(0, 6, 50)
    ge = oracle.exact_dedup(E)[0]
    ref = [i * 50 + j for i in range(50) for j in range(i) if E[i] == E[j]]
    np.testing.assert_array_equal(ge, np.array(ref, dtype=np.int64))


def _mixture_counts(levels, lam, Pi, N):
    G = list(itertools.product(*(range(L) for L in levels)))
    p = []
    for g in G:
        p1 = np.prod([Pi[1][k][g[k]] for k in range(len(levels))])
        p0 = np.prod([Pi[0][k][g[k]] for k in range(len(levels))])
        p.append(lam * p1 + (1 - lam) * p0)
    return np.round(np.array(p) * N).astype(np.int64)


def test_em_invariants(oracle):
    """EM on an exact mixture histogram: recovers the generating parameters,
    ends at a fixed point, never decreases the log-likelihood, and matches a
    scipy L-BFGS maximum of the same likelihood."""
    levels = [3, 3, 2]
    true_lam = 0.02
    Pi = [[np.array([0.9, 0.07, 0.03]), np.array([0.85, 0.1, 0.05]), np.array([0.95, 0.05])],
          [np.array([0.05, 0.15, 0.8]), np.array([0.1, 0.2, 0.7]), np.array([0.2, 0.8])]]
    counts = _mixture_counts(levels, true_lam, Pi, 10 ** 12)
    np.random.seed(0)
    init = oracle.init_params(levels)
    lls = []
    lam, P = init[0], init[1]
    G = oracle.gamma(levels)
    for _ in range(30):
        lls.append(oracle.log_likelihood(levels, counts, lam, P))
        lam, P, _, _, _ = oracle.em_fit(levels, counts, tol=0.0, max_iter=1, init=(lam, P))
    assert all(b >= a - 1e-6 * abs(a) for a, b in zip(lls, lls[1:]))
    lam, P, it, conv, ksi = oracle.em_fit(levels, counts, tol=1e-12, max_iter=20000, init=init)
    assert conv
    assert lam == pytest.approx(true_lam, rel=1e-6)
    for m in range(2):
        for k in range(3):
            np.testing.assert_allclose(P[m][k], Pi[m][k], rtol=1e-5, atol=1e-9)
    lam2, P2, _, _, _ = oracle.em_fit(levels, counts, tol=0.0, max_iter=1, init=(lam, P))
    assert lam2 == pytest.approx(lam, rel=1e-9)
    # scipy optimum of the same likelihood over a softmax parameterisation.
    from scipy.optimize import minimize

    def unpack(x):
        lam_ = 1 / (1 + np.exp(-x[0]))
        off = 1
        out = [[], []]
        for m in range(2):
            for L in levels:
                z = np.concatenate([[0.0], x[off:off + L - 1]])
                out[m].append(np.exp(z - z.max()) / np.exp(z - z.max()).sum())
                off += L - 1
        return lam_, out

    def nll(x):
        lam_, P_ = unpack(x)
        return -oracle.log_likelihood(levels, counts, lam_, P_) / counts.sum()

    x0 = np.zeros(1 + 2 * sum(L - 1 for L in levels))
    x0[0] = np.log(lam / (1 - lam))
    off = 1
    for m in range(2):
        for k, L in enumerate(levels):
            x0[off:off + L - 1] = np.log(P[m][k][1:] / P[m][k][0])
            off += L - 1
    r = minimize(nll, x0 + 0.05, method="L-BFGS-B", options={"ftol": 1e-15, "gtol": 1e-12, "maxiter": 5000})
    assert nll(x0) <= r.fun + 1e-9


def test_linkage_pattern_zero_quirk(oracle):
    """Pattern k reads Indices[k - 1], so pattern 0 selects the last list
    (faster/linkage.py:49-54)."""
    idx = [np.array([1, 2]), np.array([5]), np.array([7, 9])]
    ksi = np.array([0.9, 0.1, 0.1, 0.1])
    a, b = oracle.linkage_pairs(idx, ksi, n_B=4, threshold=0.85)
    np.testing.assert_array_equal(a * 4 + b, [7, 9])
    with pytest.raises(Exception, match="No pair of observations"):
        oracle.linkage_pairs(idx, np.zeros(4), n_B=4)


def test_group_strings_order_and_rows(oracle):
    """Unique values in Python's str order (shorter prefix first, NUL after the
    bare prefix, code points for non-ASCII) with ascending rows: a hand-worked
    case and a brute-force check of every group."""
    vals = ["b", "a", "b", "", "ab", "a\x00", "a", "é", "z"]
    uniq, starts, rows = oracle.group_strings(vals)
    assert uniq == [b"", b"a", b"a\x00", b"ab", b"b", b"z", "é".encode()]
    assert starts.tolist() == [0, 1, 3, 4, 5, 7, 8, 9]
    assert rows.tolist() == [3, 1, 6, 5, 4, 0, 2, 8, 7]
    rng = np.random.default_rng(5)
    pool = ["", "a", "ab", "abc", "b", "ba", "a\x00b", "zz", "ü", "日本"]
    vals = [pool[i] for i in rng.integers(0, len(pool), 300)]
    uniq, starts, rows = oracle.group_strings(vals)
    assert uniq == sorted({v.encode() for v in vals})  # UTF-8 byte order = code-point order
    for u, s0, s1 in zip(uniq, starts[:-1], starts[1:]):
        assert rows[s0:s1].tolist() == [i for i, v in enumerate(vals) if v.encode() == u]
    assert oracle.group_strings([])[0] == []


# ---------------------- numeric distance and missing levels -------------------
def test_numeric_levels_pins(oracle):
    """Numeric-distance levels: with d2 = d1 = 0 they are the exact-equality
    pairs of oracle.exact / oracle.exact_dedup; on integer data (exact in
    float64) level >= 1 iff |x - y| <= d1 and level 2 iff |x - y| <= d2, by
    an independent loop over pairs; NaN pairs with nothing; the relation is
    symmetric."""
    rng = np.random.default_rng(3)
    xa = rng.integers(0, 40, 150).astype(np.float64)
    xb = rng.integers(0, 40, 170).astype(np.float64)
    l1, l2 = oracle.numeric_link(xa, xb, 0.0, 0.0)
    assert len(l1) == 0
    np.testing.assert_array_equal(l2, oracle.exact(xa.astype(np.int64), xb.astype(np.int64))[0])
    d1, d2 = oracle.numeric_dedup(xa, 0.0, 0.0)
    assert len(d1) == 0
    np.testing.assert_array_equal(d2, oracle.exact_dedup(xa.astype(np.int64))[0])
    xa[::7] = np.nan
    xb[::5] = np.nan
    l1, l2 = oracle.numeric_link(xa, xb, 1.0, 3.0)
    want1, want2 = [], []
    for i in range(len(xa)):
        for j in range(len(xb)):
            if np.isnan(xa[i]) or np.isnan(xb[j]):
                continue
            dist = abs(int(xa[i]) - int(xb[j]))
            if dist <= 1:
                want2.append(i * len(xb) + j)
            elif dist <= 3:
                want1.append(i * len(xb) + j)
    np.testing.assert_array_equal(l1, want1)
    np.testing.assert_array_equal(l2, want2)
    r1, r2 = oracle.numeric_link(xb, xa, 1.0, 3.0)
    t = lambda ids: np.sort((ids % len(xa)) * len(xb) + ids // len(xa))
    np.testing.assert_array_equal(t(r1), l1)
    np.testing.assert_array_equal(t(r2), l2)


def test_missing_levels_pins(oracle):
    """Missing levels: on data without missing values, Missing=True lists the
    same pairs as the reference-restating comparison_fit at codes whose
    digits are unchanged (the extra level is empty); with missing values the
    pairs whose value is missing on variable v, counted over all patterns,
    equal the closed form n_A n_B - (n_A - m_A)(n_B - m_B), and the product's
    P(M) (faster.missing.pair_totals) equals a count by enumeration."""
    import pandas as pd

    from faster.missing import pair_totals, row_masks
    from faster.synthetic import add_numeric_and_missing, linkage_tables

    dfa, dfb, truth = linkage_tables(120, 140, seed=4)
    vf = ["first_name", "last_name"]
    _, idx_ref, counts_ref = oracle.comparison_fit([dfa[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf],
                                                      [dfa["house_number"].to_numpy()],
                                                      [dfb["house_number"].to_numpy()])
    lv, _, idx, counts = oracle.comparison_fit_full([dfa[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf],
                                                    [dfa["house_number"].to_numpy()], [dfb["house_number"].to_numpy()],
                                                    missing=True)
    assert lv == [4, 4, 3]
    G_old = np.array(list(itertools.product(range(3), range(3), range(2))))
    for k, g in enumerate(G_old[1:], start=1):
        new = g[0] * 12 + g[1] * 3 + g[2]
        np.testing.assert_array_equal(idx[new - 1], idx_ref[k - 1])
        assert counts[new] == counts_ref[k]
    assert counts.sum() == 120 * 140 and counts[0] == counts_ref[0]
    fa, fb = add_numeric_and_missing(dfa, dfb, truth, seed=4, missing=0.2)
    lv, _, idx, counts = oracle.comparison_fit_full([fa[v].to_numpy() for v in vf], [fb[v].to_numpy() for v in vf],
                                                    [fa["house_number"].to_numpy()], [fb["house_number"].to_numpy()],
                                                    [fa["birth_day"].to_numpy()], [fb["birth_day"].to_numpy()],
                                                    0.0, 3.0, missing=True)
    assert lv == [4, 4, 3, 4] and counts.sum() == 120 * 140
    G = np.array(list(itertools.product(*(range(L) for L in lv))))
    names = vf + ["house_number", "birth_day"]
    for v, name in enumerate(names):
        ma, mb = int(pd.isna(fa[name]).sum()), int(pd.isna(fb[name]).sum())
        assert counts[G[:, v] == lv[v] - 1].sum() == 120 * 140 - (120 - ma) * (140 - mb), name
    miss_a = [pd.isna(fa[v]).to_numpy() for v in names]
    miss_b = [pd.isna(fb[v]).to_numpy() for v in names]
    ma, mb = row_masks(miss_a, 120), row_masks(miss_b, 140)
    P = pair_totals(ma, mb, False)
    enum = {}
    for x in ma.tolist():
        for y in mb.tolist():
            enum[x | y] = enum.get(x | y, 0) + 1
    assert P == enum
    Pd = pair_totals(ma, ma, True)
    enum = {}
    for i in range(120):
        for j in range(i):
            enum[int(ma[i] | ma[j])] = enum.get(int(ma[i] | ma[j]), 0) + 1
    assert Pd == enum
