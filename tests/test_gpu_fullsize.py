"""Parity at the benchmark sizes.

* Plain linkage 10,000 x 10,000 (3 string variables + house number): every
  one of the 53 pattern lists, the Counts and EM at tol 1e-6 against the
  oracle run in full (3e8 JW comparisons on the host cores).
* Plain linkage 1,000,000 x 1,000,000 through the public Comparison class.
  The oracle cannot scan 1e12 pairs, so the check is split:
  - a seeded sample of A rows is scored by the oracle against every B value
    and must equal the GPU lists restricted to those rows, bit for bit;
  - size-independent invariants: strictly ascending lists, disjoint levels,
    the exact-match list size equals its closed form sum_g n_A(g) * n_B(g);
  - all 54 Counts re-derived on the CPU from the GPU's per-variable lists
    (pattern code of every listed pair by a merge of the lists, independent
    of the device merge), and the pattern list holding sampled union pairs;
  - EM (tol 1e-6) and the per-pair posteriors against the oracle's FP64 EM
    on the same histogram, to 1e-9 relative.
* Deduplication of 2,000,000 records (birth year kept as row ids): Counts sum
  to n(n+1)/2, the pairs with equal birth years add up to the closed form
  sum_y c_y(c_y - 1)/2, and every fuzzy level list restricted to 32 sampled
  rows equals the oracle's scores of those rows against all rows, in the
  reference's orientation (the lexicographically larger value is string 1).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N = 1_000_000
VF = ["first_name", "last_name", "street"]


def _np(t):
    return t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


# ------------------------------ 10k x 10k ------------------------------------
def test_config1_10k_all_patterns_and_em(oracle):
    import faster
    from faster.synthetic import linkage_tables

    n = 10_000
    dfa, dfb, _ = linkage_tables(n, n, seed=0)
    print("10k x 10k Comparison.fit", flush=True)
    comp = faster.Comparison(dfa, dfb, VF, VF, ["house_number"], ["house_number"])
    comp.fit()
    print("10k x 10k oracle", flush=True)
    _, indices, counts = oracle.comparison_fit([dfa[v].to_numpy() for v in VF], [dfb[v].to_numpy() for v in VF],
                                               [dfa["house_number"].to_numpy()], [dfb["house_number"].to_numpy()])
    assert len(comp.Indices) == len(indices) == 53
    for k, (g, r) in enumerate(zip(comp.Indices, indices)):
        np.testing.assert_array_equal(_np(g), np.asarray(r, dtype=np.int64), err_msg=f"pattern {k + 1}")
    np.testing.assert_array_equal(comp.Counts, counts)
    np.random.seed(0)
    est = faster.Estimation(3, 1, comp.Counts)
    est.fit(Tolerance=1e-6)
    np.random.seed(0)
    lam, Pi, it, conv, ksi = oracle.em_fit([3, 3, 3, 2], counts, 1e-6, 5000)
    assert est.Iterations == it and conv
    assert est.Lambda == pytest.approx(lam, rel=1e-9)
    for c in range(2):
        for v in range(4):
            np.testing.assert_allclose(est.Pi[c][v], Pi[c][v], rtol=1e-9, atol=0)
    np.testing.assert_allclose(est.Ksi, ksi, rtol=1e-9, atol=1e-300)


# ------------------------------ 1M x 1M --------------------------------------
@pytest.fixture(scope="module")
def fullsize():
    import faster
    from faster.synthetic import linkage_tables

    dfa, dfb, _ = linkage_tables(N, N, seed=0)
    print("1M x 1M Comparison.fit", flush=True)
    comp = faster.Comparison(dfa, dfb, VF, VF, ["house_number"], ["house_number"])
    comp.fit()
    assert comp.Exact_Mode == "list"
    keys = [_np(k) for k in comp._var_keys]
    return dfa, dfb, comp, keys, _np(comp._pattern_ids), comp._pattern_offsets


def test_sampled_rows_match_oracle(oracle, fullsize):
    dfa, dfb, comp, keys, _, _ = fullsize
    rng = np.random.default_rng(123)
    rows = np.sort(rng.choice(N, 48, replace=False))
    for v, name in enumerate(VF):
        ref1, ref2 = oracle.jaro_winkler_all_pairs(list(dfa[name].to_numpy()[rows]), list(dfb[name].to_numpy()))
        k = keys[v]
        a = (k >> 2) // N
        got = k[np.isin(a, rows)]
        for lvl, ref in ((1, ref1), (2, ref2)):
            g = got[(got & 3) == lvl] >> 2
            r = rows[ref // N] * N + ref % N
            np.testing.assert_array_equal(np.sort(g), np.sort(r), err_msg=f"{name} level {lvl}")
    # Levels (the per-variable, per-level lists) are the keys split by level.
    lv = comp.Levels
    for v in range(3):
        for lvl in (1, 2):
            np.testing.assert_array_equal(_np(lv[v][lvl - 1]), keys[v][(keys[v] & 3) == lvl] >> 2)


def _union_codes(keys, strides):
    """Pattern code of every pair listed by some variable, from the
    per-variable lists alone: the fuzzy lists are merged by a sort, the large
    exact list is looked up by binary search."""
    fz_ids = np.concatenate([k[:] >> 2 for k in keys[:3]])
    fz_add = np.concatenate([(k & 3) * s for k, s in zip(keys[:3], strides[:3])])
    order = np.argsort(fz_ids, kind="stable")
    fz_ids, fz_add = fz_ids[order], fz_add[order]
    heads = np.concatenate([[True], fz_ids[1:] != fz_ids[:-1]])
    start = np.nonzero(heads)[0]
    uid = fz_ids[start]
    code = np.add.reduceat(fz_add, start)
    ex_ids = keys[3] >> 2
    j = np.minimum(np.searchsorted(ex_ids, uid), len(ex_ids) - 1)
    in_ex = ex_ids[j] == uid
    code = code + np.where(in_ex, (keys[3][j] & 3) * strides[3], 0)
    return uid, code, len(ex_ids) - int(in_ex.sum())


def test_invariants_and_counts(fullsize):
    dfa, dfb, comp, keys, ids, offs = fullsize
    for k in keys:
        idv = k >> 2
        assert np.all(np.diff(idv) > 0)
        assert idv.min() >= 0 and idv.max() < N * N
        assert set(np.unique(k & 3).tolist()) <= {1, 2}
    ha = dfa["house_number"].to_numpy()
    hb = dfb["house_number"].to_numpy()
    ca = np.bincount(ha, minlength=10000).astype(np.int64)
    cb = np.bincount(hb, minlength=10000).astype(np.int64)
    assert len(keys[3]) == int(np.dot(ca, cb))
    strides = [18, 6, 2, 1]
    uid, code, exact_only = _union_codes(keys, strides)
    ref = np.bincount(code, minlength=54).astype(np.int64)
    ref[1] += exact_only
    n_union = len(uid) + exact_only
    ref[0] = N * N - n_union
    np.testing.assert_array_equal(comp.Counts, ref)
    assert comp.Counts.sum() == N * N
    # The pattern list holding sampled union pairs.
    rng = np.random.default_rng(7)
    pick = np.sort(rng.choice(len(ids), 4000, replace=False))
    pid = ids[pick]
    owner = np.searchsorted(np.asarray(offs), pick, side="right") - 1
    j = np.minimum(np.searchsorted(uid, pid), len(uid) - 1)
    expect = np.where(uid[j] == pid, code[j], 1)
    np.testing.assert_array_equal(owner, expect)
    for c in range(1, len(offs) - 1):
        seg = ids[offs[c]:offs[c + 1]]
        assert np.all(np.diff(seg) > 0)


def test_em_and_posteriors_match_oracle(oracle, fullsize):
    import faster

    _, _, comp, keys, ids, offs = fullsize
    np.random.seed(0)
    est = faster.Estimation(3, 1, comp.Counts)
    est.fit(Tolerance=1e-6)
    np.random.seed(0)
    lam, Pi, it, conv, ksi = oracle.em_fit([3, 3, 3, 2], comp.Counts, 1e-6, 5000)
    assert est.Iterations == it and conv
    assert est.Lambda == pytest.approx(lam, rel=1e-9)
    for c in range(2):
        for v in range(4):
            np.testing.assert_allclose(est.Pi[c][v], Pi[c][v], rtol=1e-9, atol=0)
    np.testing.assert_allclose(est.Ksi, ksi, rtol=1e-9, atol=1e-300)
    post = _np(comp.posteriors(est.Ksi))
    assert len(post) == len(ids)
    rng = np.random.default_rng(9)
    pick = rng.choice(len(ids), 200_000, replace=False)
    pattern = np.searchsorted(np.asarray(offs), pick, side="right") - 1
    np.testing.assert_allclose(post[pick], ksi[pattern], rtol=1e-9, atol=1e-300)


# ------------------------------ 2M dedup --------------------------------------
def test_config3_dedup_2m(oracle):
    import faster
    from faster.synthetic import dedup_table

    n = 2_000_000
    df, _ = dedup_table(n, seed=0)
    print("2M-row deduplication", flush=True)
    dd = faster.Deduplication(df, VF, ["birth_year"])
    dd.fit()
    assert dd.Exact_Mode == "rowid"
    counts = dd.Counts
    assert counts.sum() == n * (n + 1) // 2
    c = np.bincount(df["birth_year"].to_numpy().astype(np.int64)).astype(np.int64)
    equal_years = int(np.sum(c * (c - 1) // 2))
    # Exact digit is the last; patterns with an odd code have equal years.
    assert int(counts[1::2].sum()) == equal_years
    rng = np.random.default_rng(31)
    rows = np.sort(rng.choice(n, 32, replace=False))
    lv = dd.Levels
    for v, name in enumerate(VF):
        vals = df[name].to_numpy()
        sub = list(vals[rows])
        as1 = oracle.jaro_winkler_all_pairs(sub, list(vals))       # sampled value as string 1
        as2 = oracle.jaro_winkler_all_pairs(list(vals), sub)       # sampled value as string 2
        cand = {}
        for lvl in (1, 2):
            for x in as1[lvl - 1].tolist():
                r, j = int(rows[x // n]), x % n
                cand.setdefault((r, j), {})["1"] = lvl
            for x in as2[lvl - 1].tolist():
                j, r = x // len(rows), int(rows[x % len(rows)])
                cand.setdefault((r, j), {})["2"] = lvl
        # Identical values of other rows are level 2 (the forced diagonal).
        _, inv = np.unique(vals, return_inverse=True)
        inv = np.asarray(inv).reshape(-1)
        for r in rows.tolist():
            for j in np.nonzero(inv == inv[r])[0].tolist():
                if j != r:
                    cand.setdefault((r, j), {})["eq"] = 2
        expect = {1: set(), 2: set()}
        for (r, j), d in cand.items():
            if r == j:
                continue
            if vals[r] == vals[j]:
                lvl = 2
            else:
                lvl = d.get("1", 0) if vals[r] > vals[j] else d.get("2", 0)
            if lvl:
                expect[lvl].add(max(r, j) * n + min(r, j))
        rset = set(rows.tolist())
        for lvl in (1, 2):
            got = _np(lv[v][lvl - 1])
            hi, lo = got // n, got % n
            mask = np.isin(hi, rows) | np.isin(lo, rows)
            assert set(got[mask].tolist()) == expect[lvl], f"{name} level {lvl}"
            assert np.all(np.diff(got) > 0)
            assert np.all(hi > lo)
        assert rset  # sampled rows exist
