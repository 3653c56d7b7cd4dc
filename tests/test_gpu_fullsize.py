"""Full-size parity (1,000,000 x 1,000,000, the benchmark workload).

The oracle cannot scan 1e12 pairs, so the check is split:
* a seeded sample of A rows is scored by the oracle against every B value and
  must equal the GPU lists restricted to those rows, bit for bit;
* size-independent invariants hold for the whole output: strictly ascending
  lists, disjoint levels, the exact-match list size equals its closed form
  sum_g n_A(g) * n_B(g), Counts sum to n_A * n_B, and the pattern code of a
  sample of union pairs re-derived from the per-variable lists matches the
  pattern list holding the pair.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N = 1_000_000
VF = ["first_name", "last_name", "street"]


@pytest.fixture(scope="module")
def fullsize():
    from faster import _ops
    from faster.comparison import exact_keys_link, fuzzy_keys_link
    from faster.encode import group_exact_link, group_strings
    from faster.synthetic import linkage_tables

    dfa, dfb, _ = linkage_tables(N, N, seed=0)
    keys = []
    for v in VF:
        ga = group_strings(dfa[v].to_numpy()).to_device()
        gb = group_strings(dfb[v].to_numpy()).to_device()
        keys.append(fuzzy_keys_link(ga, gb))
    ex = group_exact_link(dfa["house_number"].to_numpy(), dfb["house_number"].to_numpy())
    ga_ex, sb_ex, rb_ex = ex.to_device()
    keys.append(exact_keys_link(ga_ex, sb_ex, rb_ex, ex.n_groups, N))
    ids, offs, n_union = _ops.pattern_merge(keys, [3, 3, 3, 2])
    return dfa, dfb, [k.cpu().numpy() for k in keys], ids.cpu().numpy(), offs, n_union


def test_sampled_rows_match_oracle(oracle, fullsize):
    dfa, dfb, keys, _, _, _ = fullsize
    rng = np.random.default_rng(123)
    rows = np.sort(rng.choice(N, 48, replace=False))
    for v, name in enumerate(VF):
        ref1, ref2 = oracle.jaro_winkler_all_pairs(list(dfa[name].to_numpy()[rows]), list(dfb[name].to_numpy()))
        k = keys[v]
        a = (k >> 2) // N
        mask = np.isin(a, rows)
        got = k[mask]
        for lvl, ref in ((1, ref1), (2, ref2)):
            g = got[(got & 3) == lvl] >> 2
            r = np.array([rows[x // N] * N + x % N for x in ref.tolist()], dtype=np.int64)
            np.testing.assert_array_equal(np.sort(g), np.sort(r), err_msg=f"{name} level {lvl}")


def test_invariants(fullsize):
    dfa, dfb, keys, ids, offs, n_union = fullsize
    idvs = [k >> 2 for k in keys]
    for k, idv in zip(keys, idvs):
        assert np.all(np.diff(idv) > 0)
        assert idv.min() >= 0 and idv.max() < N * N
        assert set(np.unique(k & 3).tolist()) <= {1, 2}
    ha = dfa["house_number"].to_numpy()
    hb = dfb["house_number"].to_numpy()
    ca = np.bincount(ha, minlength=10000).astype(np.int64)
    cb = np.bincount(hb, minlength=10000).astype(np.int64)
    assert len(keys[3]) == int(np.dot(ca, cb))
    counts = np.concatenate([[N * N - n_union], np.diff(np.asarray(offs))[1:]])
    assert counts.sum() == N * N
    assert n_union == len(np.unique(np.concatenate(idvs)))
    # Re-derive the pattern code of sampled union pairs from the per-variable lists.
    rng = np.random.default_rng(7)
    pick = np.sort(rng.choice(len(ids), 2000, replace=False))
    pid = ids[pick]
    strides = [18, 6, 2, 1]
    code = np.zeros(len(pick), dtype=np.int64)
    for v, (k, idv) in enumerate(zip(keys, idvs)):
        j = np.minimum(np.searchsorted(idv, pid), len(k) - 1)
        hit = idv[j] == pid
        code += np.where(hit, (k[j] & 3) * strides[v], 0)
    owner = np.searchsorted(np.asarray(offs), pick, side="right") - 1
    np.testing.assert_array_equal(owner, code)
    # pattern lists ascending
    for c in range(1, len(offs) - 1):
        seg = ids[offs[c]:offs[c + 1]]
        assert np.all(np.diff(seg) > 0)
