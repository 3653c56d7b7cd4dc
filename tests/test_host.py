"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared symbol, column encoding, threshold semantics, class
validation and error paths."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pandas as pd
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "faster_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(fh_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import faster._lib as L

    lib = L.load(require_gpu=False)
    decl = _declared_symbols()
    assert len(decl) >= 14
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(L.EXPORTED)
    assert lib.fh_version() == 100
    assert lib.fh_last_error() == b""
    # Argument checks run before any device work.
    buf = ctypes.c_void_p(0x1000)
    assert lib.fh_sort_u64_into(buf, buf, ctypes.c_int64(8), 0, 36, None) != 0
    assert b"distinct buffers" in lib.fh_last_error()
    assert lib.fh_sort_u64_into(None, None, ctypes.c_int64(-1), 0, 36, None) != 0
    assert b"n < 0" in lib.fh_last_error()


def test_ops_fail_loudly_without_gpu():
    import torch

    import faster._lib as L

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        L.load()


def test_group_strings_matches_numpy_unique():
    from faster.encode import group_strings

    rng = np.random.default_rng(0)
    vals = ["".join("abé"[k] for k in rng.integers(0, 3, rng.integers(0, 6))) for _ in range(500)]
    g = group_strings(vals)
    u, inv, cnt = np.unique(np.array(vals, dtype=object), return_inverse=True, return_counts=True)
    assert g.n_unique == len(u)
    for k in range(len(u)):
        b = g.chars[g.offsets[k]:g.offsets[k + 1]].tobytes().decode("utf-8")
        assert b == u[k]
        rows = g.rows[g.starts[k]:g.starts[k + 1]]
        np.testing.assert_array_equal(rows, np.nonzero(np.asarray(inv).reshape(-1) == k)[0])
    g2 = group_strings(["a\x00b", "a", "a\x00b"])
    assert g2.n_unique == 2


def test_group_exact():
    from faster.encode import group_exact_dedup, group_exact_link

    a = np.array([3, 1, 4, 1, 5])
    b = np.array([1, 5, 9, 1])
    g = group_exact_link(a, b)
    pairs = []
    for i in range(len(a)):
        gid = g.group_a[i]
        for j in g.rows_b[g.starts_b[gid]:g.starts_b[gid + 1]]:
            pairs.append(i * len(b) + j)
    ref = [i * len(b) + j for i in range(len(a)) for j in range(len(b)) if a[i] == b[j]]
    assert sorted(pairs) == ref
    d = group_exact_dedup(a)
    assert d.n_groups == 4
    assert list(d.rows[d.starts[d.group[1]]:d.starts[d.group[1] + 1]]) == [1, 3]


def test_threshold_semantics_fp32():
    from faster.levels import level_of_score

    f88 = float(np.float32(0.88))
    assert level_of_score(f88, 0.88, 0.94) == 1
    assert level_of_score(np.nextafter(np.float32(f88), np.float32(0)), 0.88, 0.94) == 0
    assert level_of_score(1.0, 0.88, 0.94) == 2
    assert level_of_score(0.0, 0.0, 0.5) == 1
    assert level_of_score(0.9, 0.94, 0.88) == 2


def test_class_validation_and_errors():
    import faster

    assert faster.__version__ == "0.3.1"
    df = pd.DataFrame({"a": ["x"], "b": ["y"], "n": [1]})
    with pytest.raises(Exception, match="lengths of Vars_Fuzzy_A and Vars_Fuzzy_B differ"):
        faster.Comparison(df, df, ["a"], [])
    with pytest.raises(Exception, match="lengths of Vars_Exact_A and Vars_Exact_B differ"):
        faster.Comparison(df, df, ["a"], ["a"], ["n"], [])
    with pytest.raises(Exception, match="Vars_Fuzzy_A and Vars_Fuzzy_B must match"):
        faster.Comparison(df, df, ["zz"], ["a"])
    with pytest.raises(Exception, match="Vars_Exact_A and Vars_Exact_B must match"):
        faster.Comparison(df, df, ["a"], ["a"], ["zz"], ["n"])
    with pytest.raises(Exception, match="must match variable names in df"):
        faster.Deduplication(df, ["zz"])
    c = faster.Comparison(df, df, ["a"], ["a"])
    with pytest.raises(Exception, match="must be fitted first"):
        c.Counts
    d = faster.Deduplication(df, ["a"])
    with pytest.raises(Exception, match="must be fitted first"):
        d.Counts
    e = faster.Estimation(3, 1, np.ones(54, dtype=np.int64))
    with pytest.raises(Exception, match="must be fitted first"):
        e.Ksi
    import itertools

    np.testing.assert_array_equal(e.Gamma, np.array(list(itertools.product(range(3), range(3), range(3), range(2)))))


def test_evaluation_host_math():
    """Values worked out by hand for Ksi = [0.01, 0.5, 0.99], Counts = [100,
    10, 5], Lambda = 0.05 (faster/evaluation.py:22-67): the expected matches
    below 0.9 are 0.01 * 100 + 0.5 * 10 = 6, so the reference's FNR is
    6 / 0.05 * 115 = 13800 (division before the product) and the rate is
    6 / (0.05 * 115). A threshold above every Ksi is a NumPy 0/0 (nan, no
    exception), Lambda = 0 gives inf (or nan for 0/0), and a list Ksi makes
    the reference's bare except return None."""
    import warnings

    import faster

    ksi = np.array([0.01, 0.5, 0.99])
    counts = np.array([100, 10, 5])
    ev = faster.Evaluation(0.05, ksi, counts)
    assert ev.FDR(0.9) == pytest.approx(0.05 / 5)
    assert ev.FDR(0.0) == pytest.approx((0.99 * 100 + 0.5 * 10 + 0.01 * 5) / 115)
    assert ev.FNR(0.9) == pytest.approx(13800.0)
    assert ev.FNR(0.0) == 0.0
    assert ev.FNR_rate(0.9) == pytest.approx(6 / 5.75)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        fdr_empty = ev.FDR(1.1)
        assert isinstance(fdr_empty, np.floating) and np.isnan(fdr_empty)
        z = faster.Evaluation(0.0, ksi, counts)
        assert np.isinf(z.FNR(0.9))
        assert np.isnan(z.FNR(0.0))
        # FDR + FNR is smallest at S = 0 (FNR is 0 there and >= 2300 above
        # 0.01); nan objectives above 0.99 count as 1 + Alpha.
        assert ev.Optimal_Threshold(1.0) == 0.0
    as_list = faster.Evaluation(0.05, [0.01, 0.5, 0.99], [100, 10, 5])
    assert as_list.FDR(0.5) is None and as_list.FNR(0.5) is None
    with pytest.raises(TypeError):
        as_list.Optimal_Threshold(1.0)


def test_synthetic_deterministic():
    from faster.synthetic import dedup_table, linkage_tables

    a1, b1, t1 = linkage_tables(200, 300, seed=1)
    a2, b2, t2 = linkage_tables(200, 300, seed=1)
    pd.testing.assert_frame_equal(a1, a2)
    pd.testing.assert_frame_equal(b1, b2)
    assert (t1 >= 0).sum() == 90
    lens = a1["first_name"].str.len()
    assert lens.min() >= 2 and lens.max() <= 20
    d, cl = dedup_table(500, seed=2)
    assert (cl >= 0).sum() >= 90


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_pack_column_matches_pack_strings():
    """Arrow packing gives the same bytes and offsets as the plain packer, for
    str columns (also a sliced pandas column) and for values that need str()."""
    import pandas as pd

    from faster.encode import pack_column, pack_strings

    rng = np.random.default_rng(2)
    words = ["".join("abcé日"[k] for k in rng.integers(0, 5, rng.integers(0, 9))) for _ in range(500)]
    for values in (words, pd.Series(words[100:]).to_numpy(), pd.Series(words).iloc[37:301].to_numpy(),
                   np.array(words, dtype=str), [1, "a", None, 2.5, "None"], []):
        c1, o1 = pack_column(values)
        c2, o2 = pack_strings([v if isinstance(v, str) else str(v) for v in values])
        np.testing.assert_array_equal(o1, o2)
        np.testing.assert_array_equal(c1, c2)
        assert o1.dtype == np.int64 and c1.dtype == np.uint8


def test_output_size_memory_ignores_much_smaller_calls():
    """The hit arena and expansion buffers are sized from earlier calls of a
    comparable size only: a 3,000-pair call where every pair is a hit (rate
    1) must not size a 1e12-pair call's arena at 1e12 entries."""
    from faster._ops import _Rates

    r = _Rates()
    r.add(3000, 1.0)
    assert r.predict(10 ** 12) == 0.0
    r.add(10 ** 12, 4.5e-6)
    assert r.predict(10 ** 12) == 4.5e-6
    assert r.predict(2 * 10 ** 12) == 4.5e-6  # within 64x
    assert r.predict(3000) == 1.0  # small calls still use everything seen at or above their size / 64
    for k in range(40):
        r.add(10 ** 6, 1e-3)
    assert len(r.seen) == 32 and r.predict(10 ** 6) == 1e-3


def test_linkage_pairs_frame_equals_the_reference_join():
    """Linkage.transform's frame (pairs_frame, row-aligned takes) equals the
    reference's construction (faster/linkage.py:44-72: suffixes, Index_A and
    Index_B columns, iloc, merge on Index_B), with non-default index labels, a
    column of df_B that becomes Index_B, repeated rows and an empty selection."""
    import warnings

    import pandas as pd

    from faster.linkage import pairs_frame

    rng = np.random.default_rng(8)
    n_a, n_b = 300, 200
    df_A = pd.DataFrame({"name": rng.integers(0, 50, n_a).astype(str), "year": rng.integers(1930, 2000, n_a),
                         "w": rng.random(n_a)}, index=rng.permutation(n_a) + 1000)
    df_B = pd.DataFrame({"name": rng.integers(0, 50, n_b).astype(str), "Index": rng.integers(0, 9, n_b),
                         "w": rng.random(n_b)}, index=rng.permutation(n_b) * 3)
    for m in (0, 1, 500):
        a = rng.integers(0, n_a, m)
        b = rng.integers(0, n_b, m)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            ra = df_A.add_suffix("_A")
            rb = df_B.add_suffix("_B")
            ra["Index_A"] = range(len(ra))
            rb["Index_B"] = range(len(rb))
            ra = ra.iloc[a, :]
            ra["Index_B"] = b
            ref = ra.merge(rb, on="Index_B")
        got = pairs_frame(df_A, df_B, a, b)
        pd.testing.assert_frame_equal(got, ref, check_index_type=False)


@pytest.mark.parametrize("k,huge", [(1, False), (2, False), (3, False), (2, True)])
def test_exact_totals_per_block_brute_force(k, huge):
    """Per-block counts of record pairs by exact-variable agreement pattern
    (BlockedLinkage's closed-form counts) equal a count over every record
    pair of every block; `huge` value ids overflow the int64 mixed-radix key
    and take the row-grouping path."""
    from faster.blocking import exact_totals_per_block

    rng = np.random.default_rng(5 + k)
    n_a, n_b, n_blocks = 260, 230, 7
    ba = rng.integers(0, n_blocks - 1, n_a).astype(np.int32)  # block n_blocks - 1 stays empty
    bb = rng.integers(0, n_blocks - 1, n_b).astype(np.int32)
    scale = (1 << 40) if huge else 1
    ids_a = [rng.integers(0, 4, n_a).astype(np.int64) * scale + q for q in range(k)]
    ids_b = [rng.integers(0, 4, n_b).astype(np.int64) * scale + q for q in range(k)]
    got = exact_totals_per_block(ids_a, ids_b, ba, bb, n_blocks)
    want = {d: np.zeros(n_blocks, dtype=np.int64) for d in range(2 ** k)}
    for i in range(n_a):
        for j in range(n_b):
            if ba[i] != bb[j]:
                continue
            d = 0
            for q in range(k):
                d = 2 * d + int(ids_a[q][i] == ids_b[q][j])
            want[d][ba[i]] += 1
    assert sorted(got) == sorted(want)
    for d in want:
        np.testing.assert_array_equal(got[d], want[d], err_msg=f"pattern {d}")
