"""Numeric-distance variables and missing levels on the device against the
oracle's brute force over every record pair (oracle.comparison_fit_full,
oracle.deduplication_fit_full, oracle.numeric_link/_dedup): level lists, all
pattern lists (the lazy missing-only lists materialised), Counts; and the
histogram of the 1M x 1M linkage with five variables and 5% missing values
(768 patterns) through invariants and sampled rows."""
import itertools

import numpy as np
import pandas as pd
import pytest
import torch

pytestmark = pytest.mark.gpu

VF = ["first_name", "last_name", "street"]


def _np(t):
    if hasattr(t, "materialize"):
        t = t.materialize()
    return t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def _dev(x):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device="cuda")


@pytest.mark.parametrize("case", ["ints", "fractions", "nan_ties", "wide"])
def test_numeric_pairs_vs_oracle(oracle, case):
    from faster import _ops

    rng = np.random.default_rng(17)
    if case == "ints":
        xa, xb = rng.integers(0, 200, 900).astype(float), rng.integers(0, 200, 1100).astype(float)
        d2, d1 = 0.0, 3.0
    elif case == "fractions":
        xa, xb = rng.normal(0, 5, 800), rng.normal(0, 5, 700)
        k = min(len(xa[::9]), len(xb[::11]))
        xa[:9 * k:9] = xb[::11][:k]  # some equal values
        d2, d1 = 0.1, 0.35
    elif case == "nan_ties":
        pool = np.array([0.0, -0.0, 1.0, 1.5, 2.0, np.nan, 3.0, 1e300, -1e300])
        xa, xb = pool[rng.integers(0, len(pool), 600)], pool[rng.integers(0, len(pool), 650)]
        d2, d1 = 0.5, 1.0
    else:
        xa, xb = rng.uniform(-1e6, 1e6, 700), rng.uniform(-1e6, 1e6, 900)
        d2, d1 = 1000.0, 20000.0
    print(f"numeric pairs {case}", flush=True)
    keys = _np(_ops.numeric_link(_dev(xa), _dev(xb), d2, d1))
    r1, r2 = oracle.numeric_link(xa, xb, d2, d1)
    assert np.all(np.diff(keys >> 2) > 0)
    np.testing.assert_array_equal(keys[(keys & 3) == 1] >> 2, r1)
    np.testing.assert_array_equal(keys[(keys & 3) == 2] >> 2, r2)
    keys = _np(_ops.numeric_dedup(_dev(xa), d2, d1))
    r1, r2 = oracle.numeric_dedup(xa, d2, d1)
    np.testing.assert_array_equal(keys[(keys & 3) == 1] >> 2, r1)
    np.testing.assert_array_equal(keys[(keys & 3) == 2] >> 2, r2)


def _tables(n_a, n_b, seed, missing):
    from faster.synthetic import add_numeric_and_missing, linkage_tables

    dfa, dfb, truth = linkage_tables(n_a, n_b, seed=seed)
    return add_numeric_and_missing(dfa, dfb, truth, seed=seed, missing=missing)


@pytest.mark.parametrize("missing", [0.0, 0.15])
def test_comparison_numeric_missing_vs_oracle(oracle, missing):
    import faster

    fa, fb = _tables(500, 600, 8, missing)
    comp = faster.Comparison(fa, fb, VF, VF, ["house_number"], ["house_number"], Vars_Numeric_A=["birth_day"],
                             Vars_Numeric_B=["birth_day"])
    comp.fit(Numeric_Dist=(0.0, 3.0), Missing=missing > 0)
    lv, per_var, idx, counts = oracle.comparison_fit_full(
        [fa[v].to_numpy() for v in VF], [fb[v].to_numpy() for v in VF], [fa["house_number"].to_numpy()],
        [fb["house_number"].to_numpy()], [fa["birth_day"].to_numpy()], [fb["birth_day"].to_numpy()], 0.0, 3.0,
        missing=missing > 0)
    assert comp.Var_Levels == lv
    np.testing.assert_array_equal(comp.Counts, counts)
    for v in range(5):
        for l, ref in enumerate(per_var[v]):
            np.testing.assert_array_equal(_np(comp.Levels[v][l]), ref, err_msg=f"variable {v} level {l + 1}")
    for k, (g, r) in enumerate(zip(comp.Indices, idx)):
        np.testing.assert_array_equal(_np(g), r, err_msg=f"pattern {k + 1}")
    # EM over the generalised levels runs on this histogram.
    np.random.seed(0)
    est = faster.Estimation(3, 1, comp.Counts, Levels=comp.Var_Levels)
    est.fit(Tolerance=1e-6)
    np.random.seed(0)
    lam, _, it, _, ksi = oracle.em_fit(lv, counts, 1e-6, 5000)
    assert est.Iterations == it
    np.testing.assert_allclose(est.Ksi, ksi, rtol=1e-9, atol=1e-300)


@pytest.mark.parametrize("missing", [0.0, 0.15])
def test_deduplication_numeric_missing_vs_oracle(oracle, missing):
    import faster
    from faster.synthetic import dedup_table

    df, _ = dedup_table(450, seed=6)
    rng = np.random.default_rng(6)
    df["age"] = rng.integers(20, 60, len(df)).astype(np.float64)
    df["birth_year"] = df["birth_year"].astype(np.float64)
    if missing:
        for v in VF + ["birth_year", "age"]:
            m = rng.random(len(df)) < missing
            col = df[v].to_numpy(dtype=object if v in VF else np.float64).copy()
            col[m] = None if v in VF else np.nan
            df[v] = col
    dd = faster.Deduplication(df, VF, ["birth_year"], Vars_Numeric=["age"])
    dd.fit(Numeric_Dist=(0.0, 2.0), Missing=missing > 0)
    lv, per_var, idx, counts = oracle.deduplication_fit_full([df[v].to_numpy() for v in VF],
                                                             [df["birth_year"].to_numpy()], [df["age"].to_numpy()],
                                                             0.0, 2.0, missing=missing > 0)
    assert dd.Var_Levels == lv
    np.testing.assert_array_equal(dd.Counts, counts)
    for v in range(5):
        for l, ref in enumerate(per_var[v]):
            np.testing.assert_array_equal(_np(dd.Levels[v][l]), ref, err_msg=f"variable {v} level {l + 1}")
    for k, (g, r) in enumerate(zip(dd.Indices, idx)):
        np.testing.assert_array_equal(_np(g), r, err_msg=f"pattern {k + 1}")


def test_all_missing_column():
    """A variable missing on every row of one side: no comparisons, every pair
    has the missing digit on that variable."""
    import faster

    fa, fb = _tables(200, 150, 9, 0.0)
    fa["street"] = None
    comp = faster.Comparison(fa, fb, VF, VF, ["house_number"], ["house_number"])
    comp.fit(Missing=True)
    G = np.array(list(itertools.product(*(range(L) for L in comp.Var_Levels))))
    assert comp.Counts[G[:, 2] == 3].sum() == 200 * 150
    assert comp.Counts.sum() == 200 * 150


# ---------------------- 1M x 1M, 5 variables, 5% missing ---------------------
N = 1_000_000


@pytest.fixture(scope="module")
def config5(config5_fit):
    return config5_fit


def test_config5_histogram_invariants(config5):
    fa, fb, comp = config5
    assert comp.Var_Levels == [4, 4, 4, 3, 4]
    counts = comp.Counts
    assert len(counts) == 768 and counts.min() >= 0 and counts.sum() == N * N
    G = np.array(list(itertools.product(*(range(L) for L in comp.Var_Levels))))
    for v, name in enumerate(VF + ["house_number", "birth_day"]):
        ma, mb = int(pd.isna(fa[name]).sum()), int(pd.isna(fb[name]).sum())
        assert counts[G[:, v] == comp.Var_Levels[v] - 1].sum() == N * N - (N - ma) * (N - mb), name
        # Pairs at each listed level of v, over all patterns, are v's level list.
        for l in range(1, comp.Var_Levels[v] - 1):
            assert counts[G[:, v] == l].sum() == comp.Levels[v][l - 1].numel(), (name, l)
    print("768-pattern histogram of the 1M x 1M records: " + " ".join(str(int(c)) for c in counts[:8]) + " ...", flush=True)


def test_config5_sampled_rows_vs_oracle(oracle, config5):
    """24 A rows against all B rows by the oracle's brute force: every pair
    with a listed pattern (some compared digit above 0) is in the GPU's list
    of that pattern, and no other pair of those rows is listed."""
    fa, fb, comp = config5
    rng = np.random.default_rng(44)
    rows = np.sort(rng.choice(N, 24, replace=False))
    sub = fa.iloc[rows].reset_index(drop=True)
    lv, _, idx, _ = oracle.comparison_fit_full(
        [sub[v].to_numpy() for v in VF], [fb[v].to_numpy() for v in VF], [sub["house_number"].to_numpy()],
        [fb["house_number"].to_numpy()], [sub["birth_day"].to_numpy()], [fb["birth_day"].to_numpy()], 0.0, 3.0,
        missing=True)
    G = np.array(list(itertools.product(*(range(L) for L in lv))))
    missing_only = np.all((G == 0) | (G == np.array(lv) - 1), axis=1)
    ids = comp._pattern_ids.cpu().numpy()
    offs = np.asarray(comp._pattern_offsets)
    a = ids // N
    sel = np.isin(a, rows)
    code = np.searchsorted(offs, np.nonzero(sel)[0], side="right") - 1
    got = ids[sel]
    for c in range(1, len(G)):
        if missing_only[c]:
            continue
        ref = idx[c - 1]
        ref = rows[ref // N] * N + ref % N
        np.testing.assert_array_equal(np.sort(got[code == c]), np.sort(ref), err_msg=f"pattern {c}")
