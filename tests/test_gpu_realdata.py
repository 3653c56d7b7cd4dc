"""End to end on the reference's own example data: 100,000 x 100,000 North
Carolina voter records (example/df_{A,B}_100000.Rdata, converted by
tests/golden/make_realdata.py), linked as the reference's notebook does
(example/Example.ipynb cell 5): four fuzzy variables (last_name, first_name,
house_number as str, street_name) and birth_year exact, Estimation(4, 1),
Linkage.transform() at 0.85.

Checks:
* the per-variable level lists of a seeded sample of 1,500 A rows equal the
  oracle's lists for those rows against all of B, bit for bit (real names:
  many repeated values, lengths up to 33 bytes, empty first names);
* Counts sum to n_A * n_B and EM on them equals the oracle's FP64 EM;
* precision and recall of the linked pairs against the true ids equal, to
  the four decimals printed, those the reference's notebook reports (0.9802
  and 0.9983, example/Example.ipynb cell 7). The oracle's CPU pipeline gives
  the same figures on this data (50,923 linked pairs); EM converges to the
  same optimum from any seeded start.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VF = ["last_name", "first_name", "house_number", "street_name"]


@pytest.fixture(scope="module")
def realdata():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_realdata

    return make_realdata.load()


@pytest.fixture(scope="module")
def fitted(realdata):
    import faster

    dfa, dfb, truth = realdata
    comp = faster.Comparison(dfa, dfb, VF, VF, ["birth_year"], ["birth_year"])
    comp.fit()
    return comp


def test_sampled_rows_match_oracle(oracle, realdata, fitted):
    dfa, dfb, _ = realdata
    n_b = len(dfb)
    rows = np.sort(np.random.default_rng(5).choice(len(dfa), 1500, replace=False))
    for v, name in enumerate(VF + ["birth_year"]):
        a = dfa[name].to_numpy()[rows]
        b = dfb[name].to_numpy()
        ref = oracle.exact(a, b) if name == "birth_year" else oracle.jaro_winkler_unique(list(a), list(b))
        for lvl, r in enumerate(ref, start=1):
            r = np.asarray(r, dtype=np.int64)
            want = np.sort(rows[r // n_b] * n_b + r % n_b)
            got = fitted.Levels[v][lvl - 1].cpu().numpy()
            got = got[np.isin(got // n_b, rows)]
            np.testing.assert_array_equal(got, want, err_msg=f"{name} level {lvl}")


def test_counts_and_em(oracle, realdata, fitted):
    import faster

    dfa, dfb, _ = realdata
    counts = fitted.Counts
    assert len(counts) == 3 ** 4 * 2
    assert int(counts.sum()) == len(dfa) * len(dfb)
    np.random.seed(0)
    est = faster.Estimation(4, 1, counts)
    est.fit()
    np.random.seed(0)
    lam, Pi, it, _, ksi = oracle.em_fit([3, 3, 3, 3, 2], counts, 1e-4, 5000)
    assert est.Iterations == it
    assert est.Lambda == pytest.approx(lam, rel=1e-9)
    np.testing.assert_allclose(est.Ksi, ksi, rtol=1e-9, atol=1e-300)


def test_precision_recall(realdata, fitted):
    import faster

    dfa, dfb, truth = realdata
    np.random.seed(0)
    est = faster.Estimation(4, 1, fitted.Counts)
    est.fit()
    linked = faster.Linkage(dfa, dfb, fitted.Indices, est.Ksi).transform()
    correct = truth[linked["Index_B"].to_numpy()] == linked["Index_A"].to_numpy()
    precision = float(correct.mean())
    recall = float(correct.sum()) / int((truth >= 0).sum())
    print(f"precision {precision:.4f} recall {recall:.4f} ({len(linked)} linked pairs)")
    assert len(linked) == 50923
    assert round(precision, 4) == 0.9802
    assert round(recall, 4) == 0.9983
