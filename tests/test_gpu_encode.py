"""Device grouping of string columns (fh_group_strings) against the oracle's
np.unique restatement: unique values byte for byte, group starts and the
ascending rows of every value; edge cases (empty column, empty strings, NUL
bytes, non-ASCII, strings longer than 64 bytes); and a 1M-row column against
the host grouping of the product."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _device_groups(values):
    from faster.encode import group_strings_device

    g = group_strings_device(values)
    offs = g.offsets.cpu().numpy()
    chars = g.chars.cpu().numpy().tobytes()
    uniq = [chars[offs[k]:offs[k + 1]] for k in range(g.n_unique)]
    return uniq, g.starts.cpu().numpy(), g.rows.cpu().numpy(), g


def _check(values, oracle):
    uniq, starts, rows, g = _device_groups(values)
    ru, rs, rr = oracle.group_strings(values)
    print(f"group_strings n={len(values)} unique={len(ru)}", flush=True)
    assert g.n_rows == len(values)
    assert g.n_unique == len(ru)
    assert uniq == ru
    np.testing.assert_array_equal(starts, rs)
    np.testing.assert_array_equal(rows, rr)


@pytest.mark.parametrize("case", ["words", "lengths", "specials", "one", "same", "empties"])
def test_group_strings_vs_oracle(case, oracle):
    rng = np.random.default_rng(11)
    if case == "words":
        pool = ["".join("abcdefgh"[k] for k in rng.integers(0, 8, rng.integers(1, 12))) for _ in range(3000)]
        values = [pool[i] for i in rng.integers(0, len(pool), 20000)]
    elif case == "lengths":
        # 0..300 bytes: every word count of the sort, partial last words.
        pool = ["".join("xy"[k] for k in rng.integers(0, 2, L)) for L in range(301)]
        values = [pool[i] for i in rng.integers(0, len(pool), 5000)]
    elif case == "specials":
        values = ["ab", "ab\x00", "a", "", "ab\x00\x00", "\x00", "é", "e", "日本", "日", "ab", "", "\x00", "zz" * 40,
                  "zz" * 40 + "a", "zz" * 39 + "z"] * 7
    elif case == "one":
        values = ["solo"]
    elif case == "same":
        values = ["smith"] * 1000
    else:
        values = [""] * 64 + ["a"]
    _check(values, oracle)


def test_group_strings_empty_column():
    uniq, starts, rows, g = _device_groups([])
    assert g.n_unique == 0 and uniq == [] and starts.tolist() == [0] and rows.size == 0


def test_group_strings_non_str_values(oracle):
    """Non-string values go through str(), as the host grouping does."""
    values = [12, 7, 12, "7", 3.5, None, "None"]
    _check(values, oracle)


def test_group_strings_1m_vs_host():
    """1M synthetic first names: device grouping equals the host np.unique
    grouping array for array."""
    from faster.encode import group_strings
    from faster.synthetic import random_strings

    rng = np.random.default_rng(3)
    values = np.array(random_strings(rng, 1_000_000), dtype=object)
    print("group_strings 1M", flush=True)
    uniq, starts, rows, g = _device_groups(values)
    h = group_strings(values)
    assert g.n_unique == h.n_unique
    np.testing.assert_array_equal(g.offsets.cpu().numpy(), h.offsets)
    np.testing.assert_array_equal(g.chars.cpu().numpy()[:h.chars.size], h.chars)
    np.testing.assert_array_equal(starts, h.starts)
    np.testing.assert_array_equal(rows, h.rows)


def _np_groups(values):
    """np.unique grouping: value id per row, rows per value (ascending)."""
    _, inv, cnt = np.unique(values, return_inverse=True, return_counts=True)
    inv = np.asarray(inv).reshape(-1)
    starts = np.concatenate([[0], np.cumsum(cnt)])
    return inv, starts, np.argsort(inv, kind="stable")


@pytest.mark.parametrize("case", ["int32", "int64_wide", "float_nan_zero", "bool", "one", "strings", "mixed_num"])
def test_group_exact_columns_vs_np_unique(case):
    """Device grouping of exact columns (fh_group_values, fh_groups_side_b;
    strings through fh_group_strings) equals np.unique's grouping of the
    concatenated A and B values: value ids, starts and ascending rows, and the
    B rows of every value."""
    from faster.encode import exact_column, group_exact_dedup_device, group_exact_link_device

    rng = np.random.default_rng(5)
    n_a, n_b = 3000, 2500
    if case == "int32":
        a = rng.integers(-50, 50, n_a).astype(np.int32)
        b = rng.integers(-50, 50, n_b).astype(np.int32)
    elif case == "int64_wide":
        pool = np.array([-(2 ** 62), -1, 0, 1, 2 ** 40, 2 ** 62 + 7], dtype=np.int64)
        a, b = pool[rng.integers(0, 6, n_a)], pool[rng.integers(0, 6, n_b)]
    elif case == "float_nan_zero":
        pool = np.array([0.0, -0.0, np.nan, -np.nan, 1.5, -1.5, np.inf, -np.inf, 1e-300], dtype=np.float64)
        a, b = pool[rng.integers(0, len(pool), n_a)], pool[rng.integers(0, len(pool), n_b)]
    elif case == "bool":
        a, b = rng.random(n_a) < 0.3, rng.random(n_b) < 0.6
    elif case == "one":
        a, b = np.full(n_a, 7), np.full(n_b, 7)
    elif case == "strings":
        pool = np.array(["x", "", "ab", "é", "ab\x00"], dtype=object)
        a, b = pool[rng.integers(0, 5, n_a)], pool[rng.integers(0, 5, n_b)]
    else:
        a = rng.integers(0, 5, n_a).astype(np.int64)
        b = rng.integers(0, 5, n_b).astype(np.float64) + 0.5 * (rng.random(n_b) < 0.2)
    print(f"exact grouping {case}", flush=True)
    both = np.concatenate([np.asarray(a, dtype=object if case == "strings" else None), b]) \
        if case != "mixed_num" else np.concatenate([a.astype(np.float64), b])
    inv, starts, rows = _np_groups(both)
    g = group_exact_link_device(exact_column(a, b), n_a)
    assert g.n_groups == len(starts) - 1
    np.testing.assert_array_equal(g.group_a.cpu().numpy(), inv[:n_a])
    np.testing.assert_array_equal(g.group_b.cpu().numpy(), inv[n_a:])
    cnt_b = np.bincount(inv[n_a:], minlength=g.n_groups)
    np.testing.assert_array_equal(g.starts_b.cpu().numpy(), np.concatenate([[0], np.cumsum(cnt_b)]))
    np.testing.assert_array_equal(g.rows_b.cpu().numpy(), np.argsort(inv[n_a:], kind="stable"))
    d = group_exact_dedup_device(exact_column(a))
    inv_a, starts_a, rows_a = _np_groups(np.asarray(a, dtype=object) if case == "strings" else a)
    assert d.n_groups == len(starts_a) - 1
    np.testing.assert_array_equal(d.group.cpu().numpy(), inv_a)
    np.testing.assert_array_equal(d.starts.cpu().numpy(), starts_a)
    np.testing.assert_array_equal(d.rows.cpu().numpy(), rows_a)


def test_group_strings_multi_equals_single_columns(oracle):
    """fh_group_strings_multi over columns of different lengths and value
    sets (empty columns first, in the middle and last; identical values; NUL
    bytes; values over 64 bytes) gives each column exactly the oracle's
    grouping (unique values, starts, rows)."""
    from faster.encode import group_packed_multi, packed_column_device

    rng = np.random.default_rng(21)
    pool = ["".join("abcd"[k] for k in rng.integers(0, 4, L)) for L in range(0, 90)]
    cols = [[],
            [pool[i] for i in rng.integers(0, 30, 4000)],
            ["same"] * 700,
            [],
            ["ab", "ab\x00", "a", "", "zz" * 40, "zz" * 40 + "a"] * 9,
            [pool[i] for i in rng.integers(0, len(pool), 2500)],
            []]
    groups = group_packed_multi([packed_column_device(c) for c in cols])
    assert len(groups) == len(cols)
    for k, (values, g) in enumerate(zip(cols, groups)):
        offs = g.offsets.cpu().numpy()
        chars = g.chars.cpu().numpy().tobytes()
        uniq = [chars[offs[i]:offs[i + 1]] for i in range(g.n_unique)]
        ru, rs, rr = oracle.group_strings(values)
        assert g.n_rows == len(values) and g.n_unique == len(ru), k
        assert uniq == ru, k
        np.testing.assert_array_equal(g.starts.cpu().numpy(), rs, err_msg=f"column {k}")
        np.testing.assert_array_equal(g.rows.cpu().numpy(), rr, err_msg=f"column {k}")
