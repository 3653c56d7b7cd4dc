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
