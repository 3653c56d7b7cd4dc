"""Seeded synthetic record tables for tests and benchmarks.

Linkage tables: three string variables (first_name, last_name, street) of
lowercase ASCII with lengths ``clip(round(lognormal(ln 7, 0.4)), 2, 20)`` and
an int32 house number uniform in [1, 9999]. A fraction of B rows (30% by
default) copy distinct A rows with 0..max_edits uniform edits per string
(substitution, insertion, deletion or adjacent transposition, uniformly) and
the same house number; the rest are independent. Deduplication tables add
geometric clusters of perturbed copies and an int16 birth year.
"""
from __future__ import annotations

import numpy as np
import pandas as pd

_ALPHA = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz", dtype=np.uint8)
STRING_VARS = ("first_name", "last_name", "street")


def random_strings(rng, n, mean_len=7.0, sigma=0.4, lo=2, hi=20):
    """``n`` random lowercase strings (list of str)."""
    lens = np.clip(np.round(rng.lognormal(np.log(mean_len), sigma, n)), lo, hi).astype(np.int64)
    buf = _ALPHA[rng.integers(0, 26, int(lens.sum()))].tobytes().decode("ascii")
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    return [buf[offs[i]:offs[i + 1]] for i in range(n)]


def _edit(s: str, kind: int, pos_u: float, ch: int) -> str:
    c = chr(97 + ch)
    n = len(s)
    if kind == 0 and n > 0:  # substitution
        i = int(pos_u * n)
        return s[:i] + c + s[i + 1:]
    if kind == 1:  # insertion
        i = int(pos_u * (n + 1))
        return s[:i] + c + s[i:]
    if kind == 2 and n > 1:  # deletion
        i = int(pos_u * n)
        return s[:i] + s[i + 1:]
    if kind == 3 and n > 1:  # adjacent transposition
        i = int(pos_u * (n - 1))
        return s[:i] + s[i + 1] + s[i] + s[i + 2:]
    return s


def perturb_all(rng, strings, min_edits, max_edits):
    """Applies a uniform number of edits in [min_edits, max_edits] to every string."""
    n = len(strings)
    k = rng.integers(min_edits, max_edits + 1, n)
    tot = int(k.sum())
    kinds = rng.integers(0, 4, tot)
    pos = rng.random(tot)
    chars = rng.integers(0, 26, tot)
    out = []
    j = 0
    for s, ke in zip(strings, k.tolist()):
        for _ in range(ke):
            s = _edit(s, int(kinds[j]), float(pos[j]), int(chars[j]))
            j += 1
        out.append(s)
    return out


def add_numeric_and_missing(dfa, dfb, truth, seed=0, missing=0.05, days=27759, p_same=0.8, max_shift=3,
                            p_error=0.02):
    """Extends linkage tables with a numeric variable, entry errors and
    missing values.

    ``birth_day`` is a day number uniform over ``days`` days (1930-01-01 to
    2005-12-31); a copied B row keeps its A row's day with probability
    ``p_same``, is shifted by 1..``max_shift`` days either way otherwise,
    except that with probability ``p_error`` the shift is 4..60 days (an
    entry error). With probability ``p_error`` a copied B row's house number
    is replaced by a uniform one. Then every value of every compared column
    (the string variables, ``house_number`` and ``birth_day``) is missing with
    probability ``missing``, independently (strings become None, numbers NaN,
    so ``house_number`` becomes float64). Draws come from their own
    generator, so the other columns stay those of ``linkage_tables``."""
    rng = np.random.default_rng(seed + 7919)
    n_a, n_b = len(dfa), len(dfb)
    a_day = rng.integers(0, days, n_a).astype(np.float64)
    b_day = rng.integers(0, days, n_b).astype(np.float64)
    cp = np.nonzero(truth >= 0)[0]
    sign = rng.choice([-1, 1], len(cp))
    shift = rng.integers(1, max_shift + 1, len(cp))
    far = rng.random(len(cp)) < p_error
    shift = np.where(far, rng.integers(max_shift + 1, 61, len(cp)), shift) * sign
    keep = (rng.random(len(cp)) < p_same) & ~far
    b_day[cp] = a_day[truth[cp]] + np.where(keep, 0, shift)
    dfa = dfa.copy()
    dfb = dfb.copy()
    wrong = cp[rng.random(len(cp)) < p_error]
    hn = dfb["house_number"].to_numpy().copy()
    hn[wrong] = rng.integers(1, 10000, len(wrong))
    dfb["house_number"] = hn
    dfa["birth_day"] = a_day
    dfb["birth_day"] = b_day
    for df in (dfa, dfb):
        df["house_number"] = df["house_number"].astype(np.float64)
        for v in list(STRING_VARS) + ["house_number", "birth_day"]:
            miss = rng.random(len(df)) < missing
            if v in STRING_VARS:
                col = df[v].to_numpy(dtype=object).copy()
                col[miss] = None
                df[v] = col
            else:
                col = df[v].to_numpy().copy()
                col[miss] = np.nan
                df[v] = col
    return dfa, dfb


def linkage_tables(n_a, n_b, seed=0, dup_frac=0.3, max_edits=2):
    """(df_A, df_B) for linkage. Returns also the A row of every copied B row
    (-1 for independent rows)."""
    rng = np.random.default_rng(seed)
    A = {v: random_strings(rng, n_a) for v in STRING_VARS}
    A["house_number"] = rng.integers(1, 10000, n_a).astype(np.int32)
    n_dup = min(int(round(dup_frac * n_b)), n_a)
    src = rng.choice(n_a, n_dup, replace=False)
    dst = rng.choice(n_b, n_dup, replace=False)
    B = {v: random_strings(rng, n_b) for v in STRING_VARS}
    B["house_number"] = rng.integers(1, 10000, n_b).astype(np.int32)
    truth = np.full(n_b, -1, dtype=np.int64)
    truth[dst] = src
    for v in STRING_VARS:
        copies = perturb_all(rng, [A[v][i] for i in src.tolist()], 0, max_edits)
        col = B[v]
        for d, s in zip(dst.tolist(), copies):
            col[d] = s
    B["house_number"][dst] = A["house_number"][src]
    return pd.DataFrame(A), pd.DataFrame(B), truth


def dedup_table(n, seed=0, dup_frac=0.2, min_edits=1, max_edits=3, p_geom=0.5):
    """One table for deduplication: ``dup_frac`` of the rows are perturbed
    copies arranged in clusters of 2..5 rows (sizes geometric, truncated)."""
    rng = np.random.default_rng(seed)
    cols = {v: random_strings(rng, n) for v in STRING_VARS}
    cols["birth_year"] = rng.integers(1930, 2006, n).astype(np.int16)
    n_dup = int(round(dup_frac * n))
    cluster = np.full(n, -1, dtype=np.int64)
    perm = rng.permutation(n)
    i = 0
    cid = 0
    while i < n_dup:
        size = int(min(5, 1 + rng.geometric(p_geom)))
        size = max(2, min(size, n_dup - i + 1, n - i))
        if size < 2:
            break
        rows = perm[i:i + size]
        root = rows[0]
        for v in STRING_VARS:
            copies = perturb_all(rng, [cols[v][root]] * (size - 1), min_edits, max_edits)
            for r, s in zip(rows[1:].tolist(), copies):
                cols[v][r] = s
        cols["birth_year"][rows[1:]] = cols["birth_year"][root]
        cluster[rows] = cid
        cid += 1
        i += size
    return pd.DataFrame(cols), cluster


def zipf_block_sizes(rng, n_rows, n_blocks, s=1.2, kmin=10, kmax=60000):
    """Block sizes drawn from a continuous power law k^-s on [kmin, kmax]
    (inverse CDF), rounded and rescaled so that they sum to ``n_rows``."""
    a = 1.0 - s
    u = rng.random(n_blocks)
    k = (kmin ** a + u * (kmax ** a - kmin ** a)) ** (1.0 / a)
    sizes = np.maximum(1, np.round(k * n_rows / k.sum())).astype(np.int64)
    sizes[np.argmax(sizes)] += n_rows - int(sizes.sum())
    return sizes


def blocked_tables(n_rows, n_blocks=4000, seed=0, s=1.2, kmin=10, kmax=60000, dup_frac=0.3, max_edits=2,
                   mean_len=8.0, sigma=0.45, max_len=32):
    """Blocked linkage tables: rows of both sides sorted by block, Zipf block
    sizes (same size on both sides), string lengths lognormal clipped to
    [2, max_len] (spanning the 2-8 / 9-16 / 17-32 buckets), and ``dup_frac`` of
    the B rows perturbed copies of A rows of the same block. Returns (df_A,
    df_B, block ids of A rows, block ids of B rows, sizes)."""
    rng = np.random.default_rng(seed)
    sizes = zipf_block_sizes(rng, n_rows, n_blocks, s, kmin, kmax)
    blk = np.repeat(np.arange(n_blocks), sizes).astype(np.int32)
    starts = np.concatenate([[0], np.cumsum(sizes)])
    A = {v: random_strings(rng, n_rows, mean_len, sigma, 2, max_len) for v in STRING_VARS}
    B = {v: random_strings(rng, n_rows, mean_len, sigma, 2, max_len) for v in STRING_VARS}
    A["house_number"] = rng.integers(1, 10000, n_rows).astype(np.int32)
    B["house_number"] = rng.integers(1, 10000, n_rows).astype(np.int32)
    dup = np.nonzero(rng.random(n_rows) < dup_frac)[0]
    src = starts[blk[dup]] + (rng.random(len(dup)) * sizes[blk[dup]]).astype(np.int64)
    for v in STRING_VARS:
        copies = perturb_all(rng, [A[v][i] for i in src.tolist()], 0, max_edits)
        col = B[v]
        for d, c in zip(dup.tolist(), copies):
            col[d] = c[:max_len] if len(c) > max_len else c
    B["house_number"][dup] = A["house_number"][src]
    return pd.DataFrame(A), pd.DataFrame(B), blk, blk.copy(), sizes
