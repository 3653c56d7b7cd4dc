"""The device algorithms' bit-level steps, restated on the CPU and checked
against independent references (no GPU):

* the guard-bit lane recurrence of the comparison kernel (2 to 4 string-2
  masks packed in one 32-bit word, one plain add per character, lane 0
  without a guard in the 3-lane layout) gives, per lane, the match count of a maximum bipartite matching of the Jaro window
  graph (scipy) -- the count the reference's greedy loop yields
  (faster/comparison.py:48-64, pinned in test_oracle.py);
* the table-based scorer of the candidate kernel (greedy match from
  string 2's position masks, transpositions by testing the k-th matched
  position in the k-th matched character's mask, prefix from masks 0..3)
  gives the oracle's score (the C restatement of faster/comparison.py:11-131)
  for every pair.

A restatement that dropped the guard, mis-shifted a lane or paired the wrong
matched positions fails here on random small-alphabet strings, where
collisions, saturated lanes and transpositions are common.
"""
import numpy as np
import pytest

from test_oracle import _max_matching

# Packed layouts of the comparison kernel: strings per dword -> (lane stride
# S in bits, bit of lane 0's first position). Lane l's positions start at
# base + l * S; the bit below is the lane's guard (none under lane 0 when
# base is 0). Strings of up to S - 1 bytes fit a lane.
LAYOUTS = {4: (8, 1), 3: (11, 0), 2: (16, 1)}


def _layout(nl):
    stride, base = LAYOUTS[nl]
    shifts = [base + l * stride for l in range(nl)]
    ones = sum(1 << max(sh - 1, 0) for sh in shifts)
    return stride, shifts, ones


def _window(i, w, lb):
    lo, hi = max(0, i - w), min(lb - 1, i + w)
    return ((1 << (hi - lo + 1)) - 1) << lo if hi >= lo else 0


def _packed_counts(s1, group, nl):
    """Lane counts of the packed recurrence: nc = mask | window | used,
    t = nc + ONES, used = nc ? used : t (bitwise), complemented masks, one
    32-bit word for all lanes."""
    stride, shifts, ones = _layout(nl)
    lb = len(group[0])
    w = max(len(s1), lb) // 2 - 1
    na = 0
    for i, c in enumerate(s1):
        word = 0xFFFFFFFF
        for k, s2 in enumerate(group):
            m = sum(1 << j for j, ch in enumerate(s2) if ch == c)
            word &= ~(m << shifts[k]) & 0xFFFFFFFF
        nwin = ~sum(_window(i, w, lb) << sh for sh in shifts) & 0xFFFFFFFF
        nc = word | nwin | na
        t = (nc + ones) & 0xFFFFFFFF
        na = ((nc & na) | (~nc & t)) & 0xFFFFFFFF
    return [bin((na >> sh) & ((1 << (stride - 1)) - 1)).count("1") for sh in shifts]


def test_layouts_fill_one_dword():
    for nl, (stride, base) in LAYOUTS.items():
        _, shifts, _ = _layout(nl)
        assert shifts[-1] + stride - 1 <= 32, nl  # last lane's positions fit
        assert base == 1 or nl * (stride - 1) + (nl - 1) == 32, nl  # guardless lane 0 only when needed


@pytest.mark.parametrize("nl", [4, 3, 2])
def test_guard_lane_recurrence_equals_maximum_matching(nl):
    rng = np.random.default_rng(40 + nl)
    maxlen = LAYOUTS[nl][0] - 1
    for _ in range(1500):
        lb = int(rng.integers(1, maxlen + 1))
        la = int(rng.integers(1, 33))
        alpha = ["ab", "abc", "abcdefghijklmnopqrstuvwxyz"][int(rng.integers(0, 3))]
        s1 = "".join(alpha[k] for k in rng.integers(0, len(alpha), la))
        group = ["".join(alpha[k] for k in rng.integers(0, len(alpha), lb)) for _ in range(nl)]
        got = _packed_counts(s1, group, nl)
        assert got == [_max_matching(s1, s2) for s2 in group], (s1, group, got)


def _table_score(s1, s2, oracle, p=0.1):
    """The candidate kernel's scorer: only string 2's per-character position
    masks and string 1's bytes are used."""
    la, lb = len(s1), len(s2)
    if la == 0 or lb == 0:
        return 0.0
    rows = {}
    for j, ch in enumerate(s2):
        rows[ch] = rows.get(ch, 0) | (1 << j)
    w = max(la, lb) // 2 - 1
    used = flag1 = 0
    for i, c in enumerate(s1):
        avail = rows.get(c, 0) & _window(i, w, lb) & ~used
        if avail:
            used |= avail & -avail
            flag1 |= 1 << i
    m = bin(used).count("1")
    if m == 0:
        return 0.0
    tt, f, u = 0, flag1, used
    for _ in range(m):
        i = (f & -f).bit_length() - 1
        j = (u & -u).bit_length() - 1
        if not (rows.get(s1[i], 0) >> j) & 1:
            tt += 1
        f &= f - 1
        u &= u - 1
    pre = 0
    for i in range(min(la, lb, 4)):
        if (rows.get(s1[i], 0) >> i) & 1:
            pre += 1
        else:
            break
    return oracle.jw_formula(la, lb, m, tt, pre, p)


def test_table_scorer_equals_oracle(oracle):
    rng = np.random.default_rng(50)
    for _ in range(4000):
        alpha = ["ab", "abc", "abcdefgh"][int(rng.integers(0, 3))]
        s1 = "".join(alpha[k] for k in rng.integers(0, len(alpha), int(rng.integers(1, 20))))
        s2 = "".join(alpha[k] for k in rng.integers(0, len(alpha), int(rng.integers(1, 20))))
        assert _table_score(s1, s2, oracle) == oracle.jw_score(s1, s2), (s1, s2)
