"""Linkage: the ``Linkage`` class (drop-in for faster/linkage.py, fast-er 0.3.1)."""
from __future__ import annotations

import numpy as np
import pandas as pd
import torch

from . import _lib


def _numel(x) -> int:
    return int(x.numel()) if isinstance(x, torch.Tensor) else int(np.asarray(x).size)


class Linkage():
    """Links records of two DataFrames whose pattern's match probability
    reaches a threshold.

    :param df_A: First DataFrame.
    :param df_B: Second DataFrame.
    :param Indices: Per-pattern pair lists (``Comparison.Indices``).
    :param Ksi: Match probability of every pattern (``Estimation.Ksi``).
    """

    def __init__(self, df_A: pd.DataFrame, df_B: pd.DataFrame, Indices, Ksi: np.array):
        self.df_A = df_A
        self.df_B = df_B
        self.Indices = Indices
        self.Ksi = Ksi

    def transform(self, Threshold=0.85, Compat=True):
        """DataFrame of all pairs whose match probability is at least
        ``Threshold`` (faster/linkage.py:26-72).

        ``Indices`` omits pattern 0 (every variable at level 0), so pattern
        ``k >= 1`` is ``Indices[k - 1]``. With ``Compat=True`` (the default)
        pattern 0 reads ``Indices[-1]`` as the reference's does
        (faster/linkage.py:54), i.e. the last pattern's list is linked when
        pattern 0's probability reaches the threshold; ``Compat=False`` skips
        pattern 0 instead (its pairs are never listed)."""
        above = np.ravel(np.argwhere(np.asarray(self.Ksi) >= Threshold))
        if not Compat:
            above = above[above > 0]
        if np.sum([_numel(self.Indices[i - 1]) for i in above]) == 0:
            raise Exception("No pair of observations has a conditional match probability exceeding the threshold.")
        dev = _lib.device()
        parts = []
        for i in above:
            x = self.Indices[i - 1]
            if hasattr(x, "materialize"):
                x = x.materialize()
            parts.append(x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, dtype=np.int64)))
        ids = torch.cat([x.to(dev, torch.int64) for x in parts])
        n_b = len(self.df_B)
        a = (ids // n_b).cpu().numpy()
        b = (ids % n_b).cpu().numpy()
        return pairs_frame(self.df_A, self.df_B, a, b)


def pairs_frame(df_A: pd.DataFrame, df_B: pd.DataFrame, a, b) -> pd.DataFrame:
    """The reference's result for the pairs (a[k], b[k]) (faster/linkage.py:
    44-72): the rows of df_A and df_B with the suffixes _A and _B, Index_A,
    Index_B, in pair order, on a fresh RangeIndex. The reference joins on
    Index_B, which is unique on the right (it overwrites any column of that
    name) and keeps the left order; row-aligned takes of both frames give the
    same frame without the join."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    left = df_A.iloc[a].add_suffix("_A").reset_index(drop=True)
    left["Index_A"] = a
    left["Index_B"] = b
    right = df_B.iloc[b].add_suffix("_B").reset_index(drop=True).drop(columns=["Index_B"], errors="ignore")
    return pd.concat([left, right], axis=1)
