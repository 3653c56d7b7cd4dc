"""Evaluation of a fitted model (drop-in for faster/evaluation.py, fast-er 0.3.1).

Host arithmetic over the pattern histogram (at most a few thousand bins).
``FNR`` keeps the reference's operator precedence,
``sum(Ksi * (Ksi < S) * Counts) / Lambda * sum(Counts)``
(faster/evaluation.py:43), so results match the reference bit for bit;
``FNR_rate`` gives the rate the docstring describes.
"""
from __future__ import annotations

import numpy as np


class Evaluation():
    """Accuracy estimates of the Fellegi-Sunter model.

    :param Lambda: Unconditional match probability.
    :param Ksi: Match probability of every pattern.
    :param Counts: Observed count of every pattern.
    """

    def __init__(self, Lambda: float, Ksi: np.array, Counts: np.array):
        self.Lambda = Lambda
        self.Ksi = Ksi
        self.Counts = Counts

    def FDR(self, S: float):
        """Share of non-matches among pairs with ``Ksi >= S`` (None when no pair qualifies)."""
        ksi = np.asarray(self.Ksi, dtype=np.float64)
        counts = np.asarray(self.Counts)
        den = np.sum((ksi >= S) * counts)
        if den == 0:
            return None
        return np.sum((1 - ksi) * (ksi >= S) * counts) / den

    def FNR(self, S: float):
        """The reference's expression for the false-negative rate."""
        ksi = np.asarray(self.Ksi, dtype=np.float64)
        counts = np.asarray(self.Counts)
        if self.Lambda == 0:
            return None
        return np.sum(ksi * (ksi < S) * counts) / self.Lambda * np.sum(counts)

    def FNR_rate(self, S: float):
        """Share of expected matches with ``Ksi < S``:
        ``sum(Ksi * (Ksi < S) * Counts) / (Lambda * sum(Counts))``."""
        ksi = np.asarray(self.Ksi, dtype=np.float64)
        counts = np.asarray(self.Counts)
        den = self.Lambda * np.sum(counts)
        if den == 0:
            return None
        return np.sum(ksi * (ksi < S) * counts) / den

    def Frontier(self):
        """Plots FDR against FNR for thresholds 0, 0.001, ..., 1."""
        import matplotlib.pyplot as plt

        plt.plot([self.FDR(s / 1000) for s in range(1001)], [self.FNR(s / 1000) for s in range(1001)], ".-")
        plt.xlabel("False Discovery Rate (FDR)")
        plt.ylabel("False Negative Rate (FNR)")
        plt.show()

    def Optimal_Threshold(self, Alpha: float):
        """Threshold in steps of 0.001 minimising ``FDR + Alpha * FNR``."""
        vals = []
        for s in range(1001):
            fdr = self.FDR(s / 1000)
            fnr = self.FNR(s / 1000)
            v = np.nan if fdr is None or fnr is None else fdr + Alpha * fnr
            vals.append(v)
        return np.argmin(np.nan_to_num(np.array(vals, dtype=np.float64), nan=1 + Alpha)) / 1000
