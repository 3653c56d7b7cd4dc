"""Evaluation of a fitted model (drop-in for faster/evaluation.py, fast-er 0.3.1).

Host arithmetic over the pattern histogram (at most a few thousand bins),
evaluated on the objects the caller passed in, as the reference does:

* a threshold that selects no pair makes ``FDR`` a NumPy 0/0, i.e. ``nan``
  (with NumPy's RuntimeWarning), and ``Lambda == 0`` makes ``FNR`` ``inf``
  or ``nan`` the same way (faster/evaluation.py:22-47 divides NumPy scalars,
  which never raises);
* any exception while evaluating (for example ``Ksi`` given as a plain list,
  which does not support ``1 - Ksi``) returns ``None`` (the bare ``except``
  of faster/evaluation.py:31-32,44-45);
* ``FNR`` keeps the reference's operator precedence,
  ``sum(Ksi * (Ksi < S) * Counts) / Lambda * sum(Counts)``
  (faster/evaluation.py:43); ``FNR_rate`` gives the rate the docstring
  describes.
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

    def _selected(self, S):
        """(pairs with Ksi >= S weighted by Counts, the same below S)."""
        above = (self.Ksi >= S) * self.Counts
        below = (self.Ksi < S) * self.Counts
        return above, below

    def FDR(self, S: float):
        """Share of non-matches among pairs with ``Ksi >= S``; ``nan`` when no
        pair qualifies, ``None`` when the inputs cannot be evaluated."""
        try:
            above, _ = self._selected(S)
            false_matches = np.sum((1 - self.Ksi) * above)
            return false_matches / np.sum(above)
        except Exception:
            return None

    def FNR(self, S: float):
        """The reference's expression for the false-negative rate (its
        operator precedence kept); ``None`` when the inputs cannot be
        evaluated."""
        try:
            _, below = self._selected(S)
            missed = np.sum(self.Ksi * below)
            return missed / self.Lambda * np.sum(self.Counts)
        except Exception:
            return None

    def FNR_rate(self, S: float):
        """Share of expected matches with ``Ksi < S``:
        ``sum(Ksi * (Ksi < S) * Counts) / (Lambda * sum(Counts))``."""
        try:
            _, below = self._selected(S)
            expected = np.float64(self.Lambda) * np.sum(self.Counts)
            return np.sum(self.Ksi * below) / expected
        except Exception:
            return None

    def Frontier(self):
        """Plots FDR against FNR for thresholds 0, 0.001, ..., 1."""
        import matplotlib.pyplot as plt

        grid = [s / 1000 for s in range(1001)]
        plt.plot([self.FDR(s) for s in grid], [self.FNR(s) for s in grid], ".-")
        plt.xlabel("False Discovery Rate (FDR)")
        plt.ylabel("False Negative Rate (FNR)")
        plt.show()

    def Optimal_Threshold(self, Alpha: float):
        """Threshold in steps of 0.001 minimising ``FDR + Alpha * FNR``;
        ``nan`` objectives count as ``1 + Alpha`` and infinite ones as the
        largest float (``np.nan_to_num``). A ``None`` from ``FDR`` or ``FNR``
        raises ``TypeError``, as in the reference."""
        objective = [self.FDR(s / 1000) + Alpha * self.FNR(s / 1000) for s in range(1001)]
        return np.argmin(np.nan_to_num(objective, nan=1 + Alpha)) / 1000
