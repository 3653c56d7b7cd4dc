"""Fellegi-Sunter estimation: the ``Estimation`` class.

Drop-in for faster/estimation.py (fast-er 0.3.1). The initial values are
drawn on the host exactly as the reference draws them (NumPy's global legacy
RNG, faster/estimation.py:112-120), so seeding ``np.random`` reproduces the
reference's starting point; every EM iteration then runs on the device in FP64
(fh_em_fit). The reference's E-step accumulates in float32
(faster/estimation.py:80); this one stays in float64 throughout.

``Levels`` (keyword-only, optional) generalises the reference's fixed
``[3] * K_Fuzzy + [2] * K_Exact`` level counts.
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

from . import _lib, _ops


class Estimation():
    """Estimates the Fellegi-Sunter model from the pattern counts.

    :param K_Fuzzy: Number of variables compared with Jaro-Winkler.
    :param K_Exact: Number of variables compared for exact equality.
    :param Counts: Observed count of every pattern.
    """

    def __init__(self, K_Fuzzy: int, K_Exact: int, Counts: np.array, *, Levels=None):
        self.K_Fuzzy = K_Fuzzy
        self.K_Exact = K_Exact
        self.Counts = Counts
        if Levels is None:
            Levels = list(np.repeat([3, 2], [K_Fuzzy, K_Exact]))
        self.Levels = [int(x) for x in Levels]
        self.Gamma = self._Gamma()
        self.Lambda = None
        self.Pi = None
        self.Iterations = None
        self._Fit_flag = False

    def _Gamma(self):
        """All patterns of levels, one row per pattern, first variable slowest
        (faster/estimation.py:56-70)."""
        return np.array(list(itertools.product(*(range(i) for i in self.Levels))))

    def _params(self):
        return np.concatenate([[self.Lambda], np.concatenate(self.Pi[0]), np.concatenate(self.Pi[1])])

    def _unpack(self, params):
        p = np.asarray(params, dtype=np.float64)
        off = 1
        pi0, pi1 = [], []
        for L in self.Levels:
            pi0.append(p[off:off + L].copy())
            off += L
        for L in self.Levels:
            pi1.append(p[off:off + L].copy())
            off += L
        return float(p[0]), [pi0, pi1]

    def _match_probability(self):
        """Match probability of every pattern under the current parameters
        (faster/estimation.py:72-96), FP64 on the device."""
        _lib.load()
        params = torch.as_tensor(self._params(), dtype=torch.float64)
        return _ops.match_probability(self.Levels, params).cpu().numpy()

    def fit(self, Tolerance=1e-4, Max_Iter=5000):
        """EM estimation of ``Lambda`` and ``Pi`` (faster/estimation.py:98-158).
        Stops when the largest change in ``Pi`` is below ``Tolerance`` or after
        ``Max_Iter`` iterations."""
        if self._Fit_flag:
            raise Exception("If the model has already been fitted, it cannot be fitted again.")
        _lib.load()
        self.Lambda = 0.1
        pi_0 = [-np.sort(-np.random.dirichlet(np.arange(1, i * 50 + 1, 50))) for i in self.Levels]
        pi_1 = [np.sort(np.random.dirichlet(np.arange(1, i * 50 + 1, 50))) for i in self.Levels]
        self.Pi = [pi_0, pi_1]
        counts = torch.as_tensor(np.asarray(self.Counts, dtype=np.int64))
        params, ksi, it, conv = _ops.em_fit(self.Levels, counts, torch.as_tensor(self._params()),
                                            float(Tolerance), int(Max_Iter))
        self.Lambda, self.Pi = self._unpack(params.cpu().numpy())
        self._Ksi = ksi.cpu().numpy()
        self.Iterations = it
        self._Fit_flag = True
        if conv:
            print("Convergence successfully achieved.")
        else:
            print("Reached the maximum number of iterations without achieving convergence.")

    @property
    def Ksi(self):
        """Match probability of every pattern under the fitted parameters."""
        if not self._Fit_flag:
            raise Exception("The model must be fitted first.")
        try:
            return self._Ksi
        except AttributeError:
            self._Ksi = self._match_probability()
            return self._Ksi
