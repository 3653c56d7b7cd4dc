"""Threshold semantics shared by the host code.

The reference compares float32 scores with Python-float thresholds, which
NumPy/CuPy evaluate in float32 (faster/comparison.py:321-329): level 1 when
``lower_thr <= s < upper_thr``, level 2 when ``s >= upper_thr``.
"""
from __future__ import annotations

import numpy as np


def level_of_score(score, lower_thr, upper_thr) -> int:
    s = np.float32(score)
    lo = np.float32(lower_thr)
    hi = np.float32(upper_thr)
    if s >= hi:
        return 2
    if s >= lo and s < hi:
        return 1
    return 0
