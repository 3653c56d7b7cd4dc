"""Sorted-set operations on device pair lists (faster/search.py:57-157).

``intersect``/``setdiff`` keep the reference's edge cases: an empty input
makes ``intersect`` return an empty uint64 array (search.py:71-72) and
``setdiff`` return ``arr1`` itself (search.py:103-104). The membership test is
the binary-search kernel of libfaster_hip (fh_sorted_member_mask); the pattern
merge in ``Comparison.fit`` does not use these functions (it computes pattern
codes in one pass), they are kept for users who build their own pipelines.
"""
from __future__ import annotations

import torch

from . import _lib, _ops


def _as_dev(a):
    if isinstance(a, torch.Tensor):
        return a.to(_lib.device())
    return torch.as_tensor(a).to(_lib.device())


def intersect(arr1, arr2, num_threads=256):
    """Elements common to ``arr1`` and ``arr2`` (both ascending)."""
    _lib.load()
    if len(arr1) == 0 or len(arr2) == 0:
        return torch.empty(0, dtype=torch.uint64, device=_lib.device())
    a1, a2 = _as_dev(arr1), _as_dev(arr2)
    return a1[_ops.member_mask(a1, a2, True)]


def setdiff(arr1, arr2, num_threads=256):
    """Elements of ``arr1`` absent from ``arr2`` (both ascending)."""
    _lib.load()
    if len(arr1) == 0 or len(arr2) == 0:
        return arr1
    a1, a2 = _as_dev(arr1), _as_dev(arr2)
    return a1[_ops.member_mask(a1, a2, False)]


def reduce(function, iterable, initial=None):
    """Left fold of ``function`` over ``iterable`` (search.py:121-157)."""
    it = iter(iterable)
    value = next(it) if initial is None else initial
    for element in it:
        value = function(value, element)
    return value
