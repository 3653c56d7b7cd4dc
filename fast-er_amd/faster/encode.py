"""Host-side column encoding: unique values, packed bytes and row groups.

The reference takes unique values with ``np.unique`` on the host before any
device work (faster/comparison.py:377,395, faster/deduplication.py:420); this
module does the same once per column and produces the flat arrays the device
stages consume:

* packed UTF-8 bytes of the unique values with int64 offsets (the layout of
  faster/comparison.py:283-302);
* for every unique value, its rows in ascending order (CSR ``starts`` /
  ``rows``), the role of the stable ``cp.argsort`` of the inverse at
  faster/comparison.py:380-391.

Unique values come out in ``np.unique`` order, i.e. lexicographic by code
point, which equals byte order of their UTF-8 encodings.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass
class StringGroups:
    n_rows: int
    chars: np.ndarray    # uint8, packed unique values
    offsets: np.ndarray  # int64[n_unique + 1]
    starts: np.ndarray   # int64[n_unique + 1]
    rows: np.ndarray     # int64[n_rows], rows of each unique value, ascending

    @property
    def n_unique(self) -> int:
        return len(self.offsets) - 1

    def to_device(self) -> "DeviceStringGroups":
        dev = _lib.device()
        chars = self.chars if self.chars.size else np.zeros(1, np.uint8)
        return DeviceStringGroups(
            n_rows=self.n_rows, n_unique=self.n_unique,
            chars=torch.from_numpy(chars).to(dev),
            offsets=torch.from_numpy(self.offsets).to(dev),
            starts=torch.from_numpy(self.starts).to(dev),
            rows=torch.from_numpy(self.rows).to(dev))


@dataclass
class DeviceStringGroups:
    n_rows: int
    n_unique: int
    chars: torch.Tensor
    offsets: torch.Tensor
    starts: torch.Tensor
    rows: torch.Tensor


def _to_str_array(values) -> np.ndarray:
    arr = np.asarray(values, dtype=object)
    return arr.reshape(-1)


def pack_strings(values) -> tuple[np.ndarray, np.ndarray]:
    """Packs strings (UTF-8) into bytes + int64 offsets, no uniquing."""
    arr = _to_str_array(values)
    bs = [s.encode("utf-8") if isinstance(s, str) else bytes(s) for s in arr]
    lens = np.fromiter((len(b) for b in bs), dtype=np.int64, count=len(bs))
    offsets = np.zeros(len(bs) + 1, dtype=np.int64)
    np.cumsum(lens, out=offsets[1:])
    chars = np.frombuffer(b"".join(bs), dtype=np.uint8).copy()
    return chars, offsets


def pack_column(values) -> tuple[np.ndarray, np.ndarray]:
    """Packs a column's values as strings (UTF-8 bytes + int64 offsets),
    through Arrow's buffers when every value is already a ``str``; other
    values go through ``str()`` first, as in group_strings."""
    arr = _to_str_array(values)
    try:
        import pyarrow as pa
    except ImportError:  # pragma: no cover - pyarrow is optional
        pa = None
    if pa is not None:
        try:
            pa_arr = pa.array(arr, type=pa.large_string(), from_pandas=False)
        except (pa.ArrowInvalid, pa.ArrowTypeError):
            pa_arr = None
        if pa_arr is not None and pa_arr.null_count == 0:
            bufs = pa_arr.buffers()
            n = len(pa_arr)
            offs = np.frombuffer(bufs[1], dtype=np.int64, count=n + 1, offset=pa_arr.offset * 8)
            base = int(offs[0])
            offsets = (offs - base).astype(np.int64)
            total = int(offsets[-1])
            chars = (np.frombuffer(bufs[2], dtype=np.uint8, count=total, offset=base).copy()
                     if total else np.zeros(0, np.uint8))
            return chars, offsets
    return pack_strings([s if isinstance(s, str) else str(s) for s in arr])


@dataclass
class PackedColumn:
    """A string column's packed bytes and int64 offsets on the device."""
    n: int
    chars: torch.Tensor
    offsets: torch.Tensor


def packed_column_device(values) -> PackedColumn:
    """Packs a string column on the host (Arrow buffers) and copies it to the
    device."""
    chars, offsets = pack_column(values)
    dev = _lib.device()
    return PackedColumn(len(offsets) - 1, torch.from_numpy(chars if chars.size else np.zeros(1, np.uint8)).to(dev),
                        torch.from_numpy(offsets).to(dev))


def group_packed(col: PackedColumn) -> DeviceStringGroups:
    """Unique values of a device-resident packed column with their rows
    (ascending), grouped on the device (fh_group_strings)."""
    from . import _ops

    uchars, uoff, starts, rows, nu = _ops.group_strings(col.chars, col.offsets, col.n)
    return DeviceStringGroups(n_rows=col.n, n_unique=nu, chars=uchars, offsets=uoff, starts=starts, rows=rows)


def group_packed_multi(cols) -> list:
    """group_packed for several device-resident packed columns in one device
    sort (fh_group_strings_multi, up to 16 columns per call)."""
    from . import _ops

    out = []
    for i in range(0, len(cols), 16):
        part = cols[i:i + 16]
        for col, (uchars, uoff, starts, rows, nu) in zip(part, _ops.group_strings_multi(
                [(c.chars, c.offsets, c.n) for c in part])):
            out.append(DeviceStringGroups(n_rows=col.n, n_unique=nu, chars=uchars, offsets=uoff, starts=starts,
                                          rows=rows))
    return out


def group_strings_device(values) -> DeviceStringGroups:
    """Unique values of a string column with their rows (ascending), grouped
    on the device (fh_group_strings): the host only packs the bytes."""
    return group_packed(packed_column_device(values))


def group_strings(values) -> StringGroups:
    """Unique values of a string column with their rows (ascending)."""
    arr = _to_str_array(values)
    n = len(arr)
    if n == 0:
        z = np.zeros(1, dtype=np.int64)
        return StringGroups(0, np.zeros(0, np.uint8), z, z.copy(), np.zeros(0, np.int64))
    as_str = [s if isinstance(s, str) else str(s) for s in arr]
    has_nul = any("\x00" in s for s in as_str)
    if not has_nul:
        enc = np.array([s.encode("utf-8") for s in as_str], dtype=object)
        fixed = enc.astype(bytes)
        uniq, inv, cnt = np.unique(fixed, return_inverse=True, return_counts=True)
        inv = np.asarray(inv).reshape(-1)
        width = uniq.dtype.itemsize
        ulens = np.char.str_len(uniq).astype(np.int64)
        if width == 0 or len(uniq) == 0:
            chars = np.zeros(0, np.uint8)
        else:
            mat = np.frombuffer(uniq.tobytes(), dtype=np.uint8).reshape(len(uniq), width)
            chars = mat[np.arange(width)[None, :] < ulens[:, None]].copy()
    else:
        uniq_obj, inv, cnt = np.unique(np.array(as_str, dtype=object), return_inverse=True,
                                       return_counts=True)
        inv = np.asarray(inv).reshape(-1)
        bs = [u.encode("utf-8") for u in uniq_obj]
        ulens = np.fromiter((len(b) for b in bs), dtype=np.int64, count=len(bs))
        chars = np.frombuffer(b"".join(bs), dtype=np.uint8).copy()
    offsets = np.zeros(len(ulens) + 1, dtype=np.int64)
    np.cumsum(ulens, out=offsets[1:])
    rows = np.argsort(inv, kind="stable").astype(np.int64)
    starts = np.zeros(len(cnt) + 1, dtype=np.int64)
    np.cumsum(cnt, out=starts[1:])
    return StringGroups(n, chars, offsets, starts, rows)


@dataclass
class ExactLinkGroups:
    """Exact-match grouping of two columns over their joint values."""
    group_a: np.ndarray   # int64[n_A]: joint value id of each A row
    starts_b: np.ndarray  # int64[n_groups + 1]
    rows_b: np.ndarray    # int64[n_B]: B rows of each joint value, ascending
    n_groups: int
    group_b: np.ndarray = None  # int64[n_B]: joint value id of each B row

    def to_device(self):
        dev = _lib.device()
        return (torch.from_numpy(self.group_a).to(dev), torch.from_numpy(self.starts_b).to(dev),
                torch.from_numpy(self.rows_b).to(dev))


def group_exact_link(values_a, values_b) -> ExactLinkGroups:
    """Joint value ids of both columns (np.unique over their concatenation, as
    faster/comparison.py:548 does over the two unique sets)."""
    a = np.asarray(values_a).reshape(-1)
    b = np.asarray(values_b).reshape(-1)
    both = np.concatenate((a, b))
    _, inv = np.unique(both, return_inverse=True)
    inv = np.asarray(inv).reshape(-1).astype(np.int64)
    inv_a, inv_b = inv[:len(a)], inv[len(a):]
    n_groups = int(inv.max()) + 1 if len(inv) else 0
    cnt_b = np.bincount(inv_b, minlength=n_groups).astype(np.int64)
    starts_b = np.zeros(n_groups + 1, dtype=np.int64)
    np.cumsum(cnt_b, out=starts_b[1:])
    rows_b = np.argsort(inv_b, kind="stable").astype(np.int64)
    return ExactLinkGroups(inv_a.copy(), starts_b, rows_b, n_groups, inv_b.copy())


@dataclass
class ExactDedupGroups:
    group: np.ndarray   # int64[n]
    starts: np.ndarray  # int64[n_groups + 1]
    rows: np.ndarray    # int64[n]
    n_groups: int

    def to_device(self):
        dev = _lib.device()
        return (torch.from_numpy(self.group).to(dev), torch.from_numpy(self.starts).to(dev),
                torch.from_numpy(self.rows).to(dev))


def group_exact_dedup(values) -> ExactDedupGroups:
    """Value ids of one column (np.unique, faster/deduplication.py:645)."""
    v = np.asarray(values).reshape(-1)
    _, inv, cnt = np.unique(v, return_inverse=True, return_counts=True)
    inv = np.asarray(inv).reshape(-1).astype(np.int64)
    starts = np.zeros(len(cnt) + 1, dtype=np.int64)
    np.cumsum(cnt, out=starts[1:])
    rows = np.argsort(inv, kind="stable").astype(np.int64)
    return ExactDedupGroups(inv, starts, rows, len(cnt))


# ---------------------------------------------------------------------------
# Exact-match grouping on the device (fh_group_values / fh_group_strings and
# fh_groups_side_b): the role of the host np.unique calls at
# faster/comparison.py:519,535,548 and faster/deduplication.py:645.
# ---------------------------------------------------------------------------
def _numeric_view(values):
    """(int64 array, value kind) for a numeric, boolean or datetime column; None
    for anything else (compared as strings)."""
    from . import _ops

    a = np.asarray(values).reshape(-1)
    if a.dtype.kind == "O":
        import pandas as pd

        inferred = pd.api.types.infer_dtype(a, skipna=False)
        if inferred in ("integer", "boolean"):
            try:
                return a.astype(np.int64), _ops.FH_VALUES_I64
            except (OverflowError, TypeError, ValueError):
                return None
        if inferred in ("floating", "mixed-integer-float"):
            return a.astype(np.float64).view(np.int64), _ops.FH_VALUES_F64
        return None
    if a.dtype.kind in "biu":
        return a.astype(np.int64, copy=False), _ops.FH_VALUES_I64
    if a.dtype.kind in "mM":
        return a.view(np.int64), _ops.FH_VALUES_I64
    if a.dtype.kind == "f":
        return a.astype(np.float64).view(np.int64), _ops.FH_VALUES_F64
    return None


@dataclass
class ExactColumn:
    """One exact-match column (or the concatenation of two) resident on the
    device: int64 values with their kind, or packed strings."""
    n: int
    values: torch.Tensor = None
    kind: int = 0
    chars: torch.Tensor = None
    offsets: torch.Tensor = None


def exact_column(values_a, values_b=None) -> ExactColumn:
    """Uploads a column (with ``values_b``: the rows of A followed by those of
    B) for device grouping. Numeric columns travel as int64 / float64 bits;
    other values are compared as their ``str()``."""
    from . import _ops

    dev = _lib.device()
    cols = [values_a] if values_b is None else [values_a, values_b]
    views = [_numeric_view(v) for v in cols]
    if all(v is not None for v in views):
        if len({k for _, k in views}) == 1:
            arr, kind = np.concatenate([a for a, _ in views]), views[0][1]
        else:  # integers on one side, floats on the other: compared as float64
            arr = np.concatenate([np.asarray(c).reshape(-1).astype(np.float64) for c in cols]).view(np.int64)
            kind = _ops.FH_VALUES_F64
        return ExactColumn(n=len(arr), values=torch.from_numpy(np.ascontiguousarray(arr)).to(dev), kind=kind)
    packed = [pack_column(c) for c in cols]
    if len(packed) > 1:
        (ca, oa), (cb, ob) = packed
        chars = np.concatenate([ca, cb])
        offs = np.concatenate([oa, ob[1:] + oa[-1]])
    else:
        chars, offs = packed[0]
    return ExactColumn(n=len(offs) - 1,
                       chars=torch.from_numpy(chars if chars.size else np.zeros(1, np.uint8)).to(dev),
                       offsets=torch.from_numpy(offs).to(dev))


def _group_column(col: ExactColumn):
    """(group per row or None, starts, rows, n_groups) of a device column."""
    from . import _ops

    if col.values is not None:
        return _ops.group_values(col.values, col.kind)
    _, _, starts, rows, k = _ops.group_strings(col.chars, col.offsets, col.n)
    return None, starts, rows, k


@dataclass
class DeviceExactLinkGroups:
    """Exact-match grouping of two columns over their joint values, on the
    device: value id of every A and B row, and B rows per value (ascending)."""
    n_groups: int
    group_a: torch.Tensor
    group_b: torch.Tensor
    starts_b: torch.Tensor
    rows_b: torch.Tensor

    def to_device(self):
        return self.group_a, self.starts_b, self.rows_b

    def host(self) -> ExactLinkGroups:
        return ExactLinkGroups(self.group_a.cpu().numpy(), self.starts_b.cpu().numpy(), self.rows_b.cpu().numpy(),
                               self.n_groups, self.group_b.cpu().numpy())


def group_exact_link_device(col: ExactColumn, n_a: int) -> DeviceExactLinkGroups:
    """Joint value ids of the rows of ``col`` (A rows then B rows) and the B
    rows of every value, all on the device."""
    from . import _ops

    n_b = col.n - n_a
    group, starts, rows, k = _group_column(col)
    starts_b, rows_b, g2 = _ops.groups_side_b(starts, rows, k, n_a, n_b, want_group=group is None)
    if group is None:
        group = g2
    return DeviceExactLinkGroups(k, group[:n_a], group[n_a:], starts_b, rows_b)


@dataclass
class DeviceExactDedupGroups:
    n_groups: int
    group: torch.Tensor
    starts: torch.Tensor
    rows: torch.Tensor

    def to_device(self):
        return self.group, self.starts, self.rows

    def host(self) -> ExactDedupGroups:
        return ExactDedupGroups(self.group.cpu().numpy(), self.starts.cpu().numpy(), self.rows.cpu().numpy(),
                                self.n_groups)


def group_exact_dedup_device(col: ExactColumn) -> DeviceExactDedupGroups:
    """Value id of every row and the rows of every value, on the device."""
    from . import _ops

    group, starts, rows, k = _group_column(col)
    if group is None:
        _, _, group = _ops.groups_side_b(starts, rows, k, col.n, 0, want_group=True)
    return DeviceExactDedupGroups(k, group, starts, rows)
