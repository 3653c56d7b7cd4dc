"""Writes tests/golden/realdata_100000.npz from the reference's example data
(example/df_A_100000.Rdata, example/df_B_100000.Rdata: two data.frames of
100,000 voter records each, R XDR serialization, gzip'd).

The files are read as data by the small reader below (no R, no code from the
files is run). The fixture keeps the columns the reference's example notebook
links on (example/Example.ipynb cells 3-5): last_name, first_name,
house_number (a real column, cast to str as the notebook does), street_name,
birth_year, and the truth for precision and recall: for every B row, the A
row with the same true record id (ncid), or -1.

Usage: python tests/golden/make_realdata.py [/path/to/reference]
"""
import gzip
import os
import struct
import sys

import numpy as np

NILVALUE, SYMSXP, LISTSXP, CHARSXP, LGLSXP, INTSXP, REALSXP, STRSXP, VECSXP = 254, 1, 2, 9, 10, 13, 14, 16, 19
REFSXP = 255
NA_INT = -2147483648


class _Reader:
    def __init__(self, data):
        self.d, self.p, self.refs = data, 0, []

    def i32(self):
        v = struct.unpack_from(">i", self.d, self.p)[0]
        self.p += 4
        return v

    def item(self):
        flags = self.i32()
        typ = flags & 0xFF
        has_attr, has_tag = bool(flags & (1 << 9)), bool(flags & (1 << 10))
        if typ == NILVALUE:
            return None
        if typ == REFSXP:
            return self.refs[(flags >> 8) - 1]
        if typ == SYMSXP:
            name = self.item()
            self.refs.append(name)
            return name
        if typ == LISTSXP:
            attr = self.item() if has_attr else None
            tag = self.item() if has_tag else None
            car = self.item()
            rest = self.item()
            out = [(tag, car)] + (rest if isinstance(rest, list) else [])
            return out
        if typ == CHARSXP:
            n = self.i32()
            if n == -1:
                return None
            s = self.d[self.p:self.p + n].decode("utf-8")
            self.p += n
            return s
        if typ in (LGLSXP, INTSXP):
            n = self.i32()
            v = np.frombuffer(self.d, dtype=">i4", count=n, offset=self.p).astype(np.int64)
            self.p += 4 * n
            val = v
        elif typ == REALSXP:
            n = self.i32()
            val = np.frombuffer(self.d, dtype=">f8", count=n, offset=self.p).astype(np.float64)
            self.p += 8 * n
        elif typ == STRSXP:
            n = self.i32()
            val = [self.item() for _ in range(n)]
        elif typ == VECSXP:
            n = self.i32()
            val = [self.item() for _ in range(n)]
        else:
            raise ValueError(f"unsupported SEXP type {typ}")
        attrs = dict(self.item()) if has_attr else {}
        return (val, attrs)


def read_rdata(path):
    """Returns {object name: {column: values}} for the data.frames in path."""
    d = gzip.open(path).read()
    assert d[:5] == b"RDX3\n" and d[5:7] == b"X\n", "not an XDR RData file"
    r = _Reader(d)
    r.p = 7
    r.i32(), r.i32(), r.i32()  # format version, writer and minimal reader versions
    r.p += 4 + r.i32()  # native encoding
    top = r.item()
    out = {}
    for name, obj in top:
        cols, attrs = obj
        out[name] = dict(zip(attrs["names"][0], [c[0] for c in cols]))
    return out


def _pack(strings):
    enc = [x.encode("utf-8") for x in strings]
    lens = np.array([len(b) for b in enc], dtype=np.int64)
    assert lens.max(initial=0) < 256
    return np.frombuffer(b"".join(enc), dtype=np.uint8), lens.astype(np.uint8)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    frames = {}
    for side in ("A", "B"):
        (df,) = read_rdata(os.path.join(ref, "example", f"df_{side}_100000.Rdata")).values()
        frames[side] = df
    # The truth: for every B row, the A row with the same ncid (-1: none).
    ida = np.array(frames["A"]["ncid"], dtype=str)
    idb = np.array(frames["B"]["ncid"], dtype=str)
    assert len(np.unique(ida)) == len(ida) and len(np.unique(idb)) == len(idb)
    order = np.argsort(ida)
    pos = np.clip(np.searchsorted(ida[order], idb), 0, len(ida) - 1)
    arrays = {"B_true_a_row": np.where(ida[order][pos] == idb, order[pos], -1).astype(np.int32)}
    for side, df in frames.items():
        for c in ("last_name", "first_name", "street_name"):
            v = df[c]
            if any(x is None for x in v):
                raise ValueError(f"{c}: NA strings are not expected")
            arrays[f"{side}_{c}_chars"], arrays[f"{side}_{c}_lengths"] = _pack(v)
        hn = np.asarray(df["house_number"])
        assert np.all(hn == np.round(hn)) and np.all(np.abs(hn) < 2 ** 31)
        arrays[f"{side}_house_number"] = hn.astype(np.int32)
        arrays[f"{side}_birth_year"] = np.asarray(df["birth_year"]).astype(np.int16)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "realdata_100000.npz")
    np.savez_compressed(out, **arrays)
    print(out, os.path.getsize(out), "bytes")


def load(path=None):
    """The fixture as two pandas DataFrames (house_number as str, as the
    reference's notebook casts it) and, for every B row, the A row of the same
    person (-1: none)."""
    import pandas as pd

    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "realdata_100000.npz")
    z = np.load(path)
    out = []
    for side in ("A", "B"):
        cols = {}
        for c in ("last_name", "first_name", "street_name"):
            ch = z[f"{side}_{c}_chars"].tobytes()
            off = np.concatenate([[0], np.cumsum(z[f"{side}_{c}_lengths"].astype(np.int64))])
            cols[c] = np.array([ch[off[k]:off[k + 1]].decode("utf-8") for k in range(len(off) - 1)], dtype=object)
        cols["house_number"] = z[f"{side}_house_number"].astype(str).astype(object)
        cols["birth_year"] = z[f"{side}_birth_year"].astype(np.int64)
        out.append(pd.DataFrame(cols))
    return out[0], out[1], z["B_true_a_row"].astype(np.int64)


if __name__ == "__main__":
    main()
