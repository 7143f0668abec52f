"""PLINK-1 binary trio (.bed/.bim/.fam), SNP-major, as SnpArrays reads it.

Only the payload layout matters to the path: after the 3-byte magic
0x6C 0x1B 0x01 come J columns of ceil(I/4) bytes, individual i in byte i>>2
at bit offset 2*(i&3), codes 00 hom A1 / 01 missing / 10 het / 11 hom A2.
``SnpArray.data`` is that payload as a ceil(I/4) x J column-major matrix, the
``g.s.data`` the reference's loops index (src/loops.jl:529).
"""
from __future__ import annotations

import os

import numpy as np

MAGIC = bytes([0x6C, 0x1B, 0x01])


def _count_lines(path):
    with open(path, "rb") as f:
        return sum(1 for line in f if line.strip())


class SnpArray:
    """Memory-mapped .bed with sizes taken from the .fam (I) and .bim (J) files."""

    def __init__(self, filename: str, I: int | None = None, J: int | None = None):
        if not filename.endswith(".bed"):
            raise ValueError("filename should end with .bed")
        stem = filename[:-4]
        if I is None:
            I = _count_lines(stem + ".fam")
        if J is None:
            J = _count_lines(stem + ".bim")
        self.I, self.J = int(I), int(J)
        self.path = filename
        nb = (self.I + 3) // 4
        size = os.path.getsize(filename)
        if size != 3 + nb * self.J:
            raise ValueError(f"{filename}: size {size} does not match I={self.I}, J={self.J}")
        with open(filename, "rb") as f:
            if f.read(3) != MAGIC:
                raise ValueError(f"{filename}: not a SNP-major PLINK-1 .bed (bad magic)")
        self.data = np.memmap(filename, dtype=np.uint8, mode="r", offset=3, shape=(nb, self.J), order="F")

    @property
    def shape(self):
        return (self.I, self.J)


class SnpLinAlg:
    """Typed wrapper of a SnpArray (the reference's SnpLinAlg{T}); the loops
    read ``g.s.data``."""

    def __init__(self, s: SnpArray, T=np.float64):
        self.s = s
        self.T = np.dtype(T)

    @property
    def shape(self):
        return self.s.shape


def write_plink(prefix: str, data: np.ndarray, I: int, J: int) -> str:
    """Write prefix.bed/.bim/.fam for a ceil(I/4) x J payload; returns the .bed path."""
    data = np.asarray(data, dtype=np.uint8)
    nb = (I + 3) // 4
    if data.shape != (nb, J):
        data = data.reshape((nb, J), order="F")
    with open(prefix + ".bed", "wb") as f:
        f.write(MAGIC)
        f.write(np.asfortranarray(data).tobytes(order="F"))
    with open(prefix + ".fam", "w") as f:
        for i in range(I):
            f.write(f"fam{i} ind{i} 0 0 0 -9\n")
    with open(prefix + ".bim", "w") as f:
        for j in range(J):
            f.write(f"1 snp{j} 0 {j + 1} A C\n")
    return prefix + ".bed"


def _read_lines(path):
    with open(path, "r") as f:
        return [line.rstrip("\r\n") + "\n" for line in f if line.strip()]


def filter_plink(src: str, rowmask=None, cols=None, des: str = "filtered") -> str:
    """SnpArrays.filter(src, rowmask, colmask; des) as src/driver.jl:97 and
    :120 use it: write des.bed/.bim/.fam holding the individuals where
    ``rowmask`` is true (None: all) and the SNPs ``cols`` (0-based indices in
    the order given; None: all), with the source's .fam/.bim lines.  Returns
    the .bed path."""
    stem = src[:-4] if src.endswith(".bed") else src
    g = SnpArray(stem + ".bed")
    I, J = g.shape
    cols = np.arange(J) if cols is None else np.asarray(cols, dtype=np.int64)
    if cols.size and (cols.min() < 0 or cols.max() >= J):
        raise ValueError("SNP index out of range")
    rows = np.ones(I, bool) if rowmask is None else np.asarray(rowmask, dtype=bool)
    if rows.shape != (I,):
        raise ValueError(f"rowmask must have {I} entries")
    keep = np.flatnonzero(rows)
    In = keep.size
    nb_out = (In + 3) // 4
    if In == I:
        payload = np.asarray(g.data[:, cols])
    else:
        payload = np.zeros((nb_out, cols.size), np.uint8)
        src_byte, src_shift = keep >> 2, 2 * (keep & 3)
        dst_byte, dst_shift = np.arange(In) >> 2, 2 * (np.arange(In) & 3)
        for t, j in enumerate(cols):
            codes = (np.asarray(g.data[src_byte, j]) >> src_shift) & 3
            np.bitwise_or.at(payload[:, t], dst_byte, (codes << dst_shift).astype(np.uint8))
    with open(des + ".bed", "wb") as f:
        f.write(MAGIC)
        f.write(np.asfortranarray(payload).tobytes(order="F"))
    fam = _read_lines(stem + ".fam")
    bim = _read_lines(stem + ".bim")
    with open(des + ".fam", "w") as f:
        f.writelines(fam[i] for i in keep)
    with open(des + ".bim", "w") as f:
        f.writelines(bim[j] for j in cols)
    return des + ".bed"
