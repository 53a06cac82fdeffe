"""Test infrastructure: the block-compressed row layout (subway_amd/csrc/blocks.hpp)
restated in Python, and the row sort that precedes it.

sort_rows(ptr, ends, w): every row in ascending destination order, equal
destinations keeping their order (the device sort is a stable radix sort by
destination, then by row).
encode(ptr, ends, w): (row_ptr u64[n+1], words u32[T], wbits, wbase) of rows
already sorted (gaps between consecutive destinations of a block; a row of
one block keeps its gap width in the first 6 bits of its payload instead of
a second descriptor word).
decode(row_ptr, words, ptr, wbits, wbase, weighted): (ends, w) back.

Plain Python integers do the bit packing, so it suits graphs up to about a
million edges."""
import numpy as np

BLK = 64
WFIELD = 6


def _bits(x):
    return int(x).bit_length()


def sort_rows(ptr, ends, w=None):
    deg = np.diff(ptr.astype(np.int64))
    row = np.repeat(np.arange(len(deg), dtype=np.int64), deg)
    order = np.lexsort((ends, row))
    return ends[order], (None if w is None else w[order])


def weight_code(w):
    if w is None or len(w) == 0:
        return 0, 0
    base = int(w.min())
    return _bits(int(w.max()) - base), base


def encode(ptr, ends, w=None):
    n = len(ptr) - 1
    wbits, wbase = weight_code(w)
    row_ptr = np.zeros(n + 1, np.uint64)
    out = []
    for v in range(n):
        a, b = int(ptr[v]), int(ptr[v + 1])
        row_ptr[v] = len(out)
        d = b - a
        if d == 0:
            continue
        nb = (d + BLK - 1) // BLK
        desc, pay = [], []
        for k in range(nb):
            e = a + BLK * k
            cnt = min(BLK, d - BLK * k)
            anchor = int(ends[e])
            gaps = [int(ends[e + j]) - int(ends[e + j - 1]) for j in range(1, cnt)]
            wd = _bits(max(gaps, default=0))
            val, pos = (wd, WFIELD) if nb == 1 else (0, 0)
            for g in gaps:
                val |= g << pos
                pos += wd
            if w is not None:
                for j in range(cnt):
                    val |= (int(w[e + j]) - wbase) << pos
                    pos += wbits
            words = (pos + 31) // 32
            desc += [anchor] if nb == 1 else [anchor, (len(pay) << WFIELD) | wd]
            pay += [(val >> (32 * i)) & 0xFFFFFFFF for i in range(words)]
        out += desc + pay
    row_ptr[n] = len(out)
    return row_ptr, np.array(out, dtype=np.uint32), wbits, wbase


def decode(row_ptr, words, ptr, wbits, wbase, weighted):
    """Every edge from its row offset, index and degree alone (the device's
    random access)."""
    m = int(ptr[-1])
    ends = np.zeros(m, np.uint32)
    w = np.zeros(m, np.uint32) if weighted else None
    W = [int(x) for x in words]

    def field(p, bit, width):
        if width == 0:
            return 0
        lo = W[p + (bit >> 5)]
        hi = W[p + (bit >> 5) + 1] if (bit & 31) + width > 32 else 0
        return ((hi << 32 | lo) >> (bit & 31)) & ((1 << width) - 1)

    for v in range(len(ptr) - 1):
        a, d = int(ptr[v]), int(ptr[v + 1] - ptr[v])
        nb = (d + BLK - 1) // BLK
        r = int(row_ptr[v])
        for i in range(d):
            k, j = i // BLK, i % BLK
            cnt = min(BLK, d - BLK * k)
            anchor, c = W[r + 2 * k], W[r + 2 * k + 1]
            if nb == 1:
                # one block: the payload starts at word 1 with the 6-bit width
                wd, p, base = c & 63, r + 1, WFIELD
            else:
                wd, p, base = c & 63, r + 2 * nb + (c >> WFIELD), 0
            ends[a + i] = anchor + sum(field(p, base + (t - 1) * wd, wd) for t in range(1, j + 1))
            if weighted:
                w[a + i] = wbase + field(p, base + (cnt - 1) * wd + j * wbits, wbits)
    return ends, w
