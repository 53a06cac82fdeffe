"""The block-compressed row layout, on the CPU: a hand-encoded known answer
pins the Python restatement (tests/_blocks.py), and encode -> decode gives
back every edge of random graphs, including rows of 0, 1, 31 .. 33, 63 .. 65,
128, 129 and 197 edges, duplicate destinations (offset width 0), a width of 32 bits, equal
weights (weight width 0) and unweighted rows."""
import numpy as np
import pytest

from tests import _blocks as B


def test_known_answer():
    # one row 5, 7, 12 with weights 3, 1, 2: wbase 1, wbits 2; one block, so
    # the anchor 5 and then the payload: the gap width 3 in bits 0-5, gaps 2
    # and 5 at 3 bits (bits 6-8, 9-11), weights 2, 0, 1 at bits 12-13, 14-15,
    # 16-17 -> 3 + 2*64 + 5*512 + 2*4096 + 0*16384 + 1*65536 = 76419
    ptr = np.array([0, 3], np.uint64)
    row_ptr, words, wbits, wbase = B.encode(ptr, np.array([5, 7, 12], np.uint32), np.array([3, 1, 2], np.uint32))
    assert (wbits, wbase) == (2, 1)
    assert row_ptr.tolist() == [0, 2]
    assert words.tolist() == [5, 76419]
    # a 1-edge unweighted row: the anchor and a payload word holding width 0
    row_ptr, words, _, _ = B.encode(np.array([0, 1], np.uint64), np.array([9], np.uint32))
    assert row_ptr.tolist() == [0, 2] and words.tolist() == [9, 0]
    # two blocks: 65 edges 0, 2, 4, .., 128, unweighted: block 0 anchor 0 and
    # 63 gaps of 2 at 2 bits (126 bits -> 4 words), block 1 the single edge
    # 128 (no payload, its payload offset 4 words)
    ptr = np.array([0, 0, 65], np.uint64)
    row_ptr, words, _, _ = B.encode(ptr, np.arange(0, 130, 2, dtype=np.uint32))
    assert row_ptr.tolist() == [0, 0, 4 + 4]
    assert words[:4].tolist() == [0, (0 << 6) | 2, 128, (4 << 6) | 0]
    bits = sum(int(words[4 + i]) << (32 * i) for i in range(4))
    assert [(bits >> (2 * j)) & 3 for j in range(63)] == [2] * 63
    assert bits >> 126 == 0


def random_graph(rng, n, degs, weighted, wmax=64, span=None):
    deg = np.array(degs, np.int64)
    ptr = np.zeros(n + 1, np.uint64)
    ptr[1:len(deg) + 1] = np.cumsum(deg)
    ptr[len(deg) + 1:] = ptr[len(deg)]
    m = int(ptr[-1])
    hi = span if span is not None else n
    ends = rng.integers(0, hi, m, dtype=np.uint64).astype(np.uint32)
    w = rng.integers(1, wmax + 1, m).astype(np.uint32) if weighted else None
    return ptr, ends, w


@pytest.mark.parametrize("weighted", [False, True])
def test_round_trip_row_lengths(weighted):
    rng = np.random.default_rng(3)
    degs = [0, 1, 31, 32, 33, 63, 64, 65, 0, 2, 128, 129, 197, 5]
    ptr, ends, w = random_graph(rng, 200, degs, weighted)
    ends[ptr[3]:ptr[3] + 10] = 7          # duplicates: a block of equal destinations
    e, ww = B.sort_rows(ptr, ends, w)
    enc = B.encode(ptr, e, ww)
    de, dw = B.decode(*enc[:2], ptr, enc[2], enc[3], weighted)
    assert np.array_equal(de, e)
    if weighted:
        assert np.array_equal(dw, ww)


def test_round_trip_wide_offsets_and_equal_weights():
    rng = np.random.default_rng(5)
    ptr, ends, w = random_graph(rng, 50, [2, 40, 3, 70], True, span=1 << 32)
    ends[0] = 0
    ends[1] = 0xFFFFFFFF                  # row 0: one gap of 2^32 - 1, width 32
    w[:] = 9                              # weight width 0
    e, ww = B.sort_rows(ptr, ends, w)
    row_ptr, words, wbits, wbase = B.encode(ptr, e, ww)
    assert (wbits, wbase) == (0, 9)
    assert words[1] & 63 == 32            # the gap 0 -> 0xFFFFFFFF needs 32 bits (row 0: one block, width in its payload)
    assert words[1] == (32 | (0xFFFFFFFF << 6)) & 0xFFFFFFFF and words[2] == 0xFFFFFFFF >> 26
    de, dw = B.decode(row_ptr, words, ptr, wbits, wbase, True)
    assert np.array_equal(de, e) and np.array_equal(dw, ww)


def test_sort_rows_is_stable_per_row():
    ptr = np.array([0, 4, 6], np.uint64)
    ends = np.array([3, 1, 3, 0, 9, 2], np.uint32)
    w = np.array([10, 11, 12, 13, 14, 15], np.uint32)
    e, ww = B.sort_rows(ptr, ends, w)
    assert e.tolist() == [0, 1, 3, 3, 2, 9]
    assert ww.tolist() == [13, 11, 10, 12, 15, 14]


def test_compression_on_skewed_rows():
    # rows of a skewed graph with sorted destinations take well under the
    # 5 bytes per weighted edge of the split layout
    rng = np.random.default_rng(11)
    n = 4096
    deg = rng.zipf(1.6, n).clip(0, 2000)
    ptr = np.zeros(n + 1, np.uint64)
    ptr[1:] = np.cumsum(deg)
    m = int(ptr[-1])
    ends = (rng.zipf(1.3, m) % n).astype(np.uint32)
    w = rng.integers(1, 65, m).astype(np.uint32)
    e, ww = B.sort_rows(ptr, ends, w)
    row_ptr, words, _, _ = B.encode(ptr, e, ww)
    assert 4 * len(words) < 4.0 * m
