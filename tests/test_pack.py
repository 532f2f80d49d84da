"""Packed partition images (lt_partition_pack) on the host, no GPU needed.

The reference copies a partition's rowptr and col slices as they are
(GPUGraph::loadAsync, src/GraphLoader.h:531-552); an engine with
pack_partitions=1 copies the col slice in the packed form that
include/lighttraffic.h specifies and unpacks it in HBM.  Here the format is
checked on its own terms:
- a decoder written from that specification alone recovers every partition's
  col slice exactly, on sorted (R-MAT), unsorted, duplicate-edge, empty-row
  and chunk-crossing inputs;
- a hand-computed image pins the byte and bit layout;
- sorted R-MAT partitions shrink.
"""
import ctypes

import numpy as np
import pytest

import lighttraffic_amd as lt
from oracle import ffi as oracle


def decode(img, rowptr, eb, V):
    """Column slice from a packed image, following the header's text: code
    u8[n] padded to 4 bytes, chunk_word u32[ceil(m/64)], then the chunks' bit
    streams (LSB first).  rowptr: the partition's n+1 global edge offsets."""
    n = len(rowptr) - 1
    m = int(rowptr[-1] - rowptr[0])
    vbits = max(1, int(V - 1).bit_length())
    code = img[:n]
    off = (n + 3) // 4 * 4
    nch = (m + 63) // 64
    cw = img[off:off + 4 * nch].view(np.uint32)
    words = img[off + 4 * nch:].view(np.uint32)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    row_of = np.repeat(np.arange(n), np.diff(rowptr).astype(np.int64))
    out = np.zeros(m, np.uint32)
    for c in range(nch):
        pos = int(cw[c]) * 32
        for e in range(64 * c, min(64 * c + 64, m)):
            r = row_of[e]
            first = e == int(rowptr[r] - eb)
            anchor = first or e == 64 * c or code[r] == 255
            ln = vbits if anchor else int(code[r])
            v = 0
            for k in range(ln):
                v |= int(bits[pos + k]) << k
            pos += ln
            out[e] = v if anchor else out[e - 1] + v
    return out


def check_graph(rowptr, col, part_bytes):
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, part_bytes)
    no, eo = part.node_offsets(), part.edge_offsets()
    sizes = []
    for p in range(part.count):
        img = part.pack(g, p)
        rp = rowptr[no[p]:no[p + 1] + 1]
        got = decode(img, rp, eo[p], g.num_node)
        assert np.array_equal(got, col[eo[p]:eo[p + 1]]), p
        sizes.append((len(img), 4 * int(eo[p + 1] - eo[p])))
    return sizes


def test_hand_computed_image():
    # one row [5, 6, 8] in a 16-vertex graph: vbits = 4, gaps 1 and 2 -> width 2;
    # bits: 5 (4 bits) | 1 (2 bits) | 2 (2 bits) = 0b10_01_0101 = 149
    V = 16
    rowptr = np.array([0, 3] + [3] * (V - 1), np.uint64)
    col = np.array([5, 6, 8], np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    assert part.count == 1
    img = part.pack(g, 0)
    # code u8[16] (row 0: width 2), one chunk word (0), words: [149, 0, 0]
    expect = np.zeros(16 + 4 + 12, np.uint8)
    expect[0] = 2
    expect[20:24] = np.array([149], np.uint32).view(np.uint8)
    assert np.array_equal(img, expect)


def test_hand_computed_unsorted_and_anchor():
    # row 0 = [3, 1] (unsorted -> code 255, both absolute); row 1 = 65 edges
    # 0..64 in order: edge 64 of the partition starts chunk 1 and is an anchor
    V = 70
    rowptr = np.array([0, 2, 67] + [67] * (V - 2), np.uint64)
    col = np.concatenate([[3, 1], np.arange(65)]).astype(np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    img = part.pack(g, 0)
    assert img[0] == 255 and img[1] == 1  # row 1's gaps are all 1: width 1
    n, vbits = V, 7
    cw = img[72:80].view(np.uint32)
    # chunk 0: 3 anchors (3, 1, first of row 1) of 7 bits + 61 gaps of 1 bit = 82 bits -> 3 words
    assert list(cw) == [0, 3]
    assert np.array_equal(decode(img, rowptr[:n + 1], 0, V), col)


@pytest.mark.parametrize("scale,ef,part_bytes", [(8, 4, 2048), (10, 8, 8192), (11, 16, 1 << 20)])
def test_rmat_round_trip_and_size(scale, ef, part_bytes):
    rowptr, col = oracle.rmat_graph(scale, ef, seed=scale)
    sizes = check_graph(rowptr, col, part_bytes)
    packed = sum(s for s, _ in sizes)
    raw = sum(r for _, r in sizes)
    assert packed < 0.8 * raw  # sorted neighbour lists: small gaps


def test_unsorted_duplicates_and_empty_rows():
    rng = np.random.default_rng(7)
    V = 300
    deg = rng.integers(0, 150, V)
    deg[rng.random(V) < 0.3] = 0  # empty rows
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = np.empty(int(rowptr[-1]), np.uint32)
    for r in range(V):
        a, b = int(rowptr[r]), int(rowptr[r + 1])
        x = np.sort(rng.integers(0, V, b - a))  # sorted, with duplicates (gap 0)
        if r % 3 == 0:
            rng.shuffle(x)  # unsorted rows are stored absolute
        col[a:b] = x
    check_graph(rowptr, col, 4096)
    check_graph(rowptr, col, 1 << 20)


@pytest.mark.parametrize("V", [1, 2, 3, 1 << 10, (1 << 10) + 1])
def test_value_widths_at_powers_of_two(V):
    # anchors use bits(V-1) (at least 1): V-1 = 2^k - 1 and 2^k as largest ids
    rowptr = np.zeros(V + 1, np.uint64)
    rowptr[1:] = 3
    col = np.array([V - 1, V - 1, V - 1], np.uint32)
    check_graph(rowptr, col, 1 << 20)


def test_pack_arguments():
    rowptr, col = oracle.rmat_graph(8, 4, seed=1)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 2048)
    with pytest.raises(lt.LTError):
        part.pack(g, part.count)
    L = lt.lib()
    n = ctypes.c_uint64()
    assert L.lt_partition_pack(g._h, part._h, 0, None, 0, ctypes.byref(n)) == 0 and n.value > 0
    buf = np.zeros(n.value - 1, np.uint8)
    assert L.lt_partition_pack(g._h, part._h, 0, buf.ctypes.data, n.value - 1, ctypes.byref(n)) == 6  # LT_E_ARG
