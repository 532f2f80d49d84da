"""Packed partition images (lt_partition_pack) on the host, no GPU needed.

The reference copies a partition's rowptr and col slices as they are
(GPUGraph::loadAsync, src/GraphLoader.h:193-215); an engine with
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
    """Column slice from a packed image, following the header's text: meta
    u32[2*ceil(m/64)] (first word of each chunk, codes of its four 16-edge
    blocks), then the chunks' bit streams (LSB first).  rowptr: the
    partition's n+1 global edge offsets."""
    n = len(rowptr) - 1
    m = int(rowptr[-1] - rowptr[0])
    vbits = max(1, int(V - 1).bit_length())
    nch = (m + 63) // 64
    meta = img[:8 * nch].view(np.uint32)
    words = img[8 * nch:].view(np.uint32)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    row_start = np.zeros(m + 1, bool)
    row_start[(rowptr[:-1] - eb).astype(np.int64)] = True
    out = np.zeros(m, np.uint32)
    for c in range(nch):
        pos = int(meta[2 * c]) * 32
        for e in range(64 * c, min(64 * c + 64, m)):
            code = (int(meta[2 * c + 1]) >> (8 * ((e - 64 * c) // 16))) & 0xFF
            anchor = row_start[e] or e == 64 * c or code == 255
            ln = vbits if anchor else code
            v = 0
            for k in range(ln):
                v |= int(bits[pos + k]) << k
            pos += ln
            out[e] = v if anchor else out[e - 1] + v
    return out


def decode_blocks(img, m):
    """Column slice from a block image, following the header's text:
    hdr u32[2*(nblk+1)] ({base, first bit} per 16-edge block, then {0, total
    bits}), then the bit streams (LSB first) and two zero words."""
    nblk = (m + 15) // 16
    hdr = img[:8 * (nblk + 1)].view(np.uint32)
    words = img[8 * (nblk + 1):].view(np.uint32)
    assert len(words) >= 2 and not words[-2:].any()
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    out = np.zeros(m, np.uint32)
    for e in range(m):
        k = e // 16
        base, off = int(hdr[2 * k]), int(hdr[2 * k + 1])
        w = (int(hdr[2 * k + 3]) - off) // 16
        pos = off + (e % 16) * w
        out[e] = base + sum(int(bits[pos + j]) << j for j in range(w))
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
        blk = part.pack(g, p, blocks=True)
        got = decode_blocks(blk, int(eo[p + 1] - eo[p]))
        assert np.array_equal(got, col[eo[p]:eo[p + 1]]), p
        sizes.append((len(img), 4 * int(eo[p + 1] - eo[p])))
    return sizes


def test_hand_computed_image():
    # one row [5, 6, 8] in a 16-vertex graph: vbits = 4; block 0's gaps 1 and 2 -> width 2;
    # bits: 5 (4 bits) | 1 (2 bits) | 2 (2 bits) = 0b10_01_0101 = 149
    V = 16
    rowptr = np.array([0, 3] + [3] * (V - 1), np.uint64)
    col = np.array([5, 6, 8], np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    assert part.count == 1
    img = part.pack(g, 0)
    # meta: chunk 0 starts at word 0, block codes (2, 0, 0, 0); words: [149, 0, 0]
    expect = np.array([0, 2, 149, 0, 0], np.uint32).view(np.uint8)
    assert np.array_equal(img, expect)


def test_hand_computed_block_image():
    # one row [5, 6, 8]: one block, base 5, values 0, 1, 3 (+ 13 padding zeros)
    # in bitlen(8 - 5) = 2 bits each: 0 | 1 << 2 | 3 << 4 = 52, 32 bits in all
    V = 16
    rowptr = np.array([0, 3] + [3] * (V - 1), np.uint64)
    col = np.array([5, 6, 8], np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    img = part.pack(g, 0, blocks=True)
    assert np.array_equal(img, np.array([5, 0, 0, 32, 52, 0, 0], np.uint32).view(np.uint8))


def test_block_image_unsorted_and_zero_width():
    # block 0 = [999, 0, 7 x 14]: base 0, width bitlen(999) = 10, unsorted is
    # fine; block 1 = [3 x 16]: width 0, no bits at all
    V = 1000
    col = np.array([999, 0] + [7] * 14 + [3] * 16, np.uint32)
    rowptr = np.array([0, 32] + [32] * (V - 1), np.uint64)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    img = part.pack(g, 0, blocks=True)
    assert list(img[:24].view(np.uint32)) == [0, 0, 3, 160, 0, 160]
    assert len(img) == 24 + 4 * (5 + 2)
    assert np.array_equal(decode_blocks(img, 32), col)


def test_hand_computed_raw_block_and_anchors():
    # row 0 = [3, 1] (descending -> block 0 stored absolute); row 1 = 65 edges
    # 0..64 in order: edge 64 of the partition starts chunk 1 and is an anchor
    V = 70
    rowptr = np.array([0, 2, 67] + [67] * (V - 2), np.uint64)
    col = np.concatenate([[3, 1], np.arange(65)]).astype(np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    img = part.pack(g, 0)
    meta = img[:16].view(np.uint32)
    vbits = 7
    # chunk 0: block 0 = edges 0..15, all absolute (16 x 7 bits); blocks 1-3 gaps of 1
    # -> 112 + 48 = 160 bits = 5 words; chunk 1: anchor 64 (7 bits) + gaps 1, 1 -> 1 word
    assert list(meta) == [0, 255 | (1 << 8) | (1 << 16) | (1 << 24), 5, 1]
    assert len(img) == 16 + 4 * (5 + 1 + 2)
    assert np.array_equal(decode(img, rowptr[:V + 1], 0, V), col)
    assert vbits == max(1, (V - 1).bit_length())


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


def test_partitions_without_edges():
    # 5000 vertices without neighbours between two R-MAT-like blocks: with 8 KB
    # partitions several partitions hold no edge (their image is two zero words)
    rng = np.random.default_rng(3)
    V = 9000
    deg = rng.integers(1, 20, V)
    deg[2000:7000] = 0
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = np.concatenate([np.sort(rng.integers(0, V, d)) for d in deg]).astype(np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 8192)
    eo = part.edge_offsets()
    empty = [p for p in range(part.count) if eo[p + 1] == eo[p]]
    assert len(empty) >= 2
    for p in empty:
        assert np.array_equal(part.pack(g, p), np.zeros(8, np.uint8))
    check_graph(rowptr, col, 8192)


def test_slot_layout_hand_computed():
    # one row 0..63 in a 64-vertex graph: 4 blocks of width bitlen(15) = 4,
    # image = 5 header pairs (40 B) + 8 stream words + 2 zero words (40 B) =
    # 80 B < 256 B raw; slot = rowptr 65 * 8 + 80 = 600 -> 768 bytes
    V = 64
    rowptr = np.array([0, 64] + [64] * (V - 1), np.uint64)
    col = np.arange(64, dtype=np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 1 << 20)
    assert len(part.pack(g, 0, blocks=True)) == 80
    assert part.slot_layout(g, 2) == (768, 1)
    assert part.slot_layout(g, 0) == part.slot_layout(g, 1) == (1 << 20, 1)
    # [5, 6, 8]: the block image (28 B) is not smaller than the columns (12 B),
    # so no partition is kept packed and slots stay raw
    rowptr = np.array([0, 3] + [3] * 15, np.uint64)
    g = lt.Graph.from_csr(rowptr, np.array([5, 6, 8], np.uint32))
    part = lt.Partition.build(g, 1 << 20)
    assert part.slot_layout(g, 2) == (1 << 20, 1)


def test_slot_layout_rmat():
    rowptr, col = oracle.rmat_graph(11, 16, seed=11)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 16 << 10)
    no, eo = part.node_offsets(), part.edge_offsets()
    need = 0
    for p in range(part.count):
        m = int(eo[p + 1] - eo[p])
        b = len(part.pack(g, p, blocks=True))
        need = max(need, 8 * int(no[p + 1] - no[p] + 1) + (b if b < 4 * m else 4 * m))
    sb, n = part.slot_layout(g, 2, 4)
    assert sb == (need + 255) // 256 * 256 and sb < 16 << 10
    assert n == min(part.count, 4 * (16 << 10) // sb) and n > 4
    assert part.slot_layout(g, 2, 0)[1] == part.count  # every partition has a slot either way
    assert part.slot_layout(g, 1, 4) == (16 << 10, 4)
    with pytest.raises(lt.LTError):
        part.slot_layout(g, 3)


def test_pack_blocks_arguments():
    rowptr, col = oracle.rmat_graph(8, 4, seed=1)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 2048)
    with pytest.raises(lt.LTError):
        part.pack(g, part.count, blocks=True)
    L = lt.lib()
    n = ctypes.c_uint64()
    assert L.lt_partition_pack_blocks(g._h, part._h, 0, None, 0, ctypes.byref(n)) == 0 and n.value > 0
    assert n.value == len(part.pack(g, 0, blocks=True))
    buf = np.zeros(n.value - 1, np.uint8)
    assert L.lt_partition_pack_blocks(g._h, part._h, 0, buf.ctypes.data, n.value - 1, ctypes.byref(n)) == 6
    assert L.lt_partition_pack_blocks(None, part._h, 0, None, 0, ctypes.byref(n)) == 6


def span_col_emulated(img, m, e):
    """The walk kernel's decode of edge indices e (u64 array) from a block
    image, in the kernel's 32-bit operations: the header pair of block e/16,
    the next block's offset, two stream words joined by alignbit, a mask."""
    nblk = (m + 15) // 16
    hdr = img.view(np.uint32)
    words = hdr[2 * (nblk + 1):]
    k = (e >> np.uint64(4)).astype(np.int64)
    hx, hy = hdr[2 * k].astype(np.uint64), hdr[2 * k + 1].astype(np.uint64)
    w = ((hdr[2 * k + 3].astype(np.uint64) - hy) & np.uint64(0xFFFFFFFF)) >> np.uint64(4)
    pos = (hy + (e & np.uint64(15)) * w) & np.uint64(0xFFFFFFFF)
    lo = words[(pos >> np.uint64(5)).astype(np.int64)].astype(np.uint64)
    hi = words[((pos >> np.uint64(5)) + np.uint64(1)).astype(np.int64)].astype(np.uint64)
    bits = (((hi << np.uint64(32)) | lo) >> (pos & np.uint64(31))) & np.uint64(0xFFFFFFFF)
    mask = np.where(w >= 32, np.uint64(0xFFFFFFFF), (np.uint64(1) << np.minimum(w, np.uint64(31))) - np.uint64(1))
    return ((hx + (bits & mask)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


@pytest.mark.parametrize("scale,ef,part_bytes", [(12, 16, 64 << 10), (10, 8, 4096)])
def test_kernel_block_decode_emulated(scale, ef, part_bytes):
    # every edge of every partition, through the kernel's arithmetic; reads stay
    # inside the image (the two zero words cover the last window)
    rowptr, col = oracle.rmat_graph(scale, ef, seed=scale + 1)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, part_bytes)
    eo = part.edge_offsets()
    for p in range(part.count):
        m = int(eo[p + 1] - eo[p])
        if m == 0:
            continue
        img = part.pack(g, p, blocks=True)
        got = span_col_emulated(img, m, np.arange(m, dtype=np.uint64))
        assert np.array_equal(got, col[eo[p]:eo[p + 1]]), p


def test_slot_layout_rejects_foreign_partition_table():
    big_rp, big_col = oracle.rmat_graph(10, 8, seed=2)
    small_rp, small_col = oracle.rmat_graph(8, 4, seed=2)
    big, small = lt.Graph.from_csr(big_rp, big_col), lt.Graph.from_csr(small_rp, small_col)
    part = lt.Partition.build(big, 4096)
    with pytest.raises(lt.LTError) as e:
        part.slot_layout(small, 2)
    assert e.value.status == 6  # LT_E_ARG
    # the other direction: a table built for the smaller graph, asked about the larger one
    part_small = lt.Partition.build(small, 4096)
    with pytest.raises(lt.LTError) as e:
        part_small.slot_layout(big, 2)
    assert e.value.status == 6
    # same vertex count, other edges: offsets no longer match the rowptr
    other_rp, other_col = oracle.rmat_graph(10, 8, seed=3)
    other = lt.Graph.from_csr(other_rp, other_col)
    if other.num_node == big.num_node:
        with pytest.raises(lt.LTError):
            part.slot_layout(other, 0)
    assert part.slot_layout(big, 0)[1] == part.count


def block_slot_need(rowptr, col, n0, n1):
    """Bytes a block slot needs for vertices [n0, n1), from the header's
    format text: the rowptr slice ((n+1) * 8) and the block image (16-edge
    blocks from the range's first edge, width bitlen(max - min) each, 8-byte
    header pairs plus one, the streams' words plus two), or the raw columns
    when the image is not smaller."""
    eb, ee = int(rowptr[n0]), int(rowptr[n1])
    m = ee - eb
    img = 8 + 4 * 2
    if m:
        c = col[eb:ee].astype(np.int64)
        starts = np.arange(0, m, 16)
        span = np.maximum.reduceat(c, starts) - np.minimum.reduceat(c, starts)
        width = np.where(span > 0, np.floor(np.log2(np.maximum(span, 1))).astype(np.int64) + 1, 0)
        bits = 16 * int(width.sum())
        img = 8 * (len(starts) + 1) + 4 * ((bits + 31) // 32 + 2)
    return 8 * (n1 - n0 + 1) + min(img, 4 * m)


def test_block_slot_partitions_hand_computed():
    # one row 0..63 and 63 empty rows: vertex 0 alone needs 16 + 80 bytes, all
    # 64 vertices 65 * 8 + 80 = 600; with 599 the first partition stops at 63
    # vertices (64 * 8 + 80 = 592) and vertex 63 (no edges: 16 bytes) is the next
    V = 64
    rowptr = np.array([0, 64] + [64] * (V - 1), np.uint64)
    g = lt.Graph.from_csr(rowptr, np.arange(64, dtype=np.uint32))
    one = lt.Partition.build_blocks(g, 600)
    assert one.count == 1 and one.slot_layout(g, 2) == (768, 1)
    two = lt.Partition.build_blocks(g, 599)
    assert list(two.node_offsets()) == [0, 63, 64] and list(two.edge_offsets()) == [0, 64, 64]
    assert two.bytes == 64 * 8 + 64 * 4  # the larger partition's raw size
    with pytest.raises(lt.LTError) as e:
        lt.Partition.build_blocks(g, 95)
    assert e.value.status == 3  # LT_E_OVERSIZE_VERTEX: vertex 0 needs 96 bytes


@pytest.mark.parametrize("scale,ef,slot", [(12, 16, 16 << 10), (14, 8, 48 << 10), (11, 16, 5000)])
def test_partitions_sized_for_block_slots(scale, ef, slot):
    """Every partition's rowptr slice + block image fits the slot, the next
    vertex would not, the engine's own image sizes agree, and the slot layout
    of the table is no larger than the slot asked for."""
    rowptr, col = oracle.rmat_graph(scale, ef, seed=scale)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build_blocks(g, slot)
    no, eo = part.node_offsets().astype(np.int64), part.edge_offsets()
    assert no[0] == 0 and no[-1] == g.num_node and np.array_equal(eo, rowptr[no])
    raw = 0
    for p in range(part.count):
        need = block_slot_need(rowptr, col, no[p], no[p + 1])
        assert need <= slot, p
        m = int(eo[p + 1] - eo[p])
        b = len(part.pack(g, p, blocks=True))
        assert need == 8 * int(no[p + 1] - no[p] + 1) + (b if b < 4 * m else 4 * m), p
        if p + 1 < part.count:
            assert block_slot_need(rowptr, col, no[p], no[p + 1] + 1) > slot, p
        raw = max(raw, 8 * int(no[p + 1] - no[p] + 1) + 4 * m)
    assert part.bytes == raw
    sb, n = part.slot_layout(g, 2, 0)
    assert sb <= (slot + 255) // 256 * 256 and n == part.count
