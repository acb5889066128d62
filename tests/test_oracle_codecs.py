"""CPU oracle codecs against the reference's codec tests:
  * StreamVInt tuple lengths (CommonTests/StreamVIntTests.cpp:27-43), the
    120-byte list (:105-128), bounds tuples (:60-64), 65536-value frontier and
    index round trips (:130-205), the uneven 3-value case (:207-234);
  * OpBits sizes and round trips (CommonTests/CompressionOpBitsTests.cpp:23-39);
  * bitmap vs StreamVInt selection and round trips, the large-range and
    out-of-order cases (CommonTests/CompressionTests.cpp:10-113).
Byte layouts are additionally pinned by decoding with an independent Python
decoder written from the format description."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

TUPLES = [((0, 1, 2, 3), 5), ((256, 257, 258, 259), 9), ((0x010000, 0x010001, 0x010002, 0x010003), 13),
          ((0x01000000, 0x01000001, 0x01000002, 0x01000003), 17), ((0xFF, 0xFFFF, 0xFFFFFF, 0xFFFFFFFF), 11),
          ((0xFF00, 0xDEFF01, 0x01020304, 0x5), 11), ((0x1E0F01, 0xFE020304, 0x5, 0x0505), 11),
          ((0x01000000, 0x00, 0x0100, 0x010000), 11), ((0, 0xFF00FF00, 1, 0xEEFFEEFF), 11),
          ((0, 0xFF44, 2, 3), 6), ((0, 1, 0x01DE65, 3), 7), ((0, 1, 2, 0x03344556), 8)]


def py_decode_tuple(buf, pos):
    """Independent decoder: descriptor bits 2j..2j+1 = byte length - 1 of value j."""
    desc = int(buf[pos])
    pos += 1
    out = []
    for j in range(4):
        n = 1 + ((desc >> (2 * j)) & 3)
        out.append(int.from_bytes(bytes(buf[pos:pos + n]), "little"))
        pos += n
    return out, pos


@pytest.mark.parametrize("vals,length", TUPLES)
def test_tuple_lengths(oracle, vals, length):
    v = np.array(vals, dtype=np.uint32)
    buf = np.zeros(17, dtype=np.uint8)
    assert O.lib().or_svint_encode_tuple(v, buf) == length
    got, end = py_decode_tuple(buf, 0)
    assert got == list(vals) and end == length
    out = np.zeros(4, dtype=np.uint32)
    assert O.lib().or_svint_decode_tuple(buf, out) == length
    assert out.tolist() == list(vals)


def test_tuple_list_120(oracle):
    v = np.array([x for t, _ in TUPLES for x in t], dtype=np.uint32)
    buf = np.zeros(17 * len(v), dtype=np.uint8)
    assert O.lib().or_svint_encode_indexes(len(v), v, buf) == 120
    out = np.zeros(len(v), dtype=np.uint32)
    assert O.lib().or_svint_decode_indexes(len(v), buf, out) == 120
    assert (out == v).all()


@pytest.mark.parametrize("b", [(0, 1, 2, 3, 4, 5, 6, 7), (8, 9, 10, 11, 12, 13, 14, 15), (0, 15, 2, 7, 5, 2, 3, 11)])
def test_bounds_tuple(oracle, b):
    src = np.array(b, dtype=np.uint8)
    buf = np.zeros(4, dtype=np.uint8)
    O.lib().or_svint_encode_bounds_tuple(src, buf)
    assert buf.tolist() == [b[j] | (b[j + 4] << 4) for j in range(4)]
    out = np.zeros(8, dtype=np.uint8)
    O.lib().or_svint_decode_bounds_tuple(buf, out)
    assert out.tolist() == list(b)


def svi_encode(values, bounds=None):
    n = len(values)
    adj = (n + 7) // 8 * 8
    v = np.zeros(max(adj, 8), dtype=np.uint32)
    v[:n] = values
    buf = np.zeros(4 + 17 * adj // 4 + adj // 2 + 32, dtype=np.uint8)
    if bounds is None:
        size = O.lib().or_svint_encode(n, v, buf)
    else:
        b = np.zeros(max(adj, 8), dtype=np.uint8)
        b[:n] = bounds
        size = O.lib().or_svint_encode_with_bounds(n, v, b, buf)
    return buf[:size].copy()


def svi_decode(buf, with_bounds=False):
    cnt = int.from_bytes(bytes(buf[:4]), "little")
    adj = (cnt + 7) // 8 * 8
    v = np.zeros(max(adj, 8), dtype=np.uint32)
    size = ctypes.c_int(len(buf))
    b = np.zeros(max(adj, 8), dtype=np.uint8)
    pad = np.concatenate([buf, np.zeros(32, dtype=np.uint8)])
    if with_bounds:
        n = O.lib().or_svint_decode_with_bounds(ctypes.byref(size), pad, v, b)
        return v[:n], b[:n], size.value
    n = O.lib().or_svint_decode(ctypes.byref(size), pad, v)
    return v[:n], None, size.value


def test_frontier_round_trip(oracle):
    # CommonTests/StreamVIntTests.cpp:130-146: i*i with bounds i%16, 65536 values
    i = np.arange(65536, dtype=np.uint64)
    vals = (i * i).astype(np.uint32)
    bnd = (i % 16).astype(np.uint8)
    buf = svi_encode(vals, bnd)
    v, b, size = svi_decode(buf, True)
    assert size == len(buf)
    assert (v == vals).all() and (b == bnd).all()
    # independent parse of the same bytes
    assert int.from_bytes(bytes(buf[:4]), "little") == 65536
    pos, prev, got = 4, 0, []
    for _ in range(65536 // 4):
        t, pos = py_decode_tuple(buf, pos)
        for d in t:
            prev = (prev + d) & 0xFFFFFFFF
            got.append(prev)
    assert got == vals.tolist()
    assert len(buf) - pos == 65536 // 2


def test_uneven_three_values(oracle):
    # CommonTests/StreamVIntTests.cpp:207-234
    buf = svi_encode([7, 17, 53], [13, 14, 9])
    v, b, _ = svi_decode(buf, True)
    assert v.tolist() == [7, 17, 53] and b.tolist() == [13, 14, 9]
    assert int.from_bytes(bytes(buf[:4]), "little") == 3
    assert len(buf) == 4 + 2 * 5 + 4  # count, two tuples (deltas 7,10,36,0 / 0,0,0,0), 8 bounds


@pytest.mark.parametrize("bits,vals,size", [
    (2, [0, 1, 2, 3, 2, 1, 0, 3], 2), (2, [3, 2, 3, 1, 1, 0, 0, 2], 2),
    (2, [0, 1, 2, 3, 2, 1, 0, 3, 3, 2, 3, 1, 1, 0, 0, 2], 4),
    (4, [0, 1, 2, 3, 15, 14, 13, 12], 4), (4, [13, 4, 12, 15, 15, 0, 0, 7], 4),
    (4, [0, 1, 2, 3, 15, 14, 13, 12, 13, 4, 12, 15, 15, 0, 0, 7], 8),
    (8, [0, 1, 2, 13, 25, 74, 143, 241], 8), (8, [255, 254, 127, 128, 19, 0, 0, 248], 8)])
def test_opbits(oracle, bits, vals, size):
    # CommonTests/CompressionOpBitsTests.cpp:23-39
    ops = np.array(vals, dtype=np.uint8)
    buf = np.zeros(len(vals), dtype=np.uint8)
    assert O.lib().or_opbits_encode(bits, len(vals), ops, buf) == size
    out = np.zeros(len(vals), dtype=np.uint8)
    assert O.lib().or_opbits_decode(bits, len(vals), buf, out) == size
    assert (out == ops).all()
    # 4-bit folding is the StreamVInt bounds layout
    if bits == 4:
        assert buf[:4].tolist() == [vals[j] | (vals[j + 4] << 4) for j in range(4)]


def fc_encode(values, check=True):
    v = np.array(values, dtype=np.uint32)
    n = len(v)
    v2 = np.zeros((n + 7) // 8 * 8 + 8, dtype=np.uint32)
    v2[:n] = v
    buf = np.zeros(1 << 20, dtype=np.uint8)
    size = (O.lib().or_fc_encode_with_check if check else O.lib().or_fc_encode)(n, v2, buf)
    return buf[:size].copy()


def fc_decode(buf):
    out = np.zeros(1 << 18, dtype=np.uint32)
    size = ctypes.c_int(len(buf))
    pad = np.concatenate([buf, np.zeros(64, dtype=np.uint8)])
    n = O.lib().or_fc_decode(ctypes.byref(size), pad, out)
    return out[:n], size.value


def test_bitmap_round_trip(oracle):
    # CommonTests/CompressionTests.cpp:10-23
    vals = np.array([7, 8, 32, 39, 177, 999, 1000, 1001, 1002, 1005], dtype=np.uint32)
    buf = np.zeros(4096, dtype=np.uint8)
    size = O.lib().or_fc_encode_bitmap(len(vals), vals, buf)
    out = np.zeros(64, dtype=np.uint32)
    sz = ctypes.c_int(size)
    n = O.lib().or_fc_decode_bitmap(ctypes.byref(sz), buf, out)
    assert n == len(vals) and (out[:n] == vals).all()
    assert int.from_bytes(bytes(buf[:4]), "little") == 0x80000000 | ((1005 - 7 + 63) // 64)


@pytest.mark.parametrize("step,expect_bitmap", [(11, False), (7, True)])
def test_selection(oracle, step, expect_bitmap):
    # CommonTests/CompressionTests.cpp:25-59: 150 values i*11 -> StreamVInt, i*7 -> bitmap
    vals = [i * step for i in range(150)]
    buf = fc_encode(vals, check=False)
    assert bool(O.lib().or_fc_is_bitmap(buf)) == expect_bitmap
    out, size = fc_decode(buf)
    assert out.tolist() == vals and size == len(buf)


def test_with_check_rejects_bitmap_for_duplicate_end(oracle):
    # CommonTests/CompressionTests.cpp:61-79: a trailing 700 (== 100*7) breaks the bitmap check
    vals = [i * 7 for i in range(150)] + [700]
    buf = fc_encode(vals, check=True)
    assert not O.lib().or_fc_is_bitmap(buf)
    out, _ = fc_decode(buf)
    assert out.tolist() == vals


def test_large_range(oracle):
    # CommonTests/CompressionTests.cpp:81-99 uses 64-bit values truncated to u32; any
    # unsorted-at-the-ends chunk must still round-trip through StreamVInt
    rng = np.random.default_rng(5)
    vals = (rng.integers(0, 1 << 34, size=300, dtype=np.uint64) & 0xFFFFFFFF).astype(np.uint32)
    buf = fc_encode(vals, check=False)
    out, _ = fc_decode(buf)
    assert (out == vals).all()


@pytest.mark.parametrize("vals", [[100, 10000, 2500, 2, 20000], [100, 10000, 2500, 2, 100]])
def test_bitmap_check_rejects_unsorted(oracle, vals):
    # CommonTests/CompressionTests.cpp:101-113
    v = np.array(vals, dtype=np.uint32)
    buf = np.zeros(1 << 16, dtype=np.uint8)
    assert O.lib().or_fc_encode_bitmap_with_check(len(v), v, buf) == -1
