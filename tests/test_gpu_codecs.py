"""Device codecs against the CPU oracle codecs, byte for byte: StreamVInt with
and without bounds (Common/StreamVInt.cpp:209-300), OpBits<2/4/8>
(Common/OpBitsCompression.cpp:10-76), FrontierCompression Encode /
EncodeWithCheck / Decode (Common/FrontierCompression.cpp:92-125), and bucketing
of cross-segment children by (op, segment) (Common/Multiplexor.cpp:200-212)
against a numpy sort.  Edge cases: empty chunks, counts not a multiple of 8,
the 2^18-value maximum chunk, unsorted chunks, the reference's bitmap
selection cases."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C(ps):
    from puzzlesearch_amd import codecs
    return codecs


def oracle_svi(chunk, bounds=None):
    n = len(chunk)
    if n == 0:
        return np.zeros(0, np.uint8)
    adj = (n + 7) // 8 * 8
    v = np.zeros(adj, np.uint32)
    v[:n] = chunk
    buf = np.zeros(4 + 17 * adj // 4 + adj // 2 + 32, np.uint8)
    if bounds is None:
        size = O.lib().or_svint_encode(n, v, buf)
    else:
        b = np.zeros(adj, np.uint8)
        b[:n] = bounds
        size = O.lib().or_svint_encode_with_bounds(n, v, b, buf)
    return buf[:size]


def oracle_fc(chunk, check):
    n = len(chunk)
    if n == 0:
        return np.zeros(0, np.uint8)
    v = np.zeros((n + 7) // 8 * 8 + 8, np.uint32)
    v[:n] = chunk
    buf = np.zeros(4 + 17 * len(v) // 4 + 64 + ((int(max(chunk)) - int(min(chunk))) // 8 + 64), np.uint8)
    size = (O.lib().or_fc_encode_with_check if check else O.lib().or_fc_encode)(n, v, buf)
    return buf[:size]


def sorted_unique(rng, n, hi):
    return np.sort(rng.choice(hi, size=n, replace=False)).astype(np.uint32)


def svi_chunks(rng):
    sizes = [0, 1, 3, 7, 8, 9, 17, 1000, 0, 4096, 1 << 18, 5]
    chunks = []
    for k, n in enumerate(sizes):
        if k % 3 == 0:
            chunks.append(sorted_unique(rng, n, 1 << 32) if n else np.zeros(0, np.uint32))
        elif k % 3 == 1:
            chunks.append(rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32))  # unsorted: wrapping deltas
        else:
            chunks.append(sorted_unique(rng, n, max(4 * n, 16)))  # small deltas
    return chunks


def test_svint_matches_oracle(C):
    rng = np.random.default_rng(11)
    chunks = svi_chunks(rng)
    data, off = C.svint_encode(chunks)
    expect = np.concatenate([oracle_svi(c) for c in chunks])
    assert len(data) == len(expect)
    assert np.array_equal(data, expect)
    for c, ch in enumerate(chunks):
        assert off[c + 1] - off[c] == len(oracle_svi(ch))
    back, _ = C.svint_decode(data, off)
    for a, b in zip(back, chunks):
        assert np.array_equal(a, b)


def test_svint_bounds_matches_oracle(C):
    rng = np.random.default_rng(12)
    chunks = svi_chunks(rng)
    bounds = [rng.integers(0, 16, size=len(c), dtype=np.uint8) for c in chunks]
    data, off = C.svint_encode(chunks, bounds)
    expect = np.concatenate([oracle_svi(c, b) for c, b in zip(chunks, bounds)])
    assert np.array_equal(data, expect)
    back, bb = C.svint_decode(data, off, with_bounds=True)
    for a, b, x, y in zip(back, chunks, bb, bounds):
        assert np.array_equal(a, b)
        assert np.array_equal(x, y)


def test_svint_reference_frontier_case(C):
    # CommonTests/StreamVIntTests.cpp:130-146
    i = np.arange(65536, dtype=np.uint64)
    vals = (i * i).astype(np.uint32)
    bnd = (i % 16).astype(np.uint8)
    data, off = C.svint_encode([vals], [bnd])
    assert np.array_equal(data, oracle_svi(vals, bnd))
    back, bb = C.svint_decode(data, off, with_bounds=True)
    assert np.array_equal(back[0], vals) and np.array_equal(bb[0], bnd)


@pytest.mark.parametrize("kind", ["uniform5", "uniform9", "uniform17", "mixed_multiple", "mixed_multiple_big"])
def test_svint_uniform_and_mixed_tuple_lengths(C, kind):
    # Chunks whose tuples all have one length are decoded by direct indexing;
    # a payload whose length is a multiple of the tuple count but whose tuples
    # differ in length (half of 5 bytes, half of 13: 9 on average) must take
    # the descriptor walk.  Both against the oracle's bytes and values.
    rng = np.random.default_rng(sum(map(ord, kind)))
    n = 70000 if kind.endswith("big") else 4096
    if kind == "uniform5":
        gaps = rng.integers(1, 256, size=n)
    elif kind == "uniform9":
        gaps = rng.integers(256, 65536, size=n)
    elif kind == "uniform17":
        gaps = rng.integers(1 << 24, 1 << 25, size=n)
    else:  # tuples alternate between four 1-byte and four 3-byte deltas
        gaps = np.where((np.arange(n) // 4) % 2 == 0, rng.integers(1, 256, size=n),
                        rng.integers(1 << 16, 1 << 17, size=n))
    vals = (np.cumsum(gaps.astype(np.uint64)) % (1 << 32)).astype(np.uint32)  # wraps: deltas keep their size
    bounds = rng.integers(0, 16, size=n, dtype=np.uint8)
    data, off = C.svint_encode([vals], [bounds])
    assert np.array_equal(data, oracle_svi(vals, bounds))
    T = (n + 7) // 8 * 2
    payload = len(data) - 4 - 2 * T
    if kind.startswith("mixed"):
        assert payload % T == 0 and payload // T == 9  # a multiple of T, tuples of two lengths
    else:
        assert payload == {"uniform5": 5, "uniform9": 9, "uniform17": 17}[kind] * T
    back, bb = C.svint_decode(data, off, with_bounds=True)
    assert np.array_equal(back[0], vals) and np.array_equal(bb[0], bounds)


@pytest.mark.parametrize("cut", [1, 5, 16, "all_but_header"])
def test_svint_decode_rejects_truncated_chunk(C, ps, cut):
    # a chunk whose bytes end before its last tuple (or inside it, or right
    # after its count) is refused (the call fails); the chunk after it is
    # intact
    rng = np.random.default_rng(13)
    chunks = [sorted_unique(rng, 5000, 1 << 32), sorted_unique(rng, 3000, 1 << 32)]
    data, off = C.svint_encode(chunks)
    if cut == "all_but_header":
        cut = int(off[1]) - 4
    bad = np.concatenate([data[:off[1] - cut], data[off[1]:]])
    boff = np.array([0, off[1] - cut, off[2] - cut], np.uint64)
    with pytest.raises(ps.PSError, match="tuples end"):
        C.svint_decode(bad, boff)
    back, _ = C.svint_decode(data, off)  # the valid stream still decodes
    assert np.array_equal(back[0], chunks[0]) and np.array_equal(back[1], chunks[1])


def test_svint_rejects_oversized_chunk(C, ps):
    with pytest.raises(ps.PSError):
        C.svint_encode([np.arange((1 << 18) + 1, dtype=np.uint32)])


def test_svint_encode_capacity_is_a_bound(C, ps):
    # an output capacity below the encoded size is refused, and no byte at or
    # past the capacity is written: the chunks that would end past it are
    # skipped on the device
    rng = np.random.default_rng(21)
    chunks = [sorted_unique(rng, 3000, 1 << 32), sorted_unique(rng, 5000, 1 << 32), sorted_unique(rng, 2000, 1 << 20)]
    data, off = C.svint_encode(chunks)
    cap = int(off[1]) + 10  # the first chunk fits, the second does not
    vals = np.concatenate(chunks)
    st = np.array([0, 3000, 8000, 10000], np.uint64)
    d_v = ps.DeviceBuffer.from_array(vals)
    d_st = ps.DeviceBuffer.from_array(st)
    canary = np.full(len(data) + 64, 0xA5, np.uint8)
    d_out = ps.DeviceBuffer.from_array(canary)
    d_off = ps.DeviceBuffer(8 * 4)
    with pytest.raises(ps.PSError, match="capacity") as e:
        ps.check(ps.lib().ps_svint_encode(d_v.ptr, None, d_st.ptr, 3, d_out.ptr, cap, d_off.ptr, None))
    assert e.value.code == ps.PS_ERR_CAPACITY
    out = d_out.download(np.uint8, len(canary))
    assert np.array_equal(out[:int(off[1])], data[:int(off[1])])
    assert np.all(out[cap:] == 0xA5)


def test_fc_encode_capacity_is_a_bound(C, ps):
    # the same bound for FrontierCompression, whose chunks here are a bitmap
    # built in LDS, a StreamVInt chunk and a bitmap wider than LDS
    rng = np.random.default_rng(23)
    chunks = [sorted_unique(rng, 4096, 16384), sorted_unique(rng, 3000, 1 << 31), sorted_unique(rng, 1 << 17, 1 << 18)]
    data, off = C.fc_encode(chunks)
    for cut in (1, 2):  # chunks [0, cut) fit
        cap = int(off[cut]) + 5
        d_v = ps.DeviceBuffer.from_array(np.concatenate(chunks))
        d_st = ps.DeviceBuffer.from_array(np.cumsum([0] + [len(c) for c in chunks]).astype(np.uint64))
        canary = np.full(len(data) + 64, 0x5A, np.uint8)
        d_out = ps.DeviceBuffer.from_array(canary)
        d_off = ps.DeviceBuffer(8 * 4)
        with pytest.raises(ps.PSError, match="capacity") as e:
            ps.check(ps.lib().ps_fc_encode(d_v.ptr, d_st.ptr, 3, 1, d_out.ptr, cap, d_off.ptr, None))
        assert e.value.code == ps.PS_ERR_CAPACITY
        out = d_out.download(np.uint8, len(canary))
        assert np.array_equal(out[:int(off[cut])], data[:int(off[cut])])
        assert np.all(out[cap:] == 0x5A)


def test_svint_decode_capacity_is_a_bound(C, ps):
    # values_cap below the decoded count is refused, no value at or past
    # values_cap is written, and the chunks that fit still decode
    rng = np.random.default_rng(22)
    chunks = [sorted_unique(rng, 3000, 1 << 32), sorted_unique(rng, 5000, 1 << 32)]
    data, off = C.svint_encode(chunks)
    d_in = ps.DeviceBuffer.from_array(np.concatenate([data, np.zeros(64, np.uint8)]))
    d_off = ps.DeviceBuffer.from_array(off)
    cap = 3000 + 4999
    d_v = ps.DeviceBuffer.from_array(np.full(8064, 0xDEADBEEF, np.uint32))
    d_vs = ps.DeviceBuffer(8 * 3)
    with pytest.raises(ps.PSError, match="capacity") as e:
        ps.check(ps.lib().ps_svint_decode(d_in.ptr, d_off.ptr, 2, 0, d_v.ptr, None, cap, d_vs.ptr, None))
    assert e.value.code == ps.PS_ERR_CAPACITY
    v = d_v.download(np.uint32, 8064)
    assert np.array_equal(v[:3000], chunks[0])
    assert np.all(v[3000:] == 0xDEADBEEF)


@pytest.mark.parametrize("bits", [2, 4, 8])
def test_opbits_matches_oracle(C, bits):
    rng = np.random.default_rng(bits)
    for n in (8, 64, 8 * 100003):
        ops = rng.integers(0, 1 << bits, size=n, dtype=np.uint8)
        enc = C.opbits_encode(bits, ops)
        ref = np.zeros(n * bits // 8, np.uint8)
        assert O.lib().or_opbits_encode(bits, n, ops, ref) == len(ref)
        assert np.array_equal(enc, ref)
        assert np.array_equal(C.opbits_decode(bits, enc, n), ops)


def fc_chunks(rng):
    chunks = [np.zeros(0, np.uint32),
              np.arange(150, dtype=np.uint32) * 11,                    # StreamVInt (CompressionTests.cpp:25)
              np.arange(150, dtype=np.uint32) * 7,                     # bitmap (:43)
              np.array([i * 7 for i in range(150)] + [700], np.uint32),  # check rejects the bitmap (:61)
              np.arange(100, dtype=np.uint32) * 2,                     # <= 128 values: StreamVInt
              sorted_unique(rng, 4096, 16384) + np.uint32(1 << 30),    # dense: bitmap
              sorted_unique(rng, 4096, 1 << 31),                       # sparse: StreamVInt
              sorted_unique(rng, 1 << 18, 1 << 19),                    # dense maximum chunk
              sorted_unique(rng, 129, 140),                            # smallest count with a bitmap
              ]
    return chunks


@pytest.mark.parametrize("check", [1, 0])
def test_fc_matches_oracle(C, check):
    rng = np.random.default_rng(21 + check)
    chunks = fc_chunks(rng)
    if not check:
        chunks = [c for k, c in enumerate(chunks) if k != 3]  # Encode without the check needs sorted chunks
    data, off = C.fc_encode(chunks, with_check=bool(check))
    expect = [oracle_fc(c, check) for c in chunks]
    for c in range(len(chunks)):
        got = data[int(off[c]):int(off[c + 1])]
        assert np.array_equal(got, expect[c]), f"chunk {c}"
    back = C.fc_decode(data, off)
    for a, b in zip(back, chunks):
        assert np.array_equal(a, b)
    # the oracle decodes the device bytes too
    for c in range(len(chunks)):
        if len(chunks[c]) == 0:
            continue
        seg = np.concatenate([data[int(off[c]):int(off[c + 1])], np.zeros(64, np.uint8)])
        out = np.zeros(len(chunks[c]) + 64, np.uint32)
        sz = ctypes.c_int(int(off[c + 1] - off[c]))
        n = O.lib().or_fc_decode(ctypes.byref(sz), seg, out)
        assert n == len(chunks[c]) and np.array_equal(out[:n], chunks[c])


def test_fc_bitmap_selected_where_reference_selects_it(C):
    data, off = C.fc_encode([np.arange(150, dtype=np.uint32) * 7, np.arange(150, dtype=np.uint32) * 11])
    assert data[int(off[0]) + 3] & 0x80
    assert not data[int(off[1]) + 3] & 0x80


@pytest.mark.parametrize("seg_bits,nseg,nops,n", [
    (20, 37, 4, 300000),      # few buckets, three 8-bit value passes
    (32, 5, 2, 100000),       # full 32-bit values: four value passes
    (7, 3000, 4, 400000),     # 12000 buckets: two grouping passes of at most 8192 keys
    (12, 20000, 15, 2000),    # more buckets than entries: most are empty
    (16, 1, 1, 1),            # one entry
    (16, 4, 2, 0),            # no entries
    (30, 2, 3, 3000000),      # several chunk sizes of the counting passes
])
def test_bucket_by_segment(C, seg_bits, nseg, nops, n):
    rng = np.random.default_rng(31 + n)
    children = rng.integers(0, nseg << seg_bits, size=n, dtype=np.uint64)
    ops = rng.integers(0, nops, size=n, dtype=np.uint8)
    idx, starts = C.bucket_by_segment(children, ops, seg_bits, nseg, nops)
    key = ops.astype(np.uint64) * np.uint64(nseg) + (children >> np.uint64(seg_bits))
    order = np.lexsort((children, key))
    expect_idx = (children[order] & np.uint64((1 << seg_bits) - 1)).astype(np.uint32)
    counts = np.bincount(key.astype(np.int64), minlength=nops * nseg)
    assert np.array_equal(np.diff(starts.astype(np.int64)), counts)
    assert np.array_equal(idx, expect_idx)


def test_bucket_rejects_out_of_range(C, ps):
    children = np.array([1, 2, 5 << 10], np.uint64)
    ops = np.array([0, 1, 0], np.uint8)
    with pytest.raises(ps.PSError):
        C.bucket_by_segment(children, ops, 10, 4, 2)
