"""Device codecs through the C ABI: StreamVInt (Common/StreamVInt.h),
OpBitsCompression (Common/OpBitsCompression.h), FrontierCompression
(Common/FrontierCompression.h) and segment bucketing of cross-segment children
(Common/Multiplexor.cpp:200-212).

Host arrays go in and come out; the work runs on the device.  Chunks follow
the reference's framing: every chunk is encoded on its own (at most 2^18
values, StreamVInt::MAX_INDEXES_COUNT) and the encoded chunks are stored back
to back, with offsets[c] the byte offset of chunk c."""
import numpy as np

from . import DeviceBuffer, check, lib

MAX_CHUNK_VALUES = 1 << 18


def _starts(chunks):
    st = np.zeros(len(chunks) + 1, dtype=np.uint64)
    st[1:] = np.cumsum([len(c) for c in chunks])
    return st


def _svi_cap(n_values, nchunks):
    # per chunk: 4-byte count, up to 17 bytes per 4 values, 1 byte per 2 bounds, padding to 8
    return int(4 * nchunks + 17 * (n_values + 8 * nchunks) // 4 + (n_values + 8 * nchunks) // 2 + 64)


def svint_encode(chunks, bounds=None):
    """StreamVInt::Encode per chunk; returns (bytes u8, offsets u64[nchunks+1])."""
    vals = np.concatenate([np.asarray(c, dtype=np.uint32) for c in chunks]) if chunks else np.zeros(0, np.uint32)
    st = _starts(chunks)
    cap = _svi_cap(len(vals), len(chunks))
    d_v = DeviceBuffer.from_array(vals)
    d_b = DeviceBuffer.from_array(np.concatenate([np.asarray(b, dtype=np.uint8) for b in bounds])) if bounds else None
    d_st = DeviceBuffer.from_array(st)
    d_out = DeviceBuffer(cap)
    d_off = DeviceBuffer(8 * (len(chunks) + 1))
    check(lib().ps_svint_encode(d_v.ptr, d_b.ptr if d_b else None, d_st.ptr, len(chunks), d_out.ptr, cap,
                                d_off.ptr, None))
    off = d_off.download(np.uint64, len(chunks) + 1)
    return d_out.download(np.uint8, int(off[-1])), off


def svint_decode(data, offsets, with_bounds=False, values_cap=None):
    """StreamVInt::Decode per chunk; returns (list of value arrays, list of bound arrays or None)."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = len(offsets) - 1
    cap = values_cap if values_cap is not None else n * MAX_CHUNK_VALUES
    d_in = DeviceBuffer.from_array(np.concatenate([data, np.zeros(64, np.uint8)]))
    d_off = DeviceBuffer.from_array(np.asarray(offsets, dtype=np.uint64))
    d_v = DeviceBuffer(4 * cap + 64)
    d_b = DeviceBuffer(cap + 64) if with_bounds else None
    d_vs = DeviceBuffer(8 * (n + 1))
    check(lib().ps_svint_decode(d_in.ptr, d_off.ptr, n, 1 if with_bounds else 0, d_v.ptr,
                                d_b.ptr if d_b else None, cap, d_vs.ptr, None))
    vs = d_vs.download(np.uint64, n + 1)
    v = d_v.download(np.uint32, int(vs[-1]))
    out = [v[int(vs[c]):int(vs[c + 1])] for c in range(n)]
    if not with_bounds:
        return out, None
    b = d_b.download(np.uint8, int(vs[-1]))
    return out, [b[int(vs[c]):int(vs[c + 1])] for c in range(n)]


def opbits_encode(bits, ops):
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    d_in = DeviceBuffer.from_array(ops)
    d_out = DeviceBuffer(max(8, len(ops) * bits // 8))
    check(lib().ps_opbits_encode(bits, d_in.ptr, len(ops), d_out.ptr, None))
    return d_out.download(np.uint8, len(ops) * bits // 8)


def opbits_decode(bits, data, count):
    d_in = DeviceBuffer.from_array(np.ascontiguousarray(data, dtype=np.uint8))
    d_out = DeviceBuffer(max(8, count))
    check(lib().ps_opbits_decode(bits, d_in.ptr, count, d_out.ptr, None))
    return d_out.download(np.uint8, count)


def fc_encode(chunks, with_check=True):
    """FrontierCompression::EncodeWithCheck (or Encode) per chunk."""
    vals = np.concatenate([np.asarray(c, dtype=np.uint32) for c in chunks]) if chunks else np.zeros(0, np.uint32)
    st = _starts(chunks)
    cap = _svi_cap(len(vals), len(chunks))
    d_v = DeviceBuffer.from_array(vals)
    d_st = DeviceBuffer.from_array(st)
    d_out = DeviceBuffer(cap)
    d_off = DeviceBuffer(8 * (len(chunks) + 1))
    check(lib().ps_fc_encode(d_v.ptr, d_st.ptr, len(chunks), 1 if with_check else 0, d_out.ptr, cap, d_off.ptr,
                             None))
    off = d_off.download(np.uint64, len(chunks) + 1)
    return d_out.download(np.uint8, int(off[-1])), off


def fc_decode(data, offsets, values_cap=None):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = len(offsets) - 1
    cap = values_cap if values_cap is not None else n * MAX_CHUNK_VALUES
    d_in = DeviceBuffer.from_array(np.concatenate([data, np.zeros(64, np.uint8)]))
    d_off = DeviceBuffer.from_array(np.asarray(offsets, dtype=np.uint64))
    d_v = DeviceBuffer(4 * cap + 64)
    d_vs = DeviceBuffer(8 * (n + 1))
    check(lib().ps_fc_decode(d_in.ptr, d_off.ptr, n, d_v.ptr, cap, d_vs.ptr, None))
    vs = d_vs.download(np.uint64, n + 1)
    v = d_v.download(np.uint32, int(vs[-1]))
    return [v[int(vs[c]):int(vs[c + 1])] for c in range(n)]


def bucket_by_segment(children, ops, segment_bits, nseg, nops):
    """Returns (sorted in-segment indexes u32, bucket starts u64[nops*nseg+1]);
    bucket op*nseg+seg holds the sorted indexes of children of that op in that segment."""
    children = np.ascontiguousarray(children, dtype=np.uint64)
    ops = np.ascontiguousarray(ops, dtype=np.uint8)
    n = len(children)
    d_c = DeviceBuffer.from_array(children)
    d_o = DeviceBuffer.from_array(ops)
    d_idx = DeviceBuffer(max(8, 4 * n))
    d_st = DeviceBuffer(8 * (nops * nseg + 1))
    check(lib().ps_bucket_by_segment(d_c.ptr, d_o.ptr, n, segment_bits, nseg, nops, d_idx.ptr, d_st.ptr, None))
    return d_idx.download(np.uint32, n), d_st.download(np.uint64, nops * nseg + 1)
