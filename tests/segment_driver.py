"""A complete disk-style frontier search assembled from segment steps, with
the disk stores replaced by dicts (Common/DiskBasedOptFrontierSearch.cpp:166-240):
SetInitialNode writes the start state with opBits 0 and buckets its
cross-segment children (:33-51); every layer runs the segment step on every
segment that has a frontier or incoming cross-segment marks (:216-218) and
stops at an empty layer.  Streams written by one layer are decoded with the
oracle codecs before the next layer reads them, so the loop also round-trips
every stream it produces.

`step(segment, idx, ops, xseg_per_op)` returns a dict with count, fr_bytes and
buckets [(op, seg, nvalues, bytes)], as tests/oracle_lib.fs_segment_step does."""
import ctypes

import numpy as np

import oracle_lib as O


def decode_frontier(data):
    """StreamVInt-with-bounds chunks back to (idx, ops)."""
    idx, ops = [], []
    pos = 0
    pad = np.concatenate([data, np.zeros(64, np.uint8)])
    while pos < len(data):
        cnt = int.from_bytes(bytes(data[pos:pos + 4]), "little")
        adj = (cnt + 7) // 8 * 8
        v = np.zeros(adj, np.uint32)
        b = np.zeros(adj, np.uint8)
        size = ctypes.c_int(len(data) - pos)
        n = O.lib().or_svint_decode_with_bounds(ctypes.byref(size), pad[pos:], v, b)
        assert n == cnt
        idx.append(v[:n])
        ops.append(b[:n])
        pos += size.value
    assert pos == len(data)
    if not idx:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint8)
    return np.concatenate(idx), np.concatenate(ops)


def decode_bucket(data, nvalues):
    """FrontierCompression chunks back to the u32 values."""
    out = []
    pos = 0
    pad = np.concatenate([data, np.zeros(64, np.uint8)])
    buf = np.zeros(1 << 18, np.uint32)
    while pos < len(data):
        size = ctypes.c_int(len(data) - pos)
        n = O.lib().or_fc_decode(ctypes.byref(size), pad[pos:], buf)
        out.append(buf[:n].copy())
        pos += size.value
    assert pos == len(data)
    vals = np.concatenate(out) if out else np.zeros(0, np.uint32)
    assert len(vals) == nvalues
    return vals


def frontier_search(p, start, segment_bits, step, max_layers=1000, check_streams=True):
    ops = int(O.lib().or_operators_count(ctypes.byref(p)))
    total = int(O.lib().or_indexes_count(ctypes.byref(p)))
    nseg = max(1, (total + (1 << segment_bits) - 1) >> segment_bits)
    mask = (1 << segment_bits) - 1
    seg0 = start >> segment_bits
    frontier = {seg0: (np.array([start & mask], np.uint32), np.array([0], np.uint8))}
    xseg = {}
    # SetInitialNode: cross-segment children of the start state
    ch = np.zeros(16, np.uint64)
    chop = np.zeros(16, np.int32)
    k = O.lib().or_expand_one(ctypes.byref(p), O.EXPAND_FULL, start, 0, ch, chop)
    for j in range(k):
        s = int(ch[j]) >> segment_bits
        if s != seg0:
            xseg.setdefault(s, [[] for _ in range(ops)])[int(chop[j])].append(int(ch[j]) & mask)
    counts = [1]
    while len(counts) < max_layers:
        new_frontier, new_x = {}, {}
        layer = 0
        for s in sorted(set(frontier) | set(xseg)):
            fi, fo = frontier.get(s, (np.zeros(0, np.uint32), np.zeros(0, np.uint8)))
            xs = xseg.get(s)
            xs_arrays = [np.asarray(x, np.uint32) for x in xs] if xs else None
            r = step(s, fi, fo, xs_arrays)
            layer += r["count"]
            if check_streams:
                di, do = decode_frontier(r["fr_bytes"])
                if "fr_idx" in r:
                    assert np.array_equal(di, r["fr_idx"]) and np.array_equal(do, r["fr_ops"])
            else:
                di, do = r["fr_idx"], r["fr_ops"]
            if len(di):
                new_frontier[s] = (di, do)
            for op, seg, nval, data in r["buckets"]:
                vals = decode_bucket(data, nval)
                new_x.setdefault(seg, [[] for _ in range(ops)])[op].extend(vals.tolist())
        if layer == 0:
            break
        counts.append(layer)
        frontier, xseg = new_frontier, new_x
    return counts


def oracle_step(p, segment_bits, nseg):
    def step(s, fi, fo, xs):
        return O.fs_segment_step(p, segment_bits, s, nseg, fi, fo, xs)
    return step


def decode_fc_stream(data):
    """FrontierCompression chunks (bitmap or StreamVInt) back to their values,
    as CompressedSegmentReader reads a segment (Common/CompressedFrontier.cpp:90-100)."""
    out = []
    pos = 0
    pad = np.concatenate([np.asarray(data, np.uint8), np.zeros(64, np.uint8)])
    buf = np.zeros((1 << 18) + 64, np.uint32)
    while pos < len(data):
        size = ctypes.c_int(len(data) - pos)
        n = O.lib().or_fc_decode(ctypes.byref(size), pad[pos:], buf)
        assert n > 0 and size.value > 0
        out.append(buf[:n].copy())
        pos += size.value
    assert pos == len(data)
    return np.concatenate(out) if out else np.zeros(0, np.uint32)


def three_bit_search(p, start, segment_bits, step, max_layers=1000, check_streams=True):
    """DiskBasedOptThreeBitBFS (Common/DiskBasedOptThreeBitBFS.cpp:164-252) with
    the stores replaced by dicts: SetInitialNode writes the start state and
    buckets its cross-segment children (:34-52); every layer runs the step on
    every segment (:216-218) with its current and old frontier and the marks
    bucketed for it by the previous layer, then rotates old <- cur <- new
    (:182-187).  The new frontier and the buckets are read back from their
    compressed streams, as the next layer's readers would.

    `step(segment, cur, old, xseg_per_op)` returns a dict with count, fr_bytes,
    buckets [(op, seg, nvalues, bytes)] and optionally fr_idx."""
    ops = int(O.lib().or_operators_count(ctypes.byref(p)))
    total = int(O.lib().or_indexes_count(ctypes.byref(p)))
    nseg = max(1, (total + (1 << segment_bits) - 1) >> segment_bits)
    mask = (1 << segment_bits) - 1
    seg0 = start >> segment_bits
    empty = np.zeros(0, np.uint32)
    cur = {seg0: np.array([start & mask], np.uint32)}
    old = {}
    xseg = {}
    ch = np.zeros(16, np.uint64)
    chop = np.zeros(16, np.int32)
    mode = O.lib().or_hint_mode(ctypes.byref(p), segment_bits, 1)
    k = O.lib().or_expand_one(ctypes.byref(p), mode, start, 0, ch, chop)
    for j in range(k):
        s = int(ch[j]) >> segment_bits
        if s != seg0:
            xseg.setdefault(s, [[] for _ in range(ops)])[int(chop[j])].append(int(ch[j]) & mask)
    counts = [1]
    while len(counts) < max_layers:
        new, new_x = {}, {}
        layer = 0
        for s in range(nseg):
            xs = xseg.get(s)
            xs_arrays = [np.asarray(x, np.uint32) for x in xs] if xs else None
            r = step(s, cur.get(s, empty), old.get(s, empty), xs_arrays)
            layer += r["count"]
            if check_streams:
                fr = decode_fc_stream(r["fr_bytes"])
                if "fr_idx" in r:
                    assert np.array_equal(fr, r["fr_idx"])
            else:
                fr = r["fr_idx"]
            assert len(fr) == r["count"]
            if len(fr):
                new[s] = fr
            for op, seg, nval, data in r["buckets"]:
                vals = decode_bucket(data, nval)
                new_x.setdefault(seg, [[] for _ in range(ops)])[op].extend(vals.tolist())
        if layer == 0:
            break
        counts.append(layer)
        old, cur, xseg = cur, new, new_x
    return counts


def oracle_step3(p, segment_bits, nseg):
    def step(s, cur, old, xs):
        return O.bfs3_segment_step(p, segment_bits, s, nseg, cur, old, xs)
    return step
