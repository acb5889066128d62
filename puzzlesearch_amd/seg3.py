"""One disk-segment step of the three-bit BFS on the device
(DB_Opt3BitBFS_Solver::Expand, Common/DiskBasedOptThreeBitBFS.cpp:54-144),
through the C ABI ps_seg3_*.

    with Seg3(puzzle, segment_bits) as w:
        r = w.step(segment, cur, old, xseg=[ids_op0, ids_op1, ...])

`step` takes host arrays, runs the step on the device and returns a dict:
count, fr_idx, fr_bytes (FrontierCompression chunks of 2^18 values), buckets
[(op, segment, nvalues, bytes)] in (op, segment) order with their ids under
"bucket_values", and the ps_seg3_result fields under "stats".  `step_device`
takes device pointers and leaves the outputs on the device."""
import ctypes

import numpy as np

from . import DeviceBuffer, Seg3Result, check, download_ptr, lib


class Seg3:
    def __init__(self, puzzle, segment_bits, stream=None):
        self.puzzle = puzzle
        self.segment_bits = segment_bits
        h = ctypes.c_void_p()
        check(lib().ps_seg3_create(ctypes.byref(puzzle), segment_bits, stream, ctypes.byref(h)))
        self.h = h.value
        nseg = ctypes.c_uint32()
        ops = ctypes.c_int32()
        fp = ctypes.c_uint64()
        check(lib().ps_seg3_info(self.h, ctypes.byref(nseg), ctypes.byref(ops), ctypes.byref(fp)))
        self.nseg, self.ops, self.footprint = nseg.value, ops.value, fp.value

    def step_device(self, segment, d_cur, ncur, d_old, nold, d_xseg=None, xseg_starts=None):
        """Inputs already on the device; returns the ps_seg3_result."""
        res = Seg3Result()
        xs = None
        if xseg_starts is not None:
            xs = np.ascontiguousarray(xseg_starts, dtype=np.uint64)
        check(lib().ps_seg3_step(self.h, segment, d_cur, ncur, d_old, nold, d_xseg,
                                 xs.ctypes.data if xs is not None else None, ctypes.byref(res)))
        return res

    def step(self, segment, cur, old, xseg=None, fetch=True):
        cur = np.ascontiguousarray(cur, dtype=np.uint32)
        old = np.ascontiguousarray(old, dtype=np.uint32)
        d_cur = DeviceBuffer.from_array(cur) if len(cur) else None
        d_old = DeviceBuffer.from_array(old) if len(old) else None
        d_x = None
        starts = None
        if xseg is not None:
            parts = [np.asarray(x, dtype=np.uint32) for x in xseg]
            assert len(parts) == self.ops
            starts = np.zeros(self.ops + 1, np.uint64)
            starts[1:] = np.cumsum([len(x) for x in parts])
            if starts[-1]:
                d_x = DeviceBuffer.from_array(np.concatenate(parts))
        res = self.step_device(segment, d_cur.ptr if d_cur else None, len(cur), d_old.ptr if d_old else None,
                               len(old), d_x.ptr if d_x else None, starts)
        out = {"count": int(res.count), "stats": res.as_dict()}
        if fetch:
            out.update(self.outputs(res))
        return out

    def frontier_device(self):
        """Device pointers of the last step's frontier ids and stream."""
        p_idx, p_bytes = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().ps_seg3_frontier(self.h, ctypes.byref(p_idx), ctypes.byref(p_bytes), None, 0))
        return p_idx.value, p_bytes.value

    def outputs(self, res):
        """The last step's frontier list, frontier stream and buckets, copied to the host."""
        p_idx, p_bytes = ctypes.c_void_p(), ctypes.c_void_p()
        chunks = np.zeros(int(res.frontier_chunks) + 1, np.uint64)
        check(lib().ps_seg3_frontier(self.h, ctypes.byref(p_idx), ctypes.byref(p_bytes), chunks.ctypes.data,
                                     len(chunks)))
        nf = int(res.count)
        out = {"fr_idx": download_ptr(p_idx.value, np.uint32, nf) if nf else np.zeros(0, np.uint32),
               "fr_bytes": download_ptr(p_bytes.value, np.uint8, int(res.frontier_bytes))
               if res.frontier_bytes else np.zeros(0, np.uint8),
               "fr_chunk_offsets": chunks}
        nb = int(res.buckets)
        op = np.zeros(max(nb, 1), np.uint32)
        seg = np.zeros(max(nb, 1), np.uint32)
        vals = np.zeros(max(nb, 1), np.uint64)
        voff = np.zeros(nb + 1, np.uint64)
        boff = np.zeros(nb + 1, np.uint64)
        p_v, p_x = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().ps_seg3_buckets(self.h, op.ctypes.data, seg.ctypes.data, vals.ctypes.data, voff.ctypes.data,
                                    boff.ctypes.data, max(nb, 1), ctypes.byref(p_v), ctypes.byref(p_x)))
        xb = download_ptr(p_x.value, np.uint8, int(boff[-1])) if nb and boff[-1] else np.zeros(0, np.uint8)
        xv = download_ptr(p_v.value, np.uint32, int(voff[-1])) if nb and voff[-1] else np.zeros(0, np.uint32)
        out["buckets"] = [(int(op[b]), int(seg[b]), int(vals[b]), xb[int(boff[b]):int(boff[b + 1])])
                          for b in range(nb)]
        out["bucket_values"] = [xv[int(voff[b]):int(voff[b + 1])] for b in range(nb)]
        return out

    def close(self):
        if self.h:
            check(lib().ps_seg3_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if self.h:
                lib().ps_seg3_destroy(self.h)
        except Exception:
            pass
