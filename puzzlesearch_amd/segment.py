"""One disk-segment step of the frontier search on the device
(DB_OptFS_Solver<4>::Expand, Common/DiskBasedOptFrontierSearch.cpp:53-132),
through the C ABI ps_fs_segment_*.

    with FsSegment(puzzle, segment_bits) as w:
        r = w.step(segment, idx, ops, xseg=[ids_op0, ids_op1, ...])

`step` takes host arrays (or DeviceBuffers with a count), runs the step on the
device and returns a dict: count, fr_idx, fr_ops, fr_bytes, buckets
[(op, segment, nvalues, bytes)] in (op, segment) order, and the ps_fs_result
fields under "stats"."""
import ctypes

import numpy as np

from . import DeviceBuffer, FsResult, check, download_ptr, lib


class FsSegment:
    def __init__(self, puzzle, segment_bits, stream=None):
        self.puzzle = puzzle
        self.segment_bits = segment_bits
        h = ctypes.c_void_p()
        check(lib().ps_fs_segment_create(ctypes.byref(puzzle), segment_bits, stream, ctypes.byref(h)))
        self.h = h.value
        nseg = ctypes.c_uint32()
        ops = ctypes.c_int32()
        fp = ctypes.c_uint64()
        check(lib().ps_fs_segment_info(self.h, ctypes.byref(nseg), ctypes.byref(ops), ctypes.byref(fp)))
        self.nseg, self.ops, self.footprint = nseg.value, ops.value, fp.value

    def step_device(self, segment, d_idx, d_ops, n, d_xseg=None, xseg_starts=None):
        """Inputs already on the device; returns the ps_fs_result."""
        res = FsResult()
        xs = None
        if xseg_starts is not None:
            xs = np.ascontiguousarray(xseg_starts, dtype=np.uint64)
        check(lib().ps_fs_segment_step(self.h, segment, d_idx, d_ops, n, d_xseg,
                                       xs.ctypes.data if xs is not None else None, ctypes.byref(res)))
        return res

    def step(self, segment, idx, ops, xseg=None, fetch=True):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        ops = np.ascontiguousarray(ops, dtype=np.uint8)
        n = len(idx)
        d_idx = DeviceBuffer.from_array(idx) if n else None
        d_ops = DeviceBuffer.from_array(ops) if n else None
        d_x = None
        starts = None
        if xseg is not None:
            parts = [np.asarray(x, dtype=np.uint32) for x in xseg]
            assert len(parts) == self.ops
            starts = np.zeros(self.ops + 1, np.uint64)
            starts[1:] = np.cumsum([len(x) for x in parts])
            if starts[-1]:
                d_x = DeviceBuffer.from_array(np.concatenate(parts))
        res = self.step_device(segment, d_idx.ptr if d_idx else None, d_ops.ptr if d_ops else None, n,
                               d_x.ptr if d_x else None, starts)
        out = {"count": int(res.count), "stats": res.as_dict()}
        if fetch:
            out.update(self.outputs(res))
        return out

    def outputs(self, res):
        """The last step's frontier list, frontier stream and buckets, copied to the host."""
        p_idx, p_ops, p_bytes = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        chunks = np.zeros(int(res.frontier_chunks) + 1, np.uint64)
        check(lib().ps_fs_segment_frontier(self.h, ctypes.byref(p_idx), ctypes.byref(p_ops), ctypes.byref(p_bytes),
                                           chunks.ctypes.data, len(chunks)))
        nf = int(res.frontier_count)
        out = {"fr_idx": download_ptr(p_idx.value, np.uint32, nf) if nf else np.zeros(0, np.uint32),
               "fr_ops": download_ptr(p_ops.value, np.uint8, nf) if nf else np.zeros(0, np.uint8),
               "fr_bytes": download_ptr(p_bytes.value, np.uint8, int(res.frontier_bytes))
               if res.frontier_bytes else np.zeros(0, np.uint8),
               "fr_chunk_offsets": chunks}
        nb = int(res.buckets)
        op = np.zeros(max(nb, 1), np.uint32)
        seg = np.zeros(max(nb, 1), np.uint32)
        vals = np.zeros(max(nb, 1), np.uint64)
        offs = np.zeros(nb + 1, np.uint64)
        p_x = ctypes.c_void_p()
        check(lib().ps_fs_segment_buckets(self.h, op.ctypes.data, seg.ctypes.data, vals.ctypes.data,
                                          offs.ctypes.data, max(nb, 1), ctypes.byref(p_x)))
        xb = download_ptr(p_x.value, np.uint8, int(offs[-1])) if nb and offs[-1] else np.zeros(0, np.uint8)
        out["buckets"] = [(int(op[b]), int(seg[b]), int(vals[b]), xb[int(offs[b]):int(offs[b + 1])])
                          for b in range(nb)]
        return out

    def close(self):
        if self.h:
            check(lib().ps_fs_segment_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if self.h:
                lib().ps_fs_segment_destroy(self.h)
        except Exception:
            pass
