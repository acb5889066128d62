"""puzzlesearch_amd -- MI355X (gfx950) state expansion and bit-array frontier
search for lightln2/PuzzleSearch's puzzles.

Python view of the C ABI in include/puzzlesearch_amd.h (ctypes).  Every call
goes to the native HIP library lib/libpuzzlesearch_amd.so; there is no
Python or CPU fallback: if the library is missing, importing its functions
raises NativeLibraryMissing.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libpuzzlesearch_amd.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "libpuzzlesearch_host.so")

# status codes
PS_OK = 0
PS_ERR_INVALID_ARGUMENT = 1
PS_ERR_HIP = 2
PS_ERR_OUT_OF_MEMORY = 3
PS_ERR_NOT_INITIALIZED = 4
PS_ERR_UNSUPPORTED = 5
PS_ERR_CAPACITY = 6
PS_ERR_INDEX_RANGE = 7
PS_ERR_CANCELLED = 8
PS_ERR_INTERNAL = 9

# puzzle kinds / flags / expand modes
PANCAKE, SLIDING, SLIDING_PACKED, HANOI4, HANOI4_OPT = 1, 2, 3, 4, 5
FLAG_SYMMETRY, FLAG_INV_INDEX = 1, 2
EXPAND_FULL, EXPAND_IN_SEGMENT, EXPAND_CROSS_SEGMENT = 0, 1, 2
BFS_BATCHED, BFS_PLAIN = 0, 1
CROSS_BYTES, CROSS_ATOMIC, CROSS_LISTS = 0, 1, 2
DIR_TOP_DOWN, DIR_AUTO, DIR_BOTTOM_UP = 0, 1, 2
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3
BATCH_STREAMS, BATCH_FUSED = 0, 1
INVALID_INDEX = (1 << 64) - 1


class NativeLibraryMissing(RuntimeError):
    pass


class PSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[ps error {code}] {msg}")
        self.code = code


class Puzzle(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("size", ctypes.c_int32),
                ("height", ctypes.c_int32), ("flags", ctypes.c_int32)]

    def __repr__(self):
        names = {1: "pancake", 2: "sliding", 3: "sliding_packed", 4: "hanoi4", 5: "hanoi4_opt"}
        return f"Puzzle({names.get(self.kind, self.kind)}, {self.size}, {self.height}, flags={self.flags})"


class PuzzleInfo(ctypes.Structure):
    _fields_ = [("operators_count", ctypes.c_int32), ("branching_factor", ctypes.c_int32),
                ("has_odd_cycles", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("indexes_count", ctypes.c_uint64)]


class BfsOptions(ctypes.Structure):
    _fields_ = [("bits", ctypes.c_int32), ("segment_bits", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("max_layers", ctypes.c_int32), ("index_count", ctypes.c_uint64),
                ("tile_bits_log2", ctypes.c_int32), ("block_threads", ctypes.c_int32),
                ("grid_blocks", ctypes.c_int32), ("hash", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("graph_layers", ctypes.c_int32), ("cross_marks", ctypes.c_int32), ("skip_visited", ctypes.c_int32),
                ("direction", ctypes.c_int32), ("fused_layers", ctypes.c_int32)]


class BfsStats(ctypes.Structure):
    _fields_ = [("layers", ctypes.c_uint64), ("expanded", ctypes.c_uint64), ("children", ctypes.c_uint64),
                ("algorithmic_bytes", ctypes.c_uint64), ("expand_bytes", ctypes.c_uint64),
                ("scan_words", ctypes.c_uint64), ("expand_launches", ctypes.c_uint64),
                ("finalize_launches", ctypes.c_uint64), ("expand_ms", ctypes.c_double),
                ("finalize_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("pull_words", ctypes.c_uint64), ("pull_probes", ctypes.c_uint64),
                ("apply_ms", ctypes.c_double), ("apply_launches", ctypes.c_uint64),
                ("list_entries", ctypes.c_uint64), ("pull_ms", ctypes.c_double),
                ("pull_launches", ctypes.c_uint64), ("pull_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def pancake(n, inv_index=False):
    return Puzzle(PANCAKE, n, 0, FLAG_INV_INDEX if inv_index else 0)


def sliding(width, height):
    return Puzzle(SLIDING, width, height, 0)


def sliding_packed(width, height):
    return Puzzle(SLIDING_PACKED, width, height, 0)


def hanoi(disks, symmetry=False, opt=False):
    return Puzzle(HANOI4_OPT if opt else HANOI4, disks, 0, FLAG_SYMMETRY if symmetry else 0)


_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_P = ctypes.POINTER(Puzzle)


class FsResult(ctypes.Structure):
    """ps_fs_result"""
    _fields_ = [(n, _u64) for n in ("count", "frontier_count", "frontier_bytes", "frontier_chunks", "children_in",
                                    "children_cross", "buckets", "cross_chunks", "cross_bytes", "scanned_words",
                                    "algorithmic_bytes")] + [("device_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Seg3Result(ctypes.Structure):
    """ps_seg3_result"""
    _fields_ = [(n, _u64) for n in ("count", "frontier_bytes", "frontier_chunks", "children_in", "children_cross",
                                    "buckets", "cross_chunks", "cross_bytes", "scanned_words",
                                    "algorithmic_bytes")] + [("device_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

class HanoiSearchStats(ctypes.Structure):
    """ps_hanoi_search_stats"""
    _fields_ = [(n, _u64) for n in ("layers", "segments_allocated", "expanded", "children", "listed_children",
                                    "algorithmic_bytes", "hbm_bytes")] + \
        [("device_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOptions(ctypes.Structure):
    """ps_batch_options"""
    _fields_ = [("max_jobs", _i32), ("segment_bits", _i32), ("hbm_budget", _u64), ("mode", _i32),
                ("job_grid_blocks", _i32), ("reserved_cus", _i32), ("shortest_first", _i32)]


class BatchJobInfo(ctypes.Structure):
    """ps_batch_job_info"""
    _fields_ = [("state", _i32), ("layers", _i32), ("footprint", _u64), ("steps", _u64), ("expanded", _u64),
                ("children", _u64), ("submit_ns", _u64), ("admit_ns", _u64), ("done_ns", _u64)]


class BatchStats(ctypes.Structure):
    """ps_batch_stats"""
    _fields_ = [(n, _u64) for n in ("steps", "jobs_done", "states_done", "expanded", "hbm_used", "hbm_budget")] + \
        [("step_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = {
    "ps_last_error": (ctypes.c_char_p, []),
    "ps_init": (_i32, [_i32]),
    "ps_shutdown": (_i32, []),
    "ps_device_count": (_i32, [ctypes.POINTER(_i32)]),
    "ps_device_props": (_i32, [ctypes.c_char_p, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                               ctypes.POINTER(_u64)]),
    "ps_puzzle_query": (_i32, [_P, ctypes.POINTER(PuzzleInfo)]),
    "ps_expand_slots": (_i32, [_P, _i32]),
    "ps_stream_create": (_i32, [ctypes.POINTER(_vp)]),
    "ps_stream_destroy": (_i32, [_vp]),
    "ps_stream_sync": (_i32, [_vp]),
    "ps_device_sync": (_i32, []),
    "ps_alloc": (_i32, [ctypes.POINTER(_vp), _u64]),
    "ps_free": (_i32, [_vp]),
    "ps_host_alloc": (_i32, [ctypes.POINTER(_vp), _u64]),
    "ps_host_free": (_i32, [_vp]),
    "ps_memcpy_h2d_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ps_memcpy_d2h_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ps_memcpy_d2d_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ps_fill_async": (_i32, [_vp, ctypes.c_uint8, _u64, _vp]),
    "ps_copy_async": (_i32, [_vp, _vp, _u64, _vp]),
    "ps_expand": (_i32, [_P, _i32, _vp, _vp, _u64, _vp]),
    "ps_pancake_xseg_postprocess": (_i32, [_i32, _i32, _vp, _u64, _vp]),
    "ps_hanoi_expand_in_segment": (_i32, [_i32, ctypes.c_uint32, _vp, _u64, _vp, _i32, _vp]),
    "ps_hanoi_expand": (_i32, [_i32, _vp, _u64, _vp, _vp]),
    "ps_hanoi_search_create": (_i32, [_i32, _i32, _u64, _vp, ctypes.POINTER(_vp)]),
    "ps_hanoi_search_destroy": (_i32, [_vp]),
    "ps_hanoi_search_run": (_i32, [_vp, _u64, _i32, _vp, _vp, _vp, _i32, ctypes.POINTER(_i32)]),
    "ps_hanoi_search_stats_get": (_i32, [_vp, _vp]),
    "ps_bits_andnot_count": (_i32, [_vp, _vp, _vp, _u64, _vp, _vp]),
    "ps_bits_count": (_i32, [_vp, _u64, _vp, _vp]),
    "ps_bits_hash": (_i32, [_vp, _u64, _u64, _vp, _vp]),
    "ps_bits_scan": (_i32, [_vp, _u64, _vp, _u64, _vp, _vp]),
    "ps_bits_scan_noalloc": (_i32, [_vp, _u64, _vp, _u64, _vp, _vp]),
    "ps_bits_set_list": (_i32, [_vp, _u64, _vp, _u64, _vp, _vp]),
    "ps_svint_encode": (_i32, [_vp, _vp, _vp, ctypes.c_uint32, _vp, _u64, _vp, _vp]),
    "ps_svint_decode": (_i32, [_vp, _vp, ctypes.c_uint32, _i32, _vp, _vp, _u64, _vp, _vp]),
    "ps_opbits_encode": (_i32, [_i32, _vp, _u64, _vp, _vp]),
    "ps_opbits_decode": (_i32, [_i32, _vp, _u64, _vp, _vp]),
    "ps_fc_encode": (_i32, [_vp, _vp, ctypes.c_uint32, _i32, _vp, _u64, _vp, _vp]),
    "ps_fc_decode": (_i32, [_vp, _vp, ctypes.c_uint32, _vp, _u64, _vp, _vp]),
    "ps_bucket_by_segment": (_i32, [_vp, _vp, _u64, _i32, ctypes.c_uint32, _i32, _vp, _vp, _vp]),
    "ps_fs_segment_create": (_i32, [_P, _i32, _vp, ctypes.POINTER(_vp)]),
    "ps_fs_segment_destroy": (_i32, [_vp]),
    "ps_fs_segment_info": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_i32), ctypes.POINTER(_u64)]),
    "ps_fs_segment_step": (_i32, [_vp, ctypes.c_uint32, _vp, _vp, _u64, _vp, _vp, ctypes.POINTER(FsResult)]),
    "ps_fs_segment_frontier": (_i32, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp,
                                      _u64]),
    "ps_fs_segment_buckets": (_i32, [_vp, _vp, _vp, _vp, _vp, _u64, ctypes.POINTER(_vp)]),
    "ps_seg3_create": (_i32, [_P, _i32, _vp, ctypes.POINTER(_vp)]),
    "ps_seg3_destroy": (_i32, [_vp]),
    "ps_seg3_info": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_i32), ctypes.POINTER(_u64)]),
    "ps_seg3_step": (_i32, [_vp, ctypes.c_uint32, _vp, _u64, _vp, _u64, _vp, _vp, ctypes.POINTER(Seg3Result)]),
    "ps_seg3_frontier": (_i32, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _u64]),
    "ps_seg3_buckets": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _u64, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "ps_batch_options_default": (None, [ctypes.POINTER(BatchOptions)]),
    "ps_batch_create": (_i32, [ctypes.POINTER(BatchOptions), _vp, ctypes.POINTER(_vp)]),
    "ps_batch_destroy": (_i32, [_vp]),
    "ps_batch_submit": (_i32, [_vp, _P, _u64, _u64, ctypes.POINTER(_u64)]),
    "ps_batch_step": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "ps_batch_finished": (_i32, [_vp, _vp, _i32, ctypes.POINTER(_i32)]),
    "ps_batch_job": (_i32, [_vp, _u64, ctypes.POINTER(BatchJobInfo), _vp, _vp, _i32]),
    "ps_batch_forget": (_i32, [_vp, _u64]),
    "ps_batch_stats_get": (_i32, [_vp, ctypes.POINTER(BatchStats)]),
    "ps_bfs_options_default": (None, [ctypes.POINTER(BfsOptions)]),
    "ps_bfs_create": (_i32, [_P, ctypes.POINTER(BfsOptions), _vp, ctypes.POINTER(_vp)]),
    "ps_bfs_destroy": (_i32, [_vp]),
    "ps_bfs_run": (_i32, [_vp, _u64, _vp, _vp, _i32, ctypes.POINTER(_i32)]),
    "ps_bfs_reset": (_i32, [_vp, _u64]),
    "ps_bfs_step": (_i32, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "ps_bfs_stats_get": (_i32, [_vp, ctypes.POINTER(BfsStats)]),
    "ps_bfs_frontier": (_i32, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_u64)]),
    "ps_bfs_footprint": (_i32, [_P, ctypes.POINTER(BfsOptions), ctypes.POINTER(_u64)]),
}

_lib = None


def lib():
    """The loaded native library (raises NativeLibraryMissing when not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m puzzlesearch_amd.build` "
                "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names declared in include/puzzlesearch_amd.h."""
    import re
    hdr = os.path.join(os.path.dirname(PKG_DIR), "include", "puzzlesearch_amd.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(ps_[a-z0-9_]+)\s*\(", text)))


def check(rc):
    if rc != PS_OK:
        raise PSError(rc, lib().ps_last_error().decode(errors="replace"))
    return rc


_initialized = False


def init(device=0):
    global _initialized
    if not _initialized:
        check(lib().ps_init(device))
        _initialized = True


def device_count():
    n = _i32(0)
    check(lib().ps_device_count(ctypes.byref(n)))
    return n.value


def device_props():
    name = ctypes.create_string_buffer(256)
    cus, lds, hbm = _i32(), _i32(), _u64()
    check(lib().ps_device_props(name, 256, ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(hbm)))
    return {"name": name.value.decode(), "cus": cus.value, "lds_per_block": lds.value, "hbm_bytes": hbm.value}


def puzzle_info(p):
    info = PuzzleInfo()
    check(lib().ps_puzzle_query(ctypes.byref(p), ctypes.byref(info)))
    return info


def expand_slots(p, mode=EXPAND_FULL):
    return lib().ps_expand_slots(ctypes.byref(p), mode)


def sync():
    check(lib().ps_device_sync())


class DeviceBuffer:
    """HBM allocation made by ps_alloc."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        ptr = _vp()
        check(lib().ps_alloc(ctypes.byref(ptr), self.nbytes))
        self.ptr = ptr.value

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(max(arr.nbytes, 8))
        buf.upload(arr)
        return buf

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().ps_memcpy_h2d_async(self.ptr, arr.ctypes.data, arr.nbytes, stream))
        check(lib().ps_stream_sync(stream))

    def download(self, dtype, count, stream=None, offset_bytes=0):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().ps_memcpy_d2h_async(out.ctypes.data, self.ptr + offset_bytes, out.nbytes, stream))
            check(lib().ps_stream_sync(stream))
        return out

    def zero(self, stream=None):
        check(lib().ps_fill_async(self.ptr, 0, self.nbytes, stream))
        check(lib().ps_stream_sync(stream))

    def free(self):
        if self.ptr:
            check(lib().ps_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.ps_free(self.ptr)
        except Exception:
            pass


def download_ptr(ptr, dtype, count):
    out = np.empty(count, dtype=dtype)
    if out.nbytes:
        check(lib().ps_memcpy_d2h_async(out.ctypes.data, ptr, out.nbytes, None))
        check(lib().ps_stream_sync(None))
    return out


def expand(p, packed, mode=EXPAND_FULL):
    """PuzzleGpu::ExpandGpu on the device: packed u64 parents -> u64 slots."""
    packed = np.ascontiguousarray(packed, dtype=np.uint64)
    s = expand_slots(p, mode)
    if s < 0:
        raise PSError(PS_ERR_INVALID_ARGUMENT, "invalid puzzle/mode")
    din = DeviceBuffer.from_array(packed)
    dout = DeviceBuffer(max(8, len(packed) * s * 8))
    check(lib().ps_expand(ctypes.byref(p), mode, din.ptr, dout.ptr, len(packed), None))
    out = dout.download(np.uint64, len(packed) * s)
    din.free()
    dout.free()
    return out


def bfs_options(**kw):
    o = BfsOptions()
    lib().ps_bfs_options_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Bfs:
    """Device-resident BFS (ps_bfs_*)."""

    def __init__(self, p, **opts):
        self.puzzle = p
        self.options = bfs_options(**opts)
        h = _vp()
        check(lib().ps_bfs_create(ctypes.byref(p), ctypes.byref(self.options), None, ctypes.byref(h)))
        self.h = h.value

    def run(self, start, cap=1 << 16):
        counts = np.zeros(cap, dtype=np.uint64)
        hashes = np.zeros(cap, dtype=np.uint64)
        n = _i32(0)
        check(lib().ps_bfs_run(self.h, start, counts.ctypes.data, hashes.ctypes.data, cap, ctypes.byref(n)))
        return [int(x) for x in counts[:n.value]], [int(x) for x in hashes[:n.value]]

    def reset(self, start):
        check(lib().ps_bfs_reset(self.h, start))

    def step(self):
        c, h = _u64(0), _u64(0)
        check(lib().ps_bfs_step(self.h, ctypes.byref(c), ctypes.byref(h)))
        return c.value, h.value

    def stats(self):
        s = BfsStats()
        check(lib().ps_bfs_stats_get(self.h, ctypes.byref(s)))
        return s

    def frontier(self):
        ptr, n = _vp(), _u64()
        check(lib().ps_bfs_frontier(self.h, ctypes.byref(ptr), ctypes.byref(n)))
        return download_ptr(ptr.value, np.uint64, n.value)

    def close(self):
        if self.h:
            check(lib().ps_bfs_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.ps_bfs_destroy(self.h)
        except Exception:
            pass


def bfs_footprint(p, **opts):
    o = bfs_options(**opts)
    b = _u64()
    check(lib().ps_bfs_footprint(ctypes.byref(p), ctypes.byref(o), ctypes.byref(b)))
    return b.value


class Batch:
    """Multi-search engine (ps_batch_*): continuous batching of complete
    three-bit BFS jobs of different puzzles."""

    def __init__(self, max_jobs=64, segment_bits=22, hbm_budget=0, stream=None, mode=0, job_grid_blocks=None,
                 reserved_cus=None, shortest_first=None):
        o = BatchOptions()
        lib().ps_batch_options_default(ctypes.byref(o))
        o.max_jobs, o.segment_bits, o.hbm_budget, o.mode = max_jobs, segment_bits, hbm_budget, mode
        if job_grid_blocks is not None:
            o.job_grid_blocks = job_grid_blocks
        if reserved_cus is not None:
            o.reserved_cus = reserved_cus
        if shortest_first is not None:
            o.shortest_first = shortest_first
        h = _vp()
        check(lib().ps_batch_create(ctypes.byref(o), stream, ctypes.byref(h)))
        self.h = h.value

    def submit(self, p, start=0, index_count=0):
        jid = _u64()
        check(lib().ps_batch_submit(self.h, ctypes.byref(p), start, index_count, ctypes.byref(jid)))
        return jid.value

    def step(self):
        r, f, q = _i32(), _i32(), _i32()
        check(lib().ps_batch_step(self.h, ctypes.byref(r), ctypes.byref(f), ctypes.byref(q)))
        return r.value, f.value, q.value

    def finished(self):
        n = _i32()
        check(lib().ps_batch_finished(self.h, None, 0, ctypes.byref(n)))
        ids = np.zeros(max(1, n.value), dtype=np.uint64)
        check(lib().ps_batch_finished(self.h, ids.ctypes.data, n.value, ctypes.byref(n)))
        return [int(x) for x in ids[:n.value]]

    def job(self, jid, cap=1 << 12):
        info = BatchJobInfo()
        counts = np.zeros(cap, dtype=np.uint64)
        hashes = np.zeros(cap, dtype=np.uint64)
        rc = lib().ps_batch_job(self.h, jid, ctypes.byref(info), counts.ctypes.data, hashes.ctypes.data, cap)
        out = {k: getattr(info, k) for k, _ in info._fields_}
        if rc != PS_OK:
            out["error"] = lib().ps_last_error().decode(errors="replace")
        n = min(info.layers, cap)
        out["counts"] = [int(x) for x in counts[:n]]
        out["hashes"] = [int(x) for x in hashes[:n]]
        return out

    def forget(self, jid):
        check(lib().ps_batch_forget(self.h, jid))

    def stats(self):
        st = BatchStats()
        check(lib().ps_batch_stats_get(self.h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            check(lib().ps_batch_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.ps_batch_destroy(self.h)
        except Exception:
            pass


def hanoi_expand_in_segment(disks, segment, idx, without_smallest=False):
    """GpuExpandInSegment[WithoutSmallest]<disks> on the device: 6 (or 3) u32
    slots per in-segment id of the 2^32-state segment."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    per = 3 if without_smallest else 6
    din = DeviceBuffer.from_array(idx)
    dout = DeviceBuffer(max(8, len(idx) * per * 4))
    check(lib().ps_hanoi_expand_in_segment(disks, int(segment), din.ptr, len(idx), dout.ptr,
                                           1 if without_smallest else 0, None))
    out = dout.download(np.uint32, len(idx) * per)
    din.free()
    dout.free()
    return out


def hanoi_expand(disks, idx):
    """HanoiTowers::Expand on the device: six u64 reduced children per index."""
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    din = DeviceBuffer.from_array(idx)
    dout = DeviceBuffer(max(8, len(idx) * 48))
    check(lib().ps_hanoi_expand(disks, din.ptr, len(idx), dout.ptr, None))
    out = dout.download(np.uint64, len(idx) * 6)
    din.free()
    dout.free()
    return out


class HanoiSearch:
    """HanoiSearch<disks> (ps_hanoi_search_*): the symmetry-reduced three-bit
    BFS of 4-peg Hanoi over segments of 2^segment_bits states held in HBM."""

    def __init__(self, disks, segment_bits=32, hbm_budget=0, stream=None):
        h = _vp()
        check(lib().ps_hanoi_search_create(disks, segment_bits, hbm_budget, stream, ctypes.byref(h)))
        self.h = h.value
        self.disks = disks

    def run(self, start=0, max_layers=4096):
        """(restored counts, reduced counts, reduced-frontier hashes) per layer."""
        restored = np.zeros(max_layers, np.uint64)
        reduced = np.zeros(max_layers, np.uint64)
        hashes = np.zeros(max_layers, np.uint64)
        n = _i32()
        check(lib().ps_hanoi_search_run(self.h, int(start), max_layers, restored.ctypes.data, reduced.ctypes.data,
                                        hashes.ctypes.data, max_layers, ctypes.byref(n)))
        k = n.value
        return ([int(x) for x in restored[:k]], [int(x) for x in reduced[:k]], [int(x) for x in hashes[:k]])

    def stats(self):
        st = HanoiSearchStats()
        check(lib().ps_hanoi_search_stats_get(self.h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            check(lib().ps_hanoi_search_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.ps_hanoi_search_destroy(self.h)
        except Exception:
            pass
