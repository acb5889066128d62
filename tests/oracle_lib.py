"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

PANCAKE, SLIDING, SLIDING_PACKED, HANOI4, HANOI4_OPT = 1, 2, 3, 4, 5
FLAG_SYMMETRY, FLAG_INV_INDEX = 1, 2
EXPAND_FULL, EXPAND_IN_SEGMENT, EXPAND_CROSS_SEGMENT = 0, 1, 2
INVALID = (1 << 64) - 1


class Puzzle(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("size", ctypes.c_int32),
                ("height", ctypes.c_int32), ("flags", ctypes.c_int32)]


def puzzle(kind, size, height=0, flags=0):
    return Puzzle(kind, size, height, flags)


def pancake(n, inv_index=False):
    return Puzzle(PANCAKE, n, 0, FLAG_INV_INDEX if inv_index else 0)


def sliding(w, h):
    return Puzzle(SLIDING, w, h, 0)


def sliding_packed(w, h):
    return Puzzle(SLIDING_PACKED, w, h, 0)


def hanoi(disks, symmetry=False, opt=False):
    return Puzzle(HANOI4_OPT if opt else HANOI4, disks, 0, FLAG_SYMMETRY if symmetry else 0)


def from_spec(spec):
    """Puzzle from a reference_vectors.json 'puzzle' entry."""
    k = spec["kind"]
    if k == "pancake":
        return pancake(spec["n"], spec.get("inv_index", False))
    if k == "sliding":
        return sliding(spec["width"], spec["height"])
    if k == "sliding_packed":
        return sliding_packed(spec["width"], spec["height"])
    if k == "hanoi":
        return hanoi(spec["disks"], spec.get("symmetry", False), spec.get("opt", False))
    raise ValueError(k)


_lib = None

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_P = ctypes.POINTER(Puzzle)
_i = ctypes.c_int
_u64 = ctypes.c_uint64

class FsIo(ctypes.Structure):
    """or_fs_io (oracle.h)."""
    _fields_ = [("fr_idx", ctypes.c_void_p), ("fr_ops", ctypes.c_void_p), ("fr_cap", _u64), ("fr_n", _u64),
                ("fr_bytes", ctypes.c_void_p), ("fr_bytes_cap", _u64), ("fr_nbytes", _u64),
                ("bk_op", ctypes.c_void_p), ("bk_seg", ctypes.c_void_p), ("bk_values", ctypes.c_void_p),
                ("bk_nbytes", ctypes.c_void_p), ("bk_cap", _u64), ("bk_n", _u64),
                ("xs_bytes", ctypes.c_void_p), ("xs_cap", _u64), ("xs_nbytes", _u64),
                ("children_in", _u64), ("children_cross", _u64), ("count", _u64)]


class Seg3Io(ctypes.Structure):
    """or_seg3_io (oracle.h)."""
    _fields_ = [("fr_idx", ctypes.c_void_p), ("fr_cap", _u64), ("fr_n", _u64),
                ("fr_bytes", ctypes.c_void_p), ("fr_bytes_cap", _u64), ("fr_nbytes", _u64),
                ("bk_op", ctypes.c_void_p), ("bk_seg", ctypes.c_void_p), ("bk_values", ctypes.c_void_p),
                ("bk_nbytes", ctypes.c_void_p), ("bk_cap", _u64), ("bk_n", _u64),
                ("xs_bytes", ctypes.c_void_p), ("xs_cap", _u64), ("xs_nbytes", _u64),
                ("children_in", _u64), ("children_cross", _u64), ("count", _u64)]


_SIGS = {
    "or_fs_segment_step": (ctypes.c_int64, [_P, _i, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u8p, _u64,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(FsIo)]),
    "or_bfs3_segment_step": (ctypes.c_int64, [_P, _i, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u64, _u32p, _u64,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Seg3Io)]),
    "or_hint_mode": (_i, [_P, _i, _i]),
    "or_splitmix64": (_u64, [_u64]),
    "or_bits_hash": (_u64, [_u64p, _u64]),
    "or_children_hash": (_u64, [_u64p, _i32p, _u64]),
    "or_perm_compact": (None, [_i32p, _i]),
    "or_perm_uncompact": (None, [_i32p, _i]),
    "or_perm_rank": (_u64, [_i32p, _i]),
    "or_perm_unrank": (None, [_u64, _i32p, _i]),
    "or_pancake_rank": (_u64, [_i32p, _i]),
    "or_pancake_unrank": (None, [_u64, _i32p, _i]),
    "or_packed_from_standard": (ctypes.c_int64, [_i, _i, _i32p]),
    "or_packed_to_standard": (None, [_i, _i, _u64, _i32p]),
    "or_operators_count": (_i, [_P]),
    "or_branching_factor": (_i, [_P]),
    "or_indexes_count": (_u64, [_P]),
    "or_has_odd_cycles": (_i, [_P]),
    "or_parse": (_u64, [_P, ctypes.c_char_p]),
    "or_to_string": (_i, [_P, _u64, ctypes.c_char_p, _i]),
    "or_expand_slots": (_i, [_P, _i, _u64p, _u64, _u64p]),
    "or_slots_per_parent": (_i, [_P, _i]),
    "or_expand_one": (_i, [_P, _i, _u64, _i, _u64p, _i32p]),
    "or_hanoi_expand_cpu_opt": (_i, [_i, _i, _u64, _i, _u64p, _i32p]),
    "or_combine_index_opbits": (None, [_i, _u64p, _i32p, _u64, _u64p]),
    "or_split_index_ops": (None, [_i, _u64p, _u64, _u64p, _i32p]),
    "or_pancake_xseg_postprocess": (None, [_i, _i, _u32p, _u64]),
    "or_bits_andnot_count": (_u64, [_u64p, _u64p, _u64]),
    "or_bits_count": (_u64, [_u64p, _u64]),
    "or_bfs3": (_i, [_P, _u64, _u64, _i, _u64p, _u64p]),
    "or_bfs2": (_i, [_P, _u64, _u64, _i, _u64p, _u64p]),
    "or_bfs3_segmented": (_i, [_P, _u64, _i, _i, _u64p]),
    "or_bfs3_expand_range": (_u64, [_P, _u64p, _u64p, _u64, _u64, ctypes.POINTER(_u64)]),
    "or_svint_encode_tuple": (_i, [_u32p, _u8p]),
    "or_svint_decode_tuple": (_i, [_u8p, _u32p]),
    "or_svint_encode_bounds_tuple": (None, [_u8p, _u8p]),
    "or_svint_decode_bounds_tuple": (None, [_u8p, _u8p]),
    "or_svint_encode_indexes": (_i, [_i, _u32p, _u8p]),
    "or_svint_decode_indexes": (_i, [_i, _u8p, _u32p]),
    "or_svint_decode_indexes_diff": (_i, [_i, _u8p, _u32p]),
    "or_svint_encode": (_i, [_i, _u32p, _u8p]),
    "or_svint_decode": (_i, [ctypes.POINTER(_i), _u8p, _u32p]),
    "or_svint_encode_with_bounds": (_i, [_i, _u32p, _u8p, _u8p]),
    "or_svint_decode_with_bounds": (_i, [ctypes.POINTER(_i), _u8p, _u32p, _u8p]),
    "or_opbits_encode": (ctypes.c_int64, [_i, _u64, _u8p, _u8p]),
    "or_opbits_decode": (ctypes.c_int64, [_i, _u64, _u8p, _u8p]),
    "or_fc_is_bitmap": (_i, [_u8p]),
    "or_fc_bitmap_size": (ctypes.c_int64, [_i, _u32p]),
    "or_fc_encode_bitmap": (_i, [_i, _u32p, _u8p]),
    "or_fc_encode_bitmap_with_check": (_i, [_i, _u32p, _u8p]),
    "or_fc_decode_bitmap": (_i, [ctypes.POINTER(_i), _u8p, _u32p]),
    "or_fc_encode": (_i, [_i, _u32p, _u8p]),
    "or_fc_encode_with_check": (_i, [_i, _u32p, _u8p]),
    "or_fc_decode": (_i, [ctypes.POINTER(_i), _u8p, _u32p]),
    "or_hv_pegs_mult": (_i, [_u64]),
    "or_hv_in_segment_only": (_i, [_u64, _i]),
    "or_hv_expand": (None, [_i, _u64, _u64p]),
    "or_hv_expand_cross_segment": (_i, [_i, _i, _u64, ctypes.c_uint32, _u64p]),
    "or_hv_expand_in_segment_slots": (None, [_i, ctypes.c_uint32, _u32p, _u64, _i, _u32p]),
    "or_hv_search": (_i, [_i, _u64, _i, _i, _u64p, _u64p, ctypes.c_void_p]),
}


def build():
    """Compile oracle/liboracle.so (and oracle/_ref where the reference exists)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ---------------------------------------------------------------- helpers

def parse(p, text):
    return lib().or_parse(ctypes.byref(p), text.encode())


def to_string(p, index):
    buf = ctypes.create_string_buffer(512)
    n = lib().or_to_string(ctypes.byref(p), index, buf, 512)
    assert n >= 0
    return buf.value.decode()


def ops_count(p):
    return lib().or_operators_count(ctypes.byref(p))


def indexes_count(p):
    return lib().or_indexes_count(ctypes.byref(p))


def slots_per_parent(p, mode=EXPAND_FULL):
    return lib().or_slots_per_parent(ctypes.byref(p), mode)


def expand_slots(p, packed, mode=EXPAND_FULL):
    packed = np.ascontiguousarray(packed, dtype=np.uint64)
    s = slots_per_parent(p, mode)
    out = np.zeros(len(packed) * max(s, 0), dtype=np.uint64)
    r = lib().or_expand_slots(ctypes.byref(p), mode, packed, len(packed), out)
    assert r == s
    return out


def expand_one(p, index, op_bits=0, mode=EXPAND_FULL):
    ch = np.zeros(16, dtype=np.uint64)
    ops = np.zeros(16, dtype=np.int32)
    n = lib().or_expand_one(ctypes.byref(p), mode, index, op_bits, ch, ops)
    assert n >= 0
    return [int(x) for x in ch[:n]], [int(x) for x in ops[:n]]


def bfs3(p, start, index_count=0, max_layers=4096, with_hashes=True):
    counts = np.zeros(max_layers, dtype=np.uint64)
    hashes = np.zeros(max_layers, dtype=np.uint64)
    n = lib().or_bfs3(ctypes.byref(p), start, index_count, max_layers, counts, hashes)
    assert n > 0, "oracle bfs3 failed"
    return [int(x) for x in counts[:n]], [int(x) for x in hashes[:n]]


def bfs2(p, start, index_count=0, max_layers=4096):
    counts = np.zeros(max_layers, dtype=np.uint64)
    hashes = np.zeros(max_layers, dtype=np.uint64)
    n = lib().or_bfs2(ctypes.byref(p), start, index_count, max_layers, counts, hashes)
    assert n > 0, "oracle bfs2 failed"
    return [int(x) for x in counts[:n]], [int(x) for x in hashes[:n]]


def bfs3_segmented(p, start, segment_bits, max_layers=4096):
    counts = np.zeros(max_layers, dtype=np.uint64)
    n = lib().or_bfs3_segmented(ctypes.byref(p), start, segment_bits, max_layers, counts)
    assert n > 0, "oracle segmented bfs failed"
    return [int(x) for x in counts[:n]]


def bits_hash(words):
    words = np.ascontiguousarray(words, dtype=np.uint64)
    return int(lib().or_bits_hash(words, len(words)))


def children_hash(children, ops):
    children = np.ascontiguousarray(children, dtype=np.uint64)
    ops = np.ascontiguousarray(ops, dtype=np.int32)
    return int(lib().or_children_hash(children, ops, len(children)))


def split_index_ops(ops, packed):
    packed = np.ascontiguousarray(packed, dtype=np.uint64)
    idx = np.zeros(len(packed), dtype=np.uint64)
    op = np.zeros(len(packed), dtype=np.int32)
    lib().or_split_index_ops(ops, packed, len(packed), idx, op)
    return idx, op


def start_index(p, start):
    """Start state text, or 'index:N'."""
    if start.startswith("index:"):
        return int(start.split(":")[1])
    r = parse(p, start)
    assert r != INVALID, start
    return r


def fs_segment_step(p, segment_bits, segment, nseg, idx, opbits, xseg=None):
    """One frontier-search segment step (or_fs_segment_step); returns a dict with
    count, the new frontier (idx, ops), its stream bytes, and the buckets
    [(op, seg, nvalues, bytes)] in (op, seg) order."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    opbits = np.ascontiguousarray(opbits, dtype=np.uint8)
    bf = int(lib().or_branching_factor(ctypes.byref(p)))
    ops = int(lib().or_operators_count(ctypes.byref(p)))
    xs_idx = xs_st = None
    nx = 0
    if xseg is not None:
        xs_idx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in xseg]), dtype=np.uint32)
        xs_st = np.zeros(ops + 1, np.uint64)
        xs_st[1:] = np.cumsum([len(x) for x in xseg])
        nx = len(xs_idx)
    cap = len(idx) * bf + nx + 8
    fr_idx = np.zeros(cap, np.uint32)
    fr_ops = np.zeros(cap, np.uint8)
    fr_bytes = np.zeros(5 * cap + 64 * (cap // (1 << 18) + 1) + 64, np.uint8)
    bk_cap = cap
    bk_op = np.zeros(bk_cap, np.uint32)
    bk_seg = np.zeros(bk_cap, np.uint32)
    bk_values = np.zeros(bk_cap, np.uint64)
    bk_nbytes = np.zeros(bk_cap, np.uint64)
    xs_bytes = np.zeros(5 * cap * bf + 64 * (cap * bf // 4096 + 1) + 64, np.uint8)
    io = FsIo(fr_idx.ctypes.data, fr_ops.ctypes.data, cap, 0, fr_bytes.ctypes.data, len(fr_bytes), 0,
              bk_op.ctypes.data, bk_seg.ctypes.data, bk_values.ctypes.data, bk_nbytes.ctypes.data, bk_cap, 0,
              xs_bytes.ctypes.data, len(xs_bytes), 0, 0, 0, 0)
    r = lib().or_fs_segment_step(ctypes.byref(p), segment_bits, segment, nseg, idx, opbits, len(idx),
                                 xs_idx.ctypes.data if xs_idx is not None else None,
                                 xs_st.ctypes.data if xs_st is not None else None, ctypes.byref(io))
    assert r >= 0, "oracle segment step failed"
    buckets = []
    off = 0
    for b in range(io.bk_n):
        nb = int(bk_nbytes[b])
        buckets.append((int(bk_op[b]), int(bk_seg[b]), int(bk_values[b]), xs_bytes[off:off + nb].copy()))
        off += nb
    return {"count": int(r), "fr_idx": fr_idx[:io.fr_n].copy(), "fr_ops": fr_ops[:io.fr_n].copy(),
            "fr_bytes": fr_bytes[:io.fr_nbytes].copy(), "buckets": buckets,
            "children_in": int(io.children_in), "children_cross": int(io.children_cross)}


def bfs3_segment_step(p, segment_bits, segment, nseg, cur, old, xseg=None):
    """One three-bit BFS segment step (or_bfs3_segment_step); returns a dict
    with count, the new frontier ids, their FrontierCompression stream bytes,
    and the buckets [(op, seg, nvalues, bytes)] in (op, seg) order."""
    cur = np.ascontiguousarray(cur, dtype=np.uint32)
    old = np.ascontiguousarray(old, dtype=np.uint32)
    bf = int(lib().or_branching_factor(ctypes.byref(p)))
    ops = int(lib().or_operators_count(ctypes.byref(p)))
    xs_idx = xs_st = None
    nx = 0
    if xseg is not None:
        xs_idx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in xseg] +
                                                     [np.zeros(0, np.uint32)]), dtype=np.uint32)
        xs_st = np.zeros(ops + 1, np.uint64)
        xs_st[1:] = np.cumsum([len(x) for x in xseg])
        nx = len(xs_idx)
    cap = len(cur) * bf + nx + 8
    fr_idx = np.zeros(cap, np.uint32)
    fr_bytes = np.zeros(5 * cap + 64 * (cap // (1 << 18) + 1) + 64, np.uint8)
    bk_cap = min(cap * bf, ops * nseg) + 8  # at most one bucket per (op, segment)
    bk_op = np.zeros(bk_cap, np.uint32)
    bk_seg = np.zeros(bk_cap, np.uint32)
    bk_values = np.zeros(bk_cap, np.uint64)
    bk_nbytes = np.zeros(bk_cap, np.uint64)
    xs_bytes = np.zeros(5 * cap * bf + 64 * (cap * bf // 4096 + 1) + 64, np.uint8)
    io = Seg3Io(fr_idx.ctypes.data, cap, 0, fr_bytes.ctypes.data, len(fr_bytes), 0,
                bk_op.ctypes.data, bk_seg.ctypes.data, bk_values.ctypes.data, bk_nbytes.ctypes.data, bk_cap, 0,
                xs_bytes.ctypes.data, len(xs_bytes), 0, 0, 0, 0)
    r = lib().or_bfs3_segment_step(ctypes.byref(p), segment_bits, segment, nseg, cur, len(cur), old, len(old),
                                   xs_idx.ctypes.data if xs_idx is not None else None,
                                   xs_st.ctypes.data if xs_st is not None else None, ctypes.byref(io))
    assert r >= 0, "oracle three-bit segment step failed"
    buckets = []
    off = 0
    for b in range(io.bk_n):
        nb = int(bk_nbytes[b])
        buckets.append((int(bk_op[b]), int(bk_seg[b]), int(bk_values[b]), xs_bytes[off:off + nb].copy()))
        off += nb
    return {"count": int(r), "fr_idx": fr_idx[:io.fr_n].copy(), "fr_bytes": fr_bytes[:io.fr_nbytes].copy(),
            "buckets": buckets, "children_in": int(io.children_in), "children_cross": int(io.children_cross)}


# ------------------------------------------------- HanoiTowers vertical solver

def hv_pegs_mult(index):
    return int(lib().or_hv_pegs_mult(int(index)))


def hv_in_segment_only(index, seg_bits=32):
    return bool(lib().or_hv_in_segment_only(int(index), seg_bits))


def hv_expand(disks, index):
    out = np.zeros(6, np.uint64)
    lib().or_hv_expand(disks, int(index), out)
    return [int(x) for x in out]


def hv_expand_cross_segment(disks, seg_bits, segment, index):
    out = np.zeros(6, np.uint64)
    n = lib().or_hv_expand_cross_segment(disks, seg_bits, int(segment), int(index), out)
    return [int(x) for x in out[:n]]


def hv_expand_in_segment_slots(disks, segment, idx, without_smallest=False):
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    out = np.zeros(len(idx) * (3 if without_smallest else 6), np.uint32)
    lib().or_hv_expand_in_segment_slots(disks, int(segment), idx, len(idx), int(without_smallest), out)
    return out


def hv_search(disks, start=0, seg_bits=32, max_layers=4096, with_hashes=True):
    """HanoiSearch restated: (restored counts, reduced counts, reduced-frontier hashes)."""
    restored = np.zeros(max_layers, np.uint64)
    reduced = np.zeros(max_layers, np.uint64)
    hashes = np.zeros(max_layers, np.uint64)
    n = lib().or_hv_search(disks, int(start), seg_bits, max_layers, restored, reduced,
                           hashes.ctypes.data if with_hashes else None)
    assert n > 0, "oracle vertical Hanoi search failed"
    return ([int(x) for x in restored[:n]], [int(x) for x in reduced[:n]],
            [int(x) for x in hashes[:n]] if with_hashes else None)
