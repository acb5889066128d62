"""The allocation-free bit scan (ps_bits_scan_noalloc) against ps_bits_scan,
through the C ABI, bit for bit: two calls in a row on one device, the second
with a smaller array (the static scratch keeps the first call's chunk counts,
which the second must not read), a capacity overflow, an empty array, and a
size beyond the scratch limit refused with a clear error."""
import numpy as np
import pytest

import puzzlesearch_amd as psa

pytestmark = pytest.mark.gpu


def _words(n, density, rng):
    bits = rng.random((n, 64)) < density
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(n)


def _count_buf(v):
    b = psa.DeviceBuffer(8)
    b.upload(np.array([v], np.uint64))
    return b


def _scan(fn, da, nbits, cap):
    out = psa.DeviceBuffer(max(8, cap * 8))
    out.zero()
    cnt = _count_buf(77)
    rc = fn(da.ptr if da else None, nbits, out.ptr, cap, cnt.ptr, None)
    psa.sync()
    return rc, out.download(np.uint64, cap), int(cnt.download(np.uint64, 1)[0])


def test_noalloc_scan_matches_scan(ps):
    L = psa.lib()
    rng = np.random.default_rng(11)
    # first call: 3 * 4096 + 5 words (four scratch chunks); second: 4097 words
    # minus a ragged tail (two chunks), sparser
    for nwords, tail, density in [(3 * 4096 + 5, 0, 0.3), (4097, 29, 0.01)]:
        a = _words(nwords, density, rng)
        nbits = nwords * 64 - tail
        da = psa.DeviceBuffer.from_array(a)
        want_rc, want, want_cnt = _scan(L.ps_bits_scan, da, nbits, nwords * 64)
        got_rc, got, got_cnt = _scan(L.ps_bits_scan_noalloc, da, nbits, nwords * 64)
        assert want_rc == got_rc == psa.PS_OK
        assert got_cnt == want_cnt > 0
        assert (got == want).all()
        ref = np.nonzero(np.unpackbits(a.view(np.uint8), bitorder="little")[:nbits])[0]
        assert (got[:got_cnt] == ref).all()
        # capacity overflow: the first cap indexes, and the full count
        cap = want_cnt // 2
        rc, part, cnt = _scan(L.ps_bits_scan_noalloc, da, nbits, cap)
        assert rc == psa.PS_OK and cnt == want_cnt
        assert (part == want[:cap]).all()


def test_noalloc_scan_empty_and_limit(ps):
    L = psa.lib()
    rc, _, cnt = _scan(L.ps_bits_scan_noalloc, None, 0, 0)
    assert rc == psa.PS_OK and cnt == 0
    one = psa.DeviceBuffer(8)
    cnt = _count_buf(5)
    rc = L.ps_bits_scan_noalloc(one.ptr, (1 << 38) + 1, one.ptr, 0, cnt.ptr, None)
    assert rc == psa.PS_ERR_INVALID_ARGUMENT
    assert b"scratch limit" in L.ps_last_error()
    psa.sync()
    assert int(cnt.download(np.uint64, 1)[0]) == 5  # nothing launched
