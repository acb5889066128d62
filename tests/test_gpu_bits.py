"""GPU parity of the standalone bit-array functions of the C ABI
(ps_bits_andnot_count / count / hash / scan / set_list) against the oracle's
BitArray restatement (Common/BitArray.h:81-145) and numpy, bit for bit: empty,
one-word and ragged sizes, tails that are not a multiple of 64 bits, arrays of
several 4096-word scan chunks, capacity overflow, and out-of-range indexes."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import puzzlesearch_amd as psa

pytestmark = pytest.mark.gpu

SIZES = [1, 3, 4095, 4097, 12289, (1 << 20) + 3]


def _words(n, density, rng):
    bits = rng.random((n, 64)) < density
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(n)


def _popcount(a):
    return int(np.unpackbits(a.view(np.uint8)).sum())


def _u64_buffer(v=0):
    b = psa.DeviceBuffer(8)
    b.upload(np.array([v], np.uint64))
    return b


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("density", [0.001, 0.5])
def test_andnot_count(ps, n, density):
    rng = np.random.default_rng(n * 10 + int(density * 100))
    a, e1, e2 = (_words(n, density, rng) for _ in range(3))
    for with_e2 in (False, True):
        da, d1, d2 = psa.DeviceBuffer.from_array(a), psa.DeviceBuffer.from_array(e1), psa.DeviceBuffer.from_array(e2)
        cnt = _u64_buffer(5)  # the count is added to what the word holds
        psa.check(psa.lib().ps_bits_andnot_count(da.ptr, d1.ptr, d2.ptr if with_e2 else None, n, cnt.ptr, None))
        got = da.download(np.uint64, n)
        want = a & ~e1
        if with_e2:
            want &= ~e2
        assert (got == want).all()
        assert int(cnt.download(np.uint64, 1)[0]) == 5 + _popcount(want)
        if not with_e2:
            ref = a.copy()
            assert O.lib().or_bits_andnot_count(ref, e1, n) == _popcount(want)
            assert (ref == got).all()


@pytest.mark.parametrize("n", SIZES)
def test_count_and_hash(ps, n):
    rng = np.random.default_rng(n)
    a = _words(n, 0.01, rng)
    a[rng.integers(0, n, size=max(1, n // 3))] = 0  # zero words hash to nothing
    da = psa.DeviceBuffer.from_array(a)
    cnt, h = _u64_buffer(), _u64_buffer()
    psa.check(psa.lib().ps_bits_count(da.ptr, n, cnt.ptr, None))
    psa.check(psa.lib().ps_bits_hash(da.ptr, n, 0, h.ptr, None))
    assert int(cnt.download(np.uint64, 1)[0]) == _popcount(a) == O.lib().or_bits_count(a, n)
    assert int(h.download(np.uint64, 1)[0]) == O.bits_hash(a)
    # word_offset: the hash of a slice is the hash of the array with the
    # slice's words at their own positions
    k = n // 2
    h2 = _u64_buffer()
    psa.check(psa.lib().ps_bits_hash(da.ptr + 8 * k, n - k, k, h2.ptr, None))
    b = a.copy()
    b[:k] = 0
    assert int(h2.download(np.uint64, 1)[0]) == O.bits_hash(b)


def _scan(da, nbits, cap):
    out = psa.DeviceBuffer(max(8, cap * 8))
    cnt = _u64_buffer(123)
    rc = psa.lib().ps_bits_scan(da.ptr if da else None, nbits, out.ptr, cap, cnt.ptr, None)
    return rc, out.download(np.uint64, cap), int(cnt.download(np.uint64, 1)[0])


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("tail", [0, 1, 37, 63])
def test_scan(ps, n, tail):
    """ScanBits: ascending set-bit indexes of [0, nbits); bits past nbits in
    the last word are ignored."""
    rng = np.random.default_rng(n * 64 + tail)
    a = _words(n, 0.05, rng)
    nbits = n * 64 - tail
    want = np.nonzero(np.unpackbits(a.view(np.uint8), bitorder="little")[:nbits])[0].astype(np.uint64)
    da = psa.DeviceBuffer.from_array(a)
    rc, got, cnt = _scan(da, nbits, len(want) + 5)
    assert rc == psa.PS_OK
    assert cnt == len(want)
    assert (got[: len(want)] == want).all()
    if len(want) > 2:
        rc, got, cnt = _scan(da, nbits, len(want) - 2)
        assert rc == psa.PS_ERR_CAPACITY
        assert cnt == len(want)
        assert (got == want[: len(want) - 2]).all()


def test_scan_empty(ps):
    rc, _, cnt = _scan(None, 0, 0)
    assert rc == psa.PS_OK and cnt == 0


@pytest.mark.parametrize("nbits", [1, 64, 1000, 1 << 22])
def test_set_list(ps, nbits):
    """BitArray::Set of a list of indexes, duplicates included; indexes past
    nbits are counted in *d_rejected and not written."""
    rng = np.random.default_rng(nbits)
    nwords = (nbits + 63) // 64
    a = _words(nwords, 0.1, rng)
    if nbits % 64:
        a[-1] &= np.uint64((1 << (nbits % 64)) - 1)
    idx = rng.integers(0, nbits, size=5000, dtype=np.uint64)
    bad = np.array([nbits, nbits + 1, nbits + 63, 2**63, 2**64 - 1], dtype=np.uint64)
    allidx = np.concatenate([idx, bad, idx[:100]])
    rng.shuffle(allidx)
    # one guard word after the array: an unchecked write would land there
    da = psa.DeviceBuffer.from_array(np.concatenate([a, np.zeros(1, np.uint64)]))
    di = psa.DeviceBuffer.from_array(allidx)
    rej = _u64_buffer(2)
    psa.check(psa.lib().ps_bits_set_list(da.ptr, nbits, di.ptr, len(allidx), rej.ptr, None))
    got = da.download(np.uint64, nwords + 1)
    bits = np.unpackbits(a.view(np.uint8), bitorder="little")
    bits[idx.astype(np.int64)] = 1
    want = np.packbits(bits, bitorder="little").view(np.uint64)
    assert (got[:nwords] == want).all()
    assert got[nwords] == 0
    assert int(rej.download(np.uint64, 1)[0]) == 2 + len(bad)
    # without a rejected counter
    psa.check(psa.lib().ps_bits_set_list(da.ptr, nbits, di.ptr, len(allidx), None, None))
    assert (da.download(np.uint64, nwords + 1) == got).all()
