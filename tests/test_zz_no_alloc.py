"""The allocation-free Hammer entry point (kc_filter_correct_reads_noalloc, include/kc.h)
against kc_filter_correct_reads on the same filter: device inputs and outputs (allocated
by the caller with hipMalloc of the HIP runtime libkc.so links), two calls in a row with the second on a smaller input, the
corrected store and the corrected-read count equal byte for byte; the size limit and a
read past data_len are rejected."""
import ctypes
import sys

import numpy as np
import pytest

from kc import ErrorCorrector, PackedReads, TwoTierFilter, _abi
from kc._abi import lib

pytestmark = pytest.mark.gpu

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc_fn():
    f = lib().kc_filter_correct_reads_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    return f


def reads_of(rng, genome, n, lo, hi, err):
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi))
        s = int(rng.integers(0, len(genome) - L))
        r = bytearray(genome[s:s + L])
        for p in np.flatnonzero(rng.random(L) < err):
            r[p] = ACGT[(int(np.flatnonzero(ACGT == r[p])[0]) + int(rng.integers(1, 4))) & 3]
        out.append(bytes(r))
    return out


def hip():
    """The HIP runtime libkc.so is linked against (already loaded with it)."""
    lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


H2D, D2H = 1, 2


class DevArray:
    """A device copy of a numpy array (the caller's allocation)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(self.a.nbytes, 1)) == 0
        if self.a.nbytes:
            assert hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        if out.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc_correct(filt, k, store):
    data, offs, lens = store.arrays()
    d_data, d_off, d_len = DevArray(data), DevArray(offs), DevArray(lens)
    d_out = DevArray(np.zeros_like(data))
    n = ctypes.c_uint64(0)
    st = noalloc_fn()(filt._h, k, d_data.p, data.shape[0], d_off.p, d_len.p, lens.shape[0], d_out.p, ctypes.byref(n))
    assert st == _abi.KC_OK, lib().kc_filter_last_error(filt._h)
    out = d_out.get()
    for b in (d_data, d_off, d_len, d_out):
        b.free()
    return out, n.value


@pytest.mark.parametrize("k", [21, 55, 127])
def test_noalloc_matches_allocating_entry_point(oracle, k):
    rng = np.random.default_rng(k)
    genome = ACGT[rng.integers(0, 4, 30000)].tobytes()
    f = TwoTierFilter(1 << 20)
    win = [genome[i:i + k] for i in range(len(genome) - k + 1)]
    keys = np.array([oracle.kmer_canonical(oracle.kmer_from_bytes(w, k), k)[0] for w in win], dtype=np.uint64)
    f.insert_kmer256_many(np.concatenate([keys, keys]))
    for n, (lo, hi) in ((3000, (k - 3, 400)), (700, (k, 3000))):  # the second call is smaller
        store = PackedReads()
        for r in reads_of(rng, genome, n, max(lo, 1), hi, 0.005):
            store.pack_clean(r)
        say(f"case noalloc k={k} reads={len(store)}")
        want, n_want = ErrorCorrector(k).correct_packed(store, f)
        got, n_got = noalloc_correct(f, k, store)
        assert n_got == n_want > 0
        assert np.array_equal(got, want)
    f.close()


def test_noalloc_rejects_oversized_and_out_of_range():
    f = TwoTierFilter(1 << 16)
    fn = noalloc_fn()
    n = ctypes.c_uint64(0)
    d, out = DevArray(np.zeros(64, np.uint8)), DevArray(np.zeros(64, np.uint8))
    off = DevArray(np.array([0, 0, 0, 60], np.uint64))  # 21 bases need 6 bytes: read 3 is past the store
    ln = DevArray(np.full(4, 21, np.uint32))
    st = fn(f._h, 21, d.p, 64, off.p, ln.p, _abi.KC_NOALLOC_MAX_READS + 1, out.p, ctypes.byref(n))
    assert st == _abi.KC_EINVAL and b"KC_NOALLOC_MAX_READS" in lib().kc_filter_last_error(f._h)
    st = fn(f._h, 21, d.p, 64, off.p, ln.p, 3, out.p, ctypes.byref(n))
    assert st == _abi.KC_OK and n.value == 0
    st = fn(f._h, 21, d.p, 64, off.p, ln.p, 4, out.p, ctypes.byref(n))
    assert st == _abi.KC_EINVAL and b"past data_len" in lib().kc_filter_last_error(f._h)
    for b in (d, out, off, ln):
        b.free()
    f.close()
