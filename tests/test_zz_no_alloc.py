"""sph_radix_sort_noalloc, the allocation-free RadixSort: through the C ABI,
two calls in a row (the second on fewer records, so the fixed scratch array
holds the first call's tables) against RadixSort on the same input and
against numpy's stable argsort, bit for bit; sizes above the maximum are
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(L):
    L.allocateCUDAarray.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int]
    L.deleteCUDAarray.argtypes = [C.c_void_p, C.c_int]
    L.copyArrayToDevice.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.copyArrayFromDevice.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    L.RadixSort.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint]
    L.sph_radix_sort_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint, C.c_void_p]
    L.sph_radix_sort_noalloc.restype = C.c_int
    L.sph_last_error.restype = C.c_char_p


def _records(n, bits, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    keys[::5] = keys[0]          # many equal keys: stability matters
    kv = np.zeros((n, 2), np.uint32)
    kv[:, 0], kv[:, 1] = keys, np.arange(n, dtype=np.uint32)
    return kv


def _sort(L, fn, kv, bits):
    """Sort a copy of kv on the device with RadixSort (fn None) or the
    no-alloc entry point; returns the records and the return code."""
    bufs = []
    try:
        for _ in range(2):
            p = C.c_void_p()
            L.allocateCUDAarray(C.byref(p), kv.nbytes, 0)
            bufs.append(p)
        L.copyArrayToDevice(bufs[0], kv.ctypes.data, kv.nbytes, 0)
        if fn is None:
            L.RadixSort(bufs[0], bufs[1], len(kv), bits)
            rc = 0
        else:
            rc = fn(bufs[0], bufs[1], len(kv), bits, None)
        out = np.zeros_like(kv)
        L.copyArrayFromDevice(out.ctypes.data, bufs[0], out.nbytes, 0)   # null stream: after the sort
        return out, rc
    finally:
        for p in bufs:
            L.deleteCUDAarray(p, 0)


def test_noalloc_sort_matches_radix_sort_on_two_calls(cs):
    L = cs.lib()
    _bind(L)
    for n, bits, seed in ((1_000_003, 32, 1), (70_001, 20, 2), (4096 * 3 + 17, 8, 3)):
        print(f"noalloc sort n={n} bits={bits}", flush=True)
        kv = _records(n, bits, seed)
        got, rc = _sort(L, L.sph_radix_sort_noalloc, kv, bits)
        assert rc == 0, L.sph_last_error()
        ref, _ = _sort(L, None, kv, bits)
        assert np.array_equal(got, ref), (n, bits)
        order = np.argsort(kv[:, 0], kind="stable")
        assert np.array_equal(got[:, 0], kv[order, 0]) and np.array_equal(got[:, 1], order), (n, bits)


def test_noalloc_sort_refuses_more_than_its_maximum(cs):
    L = cs.lib()
    _bind(L)
    p = C.c_void_p()
    L.allocateCUDAarray(C.byref(p), 64, 0)
    try:
        rc = L.sph_radix_sort_noalloc(p, p, (1 << 24) + 1, 32, None)
        assert rc == -1
        assert b"exceed the maximum" in L.sph_last_error()
        rc = L.sph_radix_sort_noalloc(p, p, 4, 33, None)
        assert rc == -1 and b"bits" in L.sph_last_error()
    finally:
        L.deleteCUDAarray(p, 0)
