"""cfgpu_table_copy_noalloc (outputs in the caller's device arrays, no
allocation in the call) against cfgpu_table_copy on the same tables, on two
calls in a row, the second with fewer buckets.  Bar: bit-exact rows."""
import ctypes

import numpy as np
import pytest

import cfgpu as C
from test_gpu_parity import rows_sorted, say

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP = L
    return _HIP


def dev_alloc(nbytes):
    p = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(p), max(nbytes, 1)) == 0
    return p


def to_host(p, dtype, n):
    a = np.zeros(max(n, 1), dtype)
    assert hip().hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0
    return a[:n]


def noalloc_tables(h, buckets):
    b = np.ascontiguousarray(buckets, dtype=np.uint32)
    n = int(h.bucket_stats()["n_vertices"][b].sum())
    kw = h.kw
    bufs = [dev_alloc(n * kw * 8), dev_alloc(n * 4), dev_alloc(n), dev_alloc(n)]
    try:
        bn = np.zeros(max(len(b), 1), np.uint32)
        got = ctypes.c_uint64()
        h._check(C.lib().cfgpu_table_copy_noalloc(h._h, b.ctypes.data_as(C.u32p), len(b), bufs[0], bufs[1], bufs[2],
                                                  bufs[3], bn.ctypes.data_as(C.u32p), n, ctypes.byref(got)),
                 "table_copy_noalloc")
        assert got.value == n
        keys = to_host(bufs[0], np.uint64, n * kw).reshape(n, kw)
        return {"g": np.repeat(b, bn[: len(b)]), "keys": keys, "e_f": to_host(bufs[1], np.uint32, n),
                "status": to_host(bufs[2], np.uint8, n), "dfa": to_host(bufs[3], np.uint8, n)}
    finally:
        for p in bufs:
            hip().hipFree(p)


def rows(t, kw):
    return rows_sorted([t["g"]] + [t["keys"][:, j] for j in range(kw - 1, -1, -1)] +
                       [t["e_f"], t["status"], t["dfa"]])


@pytest.mark.parametrize("k", [31, 63])
def test_table_copy_noalloc_matches_table_copy(k):
    g = C.synth_genome(60000, 0.1, 300, 2000, 0.02, seed=81)
    reads = C.synth_reads(g, 8000, 150, 0.01, seed=82)
    inp = C.Input.from_records(reads, np.arange(8001, dtype=np.uint64) * np.uint64(150), k)
    with C.Handle(k, 20, 2) as h:
        h.partition(inp)
        h.construct("learned")
        nz = np.nonzero(h.bucket_stats()["n_records"])[0].astype(np.uint32)
        for call, buckets in enumerate((nz, nz[::7][:300])):
            say(f"[case] noalloc k={k} call={call} buckets={len(buckets)}")
            want = h.tables(buckets)
            got = noalloc_tables(h, buckets)
            assert len(got["g"]) == len(want["g"]) > 0
            np.testing.assert_array_equal(rows(got, h.kw), rows(want, h.kw))
        too_many = np.zeros(16385, np.uint32)
        r = C.lib().cfgpu_table_copy_noalloc(h._h, too_many.ctypes.data_as(C.u32p), 16385, None, None, None, None,
                                             None, 0, None)
        assert r == -1
