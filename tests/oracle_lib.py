"""ctypes wrapper over oracle/libcforacle.so (the CPU oracle).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
u16p = ctypes.POINTER(ctypes.c_uint16)
u8p = ctypes.POINTER(ctypes.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(REPO, "oracle", "libcforacle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.cfo_wyhash.restype = ctypes.c_uint64
        L.cfo_wyhash.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint64]
        L.cfo_lmer_hash.restype = ctypes.c_uint64
        L.cfo_lmer_hash.argtypes = [ctypes.c_uint64]
        L.cfo_minimizers.restype = ctypes.c_int64
        L.cfo_minimizers.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, u64p, ctypes.c_size_t]
        L.cfo_parse.restype = ctypes.c_int64
        L.cfo_parse.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_char_p,
                                u64p, ctypes.c_size_t, ctypes.c_size_t]
        L.cfo_fragments.restype = ctypes.c_int64
        L.cfo_fragments.argtypes = [ctypes.c_char_p, u64p, ctypes.c_size_t, ctypes.c_int, u64p,
                                    u64p, ctypes.c_size_t]
        L.cfo_partition.restype = ctypes.c_int64
        L.cfo_partition.argtypes = [ctypes.c_char_p, u64p, ctypes.c_size_t, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_uint32, ctypes.c_int, u32p, u16p, u64p,
                                    u64p, ctypes.c_size_t]
        L.cfo_construct.restype = ctypes.c_int64
        L.cfo_construct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, u64p,
                                    u16p, u64p, ctypes.c_int, u32p, u64p, u32p, u8p,
                                    ctypes.c_size_t]
        L.cfo_pipeline.restype = ctypes.c_int
        L.cfo_pipeline.argtypes = [ctypes.c_char_p, u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.cfo_partition_hash.restype = ctypes.c_int
        L.cfo_partition_hash.argtypes = [ctypes.c_char_p, u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, u64p, u64p]
        L.cfo_records_hash.restype = ctypes.c_int
        L.cfo_records_hash.argtypes = [u32p, u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, u64p, u64p]
        L.cfo_states_hash.restype = ctypes.c_int
        L.cfo_states_hash.argtypes = [u32p, u64p, ctypes.c_int, u32p, u8p, ctypes.c_uint64, ctypes.c_int, u64p,
                                      u64p]
        L.cfo_baseline.restype = ctypes.c_int
        L.cfo_baseline.argtypes = [ctypes.c_char_p, u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.cfo_edge_at.restype = ctypes.c_uint32
        L.cfo_edge_at.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
        L.cfo_contract.restype = ctypes.c_int64
        L.cfo_contract.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_size_t, u32p, u64p,
                                   u32p, u8p, ctypes.c_int, ctypes.c_int, u64p, u32p, u32p, u64p, u32p,
                                   u64p, u8p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t, u64p]
        _LIB = L
    return _LIB


def label_words(k, l):
    return ((2 * (k - 1) - l + 2) + 31) // 32


def parse(buf: bytes, fastq: bool):
    """Records of a FASTA/FASTQ buffer -> (seq bytes, rec_off[n+1])."""
    L = lib()
    seq = ctypes.create_string_buffer(len(buf) + 1)
    off = np.zeros(len(buf) // 2 + 2, dtype=np.uint64)
    n = L.cfo_parse(buf, len(buf), int(fastq), seq, _ptr(off, u64p), len(buf), len(off) - 1)
    assert n >= 0
    tot = int(off[n])
    return seq.raw[:tot], off[: n + 1].copy()


def fragments(seq: bytes, rec_off, k):
    L = lib()
    cap = len(seq) // (k + 1) + 2
    beg = np.zeros(cap, dtype=np.uint64)
    ln = np.zeros(cap, dtype=np.uint64)
    ro = np.ascontiguousarray(rec_off, dtype=np.uint64)
    n = L.cfo_fragments(seq, _ptr(ro, u64p), len(ro) - 1, k, _ptr(beg, u64p), _ptr(ln, u64p), cap)
    assert n >= 0
    return beg[:n].copy(), ln[:n].copy()


def minimizers(frag: bytes, km1, l, brute=False):
    L = lib()
    out = np.zeros(max(1, len(frag)), dtype=np.uint64)
    n = L.cfo_minimizers(frag, len(frag), km1, l, int(brute), _ptr(out, u64p), len(out))
    assert n >= 0
    return out[:n].copy()


def partition(seq: bytes, rec_off, k, l, G=16384, scan_form=False):
    """Weak super k-mers of all records: dict of arrays g, att, labels (n x W), pos."""
    L = lib()
    W = label_words(k, l)
    cap = max(16, len(seq))
    g = np.zeros(cap, dtype=np.uint32)
    att = np.zeros(cap, dtype=np.uint16)
    lab = np.zeros(cap * W, dtype=np.uint64)
    pos = np.zeros(cap, dtype=np.uint64)
    ro = np.ascontiguousarray(rec_off, dtype=np.uint64)
    n = L.cfo_partition(seq, _ptr(ro, u64p), len(ro) - 1, k, l, G, int(scan_form), _ptr(g, u32p),
                        _ptr(att, u16p), _ptr(lab, u64p), _ptr(pos, u64p), cap)
    assert n >= 0, n
    return {"g": g[:n].copy(), "att": att[:n].copy(), "labels": lab[: n * W].reshape(n, W).copy(),
            "pos": pos[:n].copy()}


def bucketize(part, G=16384):
    """Group partition records by g (stable) -> (bucket_off[G+1], att, labels)."""
    order = np.argsort(part["g"], kind="stable")
    counts = np.bincount(part["g"], minlength=G)
    off = np.zeros(G + 1, dtype=np.uint64)
    off[1:] = np.cumsum(counts)
    return off, part["att"][order].copy(), part["labels"][order].copy()


def construct(k, l, bucket_off, att, labels, G=16384, mode=0, threads=8):
    """Per-(g, canonical k-mer) states sorted by (g, key): dict g, keys(n x 4), e_f, status."""
    L = lib()
    n_inst = int(np.sum((att & 0xFF).astype(np.int64) - (k - 1))) if len(att) else 0
    cap = max(16, n_inst)
    g = np.zeros(cap, dtype=np.uint32)
    keys = np.zeros(cap * 4, dtype=np.uint64)
    ef = np.zeros(cap, dtype=np.uint32)
    st = np.zeros(cap, dtype=np.uint8)
    bo = np.ascontiguousarray(bucket_off, dtype=np.uint64)
    a = np.ascontiguousarray(att, dtype=np.uint16)
    lb = np.ascontiguousarray(labels, dtype=np.uint64).reshape(-1)
    n = L.cfo_construct(k, l, G, mode, _ptr(bo, u64p), _ptr(a, u16p), _ptr(lb, u64p), threads,
                        _ptr(g, u32p), _ptr(keys, u64p), _ptr(ef, u32p), _ptr(st, u8p), cap)
    assert n >= 0, n
    return {"g": g[:n].copy(), "keys": keys[: 4 * n].reshape(n, 4).copy(), "e_f": ef[:n].copy(),
            "status": st[:n].copy()}


CONTRACT_COUNTS = ("isolated", "phantom", "trivial", "icc", "disc_edges", "vertex_count", "label_sz")


def contract(k, f_th, states, G=16384, normalise=True, threads=8):
    """Subgraph::contract of every subgraph over `states` (as construct returns
    them): per-bucket counters (G x 7, CONTRACT_COUNTS) and the unitigs: g,
    flags (1 cycle, 2 left exit, 4 right exit), labels (str), ends (n x 8 key
    words, left then right) and sides (n x 2; 0 front, 1 back, 2 phi)."""
    L = lib()
    n = len(states["g"])
    g = np.ascontiguousarray(states["g"], dtype=np.uint32)
    keys = np.ascontiguousarray(states["keys"], dtype=np.uint64).reshape(-1)
    ef = np.ascontiguousarray(states["e_f"], dtype=np.uint32)
    st = np.ascontiguousarray(states["status"], dtype=np.uint8)
    counts = np.zeros(G * 7, dtype=np.uint64)
    cap = max(16, n)
    rg = np.zeros(cap, np.uint32)
    rf = np.zeros(cap, np.uint32)
    ro = np.zeros(cap, np.uint64)
    rl = np.zeros(cap, np.uint32)
    ends = np.zeros(cap * 8, np.uint64)
    sides = np.zeros(cap * 2, np.uint8)
    lab_cap = max(64, n * k)
    labels = ctypes.create_string_buffer(lab_cap)
    need = ctypes.c_uint64()
    r = L.cfo_contract(k, f_th, G, n, _ptr(g, u32p), _ptr(keys, u64p), _ptr(ef, u32p), _ptr(st, u8p),
                       int(normalise), threads, _ptr(counts, u64p), _ptr(rg, u32p), _ptr(rf, u32p),
                       _ptr(ro, u64p), _ptr(rl, u32p), _ptr(ends, u64p), _ptr(sides, u8p), cap, labels,
                       lab_cap, ctypes.byref(need))
    assert r >= 0, (r, need.value)
    raw = labels.raw
    return {"counts": counts.reshape(G, 7), "g": rg[:r].copy(), "flags": rf[:r].copy(),
            "labels": [raw[int(o):int(o) + int(ln)].decode() for o, ln in zip(ro[:r], rl[:r])],
            "ends": ends[: 8 * r].reshape(r, 8).copy(), "sides": sides[: 2 * r].reshape(r, 2).copy()}


def edge_at(e_f, side, f_th):
    return lib().cfo_edge_at(int(e_f), side, f_th)


def pipeline(seq: bytes, rec_off, k, l, threads):
    """Timed CPU pipeline (partition, bucketize, construct) on `threads` workers."""
    ro = np.ascontiguousarray(rec_off, dtype=np.uint64)
    out = (ctypes.c_double * 6)()
    r = lib().cfo_pipeline(seq, _ptr(ro, u64p), len(ro) - 1, k, l, threads, out)
    assert r == 0
    return {"partition_ms": out[0], "bucketize_ms": out[1], "construct_ms": out[2],
            "n_kmers": int(out[3]), "n_records": int(out[4]), "n_states": int(out[5])}


def partition_hash(seq: bytes, rec_off, k, l, threads=16):
    """Per-bucket (checksum, count) of the oracle's weak super k-mers."""
    ro = np.ascontiguousarray(rec_off, dtype=np.uint64)
    s, c = np.zeros(16384, np.uint64), np.zeros(16384, np.uint64)
    lib().cfo_partition_hash(seq, _ptr(ro, u64p), len(ro) - 1, k, l, threads, _ptr(s, u64p), _ptr(c, u64p))
    return s, c


def records_hash(att, labels, W, threads=16):
    """Per-bucket (checksum, count) of records as the device returns them."""
    a = np.ascontiguousarray(att, dtype=np.uint32)
    lb = np.ascontiguousarray(labels, dtype=np.uint64).reshape(-1)
    s, c = np.zeros(16384, np.uint64), np.zeros(16384, np.uint64)
    lib().cfo_records_hash(_ptr(a, u32p), _ptr(lb, u64p), len(a), W, threads, _ptr(s, u64p), _ptr(c, u64p))
    return s, c


def states_hash(t, threads=16):
    """Per-bucket (checksum, count) of vertex states (dict g, keys, e_f, status)."""
    g = np.ascontiguousarray(t["g"], dtype=np.uint32)
    keys = np.ascontiguousarray(t["keys"], dtype=np.uint64)
    ef = np.ascontiguousarray(t["e_f"], dtype=np.uint32)
    st = np.ascontiguousarray(t["status"], dtype=np.uint8)
    s, c = np.zeros(16384, np.uint64), np.zeros(16384, np.uint64)
    lib().cfo_states_hash(_ptr(g, u32p), _ptr(keys, u64p), keys.shape[1] if keys.ndim == 2 else 1, _ptr(ef, u32p),
                          _ptr(st, u8p), len(g), threads, _ptr(s, u64p), _ptr(c, u64p))
    return s, c


def baseline(seq: bytes, rec_off, k, l, cutoff, part_threads, threads):
    """CPU baseline in the reference's shape (partition on part_threads workers,
    bucketize, per-subgraph construct with a reusable open-addressing map and
    contract, largest buckets first, on `threads` workers); times in ms."""
    ro = np.ascontiguousarray(rec_off, dtype=np.uint64)
    out = (ctypes.c_double * 9)()
    r = lib().cfo_baseline(seq, _ptr(ro, u64p), len(ro) - 1, k, l, cutoff, part_threads, threads, out)
    if r != 0:
        raise RuntimeError("cfo_baseline failed")
    keys = ("partition_ms", "bucketize_ms", "construct_ms", "contract_ms", "n_kmers", "n_superkmers",
            "n_states", "n_unitigs", "construct_contract_ms")
    d = dict(zip(keys, list(out)))
    for kk in ("n_kmers", "n_superkmers", "n_states", "n_unitigs"):
        d[kk] = int(d[kk])
    return d

