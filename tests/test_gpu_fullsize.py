"""The benchmark workload at full size, checked through properties that hold at
any size: 10,000,000 reads of 150 bp from a 50,000,000 bp iid genome (both
strands, 0.5 % substitutions), k = 31, l = 20, cutoff 2, 16384 subgraphs.

The oracle cannot run at this size, so the checks are ones the domain fixes.
(1) A k-mer is a vertex of every subgraph one of its weak super k-mers reaches
(src/Subgraph.cpp:57-105 inserts each instance into the bucket's table), so
the union of all subgraphs' vertex keys is exactly the set of distinct
canonical 31-mers of the reads.  That set is computed independently here with
torch on the device.  (2) No subgraph holds a key twice.  (3) The learned
batched path and the plain per-bucket path build the same number of vertices
in every subgraph."""
import numpy as np
import pytest

import cfgpu as C
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

K, L, CUTOFF = 31, 20, 2
N_READS, READ_LEN = 10_000_000, 150


def canonical_kmers_device(reads: bytes, n_reads, read_len, k, chunk=2_000_000, device="cuda"):
    """Sorted distinct canonical k-mers of fixed-length ACGT reads (no k-mer
    spans two reads), as an int64 tensor on `device`."""
    import torch
    lut = torch.full((256,), 255, dtype=torch.uint8)
    for ch, v in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3)):
        lut[ch[0]] = v
    lut = lut.to(device)
    n = read_len - k + 1
    parts = []
    for r0 in range(0, n_reads, chunk):
        r1 = min(n_reads, r0 + chunk)
        raw = torch.frombuffer(bytearray(reads[r0 * read_len:r1 * read_len]), dtype=torch.uint8)
        code = lut[raw.to(device).long()].view(r1 - r0, read_len)
        assert int(code.max()) <= 3, "reads must be ACGT only"
        code = code.long()
        fwd = torch.zeros((r1 - r0, n), dtype=torch.int64, device=device)
        rev = torch.zeros_like(fwd)
        for j in range(k):
            fwd = (fwd << 2) | code[:, j:j + n]
            rev = rev | ((3 - code[:, j:j + n]) << (2 * j))
        parts.append(torch.unique(torch.minimum(fwd, rev)))
        del fwd, rev, code
    return torch.unique(torch.cat(parts))


def test_bench_workload_vertices_are_the_distinct_kmers():
    import torch
    g = C.synth_genome(50_000_000, 0.0, 300, 6000, 0.0, seed=2026)
    reads = C.synth_reads(g, N_READS, READ_LEN, 0.005, seed=2027)
    rec_off = np.arange(N_READS + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    inp = C.Input.from_records(reads, rec_off, K)
    say("[case] full-size bench workload: partition + learned construct")
    with C.Handle(K, L, CUTOFF) as h:
        h.partition(inp)
        h.construct("learned")
        nv_learned = h.bucket_stats()["n_vertices"].astype(np.int64)
        t = h.tables()
        say("[case] full-size bench workload: plain construct")
        h.construct("plain")
        nv_plain = h.bucket_stats()["n_vertices"].astype(np.int64)
    assert (nv_learned == nv_plain).all()
    assert len(t["g"]) == int(nv_learned.sum())
    say(f"[case] full-size: {len(t['g'])} vertices; distinct canonical k-mers on the device")

    keys = torch.from_numpy(t["keys"][:, 0].astype(np.int64)).cuda()
    gs = torch.from_numpy(t["g"].astype(np.int64)).cuda()
    # (2) no key twice in one subgraph: order by (g, key), no equal neighbours
    o = torch.sort(keys, stable=True).indices
    o = o[torch.sort(gs[o], stable=True).indices]
    kk, gg = keys[o], gs[o]
    assert not bool(((kk[1:] == kk[:-1]) & (gg[1:] == gg[:-1])).any())
    # (1) the union of the vertex keys is the set of distinct canonical k-mers
    got = torch.unique(keys)
    want = canonical_kmers_device(reads, N_READS, READ_LEN, K)
    say(f"[case] full-size: union {got.numel()} distinct keys, reads {want.numel()} canonical {K}-mers")
    assert got.numel() == want.numel()
    assert torch.equal(got, want)
