"""The benchmark workload at full size, checked through properties that hold at
any size: 10,000,000 reads of 150 bp from a 50,000,000 bp iid genome (both
strands, 0.5 % substitutions), k = 31, l = 20, cutoff 2, 16384 subgraphs.

The oracle cannot run at this size, so the checks are ones the domain fixes.
(1) A k-mer is a vertex of every subgraph one of its weak super k-mers reaches
(src/Subgraph.cpp:57-105 inserts each instance into the bucket's table), so
the union of all subgraphs' vertex keys is exactly the set of distinct
canonical 31-mers of the reads.  (2) No subgraph holds a key twice.  (3) The
learned batched path and the plain per-bucket path build the same number of
vertices in every subgraph.  (1) and (2) are counted independently with torch
on the device, in a child process (fullsize_count.py) so that torch is the
only GPU user there."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import cfgpu as C
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

K, L, CUTOFF = 31, 20, 2
N_READS, READ_LEN = 10_000_000, 150
HERE = os.path.dirname(os.path.abspath(__file__))


def test_bench_workload_vertices_are_the_distinct_kmers(tmp_path):
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

    reads_path, vert_path = str(tmp_path / "reads.npy"), str(tmp_path / "vertices.npz")
    np.save(reads_path, np.frombuffer(reads, np.uint8))
    np.savez(vert_path, keys=t["keys"][:, 0].astype(np.int64), g=t["g"].astype(np.int32))
    del t
    say(f"[case] full-size: {int(nv_learned.sum())} vertices; independent count in a child process")
    r = subprocess.run([sys.executable, os.path.join(HERE, "fullsize_count.py"), reads_path, vert_path,
                        str(N_READS), str(READ_LEN), str(K)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    say(f"[case] full-size: {res}")
    assert res["duplicate_pairs"] == 0
    assert res["vertex_keys"] == res["read_kmers"]
    assert res["equal"]
