"""The independent distinct-k-mer count of the full-size GPU test (fullsize_count.py), checked on
the CPU against a plain numpy restatement on a small read set (both strands
rolled base by base; min of forward and reverse complement)."""
import numpy as np

import cfgpu as C
from fullsize_count import canonical_kmers_device


def numpy_canonical(reads: bytes, n_reads, read_len, k):
    c = np.frombuffer(reads, np.uint8).reshape(n_reads, read_len)
    code = np.select([c == 65, c == 67, c == 71], [0, 1, 2], 3).astype(np.uint64)
    n = read_len - k + 1
    fwd = np.zeros((n_reads, n), np.uint64)
    rev = np.zeros((n_reads, n), np.uint64)
    for j in range(k):
        fwd = (fwd << np.uint64(2)) | code[:, j:j + n]
        rev = rev | ((np.uint64(3) - code[:, j:j + n]) << np.uint64(2 * j))
    return np.unique(np.minimum(fwd, rev))


def test_canonical_kmer_count_matches_numpy():
    g = C.synth_genome(100_000, 0.0, 300, 6000, 0.0, seed=5)
    reads = C.synth_reads(g, 3_000, 150, 0.005, seed=6)
    for k in (31, 21):
        got = canonical_kmers_device(reads, 3_000, 150, k, chunk=1_100, device="cpu").numpy()
        want = numpy_canonical(reads, 3_000, 150, k)
        assert np.array_equal(got.astype(np.uint64), want)


def test_canonical_kmer_count_plants_a_mistake():
    # a reverse complement that forgets the complement must not agree
    reads = b"ACGTTGCAAC" * 30
    want = numpy_canonical(reads, 2, 150, 31)
    c = np.frombuffer(reads, np.uint8).reshape(2, 150)
    code = np.select([c == 65, c == 67, c == 71], [0, 1, 2], 3).astype(np.uint64)
    fwd = np.zeros((2, 120), np.uint64)
    rev = np.zeros((2, 120), np.uint64)
    for j in range(31):
        fwd = (fwd << np.uint64(2)) | code[:, j:j + 120]
        rev = rev | (code[:, j:j + 120] << np.uint64(2 * j))
    assert not np.array_equal(np.unique(np.minimum(fwd, rev)), want)
    got = canonical_kmers_device(reads, 2, 150, 31, device="cpu").numpy()
    assert np.array_equal(got.astype(np.uint64), want)
