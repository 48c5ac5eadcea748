"""Independent counts for the full-size test, run in a child process in which
torch is the only user of the GPU (the library under test runs in the parent).

    python fullsize_count.py READS.npy VERTICES.npz N_READS READ_LEN K

READS.npy holds the reads' bytes (fixed-length ACGT reads, concatenated);
VERTICES.npz holds `keys` (int64 canonical k-mer of every vertex) and `g` (its
subgraph).  Prints one JSON line: the number of distinct vertex keys, the
number of distinct canonical k-mers of the reads, whether the two sets are
equal, and the number of (subgraph, key) pairs that occur more than once."""
import json
import sys

import numpy as np


def canonical_kmers_device(reads: bytes, n_reads, read_len, k, chunk=2_000_000, device="cuda"):
    """Sorted distinct canonical k-mers of fixed-length ACGT reads (no k-mer
    spans two reads), as an int64 tensor on `device`: forward and reverse
    complement rolled base by base, the smaller of the two kept."""
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


def main(reads_path, vert_path, n_reads, read_len, k):
    import torch
    reads = np.load(reads_path).tobytes()
    v = np.load(vert_path)
    keys = torch.from_numpy(v["keys"]).cuda()
    gs = torch.from_numpy(v["g"].astype(np.int64)).cuda()
    # (subgraph, key) pairs that occur twice: order by (g, key), compare neighbours
    o = torch.sort(keys, stable=True).indices
    o = o[torch.sort(gs[o], stable=True).indices]
    kk, gg = keys[o], gs[o]
    dup = int(((kk[1:] == kk[:-1]) & (gg[1:] == gg[:-1])).sum())
    got = torch.unique(keys)
    want = canonical_kmers_device(reads, n_reads, read_len, k)
    same = got.numel() == want.numel() and bool(torch.equal(got, want))
    print(json.dumps({"vertex_keys": got.numel(), "read_kmers": want.numel(), "equal": same,
                      "duplicate_pairs": dup}), flush=True)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5]))
