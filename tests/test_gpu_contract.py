"""GPU parity of Subgraph::contract on the device-built tables (cfgpu_contract)
against the oracle's restatement (oracle/cf_oracle_contract.cpp), itself
pinned against brute-force global unitigs in test_oracle_contract.py.

Bar: bit-exact.  Per subgraph, the multiset of (label, flags) equals the
oracle's (trivial unitigs in the reference's output orientation, lm-tigs
oriented canonically, cycles rotated to their smallest vertex), and the
counters (isolated, phantom, trivial, cycles, discontinuity edges, vertices in
unitigs, label bytes) are equal."""
import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O
from test_gpu_parity import fastq, oracle_records, rows_sorted, say
from test_oracle import rand_seq, rc
from test_oracle_contract import inputs as mixed_input

pytestmark = pytest.mark.gpu


def compare(name, got, stats, want):
    np.testing.assert_array_equal(stats[:, :7].astype(np.uint64), want["counts"][:, :7])
    built = stats[:, 7] != 0
    assert not (stats[:, 7] & 4).any(), "inconsistent table flagged"
    assert built[want["counts"][:, 5] + want["counts"][:, 0] > 0].all()
    gl = sorted(zip(got["g"].tolist(), got["labels"], got["flags"].tolist()))
    wl = sorted(zip(want["g"].tolist(), want["labels"], want["flags"].tolist()))
    say(f"[case] {name}: unitigs={len(gl)} oracle={len(wl)}")
    assert len(gl) == len(wl)
    assert gl == wl


def run(name, buf, fq, k, l, cutoff, modes=("learned",)):
    say(f"[case] {name}: k={k} l={l} cutoff={cutoff} bytes={len(buf)}")
    seq, ro, part = oracle_records(buf, fq, k, l)
    off, att, lab = O.bucketize(part)
    want = O.contract(k, cutoff, O.construct(k, l, off, att, lab), normalise=True)
    inp = C.Input.from_buffer(buf, fq, k)
    with C.Handle(k, l, cutoff) as h:
        h.partition(inp)
        for mode in modes:
            say(f"[case] {name}: construct {mode} + contract")
            h.construct(mode)
            h.contract()
            compare(name, h.unitigs(), h.contract_stats(), want)
            rs = h.run_stats()
            assert rs["n_unitigs"] == len(want["labels"])
    return want


@pytest.mark.parametrize("k,l,cutoff", [(31, 20, 1), (31, 20, 2), (15, 7, 1), (63, 20, 1)])
def test_contract_mixed_small(k, l, cutoff):
    rng = np.random.default_rng(7 + k + cutoff)
    run("mixed", mixed_input(rng), False, k, l, cutoff, modes=("plain", "batched", "learned"))


def test_contract_tandem_cycles():
    rng = np.random.default_rng(21)
    recs = [rand_seq(rng, p) * (1 + 100 // p) for p in range(2, 10) for _ in range(4)]
    buf = "".join(f">c{i}\n{r}\n" for i, r in enumerate(recs)).encode()
    for k, l in ((15, 7), (31, 20)):
        want = run("cycles", buf, False, k, l, 1)
        assert (want["flags"] & 1).sum() >= 20


def test_contract_reads_cutoff2():
    g = C.synth_genome(300_000, 0.2, 300, 3000, 0.02, seed=31)
    reads = C.synth_reads(g, 60_000, 150, 0.005, seed=32)
    run("reads", fastq(reads, 150), True, 31, 20, 2, modes=("batched", "learned"))


def test_contract_reference_k127():
    base = C.synth_genome(200_000, 0.1, 300, 2000, 0.02, seed=33)
    recs = [base] + [C.synth_mutate(base, 0.005, 0.01, 10, seed=34 + i) for i in range(3)]
    buf = b"".join(b">r%d\n%s\n" % (i, r) for i, r in enumerate(recs))
    run("k127", buf, False, 127, 20, 1)


@pytest.mark.parametrize("wide", [0, 1])
def test_contract_large_subgraphs_use_dense_and_global_tables(wide):
    # Zipf-sized buckets of random records: the large ones hold more vertices
    # than an LDS table (dense tables: vertex arrays in global scratch, index
    # in LDS), the largest more than a dense table: wide dense tables (32-bit
    # states, index in global scratch) or, with those off, the global-scratch
    # walk.
    for k, l in ((31, 20), (63, 20)):
        att, lab, off = C.synth_bucket_workload(256, 150000, 1.1, k, l, seed=40 + k)
        st = O.construct(k, l, off, (att & 0xFFFF).astype(np.uint16), lab)
        want = O.contract(k, 1, st, normalise=True)
        say(f"[case] zipf-contract k={k} wide={wide} vertices={len(st['g'])}")
        with C.Handle(k, l, 1) as h:
            h.set_tuning(contract_wide=wide)
            h.load_buckets(att, lab, off)
            h.construct("batched")
            h.contract()
            stats = h.contract_stats()
            compare("zipf", h.unitigs(), stats, want)
            kinds = {f: int((stats[:, 7] == f).sum()) for f in (1, 8, 16, 2)}
            say(f"[case] zipf-contract k={k} wide={wide} subgraphs by table: {kinds}")
            assert kinds[1] > 0 and kinds[8] > 0
            assert (kinds[16] > 0 and kinds[2] == 0) if wide else (kinds[2] > 0 and kinds[16] == 0)
            rs = h.run_stats()
            assert rs["n_contract_global"] == kinds[2] and rs["n_contract_dense"] == kinds[8]
            assert rs["n_contract_wide"] == kinds[16]


def test_contract_after_release_skips_subgraph():
    rng = np.random.default_rng(3)
    buf = mixed_input(rng)
    k, l = 31, 20
    inp = C.Input.from_buffer(buf, False, k)
    with C.Handle(k, l, 1) as h:
        h.partition(inp)
        h.construct("batched")
        h.contract()
        before = h.contract_stats()
        g = int(np.nonzero(before[:, 5])[0][0])
        h.release(g)
        h.contract()
        after = h.contract_stats()
        assert after[g].sum() == 0
        keep = np.ones(len(before), bool)
        keep[g] = False
        np.testing.assert_array_equal(after[keep], before[keep])
        got = h.unitigs()
        assert g not in set(got["g"].tolist())


def _records_one_bucket(seqs, k, l, rng, disc_rate=0.0):
    """Overlapping records (length k+1 .. 2k-l, consecutive ones sharing k
    bases, so every (k+1)-mer of a sequence lies in one record) of every
    sequence, all in bucket 0, in the reference's record layout."""
    W = O.label_words(k, l)
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    att, lab = [], []
    for s in seqs:
        pos = 0
        while pos + k < len(s):
            L = min(int(rng.integers(k + 1, 2 * k - l + 1)), len(s) - pos)
            rec = s[pos:pos + L]
            words = [0] * W
            for i, ch in enumerate(rec):
                words[W - 1 - i // 32] |= code[ch] << (62 - 2 * (i % 32))
            ld = int(rng.random() < disc_rate) if pos else 0
            rd = int(rng.random() < disc_rate)
            att.append(L | (ld << 8) | (rd << 9))
            lab.append(words)
            pos += L - k
    off = np.zeros(C.G + 1, np.uint64)
    off[1:] = len(att)
    return np.array(att, np.uint32), np.array(lab, np.uint64).reshape(len(att), W), off


@pytest.mark.parametrize("n_random", [8000, 24000])
def test_contract_dense_tables_with_cycles(n_random):
    # One subgraph past an LDS table, holding a long random path and tandem
    # repeats (each a cycle of its period's k-mers): ranked in LDS (8000
    # vertices) or in global scratch (24000), cycles walked by successor words.
    k, l = 31, 20
    rng = np.random.default_rng(n_random)
    seqs = [rand_seq(rng, n_random)]
    for p in rng.integers(40, 200, size=12):
        u = rand_seq(rng, int(p))
        seqs.append(u * (2 + (k + 1) // int(p) + 1))
    att, lab, off = _records_one_bucket(seqs, k, l, rng, disc_rate=0.05)
    st = O.construct(k, l, off, (att & 0xFFFF).astype(np.uint16), lab)
    want = O.contract(k, 1, st, normalise=True)
    assert (want["flags"] & 1).sum() >= 10
    say(f"[case] dense-cycles n={n_random} vertices={len(st['g'])} cycles={int((want['flags'] & 1).sum())}")
    with C.Handle(k, l, 1) as h:
        h.load_buckets(att, lab, off)
        h.construct("batched")
        h.contract()
        stats = h.contract_stats()
        assert stats[0, 7] == 8  # the dense tables took it
        compare("dense-cycles", h.unitigs(), stats, want)


def test_construct_in_bucket_batches_glues_to_the_whole_graph():
    # cfgpu_construct_buckets: two batches of buckets, each contracted on its
    # own; each batch's tables equal the oracle's for its buckets, the other
    # buckets have none, and the two batches' unitigs glued on the host give
    # the compacted graph of a single construct.
    g = C.synth_genome(300_000, 0.2, 300, 3000, 0.02, seed=51)
    reads = C.synth_reads(g, 40_000, 150, 0.005, seed=52)
    k, l, cutoff = 31, 20, 2
    seq, ro, part = oracle_records(fastq(reads, 150), True, k, l)
    off, att, lab = O.bucketize(part)
    want = O.construct(k, l, off, att, lab)
    inp = C.Input.from_records(reads, np.arange(40_001, dtype=np.uint64) * np.uint64(150), k)
    rng = np.random.default_rng(5)
    with C.Handle(k, l, cutoff) as h:
        h.partition(inp)
        h.construct("learned")
        h.contract()
        whole = h.cdbg_device()
        nz = np.nonzero(h.bucket_stats()["n_records"])[0].astype(np.uint32)
        first = np.sort(rng.choice(nz, len(nz) // 2, replace=False))
        second = np.setdiff1d(nz, first)
        recs_all, labels_all = [], []
        for part_b in (first, second):
            for mode in ("batched", "learned"):
                h.construct_buckets(part_b, mode)
                t = h.tables()
                keep = np.isin(want["g"], part_b)
                assert set(np.unique(t["g"]).tolist()) <= set(part_b.tolist())
                np.testing.assert_array_equal(
                    rows_sorted([t["g"], t["keys"][:, 0], t["e_f"], t["status"]]),
                    rows_sorted([want["g"][keep], want["keys"][keep, 0], want["e_f"][keep], want["status"][keep]]))
            other = np.setdiff1d(nz, part_b)
            assert h.table_view(int(other[0])).capacity == 0
            h.contract()
            u = h.unitigs()
            recs_all += list(zip(u["g"].tolist(), u["flags"].tolist()))
            labels_all += u["labels"]
        got = C.cdbg_from_unitigs(k, [r[0] for r in recs_all], [r[1] for r in recs_all], labels_all)
        say(f"[case] batches: {len(got)} maximal unitigs from two batches, {len(whole)} from one construct")
        assert got == whole
