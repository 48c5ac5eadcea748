"""GPU parity: the HIP partition and construct, called through the C ABI,
against the CPU oracle on the same seeded inputs.

Bar: bit-exact.  Partition: the per-bucket multiset of records (attribute pack
and label) equals the oracle's.  Construct: the sorted (g, canonical k-mer,
Edge_Frequency word, status) rows equal the oracle's literal restatement of
Subgraph::construct, and the exported DFA byte equals edge_at/disc under the
cutoff.  All three construct modes must agree."""
import ctypes
import sys

import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rows_sorted(cols):
    m = np.stack([np.asarray(c, dtype=np.uint64) for c in cols], axis=1)
    order = np.lexsort(m.T[::-1])
    return m[order]


def oracle_records(buf, fastq, k, l):
    seq, ro = O.parse(buf, fastq)
    part = O.partition(seq, ro, k, l)
    return seq, ro, part


def check_records(h, part, k, l):
    att, lab, off = h.records()
    W = O.label_words(k, l)
    n = len(part["g"])
    assert len(att) == n, (len(att), n)
    g_gpu = (att >> 16).astype(np.uint64)
    # records sit in their own bucket
    if n:
        bucket_of = np.repeat(np.arange(C.G), np.diff(off.astype(np.int64)))
        np.testing.assert_array_equal(bucket_of, g_gpu)
    gpu = rows_sorted([g_gpu, att & 0xFFFF] + [lab[:, j] for j in range(W)])
    ora = rows_sorted([part["g"], part["att"]] + [part["labels"][:, j] for j in range(W)])
    np.testing.assert_array_equal(gpu, ora)
    return att, lab, off


def dfa_from(ef, st, cutoff):
    def side(off):
        f = np.stack([(ef >> (off + 4 * b)) & 0xF for b in range(4)], axis=1) >= cutoff
        c = f.sum(axis=1)
        which = np.argmax(f, axis=1)
        return np.where(c == 0, 5, np.where(c == 1, which, 4)).astype(np.uint32)
    return (side(0) | (side(16) << 3) | ((st.astype(np.uint32) & 2) << 5) |
            ((st.astype(np.uint32) & 4) << 5)).astype(np.uint8)


def check_tables(h, part, k, l, cutoff, want=None):
    kw = (k + 31) // 32
    t = h.tables()
    if want is None:
        off, att, lab = O.bucketize(part)
        want = O.construct(k, l, off, att, lab, mode=0)
    n = len(want["g"])
    assert len(t["g"]) == n, (len(t["g"]), n)
    gpu = rows_sorted([t["g"]] + [t["keys"][:, j] for j in range(kw - 1, -1, -1)] + [t["e_f"], t["status"], t["dfa"]])
    ora = rows_sorted([want["g"]] + [want["keys"][:, j] for j in range(kw - 1, -1, -1)] +
                      [want["e_f"], want["status"], dfa_from(want["e_f"], want["status"], cutoff)])
    np.testing.assert_array_equal(gpu, ora)
    stats = h.bucket_stats()
    nz = stats["n_records"] > 0
    assert ((stats["flags"][nz] & 1) == 1).all()
    np.testing.assert_array_equal(np.bincount(want["g"], minlength=C.G)[nz], stats["n_vertices"][nz])
    return want


def run_case(name, buf, fastq, k, l, cutoff, modes=("plain", "batched", "learned")):
    say(f"[case] {name}: k={k} l={l} cutoff={cutoff} bytes={len(buf)}")
    seq, ro, part = oracle_records(buf, fastq, k, l)
    inp = C.Input.from_buffer(buf, fastq, k)
    want = None
    with C.Handle(k, l, cutoff) as h:
        say(f"[case] {name}: partition n_bases={inp.n_bases} n_frag={inp.n_frag}")
        h.partition(inp)
        check_records(h, part, k, l)
        rs = h.run_stats()
        assert rs["n_records"] == len(part["g"])
        lens = (part["att"] & 0xFF).astype(np.int64)
        assert rs["n_kmers"] == int((lens - (k - 1)).sum())
        beg, ln = O.fragments(seq, ro, k)
        assert rs["n_kmers"] == int((ln.astype(np.int64) - k + 1).sum()) + int((lens > 0).sum()) - len(ln)
        for mode in modes:
            say(f"[case] {name}: construct mode={mode}")
            h.construct(mode)
            want = check_tables(h, part, k, l, cutoff, want)
    return want


def fasta(recs, width=80):
    return b"".join(b">r%d\n%s\n" % (i, b"\n".join(r[j:j + width] for j in range(0, len(r), width)))
                    for i, r in enumerate(recs))


def fastq(reads, read_len):
    n = len(reads) // read_len
    q = b"I" * read_len
    return b"".join(b"@q%d\n%s\n+\n%s\n" % (i, reads[i * read_len:(i + 1) * read_len], q) for i in range(n))


# --------------------------------------------------------------------------
def test_reference_mode_repeats_k31():
    base = C.synth_genome(200000, 0.0, 300, 6000, 0.0, seed=11)
    g = C.synth_insert_repeats(base, 40, 1000, 0.01, seed=12)
    run_case("ref-repeats-k31", fasta([g]), False, 31, 20, 1)


def test_read_mode_k31_cutoff2():
    g = C.synth_genome(60000, 0.1, 300, 2000, 0.02, seed=21)
    reads = C.synth_reads(g, 8000, 150, 0.005, seed=22)
    run_case("reads-k31-c2", fastq(reads, 150), True, 31, 20, 2)


def test_read_mode_k63_cutoff2():
    g = C.synth_genome(40000, 0.2, 300, 2000, 0.01, seed=31)
    reads = C.synth_reads(g, 5000, 150, 0.01, seed=32)
    run_case("reads-k63-c2", fastq(reads, 150), True, 63, 20, 2)


def test_multi_reference_k127():
    base = C.synth_genome(30000, 0.0, 300, 6000, 0.0, seed=41)
    gens = [C.synth_mutate(base, 0.005, 0.01, 10, seed=42 + i) for i in range(4)]
    run_case("multiref-k127", fasta(gens), False, 127, 20, 1)


@pytest.mark.parametrize("k,l", [(45, 20), (95, 24), (21, 11)])
def test_other_word_counts(k, l):
    g = C.synth_genome(20000, 0.2, 200, 1500, 0.02, seed=k)
    reads = C.synth_reads(g, 2000, 150, 0.01, seed=k + 1)
    run_case(f"reads-k{k}", fastq(reads, 150), True, k, l, 2)


def test_edge_cases_homopolymers_tandem_lowercase_N():
    rng = np.random.default_rng(5)
    r = bytes(rng.choice(list(b"ACGT"), 3000))
    recs = [r[:1000] + b"A" * 5000 + r[1000:1500] + b"T" * 400 + (b"CAG" * 300) + b"N" * 5 + r[1500:].lower(),
            b"G" * 40, r[:31], r[:32], b"AACTGACATGTCAGTT" * 20, b"NNNN", b""]
    run_case("edges-k31", fasta(recs), False, 31, 20, 1)
    run_case("edges-k15", fasta(recs), False, 15, 7, 3)


def test_long_homopolymer_run_crosses_many_tiles():
    # one g-run longer than 64 partition tiles: the carry-in look-back takes
    # more than one step
    g = C.synth_genome(2000, 0.0, 300, 6000, 0.0, seed=3)
    run_case("long-run", fasta([g + b"A" * 150000 + g]), False, 31, 20, 1, modes=("learned",))


def test_reference_fixtures():
    import os
    d = os.path.join(os.path.dirname(__file__), "golden", "ref_data")
    for name, fq, k, l in (("rob-palindrome.fa", False, 15, 7), ("refs2.fa", False, 9, 5),
                           ("reads.fq", True, 5, 3), ("cycles.fa", False, 3, 2)):
        with open(os.path.join(d, name), "rb") as f:
            run_case(name, f.read(), fq, k, l, 1)


def test_empty_input():
    say("[case] empty")
    inp = C.Input.from_buffer(b">x\nNNNN\n", False, 31)
    assert inp.n_frag == 0
    with C.Handle(31, 20, 1) as h:
        h.partition(inp)
        for mode in ("plain", "batched", "learned"):
            h.construct(mode)
            t = h.tables()
            assert len(t["g"]) == 0


def test_overflow_fallback_is_exact():
    # margin < 1 under-sizes the learned tables on purpose: every overflow is
    # detected and the bucket rebuilt, so the result cannot change.
    g = C.synth_genome(80000, 0.1, 300, 2000, 0.02, seed=51)
    reads = C.synth_reads(g, 10000, 150, 0.005, seed=52)
    buf = fastq(reads, 150)
    k, l = 31, 20
    seq, ro, part = oracle_records(buf, True, k, l)
    inp = C.Input.from_buffer(buf, True, k)
    say("[case] overflow-fallback")
    with C.Handle(k, l, 2) as h:
        h.partition(inp)
        # global-atomic tables only (every LDS pass table of these small
        # buckets would have room whatever the estimate)
        h.set_tuning(margin=0.25, load_factor=0.8, max_passes=0, predictor=1)
        h.construct("learned")
        rs = h.run_stats()
        assert rs["predictor"] == 1 and np.isfinite(rs["mlp_loss"]) and rs["t_train_ms"] > 0
        say(f"[case] overflowed={rs['n_overflowed']} rounds={rs['n_fallback_rounds']}")
        assert rs["n_overflowed"] > 0 and rs["n_fallback_rounds"] >= 1
        assert rs["n_lds_buckets"] == 0
        check_tables(h, part, k, l, 2)
        stats = h.bucket_stats()
        assert ((stats["flags"] & 2) > 0).sum() == rs["n_overflowed"]


def test_lds_overflow_fallback_is_exact():
    # Large Zipf buckets with under-estimated distinct counts get too few LDS
    # passes; the full pass table is detected and the bucket rebuilt on the
    # global path.
    # (120,000 records: the second and third buckets, outside the dedup probe,
    # hold more distinct records than a dedup pass table takes)
    k, l = 31, 20
    att, lab, off = C.synth_bucket_workload(256, 120000, 1.1, k, l, seed=7)
    want = O.construct(k, l, off, (att & 0xFFFF).astype(np.uint16), lab)
    say("[case] lds-overflow")
    with C.Handle(k, l, 1) as h:
        h.load_buckets(att, lab, off)
        h.set_tuning(margin=0.2, max_passes=4096, predictor=2, dedup=2)
        h.construct("learned")
        rs = h.run_stats()
        assert rs["predictor"] == 2
        # random labels rarely repeat, and dedup passes sized at 20 % overflow
        # for the large buckets, which keep their records (copy fallback)
        say(f"[case] lds-overflow dedup fallback={rs['n_dedup_fallback']} walked={rs['n_records_walked']}")
        assert rs["n_dedup_fallback"] > 0 and rs["n_records_walked"] <= len(att)
        say(f"[case] lds-overflow overflowed={rs['n_overflowed']} lds={rs['n_lds_buckets']}")
        assert rs["n_lds_buckets"] > 0 and rs["n_overflowed"] > 0 and rs["n_fallback_rounds"] >= 1
        check_tables(h, None, k, l, 1, want)
        stats = h.bucket_stats()
        assert ((stats["flags"] & 2) > 0).sum() == rs["n_overflowed"]


def test_repeat_calls_are_idempotent_and_release():
    g = C.synth_genome(50000, 0.2, 300, 2000, 0.02, seed=61)
    reads = C.synth_reads(g, 6000, 150, 0.01, seed=62)
    inp = C.Input.from_buffer(fastq(reads, 150), True, 31)
    say("[case] idempotent")
    with C.Handle(31, 20, 2) as h:
        h.partition(inp)
        h.construct("learned")
        a = h.tables()
        h.partition(inp)
        h.construct("learned")
        b = h.tables()
        np.testing.assert_array_equal(rows_sorted([a["g"], a["keys"][:, 0], a["e_f"], a["status"]]),
                                      rows_sorted([b["g"], b["keys"][:, 0], b["e_f"], b["status"]]))
        nz = np.nonzero(h.bucket_stats()["n_records"])[0]
        v = h.table_view(int(nz[0]))
        assert v.capacity >= v.n_vertices > 0 and v.key_words == 1 and v.entry_bytes == 16
        # the view's device entries: vertices or empty, the vertices those of table copy
        ent = np.zeros(2 * v.capacity, np.uint64)
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(ent.ctypes.data, v.entries, ent.nbytes, 2) == 0
        keys = ent[0::2]
        occ = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
        assert int(occ.sum()) == v.n_vertices
        one = h.tables(buckets=[int(nz[0])])
        np.testing.assert_array_equal(np.sort(keys[occ]), np.sort(one["keys"][:, 0]))
        h.release(int(nz[0]))
        assert h.table_view(int(nz[0])).capacity == 0
        assert len(h.tables(buckets=[int(nz[0])])["g"]) == 0
        assert len(h.tables(buckets=[int(nz[1])])["g"]) > 0


@pytest.mark.parametrize("k,l", [(31, 20), (37, 20), (21, 11)])
def test_record_dedup_is_exact(k, l):
    # High-coverage reads repeat most weak super k-mers: the deduplicated walk
    # (always / auto) must give the same tables as walking every record.
    # k=31, l=20: records of at most 42 bases leave at most 10 in the label's
    # last word, so the dedup key packs len, disc bits and hash bits into it;
    # k=37 (22 bases there, 2 words) and k=21, l=11 (31 bases, 1 word) use the
    # key with a separate length / hash word.
    g = C.synth_genome(40000, 0.1, 300, 2000, 0.02, seed=71)
    reads = C.synth_reads(g, 20000, 150, 0.002, seed=72)
    buf = fastq(reads, 150)
    seq, ro, part = oracle_records(buf, True, k, l)
    inp = C.Input.from_buffer(buf, True, k)
    say("[case] dedup")
    with C.Handle(k, l, 2) as h:
        h.partition(inp)
        want = None
        for dd in (0, 2, 1):
            h.set_tuning(dedup=dd)
            for mode in ("batched", "learned"):
                say(f"[case] dedup={dd} mode={mode}")
                h.construct(mode)
                want = check_tables(h, part, k, l, 2, want)
                rs = h.run_stats()
                say(f"[case] dedup k={k} walked {rs['n_records_walked']} of {rs['n_records']}")
                if dd == 0:
                    assert rs["n_records_walked"] == rs["n_records"]
                else:
                    assert rs["n_records_walked"] < rs["n_records"] / 2, rs["n_records_walked"]


def test_zipf_bucket_workload_loaded_records():
    # Zipf(1.1)-sized buckets loaded directly (the reader side of the bucket
    # format): the persistent queue must split the giant buckets and still
    # produce the oracle's states.
    for k, l in ((31, 20), (63, 20)):
        att, lab, off = C.synth_bucket_workload(512, 60000, 1.1, k, l, seed=k)
        say(f"[case] zipf-loaded k={k} records={len(att)} largest={int(np.diff(off.astype(np.int64)).max())}")
        want = O.construct(k, l, off, (att & 0xFFFF).astype(np.uint16), lab)
        with C.Handle(k, l, 1) as h:
            h.load_buckets(att, lab, off)
            h.construct("plain")
            check_tables(h, None, k, l, 1, want)
            # max_passes routes buckets between the LDS kernel and the
            # global-atomic kernel: all LDS, mixed, all global
            for mp in (32, 2, 0):
                h.set_tuning(max_passes=mp)
                for mode in ("batched", "learned"):
                    say(f"[case] zipf-loaded k={k} mode={mode} max_passes={mp}")
                    h.construct(mode)
                    check_tables(h, None, k, l, 1, want)
                    rs = h.run_stats()
                    if mode == "learned":
                        assert rs["predictor"] in (1, 2) and rs["n_samples"] >= 8
                    if mp == 0:
                        assert rs["n_lds_buckets"] == 0
                    if mp == 32 and mode == "batched":
                        assert rs["n_lds_buckets"] > 0
                        # multi-pass buckets walk their records once (per-pass sub-lists)
                        assert rs["n_lds_all_pass"] > 0


LDS_CAP = {1: 12544, 2: 7744, 3: 5632, 4: 4416}  # LDS pass-table slots per key-word count


@pytest.mark.parametrize("k", [31, 63, 127])
def test_many_pass_buckets_walk_once(k):
    # Buckets of 9..64 LDS passes (the read sets of deep coverage give 20-30 at
    # k = 63): one item walks the records once and spills every instance to
    # its pass's sub-list.  Random labels make nearly every instance a
    # distinct vertex, so the passes are as many as the instances need.
    l = 20
    att, lab, off = C.synth_bucket_workload(16, {31: 150000, 63: 60000, 127: 30000}[k], 1.1, k, l, seed=300 + k)
    want = O.construct(k, l, off, (att & 0xFFFF).astype(np.uint16), lab)
    cap = LDS_CAP[(k + 31) // 32] * 0.8
    with C.Handle(k, l, 1) as h:
        h.load_buckets(att, lab, off)
        km = h.bucket_stats()["n_kmers"].astype(np.float64)
        passes = np.ceil(km / cap)
        say(f"[case] many-pass k={k} passes={sorted(passes[passes > 0].astype(int).tolist())}")
        assert ((passes > 8) & (passes <= 64)).sum() >= 2
        for mode in ("batched", "learned"):
            say(f"[case] many-pass k={k} mode={mode}")
            h.construct(mode)
            check_tables(h, None, k, l, 1, want)
            rs = h.run_stats()
            assert rs["n_overflowed"] == 0
            if mode == "batched":
                assert rs["n_lds_all_pass"] >= int(((passes >= 2) & (passes <= 64)).sum())


def test_autotune_keeps_results_and_keeps_the_fastest():
    # Zipf buckets with max_passes 1: the largest take the global-atomic kernel,
    # so every part of the search runs (record dedup auto / on, then LDS load x
    # margin and global-kernel shapes around the faster of the two).
    # Every trial must build the same vertex counts, the kept tuning must be the
    # fastest trial's, and the final tables must equal the oracle's.
    k, l = 31, 20
    att, lab, off = C.synth_bucket_workload(512, 80000, 1.1, k, l, seed=17)
    want = O.construct(k, l, off, (att & 0xFFFF).astype(np.uint16), lab)
    with C.Handle(k, l, 1) as h:
        h.load_buckets(att, lab, off)
        h.set_tuning(max_passes=1)
        trials = h.autotune("learned")
        say(f"[case] autotune: {len(trials)} trials, fastest {min(t['ms'] for t in trials):.3f} ms")
        assert len(trials) == 2 + 9 + 16                 # dedup auto / on, then the grids
        assert [t["dedup"] for t in trials[:2]] == [1, 2]
        assert all(t["identical"] for t in trials)
        best = min(trials, key=lambda t: t["ms"])
        tun = h.tuning()
        for key in ("dedup", "lds_load", "margin", "load_factor", "workgroup", "chunk_records", "wgs_per_cu"):
            assert tun[key] == pytest.approx(best[key]), key
        rs = h.run_stats()
        assert rs["n_lds_buckets"] + rs["n_samples"] < rs["n_buckets_nonempty"]  # main-round buckets went global
        check_tables(h, None, k, l, 1, want)


def test_load_buckets_rejects_malformed_records():
    att, lab, off = C.synth_bucket_workload(8, 100, 1.1, 31, 20, seed=1)
    with C.Handle(31, 20, 1) as h:
        bad = att.copy()
        bad[5] = (bad[5] & 0xFFFF) | (9 << 16)         # wrong bucket id
        with pytest.raises(C.CfgpuError):
            h.load_buckets(bad, lab, off)
        bad = att.copy()
        bad[7] = (int(bad[7]) & 0xFFFFFF00) | 30         # shorter than k
        with pytest.raises(C.CfgpuError):
            h.load_buckets(bad, lab, off)
        with pytest.raises(C.CfgpuError):
            h.construct("learned")                       # nothing loaded
        with pytest.raises(C.CfgpuError):
            h.set_tuning(predictor=3)                    # no such predictor
