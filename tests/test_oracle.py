"""Pins the CPU oracle before it is trusted as the parity checker.

* wyhash / l-mer hash: golden vectors written by the reference's own wyhash.h.
* minimizer window: two-stack restatement vs brute force over all l-mers.
* partition: sequential restatement vs scan formulation, plus invariants that a
  wrong index or flag would break.
* construct: literal restatement vs closed-form (k+1)-mer counting vs a global
  brute force written here in Python, plus hand-computed homopolymer and
  palindrome cases.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}
CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def rc(s):
    return "".join(COMP[c] for c in reversed(s.upper()))


def lmer_val(s):
    x = 0
    for c in s.upper():
        x = (x << 2) | CODE[c]
    return x


def g_of(km1mer, l, G=16384):
    """Brute-force g of a (k-1)-mer in pure Python on top of the pinned hash."""
    best = None
    for i in range(len(km1mer) - l + 1):
        f = km1mer[i:i + l]
        h = min(O.lib().cfo_lmer_hash(lmer_val(f)), O.lib().cfo_lmer_hash(lmer_val(rc(f))))
        best = h if best is None else min(best, h)
    return best & (G - 1)


def rand_seq(rng, n):
    return "".join("ACGT"[b] for b in rng.integers(0, 4, n))


def decode_label(words, W, n):
    out = []
    for i in range(n):
        w = int(words[W - 1 - i // 32])
        out.append("ACGT"[(w >> (62 - 2 * (i % 32))) & 3])
    return "".join(out)


# --------------------------------------------------------------------------
def test_wyhash_golden_vectors():
    with open(os.path.join(HERE, "golden", "wyhash_vectors.json")) as f:
        v = json.load(f)
    L = O.lib()
    for lmer, h in v["lmer_hash_seed0"]:
        assert L.cfo_lmer_hash(int(lmer)) == int(h)
    for data, seed, h in v["wyhash_bytes"]:
        b = bytes.fromhex(data)
        assert L.cfo_wyhash(b, len(b), int(seed)) == int(h)


@pytest.mark.parametrize("k,l", [(31, 20), (63, 20), (127, 20), (15, 7), (21, 20), (33, 32)])
def test_minimizer_window_vs_brute(k, l):
    rng = np.random.default_rng(k * 100 + l)
    seqs = [rand_seq(rng, 3000), "A" * 400, "ACGT" * 150, (rand_seq(rng, 7) * 80)]
    for s in seqs:
        s = s.encode()
        a = O.minimizers(s, k - 1, l)
        b = O.minimizers(s, k - 1, l, brute=True)
        assert len(a) == len(s) - (k - 1) + 1
        np.testing.assert_array_equal(a, b)


def test_minimizer_hash_is_strand_symmetric():
    rng = np.random.default_rng(7)
    s = rand_seq(rng, 500)
    a = O.minimizers(s.encode(), 30, 20, brute=True)
    b = O.minimizers(rc(s).encode(), 30, 20, brute=True)
    np.testing.assert_array_equal(a, b[::-1])


def test_parse_and_fragments_closed_form():
    fa = b">a desc\nACGTN\nacgtACGTAC\n>b\n\nNNNNACGTACGTACGTAAA\n>c\nAC\n"
    seq, ro = O.parse(fa, fastq=False)
    assert seq == b"ACGTNacgtACGTACNNNNACGTACGTACGTAAAAC"
    assert list(ro) == [0, 15, 34, 36]
    beg, ln = O.fragments(seq, ro, 3)  # fragments: maximal ACGT runs >= k+1 = 4
    assert [(int(b), int(n)) for b, n in zip(beg, ln)] == [(0, 4), (5, 10), (19, 15)]
    beg, ln = O.fragments(seq, ro, 9)
    assert [(int(b), int(n)) for b, n in zip(beg, ln)] == [(5, 10), (19, 15)]
    fq = b"@r1\nACGTNACGT\n+\nIIIIIIIII\n@r2\nacgtt\n+\nIIIII\n"
    seq, ro = O.parse(fq, fastq=True)
    assert seq == b"ACGTNACGTacgtt" and list(ro) == [0, 9, 14]
    # '\r' is a character, not a line break: it splits fragments
    seq, ro = O.parse(b">x\r\nACGTACGT\r\nACGTACGT\n", fastq=False)
    assert seq == b"ACGTACGT\rACGTACGT"


def _partition_invariants(seq, ro, k, l, part):
    W = O.label_words(k, l)
    w = k - l
    beg, ln = O.fragments(seq, ro, k)
    n_sk = len(part["g"])
    lens = (part["att"] & 0xFF).astype(np.int64)
    ldisc = (part["att"] >> 8) & 1
    rdisc = (part["att"] >> 9) & 1
    assert ((part["att"] >> 10) == 0).all()
    assert (lens >= k).all() and (lens <= 2 * k - l).all()
    # labels decode to the input substring (first len bases only)
    for i in range(n_sk):
        p = int(part["pos"][i])
        assert decode_label(part["labels"][i], W, lens[i]) == seq[p:p + lens[i]].decode().upper()
    # every (k+1)-mer instance is covered by exactly one record whose g is the
    # g of its middle (k-1)-mer; every k-mer instance lies in 1 or 2 records
    cover = {}
    for i in range(n_sk):
        p = int(part["pos"][i])
        for q in range(p, p + lens[i] - k):
            cover.setdefault(q, []).append(i)
    n_edges = 0
    for b, n in zip(beg, ln):
        b, n = int(b), int(n)
        frag = seq[b:b + n].decode()
        gs = [g_of(frag[p:p + k - 1], l) for p in range(n - k + 2)]
        for q in range(n - k):
            assert len(cover[b + q]) == 1
            assert part["g"][cover[b + q][0]] == gs[q + 1]
            n_edges += 1
        # pieces: consecutive (k-1)-mers with the same g, at most w per piece
        # records of this fragment, in emission order: only the first is not
        # joined on the left, only the last is not joined on the right
        recs = [i for i in range(n_sk) if b <= part["pos"][i] < b + n]
        covered = 0
        for j, i in enumerate(recs):
            p = int(part["pos"][i]) - b
            lj, rj = int(j > 0), int(j + 1 < len(recs))
            first = p + lj                                   # first (k-1)-mer of the piece
            m = int(lens[i]) - lj - rj - (k - 2)             # (k-1)-mers in the piece
            assert 1 <= m <= w
            assert first == covered
            covered += m
            piece_g = {gs[x] for x in range(first, first + m)}
            assert piece_g == {int(part["g"][i])}
            assert bool(ldisc[i]) == (first > 0 and gs[first - 1] != gs[first])
            nxt = first + m
            assert bool(rdisc[i]) == (nxt < len(gs) and gs[nxt] != gs[first])
            # a piece is cut early only at a g change or after w (k-1)-mers of a run
            if nxt < len(gs) and gs[nxt] == gs[first]:
                run_start = first
                while run_start > 0 and gs[run_start - 1] == gs[first]:
                    run_start -= 1
                assert (nxt - run_start) % w == 0
        assert covered == len(gs)
    assert n_edges == sum(int(n) - k for n in ln)


@pytest.mark.parametrize("k,l", [(31, 20), (15, 7), (63, 20)])
def test_partition_invariants_and_two_formulations(k, l):
    rng = np.random.default_rng(k)
    s = rand_seq(rng, 1500) + "A" * 90 + rand_seq(rng, 300).lower() + "NN" + "CG" * 60 + rand_seq(rng, 400)
    recs = [s, rand_seq(rng, k), rand_seq(rng, k + 1), rand_seq(rng, 150)]
    fa = "".join(f">r{i}\n{r}\n" for i, r in enumerate(recs)).encode()
    seq, ro = O.parse(fa, fastq=False)
    p1 = O.partition(seq, ro, k, l)
    p2 = O.partition(seq, ro, k, l, scan_form=True)
    for key in ("g", "att", "labels", "pos"):
        np.testing.assert_array_equal(p1[key], p2[key])
    _partition_invariants(seq, ro, k, l, p1)


# --------------------------------------------------------------------------
def global_side_counts(seq, ro, k):
    """Brute force over all (k+1)-mer instances of all fragments, without any
    partitioning: count per (canonical vertex, side, base) slot using the
    reference's per-instance rule (an instance updates its prefix k-mer's succ
    side and, unless it is a self-loop, its suffix k-mer's pred side)."""
    beg, ln = O.fragments(seq, ro, k)
    cnt = {}
    verts = set()

    def bump(v, side, base):
        cnt[(v, side, base)] = cnt.get((v, side, base), 0) + 1

    for b, n in zip(beg, ln):
        frag = seq[int(b):int(b) + int(n)].decode().upper()
        for i in range(len(frag) - k + 1):
            x = frag[i:i + k]
            verts.add(min(x, rc(x)))
        for i in range(len(frag) - k):
            u, v = frag[i:i + k], frag[i + 1:i + k + 1]
            uc, vc = min(u, rc(u)), min(v, rc(v))
            succ, pred = frag[i + k], frag[i]
            bump(uc, 1, succ) if u == uc else bump(uc, 0, COMP[succ])
            if uc != vc:
                bump(vc, 0, pred) if v == vc else bump(vc, 1, COMP[pred])
    return cnt, verts


def key_str(words, k):
    x = 0
    for i in range(3, -1, -1):
        x = (x << 64) | int(words[i])
    return "".join("ACGT"[(x >> (2 * (k - 1 - i))) & 3] for i in range(k))


def _stitch_check(seq, ro, k, l, states):
    cnt, verts = global_side_counts(seq, ro, k)
    seen = set()
    for g, kw, ef, st in zip(states["g"], states["keys"], states["e_f"], states["status"]):
        v = key_str(kw, k)
        seen.add(v)
        gf, gb = g_of(v[:k - 1], l), g_of(v[1:], l)   # front / back (k-1)-mers
        assert g in (gf, gb)
        for side, gside in ((0, gf), (1, gb)):
            disc = bool(st & (2 if side == 0 else 4))
            field = [(int(ef) >> ((0 if side == 0 else 16) + 4 * b)) & 0xF for b in range(4)]
            if gside == g:
                assert not disc
                assert field == [min(15, cnt.get((v, side, c), 0)) for c in "ACGT"]
            else:
                assert disc and field == [0, 0, 0, 0]
        assert (st & 1) == 0
    assert seen == verts


@pytest.mark.parametrize("k,l", [(31, 20), (15, 7), (63, 20), (95, 20), (127, 20)])
def test_construct_three_ways(k, l):
    rng = np.random.default_rng(100 + k)
    base = rand_seq(rng, 800)
    # repeats (counts up to > 15), a homopolymer in both orientations, a tandem
    # repeat and a palindrome give saturation, self-loops and orientation splits
    pal = "AACTGACATGTCAGTT"
    s = base + base[:300] + "A" * 40 + rand_seq(rng, 50) + "T" * 45 + ("CAG" * 30) + pal * 3
    recs = [s] + [base[100:400]] * 18 + [rc(base[200:500])] * 3
    fa = "".join(f">r{i}\n{r}\n" for i, r in enumerate(recs)).encode()
    seq, ro = O.parse(fa, fastq=False)
    part = O.partition(seq, ro, k, l)
    off, att, lab = O.bucketize(part)
    lit = O.construct(k, l, off, att, lab, mode=0)
    cf = O.construct(k, l, off, att, lab, mode=1)
    flat = O.construct(k, l, off, att, lab, mode=2)  # the CPU baseline's open-addressing map
    for key in ("g", "keys", "e_f", "status"):
        np.testing.assert_array_equal(lit[key], cf[key])
        np.testing.assert_array_equal(lit[key], flat[key])
    _stitch_check(seq, ro, k, l, lit)


def test_homopolymer_orientation_split_hand_case():
    # A^{k+3} has three A^{k+1} instances -> A^k.back[A] = 3; its reverse
    # complement T^{k+3} puts 3 on A^k.front[A].  A canonical-(k+1)-mer counter
    # would give 6 on both sides.
    k, l = 15, 7
    for s, want in (("A" * (k + 3), (0, 3)), ("T" * (k + 3), (3, 0)), ("A" * (k + 3) + "N" + "T" * (k + 3), (3, 3))):
        seq, ro = O.parse(f">x\n{s}\n".encode(), fastq=False)
        part = O.partition(seq, ro, k, l)
        off, att, lab = O.bucketize(part)
        st = O.construct(k, l, off, att, lab)
        assert len(st["g"]) == 1 and key_str(st["keys"][0], k) == "A" * k
        ef = int(st["e_f"][0])
        assert (ef & 0xF, (ef >> 16) & 0xF) == want
        assert ef & ~0x000F000F == 0


def test_palindrome_fixture_single_slot():
    # tests/golden/ref_data/rob-palindrome.fa (the reference's data/): a 16-base
    # reverse-complement palindrome; at k = 15 its only (k+1)-mer is a self-loop
    # between the two 15-mers, which are each other's reverse complement.
    with open(os.path.join(HERE, "golden", "ref_data", "rob-palindrome.fa"), "rb") as f:
        seq, ro = O.parse(f.read(), fastq=False)
    k, l = 15, 7
    part = O.partition(seq, ro, k, l)
    off, att, lab = O.bucketize(part)
    st = O.construct(k, l, off, att, lab)
    assert len(st["g"]) == 1
    v = key_str(st["keys"][0], k)
    assert v == min("AACTGACATGTCAGT", rc("AACTGACATGTCAGT"))
    ef = int(st["e_f"][0])
    # exactly one saturating counter is 1, all others 0
    fields = [(ef >> (4 * i)) & 0xF for i in range(8)]
    assert sorted(fields) == [0] * 7 + [1]
    # and it is the succ side of the observed prefix k-mer
    u = "AACTGACATGTCAGT"
    if u == v:
        assert (ef >> (16 + 4 * CODE["T"])) & 0xF == 1
    else:
        assert (ef >> (4 * CODE[COMP["T"]])) & 0xF == 1


def test_edge_at_cutoff():
    # f_th semantics (State_Config.hpp:220-252): E / single base / N
    ef = (2 << 0) | (1 << 4) | (15 << 16) | (15 << 20)
    assert O.edge_at(ef, 0, 1) == 4          # A(2), C(1) >= 1 -> N
    assert O.edge_at(ef, 0, 2) == 0          # only A -> A
    assert O.edge_at(ef, 0, 3) == 5          # none -> E
    assert O.edge_at(ef, 1, 15) == 4
    assert O.edge_at(0, 1, 1) == 5


@pytest.mark.parametrize("k", [31, 63])
def test_bucket_checksums_pin_record_multisets(k):
    # The full-size tests compare per-bucket checksums of records: the oracle's
    # streaming partition checksum must equal the checksum of its own records,
    # ignore their order, and change when one label bit changes.
    import cfgpu as C
    g = C.synth_genome(100000, 0.1, 300, 2000, 0.02, seed=5)
    reads = C.synth_reads(g, 6000, 150, 0.01, seed=6)
    ro = np.arange(6001, dtype=np.uint64) * np.uint64(150)
    part = O.partition(reads, ro, k, 20)
    W = O.label_words(k, 20)
    att = part["att"].astype(np.uint32) | (part["g"].astype(np.uint32) << 16)
    s1, c1 = O.partition_hash(reads, ro, k, 20, threads=3)
    s2, c2 = O.records_hash(att, part["labels"], W, threads=2)
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_array_equal(c1, c2)
    assert int(c1.sum()) == len(att)
    p = np.random.default_rng(1).permutation(len(att))
    s3, _ = O.records_hash(att[p], part["labels"][p], W)
    np.testing.assert_array_equal(s2, s3)
    bad = part["labels"].copy()
    bad[7, 0] ^= np.uint64(4)
    s4, _ = O.records_hash(att, bad, W)
    assert int((s4 != s2).sum()) == 1
    badatt = att.copy()
    badatt[9] ^= np.uint32(1 << 9)  # a flipped r_disc bit
    s5, _ = O.records_hash(badatt, part["labels"], W)
    assert int((s5 != s2).sum()) == 1
