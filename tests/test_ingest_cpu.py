"""Host ingest without a GPU: the parallel parse of FASTA / FASTQ text, its 2-bit
packing and the chunk boundaries of the pipelined ingest, against a plain
sequential restatement of the reference's record rules (RabbitFX
include/RabbitFX/io/Formater.cpp:87-130 FASTA, :238-263 FASTQ; fragments =
maximal A/C/G/T runs of length >= k+1, src/Graph_Partitioner.cpp:476-501)."""
import random

import numpy as np
import pytest

import cfgpu as C

K = 21


def records_seq(text: bytes, fastq: bool):
    """Sequential parse: '\\n' ends lines; FASTQ sequence = line 2 of every 4;
    FASTA record = the lines after a '>' line, concatenated."""
    lines, pos = [], 0
    while pos < len(text):
        e = text.find(b"\n", pos)
        e = len(text) if e < 0 else e
        lines.append(text[pos:e])
        pos = e + 1
    if fastq:
        return [lines[i] for i in range(1, len(lines), 4)]
    recs, cur = [], None
    for ln in lines:
        if ln[:1] == b">":
            if cur is not None:
                recs.append(b"".join(cur))
            cur = []
        elif cur is not None:
            cur.append(ln)
    if cur is not None:
        recs.append(b"".join(cur))
    return recs


def fragments(recs, k):
    out = []
    for r in recs:
        i = 0
        while i < len(r):
            while i < len(r) and chr(r[i]).upper() not in "ACGT":
                i += 1
            s = i
            while i < len(r) and chr(r[i]).upper() in "ACGT":
                i += 1
            if i - s >= k + 1:
                out.append(r[s:i].upper())
    return out


def decode(inp):
    codes = []
    for w in inp.packed[: (inp.n_bases + 31) // 32]:
        w = int(w)
        codes.extend((w >> (62 - 2 * j)) & 3 for j in range(32))
    s = bytes(b"ACGT"[c] for c in codes[: inp.n_bases])
    off = inp.frag_off.astype(np.int64)
    return [s[off[i]:off[i + 1]] for i in range(inp.n_frag)]


def fastq_text(rng, n_rec):
    out = []
    for i in range(n_rec):
        L = rng.choice([5, 20, 22, 60, 150, 151])
        seq = "".join(rng.choice("ACGTACGTACGTNacgt") for _ in range(L))
        qual = "".join(rng.choice("@+I#5") for _ in range(L))  # lines that start with '@' or '+'
        out.append(f"@r{i} x\n{seq}\n+{'' if i % 3 else 'r' + str(i)}\n{qual}\n")
    return "".join(out).encode()


def fasta_text(rng, n_rec):
    out = ["junk before the first header\nACGTACGTACGTACGTACGTACGTACGT\n"]
    for i in range(n_rec):
        out.append(f">rec{i}\n")
        for _ in range(rng.choice([0, 1, 3, 40])):  # empty records and long ones
            ln = "".join(rng.choice("ACGTACGTNacgt") for _ in range(rng.choice([0, 30, 80])))
            if rng.random() < 0.1:
                ln = ln[:10] + ">" + ln[10:]  # '>' inside a line is not a header
            out.append(ln + "\n")
    return "".join(out).encode()


@pytest.mark.parametrize("fastq", [True, False])
@pytest.mark.parametrize("trail", [True, False])
def test_parallel_parse_equals_sequential(fastq, trail):
    rng = random.Random(7 + fastq + 2 * trail)
    text = fastq_text(rng, 6000) if fastq else fasta_text(rng, 1800)
    if not trail:
        text = text.rstrip(b"\n")
    assert len(text) > 8 << 16  # every one of 8 threads gets a share
    want = fragments(records_seq(text, fastq), K)
    for threads in (1, 3, 8):
        inp = C.Input.from_buffer(text, fastq, K, threads=threads)
        assert inp.n_records == len(records_seq(text, fastq))
        assert decode(inp) == want, threads
        assert int(inp.packed[-1]) == 0 and int(inp.packed[-2]) == 0


@pytest.mark.parametrize("fastq", [True, False])
def test_chunks_parse_to_the_whole(fastq):
    rng = random.Random(11 + fastq)
    text = fastq_text(rng, 4000) if fastq else fasta_text(rng, 600)
    want = fragments(records_seq(text, fastq), K)
    for chunk in (1, 997, 1 << 14, 1 << 30):
        off = C.split_records(text, fastq, chunk, threads=4).astype(np.int64)
        assert off[0] == 0 and off[-1] == len(text) and (np.diff(off) > 0).all()
        for o in off[1:-1]:
            assert text[o - 1:o] == b"\n"
            if not fastq:
                assert text[o:o + 1] == b">"
            else:
                assert text[:o].count(b"\n") % 4 == 0
        got = []
        for a, b in zip(off[:-1], off[1:]):
            got += decode(C.Input.from_buffer(text[a:b], fastq, K, threads=2))
        assert got == want, chunk


def test_split_records_edge_cases():
    assert list(C.split_records(b"", True, 100)) == [0, 0]
    one = b">only\n" + b"ACGT" * 1000 + b"\n"
    assert list(C.split_records(one, False, 10)) == [0, len(one)]  # one record: one chunk
