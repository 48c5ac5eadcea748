"""Chunked input on the device: cfgpu_partition with CFGPU_PARTITION_APPEND +
cfgpu_partition_finish, and the pipelined host ingest (cfhost_partition_buffer /
_file: parse + pack of chunk i+1 overlapped with chunk i's copy and partition),
against one cfgpu_partition of the whole input.  Records compare as per-bucket
multisets (their order inside a bucket is unspecified, as in the reference,
whose partition workers append chunk by chunk)."""
import numpy as np
import pytest

import cfgpu as C
from test_gpu_parity import fastq, rows_sorted, say

pytestmark = pytest.mark.gpu

K, L = 31, 20


def fasta(reads: bytes, read_len: int, per_record: int = 40) -> bytes:
    out = []
    n = len(reads) // read_len
    for r in range(0, n, per_record):
        out.append(b">r%d\n" % r)
        block = reads[r * read_len:(r + per_record) * read_len]
        out += [block[i:i + 70] + b"\n" for i in range(0, len(block), 70)]
    return b"".join(out)


def records_rows(h):
    att, lab, off = h.records()
    return rows_sorted([att] + [lab[:, w] for w in range(lab.shape[1])]), off


def bucket_counts(h):
    s = h.bucket_stats()
    return np.stack([s["n_records"], s["n_kmers"], s["n_bases"]])


@pytest.fixture(scope="module")
def text():
    g = C.synth_genome(2_000_000, 0.2, 300, 2000, 0.02, seed=91)
    reads = C.synth_reads(g, 100_000, 150, 0.01, seed=92)
    return {"fastq": fastq(reads, 150), "fasta": fasta(reads, 150)}


@pytest.mark.parametrize("fmt", ["fastq", "fasta"])
def test_appended_chunks_equal_one_partition(text, fmt):
    buf, fq = text[fmt], fmt == "fastq"
    say(f"[case] append+finish {fmt}: {len(buf)} bytes")
    with C.Handle(K, L, 2) as h:
        h.partition(C.Input.from_buffer(buf, fq, K))
        want, want_off = records_rows(h)
        want_cnt = bucket_counts(h)
        off = C.split_records(buf, fq, 700_000).astype(np.int64)
        assert len(off) > 5
        for a, b in zip(off[:-1], off[1:]):
            h.partition(C.Input.from_buffer(buf[a:b], fq, K), append=True)
        with pytest.raises(C.CfgpuError):
            h.construct("learned")  # no buckets until the finish
        h.partition_finish()
        assert h.run_stats()["n_partition_chunks"] == len(off) - 1
        got, got_off = records_rows(h)
        np.testing.assert_array_equal(got_off, want_off)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(bucket_counts(h), want_cnt)


@pytest.mark.parametrize("fmt", ["fastq", "fasta"])
def test_pipelined_ingest_equals_one_partition(text, fmt, tmp_path):
    buf, fq = text[fmt], fmt == "fastq"
    with C.Handle(K, L, 2) as h:
        h.partition(C.Input.from_buffer(buf, fq, K))
        want, want_off = records_rows(h)
        h.construct("learned")
        t_want = h.tables()
        say(f"[case] pipelined ingest {fmt}: buffer, 256 KiB chunks")
        st = h.ingest_buffer(buf, fq, threads=4, chunk_bytes=256 << 10)
        assert st["n_chunks"] > 10 and st["n_bytes"] == len(buf)
        got, got_off = records_rows(h)
        np.testing.assert_array_equal(got_off, want_off)
        np.testing.assert_array_equal(got, want)
        h.construct("learned")
        t_got = h.tables()
        cols = lambda t: rows_sorted([t["g"], t["keys"][:, 0], t["e_f"], t["status"]])
        np.testing.assert_array_equal(cols(t_got), cols(t_want))
        say(f"[case] pipelined ingest {fmt}: file, one chunk")
        p = tmp_path / f"in.{fmt}"
        p.write_bytes(buf)
        st = h.ingest_file(str(p), fq)
        assert st["n_chunks"] == 1
        got, _ = records_rows(h)
        np.testing.assert_array_equal(got, want)


def test_append_edge_cases():
    with C.Handle(K, L, 1) as h:
        with pytest.raises(C.CfgpuError):
            h.partition_finish()  # nothing appended
        empty = C.Input.from_buffer(b"", True, K)
        h.partition(empty, append=True)
        h.partition(C.Input.from_buffer(b"@a\n" + b"ACGT" * 20 + b"\n+\n" + b"I" * 80 + b"\n", True, K), append=True)
        h.partition(empty, append=True)
        h.partition_finish()
        assert h.run_stats()["n_records"] > 0 and h.run_stats()["n_partition_chunks"] == 3
        # a partition without APPEND drops appended chunks
        h.partition(empty, append=True)
        h.partition(empty)
        with pytest.raises(C.CfgpuError):
            h.partition_finish()
