"""External-memory buckets: chunks partitioned with CFGPU_PARTITION_SPILL go to
the bucket store (pinned host memory, or one file per chunk), the statistics
and features are computed from the store, and cfgpu_buckets_restore brings
buckets back for cfgpu_construct_buckets.  Checked against one partition of
all the reads on the same handle type: bucket statistics, the MLP features,
every restored bucket's records (per-bucket checksums), the vertex tables, and
the compacted graph glued from batches built one restore at a time."""
import os

import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O
from test_gpu_parity import rows_sorted, say

pytestmark = pytest.mark.gpu

K, L, CUTOFF = 31, 20, 2
N_READS, READ_LEN = 40_000, 150


@pytest.fixture(scope="module")
def reads():
    g = C.synth_genome(300_000, 0.2, 300, 3000, 0.02, seed=61)
    return C.synth_reads(g, N_READS, READ_LEN, 0.005, seed=62)


def chunk_inputs(reads, n_chunks):
    bounds = np.linspace(0, N_READS, n_chunks + 1).astype(int)
    for a, b in zip(bounds[:-1], bounds[1:]):
        ro = np.arange(b - a + 1, dtype=np.uint64) * np.uint64(READ_LEN)
        yield C.Input.from_records(reads[a * READ_LEN:b * READ_LEN], ro, K)


def whole_input(reads):
    return C.Input.from_records(reads, np.arange(N_READS + 1, dtype=np.uint64) * np.uint64(READ_LEN), K)


def bucket_sums(h):
    att, lab, off = h.records()
    return O.records_hash(att, lab, O.label_words(K, L))


def table_rows(h, buckets=None):
    t = h.tables(buckets=buckets)
    return rows_sorted([t["g"], t["keys"][:, 0], t["e_f"], t["status"], t["dfa"]])


@pytest.mark.parametrize("where", ["host", "files"])
def test_spilled_buckets_restore_to_the_same_graph(reads, where, tmp_path):
    say(f"[case] bucket store in {where}: one partition vs 3 spilled chunks")
    with C.Handle(K, L, CUTOFF) as ref:
        ref.partition(whole_input(reads))
        st_ref = ref.bucket_stats()
        s_ref, c_ref = bucket_sums(ref)
        ref.set_tuning(predictor=1)  # train the MLP: its features are readable
        ref.construct("learned")
        feat_ref = ref.sizing_model()["features"]
        rows_ref = table_rows(ref)
        ref.contract()
        whole = ref.cdbg_device()

    with C.Handle(K, L, CUTOFF) as h:
        if where == "files":
            h.spill_dir(tmp_path)
        for inp in chunk_inputs(reads, 3):
            h.partition(inp, spill=True)
        sp = h.spill_stats()
        assert sp["n_segments"] == 3 and sp["in_files"] == (where == "files")
        if where == "files":
            assert len([f for f in os.listdir(tmp_path) if f.startswith("cfgpu_bucket_segment_")]) == 3
        h.partition_finish()
        st = h.bucket_stats()
        for f in ("n_records", "n_kmers", "n_bases"):
            np.testing.assert_array_equal(st[f], st_ref[f])
        assert h.spill_stats()["n_records"] == int(st_ref["n_records"].sum())
        nz = np.nonzero(st["n_records"])[0].astype(np.uint32)
        # nothing is in device memory until a restore
        with pytest.raises(C.CfgpuError):
            h.construct("learned")
        h.restore(nz)
        assert h.spill_stats()["n_resident"] == len(nz)
        s, c = bucket_sums(h)
        np.testing.assert_array_equal(c, c_ref)
        np.testing.assert_array_equal(s, s_ref)
        h.set_tuning(predictor=1)
        h.construct("learned")
        np.testing.assert_array_equal(h.sizing_model()["features"], feat_ref)
        np.testing.assert_array_equal(table_rows(h), rows_ref)

        # batches: each restore replaces the device buckets; a bucket not
        # restored cannot be built
        rng = np.random.default_rng(7)
        first = np.sort(rng.choice(nz, len(nz) // 2, replace=False)).astype(np.uint32)
        second = np.setdiff1d(nz, first).astype(np.uint32)
        recs, labels = [], []
        for part in (first, second):
            h.restore(part)
            other = second if part is first else first
            with pytest.raises(C.CfgpuError):
                h.construct_buckets(other[:1], "learned")
            h.construct_buckets(part, "learned")
            keep = np.isin(rows_ref[:, 0], part)
            np.testing.assert_array_equal(table_rows(h), rows_ref[keep])
            h.contract()
            u = h.unitigs()
            recs += list(zip(u["g"].tolist(), u["flags"].tolist()))
            labels += u["labels"]
        got = C.cdbg_from_unitigs(K, [r[0] for r in recs], [r[1] for r in recs], labels)
        say(f"[case] bucket store in {where}: {len(got)} maximal unitigs from two restored batches, "
            f"{len(whole)} from one partition")
        assert got == whole
        # a new partition empties the store (and removes its files)
        h.partition(whole_input(reads))
        assert h.spill_stats()["n_segments"] == 0
        if where == "files":
            assert not [f for f in os.listdir(tmp_path) if f.startswith("cfgpu_bucket_segment_")]


def test_spill_rejects_mixing_and_early_restore(reads):
    with C.Handle(K, L, CUTOFF) as h:
        inputs = list(chunk_inputs(reads, 2))
        h.partition(inputs[0], append=True)
        with pytest.raises(C.CfgpuError):  # kept chunks pending
            h.partition(inputs[1], spill=True)
    with C.Handle(K, L, CUTOFF) as h:
        inputs = list(chunk_inputs(reads, 2))
        h.partition(inputs[0], spill=True)
        with pytest.raises(C.CfgpuError):  # not finished
            h.restore(np.arange(4, dtype=np.uint32))
        with pytest.raises(C.CfgpuError):  # spilled chunks pending
            h.partition(inputs[1], append=True)
        h.partition(inputs[1], spill=True)
        h.partition_finish()
        with pytest.raises(C.CfgpuError):
            h.restore(np.array([C.G], np.uint32))


def test_pipelined_ingest_into_the_store(reads, tmp_path):
    # FASTQ text through the chunked ingest pipeline, every chunk's buckets to
    # files: after a restore of every bucket the records are one partition's
    text = b"".join(b"@r\n" + reads[i * READ_LEN:(i + 1) * READ_LEN] + b"\n+\n" + b"I" * READ_LEN + b"\n"
                    for i in range(N_READS))
    with C.Handle(K, L, CUTOFF) as ref:
        ref.partition(whole_input(reads))
        s_ref, c_ref = bucket_sums(ref)
    with C.Handle(K, L, CUTOFF) as h:
        h.spill_dir(tmp_path)
        st = h.ingest_buffer(text, True, threads=4, chunk_bytes=256 << 10, spill=True)
        sp = h.spill_stats()
        say(f"[case] pipelined ingest into the store: {st['n_chunks']} chunks, {sp['n_segments']} segments")
        assert st["n_chunks"] > 4 and sp["n_segments"] == st["n_chunks"] and sp["in_files"]
        h.restore(np.arange(C.G, dtype=np.uint32))
        s, c = bucket_sums(h)
        np.testing.assert_array_equal(c, c_ref)
        np.testing.assert_array_equal(s, s_ref)
