"""The 50M-read workload at full size: 50,000,000 reads of 150 bp at 25x from
a 300,000,000 bp genome with 20 % repeats (families of 300 bp - 6 kb, 2 %
diverged copies), 1 % substitutions, k = 63 (128-bit keys), l = 20, cutoff 2.

Checks that hold at this size:
(1) every bucket's multiset of weak super k-mers equals the oracle's
    (per-bucket order-independent checksums, the oracle partitioning every
    read on the host's cores);
(2) 96 sampled buckets: every (g, key, e_f, status, dfa) row of the learned
    batched path equals the oracle's literal construct of the same records;
(3) 24 of them: the per-subgraph unitigs (label, flags) and counters of
    cfgpu_contract equal the oracle's Subgraph::contract of its own tables;
(4) in every subgraph, vertices in unitigs + isolated vertices = vertices
    (the reference's assert, src/Subgraph.cpp:268);
(5) the maximal unitigs glued on the device equal those the host routine
    (cfhost_cdbg_build) glues from the same contract output."""
import ctypes

import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O
from test_gpu_contract import compare
from test_gpu_parity import dfa_from, rows_sorted, say

pytestmark = pytest.mark.gpu

K, L, CUTOFF = 63, 20, 2
N_READS, READ_LEN = 50_000_000, 150


@pytest.fixture(scope="module")
def run():
    g = C.synth_genome(300_000_000, 0.2, 300, 6000, 0.02, seed=2026)
    reads = C.synth_reads(g, N_READS, READ_LEN, 0.01, seed=2027)
    del g
    rec_off = np.arange(N_READS + 1, dtype=np.uint64) * np.uint64(READ_LEN)
    inp = C.Input.from_records(reads, rec_off, K)
    say("[case] full-size 50M reads k=63: partition, records, learned construct, contract")
    h = C.Handle(K, L, CUTOFF)
    h.partition(inp)
    del inp
    att, lab, off = h.records()
    h.construct("learned")
    h.contract()
    yield {"reads": reads, "rec_off": rec_off, "h": h, "att": att, "lab": lab, "off": off}
    h.close()


def test_every_bucket_holds_the_oracles_records(run):
    say("[case] full-size k=63: oracle partition checksums of every bucket")
    s_gpu, c_gpu = O.records_hash(run["att"], run["lab"], O.label_words(K, L))
    s_ora, c_ora = O.partition_hash(run["reads"], run["rec_off"], K, L, threads=16)
    np.testing.assert_array_equal(c_gpu, c_ora)
    np.testing.assert_array_equal(s_gpu, s_ora)


def _sample_states(run, sample):
    off = run["off"].astype(np.int64)
    idx = np.concatenate([np.arange(off[b], off[b + 1]) for b in sample])
    sub_off = np.zeros(C.G + 1, np.uint64)
    sizes = np.zeros(C.G, np.uint64)
    sizes[sample] = np.diff(off)[sample]
    sub_off[1:] = np.cumsum(sizes)
    return O.construct(K, L, sub_off, (run["att"][idx] & 0xFFFF).astype(np.uint16), run["lab"][idx], threads=16)


def test_sampled_buckets_states_and_unitigs_equal_oracle(run):
    h = run["h"]
    nz = np.nonzero(np.diff(run["off"].astype(np.int64)))[0]
    sample = nz[np.linspace(0, len(nz) - 1, 96).astype(int)]
    say(f"[case] full-size k=63: {len(sample)} sampled buckets against the oracle construct")
    want = _sample_states(run, sample)
    t = h.tables(buckets=sample)
    gpu = rows_sorted([t["g"], t["keys"][:, 1], t["keys"][:, 0], t["e_f"], t["status"], t["dfa"]])
    ora = rows_sorted([want["g"], want["keys"][:, 1], want["keys"][:, 0], want["e_f"], want["status"],
                       dfa_from(want["e_f"], want["status"], CUTOFF)])
    np.testing.assert_array_equal(gpu, ora)

    sub = sample[::4]
    say(f"[case] full-size k=63: unitigs of {len(sub)} subgraphs against the oracle contract")
    keep = np.isin(want["g"], sub)
    states = {c: want[c][keep] for c in ("g", "keys", "e_f", "status")}
    wu = O.contract(K, CUTOFF, states, normalise=True, threads=16)
    recs, n, buf, nb = h.unitigs_raw()
    r = np.frombuffer(recs, dtype=[("off", "<u8"), ("len", "<u4"), ("g", "<u2"), ("flags", "<u2")], count=n)
    sel = np.nonzero(np.isin(r["g"], sub))[0]
    raw = buf.raw
    got = {"g": r["g"][sel].astype(np.uint32), "flags": r["flags"][sel].astype(np.uint32),
           "labels": [raw[int(r["off"][i]):int(r["off"][i] + r["len"][i])].decode() for i in sel]}
    stats = h.contract_stats()
    only = np.zeros(C.G, bool)
    only[sub] = True
    compare("full-size k=63", got, np.where(only[:, None], stats, 0), wu)


def test_contract_counts_every_vertex(run):
    h = run["h"]
    nv = h.bucket_stats()["n_vertices"].astype(np.int64)
    st = h.contract_stats().astype(np.int64)
    assert not (st[:, 7] & 4).any()
    np.testing.assert_array_equal(st[:, 5] + st[:, 0], nv)


def test_device_gluing_equals_host_gluing(run):
    h = run["h"]
    say("[case] full-size k=63: maximal unitigs, device and host gluing")
    n_d, nb_d = h.cdbg_device(decode=False)
    seq = ctypes.create_string_buffer(max(nb_d, 1))
    rec = np.zeros(2 * max(n_d, 1), np.uint64)
    h._check(C.lib().cfgpu_cdbg_copy(h._h, seq, nb_d, rec.ctypes.data_as(C.u64p), n_d, None, None), "cdbg_copy")
    raw = seq.raw
    dev = sorted(raw[int(rec[2 * i]):int(rec[2 * i] + rec[2 * i + 1])] for i in range(n_d))
    del seq, raw
    recs, n, buf, nb = h.unitigs_raw()
    host = C.cdbg_from_records(K, recs, n, buf, nb, sorted=True)
    say(f"[case] full-size k=63: {n_d} maximal unitigs on the device, {len(host)} on the host")
    assert len(dev) == len(host)
    assert all(a == b.encode() for a, b in zip(dev, host))
