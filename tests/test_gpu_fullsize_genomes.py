"""The two reference-genome workloads at full size:
  * one 500,000,000 bp genome with 30 % repeats (families of 300 bp - 6 kb,
    2 % diverged copies), one FASTA record, k = 31, l = 20, cutoff 1;
  * 64 genomes derived from one 5,000,000 bp base genome (0.5 % SNPs, 1 %
    indels of 1-10 bp), one record each, k = 127 (256-bit keys), l = 20,
    cutoff 1.

Checks that hold at this size, as for the read workloads:
(1) every bucket's multiset of weak super k-mers equals the oracle's
    (per-bucket order-independent checksums);
(2) 64 sampled buckets: every (g, key, e_f, status, dfa) row of the learned
    batched path equals the oracle's literal construct of the same records;
(3) 16 of them: per-subgraph unitigs (label, flags) and counters equal the
    oracle's Subgraph::contract of its own tables;
(4) in every subgraph, vertices in unitigs + isolated vertices = vertices
    (src/Subgraph.cpp:268);
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

WORKLOADS = {
    "genome500M_k31": dict(k=31, l=20, cutoff=1),
    "pangenome64x5M_k127": dict(k=127, l=20, cutoff=1),
}


def make(name):
    if name == "genome500M_k31":
        seq = C.synth_genome(500_000_000, 0.3, 300, 6000, 0.02, seed=2026)
        return seq, np.array([0, len(seq)], np.uint64)
    base = C.synth_genome(5_000_000, 0.0, 300, 6000, 0.0, seed=2026)
    recs = [C.synth_mutate(base, 0.005, 0.01, 10, seed=2027 + i) for i in range(64)]
    off = np.zeros(len(recs) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in recs])
    return b"".join(recs), off


@pytest.fixture(scope="module", params=sorted(WORKLOADS))
def run(request):
    name = request.param
    p = WORKLOADS[name]
    seq, rec_off = make(name)
    say(f"[case] full-size {name}: {len(seq)} bases in {len(rec_off) - 1} records: partition, learned construct, "
        "contract")
    h = C.Handle(p["k"], p["l"], p["cutoff"])
    h.partition(C.Input.from_records(seq, rec_off, p["k"]))
    att, lab, off = h.records()
    h.construct("learned")
    h.contract()
    yield dict(p, name=name, seq=seq, rec_off=rec_off, h=h, att=att, lab=lab, off=off)
    h.close()


def test_every_bucket_holds_the_oracles_records(run):
    k, l = run["k"], run["l"]
    say(f"[case] full-size {run['name']}: oracle partition checksums of every bucket")
    s_gpu, c_gpu = O.records_hash(run["att"], run["lab"], O.label_words(k, l))
    s_ora, c_ora = O.partition_hash(run["seq"], run["rec_off"], k, l, threads=16)
    np.testing.assert_array_equal(c_gpu, c_ora)
    np.testing.assert_array_equal(s_gpu, s_ora)


def test_sampled_buckets_states_and_unitigs_equal_oracle(run):
    h, k, l, cutoff = run["h"], run["k"], run["l"], run["cutoff"]
    off = run["off"].astype(np.int64)
    nz = np.nonzero(np.diff(off))[0]
    sample = nz[np.linspace(0, len(nz) - 1, 64).astype(int)]
    say(f"[case] full-size {run['name']}: {len(sample)} sampled buckets against the oracle construct")
    idx = np.concatenate([np.arange(off[b], off[b + 1]) for b in sample])
    sizes = np.zeros(C.G, np.uint64)
    sizes[sample] = np.diff(off)[sample]
    sub_off = np.zeros(C.G + 1, np.uint64)
    sub_off[1:] = np.cumsum(sizes)
    want = O.construct(k, l, sub_off, (run["att"][idx] & 0xFFFF).astype(np.uint16), run["lab"][idx], threads=16)
    t = h.tables(buckets=sample)
    kw = t["keys"].shape[1]
    gpu = rows_sorted([t["g"]] + [t["keys"][:, j] for j in range(kw - 1, -1, -1)] + [t["e_f"], t["status"], t["dfa"]])
    ora = rows_sorted([want["g"]] + [want["keys"][:, j] for j in range(kw - 1, -1, -1)] +
                      [want["e_f"], want["status"], dfa_from(want["e_f"], want["status"], cutoff)])
    np.testing.assert_array_equal(gpu, ora)

    sub = sample[::4]
    say(f"[case] full-size {run['name']}: unitigs of {len(sub)} subgraphs against the oracle contract")
    keep = np.isin(want["g"], sub)
    states = {c: want[c][keep] for c in ("g", "keys", "e_f", "status")}
    wu = O.contract(k, cutoff, states, normalise=True, threads=16)
    recs, n, buf, nb = h.unitigs_raw()
    r = np.frombuffer(recs, dtype=[("off", "<u8"), ("len", "<u4"), ("g", "<u2"), ("flags", "<u2")], count=n)
    sel = np.nonzero(np.isin(r["g"], sub))[0]
    raw = buf.raw
    got = {"g": r["g"][sel].astype(np.uint32), "flags": r["flags"][sel].astype(np.uint32),
           "labels": [raw[int(r["off"][i]):int(r["off"][i] + r["len"][i])].decode() for i in sel]}
    stats = h.contract_stats()
    only = np.zeros(C.G, bool)
    only[sub] = True
    compare(f"full-size {run['name']}", got, np.where(only[:, None], stats, 0), wu)


def test_contract_counts_every_vertex(run):
    h = run["h"]
    nv = h.bucket_stats()["n_vertices"].astype(np.int64)
    st = h.contract_stats().astype(np.int64)
    assert not (st[:, 7] & 4).any()
    np.testing.assert_array_equal(st[:, 5] + st[:, 0], nv)


def test_device_gluing_equals_host_gluing(run):
    h, k = run["h"], run["k"]
    say(f"[case] full-size {run['name']}: maximal unitigs, device and host gluing")
    n_d, nb_d = h.cdbg_device(decode=False)
    seq = ctypes.create_string_buffer(max(nb_d, 1))
    rec = np.zeros(2 * max(n_d, 1), np.uint64)
    h._check(C.lib().cfgpu_cdbg_copy(h._h, seq, nb_d, rec.ctypes.data_as(C.u64p), n_d, None, None), "cdbg_copy")
    raw = seq.raw
    dev = sorted(raw[int(rec[2 * i]):int(rec[2 * i] + rec[2 * i + 1])] for i in range(n_d))
    del seq, raw
    recs, n, buf, nb = h.unitigs_raw()
    host = C.cdbg_from_records(k, recs, n, buf, nb, sorted=True)
    say(f"[case] full-size {run['name']}: {n_d} maximal unitigs on the device, {len(host)} on the host")
    assert len(dev) == len(host)
    assert all(a == b.encode() for a, b in zip(dev, host))
