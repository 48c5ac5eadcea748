"""The compacted de Bruijn graph end to end on the GPU: partition, construct,
contract and the gluing of lm-tigs on the device (cfgpu_cdbg), which must also
agree with the host gluing (cfhost_cdbg_build).  The
sorted set of canonical maximal unitigs must equal the brute-force global
graph's (unitig_ref.py) on small inputs, and the oracle pipeline's (partition,
construct and contract restated on the CPU, glued the same way) on inputs
shaped like the benchmark configurations."""
import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O
import unitig_ref as U
from test_gpu_parity import fasta, fastq, say
from test_oracle_contract import inputs as mixed_input

pytestmark = pytest.mark.gpu


def gpu_cdbg(buf, fq, k, l, cutoff, mode="learned"):
    """Glued on the device (cfgpu_cdbg); the host gluing of the same unitigs must agree."""
    with C.Handle(k, l, cutoff) as h:
        h.partition(C.Input.from_buffer(buf, fq, k))
        h.construct(mode)
        h.contract()
        dev = h.cdbg_device()
        rs = h.run_stats()
        say(f"[case] cdbg device: unitigs={rs['n_cdbg_unitigs']} rounds={rs['n_cdbg_rounds']} "
            f"cycle_pieces={rs['n_cdbg_cycle_pieces']}")
        assert dev == h.cdbg()
        assert rs["n_cdbg_unitigs"] == len(dev)
        gpu_cdbg.last_stats = rs
        return dev


def oracle_cdbg(buf, fq, k, l, cutoff):
    seq, ro = O.parse(buf, fq)
    part = O.partition(seq, ro, k, l)
    off, att, lab = O.bucketize(part)
    res = O.contract(k, cutoff, O.construct(k, l, off, att, lab), normalise=True)
    return seq, ro, C.cdbg_from_unitigs(k, res["g"], res["flags"], res["labels"])


@pytest.mark.parametrize("k,l,cutoff", [(31, 20, 1), (31, 20, 2), (15, 7, 1), (63, 20, 1)])
def test_cdbg_small_equals_bruteforce(k, l, cutoff):
    rng = np.random.default_rng(90 + k + cutoff)
    buf = mixed_input(rng)
    say(f"[case] cdbg-small k={k} cutoff={cutoff}")
    seq, ro = O.parse(buf, False)
    got = gpu_cdbg(buf, False, k, l, cutoff, mode="batched")
    assert got == U.unitigs(seq, ro, k, cutoff)


def test_cdbg_cycles_within_and_across_subgraphs():
    # short-period tandem repeats cycle inside one subgraph; long periods pass
    # through several subgraphs and close on the host
    rng = np.random.default_rng(95)
    recs = ["".join("ACGT"[b] for b in rng.integers(0, 4, p)) for p in (3, 4, 7, 40, 55, 70, 90, 120)]
    recs = [(u * (2 + 300 // len(u))).encode() for u in recs]
    buf = fasta(recs)
    seq, ro = O.parse(buf, False)
    for k, l in ((31, 20), (21, 11)):
        say(f"[case] cdbg-cycles k={k}")
        got = gpu_cdbg(buf, False, k, l, 1, mode="batched")
        assert got == U.unitigs(seq, ro, k, 1)
        assert gpu_cdbg.last_stats["n_cdbg_cycle_pieces"] > 0


def test_cdbg_two_piece_cycles():
    # period-16 tandem repeats at k=15, l=7 whose cycle is glued from exactly two
    # lm-tigs (the shortest cycle through two subgraphs): pointer jumping can
    # land a state of such a cycle on itself, and it must still read as a cycle
    units = ["GACCTCTGACGATTGA", "TCTACACTCATTGGAT", "AATCCAAACGCCAGTT",
             "GCCTAACATCGTTACG", "ACTCTAATCACGAGAG", "CTGGATGGCTCTTTTT"]
    buf = fasta([(u * 14).encode() for u in units])
    seq, ro = O.parse(buf, False)
    say("[case] cdbg-two-piece-cycles k=15")
    got = gpu_cdbg(buf, False, 15, 7, 1, mode="batched")
    assert got == U.unitigs(seq, ro, 15, 1)
    assert gpu_cdbg.last_stats["n_cdbg_cycle_pieces"] >= 2 * len(units)


def test_cdbg_reference_with_repeat_copies():
    # 4,000,000 iid bases plus 200 copies of a 1 kb repeat (1 % substitutions), k = 31, cutoff 1
    base = C.synth_genome(4_000_000, 0.0, 300, 6000, 0.0, seed=101)
    g = C.synth_insert_repeats(base, 200, 1000, 0.01, seed=102)
    buf = fasta([g])
    say("[case] cdbg-C1")
    _, _, want = oracle_cdbg(buf, False, 31, 20, 1)
    got = gpu_cdbg(buf, False, 31, 20, 1)
    say(f"[case] cdbg-C1 unitigs gpu={len(got)} oracle={len(want)}")
    assert got == want
    # with cutoff 1 every distinct canonical 31-mer of the genome lies in exactly one unitig
    assert sum(len(u) - 30 for u in got) == distinct_canonical_kmers(g, 31)


def distinct_canonical_kmers(seq, k):
    c = np.frombuffer(seq, np.uint8)
    code = np.select([c == 65, c == 67, c == 71], [0, 1, 2], 3).astype(np.uint64)
    n = len(code) - k + 1
    fwd = np.zeros(n, np.uint64)
    rev = np.zeros(n, np.uint64)
    for j in range(k):
        fwd = (fwd << np.uint64(2)) | code[j:j + n]
        rev = rev | ((np.uint64(3) - code[j:j + n]) << np.uint64(2 * j))
    return len(np.unique(np.minimum(fwd, rev)))


def test_cdbg_reads_cutoff2():
    g = C.synth_genome(400_000, 0.2, 300, 3000, 0.02, seed=91)
    reads = C.synth_reads(g, 80_000, 150, 0.005, seed=92)
    buf = fastq(reads, 150)
    say("[case] cdbg-reads")
    _, _, want = oracle_cdbg(buf, True, 31, 20, 2)
    for mode in ("plain", "learned"):
        got = gpu_cdbg(buf, True, 31, 20, 2, mode=mode)
        say(f"[case] cdbg-reads mode={mode} unitigs={len(got)}")
        assert got == want


def test_cdbg_pangenome_k127():
    base = C.synth_genome(300_000, 0.0, 300, 6000, 0.0, seed=93)
    recs = [base] + [C.synth_mutate(base, 0.005, 0.01, 10, seed=94 + i) for i in range(3)]
    buf = fasta(recs)
    say("[case] cdbg-k127")
    _, _, want = oracle_cdbg(buf, False, 127, 20, 1)
    got = gpu_cdbg(buf, False, 127, 20, 1)
    assert got == want


def test_cf_build_cli_writes_unitigs(tmp_path):
    # the `cuttlefish build` front end: -o PREFIX writes PREFIX.fa, one record per maximal unitig
    import os
    import subprocess
    rng = np.random.default_rng(97)
    buf = mixed_input(rng)
    (tmp_path / "in.fa").write_bytes(buf)
    exe = os.path.join(os.path.dirname(C.LIB_PATH), "cf_build")
    say("[case] cf_build CLI")
    r = subprocess.run([exe, "-s", str(tmp_path / "in.fa"), "-k", "31", "--ref", "-o", str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "out.fa").read_text().split("\n")
    got = sorted(lines[i + 1] for i in range(0, len(lines) - 1, 2))
    assert all(lines[i].startswith(">") for i in range(0, len(lines) - 1, 2))
    seq, ro = O.parse(buf, False)
    assert got == U.unitigs(seq, ro, 31, 1)
