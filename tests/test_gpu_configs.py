"""GPU parity on inputs shaped like the benchmark configurations (the reference's
headline workloads), checked record for record and state for state against the
oracle in every construct mode: C1 at full size, C3 / C4 / C5 scaled down so
the CPU oracle finishes in seconds (same k, cutoff, read model and repeat
structure)."""
import numpy as np
import pytest

import cfgpu as C
from test_gpu_parity import fasta, fastq, run_case, say

pytestmark = pytest.mark.gpu


def test_c1_reference_with_repeat_copies():
    # 4,000,000 iid bases plus 200 copies of one 1 kb repeat at 1 % substitutions,
    # one FASTA record, k = 31, cutoff 1
    base = C.synth_genome(4_000_000, 0.0, 300, 6000, 0.0, seed=101)
    g = C.synth_insert_repeats(base, 200, 1000, 0.01, seed=102)
    say(f"[case] C1 bases={len(g)}")
    run_case("C1", fasta([g]), False, 31, 20, 1, modes=("batched", "learned"))


def test_c3_repeat_rich_reference_scaled():
    # 30 % repeats from families of 300 bp .. 6 kb, 2 % diverged copies, k = 31,
    # cutoff 1 (the full configuration is 500 Mbp)
    g = C.synth_genome(12_000_000, 0.3, 300, 6000, 0.02, seed=103)
    say(f"[case] C3 bases={len(g)}")
    run_case("C3", fasta([g]), False, 31, 20, 1, modes=("batched", "learned"))


def test_c4_reads_k63_scaled():
    # 150 bp reads at 25x of a 20 %-repeat genome, 1 % substitutions, k = 63,
    # cutoff 2 (the full configuration is 300 Mbp / 50 M reads)
    g = C.synth_genome(600_000, 0.2, 300, 6000, 0.02, seed=104)
    n = len(g) * 25 // 150
    reads = C.synth_reads(g, n, 150, 0.01, seed=105)
    say(f"[case] C4 reads={n}")
    run_case("C4", fastq(reads, 150), True, 63, 20, 2, modes=("batched", "learned"))


def test_c5_pangenome_k127_scaled():
    # a base genome and derivatives with 0.5 % SNPs and 1 % indels of 1..10 bp,
    # one FASTA record each, k = 127, cutoff 1 (the full configuration is
    # 64 x 5 Mbp)
    base = C.synth_genome(1_000_000, 0.0, 300, 6000, 0.0, seed=106)
    recs = [base] + [C.synth_mutate(base, 0.005, 0.01, 10, seed=200 + i) for i in range(7)]
    say(f"[case] C5 records={len(recs)} bases={sum(map(len, recs))}")
    want = run_case("C5", fasta(recs), False, 127, 20, 1, modes=("batched", "learned"))
    # k-mers shared between the derivatives collapse into one vertex
    assert len(want["g"]) < 0.95 * sum(len(r) for r in recs)
