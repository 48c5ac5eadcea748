"""The compacted graph on the host (cfhost_cdbg_build, no GPU): per-subgraph
unitigs from the oracle's restatement of Subgraph::contract, glued by the
library, must equal the maximal unitigs of the brute-force global graph
(unitig_ref.py) as a sorted set of canonical strings."""
import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O
import unitig_ref as U
from test_oracle import rand_seq, rc
from test_oracle_contract import inputs as mixed_input


def oracle_unitigs(seq, ro, k, l, f_th):
    part = O.partition(seq, ro, k, l)
    off, att, lab = O.bucketize(part)
    return O.contract(k, f_th, O.construct(k, l, off, att, lab), normalise=True)


@pytest.mark.parametrize("k,l,f_th", [(31, 20, 1), (31, 20, 2), (15, 7, 1), (15, 7, 3), (21, 11, 1), (63, 20, 1),
                                      (95, 20, 1), (95, 20, 2), (127, 20, 1), (127, 20, 2)])
def test_cdbg_equals_bruteforce_unitigs(k, l, f_th):
    rng = np.random.default_rng(70 + k + f_th)
    seq, ro = O.parse(mixed_input(rng), fastq=False)
    res = oracle_unitigs(seq, ro, k, l, f_th)
    got = C.cdbg_from_unitigs(k, res["g"], res["flags"], res["labels"])
    assert got == U.unitigs(seq, ro, k, f_th)


def test_cdbg_reads_and_cycles():
    rng = np.random.default_rng(12)
    g = rand_seq(rng, 2000)
    reads = []
    for _ in range(150):
        p = int(rng.integers(0, len(g) - 100))
        r = list(g[p:p + 100])
        for j in range(100):
            if rng.random() < 0.01:
                r[j] = "ACGT"[int(rng.integers(0, 4))]
        r = "".join(r)
        reads.append(r if rng.random() < 0.5 else rc(r))
    # tandem repeats give cycles inside one subgraph and across several
    reads += [rand_seq(rng, p) * (1 + 120 // p) for p in (3, 5, 7, 40, 45, 60)]
    buf = "".join(f">r{i}\n{r}\n" for i, r in enumerate(reads)).encode()
    seq, ro = O.parse(buf, fastq=False)
    for k, l, f_th in ((21, 11, 1), (21, 11, 2), (31, 20, 1), (63, 20, 1), (95, 20, 2)):
        res = oracle_unitigs(seq, ro, k, l, f_th)
        got = C.cdbg_from_unitigs(k, res["g"], res["flags"], res["labels"])
        want = U.unitigs(seq, ro, k, f_th)
        assert got == want
        assert len(got) == len(set(got))


def test_cdbg_rejects_bad_records():
    # a label shorter than k and a label beyond the buffer are refused
    with pytest.raises(RuntimeError):
        C.cdbg_from_unitigs(31, [0], [0], ["ACGT"])
