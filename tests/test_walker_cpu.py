"""The construct kernels' per-record walker (RecordWalker, rolled k-mer and
reverse complement) and counter update, built for the host, against the
oracle's literal Subgraph::construct restatement (src/Subgraph.cpp:44-108) on
adversarial records: repeats past counter saturation, homopolymers in both
orientations, tandem repeats, the reference's palindrome fixture, N runs, and
every label-word count."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("walker") / "walker_host")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", os.path.join(HERE, "walker_host.hip"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def run_harness(exe, k, W, g, att16, labels):
    att = (att16.astype(np.uint32) | (g.astype(np.uint32) << 16)).astype(np.uint32)
    lab = np.ascontiguousarray(labels, dtype=np.uint64).reshape(-1)
    assert lab.size == len(att) * W
    blob = (np.array([k, W], np.uint32).tobytes() + np.array([len(att)], np.uint64).tobytes() +
            att.tobytes() + lab.tobytes())
    out = subprocess.run([exe], input=blob, capture_output=True, check=True).stdout
    n = int(np.frombuffer(out[:8], np.uint64)[0])
    rec = np.dtype([("g", "<u4"), ("key", "<u8", (4,)), ("ef", "<u4"), ("st", "<u4")])
    return np.frombuffer(out[8:], rec, count=n)


def rand_seq(rng, n):
    return "".join("ACGT"[i] for i in rng.integers(0, 4, n))


def rc(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def adversarial_fasta(k, seed):
    rng = np.random.default_rng(seed)
    base = rand_seq(rng, 1500)
    pal = "AACTGACATGTCAGTT"
    s = (base + base[:400] + "A" * (k + 9) + rand_seq(rng, 60) + "T" * (k + 14) + "CAG" * (k // 2 + 20) +
         pal * 3 + "NNNN" + "acgt" * 30 + "C" * (k + 2) + "G" * (k + 5))
    recs = [s] + [base[100:600]] * 20 + [rc(base[300:800])] * 4 + ["AT" * (k + 3), "N" * 7 + base[:k]]
    return "".join(f">r{i}\n{r}\n" for i, r in enumerate(recs)).encode()


def compare(exe, seq, ro, k, l):
    part = O.partition(seq, ro, k, l)
    off, att, lab = O.bucketize(part)
    g = np.repeat(np.arange(len(off) - 1, dtype=np.uint32), np.diff(off.astype(np.int64)))
    got = run_harness(exe, k, O.label_words(k, l), g, att, lab)
    want = O.construct(k, l, off, att, lab, mode=0)
    assert len(got) == len(want["g"])

    def rows(g, keys, ef, st):
        cols = [g.astype(np.uint64)] + [keys[:, j].astype(np.uint64) for j in range(4)] + \
               [ef.astype(np.uint64), st.astype(np.uint64)]
        m = np.stack(cols, axis=1)
        return m[np.lexsort(cols[::-1])]

    np.testing.assert_array_equal(rows(got["g"], got["key"], got["ef"], got["st"]),
                                  rows(want["g"], want["keys"], want["e_f"], want["status"]))
    return len(got)


@pytest.mark.parametrize("k,l", [(31, 20), (15, 7), (21, 11), (3, 1), (33, 1), (45, 20), (63, 20),
                                 (95, 24), (127, 20), (127, 1)])
def test_walker_matches_oracle(harness, k, l):
    seq, ro = O.parse(adversarial_fasta(k, seed=k * 7 + l), fastq=False)
    assert compare(harness, seq, ro, k, l) > 0


def test_walker_palindrome_fixture(harness):
    with open(os.path.join(HERE, "golden", "ref_data", "rob-palindrome.fa"), "rb") as f:
        seq, ro = O.parse(f.read(), fastq=False)
    assert compare(harness, seq, ro, 15, 7) == 1


def test_ef_active_matches_edge_at(harness):
    # the write-out's active test equals "edge_at is not E on some side" for
    # every cutoff and every counter word built from boundary field values
    out = subprocess.run([harness, "ef_active"], capture_output=True, check=True, text=True).stdout.split()
    bad, n = int(out[0]), int(out[1])
    assert n == 15 * 5 ** 8
    assert bad == 0
