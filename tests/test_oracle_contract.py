"""Pins the oracle's restatement of Subgraph::contract (src/Subgraph.cpp:184-270,
include/Subgraph.hpp:414-543) before it is trusted as the checker of the
device contract.

* Gluing: the lm-tigs of all subgraphs joined at their shared discontinuity
  vertices, plus the trivial unitigs, are exactly the maximal unitigs of the
  global graph built by brute force over all (k+1)-mers (unitig_ref.py), as
  canonical strings (Unitig_Collator.hpp:308-400).
* Endpoints: each discontinuity edge (u, s_u, v, s_v) the walk reports is the
  one the lm-tig's label implies (left end = its first k-mer, right end = its
  last; the side is the one facing the lm-tig).
* Counters: vertices in unitigs + isolated vertices = the subgraph's vertices
  (the reference asserts this, src/Subgraph.cpp:268), one unitig per trivial
  unitig or discontinuity edge, label bytes = sum of label lengths.
"""
import numpy as np
import pytest

import oracle_lib as O
import unitig_ref as U
from test_oracle import key_str, rand_seq, rc


def canon(x):
    return min(x, rc(x))


def label_ends(lab, k):
    """(canonical k-mer, side facing the lm-tig) of the left and right ends."""
    x, y = lab[:k], lab[-k:]
    return (canon(x), 1 if x == canon(x) else 0), (canon(y), 0 if y == canon(y) else 1)


def glue(k, res):
    """Global maximal unitigs from per-subgraph contract output."""
    out, lm = [], []
    for lab, fl in zip(res["labels"], res["flags"]):
        if fl & 6:
            lm.append(lab)
        else:
            out.append(lab)
    flags = {i: fl for i, fl in enumerate(f for f in res["flags"] if f & 6)}
    at = {}
    for i, lab in enumerate(lm):
        (xl, sl), (xr, sr) = label_ends(lab, k)
        for key, end, bit in (((xl, sl), "L", 2), ((xr, sr), "R", 4)):
            if flags[i] & bit:
                assert key not in at, key  # a discontinuity vertex side faces one lm-tig
                at[key] = (i, end)
    used = [False] * len(lm)

    def oriented(i, left_end):
        return lm[i] if left_end == "L" else rc(lm[i])

    def exits(i, end):
        return bool(flags[i] & (2 if end == "L" else 4))

    def follow(i, start_end):
        """Walk the chain from lm-tig i entered at `start_end`; returns (string, closed)."""
        s = oriented(i, start_end)
        used[i] = True
        cur, far = i, "R" if start_end == "L" else "L"
        while exits(cur, far):
            y = s[-k:]
            yc = canon(y)
            s_int = 0 if y == yc else 1
            nxt = at.get((yc, 1 - s_int))
            if nxt is None:
                break
            j, jend = nxt
            if j == i:
                return s[:-1], True  # back at the start: a cycle through several subgraphs
            assert not used[j]
            used[j] = True
            t = oriented(j, jend)
            assert t[:k] == y
            s += t[k:]
            cur, far = j, "R" if jend == "L" else "L"
        return s, False

    for i in range(len(lm)):
        if used[i]:
            continue
        for end in ("L", "R"):
            if not exits(i, end):
                out.append(U.canonical_linear(follow(i, end)[0], k))
                break
            (x, s) = label_ends(lm[i], k)[0 if end == "L" else 1]
            if (x, 1 - s) not in at:
                out.append(U.canonical_linear(follow(i, end)[0], k))
                break
    for i in range(len(lm)):
        if not used[i]:
            s, closed = follow(i, "L")
            assert closed
            out.append(U.canonical_cycle(s, k))
    return sorted(out)


def check_contract(seq, ro, k, l, f_th):
    part = O.partition(seq, ro, k, l)
    off, att, lab = O.bucketize(part)
    st = O.construct(k, l, off, att, lab)
    raw = O.contract(k, f_th, st, normalise=False)
    res = O.contract(k, f_th, st, normalise=True)
    # counters
    c = res["counts"]
    nv = np.bincount(st["g"], minlength=16384)
    np.testing.assert_array_equal(c[:, 5] + c[:, 0], nv)
    per_g = np.bincount(res["g"], minlength=16384)
    np.testing.assert_array_equal(c[:, 2] + c[:, 4], per_g)
    assert (c[:, 1] <= c[:, 0]).all()
    assert int(c[:, 3].sum()) == int(((res["flags"] & 1) == 1).sum())
    assert int(c[:, 6].sum()) == sum(len(x) for x in res["labels"])
    # endpoints reported by the walk match the label, in either orientation
    for r in (raw, res):
        for lab_, fl, e, sd in zip(r["labels"], r["flags"], r["ends"], r["sides"]):
            if not fl & 6:
                assert tuple(sd) == (2, 2)
                continue
            assert not fl & 1
            (xl, sl), (xr, sr) = label_ends(lab_, k)
            if fl & 2:
                assert (key_str(e[:4], k), sd[0]) == (xl, sl)
            else:
                assert sd[0] == 2 and not e[:4].any()
            if fl & 4:
                assert (key_str(e[4:], k), sd[1]) == (xr, sr)
            else:
                assert sd[1] == 2 and not e[4:].any()
    # trivial unitigs come out in the reference's orientation
    for lab_, fl in zip(res["labels"], res["flags"]):
        if fl & 1:
            assert U.canonical_cycle(lab_, k) == lab_
        elif not fl & 6:
            assert U.canonical_linear(lab_, k) == lab_
    # normalising only re-orients lm-tigs
    assert sorted(raw["labels"][i] if not raw["flags"][i] & 6 else canon_lm(raw["labels"][i], k)
                  for i in range(len(raw["labels"]))) == sorted(res["labels"])
    want = U.unitigs(seq, ro, k, f_th)
    got = glue(k, res)
    assert got == want
    return res


def canon_lm(lab, k):
    return U.canonical_linear(lab, k)


def inputs(rng):
    base = rand_seq(rng, 700)
    pal = "AACTGACATGTCAGTT"
    unit = rand_seq(rng, 45)
    s = base + base[:300] + "A" * 40 + rand_seq(rng, 50) + "T" * 45 + ("CAG" * 30) + pal * 3
    recs = [s, unit * 6, "AC" * 40, "A" * 60] + [base[100:400]] * 3 + [rc(base[200:500])] * 2
    recs += [base[350:380] + "G" + base[381:600]]
    return "".join(f">r{i}\n{r}\n" for i, r in enumerate(recs)).encode()


@pytest.mark.parametrize("k,l,f_th", [(31, 20, 1), (31, 20, 2), (15, 7, 1), (15, 7, 3), (21, 11, 1), (63, 20, 1),
                                      (95, 20, 1), (95, 20, 2), (127, 20, 1), (127, 20, 2)])
def test_contract_glued_equals_global_unitigs(k, l, f_th):
    rng = np.random.default_rng(7 + k + f_th)
    seq, ro = O.parse(inputs(rng), fastq=False)
    res = check_contract(seq, ro, k, l, f_th)
    assert (res["flags"] & 6).any()  # lm-tigs present


def test_contract_reads_with_errors():
    # reads over a small genome with substitutions: tips and bubbles, cutoff 2
    rng = np.random.default_rng(11)
    g = rand_seq(rng, 1500)
    reads = []
    for _ in range(120):
        p = int(rng.integers(0, len(g) - 100))
        r = list(g[p:p + 100])
        for j in range(100):
            if rng.random() < 0.01:
                r[j] = "ACGT"[int(rng.integers(0, 4))]
        r = "".join(r)
        reads.append(r if rng.random() < 0.5 else rc(r))
    fq = "".join(f"@r{i}\n{r}\n+\n{'I' * len(r)}\n" for i, r in enumerate(reads)).encode()
    seq, ro = O.parse(fq, fastq=True)
    for f_th in (1, 2):
        check_contract(seq, ro, 21, 11, f_th)


def test_contract_single_subgraph_cycle():
    # the homopolymer A^k alone forms a one-vertex cycle through its own two
    # sides (A^{k+1} on back, T^{k+1} on front); k = 15
    k, l = 15, 7
    seq, ro = O.parse(b">a\n" + b"A" * 30 + b"\n>t\n" + b"T" * 30 + b"\n", fastq=False)
    res = check_contract(seq, ro, k, l, 1)
    assert list(res["labels"]) == ["A" * k] and list(res["flags"]) == [1]


@pytest.mark.parametrize("k,l", [(15, 7), (31, 20)])
def test_contract_tandem_repeat_cycles(k, l):
    # a tandem repeat of period p <= w has the same l-mers in every (k-1)-mer
    # window, so its p vertices share one subgraph and form a cycle there;
    # random units put the minimum vertex at every rotation and strand
    rng = np.random.default_rng(5 + k)
    recs = []
    for p in range(2, 10):
        for _ in range(4):
            u = rand_seq(rng, p)
            recs.append(u * (1 + 3 * k // p))
    fa = "".join(f">c{i}\n{r}\n" for i, r in enumerate(recs)).encode()
    seq, ro = O.parse(fa, fastq=False)
    res = check_contract(seq, ro, k, l, 1)
    cyc = [x for x, f in zip(res["labels"], res["flags"]) if f & 1]
    assert len(cyc) >= 20
    # the minimum vertex is not the walk's first one for some cycles
    assert len({c[:k] for c in cyc}) == len(cyc)
