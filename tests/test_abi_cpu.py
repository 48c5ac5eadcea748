"""CPU-only checks of the product library: it loads, exports exactly what
include/*.h declares, validates arguments, and its host ingest reproduces the
oracle's records -> fragments -> 2-bit encoding.  No kernel is launched."""
import ctypes
import os
import re

import numpy as np
import pytest

import cfgpu as C
import oracle_lib as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("cfgpu.h", "cfhost.h"):
        with open(os.path.join(REPO, "include", h)) as f:
            text = f.read()
        names |= set(re.findall(r"^\s*(?:int|void)\s+(cf(?:gpu|host)_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 20
    L = C.lib()
    for n in names:
        assert hasattr(L, n), n
    assert names == set(C.SIGNATURES), names ^ set(C.SIGNATURES)


@pytest.mark.parametrize("k,l,c", [(30, 20, 1), (31, 31, 1), (31, 20, 0), (31, 20, 16), (129, 20, 1),
                                   (31, 33, 1), (1, 1, 1)])
def test_open_rejects_bad_params(k, l, c):
    p = C.Params(k, l, c, C.G, 0, 0)
    h = ctypes.c_void_p()
    assert C.lib().cfgpu_open(ctypes.byref(p), ctypes.byref(h)) == -1


def test_open_rejects_other_subgraph_counts():
    p = C.Params(31, 20, 1, 8192, 0, 0)
    h = ctypes.c_void_p()
    assert C.lib().cfgpu_open(ctypes.byref(p), ctypes.byref(h)) == -1


def _decode(inp):
    n = inp.n_bases
    idx = np.arange(n, dtype=np.uint64)
    sh = (np.uint64(62) - np.uint64(2) * (idx % np.uint64(32))).astype(np.uint64)
    return ((inp.packed[idx // np.uint64(32)] >> sh) & np.uint64(3)).astype(np.uint8)


LUT = np.full(256, 255, np.uint8)
for _c, _v in zip(b"ACGTacgt", (0, 1, 2, 3, 0, 1, 2, 3)):
    LUT[_c] = _v


@pytest.mark.parametrize("k", [3, 15, 31, 63, 127])
@pytest.mark.parametrize("fastq", [False, True])
def test_ingest_matches_oracle(k, fastq):
    rng = np.random.default_rng(k)
    g = C.synth_genome(40000, 0.2, 100, 1500, 0.02, seed=k)
    parts = [g[:9000], b"NN" + g[9000:9100].lower() + b"N" + g[9100:12000], b"ACGTN", g[12000:12000 + k],
             g[13000:13000 + k + 1], b"A" * 300 + b"\r" + g[20000:30000]]
    if fastq:
        buf = b"".join(b"@r%d\n%s\n+\n%s\n" % (i, p, b"I" * len(p)) for i, p in enumerate(parts))
    else:
        buf = b"".join(b">r%d x\n%s\n" % (i, b"\n".join(p[j:j + 70] for j in range(0, len(p), 70)))
                       for i, p in enumerate(parts))
    for threads in (1, 3, 8):
        inp = C.Input.from_buffer(buf, fastq, k, threads=threads)
        seq, ro = O.parse(buf, fastq)
        beg, ln = O.fragments(seq, ro, k)
        assert inp.n_records == len(ro) - 1
        assert inp.n_frag == len(beg)
        np.testing.assert_array_equal(np.diff(inp.frag_off.astype(np.int64)), ln.astype(np.int64))
        want = np.concatenate([LUT[np.frombuffer(seq[b:b + n], np.uint8)] for b, n in zip(beg, ln)]) \
            if len(beg) else np.zeros(0, np.uint8)
        np.testing.assert_array_equal(_decode(inp), want)
        assert inp.n_words == (inp.n_bases + 31) // 32 + 2
        assert (inp.packed[-2:] == 0).all()


def test_ingest_from_file_and_reference_fixtures(tmp_path):
    for name, fq in (("reads.fq", True), ("refs1.fa", False), ("refs2.fa", False), ("cycles.fa", False),
                     ("rob-palindrome.fa", False)):
        path = os.path.join(REPO, "tests", "golden", "ref_data", name)
        with open(path, "rb") as f:
            buf = f.read()
        a = C.Input.from_file(path, fq, 5)
        b = C.Input.from_buffer(buf, fq, 5)
        np.testing.assert_array_equal(a.packed, b.packed)
        np.testing.assert_array_equal(a.frag_off, b.frag_off)
        assert a.n_records >= 1


def test_synth_is_seeded_and_thread_independent():
    g1 = C.synth_genome(100000, 0.3, 300, 6000, 0.02, seed=9)
    g2 = C.synth_genome(100000, 0.3, 300, 6000, 0.02, seed=9)
    assert g1 == g2 and set(g1) <= set(b"ACGT")
    assert C.synth_genome(100000, 0.3, 300, 6000, 0.02, seed=10) != g1
    r1 = C.synth_reads(g1, 3000, 150, 0.01, 3, threads=1)
    r2 = C.synth_reads(g1, 3000, 150, 0.01, 3, threads=7)
    assert r1 == r2 and len(r1) == 3000 * 150
    m = C.synth_mutate(g1[:50000], 0.005, 0.01, 10, 4)
    assert 0.8 * 50000 < len(m) < 1.2 * 50000
    rep = C.synth_insert_repeats(g1[:40000], 200, 1000, 0.01, 5)
    assert len(rep) == 40000 + 200 * 1000


def _c_struct_fields(name):
    """Field names, in order, of `typedef struct name {...} name;` in include/cfgpu.h."""
    import re
    text = open(os.path.join(REPO, "include", "cfgpu.h")).read()
    body = text[text.index("typedef struct " + name + " {"):]
    body = body[body.index("{") + 1:body.index("} " + name + ";")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        _, names = decl.split(None, 1) if not decl.startswith(("const ", "struct ")) else decl.rsplit(None, 1)
        fields += [n.strip().lstrip("*") for n in names.split(",")]
    return fields


@pytest.mark.parametrize("cname,pyname", [("cfgpu_run_stats", "RunStats"), ("cfgpu_bucket_stat", "BucketStat"),
                                          ("cfgpu_tuning", "Tuning"), ("cfgpu_params", "Params"),
                                          ("cfgpu_table_view", "TableView"), ("cfgpu_unitig", "Unitig"),
                                          ("cfgpu_contract_stat", "ContractStat")])
def test_ctypes_structs_mirror_the_header(cname, pyname):
    py = [n for n, _ in getattr(C, pyname)._fields_]
    assert py == _c_struct_fields(cname)
