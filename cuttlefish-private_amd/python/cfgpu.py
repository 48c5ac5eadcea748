"""ctypes binding of libcfgpu (include/cfgpu.h, include/cfhost.h).

The binding is thin: every call goes straight to the C ABI, which runs the
gfx950 kernels.  There is no CPU fallback; a missing library or GPU raises.
"""
import builtins
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libcfgpu.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)

MODE_PLAIN, MODE_BATCHED, MODE_BATCHED_LEARNED = 0, 1, 2
MODES = {"plain": MODE_PLAIN, "batched": MODE_BATCHED, "learned": MODE_BATCHED_LEARNED}
G = 16384


class Params(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("l", ctypes.c_uint32), ("cutoff", ctypes.c_uint32),
                ("subgraph_count", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("reserved", ctypes.c_uint32)]


class BucketStat(ctypes.Structure):
    _fields_ = [("n_records", ctypes.c_uint64), ("n_kmers", ctypes.c_uint64),
                ("n_bases", ctypes.c_uint64), ("n_vertices", ctypes.c_uint64),
                ("capacity", ctypes.c_uint64), ("predicted", ctypes.c_float),
                ("flags", ctypes.c_uint32), ("n_records_walked", ctypes.c_uint64),
                ("n_kmers_walked", ctypes.c_uint64)]


class TableView(ctypes.Structure):
    _fields_ = [("entries", ctypes.c_void_p), ("capacity", ctypes.c_uint64),
                ("entry_bytes", ctypes.c_uint32), ("key_words", ctypes.c_uint32),
                ("n_vertices", ctypes.c_uint64)]


class RunStats(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint64) for n in ("n_bases", "n_fragments", "n_records", "n_kmers",
                                                 "n_edges", "n_vertices", "table_slots",
                                                 "table_bytes")]
                + [(n, ctypes.c_uint32) for n in ("n_buckets_nonempty", "n_overflowed",
                                                  "n_fallback_rounds", "n_samples")]
                + [(n, ctypes.c_double) for n in ("t_minimizer_ms", "t_count_ms", "t_emit_ms",
                                                  "t_features_ms", "t_partition_ms", "t_sample_ms",
                                                  "t_train_ms", "t_infer_ms",
                                                  "t_construct_main_ms", "t_fallback_ms",
                                                  "t_construct_ms", "t_construct_kernel_ms")]
                + [(n, ctypes.c_uint32) for n in ("n_lds_buckets", "predictor")]
                + [(n, ctypes.c_double) for n in ("mlp_loss", "mispredict_mean_abs_log",
                                                  "mispredict_under_frac", "fallback_rate",
                                                  "holdout_err_mlp", "holdout_err_const")]
                + [("n_records_walked", ctypes.c_uint64), ("n_dedup_fallback", ctypes.c_uint32),
                   ("n_lds_all_pass", ctypes.c_uint32), ("t_dedup_ms", ctypes.c_double),
                   ("t_contract_ms", ctypes.c_double), ("n_unitigs", ctypes.c_uint64),
                   ("n_unitig_bytes", ctypes.c_uint64), ("n_contract_global", ctypes.c_uint32),
                   ("n_contract_runs", ctypes.c_uint32), ("t_cdbg_ms", ctypes.c_double),
                   ("n_cdbg_unitigs", ctypes.c_uint64), ("n_cdbg_bytes", ctypes.c_uint64),
                   ("n_cdbg_rounds", ctypes.c_uint32), ("n_cdbg_cycle_pieces", ctypes.c_uint32),
                   ("n_contract_dense", ctypes.c_uint32), ("n_partition_chunks", ctypes.c_uint32),
                   ("t_partition_merge_ms", ctypes.c_double), ("n_contract_wide", ctypes.c_uint32)])

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Unitig(ctypes.Structure):
    _fields_ = [("label_off", ctypes.c_uint64), ("label_len", ctypes.c_uint32),
                ("g", ctypes.c_uint16), ("flags", ctypes.c_uint16)]


CONTRACT_STAT_FIELDS = ("isolated", "phantom", "trivial", "icc", "disc_edges", "vertices",
                        "label_bytes", "flags")


class ContractStat(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in CONTRACT_STAT_FIELDS]


class Tuning(ctypes.Structure):
    _fields_ = [("load_factor", ctypes.c_float), ("workgroup", ctypes.c_uint32),
                ("chunk_records", ctypes.c_uint32), ("wgs_per_cu", ctypes.c_uint32),
                ("margin", ctypes.c_float), ("use_graph", ctypes.c_uint32),
                ("max_passes", ctypes.c_uint32), ("predictor", ctypes.c_uint32),
                ("dedup", ctypes.c_uint32), ("mlp_epochs", ctypes.c_uint32), ("mlp_lr", ctypes.c_float),
                ("lds_load", ctypes.c_float), ("contract_wide", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SizingModel(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("n_features", "hidden", "n_params", "n_train", "epochs", "seed")] + \
               [("lr", ctypes.c_float), ("loss", ctypes.c_float)]


class SpillStat(ctypes.Structure):
    _fields_ = [("n_segments", ctypes.c_uint64), ("n_records", ctypes.c_uint64), ("bytes", ctypes.c_uint64),
                ("in_files", ctypes.c_uint32), ("n_resident", ctypes.c_uint32),
                ("n_resident_records", ctypes.c_uint64), ("t_spill_ms", ctypes.c_double),
                ("t_restore_ms", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class IngestStats(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint64) for n in ("n_bytes", "n_records", "n_bases", "n_fragments")]
                + [("n_chunks", ctypes.c_uint32), ("threads", ctypes.c_uint32)]
                + [(n, ctypes.c_double) for n in ("t_total_ms", "t_host_ms", "t_device_ms",
                                                  "t_finish_ms")])

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TuneTrial(ctypes.Structure):
    _fields_ = [("tuning", Tuning), ("ms", ctypes.c_double), ("identical", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


# name -> (restype, argtypes); the set of names must equal the header's.
SIGNATURES = {
    "cfgpu_open": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]),
    "cfgpu_close": (ctypes.c_int, [ctypes.c_void_p]),
    "cfgpu_last_error": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    "cfgpu_params_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Params)]),
    "cfgpu_partition": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32]),
    "cfgpu_partition_finish": (ctypes.c_int, [ctypes.c_void_p]),
    "cfgpu_spill_dir": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "cfgpu_buckets_restore": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32]),
    "cfgpu_spill_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SpillStat)]),
    "cfgpu_bucket_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BucketStat),
                                          ctypes.c_uint32]),
    "cfgpu_records_copy": (ctypes.c_int, [ctypes.c_void_p, u32p, u64p, u64p, ctypes.c_uint64,
                                          u64p]),
    "cfgpu_load_buckets": (ctypes.c_int, [ctypes.c_void_p, u32p, u64p, u64p]),
    "cfgpu_construct": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "cfgpu_construct_buckets": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, u32p, ctypes.c_uint32]),
    "cfgpu_table_view_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32,
                                            ctypes.POINTER(TableView)]),
    "cfgpu_table_copy": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32, u64p, u32p, u8p,
                                        u8p, u32p, ctypes.c_uint64, u64p]),
    "cfgpu_table_copy_noalloc": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u32p,
                                                ctypes.c_uint64, u64p]),
    "cfgpu_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "cfgpu_contract": (ctypes.c_int, [ctypes.c_void_p]),
    "cfgpu_contract_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ContractStat),
                                            ctypes.c_uint32]),
    "cfgpu_unitigs_view": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(ctypes.c_void_p), u64p, u64p]),
    "cfgpu_unitigs_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Unitig), ctypes.c_uint64,
                                          ctypes.c_char_p, ctypes.c_uint64, u64p, u64p]),
    "cfgpu_run_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RunStats)]),
    "cfgpu_sizing_model_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SizingModel), ctypes.c_void_p,
                                              ctypes.c_void_p, u32p, ctypes.c_void_p, ctypes.c_void_p]),
    "cfgpu_tuning_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Tuning)]),
    "cfgpu_tuning_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Tuning)]),
    "cfgpu_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.POINTER(TuneTrial), ctypes.c_uint32, u32p]),
    "cfgpu_cdbg": (ctypes.c_int, [ctypes.c_void_p]),
    "cfgpu_cdbg_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, u64p, ctypes.c_uint64,
                                       u64p, u64p]),
    "cfhost_cdbg_build": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(Unitig), ctypes.c_uint64,
                                         ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "cfhost_cdbg_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(u64p), u64p]),
    "cfhost_cdbg_write_fasta": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "cfhost_cdbg_free": (None, [ctypes.c_void_p]),
    "cfhost_input_from_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    "cfhost_input_from_buffer": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
                                                ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.POINTER(ctypes.c_void_p)]),
    "cfhost_input_from_records": (ctypes.c_int, [ctypes.c_char_p, u64p, ctypes.c_uint64,
                                                 ctypes.c_uint32, ctypes.c_uint32,
                                                 ctypes.POINTER(ctypes.c_void_p)]),
    "cfhost_input_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(u64p), u64p, u64p,
                                        ctypes.POINTER(u64p), u64p, u64p]),
    "cfhost_input_free": (None, [ctypes.c_void_p]),
    "cfhost_split_records": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                            ctypes.c_uint32, u64p, ctypes.c_uint64, u64p]),
    "cfhost_partition_buffer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
                                               ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(IngestStats)]),
    "cfhost_partition_file": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_uint32,
                                             ctypes.c_uint64, ctypes.POINTER(IngestStats)]),
    "cfhost_partition_buffer_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
                                                  ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                                  ctypes.POINTER(IngestStats)]),
    "cfhost_partition_file_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_uint32,
                                                ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(IngestStats)]),
    "cfhost_synth_genome": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                           ctypes.c_void_p]),
    "cfhost_synth_insert_repeats": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64,
                                                   ctypes.c_uint32, ctypes.c_uint32,
                                                   ctypes.c_double, ctypes.c_uint64,
                                                   ctypes.c_void_p]),
    "cfhost_synth_reads": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                          ctypes.c_uint32, ctypes.c_void_p]),
    "cfhost_synth_bucket_workload": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64,
                                                    ctypes.c_double, ctypes.c_uint32,
                                                    ctypes.c_uint32, ctypes.c_uint64, u32p, u64p,
                                                    u64p]),
    "cfhost_synth_mutate": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64,
                                           ctypes.c_void_p, ctypes.c_uint64, u64p]),
}

_LIB = None
sorted_fn = builtins.sorted


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libcfgpu.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def split_records(buf: bytes, fastq: bool, chunk_bytes: int, threads: int = 0):
    """cfhost_split_records: chunk boundaries (byte offsets) at record starts."""
    n = ctypes.c_uint64()
    if lib().cfhost_split_records(buf, len(buf), int(fastq), chunk_bytes, threads, None, 0, ctypes.byref(n)) != 0:
        raise RuntimeError("cfhost_split_records failed")
    off = np.zeros(n.value + 1, np.uint64)
    if lib().cfhost_split_records(buf, len(buf), int(fastq), chunk_bytes, threads, _p(off, u64p), off.size,
                                  ctypes.byref(n)) != 0:
        raise RuntimeError("cfhost_split_records failed")
    return off


# ---------------------------------------------------------------------------
class Input:
    """Packed fragment stream built by the host ingest (cfhost_input)."""

    def __init__(self, handle):
        self._h = handle
        L = lib()
        packed, frag = u64p(), u64p()
        nw, nb, nf, nr = (ctypes.c_uint64() for _ in range(4))
        L.cfhost_input_get(handle, ctypes.byref(packed), ctypes.byref(nw), ctypes.byref(nb),
                           ctypes.byref(frag), ctypes.byref(nf), ctypes.byref(nr))
        self.n_words, self.n_bases, self.n_frag, self.n_records = nw.value, nb.value, nf.value, nr.value
        self.packed = np.ctypeslib.as_array(packed, shape=(self.n_words,)) if self.n_words else np.zeros(2, np.uint64)
        self.frag_off = np.ctypeslib.as_array(frag, shape=(self.n_frag + 1,))

    @classmethod
    def from_buffer(cls, buf: bytes, fastq: bool, k: int, threads: int = 0):
        h = ctypes.c_void_p()
        r = lib().cfhost_input_from_buffer(buf, len(buf), int(fastq), k, threads, ctypes.byref(h))
        if r != 0:
            raise RuntimeError(f"cfhost_input_from_buffer failed: {r}")
        return cls(h)

    @classmethod
    def from_file(cls, path: str, fastq: bool, k: int, threads: int = 0):
        h = ctypes.c_void_p()
        r = lib().cfhost_input_from_file(path.encode(), int(fastq), k, threads, ctypes.byref(h))
        if r != 0:
            raise RuntimeError(f"cfhost_input_from_file failed: {r}")
        return cls(h)

    @classmethod
    def from_records(cls, seq: bytes, rec_off, k: int, threads: int = 0):
        ro = np.ascontiguousarray(rec_off, dtype=np.uint64)
        h = ctypes.c_void_p()
        r = lib().cfhost_input_from_records(seq, _p(ro, u64p), len(ro) - 1, k, threads,
                                            ctypes.byref(h))
        if r != 0:
            raise RuntimeError(f"cfhost_input_from_records failed: {r}")
        return cls(h)

    def close(self):
        if self._h:
            lib().cfhost_input_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cdbg_from_records(k, recs, n, buf, nb, sorted=True, decode=True):
    """cfhost_cdbg_build over Unitig records and their label buffer -> list of str
    (decode=False: (unitig count, total label bytes))."""
    c = ctypes.c_void_p()
    r = lib().cfhost_cdbg_build(k, recs, n, buf, nb, int(sorted), ctypes.byref(c))
    if r != 0:
        raise RuntimeError(f"cfhost_cdbg_build failed ({r})")
    try:
        seq, off, m = ctypes.c_void_p(), u64p(), ctypes.c_uint64()
        lib().cfhost_cdbg_get(c, ctypes.byref(seq), ctypes.byref(off), ctypes.byref(m))
        m = m.value
        if not decode:
            return m, int(off[m]) if m else 0
        o = np.ctypeslib.as_array(off, shape=(m + 1,)).copy() if m else np.zeros(1, np.uint64)
        raw = ctypes.string_at(seq, int(o[-1])) if m else b""
        return [raw[int(o[i]):int(o[i + 1])].decode() for i in range(m)]
    finally:
        lib().cfhost_cdbg_free(c)


def cdbg_from_unitigs(k, g, flags, labels, sorted=True):
    """cfhost_cdbg_build over unitigs given as arrays (g, flags) and a list of labels."""
    n = len(labels)
    recs = (Unitig * max(n, 1))()
    off = 0
    for i, lab in enumerate(labels):
        recs[i].label_off, recs[i].label_len = off, len(lab)
        recs[i].g, recs[i].flags = int(g[i]), int(flags[i])
        off += len(lab)
    buf = "".join(labels).encode()
    return cdbg_from_records(k, recs, n, buf, len(buf), sorted)


def synth_genome(n, repeat_frac=0.0, fam_min=300, fam_max=6000, divergence=0.02, seed=1):
    out = ctypes.create_string_buffer(n)
    r = lib().cfhost_synth_genome(n, repeat_frac, fam_min, fam_max, divergence, seed, out)
    if r != 0:
        raise RuntimeError("synth_genome failed")
    return out.raw[:n]


def synth_insert_repeats(base: bytes, n_copies, rep_len, sub_rate, seed):
    n = len(base) + n_copies * rep_len
    out = ctypes.create_string_buffer(n)
    r = lib().cfhost_synth_insert_repeats(base, len(base), n_copies, rep_len, sub_rate, seed, out)
    if r != 0:
        raise RuntimeError("synth_insert_repeats failed")
    return out.raw[:n]


def synth_reads(genome: bytes, n_reads, read_len, sub_rate, seed, threads=0):
    out = ctypes.create_string_buffer(n_reads * read_len)
    r = lib().cfhost_synth_reads(genome, len(genome), n_reads, read_len, sub_rate, seed, threads, out)
    if r != 0:
        raise RuntimeError("synth_reads failed")
    return out.raw[: n_reads * read_len]


def synth_mutate(base: bytes, snp_rate, indel_rate, indel_max, seed):
    cap = int(len(base) * (1 + indel_rate * indel_max) + 1024)
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_uint64()
    r = lib().cfhost_synth_mutate(base, len(base), snp_rate, indel_rate, indel_max, seed, out, cap,
                                  ctypes.byref(n))
    if r != 0:
        raise RuntimeError("synth_mutate failed")
    return out.raw[: n.value]


def synth_bucket_workload(n_buckets, n_records, zipf_s, k, l, seed):
    W = ((2 * (k - 1) - l + 2) + 31) // 32
    att = np.zeros(max(n_records, 1), np.uint32)
    lab = np.zeros(max(n_records, 1) * W, np.uint64)
    off = np.zeros(G + 1, np.uint64)
    r = lib().cfhost_synth_bucket_workload(n_buckets, n_records, zipf_s, k, l, seed, _p(att, u32p),
                                           _p(lab, u64p), _p(off, u64p))
    if r != 0:
        raise RuntimeError("synth_bucket_workload failed")
    return att[:n_records], lab[: n_records * W].reshape(n_records, W), off


# ---------------------------------------------------------------------------
class CfgpuError(RuntimeError):
    pass


class Handle:
    """One cfgpu handle: partition -> construct -> table views."""

    def __init__(self, k, l=20, cutoff=1, device=0):
        self.k, self.l, self.cutoff = k, l, cutoff
        self.kw = (k + 31) // 32
        self.W = ((2 * (k - 1) - l + 2) + 31) // 32
        p = Params(k, l, cutoff, G, device, 0)
        h = ctypes.c_void_p()
        r = lib().cfgpu_open(ctypes.byref(p), ctypes.byref(h))
        if r != 0:
            raise CfgpuError(f"cfgpu_open failed ({r}) for k={k} l={l} cutoff={cutoff}")
        self._h = h

    def _check(self, r, what):
        if r != 0:
            buf = ctypes.create_string_buffer(512)
            lib().cfgpu_last_error(self._h, buf, 512)
            raise CfgpuError(f"{what} failed ({r}): {buf.value.decode(errors='replace')}")

    def close(self):
        if getattr(self, "_h", None):
            lib().cfgpu_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def partition(self, inp: Input, append: bool = False, spill: bool = False):
        """append: keep the chunk's buckets aside; spill: move them to the bucket store."""
        flags = (2 if append else 0) | (4 if spill else 0)
        self._check(lib().cfgpu_partition(self._h, inp.packed.ctypes.data, inp.n_bases,
                                          inp.frag_off.ctypes.data, inp.n_frag, flags), "partition")

    def partition_finish(self):
        self._check(lib().cfgpu_partition_finish(self._h), "partition_finish")

    def spill_dir(self, path):
        """Bucket store location: None for pinned host memory, else a directory."""
        self._check(lib().cfgpu_spill_dir(self._h, None if path is None else str(path).encode()), "spill_dir")

    def restore(self, buckets):
        """cfgpu_buckets_restore: the listed buckets into device memory."""
        b = np.ascontiguousarray(buckets, dtype=np.uint32)
        self._check(lib().cfgpu_buckets_restore(self._h, _p(b, u32p) if b.size else None, len(b)), "restore")

    def spill_stats(self):
        st = SpillStat()
        self._check(lib().cfgpu_spill_stats(self._h, ctypes.byref(st)), "spill_stats")
        return st.as_dict()

    def ingest_buffer(self, buf: bytes, fastq: bool, threads: int = 0, chunk_bytes: int = 0, spill: bool = False):
        """cfhost_partition_buffer_ex: chunked parse + pack overlapped with the device partition
        (spill: every chunk's buckets to the bucket store)."""
        st = IngestStats()
        r = lib().cfhost_partition_buffer_ex(self._h, buf, len(buf), int(fastq), threads, chunk_bytes,
                                             4 if spill else 0, ctypes.byref(st))
        self._check(r, "ingest_buffer")
        return st.as_dict()

    def ingest_file(self, path: str, fastq: bool, threads: int = 0, chunk_bytes: int = 0):
        st = IngestStats()
        r = lib().cfhost_partition_file(self._h, path.encode(), int(fastq), threads, chunk_bytes, ctypes.byref(st))
        self._check(r, "ingest_file")
        return st.as_dict()

    def partition_device(self, packed_ptr, n_bases, frag_ptr, n_frag):
        self._check(lib().cfgpu_partition(self._h, packed_ptr, n_bases, frag_ptr, n_frag, 1),
                    "partition")

    def load_buckets(self, att, labels, bucket_off):
        a = np.ascontiguousarray(att, dtype=np.uint32)
        lb = np.ascontiguousarray(labels, dtype=np.uint64).reshape(-1)
        if lb.size == 0:
            lb = np.zeros(1, np.uint64)
        if a.size == 0:
            a = np.zeros(1, np.uint32)
        o = np.ascontiguousarray(bucket_off, dtype=np.uint64)
        self._check(lib().cfgpu_load_buckets(self._h, _p(a, u32p), _p(lb, u64p), _p(o, u64p)),
                    "load_buckets")

    def construct(self, mode="learned"):
        self._check(lib().cfgpu_construct(self._h, MODES.get(mode, mode)), "construct")

    def construct_buckets(self, buckets, mode="learned"):
        b = np.ascontiguousarray(buckets, dtype=np.uint32)
        if b.size == 0:
            b = np.zeros(1, np.uint32)[:0]
        self._check(lib().cfgpu_construct_buckets(self._h, MODES.get(mode, mode), _p(b, u32p) if b.size else None,
                                                  len(b)), "construct_buckets")

    def bucket_stats(self):
        arr = (BucketStat * G)()
        self._check(lib().cfgpu_bucket_stats(self._h, arr, G), "bucket_stats")
        out = {n: np.array([getattr(s, n) for s in arr]) for n, _ in BucketStat._fields_}
        return out

    def records(self):
        n = ctypes.c_uint64()
        lib().cfgpu_records_copy(self._h, None, None, None, 0, ctypes.byref(n))
        n = n.value
        att = np.zeros(max(n, 1), np.uint32)
        lab = np.zeros(max(n, 1) * self.W, np.uint64)
        off = np.zeros(G + 1, np.uint64)
        m = ctypes.c_uint64()
        self._check(lib().cfgpu_records_copy(self._h, _p(att, u32p), _p(lab, u64p), _p(off, u64p),
                                             n, ctypes.byref(m)), "records_copy")
        return att[:n], lab[: n * self.W].reshape(n, self.W), off

    def tables(self, buckets=None):
        """All occupied entries: dict g, keys (n x kw), e_f, status, dfa."""
        stats = self.bucket_stats()
        if buckets is None:
            buckets = np.nonzero(stats["n_records"])[0]
        b = np.ascontiguousarray(buckets, dtype=np.uint32)
        total = int(stats["n_vertices"][b].sum()) if len(b) else 0
        keys = np.zeros(max(total, 1) * self.kw, np.uint64)
        ef = np.zeros(max(total, 1), np.uint32)
        st = np.zeros(max(total, 1), np.uint8)
        dfa = np.zeros(max(total, 1), np.uint8)
        bn = np.zeros(max(len(b), 1), np.uint32)
        n = ctypes.c_uint64()
        self._check(lib().cfgpu_table_copy(self._h, _p(b, u32p), len(b), _p(keys, u64p),
                                           _p(ef, u32p), _p(st, u8p), _p(dfa, u8p), _p(bn, u32p),
                                           max(total, 1), ctypes.byref(n)), "table_copy")
        total = n.value
        g = np.repeat(b, bn[: len(b)])
        return {"g": g, "keys": keys[: total * self.kw].reshape(total, self.kw), "e_f": ef[:total],
                "status": st[:total], "dfa": dfa[:total]}

    def table_view(self, g):
        v = TableView()
        self._check(lib().cfgpu_table_view_get(self._h, g, ctypes.byref(v)), "table_view_get")
        return v

    def release(self, g):
        self._check(lib().cfgpu_release(self._h, g), "release")

    def contract(self):
        """Subgraph::contract of every built subgraph on the device."""
        self._check(lib().cfgpu_contract(self._h), "contract")

    def contract_stats(self):
        """Per-subgraph counters (G x 8, CONTRACT_STAT_FIELDS order)."""
        arr = np.zeros((G, 8), np.uint32)
        self._check(lib().cfgpu_contract_stats(self._h, arr.ctypes.data_as(ctypes.POINTER(ContractStat)), G),
                    "contract_stats")
        return arr

    def unitigs_raw(self):
        """(Unitig records array, label bytes) of the last contract."""
        n, nb = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().cfgpu_unitigs_copy(self._h, None, 0, None, 0, ctypes.byref(n), ctypes.byref(nb)),
                    "unitigs_copy")
        recs = (Unitig * max(n.value, 1))()
        buf = ctypes.create_string_buffer(max(nb.value, 1))
        self._check(lib().cfgpu_unitigs_copy(self._h, recs, n.value, buf, nb.value, ctypes.byref(n),
                                             ctypes.byref(nb)), "unitigs_copy")
        return recs, n.value, buf, nb.value

    def cdbg_device(self, sorted=True, decode=True):
        """Maximal unitigs of the whole graph glued on the device (cfgpu_cdbg)."""
        self._check(lib().cfgpu_cdbg(self._h), "cdbg")
        n, nb = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().cfgpu_cdbg_copy(self._h, None, 0, None, 0, ctypes.byref(n), ctypes.byref(nb)), "cdbg_copy")
        if not decode:
            return n.value, nb.value
        buf = ctypes.create_string_buffer(max(nb.value, 1))
        rec = np.zeros(max(2 * n.value, 2), np.uint64)
        self._check(lib().cfgpu_cdbg_copy(self._h, buf, nb.value, _p(rec, u64p), n.value, ctypes.byref(n),
                                          ctypes.byref(nb)), "cdbg_copy")
        raw = buf.raw
        out = [raw[int(rec[2 * i]):int(rec[2 * i] + rec[2 * i + 1])].decode() for i in range(n.value)]
        return sorted_fn(out) if sorted else out

    def cdbg(self, sorted=True):
        """Maximal unitigs of the whole graph (cfhost_cdbg_build over the last contract)."""
        recs, n, buf, nb = self.unitigs_raw()
        return cdbg_from_records(self.k, recs, n, buf, nb, sorted)

    def unitigs(self):
        """Unitigs of the last contract: dict g, flags (1 cycle, 2 left exit, 4 right exit), labels (str)."""
        n, nb = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().cfgpu_unitigs_copy(self._h, None, 0, None, 0, ctypes.byref(n), ctypes.byref(nb)),
                    "unitigs_copy")
        recs = (Unitig * max(n.value, 1))()
        buf = ctypes.create_string_buffer(max(nb.value, 1))
        self._check(lib().cfgpu_unitigs_copy(self._h, recs, n.value, buf, nb.value, ctypes.byref(n),
                                             ctypes.byref(nb)), "unitigs_copy")
        raw = buf.raw
        m = n.value
        return {"g": np.array([recs[i].g for i in range(m)], np.uint32),
                "flags": np.array([recs[i].flags for i in range(m)], np.uint32),
                "labels": [raw[recs[i].label_off:recs[i].label_off + recs[i].label_len].decode() for i in range(m)]}

    def run_stats(self):
        s = RunStats()
        self._check(lib().cfgpu_run_stats_get(self._h, ctypes.byref(s)), "run_stats")
        return s.as_dict()

    def sizing_model(self):
        """The trained MLP predictor: dict of its settings and arrays."""
        info = SizingModel()
        self._check(lib().cfgpu_sizing_model_get(self._h, ctypes.byref(info), None, None, None, None, None),
                    "sizing_model_get")
        params = np.zeros(info.n_params, np.float32)
        feat = np.zeros((G, info.n_features), np.float32)
        idx = np.zeros(max(info.n_train, 1), np.uint32)
        tgt = np.zeros(max(info.n_train, 1), np.float32)
        pred = np.zeros(G, np.float32)
        self._check(lib().cfgpu_sizing_model_get(self._h, ctypes.byref(info), params.ctypes.data, feat.ctypes.data,
                                                 _p(idx, u32p), tgt.ctypes.data, pred.ctypes.data),
                    "sizing_model_get")
        out = {n: getattr(info, n) for n, _ in SizingModel._fields_}
        out.update(params=params, features=feat, train_buckets=idx[:info.n_train], train_targets=tgt[:info.n_train],
                   predictions=pred)
        return out

    def tuning(self):
        t = Tuning()
        self._check(lib().cfgpu_tuning_get(self._h, ctypes.byref(t)), "tuning_get")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        lib().cfgpu_tuning_get(self._h, ctypes.byref(t))
        for k_, v in kw.items():
            setattr(t, k_, v)
        self._check(lib().cfgpu_tuning_set(self._h, ctypes.byref(t)), "tuning_set")

    def autotune(self, mode="learned", n=64):
        arr = (TuneTrial * n)()
        w = ctypes.c_uint32()
        self._check(lib().cfgpu_autotune(self._h, MODES.get(mode, mode), arr, n, ctypes.byref(w)),
                    "autotune")
        return [dict(arr[i].tuning.as_dict(), ms=arr[i].ms, identical=bool(arr[i].identical)) for i in range(w.value)]

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
