"""Benchmark: k-mers/s from device-resident packed reads to per-subgraph vertex
states (partition + construct) on one MI355X.

Workload (default): 150 bp reads sampled uniformly from both strands of a 50 Mbp
iid-ACGT genome at 30x (10,000,000 reads), 0.5% iid substitutions, k=31,
cutoff 2, l=20 — synthetic, seeded.  A step = cfgpu_partition over the packed
reads already in HBM + cfgpu_construct (batched persistent launch, learned
table sizing).  The plain GPU path and the batched path without the learned
sizing are timed beside it; the CPU baseline is the oracle pipeline on a
bounded sample of the same reads on the host's cores.

Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cuttlefish-private_amd", "python"))

import cfgpu as C  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "partition_construct_reads10M_k31_hbm_traffic.txt")
ROOFLINE_KERNEL = "k_construct_lds<1, 2>"


def measured_traffic():
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch " + ROOFLINE_KERNEL + ":"):
                    return float(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class Hip:
    """hipMalloc / hipMemcpy through the HIP runtime (device input buffers)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipSetDevice.argtypes = [ctypes.c_int]

    def upload(self, arr):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), arr.nbytes) == 0, "hipMalloc failed"
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0, "hipMemcpy failed"
        return p

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p):
        self.L.hipFree(p)

    def copy_d2d(self, nbytes):
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(a), nbytes) == 0 and self.L.hipMalloc(ctypes.byref(b), nbytes) == 0
        assert self.L.hipMemcpy(b, a, nbytes, 3) == 0
        self.sync()
        self.free(a)
        self.free(b)


def algorithmic_bytes(k, l, n_sk, n_e, n_km):
    W = ((2 * (k - 1) - l + 2) + 31) // 32
    S = 4 + 8 * W
    Ke = 8 * ((k + 1 + 31) // 32)
    Kv = 8 * ((k + 31) // 32)
    return n_sk * S + n_e * 2 * (Ke + 4) + n_km * 2 * (Kv + 4)


def make_workload(args):
    t = time.time()
    g = C.synth_genome(args.genome, 0.0, 300, 6000, 0.0, seed=args.seed)
    reads = C.synth_reads(g, args.reads, args.read_len, args.sub_rate, seed=args.seed + 1)
    rec_off = np.arange(args.reads + 1, dtype=np.uint64) * np.uint64(args.read_len)
    inp = C.Input.from_records(reads, rec_off, args.k)
    log(f"workload ready in {time.time() - t:.1f}s: {inp.n_bases} bases in {inp.n_frag} fragments")
    return reads, rec_off, inp


def timed_steps(hip, h, dev, inp, mode, steps):
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.partition_device(dev[0].value, inp.n_bases, dev[1].value, inp.n_frag)
        h.construct(mode)
    hip.sync()
    return (time.perf_counter() - t0) * 1000.0 / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--genome", type=int, default=50_000_000)
    ap.add_argument("--reads", type=int, default=10_000_000)
    ap.add_argument("--read-len", type=int, default=150)
    ap.add_argument("--sub-rate", type=float, default=0.005)
    ap.add_argument("--k", type=int, default=31)
    ap.add_argument("--l", type=int, default=20)
    ap.add_argument("--cutoff", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--compare-steps", type=int, default=2)
    ap.add_argument("--cpu-sample-reads", type=int, default=3_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before the steps, copy this many MiB device-to-device (a known-size "
                         "copy for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU design shards buckets by subgraph id, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    reads, rec_off, inp = make_workload(args)
    hip = Hip()
    hip.L.hipSetDevice(0)
    dev = (hip.upload(inp.packed), hip.upload(inp.frag_off))

    h = C.Handle(args.k, args.l, args.cutoff)
    if args.calib_copy_mib:
        log(f"calibration copy {args.calib_copy_mib} MiB")
        hip.copy_d2d(args.calib_copy_mib << 20)
    log("warmup")
    for _ in range(args.warmup):
        h.partition_device(dev[0].value, inp.n_bases, dev[1].value, inp.n_frag)
        h.construct("learned")
    log("timed learned steps")
    ms = timed_steps(hip, h, dev, inp, "learned", args.steps)
    rs = h.run_stats()
    st = h.bucket_stats()
    n_vert_learned = st["n_vertices"].copy()
    n_km = rs["n_kmers"]
    value = n_km / (ms / 1000.0)

    # kernel-only timing of the main construct launch (direct launch, events)
    h.set_tuning(use_graph=0)
    h.partition_device(dev[0].value, inp.n_bases, dev[1].value, inp.n_frag)
    h.construct("learned")
    rk = h.run_stats()
    stk = h.bucket_stats()
    h.set_tuning(use_graph=1)
    main = (stk["flags"] & 4) == 0
    # the records the main round walks (distinct records after dedup)
    n_sk_m = int(stk["n_records_walked"][main].sum())
    n_km_m = int(stk["n_kmers_walked"][main].sum())
    n_e_m = n_km_m - n_sk_m
    kbytes = algorithmic_bytes(args.k, args.l, n_sk_m, n_e_m, n_km_m)
    kms = rk["t_construct_kernel_ms"]
    achieved = kbytes / (kms / 1000.0) / 1e9 if kms > 0 else 0.0
    pbytes = inp.n_bases / 4 + rs["n_records"] * (4 + 8 * (((2 * (args.k - 1) - args.l + 2) + 31) // 32))
    part_ms = rs["t_minimizer_ms"] + rs["t_count_ms"] + rs["t_emit_ms"]

    # Subgraph::contract of every subgraph on the tables just built (the first
    # call sizes its output buffers)
    log("contract")
    h.contract()
    ct = []
    for _ in range(3):
        h.contract()
        ct.append(h.run_stats()["t_contract_ms"])
    rc = h.run_stats()
    cst = h.contract_stats()
    ct_ms = sorted(ct)[len(ct) // 2]
    # the compacted graph: lm-tigs glued on the device (the first call sizes its
    # buffers), then the maximal unitigs copied to the host
    h.cdbg_device(decode=False)
    cd = []
    for _ in range(3):
        h.cdbg_device(decode=False)
        cd.append(h.run_stats()["t_cdbg_ms"])
    rcd = h.run_stats()
    cdbg_ms = sorted(cd)[len(cd) // 2]
    hip.sync()
    t0 = time.perf_counter()
    n_cdbg, cdbg_bytes = rcd["n_cdbg_unitigs"], rcd["n_cdbg_bytes"]
    seq_buf = ctypes.create_string_buffer(max(cdbg_bytes, 1))
    rec_buf = np.zeros(max(2 * n_cdbg, 2), np.uint64)
    C.lib().cfgpu_cdbg_copy(h._h, seq_buf, cdbg_bytes, rec_buf.ctypes.data_as(C.u64p), n_cdbg, None, None)
    copy_ms = (time.perf_counter() - t0) * 1000.0
    del seq_buf, rec_buf
    contract = {"ms": ct_ms, "unitigs": rc["n_unitigs"], "label_bytes": rc["n_unitig_bytes"],
                "global_table_subgraphs": rc["n_contract_global"], "runs": rc["n_contract_runs"],
                "isolated": int(cst[:, 0].sum()), "phantom": int(cst[:, 1].sum()),
                "trivial": int(cst[:, 2].sum()), "cycles": int(cst[:, 3].sum()),
                "lmtigs": int(cst[:, 4].sum()), "vertices_in_unitigs": int(cst[:, 5].sum()),
                "cdbg": {"maximal_unitigs": n_cdbg, "label_bytes": cdbg_bytes, "device_ms": cdbg_ms,
                         "rounds": rcd["n_cdbg_rounds"], "cycle_pieces_on_host": rcd["n_cdbg_cycle_pieces"],
                         "copy_to_host_ms": copy_ms}}

    # host buffers in (the packed reads cross PCIe inside the step)
    hip.sync()
    t0 = time.perf_counter()
    h.partition(inp)
    h.construct("learned")
    hip.sync()
    ms_host = (time.perf_counter() - t0) * 1000.0

    ms_plain = ms_batched = float("nan")
    rs_plain = {"n_overflowed": None, "n_fallback_rounds": None}
    same_plain = same_batched = None
    if args.compare_steps > 0:
        log("plain and batched comparison")
        ms_plain = timed_steps(hip, h, dev, inp, "plain", args.compare_steps)
        rs_plain = h.run_stats()
        same_plain = bool((h.bucket_stats()["n_vertices"] == n_vert_learned).all())
        ms_batched = timed_steps(hip, h, dev, inp, "batched", args.compare_steps)
        same_batched = bool((h.bucket_stats()["n_vertices"] == n_vert_learned).all())

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import oracle_lib as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        nr = min(args.cpu_sample_reads, args.reads)
        sample = reads[: nr * args.read_len]
        log(f"cpu baseline: {nr} reads on {threads} threads")
        t0 = time.perf_counter()
        r = O.pipeline(sample, rec_off[: nr + 1], args.k, args.l, threads)
        dt = time.perf_counter() - t0
        cpu = {"value": r["n_kmers"] / dt, "unit": "k-mers/s", "cores": threads, "kind": "port",
               "sample": f"first {nr} of the {args.reads} reads ({r['n_kmers']} k-mer instances): "
                         f"oracle partition + bucketize + per-subgraph construct, "
                         f"{r['partition_ms']:.0f}/{r['bucketize_ms']:.0f}/{r['construct_ms']:.0f} ms"}

    hip.free(dev[0])
    hip.free(dev[1])
    h.close()
    traffic = measured_traffic()
    default_workload = (args.genome, args.reads, args.read_len, args.sub_rate, args.k, args.l, args.cutoff,
                        args.seed) == (50_000_000, 10_000_000, 150, 0.005, 31, 20, 2, 2026)
    out = {
        "metric": "k-mers/sec to per-subgraph vertex states (partition + construct), batched+learned",
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"reads: {args.reads} x {args.read_len} bp from a {args.genome} bp iid genome "
                               f"(both strands, {args.sub_rate * 100:.1f}% substitutions), k={args.k}, "
                               f"l={args.l}, cutoff={args.cutoff}, 16384 subgraphs",
                   "n_kmers": n_km, "n_records": rs["n_records"], "n_vertices": rs["n_vertices"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic if default_workload else None,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, REPO) if traffic and default_workload else None,
                     "kernel": "construct main round (k_construct_lds, plus k_construct for buckets "
                               "too large for LDS passes)", "kernel_ms": kms,
                     "algorithmic_bytes": kbytes},
        "cpu_baseline": cpu,
        "paths": {
            "batched_learned": {"ms_per_step": ms, "kmers_per_s": value},
            "plain": {"ms_per_step": ms_plain, "kmers_per_s": n_km / (ms_plain / 1000.0),
                      "same_vertex_counts": same_plain, "overflowed": rs_plain["n_overflowed"],
                      "retry_rounds": rs_plain["n_fallback_rounds"]},
            "batched_upper_bound": {"ms_per_step": ms_batched, "kmers_per_s": n_km / (ms_batched / 1000.0),
                                    "same_vertex_counts": same_batched},
            "speedup_vs_plain": ms_plain / ms,
            "with_contract": {"ms_per_step": ms + ct_ms, "kmers_per_s": n_km / ((ms + ct_ms) / 1000.0),
                              "note": "the step plus Subgraph::contract of every subgraph (unitig labels on "
                                      "the device)"},
            "to_cdbg": {"ms_per_step": ms + ct_ms + cdbg_ms,
                        "kmers_per_s": n_km / ((ms + ct_ms + cdbg_ms) / 1000.0),
                        "with_copy_to_host_ms": ms + ct_ms + cdbg_ms + copy_ms,
                        "note": "the step, contract and the gluing of lm-tigs into every maximal unitig of the "
                                "graph, all on the device (the copy of the unitigs to the host apart)"},
            "host_input": {"ms_per_step": ms_host, "kmers_per_s": n_km / (ms_host / 1000.0),
                           "note": "cfgpu_partition from host buffers: includes the H2D copy of the packed reads"},
        },
        "phases_ms": {"partition": rs["t_partition_ms"], "minimizer": rs["t_minimizer_ms"],
                      "count": rs["t_count_ms"], "emit": rs["t_emit_ms"], "features": rs["t_features_ms"],
                      "construct": rs["t_construct_ms"], "sample": rs["t_sample_ms"],
                      "train": rs["t_train_ms"], "infer": rs["t_infer_ms"],
                      "construct_main": rs["t_construct_main_ms"], "fallback": rs["t_fallback_ms"],
                      "dedup": rs["t_dedup_ms"]},
        "contract": contract,
        "partition_roofline": {"achieved": pbytes / (part_ms / 1000.0) / 1e9 if part_ms else 0,
                               "unit": "GB/s", "algorithmic_bytes": pbytes},
        "learned": {"samples": rs["n_samples"], "mlp_loss": rs["mlp_loss"],
                    "mispredict_mean_abs_log": rs["mispredict_mean_abs_log"],
                    "mispredict_under_frac": rs["mispredict_under_frac"],
                    "fallback_rate": rs["fallback_rate"], "overflowed": rs["n_overflowed"],
                    "lds_buckets": rs["n_lds_buckets"],
                    "records_walked": rs["n_records_walked"], "dedup_ms": rs["t_dedup_ms"],
                    "dedup_fallback_buckets": rs["n_dedup_fallback"],
                    "predictor": {0: "none", 1: "mlp", 2: "constant ratio"}.get(rs["predictor"]),
                    "holdout_err_mlp": rs["holdout_err_mlp"], "holdout_err_const": rs["holdout_err_const"],
                    "table_bytes": rs["table_bytes"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
