"""Benchmark of hora's BruteForceIndex search on one MI355X.

Workload (the configuration the headline QPS is quoted on): BruteForceIndex,
60,000 base x d=784, integer values 0-255 stored as f32 (each element 0 with
probability 0.5, else uniform in 1..255), 10,000 queries of the same form,
k=10.  One step = one search of all 10,000 queries (node_search_k for each)
through the C ABI with queries and outputs already resident in HBM.

Printed: one JSON line with QPS, per-metric QPS, the roofline of the dominant
kernel (timed live with HIP events on its launch stream) and the CPU
baseline (the oracle, a C port of the reference's scalar loop, timed on this
host on a bounded query sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_BASE, N_QUERY, DIM, K = 60_000, 10_000, 784, 10
VALU_PEAK_TOPS = 157.3  # MI355X published FP32 vector/matrix peak, TFLOP/s
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="exact", choices=["exact", "bf16", "fp8"])
    ap.add_argument("--cpu-sample", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--all-metrics", action="store_true", default=True)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import hora_amd
    from hora_amd import BruteForceIndex, Metric
    from hora_amd._lib import profile_enable, profile_query, profile_reset
    from hora_amd.datasets import bf_sparse_u8

    t0 = time.time()
    base = bf_sparse_u8(N_BASE, DIM, seed=20260101)
    queries = bf_sparse_u8(N_QUERY, DIM, seed=20260102)
    print(f"[bench] data generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    dq = torch.from_numpy(queries).to(dev)
    d_idx = torch.empty((N_QUERY, K), dtype=torch.int64, device=dev)
    d_dist = torch.empty((N_QUERY, K), dtype=torch.float32, device=dev)
    d_cnt = torch.empty((N_QUERY,), dtype=torch.int32, device=dev)

    mode_code = {"exact": hora_amd.HM_SEARCH_EXACT, "bf16": hora_amd.HM_SEARCH_BF16_RERANK,
                 "fp8": hora_amd.HM_SEARCH_FP8_RERANK}[args.mode]

    def make(mt):
        ix = BruteForceIndex(DIM)
        ix.add_batch(base)
        ix.build(mt)
        if mode_code != hora_amd.HM_SEARCH_EXACT:
            ix.set_search_mode(mode_code, 0)
        return ix

    def step(ix):
        ix.search_device(dq.data_ptr(), N_QUERY, K, d_idx.data_ptr(), d_dist.data_ptr(),
                         d_cnt.data_ptr(), stream.cuda_stream)

    def timed(ix, steps, warmup):
        for _ in range(warmup):
            step(ix)
        torch.cuda.synchronize()
        profile_reset()
        profile_enable(True)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record(stream)
        for i in range(steps):
            step(ix)
            print(f"[bench] step {i + 1}/{steps} done", file=sys.stderr, flush=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        profile_enable(False)
        return ev0.elapsed_time(ev1) / 1e3, wall

    # ---- headline: Euclidean
    ix = make(Metric.Euclidean)
    gpu_s, wall_s = timed(ix, args.steps, args.warmup)
    qps = args.steps * N_QUERY / wall_s
    ms_step = wall_s / args.steps * 1e3
    kern_ms, kern_n = profile_query("exact_tile")
    sel_ms, sel_n = profile_query("row_select")
    fin_ms, fin_n = profile_query("bf_finalize")
    gpu_ids = d_idx.cpu().numpy().view(np.uint64).copy()
    fallback = ix.search_stats() if args.mode != "exact" else (0, 0)
    print(f"[bench] Euclidean: {qps:.0f} QPS, {ms_step:.2f} ms/step", file=sys.stderr, flush=True)

    gemm_ms, gemm_n = profile_query("score_gemm")
    rr_ms, rr_n = profile_query("rerank")
    roofline = None
    if args.mode == "exact" and kern_n:
        avg_s = kern_ms / kern_n / 1e3
        rows_per_launch = N_QUERY * args.steps / kern_n
        flops = 3.0 * rows_per_launch * N_BASE * DIM  # sub, mul, add per element pair
        achieved = flops / avg_s / 1e12
        roofline = {"bound": "valu", "kernel": "exact_tile_kernel<Euclidean>",
                    "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                    "traffic": None, "avg_launch_ms": round(kern_ms / kern_n, 3),
                    "launches": kern_n,
                    "note": "exact reference-order f32: 3 non-FMA VALU ops per element pair; "
                            "without FMA or packed-f32 the VALU issue ceiling is 39.3 T op/s"}
    elif gemm_n:
        avg_s = gemm_ms / gemm_n / 1e3
        rows_per_launch = N_QUERY * args.steps / gemm_n
        dk = (DIM + 31) // 32 * 32 if args.mode == "bf16" else (DIM + 63) // 64 * 64
        flops = 2.0 * rows_per_launch * N_BASE * dk
        peak = 2500.0 if args.mode == "bf16" else 5000.0
        achieved = flops / avg_s / 1e12
        roofline = {"bound": "mfma", "kernel": f"score_gemm_kernel<{args.mode}>",
                    "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(achieved / peak, 4), "traffic": None,
                    "avg_launch_ms": round(gemm_ms / gemm_n, 3), "launches": gemm_n}

    per_metric = {"Euclidean": round(qps, 1)}
    if args.all_metrics:
        for mt in (Metric.Manhattan, Metric.DotProduct, Metric.CosineSimilarity):
            ixm = make(mt)
            _, w = timed(ixm, 1, 1)
            per_metric[mt.name] = round(N_QUERY / w, 1)
            print(f"[bench] {mt.name}: {per_metric[mt.name]} QPS", file=sys.stderr, flush=True)
            del ixm

    cpu = None
    recall = None
    if not args.no_cpu_baseline:
        import oracle

        nthreads = min(os.cpu_count() or 1, 16)
        s = args.cpu_sample
        c0 = time.perf_counter()
        ri, rd, rc = oracle.bf_search(base, queries[:s], K, oracle.EUCLIDEAN, nthreads=nthreads)
        c_all = time.perf_counter() - c0
        s1 = max(4, s // 20)
        c0 = time.perf_counter()
        oracle.bf_search(base, queries[:s1], K, oracle.EUCLIDEAN, nthreads=1)
        c_one = time.perf_counter() - c0
        cpu = {"value": round(s / c_all, 2), "unit": "QPS", "cores": nthreads, "kind": "port",
               "sample": f"first {s} of the 10,000 queries, Euclidean, k=10, full 60,000-row "
                         f"scan each ({nthreads} threads, query-parallel)",
               "single_thread_qps": round(s1 / c_one, 2),
               "single_thread_sample": f"first {s1} queries, 1 thread"}
        recall = float(np.mean([len(set(gpu_ids[r].tolist()) & set(ri[r].tolist())) / K
                                for r in range(s)]))
        exact_ids = bool(np.array_equal(gpu_ids[:s], ri))
        cpu["gpu_ids_bit_exact_on_sample"] = exact_ids

    line = {
        "metric": "QPS@recall10>=0.95, 1 MI355X: FP32 vs BF16/FP8+rerank vs learned probe; "
                  "%roofline",
        "value": round(qps, 1),
        "unit": "QPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded): 0-255 integers as f32, 50% zeros",
        "config": {"workload": "BruteForceIndex Euclidean 60,000x784 base, 10,000 queries, "
                               "k=10", "form": f"{args.mode}", "global_batch": N_QUERY,
                   "parallelism": "single GPU"},
        "recall_at_10": recall,
        "per_metric_qps": per_metric,
        "exact_fallback_rows": {"rerun": fallback[0], "uncertified": fallback[1]},
        "kernel_ms_per_step": {"exact_tile": round(kern_ms / max(args.steps, 1), 3),
                               "score_gemm": round(gemm_ms / max(args.steps, 1), 3),
                               "rerank": round(rr_ms / max(args.steps, 1), 3),
                               "row_select": round(sel_ms / max(args.steps, 1), 3),
                               "bf_finalize": round(fin_ms / max(args.steps, 1), 3)},
        "gpu_event_s": round(gpu_s, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
