"""Benchmarks of hora's distance-and-scan path on one MI355X.

Default workload (the configuration the headline QPS is quoted on):
BruteForceIndex, 60,000 base x d=784, integer values 0-255 stored as f32 (each
element 0 with probability 0.5, else uniform in 1..255), 10,000 queries of the
same form, k=10.  One step = one search of all 10,000 queries through the C
ABI (hm_index_search_device) with queries and outputs resident in HBM.  The
three search forms are timed -- exact FP32, BF16 MFMA shortlist + exact
rerank, FP8 shortlist + exact rerank -- and `value` is the fastest form whose
recall@10 on the checked sample is >= 0.95.

Other workloads (--workload): ivfpq (IVFPQIndex 1M x 128), ragged (ragged
query batches over 1M x 768), probe (learned nprobe over 5M x 96) and pq
(standalone PQIndex 100k x 128); each prints its own JSON line.

The roofline of the dominant kernel is timed live with HIP events that the
library records on the launch stream; the CPU baseline is the oracle (a C
port of the reference's scalar loop) on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS@recall10>=0.95, 1 MI355X: FP32 vs BF16/FP8+rerank vs learned probe; %roofline"
FP32_PEAK = 157.3    # TFLOP/s, MI355X published FP32
BF16_PEAK = 2500.0   # TFLOP/s dense
FP8_PEAK = 5000.0
HBM_PEAK = 8000.0    # GB/s
LDS_PEAK = 256 * 128 * 2.4  # GB/s: 256 CUs x 128 B/clk x 2.4 GHz


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bf",
                    choices=["bf", "ivfpq", "ragged", "probe", "copy", "pq", "knn_graph",
                             "distances"])
    ap.add_argument("--forms", default="exact,bf16,fp8")
    ap.add_argument("--cpu-sample", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-metrics", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="bf: run hm_index_autotune on 2,000 of the queries before timing")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Timer:
    """Times steps on the caller's stream with events, plus wall clock."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def run(self, fn, steps, warmup, label, only=None):
        """Warm-up, then `steps` timed calls; the library's HIP events time
        only the `only` family inside the timed region (two events per
        launch of it, so the step time carries no instrumentation of the
        other kernels).  One more call with every family timed then gives
        the per-kernel breakdown (kernel_ms reports it per step)."""
        from hora_amd._lib import profile_enable, profile_only, profile_reset
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        profile_reset()
        profile_only(only)
        profile_enable(only is not None)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(self.stream)
        for i in range(steps):
            fn()
            log(f"{label}: step {i + 1}/{steps} done")
        e1.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        profile_enable(False)
        self.timed = {only: kernel_ms([only])[only]} if only else {}
        # the breakdown: one more call, every family timed, scaled to steps
        profile_reset()
        profile_only(None)
        profile_enable(True)
        fn()
        torch.cuda.synchronize()
        profile_enable(False)
        self.scale = steps
        return wall, e0.elapsed_time(e1) / 1e3


def measured_traffic(form):
    """HBM bytes per launch of the filtered scan, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in profiles/score_filter_hbm_traffic.md (None if
    absent)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "score_filter_hbm_traffic.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == form:
                return int(f[2])
    except OSError:
        pass
    return None


def kernel_ms(names, timer=None):
    """(total ms, launches) per family.  With a Timer: the breakdown call
    after its timed region, scaled to the region's steps, and the timed
    family's own figures from the timed region itself."""
    from hora_amd._lib import profile_query
    out = {}
    for n in names:
        ms, cnt = profile_query(n)
        if timer is not None:
            ms, cnt = ms * timer.scale, cnt * timer.scale
        out[n] = (ms, cnt)
    if timer is not None:
        out.update(timer.timed)
    return out


def main_bf(args):
    import numpy as np
    import torch

    import hora_amd
    from hora_amd import BruteForceIndex, Metric
    from hora_amd.datasets import bf_sparse_u8

    N, NQ, D, K = 60_000, 10_000, 784, 10
    t0 = time.time()
    base = bf_sparse_u8(N, D, seed=20260101)
    queries = bf_sparse_u8(NQ, D, seed=20260102)
    log(f"data generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    dq = torch.from_numpy(queries).to(dev)
    d_idx = torch.empty((NQ, K), dtype=torch.int64, device=dev)
    d_dist = torch.empty((NQ, K), dtype=torch.float32, device=dev)
    d_cnt = torch.empty((NQ,), dtype=torch.int32, device=dev)
    timer = Timer(torch, stream)
    codes = {"exact": hora_amd.HM_SEARCH_EXACT, "bf16": hora_amd.HM_SEARCH_BF16_RERANK,
             "fp8": hora_amd.HM_SEARCH_FP8_RERANK}

    def make(mt, form):
        ix = BruteForceIndex(D)
        ix.add_batch(base)
        ix.build(mt)
        ix.set_search_mode(codes[form], 0)
        return ix

    def step_fn(ix):
        return lambda: ix.search_device(dq.data_ptr(), NQ, K, d_idx.data_ptr(),
                                        d_dist.data_ptr(), d_cnt.data_ptr(), stream.cuda_stream)

    forms = {}
    ids_by_form = {}
    for form in args.forms.split(","):
        ix = make(Metric.Euclidean, form)
        tuning = None
        if args.autotune and form != "exact":
            # hm_index_autotune on the first 2,000 queries: depth, then tile order
            ix.autotune(queries[:2000], K)
            tuning = ix.tuning()
            log(f"Euclidean/{form}: tuned {tuning}")
        wall, ev = timer.run(step_fn(ix), args.steps, args.warmup, f"Euclidean/{form}",
                             only="exact_tile" if form == "exact" else "score_filter")
        km = kernel_ms(["exact_tile", "exact_rows", "row_select", "kth_smallest", "bf_finalize",
                        "score_gemm", "score_filter", "list_gather", "list_prepare", "rerank"], timer)
        fb = ix.search_stats()
        ids_by_form[form] = (d_idx.cpu().numpy().view(np.uint64).copy(),
                             d_dist.cpu().numpy().copy(),
                             d_cnt.cpu().numpy().view(np.uint32).copy())
        forms[form] = {"qps": round(args.steps * NQ / wall, 1),
                       "ms_per_step": round(wall / args.steps * 1e3, 3),
                       "kernel_ms_per_step": {k: round(v[0] / args.steps, 3)
                                              for k, v in km.items() if v[1]},
                       "exact_fallback_rows": {"rerun": fb[0], "uncertified": fb[1]},
                       "_km": km}
        if tuning is not None:
            forms[form]["tuning"] = tuning
        log(f"Euclidean/{form}: {forms[form]['qps']} QPS")
        del ix

    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        hc = host_cpu()
        nthreads = hc["threads"]
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
               "single_thread_sample": f"first {s1} queries, 1 thread",
               "host": hc}
        for form, (ids, dists, cnts) in ids_by_form.items():
            rec = float(np.mean(
                [len(set(ids[r].tolist()) & set(ri[r].tolist())) / K for r in range(s)]))
            forms[form]["recall_at_10_sample"] = rec
            forms[form]["recall_loss_sample"] = round(1.0 - rec, 6)
            forms[form]["ids_bit_exact_on_sample"] = bool(np.array_equal(ids[:s], ri))
            forms[form]["fp64_error_sample"] = fp64_error(base, queries[:s], ids[:s], dists[:s],
                                                          cnts[:s], Metric.Euclidean)

    ok = [f for f in forms if forms[f].get("recall_at_10_sample", 1.0) >= 0.95]
    best = max(ok or forms, key=lambda f: forms[f]["qps"])
    b = forms[best]
    km = b["_km"]
    roofline = None
    if best == "exact" and km["exact_tile"][1]:
        avg = km["exact_tile"][0] / km["exact_tile"][1] / 1e3
        work = 3.0 * (NQ * args.steps / km["exact_tile"][1]) * N * D
        ach = work / avg / 1e12
        roofline = {"bound": "valu", "kernel": "exact_tile_kernel<Euclidean>",
                    "achieved": round(ach, 3), "peak": FP32_PEAK, "unit": "TFLOP/s",
                    "frac": round(ach / FP32_PEAK, 4), "traffic": None,
                    "avg_launch_ms": round(avg * 1e3, 3)}
    else:
        # the dominant MFMA kernel: the filtered full scan when the filter
        # form ran, else the full key-matrix GEMM; algorithmic work is
        # 2 * queries * rows * D (the unpadded dimension)
        kname = "score_filter" if km["score_filter"][1] else "score_gemm"
        if km[kname][1]:
            avg = km[kname][0] / km[kname][1] / 1e3
            work = 2.0 * (NQ * args.steps / km[kname][1]) * N * D
            peak = BF16_PEAK if best == "bf16" else FP8_PEAK
            ach = work / avg / 1e12
            roofline = {"bound": "mfma", "kernel": f"{kname}_kernel<{best}>",
                        "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
                        "frac": round(ach / peak, 4),
                        "traffic": measured_traffic(best) if kname == "score_filter" else None,
                        "avg_launch_ms": round(avg * 1e3, 3)}

    per_metric = {"Euclidean": b["qps"]}
    if not args.no_metrics:
        for mt in (Metric.Manhattan, Metric.DotProduct, Metric.CosineSimilarity, Metric.Angular):
            ixm = make(mt, best)
            wall, _ = timer.run(step_fn(ixm), 1, 1, f"{mt.name}/{best}")
            per_metric[mt.name] = round(NQ / wall, 1)
            del ixm
    for f in forms.values():
        f.pop("_km", None)
    line = {
        "metric": METRIC, "value": b["qps"], "unit": "QPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if best == "exact" else
        f"{best} shortlist + f32 rerank", "data": "synthetic (seeded): 0-255 integers as f32, "
        "50% zeros",
        "config": {"workload": "BruteForceIndex Euclidean 60,000x784 base, 10,000 queries, k=10",
                   "form": best, "global_batch": NQ, "parallelism": "single GPU"},
        "forms": forms, "per_metric_qps": per_metric, "roofline": roofline,
        "cpu_baseline": cpu,
        "reference_published_qps": {"value": 33, "note": "hora BruteForceIndex on real "
                                    "fashion-mnist-784, aws t2.medium (2 vCPU), read off the "
                                    "README plot; not the same data"},
    }
    print(json.dumps(line), flush=True)


def main_ivfpq(args):
    import numpy as np
    import torch

    from hora_amd import BruteForceIndex, IVFPQIndex, IVFPQParams, Metric
    import hora_amd
    from hora_amd.datasets import gaussian_mixture

    N, NQ, D, K = 1_000_000, 10_000, 128, 100
    epochs = int(os.environ.get("HM_IVF_EPOCHS", "10"))
    t0 = time.time()
    x, cen, _ = gaussian_mixture(N, D, 1024, seed=303, sigma=0.3)
    q, _, _ = gaussian_mixture(NQ, D, 1024, seed=304, sigma=0.3, centres=cen)
    log(f"data generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    p = (IVFPQParams().n_sub(16).sub_bits(8).n_kmeans_center(1024).search_n_center(8)
         .train_epoch(epochs).seed(7).kmeans_accumulate(1))
    ix = IVFPQIndex(D, p)
    ix.add_batch(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.build(Metric.Euclidean)
    build_s = time.perf_counter() - t0
    log(f"IVFPQ build (nlist 1024, M 16x256, {epochs} epochs): {build_s:.1f}s")
    # ground truth: exact top-100 by the certified BF16 brute force
    bf = BruteForceIndex(D)
    bf.add_batch(x)
    bf.build(Metric.Euclidean)
    bf.set_search_mode(hora_amd.HM_SEARCH_BF16_RERANK, 0)
    truth, _, _ = bf.search_batch(q, K)
    del bf
    dq = torch.from_numpy(q).to(dev)
    d_idx = torch.empty((NQ, K), dtype=torch.int64, device=dev)
    d_dist = torch.empty((NQ, K), dtype=torch.float32, device=dev)
    d_cnt = torch.empty((NQ,), dtype=torch.int32, device=dev)
    timer = Timer(torch, stream)
    res = {}
    R = int(os.environ.get("HM_RERANK", "200"))
    sweep = [(0, npb) for npb in (8, 16, 32, 64)] + [(R, npb) for npb in (1, 2, 4, 8, 16)]
    for rr, nprobe in sweep:
        ix.set_rerank(rr)
        ix.set_search_n_center(nprobe)
        wall, _ = timer.run(lambda: ix.search_device(dq.data_ptr(), NQ, K, d_idx.data_ptr(),
                                                     d_dist.data_ptr(), d_cnt.data_ptr(),
                                                     stream.cuda_stream),
                            args.steps, args.warmup, f"rerank={rr} nprobe={nprobe}",
                            only="ivf_scan")
        km = kernel_ms(["ivf_scan", "ivf_merge", "ivf_replay", "exact_tile", "row_select",
                        "rerank"], timer)
        replayed = int(ix.search_stats()[0])
        probes, counts = ix.last_probes(NQ, nprobe + 1)
        got = d_idx.cpu().numpy().view(np.uint64)
        r10 = float(np.mean([len(set(got[i, :10].tolist()) & set(truth[i, :10].tolist())) / 10
                             for i in range(NQ)]))
        r100 = float(np.mean([len(set(got[i].tolist()) & set(truth[i].tolist())) / K
                              for i in range(NQ)]))
        key = f"rerank{rr}/nprobe{nprobe}" if rr else f"adc/nprobe{nprobe}"
        res[key] = {"qps": round(args.steps * NQ / wall, 1), "recall_at_10": round(r10, 4),
                       "recall_at_100": round(r100, 4),
                       "kernel_ms_per_step": {k: round(v[0] / args.steps, 3)
                                              for k, v in km.items() if v[1]},
                       "tie_replayed_queries": replayed,
                       "_scan": km["ivf_scan"], "_probes": probes[:, :nprobe]}
        log(f"{key}: {res[key]['qps']} QPS, recall@10 {r10:.3f}")
    ok = [p_ for p_ in res if res[p_]["recall_at_10"] >= 0.95]
    if ok:
        pick = max(ok, key=lambda p_: res[p_]["qps"])
    else:  # the fastest setting at the best recall
        best_r = max(v["recall_at_10"] for v in res.values())
        pick = max((p_ for p_ in res if res[p_]["recall_at_10"] >= best_r - 1e-4),
                   key=lambda p_: res[p_]["qps"])
    scan_ms, scan_n = res[pick]["_scan"]
    roofline = None
    if scan_n:
        info = ix.export()
        sizes = np.diff(info["list_offsets"].astype(np.int64))
        pr = res[pick]["_probes"]
        pairs = float(pr.size)
        per_launch = NQ * args.steps / scan_n / NQ
        # the row-form scan is bound by the LUT fold on the VALU: per (query,
        # probed list) pair, 16 x 256 codewords x 8 dims x 3 non-FMA ops
        # (sub, mul, add in the reference's order); the codes it streams
        # (16 B per item, plus 4 B of ADC distance written) are far below HBM
        ops = per_launch * pairs * 16 * 256 * 8 * 3
        avg_s = scan_ms / scan_n / 1e3
        ach = ops / avg_s / 1e12
        bytes_ = per_launch * float(sizes[pr].sum()) * (16 + 4)
        roofline = {"bound": "valu", "kernel": "ivf_adc_rows_kernel", "achieved": round(ach, 2),
                    "peak": 78.6, "unit": "T op/s (non-FMA f32)", "frac": round(ach / 78.6, 4),
                    "hbm_gbs_of_codes": round(bytes_ / avg_s / 1e9, 1), "traffic": None,
                    "avg_launch_ms": round(scan_ms / scan_n, 3)}
    err = None
    if pick.startswith("rerank"):
        ix.set_rerank(R)
        ix.set_search_n_center(int(pick.split("nprobe")[1]))
        gi, gd, gc = ix.search_batch(q[:50], 10)
        err = fp64_error(x, q[:50], gi, gd, gc, Metric.Euclidean)
    for v in res.values():
        v.pop("_scan")
        v.pop("_probes")
    print(json.dumps({"metric": METRIC, "value": res[pick]["qps"], "unit": "QPS", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32 (u8 PQ codes)",
                      "data": "synthetic Gaussian mixture, 1,024 centres, sigma 0.3",
                      "config": {"workload": "PQIVFIndex Euclidean 1,000,000x128, nlist 1024, "
                                             "16x256 PQ, 10,000 queries, k=100",
                                 "setting": pick, "train_epoch": epochs},
                      "fp64_error_pick": err,
                      "recall_loss_pick": round(1 - res[pick]["recall_at_10"], 4),
                      "build_s": round(build_s, 2), "by_setting": res, "roofline": roofline}),
          flush=True)


def main_pq(args):
    """Standalone PQIndex (pq_idx.rs:21-265) at the reference's item limit:
    100,000 x 128 Gaussian mixture, 16 sub-spaces x 256 codewords, 10,000
    queries, k=10; recall@10 against the exact brute force."""
    import numpy as np
    import torch

    import oracle
    from hora_amd import BruteForceIndex, Metric, PQIndex, PQParams
    import hora_amd
    from hora_amd.datasets import gaussian_mixture

    N, NQ, D, K = 100_000, 10_000, 128, 10
    epochs = int(os.environ.get("HM_PQ_EPOCHS", "100"))
    # sigma 1.0 by default; at sigma 0.3 the 98 items of a cluster share
    # most of their codes, so most query heads are tied and go through the
    # heap replay
    sigma = float(os.environ.get("HM_PQ_SIGMA", "1.0"))
    x, cen, _ = gaussian_mixture(N, D, 1024, seed=501, sigma=sigma)
    q, _, _ = gaussian_mixture(NQ, D, 1024, seed=502, sigma=sigma, centres=cen)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ix = PQIndex(D, PQParams().n_sub(16).sub_bits(8).train_epoch(epochs).seed(7))
    ix.add_batch(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.build(Metric.Euclidean)
    build_s = time.perf_counter() - t0
    log(f"PQIndex build (16x256, {epochs} epochs): {build_s:.2f}s")
    bf = BruteForceIndex(D)
    bf.add_batch(x)
    bf.build(Metric.Euclidean)
    bf.set_search_mode(hora_amd.HM_SEARCH_BF16_RERANK, 0)
    truth, _, _ = bf.search_batch(q, K)
    del bf
    dq = torch.from_numpy(q).to(dev)
    d_idx = torch.empty((NQ, K), dtype=torch.int64, device=dev)
    d_dist = torch.empty((NQ, K), dtype=torch.float32, device=dev)
    d_cnt = torch.empty((NQ,), dtype=torch.int32, device=dev)
    timer = Timer(torch, stream)
    wall, ev = timer.run(lambda: ix.search_device(dq.data_ptr(), NQ, K, d_idx.data_ptr(),
                                                  d_dist.data_ptr(), d_cnt.data_ptr(),
                                                  stream.cuda_stream),
                         args.steps, args.warmup, "pq", only="pq_scan")
    km = kernel_ms(["pq_scan", "pq_list_select", "row_select", "pq_finalize", "pq_replay"],
                   timer)
    replayed = int(ix.search_stats()[0])
    got = d_idx.cpu().numpy().view(np.uint64)
    r10 = float(np.mean([len(set(got[i].tolist()) & set(truth[i].tolist())) / K
                         for i in range(NQ)]))
    # parity on a sample: the oracle's search over the exported codebook/codes
    ex = ix.export()
    samp = 64
    ri, rd, _ = oracle.pqindex_search(q[:samp], oracle.EUCLIDEAN, ex["codebook"], ex["codes"], K)
    gd = d_dist.cpu().numpy()
    parity = bool(np.array_equal(ri, got[:samp]) and np.array_equal(rd, gd[:samp]))
    scan_ms, scan_n = km["pq_scan"]
    roofline = None
    if scan_n:
        # The scan gathers LUT entries from LDS: one 16-byte read serves 4
        # queries, so each query's lookup moves 4 bytes; per scan scope,
        # rows x (N + 4,096 sampled items in the filter form) x 16 lookups.
        # The distance block (rows x N x 4 B, whole-row form only) and the
        # codes are the HBM side.
        rows = NQ * args.steps / scan_n
        filt = km["row_select"][1] == 0
        items = N + (4096 if filt else 0)
        t = scan_ms / scan_n / 1e3
        lds_bytes = rows * items * 16 * 4
        ach = lds_bytes / t / 1e9
        roofline = {"bound": "lds", "kernel": "pq_scan_kernel" + (" (filter form)" if filt else ""),
                    "achieved": round(ach, 1), "peak": LDS_PEAK, "unit": "GB/s",
                    "frac": round(ach / LDS_PEAK, 4), "traffic": None,
                    "avg_launch_ms": round(scan_ms / scan_n, 3),
                    "lds_lookups_per_s": round(rows * items * 16 / t, 1),
                    "hbm_bytes_per_launch": N * 16 + (0 if filt else rows * N * 4)}
    cpu = None
    if not args.no_cpu_baseline:
        ns = min(args.cpu_sample, NQ)
        hc = host_cpu()
        t0 = time.perf_counter()
        oracle.pqindex_search(q[:ns], oracle.EUCLIDEAN, ex["codebook"], ex["codes"], K,
                              nthreads=hc["threads"])
        t_all = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.pqindex_search(q[:10], oracle.EUCLIDEAN, ex["codebook"], ex["codes"], K,
                              nthreads=1)
        t_one = time.perf_counter() - t0
        cpu = {"value": round(ns / t_all, 1), "unit": "QPS", "cores": hc["threads"],
               "kind": "port",
               "sample": f"first {ns} queries on {hc['threads']} threads (single thread: "
                         f"{round(10 / t_one, 1)} QPS on 10 queries)", "host": hc}
    print(json.dumps({"metric": "QPS, 1 MI355X, PQIndex ADC", "value": round(args.steps * NQ / wall, 1),
                      "unit": "QPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(ev * 1e3 / args.steps, 3), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32 (u8 PQ codes)",
                      "data": f"synthetic Gaussian mixture, 1,024 centres, sigma {sigma}",
                      "config": {"workload": "PQIndex Euclidean 100,000x128, 16x256 PQ, "
                                             "10,000 queries, k=10", "train_epoch": epochs},
                      "recall_at_10": round(r10, 4), "parity_sample_bit_exact": parity,
                      "build_s": round(build_s, 3), "tie_replayed_queries": replayed,
                      "kernel_ms_per_step": {k: round(v[0] / args.steps, 3)
                                             for k, v in km.items() if v[1]},
                      "roofline": roofline, "cpu_baseline": cpu}), flush=True)


def main_knn_graph(args):
    """naive_build_knn_graph (knn.rs:14-41) over the headline base: the
    60,000 x 784 integer-valued rows, Euclidean, k=10, every row a query
    against all others (exact FP32, bit-exact to the reference)."""
    import numpy as np
    import torch

    import oracle
    from hora_amd import BruteForceIndex, Metric
    from hora_amd.datasets import bf_sparse_u8

    N, D, K = 60_000, 784, 10
    x = bf_sparse_u8(N, D, seed=1)
    ix = BruteForceIndex(D)
    ix.add_batch(x)
    ix.build(Metric.Euclidean)
    for _ in range(args.warmup):
        ix.build_knn_graph(K)
    torch.cuda.synchronize()
    from hora_amd._lib import profile_enable, profile_reset
    profile_reset()
    profile_enable(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        gi, gd, gc = ix.build_knn_graph(K)
        log(f"knn_graph: step {i + 1}/{args.steps} done")
    wall = (time.perf_counter() - t0) / args.steps
    profile_enable(False)
    km = kernel_ms(["exact_tile", "row_select", "graph_finalize", "bf_finalize"])
    tied = int(ix.search_stats()[0])
    # parity on a sample of rows: the oracle's graph over the same rows is
    # the full O(n^2) loop, so check the sampled rows with bf_search(k+1)
    # minus the row itself on tie-free rows, and count agreement
    samp = list(range(0, N, N // 64))
    ri, rd, _ = oracle.bf_search(x, x[samp], K + 1, oracle.EUCLIDEAN, nthreads=16)
    agree = 0
    for j, r in enumerate(samp):
        row = [(int(i), float(d)) for i, d in zip(ri[j], rd[j]) if int(i) != r][:K]
        agree += [i for i, _ in row][::-1] == gi[r].tolist()
    et_ms, et_n = km["exact_tile"]
    ops = 3.0 * N * N * D  # sub, mul, add per pair-dimension
    roofline = None
    if et_n:
        ach = ops / (et_ms / args.steps / 1e3) / 1e12
        roofline = {"bound": "valu", "kernel": "exact_tile_kernel", "achieved": round(ach, 1),
                    "peak": FP32_PEAK, "unit": "T op/s", "frac": round(ach / FP32_PEAK, 4),
                    "traffic": None, "avg_launch_ms": round(et_ms / et_n, 3)}
    cpu = None
    if not args.no_cpu_baseline:
        ns = 3000
        hc = host_cpu()
        t0 = time.perf_counter()
        oracle.knn_graph(x[:ns], K, oracle.EUCLIDEAN, nthreads=hc["threads"])
        t = time.perf_counter() - t0
        pairs_per_s = ns * (ns - 1) / t
        cpu = {"value": round(pairs_per_s / (N - 1), 1), "unit": "graph rows/s",
               "cores": hc["threads"], "kind": "port",
               "sample": f"graph of the first {ns} rows on {hc['threads']} threads "
                         f"({pairs_per_s:.3g} pairs/s), scaled to rows of {N - 1} pairs",
               "host": hc}
    print(json.dumps({"metric": "graph rows/s, 1 MI355X, naive_build_knn_graph",
                      "value": round(N / wall, 1), "unit": "graph rows/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic: integer values 0-255 as f32, half zeros",
                      "config": {"workload": "kNN graph, 60,000x784, Euclidean, k=10"},
                      "tie_replayed_rows": tied, "sample_rows_matching_oracle":
                          f"{agree}/{len(samp)} (rows with ties may differ from a sort)",
                      "kernel_ms_per_step": {k: round(v[0] / args.steps, 3)
                                             for k, v in km.items() if v[1]},
                      "roofline": roofline, "cpu_baseline": cpu}), flush=True)


def main_distances(args):
    """Batched distance service of the graph searches over the headline base
    (60,000 x 784, Euclidean): 10,000 queries x 64 random node ids per call,
    and the latency of one query x 32 ids."""
    import numpy as np
    import torch

    import oracle
    from hora_amd import BruteForceIndex, Metric
    from hora_amd.datasets import bf_sparse_u8

    N, D, NQ, L = 60_000, 784, 10_000, 64
    x = bf_sparse_u8(N, D, seed=1)
    q = bf_sparse_u8(NQ, D, seed=2)
    ix = BruteForceIndex(D)
    ix.add_batch(x)
    ix.build(Metric.Euclidean)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(5)
    ids = rng.integers(0, N, NQ * L).astype(np.uint32)
    off = np.arange(NQ + 1, dtype=np.uint64) * L
    dq = torch.from_numpy(q).to(dev)
    di = torch.from_numpy(ids.view(np.int32)).to(dev)
    do = torch.from_numpy(off.view(np.int64)).to(dev)
    out = torch.empty(NQ * L, dtype=torch.float32, device=dev)
    timer = Timer(torch, stream)
    wall, ev = timer.run(lambda: ix.distances_device(dq.data_ptr(), NQ, di.data_ptr(),
                                                     do.data_ptr(), NQ * L, out.data_ptr(),
                                                     stream.cuda_stream),
                         args.steps, args.warmup, "distances", only="pair_dist")
    km = kernel_ms(["pair_dist"], timer)
    got = out.cpu().numpy()
    ref = oracle.distance_rows(x, q[:16], oracle.EUCLIDEAN, nthreads=16)
    parity = bool(all(np.array_equal(got[r * L:(r + 1) * L], ref[r, ids[r * L:(r + 1) * L]])
                      for r in range(16)))
    pd_ms, pd_n = km["pair_dist"]
    t = pd_ms / pd_n / 1e3
    pairs = NQ * L
    ach = pairs * D * 4 / t / 1e9  # the gathered node rows
    # latency: one query, 32 ids, through the device entry
    lat = []
    for i in range(200):
        t0 = time.perf_counter()
        ix.distances_device(dq.data_ptr(), 1, di.data_ptr(), do.data_ptr(), 32, out.data_ptr(),
                            stream.cuda_stream)
        lat.append(time.perf_counter() - t0)
    cpu = None
    if not args.no_cpu_baseline:
        ns = 400
        hc = host_cpu()
        t0 = time.perf_counter()
        oracle.distance_rows(x, q[:ns], oracle.EUCLIDEAN, nthreads=hc["threads"])
        tc = time.perf_counter() - t0
        cpu = {"value": round(ns * N / tc, 1), "unit": "distance pairs/s",
               "cores": hc["threads"], "kind": "port",
               "sample": f"{ns} queries x all {N} nodes on {hc['threads']} threads",
               "host": hc}
    print(json.dumps({"metric": "distance pairs/s, 1 MI355X, graph-search distance service",
                      "value": round(args.steps * pairs / wall, 1), "unit": "distance pairs/s",
                      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(ev * 1e3 / args.steps, 3), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic: integer values 0-255 as f32, half zeros; random ids",
                      "config": {"workload": "60,000x784 Euclidean, 10,000 queries x 64 node ids"},
                      "parity_sample_bit_exact": parity,
                      "latency_1x32_us_median": round(float(np.median(lat)) * 1e6, 1),
                      "roofline": {"bound": "hbm", "kernel": "pair_dist_kernel",
                                   "achieved": round(ach, 1), "peak": HBM_PEAK, "unit": "GB/s",
                                   "frac": round(ach / HBM_PEAK, 4), "traffic": None,
                                   "avg_launch_ms": round(pd_ms / pd_n, 3)},
                      "cpu_baseline": cpu}), flush=True)


def main_ragged(args):
    import numpy as np
    import torch

    import hora_amd
    from hora_amd import BruteForceIndex, Metric

    N, D = 1_000_000, 768
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(404)
    xb = torch.randn((N, D), generator=g, device=dev, dtype=torch.float32)
    xb /= torch.linalg.norm(xb, dim=1, keepdim=True)
    rng = np.random.default_rng(405)
    nb = 64
    sizes = rng.integers(1, 4097, nb)
    ks = rng.choice([1, 10, 100], nb)
    Q = int(sizes.sum())
    xq = torch.randn((Q, D), generator=g, device=dev, dtype=torch.float32)
    xq /= torch.linalg.norm(xq, dim=1, keepdim=True)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    tot = int((sizes * ks).sum())
    d_idx = torch.empty(tot, dtype=torch.int64, device=dev)
    d_dist = torch.empty(tot, dtype=torch.float32, device=dev)
    d_cnt = torch.empty(Q, dtype=torch.int32, device=dev)
    timer = Timer(torch, stream)
    out = {}
    # the configuration's shortlist of 4,096 rows, and the automatic depth
    # (max(64, 4 (k + 1)) per query) for comparison
    S_cfg = int(os.environ.get("HM_SHORTLIST", "4096"))
    for mt in (Metric.DotProduct, Metric.CosineSimilarity):
        for form, code in (("bf16", hora_amd.HM_SEARCH_BF16_RERANK),
                           ("fp8", hora_amd.HM_SEARCH_FP8_RERANK)):
            for S in (S_cfg, 0):
                ix = BruteForceIndex(D)
                ix.add_device(xb.data_ptr(), N)
                ix.build(mt)
                ix.set_search_mode(code, S)
                key = f"{mt.name}/{form}/shortlist_{S or 'auto'}"
                # one captured graph per call (every chunk and the device
                # rerun of tied rows), so the library's per-kernel events,
                # which are off under capture, time only the host reruns
                wall, _ = timer.run(lambda: ix.search_ragged_device(
                    xq.data_ptr(), off, ks, d_idx.data_ptr(), d_dist.data_ptr(),
                    d_cnt.data_ptr(), stream.cuda_stream), args.steps, args.warmup, key)
                km = kernel_ms(["exact_tile", "row_select", "bf_finalize"], timer)
                fb = ix.search_stats()
                out[key] = {
                    "qps": round(args.steps * Q / wall, 1),
                    "ms_per_call": round(wall / args.steps * 1e3, 2),
                    "exact_fallback_rows": {"rerun": fb[0], "uncertified": fb[1]},
                    "host_rerun_kernel_ms_per_call": {k: round(v[0] / args.steps, 3)
                                                      for k, v in km.items() if v[1]}}
                log(f"{key}: {out[key]['qps']} QPS")
                del ix
    cfg = [k_ for k_ in out if k_.endswith(f"shortlist_{S_cfg}")]
    best = max(cfg, key=lambda k_: out[k_]["qps"])
    print(json.dumps({"metric": METRIC, "value": out[best]["qps"], "unit": "QPS", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None,
                      "dtype": best.split("/")[1] + " shortlist + f32 rerank",
                      "data": "synthetic N(0,1) rows L2-normalised (torch, seeded)",
                      "config": {"workload": "BruteForceIndex 1,000,000x768, 64 ragged batches "
                                             "of 1-4096 queries, k in {1,10,100}",
                                 "queries": Q, "shortlist": S_cfg, "metric_form": best,
                                 "graph_launches_per_call": 1},
                      "results": out}), flush=True)


def fp64_error(base, queries, ids, dists, counts, mt):
    """Largest |returned distance - the FP64 value of the same metric| over
    the returned (query, id) pairs, absolute and relative to |FP64 value|."""
    import numpy as np

    import oracle

    worst_abs = worst_rel = 0.0
    for r in range(queries.shape[0]):
        for j in range(int(counts[r])):
            e = oracle.metric_f64(queries[r], base[int(ids[r, j])], int(mt))
            a = abs(float(dists[r, j]) - e)
            worst_abs = max(worst_abs, a)
            worst_rel = max(worst_rel, a / max(abs(e), 1e-30))
    return {"max_abs": float(f"{worst_abs:.4g}"), "max_rel": float(f"{worst_rel:.4g}"),
            "pairs": int(np.sum(counts))}


def host_cpu():
    """CPU model and core counts of this host, and the threads the oracle
    may use (this process's CPU share, capped by OMP_NUM_THREADS)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(share, omp) if omp > 0 else share
    return {"model": model, "nproc": os.cpu_count(), "process_cpu_share": share,
            "threads": threads}


def main_probe(args):
    """Learned nprobe (5,000,000 x 96 anisotropic mixture, nlist 4,096,
    12 x 256 PQ, MLP 96-512-4,096 trained on 200,000 queries' GPU brute-force
    labels) against fixed nprobe, both with the exact rerank of the ADC
    top-R (recall is otherwise capped by the PQ codes).  One step = all
    10,000 test queries through hm_index_search_device, inputs in HBM."""
    import numpy as np
    import torch

    import hora_amd
    from hora_amd import BruteForceIndex, IVFPQIndex, IVFPQParams, Metric
    from hora_amd.datasets import gaussian_mixture

    N = int(os.environ.get("HM_PROBE_N", "5000000"))
    D, NL, K = 96, 4096, 10
    NTRAIN, NTEST = int(os.environ.get("HM_PROBE_TRAIN", "200000")), 10_000
    R = int(os.environ.get("HM_RERANK", "100"))
    PMAX = int(os.environ.get("HM_PMAX", "16"))  # the learned form's largest nprobe
    t0 = time.time()
    x, cen, _ = gaussian_mixture(N, D, 4096, seed=505, aniso=(0.1, 0.5))
    tr, _, _ = gaussian_mixture(NTRAIN, D, 4096, seed=506, aniso=(0.1, 0.5), centres=cen)
    te, _, _ = gaussian_mixture(NTEST, D, 4096, seed=507, aniso=(0.1, 0.5), centres=cen)
    log(f"data generated in {time.time() - t0:.1f}s")
    epochs = int(os.environ.get("HM_IVF_EPOCHS", "6"))
    p = (IVFPQParams().n_sub(12).sub_bits(8).n_kmeans_center(NL).search_n_center(1)
         .train_epoch(epochs).seed(9).kmeans_accumulate(1))
    ix = IVFPQIndex(D, p)
    ix.add_batch(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.build(Metric.Euclidean)
    build_s = time.perf_counter() - t0
    log(f"build {build_s:.1f}s")
    t0 = time.perf_counter()
    pep = int(os.environ.get("HM_PROBE_EP", "3"))
    loss = ix.train_probe(tr, label_k=K, hidden=512, epochs=pep, batch=1024, lr=1e-3, seed=1)
    train_s = time.perf_counter() - t0
    log(f"probe training {train_s:.1f}s, loss {loss:.3f}")
    bf = BruteForceIndex(D)
    bf.add_batch(x)
    bf.build(Metric.Euclidean)
    bf.set_search_mode(hora_amd.HM_SEARCH_BF16_RERANK, 0)
    truth, _, _ = bf.search_batch(te, K)
    del bf
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    dq = torch.from_numpy(te).to(dev)
    d_idx = torch.empty((NTEST, K), dtype=torch.int64, device=dev)
    d_dist = torch.empty((NTEST, K), dtype=torch.float32, device=dev)
    d_cnt = torch.empty((NTEST,), dtype=torch.int32, device=dev)
    timer = Timer(torch, stream)
    fams = ["mlp_gemm", "ivf_scan", "row_select", "ivf_merge", "ivf_replay", "rerank",
            "exact_tile"]

    def measure(label):
        wall, _ = timer.run(lambda: ix.search_device(dq.data_ptr(), NTEST, K, d_idx.data_ptr(),
                                                     d_dist.data_ptr(), d_cnt.data_ptr(),
                                                     stream.cuda_stream),
                            args.steps, args.warmup, label, only="ivf_scan")
        km = kernel_ms(fams, timer)
        got = d_idx.cpu().numpy().view(np.uint64)
        rec = float(np.mean([len(set(got[i].tolist()) & set(truth[i].tolist())) / K
                             for i in range(NTEST)]))
        return {"qps": round(args.steps * NTEST / wall, 1), "recall_at_10": round(rec, 4),
                "mean_nprobe": round(ix.mean_nprobe(), 3),
                "kernel_ms_per_step": {k_: round(v[0] / args.steps, 3) for k_, v in km.items()
                                       if v[1]}}

    res = {}
    for rr in (0, R):
        ix.set_rerank(rr)
        tag = f"rerank{rr}" if rr else "adc"
        ix.set_probe_policy(0, pmin=1, pmax=64)
        res[f"fixed/{tag}"] = {}
        for npb in (1, 2, 3, 4, 6, 8, 16):
            ix.set_search_n_center(npb)
            res[f"fixed/{tag}"][npb] = measure(f"fixed {tag} nprobe {npb}")
        res[f"learned/{tag}"] = {}
        for mass in (0.5, 0.7, 0.8, 0.9, 0.95, 0.98):
            ix.set_probe_policy(1, mass=mass, pmin=1, pmax=PMAX)
            res[f"learned/{tag}"][mass] = measure(f"learned {tag} mass {mass}")
    picks = {}
    for form in ("fixed", "learned"):
        rows = res[f"{form}/rerank{R}"]
        ok = [s_ for s_ in rows if rows[s_]["recall_at_10"] >= 0.95]
        if ok:
            s_ = max(ok, key=lambda v: rows[v]["qps"])
            picks[form] = {"setting": s_, **rows[s_]}
        log(f"{form}: {picks.get(form)}")
    # the chosen learned setting once more, for its distance error on a sample
    err = None
    if "learned" in picks:
        ix.set_rerank(R)
        ix.set_probe_policy(1, mass=picks["learned"]["setting"], pmin=1, pmax=PMAX)
        gi, gd, gc = ix.search_batch(te[:100], K)
        err = fp64_error(x, te[:100], gi, gd, gc, Metric.Euclidean)
    value = picks.get("learned", {}).get("qps")
    line = {"metric": METRIC, "value": value, "unit": "QPS", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (u8 PQ codes, bf16 MLP) + f32 exact rerank",
            "data": "synthetic anisotropic Gaussian mixture (per-cluster sigma in 0.1-0.5)",
            "config": {"workload": f"PQIVFIndex {N}x96, nlist 4096, 12x256 PQ, MLP 96-512-4096 "
                                   f"learned nprobe, 10,000 queries, k=10, exact rerank of the "
                                   f"ADC top-{R}", "train_queries": NTRAIN,
                       "train_epoch": epochs, "probe_epochs": pep},
            "picked_at_recall_0.95": picks,
            "learned_over_fixed_qps": (round(picks["learned"]["qps"] / picks["fixed"]["qps"], 3)
                                       if len(picks) == 2 else None),
            "fp64_error_learned_pick": err,
            "recall_loss_learned_pick": (round(1 - picks["learned"]["recall_at_10"], 4)
                                         if "learned" in picks else None),
            "build_s": round(build_s, 2), "probe_train_s": round(train_s, 1),
            "probe_loss": round(loss, 4), "results": res}
    print(json.dumps(line), flush=True)


def main_copy(args):
    """Calibration for the HBM byte counters: a plain device copy of known
    size (512 MiB read + 512 MiB written per launch)."""
    import torch

    n = 512 << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(max(args.steps, 1)):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": n,
                      "bytes_written_per_launch": n, "launches": max(args.steps, 1)}),
          flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (each GPU would "
              "hold its own shard of the base); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401  (load torch's HIP runtime before the engine's library)

    {"bf": main_bf, "ivfpq": main_ivfpq, "ragged": main_ragged, "probe": main_probe,
     "copy": main_copy, "pq": main_pq, "knn_graph": main_knn_graph,
     "distances": main_distances}[args.workload](args)


if __name__ == "__main__":
    main()
