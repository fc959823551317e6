"""Headline benchmark: points embedded per second by the progressive engine.

A step is one complete progressive session over the workload, from an empty
graph until (i) every row is inserted, (ii) the update has converged
(update_completed) and (iii) one full embedding sweep has run after (ii).
The op plan is the reference's fixed plan: run(500, 500, 500) per 500-row
chunk, then run(0, N, N) until the update completes.  ms_per_step is the sum
of the run() wall times of the step (append_data is outside run(), and the
rows are staged in HBM before the timed region, so no PCIe is timed); value =
N / that time.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload image_like]

One JSON line on stdout.  Only one GPU is used: --gpus above 1 exits with a
message.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
FP32_PEAK = 157.3e12       # MI355X FP32 vector, flop/s
CHUNK = 500

WORKLOAD_DESC = {
    "image_like": "70,000x784 float32 image-like 10-cluster mixture, sparse "
                  "means, Beta(2,5) noise quantised to 1/255; euclidean, k=15, "
                  "2-D, 200 epochs",
    "gaussian10": "10,000x64 float32 mixture of 10 isotropic Gaussians; "
                  "euclidean, k=15, 2-D, 200 epochs",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def one_session(capi, xdev, n, d, profile=False, knn_mode=1):
    e = capi.Engine(n, d, knn_mode=knn_mode)
    if profile:
        e.set_profiling(True)
    base = xdev.data_ptr()
    lat, pos, cycles = [], 0, 0
    swept = False
    while True:
        ins = min(CHUNK, n - pos)
        if ins:
            e.append_device(base + pos * d * 4, ins, d)
        plan = (ins, CHUNK, CHUNK) if ins else (0, n, n)
        t0 = time.perf_counter()
        r = e.run(*plan)
        lat.append(time.perf_counter() - t0)
        pos += ins
        cycles += 1
        if pos == n and r.update_completed:
            if plan[2] >= n:
                swept = True
            else:
                t0 = time.perf_counter()
                r = e.run(0, 0, n)
                lat.append(time.perf_counter() - t0)
                cycles += 1
                swept = True
        if swept or cycles > 5000:
            break
    return e, r, lat


def roofline_numbers(e, capi, d, c):
    out = {}
    js = e.kernel_stats(capi.KSTAT_JOIN)
    if js.launches:
        members, pairs, offers = js.work[0], js.work[1], js.work[2]
        # each list member's vector and 32-slot pool are gathered once per
        # visit; each queued offer reads one 32-slot pool row (8 B / slot)
        bytes_ = members * (4 * d + 8 * 32) + offers * 8 * 32
        flops = pairs * 3 * d  # subtract + fused multiply-add per feature
        ms = js.device_ms / js.launches
        out["k_join"] = {
            "launches": int(js.launches), "avg_ms": ms,
            "algo_bytes_per_launch": bytes_ / js.launches,
            "achieved_GBps": bytes_ / (js.device_ms * 1e-3) / 1e9,
            "achieved_TFLOPs": flops / (js.device_ms * 1e-3) / 1e12,
            "total_ms": js.device_ms}
    ss = e.kernel_stats(capi.KSTAT_SGD)
    if ss.launches:
        pos_, neg_, scans = ss.work[0], ss.work[1], ss.work[2]
        # positive sample: head/tail ids 8 B, eps/next/next-neg 12 B read +
        # 8 B written, both endpoints read 8c and atomically added 8c;
        # negative sample: one 4c read; every edge's schedule check reads 4 B
        bytes_ = pos_ * (28 + 16 * c) + neg_ * 4 * c + scans * 4
        ms = ss.device_ms / ss.launches
        out["k_sgd_epoch"] = {
            "launches": int(ss.launches), "avg_ms": ms,
            "algo_bytes_per_launch": bytes_ / ss.launches,
            "achieved_GBps": bytes_ / (ss.device_ms * 1e-3) / 1e9,
            "total_ms": ss.device_ms}
    ins = e.kernel_stats(capi.KSTAT_INSERT)
    if ins.launches:
        pairs, survivors, redone = ins.work[0], ins.work[1], ins.work[2]
        ms = ins.device_ms / ins.launches
        # algorithmic work of an exact insertion: one multiply-add per
        # feature for every (batch row, inserted row) pair
        flops = 2.0 * d * pairs
        out["insertion"] = {
            "launches": int(ins.launches), "avg_ms": ms,
            "pairs": int(pairs), "survivors_reranked": int(survivors),
            "survivors_per_row": survivors / max(1, e.n),
            "insertions_redone_fp32": int(redone),
            "algo_flops_per_launch": flops / ins.launches,
            "achieved_TFLOPs": flops / (ins.device_ms * 1e-3) / 1e12,
            "total_ms": ins.device_ms}
    return out


def cpu_baseline(x, budget_rows, knn_mode=1):
    import oracle

    xs = np.ascontiguousarray(x[:budget_rows])
    n, d = xs.shape
    threads = os.cpu_count() or 1
    threads = min(threads, 64)
    o = oracle.Oracle(n, d, n_threads=threads, knn_mode=min(knn_mode, 1))
    t_run, pos = 0.0, 0
    while True:
        ins = min(CHUNK, n - pos)
        if ins:
            o.append(xs[pos:pos + ins])
        plan = (ins, CHUNK, CHUNK) if ins else (0, n, n)
        t0 = time.perf_counter()
        r = o.run(*plan)
        t_run += time.perf_counter() - t0
        pos += ins
        if pos == n and r.update_completed:
            if plan[2] < n:
                t0 = time.perf_counter()
                o.run(0, 0, n)
                t_run += time.perf_counter() - t0
            break
    return {"value": n / t_run, "unit": "points/s", "cores": threads,
            "kind": "port",
            "sample": f"first {n} rows of the same workload, same op plan and "
                      f"completion rule, OpenMP oracle ({t_run:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="image_like",
                    choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--cpu-rows", type=int, default=None)
    ap.add_argument("--knn-mode", default="exact_fp32",
                    choices=["nndescent", "exact_fp32", "exact_bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the progressive engine runs one "
              "session per GPU and this build has no multi-GPU path; rerun "
              "with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from gpumap_amd import capi
    from gpumap_amd import evaluation as ev

    x = ev.WORKLOADS[args.workload]()
    n, d = x.shape
    xdev = torch.from_numpy(x).to("cuda:0")
    torch.cuda.synchronize()
    log(f"workload {args.workload} {x.shape}")

    mode = {"nndescent": 0, "exact_fp32": 1, "exact_bf16": 2}[args.knn_mode]
    for i in range(args.warmup):
        e, r, lat = one_session(capi, xdev, n, d, knn_mode=mode)
        log(f"warmup {i}: {sum(lat):.3f} s over {len(lat)} cycles")
        e.close()

    torch.cuda.synchronize()
    totals, lats, last = [], [], None
    for i in range(args.steps):
        e, r, lat = one_session(capi, xdev, n, d, profile=True, knn_mode=mode)
        totals.append(sum(lat))
        lats.extend(lat)
        log(f"step {i}: {sum(lat):.3f} s over {len(lat)} cycles")
        if last is not None:
            last.close()
        last = e
    torch.cuda.synchronize()

    ms_per_step = 1e3 * float(np.mean(totals))
    value = n / (ms_per_step / 1e3)
    kern = roofline_numbers(last, capi, d, 2)
    dom_name = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None
    roof = None
    if dom_name:
        kd = kern[dom_name]
        achieved = kd["achieved_GBps"]
        roof = {"kernel": dom_name, "bound": "hbm", "achieved": achieved,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / (HBM_PEAK / 1e9), "traffic": None}

    quality = {}
    if not args.no_quality:
        idx, dist = last.knn()
        q = np.random.Generator(np.random.PCG64(7)).choice(n, 2000, replace=False)
        ei, ed = ev.exact_knn(x, 15, queries=q, device="cuda:0")
        quality["recall_at_15"] = ev.recall_at_k(idx, x, ei, ed, queries=q)
        quality["trustworthiness_at_15"] = ev.trustworthiness(
            x, last.embedding(), sample=5000)
        log(f"quality {quality}")
    last.close()

    cpu = None
    if not args.no_cpu_baseline:
        rows = args.cpu_rows or (4000 if d > 256 else 10_000)
        cpu = cpu_baseline(x, min(rows, n), knn_mode=mode)
        log(f"cpu baseline {cpu}")

    line = {
        "metric": "points embedded/s @recall@15>=0.95 (FP32 path); p50/p99 run() latency",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_DESC[args.workload], "n_points": n,
                   "n_features": d, "knn_mode": args.knn_mode, "op_plan": "run(500,500,500) per 500-row "
                   "chunk, then run(0,N,N) until update_completed, then one "
                   "full embedding sweep"},
        "latency_p50_ms": 1e3 * float(np.percentile(lats, 50)),
        "latency_p99_ms": 1e3 * float(np.percentile(lats, 99)),
        "quality": quality,
        "kernels": kern,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
