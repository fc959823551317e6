"""Headline benchmark: points embedded per second by the progressive engine.

A step is one complete progressive session over the workload, from an empty
graph until (i) every row is inserted, (ii) the update has converged
(update_completed) and (iii) one full embedding sweep has run after (ii).
The op plan is the reference's fixed plan: run(500, 500, 500) per 500-row
chunk, then run(0, N, N) until the update completes, then run(0, 0, N).
ms_per_step is the sum of the run() wall times of the step (append is
outside run(), and the rows are staged in HBM before the timed region, so no
PCIe is timed); value = N / that time.

The headline path is the BF16-MFMA screened exact insertion (knn_mode
exact_bf16); the same run also times the plain FP32 path (exact FP32
distance tiles, knn_mode exact_fp32) on the same workload and reports the
ratio.  Both paths produce the same kNN graph bit for bit.

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
BF16_PEAK = 2.5e15         # MI355X dense BF16 MFMA, flop/s
FP8_PEAK = 5.0e15          # MI355X dense FP8 MFMA, flop/s
CHUNK = 500  # the workload's insertion chunk (set from WORKLOAD_CFG)
METRIC = ("points embedded/s @recall@15≥0.95, BF16-MFMA vs FP32; "
          "p50/p99 run() latency")
MODES = {"nndescent": 0, "exact_fp32": 1, "exact_bf16": 2, "exact_fp8": 3,
         "nndescent_bf16": 4}
PRECISION = {
    0: "NN-Descent candidates and local join in FP32",
    1: "FP32",
    2: ("distance tiles BF16 MFMA (hi/lo split, FP32 accumulation) screened "
        "by a rigorous error bound and re-ranked in exact FP32; sigma/rho "
        "FP64; SGD FP32"),
    3: ("distance tiles FP8 E4M3 MFMA (per-row power-of-two scales, hi/lo "
        "split, FP32 accumulation) screened by a rigorous error bound and "
        "re-ranked in exact FP32; sigma/rho FP64; SGD FP32"),
    4: ("NN-Descent candidates in FP32; local-join distance blocks BF16 MFMA "
        "(hi/lo split, FP32 accumulation) screened by a rigorous error bound "
        "and re-ranked in exact FP32; sigma/rho FP64; SGD FP32"),
}
DTYPE = {0: "f32", 1: "f32", 2: "bf16", 3: "fp8_e4m3", 4: "bf16"}

WORKLOAD_DESC = {
    "image_like": "70,000x784 float32 image-like 10-cluster mixture, sparse "
                  "means, Beta(2,5) noise quantised to 1/255; euclidean, k=15, "
                  "2-D, 200 epochs",
    "gaussian10": "10,000x64 float32 mixture of 10 isotropic Gaussians; "
                  "euclidean, k=15, 2-D, 200 epochs",
    "vmf50": "200,000x300 float32 unit vectors, 50 von Mises-Fisher clusters "
             "(kappa 50, Zipf(1.2) sizes); cosine, k=30, 2-D, 200 epochs",
    "stream100": "2,000,000x128 float32 100-cluster Gaussian mixture (sigma "
                 "U[0.5,2], centres U[-20,20]); euclidean, k=15, 3-D, 200 "
                 "epochs, 10,000-row chunks",
}
# per workload: n_neighbors, n_components, metric (0 euclidean, 1 cosine) and
# the insertion chunk of the op plan
WORKLOAD_CFG = {
    "image_like": (15, 2, 0, 500),
    "gaussian10": (15, 2, 0, 500),
    "vmf50": (30, 2, 1, 500),
    "stream100": (15, 3, 0, 10_000),
}
CFG = WORKLOAD_CFG["image_like"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def one_session(capi, xdev, n, d, profile=False, knn_mode=2):
    k, c, metric, _ = CFG
    e = capi.Engine(n, d, n_neighbors=k, n_components=c, metric=metric,
                    knn_mode=knn_mode)
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


def kernel_numbers(e, capi, d, c):
    """Per-kernel figures from the engine's HIP events (on its own stream)
    and device work counters."""
    out = {}
    ss = e.kernel_stats(capi.KSTAT_SGD)
    if ss.launches:
        pos_, neg_ = ss.work[0], ss.work[1]
        # per positive sample: head/tail ids 8 B, both endpoints read 8c and
        # atomically written 8c, schedule state 16 B, nsr = 5 negative reads
        # of 4c: 24 + 36c bytes
        bytes_ = pos_ * (24 + 36 * c)
        ms = ss.device_ms / ss.launches
        waves = int(ss.work[3])
        per_wave = ("k_sgd_waves" if os.environ.get("GPUMAP_AMD_SGD_PIPE") ==
                    "0" else "k_sgd_pipe")
        name = (per_wave if waves == ss.launches else
                "k_sgd_fused" if waves == 0 else per_wave + "+k_sgd_fused")
        out[name] = {
            "launches": int(ss.launches), "avg_ms": ms,
            "positive_samples": int(pos_), "negative_samples": int(neg_),
            "algo_bytes_per_launch": bytes_ / ss.launches,
            "achieved_GBps": bytes_ / (ss.device_ms * 1e-3) / 1e9,
            "total_ms": ss.device_ms}
    st = e.kernel_stats(capi.KSTAT_SGD_STREAM)
    if st.launches:
        bytes_ = st.work[0] * (24 + 36 * c)
        out["k_sgd_stream"] = {
            "launches": int(st.launches), "avg_ms": st.device_ms / st.launches,
            "positive_samples": int(st.work[0]),
            "algo_bytes_per_launch": bytes_ / st.launches,
            "achieved_GBps": bytes_ / (st.device_ms * 1e-3) / 1e9,
            "total_ms": st.device_ms}
    ins = e.kernel_stats(capi.KSTAT_INSERT)
    if ins.launches:
        pairs, survivors, redone = ins.work[0], ins.work[1], ins.work[2]
        # algorithmic work of an exact insertion: one multiply-add per
        # feature for every (batch row, inserted row) pair
        flops = 2.0 * d * pairs
        out["insertion"] = {
            "launches": int(ins.launches),
            "avg_ms": ins.device_ms / ins.launches,
            "pairs": int(pairs), "survivors_reranked": int(survivors),
            "survivors_per_row": survivors / max(1, e.n),
            "insertions_redone_fp32": int(redone),
            "algo_flops_per_launch": flops / ins.launches,
            "achieved_TFLOPs": flops / (ins.device_ms * 1e-3) / 1e12,
            "total_ms": ins.device_ms}
    js = e.kernel_stats(capi.KSTAT_JOIN)
    if js.launches:
        members, pairs, offers = js.work[0], js.work[1], js.work[2]
        bytes_ = members * (4 * d + 8 * 32) + offers * 8 * 32
        # distance work: one multiply-add per feature and pair of the lists
        flops = 2.0 * d * pairs
        out["k_join"] = {
            "launches": int(js.launches), "avg_ms": js.device_ms / js.launches,
            "algo_bytes_per_launch": bytes_ / js.launches,
            "achieved_GBps": bytes_ / (js.device_ms * 1e-3) / 1e9,
            "pairs": int(pairs), "offers": int(offers),
            "pairs_reranked_fp32": int(js.work[3]),
            "algo_flops_per_launch": flops / js.launches,
            "achieved_TFLOPs": flops / (js.device_ms * 1e-3) / 1e12,
            "total_ms": js.device_ms}
    return out


def measure(capi, xdev, n, d, mode, steps, warmup):
    import torch

    for i in range(warmup):
        e, r, lat = one_session(capi, xdev, n, d, knn_mode=mode)
        log(f"[{mode}] warmup {i}: {sum(lat):.3f} s over {len(lat)} cycles")
        e.close()
    torch.cuda.synchronize()
    totals, lats, last = [], [], None
    for i in range(steps):
        e, r, lat = one_session(capi, xdev, n, d, profile=True, knn_mode=mode)
        totals.append(sum(lat))
        lats.extend(lat)
        log(f"[{mode}] step {i}: {sum(lat):.3f} s over {len(lat)} cycles")
        if last is not None:
            last.close()
        last = e
    torch.cuda.synchronize()
    return last, totals, lats


def served(xdevs, n, d, sessions, window=0.001, clients="loop"):
    """`sessions` sessions of the workload served together by the Server,
    session i on its own rows xdevs[i] (distinct draws of the workload), each
    running the headline's op plan through requests of run_with_n_ops; the
    Server batches the pending requests into cycles.
    clients="loop": one front-end thread submits every session's next
    request as soon as its previous result is back (the event-loop form of a
    serving front end); "threads": one blocking client thread per session.
    Returns the wall time from the first request to the last result, every
    request's latency, the cycle count and one session's model."""
    import threading
    from concurrent.futures import FIRST_COMPLETED, wait

    import gpumap_amd as gm
    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import Server

    def client(sess, lat):
        pos = 0
        while True:
            ins = min(CHUNK, n - pos)
            plan = (ins, CHUNK, CHUNK) if ins else (0, n, n)
            t0 = time.perf_counter()
            r = sess.run_with_n_ops(*plan)
            lat.append(time.perf_counter() - t0)
            pos += ins
            if pos == n and r.update_completed:
                if plan[2] < n:
                    t0 = time.perf_counter()
                    sess.run_with_n_ops(0, 0, n)
                    lat.append(time.perf_counter() - t0)
                return

    srv = Server(batch_window=window)
    ss = []
    k, c, metric, _ = CFG
    for i in range(sessions):
        p = gm.GPUMAPParams()
        p.knn_mode = "exact_bf16"
        p.n_neighbors, p.n_components, p.metric = k, c, metric
        ss.append(srv.open(ArrayDataLoader(xdevs[i % len(xdevs)]), p))
    lats = [[] for _ in ss]
    c0 = srv.cycles
    t0 = time.perf_counter()
    if clients == "threads":
        th = [threading.Thread(target=client, args=(s, l))
              for s, l in zip(ss, lats)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    else:
        pos = [0] * len(ss)
        swept = [False] * len(ss)
        inflight = {}

        def nxt(i):
            ins = min(CHUNK, n - pos[i])
            plan = (ins, CHUNK, CHUNK) if ins else (0, n, n)
            pos[i] += ins
            f = ss[i].submit(n_ops=plan)
            inflight[f] = (i, plan, time.perf_counter())

        for i in range(len(ss)):
            nxt(i)
        while inflight:
            done, _ = wait(list(inflight), return_when=FIRST_COMPLETED)
            for f in done:
                i, plan, ts = inflight.pop(f)
                r = f.result()
                lats[i].append(time.perf_counter() - ts)
                if pos[i] == n and r.update_completed:
                    if plan[2] >= n or swept[i]:
                        continue
                    swept[i] = True
                    ts2 = time.perf_counter()
                    f2 = ss[i].submit(n_ops=(0, 0, n))
                    inflight[f2] = (i, (0, 0, n), ts2)
                    continue
                nxt(i)
    wall = time.perf_counter() - t0
    cycles = srv.cycles - c0
    keep = ss[0]
    for s in ss[1:]:
        s.close()
    return srv, keep, wall, np.concatenate([np.array(l) for l in lats]), cycles


def served_entry(ev, x, xdev, n, d, sessions, steps, fp32_value, no_quality,
                 workload):
    import torch

    # session 0 on the headline's rows, the others on their own draws of the
    # same workload
    xdevs = [xdev] + [
        torch.from_numpy(ev.WORKLOADS[workload](n=n, d=d, index=1000 + i)).to(
            xdev.device) for i in range(1, sessions)]
    torch.cuda.synchronize()
    srv, keep, wall, lat, cycles = served(xdevs, n, d, sessions)  # warmup
    keep.close()
    srv.close()
    # the same sessions with one blocking client thread each
    srv, keep, t_wall, t_lat, _ = served(xdevs, n, d, sessions,
                                         clients="threads")
    keep.close()
    srv.close()
    walls, lats = [], []
    for i in range(steps):
        srv, keep, wall, lat, cycles = served(xdevs, n, d, sessions)
        walls.append(wall)
        lats.append(lat)
        if i < steps - 1:
            keep.close()
            srv.close()
    lat = np.concatenate(lats)
    value = sessions * n / float(np.mean(walls))
    entry = {"sessions": sessions, "value": value, "unit": "points/s",
             "data": "session 0 the headline's rows, each other session its "
                     "own draw of the workload",
             "wall_s": float(np.mean(walls)),
             "request_latency_p50_ms": 1e3 * float(np.percentile(lat, 50)),
             "request_latency_p99_ms": 1e3 * float(np.percentile(lat, 99)),
             "cycles": cycles, "batch_window_ms": 1.0,
             "clients": "one front-end thread submitting each session's next "
                        "request when its result is back",
             "vs_fp32_plain": value / fp32_value if fp32_value else None,
             "threads_clients": {
                 "value": sessions * n / t_wall,
                 "request_latency_p50_ms": 1e3 * float(np.percentile(t_lat, 50)),
                 "request_latency_p99_ms": 1e3 * float(np.percentile(t_lat, 99)),
                 "vs_fp32_plain": (sessions * n / t_wall / fp32_value
                                   if fp32_value else None)}}
    if not no_quality:
        emb = keep.get_embedding().cpu().numpy()
        entry["trustworthiness_at_15_session0"] = ev.trustworthiness(
            x, emb, sample=5000, metric="cosine" if CFG[2] == 1 else
            "euclidean")
    keep.close()
    srv.close()
    log(f"[served x{sessions}] {value:.0f} points/s"
        + (f", {entry['vs_fp32_plain']:.2f}x the FP32 path" if fp32_value
           else ""))
    return entry


def serving_mix(ev, n_sessions=64, rate=4.0, budget=0.5, seed=1003,
                deadline=240.0, burst=False):
    """The serving mix: `n_sessions` independent sessions arriving as a
    Poisson process of `rate` per second, sizes log-uniform in [1e3, 1e5],
    dims uniform in {32, 128, 784}, 10-cluster Gaussian mixtures; each
    session runs run(target_latency=budget) until its insertion, update and
    embedding have completed, served by one Server.  Sessions still running
    `deadline` seconds after the start stop there and count as incomplete.
    `burst` makes every session arrive at the start (the server's capacity
    on the mix rather than the arrival rate)."""
    import threading

    import torch

    import gpumap_amd as gm
    from gpumap_amd.data_loader import ArrayDataLoader
    from gpumap_amd.serving import Server

    g = np.random.default_rng(seed)
    sizes = np.exp(g.uniform(np.log(1e3), np.log(1e5), n_sessions)).astype(int)
    dims = g.choice([32, 128, 784], n_sessions)
    arrive = np.cumsum(g.exponential(1.0 / rate, n_sessions))
    if burst:
        arrive[:] = 0.0  # the same sessions, all arriving at once
    data = [torch.from_numpy(ev.gaussian10(n=int(s), d=int(dd), index=100 + i))
            .to("cuda:0") for i, (s, dd) in enumerate(zip(sizes, dims))]
    torch.cuda.synchronize()
    log(f"[serving mix] {n_sessions} sessions, {int(sizes.sum())} points "
        f"staged in HBM")
    srv = Server(batch_window=0.001)
    lat = [[] for _ in range(n_sessions)]
    span = [None] * n_sessions
    err = []
    complete = [False] * n_sessions

    def session(i, t_start):
        time.sleep(max(0.0, t_start + arrive[i] - time.perf_counter()))
        t_arr = time.perf_counter()
        try:
            p = gm.GPUMAPParams()
            p.knn_mode = "exact_bf16"
            s = srv.open(ArrayDataLoader(data[i]), p)
            while time.perf_counter() - t_start < deadline:
                t0 = time.perf_counter()
                r = s.run(target_latency=budget)
                lat[i].append(time.perf_counter() - t0)
                if (r.insertion_completed and r.update_completed
                        and r.embedding_completed):
                    complete[i] = True
                    break
            s.close()
        except BaseException as e:  # reported in the JSON line
            err.append(f"session {i}: {e!r}")
        span[i] = (t_arr, time.perf_counter())
        log(f"[serving mix] session {i} ({sizes[i]}x{dims[i]}) "
            f"{'done' if complete[i] else 'stopped'} after {len(lat[i])} "
            f"requests, {span[i][1] - span[i][0]:.2f} s; at "
            f"{span[i][1] - t_start:.1f} s")

    t_start = time.perf_counter()
    th = [threading.Thread(target=session, args=(i, t_start))
          for i in range(n_sessions)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    t_end = time.perf_counter()
    srv.close()
    allat = np.concatenate([np.array(v) for v in lat if v])
    makespan = t_end - t_start
    total = int(sizes.sum())
    done = np.array([b - a for a, b in span])
    return {
        "sessions": n_sessions,
        "arrival_rate_per_s": "all at once" if burst else rate,
        "points": total, "dims": {str(k): int((dims == k).sum())
                                  for k in (32, 128, 784)},
        "size_min": int(sizes.min()), "size_max": int(sizes.max()),
        "budget_ms": 1e3 * budget, "makespan_s": makespan,
        "last_arrival_s": float(arrive[-1]),
        "value": total / makespan, "unit": "points/s",
        "request_latency_p50_ms": 1e3 * float(np.percentile(allat, 50)),
        "request_latency_p99_ms": 1e3 * float(np.percentile(allat, 99)),
        "requests": int(allat.size),
        "requests_over_budget": int((allat > budget).sum()),
        "session_time_p50_s": float(np.percentile(done, 50)),
        "session_time_p99_s": float(np.percentile(done, 99)),
        "incomplete": int(n_sessions - sum(complete)),
        "cycles": srv.cycles, "max_sessions_per_cycle": srv.max_batch,
        "mean_sessions_per_cycle": srv.requests_run / max(srv.cycles, 1),
        "errors": err,
    }


def dist_error_vs_fp64(x, idx, dist, q, metric):
    """Published neighbour distances of the rows q against the same pairs'
    squared distances in FP64 (cosine: of the FP64-normalised rows):
    relative error max and mean."""
    xd = np.asarray(x, dtype=np.float64)
    if metric == 1:
        xd = xd / np.maximum(np.linalg.norm(xd, axis=1, keepdims=True), 1e-300)
    a = np.asarray(idx)[q].astype(np.int64)
    ok = (a >= 0) & (a < xd.shape[0])
    a = np.where(ok, a, 0)
    diff = xd[a] - xd[q][:, None, :]
    d64 = (diff * diff).sum(-1)
    d32 = np.asarray(dist, dtype=np.float64)[q]
    rel = np.abs(d32 - d64) / np.maximum(d64, 1e-30)
    rel = rel[ok & (d64 > 0)]
    return {"pairs": int(rel.size), "max_rel": float(rel.max()),
            "mean_rel": float(rel.mean())}


def quality(ev, x, e, n):
    k, _, metric, _ = CFG
    mname = "cosine" if metric == 1 else "euclidean"
    idx, dist = e.knn()
    q = np.random.Generator(np.random.PCG64(7)).choice(n, 2000, replace=False)
    ei, ed = ev.exact_knn(x, k, queries=q, metric=mname, device="cuda:0")
    return {f"recall_at_{k}": ev.recall_at_k(idx, x, ei, ed, queries=q,
                                             metric=mname),
            "distance_error_vs_fp64": dist_error_vs_fp64(x, idx, dist, q,
                                                         metric),
            "trustworthiness_at_15": ev.trustworthiness(
                x, e.embedding(), sample=5000, metric=mname)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_session(xs, threads, knn_mode):
    """Wall time of the oracle's run() calls over one session of the op plan
    on the rows `xs`."""
    import oracle

    n, d = xs.shape
    k, c, metric, _ = CFG
    o = oracle.Oracle(n, d, n_neighbors=k, n_components=c, metric=metric,
                      n_threads=threads, knn_mode=knn_mode)
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
    return t_run


def cpu_baseline(x, budget_rows, knn_mode=1):
    """The oracle (OpenMP C++ restatement, exact FP32 insertion as the
    headline path) on two prefixes of the workload, budget_rows / 2 and
    budget_rows.  An exact insertion's cost per row grows with the rows
    already inserted, so the two session times also fix t(n) = a n + b n^2,
    whose value at the full row count is reported beside the measured
    rate."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover - Linux has it
        nproc = os.cpu_count() or 1
    kmode = min(knn_mode, 1)
    sizes = (budget_rows // 2, budget_rows)
    times = [_oracle_session(np.ascontiguousarray(x[:m]), threads, kmode)
             for m in sizes]
    n_full = x.shape[0]
    (n1, n2), (t1, t2) = sizes, times
    # t = a n + b n^2 through both points
    b = (t2 / n2 - t1 / n1) / (n2 - n1)
    a = t1 / n1 - b * n1
    t_full = a * n_full + b * n_full * n_full
    return {"value": n2 / t2, "unit": "points/s", "cores": threads,
            "kind": "port",
            "cpu_model": _cpu_model(), "nproc": nproc,
            "sample": f"first {n2} of the workload's rows, same op plan and "
                      f"completion rule, exact FP32 insertion, OpenMP C++ "
                      f"restatement on {threads} threads ({t2:.1f} s; the "
                      f"first {n1} rows took {t1:.1f} s)",
            "fitted_full_config": {
                "rows": n_full, "seconds": t_full,
                "points_per_s": n_full / t_full if t_full > 0 else None,
                "model": "t(n) = a n + b n^2 through the two samples"}}


def hbm_copy_rate(mib=2048, reps=5):
    """Streaming-copy bandwidth of this GPU in this run: torch's device copy
    of `mib` MiB (read + write counted), the best of `reps`."""
    import torch

    src = torch.empty(mib << 18, dtype=torch.float32, device="cuda:0")
    src.fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    best = 0.0
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        best = max(best, 2.0 * src.numel() * 4 / (a.elapsed_time(b) * 1e-3))
    del src, dst
    torch.cuda.empty_cache()
    return best


def strategy_leg(ev, x, xdev, n, d, knn_mode, budget=0.5, max_cycles=2000):
    """The workload under the facade's default WeightedStrategy: run(budget)
    until insertion, update and embedding have completed.  Reports the
    points/s, every run()'s latency against the budget and the on-chip
    per-op cost model, the seconds per op of each operation measured per
    cycle, tabulated over the inserted row count."""
    import gpumap_amd as gm
    from gpumap_amd.data_loader import ArrayDataLoader

    k, c, metric, _ = CFG
    p = gm.GPUMAPParams()
    p.n_instances, p.n_features, p.n_neighbors = n, d, k
    p.n_components = c
    p.metric = "cosine" if metric == 1 else "euclidean"
    p.knn_mode = knn_mode
    m = gm.GPUMAP(ArrayDataLoader(xdev), p)
    lat, rows = [], []
    for _ in range(max_cycles):
        t0 = time.perf_counter()
        r = m.run(budget)
        lat.append(time.perf_counter() - t0)
        rows.append((r.n_inserted_instances, r.n_insertion_ops,
                     r.insertion_time, r.n_update_ops, r.update_time,
                     r.n_embedding_ops, r.embedding_time))
        if r.insertion_completed and r.update_completed and \
                r.embedding_completed:
            break
    emb = m.get_embedding().cpu().numpy()
    done = bool(r.insertion_completed and r.update_completed and
                r.embedding_completed)
    m.close()
    # per-op seconds of each cycle, against the rows inserted at its end
    table = []
    for q in rows:
        ent = {"rows": int(q[0])}
        for name, i_ops, i_t in (("insertion", 1, 2), ("update", 3, 4),
                                 ("embedding", 5, 6)):
            ent[f"{name}_ops"] = int(q[i_ops])
            ent[f"{name}_s_per_op"] = q[i_t] / q[i_ops] if q[i_ops] else None
        table.append(ent)
    lat_ms = 1e3 * np.asarray(lat)
    out = {"value": n / float(np.sum(lat)), "unit": "points/s",
           "budget_s": budget, "cycles": len(lat), "completed": done,
           "latency_p50_ms": float(np.percentile(lat_ms, 50)),
           "latency_p99_ms": float(np.percentile(lat_ms, 99)),
           "latency_max_ms": float(lat_ms.max()),
           "over_budget": int((lat_ms > 1e3 * budget).sum()),
           "cost_model": table}
    out["trustworthiness_at_15"] = ev.trustworthiness(
        x, emb, sample=5000, metric=p.metric)
    return out


def load_traffic(kernel, workload, knn_mode):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters
    on this workload and knn mode, from the committed profile summary
    (profiles/traffic.json), or None when it was measured on another."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("_workload") != workload or t.get("_knn_mode") != knn_mode:
        return None
    return t.get(kernel, {}).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="image_like",
                    choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--knn-mode", default="exact_bf16", choices=sorted(MODES))
    ap.add_argument("--compare", default="exact_fp32",
                    help="comma list of knn modes timed for comparison "
                         "('' for none)")
    ap.add_argument("--cpu-rows", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--served", default="4,8",
                    help="comma list of session counts served together on "
                         "the workload ('' for none)")
    ap.add_argument("--no-serving-mix", action="store_true",
                    help="skip the 64-session serving mix")
    ap.add_argument("--no-strategy", action="store_true",
                    help="skip the WeightedStrategy run(target_latency) leg")
    ap.add_argument("--budget", type=float, default=0.5,
                    help="run(target_latency) budget of the strategy leg, s")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device copy of MIB MiB (a kernel of "
                         "known traffic that scales rocprofv3 PMC counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the progressive engine runs one "
              "session per GPU and this build has no multi-GPU path; rerun "
              "with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from gpumap_amd import capi
    from gpumap_amd import evaluation as ev

    global CFG, CHUNK
    CFG = WORKLOAD_CFG[args.workload]
    CHUNK = CFG[3]
    x = ev.WORKLOADS[args.workload]()
    n, d = x.shape
    xdev = torch.from_numpy(x).to("cuda:0")
    torch.cuda.synchronize()
    log(f"workload {args.workload} {x.shape}")

    mode = MODES[args.knn_mode]
    last, totals, lats = measure(capi, xdev, n, d, mode, args.steps,
                                 args.warmup)
    ms_per_step = 1e3 * float(np.mean(totals))
    value = n / (ms_per_step / 1e3)
    kern = kernel_numbers(last, capi, d, CFG[1])
    qual = {} if args.no_quality else quality(ev, x, last, n)
    if qual:
        log(f"[{args.knn_mode}] quality {qual}")
    last.close()

    # dominant kernel: the most device time over the timed session
    timed = {k: v for k, v in kern.items() if k.startswith("k_")}
    roof = None
    if timed:
        name = max(timed, key=lambda k: timed[k]["total_ms"])
        kd = timed[name]
        roof = {"kernel": name, "bound": "hbm",
                "achieved": kd["achieved_GBps"], "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": kd["achieved_GBps"] / (HBM_PEAK / 1e9),
                "traffic": load_traffic(name, args.workload, args.knn_mode),
                "algo_bytes_per_launch": kd["algo_bytes_per_launch"],
                "avg_launch_ms": kd["avg_ms"]}
    if "insertion" in kern and mode in (MODES["exact_bf16"],
                                        MODES["exact_fp8"]):
        kern["insertion"]["mfma_frac_algorithmic"] = (
            kern["insertion"]["achieved_TFLOPs"] * 1e12 /
            (FP8_PEAK if mode == MODES["exact_fp8"] else BF16_PEAK))
    if "k_join" in kern and mode == MODES["nndescent_bf16"]:
        kern["k_join"]["mfma_frac_algorithmic"] = (
            kern["k_join"]["achieved_TFLOPs"] * 1e12 / BF16_PEAK)

    if args.copy_probe:
        src = torch.empty(args.copy_probe << 18, dtype=torch.float32,
                          device="cuda:0")
        src.fill_(1.0)
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    compare = {}
    for m in [c for c in args.compare.split(",") if c and c != args.knn_mode]:
        e2, tot2, lat2 = measure(capi, xdev, n, d, MODES[m], args.steps,
                                 min(args.warmup, 1))
        ms2 = 1e3 * float(np.mean(tot2))
        entry = {"value": n / (ms2 / 1e3), "ms_per_step": ms2,
                 "latency_p50_ms": 1e3 * float(np.percentile(lat2, 50)),
                 "latency_p99_ms": 1e3 * float(np.percentile(lat2, 99)),
                 "kernels": kernel_numbers(e2, capi, d, CFG[1])}
        if not args.no_quality:
            entry["quality"] = quality(ev, x, e2, n)
        e2.close()
        entry["headline_speedup"] = value / entry["value"]
        compare[m] = entry
        log(f"[{m}] {entry['value']:.0f} points/s, headline is "
            f"{entry['headline_speedup']:.2f}x")

    fp32 = compare.get("exact_fp32", {}).get("value")
    served_out = {}
    for k in [int(v) for v in args.served.split(",") if v]:
        served_out[str(k)] = served_entry(ev, x, xdev, n, d, k, args.steps,
                                          fp32, args.no_quality, args.workload)
    mix = None
    if not args.no_serving_mix:
        mix = serving_mix(ev)
        mix["burst"] = serving_mix(ev, burst=True)
        for m in (mix, mix["burst"]):
            log(f"[serving mix] {m['value']:.0f} points/s over "
                f"{m['makespan_s']:.1f} s, request p50 "
                f"{m['request_latency_p50_ms']:.1f} ms p99 "
                f"{m['request_latency_p99_ms']:.1f} ms")

    strat = None
    if not args.no_strategy:
        # the requested budget, then a tenth of it (more cycles: a finer
        # per-op cost model over the row count)
        strat = strategy_leg(ev, x, xdev, n, d, args.knn_mode,
                             budget=args.budget)
        strat["tenth_budget"] = strategy_leg(ev, x, xdev, n, d,
                                             args.knn_mode,
                                             budget=args.budget / 10)
        for s_ in (strat, strat["tenth_budget"]):
            log(f"[strategy run({s_['budget_s']})] {s_['value']:.0f} "
                f"points/s over {s_['cycles']} cycles, p50 "
                f"{s_['latency_p50_ms']:.1f} ms p99 "
                f"{s_['latency_p99_ms']:.1f} ms, "
                f"{s_['over_budget']} over budget")
    copy_rate = hbm_copy_rate()
    log(f"streaming copy {copy_rate / 1e9:.0f} GB/s")
    if roof:
        roof["peak_copy_measured"] = copy_rate / 1e9
        roof["frac_of_copy"] = roof["achieved"] / (copy_rate / 1e9)

    cpu = None
    if not args.no_cpu_baseline:
        # a sample of about 10-30 s of CPU work on 16 cores
        rows = args.cpu_rows or (15_000 if d > 256 else 10_000)
        cpu = cpu_baseline(x, min(rows, n), knn_mode=mode)
        log(f"cpu baseline {cpu}")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[mode],
        "data": "synthetic",
        "config": {"workload": WORKLOAD_DESC[args.workload], "n_points": n,
                   "n_features": d, "knn_mode": args.knn_mode,
                   "precision": PRECISION[mode],
                   "op_plan": f"run({CHUNK},{CHUNK},{CHUNK}) per "
                              f"{CHUNK}-row chunk, then run(0,N,N) until "
                              "update_completed, then one full embedding "
                              "sweep"},
        "latency_p50_ms": 1e3 * float(np.percentile(lats, 50)),
        "latency_p99_ms": 1e3 * float(np.percentile(lats, 99)),
        "quality": qual,
        "kernels": kern,
        "roofline": roof,
        "compare": compare,
        "served": served_out,
        "serving_mix": mix,
        "strategy": strat,
        "screen_shapes": capi.screen_shapes(),
        "hbm_copy_GBps": copy_rate / 1e9,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
