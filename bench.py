"""Headline benchmark: modern Adam layout edge-epochs/s at 1M rows.

Workload (1,000,000 x 784 image-like rows): a 10-cluster Gaussian mixture
(centres U[32,224]^784, sigma 24, seed 42), rounded and clipped to 0..255,
as float32. Exact k=15 fuzzy kNN graph (about 15M directed edges),
pruned at max/200, compacted; random init (seed 42) with isotropic
rescale; 200 Adam epochs with negative_selection_range=200000 (so negatives
come from per-epoch random-projection windows) and scale adaptation.

A step is one complete 200-epoch layout from the same resident initial
state (graph, clocks, init and directions already in HBM). value =
E * 200 * steps / seconds. The roofline object prices the dominant kernels
(the edge-balanced unit sweep + its Adam step) with the algorithmic stream
bytes of each epoch, divided by their average duration per epoch measured
with device events on the plan's stream; `traffic` is the measured HBM
bytes per epoch of the same kernels from the committed rocprofv3 counter
summary. The CPU baseline times the oracle (C restatement, OpenMP over
rows) on the whole 200-epoch layout from the same graph and init;
`quality` compares trustworthiness@15 and kNN preservation of the device
layout and of that oracle layout on a fixed 10,000-row sample (the gate:
within 0.01).

Side legs (each skippable): `knn_batches` times the exact kNN of query
batches of 1 to 16,384 rows against the 1M references and prices each
against its roofline (HBM for the BF16 stream below about 310 rows, MFMA
above); `batched` runs 256 independent layouts as one hipGraph launch;
`serving` runs the transform server under Poisson arrivals; `fit_2k` times
the whole CPU fit of the 2,000 x 10 blob data (and the device fit), with
trustworthiness; `layout_100k` times the 100,000 x 50 global-negative
layout with exclusion on the device and in the oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_data(rows, dims, clusters, seed, device, row_seed=None):
    """Rows of the 10-cluster 0..255 mixture; centres come from `seed`, the
    rows from `row_seed` (default: the same generator as the centres)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    centers = 32.0 + 192.0 * torch.rand((clusters, dims), generator=g, device=device)
    if row_seed is not None:
        g = torch.Generator(device=device)
        g.manual_seed(row_seed)
    labels = torch.randint(0, clusters, (rows,), generator=g, device=device)
    x = torch.empty((rows, dims), dtype=torch.float32, device=device)
    step = 1 << 17
    for s in range(0, rows, step):
        e = min(rows, s + step)
        noise = torch.randn((e - s, dims), generator=g, device=device) * 24.0
        x[s:e] = torch.clamp(torch.round(centers[labels[s:e]] + noise), 0.0, 255.0)
    return x


def stream_bytes(graph, epochs, h, d, windowed):
    """Sum over epochs of the algorithmic stream bytes B_stream(t):
    8E + F_t*44 + S_t*4 + H*(4 + 4*windowed) + 4H + H*d*4*6, with F_t (firing
    edges) and S_t (negatives drawn) replayed exactly from the clocks.
    Returns (total bytes, sum F_t, sum S_t, per-epoch F_t + S_t)."""
    import torch

    dev = torch.device("cuda:0")
    npos = torch.as_tensor(graph["next_positive"], device=dev).clone()
    nneg = torch.as_tensor(graph["next_negative"], device=dev).clone()
    per = torch.as_tensor(graph["periods"], device=dev)
    nper = torch.as_tensor(graph["negative_periods"], device=dev)
    E = npos.numel()
    total = 0
    F = S = 0
    terms = []
    for t in range(epochs):
        fire = npos <= float(t)
        k = torch.trunc((float(t) - nneg) / nper)
        cnt = torch.where(fire, k, torch.zeros_like(k))
        f_t = int(fire.sum().item())
        s_t = int(cnt.clamp_min(0).sum().item())
        npos = torch.where(fire, npos + per, npos)
        nneg = torch.where(fire, nneg + cnt * nper, nneg)
        F += f_t
        S += s_t
        terms.append(f_t + s_t)
        total += 8 * E + f_t * 44 + s_t * 4 + h * (4 + 4 * int(windowed)) + 4 * h + h * d * 24
    return total, F, S, terms


def build_workload(rows, dims, clusters, seed, k, epochs):
    import torch

    from cuscan_amd import graph as G
    from cuscan_amd.knn import KnnIndex

    dev = torch.device("cuda:0")
    t0 = time.time()
    x = make_data(rows, dims, clusters, seed, dev)
    torch.cuda.synchronize()
    log(f"data {rows}x{dims} in {time.time() - t0:.1f}s")
    t0 = time.time()
    chunk = 131072
    index = KnnIndex(x, max_queries=min(rows, chunk))
    idx = np.empty((rows, k - 1), np.int32)
    dist = np.empty((rows, k - 1), np.float32)
    for s in range(0, rows, chunk):
        e = min(rows, s + chunk)
        i_d, d_d = index.query_device(x[s:e], k - 1, exclude_self=True, query_offset=s)
        idx[s:e] = i_d.cpu().numpy()
        dist[s:e] = d_d.cpu().numpy()
        log(f"  kNN rows {e}/{rows} ({time.time() - t0:.1f}s)")
    knn_s = time.time() - t0
    log(f"exact kNN ({k - 1} non-self) in {knn_s:.1f}s")
    t0 = time.time()
    indices = np.concatenate([np.arange(rows, dtype=np.int32)[:, None], idx], axis=1)
    distances = np.concatenate([np.zeros((rows, 1), np.float32), dist], axis=1)
    compact = G.fit_graph_device(indices, distances, epochs=epochs, rate=5.0,
                                 prune_divisor=float(epochs))
    a, b = G.fit_curve(0.1, 1.0)
    init = G.random_init(rows, 2, seed)
    graph_s = time.time() - t0
    log(f"fuzzy graph + compaction (device): E={len(compact['indices'])} in {graph_s:.2f}s")
    return x, compact, init, a, b, knn_s, graph_s, index


def recursive_benchmark(x, compact, a, b, args):
    """init='auto' on the benchmark graph (defaults: ratio 4, base 1024, 50
    coarse epochs, degree hubs, 16,384-row PCA anchor), then isotropic
    rescale; with --quality the 200-epoch fit from it (good_initialization)
    and its trustworthiness on the same 10k sample."""
    import torch

    from cuscan_amd import layout as L
    from cuscan_amd import recursive as R

    xs = x.cpu().numpy()
    t0 = time.perf_counter()
    values, st = R.recursive_init(compact["offsets"], compact["indices"], compact["weights"], xs,
                                  2, a, b, 42, R.params(), return_stats=True)
    R.init_rescale(values, "isotropic", offsets=compact["offsets"], indices=compact["indices"],
                   weights=compact["weights"])
    secs = time.perf_counter() - t0
    res = {"seconds": round(secs, 2), "host_ms": round(st["host_ms"], 1),
           "coarse_layout_ms": round(st["layout_ms"], 1), "level_rows": st["rows"],
           "level_edges": st["edges"], "propagation_sweeps": st["sweeps"],
           "host_stage_ms": st["stage_ms"]}
    log(f"recursive init {secs:.2f}s (host {st['host_ms']:.0f} ms, coarse layouts "
        f"{st['layout_ms']:.0f} ms), levels {st['rows']}")
    if args.quality:
        from cuscan_amd.quality import trustworthiness_device

        params = L.LayoutParams(epochs=args.epochs, a=a, b=b, seed=42, good_initialization=True)
        plan = L.LayoutPlan(values, None, compact, params, shared=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plan.run()
        emb, _ = plan.download()
        res["layout_seconds"] = round(time.perf_counter() - t1, 3)
        tw, pres = trustworthiness_device(xs, emb, k=15, sample=10_000)
        res["trustworthiness@15"] = round(tw, 4)
        res["knn_preservation@15"] = round(pres, 4)
        tw0, _ = trustworthiness_device(xs, values, k=15, sample=10_000)
        res["init_trustworthiness@15"] = round(tw0, 4)
        log(f"recursive fit quality {tw:.4f}/{pres:.4f} (init alone {tw0:.4f})")
    return res


def knn_batch_benchmark(index, args, sizes=(1, 64, 512, 2048, 4096, 16384), reps=5):
    """Exact k = 15 kNN of query batches (rows of the same mixture, row seed
    43) against the resident 1M references: the BF16 prefilter streams the
    [N, 800] BF16 copy once per batch (HBM bound below about 310 rows, where
    2 q FLOP per streamed byte meets the 2.5 PFLOP/s / 8 TB/s ridge), then
    the merge and the FP32 rerank. Mean of `reps` timed calls after two
    warm-up calls."""
    import torch

    n, d = index.n, index.d
    dp = (d + 31) // 32 * 32
    stream_bytes = n * dp * 2 + n * 4            # BF16 rows + norms
    xq = make_data(max(sizes), args.dims, 10, 42, torch.device("cuda:0"), row_seed=43)
    out = []
    for q in sizes:
        qs = xq[:q].contiguous()
        for _ in range(2):
            index.query_device(qs, 15)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            index.query_device(qs, 15)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        flops = 2.0 * q * n * dp
        hbm = q < 310
        achieved = stream_bytes / t / 1e9 if hbm else flops / t / 1e12
        peak = 8000.0 if hbm else 2500.0
        out.append({"rows": q, "ms": round(t * 1e3, 3), "bound": "hbm" if hbm else "mfma",
                    "achieved": round(achieved, 1), "unit": "GB/s" if hbm else "TFLOP/s",
                    "peak": peak, "frac": round(achieved / peak, 4)})
        log(f"knn batch {q}: {t * 1e3:.3f} ms, {achieved:.1f} {'GB/s' if hbm else 'TFLOP/s'}")
    return out


def serve_benchmark(x, emb, compact_index, a, b, args):
    """Transform serving on the resident 1M-row model: requests of
    round(lognormal(ln 64, 1)) rows clipped to 1..4096 from the same mixture
    (rows seed 43), 100-epoch Adam transform (lr/4, transform_seed 42)."""
    import threading

    import torch

    from cuscan_amd import graph as G
    from cuscan_amd import layout as L
    from cuscan_amd.serve import TransformServer

    rng = np.random.default_rng(43)
    n_plain, n_served = args.serve_plain, args.serve_requests
    sizes = np.clip(np.round(rng.lognormal(np.log(64), 1.0, n_served + n_plain)), 1, 4096).astype(int)
    pool = make_data(int(sizes.sum()), x.shape[1], 10, 42, x.device, row_seed=43).cpu().numpy()
    starts = np.concatenate([[0], np.cumsum(sizes)])
    req = [pool[starts[i]:starts[i + 1]] for i in range(len(sizes))]
    epochs, k = 100, 15
    emb = np.ascontiguousarray(emb, np.float32)

    # ---- plain per-request transform (the UMAP.transform call chain)
    params = L.LayoutParams(epochs=epochs, a=a, b=b, initial_alpha=0.25, seed=42)

    emb_dev = torch.as_tensor(emb, device=x.device)

    def plain(xq):
        idx, dist = compact_index.query(xq, k)
        cg = G.transform_graph_device(idx, dist, emb_dev, local_connectivity=0.0, epochs=epochs,
                                      rate=5.0)
        out, _, _, _ = L.optimize_compact(cg["init"], emb, cg, params, shared=False)
        return out

    plain(req[-1])  # warm
    t_plain, plain_out = [], []
    for i in range(n_plain):
        t0 = time.perf_counter()
        plain_out.append(plain(req[n_served + i]))
        t_plain.append((time.perf_counter() - t0) * 1e3)
    plain_rows = int(sizes[n_served:].sum())
    plain_rps = plain_rows / (sum(t_plain) / 1e3)
    log(f"plain transform: {n_plain} requests, {plain_rows} rows, {plain_rps:.0f} rows/s, "
        f"p50 {np.percentile(t_plain, 50):.2f} ms")

    t0 = time.time()
    srv = TransformServer(x.cpu().numpy(), emb, a=a, b=b, n_neighbors=k, epochs=epochs,
                          learning_rate=1.0, negative_selection_range=200_000, transform_seed=42,
                          max_batch_rows=8192)
    setup_s = time.time() - t0
    log(f"server setup {setup_s:.1f}s")
    # equal accuracy: the served coordinates of the plain requests
    same = True
    for i in range(n_plain):
        out, _ = srv.wait(srv.submit(req[n_served + i]))
        same = same and bool((out == plain_out[i]).all())

    def run(rate):
        """Poisson arrivals at `rate` requests/s (inf: all at once)."""
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_served)) if np.isfinite(rate) else \
            np.zeros(n_served)
        tickets = [None] * n_served
        t_start = time.perf_counter()

        def submitter():
            for i in range(n_served):
                delay = t_start + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(req[i])

        th = threading.Thread(target=submitter)
        th.start()
        lat = []
        for i in range(n_served):
            while tickets[i] is None:
                time.sleep(0.0002)
            res = srv.wait(tickets[i], 120.0)
            lat.append(res[1])
        t_end = time.perf_counter()
        th.join()
        rows = int(sizes[:n_served].sum())
        return {"rate_req_s": None if not np.isfinite(rate) else rate,
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "rows_per_s": round(rows / (t_end - t_start), 1),
                "requests": n_served}

    def stage_means(before, after):
        n = max(after["requests"] - before["requests"], 1)
        return {k: round((after[k] * after["requests"] - before[k] * before["requests"]) / n, 3)
                for k in ("mean_queue_ms", "mean_knn_ms", "mean_graph_ms", "mean_layout_ms")}

    sweep = []
    for rate in args.serve_rates:
        before = srv.stats()
        sweep.append(run(rate))
        sweep[-1]["stage_ms"] = stage_means(before, srv.stats())
        log(f"served rate {rate}: {sweep[-1]}")
    sat = run(float("inf"))
    log(f"served saturation: {sat}")
    stats = srv.stats()
    srv.close()

    # BF16 prefilter recall@15 against fp32 brute force on 256 query rows
    qs = torch.from_numpy(pool[:256]).to(x.device)
    cand, counts = compact_index.candidates_device(qs)
    d2 = torch.cdist(qs, x) ** 2
    exact = torch.topk(d2, k, largest=False).indices.to(torch.int32)
    top15 = cand[:, :k]
    recall = (top15.unsqueeze(2) == exact.unsqueeze(1)).any(2).float().mean().item()
    coverage = (cand.unsqueeze(2) == exact.unsqueeze(1)).any(1).float().mean().item()
    return {
        "metric": "transform p50/p99 ms and rows/s, served vs plain",
        "workload": "resident 1M x 784 model; requests of round(lognormal(ln 64, 1)) rows in "
                    "1..4096 from the same mixture (rows seed 43), Poisson arrivals; exact k=15 "
                    "kNN (BF16 MFMA prefilter + FP32 rerank), 100-epoch Adam transform",
        "plain": {"requests": n_plain, "rows_per_s": round(plain_rps, 1),
                  "p50_ms": round(float(np.percentile(t_plain, 50)), 3),
                  "p99_ms": round(float(np.percentile(t_plain, 99)), 3)},
        "served_sweep": sweep, "served_saturation": sat,
        "speedup_rows_per_s": round(sat["rows_per_s"] / plain_rps, 2),
        "served_equals_plain": same, "server_setup_s": round(setup_s, 2),
        "layout_launches": stats["steps"],
        "requests_per_layout_launch": round(stats["mean_active"], 2),
        "peak_requests_in_flight": stats["peak_active"],
        "prefilter_recall_at_15": round(recall, 5), "prefilter_candidate_coverage": round(coverage, 5),
    }


def batch_benchmark(args):
    """256 independent fits: rows round(lognormal(ln 5000 - 0.5, 1)) clipped to
    500..50000, 32-D mixtures of U{5..20} unit-variance clusters (centres in
    [-10, 10]), exact k=15 graphs, seeds 42+g, 200 Adam epochs (global
    negatives). Batched (autotuned tile shape, the 200 per-epoch launches
    replayed as one hipGraph) vs the same batch in one persistent cooperative
    launch (grid barrier between epochs) and vs the same layouts as
    sequential single-graph plans."""
    import torch

    from cuscan_amd import graph as G
    from cuscan_amd import layout as L
    from cuscan_amd.batch import SHAPES, LayoutBatch
    from cuscan_amd.knn import KnnIndex

    n_graphs, epochs = args.batch_graphs, 200
    t0 = time.time()
    graphs, heads, seeds = [], [], []
    size_rng = np.random.default_rng(42)
    sizes = np.clip(np.round(size_rng.lognormal(np.log(5000) - 0.5, 1.0, n_graphs)), 500,
                    50_000).astype(int)
    for g in range(n_graphs):
        rng = np.random.default_rng(42 + g)
        c = int(rng.integers(5, 21))
        centers = rng.uniform(-10, 10, (c, 32))
        x = (centers[rng.integers(0, c, sizes[g])] + rng.normal(0, 1, (sizes[g], 32))).astype(np.float32)
        index = KnnIndex(torch.from_numpy(x).cuda(), max_queries=int(sizes[g]))
        idx, dist = index.query(x, 14, exclude_self=True)
        index.close()
        n = sizes[g]
        indices = np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1)
        distances = np.concatenate([np.zeros((n, 1), np.float32), dist], 1)
        fg = G.fuzzy_graph(indices, distances)
        graphs.append(G.compact_graph(n, n, fg["heads"], fg["tails"], fg["weights"], epochs=epochs,
                                      rate=5.0, prune_divisor=float(epochs)))
        heads.append(G.random_init(n, 2, 42 + g))
        seeds.append(42 + g)
    a, b = G.fit_curve(0.1, 1.0)
    build_s = time.time() - t0
    total_e = sum(len(gr["indices"]) for gr in graphs)
    log(f"batched workload: {n_graphs} graphs, {int(sizes.sum())} rows, {total_e} edges "
        f"({build_s:.1f}s)")
    params = L.LayoutParams(epochs=epochs, a=a, b=b)
    batch = LayoutBatch(graphs, heads, params, seeds, persistent=False)

    def timed(runs):
        batch.run(use_graph=True)   # capture (per-epoch form) + warm
        first, _ = batch.download()
        times = []
        for _ in range(runs):
            t1 = time.perf_counter()
            batch.run(use_graph=True)
            out, dev_ms = batch.download()
            times.append(time.perf_counter() - t1)
        return min(times), dev_ms, out

    shape, shape_ms = batch.tune(probe_epochs=20)
    batched_s, dev_ms, emb_g = timed(args.batch_steps)
    batch.set_persistent(True)
    shape_p, shape_ms_p = batch.tune(probe_epochs=20)
    persist_s, persist_dev_ms, emb_p = timed(args.batch_steps)
    same = shape_p != shape or all((x == y).all() for x, y in zip(emb_p, emb_g))
    batch.close()
    # the same layouts one graph at a time (resident plans; runs timed)
    plans = [L.LayoutPlan(h, None, gr, L.LayoutParams(epochs=epochs, a=a, b=b, seed=sd), shared=True)
             for gr, h, sd in zip(graphs, heads, seeds)]
    plans[0].run()
    plans[0].download()
    t1 = time.perf_counter()
    for pl in plans:
        pl.run()
        pl.download()
    seq_s = time.perf_counter() - t1
    for pl in plans:
        pl.close()
    value = total_e * epochs / batched_s
    seq = total_e * epochs / seq_s
    log(f"batched {batched_s * 1e3:.1f} ms ({value:.3e} edge-epochs/s, shape {SHAPES[shape]}, "
        f"one hipGraph of per-epoch launches) vs one persistent launch {persist_s * 1e3:.1f} ms "
        f"(shape {SHAPES[shape_p]}) vs sequential {seq_s * 1e3:.1f} ms ({seq:.3e})")
    return {"metric": "batched layouts edge-epochs/s (sum E_g * 200 / t)",
            "workload": f"{n_graphs} graphs, rows lognormal(ln 5000 - 0.5, 1) in 500..50000 "
                        "(32-D, U{5..20} clusters, k=15, seeds 42+g), 200 Adam epochs",
            "graphs": n_graphs, "rows": int(sizes.sum()), "edges": int(total_e),
            "value": value, "ms": round(batched_s * 1e3, 2), "device_ms": round(dev_ms, 2),
            "sequential_value": seq, "sequential_ms": round(seq_s * 1e3, 2),
            "speedup": round(seq_s / batched_s, 2), "tile_shape": SHAPES[shape],
            "tune_ms": [round(t, 3) for t in shape_ms],
            "launch": "one hipGraph launch replaying the 200 per-epoch segment sweeps",
            "persistent_ms": round(persist_s * 1e3, 2),
            "persistent_device_ms": round(persist_dev_ms, 2),
            "persistent_tile_shape": SHAPES[shape_p],
            "persistent_tune_ms": [round(t, 3) for t in shape_ms_p],
            "persistent_equals_graph": bool(same)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def oracle_layout(compact, init, a, b, epochs):
    """The oracle's whole layout (all host threads) from the same graph and
    init as the device plan: (embedding, seconds, threads)."""
    from oracle import oracle as O

    threads = host_threads()
    t0 = time.time()
    emb = O.optimize_compact(init, None, compact, epochs=epochs, a=a, b=b,
                             negative_selection_range=200_000, seed=42, threads=threads,
                             collect_stats=False)[0]
    return emb, time.time() - t0, threads


def quality_leg(x, emb, emb_oracle):
    """trustworthiness@15 and kNN preservation@15 of the device and oracle
    layouts on the same fixed 10,000-row sample (comparison/umap06/
    quality.py semantics)."""
    from cuscan_amd.quality import trustworthiness_device

    xs = x.cpu().numpy()
    tw, pres = trustworthiness_device(xs, emb, k=15, sample=10_000)
    tw_c, pres_c = trustworthiness_device(xs, emb_oracle, k=15, sample=10_000)
    log(f"quality: device {tw:.4f}/{pres:.4f}, oracle {tw_c:.4f}/{pres_c:.4f}")
    return {"trustworthiness@15": round(tw, 4), "oracle_trustworthiness@15": round(tw_c, 4),
            "delta": round(tw - tw_c, 4), "knn_preservation@15": round(pres, 4),
            "oracle_knn_preservation@15": round(pres_c, 4), "sample_rows": 10_000,
            "gate": "|trustworthiness - oracle| <= 0.01", "pass": abs(tw - tw_c) <= 0.01}


def fit_2k_leg():
    """2,000 x 10 rows from 10 isotropic Gaussian blobs (centres U[-10, 10]^10,
    sigma 1, 200 rows each, numpy seed 42); UMAP(n_neighbors=15,
    n_components=2, optimizer=adam, init=random, 200 epochs,
    negative_sample_rate=5): the whole CPU fit of the oracle (kNN, graph,
    layout) on all host threads, and the device fit of the same data."""
    import torch

    from cuscan_amd.quality import trustworthiness
    from cuscan_amd.umap import UMAP
    from oracle import oracle as O

    rng = np.random.default_rng(42)
    centers = rng.uniform(-10, 10, (10, 10))
    x = np.concatenate([centers[c] + rng.normal(0, 1.0, (200, 10)) for c in range(10)])
    x = x.astype(np.float32)
    threads = host_threads()
    t0 = time.time()
    emb_c = O.fit(x, n_neighbors=15, n_epochs=200, seed=42, threads=threads)
    cpu_s = time.time() - t0
    model = UMAP(n_neighbors=15, n_components=2, optimizer="adam", init="random", n_epochs=200,
                 negative_sample_rate=5, random_state=42)
    model.fit(x)   # warm
    torch.cuda.synchronize()
    t0 = time.time()
    emb_d = UMAP(n_neighbors=15, n_components=2, optimizer="adam", init="random", n_epochs=200,
                 negative_sample_rate=5, random_state=42).fit_transform(x)
    dev_s = time.time() - t0
    tw_c, _ = trustworthiness(x, emb_c, k=15)
    tw_d, _ = trustworthiness(x, emb_d, k=15)
    log(f"2k fit: cpu {cpu_s:.3f}s tw {tw_c:.4f}; device {dev_s:.3f}s tw {tw_d:.4f}")
    return {"workload": "2000x10 blobs (10 clusters, sigma 1, seed 42), n_neighbors=15, adam, "
                        "init=random, 200 epochs, negative_sample_rate=5",
            "cpu_fit_seconds": round(cpu_s, 3), "cpu_trustworthiness@15": round(tw_c, 4),
            "cpu_cores": threads, "device_fit_seconds": round(dev_s, 3),
            "device_trustworthiness@15": round(tw_d, 4)}


def layout_100k_leg(steps=5):
    """100,000 x 50 rows of 50 clusters (centres U[-10, 10]^50, sigma
    U[0.5, 2], seed 42), exact k=15 fuzzy graph, 200 Adam epochs, global
    negatives with exclude_graph_neighbors: edge-epochs/s of the device plan
    (mean of `steps` layouts) and of the oracle's whole layout."""
    import torch

    from cuscan_amd import graph as G
    from cuscan_amd import layout as L
    from cuscan_amd.knn import KnnIndex
    from oracle import oracle as O

    rng = np.random.default_rng(42)
    centers = rng.uniform(-10, 10, (50, 50))
    sig = rng.uniform(0.5, 2.0, 50)
    lab = rng.integers(0, 50, 100_000)
    x = (centers[lab] + rng.normal(0, 1, (100_000, 50)) * sig[lab, None]).astype(np.float32)
    n = len(x)
    index = KnnIndex(torch.from_numpy(x).cuda(), max_queries=n)
    idx, dist = index.query(x, 14, exclude_self=True)
    index.close()
    g = G.fit_graph_device(np.concatenate([np.arange(n, dtype=np.int32)[:, None], idx], 1),
                           np.concatenate([np.zeros((n, 1), np.float32), dist], 1), epochs=200,
                           rate=5.0, prune_divisor=200.0)
    a, b = G.fit_curve(0.1, 1.0)
    head = G.random_init(n, 2, 42)
    E = len(g["indices"])
    params = L.LayoutParams(epochs=200, a=a, b=b, exclude_graph_neighbors=True)
    plan = L.LayoutPlan(head, None, g, params, shared=True)
    plan.run()
    plan.download()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    plan.download()
    dev_s = (time.perf_counter() - t0) / steps
    plan.close()
    threads = host_threads()
    t0 = time.time()
    O.optimize_compact(head, None, g, epochs=200, a=a, b=b, exclude=True, threads=threads,
                       collect_stats=False)
    cpu_s = time.time() - t0
    log(f"100k layout: device {dev_s * 1e3:.1f} ms ({E * 200 / dev_s:.3e} edge-epochs/s), "
        f"oracle {cpu_s:.2f}s ({E * 200 / cpu_s:.3e})")
    return {"workload": "100000x50, 50 clusters, exact k=15 fuzzy graph, 200 Adam epochs, global "
                        "negatives, exclude_graph_neighbors", "edges": E,
            "device_ms": round(dev_s * 1e3, 2), "device_value": E * 200 / dev_s,
            "cpu_seconds": round(cpu_s, 2), "cpu_value": E * 200 / cpu_s, "cpu_cores": threads,
            "unit": "edge-epochs/s", "speedup": round(cpu_s / dev_s, 1)}


def cpu_baseline(compact, init, a, b, epochs, terms, oracle_run):
    """The oracle's whole layout (C restatement, OpenMP over rows, all host
    threads) on the same graph and init: E * epochs / seconds. The
    one-thread figure times epochs 0..3 on one thread and scales them to the
    layout average by the exact per-epoch force-term counts (the oracle's
    cost follows the firing edges and negatives drawn, which are data
    independent)."""
    from oracle import oracle as O

    _, secs, threads = oracle_run
    E = len(compact["indices"])
    mean_terms = sum(terms) / len(terms)
    one_epochs = 4
    t0 = time.time()
    O.optimize_compact(init, None, compact, epochs=epochs, a=a, b=b, negative_selection_range=200_000,
                       seed=42, threads=1, stop_epoch=one_epochs, collect_stats=False)
    one_secs = time.time() - t0
    one_ratio = mean_terms / (sum(terms[:one_epochs]) / one_epochs)
    one = E * one_epochs / one_secs / one_ratio
    return {"value": E * epochs / secs, "unit": "edge-epochs/s", "cores": threads, "kind": "port",
            "one_thread_value": one,
            "sample": f"oracle C restatement (OpenMP, {threads} threads, {cpu_model()}, "
                      f"nproc {os.cpu_count()}): the whole {epochs}-epoch layout from the same "
                      f"graph and init ({secs:.1f}s); one_thread_value: 1 thread on epochs "
                      f"0..{one_epochs - 1} ({one_secs:.1f}s), scaled to the layout average by "
                      f"force terms per epoch"}


def copy_calibration(gib):
    """A plain device copy of a known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch

    n = int(gib * (1 << 30)) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4}), flush=True)


def measured_traffic():
    """HBM bytes per epoch of the edge pass + Adam step from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch,
    scaled by the copy calibration), or None."""
    path = ROOT / "profiles" / "unit_sweep_1m_traffic.json"
    if not path.exists():
        return None
    t = json.loads(path.read_text())
    return int(t["bytes_per_epoch"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--dims", type=int, default=784)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quality", action="store_true",
                    help="skip the trustworthiness@15 comparison with the oracle's layout")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 2,000 x 10 fit and 100,000 x 50 layout legs")
    ap.add_argument("--recursive", action="store_true",
                    help="also time the recursive initialisation (init='auto') on the same "
                         "graph, and with --quality the fit that starts from it")
    ap.add_argument("--no-serve", action="store_true", help="skip the transform serving leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-layouts leg")
    ap.add_argument("--no-knn", action="store_true", help="skip the kNN batch-size leg")
    ap.add_argument("--batch-graphs", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-plain", type=int, default=40)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[50.0, 200.0, 800.0])
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB GiB (counter calibration)")
    args = ap.parse_args()
    args.quality = not args.no_quality
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from cuscan_amd import layout as L

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    x, compact, init, a, b, knn_s, graph_s, index = build_workload(args.rows, args.dims, 10, 42,
                                                                   15, args.epochs)
    E = len(compact["indices"])
    H = args.rows
    params = L.LayoutParams(epochs=args.epochs, a=a, b=b, seed=42)
    plan = L.LayoutPlan(init, None, compact, params, shared=True)
    for _ in range(args.warmup):
        plan.run()
        plan.download()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()
    emb, stats = plan.download()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    value = E * args.epochs * args.steps / secs
    log(f"{args.steps} layouts in {secs:.3f}s: {value:.3e} edge-epochs/s "
        f"(device {stats['device_ms']:.1f} ms/layout)")
    # per-kernel durations from device events (one more profiled run)
    plan.set_profiling(True)
    plan.run()
    _, pstats = plan.download()
    plan.set_profiling(False)
    sweep_avg_ms = pstats["sweep_ms"] / max(pstats["sweep_launches"], 1)
    total_bytes, F, S, terms = stream_bytes(compact, args.epochs, H, 2, windowed=True)
    bytes_per_launch = total_bytes / args.epochs
    achieved = bytes_per_launch / (sweep_avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": measured_traffic() if args.rows == 1_000_000 else None,
                "kernel": "unit_schedule_kernel (its passes spread over the epochs) + rank "
                          "prefix + unit_sweep_kernel + unit_adam_ranked_kernel: the edge pass "
                          "and lazy Adam of one epoch",
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(sweep_avg_ms, 4),
                "firing_edges_per_epoch": F / args.epochs, "negatives_per_epoch": S / args.epochs,
                # B_gather: the 8-byte position gathers (tails and negative table rows),
                # reported beside the stream bytes and not in the HBM fraction
                "gather_bytes_per_epoch": int(8 * (F + S) / args.epochs),
                "refresh_avg_ms": round(pstats["refresh_ms"] / max(pstats["refresh_launches"], 1), 4),
                "rank_ordered_epochs": pstats["rank_ordered_epochs"],
                "compact_records": pstats["compact_records"],
                # epochs per clock-schedule pass (its launches are inside the averaged time)
                "clock_schedule": pstats["clock_schedule"]}
    out = {
        "metric": "layout edge-epochs/s at 1M rows", "value": value, "unit": "edge-epochs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"modern Adam layout, {args.rows}x{args.dims} 10-cluster 0..255 "
                               f"mixture, exact k=15 fuzzy graph, {args.epochs} epochs, "
                               "projection windows (Q=200000), 2-D",
                   "rows": args.rows, "edges": E, "epochs": args.epochs,
                   "negative_selection_range": 200_000, "knn_seconds": round(knn_s, 2),
                   "graph_seconds": round(graph_s, 2),
                   "final_scale": round(stats["final_scale"], 6)},
        "roofline": roofline,
    }
    oracle_run = None
    if args.quality or not args.no_cpu:
        log("oracle layout (whole) ...")
        oracle_run = oracle_layout(compact, init, a, b, args.epochs)
        log(f"oracle layout {oracle_run[1]:.1f}s on {oracle_run[2]} threads")
    if args.quality:
        out["quality"] = quality_leg(x, emb, oracle_run[0])
    if args.recursive:
        log("recursive initialisation ...")
        out["recursive_init"] = recursive_benchmark(x, compact, a, b, args)
    if not args.no_knn and args.rows >= 16384:
        log("kNN batches ...")
        out["knn_batches"] = knn_batch_benchmark(index, args)
    if not args.no_batch:
        log("batched layouts ...")
        out["batched"] = batch_benchmark(args)
    if not args.no_serve:
        log("transform serving ...")
        out["serving"] = serve_benchmark(x, emb, index, a, b, args)
    if not args.no_small:
        log("2k fit ...")
        out["fit_2k"] = fit_2k_leg()
        log("100k layout ...")
        out["layout_100k"] = layout_100k_leg()
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(compact, init, a, b, args.epochs, terms, oracle_run)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
