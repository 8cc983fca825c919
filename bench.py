#!/usr/bin/env python3
"""Benchmark of the dual-mode edge engine on one MI355X.

A step is one PageRank pull iteration (gather over every in-edge + fused
apply) over an RMAT graph resident in HBM: scale 27, edge factor 16,
a/b/c/d = .57/.19/.19/.05, 2^27 ids, 2^31 directed edges, FP32 state with
exact (binned) or compensated (merge-path) sums.  Both pull layouts are timed;
`value` is GTEPS = edges / step time of the faster one.

Printed as one JSON line (rank 0), with
  roofline     achieved HBM GB/s of the pull kernels of the headline layout
               (algorithmic bytes per iteration / average kernel time per
               iteration, HIP events on the library stream)
  cpu_baseline the FP64 CPU oracle (OpenMP restatement of the reference's pull
               loop) timed on a bounded slice of the same graph on the host
  accuracy     L1 / L-infinity of a full FP32 run against the FP64 run
  bfs          BFS on RMAT scale 24 symmetrized, 64 roots (GTEPS, Graph500 style)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS + % HBM roofline; batched PPR queries/s & p99 vs 1-seed loop @L1 1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu_info():
    """The host the CPU baseline runs on: cores this process may use (its affinity set, the
    share a GPU box gives one GPU), all cores, CPU model and NUMA nodes."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and d[4:].isdigit()])
    except OSError:
        numa = None
    return {"usable_cores": usable, "cpu_count": os.cpu_count(), "model": model, "numa_nodes": numa}


def cpu_baseline(g, seconds_target=10.0, threads=None):
    """Time the oracle's pull iteration (FP64, OpenMP over every core this process may use)
    on a sample of the graph's columns: 16 evenly spaced ranges of V/256 columns each."""
    import oracle
    info = host_cpu_info()
    threads = threads or info["usable_cores"]
    cp, ri = g.download_csc()
    od = g.out_degrees()
    if g.info.relabeled:  # out_degrees() is in input order; the CSC is internal
        od = od[np.argsort(g.permutation())]

    class H:  # host CSC view for the oracle binding
        pass
    h = H()
    h.V, h.col_ptr, h.row_idx, h.od = g.V, cp, ri, od
    x = np.where(od > 0, 1.0 / np.maximum(od, 1), 1.0)
    y = np.zeros(g.V, np.float64)
    width = max(1, g.V // 256)
    ranges = [(int(k * g.V // 16), int(min(g.V, k * g.V // 16 + width))) for k in range(16)]
    edges = int(sum(cp[hi] - cp[lo] for lo, hi in ranges))
    done_edges, spent, passes = 0, 0.0, 0
    while spent < seconds_target and passes < 50:
        t0 = time.perf_counter()
        for lo, hi in ranges:
            oracle.pagerank_step(h, x, 0.85, False, lo, hi, threads, y=y)
        spent += time.perf_counter() - t0
        done_edges += edges
        passes += 1
    return {"value": done_edges / spent / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port", "host": info,
            "sample": f"{passes} passes of one PageRank pull iteration (FP64, OpenMP) over 16 evenly "
                      f"spaced ranges of {width} columns ({edges} in-edges) of the same graph, "
                      f"{spent:.1f} s"}


def profiled_traffic(kernel, scale):
    """HBM bytes per pull iteration from the committed rocprofv3 summary of this workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{kernel}_rmat{scale}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def measure_pull(ctx, g, layout, steps, warmup, weighted=False):
    """Time `steps` FP32 PageRank iterations on one pull layout: per-kernel HIP events on the
    library stream (timed run), then the same iterations replayed from a captured graph."""
    from plato_amd import _lib as L
    from plato_amd.algo import pagerank as plain, pagerank_opts_t, weighted_pagerank
    pagerank = weighted_pagerank if weighted else plain
    t0 = time.time()
    g.set_pull_layout(layout)
    layout_s = time.time() - t0
    if warmup:
        pagerank(g, pagerank_opts_t(warmup, 0.85, 0.0))
    kts = ([L.PE_KT_PULL_TILE, L.PE_KT_PULL_FIXUP] if layout == "merge_path"
           else [L.PE_KT_PULL_BIN_SCATTER, L.PE_KT_PULL_BIN_GATHER])
    ctx.set_kernel_timing(True)
    for k in kts:
        ctx.kernel_time(k)
    ctx.sync()
    res = pagerank(g, pagerank_opts_t(steps, 0.85, 0.0))
    ctx.sync()
    times = [ctx.kernel_time(k) for k in kts]
    ctx.set_kernel_timing(False)
    avg = [ms / max(n, 1) for ms, n in times]
    res_g = pagerank(g, pagerank_opts_t(steps, 0.85, 0.0))
    out = {"ms_per_step": res.seconds * 1e3 / steps, "graph_ms": res_g.seconds * 1e3 / steps}
    if layout == "merge_path":
        out.update(kernel="k_pull_tiles", kernel_ms=avg[0],
                   extra={"fixup_avg_ms": round(avg[1], 4)})
    else:
        info = g.info
        out.update(kernel="k_pb_scatter+k_pb_gather", kernel_ms=avg[0] + avg[1],
                   extra={"scatter_avg_ms": round(avg[0], 4), "gather_avg_ms": round(avg[1], 4),
                          "layout_build_s": round(layout_s, 2), "run_records": int(info.binned_runs),
                          "layout_bytes": int(info.binned_bytes)})
    log(f"[bench] {'weighted ' if weighted else ''}{layout}: {out['ms_per_step']:.3f} ms/iteration (timed), {out['graph_ms']:.3f} ms "
        f"(graph replay), kernel {out['kernel_ms']:.3f} ms")
    return out


def run_c1(ctx, scale=16, iters=20, seed=1):
    """The CPU-runnable configuration: PageRank, RMAT scale 16, 20 iterations, eps 0, FP64
    ranks (example/pagerank_simple.cc:67-112 with --is_directed=true --iterations 20
    --eps 0).  The device's FP64 run against the oracle's full-graph FP64 run on every
    usable host core, both timed, and their L1 / L-infinity difference."""
    import oracle  # CPU baseline and checker
    from plato_amd import Graph
    from plato_amd.algo import pagerank, pagerank_opts_t
    g = Graph.rmat(ctx, scale, 16, seed=seed)
    E = g.nnz
    pagerank(g, pagerank_opts_t(iters, 0.85, 0.0), precision=64)  # warm-up
    best = None
    for _ in range(5):
        r = pagerank(g, pagerank_opts_t(iters, 0.85, 0.0), precision=64)
        best = r if best is None or r.seconds < best.seconds else best
    g.close()
    s, d = oracle.rmat_edges(scale, 16, seed)
    og = oracle.CSC(1 << scale, s, d, csr=False)
    info = host_cpu_info()
    oracle.pagerank(og, iters, 0.85, 0.0, nthreads=info["usable_cores"])  # warm-up
    t_cpu, reps = 0.0, 0
    while t_cpu < 3.0 and reps < 50:
        t0 = time.perf_counter()
        ref, _ = oracle.pagerank(og, iters, 0.85, 0.0, nthreads=info["usable_cores"])
        t_cpu += time.perf_counter() - t0
        reps += 1
    diff = np.abs(best.rank - ref)
    gpu_ms = best.seconds * 1e3 / iters
    cpu_ms = t_cpu * 1e3 / reps / iters
    return {"workload": f"pagerank_rmat{scale}_fp64_{iters}_iterations", "E": int(E),
            "gpu_ms_per_iteration": round(gpu_ms, 4), "gpu_gteps": round(E / gpu_ms / 1e6, 2),
            "cpu_ms_per_iteration": round(cpu_ms, 3), "cpu_gteps": round(E / cpu_ms / 1e6, 3),
            "cpu_cores": info["usable_cores"], "cpu_kind": "port", "cpu_runs": reps,
            "l1_vs_oracle": float(diff.sum()), "linf_rel_vs_oracle": float(np.max(diff / np.abs(ref)))}


def run_weighted(ctx, args, layouts):
    """weighted_pagerank on the same RMAT graph with FP32 edge weights uniform in (0, 1],
    normalised per source: both pull layouts, algorithmic bytes 8(V+1) + 8E + 12V."""
    from plato_amd import Graph
    from plato_amd.algo import pagerank_opts_t, weighted_pagerank
    log(f"[bench] weighted RMAT scale {args.scale}")
    t0 = time.time()
    g = Graph.rmat(ctx, args.scale, args.edge_factor, seed=args.seed, weighted=True,
                   relabel=not args.no_relabel)
    build_s = time.time() - t0
    V, E = g.V, g.nnz
    runs = {lay: measure_pull(ctx, g, lay, args.steps, args.warmup, weighted=True) for lay in layouts}
    best = min(runs, key=lambda k: runs[k]["ms_per_step"])
    head = runs[best]
    alg_bytes = 8 * (V + 1) + 8 * E + 12 * V
    achieved = alg_bytes / (head["kernel_ms"] * 1e-3) / 1e9
    out = {"gteps": round(E / (head["ms_per_step"] * 1e-3) / 1e9, 2),
           "ms_per_step": round(head["ms_per_step"], 4), "pull_layout": best, "graph_build_s": round(build_s, 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": head["kernel"],
                        "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": round(head["kernel_ms"], 4)},
           "pull_layouts": {k: {kk: vv for kk, vv in r.items() if kk != "extra"} | r["extra"]
                            for k, r in runs.items()}}
    if not args.no_accuracy:
        g.set_pull_layout(best)
        r32 = weighted_pagerank(g, pagerank_opts_t(1000, 0.85, 0.0), precision=32, l1_norm_tol=1e-7)
        r64 = weighted_pagerank(g, pagerank_opts_t(1000, 0.85, 0.0), precision=64, l1_norm_tol=1e-12)
        diff = np.abs(r32.rank - r64.rank)
        out["accuracy"] = {"fp32_iterations": r32.iterations, "fp64_iterations": r64.iterations,
                           "l1_normalised": float(diff.sum() / np.abs(r64.rank).sum()),
                           "linf_rel": float(np.max(diff / np.maximum(np.abs(r64.rank), 1e-300)))}
        del r32, r64, diff
    g.close()
    return out


def run_bfs(ctx, scale, roots, seed=11):
    from plato_amd import Graph
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    t0 = time.time()
    g = Graph.rmat(ctx, scale, 16, seed=seed, directed=False, relabel=True)
    build_s = time.time() - t0
    od = g.out_degrees()
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(od > 0)
    rs = rng.choice(cand, roots, replace=False)
    teps, secs, alg = [], [], []
    for r in rs:
        res = breadth_first_search(g, bfs_opts_t(int(r)))
        m = res.reached_edges / 2  # undirected edges in the component (Graph500)
        teps.append(m / res.seconds / 1e9)
        secs.append(res.seconds)
        alg.append(res.alg_bytes)
    teps = np.array(teps)
    # algorithmic bytes per level (push 8|F| + 4 sum_F deg + 4|new|, pull 8|U| + 4 examined
    # + V/8), summed over the levels of each search, against the search's time
    rate = float(np.sum(alg) / np.sum(secs) / 1e9)
    out = {"scale": scale, "roots": roots, "harmonic_mean_gteps": float(len(teps) / np.sum(1 / teps)),
           "min_gteps": float(teps.min()), "max_gteps": float(teps.max()),
           "mean_ms": float(np.mean(secs) * 1e3), "build_s": round(build_s, 2),
           "relabel_by_degree": True,
           "roofline": {"bound": "hbm", "achieved": round(rate, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(rate / 8000.0, 4), "alg_bytes_per_root_mean": int(np.mean(alg)),
                        "definition": "per level: push 8|F| + 4 sum_F deg + 4|new|; pull 8|U| + "
                                      "4 examined in-edges (each list up to its first frontier member) + V/8"}}
    g.close()
    return out


def run_kcore(ctx, scale, cpu_scale=20, seed=11):
    """k-core (kcore_algo_t::compute_shell_index) on the symmetrized RMAT graph of the BFS
    leg; the oracle's sequential restatement is timed on a smaller graph of the same
    family beside the device run on that graph."""
    import oracle  # CPU baseline (and the check of the device result on its graph)
    from plato_amd import Graph
    from plato_amd.algo import kcore
    g = Graph.rmat(ctx, scale, 16, seed=seed, directed=False, relabel=True)
    kcore(g)  # warm-up
    res = kcore(g)
    out = {"scale": scale, "nnz": int(g.nnz), "epochs": res.epochs, "seconds": round(res.seconds, 4),
           "ms_per_epoch": round(res.seconds * 1e3 / max(res.epochs, 1), 3),
           "max_coreness": int(res.coreness.max()), "modified_total": int(res.modified.sum())}
    g.close()
    s, d = oracle.rmat_edges(cpu_scale, 16, seed)
    og = oracle.CSC(1 << cpu_scale, s, d, symmetrize=True)
    t0 = time.perf_counter()
    core, _, _ = oracle.kcore(og)
    cpu_s = time.perf_counter() - t0
    gs = Graph.rmat(ctx, cpu_scale, 16, seed=seed, directed=False)
    rs = kcore(gs)
    gs.close()
    assert np.array_equal(rs.coreness, core)
    out["cpu_baseline"] = {"scale": cpu_scale, "seconds": round(cpu_s, 3), "cores": 1, "kind": "port",
                           "gpu_seconds_same_graph": round(rs.seconds, 4)}
    return out


def run_cc(ctx, scale, cpu_scale=22, seed=11):
    """Connected components (connected_component_t::compute) on the symmetrized RMAT graph
    of the BFS leg; the oracle's search-based labelling is timed on a smaller graph of
    the same family and checks the device labels there."""
    import oracle  # CPU baseline (and the check of the device result on its graph)
    from plato_amd import Graph
    from plato_amd.algo import connected_components
    g = Graph.rmat(ctx, scale, 16, seed=seed, directed=False, build_csr=True, relabel=True)
    connected_components(g)  # warm-up
    res = connected_components(g)
    out = {"scale": scale, "nnz": int(g.nnz), "epochs": res.epochs, "seconds": round(res.seconds, 4),
           "gteps": round(g.nnz / 2 / res.seconds / 1e9, 2), "components": res.components,
           "modes": res.modes.tolist()}
    g.close()
    s, d = oracle.rmat_edges(cpu_scale, 16, seed)
    og = oracle.CSC(1 << cpu_scale, s, d, symmetrize=True)
    t0 = time.perf_counter()
    lab = oracle.cc_labels(og)
    cpu_s = time.perf_counter() - t0
    gs = Graph.rmat(ctx, cpu_scale, 16, seed=seed, directed=False, build_csr=True)
    rs = connected_components(gs)
    gs.close()
    assert np.array_equal(rs.labels, lab)
    out["cpu_baseline"] = {"scale": cpu_scale, "seconds": round(cpu_s, 3), "cores": 1, "kind": "port",
                           "gpu_seconds_same_graph": round(rs.seconds, 4)}
    return out


def run_hyperanf(ctx, scale, bits=6, cpu_scale=16, seed=11):
    """HyperANF (plato::algo::hyperanf) on the symmetrized RMAT graph of the BFS leg with
    2^bits-register counters (the reference driver's default width); the oracle runs the
    same iteration on a smaller graph of the same family and checks the device there."""
    import oracle  # CPU baseline (and the check of the device result on its graph)
    from plato_amd import Graph
    from plato_amd.algo import hyperanf, hyperanf_opts_t
    g = Graph.rmat(ctx, scale, 16, seed=seed, directed=False, relabel=True)
    hyperanf(g, hyperanf_opts_t(20, bits))  # warm-up
    res = hyperanf(g, hyperanf_opts_t(20, bits))
    iters = max(res.anfs.size, 1)
    out = {"scale": scale, "bits": bits, "nnz": int(g.nnz), "anfs": int(res.anfs.size),
           "avg_distance": res.avg_distance, "seconds": round(res.seconds, 4),
           "counter_bytes_per_edge_iteration": 1 << bits,
           "gteps": round(g.nnz * iters / res.seconds / 1e9, 2)}
    g.close()
    s, d = oracle.rmat_edges(cpu_scale, 16, seed)
    og = oracle.CSC(1 << cpu_scale, s, d, symmetrize=True)
    t0 = time.perf_counter()
    anfs, avg, regs = oracle.hyperanf(og, bits=bits, iters=20)
    cpu_s = time.perf_counter() - t0
    gs = Graph.rmat(ctx, cpu_scale, 16, seed=seed, directed=False)
    rs = hyperanf(gs, hyperanf_opts_t(20, bits), registers=True)
    gs.close()
    assert np.array_equal(rs.registers, regs)
    out["cpu_baseline"] = {"scale": cpu_scale, "seconds": round(cpu_s, 3), "cores": 1, "kind": "port",
                           "gpu_seconds_same_graph": round(rs.seconds, 4)}
    return out


def csv_text(src, dst):
    """`src,dst\n` lines with 10-digit zero-padded ids (22 bytes per edge), built with numpy."""
    E = src.size
    out = np.empty((E, 22), np.uint8)
    for col, arr in ((0, src), (11, dst)):
        v = arr.astype(np.uint64)
        for k in range(9, -1, -1):
            out[:, col + k] = (v % 10).astype(np.uint8) + 48
            v //= 10
    out[:, 10] = ord(",")
    out[:, 21] = ord("\n")
    return out.tobytes()


def run_ingest(ctx, scale, seed=21):
    """Edge-list text -> device graph (pe_graph_create_csv: H2D copy, parse, CSC/CSR build)
    against the reference parser's rules restated on the CPU (oracle, one thread, sample)."""
    from plato_amd import Graph
    src, dst = ctx.rmat_edges(scale, seed, 0, 16 << scale)
    text = csv_text(src, dst)
    del src, dst
    Graph.from_csv_bytes(ctx, text[: 22 * 4096], build_csr=True).close()  # warm-up
    best = 1e30
    for _ in range(2):
        ctx.sync()
        t0 = time.perf_counter()
        gg = Graph.from_csv_bytes(ctx, text, build_csr=True)
        ctx.sync()
        best = min(best, time.perf_counter() - t0)
        E = gg.info.edges
        gg.close()
    import oracle  # CPU baseline: the reference parser's rules on a sample of the same text
    sample = text[: 22 * (1 << 21)]
    t0 = time.perf_counter()
    ns = oracle.parse_csv(sample)[0].size
    t_cpu = time.perf_counter() - t0
    return {"scale": scale, "edges": int(E), "text_bytes": len(text), "seconds": round(best, 4),
            "text_gb_per_s": round(len(text) / best / 1e9, 2), "medges_per_s": round(E / best / 1e6, 1),
            "includes": "host-to-device copy of the text, parse, CSC + CSR build",
            "cpu_reference_parser": {"medges_per_s": round(ns / t_cpu / 1e6, 2), "cores": 1,
                                     "sample": f"first {ns} lines of the same text"}}


def ppr_seeds(od, rng, n=64):
    """Half from the top 1% by out-degree, half with 1 <= od <= median (uneven convergence)."""
    cand = np.flatnonzero(od > 0)
    order = cand[np.argsort(-od[cand], kind="stable")]
    hi = order[: max(1, cand.size // 100)]
    lo = cand[od[cand] <= np.median(od[cand])]
    return [int(rng.choice(hi)) if i % 2 == 0 else int(rng.choice(lo)) for i in range(n)]


def partial_passes(ctx, g, seeds, counts=(64, 48, 32, 16, 8, 1), passes=8):
    """Kernel time of one block pass with n columns active (tolerance 0, so no column
    retires), after two passes that load them."""
    from plato_amd import _lib as L
    from plato_amd.batch import PprBatch
    out = {}
    for n in counts:
        b = PprBatch(g, 0.85, 0.0, 100000)
        for slot in range(n):
            b.admit(slot, [seeds[slot]])
        b.step(2)
        ctx.sync()
        ctx.set_kernel_timing(True)
        ctx.kernel_time(L.PE_KT_PPR_BLOCK)
        b.step(passes)
        ctx.sync()
        ms, k = ctx.kernel_time(L.PE_KT_PPR_BLOCK)
        ctx.set_kernel_timing(False)
        b.close()
        out[str(n)] = round(ms / max(k, 1), 4)
    return out


def run_batched_ppr(ctx, scale, tol=1e-7, max_iters=200, seed=24, single_runs=True):
    """64 PPR queries as one 64-column block vs the same queries one at a time."""
    from plato_amd import Graph
    from plato_amd.algo import (multi_seed_ppr, personalized_pagerank_opts_t,
                                personalized_pagerank_pushpull)
    from plato_amd.batch import PprBatch
    g = Graph.rmat(ctx, scale, 16, seed=seed, relabel=True, build_csr=True)
    od = g.out_degrees()
    seeds = ppr_seeds(od, np.random.default_rng(seed))
    b = PprBatch(g, 0.85, tol, max_iters)
    # warm-up: one query through the captured iteration graph
    b.admit(0, [seeds[0]])
    b.run_until_done([0], chunk=4)
    b.release(0)
    ctx.sync()
    t0 = time.perf_counter()
    for slot, sd in enumerate(seeds):
        b.admit(slot, [sd])
    passes = b.run_until_done(range(64), chunk=4)
    ctx.sync()
    t_batch = time.perf_counter() - t0
    _, iters, _ = b.poll()
    sample = [0, 1, 2, 3, 32, 33, 62, 63]
    batch_ranks = b.results(sample)
    b.close()
    partial = partial_passes(ctx, g, seeds)
    # the plain one-seed GPU engine: pull PPR, same tolerance rule, one query at a time,
    # on each pull layout (the binned one is the faster one-seed engine)
    single = {}
    ref, l1_single, single_iters = {}, [], []
    for layout in (("merge_path", "binned") if single_runs else ()):
        g.set_pull_layout(layout)
        multi_seed_ppr(g, [seeds[0]], max_iters, 0.85, tol, precision=32)  # warm-up
        t_single = 0.0
        for j, sd in enumerate(seeds):
            t0 = time.perf_counter()
            r = multi_seed_ppr(g, [sd], max_iters, 0.85, tol, precision=32)
            t_single += time.perf_counter() - t0
            if layout == "binned":
                single_iters.append(r.iterations)
                if j in sample:
                    ref[j] = multi_seed_ppr(g, [sd], 1000, 0.85, 0.0, precision=64, l1_norm_tol=1e-13).rank
                    l1_single.append(float(np.abs(r.rank - ref[j]).sum()))
        single[layout] = t_single
    g.set_pull_layout("merge_path")
    if not single_runs:  # profiling runs: the block alone
        g.close()
        return {"scale": scale, "k": 64, "tol": tol, "batched_qps": 64 / t_batch,
                "batched_seconds": t_batch, "passes": passes, "pass_ms_by_active_columns": partial}
    l1_batch = [float(np.abs(batch_ranks[i] - ref[j]).sum()) for i, j in enumerate(sample)]
    # push/pull single-seed variant on 8 of the seeds
    t_pp = 0.0
    for sd in seeds[:8]:
        t0 = time.perf_counter()
        personalized_pagerank_pushpull(g, personalized_pagerank_opts_t(max_iters, 0.85, sd, tol))
        t_pp += time.perf_counter() - t0
    g.close()
    best = min(single.values())
    return {"scale": scale, "k": 64, "tol": tol, "batched_qps": 64 / t_batch,
            "batched_seconds": t_batch, "passes": passes,
            "single_seed_qps": 64 / best, "speedup_vs_single": best / t_batch,
            "single_seed_engine": min(single, key=single.get),
            "single_seed_qps_by_layout": {k: 64 / v for k, v in single.items()},
            "speedup_vs_single_by_layout": {k: v / t_batch for k, v in single.items()},
            "pushpull_single_qps": 8 / t_pp,
            "mean_iterations": float(np.mean(iters)), "max_iterations": int(np.max(iters)),
            "single_mean_iterations": float(np.mean(single_iters)),
            "max_l1_vs_fp64_batched": max(l1_batch), "max_l1_vs_fp64_single": max(l1_single),
            "l1_sample_columns": len(sample), "pass_ms_by_active_columns": partial}


def run_serving(ctx, scale, qps_list, n_queries, seed=25):
    """Poisson request stream against a resident graph, 1-256 seeds per query, top-100.
    Per point: HBM admitted (blocks opened x one block's device bytes).  After the sweep:
    the top-100 of sampled queries of the busiest point run again alone in a block (max L1
    difference of the top-100 values, and whether the ids agree), and the latency of the
    one-seed engine (binned pull) on the same graph for comparison."""
    from plato_amd import Graph
    from plato_amd.algo import multi_seed_ppr
    from plato_amd.batch import PprBatch
    from plato_amd.serve import Server, poisson_trace
    g = Graph.rmat(ctx, scale, 16, seed=seed, relabel=True)
    cand = np.flatnonzero(g.out_degrees() > 0)
    alone = PprBatch(g, 0.85, 1e-7, 200)
    block_bytes = alone.device_bytes()
    out = []
    last = None
    for qps in qps_list:
        rng = np.random.default_rng(seed + int(qps))
        arrivals, queries = poisson_trace(rng, n_queries, qps, cand)
        srv = Server(g, 0.85, 1e-7, 200, topk=100, iters_per_pass=2, max_batches=4,
                     hbm_reserve_bytes=16 << 30)
        res = srv.run_trace(arrivals, queries)
        srv.close()
        row = {"offered_qps": qps}
        row.update(res.summary())
        row["admitted_hbm_bytes"] = int(row["blocks_used"] * block_bytes)
        out.append(row)
        last = (queries, res)
        log(f"[bench] serving qps={qps}: {row}")
    queries, res = last
    sample = sorted({0, len(queries) // 3, 2 * len(queries) // 3, len(queries) - 1})
    l1, ids_equal = 0.0, True
    for qi in sample:
        alone.admit(0, queries[qi])
        alone.run_until_done([0], chunk=4)
        ids, vals = alone.topk([0], 100)
        alone.release(0)
        ids_equal &= bool(np.array_equal(ids[0], res.topk_ids[qi]))
        l1 = max(l1, float(np.abs(vals[0].astype(np.float64) - res.topk_vals[qi].astype(np.float64)).sum()))
    alone.close()
    g.set_pull_layout("binned")
    one = []
    for sd in cand[np.random.default_rng(seed).choice(cand.size, 6, replace=False)]:
        r = multi_seed_ppr(g, [int(sd)], 200, 0.85, 1e-7, precision=32)
        one.append(r.seconds)
    g.close()
    return {"scale": scale, "queries_per_point": n_queries, "points": out, "block_bytes": int(block_bytes),
            "alone_vs_mix": {"queries": len(sample), "max_l1_top100": l1, "top100_ids_equal": ids_equal,
                             "offered_qps": qps_list[-1]},
            "one_seed_engine_ms": {"mean": float(np.mean(one[1:]) * 1e3), "runs": len(one) - 1,
                                   "engine": "binned pull, FP32, tol 1e-7"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=27)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tile", type=int, default=0, help="merge-path items per tile (0 = default)")
    ap.add_argument("--no-relabel", action="store_true", help="keep input ids (no degree renumbering)")
    ap.add_argument("--pull-layout", choices=["both", "merge_path", "binned"], default="both",
                    help="pull layouts to time; the faster one is the headline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (reference dispatch for traffic counters)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-weighted", action="store_true", help="skip the weighted PageRank leg")
    ap.add_argument("--bfs-scale", type=int, default=24)
    ap.add_argument("--bfs-roots", type=int, default=64)
    ap.add_argument("--no-bfs", action="store_true")
    ap.add_argument("--no-kcore", action="store_true")
    ap.add_argument("--no-cc", action="store_true")
    ap.add_argument("--no-hyperanf", action="store_true")
    ap.add_argument("--ppr-scale", type=int, default=24)
    ap.add_argument("--no-ppr", action="store_true")
    ap.add_argument("--ppr-block-only", action="store_true",
                    help="time the 64-column block without the one-seed baselines (profiling)")
    ap.add_argument("--serve-scale", type=int, default=25)
    ap.add_argument("--serve-qps", default="4,16,64")
    ap.add_argument("--serve-queries", type=int, default=128)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--ingest-scale", type=int, default=22)
    ap.add_argument("--no-ingest", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from plato_amd import Context, Graph
    from plato_amd.algo import pagerank, pagerank_opts_t

    ctx = Context(0)
    # what the device streams (the ceiling the roofline fraction is read against)
    stream = ctx.stream_rates(1 << 31, 5)
    log(f"[bench] stream rates over 2 GiB: read {stream['read_gbs']:.0f} GB/s, copy {stream['copy_gbs']:.0f} GB/s")
    log(f"[bench] building RMAT scale {args.scale} ef {args.edge_factor} on the device")
    t0 = time.time()
    g = Graph.rmat(ctx, args.scale, args.edge_factor, seed=args.seed, relabel=not args.no_relabel)
    build_s = time.time() - t0
    if args.calibrate_copy:
        from plato_amd import DenseState
        a = DenseState(ctx, 1 << 28, np.float32)
        b = DenseState(ctx, 1 << 28, np.float32)
        a.fill(1.0)
        b.copy_from(a)
        ctx.sync()
        a.close()
        b.close()
    if args.tile:
        g.set_tile(args.tile)
    V, E = g.V, g.nnz
    log(f"[bench] V={V} E={E} built in {build_s:.1f}s tiles={g.info.num_tiles} "
        f"fixups={g.info.num_fixup_rows}")

    layouts = ["merge_path", "binned"] if args.pull_layout == "both" else [args.pull_layout]
    runs = {lay: measure_pull(ctx, g, lay, args.steps, args.warmup) for lay in layouts}
    best = min(runs, key=lambda k: runs[k]["ms_per_step"])
    head = runs[best]
    g.set_pull_layout(best)
    ms_per_step = head["ms_per_step"]
    gteps = E / (ms_per_step * 1e-3) / 1e9
    # algorithmic bytes of one pull iteration: col_ptr 8(V+1), row_idx 4E,
    # x read 4V, out-degree 4V, new state 4V
    alg_bytes = 8 * (V + 1) + 4 * E + 12 * V
    kern_s = head["kernel_ms"] * 1e-3
    achieved = alg_bytes / kern_s / 1e9
    if best == "merge_path":
        traffic = profiled_traffic("k_pull_tiles", args.scale) if (not args.tile and not args.no_relabel) else None
        tsrc = f"profiles/k_pull_tiles_rmat{args.scale}.md (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    else:
        traffic = profiled_traffic("k_pb_scatter_gather", args.scale) if not args.no_relabel else None
        tsrc = f"profiles/k_pb_scatter_gather_rmat{args.scale}.md (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
            "kernel": head["kernel"], "alg_bytes_per_launch": alg_bytes,
            "avg_launch_ms": round(head["kernel_ms"], 4)}
    roof.update(head["extra"])
    roof["stream"] = {"read_gbs": round(stream["read_gbs"], 1), "copy_gbs": round(stream["copy_gbs"], 1),
                      "kernels": "k_read_words, k_copy_words over 2 GiB, 8 blocks of 256 per CU"}
    if traffic:
        roof["traffic_gbs"] = round(traffic / kern_s / 1e9, 1)
    out = {
        "metric": METRIC,
        "value": round(gteps, 2),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"pagerank_pull_iteration_rmat{args.scale}", "scale": args.scale,
                   "edge_factor": args.edge_factor, "V": V, "E": E, "seed": args.seed,
                   "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "damping": 0.85,
                   "state": "fp32 rank/od, fp64 apply",
                   "pull_layout": best, "tile_items": g.info.tile_items,
                   "relabel_by_degree": bool(g.info.relabeled), "graph_build_s": round(build_s, 2)},
        "roofline": roof,
        "graph_replay_ms_per_step": round(head["graph_ms"], 4),
        "pull_layouts": {k: {kk: vv for kk, vv in r.items() if kk != "extra"} | r["extra"]
                         for k, r in runs.items()},
    }
    if not args.no_accuracy:
        log("[bench] accuracy: FP32 vs FP64 runs to delta/sum < 1e-7 and 1e-12")
        r32 = pagerank(g, pagerank_opts_t(1000, 0.85, 0.0), precision=32, l1_norm_tol=1e-7)
        r64 = pagerank(g, pagerank_opts_t(1000, 0.85, 0.0), precision=64, l1_norm_tol=1e-12)
        diff = np.abs(r32.rank - r64.rank)
        out["accuracy"] = {
            "fp32_iterations": r32.iterations, "fp32_seconds": round(r32.seconds, 4),
            "fp64_iterations": r64.iterations,
            "l1_normalised": float(diff.sum() / np.abs(r64.rank).sum()),
            "linf_abs": float(diff.max()),
            "linf_rel": float(np.max(diff / np.maximum(np.abs(r64.rank), 1e-300)))}
        del r32, r64, diff
    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(g)
        log("[bench] C1: PageRank RMAT-16 FP64, 20 iterations, device and oracle")
        out["c1_pagerank_s16_fp64"] = run_c1(ctx)
    g.close()
    if not args.no_weighted:
        out["weighted"] = run_weighted(ctx, args, layouts)
    if not args.no_bfs:
        log(f"[bench] bfs scale {args.bfs_scale}, {args.bfs_roots} roots")
        out["bfs"] = run_bfs(ctx, args.bfs_scale, args.bfs_roots)
    if not args.no_cc:
        log(f"[bench] connected components scale {args.bfs_scale}")
        out["connected_components"] = run_cc(ctx, args.bfs_scale)
    if not args.no_hyperanf:
        log(f"[bench] hyperanf scale {args.bfs_scale}")
        out["hyperanf"] = run_hyperanf(ctx, args.bfs_scale)
    if not args.no_kcore:
        log(f"[bench] k-core scale {args.bfs_scale}")
        out["kcore"] = run_kcore(ctx, args.bfs_scale)
    if not args.no_ppr:
        log(f"[bench] batched PPR scale {args.ppr_scale}, k = 64")
        out["ppr_batched"] = run_batched_ppr(ctx, args.ppr_scale, single_runs=not args.ppr_block_only)
    if not args.no_ingest:
        log(f"[bench] CSV ingestion scale {args.ingest_scale}")
        out["ingest"] = run_ingest(ctx, args.ingest_scale)
    if not args.no_serve:
        log(f"[bench] serving scale {args.serve_scale}")
        out["serving"] = run_serving(ctx, args.serve_scale,
                                     [float(q) for q in args.serve_qps.split(",")],
                                     args.serve_queries)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
