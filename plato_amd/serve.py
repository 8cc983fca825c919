"""PPR serving loop over the C-ABI (pe_server_*): continuous batching of
queries into 64-column rank blocks, Poisson traces, latency statistics."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check
from .engine import Graph


@dataclass
class TraceResult:
    arrival: np.ndarray
    admit: np.ndarray
    done: np.ndarray
    iterations: np.ndarray
    batch: np.ndarray
    slot: np.ndarray
    topk_ids: np.ndarray
    topk_vals: np.ndarray
    wall_s: float

    @property
    def latency(self):
        return self.done - self.arrival

    def summary(self):
        lat = self.latency
        return {"queries": int(lat.size), "queries_per_s": float(lat.size / self.wall_s),
                "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_iterations": float(self.iterations.mean()),
                "blocks_used": int(self.batch.max()) + 1 if self.batch.size else 0,
                "wall_s": self.wall_s}


class Server:
    def __init__(self, graph: Graph, damping=0.85, tol=1e-7, max_iters=200, topk=100,
                 iters_per_pass=2, max_batches=4, hbm_reserve_bytes=8 << 30):
        self.graph, self.ctx, self.lib = graph, graph.ctx, graph.lib
        o = L.ServerOpts()
        o.damping, o.tol, o.max_iters, o.topk = damping, tol, max_iters, topk
        o.iters_per_pass, o.max_batches, o.hbm_reserve_bytes = iters_per_pass, max_batches, hbm_reserve_bytes
        self.topk = topk
        h = C.c_void_p()
        check(self.lib.pe_server_create(graph.h, C.byref(o), C.byref(h)), self.ctx.h)
        self.h = h
        self.ctx._adopt(self)
        graph._depend(self)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.pe_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_trace(self, arrivals, queries) -> TraceResult:
        n = len(queries)
        ids = np.arange(n, dtype=np.uint64)
        arr = np.ascontiguousarray(arrivals, np.float64)
        offs = np.zeros(n + 1, np.uint64)
        offs[1:] = np.cumsum([len(q) for q in queries])
        seeds = np.ascontiguousarray(np.concatenate([np.asarray(q, np.uint32) for q in queries]))
        stats = (L.RequestStats * n)()
        tid = np.zeros((n, self.topk), np.uint32)
        tval = np.zeros((n, self.topk), np.float32)
        wall = C.c_double()
        check(self.lib.pe_server_run_trace(self.h, n, ids.ctypes.data, arr.ctypes.data,
                                           offs.ctypes.data, seeds.ctypes.data, stats,
                                           tid.ctypes.data, tval.ctypes.data, C.byref(wall)),
              self.ctx.h)
        f = lambda name, dt: np.array([getattr(s, name) for s in stats], dt)  # noqa: E731
        return TraceResult(f("arrival_s", np.float64), f("admit_s", np.float64),
                           f("done_s", np.float64), f("iterations", np.uint32),
                           f("batch", np.uint32), f("slot", np.uint32), tid, tval, wall.value)


def poisson_trace(rng, n, qps, candidates, max_seeds=256):
    """Poisson arrivals at `qps`; each query has |S| = floor(exp(U(0, ln 257))) in [1, 256]
    seeds drawn uniformly from `candidates`."""
    gaps = rng.exponential(1.0 / qps, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    queries = []
    for _ in range(n):
        k = int(np.clip(np.floor(np.exp(rng.uniform(0, np.log(max_seeds + 1)))), 1, max_seeds))
        queries.append(rng.choice(candidates, k).astype(np.uint32))
    return arrivals, queries
