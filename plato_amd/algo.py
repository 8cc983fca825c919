"""Algorithms on the device engine, named after plato::algo (plato/algo/...).

Each function runs the whole reference loop on the GPU through one C-ABI call
and returns FP64 numpy arrays, like the reference's dense_state_t<double>.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check
from .engine import Graph


@dataclass
class pagerank_opts_t:
    """plato/algo/pagerank/pagerank.hpp:39-46"""
    iteration_: int = 100
    damping_: float = 0.85
    eps_: float = 0.001


@dataclass
class personalized_pagerank_opts_t:
    """plato/algo/ppr/personalized_pagerank.hpp:18-26"""
    iteration_: int = 100
    damping_: float = 0.85
    src_: int = 0
    eps_: float = 0.001


@dataclass
class bfs_opts_t:
    """plato/algo/bfs/bfs.hpp:37-39"""
    root_: int = 0


@dataclass
class RankResult:
    rank: np.ndarray
    deltas: np.ndarray
    iterations: int
    seconds: float
    gteps: float
    modes: np.ndarray | None = None


def _opts(iters, damping, eps, precision, weighted, l1_norm_tol):
    o = L.RankOpts()
    o.iterations, o.damping, o.eps = int(iters), float(damping), float(eps)
    o.precision, o.weighted, o.l1_norm_tol = int(precision), int(weighted), float(l1_norm_tol)
    return o


def _run(graph, fn, *pre, iters, **kw):
    out = np.zeros(graph.V, np.float64)
    deltas = np.zeros(max(int(iters), 1), np.float64)
    st = L.RankStats()
    check(fn(graph.h, *pre, C.byref(kw["opts"]), out.ctypes.data, deltas.ctypes.data,
             *kw.get("post", ()), C.byref(st)), graph.ctx.h)
    n = st.iterations_run
    return RankResult(out, deltas[:n], n, st.seconds, st.gteps)


def pagerank(graph: Graph, opts: pagerank_opts_t = pagerank_opts_t(), precision: int = 32,
             l1_norm_tol: float = 0.0) -> RankResult:
    """plato::algo::pagerank (plato/algo/pagerank/pagerank.hpp:60-149)."""
    o = _opts(opts.iteration_, opts.damping_, opts.eps_, precision, False, l1_norm_tol)
    return _run(graph, graph.lib.pe_pagerank, iters=opts.iteration_, opts=o)


def weighted_pagerank(graph: Graph, opts: pagerank_opts_t = pagerank_opts_t(),
                      precision: int = 32, l1_norm_tol: float = 0.0) -> RankResult:
    """plato::algo::weighted_pagerank (plato/algo/pagerank/weighted_pagerank.hpp:60-157)."""
    o = _opts(opts.iteration_, opts.damping_, opts.eps_, precision, True, l1_norm_tol)
    return _run(graph, graph.lib.pe_pagerank, iters=opts.iteration_, opts=o)


def personalized_pagerank(graph: Graph, opts: personalized_pagerank_opts_t =
                          personalized_pagerank_opts_t(), precision: int = 32,
                          l1_norm_tol: float = 0.0) -> RankResult:
    """plato::algo::personalized_pagerank, pull form
    (plato/algo/ppr/personalized_pagerank.hpp:40-134)."""
    return multi_seed_ppr(graph, [opts.src_], opts.iteration_, opts.damping_, opts.eps_,
                          precision, l1_norm_tol)


def multi_seed_ppr(graph: Graph, seeds, iteration=100, damping=0.85, eps=0.001, precision=32,
                   l1_norm_tol=0.0) -> RankResult:
    """PPR with restart mass (1-d)/|S| on each seed, the TrustRank restart
    (plato/algo/trustrank/trustrank.hpp:197-250); one seed is plain PPR."""
    s = np.ascontiguousarray(np.atleast_1d(seeds), np.uint32)
    o = _opts(iteration, damping, eps, precision, False, l1_norm_tol)
    return _run(graph, graph.lib.pe_ppr, s.ctypes.data, s.size, iters=iteration, opts=o)


def personalized_pagerank_pushpull(graph: Graph, opts: personalized_pagerank_opts_t =
                                   personalized_pagerank_opts_t(), precision: int = 32
                                   ) -> RankResult:
    """Push/pull PPR (plato/algo/ppr/personalized_pagerank.hpp:137-275)."""
    o = _opts(opts.iteration_, opts.damping_, opts.eps_, precision, False, 0.0)
    modes = np.zeros(max(opts.iteration_, 1), np.uint8)
    res = _run(graph, graph.lib.pe_ppr_pushpull, opts.src_, iters=opts.iteration_, opts=o,
               post=(modes.ctypes.data,))
    res.modes = modes[: res.iterations]
    return res


@dataclass
class BfsResult:
    levels: np.ndarray
    modes: np.ndarray
    frontier: np.ndarray
    active_edges: np.ndarray
    epochs: int
    visited: int
    reached_edges: int
    seconds: float


def breadth_first_search(graph: Graph, opts: bfs_opts_t = bfs_opts_t(),
                         push_threshold: float = 0.05, max_epochs: int = 4096) -> BfsResult:
    """plato::algo::breadth_first_search (plato/algo/bfs/bfs.hpp:55-155) with levels."""
    levels = np.zeros(graph.V, np.int32)
    modes = np.zeros(max_epochs, np.uint8)
    frontier = np.zeros(max_epochs, np.uint64)
    active = np.zeros(max_epochs, np.uint64)
    st = L.BfsStats()
    check(graph.lib.pe_bfs(graph.h, int(opts.root_), float(push_threshold), levels.ctypes.data,
                           modes.ctypes.data, frontier.ctypes.data, active.ctypes.data,
                           max_epochs, C.byref(st)), graph.ctx.h)
    k = min(st.epochs, max_epochs)
    return BfsResult(levels, modes[:k], frontier[:k], active[:k], st.epochs, st.visited,
                     st.reached_edges, st.seconds)
