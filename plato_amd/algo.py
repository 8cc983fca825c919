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
class trustrank_opts_t:
    """plato/algo/trustrank/trustrank.hpp:23-33 (random_seed_ replaces std::random_device)"""
    iteration_: int = 100
    damping_: float = 0.85
    seed_num_: int = 100
    select_method_: int = 0   # 0 inverse PageRank, 1 PageRank, 2 random
    eps_: float = 0.001
    random_seed_: int = 0


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
    alg_bytes: int = 0  # per level: push 8|F| + 4 sum_F deg + 4|new|, pull 8|U| + 4 examined + V/8


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
                     st.reached_edges, st.seconds, st.alg_bytes)


def _tr_opts(opts, precision):
    o = L.TrustRankOpts()
    o.iterations, o.damping, o.seed_num = opts.iteration_, opts.damping_, opts.seed_num_
    o.select_method, o.eps, o.random_seed, o.precision = (opts.select_method_, opts.eps_,
                                                          opts.random_seed_, precision)
    return o


def _good_words(graph, good):
    if good is None:
        return None
    g = np.zeros(graph.V, bool)
    g[np.asarray(good, np.int64)] = True
    words = np.zeros((graph.V >> 6) + 1, np.uint64)
    idx = np.flatnonzero(g)
    np.bitwise_or.at(words, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    return words


def _words_to_ids(words, V):
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:V]
    return np.flatnonzero(bits).astype(np.uint32)


def trustrank_seeds(graph: Graph, opts: trustrank_opts_t = trustrank_opts_t(), good=None,
                    reverse: Graph = None) -> np.ndarray:
    """select_good_seeds (trustrank.hpp:80-145): the chosen seed ids, ascending."""
    o = _tr_opts(opts, 64)
    gw = _good_words(graph, good)
    words = np.zeros((graph.V >> 6) + 1, np.uint64)
    n = C.c_uint32()
    check(graph.lib.pe_trustrank_select(graph.h, reverse.h if reverse else None,
                                        gw.ctypes.data if gw is not None else None, C.byref(o),
                                        words.ctypes.data, C.byref(n)), graph.ctx.h)
    return _words_to_ids(words, graph.V)


def trustrank(graph: Graph, opts: trustrank_opts_t = trustrank_opts_t(), good=None,
              reverse: Graph = None, precision: int = 64):
    """plato::algo::trustrank (trustrank.hpp:163-259).  Returns (RankResult, seed ids).
    good: ids of the good vertices (None: all); reverse: the reversed graph, needed by
    inverse PageRank (select_method_ 0) on a directed graph."""
    o = _tr_opts(opts, precision)
    gw = _good_words(graph, good)
    words = np.zeros((graph.V >> 6) + 1, np.uint64)
    out = np.zeros(graph.V, np.float64)
    st = L.RankStats()
    check(graph.lib.pe_trustrank(graph.h, reverse.h if reverse else None,
                                 gw.ctypes.data if gw is not None else None, C.byref(o),
                                 out.ctypes.data, words.ctypes.data, C.byref(st)), graph.ctx.h)
    res = RankResult(out, np.zeros(0), st.iterations_run, st.seconds, st.gteps)
    return res, _words_to_ids(words, graph.V)


@dataclass
class KcoreResult:
    coreness: np.ndarray      # uint32 per vertex (input ids)
    modified: np.ndarray      # per epoch: vertices whose coreness dropped
    next_active: np.ndarray   # per epoch: size of the next active set
    epochs: int
    seconds: float


def kcore(graph: Graph, out_edges: bool = False, max_epochs: int = 1 << 16) -> KcoreResult:
    """kcore_algo_t::compute_shell_index (kcore.hpp:80-268) over the in-edges, or the
    out-edges (a BCSR passed as the reference's "incomings")."""
    core = np.zeros(graph.V, np.uint32)
    mod = np.zeros(max_epochs, np.uint64)
    nxt = np.zeros(max_epochs, np.uint64)
    st = L.KcoreStats()
    check(graph.lib.pe_kcore(graph.h, 1 if out_edges else 0, core.ctypes.data, mod.ctypes.data,
                             nxt.ctypes.data, max_epochs, C.byref(st)), graph.ctx.h)
    k = min(st.epochs, max_epochs)
    return KcoreResult(core, mod[:k], nxt[:k], st.epochs, st.seconds)


@dataclass
class CcResult:
    labels: np.ndarray     # uint32 per vertex: smallest input id that reaches it
    modes: np.ndarray      # per epoch: 1 push, 0 pull
    actives: np.ndarray    # per epoch: sum of the slots' returns
    epochs: int
    components: int
    seconds: float


def connected_components(graph: Graph, max_epochs: int = 1 << 16) -> CcResult:
    """connected_component_t::compute (cgm/connected_component.hpp:170-234)."""
    labels = np.zeros(graph.V, np.uint32)
    modes = np.zeros(max_epochs, np.uint8)
    act = np.zeros(max_epochs, np.uint64)
    st = L.CcStats()
    check(graph.lib.pe_connected_components(graph.h, labels.ctypes.data, modes.ctypes.data, act.ctypes.data,
                                            max_epochs, C.byref(st)), graph.ctx.h)
    k = min(st.epochs, max_epochs)
    return CcResult(labels, modes[:k], act[:k], st.epochs, st.components, st.seconds)


@dataclass
class hyperanf_opts_t:
    """plato/algo/hyperanf/hyperanf.hpp:39-42 (bits: the BitWidth template argument)."""
    iteration_: int = 20
    bits: int = 12


@dataclass
class HyperAnfResult:
    avg_distance: float
    anfs: np.ndarray
    registers: np.ndarray | None  # [V][2^bits] uint8 in input ids, when asked for
    seconds: float


def hyperanf(graph: Graph, opts: hyperanf_opts_t = hyperanf_opts_t(), registers: bool = False,
             max_anfs: int = 4096) -> HyperAnfResult:
    """plato::algo::hyperanf (hyperanf.hpp:80-197): the average distance estimate."""
    o = L.HyperAnfOpts(opts.iteration_, opts.bits)
    anfs = np.zeros(max_anfs, np.float64)
    n = C.c_uint32()
    avg = C.c_double()
    sec = C.c_double()
    regs = np.zeros((graph.V, 1 << opts.bits), np.uint8) if registers else None
    check(graph.lib.pe_hyperanf(graph.h, C.byref(o), anfs.ctypes.data, max_anfs, C.byref(n), C.byref(avg),
                                regs.ctypes.data if registers else None, C.byref(sec)), graph.ctx.h)
    return HyperAnfResult(avg.value, anfs[:min(n.value, max_anfs)], regs, sec.value)


@dataclass
class nstepdegree_opts_t:
    """plato/algo/nstepdegrees/nstepdegrees.hpp:49-53 (+ the HyperLogLog width, a
    template argument there)"""
    step: int = 20
    type: str = "both"
    is_directed: bool = False
    bits: int = 12


@dataclass
class NstepResult:
    in_degrees: np.ndarray | None
    out_degrees: np.ndarray | None
    seconds: float


def nstep_degrees(graph: Graph, opts: nstepdegree_opts_t = nstepdegree_opts_t(), targets=None) -> NstepResult:
    """nstepdegrees_t::compute (plato/algo/nstepdegrees/nstepdegrees.hpp:437-466): n-step
    in/out degrees of the target vertices (all when targets is None), 0 elsewhere."""
    t = {"in": L.PE_NSTEP_IN, "out": L.PE_NSTEP_OUT, "both": L.PE_NSTEP_BOTH}[opts.type]
    V = graph.V
    din = np.zeros(V, np.uint32) if t & L.PE_NSTEP_IN else None
    dout = np.zeros(V, np.uint32) if t & L.PE_NSTEP_OUT else None
    words = None
    if targets is not None:
        ids = np.asarray(targets, np.uint64)
        words = np.zeros((V >> 6) + 1, np.uint64)
        np.bitwise_or.at(words, (ids >> np.uint64(6)).astype(np.int64), np.left_shift(np.uint64(1), ids & np.uint64(63)))
    secs = C.c_double()
    check(graph.lib.pe_nstep_degrees(graph.h, None if words is None else words.ctypes.data,
                                     int(opts.step) & 0xFFFFFFFF, int(opts.bits), t,
                                     None if din is None else din.ctypes.data,
                                     None if dout is None else dout.ctypes.data, C.byref(secs)), graph.ctx.h)
    return NstepResult(din, dout, secs.value)
