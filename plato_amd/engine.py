"""Python host API over the C-ABI: context, graph, engine state and the
dual-mode engine operators.

Names follow the reference's C++ API (plato/engine/dualmode.hpp,
plato/graph/state/dense_state.hpp, plato/util/bitmap.hpp) so tests read like
the reference's own; every call goes to libplato_hip.so.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Context:
    """One GPU, one stream, an allocation ledger (cluster_info_t for one device)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.pe_ctx_create(device, C.byref(h))
        if rc != L.PE_OK:
            raise L.PlatoHipError(rc, "pe_ctx_create failed (no GPU?)")
        self.h = h
        # device objects created on this context; destroyed before the context
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if getattr(self, "h", None):
            for child in list(self._children):
                child.close()
            self.lib.pe_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(self.lib.pe_ctx_sync(self.h), self.h)

    def mem_info(self):
        f, t, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(self.lib.pe_ctx_mem_info(self.h, C.byref(f), C.byref(t), C.byref(c)), self.h)
        return dict(free=f.value, total=t.value, ctx=c.value)

    def num_cus(self):
        n = C.c_int()
        check(self.lib.pe_ctx_num_cus(self.h, C.byref(n)), self.h)
        return n.value

    def stream_rates(self, nbytes: int = 1 << 31, reps: int = 5):
        """Read-only and copy streaming rates of the device (GB/s) over nbytes buffers."""
        r, c = C.c_double(), C.c_double()
        check(self.lib.pe_ctx_stream_rates(self.h, int(nbytes), int(reps), C.byref(r), C.byref(c)), self.h)
        return dict(read_gbs=r.value, copy_gbs=c.value, bytes=int(nbytes))

    def set_kernel_timing(self, on: bool):
        check(self.lib.pe_ctx_set_kernel_timing(self.h, int(on)), self.h)

    def kernel_time(self, which: int):
        ms, n = C.c_double(), C.c_uint64()
        check(self.lib.pe_ctx_kernel_time(self.h, which, C.byref(ms), C.byref(n)), self.h)
        return ms.value, n.value

    def timer_start(self):
        check(self.lib.pe_ctx_timer_start(self.h), self.h)

    def timer_stop(self) -> float:
        ms = C.c_double()
        check(self.lib.pe_ctx_timer_stop(self.h, C.byref(ms)), self.h)
        return ms.value

    def rmat_edges(self, scale, seed, e0, n, a=0.57, b=0.19, c=0.19, weights=False):
        src = np.zeros(n, np.uint32)
        dst = np.zeros(n, np.uint32)
        w = np.zeros(n, np.float32) if weights else None
        check(self.lib.pe_rmat_edges(self.h, scale, seed, a, b, c, e0, n, _ptr(src), _ptr(dst),
                                     _ptr(w)), self.h)
        return (src, dst, w) if weights else (src, dst)


@dataclass
class graph_info_t:
    """plato/graph/base.hpp:241-257"""
    is_directed_: bool
    vertices_: int
    edges_: int
    max_v_i_: int


class Graph:
    """Device CSC (+ optional CSR) of one graph (dcsc_t / bcsr_t for one GPU)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = handle
        self.info = L.GraphInfo()
        check(self.lib.pe_graph_get_info(self.h, C.byref(self.info)), ctx.h)
        ctx._adopt(self)
        # objects built on this graph (BFS states, PPR blocks, servers): closed before it
        self._dependents = weakref.WeakSet()

    def _depend(self, obj):
        self._dependents.add(obj)

    @classmethod
    def from_edges(cls, ctx, src, dst, V, weights=None, directed=True, build_csr=False,
                   relabel=False):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        flags = 0 if directed else L.PE_GRAPH_SYMMETRIZE
        if build_csr:
            flags |= L.PE_GRAPH_BUILD_CSR
        if relabel:
            flags |= L.PE_GRAPH_RELABEL_DEGREE
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, np.float32)
            flags |= L.PE_GRAPH_WEIGHTED
        h = C.c_void_p()
        check(ctx.lib.pe_graph_create(ctx.h, _ptr(src), _ptr(dst), _ptr(w), src.size, V, flags,
                                      C.byref(h)), ctx.h)
        return cls(ctx, h)

    @classmethod
    def from_csv_bytes(cls, ctx, data: bytes, weighted=False, directed=True, build_csr=False,
                       relabel=False, encode_ids=False):
        """Edges parsed on the device from CSV text (`src,dst[,w]` lines, the reference's
        csv_parser rules); encode_ids compacts the ids that occur to 0..n-1
        (vertex_ids() maps them back)."""
        flags = 0 if directed else L.PE_GRAPH_SYMMETRIZE
        if build_csr:
            flags |= L.PE_GRAPH_BUILD_CSR
        if relabel:
            flags |= L.PE_GRAPH_RELABEL_DEGREE
        if weighted:
            flags |= L.PE_GRAPH_WEIGHTED
        if encode_ids:
            flags |= L.PE_GRAPH_ENCODE_IDS
        h = C.c_void_p()
        check(ctx.lib.pe_graph_create_csv(ctx.h, data, len(data), flags, C.byref(h)), ctx.h)
        return cls(ctx, h)

    @classmethod
    def from_csv(cls, ctx, path, **kw):
        with open(path, "rb") as f:
            return cls.from_csv_bytes(ctx, f.read(), **kw)

    def reverse(self, build_csr=False, relabel=False, same_ids=False):
        """A new graph with every edge reversed (pe_graph_create_reverse), same input ids.
        same_ids keeps this graph's internal ids too, so its device states, bitmaps and
        BFS states stay valid for the reverse (excludes relabel)."""
        flags = (L.PE_GRAPH_BUILD_CSR if build_csr else 0) | (L.PE_GRAPH_RELABEL_DEGREE if relabel else 0) | \
            (L.PE_GRAPH_SAME_IDS if same_ids else 0)
        h = C.c_void_p()
        check(self.lib.pe_graph_create_reverse(self.h, flags, C.byref(h)), self.ctx.h)
        return Graph(self.ctx, h)

    def vertex_ids(self):
        """Original id of every (dense) vertex id: identity unless the ids were encoded."""
        out = np.zeros(self.V, np.uint64)
        check(self.lib.pe_graph_vertex_ids(self.h, out.ctypes.data), self.ctx.h)
        return out

    @classmethod
    def rmat(cls, ctx, scale, edge_factor=16, seed=1, a=0.57, b=0.19, c=0.19, directed=True,
             build_csr=False, weighted=False, relabel=False):
        flags = 0 if directed else L.PE_GRAPH_SYMMETRIZE
        if relabel:
            flags |= L.PE_GRAPH_RELABEL_DEGREE
        if build_csr:
            flags |= L.PE_GRAPH_BUILD_CSR
        if weighted:
            flags |= L.PE_GRAPH_WEIGHTED
        h = C.c_void_p()
        check(ctx.lib.pe_graph_create_rmat(ctx.h, scale, edge_factor, seed, a, b, c, flags,
                                           C.byref(h)), ctx.h)
        return cls(ctx, h)

    def close(self):
        if getattr(self, "h", None):
            for dep in list(getattr(self, "_dependents", ())):
                dep.close()
            if self.ctx.h:
                self.lib.pe_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def V(self):
        return self.info.V

    @property
    def nnz(self):
        return self.info.nnz

    @property
    def edges_prime(self):
        """E' of dualmode.hpp:261-262"""
        return self.info.edges if self.info.is_directed else 2 * self.info.edges

    def graph_info(self) -> graph_info_t:
        return graph_info_t(bool(self.info.is_directed), self.info.vertices, self.info.edges,
                            self.info.max_v_i)

    def permutation(self):
        """rank_of[input id] = internal id."""
        r = np.zeros(self.V, np.uint32)
        check(self.lib.pe_graph_permutation(self.h, r.ctypes.data), self.ctx.h)
        return r

    def set_tile(self, items):
        check(self.lib.pe_graph_set_tile(self.h, items), self.ctx.h)
        check(self.lib.pe_graph_get_info(self.h, C.byref(self.info)), self.ctx.h)

    def set_pull_layout(self, layout, src_bits=0, dst_bits=0, run_cap=0, unit_log2=0):
        """Pull layout of the FP32 rank iterations: "merge_path" (CSC gather) or
        "binned" (partition-centric, exact sums; builds a bin-ordered copy of the
        edges).  Zero options keep the library defaults."""
        code = {"merge_path": L.PE_PULL_MERGE_PATH, "binned": L.PE_PULL_BINNED}[layout]
        o = L.BinnedOpts(src_bits, dst_bits, run_cap, unit_log2)
        check(self.lib.pe_graph_set_pull_layout(self.h, code, C.byref(o)), self.ctx.h)
        check(self.lib.pe_graph_get_info(self.h, C.byref(self.info)), self.ctx.h)

    def autotune(self, pull=True, batch=False, bfs=False, binned=False):
        """Time the engine's tunables on this graph (pe_autotune); the pull tile choice
        and the best binned layout (binned=True) are applied to the graph.  Returns the
        chosen values."""
        what = (L.PE_TUNE_PULL if pull else 0) | (L.PE_TUNE_BATCH if batch else 0) | \
            (L.PE_TUNE_BFS if bfs else 0) | (L.PE_TUNE_BINNED if binned else 0)
        r = L.TuneResult()
        check(self.lib.pe_autotune(self.h, what, C.byref(r)), self.ctx.h)
        check(self.lib.pe_graph_get_info(self.h, C.byref(self.info)), self.ctx.h)
        out = {}
        for f, _ in L.TuneResult._fields_:
            v = getattr(r, f)
            out[f] = list(v) if isinstance(v, C.Array) else v
        return out

    def download_csc(self, weights=False):
        cp = np.zeros(self.V + 1, np.uint64)
        ri = np.zeros(max(self.nnz, 1), np.uint32)
        wn = np.zeros(max(self.nnz, 1), np.float32) if weights else None
        check(self.lib.pe_graph_download_csc(self.h, _ptr(cp), _ptr(ri), _ptr(wn)), self.ctx.h)
        ri = ri[: self.nnz]
        return (cp, ri, wn[: self.nnz]) if weights else (cp, ri)

    def download_csr(self):
        rp = np.zeros(self.V + 1, np.uint64)
        ci = np.zeros(max(self.nnz, 1), np.uint32)
        check(self.lib.pe_graph_download_csr(self.h, _ptr(rp), _ptr(ci)), self.ctx.h)
        return rp, ci[: self.nnz]

    def out_degrees(self):
        od = np.zeros(self.V, np.uint32)
        check(self.lib.pe_graph_out_degrees(self.h, _ptr(od)), self.ctx.h)
        return od

    def in_degrees(self):
        d = np.zeros(self.V, np.uint32)
        check(self.lib.pe_graph_in_degrees(self.h, _ptr(d)), self.ctx.h)
        return d


class DenseState:
    """dense_state_t on the device (plato/graph/state/dense_state.hpp:43-175)."""

    _DT = {np.float32: L.PE_F32, np.float64: L.PE_F64, np.uint32: L.PE_U32}

    def __init__(self, ctx: Context, n: int, dtype=np.float32):
        self.ctx, self.lib, self.n = ctx, ctx.lib, int(n)
        self.dtype = np.dtype(dtype).type
        h = C.c_void_p()
        check(self.lib.pe_vec_create(ctx.h, self._DT[self.dtype], self.n, C.byref(h)), ctx.h)
        self.h = h
        ctx._adopt(self)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.pe_vec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fill(self, value):
        check(self.lib.pe_vec_fill(self.h, float(value)), self.ctx.h)

    def copy_from(self, other: "DenseState"):
        check(self.lib.pe_vec_copy(self.h, other.h), self.ctx.h)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        check(self.lib.pe_vec_upload(self.h, _ptr(arr), arr.size), self.ctx.h)

    def numpy(self):
        out = np.zeros(self.n, self.dtype)
        check(self.lib.pe_vec_download(self.h, _ptr(out), self.n), self.ctx.h)
        return out

    def to_f64(self):
        out = np.zeros(self.n, np.float64)
        check(self.lib.pe_vec_download_f64(self.h, _ptr(out), self.n), self.ctx.h)
        return out

    def __setitem__(self, i, value):
        check(self.lib.pe_vec_set(self.h, int(i), float(value)), self.ctx.h)

    def sum(self):
        s = C.c_double()
        check(self.lib.pe_vec_sum(self.h, C.byref(s)), self.ctx.h)
        return s.value


class Bitmap:
    """bitmap_t on the device (plato/util/bitmap.hpp:43-350), same word layout."""

    def __init__(self, ctx: Context, size: int):
        self.ctx, self.lib, self.size = ctx, ctx.lib, int(size)
        self.words = (self.size >> 6) + 1
        h = C.c_void_p()
        check(self.lib.pe_bitmap_create(ctx.h, self.size, C.byref(h)), ctx.h)
        self.h = h
        ctx._adopt(self)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.pe_bitmap_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        check(self.lib.pe_bitmap_clear(self.h), self.ctx.h)

    def fill(self):
        check(self.lib.pe_bitmap_fill(self.h), self.ctx.h)

    def set_bit(self, i):
        check(self.lib.pe_bitmap_set_bit(self.h, int(i)), self.ctx.h)

    def clr_bit(self, i):
        check(self.lib.pe_bitmap_clr_bit(self.h, int(i)), self.ctx.h)

    def get_bit(self, i):
        o = C.c_int()
        check(self.lib.pe_bitmap_get_bit(self.h, int(i), C.byref(o)), self.ctx.h)
        return o.value

    def count(self):
        o = C.c_uint64()
        check(self.lib.pe_bitmap_count(self.h, C.byref(o)), self.ctx.h)
        return o.value

    def msb(self):
        o = C.c_uint64()
        check(self.lib.pe_bitmap_msb(self.h, C.byref(o)), self.ctx.h)
        return o.value

    def copy_from(self, other):
        check(self.lib.pe_bitmap_copy_from(self.h, other.h), self.ctx.h)

    def words_numpy(self):
        w = np.zeros(self.words, np.uint64)
        check(self.lib.pe_bitmap_download(self.h, _ptr(w), self.words), self.ctx.h)
        return w

    def upload_words(self, w):
        w = np.ascontiguousarray(w, np.uint64)
        check(self.lib.pe_bitmap_upload(self.h, _ptr(w), w.size), self.ctx.h)

    def to_vector(self):
        n = C.c_uint64()
        check(self.lib.pe_bitmap_to_vector(self.h, None, 0, C.byref(n)), self.ctx.h)
        ids = np.zeros(max(n.value, 1), np.uint32)
        check(self.lib.pe_bitmap_to_vector(self.h, _ptr(ids), n.value, C.byref(n)), self.ctx.h)
        return ids[: n.value]


def spread_message(actives: Bitmap, sink: DenseState, key: DenseState | None = None,
                   value: DenseState | None = None) -> int:
    """spread_message (plato/graph/message_passing.hpp:152-249) with a device sink: each
    vertex set in `actives` sends one message to key[v] (or v); a uint32 sink counts the
    messages per target, a float sink adds value[v] (or 1).  Returns the number of
    messages (the sum of the sink's returns)."""
    op = L.PE_SPREAD_COUNT if sink.dtype == np.uint32 else L.PE_SPREAD_ADD
    n = C.c_uint64()
    check(actives.lib.pe_spread_message(actives.h, op, key.h if key is not None else None,
                                        value.h if value is not None else None, sink.h, C.byref(n)),
          actives.ctx.h)
    return n.value


class dualmode_engine_t:
    """The engine of plato/engine/dualmode.hpp:75-166 over one device graph.

    The reference takes host lambdas as signal/slot functors; here the operator
    set is closed and chosen by method: pull gather-sum (aggregate_message with
    the write_add slot), restart-apply, fused iterate, and the push/pull choice.
    """

    def __init__(self, graph: Graph, push_threshold: float = 0.05):
        self.graph = graph
        self.ctx = graph.ctx
        self.lib = graph.lib
        self.push_threshold_ = push_threshold  # dualmode_engine_opts_t (dualmode.hpp:41-43)
        self.reversed_ = False
        self._rev = None  # the directed graph's reverse over the same ids, built on first use

    def reverse(self):
        """Walk the edges backwards from now on (dualmode.hpp:151-153): pulls sum over
        out-edges and pushes follow in-edges (dualmode.hpp:299-322).  The push/pull choice
        keeps the graph's own out-degrees, as the reference's does (dualmode.hpp:250-278).
        A symmetrized graph is its own reverse; a directed one is reversed on the device
        once, keeping its ids, on the first reversed call."""
        self.reversed_ = not self.reversed_

    def is_reversed(self):
        return self.reversed_

    def walk_graph(self) -> Graph:
        """The graph whose in-edges a pull and whose out-edges a push walk."""
        if not self.reversed_ or not self.graph.info.is_directed:
            return self.graph
        if self._rev is None:
            self._rev = self.graph.reverse(build_csr=True, same_ids=True)
        return self._rev

    def close(self):
        if self._rev is not None:
            self._rev.close()
            self._rev = None

    def alloc_v_state(self, dtype=np.float32) -> DenseState:
        return DenseState(self.ctx, self.graph.V, dtype)

    def alloc_v_subset(self) -> Bitmap:
        return Bitmap(self.ctx, self.graph.V)

    def active_edges(self, actives: Bitmap) -> int:
        o = C.c_uint64()
        check(self.lib.pe_active_edges(self.graph.h, actives.h, C.byref(o)), self.ctx.h)
        return o.value

    def choose_push(self, actives: Bitmap):
        p, a = C.c_int(), C.c_uint64()
        check(self.lib.pe_choose_mode(self.graph.h, actives.h, self.push_threshold_, C.byref(p),
                                      C.byref(a)), self.ctx.h)
        return bool(p.value), a.value

    def pull_sum(self, x: DenseState, y: DenseState, weighted=False):
        check(self.lib.pe_pull_sum(self.walk_graph().h, x.h, y.h, int(weighted)), self.ctx.h)

    @staticmethod
    def _params(damping, base, seed_mass, seeds, divide_by_od, final_iter):
        p = L.ApplyParams()
        p.damping, p.base, p.seed_mass = damping, base, seed_mass
        p.seeds = seeds.h if seeds is not None else None
        p.divide_by_od, p.final_iter = int(divide_by_od), int(final_iter)
        return p

    def rank_apply(self, x, y, damping, base, seed_mass=0.0, seeds=None, divide_by_od=True,
                   final_iter=False):
        p = self._params(damping, base, seed_mass, seeds, divide_by_od, final_iter)
        d = C.c_double()
        check(self.lib.pe_rank_apply(self.walk_graph().h, x.h, y.h, C.byref(p), C.byref(d)), self.ctx.h)
        return d.value

    def rank_iterate(self, x, y, damping, base, seed_mass=0.0, seeds=None, divide_by_od=True,
                     final_iter=False, weighted=False):
        p = self._params(damping, base, seed_mass, seeds, divide_by_od, final_iter)
        d = C.c_double()
        check(self.lib.pe_rank_iterate(self.walk_graph().h, x.h, y.h, int(weighted), C.byref(p),
                                       C.byref(d)), self.ctx.h)
        return d.value


    def bfs_step(self, state: "BfsEngineState", push: bool, current: Bitmap, nxt: Bitmap):
        """One BFS epoch over the edges this engine walks (reversed after reverse())."""
        state.bind(self.walk_graph())
        return state.step(push, current, nxt)

    def foreach_edges_bfs(self, state: "BfsEngineState", current: Bitmap, nxt: Bitmap):
        """The 5-argument foreach_edges with the BFS operators (bfs.hpp:95-137): push or
        pull by the active-edge rule, then one epoch.  Returns (push, activated)."""
        push, _ = self.choose_push(current)
        activated, _ = self.bfs_step(state, push, current, nxt)
        return push, activated


class BfsEngineState:
    """Device BFS state for engine-level stepping (pe_bfs_*): one epoch of the
    reference's 5-argument foreach_edges with the BFS operators per step()."""

    def __init__(self, graph: Graph):
        self.graph, self.ctx, self.lib = graph, graph.ctx, graph.lib
        h = C.c_void_p()
        check(self.lib.pe_bfs_create(graph.h, C.byref(h)), self.ctx.h)
        self.h = h
        self.ctx._adopt(self)
        graph._depend(self)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.pe_bfs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, root, frontier: Bitmap):
        check(self.lib.pe_bfs_reset(self.h, int(root), frontier.h), self.ctx.h)

    def bind(self, graph: Graph):
        """Step over `graph`'s edges from now on: this search's graph or its reverse made
        with same_ids (pe_bfs_bind)."""
        check(self.lib.pe_bfs_bind(self.h, graph.h), self.ctx.h)

    def step(self, push: bool, current: Bitmap, nxt: Bitmap):
        a, n = C.c_uint64(), C.c_uint64()
        check(self.lib.pe_bfs_step(self.h, int(push), current.h, nxt.h, C.byref(a), C.byref(n)),
              self.ctx.h)
        return a.value, n.value

    def levels(self):
        lv = np.zeros(self.graph.V, np.int32)
        check(self.lib.pe_bfs_get_levels(self.h, lv.ctypes.data), self.ctx.h)
        return lv
