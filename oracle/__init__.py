"""CPU oracle for the dual-mode edge engine: ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY.  tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use it as the checker; the product package
``plato_amd`` never imports it.

Each wrapper returns numpy arrays; see oracle/oracle.h for the reference
file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile oracle/liboracle.so (gcc, OpenMP)."""
    root = os.path.dirname(_HERE)
    subprocess.run(["make", "-s", "-f", "oracle/Makefile", "oracle/liboracle.so"],
                   cwd=root, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.orc_mix64.argtypes = [C.c_uint64]
    L.orc_mix64.restype = C.c_uint64
    L.orc_rmat_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, _u32p]
    L.orc_rmat_edges.argtypes = [C.c_uint32, C.c_uint64, _u32p, C.c_uint64, C.c_uint64, _u32p, _u32p]
    L.orc_rmat_weight.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_rmat_weight.restype = C.c_float
    L.orc_build_csc.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_void_p, C.c_int,
                                _u64p, _u32p, C.c_void_p]
    L.orc_build_csr.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_int, _u64p, _u32p]
    L.orc_out_degrees.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p]
    L.orc_pagerank.argtypes = [C.c_uint32, _u64p, _u32p, _u32p, C.c_uint32, C.c_double,
                               C.c_double, _f64p, _f64p, C.c_int]
    L.orc_pagerank_step.argtypes = [C.c_uint32, _u64p, _u32p, _u32p, C.c_double, C.c_int,
                                    _f64p, _f64p, C.c_uint32, C.c_uint32, C.c_int]
    L.orc_pagerank_step.restype = C.c_double
    L.orc_weighted_pagerank.argtypes = [C.c_uint32, _u64p, _u32p, _f32p, C.c_uint32,
                                        C.c_double, C.c_double, _f64p, _f64p]
    L.orc_ppr.argtypes = [C.c_uint32, _u64p, _u32p, _u32p, _u32p, C.c_uint32, C.c_uint32,
                          C.c_double, C.c_double, _f64p, _f64p]
    L.orc_ppr_pushpull.argtypes = [C.c_uint32, _u64p, _u32p, _u64p, _u32p, _u32p, C.c_uint64,
                                   C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                   _f64p, _f64p, _u8p]
    L.orc_bfs.argtypes = [C.c_uint32, _u64p, _u32p, _u64p, _u32p, _u32p, C.c_uint64,
                          C.c_double, C.c_uint32, _i32p, _u8p, _u64p, _u64p, C.c_uint32,
                          C.POINTER(C.c_uint32)]
    L.orc_murmur3_32.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32]
    L.orc_murmur3_32.restype = C.c_uint32
    L.orc_hll_estimate.argtypes = [_u8p, C.c_uint32]
    L.orc_hll_estimate.restype = C.c_double
    L.orc_hyperanf.argtypes = [C.c_uint32, _u64p, _u32p, C.c_uint32, C.c_uint32, _u8p, _f64p, C.c_uint32]
    L.orc_hyperanf.restype = C.c_int
    L.orc_anf_avg_distance.argtypes = [_f64p, C.c_uint32]
    L.orc_anf_avg_distance.restype = C.c_double
    L.orc_cc_labels.argtypes = [C.c_uint32, _u64p, _u32p, _u32p]
    L.orc_cc_labels.restype = None
    L.orc_kcore.argtypes = [C.c_uint32, _u64p, _u32p, _u32p, _u64p, _u64p, C.c_uint32]
    L.orc_kcore.restype = C.c_int
    L.orc_parse_csv.argtypes = [C.c_char_p, C.c_uint64, C.c_int, C.c_uint64, _u32p, _u32p, C.c_void_p]
    L.orc_parse_csv.restype = C.c_int64
    L.orc_format_csv.argtypes = [C.c_uint64, _u32p, _u32p, C.c_void_p, C.c_uint64]
    L.orc_format_csv.restype = C.c_int64
    L.orc_topk.argtypes = [C.c_uint32, _f64p, C.c_uint32, _u32p, _f64p]
    L.orc_nstep_degrees.argtypes = [C.c_uint32, _u64p, _u32p, C.c_uint32, C.c_uint32, _u32p]
    L.orc_nstep_degrees.restype = C.c_int
    L.orc_rmat_weights.argtypes = [C.c_uint64, C.c_uint64, _u64p, _f32p]
    L.orc_rmat_edges_at.argtypes = [C.c_uint32, C.c_uint64, _u32p, C.c_uint64, _u64p, _u32p, _u32p]
    L.orc_rmat_scan.argtypes = [C.c_uint32, C.c_uint64, _u32p, C.c_uint64, C.c_uint32, _u32p, _u32p,
                                _u32p, C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64, _u64p]
    L.orc_rmat_scan.restype = C.c_int64
    L.orc_edges_digest.argtypes = [C.c_uint64, _u32p, _u32p]
    L.orc_edges_digest.restype = C.c_uint64
    L.orc_csc_digest.argtypes = [C.c_uint32, _u64p, _u32p, C.c_void_p]
    L.orc_csc_digest.restype = C.c_uint64
    L.orc_weighted_pagerank_step.argtypes = [C.c_uint32, _u64p, _u32p, _f32p, _f64p, C.c_double,
                                             _f64p, _f64p, C.c_uint32, C.c_uint32]
    L.orc_weighted_pagerank_step.restype = C.c_double
    _lib = L
    return L


# ---- synthetic graph --------------------------------------------------------

def mix64(x: int) -> int:
    return int(lib().orc_mix64(x))


def rmat_thresholds(a=0.57, b=0.19, c=0.19) -> np.ndarray:
    out = np.zeros(3, np.uint32)
    lib().orc_rmat_thresholds(a, b, c, out)
    return out


def rmat_edges(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19, c=0.19,
               e0: int = 0, n: int | None = None):
    E = edge_factor << scale
    if n is None:
        n = E - e0
    src = np.zeros(n, np.uint32)
    dst = np.zeros(n, np.uint32)
    lib().orc_rmat_edges(scale, seed, rmat_thresholds(a, b, c), e0, n, src, dst)
    return src, dst


def rmat_edges_at(scale: int, seed: int, eids, a=0.57, b=0.19, c=0.19):
    """The RMAT edges with the given edge ids."""
    eids = np.ascontiguousarray(eids, np.uint64)
    src = np.zeros(eids.size, np.uint32)
    dst = np.zeros(eids.size, np.uint32)
    lib().orc_rmat_edges_at(scale, seed, rmat_thresholds(a, b, c), eids.size, eids, src, dst)
    return src, dst


def rmat_weights(seed: int, E: int = 0, eids=None) -> np.ndarray:
    """Weights of edges [0, E), or of the edge ids in `eids`."""
    if eids is None:
        eids = np.arange(E, dtype=np.uint64)
    eids = np.ascontiguousarray(eids, np.uint64)
    out = np.zeros(eids.size, np.float32)
    lib().orc_rmat_weights(seed, eids.size, eids, out)
    return out


def edges_digest(src, dst) -> int:
    """Order-independent digest of an edge multiset (orc_edges_digest)."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    return int(lib().orc_edges_digest(src.size, src, dst))


def csc_digest(V, col_ptr, row_idx, perm=None) -> int:
    """The same digest over a CSC's entries, both ends mapped through perm (internal ->
    input ids) when given."""
    cp = np.ascontiguousarray(col_ptr, np.uint64)
    ri = np.ascontiguousarray(row_idx, np.uint32)
    pm = None if perm is None else np.ascontiguousarray(perm, np.uint32)
    return int(lib().orc_csc_digest(int(V), cp, ri, None if pm is None else pm.ctypes.data))


class RmatScan:
    """One pass over an RMAT graph too large to hold twice on the host: out-degrees,
    per-source weight sums, the edge-multiset digest, and the in-edges of the columns in
    `ranges` (a list of [lo, hi) pairs) as a CSC over all V columns (empty elsewhere)."""

    def __init__(self, scale, edge_factor, seed, ranges, weighted=False, a=0.57, b=0.19, c=0.19):
        V, E = 1 << scale, edge_factor << scale
        lo = np.array([r[0] for r in ranges], np.uint32)
        hi = np.array([r[1] for r in ranges], np.uint32)
        self.V, self.E, self.ranges = V, E, [(int(x), int(y)) for x, y in ranges]
        self.od = np.zeros(V, np.uint32)
        self.W = np.zeros(V, np.float64) if weighted else None
        dg = C.c_uint64()
        thr = rmat_thresholds(a, b, c)
        cap = max(1 << 20, E // 4)
        while True:
            sel = np.zeros(cap, np.uint64)
            n = lib().orc_rmat_scan(scale, seed, thr, E, lo.size, lo, hi, self.od,
                                    None if self.W is None else self.W.ctypes.data, C.byref(dg), cap, sel)
            if n <= cap:
                break
            cap = int(n)
        self.digest = int(dg.value)
        eids = np.sort(sel[:n])  # edge order, as the structure builders keep it
        self.eids = eids
        ss, dd = rmat_edges_at(scale, seed, eids, a, b, c)
        w = rmat_weights(seed, eids=eids) if weighted else None
        self.slice = CSC(V, ss, dd, w=w, csr=False)


# ---- structures --------------------------------------------------------------

class CSC:
    """Host CSC (+ optional CSR) of an edge list, built by the oracle."""

    def __init__(self, V, src, dst, w=None, symmetrize=False, csr=True):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        self.V = int(V)
        self.E = int(src.size)                      # input edges (Plato edges_)
        self.directed = not symmetrize
        self.nnz = self.E * (2 if symmetrize else 1)
        self.col_ptr = np.zeros(self.V + 1, np.uint64)
        self.row_idx = np.zeros(max(self.nnz, 1), np.uint32)
        self.w = None
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, np.float32)
            self.w = np.zeros(max(self.nnz, 1), np.float32)
            wp = w.ctypes.data
        rc = lib().orc_build_csc(self.V, self.E, src, dst, wp, int(symmetrize), self.col_ptr,
                                 self.row_idx, self.w.ctypes.data if self.w is not None else None)
        if rc != 0:
            raise ValueError(f"orc_build_csc failed rc={rc}")
        self.row_idx = self.row_idx[: self.nnz]
        if self.w is not None:
            self.w = self.w[: self.nnz]
        self.od = np.zeros(self.V, np.uint32)
        lib().orc_out_degrees(self.V, self.nnz, np.ascontiguousarray(self.row_idx), self.od)
        self.row_ptr = self.col_idx = None
        if csr:
            self.row_ptr = np.zeros(self.V + 1, np.uint64)
            self.col_idx = np.zeros(max(self.nnz, 1), np.uint32)
            rc = lib().orc_build_csr(self.V, self.E, src, dst, int(symmetrize), self.row_ptr,
                                     self.col_idx)
            if rc != 0:
                raise ValueError(f"orc_build_csr failed rc={rc}")
            self.col_idx = self.col_idx[: self.nnz]

    @property
    def edges_prime(self) -> int:
        """E' of dualmode.hpp:261-262."""
        return self.E if self.directed else 2 * self.E

    def in_degrees(self) -> np.ndarray:
        return np.diff(self.col_ptr).astype(np.uint32)


# ---- algorithms ----------------------------------------------------------------

def pagerank(g: CSC, iters=20, damping=0.85, eps=0.0, nthreads=0):
    out = np.zeros(g.V, np.float64)
    deltas = np.zeros(max(iters, 1), np.float64)
    n = lib().orc_pagerank(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx), g.od, iters,
                           damping, eps, out, deltas, nthreads)
    if n < 0:
        raise RuntimeError("orc_pagerank failed")
    return out, deltas[:n]


def pagerank_step(g: CSC, x, damping, final_iter, v_lo=0, v_hi=None, nthreads=0, y=None):
    if v_hi is None:
        v_hi = g.V
    if y is None:
        y = np.zeros(g.V, np.float64)
    d = lib().orc_pagerank_step(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx), g.od, damping,
                                int(final_iter), x, y, v_lo, v_hi, nthreads)
    return y, d


def weighted_pagerank(g: CSC, iters=20, damping=0.85, eps=0.0):
    out = np.zeros(g.V, np.float64)
    deltas = np.zeros(max(iters, 1), np.float64)
    n = lib().orc_weighted_pagerank(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx),
                                    np.ascontiguousarray(g.w), iters, damping, eps, out, deltas)
    if n < 0:
        raise RuntimeError("orc_weighted_pagerank failed")
    return out, deltas[:n]


def weighted_pagerank_step(g: CSC, W, x, damping, v_lo=0, v_hi=None, y=None):
    """One weighted PageRank iteration over columns [v_lo, v_hi) (x: undivided ranks,
    W: FP64 out-weight sums)."""
    if v_hi is None:
        v_hi = g.V
    if y is None:
        y = np.zeros(g.V, np.float64)
    lib().orc_weighted_pagerank_step(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx),
                                     np.ascontiguousarray(g.w), np.ascontiguousarray(W, np.float64),
                                     damping, np.ascontiguousarray(x, np.float64), y, v_lo, v_hi)
    return y


def ppr(g: CSC, seeds, iters=100, damping=0.85, eps=0.0):
    seeds = np.ascontiguousarray(np.atleast_1d(seeds), np.uint32)
    out = np.zeros(g.V, np.float64)
    deltas = np.zeros(max(iters, 1), np.float64)
    n = lib().orc_ppr(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx), g.od, seeds, seeds.size,
                      iters, damping, eps, out, deltas)
    if n < 0:
        raise RuntimeError("orc_ppr failed")
    return out, deltas[:n]


def ppr_pushpull(g: CSC, src, iters=100, damping=0.85, eps=0.0):
    out = np.zeros(g.V, np.float64)
    deltas = np.zeros(max(iters, 1), np.float64)
    modes = np.zeros(max(iters, 1), np.uint8)
    n = lib().orc_ppr_pushpull(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx), g.row_ptr,
                               np.ascontiguousarray(g.col_idx), g.od, g.edges_prime, src, iters,
                               damping, eps, out, deltas, modes)
    if n < 0:
        raise RuntimeError("orc_ppr_pushpull failed")
    return out, deltas[:n], modes[:n]


def bfs(g: CSC, root, threshold=0.05, max_epochs=4096):
    levels = np.zeros(g.V, np.int32)
    modes = np.zeros(max_epochs, np.uint8)
    frontier = np.zeros(max_epochs, np.uint64)
    active = np.zeros(max_epochs, np.uint64)
    visited = C.c_uint32(0)
    n = lib().orc_bfs(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx), g.row_ptr,
                      np.ascontiguousarray(g.col_idx), g.od, g.edges_prime, threshold, root,
                      levels, modes, frontier, active, max_epochs, C.byref(visited))
    if n < 0:
        raise RuntimeError("orc_bfs failed")
    k = min(n, max_epochs)
    return dict(levels=levels, modes=modes[:k], frontier=frontier[:k], active_edges=active[:k],
                epochs=n, visited=int(visited.value))


def murmur3_32(data: bytes, seed: int) -> int:
    return int(lib().orc_murmur3_32(data, len(data), seed))


def hll_estimate(regs) -> float:
    regs = np.ascontiguousarray(regs, np.uint8)
    return float(lib().orc_hll_estimate(regs, int(regs.size).bit_length() - 1))


def hyperanf(g: CSC, bits=12, iters=20, max_anfs=4096):
    """hyperanf over the in-edges: (anfs, average distance, registers[V][2^bits])."""
    regs = np.zeros((g.V, 1 << bits), np.uint8)
    anfs = np.zeros(max_anfs, np.float64)
    n = lib().orc_hyperanf(g.V, g.col_ptr, np.ascontiguousarray(g.row_idx), bits, iters, regs, anfs, max_anfs)
    anfs = anfs[:min(n, max_anfs)]
    return anfs, float(lib().orc_anf_avg_distance(anfs, anfs.size)), regs


def nstep_degrees(g: CSC, step=20, bits=6, direction="in"):
    """n-step degrees of every vertex: direction "in" over in-edges (CSC), "out" over
    out-edges (CSR).  Returns (degrees, iterations run)."""
    out = np.zeros(g.V, np.uint32)
    ptr, idx = (g.col_ptr, g.row_idx) if direction == "in" else (g.row_ptr, g.col_idx)
    n = lib().orc_nstep_degrees(g.V, ptr, np.ascontiguousarray(idx), bits, step & 0xFFFFFFFF, out)
    if n < 0:
        raise RuntimeError("orc_nstep_degrees failed")
    return out, n


def cc_labels(g: CSC):
    """connected_component_t's fixed point: the smallest id that reaches each vertex."""
    labels = np.zeros(g.V, np.uint32)
    lib().orc_cc_labels(g.V, g.row_ptr, np.ascontiguousarray(g.col_idx), labels)
    return labels


def kcore(g: CSC, out_edges=False, max_epochs=4096):
    """kcore_algo_t::compute_shell_index over the in-edges (CSC), or the out-edges (CSR)."""
    ptr, idx = (g.row_ptr, g.col_idx) if out_edges else (g.col_ptr, g.row_idx)
    core = np.zeros(g.V, np.uint32)
    mod = np.zeros(max_epochs, np.uint64)
    nxt = np.zeros(max_epochs, np.uint64)
    n = lib().orc_kcore(g.V, ptr, np.ascontiguousarray(idx), core, mod, nxt, max_epochs)
    k = min(n, max_epochs)
    return core, mod[:k], nxt[:k]


def topk(r, k):
    r = np.ascontiguousarray(r, np.float64)
    ids = np.zeros(k, np.uint32)
    vals = np.zeros(k, np.float64)
    lib().orc_topk(r.size, r, k, ids, vals)
    kk = min(k, r.size)
    return ids[:kk], vals[:kk]


def parse_csv(data: bytes, weighted=False):
    """Plato's csv_parser semantics over a byte buffer (orc_parse_csv).  Returns
    (src, dst[, w]) in file order; raises ValueError(line) where the reference CHECK-fails."""
    n = len(data)
    cap = data.count(b"\n") + 1
    s = np.zeros(cap, np.uint32)
    d = np.zeros(cap, np.uint32)
    w = np.zeros(cap, np.float32)
    r = lib().orc_parse_csv(data, n, int(weighted), cap, s, d, w.ctypes.data if weighted else None)
    if r < 0:
        raise ValueError(f"id above 2^32 - 1 on line {-r - 1}")
    return (s[:r], d[:r], w[:r]) if weighted else (s[:r], d[:r])


def format_csv(src, dst) -> bytes:
    """`src,dst` lines for the given edges (test data for the parser)."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    cap = 22 * src.size + 1
    buf = np.empty(cap, np.uint8)
    n = lib().orc_format_csv(src.size, src, dst, buf.ctypes.data, cap)
    assert n >= 0
    return buf[:n].tobytes()


def read_edge_csv(path, weighted=False):
    """Parse Plato's `src,dst[,w]` CSV (parsers.hpp:225-293; whitespace tolerated)."""
    src, dst, w = [], [], []
    with open(path) as f:
        for line in f:
            parts = [p.strip() for p in line.strip().split(",")]
            if len(parts) < 2 or parts[0] == "":
                continue
            src.append(int(parts[0]))
            dst.append(int(parts[1]))
            if weighted:
                w.append(float(parts[2]) if len(parts) > 2 else 1.0)
    s = np.array(src, np.uint32)
    d = np.array(dst, np.uint32)
    return (s, d, np.array(w, np.float32)) if weighted else (s, d)
