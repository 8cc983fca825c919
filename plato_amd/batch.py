"""Batched PPR (k = 64 query columns) over the C-ABI: pe_ppr_batch_*."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check
from .engine import Graph

EMPTY, LOADING, ACTIVE, FINAL, DONE = 0, 1, 2, 3, 4
K = 64


class PprBatch:
    """A resident rank block of 64 query columns on one graph."""

    def __init__(self, graph: Graph, damping=0.85, tol=1e-7, max_iters=200, task_edges=0, plan=(0, 0, 0)):
        """plan: (packed, edges4, edges2) limits of the pass plan (pe_ppr_batch_opts_t): 0 the
        kernel's widest, L.PE_BATCH_PLAN_OFF never."""
        self.graph, self.ctx, self.lib = graph, graph.ctx, graph.lib
        o = L.PprBatchOpts()
        o.k, o.damping, o.tol, o.max_iters, o.task_edges = K, damping, tol, max_iters, task_edges
        o.plan_packed, o.plan_edges4, o.plan_edges2 = plan
        h = C.c_void_p()
        check(self.lib.pe_ppr_batch_create(graph.h, C.byref(o), C.byref(h)), self.ctx.h)
        self.h = h
        self.ctx._adopt(self)
        graph._depend(self)

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.pe_ppr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        b = C.c_uint64()
        check(self.lib.pe_ppr_batch_bytes(self.h, C.byref(b)), self.ctx.h)
        return b.value

    def admit(self, slot, seeds):
        s = np.ascontiguousarray(np.atleast_1d(seeds), np.uint32)
        check(self.lib.pe_ppr_batch_admit(self.h, int(slot), s.ctypes.data, s.size), self.ctx.h)

    def release(self, slot):
        check(self.lib.pe_ppr_batch_release(self.h, int(slot)), self.ctx.h)

    def step(self, iterations=1):
        check(self.lib.pe_ppr_batch_step(self.h, int(iterations)), self.ctx.h)

    def poll(self):
        st = np.zeros(K, np.uint32)
        it = np.zeros(K, np.uint32)
        dl = np.zeros(K, np.float64)
        check(self.lib.pe_ppr_batch_poll(self.h, st.ctypes.data, it.ctypes.data, dl.ctypes.data),
              self.ctx.h)
        return st, it, dl

    def results(self, slots):
        slots = np.ascontiguousarray(np.atleast_1d(slots), np.uint32)
        out = np.zeros((slots.size, self.graph.V), np.float64)
        check(self.lib.pe_ppr_batch_results(self.h, slots.ctypes.data, slots.size, out.ctypes.data),
              self.ctx.h)
        return out

    def topk(self, slots, k=100):
        slots = np.ascontiguousarray(np.atleast_1d(slots), np.uint32)
        ids = np.zeros((slots.size, k), np.uint32)
        vals = np.zeros((slots.size, k), np.float32)
        check(self.lib.pe_ppr_batch_topk(self.h, slots.ctypes.data, slots.size, k, ids.ctypes.data,
                                         vals.ctypes.data), self.ctx.h)
        return ids, vals

    def run_until_done(self, slots, chunk=8, max_passes=100000):
        """Step until every slot in `slots` is DONE."""
        slots = list(slots)
        passes = 0
        while passes < max_passes:
            self.step(chunk)
            passes += chunk
            st, _, _ = self.poll()
            if all(st[s] == DONE for s in slots):
                return passes
        raise RuntimeError("batch did not finish")


def run_queries(graph: Graph, queries, damping=0.85, tol=1e-7, max_iters=200, chunk=4,
                want_ranks=False, topk=0):
    """Run a list of queries (each a seed list) through one 64-column batch with
    continuous refill: a finished column is read out and its slot reused at once.
    Returns per-query iteration counts (and ranks / top-k when asked)."""
    b = PprBatch(graph, damping, tol, max_iters)
    pending = list(range(len(queries)))
    slot_q = [-1] * K
    iters = np.zeros(len(queries), np.uint32)
    ranks = [None] * len(queries) if want_ranks else None
    tops = [None] * len(queries) if topk else None
    for s in range(K):
        if not pending:
            break
        q = pending.pop(0)
        b.admit(s, queries[q])
        slot_q[s] = q
    while any(q >= 0 for q in slot_q):
        b.step(chunk)
        st, it, _ = b.poll()
        done = [s for s in range(K) if slot_q[s] >= 0 and st[s] == DONE]
        if not done:
            continue
        if want_ranks:
            for s, r in zip(done, b.results(done)):
                ranks[slot_q[s]] = r
        if topk:
            ids, vals = b.topk(done, topk)
            for i, s in enumerate(done):
                tops[slot_q[s]] = (ids[i], vals[i])
        for s in done:
            iters[slot_q[s]] = it[s]
            b.release(s)
            slot_q[s] = -1
            if pending:
                q = pending.pop(0)
                b.admit(s, queries[q])
                slot_q[s] = q
    b.close()
    return iters, ranks, tops
