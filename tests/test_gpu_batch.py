"""GPU: batched PPR (64 query columns) against the FP64 oracle.

Tolerances: each column stops per the reference's rule with tol = 1e-7 on its
delta; against the FP64 oracle run to convergence (2000 iterations, eps 0) the
L1 error of every column is <= 1e-6.  A column's bits do not depend on its
slot or on the other queries in the block (checked with array_equal).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
D = 0.85


@pytest.fixture(scope="module")
def s13(ctx):
    from plato_amd import Graph
    g = Graph.rmat(ctx, 13, 16, seed=21, relabel=True)
    s, d = oracle.rmat_edges(13, 16, 21)
    return g, oracle.CSC(1 << 13, s, d, csr=False)


def _seeds(og, n, rng):
    cand = np.flatnonzero(og.od > 0)
    hi = cand[np.argsort(-og.od[cand])][: max(1, cand.size // 100)]
    lo = cand[og.od[cand] <= np.median(og.od[cand])]
    return [int(rng.choice(hi)) if i % 2 == 0 else int(rng.choice(lo)) for i in range(n)]


def test_batch_64_columns_vs_oracle(s13):
    from plato_amd.batch import DONE, PprBatch
    g, og = s13
    rng = np.random.default_rng(1)
    seeds = _seeds(og, 64, rng)
    b = PprBatch(g, D, tol=1e-7, max_iters=300)
    for slot, s in enumerate(seeds):
        b.admit(slot, [s])
    b.run_until_done(range(64))
    st, it, _ = b.poll()
    assert (st == DONE).all()
    res = b.results(list(range(64)))
    for j, s in enumerate(seeds):
        ref, _ = oracle.ppr(og, [s], 2000, D, 0.0)
        l1 = float(np.abs(res[j] - ref).sum())
        assert l1 <= 1e-6, (j, s, l1)
        # same stopping rule as the reference run with eps = tol
        _, rdel = oracle.ppr(og, [s], 300, D, 1e-7)
        assert abs(int(it[j]) - len(rdel)) <= 1, (j, it[j], len(rdel))


def test_alone_equals_in_mix_bitwise(s13):
    from plato_amd.batch import PprBatch
    g, og = s13
    rng = np.random.default_rng(2)
    seeds = _seeds(og, 64, rng)
    q = [seeds[0], seeds[1], 77]
    alone = PprBatch(g, D, tol=1e-7, max_iters=300)
    alone.admit(5, q)
    alone.run_until_done([5])
    r_alone = alone.results([5])[0]
    mix = PprBatch(g, D, tol=1e-7, max_iters=300)
    for slot in range(64):
        mix.admit(slot, q if slot == 40 else [seeds[slot]])
    mix.run_until_done(range(64))
    r_mix = mix.results([40])[0]
    assert np.array_equal(r_alone, r_mix)
    ref, _ = oracle.ppr(og, q, 2000, D, 0.0)
    assert float(np.abs(r_mix - ref).sum()) <= 1e-6


def _alone(g, q, slot=7):
    from plato_amd.batch import PprBatch
    a = PprBatch(g, D, tol=1e-7, max_iters=300)
    a.admit(slot, q)
    a.run_until_done([slot])
    r = a.results([slot])[0]
    a.close()
    return r


def test_pass_modes_bitwise(s13):
    """Admissions staggered so that passes run on the packed block (at most 8 columns),
    with 4 and 2 edges per row load (at most 16 and 32 columns) and with one (more), and
    back down as columns finish: every column equals its query run alone, bit for bit."""
    from plato_amd.batch import PprBatch
    g, og = s13
    seeds = _seeds(og, 48, np.random.default_rng(4))
    b = PprBatch(g, D, tol=1e-7, max_iters=300)
    for a0, a1, n in [(0, 3, 6), (3, 12, 6), (12, 30, 6), (30, 48, 6)]:
        for slot in range(a0, a1):
            b.admit(slot, [seeds[slot]])
        b.step(n)
    b.run_until_done(range(48))
    res = b.results(list(range(48)))
    b.close()
    for slot in (0, 2, 5, 11, 12, 29, 30, 47):
        assert np.array_equal(res[slot], _alone(g, [seeds[slot]])), slot
    ref, _ = oracle.ppr(og, [seeds[0]], 2000, D, 0.0)
    assert float(np.abs(res[0] - ref).sum()) <= 1e-6


def test_plan_limits_do_not_change_bits(s13):
    """The pass plan's limits (pe_ppr_batch_opts_t) pick kernels, never arithmetic: the same
    staggered queries with every narrow kernel off, with only the packed block off, and with
    the default plan give the same bits."""
    from plato_amd import _lib as L
    from plato_amd.batch import PprBatch
    g, og = s13
    seeds = _seeds(og, 20, np.random.default_rng(6))
    OFF = L.PE_BATCH_PLAN_OFF
    out = []
    for plan in ((0, 0, 0), (OFF, OFF, OFF), (OFF, 0, 0), (0, OFF, OFF)):
        b = PprBatch(g, D, tol=1e-7, max_iters=300, plan=plan)
        for a0, a1, n in [(0, 2, 5), (2, 12, 5), (12, 20, 3)]:
            for slot in range(a0, a1):
                b.admit(slot, [seeds[slot]])
            b.step(n)
        b.run_until_done(range(20))
        out.append(b.results(list(range(20))))
        b.close()
    for r in out[1:]:
        assert np.array_equal(r, out[0])


def test_packed_slot_refilled_in_place(s13):
    """A column finishes and its slot is refilled before the next pass while another column
    keeps running: the pass after sees the same covered slots as the last packed pass and
    reads the packed block, whose entries for the slot still hold the old query."""
    from plato_amd.batch import DONE, PprBatch
    g, og = s13
    seeds = _seeds(og, 3, np.random.default_rng(5))
    b = PprBatch(g, D, tol=1e-7, max_iters=300)
    b.admit(0, [seeds[0]])
    b.step(20)
    b.admit(1, [seeds[1]])
    for _ in range(400):
        b.step(1)
        st, _, _ = b.poll()
        if st[0] == DONE:
            break
    assert st[0] == DONE and st[1] != DONE
    first = b.results([0])[0]
    b.release(0)
    b.admit(0, [seeds[2]])
    b.run_until_done([0, 1])
    res = b.results([0, 1])
    b.close()
    assert np.array_equal(first, _alone(g, [seeds[0]]))
    assert np.array_equal(res[0], _alone(g, [seeds[2]]))
    assert np.array_equal(res[1], _alone(g, [seeds[1]]))


def test_continuous_refill_and_topk(s13):
    from plato_amd.batch import run_queries
    g, og = s13
    rng = np.random.default_rng(3)
    queries = []
    for i in range(150):
        n = int(np.clip(np.floor(np.exp(rng.uniform(0, np.log(257)))), 1, 256))
        queries.append(rng.choice(np.flatnonzero(og.od > 0), n).astype(np.uint32))
    iters, ranks, tops = run_queries(g, queries, D, 1e-7, 300, chunk=3, want_ranks=True, topk=100)
    for qi in range(0, 150, 7):
        ref, _ = oracle.ppr(og, queries[qi], 2000, D, 0.0)
        assert float(np.abs(ranks[qi] - ref).sum()) <= 1e-6, qi
        if qi % 21 == 0:  # the same bits as the query alone, whatever the mix around it
            assert np.array_equal(ranks[qi], _alone(g, list(queries[qi]))), qi
        ids, vals = tops[qi]
        order = np.lexsort((np.arange(g.V), -ranks[qi].astype(np.float32)))[:100]
        assert np.array_equal(ids, order.astype(np.uint32))
        assert np.array_equal(vals, ranks[qi].astype(np.float32)[order])


def test_batch_outlives_its_graph(ctx):
    """A block (or BFS state) destroyed after its graph: destroy touches only the context.
    The Python wrappers close a graph's dependents before the graph itself."""
    from plato_amd import Graph
    from plato_amd.batch import PprBatch
    from plato_amd.engine import BfsEngineState
    g = Graph.rmat(ctx, 10, 8, seed=1, build_csr=True)
    b = PprBatch(g, D, tol=1e-7, max_iters=50)
    st = BfsEngineState(g)
    b.admit(0, [1])
    b.step(2)
    hb, hs = b.h, st.h
    assert g.lib.pe_graph_destroy(g.h) == 0
    g.h = None
    assert g.lib.pe_ppr_batch_destroy(hb) == 0
    assert g.lib.pe_bfs_destroy(hs) == 0
    b.h = st.h = None
    g2 = Graph.rmat(ctx, 10, 8, seed=2)
    b2 = PprBatch(g2, D, tol=1e-7, max_iters=50)
    g2.close()
    assert b2.h is None


def test_slot_errors(s13):
    from plato_amd._lib import PlatoHipError
    from plato_amd.batch import PprBatch
    g, _ = s13
    b = PprBatch(g)
    b.admit(0, [1])
    with pytest.raises(PlatoHipError):
        b.admit(0, [2])          # occupied
    with pytest.raises(PlatoHipError):
        b.results([0])           # not finished
    with pytest.raises(PlatoHipError):
        b.admit(64, [1])         # out of range
    b.release(0)
    b.admit(0, [2])
