"""GPU: TrustRank (trustrank.hpp:23-259) and the reversed graph, against the oracle.

Seed selection ranks vertices with PageRank (100 iterations, damping 0.85, eps 0.001,
FP64) on the graph (method 1) or on its reverse (method 0, inverse PageRank) and keeps
the top seed_num by (rank desc, id asc) that are good; the oracle does the same with
orc_pagerank and orc_topk.  Ranks that differ from the cut-off by less than 1e-12
relative may order either way (summation order), so the sets are compared outside that
band.  The TrustRank iteration is the seed-set PPR: FP64 relative L-inf <= 1e-12
against orc_ppr with the same seeds."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def s12(ctx):
    from plato_amd import Graph
    s, d = oracle.rmat_edges(12, 16, 8)
    g = Graph.rmat(ctx, 12, 16, seed=8, build_csr=True)
    return g, s, d


def _expected_seeds(s, d, V, seed_num, good, method):
    a, b = (d, s) if method == 0 else (s, d)
    og = oracle.CSC(V, a, b)
    r, _ = oracle.pagerank(og, 100, 0.85, 0.001)
    n = int(max(s.max(), d.max())) + 1
    ids, vals = oracle.topk(r[:n], seed_num)
    cut = vals[-1]
    sure = ids[np.abs(vals - cut) > 1e-12 * cut]
    return r, set(int(v) for v in sure if good[v]), cut


def test_reverse_graph(ctx, s12):
    from plato_amd import Graph
    g, s, d = s12
    rg = g.reverse(build_csr=True)
    og = oracle.CSC(g.V, d, s)
    cp, ri = rg.download_csc()
    assert np.array_equal(cp, og.col_ptr) and np.array_equal(ri, og.row_idx)
    # reverse of a relabeled graph keeps input ids
    gr = Graph.rmat(ctx, 12, 16, seed=8, relabel=True)
    rr = gr.reverse()
    assert np.array_equal(rr.download_csc()[1], ri) and np.array_equal(rr.out_degrees(), og.od)


@pytest.mark.parametrize("method", [0, 1])
def test_seed_selection_by_pagerank(ctx, s12, method):
    from plato_amd.algo import trustrank_opts_t, trustrank_seeds
    g, s, d = s12
    rng = np.random.default_rng(method)
    good_mask = rng.random(g.V) < 0.6
    good = np.flatnonzero(good_mask)
    rg = g.reverse() if method == 0 else None
    seeds = trustrank_seeds(g, trustrank_opts_t(seed_num_=100, select_method_=method), good=good, reverse=rg)
    r, sure, cut = _expected_seeds(s, d, g.V, 100, good_mask, method)
    got = set(seeds.tolist())
    assert got <= set(good.tolist()) and len(got) <= 100
    assert sure <= got
    extra = got - sure
    assert all(abs(r[v] - cut) <= 1e-12 * cut for v in extra)


def test_trustrank_matches_seed_set_ppr(ctx, s12):
    from plato_amd.algo import trustrank, trustrank_opts_t
    g, s, d = s12
    opts = trustrank_opts_t(iteration_=60, seed_num_=50, select_method_=1, eps_=0.0)
    res, seeds = trustrank(g, opts, good=None)
    ref, _ = oracle.ppr(oracle.CSC(g.V, s, d), seeds.tolist(), 60, 0.85, 0.0)
    assert seeds.size == 50
    np.testing.assert_allclose(res.rank, ref, rtol=1e-12, atol=1e-300)
    # eps semantics: the reference's jump to one final iteration
    o2 = trustrank_opts_t(iteration_=100, seed_num_=50, select_method_=1, eps_=1e-3)
    r2, s2 = trustrank(g, o2)
    ref2, d2 = oracle.ppr(oracle.CSC(g.V, s, d), s2.tolist(), 100, 0.85, 1e-3)
    assert r2.iterations == len(d2)
    np.testing.assert_allclose(r2.rank, ref2, rtol=1e-12, atol=1e-300)


def test_random_selection_and_small_graph(ctx, s12):
    from plato_amd import Graph
    from plato_amd.algo import trustrank_opts_t, trustrank_seeds
    g, s, d = s12
    good = np.flatnonzero(np.arange(g.V) % 3 != 0)
    o = trustrank_opts_t(seed_num_=200, select_method_=2, random_seed_=42)
    a = trustrank_seeds(g, o, good=good)
    b = trustrank_seeds(g, o, good=good)
    assert np.array_equal(a, b) and 0 < a.size <= 200 and set(a.tolist()) <= set(good.tolist())
    all_good = trustrank_seeds(g, o)
    assert all_good.size == 200
    # fewer vertices than seed_num: every good vertex is a seed (trustrank.hpp:95-98)
    small = Graph.from_edges(ctx, np.array([0, 1, 2], np.uint32), np.array([1, 2, 0], np.uint32), 3)
    assert trustrank_seeds(small, trustrank_opts_t(seed_num_=10), good=[0, 2]).tolist() == [0, 2]
