"""GPU: PageRank / weighted PageRank / PPR on the pull engine against the FP64 oracle.

Tolerances (FP32 state, compensated sums, FP64 apply):
  normalised L1 = sum|r - r_oracle| / sum|r_oracle| <= 2e-7 after 20 iterations,
  relative L-infinity <= 2e-6.
FP64 state: relative L-infinity <= 1e-12 (only the summation order differs).
"""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu
D = 0.85


def l1n(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def linf_rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.fixture(scope="module")
def s16(ctx):
    """RMAT scale 16, edge factor 16 (the reference's CPU-runnable PageRank case)."""
    from plato_amd import Graph
    g = Graph.rmat(ctx, 16, 16, seed=1, build_csr=True)
    s, d = oracle.rmat_edges(16, 16, 1)
    return g, oracle.CSC(1 << 16, s, d)


def test_pagerank_s16_fp32_vs_fp64_oracle(s16):
    from plato_amd.algo import pagerank, pagerank_opts_t
    g, og = s16
    print("case: pagerank s16 fp32 20 iterations", flush=True)
    ref, rdel = oracle.pagerank(og, 20, D, 0.0)
    res = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32)
    assert res.iterations == 20
    assert l1n(res.rank, ref) <= 2e-7, l1n(res.rank, ref)
    assert linf_rel(res.rank, ref) <= 2e-6, linf_rel(res.rank, ref)
    np.testing.assert_allclose(res.deltas[:-1], rdel[:-1], rtol=1e-4)


def test_pagerank_s16_fp64(s16):
    from plato_amd.algo import pagerank, pagerank_opts_t
    g, og = s16
    ref, rdel = oracle.pagerank(og, 20, D, 0.0)
    res = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=64)
    assert linf_rel(res.rank, ref) <= 1e-12
    np.testing.assert_allclose(res.deltas, rdel, rtol=1e-10)


@pytest.mark.parametrize("tile", [1024, 2048, 4096])
def test_pagerank_every_tiling(ctx, tile):
    from plato_amd import Graph
    from plato_amd.algo import pagerank, pagerank_opts_t
    print(f"case: pagerank s13 tile={tile}", flush=True)
    g = Graph.rmat(ctx, 13, 16, seed=7)
    g.set_tile(tile)
    s, d = oracle.rmat_edges(13, 16, 7)
    ref, _ = oracle.pagerank(oracle.CSC(1 << 13, s, d), 20, D, 0.0)
    res = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=64)
    assert linf_rel(res.rank, ref) <= 1e-12
    res32 = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32)
    assert l1n(res32.rank, ref) <= 2e-7


def test_pagerank_chain_closed_form_exact(ctx):
    from plato_amd import Graph
    from plato_amd.algo import pagerank, pagerank_opts_t
    s, d, _ = oracle.read_edge_csv(golden("graph_10_9.csv"), weighted=True)
    g = Graph.from_edges(ctx, s, d, 10)
    res = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=64)
    exp = np.array([0.15 * sum(D ** i for i in range(k + 1)) for k in range(10)])
    np.testing.assert_allclose(res.rank, exp, rtol=1e-15)


def test_pagerank_eps_semantics(s16):
    from plato_amd.algo import pagerank, pagerank_opts_t
    g, og = s16
    ref, rdel = oracle.pagerank(og, 100, D, 1e-2)
    res = pagerank(g, pagerank_opts_t(100, D, 1e-2), precision=64)
    assert res.iterations == len(rdel)          # same early stop + one final iteration
    assert linf_rel(res.rank, ref) <= 1e-12


def test_pagerank_deterministic(s16):
    from plato_amd.algo import pagerank, pagerank_opts_t
    g, _ = s16
    a = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32).rank
    b = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32).rank
    assert np.array_equal(a, b)


def test_weighted_pagerank(ctx):
    from plato_amd import Graph
    from plato_amd.algo import pagerank_opts_t, weighted_pagerank
    s, d = oracle.rmat_edges(12, 16, 9)
    w = oracle.rmat_weights(9, s.size)
    g = Graph.rmat(ctx, 12, 16, seed=9, weighted=True)
    ref, _ = oracle.weighted_pagerank(oracle.CSC(1 << 12, s, d, w=w), 20, D, 0.0)
    res = weighted_pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32)
    assert l1n(res.rank, ref) <= 5e-7, l1n(res.rank, ref)
    res64 = weighted_pagerank(g, pagerank_opts_t(20, D, 0.0), precision=64)
    # FP32 normalised weights bound the FP64 run
    assert linf_rel(res64.rank, ref) <= 1e-6


def test_ppr_single_and_multi_seed(s16):
    from plato_amd.algo import multi_seed_ppr, personalized_pagerank, personalized_pagerank_opts_t
    g, og = s16
    src = int(np.flatnonzero(og.od > 0)[5])
    ref, rdel = oracle.ppr(og, [src], 60, D, 0.0)
    res = personalized_pagerank(g, personalized_pagerank_opts_t(60, D, src, 0.0))
    assert float(np.abs(res.rank - ref).sum()) <= 1e-6
    res64 = personalized_pagerank(g, personalized_pagerank_opts_t(60, D, src, 0.0), precision=64)
    assert float(np.abs(res64.rank - ref).sum()) <= 1e-13
    seeds = [src, 17, 17, 4000]
    refm, _ = oracle.ppr(og, seeds, 40, D, 0.0)
    rm = multi_seed_ppr(g, seeds, 40, D, 0.0, precision=64)
    assert float(np.abs(rm.rank - refm).sum()) <= 1e-13


def test_ppr_pushpull_matches_oracle_modes(s16):
    from plato_amd.algo import personalized_pagerank_opts_t, personalized_pagerank_pushpull
    g, og = s16
    src = int(np.flatnonzero((og.od > 0) & (og.od < 4))[3])
    ref, rdel, rmodes = oracle.ppr_pushpull(og, src, 30, D, 0.0)
    res = personalized_pagerank_pushpull(g, personalized_pagerank_opts_t(30, D, src, 0.0),
                                         precision=64)
    assert np.array_equal(res.modes, rmodes)
    assert float(np.abs(res.rank - ref).sum()) <= 1e-13
    assert rmodes[0] == 1 and rmodes[-1] == 0


def test_engine_level_operators_match_fused(ctx, s16):
    """pull_sum + rank_apply == rank_iterate == one step of the oracle."""
    from plato_amd import dualmode_engine_t
    g, og = s16
    eng = dualmode_engine_t(g)
    x = eng.alloc_v_state(np.float64)
    y = eng.alloc_v_state(np.float64)
    z = eng.alloc_v_state(np.float64)
    x0 = np.where(og.od > 0, 1.0 / np.maximum(og.od, 1), 1.0)
    x.upload(x0)
    eng.pull_sum(x, y)
    sums = y.to_f64()
    cols = np.repeat(np.arange(og.V), np.diff(og.col_ptr).astype(np.int64))
    exp_sum = np.bincount(cols, weights=x0[og.row_idx], minlength=og.V)
    np.testing.assert_allclose(sums, exp_sum, rtol=1e-13)
    d1 = eng.rank_apply(x, y, D, 1 - D)
    d2 = eng.rank_iterate(x, z, D, 1 - D)
    yr, dr = oracle.pagerank_step(og, x0, D, False)
    np.testing.assert_allclose(y.to_f64(), yr, rtol=1e-13)
    np.testing.assert_allclose(z.to_f64(), yr, rtol=1e-13)
    assert abs(d1 - dr) <= 1e-9 * dr and abs(d2 - dr) <= 1e-9 * dr


def test_aggregate_message_known_answer(ctx):
    """message_passing.cc:48-88 through the pull operator: sum of ones per column."""
    from plato_amd import Graph, dualmode_engine_t
    s, d, w = oracle.read_edge_csv(golden("graph_10_9.csv"), weighted=True)
    g = Graph.from_edges(ctx, s, d, 10)
    eng = dualmode_engine_t(g)
    x = eng.alloc_v_state(np.float32)
    y = eng.alloc_v_state(np.float32)
    x.fill(1.0)
    eng.pull_sum(x, y)
    vals = y.numpy()
    assert vals.sum() == 9 and vals[0] == 0 and (vals[1:] == 1).all()


def test_relabeled_graph_gives_same_ranks(ctx, s16):
    """Degree relabeling is internal: results come back in input id order."""
    from plato_amd import Graph
    from plato_amd.algo import (multi_seed_ppr, pagerank, pagerank_opts_t,
                                personalized_pagerank_opts_t, personalized_pagerank_pushpull)
    _, og = s16
    g = Graph.rmat(ctx, 16, 16, seed=1, build_csr=True, relabel=True)
    assert g.info.relabeled == 1
    rank_of = g.permutation()
    assert np.array_equal(np.sort(rank_of), np.arange(og.V, dtype=np.uint32))
    # internal ids ordered by out-degree, descending
    od_internal = np.zeros(og.V, np.uint32)
    od_internal[rank_of] = og.od
    assert (np.diff(od_internal.astype(np.int64)) <= 0).all()
    # internal CSC is the oracle CSC of the renumbered edge list
    s, d = oracle.rmat_edges(16, 16, 1)
    orl = oracle.CSC(og.V, rank_of[s], rank_of[d])
    cp, ri = g.download_csc()
    assert np.array_equal(cp, orl.col_ptr) and np.array_equal(ri, orl.row_idx)
    assert np.array_equal(g.out_degrees(), og.od)
    gi = g.graph_info()
    present = (og.od > 0) | (og.in_degrees() > 0)
    assert gi.max_v_i_ == int(np.flatnonzero(present).max()) and gi.vertices_ == int(present.sum())
    ref, _ = oracle.pagerank(og, 20, D, 0.0)
    res = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=64)
    assert linf_rel(res.rank, ref) <= 1e-12
    res32 = pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32)
    assert l1n(res32.rank, ref) <= 2e-7
    seeds = [5, 900, 31000]
    refm, _ = oracle.ppr(og, seeds, 40, D, 0.0)
    rm = multi_seed_ppr(g, seeds, 40, D, 0.0, precision=64)
    assert float(np.abs(rm.rank - refm).sum()) <= 1e-13
    src = int(np.flatnonzero((og.od > 0) & (og.od < 4))[3])
    refp, _, modes = oracle.ppr_pushpull(og, src, 30, D, 0.0)
    rp = personalized_pagerank_pushpull(g, personalized_pagerank_opts_t(30, D, src, 0.0), precision=64)
    assert np.array_equal(rp.modes, modes) and float(np.abs(rp.rank - refp).sum()) <= 1e-13


def test_ppr_pushpull_hub_seed(s16):
    """Seeded at the vertex of most out-edges (more than 4096: its push is cut into chunks
    over all waves), in both precisions."""
    from plato_amd.algo import personalized_pagerank_opts_t, personalized_pagerank_pushpull
    g, og = s16
    src = int(np.argmax(og.od))
    assert og.od[src] > 4096
    ref, rdel, rmodes = oracle.ppr_pushpull(og, src, 30, D, 0.0)
    assert rmodes[0] == 1
    res = personalized_pagerank_pushpull(g, personalized_pagerank_opts_t(30, D, src, 0.0), precision=64)
    assert np.array_equal(res.modes, rmodes)
    assert float(np.abs(res.rank - ref).sum()) <= 1e-13
    res32 = personalized_pagerank_pushpull(g, personalized_pagerank_opts_t(30, D, src, 0.0))
    assert np.array_equal(res32.modes, rmodes)
    assert float(np.abs(res32.rank - ref).sum()) <= 1e-5
