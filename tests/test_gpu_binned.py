"""GPU: the binned (partition-centric) pull layout of the FP32 rank iterations.

The binned gather sums every column exactly in fixed point (each term truncated at
2^-F with F set by the rank mass, then 64-bit integer sums, independent of order), so
  * against the FP64 oracle it meets the FP32 tolerances of the merge-path
    gather: normalised L1 <= 2e-7, relative L-infinity <= 2e-6 after 20 PageRank
    iterations (weighted: L1 <= 5e-7; PPR: absolute L1 <= 1e-6);
  * its results are bit-identical under every layout choice (chunk and bin
    sizes, run records, split bins) and from run to run: the sum into a column
    does not depend on the order of its terms.
"""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu
D = 0.85

# layouts: library default, and a tiny one that makes many chunks, bins, run
# records and split bins even on small graphs
TINY = dict(src_bits=4, dst_bits=4, run_cap=64, unit_log2=10)
MID = dict(src_bits=9, dst_bits=7, run_cap=256, unit_log2=12)


def l1n(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def linf_rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.fixture(scope="module")
def s14(ctx):
    from plato_amd import Graph
    g = Graph.rmat(ctx, 14, 16, seed=3, build_csr=True)
    s, d = oracle.rmat_edges(14, 16, 3)
    return g, oracle.CSC(1 << 14, s, d)


def _pr(g, layout, iters=20, eps=0.0, **opts):
    from plato_amd.algo import pagerank, pagerank_opts_t
    g.set_pull_layout(layout, **opts)
    try:
        return pagerank(g, pagerank_opts_t(iters, D, eps), precision=32)
    finally:
        g.set_pull_layout("merge_path")


def test_binned_pagerank_vs_oracle_and_layout_independent(s14):
    g, og = s14
    ref, rdel = oracle.pagerank(og, 20, D, 0.0)
    results = []
    for opts in ({}, TINY, MID):
        print(f"case: binned pagerank s14 layout={opts}", flush=True)
        res = _pr(g, "binned", **opts)
        assert res.iterations == 20
        assert l1n(res.rank, ref) <= 2e-7, l1n(res.rank, ref)
        assert linf_rel(res.rank, ref) <= 2e-6, linf_rel(res.rank, ref)
        np.testing.assert_allclose(res.deltas[:-1], rdel[:-1], rtol=1e-4)
        results.append(res.rank)
    for r in results[1:]:
        assert np.array_equal(r, results[0])
    # and run to run
    assert np.array_equal(_pr(g, "binned", **TINY).rank, results[0])


def test_binned_layout_info_and_release(s14):
    g, _ = s14
    g.set_pull_layout("binned", **TINY)
    info = g.info
    assert info.pull_layout == 1 and info.binned_runs > 0 and info.binned_bytes >= 2 * info.nnz
    g.set_pull_layout("merge_path")
    assert g.info.pull_layout == 0 and g.info.binned_bytes == 0


def test_binned_rejects_bad_options(s14):
    from plato_amd._lib import PlatoHipError
    g, _ = s14
    for bad in (dict(src_bits=16), dict(dst_bits=15), dict(src_bits=3), dict(run_cap=32), dict(run_cap=512),
                dict(unit_log2=31)):
        with pytest.raises(PlatoHipError):
            g.set_pull_layout("binned", **bad)
    assert g.info.pull_layout == 0


def test_binned_eps_semantics(s14):
    g, og = s14
    ref, rdel = oracle.pagerank(og, 100, D, 1e-2)
    res = _pr(g, "binned", iters=100, eps=1e-2, **MID)
    assert res.iterations == len(rdel)
    assert l1n(res.rank, ref) <= 2e-7


def test_binned_golden_graphs(ctx):
    from plato_amd import Graph
    # chain closed form, rank_k = 0.15 * sum_{i<=k} 0.85^i
    s, d, _ = oracle.read_edge_csv(golden("graph_10_9.csv"), weighted=True)
    g = Graph.from_edges(ctx, s, d, 10)
    res = _pr(g, "binned", **TINY)
    exp = np.array([0.15 * sum(D ** i for i in range(k + 1)) for k in range(10)])
    np.testing.assert_allclose(res.rank, exp, rtol=2e-7)
    for name in ("v100_e2150_ua_c3.csv", "v100_e4300_da_c3.csv"):
        s, d = oracle.read_edge_csv(golden(name))
        g = Graph.from_edges(ctx, s, d, 100)
        ref, _ = oracle.pagerank(oracle.CSC(100, s, d), 20, D, 0.0)
        for opts in ({}, TINY):
            res = _pr(g, "binned", **opts)
            assert l1n(res.rank, ref) <= 2e-7, (name, opts)


def test_binned_weighted_pagerank(ctx):
    from plato_amd import Graph
    from plato_amd.algo import pagerank_opts_t, weighted_pagerank
    s, d = oracle.rmat_edges(12, 16, 9)
    w = oracle.rmat_weights(9, s.size)
    g = Graph.rmat(ctx, 12, 16, seed=9, weighted=True)
    ref, _ = oracle.weighted_pagerank(oracle.CSC(1 << 12, s, d, w=w), 20, D, 0.0)
    out = []
    for opts in ({}, TINY):
        print(f"case: binned weighted pagerank s12 layout={opts}", flush=True)
        g.set_pull_layout("binned", **opts)
        res = weighted_pagerank(g, pagerank_opts_t(20, D, 0.0), precision=32)
        assert l1n(res.rank, ref) <= 5e-7, l1n(res.rank, ref)
        out.append(res.rank)
    g.set_pull_layout("merge_path")
    assert np.array_equal(out[0], out[1])


def test_binned_ppr_and_pushpull(s14):
    from plato_amd.algo import (multi_seed_ppr, personalized_pagerank, personalized_pagerank_opts_t,
                                personalized_pagerank_pushpull)
    g, og = s14
    src = int(np.flatnonzero(og.od > 0)[5])
    ref, _ = oracle.ppr(og, [src], 60, D, 0.0)
    g.set_pull_layout("binned", **MID)
    try:
        res = personalized_pagerank(g, personalized_pagerank_opts_t(60, D, src, 0.0))
        assert float(np.abs(res.rank - ref).sum()) <= 1e-6
        seeds = [src, 17, 17, 4000]
        refm, _ = oracle.ppr(og, seeds, 40, D, 0.0)
        rm = multi_seed_ppr(g, seeds, 40, D, 0.0)
        assert float(np.abs(rm.rank - refm).sum()) <= 1e-6
        lo = int(np.flatnonzero((og.od > 0) & (og.od < 4))[3])
        refp, _, modes = oracle.ppr_pushpull(og, lo, 30, D, 0.0)
        rp = personalized_pagerank_pushpull(g, personalized_pagerank_opts_t(30, D, lo, 0.0))
        assert np.array_equal(rp.modes, modes)
        assert float(np.abs(rp.rank - refp).sum()) <= 1e-6
    finally:
        g.set_pull_layout("merge_path")


def test_binned_relabeled_graph(ctx, s14):
    from plato_amd import Graph
    _, og = s14
    g = Graph.rmat(ctx, 14, 16, seed=3, relabel=True)
    ref, _ = oracle.pagerank(og, 20, D, 0.0)
    a = _pr(g, "binned").rank
    b = _pr(g, "binned", **TINY).rank
    assert np.array_equal(a, b)
    assert l1n(a, ref) <= 2e-7


def test_binned_engine_iterate_and_kernel_timing(ctx, s14):
    """pe_rank_iterate on the binned layout == one oracle step; timed launches agree."""
    from plato_amd import dualmode_engine_t
    from plato_amd import _lib as L
    g, og = s14
    g.set_pull_layout("binned", **MID)
    try:
        eng = dualmode_engine_t(g)
        x = eng.alloc_v_state(np.float32)
        z = eng.alloc_v_state(np.float32)
        x0 = np.where(og.od > 0, 1.0 / np.maximum(og.od, 1), 1.0).astype(np.float32)
        x.upload(x0)
        d2 = eng.rank_iterate(x, z, D, 1 - D)
        yr, dr = oracle.pagerank_step(og, x0.astype(np.float64), D, False)
        np.testing.assert_allclose(z.to_f64(), yr, rtol=1e-6)
        assert abs(d2 - dr) <= 1e-5 * dr
        untimed = _pr(g, "binned", **MID).rank
        g.set_pull_layout("binned", **MID)
        ctx.set_kernel_timing(True)
        try:
            timed = _pr(g, "binned", **MID).rank
            ms, n = ctx.kernel_time(L.PE_KT_PULL_BIN_GATHER)
            assert n == 20 and ms > 0
            ms, n = ctx.kernel_time(L.PE_KT_PULL_BIN_SCATTER)
            assert n == 20 and ms > 0
        finally:
            ctx.set_kernel_timing(False)
        assert np.array_equal(untimed, timed)
    finally:
        g.set_pull_layout("merge_path")


def test_binned_s20_matches_merge_path(ctx):
    """At 2^20 vertices (16.8M edges): binned and merge-path agree, and rank mass
    is conserved the same way (sum of ranks equal to FP32 rounding)."""
    from plato_amd import Graph
    g = Graph.rmat(ctx, 20, 16, seed=5, relabel=True)
    print("case: binned vs merge-path s20", flush=True)
    mp = _pr(g, "merge_path").rank
    bn = _pr(g, "binned").rank
    assert l1n(bn, mp) <= 2e-7, l1n(bn, mp)
    assert abs(bn.sum() - mp.sum()) <= 1e-6 * mp.sum()
    assert np.array_equal(_pr(g, "binned", **MID).rank, bn)
