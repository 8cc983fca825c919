"""GPU parity at the benchmark configurations' own sizes.

The other GPU tests stay at sizes where every device path is already covered in
seconds; these reach the code that only runs on large graphs:

* BFS on symmetrized RMAT-22 (65 537 frontier words, seven in eight of them past the
  8192-word frontier head the pull step keeps in LDS) and RMAT-24, the BFS
  configuration itself: levels, per-epoch modes, frontier sizes and active edges
  bit-exact against the oracle (bfs.hpp:79-154, dualmode.hpp:264-279).
* PageRank and weighted PageRank on RMAT-27 (2^31 edges, uint64 offsets), both pull
  layouts: the whole CSC against the generator's edge multiset (digest), out-degrees
  against the generator, one iteration on 16 column ranges against the oracle's
  step, and the rank-mass identity over the whole vector (pagerank.hpp:100-146,
  weighted_pagerank.hpp:110-143).
* Batched PPR on RMAT-24, 64 columns: columns against the converged FP64 oracle
  (personalized_pagerank.hpp:40-134).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

D = 0.85


def _say(msg):
    print(msg, flush=True)


def _check_bfs(g, og, root, threshold=0.05):
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    ref = oracle.bfs(og, int(root), threshold=threshold)
    res = breadth_first_search(g, bfs_opts_t(int(root)), push_threshold=threshold)
    assert np.array_equal(res.levels, ref["levels"]), f"root {root} threshold {threshold}"
    assert np.array_equal(res.modes, ref["modes"]), (res.modes, ref["modes"])
    assert np.array_equal(res.frontier, ref["frontier"])
    assert np.array_equal(res.active_edges, ref["active_edges"])
    assert res.visited == ref["visited"]
    assert res.reached_edges == int(og.od[ref["levels"] >= 0].sum())
    return res, ref


def _sym_oracle(scale, seed):
    s, d = oracle.rmat_edges(scale, 16, seed)
    og = oracle.CSC(1 << scale, s, d, symmetrize=True, csr=False)
    og.row_ptr, og.col_idx = og.col_ptr, og.row_idx  # symmetrized: CSR == CSC
    return og


@pytest.mark.timeout(600)
def test_bfs_rmat22_past_the_lds_frontier_head(ctx):
    from plato_amd import Graph
    scale, seed = 22, 5
    words = ((1 << scale) >> 6) + 1
    assert (words - 8192) / words > 0.75
    _say(f"case: bfs rmat s{scale} symmetrized relabeled, {words} frontier words")
    og = _sym_oracle(scale, seed)
    g = Graph.rmat(ctx, scale, 16, seed=seed, directed=False, relabel=True)
    cand = np.flatnonzero(og.od > 0)
    order = cand[np.argsort(og.od[cand], kind="stable")]
    rng = np.random.default_rng(22)
    roots = [int(order[0]), int(order[len(order) // 2]), int(order[-1])] + \
        [int(r) for r in rng.choice(cand, 3, replace=False)]
    internal = g.permutation()  # internal id of every input id
    pulled_far = 0
    for root in roots:
        _say(f"  root {root} (degree {og.od[root]})")
        res, ref = _check_bfs(g, og, root)
        # pull levels whose frontier holds vertices past the LDS head (internal ids >= 2^19)
        lv = ref["levels"]
        for ep, m in enumerate(ref["modes"].tolist()):
            if m == 0 and np.any(internal[lv == ep] >= (1 << 19)):
                pulled_far += 1
    assert pulled_far > 0
    for threshold in (1e9, 0.0):  # every level top-down, every level bottom-up
        for root in roots[:2] + roots[-1:]:
            _say(f"  root {root} forced threshold {threshold}")
            res, _ = _check_bfs(g, og, root, threshold)
            assert set(res.modes.tolist()) == ({1} if threshold > 1 else {0})
    g.close()


@pytest.mark.timeout(900)
def test_bfs_rmat24_configuration(ctx):
    """The BFS configuration's graph (scale 24, symmetrized, degree-relabeled) and its
    first roots as the benchmark draws them."""
    from plato_amd import Graph
    scale, seed = 24, 11
    _say(f"case: bfs rmat s{scale} symmetrized relabeled (benchmark graph)")
    og = _sym_oracle(scale, seed)
    g = Graph.rmat(ctx, scale, 16, seed=seed, directed=False, relabel=True)
    od = g.out_degrees()
    assert np.array_equal(od, og.od)
    rng = np.random.default_rng(seed)
    roots = rng.choice(np.flatnonzero(od > 0), 64, replace=False)[:3]
    for root in roots:
        _say(f"  root {root} (degree {od[root]})")
        _check_bfs(g, og, int(root))
    g.close()


def _ranges(V, n=16):
    width = V // 256
    return [(k * V // n, k * V // n + width) for k in range(n)]


def _slice_mask(V, ranges):
    m = np.zeros(V, bool)
    for lo, hi in ranges:
        m[lo:hi] = True
    return m


def _structure_checks(g, scan):
    """Out-degrees and the whole CSC's edge multiset against the generator."""
    od = g.out_degrees()
    assert np.array_equal(od, scan.od), "out-degrees differ from the generator's"
    cp, ri = g.download_csc()
    perm = np.argsort(g.permutation()).astype(np.uint32)  # internal -> input id
    assert int(cp[-1]) == scan.E
    assert oracle.csc_digest(g.V, cp, ri, perm) == scan.digest, "CSC edge multiset differs"
    del cp, ri


def _compare(gpu, ref, mask, what, rtol=2e-6):
    a, b = gpu[mask], ref[mask]
    err = np.abs(a - b) / np.abs(b)
    worst = int(np.argmax(err))
    assert err[worst] <= rtol, f"{what}: relative error {err[worst]:.3g} at slice index {worst}"
    return float(err.max())


@pytest.mark.timeout(1200)
def test_pagerank_rmat27_against_oracle(ctx):
    from plato_amd import Graph
    from plato_amd.algo import pagerank, pagerank_opts_t
    scale, seed = 27, 1
    V = 1 << scale
    ranges = _ranges(V)
    mask = _slice_mask(V, ranges)
    _say(f"case: pagerank rmat s{scale}: oracle scan of 2^31 edges")
    scan = oracle.RmatScan(scale, 16, seed, ranges, weighted=True)  # weights for the second half
    og = scan.slice
    od = scan.od.astype(np.float64)
    _say(f"  {scan.eids.size} in-edges in the 16 column ranges")
    g = Graph.rmat(ctx, scale, 16, seed=seed, relabel=True)
    _structure_checks(g, scan)
    x0 = np.where(od > 0, 1.0 / np.maximum(od, 1), 1.0)
    ref1, _ = oracle.pagerank_step(_Host(og, scan.od), x0, D, True, 0, V)
    for layout in ("binned", "merge_path"):
        _say(f"  layout {layout}")
        g.set_pull_layout(layout)
        r1 = pagerank(g, pagerank_opts_t(1, D, 0.0)).rank
        e1 = _compare(r1, ref1, mask, f"{layout} iteration 1")
        r3 = pagerank(g, pagerank_opts_t(3, D, 0.0)).rank
        r4 = pagerank(g, pagerank_opts_t(4, D, 0.0)).rank
        x3 = np.where(od > 0, r3 / np.maximum(od, 1), r3)
        ref4, _ = oracle.pagerank_step(_Host(og, scan.od), x3, D, True, 0, V)
        e4 = _compare(r4, ref4, mask, f"{layout} iteration 4")
        # rank mass: sum r4 = (1-d) V + d * sum of r3 over vertices with out-edges
        mass = (1 - D) * V + D * float(r3[od > 0].sum())
        assert abs(float(r4.sum()) - mass) <= 1e-6 * mass, (float(r4.sum()), mass)
        _say(f"  {layout}: max relative error {e1:.2e} (iteration 1), {e4:.2e} (iteration 4)")
    g.close()

    _say(f"case: weighted pagerank rmat s{scale}")
    g = Graph.rmat(ctx, scale, 16, seed=seed, weighted=True, relabel=True)
    _structure_checks(g, scan)
    from plato_amd.algo import weighted_pagerank
    ref1 = oracle.weighted_pagerank_step(og, scan.W, np.ones(V), D)
    has_out = scan.W > 0
    for layout in ("binned", "merge_path"):
        _say(f"  layout {layout}")
        g.set_pull_layout(layout)
        r1 = weighted_pagerank(g, pagerank_opts_t(1, D, 0.0)).rank
        e1 = _compare(r1, ref1, mask, f"weighted {layout} iteration 1")
        r3 = weighted_pagerank(g, pagerank_opts_t(3, D, 0.0)).rank
        r4 = weighted_pagerank(g, pagerank_opts_t(4, D, 0.0)).rank
        ref4 = oracle.weighted_pagerank_step(og, scan.W, r3, D)
        e4 = _compare(r4, ref4, mask, f"weighted {layout} iteration 4")
        mass = (1 - D) * V + D * float(r3[has_out].sum())
        assert abs(float(r4.sum()) - mass) <= 1e-6 * mass, (float(r4.sum()), mass)
        _say(f"  weighted {layout}: max relative error {e1:.2e}, {e4:.2e}")
    g.close()


class _Host:
    """Oracle view of a slice CSC with the whole graph's out-degrees."""

    def __init__(self, og, od):
        self.V, self.col_ptr, self.row_idx, self.od = og.V, og.col_ptr, og.row_idx, od


@pytest.mark.timeout(900)
def test_batched_ppr_rmat24_against_converged_oracle(ctx):
    from plato_amd import Graph
    from plato_amd.batch import PprBatch
    scale, seed, tol = 24, 24, 1e-7
    _say(f"case: batched ppr rmat s{scale}, 64 columns")
    s, d = oracle.rmat_edges(scale, 16, seed)
    og = oracle.CSC(1 << scale, s, d, csr=False)
    del s, d
    g = Graph.rmat(ctx, scale, 16, seed=seed, relabel=True, build_csr=True)
    od = g.out_degrees()
    assert np.array_equal(od, og.od)
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(od > 0)
    order = cand[np.argsort(-od[cand], kind="stable")]
    hi = order[: cand.size // 100]
    lo = cand[od[cand] <= np.median(od[cand])]
    seeds = [int(rng.choice(hi)) if i % 2 == 0 else int(rng.choice(lo)) for i in range(64)]
    b = PprBatch(g, D, tol, 200)
    for slot, sd in enumerate(seeds):
        b.admit(slot, [sd])
    b.run_until_done(range(64), chunk=4)
    check = [0, 63]  # a high-degree and a low-degree seed
    got = b.results(check)
    b.close()
    g.close()
    for i, slot in enumerate(check):
        ref, _ = oracle.ppr(og, [seeds[slot]], 1000, D, 1e-12)
        l1 = float(np.abs(got[i] - ref).sum())
        _say(f"  slot {slot} seed {seeds[slot]} (degree {od[seeds[slot]]}): L1 {l1:.2e}")
        assert l1 <= 1e-6, (slot, l1)
