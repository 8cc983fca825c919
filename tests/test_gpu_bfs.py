"""GPU: direction-switching BFS.  Levels are bit-exact to the oracle (integer
work), and so are the per-epoch push/pull decisions, frontier sizes and active
edge counts."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("directed", [False, True])
def test_bfs_rmat_bit_exact(ctx, directed):
    from plato_amd import Graph
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    scale = 13
    print(f"case: bfs rmat s{scale} directed={directed}", flush=True)
    g = Graph.rmat(ctx, scale, 16, seed=77, directed=directed, build_csr=True)
    s, d = oracle.rmat_edges(scale, 16, 77)
    og = oracle.CSC(1 << scale, s, d, symmetrize=not directed)
    rng = np.random.default_rng(5)
    cand = np.flatnonzero(og.od > 0)
    roots = list(rng.choice(cand, 6, replace=False)) + [int(np.argmax(og.od))]
    for root in roots:
        ref = oracle.bfs(og, int(root))
        res = breadth_first_search(g, bfs_opts_t(int(root)))
        assert np.array_equal(res.levels, ref["levels"]), f"root {root}"
        assert res.epochs == ref["epochs"]
        assert np.array_equal(res.modes, ref["modes"])
        assert np.array_equal(res.frontier, ref["frontier"])
        assert np.array_equal(res.active_edges, ref["active_edges"])
        assert res.visited == ref["visited"]
        assert res.reached_edges == int(og.od[ref["levels"] >= 0].sum())
        assert res.alg_bytes == _alg_bytes(og, ref["levels"], ref["modes"], ref["frontier"],
                                           ref["active_edges"]), f"root {root}"


def test_bfs_modes_cover_push_and_pull(ctx):
    """A low-degree root on a symmetrized RMAT goes push -> pull -> push."""
    from plato_amd import Graph
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    g = Graph.rmat(ctx, 14, 16, seed=2, directed=False)
    s, d = oracle.rmat_edges(14, 16, 2)
    og = oracle.CSC(1 << 14, s, d, symmetrize=True)
    root = int(np.flatnonzero(og.od == 1)[0])
    ref = oracle.bfs(og, root)
    res = breadth_first_search(g, bfs_opts_t(root))
    assert np.array_equal(res.levels, ref["levels"])
    m = res.modes.tolist()
    assert 1 in m and 0 in m and m[0] == 1


def test_bfs_fixtures(ctx):
    from plato_amd import Graph
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    s, d = oracle.read_edge_csv(golden("complete_binary_tree_10_9.csv"))
    g = Graph.from_edges(ctx, s, d, 10, build_csr=True)
    res = breadth_first_search(g, bfs_opts_t(0))
    assert np.array_equal(res.levels, np.floor(np.log2(np.arange(10) + 1)).astype(np.int32))
    assert res.visited == 10
    s, d, _ = oracle.read_edge_csv(golden("graph_10_9.csv"), weighted=True)
    g = Graph.from_edges(ctx, s, d, 10, build_csr=True)
    res = breadth_first_search(g, bfs_opts_t(4))
    assert res.levels.tolist() == [-1] * 4 + list(range(6)) and res.visited == 6


def test_bfs_isolated_root(ctx):
    from plato_amd import Graph
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    g = Graph.from_edges(ctx, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32), 8,
                         build_csr=True)
    res = breadth_first_search(g, bfs_opts_t(5))
    assert res.visited == 1 and res.epochs == 1 and res.levels[5] == 0
    assert (np.delete(res.levels, 5) == -1).all()


def test_bfs_relabeled_graph(ctx):
    from plato_amd import Graph
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    g = Graph.rmat(ctx, 13, 16, seed=77, directed=False, relabel=True)
    s, d = oracle.rmat_edges(13, 16, 77)
    og = oracle.CSC(1 << 13, s, d, symmetrize=True)
    for root in np.flatnonzero(og.od > 0)[[0, 50, 999]]:
        ref = oracle.bfs(og, int(root))
        res = breadth_first_search(g, bfs_opts_t(int(root)))
        assert np.array_equal(res.levels, ref["levels"])
        assert np.array_equal(res.modes, ref["modes"]) and res.visited == ref["visited"]


def test_engine_step_loop_mirrors_reference(ctx):
    """bfs.hpp:88-151 driven step by step: choose mode from the active set, one
    foreach_edges epoch, swap; every decision and frontier equals the oracle's."""
    from plato_amd import Graph, dualmode_engine_t
    from plato_amd.engine import BfsEngineState
    for directed in (False, True):
        g = Graph.rmat(ctx, 12, 16, seed=31, directed=directed, build_csr=True)
        s, d = oracle.rmat_edges(12, 16, 31)
        og = oracle.CSC(1 << 12, s, d, symmetrize=not directed)
        root = int(np.flatnonzero(og.od == og.od[og.od > 0].min())[0])
        ref = oracle.bfs(og, root)
        eng = dualmode_engine_t(g)
        cur, nxt = eng.alloc_v_subset(), eng.alloc_v_subset()
        st = BfsEngineState(g)
        st.reset(root, cur)
        actives, epoch, modes, visited = 1, 0, [], 1
        while actives:
            push, A = eng.choose_push(cur)
            assert A == ref["active_edges"][epoch] and cur.count() == ref["frontier"][epoch]
            modes.append(int(push))
            actives, _ = st.step(push, cur, nxt)
            visited += actives
            cur, nxt = nxt, cur
            epoch += 1
        assert modes == ref["modes"].tolist() and visited == ref["visited"]
        assert np.array_equal(st.levels(), ref["levels"])


def _alg_bytes(og, levels, modes, frontier, active):
    """Algorithmic bytes of a search from its levels, counted per level as the bench's BFS
    roofline counts them: push 8|F| + 4 sum_F deg + 4|new|; pull 8|U| + 4 examined + V/8,
    where examined is what a sequential scan of each unvisited vertex's ascending in-edge
    list reads up to its first frontier member (the whole list if there is none)."""
    V = og.V
    cp = og.col_ptr.astype(np.int64)
    ri = og.row_idx[: int(cp[-1])]
    deg = np.diff(cp)
    pos = np.arange(ri.size, dtype=np.int64) - np.repeat(cp[:-1], deg)
    nz = np.flatnonzero(deg > 0)
    total = 0
    for e, m in enumerate(modes.tolist()):
        n = int(frontier[e])
        nxt = int(frontier[e + 1]) if e + 1 < len(frontier) else 0
        if m == 1:
            total += 8 * n + 4 * int(active[e]) + 4 * nxt
            continue
        cand = (levels > e) | (levels < 0)
        hit = levels[ri] == e
        first = np.minimum.reduceat(np.where(hit, pos, np.repeat(deg, deg)), cp[nz])
        ex = np.where(first < deg[nz], first + 1, deg[nz])
        total += 8 * int(cand.sum()) + 4 * int(ex[cand[nz]].sum()) + V // 8
    return total


def _check_bfs(g, og, root, threshold):
    from plato_amd.algo import bfs_opts_t, breadth_first_search
    ref = oracle.bfs(og, int(root), threshold=threshold)
    res = breadth_first_search(g, bfs_opts_t(int(root)), push_threshold=threshold)
    assert np.array_equal(res.levels, ref["levels"]), f"root {root} threshold {threshold}"
    assert np.array_equal(res.modes, ref["modes"])
    assert np.array_equal(res.frontier, ref["frontier"])
    assert np.array_equal(res.active_edges, ref["active_edges"])
    assert res.visited == ref["visited"]
    assert res.alg_bytes == _alg_bytes(og, ref["levels"], ref["modes"], ref["frontier"], ref["active_edges"])
    return res


def test_bfs_edge_balanced_push_zero_degree_runs(ctx):
    """Top-down steps over a queue with long runs of zero-degree vertices between
    vertices of 300 out-edges (more than 4096 active edges: the edge-balanced step,
    whose 64-vertex window then holds fewer than 64 edges), push forced."""
    from plato_amd import Graph
    n, hub_every, hub_deg = 20000, 97, 300
    src = [np.zeros(n, np.uint32)]
    dst = [np.arange(1, n + 1, dtype=np.uint32)]
    nxt = n + 1
    for v in range(1, n + 1):
        if v % hub_every == 0:
            src.append(np.full(hub_deg, v, np.uint32))
            dst.append(np.arange(nxt, nxt + hub_deg, dtype=np.uint32))
            nxt += hub_deg
    s, d = np.concatenate(src), np.concatenate(dst)
    V = nxt
    print(f"case: bfs zero-degree runs V={V} E={len(s)}", flush=True)
    g = Graph.from_edges(ctx, s, d, V, build_csr=True)
    og = oracle.CSC(V, s, d, symmetrize=False)
    res = _check_bfs(g, og, 0, 1e9)
    assert res.modes.tolist() == [1, 1, 1] and int(res.active_edges[1]) > 4096


@pytest.mark.parametrize("threshold", [1e9, 0.0])
def test_bfs_forced_direction_rmat(ctx, threshold):
    """Every level top-down (threshold 1e9: edge-balanced steps over hub frontiers) or
    every level bottom-up (threshold 0), directed and symmetrized."""
    from plato_amd import Graph
    for directed in (True, False):
        print(f"case: bfs forced threshold={threshold} directed={directed}", flush=True)
        g = Graph.rmat(ctx, 13, 16, seed=19, directed=directed, build_csr=True)
        s, d = oracle.rmat_edges(13, 16, 19)
        og = oracle.CSC(1 << 13, s, d, symmetrize=not directed)
        cand = np.flatnonzero(og.od > 0)
        for root in [int(cand[0]), int(cand[len(cand) // 2]), int(np.argmax(og.od))]:
            res = _check_bfs(g, og, root, threshold)
            assert set(res.modes.tolist()) == ({1} if threshold > 1 else {0})
