"""GPU: k-core (kcore_algo_t::compute_shell_index, kcore.hpp:80-268) against the oracle.

Integer work: coreness and the per-epoch log (modified vertices, size of the next
active set) must be bit-exact.  The iteration is Jacobi, so neither depends on the
order in which the device visits vertices; relabeling the graph must not change them
either."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def _check(g, og, out_edges=False):
    from plato_amd.algo import kcore
    res = kcore(g, out_edges=out_edges)
    core, mod, nxt = oracle.kcore(og, out_edges=out_edges)
    assert np.array_equal(res.coreness, core)
    assert np.array_equal(res.modified, mod) and np.array_equal(res.next_active, nxt)
    assert res.epochs == len(mod)
    return res


def test_hand_worked(ctx):
    from plato_amd import Graph
    s = np.array([0, 1, 2, 2], np.uint32)
    d = np.array([1, 2, 0, 3], np.uint32)
    g = Graph.from_edges(ctx, s, d, 4, directed=False)
    assert _check(g, oracle.CSC(4, s, d, symmetrize=True)).coreness.tolist() == [2, 2, 2, 1]
    s = np.array([0, 1, 2, 4, 5], np.uint32)
    d = np.array([2, 2, 3, 3, 4], np.uint32)
    g = Graph.from_edges(ctx, s, d, 6, build_csr=True)
    og = oracle.CSC(6, s, d)
    assert _check(g, og).coreness.tolist() == [0, 0, 2, 1, 1, 0]
    assert _check(g, og, out_edges=True).coreness.tolist() == [1, 1, 1, 0, 1, 1]


def test_reference_fixtures(ctx):
    from plato_amd import Graph
    for name in ("v100_e2150_ua_c3.csv", "v100_e4300_da_c3.csv", "graph_10_9.csv",
                 "complete_binary_tree_10_9.csv", "kcore_test_7_11.csv"):
        s, d = oracle.read_edge_csv(golden(name))
        V = int(max(s.max(), d.max())) + 1
        for directed in (True, False):
            print(f"case: kcore {name} directed={directed}", flush=True)
            g = Graph.from_edges(ctx, s, d, V, directed=directed, build_csr=True)
            og = oracle.CSC(V, s, d, symmetrize=not directed)
            _check(g, og)
            if directed:
                _check(g, og, out_edges=True)
            g.close()


@pytest.mark.parametrize("scale,directed", [(12, True), (14, False), (16, False), (18, False)])
def test_rmat_bit_exact(ctx, scale, directed):
    """RMAT graphs with their duplicates and self-loops; hubs go through the wave-per-vertex
    pass (and, at scale 18, its search above the 1024-bin histogram in the first epochs)."""
    from plato_amd import Graph
    print(f"case: kcore rmat s{scale} directed={directed}", flush=True)
    s, d = oracle.rmat_edges(scale, 16, 7)
    g = Graph.rmat(ctx, scale, 16, seed=7, directed=directed, build_csr=True)
    og = oracle.CSC(1 << scale, s, d, symmetrize=not directed)
    res = _check(g, og)
    if directed:
        _check(g, og, out_edges=True)
    assert res.coreness.max() > 0
    g.close()
    gr = Graph.rmat(ctx, scale, 16, seed=7, directed=directed, relabel=True)
    assert np.array_equal(_check(gr, og).coreness, res.coreness)
    gr.close()


def test_big_vertex_search_above_histogram(ctx):
    """A clique of 1100 vertices (coreness 1099, above the wave pass's 1024 histogram
    bins, so the block pass takes them) plus a tail."""
    from plato_amd import Graph
    n = 1100
    a, b = np.triu_indices(n, 1)
    s = np.concatenate([a, [0, n, n + 1]]).astype(np.uint32)
    d = np.concatenate([b, [n, n + 1, n + 2]]).astype(np.uint32)
    V = n + 3
    g = Graph.from_edges(ctx, s, d, V, directed=False)
    res = _check(g, oracle.CSC(V, s, d, symmetrize=True))
    assert res.coreness[:n].tolist() == [n - 1] * n and res.coreness[n:].tolist() == [1, 1, 1]


def test_giant_hub(ctx):
    """A vertex of 70000 neighbours (above the 65536 at which its adjacency is cut into
    chunks whose histograms add up in global memory), on a wheel whose rim carries random
    chords, so the hub's h-index moves over several epochs; directed and relabeled too."""
    from plato_amd import Graph
    rim = 70000
    rng = np.random.default_rng(8)
    r = np.arange(1, rim + 1, dtype=np.uint32)
    chord_a = rng.integers(1, rim + 1, 3 * rim).astype(np.uint32)
    chord_b = rng.integers(1, rim + 1, 3 * rim).astype(np.uint32)
    s = np.concatenate([np.zeros(rim, np.uint32), r, chord_a])
    d = np.concatenate([r, r % rim + 1, chord_b])
    V = rim + 1
    print(f"case: kcore giant hub V={V} E={len(s)}", flush=True)
    og = oracle.CSC(V, s, d, symmetrize=True)
    assert int(np.diff(og.col_ptr)[0]) > 65536
    res = _check(Graph.from_edges(ctx, s, d, V, directed=False), og)
    assert res.coreness[0] >= 3
    _check(Graph.from_edges(ctx, s, d, V, directed=False, relabel=True), og)
    # directed: the hub's 70000 in-edges come from the rim's reversed edges
    s2, d2 = np.concatenate([s, r]), np.concatenate([d, np.zeros(rim, np.uint32)])
    _check(Graph.from_edges(ctx, s2, d2, V, build_csr=True), oracle.CSC(V, s2, d2))
