"""GPU: HyperANF (hyperanf.hpp:80-197) against the oracle.

The counters are integer (register-wise max of HyperLogLog sketches) and must be
bit-exact; the neighbourhood-function estimates are FP64 sums of per-vertex estimates
whose order differs from the oracle's, compared at relative 1e-12, and the reference
test's known answer (average distance within 0.05 of the exact BFS value) must hold."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def _run(g, og, bits, iters):
    from plato_amd.algo import hyperanf, hyperanf_opts_t
    res = hyperanf(g, hyperanf_opts_t(iters, bits), registers=True)
    anfs, avg, regs = oracle.hyperanf(og, bits=bits, iters=iters)
    assert np.array_equal(res.registers, regs)
    assert res.anfs.size == anfs.size
    np.testing.assert_allclose(res.anfs, anfs, rtol=1e-12)
    assert abs(res.avg_distance - avg) <= 1e-9 * abs(avg)
    return res


def test_reference_test_case(ctx):
    """hyperanf/test/test_hyperanf.cc:133-219 on raw_graph_7_7.csv (undirected, bits 6)."""
    from plato_amd import Graph
    s, d = oracle.read_edge_csv(golden("raw_graph_7_7.csv"))
    V = int(max(s.max(), d.max())) + 1
    g = Graph.from_edges(ctx, s, d, V, directed=False)
    res = _run(g, oracle.CSC(V, s, d, symmetrize=True), 6, 20)
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import shortest_path
    dist = shortest_path(coo_matrix((np.ones(s.size), (s, d)), shape=(V, V)), directed=False, unweighted=True)
    real = 0.0
    for root in range(V):
        dd = dist[root]
        mx = int(dd.max())
        friends = np.array([np.sum(dd <= i) for i in range(mx + 1)], np.float64)
        real += sum(i * (friends[i] - friends[i - 1]) for i in range(1, mx + 1)) / (friends[-1] - friends[0])
    assert abs(real / V - res.avg_distance) <= 0.05


@pytest.mark.parametrize("scale,bits,directed,relabel", [
    (12, 6, True, False), (12, 7, False, True), (12, 8, True, True), (10, 12, False, False), (8, 14, True, False),
    # symmetrized scale 14 has a vertex of 5561 in-edges: the hub pass (chunks of in-edges
    # merged with compare-and-swap), for rows of 16, 32, 64 and 128 words
    (14, 6, False, True), (14, 7, False, False), (14, 8, False, True), (14, 9, False, False)])
def test_rmat(ctx, scale, bits, directed, relabel):
    from plato_amd import Graph
    print(f"case: hyperanf rmat s{scale} bits={bits} directed={directed} relabel={relabel}", flush=True)
    s, d = oracle.rmat_edges(scale, 8, 21)
    g = Graph.rmat(ctx, scale, 8, seed=21, directed=directed, relabel=relabel)
    _run(g, oracle.CSC(1 << scale, s, d, symmetrize=not directed), bits, 20)
    g.close()
