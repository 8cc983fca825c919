"""GPU: connected components (connected_component_t::compute,
cgm/connected_component.hpp:170-234) against the oracle.

The labels are integer results and must be bit-exact: every vertex gets the smallest
input id that reaches it (the smallest id of its component on a symmetrized graph),
whatever the push/pull sequence and the order in which labels spread.  The per-epoch
activity is not compared: it depends on that order, in the reference too."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def test_reference_test_case(ctx):
    """cgm/test/test_cgm.cc: v100_e2150_ua_c3.csv, undirected, three components."""
    from plato_amd import Graph
    from plato_amd.algo import connected_components
    s, d = oracle.read_edge_csv(golden("v100_e2150_ua_c3.csv"))
    og = oracle.CSC(100, s, d, symmetrize=True)
    for relabel in (False, True):
        g = Graph.from_edges(ctx, s, d, 100, directed=False, relabel=relabel)
        res = connected_components(g)
        assert np.array_equal(res.labels, oracle.cc_labels(og))
        assert res.components == 3 and res.modes[0] == 0  # epoch 0: all active, pull
        g.close()


@pytest.mark.parametrize("scale,ef,directed,csr,relabel", [
    (12, 2, False, True, False), (14, 4, False, True, True), (16, 16, False, True, False),
    (12, 2, True, True, False), (14, 1, True, True, True), (14, 2, True, False, False)])
def test_rmat_labels(ctx, scale, ef, directed, csr, relabel):
    """Sparse RMAT graphs (edge factor 1-4) leave many components and long chains, so
    the late epochs push; directed graphs propagate along edge direction only."""
    from plato_amd import Graph
    from plato_amd.algo import connected_components
    print(f"case: cc rmat s{scale} ef{ef} directed={directed} csr={csr} relabel={relabel}", flush=True)
    s, d = oracle.rmat_edges(scale, ef, 13)
    g = Graph.rmat(ctx, scale, ef, seed=13, directed=directed, build_csr=csr, relabel=relabel)
    og = oracle.CSC(1 << scale, s, d, symmetrize=not directed)
    res = connected_components(g)
    exp = oracle.cc_labels(og)
    assert np.array_equal(res.labels, exp)
    assert res.components == len(np.unique(exp))
    if csr and not directed:
        assert 1 in res.modes.tolist()  # the sparse tail ran in push mode
    if not csr:
        assert set(res.modes.tolist()) == {0}
    g.close()


def test_large_symmetrized(ctx):
    from plato_amd import Graph
    from plato_amd.algo import connected_components
    s, d = oracle.rmat_edges(20, 16, 2)
    g = Graph.rmat(ctx, 20, 16, seed=2, directed=False, relabel=True)
    res = connected_components(g)
    exp = oracle.cc_labels(oracle.CSC(1 << 20, s, d, symmetrize=True))
    assert np.array_equal(res.labels, exp)
    g.close()
