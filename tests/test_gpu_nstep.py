"""GPU: n-step degrees (plato/algo/nstepdegrees/nstepdegrees.hpp:166-489) against the
reference test's known answers and the oracle (bit-exact: integer register maxima, the
same FP64 estimate)."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def test_nstep_reference_known_answers(ctx):
    """test_nstepdegrees.cc: raw_graph_7_7.csv, directed, step 1, bits 6, type both."""
    from plato_amd import Graph
    from plato_amd.algo import nstep_degrees, nstepdegree_opts_t
    s, d = oracle.read_edge_csv(golden("raw_graph_7_7.csv"))
    V = int(max(s.max(), d.max())) + 1
    g = Graph.from_edges(ctx, s, d, V, build_csr=True)
    r = nstep_degrees(g, nstepdegree_opts_t(step=1, type="both", is_directed=True, bits=6))
    assert (r.in_degrees[0], r.out_degrees[0], r.in_degrees[4], r.out_degrees[4]) == (3, 1, 1, 1)


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("bits", [6, 8, 12])
def test_nstep_rmat_matches_oracle(ctx, directed, bits):
    from plato_amd import Graph
    from plato_amd.algo import nstep_degrees, nstepdegree_opts_t
    scale = 12
    s, d = oracle.rmat_edges(scale, 8, 71)
    og = oracle.CSC(1 << scale, s, d, symmetrize=not directed)
    for relabel in (False, True):
        print(f"case: nstep rmat s{scale} directed={directed} bits={bits} relabel={relabel}", flush=True)
        g = Graph.rmat(ctx, scale, 8, seed=71, directed=directed, build_csr=True, relabel=relabel)
        for step in (1, 3, -1):
            r = nstep_degrees(g, nstepdegree_opts_t(step=step, type="both", is_directed=directed, bits=bits))
            din, _ = oracle.nstep_degrees(og, step=step, bits=bits, direction="in")
            dout, _ = oracle.nstep_degrees(og, step=step, bits=bits, direction="out")
            assert np.array_equal(r.in_degrees, din), (step, relabel)
            assert np.array_equal(r.out_degrees, dout), (step, relabel)
        g.close()


def test_nstep_targets_only(ctx):
    from plato_amd import Graph
    from plato_amd.algo import nstep_degrees, nstepdegree_opts_t
    s, d = oracle.rmat_edges(12, 8, 73)
    og = oracle.CSC(1 << 12, s, d)
    g = Graph.rmat(ctx, 12, 8, seed=73, build_csr=True, relabel=True)
    targets = [0, 5, 17, 4095]
    r = nstep_degrees(g, nstepdegree_opts_t(step=2, type="in", is_directed=True, bits=7), targets=targets)
    din, _ = oracle.nstep_degrees(og, step=2, bits=7, direction="in")
    assert r.out_degrees is None
    want = np.zeros_like(din)
    want[targets] = din[targets]
    assert np.array_equal(r.in_degrees, want)
