"""reverse() of the dual-mode engine on directed graphs (dualmode.hpp:151-153, 299-322):
after reverse() a pull sums over out-edges and a push follows in-edges, while the
push/pull choice keeps the graph's own out-degrees (dualmode.hpp:250-278).

Checked against a float64 sum over the generator's edge list and the oracle's BFS over
the reversed edge list, with and without degree relabeling, in push-only, pull-only
and switching runs."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _graph(ctx, scale, seed, relabel, build_csr=False):
    from plato_amd import Graph
    return Graph.rmat(ctx, scale, 8, seed=seed, relabel=relabel, build_csr=build_csr)


@pytest.mark.parametrize("relabel", [False, True])
def test_reversed_pull_sums_over_out_edges(ctx, relabel):
    from plato_amd.engine import dualmode_engine_t
    scale, seed = 12, 3
    V = 1 << scale
    s, d = oracle.rmat_edges(scale, 8, seed)
    g = _graph(ctx, scale, seed, relabel)
    eng = dualmode_engine_t(g)
    p = g.permutation()  # internal id of every input id
    x = np.random.default_rng(1).random(V)
    xs, ys = eng.alloc_v_state(np.float64), eng.alloc_v_state(np.float64)
    xi = np.zeros(V)
    xi[p] = x
    xs.upload(xi)
    eng.reverse()
    assert eng.is_reversed()
    eng.pull_sum(xs, ys)
    back = np.zeros(V)  # y[u] = sum of x[v] over the edges u -> v
    np.add.at(back, s, x[d])
    assert np.allclose(ys.numpy()[p], back, rtol=1e-12, atol=0)
    eng.reverse()  # forward again: y[v] = sum of x[u] over u -> v
    assert not eng.is_reversed()
    eng.pull_sum(xs, ys)
    fwd = np.zeros(V)
    np.add.at(fwd, d, x[s])
    assert np.allclose(ys.numpy()[p], fwd, rtol=1e-12, atol=0)
    assert not np.allclose(back, fwd)
    eng.close()
    g.close()


@pytest.mark.parametrize("relabel", [False, True])
def test_reversed_search_follows_in_edges(ctx, relabel):
    from plato_amd.engine import BfsEngineState, dualmode_engine_t
    scale, seed = 13, 4
    V = 1 << scale
    s, d = oracle.rmat_edges(scale, 8, seed)
    og_rev = oracle.CSC(V, d, s)  # the reversed edges, with their CSR
    g = _graph(ctx, scale, seed, relabel, build_csr=True)
    od = g.out_degrees().astype(np.float64)  # forward out-degrees, input order
    E = float(s.size)
    cand = np.flatnonzero(og_rev.od > 0)  # vertices with in-edges
    order = cand[np.argsort(og_rev.od[cand], kind="stable")]
    roots = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for threshold in (0.05, 1e9, 0.0):
        eng = dualmode_engine_t(g, push_threshold=threshold)
        eng.reverse()
        st = BfsEngineState(g)
        for root in roots:
            print(f"  root {root} threshold {threshold} relabel {relabel}", flush=True)
            cur, nxt = eng.alloc_v_subset(), eng.alloc_v_subset()
            st.reset(root, cur)
            modes = []
            while True:
                push, activated = eng.foreach_edges_bfs(st, cur, nxt)
                modes.append(push)
                cur, nxt = nxt, cur
                if activated == 0:
                    break
            lv = st.levels()
            ref = oracle.bfs(og_rev, root, threshold)["levels"]
            assert np.array_equal(lv, ref), (root, threshold)
            # the choice of each epoch: forward out-degrees over that epoch's frontier
            want = [bool(od[lv == e].sum() / E < threshold) for e in range(int(lv.max()) + 1)]
            assert modes == want, (modes, want)
            cur.close()
            nxt.close()
        st.close()
        eng.close()
    g.close()


def test_bind_needs_the_same_ids(ctx):
    from plato_amd._lib import PlatoHipError
    from plato_amd.engine import BfsEngineState
    g = _graph(ctx, 10, 5, True, build_csr=True)
    st = BfsEngineState(g)
    other = g.reverse(build_csr=True)  # input ids, renumbered afresh: not the same ids
    with pytest.raises(PlatoHipError):
        st.bind(other)
    same = g.reverse(build_csr=True, same_ids=True)
    st.bind(same)
    st.bind(g)
    with pytest.raises(PlatoHipError):
        g.reverse(relabel=True, same_ids=True)
    st.close()
    other.close()
    same.close()
    g.close()
