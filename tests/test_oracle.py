"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each test checks an oracle function against something that does not come from
the oracle's own formula: the reference tests' known answers on the reference's
data files, closed forms, dense linear solves of the fixed point, scipy's BFS,
published hash vectors, or algebraic invariants.  CPU only.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import oracle
import oracle.rmat as R
from conftest import golden

D = 0.85


# ---- generator ------------------------------------------------------------------

def test_mix64_published_vector():
    # SplitMix64 from state 0: first outputs (Vigna's reference implementation)
    assert oracle.mix64(0) == 0xE220A8397B1DCDAF
    assert oracle.mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    assert oracle.mix64(2 * 0x9E3779B97F4A7C15 % 2**64) == 0x06C45D188009454F


def test_rmat_c_matches_numpy_twin():
    s, d = oracle.rmat_edges(12, 16, 42)
    s2, d2 = R.edges(12, 16, 42)
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    # an offset window equals the same slice of the full list
    s3, d3 = oracle.rmat_edges(12, 16, 42, e0=1000, n=500)
    assert np.array_equal(s3, s[1000:1500]) and np.array_equal(d3, d[1000:1500])
    w = oracle.rmat_weights(42, 300)
    assert np.array_equal(w, R.weights(42, 0, 300))
    assert w.min() > 0 and w.max() <= 1


def test_rmat_bit_marginals():
    # each level: P(src bit) = c + d = 0.24, P(dst bit) = b + d = 0.24, P(both) = d = 0.05
    scale = 14
    s, d = oracle.rmat_edges(scale, 16, 7)
    n = s.size
    for bit in (0, 5, scale - 1):
        sb = (s >> bit) & 1
        db = (d >> bit) & 1
        tol = 5 * np.sqrt(0.25 / n)
        assert abs(sb.mean() - 0.24) < tol
        assert abs(db.mean() - 0.24) < tol
        assert abs((sb & db).mean() - 0.05) < tol


# ---- structures (reference known answers) ----------------------------------------

def _csv_graph(name, directed, weighted=False):
    if weighted:
        s, d, w = oracle.read_edge_csv(golden(name), weighted=True)
    else:
        s, d = oracle.read_edge_csv(golden(name))
        w = None
    V = int(max(s.max(), d.max())) + 1
    return oracle.CSC(V, s, d, w=w, symmetrize=not directed), s, d


def test_dcsc_graph_10_9():
    # dcsc.cc:138-191: weighted directed chain -> 9 columns, 9 entries
    g, s, d = _csv_graph("graph_10_9.csv", True, weighted=True)
    indeg = g.in_degrees()
    assert g.nnz == 9 and int((indeg > 0).sum()) == 9
    # message_passing.cc:48-88: one message per in-edge -> sum 9, values[0] = 0, others 1
    assert int(indeg.sum()) == 9 and indeg[0] == 0 and (indeg[1:] == 1).all()
    # weights survive the CSV parser's whitespace tolerance (graph_10_9.csv lines 6, 9)
    assert np.isclose(g.w[g.col_ptr[6]], 102.1) and np.isclose(g.w[g.col_ptr[9]], 1.111)


def test_dcsc_undirected_v100():
    # dcsc.cc:81-136: 100 vertices, 4300 entries, every entry an input edge (either way)
    g, s, d = _csv_graph("v100_e2150_ua_c3.csv", False)
    assert g.V == 100 and g.nnz == 4300
    edges = set(zip(s.tolist(), d.tolist()))
    for v in range(g.V):
        for e in range(int(g.col_ptr[v]), int(g.col_ptr[v + 1])):
            u = int(g.row_idx[e])
            assert (u, v) in edges or (v, u) in edges
    # structure.cc:137-151: degrees 19 / 59 / 19
    assert (g.od[0], g.od[99], g.od[28]) == (19, 59, 19)


def test_out_degrees_from_in_graph_directed_v100():
    # structure.cc:205-218 (generate_dense_out_degrees_fg on the DCSC)
    g, s, d = _csv_graph("v100_e4300_da_c3.csv", True)
    assert (g.od[0], g.od[99], g.od[28]) == (19, 59, 19)
    assert np.array_equal(g.od, np.bincount(s, minlength=100).astype(np.uint32))


def test_csc_sorted_and_complete():
    s, d = oracle.rmat_edges(9, 8, 3)
    g = oracle.CSC(512, s, d, symmetrize=True)
    for v in range(g.V):
        col = g.row_idx[int(g.col_ptr[v]):int(g.col_ptr[v + 1])]
        assert (np.diff(col.astype(np.int64)) >= 0).all()
    # multiset of entries = edges + reversals
    ent = sorted(zip(np.repeat(np.arange(g.V), np.diff(g.col_ptr).astype(np.int64)).tolist(),
                     g.row_idx.tolist()))
    exp = sorted(list(zip(d.tolist(), s.tolist())) + list(zip(s.tolist(), d.tolist())))
    assert ent == exp
    # symmetrized CSR is the CSC
    assert np.array_equal(g.row_ptr, g.col_ptr) and np.array_equal(g.col_idx, g.row_idx)


# ---- PageRank ------------------------------------------------------------------------

def _dense_M(g, weighted=False):
    """M[v, u] = sum over entries u->v of (1/od(u) or w/W(u))."""
    V = g.V
    M = np.zeros((V, V))
    cols = np.repeat(np.arange(V), np.diff(g.col_ptr).astype(np.int64))
    if weighted:
        W = np.zeros(V)
        np.add.at(W, g.row_idx, g.w.astype(np.float64))
        np.add.at(M, (cols, g.row_idx), g.w.astype(np.float64) / W[g.row_idx])
    else:
        od = g.od.astype(np.float64)
        np.add.at(M, (cols, g.row_idx), 1.0 / od[g.row_idx])
    return M


def test_pagerank_chain_closed_form():
    g, _, _ = _csv_graph("graph_10_9.csv", True)
    v = np.arange(10)
    for iters in (3, 20):
        r, _ = oracle.pagerank(g, iters, D, 0.0)
        # r_t(v) = 0.15 sum_{i<min(v+1,t)} 0.85^i + [t <= v] 0.85^t
        exp = np.array([0.15 * sum(D ** i for i in range(min(k + 1, iters))) +
                        (D ** iters if iters <= k else 0.0) for k in v])
        np.testing.assert_allclose(r, exp, rtol=1e-14)


def test_pagerank_fixed_point_linear_solve():
    s, d = oracle.rmat_edges(8, 8, 11)
    g = oracle.CSC(256, s, d)
    r, deltas = oracle.pagerank(g, 400, D, 0.0)
    M = _dense_M(g)
    exact = np.linalg.solve(np.eye(g.V) - D * M, (1 - D) * np.ones(g.V))
    np.testing.assert_allclose(r, exact, rtol=1e-10)
    assert deltas[-1] == 0.0 and deltas[0] > deltas[-2]


def test_pagerank_eps_jumps_to_one_final_iteration():
    s, d = oracle.rmat_edges(8, 8, 11)
    g = oracle.CSC(256, s, d)
    r, deltas = oracle.pagerank(g, 100, D, 1e-3)
    k = int(np.argmax(deltas[:-1] < 1e-3))
    # the first delta below eps is followed by exactly one more (final) iteration
    assert len(deltas) == k + 2 and (deltas[:k] >= 1e-3).all()
    # the final state is the undivided rank, i.e. one more step from the stored x
    r_full, _ = oracle.pagerank(g, k + 2, D, 0.0)
    np.testing.assert_array_equal(r, r_full)


def test_pagerank_step_threads_independent():
    s, d = oracle.rmat_edges(12, 16, 5)
    g = oracle.CSC(4096, s, d)
    x = np.where(g.od > 0, 1.0 / np.maximum(g.od, 1), 1.0)
    y1, d1 = oracle.pagerank_step(g, x, D, False, nthreads=1)
    y4, d4 = oracle.pagerank_step(g, x, D, False, nthreads=4)
    assert np.array_equal(y1, y4) and d1 == d4


def test_weighted_pagerank_chain_and_solve():
    g, _, _ = _csv_graph("graph_10_9.csv", True, weighted=True)
    r, _ = oracle.weighted_pagerank(g, 20, D, 0.0)
    gu, _, _ = _csv_graph("graph_10_9.csv", True)
    ru, _ = oracle.pagerank(gu, 20, D, 0.0)
    np.testing.assert_allclose(r, ru, rtol=1e-14)   # one out-edge per source: beta = 1
    s, d = oracle.rmat_edges(8, 8, 13)
    w = oracle.rmat_weights(13, s.size)
    g = oracle.CSC(256, s, d, w=w)
    r, _ = oracle.weighted_pagerank(g, 400, D, 0.0)
    exact = np.linalg.solve(np.eye(g.V) - D * _dense_M(g, weighted=True), (1 - D) * np.ones(g.V))
    np.testing.assert_allclose(r, exact, rtol=1e-10)


# ---- PPR -----------------------------------------------------------------------------

def test_ppr_chain_closed_form():
    g, _, _ = _csv_graph("graph_10_9.csv", True)
    r, _ = oracle.ppr(g, [0], 20, D, 0.0)
    np.testing.assert_allclose(r, 0.15 * D ** np.arange(10), rtol=1e-14)


def test_ppr_linear_solve_and_linearity():
    s, d = oracle.rmat_edges(8, 8, 17)
    g = oracle.CSC(256, s, d)
    M = _dense_M(g)
    seeds = [3, 40, 40, 200]   # duplicate counts once
    r, _ = oracle.ppr(g, seeds, 400, D, 0.0)
    b = np.zeros(g.V)
    b[[3, 40, 200]] = (1 - D) / 3
    exact = np.linalg.solve(np.eye(g.V) - D * M, b)
    np.testing.assert_allclose(r, exact, rtol=1e-9, atol=1e-15)
    # linear in the seed set, at a finite iteration count too
    parts = [oracle.ppr(g, [v], 30, D, 0.0)[0] for v in (3, 40, 200)]
    multi, _ = oracle.ppr(g, [3, 40, 200], 30, D, 0.0)
    np.testing.assert_allclose(multi, sum(parts) / 3, rtol=1e-12, atol=1e-16)


def test_ppr_pushpull_equals_pull_and_switch_is_sticky():
    s, d = oracle.rmat_edges(10, 16, 19)
    g = oracle.CSC(1024, s, d)
    src = int(np.argmax((g.od > 0) & (g.od < 3)))
    pull, dp = oracle.ppr(g, [src], 30, D, 0.0)
    pp, dpp, modes = oracle.ppr_pushpull(g, src, 30, D, 0.0)
    assert np.array_equal(pull, pp) and np.array_equal(dp, dpp)
    assert modes[0] == 1                      # a low-degree source starts in push
    first_pull = int(np.argmin(modes))
    assert modes[first_pull] == 0 and (modes[first_pull:] == 0).all()


# ---- BFS -------------------------------------------------------------------------------

def _scipy_levels(g, root):
    V = g.V
    cols = np.repeat(np.arange(V), np.diff(g.col_ptr).astype(np.int64))
    A = sp.csr_matrix((np.ones(g.nnz), (g.row_idx, cols)), shape=(V, V))  # u -> v
    dist = csg.shortest_path(A, directed=True, unweighted=True, indices=root)
    return np.where(np.isinf(dist), -1, dist).astype(np.int32)


@pytest.mark.parametrize("directed", [True, False])
def test_bfs_levels_match_scipy(directed):
    s, d = oracle.rmat_edges(11, 16, 23)
    g = oracle.CSC(2048, s, d, symmetrize=not directed)
    roots = np.flatnonzero(g.od > 0)[[0, 7, 300]]
    for root in roots:
        b = oracle.bfs(g, int(root))
        lv = _scipy_levels(g, int(root))
        assert np.array_equal(b["levels"], lv)
        # frontier sizes are the level histogram; visited is the reached count
        hist = np.bincount(lv[lv >= 0])
        assert np.array_equal(b["frontier"], hist.astype(np.uint64))
        assert b["visited"] == int((lv >= 0).sum())
        # decision per epoch from the level sets alone (dualmode.hpp:279)
        for ep in range(b["epochs"]):
            A = int(g.od[lv == ep].sum())
            assert b["active_edges"][ep] == A
            assert b["modes"][ep] == int(A / g.edges_prime < 0.05)


def test_bfs_tree_and_chain():
    g, _, _ = _csv_graph("complete_binary_tree_10_9.csv", True)
    b = oracle.bfs(g, 0)
    assert np.array_equal(b["levels"], np.floor(np.log2(np.arange(10) + 1)).astype(np.int32))
    g, _, _ = _csv_graph("graph_10_9.csv", True)
    b = oracle.bfs(g, 0)
    assert np.array_equal(b["levels"], np.arange(10, dtype=np.int32)) and b["visited"] == 10


def test_topk_tie_break():
    r = np.array([0.5, 0.9, 0.5, 0.1, 0.9, 0.5])
    ids, vals = oracle.topk(r, 4)
    order = np.lexsort((np.arange(r.size), -r))[:4]   # rank desc, id asc
    assert np.array_equal(ids, order) and np.array_equal(vals, r[order])


# ---- edge-list ingestion (parsers.hpp:225-293) ---------------------------------------------------
def test_parse_csv_reference_files():
    """The reference's data files parse to their known contents (same edges as the
    line-by-line reader the other pins use; counts from dcsc.cc / structure.cc)."""
    for name, n in (("graph_10_9.csv", 9), ("v100_e2150_ua_c3.csv", 2150), ("v100_e4300_da_c3.csv", 4300),
                    ("complete_binary_tree_10_9.csv", 9)):
        s, d = oracle.parse_csv(open(golden(name), "rb").read())
        s2, d2 = oracle.read_edge_csv(golden(name))
        assert s.size == n and np.array_equal(s, s2) and np.array_equal(d, d2)
    s, d, w = oracle.parse_csv(open(golden("graph_10_9.csv"), "rb").read(), weighted=True)
    assert np.array_equal(s, np.arange(9)) and np.array_equal(d, np.arange(1, 10))
    assert w[0] == np.float32(0.123) and w[1] == np.float32(1.779)


def test_parse_csv_strtok_strtoul_rules():
    """Hand-worked lines: strtok_r on ',' (repeated and leading commas collapse), strtoul
    (leading blanks, sign, digits up to the first other byte), empty and short lines skipped,
    CRLF, last line without a newline."""
    s, d = oracle.parse_csv(b" 1, 2\n\n,,3,,4,\r\n5\n+6,-0\n7,8xyz\n9,10")
    assert s.tolist() == [1, 3, 6, 7, 9] and d.tolist() == [2, 4, 0, 8, 10]
    s, d, w = oracle.parse_csv(b"1,2,0.5\n3,4\n5,6,abc\n7,8, 1e-3\n", weighted=True)
    assert s.tolist() == [1, 5, 7] and w.tolist() == [0.5, 0.0, np.float32(1e-3)]
    assert oracle.parse_csv(b"")[0].size == 0
    for bad, line in ((b"1,2\n4294967296,3\n", 1), (b"4294967296\n", 0), (b"1,-1\n", 0)):
        with pytest.raises(ValueError, match=f"line {line}"):
            oracle.parse_csv(bad)


def test_format_csv_round_trip():
    s, d = oracle.rmat_edges(12, 8, 2)
    s2, d2 = oracle.parse_csv(oracle.format_csv(s, d))
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    assert oracle.format_csv(np.array([0, 4294967295], np.uint32), np.array([10, 7], np.uint32)) == \
        b"0,10\n4294967295,7\n"


# ---- k-core (kcore.hpp:80-268) ---------------------------------------------------------------------
def _simple_undirected(s, d):
    """Distinct undirected pairs without self-loops, as src < dst arrays."""
    a = np.minimum(s, d).astype(np.uint64)
    b = np.maximum(s, d).astype(np.uint64)
    keep = a != b
    k = np.unique((a[keep] << np.uint64(32)) | b[keep])
    return (k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def test_kcore_matches_networkx_core_number():
    """On simple undirected graphs the reference's h-index iteration converges to the
    textbook coreness (networkx core_number, Batagelj-Zaversnik peeling)."""
    import networkx as nx
    cases = [oracle.rmat_edges(10, 8, 4), oracle.rmat_edges(12, 16, 5),
             oracle.read_edge_csv(golden("v100_e2150_ua_c3.csv")), oracle.read_edge_csv(golden("kcore_test_7_11.csv"))]
    for s, d in cases:
        a, b = _simple_undirected(s, d)
        V = int(max(s.max(), d.max())) + 1
        core, mod, nxt = oracle.kcore(oracle.CSC(V, a, b, symmetrize=True))
        G = nx.Graph()
        G.add_nodes_from(range(V))
        G.add_edges_from(zip(a.tolist(), b.tolist()))
        ref = nx.core_number(G)
        assert np.array_equal(core, np.array([ref[v] for v in range(V)], np.uint32))
        assert nxt[-1] == 0 and np.all(nxt[:-1] > 0) and mod[-1] == 0


def test_kcore_hand_worked_cases():
    # triangle 0-1-2 with a pendant 3 on vertex 2: cores 2, 2, 2, 1
    s = np.array([0, 1, 2, 2], np.uint32)
    d = np.array([1, 2, 0, 3], np.uint32)
    core, mod, nxt = oracle.kcore(oracle.CSC(4, s, d, symmetrize=True))
    assert core.tolist() == [2, 2, 2, 1]
    # directed, in-edges: 0->2, 1->2, 2->3, 4->3, 5->4.  Epoch 0: 0, 1 and 5 have no
    # in-edges (-> 0); vertex 3 (in-neighbours 2 with estimate 2, 4 with estimate 1)
    # drops to 1 and activates 2 and 4; 2 and 4 keep 2 and 1 (their in-neighbours start
    # at 0xFFFFFFFF).  Epoch 1: the in-neighbours of 2 and 4 now have estimate 0, so no
    # i >= 1 qualifies and both keep their value; nothing is modified, the loop ends.
    s = np.array([0, 1, 2, 4, 5], np.uint32)
    d = np.array([2, 2, 3, 3, 4], np.uint32)
    g = oracle.CSC(6, s, d)
    core, mod, nxt = oracle.kcore(g)
    assert core.tolist() == [0, 0, 2, 1, 1, 0]
    assert mod.tolist() == [1, 0] and nxt.tolist() == [2, 0]
    # the same edges read as out-adjacency (a BCSR "incomings", as kcore_simple passes):
    # 3 has none (-> 0); 2's only out-neighbour 3 starts at 0xFFFFFFFF, so 2 keeps 1.
    core, mod, nxt = oracle.kcore(g, out_edges=True)
    assert core.tolist() == [1, 1, 1, 0, 1, 1] and mod.tolist() == [0] and nxt.tolist() == [0]


# ---- connected components (cgm/connected_component.hpp) --------------------------------------------
def _bfs_labels_like_reference_test(s, d, V):
    """cgm/test/test_cgm.cc:83-104: undirected adjacency, BFS from each unlabelled id in
    ascending order, every reached vertex labelled with the root."""
    adj = [[] for _ in range(V)]
    for a, b in zip(s.tolist(), d.tolist()):
        adj[a].append(b)
        adj[b].append(a)
    label = [-1] * V
    for i in range(V):
        if label[i] != -1:
            continue
        label[i] = i
        q = [i]
        while q:
            u = q.pop(0)
            for v in adj[u]:
                if label[v] == -1:
                    label[v] = i
                    q.append(v)
    return np.array(label, np.uint32)


def test_cc_labels_match_reference_test_and_scipy():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    s, d = oracle.read_edge_csv(golden("v100_e2150_ua_c3.csv"))
    g = oracle.CSC(100, s, d, symmetrize=True)
    lab = oracle.cc_labels(g)
    assert np.array_equal(lab, _bfs_labels_like_reference_test(s, d, 100))
    assert len(set(lab.tolist())) == 3  # the file's three components
    for scale in (10, 14):
        s, d = oracle.rmat_edges(scale, 4, 3)
        V = 1 << scale
        lab = oracle.cc_labels(oracle.CSC(V, s, d, symmetrize=True))
        n, comp = connected_components(coo_matrix((np.ones(s.size), (s, d)), shape=(V, V)), directed=False)
        assert len(set(lab.tolist())) == n
        # the label is the smallest id of the component
        mins = {}
        for v, c in enumerate(comp.tolist()):
            mins.setdefault(c, v)
        assert np.array_equal(lab, np.array([mins[c] for c in comp.tolist()], np.uint32))


def test_cc_labels_directed_reachability():
    """Directed: the smallest id that reaches v, by brute force over reachability."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import shortest_path
    s, d = oracle.rmat_edges(7, 2, 5)
    V = 1 << 7
    lab = oracle.cc_labels(oracle.CSC(V, s, d))
    reach = np.isfinite(shortest_path(coo_matrix((np.ones(s.size), (s, d)), shape=(V, V)), directed=True,
                                      unweighted=True))
    exp = np.array([np.flatnonzero(reach[:, v])[0] for v in range(V)], np.uint32)
    assert np.array_equal(lab, exp)


# ---- HyperANF (hyperanf.hpp:80-197, hyperloglog.hpp, hash.hpp) ----------------------------------------
def test_murmur3_published_vectors():
    """MurmurHash3_x86_32's published test vectors (empty input, every tail length, seeds)."""
    vec = [(b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xffffffff, 0x81F16F39),
           (b"\xff\xff\xff\xff", 0, 0x76293B50), (b"!Ce\x87", 0, 0xF55B516B), (b"!Ce\x87", 0x5082EDEE, 0x2362F9DE),
           (b"!Ce", 0, 0x7E4A8634), (b"!C", 0, 0xA0F7B07A), (b"!", 0, 0x72661CF4),
           (b"\x00\x00\x00\x00", 0, 0x2362F9DE), (b"\x00\x00\x00", 0, 0x85F0B427), (b"\x00\x00", 0, 0x30F4C306),
           (b"\x00", 0, 0x514E28B7), (b"Hello, world!", 0x9747b28c, 0x24884CBA),
           (b"The quick brown fox jumps over the lazy dog", 0x9747b28c, 0x2FA826CD), (b"aaaa", 0x9747b28c, 0x5A97808A)]
    for data, seed, exp in vec:
        assert oracle.murmur3_32(data, seed) == exp, (data, seed)


def _hll_of(ids, bits):
    """hyperloglog_t::add over the ids (4-byte little-endian, seed 0x5f61767a)."""
    regs = np.zeros(1 << bits, np.uint8)
    for v in ids:
        h = oracle.murmur3_32(int(v).to_bytes(4, "little"), 0x5f61767a)
        idx = h >> (32 - bits)
        w = (h << bits) & 0xFFFFFFFF
        rank = 1
        while rank <= 32 - bits and not (w & 0x80000000):
            rank += 1
            w = (w << 1) & 0xFFFFFFFF
        regs[idx] = max(regs[idx], rank)
    return regs


def test_hyperanf_counters_are_hll_of_balls():
    """After t iterations a vertex's counter is the HLL of the vertices that reach it in
    at most t steps (register-wise max is the union of sketches)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import shortest_path
    s, d = oracle.rmat_edges(6, 4, 9)
    V = 1 << 6
    g = oracle.CSC(V, s, d)
    dist = shortest_path(coo_matrix((np.ones(s.size), (s, d)), shape=(V, V)), directed=True, unweighted=True)
    for t in (0, 1, 3):
        anfs, _, regs = oracle.hyperanf(g, bits=6, iters=t)
        if len(anfs) <= t:  # stopped early: nothing changes any more
            t = len(anfs) - 1
        for v in range(V):
            ball = np.flatnonzero(dist[:, v] <= t)
            assert np.array_equal(regs[v], _hll_of(ball, 6)), (t, v)
    # one element: one register set, so the small-range estimate m * ln(m / (m - 1))
    r = _hll_of([7], 6)
    assert abs(oracle.hll_estimate(r) - 64 * np.log(64 / 63)) < 1e-12


def test_hyperanf_reference_test_average_distance():
    """hyperanf/test/test_hyperanf.cc:133-219: raw_graph_7_7.csv undirected, bits 6, 20
    iterations; the estimate is within 0.05 of the exact BFS average distance."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import shortest_path
    s, d = oracle.read_edge_csv(golden("raw_graph_7_7.csv"))
    V = int(max(s.max(), d.max())) + 1
    anfs, avg, _ = oracle.hyperanf(oracle.CSC(V, s, d, symmetrize=True), bits=6, iters=20)
    dist = shortest_path(coo_matrix((np.ones(s.size), (s, d)), shape=(V, V)), directed=False, unweighted=True)
    real = 0.0
    for root in range(V):  # number_of_friends + get_avg_distance_by_src of the test
        dd = dist[root]
        mx = int(dd[np.isfinite(dd)].max())
        friends = np.array([np.sum(dd <= i) for i in range(mx + 1)], np.float64)
        real += sum(i * (friends[i] - friends[i - 1]) for i in range(1, mx + 1)) / (friends[-1] - friends[0])
    real /= V
    assert abs(real - avg) <= 0.05, (real, avg)


# ---- parallel structure builder, edge digests, the one-pass scan ---------------------------

def _lexsort_csc(V, s, d, w=None, symmetrize=False):
    """Independent restatement of the CSC: numpy's stable lexsort of the entries by
    (destination, source), reversals appended for undirected input."""
    if symmetrize:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
        w = None if w is None else np.concatenate([w, w])
    order = np.lexsort((s, d))
    ptr = np.zeros(V + 1, np.uint64)
    ptr[1:] = np.cumsum(np.bincount(d, minlength=V))
    return ptr, s[order], (None if w is None else w[order])


@pytest.mark.parametrize("symmetrize", [False, True])
def test_csc_and_csr_equal_stable_lexsort(symmetrize):
    s, d = oracle.rmat_edges(12, 8, 29)
    w = oracle.rmat_weights(29, s.size)
    g = oracle.CSC(1 << 12, s, d, w=w, symmetrize=symmetrize)
    ptr, idx, ww = _lexsort_csc(1 << 12, s, d, w, symmetrize)
    assert np.array_equal(g.col_ptr, ptr) and np.array_equal(g.row_idx, idx)
    assert np.array_equal(g.w, ww)  # equal (dst, src) keep input order: weights pin stability
    rp, ci, _ = _lexsort_csc(1 << 12, d, s, None, symmetrize)
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, ci)
    # V not a power of two, empty leading and trailing columns
    g2 = oracle.CSC(5000, s + 3, d + 3)
    p2, i2, _ = _lexsort_csc(5000, s + 3, d + 3)
    assert np.array_equal(g2.col_ptr, p2) and np.array_equal(g2.row_idx, i2)


def test_csc_independent_of_thread_count():
    s, d = oracle.rmat_edges(13, 8, 31)
    ref = oracle.CSC(1 << 13, s, d, symmetrize=True)
    g = oracle.CSC(1 << 13, s, d)
    x = np.ones(g.V)
    for t in (1, 3):
        oracle.pagerank_step(g, x, D, False, 0, 1, nthreads=t)  # sets the OpenMP team size
        h = oracle.CSC(1 << 13, s, d, symmetrize=True)
        assert np.array_equal(h.col_ptr, ref.col_ptr) and np.array_equal(h.row_idx, ref.row_idx)
    oracle.pagerank_step(g, x, D, False, 0, 1, nthreads=0)


def test_edges_digest_matches_numpy_twin_and_csc():
    s, d = oracle.rmat_edges(11, 16, 37)
    twin = int(np.sum(R.mix64((s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)),
                      dtype=np.uint64))
    assert oracle.edges_digest(s, d) == twin
    g = oracle.CSC(1 << 11, s, d)
    assert oracle.csc_digest(g.V, g.col_ptr, g.row_idx) == twin
    # through a relabeling map: internal ids v' = rank_of[v], map = perm (internal -> input)
    perm = np.random.default_rng(1).permutation(g.V).astype(np.uint32)
    rank_of = np.argsort(perm).astype(np.uint32)
    h = oracle.CSC(g.V, rank_of[s], rank_of[d])
    assert oracle.csc_digest(h.V, h.col_ptr, h.row_idx, perm) == twin
    # a dropped edge, a moved endpoint and a reversed edge each change it
    assert oracle.edges_digest(s[1:], d[1:]) != twin
    d2 = d.copy()
    d2[5] ^= 1
    assert oracle.edges_digest(s, d2) != twin
    k = int(np.flatnonzero(s != d)[0])
    s3, d3 = s.copy(), d.copy()
    s3[k], d3[k] = d[k], s[k]
    assert oracle.edges_digest(s3, d3) != twin


def test_rmat_scan_matches_full_generation():
    scale, ef, seed = 12, 16, 41
    ranges = [(0, 40), (1000, 1100), (4000, 4096)]
    sc = oracle.RmatScan(scale, ef, seed, ranges, weighted=True)
    s, d = oracle.rmat_edges(scale, ef, seed)
    w = oracle.rmat_weights(seed, s.size)
    assert np.array_equal(sc.od, np.bincount(s, minlength=1 << scale).astype(np.uint32))
    np.testing.assert_allclose(sc.W, np.bincount(s, weights=w.astype(np.float64), minlength=1 << scale),
                               rtol=1e-13)
    assert sc.digest == oracle.edges_digest(s, d)
    keep = np.zeros(s.size, bool)
    for lo, hi in ranges:
        keep |= (d >= lo) & (d < hi)
    assert np.array_equal(sc.eids, np.flatnonzero(keep).astype(np.uint64))
    ptr, idx, ww = _lexsort_csc(1 << scale, s[keep], d[keep], w[keep])
    assert np.array_equal(sc.slice.col_ptr, ptr) and np.array_equal(sc.slice.row_idx, idx)
    assert np.array_equal(sc.slice.w, ww)


def test_weighted_pagerank_step_equals_one_iteration():
    s, d = oracle.rmat_edges(10, 16, 43)
    w = oracle.rmat_weights(43, s.size)
    g = oracle.CSC(1 << 10, s, d, w=w)
    r1, _ = oracle.weighted_pagerank(g, 1, D, 0.0)
    W = np.bincount(g.row_idx, weights=g.w.astype(np.float64), minlength=g.V)
    y = oracle.weighted_pagerank_step(g, W, np.ones(g.V), D)
    np.testing.assert_allclose(y, r1, rtol=1e-14)
    # two iterations: the step applied to the first iteration's state
    r2, _ = oracle.weighted_pagerank(g, 2, D, 0.0)
    np.testing.assert_allclose(oracle.weighted_pagerank_step(g, W, r1, D), r2, rtol=1e-14)


# ---- the push/pull PPR switch, pinned by hop balls ---------------------------------------------

def _expected_ppr_modes(g, src, iters):
    """personalized_pagerank.hpp:188-205: push while the accumulated active set (every
    vertex within t out-hops of src at epoch t) has active edges <= E'/20 (integer
    division), then pull for good.  Hop distances from scipy, independent of the oracle."""
    V = g.V
    cols = np.repeat(np.arange(V), np.diff(g.col_ptr).astype(np.int64))
    A = sp.csr_matrix((np.ones(g.nnz), (g.row_idx, cols)), shape=(V, V))  # u -> v
    dist = csg.shortest_path(A, directed=True, unweighted=True, indices=src)
    modes, pull = [], False
    for t in range(iters):
        if not pull:
            active_edges = int(g.od[dist <= t].sum())
            pull = active_edges > g.edges_prime // 20
        modes.append(0 if pull else 1)
    return np.array(modes, np.uint8)


@pytest.mark.parametrize("directed", [True, False])
def test_ppr_pushpull_switch_follows_hop_balls(directed):
    s, d = oracle.rmat_edges(11, 16, 53)
    g = oracle.CSC(1 << 11, s, d, symmetrize=not directed)
    cand = np.flatnonzero(g.od > 0)
    order = cand[np.argsort(g.od[cand], kind="stable")]
    srcs = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    seen = set()
    for src in srcs:
        _, _, modes = oracle.ppr_pushpull(g, src, 12, D, 0.0)
        exp = _expected_ppr_modes(g, src, 12)
        assert np.array_equal(modes, exp), (src, modes, exp)
        seen.add(int(np.argmin(exp)) if (exp == 0).any() else 99)
    assert len(seen) >= 2  # the sources switch at different epochs


def test_ppr_pushpull_switch_boundary():
    """A_t == E'/20 stays in push; the first A_t above it pulls (strict '>')."""
    # 40 edges: a source 0 with two out-edges to 1 and 2, vertex 1 with 3 out-edges,
    # the rest a chain among vertices 10..45 that the source never reaches
    s = [0, 0, 1, 1, 1] + list(range(10, 45))
    d = [1, 2, 3, 4, 5] + list(range(11, 46))
    g = oracle.CSC(46, np.array(s, np.uint32), np.array(d, np.uint32))
    assert g.edges_prime // 20 == 2
    _, _, modes = oracle.ppr_pushpull(g, 0, 4, D, 0.0)
    # epoch 0: A = od(0) = 2, not above 2 -> push; epoch 1: A = 2 + 3 + 0 = 5 -> pull
    assert modes.tolist() == [1, 0, 0, 0]
    assert np.array_equal(modes, _expected_ppr_modes(g, 0, 4))


# ---- n-step degrees (nstepdegrees.hpp:166-489) -------------------------------------------------

def test_nstep_degrees_reference_known_answers():
    """test_nstepdegrees.cc: raw_graph_7_7.csv, directed, step 1, bits 6, type both:
    in(0) = 3, out(0) = 1, in(4) = 1, out(4) = 1."""
    s, d = oracle.read_edge_csv(golden("raw_graph_7_7.csv"))
    V = int(max(s.max(), d.max())) + 1
    g = oracle.CSC(V, s, d)
    din, _ = oracle.nstep_degrees(g, step=1, bits=6, direction="in")
    dout, _ = oracle.nstep_degrees(g, step=1, bits=6, direction="out")
    assert (din[0], dout[0], din[4], dout[4]) == (3, 1, 1, 1)


def _ball_hll_degree(g, v, step, bits, direction):
    """From scratch: the set of vertices within `step` hops of v (against the edges for
    in-degrees), each id hashed into a fresh HyperLogLog, estimate - 1."""
    V = g.V
    ptr, idx = (g.col_ptr, g.row_idx) if direction == "in" else (g.row_ptr, g.col_idx)
    ball, frontier = {v}, {v}
    for _ in range(step):
        nxt = set()
        for u in frontier:
            nxt.update(int(w) for w in idx[int(ptr[u]):int(ptr[u + 1])])
        nxt -= ball
        ball |= nxt
        frontier = nxt
        if not frontier:
            break
    m = 1 << bits
    regs = np.zeros(m, np.uint8)
    for u in ball:
        h = oracle.murmur3_32(int(u).to_bytes(4, "little"), 0x5F61767A)
        r, w = 1, (h << bits) & 0xFFFFFFFF
        while r <= 32 - bits and not (w & 0x80000000):
            r += 1
            w = (w << 1) & 0xFFFFFFFF
        regs[h >> (32 - bits)] = max(regs[h >> (32 - bits)], r)
    return int(oracle.hll_estimate(regs)) - 1


@pytest.mark.parametrize("direction", ["in", "out"])
def test_nstep_degrees_equal_hll_of_balls(direction):
    s, d = oracle.rmat_edges(9, 4, 61)
    g = oracle.CSC(1 << 9, s, d)
    for step in (1, 2, 3):
        deg, _ = oracle.nstep_degrees(g, step=step, bits=7, direction=direction)
        for v in (0, 1, 5, 77, 300, 511):
            assert deg[v] == _ball_hll_degree(g, v, step, 7, direction), (step, v)
    # unbounded steps: the closure; the loop stops once nothing grows
    deg, n = oracle.nstep_degrees(g, step=-1, bits=7, direction=direction)
    assert n < 64
    for v in (0, 300):
        assert deg[v] == _ball_hll_degree(g, v, 1 << 20, 7, direction)
