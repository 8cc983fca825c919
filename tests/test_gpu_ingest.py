"""GPU: CSV ingestion on the device (pe_graph_create_csv) against the oracle's
restatement of the reference parser (orc_parse_csv: strtok_r / strtoul / strtof).

Bar: integer work is bit-exact -- the device-built CSC equals the oracle CSC of the
oracle-parsed edges; weights must equal strtof bit for bit, checked end to end (the
weighted PageRank of the device-parsed graph equals that of the graph built from the
oracle-parsed weights, bitwise, since the engine is deterministic)."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu

TRICKY = b" 1, 2\n\n,,3,,4,\r\n5\n+6,-0\n7,8xyz\n9,10"


def _same_csc(g, s, d, V, directed=True):
    og = oracle.CSC(V, s, d, symmetrize=not directed)
    cp, ri = g.download_csc()
    return np.array_equal(cp, og.col_ptr) and np.array_equal(ri, og.row_idx)


def test_reference_files_and_tricky_lines(ctx):
    from plato_amd import Graph
    for name in ("graph_10_9.csv", "v100_e2150_ua_c3.csv", "v100_e4300_da_c3.csv",
                 "complete_binary_tree_10_9.csv"):
        data = open(golden(name), "rb").read()
        s, d = oracle.parse_csv(data)
        for directed in (True, False):
            print(f"case: csv {name} directed={directed}", flush=True)
            g = Graph.from_csv_bytes(ctx, data, directed=directed)
            V = int(max(s.max(), d.max())) + 1
            assert g.V == V and g.info.edges == s.size
            assert _same_csc(g, s, d, V, directed)
            g.close()
    s, d = oracle.parse_csv(TRICKY)
    g = Graph.from_csv_bytes(ctx, TRICKY)
    assert g.info.edges == 5 and _same_csc(g, s, d, 11)


def test_errors_name_the_line(ctx):
    from plato_amd import Graph
    from plato_amd._lib import PlatoHipError
    for text, line in ((b"1,2\n4294967296,3\n", 2), (b"4294967296\n", 1), (b"1,-1\n", 1)):
        with pytest.raises(PlatoHipError, match=f"line {line}"):
            Graph.from_csv_bytes(ctx, text)
    for text in (b"1,2,inf\n", b"1,2,0x1p3\n", b"1,2,1.2345678901234567\n", b"1,2,1e-40\n"):
        with pytest.raises(PlatoHipError, match="line 1"):
            Graph.from_csv_bytes(ctx, text, weighted=True)
    with pytest.raises(PlatoHipError):
        Graph.from_csv_bytes(ctx, b"\n\n5\n")  # no edges


def test_weights_match_strtof(ctx):
    """Weights written with %.9g (RMAT weights), plus exact float midpoints that strtof
    rounds to even (16777217, 16777219) and plain literals, parse bit-for-bit like
    strtof: the weighted PageRank of both graphs is identical."""
    from plato_amd import Graph
    from plato_amd.algo import pagerank_opts_t, weighted_pagerank
    s, d = oracle.rmat_edges(12, 16, 3)
    w = oracle.rmat_weights(3, s.size).astype(np.float64)
    extra = ["16777217", "16777219", "0.5", "1e-3", "  2.50", "+7", "123456789012345", "0.000001",
             "3.4028234e38", "1.17549435e-38"]
    lines = [f"{a},{b},{x:.9g}" for a, b, x in zip(s, d, w)]
    lines += [f"{i},{(i * 7) % 4096},{lit}" for i, lit in enumerate(extra)]
    data = ("\n".join(lines) + "\n").encode()
    ps, pd, pw = oracle.parse_csv(data, weighted=True)
    assert pw[-10] == np.float32(16777216) and pw[-9] == np.float32(16777220)
    V = int(max(ps.max(), pd.max())) + 1
    g1 = Graph.from_csv_bytes(ctx, data, weighted=True)
    g2 = Graph.from_edges(ctx, ps, pd, V, weights=pw)
    assert g1.V == V and _same_csc(g1, ps, pd, V)
    r1 = weighted_pagerank(g1, pagerank_opts_t(20, 0.85, 0.0), precision=64).rank
    r2 = weighted_pagerank(g2, pagerank_opts_t(20, 0.85, 0.0), precision=64).rank
    assert np.array_equal(r1, r2)
    c1 = g1.download_csc(weights=True)[2]
    c2 = g2.download_csc(weights=True)[2]
    assert np.array_equal(c1, c2)


def test_encode_ids(ctx):
    """64-bit ids compacted to 0..n-1 in ascending order; edges map back exactly."""
    from plato_amd import Graph
    big = [10 ** 12, 5, 10 ** 18, 77, 2 ** 63 + 1]
    rng = np.random.default_rng(1)
    pairs = [(big[rng.integers(5)], big[rng.integers(5)]) for _ in range(200)]
    data = "".join(f"{a},{b}\n" for a, b in pairs).encode()
    g = Graph.from_csv_bytes(ctx, data, encode_ids=True, build_csr=True)
    ids = g.vertex_ids()
    used = sorted({x for p in pairs for x in p})
    assert ids.tolist() == used and g.V == len(used)
    cp, ri = g.download_csc()
    cols = np.repeat(np.arange(g.V), np.diff(cp).astype(np.int64))
    got = sorted(zip(ids[ri].tolist(), ids[cols].tolist()))
    assert got == sorted(pairs)


def test_rmat_csv_matches_generated_graph(ctx):
    """RMAT scale 16 written as text (with CRLF line ends) and parsed on the device gives
    the same CSC as the graph built from the generated arrays."""
    from plato_amd import Graph
    s, d = oracle.rmat_edges(16, 16, 5)
    data = ("\r\n".join(f"{a},{b}" for a, b in zip(s, d)) + "\r\n").encode()
    g = Graph.from_csv_bytes(ctx, data, build_csr=True)
    V = int(max(s.max(), d.max())) + 1
    assert g.V == V and g.info.edges == s.size
    assert _same_csc(g, s, d, V)
