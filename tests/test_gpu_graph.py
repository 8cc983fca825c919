"""GPU: device graph construction, engine state and bitmap operators against the
oracle and the reference tests' known answers (bit-exact: integer work)."""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def _csv(name, weighted=False):
    if weighted:
        s, d, w = oracle.read_edge_csv(golden(name), weighted=True)
        return s, d, w
    s, d = oracle.read_edge_csv(golden(name))
    return s, d, None


def test_device_rmat_generator_matches_oracle(ctx):
    print("case: rmat generator scale 14 seed 99", flush=True)
    s, d, w = ctx.rmat_edges(14, 99, 0, 1 << 16, weights=True)
    so, do = oracle.rmat_edges(14, 16, 99, e0=0, n=1 << 16)
    assert np.array_equal(s, so) and np.array_equal(d, do)
    assert np.array_equal(w, oracle.rmat_weights(99, 1 << 16))
    s2, d2 = ctx.rmat_edges(27, 5, 123456789, 4)
    so2, do2 = oracle.rmat_edges(27, 16, 5, e0=123456789, n=4)
    assert np.array_equal(s2, so2) and np.array_equal(d2, do2)


@pytest.mark.parametrize("directed", [True, False])
def test_rmat_csc_csr_match_oracle(ctx, directed):
    from plato_amd import Graph
    print(f"case: rmat s11 csc/csr directed={directed}", flush=True)
    g = Graph.rmat(ctx, 11, 16, seed=3, directed=directed, build_csr=True)
    s, d = oracle.rmat_edges(11, 16, 3)
    og = oracle.CSC(2048, s, d, symmetrize=not directed)
    cp, ri = g.download_csc()
    assert np.array_equal(cp, og.col_ptr) and np.array_equal(ri, og.row_idx)
    rp, ci = g.download_csr()
    assert np.array_equal(rp, og.row_ptr) and np.array_equal(ci, og.col_idx)
    assert np.array_equal(g.out_degrees(), og.od)
    assert np.array_equal(g.in_degrees(), og.in_degrees())
    gi = g.graph_info()
    present = (og.od > 0) | (og.in_degrees() > 0)
    assert gi.vertices_ == int(present.sum()) and gi.max_v_i_ == int(np.flatnonzero(present).max())
    assert gi.edges_ == s.size and gi.is_directed_ == directed


def test_fixture_structures(ctx):
    from plato_amd import Graph
    # dcsc.cc:138-191 and message_passing.cc:48-88 on graph_10_9
    s, d, w = _csv("graph_10_9.csv", weighted=True)
    g = Graph.from_edges(ctx, s, d, 10, weights=w)
    indeg = g.in_degrees()
    assert g.nnz == 9 and int((indeg > 0).sum()) == 9
    assert int(indeg.sum()) == 9 and indeg[0] == 0 and (indeg[1:] == 1).all()
    # structure.cc:137-151 / 205-218 degrees on the v100 files
    s, d, _ = _csv("v100_e2150_ua_c3.csv")
    g = Graph.from_edges(ctx, s, d, 100, directed=False)
    assert g.V == 100 and g.nnz == 4300
    od = g.out_degrees()
    assert (od[0], od[99], od[28]) == (19, 59, 19)
    s, d, _ = _csv("v100_e4300_da_c3.csv")
    g = Graph.from_edges(ctx, s, d, 100, build_csr=True)
    od = g.out_degrees()
    assert (od[0], od[99], od[28]) == (19, 59, 19)
    og = oracle.CSC(100, s, d)
    cp, ri = g.download_csc()
    assert np.array_equal(cp, og.col_ptr) and np.array_equal(ri, og.row_idx)


def test_weights_normalised_per_source(ctx):
    from plato_amd import Graph
    s, d = oracle.rmat_edges(10, 8, 31)
    w = oracle.rmat_weights(31, s.size)
    g = Graph.from_edges(ctx, s, d, 1024, weights=w)
    cp, ri, wn = g.download_csc(weights=True)
    og = oracle.CSC(1024, s, d, w=w)
    assert np.array_equal(ri, og.row_idx)
    W = np.zeros(1024)
    np.add.at(W, og.row_idx, og.w.astype(np.float64))
    exp = (og.w.astype(np.float64) / W[og.row_idx]).astype(np.float32)
    # the source sums are taken in a different order: allow one FP32 ulp
    np.testing.assert_allclose(wn, exp, rtol=2.5e-7)
    # device rmat weights equal host-list weights
    g2 = Graph.rmat(ctx, 10, 8, seed=31, weighted=True)
    _, ri2, wn2 = g2.download_csc(weights=True)
    assert np.array_equal(ri2, ri) and np.array_equal(wn2, wn)


def test_bitmap_known_answers(ctx):
    from plato_amd import Bitmap
    # bitmap.cc:31-50 (msb)
    b = Bitmap(ctx, 40988)
    b.set_bit(0)
    assert b.msb() == 0
    b.set_bit(13)
    assert b.msb() == 13
    b.set_bit(1982)
    assert b.msb() == 1982
    b.clear()
    b.set_bit(40987)
    assert b.msb() == 40987
    b.set_bit(0)
    b.set_bit(40986)
    assert b.msb() == 40987
    # bitmap.cc:52-103 (traversal in ascending order)
    b.clear()
    for v in (0, 19, 40987):
        b.set_bit(v)
    assert b.to_vector().tolist() == [0, 19, 40987]
    assert b.count() == 3 and b.get_bit(19) == 1 and b.get_bit(20) == 0
    b.clr_bit(19)
    assert b.to_vector().tolist() == [0, 40987]
    # fill masks the tail (bitmap.hpp:168-179); message_passing.cc:90-123 full set = 10
    f = Bitmap(ctx, 10)
    f.fill()
    assert f.count() == 10 and f.words_numpy()[0] == (1 << 10) - 1
    c = Bitmap(ctx, 10)
    c.copy_from(f)
    assert c.count() == 10
    sp = Bitmap(ctx, 10)
    for v in (0, 5, 9):
        sp.set_bit(v)
    assert sp.count() == 3 and sp.to_vector().tolist() == [0, 5, 9]


def test_spread_message_known_answers(ctx):
    """message_passing.cc:90-169: every active vertex sends itself, the sink counts per
    vertex and returns 1: 10 over the full set of graph_10_9.csv's vertices, 3 over
    {0, 5, 9}, one count at each of those."""
    from plato_amd import Bitmap, DenseState, spread_message
    full = Bitmap(ctx, 10)
    full.fill()
    counts = DenseState(ctx, 10, np.uint32)
    counts.fill(0)
    assert spread_message(full, counts) == 10
    assert counts.numpy().tolist() == [1] * 10
    sp = Bitmap(ctx, 10)
    for v in (0, 5, 9):
        sp.set_bit(v)
    counts.fill(0)
    assert spread_message(sp, counts) == 3
    c = counts.numpy()
    assert c[[0, 5, 9]].tolist() == [1, 1, 1] and int(c.sum()) == 3


@pytest.mark.parametrize("dtype", [np.uint32, np.float64, np.float32])
def test_spread_message_keys_and_values(ctx, dtype):
    """A random active set over 2^20 + 37 ids (a partial last word), messages sent to
    key[v] < 5000: counts against numpy's bincount, sums of value[v] against numpy."""
    from plato_amd import Bitmap, DenseState, spread_message
    n, m = (1 << 20) + 37, 5000
    rng = np.random.default_rng(4)
    act = rng.random(n) < 0.3
    words = np.zeros((n >> 6) + 1, np.uint64)
    ids = np.flatnonzero(act)
    np.bitwise_or.at(words, ids >> 6, np.left_shift(np.uint64(1), (ids & 63).astype(np.uint64)))
    b = Bitmap(ctx, n)
    b.upload_words(words)
    keys = rng.integers(0, m, n).astype(np.uint32)
    kd = DenseState(ctx, n, np.uint32)
    kd.upload(keys)
    sink = DenseState(ctx, m, dtype)
    sink.fill(0)
    if dtype == np.uint32:
        assert spread_message(b, sink, key=kd) == ids.size
        assert np.array_equal(sink.numpy(), np.bincount(keys[ids], minlength=m).astype(np.uint32))
        return
    vals = rng.random(n).astype(dtype)
    vd = DenseState(ctx, n, dtype)
    vd.upload(vals)
    assert spread_message(b, sink, key=kd, value=vd) == ids.size
    want = np.bincount(keys[ids], weights=vals[ids].astype(np.float64), minlength=m)
    np.testing.assert_allclose(sink.numpy().astype(np.float64), want,
                               rtol=1e-12 if dtype == np.float64 else 2e-5)


def test_spread_message_target_out_of_range(ctx):
    from plato_amd import Bitmap, DenseState, spread_message
    from plato_amd._lib import PlatoHipError
    b = Bitmap(ctx, 100)
    b.fill()
    sink = DenseState(ctx, 50, np.uint32)
    sink.fill(0)
    with pytest.raises(PlatoHipError, match="beyond the sink"):
        spread_message(b, sink)


def test_dense_state_ops(ctx):
    from plato_amd import DenseState
    x = DenseState(ctx, 1000, np.float32)
    x.fill(0.5)
    assert x.sum() == 500.0
    x[7] = 2.0
    v = x.numpy()
    assert v[7] == 2.0 and v[6] == 0.5
    y = DenseState(ctx, 1000, np.float64)
    y.upload(np.arange(1000, dtype=np.float64))
    assert y.sum() == 499500.0 and np.array_equal(y.to_f64(), np.arange(1000.0))
    z = DenseState(ctx, 1000, np.float64)
    z.copy_from(y)
    assert np.array_equal(z.numpy(), np.arange(1000.0))
    odd = DenseState(ctx, 1001, np.float32)      # length not a multiple of 16 bytes
    src = DenseState(ctx, 1001, np.float32)
    src.upload(np.arange(1001, dtype=np.float32))
    odd.copy_from(src)
    assert np.array_equal(odd.numpy(), np.arange(1001, dtype=np.float32))


def test_active_edges_and_mode_choice(ctx):
    from plato_amd import Graph, dualmode_engine_t
    s, d = oracle.rmat_edges(11, 16, 41)
    g = Graph.rmat(ctx, 11, 16, seed=41, build_csr=True)
    og = oracle.CSC(2048, s, d)
    eng = dualmode_engine_t(g)
    act = eng.alloc_v_subset()
    rng = np.random.default_rng(0)
    chosen = rng.choice(2048, 300, replace=False)
    for v in chosen:
        act.set_bit(int(v))
    A = int(og.od[chosen].sum())
    assert eng.active_edges(act) == A
    push, a2 = eng.choose_push(act)
    assert a2 == A and push == (A / og.edges_prime < 0.05)


def test_errors_are_status_codes(ctx):
    from plato_amd import Graph
    from plato_amd._lib import PlatoHipError
    with pytest.raises(PlatoHipError):
        Graph.from_edges(ctx, np.array([0, 5], np.uint32), np.array([1, 2], np.uint32), 4)
    g = Graph.rmat(ctx, 8, 4, seed=1)
    with pytest.raises(PlatoHipError):
        g.download_csr()          # no CSR was built
    with pytest.raises(PlatoHipError):
        g.set_tile(3000)


def test_autotune_choices_are_valid_and_keep_results(ctx, tmp_path):
    from plato_amd import Graph
    from plato_amd.algo import pagerank, pagerank_opts_t
    g = Graph.rmat(ctx, 12, 16, seed=8, build_csr=True, relabel=True)
    before = pagerank(g, pagerank_opts_t(10, 0.85, 0.0), precision=64).rank
    r = g.autotune(pull=True, batch=True, bfs=True)
    assert r["tile_items"] in (1024, 2048, 4096) and g.info.tile_items == r["tile_items"]
    assert r["batch_task_edges"] in (256, 1024, 4096, 16384)
    assert r["bfs_push_threshold"] in (0.01, 0.02, 0.05, 0.1, 0.2)
    assert r["pull_ms"] > 0 and r["batch_ms"] > 0 and r["bfs_ms"] > 0
    after = pagerank(g, pagerank_opts_t(10, 0.85, 0.0), precision=64).rank
    np.testing.assert_allclose(after, before, rtol=1e-12)
    import ctypes as C
    from plato_amd import _lib as L
    res = L.TuneResult(**{k: v for k, v in r.items() if k != "batch_plan_ms"})
    path = str(tmp_path / "tune.txt").encode()
    assert g.lib.pe_autotune_save(g.h, C.byref(res), path) == 0
    out, found = L.TuneResult(), C.c_int()
    assert g.lib.pe_autotune_load(g.h, path, C.byref(out), C.byref(found)) == 0
    assert found.value == 1 and out.tile_items == r["tile_items"]
