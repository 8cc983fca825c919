"""Host C++ layer (plato_amd/include/plato/...): it compiles with g++ on the CPU,
and on the GPU the C++ API test and the example drivers agree with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "plato_amd", "bin")


def test_host_programs_build():
    r = subprocess.run(["make", "-s", "-f", "plato_amd/host/Makefile", f"ROOT={ROOT}"], cwd=ROOT,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for p in ("pagerank_simple", "bfs_simple", "personalized_pagerank", "trustrank", "kcore_simple", "cgm_simple", "hyperanf", "nstepdegrees",
              "test_plato_api"):
        assert os.access(os.path.join(BIN, p), os.X_OK)


def test_deferred_pull_call_sequences(tmp_path):
    """The engine's deferred pull against a call-logging double of the C-ABI (CPU only)."""
    exe = str(tmp_path / "test_deferred_pull")
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", f"-I{ROOT}/plato_amd/include",
                        os.path.join(ROOT, "tests", "cpp", "test_deferred_pull.cc"),
                        os.path.join(ROOT, "tests", "cpp", "abi_call_log.cc"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failure(s)" in r.stdout


def test_unsupported_functor_is_a_compile_error():
    src = ('#include "plato/engine/dualmode.hpp"\n'
           'int main(){ plato::graph_info_t gi(true);\n'
           ' auto g = std::make_shared<plato::dcsc_t<plato::empty_t>>();\n'
           ' plato::dualmode_engine_t<plato::dcsc_t<plato::empty_t>, std::nullptr_t> e(g, nullptr, gi);\n'
           ' e.foreach_edges<double,int>([](int){}, [](int){ return 0; }); }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-",
                        f"-I{ROOT}/include", f"-I{ROOT}/plato_amd/include"], input=src, text=True,
                       capture_output=True)
    assert r.returncode != 0 and "foreach_edges (pull)" in r.stderr


@pytest.mark.gpu
def test_cpp_api_known_answers():
    r = subprocess.run([os.path.join(BIN, "test_plato_api"), GOLDEN], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("  OK  ") == 11


@pytest.mark.gpu
def test_pagerank_driver_output_matches_oracle(tmp_path):
    s, d = oracle.read_edge_csv(os.path.join(GOLDEN, "v100_e4300_da_c3.csv"))
    r = subprocess.run([os.path.join(BIN, "pagerank_simple"), "--input",
                        os.path.join(GOLDEN, "v100_e4300_da_c3.csv"), "--is_directed=true",
                        "--iterations", "20", "--eps", "0", "--precision", "64",
                        "--output", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "0000_0.csv", delimiter=",")
    ref, _ = oracle.pagerank(oracle.CSC(100, s, d), 20, 0.85, 0.0)
    assert np.array_equal(got[:, 0], np.arange(100))
    np.testing.assert_allclose(got[:, 1], ref, rtol=1e-12)


@pytest.mark.gpu
def test_pagerank_driver_binned_layout(tmp_path):
    """--pull_layout binned (FP32): normalised L1 <= 2e-7 against the FP64 oracle."""
    s, d = oracle.read_edge_csv(os.path.join(GOLDEN, "v100_e4300_da_c3.csv"))
    r = subprocess.run([os.path.join(BIN, "pagerank_simple"), "--input",
                        os.path.join(GOLDEN, "v100_e4300_da_c3.csv"), "--is_directed=true",
                        "--iterations", "20", "--eps", "0", "--precision", "32",
                        "--pull_layout", "binned", "--output", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "0000_0.csv", delimiter=",")
    ref, _ = oracle.pagerank(oracle.CSC(100, s, d), 20, 0.85, 0.0)
    assert np.abs(got[:, 1] - ref).sum() / ref.sum() <= 2e-7


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_ppr_driver_forms_match_oracle(tmp_path, precision):
    """Both PPR forms of the C++ API (the reference's control flow over the engine's
    operators: pull-only, and push/pull with the sticky switch) on a directed RMAT-11
    from a CSV file, against the oracle's pull PPR.  FP64: relative 1e-10 (the push steps
    add with FP64 atomics, so the last bits vary); FP32: L1 <= 1e-6."""
    s, d = oracle.rmat_edges(11, 8, 3)
    V = int(max(s.max(), d.max())) + 1
    path = tmp_path / "g.csv"
    path.write_bytes(oracle.format_csv(s, d))
    og = oracle.CSC(V, s, d)
    src = int(np.argmax(og.od))
    ref, _ = oracle.ppr(og, [src], 40, 0.85, 0.0)
    for form in ("pull", "pushpull"):
        out = tmp_path / form
        out.mkdir()
        r = subprocess.run([os.path.join(BIN, "personalized_pagerank"), "--input", str(path), "--is_directed=true",
                            "--src", str(src), "--iterations", "40", "--eps", "0", "--type", form,
                            "--precision", str(precision), "--output", str(out)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = np.loadtxt(out / "0000_0.csv", delimiter=",")
        assert np.array_equal(got[:, 0], np.arange(V))
        if precision == 64:
            np.testing.assert_allclose(got[:, 1], ref, rtol=1e-10, atol=1e-18)
        else:
            assert np.abs(got[:, 1] - ref).sum() <= 1e-6, form


@pytest.mark.gpu
def test_bfs_driver_levels(tmp_path):
    path = os.path.join(GOLDEN, "v100_e2150_ua_c3.csv")
    s, d = oracle.read_edge_csv(path)
    r = subprocess.run([os.path.join(BIN, "bfs_simple"), "--input", path, "--root", "7",
                        "--output", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "0000_0.csv", delimiter=",", dtype=np.int64)
    ref = oracle.bfs(oracle.CSC(100, s, d, symmetrize=True), 7)
    assert np.array_equal(got[:, 1], ref["levels"])
    assert f"visited: {ref['visited']}" in r.stdout


@pytest.mark.gpu
def test_trustrank_driver_matches_oracle(tmp_path):
    """select_method 1 (PageRank seeds) on a directed RMAT scale-10 graph written as CSV,
    good = even ids: ranks equal the oracle's seed-set PPR over the oracle's top-30 good
    seeds (FP64, rtol 1e-12).  The test requires a gap of more than 1e-12 relative
    between consecutive top ranks, so summation order cannot reorder the seeds."""
    s, d = oracle.rmat_edges(10, 16, 4)
    path = tmp_path / "edges.csv"
    path.write_text("".join(f"{a},{b}\n" for a, b in zip(s.tolist(), d.tolist())))
    V = int(max(s.max(), d.max())) + 1
    g = oracle.CSC(V, s, d)
    pr, _ = oracle.pagerank(g, 100, 0.85, 0.001)
    ids, vals = oracle.topk(pr, 31)
    assert np.all(np.abs(np.diff(vals)) > 1e-12 * vals[0])
    seeds = [int(v) for v in ids[:30] if v % 2 == 0]
    goods = tmp_path / "goods.csv"
    goods.write_text("".join(f"{v}\n" for v in range(0, V, 2)))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([os.path.join(BIN, "trustrank"), "--input_edges", str(path), "--input_goods", str(goods),
                        "--is_directed=true", "--seed_num", "30", "--select_method", "1",
                        "--iterations", "30", "--eps", "0", "--output", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out / "0000_0.csv", delimiter=",")
    ref, _ = oracle.ppr(g, seeds, 30, 0.85, 0.0)
    assert np.array_equal(got[:, 0], np.arange(V))
    np.testing.assert_allclose(got[:, 1], ref, rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
def test_kcore_driver_matches_oracle(tmp_path):
    """kcore_simple reads its input as a BCSR over incoming edges, like the reference
    driver: vertex mode equals the oracle's coreness over the in-edges, and subgraph mode
    writes the in-adjacency of the vertices of each coreness under <k>_core/."""
    path = os.path.join(GOLDEN, "v100_e4300_da_c3.csv")
    s, d = oracle.read_edge_csv(path)
    g = oracle.CSC(100, s, d)
    core, _, _ = oracle.kcore(g)
    vdir = tmp_path / "v"
    vdir.mkdir()
    r = subprocess.run([os.path.join(BIN, "kcore_simple"), "--input", path, "--is_directed=true",
                        "--type", "vertex", "--output", str(vdir)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(vdir / "0000_0.csv", delimiter=",", dtype=np.int64)
    assert np.array_equal(got[:, 0], np.arange(100)) and np.array_equal(got[:, 1], core)
    sdir = tmp_path / "s"
    sdir.mkdir()
    r = subprocess.run([os.path.join(BIN, "kcore_simple"), "--input", path, "--is_directed=true",
                        "--output", str(sdir)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for k in range(int(core.max()) + 1):
        f = sdir / f"{k}_core" / "0000_0.csv"
        lines = f.read_text().split()
        exp = [f"{v},{u}" for v in range(100) if core[v] == k
               for u in g.row_idx[int(g.col_ptr[v]):int(g.col_ptr[v + 1])]]
        assert lines == exp


@pytest.mark.gpu
def test_cgm_driver_outputs(tmp_path):
    """cgm_simple: all_vertices writes `v,label` (the oracle's labels); the default
    sub_graph_by_label writes the largest component's edges, src < dst."""
    path = os.path.join(GOLDEN, "v100_e2150_ua_c3.csv")
    s, d = oracle.read_edge_csv(path)
    g = oracle.CSC(100, s, d, symmetrize=True)
    lab = oracle.cc_labels(g)
    r = subprocess.run([os.path.join(BIN, "cgm_simple"), "--input", path, "--output", str(tmp_path / "v"),
                        "--output_method", "all_vertices"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "There are 3 components." in r.stdout
    got = np.loadtxt(tmp_path / "v" / "part_0000_0.csv", delimiter=",", dtype=np.int64)
    assert np.array_equal(got[:, 0], np.arange(100)) and np.array_equal(got[:, 1], lab)
    r = subprocess.run([os.path.join(BIN, "cgm_simple"), "--input", path, "--output", str(tmp_path / "c")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    major = np.bincount(lab).argmax()
    lines = (tmp_path / "c" / "part_0000_0.csv").read_text().split()
    exp = [f"{int(u)},{v}" for v in range(100) for u in g.row_idx[int(g.col_ptr[v]):int(g.col_ptr[v + 1])]
           if lab[u] == major and u < v]
    assert lines == exp


@pytest.mark.gpu
def test_hyperanf_driver_reference_case():
    """The reference test's case through the driver: raw_graph_7_7.csv, bits 6, 20
    iterations; the printed average distance is the oracle's (FP64, 1e-9)."""
    path = os.path.join(GOLDEN, "raw_graph_7_7.csv")
    s, d = oracle.read_edge_csv(path)
    V = int(max(s.max(), d.max())) + 1
    _, avg, _ = oracle.hyperanf(oracle.CSC(V, s, d, symmetrize=True), bits=6, iters=20)
    r = subprocess.run([os.path.join(BIN, "hyperanf"), "--input", path, "--bits", "6"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = float(r.stdout.split("avg_distance: ")[1].split(",")[0])
    assert abs(got - avg) <= 1e-9 * abs(avg)


@pytest.mark.gpu
def test_nstepdegrees_driver_matches_oracle(tmp_path):
    """example/nstepdegrees.cc's command line on raw_graph_7_7.csv and an RMAT edge list:
    `v,in,out` for every vertex against the oracle."""
    path = os.path.join(GOLDEN, "raw_graph_7_7.csv")
    r = subprocess.run([os.path.join(BIN, "nstepdegrees"), "--input", path, "--output", str(tmp_path),
                        "--is_directed=true", "--step", "1", "--bits", "6", "--type", "both"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(tmp_path / "0000_0.csv", delimiter=",", dtype=np.int64, ndmin=2)
    s, d = oracle.read_edge_csv(path)
    g = oracle.CSC(int(max(s.max(), d.max())) + 1, s, d)
    din, _ = oracle.nstep_degrees(g, step=1, bits=6, direction="in")
    dout, _ = oracle.nstep_degrees(g, step=1, bits=6, direction="out")
    assert np.array_equal(got[:, 0], np.arange(g.V))
    assert np.array_equal(got[:, 1], din) and np.array_equal(got[:, 2], dout)
    assert (got[0, 1], got[0, 2], got[4, 1], got[4, 2]) == (3, 1, 1, 1)
