// The host C++ API against the reference tests' known answers, on the device.
// Each case names the reference test it restates.  Run from the repository
// root: plato_amd/bin/test_plato_api tests/golden
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "plato/algo/bfs/bfs.hpp"
#include "plato/algo/pagerank/pagerank.hpp"
#include "plato/algo/ppr/personalized_pagerank.hpp"
#include "plato/algo/trustrank/trustrank.hpp"
#include "plato/algo/cgm/connected_component.hpp"
#include "plato/util/foutput.h"
#include "plato/graph/message_passing.hpp"

#include <zlib.h>

static int failures = 0;
#define EXPECT(cond)                                                   \
  do {                                                                 \
    if (!(cond)) {                                                     \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);  \
      ++failures;                                                      \
    }                                                                  \
  } while (0)

static std::string dir;

// plato/util/test/bitmap.cc:31-50 (MSB) and 52-76 (ORIGIN traversal)
static void Bitmap_MSB_Traverse() {
  plato::bitmap_t<> bitmap(40988);
  bitmap.set_bit(0);
  EXPECT(bitmap.msb() == 0);
  bitmap.set_bit(13);
  EXPECT(bitmap.msb() == 13);
  bitmap.set_bit(1982);
  EXPECT(bitmap.msb() == 1982);
  bitmap.clear();
  bitmap.set_bit(40987);
  EXPECT(bitmap.msb() == 40987);
  bitmap.set_bit(0);
  bitmap.set_bit(40986);
  EXPECT(bitmap.msb() == 40987);
  plato::bitmap_t<> b2(40988);
  b2.set_bit(0);
  b2.set_bit(19);
  b2.set_bit(40987);
  auto v = b2.to_vector();
  EXPECT(v.size() == 3 && v[0] == 0 && v[1] == 19 && v[2] == 40987);
}

// plato/graph/test/message_passing.cc:48-88 through the engine's pull operators
static void MessagePassing_AggregateMessage() {
  plato::graph_info_t graph_info(true);
  auto pdcsc = plato::create_dcsc_seqs_from_path<float>(&graph_info, dir + "/graph_10_9.csv");
  plato::dualmode_engine_t<plato::dcsc_t<float>, std::nullptr_t> engine(pdcsc, nullptr, graph_info);
  auto ones = engine.alloc_v_state<float>();
  auto values = engine.alloc_v_state<float>();
  ones.fill(1.0f);
  engine.foreach_edges<float, int>(plato::ops::gather_sum<float>{ones}, plato::ops::write_add<float>{values});
  EXPECT(values.sum() == 9.0);
  EXPECT(values[0] == 0.0f);
  for (plato::vid_t i = 1; i < 10; ++i) EXPECT(values[i] == 1.0f);
  EXPECT(graph_info.edges_ == 9 && graph_info.max_v_i_ == 9);
}

// plato/graph/test/structure.cc:137-151 and 205-218
static void GraphOps_Degrees() {
  plato::graph_info_t undirected(false);
  auto g1 = plato::create_dcsc_seqs_from_path<plato::empty_t>(&undirected, dir + "/v100_e2150_ua_c3.csv");
  auto d1 = plato::generate_dense_out_degrees_fg<uint32_t>(undirected, *g1, false);
  EXPECT(d1[0] == 19 && d1[99] == 59 && d1[28] == 19);
  EXPECT(undirected.vertices_ == 100);
  plato::graph_info_t directed(true);
  auto g2 = plato::create_dcsc_seqs_from_path<plato::empty_t>(&directed, dir + "/v100_e4300_da_c3.csv");
  auto d2 = plato::generate_dense_out_degrees_fg<uint32_t>(directed, *g2, false);
  EXPECT(d2[0] == 19 && d2[99] == 59 && d2[28] == 19);
}

// chain 0->1->...->9: r(v) = 0.15 * sum_{i<=v} 0.85^i after 20 iterations
static void PageRank_ChainClosedForm() {
  plato::graph_info_t graph_info(true);
  auto pdcsc = plato::create_dcsc_seqs_from_path<plato::empty_t>(&graph_info, dir + "/graph_10_9.csv");
  plato::algo::pagerank_opts_t opts;
  opts.iteration_ = 20;
  opts.eps_ = 0.0;
  opts.precision_ = 64;
  auto ranks = plato::algo::pagerank(*pdcsc, graph_info, opts);
  double expect = 0.0, p = 1.0;
  for (plato::vid_t v = 0; v < 10; ++v) {
    expect += 0.15 * p;
    p *= 0.85;
    EXPECT(std::fabs(ranks[v] - expect) <= 1e-14 * expect);
  }
  auto w = plato::create_dcsc_seqs_from_path<float>(&graph_info, dir + "/graph_10_9.csv");
  auto wranks = plato::algo::weighted_pagerank(*w, graph_info, opts);
  for (plato::vid_t v = 0; v < 10; ++v) EXPECT(std::fabs(wranks[v] - ranks[v]) <= 1e-6 * ranks[v]);
}

// PPR from 0 on the chain: r(v) = 0.15 * 0.85^v (pull and push/pull forms)
static void PPR_ChainClosedForm() {
  plato::graph_info_t graph_info(true);
  auto g = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, dir + "/graph_10_9.csv");
  plato::algo::personalized_pagerank_opts_t opts;
  opts.iteration_ = 20;
  opts.eps_ = 0.0;
  opts.src_ = 0;
  opts.precision_ = 64;
  auto pull = plato::algo::personalized_pagerank(g.second, graph_info, opts);
  auto pp = plato::algo::personalized_pagerank(g.second, g.first, graph_info, opts);
  double expect = 0.15;
  for (plato::vid_t v = 0; v < 10; ++v) {
    EXPECT(std::fabs(pull[v] - expect) <= 1e-14 * expect);
    EXPECT(std::fabs(pp[v] - expect) <= 1e-14 * expect);
    expect *= 0.85;
  }
}

// complete binary tree: visited 10, level(v) = floor(log2(v + 1))
static void BFS_BinaryTree() {
  plato::graph_info_t graph_info(true);
  auto g = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, dir + "/complete_binary_tree_10_9.csv");
  plato::algo::bfs_opts_t opts;
  opts.root_ = 0;
  std::vector<int32_t> lv;
  const plato::vid_t visited = plato::algo::breadth_first_search(g.second, g.first, graph_info, opts, &lv);
  EXPECT(visited == 10);
  for (int v = 0; v < 10; ++v) EXPECT(lv[v] == (int)std::floor(std::log2(v + 1)));
  plato::graph_info_t undirected(false);
  auto u = plato::create_dualmode_seq_from_path<plato::empty_t>(&undirected, dir + "/v100_e2150_ua_c3.csv");
  // the graph has 3 connected components; vertex 0's holds 20 of the 100 vertices
  // (scipy.sparse.csgraph.connected_components on the symmetrized edges)
  opts.root_ = 0;
  EXPECT(plato::algo::breadth_first_search(u.second, u.first, undirected, opts) == 20);
}

// TrustRank on the chain 0->1->...->9 (trustrank.hpp:80-259).  PageRank grows along
// the chain, so its single top seed is 9; inverse PageRank ranks the reversed chain,
// whose top is 0.  With good = {0} and seed_num >= 10 every good vertex is a seed and
// TrustRank is the PPR from 0: r(v) = 0.15 * 0.85^v.
static void TrustRank_Chain() {
  plato::graph_info_t graph_info(true);
  auto g = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, dir + "/graph_10_9.csv");
  plato::bitmap_t<> all(graph_info.max_v_i_ + 1);
  all.fill();
  plato::algo::trustrank_opts_t opts;
  opts.seed_num_ = 1;
  opts.select_method_ = 1;
  auto s1 = plato::algo::select_good_seeds(g.second, g.first, graph_info, all, opts);
  auto v1 = s1.to_vector();
  EXPECT(v1.size() == 1 && v1[0] == 9);
  opts.select_method_ = 0;
  auto s0 = plato::algo::select_good_seeds(g.second, g.first, graph_info, all, opts);
  auto v0 = s0.to_vector();
  EXPECT(v0.size() == 1 && v0[0] == 0);
  plato::bitmap_t<> good(graph_info.max_v_i_ + 1);
  good.set_bit(0);
  opts.seed_num_ = 100;
  opts.iteration_ = 20;
  opts.eps_ = 0.0;
  auto r = plato::algo::trustrank(g.second, g.first, graph_info, good, opts);
  double expect = 0.15;
  for (plato::vid_t v = 0; v < 10; ++v) {
    EXPECT(std::fabs(r[v] - expect) <= 1e-14 * expect);
    expect *= 0.85;
  }
  auto pairs = plato::algo::generate_rank_pairs(g.second, graph_info);
  EXPECT(pairs.size() == 10 && pairs[9].v_i == 9 && pairs[9].rank > pairs[0].rank);
}

// plato/algo/cgm/test/test_cgm.cc:83-139: on v100_e2150_ua_c3.csv (undirected) every
// label equals the BFS label of the test (the smallest id of the component); the file
// has three components, and the edges of all components add up to the input edges.
static void CGM_CalcComponent() {
  plato::graph_info_t graph_info(false);
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, dir + "/v100_e2150_ua_c3.csv");
  using dcsc_spec_t = plato::dcsc_t<plato::empty_t>;
  using bcsr_spec_t = plato::bcsr_t<plato::empty_t>;
  plato::dualmode_engine_t<dcsc_spec_t, bcsr_spec_t> engine(
      std::shared_ptr<dcsc_spec_t>(&graph.second, [](dcsc_spec_t*) {}),
      std::shared_ptr<bcsr_spec_t>(&graph.first, [](bcsr_spec_t*) {}), graph_info);
  plato::algo::connected_component_t<dcsc_spec_t, bcsr_spec_t> cc(&engine, graph_info);
  cc.compute();
  // the test's reference labels, by BFS over the file's edges
  std::vector<std::vector<int>> adj(100);
  {
    FILE* f = std::fopen((dir + "/v100_e2150_ua_c3.csv").c_str(), "r");
    int a, b;
    while (f && std::fscanf(f, "%d,%d", &a, &b) == 2) {
      adj[a].push_back(b);
      adj[b].push_back(a);
    }
    if (f) std::fclose(f);
  }
  std::vector<int> label(100, -1);
  for (int i = 0; i < 100; ++i) {
    if (label[i] != -1) continue;
    label[i] = i;
    std::vector<int> q{i};
    for (size_t h = 0; h < q.size(); ++h)
      for (int v : adj[q[h]])
        if (label[v] == -1) {
          label[v] = i;
          q.push_back(v);
        }
  }
  auto* labels = cc.get_labels();
  for (int v = 0; v < 100; ++v) EXPECT((int)(*labels)[v] == label[v]);
  EXPECT(cc.get_num_components() == 3);
  uint64_t edges = 0;
  for (auto& c : *cc.get_components_info()) edges += c.second.edges_;
  EXPECT(edges == 2150);
  EXPECT(cc.get_major_vertices() == cc.get_components_info()->at(cc.get_major_label()).vertices_);
}

// plato/graph/test/message_passing.cc:90-169: spread_message over a full and a sparse
// bitmap of graph_10_9.csv's 10 vertices, each active vertex sending itself to a sink
// that counts per vertex and returns 1.
static void MessagePassing_SpreadMessage() {
  plato::graph_info_t graph_info(true);
  auto pbcsr = plato::create_bcsr_seqd_from_path<float>(&graph_info, dir + "/graph_10_9.csv",
                                                        plato::edge_format_t::CSV, plato::float_decoder);
  using part_t = plato::sequence_balanced_by_destination_t;
  {
    plato::bitmap_t<> bitmap(graph_info.vertices_);
    bitmap.fill();
    plato::dense_state_t<uint32_t, part_t> values((uint64_t)10);
    values.fill(0);
    const int sum = plato::spread_message<plato::vid_t, int>(bitmap, plato::ops::send_self{},
                                                            plato::ops::count_into(values));
    EXPECT(sum == 10);
    for (plato::vid_t i = 0; i < 10; ++i) EXPECT(values[i] == 1u);
  }
  {
    plato::bitmap_t<> bitmap(graph_info.vertices_);
    bitmap.set_bit(0);
    bitmap.set_bit(5);
    bitmap.set_bit(9);
    plato::dense_state_t<uint32_t, part_t> values((uint64_t)10);
    values.fill(0);
    const int sum = plato::spread_message<plato::vid_t, int>(bitmap, plato::ops::send_self{},
                                                            plato::ops::count_into(values));
    EXPECT(sum == 3);
    EXPECT(values[0] == 1u && values[5] == 1u && values[9] == 1u);
    uint32_t total = 0;
    for (plato::vid_t i = 0; i < 10; ++i) total += values[i];
    EXPECT(total == 3u);
  }
}

// The plato:: calls of the reference driver example/pagerank_simple.cc:67-106 in the
// form that file writes them: the three-argument factory template with six arguments,
// pagerank<graph_spec_t> returning a dense state, its mutable foreach, and
// thread_local_fs_output with a gzip stream (the driver's gflags, glog and
// boost::format are its own dependencies and are replaced by constants here).
static void Reference_PagerankSimpleCallSites() {
  auto& cluster_info = plato::cluster_info_t::get_instance();
  plato::graph_info_t graph_info(true);
  auto pdcsc = plato::create_dcsc_seqs_from_path<plato::empty_t, plato::vid_t, plato::edge_file_cache_t>(
    &graph_info, dir + "/graph_10_9.csv", plato::edge_format_t::CSV,
    plato::dummy_decoder<plato::empty_t>, -1, false
  );
  using graph_spec_t = std::remove_reference<decltype(*pdcsc)>::type;
  plato::algo::pagerank_opts_t opts;
  opts.iteration_ = 20;
  opts.damping_   = 0.85;
  opts.eps_       = 0.0;
  auto ranks = plato::algo::pagerank<graph_spec_t>(*pdcsc, graph_info, opts);
  char tmpl[] = "/tmp/plato_api_XXXXXX";
  const std::string out = mkdtemp(tmpl);
  {
    plato::thread_local_fs_output os(out, "0000_", true);
    ranks.foreach<int> (
      [&](plato::vid_t v_i, double* pval) {
        auto& fs_output = os.local();
        fs_output << v_i << "," << *pval << "\n";
        return 0;
      }
    );
  }
  // the file holds the chain's closed form r_v = 0.15 * sum_{i<=v} 0.85^i
  gzFile f = gzopen((out + "/0000_0.csv.gz").c_str(), "rb");
  EXPECT(f != nullptr);
  if (f) {
    char line[128];
    int n = 0;
    while (gzgets(f, line, sizeof(line))) {
      unsigned v = 0;
      double r = 0;
      EXPECT(std::sscanf(line, "%u,%lf", &v, &r) == 2);
      double want = 0, p = 1;
      for (unsigned i = 0; i <= v; ++i, p *= 0.85) want += 0.15 * p;
      EXPECT(v == (unsigned)n && std::fabs(r - want) <= 1e-5 * want);
      ++n;
    }
    gzclose(f);
    EXPECT(n == 10);
  }
  std::remove((out + "/0000_0.csv.gz").c_str());
  std::remove(out.c_str());
  // writes through the pointer reach the device state
  const double before = ranks.sum();
  const int touched = ranks.foreach<int>([](plato::vid_t, double* pval) { *pval *= 2.0; return 1; });
  EXPECT(touched == 10);
  EXPECT(std::fabs(ranks.sum() - 2.0 * before) <= 1e-12 * before);
  // the dual-mode factory and engine as example/nstepdegrees.cc:146-163 write them
  auto graph = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, dir + "/graph_10_9.csv",
      plato::edge_format_t::CSV, plato::dummy_decoder<plato::empty_t>, -1, true);
  using bcsr_spec_t = plato::bcsr_t<plato::empty_t>;
  using dcsc_spec_t = plato::dcsc_t<plato::empty_t>;
  plato::dualmode_engine_t<dcsc_spec_t, bcsr_spec_t> engine (
    std::shared_ptr<dcsc_spec_t>(&graph.second,  [](dcsc_spec_t*) { }),
    std::shared_ptr<bcsr_spec_t>(&graph.first, [](bcsr_spec_t*) { }),
    graph_info);
  auto actives = engine.alloc_v_subset();
  actives.fill();
  EXPECT(actives.count() == 10);
  (void)cluster_info;
}

// dualmode.hpp:151-153, 299-322: after reverse() the pull sums over out-edges and the
// push follows in-edges.  Chain 0->1->...->9: every vertex but 9 has one out-edge, and a
// search from 9 over the reversed edges reaches v at level 9 - v, in push mode and in
// pull mode; reversing again restores the forward edges.
static void Engine_Reverse() {
  using in_t = plato::dcsc_t<plato::empty_t>;
  using out_t = plato::bcsr_t<plato::empty_t>;
  plato::graph_info_t graph_info(true);
  auto g = plato::create_dualmode_seq_from_path<plato::empty_t>(&graph_info, dir + "/graph_10_9.csv");
  for (double threshold : {1e9, 0.0}) {  // every epoch push, then every epoch pull
    plato::dualmode_engine_opts_t eo;
    eo.push_threshold_ = threshold;
    plato::dualmode_engine_t<in_t, out_t> engine(std::shared_ptr<in_t>(&g.second, [](in_t*) {}),
                                                 std::shared_ptr<out_t>(&g.first, [](out_t*) {}), graph_info, eo);
    engine.reverse();
    EXPECT(engine.is_reversed());
    auto ones = engine.alloc_v_state<float>();
    auto values = engine.alloc_v_state<float>();
    ones.fill(1.0f);
    engine.foreach_edges<float, int>(plato::ops::gather_sum<float>{ones}, plato::ops::write_add<float>{values});
    for (plato::vid_t v = 0; v < 9; ++v) EXPECT(values[v] == 1.0f);
    EXPECT(values[9] == 0.0f);

    plato::ops::bfs_state st;
    plato::check_status(pe_bfs_create(g.second.handle(), &st.h), "pe_bfs_create");
    auto current = engine.alloc_v_subset();
    auto next = engine.alloc_v_subset();
    plato::check_status(pe_bfs_reset(st.h, 9, current.handle()), "pe_bfs_reset");
    plato::vid_t visited = 1, actives = 1;
    while (actives) {
      st.next = &next;
      actives = engine.foreach_edges<plato::vid_t, plato::vid_t>(
          plato::ops::bfs_push_signal{&st}, plato::ops::bfs_push_slot{&st}, plato::ops::bfs_pull_signal{&st},
          plato::ops::bfs_pull_slot{&st}, current);
      EXPECT(engine.last_was_push() == (threshold > 1.0));
      visited += actives;
      std::swap(current, next);
    }
    EXPECT(visited == 10);
    std::vector<int32_t> lv(10, -2);
    plato::check_status(pe_bfs_get_levels(st.h, lv.data()), "pe_bfs_get_levels");
    for (int v = 0; v < 10; ++v) EXPECT(lv[v] == 9 - v);
    pe_bfs_destroy(st.h);

    engine.reverse();
    EXPECT(!engine.is_reversed());
    engine.foreach_edges<float, int>(plato::ops::gather_sum<float>{ones}, plato::ops::write_add<float>{values});
    EXPECT(values[0] == 0.0f);
    for (plato::vid_t v = 1; v < 10; ++v) EXPECT(values[v] == 1.0f);
  }
}

int main(int argc, char** argv) {
  dir = argc > 1 ? argv[1] : "tests/golden";
  struct { const char* name; void (*fn)(); } cases[] = {
      {"Bitmap.MSB_Traverse", Bitmap_MSB_Traverse},
      {"MessagePassing.AggregateMessage", MessagePassing_AggregateMessage},
      {"GraphOps.Degrees", GraphOps_Degrees},
      {"PageRank.ChainClosedForm", PageRank_ChainClosedForm},
      {"PPR.ChainClosedForm", PPR_ChainClosedForm},
      {"BFS.BinaryTree", BFS_BinaryTree},
      {"TrustRank.Chain", TrustRank_Chain},
      {"CGM.CalcComponent", CGM_CalcComponent},
      {"MessagePassing.SpreadMessage", MessagePassing_SpreadMessage},
      {"Reference.PagerankSimpleCallSites", Reference_PagerankSimpleCallSites},
      {"Engine.Reverse", Engine_Reverse},
  };
  for (auto& c : cases) {
    const int before = failures;
    c.fn();
    std::printf("[%s] %s\n", failures == before ? "  OK  " : "FAILED", c.name);
  }
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
