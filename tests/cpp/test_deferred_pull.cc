// The engine defers a pull (foreach_edges with gather_sum / write_add) until the written
// state is next used: a restart-apply on it right after runs one fused pass
// (pe_rank_iterate); any other use runs the pull alone first (pe_pull_sum); fill drops it;
// moves and std::swap carry it.  Checked against the call log of a test double of the
// C-ABI (tests/cpp/abi_call_log.cc), on the CPU.
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "plato/engine/dualmode.hpp"

extern std::vector<std::string> g_calls;

namespace {
int failures = 0;
void expect(const std::vector<std::string>& want, const char* what) {
  if (g_calls != want) {
    ++failures;
    std::printf("FAILED %s\n  got:", what);
    for (auto& c : g_calls) std::printf(" [%s]", c.c_str());
    std::printf("\n  want:");
    for (auto& c : want) std::printf(" [%s]", c.c_str());
    std::printf("\n");
  } else {
    std::printf("ok %s\n", what);
  }
  g_calls.clear();
}
std::string id(const void* p) { return std::to_string(reinterpret_cast<uintptr_t>(p)); }

// the smallest graph facade the pull-only engine needs
struct fake_graph {
  using partition_t = plato::sequence_balanced_by_destination_t;
  pe_graph_t* handle() const { return reinterpret_cast<pe_graph_t*>(uintptr_t(7)); }
};
}  // namespace

int main() {
  using state_t = plato::dense_state_t<float, plato::sequence_balanced_by_destination_t>;
  plato::graph_info_t info(true);
  info.max_v_i_ = 9;
  fake_graph fg;
  plato::dualmode_engine_t<fake_graph, std::nullptr_t> engine(std::shared_ptr<fake_graph>(&fg, [](fake_graph*) {}),
                                                              nullptr, info);
  state_t a = engine.alloc_v_state<float>();
  state_t b = engine.alloc_v_state<float>();
  pe_graph_t* g = engine.walk_graph();
  const std::string A = id(a.handle()), B = id(b.handle());
  g_calls.clear();

  auto pull = [&](state_t& src, state_t& dst) {
    engine.foreach_edges<float, int>(plato::ops::gather_sum<float>{src}, plato::ops::write_add<float>{dst});
  };
  auto apply_op = [&](state_t* curt) {
    return plato::ops::rank_apply<state_t>{g, curt, 0.85, 0.15, 0.0, nullptr, true, false};
  };

  pull(a, b);
  expect({}, "the pull itself issues nothing");
  b.foreach<double>(apply_op(&a));
  expect({"rank_iterate " + A + " " + B}, "pull then apply from the same state: one fused pass");

  pull(a, b);
  (void)b.sum();
  expect({"pull_sum " + A + " " + B, "sum " + B}, "pull then another use: the pull runs first");

  pull(a, b);
  b.fill(0.0f);
  expect({"fill " + B}, "pull then fill: the pull is dropped");

  pull(a, b);
  (void)b[3];
  expect({"pull_sum " + A + " " + B, "download " + B}, "pull then an element read");

  pull(a, b);
  (void)b.handle();
  expect({"pull_sum " + A + " " + B}, "pull then the raw handle");

  state_t c = engine.alloc_v_state<float>();
  const std::string C = id(c.handle());
  g_calls.clear();
  pull(a, b);
  b.foreach<double>(apply_op(&c));
  expect({"pull_sum " + A + " " + B, "rank_apply " + C + " " + B},
         "apply against another state: pull alone, then the apply");

  pull(a, b);
  std::swap(a, b);  // a now holds B's vector and its pending pull from A's vector, now in b
  a.foreach<double>(apply_op(&b));
  expect({"rank_iterate " + A + " " + B}, "std::swap carries the pending pull");

  pull(b, a);
  pull(a, c);  // a (B's vector) is read by the second pull: its own pull runs first
  expect({"pull_sum " + A + " " + B}, "a pending state used as a pull source is brought up to date");
  c.foreach<double>(apply_op(&a));
  expect({"rank_iterate " + B + " " + C}, "then the second pull fuses with its apply");

  state_t moved(std::move(c));
  pull(b, moved);
  state_t moved2(std::move(moved));
  moved2.foreach<double>(apply_op(&b));
  expect({"rank_iterate " + A + " " + C}, "a move carries the pending pull");

  // reverse(): the pull walks the directed graph's reverse, built once; an apply must
  // name the same graph to fuse with it
  engine.reverse();
  pe_graph_t* rev = engine.walk_graph();
  expect({"create_reverse"}, "reverse() builds the device reverse on first use");
  pull(b, moved2);
  moved2.foreach<double>(plato::ops::rank_apply<state_t>{rev, &b, 0.85, 0.15, 0.0, nullptr, true, false});
  expect({"rank_iterate " + A + " " + C}, "reversed pull and apply on the reverse: fused");
  pull(b, moved2);
  moved2.foreach<double>(apply_op(&b));  // names the forward graph
  expect({"pull_sum " + A + " " + C, "rank_apply " + A + " " + C}, "an apply on another graph does not fuse");

  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
