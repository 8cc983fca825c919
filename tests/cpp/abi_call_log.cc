// A test double of the C-ABI entry points the engine's deferred pull uses: each call is
// appended to a log (name and the vectors it names) and returns PE_OK; no values are
// computed.  Linked only into tests/cpp/test_deferred_pull.cc, which checks the call
// sequences the host headers produce.  Opaque handles are small counters.
#include <cstdint>
#include <string>
#include <vector>

#include "plato_hip.h"

std::vector<std::string> g_calls;

namespace {
uintptr_t g_next = 1;
std::string id(const void* p) { return std::to_string(reinterpret_cast<uintptr_t>(p)); }
void log(const std::string& s) { g_calls.push_back(s); }
}  // namespace

extern "C" {
int pe_ctx_create(int, pe_ctx_t** out) {
  *out = reinterpret_cast<pe_ctx_t*>(g_next++);
  return PE_OK;
}
int pe_ctx_destroy(pe_ctx_t*) { return PE_OK; }
const char* pe_ctx_last_error(const pe_ctx_t*) { return ""; }
int pe_vec_create(pe_ctx_t*, int, uint64_t, pe_vec_t** out) {
  *out = reinterpret_cast<pe_vec_t*>(g_next++);
  return PE_OK;
}
int pe_vec_destroy(pe_vec_t*) { return PE_OK; }
int pe_vec_fill(pe_vec_t* v, double) {
  log("fill " + id(v));
  return PE_OK;
}
int pe_vec_set(pe_vec_t* v, uint64_t, double) {
  log("set " + id(v));
  return PE_OK;
}
int pe_vec_upload(pe_vec_t* v, const void*, uint64_t) {
  log("upload " + id(v));
  return PE_OK;
}
int pe_vec_download(pe_vec_t* v, void*, uint64_t) {
  log("download " + id(v));
  return PE_OK;
}
int pe_vec_sum(pe_vec_t* v, double* out) {
  *out = 0.0;
  log("sum " + id(v));
  return PE_OK;
}
int pe_pull_sum(pe_graph_t*, pe_vec_t* x, pe_vec_t* y, int) {
  log("pull_sum " + id(x) + " " + id(y));
  return PE_OK;
}
int pe_rank_apply(pe_graph_t*, pe_vec_t* x, pe_vec_t* y, const pe_apply_params_t*, double* d) {
  *d = 0.0;
  log("rank_apply " + id(x) + " " + id(y));
  return PE_OK;
}
int pe_rank_iterate(pe_graph_t*, pe_vec_t* x, pe_vec_t* y, int, const pe_apply_params_t*, double* d) {
  *d = 0.0;
  log("rank_iterate " + id(x) + " " + id(y));
  return PE_OK;
}
int pe_graph_get_info(const pe_graph_t*, pe_graph_info_t* info) {
  *info = pe_graph_info_t{};
  info->is_directed = 1;
  return PE_OK;
}
int pe_graph_create_reverse(pe_graph_t*, uint32_t, pe_graph_t** out) {
  *out = reinterpret_cast<pe_graph_t*>(g_next++);
  log("create_reverse");
  return PE_OK;
}
int pe_graph_destroy(pe_graph_t*) { return PE_OK; }
int pe_rank_init(pe_graph_t*, pe_vec_t* x, const pe_bitmap_t*, double, int) {
  log("rank_init " + id(x));
  return PE_OK;
}
}  // extern "C"
