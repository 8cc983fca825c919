// spread_message (plato/graph/message_passing.hpp:145-249) over a device bitmap.
//
// The reference's spread functor sends messages to explicit ranks and a sink functor
// consumes them; on one device every message is local, and both functors come from a
// closed set of operator objects (host lambdas cannot run on the GPU):
//   spread  plato::ops::send_self{}            message = the active vertex v
//           plato::ops::send_key{key}          message = key[v] (a uint32 state)
//   sink    plato::ops::count_into{counts}     counts[msg] += 1, returns 1
//           plato::ops::add_into<T>{acc, val}  acc[msg] += val[v] (or 1), returns 1
// The result is the sum of the sink's returns (the number of messages).  Any other
// functor is a compile error.
#pragma once

#include <type_traits>

#include "plato/graph/state/dense_state.hpp"
#include "plato/util/bitmap.hpp"

namespace plato {

template <typename MSG>
struct mepa_sd_context_t {};  // the device operators send without a context object

namespace ops {
struct send_self {};
template <typename PART>
struct send_key_t {
  dense_state_t<uint32_t, PART>& key;
};
template <typename PART>
send_key_t<PART> send_key(dense_state_t<uint32_t, PART>& key) { return {key}; }

template <typename PART>
struct count_into_t {
  dense_state_t<uint32_t, PART>& counts;
};
template <typename PART>
count_into_t<PART> count_into(dense_state_t<uint32_t, PART>& counts) { return {counts}; }

template <typename T, typename PART>
struct add_into_t {
  dense_state_t<T, PART>& acc;
  dense_state_t<T, PART>* value;
};
template <typename T, typename PART>
add_into_t<T, PART> add_into(dense_state_t<T, PART>& acc, dense_state_t<T, PART>* value = nullptr) {
  return {acc, value};
}

template <typename...>
struct unsupported_spread : std::false_type {};
}  // namespace ops

namespace detail {
inline pe_vec_t* spread_key_handle(const ops::send_self&) { return nullptr; }
template <typename PART>
pe_vec_t* spread_key_handle(const ops::send_key_t<PART>& k) { return k.key.handle(); }

template <typename SPREAD>
uint64_t spread_run(bitmap_t<>& actives, const SPREAD& sp, int op, pe_vec_t* value, pe_vec_t* sink) {
  uint64_t n = 0;
  check_status(pe_spread_message(actives.handle(), op, spread_key_handle(sp), value, sink, &n), "pe_spread_message");
  return n;
}
}  // namespace detail

template <typename MSG, typename R, typename SPREAD, typename PART>
R spread_message(bitmap_t<>& actives, SPREAD&& spread, ops::count_into_t<PART> sink) {
  const R r = (R)detail::spread_run(actives, spread, PE_SPREAD_COUNT, nullptr, sink.counts.handle());
  sink.counts.invalidate();
  return r;
}

template <typename MSG, typename R, typename SPREAD, typename T, typename PART>
R spread_message(bitmap_t<>& actives, SPREAD&& spread, ops::add_into_t<T, PART> sink) {
  const R r = (R)detail::spread_run(actives, spread, PE_SPREAD_ADD, sink.value ? sink.value->handle() : nullptr,
                                    sink.acc.handle());
  sink.acc.invalidate();
  return r;
}

template <typename MSG, typename R, typename SPREAD, typename SINK>
R spread_message(bitmap_t<>&, SPREAD&&, SINK&&) {
  static_assert(ops::unsupported_spread<SPREAD, SINK>::value,
                "spread_message: use plato::ops::send_self / send_key with count_into / add_into");
  return R();
}

}  // namespace plato
