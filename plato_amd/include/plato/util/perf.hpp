// stop_watch_t (plato/util/perf.hpp:71-90): named wall-clock marks in milliseconds.
#pragma once

#include <chrono>
#include <map>
#include <string>

namespace plato {

class stop_watch_t {
 public:
  void mark(const std::string& smark) { mark_[smark] = now_ms(); }
  // milliseconds since the mark
  double show(const std::string& smark) { return now_ms() - mark_[smark]; }

 private:
  static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
  std::map<std::string, double> mark_;
};

}  // namespace plato
