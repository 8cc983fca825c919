// Minimal `--name=value` / `--name value` flag reader for the example drivers
// (the reference's drivers use gflags with the same flag names).
#pragma once

#include <cstdlib>
#include <cstring>
#include <string>

namespace plato_examples {

inline const char* flag(int argc, char** argv, const char* name, const char* dflt) {
  const size_t n = std::strlen(name);
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    if (std::strncmp(a, "--", 2) != 0 || std::strncmp(a + 2, name, n) != 0) continue;
    if (a[2 + n] == '=') return a + 3 + n;
    if (a[2 + n] == '\0' && i + 1 < argc) return argv[i + 1];
  }
  return dflt;
}

inline bool flag_bool(int argc, char** argv, const char* name, bool dflt) {
  const std::string v = flag(argc, argv, name, dflt ? "true" : "false");
  return v == "true" || v == "1";
}

}  // namespace plato_examples
