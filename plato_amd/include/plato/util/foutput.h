// thread_local_fs_output (plato/util/foutput.h:57-90): one output stream per
// process for results written by a host traversal, `<path>/<prefix>0.csv`, or
// `<path>/<prefix>0.csv.gz` when compressed (gzip through zlib).  The reference
// keeps one file per OpenMP thread; the device engine's host traversals run on
// one thread, so there is one file.  HDFS paths are not supported.
#pragma once

#include <zlib.h>

#include <cstdio>
#include <memory>
#include <ostream>
#include <streambuf>
#include <string>

#include "plato/graph/base.hpp"

namespace plato {

namespace detail {
// std::streambuf over a gzip or plain FILE sink
class file_sink_buf : public std::streambuf {
 public:
  file_sink_buf(const std::string& name, bool gz) {
    if (gz) {
      gz_ = gzopen(name.c_str(), "wb");
      if (!gz_) fatal("gzopen", -1, name.c_str());
    } else {
      f_ = std::fopen(name.c_str(), "w");
      if (!f_) fatal("fopen", -1, name.c_str());
    }
    setp(buf_, buf_ + sizeof(buf_));
  }
  ~file_sink_buf() override {
    flush_out();
    if (gz_) gzclose(gz_);
    if (f_) std::fclose(f_);
  }

 protected:
  int_type overflow(int_type c) override {
    flush_out();
    if (c != traits_type::eof()) {
      *pptr() = (char)c;
      pbump(1);
    }
    return traits_type::not_eof(c);
  }
  int sync() override {
    flush_out();
    return 0;
  }

 private:
  void flush_out() {
    const long n = pptr() - pbase();
    if (n > 0) {
      if (gz_) gzwrite(gz_, pbase(), (unsigned)n);
      else std::fwrite(pbase(), 1, (size_t)n, f_);
    }
    setp(buf_, buf_ + sizeof(buf_));
  }
  gzFile gz_ = nullptr;
  FILE* f_ = nullptr;
  char buf_[1 << 16];
};
}  // namespace detail

class thread_local_fs_output {
 public:
  thread_local_fs_output(const std::string& path, const std::string& prefix, bool compressed = true)
      : buf_(new detail::file_sink_buf(path + "/" + prefix + "0.csv" + (compressed ? ".gz" : ""), compressed)),
        os_(buf_.get()) {}
  thread_local_fs_output(const thread_local_fs_output&) = delete;
  thread_local_fs_output& operator=(const thread_local_fs_output&) = delete;
  ~thread_local_fs_output() { os_.flush(); }

  std::ostream& local() { return os_; }

 private:
  std::unique_ptr<detail::file_sink_buf> buf_;
  std::ostream os_;
};

// fs_mt_omp_output_t (plato/util/foutput.h): the writer the reference's n-step degrees
// driver saves through; one file `<path>/<prefix>0.csv[.gz]`, written via ostream().
class fs_mt_omp_output_t {
 public:
  fs_mt_omp_output_t(const std::string& path, const std::string& prefix, bool compressed = true)
      : out_(path, prefix, compressed) {}
  std::ostream& ostream() { return out_.local(); }

 private:
  thread_local_fs_output out_;
};

}  // namespace plato
