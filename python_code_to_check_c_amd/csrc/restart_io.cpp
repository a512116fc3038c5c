// Restart blocks in the reference's text layout (CCSD_Helper.py:1107-1125):
// one row per element in C order, "i<TAB>a<TAB>value" or
// "i<TAB>j<TAB>a<TAB>b<TAB>value", newline-terminated, where value is
// numpy.around(x, precs) printed the way Python's csv module prints a numpy
// float64 (str(), i.e. the shortest digits that read back to the same double,
// positional for decimal exponents -4 .. 15, otherwise d[.ddd]e+XX).  A
// C++ writer because at o = 20, v = 160 each doubles file has 10.24M rows.
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tdcc.h"

namespace {

thread_local std::string io_err;

// numpy.around(x, precs) for precs >= 0: rint(x * 10^precs) / 10^precs
double around(double x, int precs) {
  if (precs == 0) return std::rint(x);
  double s = 1.0;
  for (int k = 0; k < precs; ++k) s *= 10.0;
  return std::rint(x * s) / s;
}

// Python's repr() of a double into out (at most 32 chars); returns the length.
int py_repr(double x, char* out) {
  if (std::isnan(x)) return std::snprintf(out, 32, "nan");
  if (std::isinf(x)) return std::snprintf(out, 32, x < 0 ? "-inf" : "inf");
  if (x == 0.0) return std::snprintf(out, 32, std::signbit(x) ? "-0.0" : "0.0");
  char sci[40];
  const auto r = std::to_chars(sci, sci + sizeof sci, x, std::chars_format::scientific);
  *r.ptr = '\0';
  // sci = [-]d[.ddd]e(+|-)XX
  const char* p = sci;
  int n = 0;
  if (*p == '-') {
    out[n++] = '-';
    ++p;
  }
  char digits[24];
  int nd = 0;
  for (; *p && *p != 'e'; ++p)
    if (*p != '.') digits[nd++] = *p;
  const int e = std::atoi(p + 1);
  if (e >= -4 && e < 16) {
    if (e >= 0) {
      for (int k = 0; k <= e; ++k) out[n++] = k < nd ? digits[k] : '0';
      out[n++] = '.';
      if (nd > e + 1)
        for (int k = e + 1; k < nd; ++k) out[n++] = digits[k];
      else
        out[n++] = '0';
    } else {
      out[n++] = '0';
      out[n++] = '.';
      for (int k = 0; k < -e - 1; ++k) out[n++] = '0';
      for (int k = 0; k < nd; ++k) out[n++] = digits[k];
    }
  } else {
    out[n++] = digits[0];
    if (nd > 1) {
      out[n++] = '.';
      for (int k = 1; k < nd; ++k) out[n++] = digits[k];
    }
    n += std::snprintf(out + n, 12, "e%c%02d", e < 0 ? '-' : '+', e < 0 ? -e : e);
  }
  out[n] = '\0';
  return n;
}

int put_int(int64_t v, char* out) {
  const auto r = std::to_chars(out, out + 24, v);
  return (int)(r.ptr - out);
}

}  // namespace

extern "C" {

int tdcc_write_restart_block(const double* x, int rank, const int64_t* dims, int precs, const char* path) {
  io_err.clear();
  if (!x || !dims || !path) {
    io_err = "x, dims and path are required";
    return 1;
  }
  if (rank != 2 && rank != 4) {
    io_err = "rank must be 2 or 4";
    return 1;
  }
  if (precs < 0 || precs > 22) {
    io_err = "precs must be in [0, 22]";
    return 1;
  }
  int64_t total = 1;
  for (int d = 0; d < rank; ++d) {
    if (dims[d] < 0) {
      io_err = "negative extent";
      return 1;
    }
    total *= dims[d];
  }
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), std::fclose);
  if (!f) {
    io_err = std::string("cannot open ") + path;
    return 1;
  }
  std::vector<char> buf;
  buf.reserve(1 << 22);
  char tmp[48];
  int64_t idx[4] = {0, 0, 0, 0};
  for (int64_t e = 0; e < total; ++e) {
    for (int d = 0; d < rank; ++d) {
      const int n = put_int(idx[d], tmp);
      buf.insert(buf.end(), tmp, tmp + n);
      buf.push_back('\t');
    }
    const int n = py_repr(around(x[e], precs), tmp);
    buf.insert(buf.end(), tmp, tmp + n);
    buf.push_back('\n');
    for (int d = rank - 1; d >= 0; --d) {
      if (++idx[d] < dims[d]) break;
      idx[d] = 0;
    }
    if (buf.size() > (1 << 22) - 128) {
      if (std::fwrite(buf.data(), 1, buf.size(), f.get()) != buf.size()) {
        io_err = std::string("write failed: ") + path;
        return 1;
      }
      buf.clear();
    }
  }
  if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f.get()) != buf.size()) {
    io_err = std::string("write failed: ") + path;
    return 1;
  }
  if (std::fclose(f.release()) != 0) {
    io_err = std::string("close failed: ") + path;
    return 1;
  }
  return 0;
}

const char* tdcc_io_error(void) { return io_err.c_str(); }

}  // extern "C"
