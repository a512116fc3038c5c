// Restart blocks in the reference's text layout (CCSD_Helper.py:1107-1125):
// one row per element in C order, "i<TAB>a<TAB>value" or
// "i<TAB>j<TAB>a<TAB>b<TAB>value", newline-terminated, where value is
// numpy.around(x, precs) printed the way Python's csv module prints a numpy
// float64 (str(), i.e. the shortest digits that read back to the same double,
// positional for decimal exponents -4 .. 15, otherwise d[.ddd]e+XX).  A
// C++ writer because at o = 20, v = 160 each doubles file has 10.24M rows.
//
// The reader parses those blocks back into dense arrays as the reference's
// convert_2data / convert_4data do (CC_Calculator.py:85-113): a zero array of
// the given extents, then for each row the indices (numbers rounded to the
// nearest integer) and the value assigned in file order.
//
// The state file is an exact binary snapshot of a stopped run (state, time,
// Fock and field parameters), for resuming bit for bit; the text blocks keep
// the reference's layout and semantics.
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <system_error>
#include <utility>
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

// Whole file into memory, NUL-terminated.
bool slurp(const char* path, std::vector<char>& buf) {
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), std::fclose);
  if (!f) {
    io_err = std::string("cannot open ") + path;
    return false;
  }
  buf.clear();
  char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof chunk, f.get())) > 0) buf.insert(buf.end(), chunk, chunk + n);
  if (std::ferror(f.get())) {
    io_err = std::string("read failed: ") + path;
    return false;
  }
  buf.push_back('\0');
  return true;
}

// One number at p (stops at TAB, newline, CR or NUL); advances p.
bool parse_number(const char*& p, const char* end, double& x) {
  const char* q = p;
  while (q < end && *q != '\t' && *q != '\n' && *q != '\r' && *q != '\0') ++q;
  if (q == p) return false;
  const char* b = p;
  if (*b == '+') ++b;
  const auto r = std::from_chars(b, q, x);
  if (r.ec != std::errc() || r.ptr != q) {
    // forms from_chars does not take (e.g. "-nan"): the C library's strtod
    std::string tok(p, q);
    char* e = nullptr;
    x = std::strtod(tok.c_str(), &e);
    if (!e || *e != '\0') return false;
  }
  p = q;
  return true;
}

constexpr char kStateMagic[16] = {'T', 'D', 'C', 'C', '-', 'S', 'T', 'A', 'T', 'E', '-', 'V', '1', 0, 0, 0};

uint64_t fnv1a(const unsigned char* p, size_t n, uint64_t h = 1469598103934665603ull) {
  for (size_t k = 0; k < n; ++k) {
    h ^= p[k];
    h *= 1099511628211ull;
  }
  return h;
}

struct StateHeader {
  char magic[16];
  int32_t o, v, method, precs;
  double t, dt, w0, A;
};
static_assert(sizeof(StateHeader) == 64, "state file header is 64 bytes");

int64_t state_doubles(int64_t o, int64_t v) {
  const int64_t n = o + v;
  return n * n + 2 * (2 * o * v + 2 * o * o * v * v);
}

bool read_state(const char* path, StateHeader& h, std::vector<double>* body) {
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), std::fclose);
  if (!f) {
    io_err = std::string("cannot open ") + path;
    return false;
  }
  if (std::fread(&h, sizeof h, 1, f.get()) != 1 || std::memcmp(h.magic, kStateMagic, 16) != 0) {
    io_err = std::string("not a TDCC state file: ") + path;
    return false;
  }
  if (h.o <= 0 || h.v <= 0 || h.o > 4096 || h.v > 4096) {
    io_err = "state file with invalid sizes";
    return false;
  }
  if (!body) return true;
  const int64_t nd = state_doubles(h.o, h.v);
  body->resize(nd);
  uint64_t sum = 0;
  if (std::fread(body->data(), sizeof(double), nd, f.get()) != (size_t)nd ||
      std::fread(&sum, sizeof sum, 1, f.get()) != 1) {
    io_err = std::string("truncated state file: ") + path;
    return false;
  }
  uint64_t hsh = fnv1a(reinterpret_cast<const unsigned char*>(&h), sizeof h);
  hsh = fnv1a(reinterpret_cast<const unsigned char*>(body->data()), nd * sizeof(double), hsh);
  if (hsh != sum) {
    io_err = std::string("checksum mismatch in state file: ") + path;
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

int tdcc_read_restart_block(const char* path, int rank, const int64_t* dims, double* out) {
  io_err.clear();
  if (!path || !dims || !out) {
    io_err = "path, dims and out are required";
    return 1;
  }
  if (rank != 2 && rank != 4) {
    io_err = "rank must be 2 or 4";
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
  std::vector<char> buf;
  if (!slurp(path, buf)) return 1;
  std::memset(out, 0, total * sizeof(double));
  const char* p = buf.data();
  const char* end = buf.data() + buf.size() - 1;
  int64_t line = 0;
  while (p < end) {
    ++line;
    if (*p == '\n' || (*p == '\r' && p + 1 < end && p[1] == '\n')) {  // blank line
      p += (*p == '\r') ? 2 : 1;
      continue;
    }
    int64_t off = 0;
    for (int d = 0; d <= rank; ++d) {
      double x;
      if (!parse_number(p, end, x)) {
        io_err = std::string(path) + ": line " + std::to_string(line) + ": expected " + std::to_string(rank + 1) +
                 " tab-separated numbers";
        return 1;
      }
      if (d < rank) {
        const double r = std::nearbyint(x);
        if (!(r >= 0.0) || r >= (double)dims[d]) {
          io_err = std::string(path) + ": line " + std::to_string(line) + ": index out of range";
          return 1;
        }
        off = off * dims[d] + (int64_t)r;
        if (*p != '\t') {
          io_err = std::string(path) + ": line " + std::to_string(line) + ": expected " + std::to_string(rank + 1) +
                   " tab-separated numbers";
          return 1;
        }
        ++p;
      } else {
        out[off] = x;
      }
    }
    if (*p == '\r') ++p;
    if (p < end && *p != '\n') {
      io_err = std::string(path) + ": line " + std::to_string(line) + ": trailing characters";
      return 1;
    }
    if (p < end) ++p;
  }
  return 0;
}

int tdcc_write_state_file(const char* path, int o, int v, int method, int precs, const double* params,
                          const double* F, const double* t1, const double* t2, const double* l1, const double* l2) {
  io_err.clear();
  if (!path || !params || !F || !t1 || !t2 || !l1 || !l2) {
    io_err = "all arguments are required";
    return 1;
  }
  if (o <= 0 || v <= 0 || o > 4096 || v > 4096) {
    io_err = "invalid sizes";
    return 1;
  }
  StateHeader h{};
  std::memcpy(h.magic, kStateMagic, 16);
  h.o = o;
  h.v = v;
  h.method = method;
  h.precs = precs;
  h.t = params[0];
  h.dt = params[1];
  h.w0 = params[2];
  h.A = params[3];
  const int64_t n = (int64_t)o + v, n1 = (int64_t)o * v, n2 = n1 * o * v;
  const std::string tmp = std::string(path) + ".tmp";
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(tmp.c_str(), "wb"), std::fclose);
  if (!f) {
    io_err = "cannot open " + tmp;
    return 1;
  }
  uint64_t hsh = fnv1a(reinterpret_cast<const unsigned char*>(&h), sizeof h);
  bool ok = std::fwrite(&h, sizeof h, 1, f.get()) == 1;
  const std::pair<const double*, int64_t> parts[5] = {{F, n * n}, {t1, 2 * n1}, {t2, 2 * n2}, {l1, 2 * n1}, {l2, 2 * n2}};
  for (const auto& pr : parts) {
    ok = ok && std::fwrite(pr.first, sizeof(double), pr.second, f.get()) == (size_t)pr.second;
    hsh = fnv1a(reinterpret_cast<const unsigned char*>(pr.first), pr.second * sizeof(double), hsh);
  }
  ok = ok && std::fwrite(&hsh, sizeof hsh, 1, f.get()) == 1;
  if (std::fclose(f.release()) != 0) ok = false;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    io_err = std::string("write failed: ") + path;
    std::remove(tmp.c_str());
    return 1;
  }
  return 0;
}

int tdcc_read_state_file_header(const char* path, int* o, int* v, int* method, int* precs, double* params) {
  io_err.clear();
  StateHeader h{};
  if (!path || !read_state(path, h, nullptr)) {
    if (!path) io_err = "path is required";
    return 1;
  }
  if (o) *o = h.o;
  if (v) *v = h.v;
  if (method) *method = h.method;
  if (precs) *precs = h.precs;
  if (params) {
    params[0] = h.t;
    params[1] = h.dt;
    params[2] = h.w0;
    params[3] = h.A;
  }
  return 0;
}

int tdcc_read_state_file(const char* path, double* F, double* t1, double* t2, double* l1, double* l2) {
  io_err.clear();
  StateHeader h{};
  std::vector<double> body;
  if (!path || !F || !t1 || !t2 || !l1 || !l2) {
    io_err = "all arguments are required";
    return 1;
  }
  if (!read_state(path, h, &body)) return 1;
  const int64_t n = (int64_t)h.o + h.v, n1 = (int64_t)h.o * h.v, n2 = n1 * h.o * h.v;
  const double* p = body.data();
  double* dst[5] = {F, t1, t2, l1, l2};
  const int64_t len[5] = {n * n, 2 * n1, 2 * n2, 2 * n1, 2 * n2};
  for (int k = 0; k < 5; ++k) {
    std::memcpy(dst[k], p, len[k] * sizeof(double));
    p += len[k];
  }
  return 0;
}

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
