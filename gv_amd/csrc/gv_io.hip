// gv_io.hip -- MatrixMarket coordinate reader (host code; include/gv.h).
//
// Replaces readFile (marketReader.cu:3-121), which reads the same format
// line by line: '%' lines skipped, a "rows cols nnz" header, then one 1-based
// "src dst [value ...]" entry per line.  Here the file is parsed in large
// buffered blocks with a hand-written integer scanner, entries are checked
// (count, range, shape) instead of trusted, and the pairs come back 0-based;
// undirected mirroring is left to the graph (gv_create(..., directed = 0)
// applies every pair both ways), as the reference's !DIRECTED branch does.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gv_internal.h"

using namespace gv;

namespace {

struct Reader {
  FILE *f = nullptr;
  std::vector<char> buf;
  size_t pos = 0, len = 0;
  uint64_t line = 1;
  bool eof = false;

  explicit Reader(FILE *fp) : f(fp), buf(1 << 22) {}
  bool fill() {
    if (eof) return false;
    len = std::fread(buf.data(), 1, buf.size(), f);
    pos = 0;
    if (len == 0) eof = true;
    return len != 0;
  }
  int peek() {
    if (pos == len && !fill()) return EOF;
    return (unsigned char)buf[pos];
  }
  int get() {
    int c = peek();
    if (c != EOF) {
      ++pos;
      if (c == '\n') ++line;
    }
    return c;
  }
  // Reads the rest of the line into s (without the newline).
  void rest_of_line(std::string &s) {
    s.clear();
    int c;
    while ((c = get()) != EOF && c != '\n') s.push_back((char)c);
  }
  void skip_line() {
    int c;
    while ((c = get()) != EOF && c != '\n') {
    }
  }
  // Skips spaces and tabs (and a carriage return) within the line.
  void skip_blank() {
    int c;
    while ((c = peek()) == ' ' || c == '\t' || c == '\r') get();
  }
  // Unsigned decimal integer; false if none is here.
  bool number(uint64_t &v) {
    skip_blank();
    int c = peek();
    if (c < '0' || c > '9') return false;
    v = 0;
    while ((c = peek()) >= '0' && c <= '9') {
      if (v > (UINT64_MAX - 9) / 10) return false;
      v = v * 10 + (uint64_t)(c - '0');
      get();
    }
    return true;
  }
};

gv_status mm_fail(const char *path, uint64_t line, const char *what) {
  set_error(std::string("matrix market ") + path + ":" + std::to_string(line) + ": " + what);
  return GV_E_INVALID;
}

}  // namespace

extern "C" gv_status gv_read_matrix_market(const char *path, uint32_t *num_vertices, uint64_t *n,
                                           uint32_t *src, uint32_t *dst, int *symmetric) {
  if (!path || !num_vertices || !n) {
    set_error("null argument");
    return GV_E_INVALID;
  }
  const bool fill = src != nullptr || dst != nullptr;
  if (fill && (!src || !dst)) {
    set_error("src and dst must both be given to read the entries");
    return GV_E_INVALID;
  }
  const uint64_t cap = fill ? *n : 0;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error(std::string("cannot open ") + path + ": " + std::strerror(errno));
    return GV_E_INVALID;
  }
  Reader r(fp);
  gv_status rc = GV_OK;
  int sym = 0;
  uint64_t rows = 0, cols = 0, nnz = 0;
  std::string s;
  // banner / comments, then the size line
  bool have_size = false;
  while (!have_size) {
    int c = r.peek();
    if (c == EOF) {
      rc = mm_fail(path, r.line, "no size line");
      break;
    }
    if (c == '%') {
      const uint64_t ln = r.line;
      r.rest_of_line(s);
      if (ln == 1 && s.rfind("%%MatrixMarket", 0) == 0) {
        for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
        if (s.find("array") != std::string::npos) {
          rc = mm_fail(path, ln, "dense (array) matrices are not graphs");
          break;
        }
        if (s.find("symmetric") != std::string::npos || s.find("hermitian") != std::string::npos)
          sym = 1;
      }
      continue;
    }
    r.skip_blank();
    if (r.peek() == '\n') {  // blank line
      r.get();
      continue;
    }
    const uint64_t ln = r.line;
    if (!r.number(rows) || !r.number(cols) || !r.number(nnz)) {
      rc = mm_fail(path, ln, "expected 'rows cols nnz'");
      break;
    }
    r.skip_line();
    have_size = true;
  }
  if (!rc) {
    const uint64_t V = rows > cols ? rows : cols;
    if (V >= 0xFFFFFFFFull) rc = mm_fail(path, r.line, "more than 2^32-2 vertices");
    *num_vertices = (uint32_t)V;
  }
  if (!rc && fill && nnz > cap) rc = mm_fail(path, r.line, "entry buffer smaller than nnz");
  uint64_t k = 0;
  while (!rc) {
    int c = r.peek();
    if (c == EOF) break;
    if (c == '%') {
      r.skip_line();
      continue;
    }
    r.skip_blank();
    c = r.peek();
    if (c == '\n' || c == EOF) {
      if (c == '\n') r.get();
      continue;
    }
    const uint64_t ln = r.line;
    uint64_t u = 0, v = 0;
    if (!r.number(u) || !r.number(v)) {
      rc = mm_fail(path, ln, "expected 'row col [value]'");
      break;
    }
    if (u < 1 || u > rows || v < 1 || v > cols) {
      rc = mm_fail(path, ln, "entry outside 1..rows x 1..cols");
      break;
    }
    if (k >= nnz) {
      rc = mm_fail(path, ln, "more entries than the size line declares");
      break;
    }
    if (fill) {
      src[k] = (uint32_t)(u - 1);
      dst[k] = (uint32_t)(v - 1);
    }
    ++k;
    r.skip_line();  // values, if any, are ignored
  }
  if (!rc && k != nnz) rc = mm_fail(path, r.line, "fewer entries than the size line declares");
  std::fclose(fp);
  if (rc) return rc;
  *n = nnz;
  if (symmetric) *symmetric = sym;
  return GV_OK;
}
