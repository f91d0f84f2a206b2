// Matrix Market (NIST MM exchange format) reader and writer: on-disk operators for the CSR path.
//
// Reads `matrix coordinate {real|integer|pattern} {general|symmetric|skew-symmetric}` and
// `matrix array {real|integer} general` (column-major).  Output is 0-based CSR with columns
// sorted in each row and duplicate entries summed in file order.  Complex and Hermitian files are
// rejected: the solver's element type is real double.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gDMR.h"
#include "gDMRC.h"

namespace {

struct Triplets {
  int rows = 0, cols = 0;
  std::vector<int> r, c;
  std::vector<double> v;
};

std::string lower(std::string s) {
  for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// Whole-file read; the parser walks the buffer with strtol / strtod.
bool slurp(const char *path, std::string &buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(len > 0 ? (size_t)len : 0);
  const size_t got = len > 0 ? std::fread(&buf[0], 1, (size_t)len, f) : 0;
  std::fclose(f);
  return got == buf.size();
}

int parse(const char *path, Triplets &T) {
  std::string buf;
  if (!path || !slurp(path, buf)) return DMR_ERR_ARG;
  const char *p = buf.c_str();
  const char *end = p + buf.size();
  // Banner: %%MatrixMarket matrix <format> <field> <symmetry>
  const char *nl = std::find(p, end, '\n');
  char w0[64] = {0}, w1[64] = {0}, w2[64] = {0}, w3[64] = {0}, w4[64] = {0};
  const std::string banner(p, nl);
  if (std::sscanf(banner.c_str(), "%63s %63s %63s %63s %63s", w0, w1, w2, w3, w4) != 5)
    return DMR_ERR_ARG;
  if (lower(w0) != "%%matrixmarket" || lower(w1) != "matrix") return DMR_ERR_ARG;
  const std::string format = lower(w2), field = lower(w3), sym = lower(w4);
  const bool coord = format == "coordinate";
  if (!coord && format != "array") return DMR_ERR_ARG;
  if (field == "complex" || sym == "hermitian") return DMR_ERR_NOTBUILT;
  if (field != "real" && field != "integer" && field != "double" && field != "pattern")
    return DMR_ERR_ARG;
  const bool pattern = field == "pattern";
  if (!coord && (pattern || sym != "general")) return DMR_ERR_ARG;
  int mirror = 0;  // 0 general, 1 symmetric, -1 skew-symmetric
  if (sym == "symmetric")
    mirror = 1;
  else if (sym == "skew-symmetric")
    mirror = -1;
  else if (sym != "general")
    return DMR_ERR_ARG;
  // Skip comments and blank lines.
  p = nl < end ? nl + 1 : end;
  for (;;) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    if (p < end && (*p == '%' || *p == '\n')) {
      p = std::find(p, end, '\n');
      if (p < end) ++p;
      continue;
    }
    break;
  }
  char *q;
  errno = 0;
  const long rows = std::strtol(p, &q, 10);
  if (q == p) return DMR_ERR_ARG;
  p = q;
  const long cols = std::strtol(p, &q, 10);
  if (q == p) return DMR_ERR_ARG;
  p = q;
  long long entries = 0;
  if (coord) {
    entries = std::strtoll(p, &q, 10);
    if (q == p) return DMR_ERR_ARG;
    p = q;
  } else {
    entries = (long long)rows * cols;
  }
  if (rows < 1 || cols < 1 || rows > 0x7fffffffL || cols > 0x7fffffffL || entries < 0)
    return DMR_ERR_ARG;
  if (mirror && rows != cols) return DMR_ERR_ARG;
  T.rows = (int)rows;
  T.cols = (int)cols;
  const long long cap = mirror ? 2 * entries : entries;
  T.r.reserve(cap);
  T.c.reserve(cap);
  T.v.reserve(cap);
  for (long long e = 0; e < entries; ++e) {
    long i, j;
    if (coord) {
      i = std::strtol(p, &q, 10);
      if (q == p) return DMR_ERR_ARG;
      p = q;
      j = std::strtol(p, &q, 10);
      if (q == p) return DMR_ERR_ARG;
      p = q;
    } else {
      i = (long)(e % rows) + 1;
      j = (long)(e / rows) + 1;
    }
    double v = 1.0;
    if (!pattern) {
      v = std::strtod(p, &q);
      if (q == p) return DMR_ERR_ARG;
      p = q;
    }
    if (i < 1 || i > rows || j < 1 || j > cols) return DMR_ERR_ARG;
    if (!coord && v == 0.0) continue;  // dense file: keep the nonzeros only
    T.r.push_back((int)(i - 1));
    T.c.push_back((int)(j - 1));
    T.v.push_back(v);
    if (mirror && i != j) {
      T.r.push_back((int)(j - 1));
      T.c.push_back((int)(i - 1));
      T.v.push_back(mirror * v);
    }
  }
  return DMR_OK;
}

// Triplets -> CSR: counting sort by row (stable, so file order is kept within a row), then a
// stable sort by column per row and a merge of duplicates.
void to_csr(const Triplets &T, std::vector<int> &rp, std::vector<int> &ci, std::vector<double> &va) {
  const size_t m = T.r.size();
  rp.assign(T.rows + 1, 0);
  for (size_t e = 0; e < m; ++e) ++rp[T.r[e] + 1];
  for (int i = 0; i < T.rows; ++i) rp[i + 1] += rp[i];
  std::vector<int> pos(rp.begin(), rp.end() - 1);
  std::vector<int> c(m);
  std::vector<double> v(m);
  for (size_t e = 0; e < m; ++e) {
    const int at = pos[T.r[e]]++;
    c[at] = T.c[e];
    v[at] = T.v[e];
  }
  ci.clear();
  va.clear();
  ci.reserve(m);
  va.reserve(m);
  std::vector<int> out(T.rows + 1, 0);
  std::vector<int> idx;
  for (int i = 0; i < T.rows; ++i) {
    idx.resize(rp[i + 1] - rp[i]);
    for (size_t t = 0; t < idx.size(); ++t) idx[t] = rp[i] + (int)t;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return c[a] < c[b]; });
    for (size_t t = 0; t < idx.size(); ++t) {
      const int e = idx[t];
      if (t > 0 && c[e] == ci.back())
        va.back() += v[e];
      else {
        ci.push_back(c[e]);
        va.push_back(v[e]);
      }
    }
    out[i + 1] = (int)ci.size();
  }
  rp.swap(out);
}

}  // namespace

extern "C" {

int dmrReadMatrixMarket(const char *path, int *rows, int *cols, long long *nnz, int *rowPtr,
                        int *colInd, double *val) {
  Triplets T;
  const int rc = parse(path, T);
  if (rc != DMR_OK) return rc;
  std::vector<int> rp, ci;
  std::vector<double> va;
  if (T.r.size() > 0x7fffffffu) return DMR_ERR_ARG;  // int32 CSR indices
  to_csr(T, rp, ci, va);
  if (rows) *rows = T.rows;
  if (cols) *cols = T.cols;
  if (nnz) *nnz = (long long)ci.size();
  if (rowPtr) std::copy(rp.begin(), rp.end(), rowPtr);
  if (colInd) std::copy(ci.begin(), ci.end(), colInd);
  if (val) std::copy(va.begin(), va.end(), val);
  return DMR_OK;
}

int dmrWriteMatrixMarket(const char *path, int n, const int *rowPtr, const int *colInd,
                         const double *val) {
  if (!path || n < 1 || !rowPtr || !colInd || !val) return DMR_ERR_ARG;
  FILE *f = std::fopen(path, "w");
  if (!f) return DMR_ERR_ARG;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", n, n, rowPtr[n]);
  for (int i = 0; i < n; ++i)
    for (int e = rowPtr[i]; e < rowPtr[i + 1]; ++e)
      std::fprintf(f, "%d %d %.17g\n", i + 1, colInd[e] + 1, val[e]);
  return std::fclose(f) == 0 ? DMR_OK : DMR_ERR_ARG;
}

}  // extern "C"
