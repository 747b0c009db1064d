// Input conversion: Matrix Market patterns and the SuperLU_DIST-style reordering.
//
// The reference consumes the matrix SuperLU_DIST factors after its own reordering, dumped to the
// reference's CSR files (README.md:27-30: "the reordered matrix by the SuperLU_DIST solver
// before symbolic factorization"). SuperLU_DIST factors Pc * Pr * A * Pc^T: perm_r[i] is the
// position of row i of A in Pr * A, perm_c[j] the position of column j in A * Pc^T, and the
// column permutation is applied to the rows as well, so that the diagonal stays in place. The
// conversion here builds that matrix's pattern from A and the two permutation vectors:
//   B(perm_c[perm_r[i]], perm_c[j]) = A(i, j).
// Matrix Market coordinate files (the format SuperLU_DIST's examples read) give A; values are
// ignored (the symbolic path has none), symmetric, skew-symmetric and Hermitian files are
// expanded to both triangles.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "gsofa_internal.h"

using namespace gsofa;

namespace {

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Reads the whole file into memory (Matrix Market files are parsed in one pass).
int slurp(const char* path, std::vector<char>& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  std::fclose(f);
  buf.push_back('\0');
  return GSOFA_OK;
}

// CSR of the pattern from (row, column) pairs, rows in input order.
void pairs_to_host(int32_t n, const std::vector<int32_t>& ri, const std::vector<int32_t>& ci, HostMatrix& m) {
  m.n = n;
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  for (int32_t r : ri) ++cnt[(size_t)r + 1];
  for (int32_t v = 0; v < n; ++v) cnt[(size_t)v + 1] += cnt[v];
  m.beg.resize(n);
  m.end.resize(n);
  for (int32_t v = 0; v < n; ++v) {
    m.beg[v] = (int32_t)cnt[v];
    m.end[v] = (int32_t)cnt[v];
  }
  m.adj.resize(ri.size());
  for (size_t k = 0; k < ri.size(); ++k) m.adj[(size_t)m.end[ri[k]]++] = ci[k];
}

int check_perm(const int32_t* p, int32_t n, const char* what) {
  std::vector<uint8_t> seen((size_t)n, 0);
  for (int32_t i = 0; i < n; ++i) {
    if (p[i] < 0 || p[i] >= n || seen[p[i]])
      return fail(GSOFA_ERR_INVALID_ARG, std::string(what) + " is not a permutation of [0, n)");
    seen[p[i]] = 1;
  }
  return GSOFA_OK;
}

}  // namespace

extern "C" {

int gsofa_read_matrix_market(const char* path, gsofa_host_matrix** out) {
  if (!path || !out) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  std::vector<char> buf;
  int rc = slurp(path, buf);
  if (rc != GSOFA_OK) return rc;
  char* p = buf.data();
  // banner: %%MatrixMarket matrix coordinate <field> <symmetry>
  char* eol = std::strchr(p, '\n');
  std::string banner(p, eol ? (size_t)(eol - p) : std::strlen(p));
  char obj[64] = {0}, fmt[64] = {0}, field[64] = {0}, sym[64] = {0};
  if (std::sscanf(banner.c_str(), "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, sym) != 4)
    return fail(GSOFA_ERR_IO, std::string(path) + ": missing %%MatrixMarket banner");
  if (lower(obj) != "matrix" || lower(fmt) != "coordinate")
    return fail(GSOFA_ERR_IO, std::string(path) + ": only coordinate matrices are supported");
  const std::string fl = lower(field), sy = lower(sym);
  int values;
  if (fl == "pattern") values = 0;
  else if (fl == "real" || fl == "integer" || fl == "double") values = 1;
  else if (fl == "complex") values = 2;
  else return fail(GSOFA_ERR_IO, std::string(path) + ": unknown field " + field);
  const bool mirror = sy == "symmetric" || sy == "skew-symmetric" || sy == "hermitian";
  if (!mirror && sy != "general") return fail(GSOFA_ERR_IO, std::string(path) + ": unknown symmetry " + sym);
  p = eol ? eol + 1 : p + std::strlen(p);
  // comments, then the size line
  while (*p == '%' || *p == '\n' || *p == '\r') {
    char* e = std::strchr(p, '\n');
    if (!e) return fail(GSOFA_ERR_IO, std::string(path) + ": no size line");
    p = e + 1;
  }
  char* e = nullptr;
  const long long M = std::strtoll(p, &e, 10);
  const long long N = std::strtoll(e, &e, 10);
  const long long NZ = std::strtoll(e, &e, 10);
  if (M <= 0 || N <= 0 || NZ < 0 || M > INT32_MAX) return fail(GSOFA_ERR_IO, std::string(path) + ": bad size line");
  if (M != N) return fail(GSOFA_ERR_IO, std::string(path) + ": the matrix must be square");
  if ((mirror ? 2 : 1) * NZ > (long long)INT32_MAX - 64) return fail(GSOFA_ERR_CAPACITY, "too many entries");
  p = e;
  std::vector<int32_t> ri, ci;
  ri.reserve((size_t)(mirror ? 2 * NZ : NZ));
  ci.reserve(ri.capacity());
  for (long long k = 0; k < NZ; ++k) {
    errno = 0;
    const long long i = std::strtoll(p, &e, 10);
    if (e == p) return fail(GSOFA_ERR_IO, std::string(path) + ": fewer entries than the size line states");
    p = e;
    const long long j = std::strtoll(p, &e, 10);
    if (e == p || errno) return fail(GSOFA_ERR_IO, std::string(path) + ": malformed entry");
    p = e;
    for (int v = 0; v < values; ++v) {  // skip the value(s)
      std::strtod(p, &e);
      if (e == p) return fail(GSOFA_ERR_IO, std::string(path) + ": missing value");
      p = e;
    }
    if (i < 1 || i > M || j < 1 || j > N) return fail(GSOFA_ERR_BAD_CSR, std::string(path) + ": index out of range");
    ri.push_back((int32_t)(i - 1));
    ci.push_back((int32_t)(j - 1));
    if (mirror && i != j) {
      ri.push_back((int32_t)(j - 1));
      ci.push_back((int32_t)(i - 1));
    }
  }
  std::unique_ptr<gsofa_host_matrix> h(new (std::nothrow) gsofa_host_matrix);
  if (!h) return fail(GSOFA_ERR_OUT_OF_MEMORY, "host matrix");
  pairs_to_host((int32_t)M, ri, ci, h->m);
  *out = h.release();
  return GSOFA_OK;
}

int gsofa_permute(const gsofa_csr* a, const int32_t* perm_r, const int32_t* perm_c, gsofa_host_matrix** out) {
  if (!out) return fail(GSOFA_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  int rc = validate_csr(a);
  if (rc != GSOFA_OK) return rc;
  const int32_t n = a->n;
  if (perm_r && (rc = check_perm(perm_r, n, "perm_r")) != GSOFA_OK) return rc;
  if (perm_c && (rc = check_perm(perm_c, n, "perm_c")) != GSOFA_OK) return rc;
  std::vector<int32_t> ri, ci;
  ri.reserve((size_t)a->nnz);
  ci.reserve((size_t)a->nnz);
  for (int32_t i = 0; i < n; ++i) {
    const int32_t r = perm_r ? perm_r[i] : i;
    const int32_t rr = perm_c ? perm_c[r] : r;
    for (int64_t k = a->beg[i]; k < a->end[i]; ++k) {
      ri.push_back(rr);
      ci.push_back(perm_c ? perm_c[a->adj[k]] : a->adj[k]);
    }
  }
  std::unique_ptr<gsofa_host_matrix> h(new (std::nothrow) gsofa_host_matrix);
  if (!h) return fail(GSOFA_ERR_OUT_OF_MEMORY, "host matrix");
  pairs_to_host(n, ri, ci, h->m);
  *out = h.release();
  return GSOFA_OK;
}

int gsofa_read_permutation(const char* path, int32_t n, int32_t base, int32_t* perm) {
  if (!path || !perm || n <= 0 || (base != 0 && base != 1)) return fail(GSOFA_ERR_INVALID_ARG, "bad argument");
  std::vector<char> buf;
  int rc = slurp(path, buf);
  if (rc != GSOFA_OK) return rc;
  char* p = buf.data();
  for (int32_t i = 0; i < n; ++i) {
    char* e = nullptr;
    errno = 0;
    const long long v = std::strtoll(p, &e, 10);
    if (e == p || errno) return fail(GSOFA_ERR_IO, std::string(path) + ": fewer than n integers");
    if (v - base < 0 || v - base >= n) return fail(GSOFA_ERR_IO, std::string(path) + ": entry out of range");
    perm[i] = (int32_t)(v - base);
    p = e;
  }
  return check_perm(perm, n, path);
}

}  // extern "C"
