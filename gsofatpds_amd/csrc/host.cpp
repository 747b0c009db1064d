// Host-side C ABI: error text, CSR validation, host matrices and their generators.
#include "gsofa_internal.h"

#include <cstring>
#include <new>

namespace gsofa {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int validate_csr(const gsofa_csr* a) {
  if (!a) return fail(GSOFA_ERR_INVALID_ARG, "csr is null");
  if (a->n <= 0) return fail(GSOFA_ERR_BAD_CSR, "n must be positive");
  if (a->nnz < 0) return fail(GSOFA_ERR_BAD_CSR, "nnz must be non-negative");
  if (!a->beg || !a->end || (a->nnz > 0 && !a->adj)) return fail(GSOFA_ERR_INVALID_ARG, "null csr array");
  for (int32_t v = 0; v < a->n; ++v) {
    const int64_t b = a->beg[v], e = a->end[v];
    if (b < 0 || e < b || e > a->nnz)
      return fail(GSOFA_ERR_BAD_CSR, "row " + std::to_string(v) + ": extent [" + std::to_string(b) + ", " +
                                         std::to_string(e) + ") outside adj of length " + std::to_string(a->nnz));
    for (int64_t k = b; k < e; ++k) {
      const int32_t c = a->adj[k];
      if (c < 0 || c >= a->n)
        return fail(GSOFA_ERR_BAD_CSR, "row " + std::to_string(v) + ": column id " + std::to_string(c) +
                                           " outside [0, " + std::to_string(a->n) + ")");
    }
  }
  return GSOFA_OK;
}

}  // namespace gsofa

using namespace gsofa;

extern "C" {

const char* gsofa_last_error(void) { return g_last_error.c_str(); }

int gsofa_generate_workload(int workload, int64_t member, gsofa_host_matrix** out) {
  if (!out) return fail(GSOFA_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  GenParams p;
  if (preset_params(workload, member, p) != GSOFA_OK)
    return fail(GSOFA_ERR_INVALID_ARG, "unknown workload " + std::to_string(workload));
  auto* h = new (std::nothrow) gsofa_host_matrix;
  if (!h) return fail(GSOFA_ERR_OUT_OF_MEMORY, "host matrix");
  int rc = generate(p, h->m);
  if (rc != GSOFA_OK) {
    delete h;
    return fail(rc, "generator rejected the workload parameters");
  }
  *out = h;
  return GSOFA_OK;
}

int gsofa_generate(const int64_t* q, uint64_t seed, gsofa_host_matrix** out) {
  if (!q || !out) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  GenParams p;
  p.kind = (int)q[0];
  p.dims = (int)q[1];
  p.points = (int)q[2];
  p.nx = q[3];
  p.ny = q[4];
  p.nz = q[5];
  p.n_keep = q[6];
  p.order = (int)q[7];
  p.nd_leaf = q[8];
  p.drop_side = (int)q[9];
  p.drop_ppm = (uint32_t)q[10];
  p.extra_every = q[11];
  p.avg_degree_x100 = q[12];
  p.seed = seed;
  if (p.kind == GSOFA_GEN_STENCIL) {
    if (!((p.dims == 2 && (p.points == 5 || p.points == 9)) || (p.dims == 3 && (p.points == 7 || p.points == 27))))
      return fail(GSOFA_ERR_INVALID_ARG, "stencil must be 2-D 5/9-point or 3-D 7/27-point");
    if (p.nx <= 0 || p.ny <= 0 || (p.dims == 3 && p.nz <= 0)) return fail(GSOFA_ERR_INVALID_ARG, "grid extents");
  }
  if (p.drop_ppm > 1000000) return fail(GSOFA_ERR_INVALID_ARG, "drop_ppm > 1e6");
  auto* h = new (std::nothrow) gsofa_host_matrix;
  if (!h) return fail(GSOFA_ERR_OUT_OF_MEMORY, "host matrix");
  int rc = generate(p, h->m);
  if (rc != GSOFA_OK) {
    delete h;
    return fail(rc, "generator rejected the parameters");
  }
  *out = h;
  return GSOFA_OK;
}

int gsofa_host_matrix_csr(const gsofa_host_matrix* m, gsofa_csr* view) {
  if (!m || !view) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  *view = m->m.view();
  return GSOFA_OK;
}

void gsofa_host_matrix_free(gsofa_host_matrix* m) { delete m; }

int gsofa_digest_fnv1a(const void* data, int64_t bytes, uint64_t* out) {
  if (!out || bytes < 0 || (bytes > 0 && !data)) return fail(GSOFA_ERR_INVALID_ARG, "bad digest arguments");
  const unsigned char* p = static_cast<const unsigned char*>(data);
  uint64_t h = 1469598103934665603ull;
  for (int64_t i = 0; i < bytes; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  *out = h;
  return GSOFA_OK;
}

}  // extern "C"
