// Internal host-side types shared by the engine's C++ translation units.
#pragma once

#include <array>
#include <climits>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gsofa.h"

namespace gsofa {

// A host-resident pattern in the reference's layout: row v is adj[beg[v] .. end[v]).
struct HostMatrix {
  int32_t n = 0;
  std::vector<int32_t> beg, end, adj;
  gsofa_csr view() const {
    gsofa_csr c;
    c.n = n;
    c.nnz = (int64_t)adj.size();
    c.beg = beg.data();
    c.end = end.data();
    c.adj = adj.data();
    return c;
  }
};

struct GenParams {
  int kind = GSOFA_GEN_STENCIL;
  int dims = 2, points = 5;
  int64_t nx = 0, ny = 0, nz = 1, n_keep = 0;
  int order = GSOFA_ORDER_NATURAL;
  int64_t nd_leaf = 64;
  int drop_side = GSOFA_DROP_ANY;
  uint32_t drop_ppm = 0;
  int64_t extra_every = 0;
  int64_t avg_degree_x100 = 600;
  uint64_t seed = 0;
};

uint64_t splitmix64(uint64_t x);
int generate(const GenParams& p, HostMatrix& M);
int preset_params(int workload, int64_t member, GenParams& p);

// Thread-local error text behind gsofa_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// Validates a caller CSR: n > 0, 0 <= beg[v] <= end[v] <= nnz, column ids in [0, n).
int validate_csr(const gsofa_csr* a);

// Supernodal export (supernodal.cpp): the partition from the supernode flags (returns the
// number of supernodes; xsup n_super + 1 entries, supno n, either may be null) and the panels
// of every supernode from the strict structure (sizes only when every output array is null).
int64_t supernode_partition(int32_t n, const uint8_t* flags, int32_t* xsup, int32_t* supno);
int supernodal_panels(int32_t n, const uint8_t* flags, const int64_t* u_ptr, const int32_t* u_idx, const int64_t* l_ptr,
                      const int32_t* l_idx, int64_t* nsuper, int64_t* nnz_lsub, int64_t* nnz_usub, int32_t* xsup,
                      int64_t* xlsub, int32_t* lsub, int64_t* xusub, int32_t* usub);

}  // namespace gsofa

struct gsofa_host_matrix {
  gsofa::HostMatrix m;
};
