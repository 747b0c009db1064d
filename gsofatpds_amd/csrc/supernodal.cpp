// Supernodal export: the supernode partition and per-supernode panel structure.
//
// The reference detects supernodes with struct Super {start, end}
// (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:341-345): a walk records
// start = end = s for each start s (:1970-1971) and moves end to every row that extends the run
// (:2039); a row that fails the second filter counts a new supernode whose start/end lines are
// left commented out (:2048-2049), and the array is never returned to the caller. With those
// lines in place every supernode is the contiguous range from the row where the count increments
// to the row before the next increment. That is the partition exported here:
//   xsup[k] = k-th row whose supernode flag is 1 (row 0 always is), xsup[nsuper] = n,
//   supno[i] = k for xsup[k] <= i < xsup[k+1].
// For a supernodal numeric factorisation each supernode k = [a, b) also gets its panels:
//   lsub(k) = sorted rows i >= b with (i, j) in F for some column j in [a, b)  (L panel rows)
//   usub(k) = sorted columns m >= b with (i, m) in F for some row i in [a, b)  (U panel columns)
// The diagonal block [a, b) x [a, b) is kept dense by such a factorisation and is not listed.
// Built on the host from the explicit structure (strict U by rows, strict L by columns) in one
// pass per supernode with a stamp array: O(nnz(L) + nnz(U)) plus the sorts of the panels.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gsofa_internal.h"

namespace gsofa {

int64_t supernode_partition(int32_t n, const uint8_t* flags, int32_t* xsup, int32_t* supno) {
  int64_t k = -1;
  for (int32_t i = 0; i < n; ++i) {
    if (i == 0 || flags[i]) {
      ++k;
      if (xsup) xsup[k] = i;
    }
    if (supno) supno[i] = (int32_t)k;
  }
  if (xsup) xsup[k + 1] = n;
  return k + 1;
}

int supernodal_panels(int32_t n, const uint8_t* flags, const int64_t* u_ptr, const int32_t* u_idx, const int64_t* l_ptr,
                      const int32_t* l_idx, int64_t* nsuper_out, int64_t* nnz_lsub, int64_t* nnz_usub, int32_t* xsup_out,
                      int64_t* xlsub, int32_t* lsub, int64_t* xusub, int32_t* usub) {
  if (n <= 0 || !flags || !u_ptr || !l_ptr) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  const int64_t ns = supernode_partition(n, flags, nullptr, nullptr);
  std::vector<int32_t> xsup((size_t)ns + 1);
  supernode_partition(n, flags, xsup.data(), nullptr);
  std::vector<int64_t> stamp((size_t)n, -1);
  std::vector<int32_t> panel;
  int64_t tl = 0, tu = 0;
  const bool fill = xlsub || lsub || xusub || usub;
  for (int64_t k = 0; k < ns; ++k) {
    const int32_t a = xsup[k], b = xsup[k + 1];
    // L panel: rows below the block in the columns of the supernode
    panel.clear();
    for (int32_t j = a; j < b; ++j)
      for (int64_t p = l_ptr[j]; p < l_ptr[j + 1]; ++p) {
        const int32_t i = l_idx[p];
        if (i >= b && stamp[i] != 2 * k) {
          stamp[i] = 2 * k;
          panel.push_back(i);
        }
      }
    if (fill) {
      std::sort(panel.begin(), panel.end());
      if (xlsub) xlsub[k] = tl;
      if (lsub) std::copy(panel.begin(), panel.end(), lsub + tl);
    }
    tl += (int64_t)panel.size();
    // U panel: columns right of the block in the rows of the supernode
    panel.clear();
    for (int32_t i = a; i < b; ++i)
      for (int64_t p = u_ptr[i]; p < u_ptr[i + 1]; ++p) {
        const int32_t m = u_idx[p];
        if (m >= b && stamp[m] != 2 * k + 1) {
          stamp[m] = 2 * k + 1;
          panel.push_back(m);
        }
      }
    if (fill) {
      std::sort(panel.begin(), panel.end());
      if (xusub) xusub[k] = tu;
      if (usub) std::copy(panel.begin(), panel.end(), usub + tu);
    }
    tu += (int64_t)panel.size();
  }
  if (xlsub) xlsub[ns] = tl;
  if (xusub) xusub[ns] = tu;
  if (xsup_out) std::copy(xsup.begin(), xsup.end(), xsup_out);
  if (nsuper_out) *nsuper_out = ns;
  if (nnz_lsub) *nnz_lsub = tl;
  if (nnz_usub) *nnz_usub = tu;
  return GSOFA_OK;
}

}  // namespace gsofa

extern "C" int gsofa_supernodal_from_structure(int32_t n, const uint8_t* flags, const int64_t* u_rowptr,
                                               const int32_t* u_col, const int64_t* l_colptr, const int32_t* l_row,
                                               int64_t* nsuper, int64_t* nnz_lsub, int64_t* nnz_usub, int32_t* xsup,
                                               int64_t* xlsub, int32_t* lsub, int64_t* xusub, int32_t* usub) {
  if (n <= 0 || !flags || !u_rowptr || !l_colptr) return gsofa::fail(GSOFA_ERR_INVALID_ARG, "bad argument");
  if (u_rowptr[n] > 0 && !u_col) return gsofa::fail(GSOFA_ERR_INVALID_ARG, "u_col is null");
  if (l_colptr[n] > 0 && !l_row) return gsofa::fail(GSOFA_ERR_INVALID_ARG, "l_row is null");
  return gsofa::supernodal_panels(n, flags, u_rowptr, u_col, l_colptr, l_row, nsuper, nnz_lsub, nnz_usub, xsup, xlsub,
                                  lsub, xusub, usub);
}
