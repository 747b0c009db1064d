// gsofa_convert: prepares the reference's input files from a Matrix Market matrix and, like the
// reference's data pipeline (README.md:27-30), applies SuperLU_DIST's row and column
// permutations first: the output is the pattern of Pc * Pr * A * Pc^T, the matrix SuperLU_DIST
// hands to its symbolic factorisation.
//
//   gsofa_convert matrix.mtx out_prefix [--perm-r FILE] [--perm-c FILE] [--one-based]
//                 [--binary INDEX_BYTES VERT_BYTES]
//
// Writes out_prefix.beg, out_prefix.end and out_prefix.csr (the ASCII triplet the entry reads,
// gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:2567-2616) or, with --binary,
// out_prefix.beg.bin and out_prefix.csr.bin (gSoFa_SOAComparison/graph.hpp layout). Permutation
// files hold n integers (0-based unless --one-based).
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/gsofa.h"

static int usage(const char* argv0) {
  std::cerr << "usage: " << argv0
            << " matrix.mtx out_prefix [--perm-r FILE] [--perm-c FILE] [--one-based] [--binary INDEX_BYTES VERT_BYTES]\n";
  return 1;
}

static int die(int rc) {
  std::cerr << "gsofa error " << rc << ": " << gsofa_last_error() << std::endl;
  return 2;
}

int main(int argc, char** argv) {
  if (argc < 3) return usage(argc > 0 ? argv[0] : "gsofa_convert");
  const char *perm_r_file = nullptr, *perm_c_file = nullptr;
  int base = 0, ib = 0, vb = 0;
  for (int k = 3; k < argc; ++k) {
    const std::string a = argv[k];
    if (a == "--perm-r" && k + 1 < argc) perm_r_file = argv[++k];
    else if (a == "--perm-c" && k + 1 < argc) perm_c_file = argv[++k];
    else if (a == "--one-based") base = 1;
    else if (a == "--binary" && k + 2 < argc) {
      ib = std::atoi(argv[++k]);
      vb = std::atoi(argv[++k]);
    } else return usage(argv[0]);
  }
  gsofa_host_matrix* m = nullptr;
  int rc = gsofa_read_matrix_market(argv[1], &m);
  if (rc != GSOFA_OK) return die(rc);
  gsofa_csr a;
  gsofa_host_matrix_csr(m, &a);
  std::vector<int32_t> pr, pc;
  if (perm_r_file) {
    pr.resize((size_t)a.n);
    if ((rc = gsofa_read_permutation(perm_r_file, a.n, base, pr.data())) != GSOFA_OK) return die(rc);
  }
  if (perm_c_file) {
    pc.resize((size_t)a.n);
    if ((rc = gsofa_read_permutation(perm_c_file, a.n, base, pc.data())) != GSOFA_OK) return die(rc);
  }
  gsofa_host_matrix* b = nullptr;
  rc = gsofa_permute(&a, perm_r_file ? pr.data() : nullptr, perm_c_file ? pc.data() : nullptr, &b);
  gsofa_host_matrix_free(m);
  if (rc != GSOFA_OK) return die(rc);
  gsofa_csr bc;
  gsofa_host_matrix_csr(b, &bc);
  const std::string prefix = argv[2];
  if (ib) rc = gsofa_write_binary(&bc, (prefix + ".beg.bin").c_str(), (prefix + ".csr.bin").c_str(), ib, vb);
  else rc = gsofa_write_ascii(&bc, (prefix + ".beg").c_str(), (prefix + ".end").c_str(), (prefix + ".csr").c_str());
  const int32_t n = bc.n;
  const int64_t nnz = bc.nnz;
  gsofa_host_matrix_free(b);
  if (rc != GSOFA_OK) return die(rc);
  std::cout << "n " << n << " nnz " << nnz << std::endl;
  return 0;
}
