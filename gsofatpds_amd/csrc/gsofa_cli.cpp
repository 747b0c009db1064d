// gsofa_cli: the reference's command line on the engine.
// Mirrors gSoFa_SOAComparison/main.cu:20-63 for one process: runs the drop-in entry, prints the
// "Final ..." lines and appends the CSV row (header, then argv[2];argv[4];argv[6];argv[7];argv[8];
// min;max) to Vary_percent_blocksi_pre2.csv in the working directory. Errors print the engine's
// message and return a nonzero status instead of exit(1) inside the library.
#include <cstdio>
#include <fstream>
#include <iostream>

#include "../../include/gsofa.h"

int main(int argc, char** argv) {
  if (argc != 11) {
    std::cerr << "usage: " << (argc > 0 ? argv[0] : "gsofa_cli")
              << " beg end csr num_process chunk_size percent_cat0 percent_cat2 N_blocks_source_cat2 N_GPU_Node "
                 "N_src_group\n";
    return 1;
  }
  std::cout << "Number of Processes: " << argv[4] << std::endl;
  uint64_t fill = 0, teps = 0;
  double ms = 0.0;
  int64_t nsup = 0;
  const int rc = gsofa_symbfact_min_id(argc, argv, 0, &fill, &ms, &nsup, &teps);
  if (rc != GSOFA_OK) {
    std::cerr << "gsofa error " << rc << ": " << gsofa_last_error() << std::endl;
    return 2;
  }
  const double max_time = ms, min_time = ms;  // one process: max == min
  std::cout << "Final Number of fill-ins detected: " << fill << std::endl;
  std::cout << "Final Number of Supernodes detected: " << nsup << std::endl;
  std::cout << "Final Maximum time reported among the processes: " << max_time << " ms" << std::endl;
  std::cout << "Final Minimum time reported among the processes: " << min_time << " ms" << std::endl;
  std::cout << "Final TEPS according to slowest process: " << (double)teps / (max_time / 1000.0) << std::endl;
  std::fstream log("Vary_percent_blocksi_pre2.csv", std::fstream::out | std::fstream::app);
  log << "vert_count;N_gpus;" << "p0;" << "p1" << ";" << "#blocks_src" << ";" << "min_time" << ";" << "max_time" << std::endl;
  log << argv[2] << ";" << argv[4] << ";" << argv[6] << ";" << argv[7] << ";" << argv[8] << ";" << min_time << ";"
      << max_time << std::endl;
  return 0;
}
