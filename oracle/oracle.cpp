// ORACLE - test infrastructure only.
//
// A host C++ restatement of gSoFa's symbolic factorisation, used as the parity checker for the
// GPU engine and as the CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load it; the engine never does.
//
// What it restates (file:line into the reference):
//   * per-source minimax fill-path traversal, label-correcting, with the reference's epoch
//     labels and "last source" fill marks:
//       seed phase      gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:821-906
//       relaxation      ...:909-1027  (cost = max(front+off, label[front]) :923;
//                                     atomicMin relax :961; enqueue only m < src :964-972;
//                                     fill test m+off > cost with mark dedupe :974-992)
//       epoch labels    ...:470-479, 2337-2338, 3116-3146
//   * nzU[i] = |{m > i : (i,m) in F}| counted as the reference does (A's entries with
//     multiplicity at :896-901, fills at :986-991)
//   * supernode rule of warpcentric_local_detection_test ...:1790-2123 (starts :1879, :1928;
//     pass-through test :2034; extra supernode on a failed pass-through :2047) and the
//     reported count + 1 of ...:3549
//
// Pins (tests/test_oracle.py): the reference's known answer 980,136 fills for its shipped
// 11,948-row input (SpaceOptimization/with_supernodes/space_optimization_log_1_264_GPU_warp_centric.dat:2),
// dense boolean Gaussian elimination on small random patterns, symbolic Cholesky through the
// elimination tree on symmetric patterns, closed forms (tridiagonal, arrows), and invariance of
// every output under chunk size and thread count. The supernode count has no value recorded in
// the reference: it is pinned by hand-worked cases only (partial pin).

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

typedef struct oracle_result {
  uint64_t fill_count;   // |F \ A| off the diagonal ("Number of fill-ins detected")
  uint64_t nnz_F;        // |F| including the diagonal
  uint64_t W;            // schedule-independent edge checks: sum_i deg(i) + sum_{v in R_i} deg(v)
  uint64_t R;            // sum_i |R_i|, R_i = {v < i reachable from i through vertices < i}
  uint64_t expansions;   // frontier entries processed by this label-correcting run (>= R)
  uint64_t edge_checks;  // neighbour inspections by this run (>= W)
  int64_t nsup;          // supernodes counted by the reference rule (reported value is nsup + 1)
} oracle_result;

}

namespace {

struct Worker {
  int32_t n;
  std::vector<uint32_t> label;   // epoch labels: value - off is the max intermediate id
  std::vector<int32_t> mark;     // last source that claimed (src, v)
  std::vector<int32_t> queue, next_queue;
  std::vector<int32_t> row;      // F row of the current source (unsorted)
  uint32_t off;
  explicit Worker(int32_t n_) : n(n_), label(n_, UINT32_MAX), mark(n_, -1) {
    off = UINT32_MAX - (uint32_t)n_;
  }
  // Start a new source: move the epoch so every stored label reads as "infinity".
  // The reference moves max_id_offset down by n per iteration and re-initialises all labels
  // when it would underflow (...:3116-3146).
  void new_epoch() {
    if (off < (uint32_t)n + (uint32_t)n) {
      std::fill(label.begin(), label.end(), UINT32_MAX);
      off = UINT32_MAX - (uint32_t)n;
    } else {
      off -= (uint32_t)n;
    }
  }
};

struct SourceStats {
  int64_t nzU;
  uint64_t fills, expansions, edge_checks, W, R, nnzF;
};

// One source row: the reference's seed phase then its level loop, restated sequentially with
// level-synchronous FIFO queues (the fixed point does not depend on the order).
SourceStats traverse(Worker& w, const int32_t* beg, const int32_t* end, const int32_t* adj, int32_t src) {
  SourceStats st{0, 0, 0, 0, 0, 0, 0};
  w.new_epoch();
  const uint32_t off = w.off;
  uint32_t* label = w.label.data();
  int32_t* mark = w.mark.data();
  w.row.clear();
  w.queue.clear();
  // seed (...:842-906)
  label[src] = off;
  if (mark[src] != src) { mark[src] = src; w.row.push_back(src); }
  const int64_t deg_src = (int64_t)end[src] - beg[src];
  st.W += (uint64_t)deg_src;
  for (int64_t k = beg[src]; k < end[src]; ++k) {
    const int32_t nb = adj[k];
    label[nb] = off;
    if (mark[nb] != src) { mark[nb] = src; w.row.push_back(nb); }
    if (nb < src) w.queue.push_back(nb);
    else if (nb > src) st.nzU++;
  }
  // Count each reached vertex of R_i once (for W and |R|): a vertex v < src is in R_i iff its
  // label is finite at the end. Seeds w < src with duplicates are counted once via the label
  // pass below.
  // level loop (...:909-1027)
  while (!w.queue.empty()) {
    w.next_queue.clear();
    for (int32_t front : w.queue) {
      st.expansions++;
      const uint32_t lf = label[front];
      const uint32_t cost = std::max((uint32_t)front + off, lf);
      for (int64_t k = beg[front]; k < end[front]; ++k) {
        const int32_t m = adj[k];
        st.edge_checks++;
        if (label[m] > cost) {
          label[m] = cost;
          if (m < src) w.next_queue.push_back(m);
          if ((uint32_t)m + off > cost) {
            if (mark[m] != src) {
              mark[m] = src;
              w.row.push_back(m);
              st.fills++;
              if (m > src) st.nzU++;
            }
          }
        }
      }
    }
    std::swap(w.queue, w.next_queue);
  }
  st.nnzF = w.row.size();
  return st;
}

// |R_i| and W_i from the final labels: every v < src with a finite label, src excluded.
// Needs the reach set, so a plain reachability pass (identical set) is used.
void reach_stats(Worker& w, const int32_t* beg, const int32_t* end, const int32_t* adj, int32_t src,
                 std::vector<int32_t>& stack, std::vector<int32_t>& seen, int32_t stamp, uint64_t& R, uint64_t& W) {
  stack.clear();
  seen[src] = stamp;
  for (int64_t k = beg[src]; k < end[src]; ++k) {
    const int32_t nb = adj[k];
    if (nb < src && seen[nb] != stamp) { seen[nb] = stamp; stack.push_back(nb); }
  }
  while (!stack.empty()) {
    const int32_t v = stack.back();
    stack.pop_back();
    R++;
    W += (uint64_t)((int64_t)end[v] - beg[v]);
    for (int64_t k = beg[v]; k < end[v]; ++k) {
      const int32_t m = adj[k];
      if (m < src && seen[m] != stamp) { seen[m] = stamp; stack.push_back(m); }
    }
  }
  (void)w;
}

}  // namespace

extern "C" {

// Full restatement over all n sources, chunked by S rows; OpenMP over chunks, dynamic.
//   nzU        out[n] or null
//   row_nnz    out[n] or null: |F row i| including the diagonal
//   F_rowptr   in[n+1] or null: if given together with F_col, the sorted F rows are written
//              at F_col[F_rowptr[i] .. F_rowptr[i+1])
//   sup_flag   out[n] or null: 1 where the supernode count increments
//   stats      1 = also compute W and |R| (a second reachability pass per source)
int oracle_symbfact(int32_t n, const int32_t* beg, const int32_t* end, const int32_t* adj, int32_t S,
                    int32_t nthreads, int32_t stats, int32_t* nzU, int64_t* row_nnz, const int64_t* F_rowptr,
                    int32_t* F_col, uint8_t* sup_flag, oracle_result* res) {
  if (n <= 0 || S <= 0 || !beg || !end || !adj || !res) return 1;
  std::memset(res, 0, sizeof(*res));
  const int64_t nchunks = ((int64_t)n + S - 1) / S;
  std::atomic<uint64_t> fills{0}, nnzF{0}, Wsum{0}, Rsum{0}, exps{0}, checks{0};
  std::atomic<int64_t> nsup{0};
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    Worker w(n);
    std::vector<int32_t> stack, seen(stats ? n : 0, -1);
    std::vector<int32_t> chunk_nzU(S);
    uint64_t lf = 0, lnnz = 0, lW = 0, lR = 0, le = 0, lc = 0;
    int64_t lsup = 0;
#pragma omp for schedule(dynamic, 1)
    for (int64_t c = 0; c < nchunks; ++c) {
      const int32_t r0 = (int32_t)(c * S);
      const int32_t r1 = (int32_t)std::min<int64_t>((int64_t)r0 + S, n);
      int32_t start = -1;
      for (int32_t r = r0; r < r1; ++r) {
        SourceStats st = traverse(w, beg, end, adj, r);
        lf += st.fills;
        lnnz += st.nnzF;
        le += st.expansions;
        lc += st.edge_checks;
        if (stats) {
          uint64_t R = 0, W = (uint64_t)((int64_t)end[r] - beg[r]);
          reach_stats(w, beg, end, adj, r, stack, seen, r, R, W);
          lR += R;
          lW += W;
        }
        if (nzU) nzU[r] = (int32_t)st.nzU;
        if (row_nnz) row_nnz[r] = (int64_t)st.nnzF;
        if (F_rowptr && F_col) {
          std::sort(w.row.begin(), w.row.end());
          std::copy(w.row.begin(), w.row.end(), F_col + F_rowptr[r]);
        }
        chunk_nzU[r - r0] = (int32_t)st.nzU;
        // supernode rule (...:1833-1950 phase 1, ...:1961-2111 phase 2)
        const int32_t offc = r - r0;
        const bool is_start = (offc % 128 == 0) || (chunk_nzU[offc - 1] - chunk_nzU[offc] != 1);
        uint8_t flag = 0;
        if (is_start) {
          start = r;
          flag = 1;
        } else if (w.mark[start] != r) {  // (r, start) not in F: one more supernode
          flag = 1;
        }
        lsup += flag;
        if (sup_flag) sup_flag[r] = flag;
      }
    }
    fills += lf; nnzF += lnnz; Wsum += lW; Rsum += lR; exps += le; checks += lc; nsup += lsup;
  }
  res->fill_count = fills;
  res->nnz_F = nnzF;
  res->W = Wsum;
  res->R = Rsum;
  res->expansions = exps;
  res->edge_checks = checks;
  res->nsup = nsup;
  return 0;
}

// The same traversal for an explicit subset of sources (bounded CPU-baseline samples).
// Returns edge checks and fills over the subset; no supernodes.
int oracle_sources(int32_t n, const int32_t* beg, const int32_t* end, const int32_t* adj, const int32_t* srcs,
                   int64_t nsrc, int32_t nthreads, uint64_t* fills_out, uint64_t* checks_out) {
  if (n <= 0 || !srcs || nsrc < 0) return 1;
  std::atomic<uint64_t> fills{0}, checks{0};
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    Worker w(n);
    uint64_t lf = 0, lc = 0;
#pragma omp for schedule(dynamic, 16)
    for (int64_t k = 0; k < nsrc; ++k) {
      SourceStats st = traverse(w, beg, end, adj, srcs[k]);
      lf += st.fills;
      lc += st.edge_checks;
    }
    fills += lf;
    checks += lc;
  }
  if (fills_out) *fills_out = fills;
  if (checks_out) *checks_out = checks;
  return 0;
}

// Schedule-independent work of a subset of sources: sum of W_i and |R_i| (reachability only).
int oracle_sources_w(int32_t n, const int32_t* beg, const int32_t* end, const int32_t* adj, const int32_t* srcs,
                     int64_t nsrc, int32_t nthreads, uint64_t* w_out, uint64_t* r_out) {
  if (n <= 0 || !srcs || nsrc < 0) return 1;
  std::atomic<uint64_t> wsum{0}, rsum{0};
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    Worker w(n);
    std::vector<int32_t> stack, seen(n, -1);
    uint64_t lw = 0, lr = 0;
#pragma omp for schedule(dynamic, 16)
    for (int64_t k = 0; k < nsrc; ++k) {
      const int32_t s = srcs[k];
      uint64_t R = 0, W = (uint64_t)((int64_t)end[s] - beg[s]);
      reach_stats(w, beg, end, adj, s, stack, seen, s, R, W);
      lw += W;
      lr += R;
    }
    wsum += lw;
    rsum += lr;
  }
  if (w_out) *w_out = wsum;
  if (r_out) *r_out = rsum;
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
