// Host engine: prepared patterns, device state of a set of matrices, and the launch sequence.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "gsofa_internal.h"
#include "kernels.h"

namespace gsofa {

// A caller pattern after the reference's input step (row sort, ...cuh:2214-2227) plus
// de-duplication: packed rowptr/adj with strictly increasing rows. dup_u[i] keeps the number of
// duplicate entries > i that the reference would still count in its nzU (...cuh:896-901).
struct Prepared {
  int32_t n = 0;
  std::vector<int32_t> rowptr;  // n + 1
  std::vector<int32_t> adj;
  std::vector<int32_t> dup_u;   // empty when the input had no duplicate upper entries
  std::vector<int32_t> dup_all; // input row length minus de-duplicated length; empty if none
  int64_t hold_prev = 0;        // rows s >= 1 whose row holds s-1
  bool symmetric = false;       // every off-diagonal (i, j) has its (j, i)
  // Subtree sizes of the elimination tree of A + A^T (|T[s]| bounds |R_s| + 1), computed on
  // first use (block cuts) and kept.
  mutable std::vector<int32_t> subtree;
};

// Rows where chain blocks are cut into two workgroup-path items: in each block the row s whose
// elimination subtree is at least 1.5 times its predecessor's and at least
// max(16 * chain_block, n / divisor) rows (the start of a separator over a large subtree: a
// traversal of that subtree begins there), the largest such; -1 where none. Returns the count.
int64_t block_cuts(const Prepared& P, int32_t chain_block, int32_t divisor, std::vector<int32_t>& cut);
constexpr int32_t kCutDivisor = 64;  // the automatic divisor
// Automatic cuts apply to matrices holding at least 1 / kCutShare of the set's rows. Measured on
// 256 matrices of 10,000-200,000 rows in one set: cutting every member 149 -> 160 ms per run.
constexpr int32_t kCutShare = 4;
// The start of a matrix's top separator among its block cuts (`cut`, from block_cuts): the cut row
// with the largest subtree if that subtree holds at least half the rows, else -1. The items of the
// rows from there on are the longest of the workgroup path (each begins with a traversal of
// nearly the whole matrix); they need no other item's results and no lower row reads their
// escalation flags, so with Options::early_top they run in a launch of their own that starts
// beside the wave path.
int32_t top_separator_rows(const Prepared& P, const std::vector<int32_t>& cut);
// Pre-escalation applies to unmirrored matrices whose off-diagonal entries (sampled) have their
// transpose at least this often (the measured stencils: 85-90 %).
constexpr double kPreSymmetry = 0.75;
// The wave path's escalation limit for a matrix of n rows under a hash-set variant.
int32_t light_limit_for(int32_t n, int light_bits);
// At most this many items go to the early launch (one workgroup each), and this many per
// direction of a matrix (each over several chain blocks).
constexpr int64_t kTopMaxItems = 32;
constexpr int64_t kTopItemsPerDir = 3;

// Whether the workgroup path runs both directions of a block together (Options::pair_directions).
bool pair_directions(const Prepared& P, int32_t mode);
struct Options;
// Whether a matrix is traversed from A only, with L mirrored from U (Options::symmetric).
bool mirrored(const Prepared& P, const Options& opt);

int prepare(const gsofa_csr* a, Prepared& out);

struct Options {
  int32_t chunk = 128;          // supernode chunk size S
  int32_t n_processes = 1;      // reported supernodes = count + n_processes (...cuh:3549)
  int32_t structure = 0;        // produce the explicit L and U structure
  int32_t device = 0;
  int32_t heavy_threads = 0;    // 0 = tuned
  int32_t heavy_wgs_per_cu = 0; // 0 = tuned
  int32_t light_wgs_per_cu = 0; // 0 = tuned
  int32_t qcap = 0;             // LDS queue entries per level buffer, 0 = tuned
  int32_t chain_block = 0;      // rows per workgroup-path item; 1 disables chaining; 0 = auto
  int32_t pair_directions = -1; // 1 on, 0 off, -1 where few rows hold their predecessor
  int32_t autotune = 0;         // pick shape and block size by timing candidates on the set
  int32_t bitmap_in_hbm = 0;    // bitmap-path bitmaps in HBM: 1 always, 0 only when LDS is too small
  int32_t cu_share = 0;         // size persistent grids for this many CUs (0 = all): lets
                                // concurrent sets on other streams overlap on the device
  int32_t split_blocks = 0;     // cut blocks at large separator starts (block_cuts): 0 automatic
                                // (kCutDivisor, matrices of at least 1 / kCutShare of the set's
                                // rows), k > 0 every matrix with subtrees of at least n / k rows,
                                // -1 never
  int32_t symmetric = 0;        // 0 traverse structurally symmetric patterns once (L mirrored from
                                // U), -1 always both directions
  int32_t pre_escalate = 0;     // 0: rows of unpaired matrices whose elimination subtree exceeds
                                // the escalation limit skip the wave path; -1 never
  int32_t early_top = 0;        // 1: a large matrix's top-separator items start beside the wave
                                // path (top_separator_rows); 0 or -1: not
};

struct MatResult {
  uint64_t fill = 0, W = 0, R = 0, nnz_u = 0, nnz_l = 0, heavy = 0;
  uint64_t heavy_w = 0, heavy_r = 0, heavy_nnz = 0;  // the share of the workgroup path
  uint64_t exec[6] = {0, 0, 0, 0, 0, 0};            // executed edges/rows/outputs: heavy, light
  uint64_t exec_top[3] = {0, 0, 0};                 // the early launch's share of exec[0..2]
  int64_t nsup = 0;
};

struct DeviceInfo {
  int cus = 0;
  size_t lds_per_block = 0;
  size_t lds_per_cu = 0;
  std::string arch;
};
int device_info(int device, DeviceInfo& info);

// Launch configuration of a set: bitmap-path workgroup shape, its LDS split, the grids and the
// block sizes of both paths.
struct LaunchConfig {
  int heavy_threads = 1024, heavy_grid = 0, heavy_wgs_per_cu = 0, light_grid = 0;
  int light_bits = kLightBits;  // wave path hash set size (kLightBits or kLightBitsBig)
  int32_t qcap = 0, ucap = 0, bw_max = 0, chain_block = 0, light_block = kLightBlockMin;
  bool hbm_bitmap = false;
  size_t heavy_lds = 0;  // dynamic LDS bytes of k_heavy
};
// Picks the configuration for a set whose largest matrix has n_max rows (nrows in all) and sets
// k_heavy's LDS limit. max_heavy_grid > 0 caps the bitmap-path workgroups (each holds a spill
// region of heavy_region_ints(n_max) ints), for callers whose regions live in a fixed buffer.
int configure_launch(int32_t n_max, int64_t nrows, bool paired_any, const Options& opt, const DeviceInfo& info,
                     int64_t max_heavy_grid, LaunchConfig& out);

// The device buffers and sizes of one launch sequence: a device set's, or the allocation-free
// entry's (which carves the same buffers from a static array).
struct Sequence {
  int64_t nrows = 0, nitems = 0, nblk_total = 0, nblocks_sup = 0, esc_words = 0;
  int nmats = 0;
  int32_t nb_max = 0, n_max = 0;
  int32_t paired_any = 0;
  int32_t prepare = 1;         // 0: the ELL rows and A^T are already built (enqueue_preparation)
  int32_t mirror_any = 0;      // some matrix is mirrored (k_mirror runs)
  int32_t mirror_all = 0;      // every matrix is mirrored (A^T is not built)
  LaunchConfig cfg;
  int32_t *rowptr_a = nullptr, *adj_a = nullptr, *ell_a = nullptr, *col_cnt = nullptr, *rowptr_t = nullptr;
  int32_t *cursor = nullptr, *adj_t = nullptr, *ell_t = nullptr, *block_sums = nullptr, *total = nullptr;
  int32_t *bucket = nullptr, *spill = nullptr;
  unsigned long long* acc = nullptr;
  WorkState* ws = nullptr;
  uint64_t* heavy_items = nullptr;
  uint8_t* esc = nullptr;
  uint32_t *blk_flag = nullptr, *hbm_bm = nullptr;
  MatDev* mats = nullptr;
  int64_t *item_base = nullptr, *row_base = nullptr, *blk_base = nullptr, *blkq_base = nullptr;
  // The early launch (n_top > 0): top_grid workgroups of k_heavy over the fixed items top_items
  // on `side`, forked from the run's stream after its clears and joined before k_mirror and
  // k_supernodes; its own work state and spill regions. top_ev (optional, 2 events) brackets it.
  int64_t n_top = 0;
  int top_grid = 0;
  const uint64_t* top_items = nullptr;
  WorkState* ws_top = nullptr;
  int32_t* spill_top = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t* top_ev = nullptr;
};
// Enqueues the whole run on `s` (kernels only: no allocation, no memset node, no host
// synchronisation; with an early launch, also on q.side, joined back into `s`). ev (optional, 5
// events): before the run, after its clears (and the preparation when q.prepare), after the wave
// path and the workgroup-path queue, after the main workgroup-path launch, after the join,
// k_mirror and supernodes.
int enqueue_sequence(const Sequence& q, hipStream_t s, hipEvent_t* ev);
// The input's device forms: ELL rows of A, and A^T with its ELL rows unless every matrix is
// mirrored. Part of enqueue_sequence when q.prepare is set.
int enqueue_preparation(const Sequence& q, hipStream_t s);

// Device state for one matrix (plain plan) or many (batch); all matrices share one launch of
// each kernel.
class DeviceSet {
 public:
  DeviceSet() = default;
  DeviceSet(const DeviceSet&) = delete;  // owns device memory, a stream and events
  DeviceSet& operator=(const DeviceSet&) = delete;
  ~DeviceSet();
  // alloc_stream != null: stream-ordered allocations and uploads on that stream (no device-wide
  // synchronisation; used by the server), freed on the same stream by the destructor.
  int create(const std::vector<const Prepared*>& mats, const Options& opt, hipStream_t alloc_stream = nullptr);
  // Device bytes create() will allocate for these matrices (admission by footprint).
  static int64_t footprint(const std::vector<const Prepared*>& mats, const Options& opt, int64_t* out_bytes);
  // Enqueue the whole symbolic factorisation on `s` (no host synchronisation, no allocation,
  // no memset nodes: safe to capture in a graph).
  int enqueue(hipStream_t s, hipEvent_t* stage_events = nullptr);
  // The launch sequence over this set's buffers (the preparation included).
  Sequence sequence() const;
  // After the stream completed: read per-matrix totals.
  int read_results(std::vector<MatResult>& out);
  int read_counts(int mat, int32_t* cnt_u, int32_t* cnt_l, uint8_t* sup_flag, int32_t* nzu);
  // Structure of matrix `mat`: strict U rows (CSR) and strict L columns (CSC), sorted.
  int read_structure(int mat, std::vector<int64_t>& u_ptr, std::vector<int32_t>& u_idx, std::vector<int64_t>& l_ptr,
                     std::vector<int32_t>& l_idx);
  bool structure_ready() const { return pools_sized_; }
  int size_pools_from_counts();
  int nmats() const { return nmats_; }
  const std::vector<int32_t>& ns() const { return ns_; }
  const Options& options() const { return opt_; }
  int heavy_threads() const { return heavy_threads_; }
  int heavy_grid() const { return heavy_grid_; }
  int light_grid() const { return light_grid_; }
  int qcap() const { return qcap_; }
  int heavy_wgs_per_cu() const { return heavy_wgs_per_cu_; }
  bool bitmap_in_hbm() const { return hbm_bitmap_; }
  // Items of the early launch (0: none) and, after a run enqueued with stage events, its time.
  int64_t top_items() const { return n_top_; }
  int top_ms(double* ms);

 private:
  template <typename T>
  int alloc(T*& p, size_t count, const char* what);
  int hip_fail(hipError_t e, const char* what);

  Options opt_;
  int nmats_ = 0;
  std::vector<int32_t> ns_;
  std::vector<int64_t> row_base_h_, item_base_h_, blk_base_h_;
  std::vector<int64_t> nnz_base_h_;
  std::vector<int32_t> dup_any_;
  int64_t nrows_ = 0, nnz_ = 0, nitems_ = 0, nblocks_sup_ = 0;
  int32_t n_max_ = 0;
  int heavy_threads_ = 1024, heavy_grid_ = 0, light_grid_ = 0;
  int32_t qcap_ = 0, bw_max_ = 0;
  bool pools_sized_ = false;

  // device buffers (owned)
  int32_t *rowptr_a_ = nullptr, *adj_a_ = nullptr, *ell_a_ = nullptr;
  int32_t *col_cnt_ = nullptr, *rowptr_t_ = nullptr, *cursor_ = nullptr, *adj_t_ = nullptr, *ell_t_ = nullptr;
  int32_t *block_sums_ = nullptr, *total_ = nullptr;
  int32_t *cnt_u_ = nullptr, *cnt_l_ = nullptr, *dup_u_ = nullptr, *dup_all_ = nullptr;
  uint32_t* win_ = nullptr;
  uint8_t* sup_flag_ = nullptr;
  unsigned long long* acc_ = nullptr;
  MatDev* mats_d_ = nullptr;
  int64_t *item_base_d_ = nullptr, *row_base_d_ = nullptr, *blk_base_d_ = nullptr;
  WorkState* ws_ = nullptr;
  uint64_t* heavy_items_ = nullptr;
  int32_t* spill_ = nullptr;
  uint8_t* esc_ = nullptr;        // 2 * nrows escalation flags (padded to a multiple of 4)
  uint32_t* blk_flag_ = nullptr;  // 2 * nblk_total
  int32_t* cut_ = nullptr;        // nblk_total: block cut rows (block_cuts), when any
  LaunchConfig cfg_;
  int64_t nblk_total_ = 0, esc_words_ = 0;
  std::vector<int64_t> blkq_base_h_;  // per matrix, first block of its chain blocks
  int64_t* blkq_base_d_ = nullptr;
  int32_t paired_any_ = 0;            // some matrix runs paired workgroup-path items
  int32_t mirror_any_ = 0, mirror_all_ = 0;  // mirrored matrices (Options::symmetric)
  std::vector<uint8_t> mirror_;       // per matrix
  int heavy_wgs_per_cu_ = 0;          // resident bitmap-path workgroups per CU
  int32_t light_block_ = kLightBlockMin;  // rows per wave-path item
  int32_t nb_max_ = 0;                // chain blocks of the largest matrix
  int32_t* bucket_ = nullptr;         // nb_max: escalated blocks per block index
  bool hbm_bitmap_ = false;           // bitmap-path bitmaps in HBM (hbm_bm_), not LDS
  uint32_t* hbm_bm_ = nullptr;        // heavy_grid * bw_max words
  int32_t ucap_ = 0;                  // bitmap-path U-list entries per direction kept in LDS
  int64_t *list_ptr_u_ = nullptr, *list_ptr_l_ = nullptr;  // nrows + nmats entries
  int32_t *pool_u_ = nullptr, *pool_l_ = nullptr;
  std::vector<int64_t> ptr_base_h_;                        // matrix b's offsets start at row_base + b
  std::vector<int64_t> pool_base_u_, pool_base_l_;        // per matrix, into the pools
  std::vector<std::vector<int64_t>> u_ptr_h_, l_ptr_h_;    // host copies of the offsets
  std::vector<MatDev> mats_h_;
  std::vector<void*> owned_;
  hipStream_t alloc_stream_ = nullptr;
  // the early launch of top-separator items (top_separator_rows)
  int64_t n_top_ = 0;
  int top_grid_ = 0;
  uint64_t* top_items_ = nullptr;
  int32_t* subtree_d_ = nullptr;  // nrows: subtree sizes for pre-escalation (when some matrix uses it)
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr, top_ev_[2] = {nullptr, nullptr};
  bool top_timed_ = false;  // the last run recorded top_ev_
  std::vector<int32_t> light_end_;  // per matrix: MatDev::light_end
  std::vector<std::vector<int32_t>> staging_;  // host copies kept alive for async uploads
  int upload(void* dst, const void* src, size_t bytes, const char* what);
  int download(void* dst, const void* src, size_t bytes, const char* what);
};

}  // namespace gsofa
