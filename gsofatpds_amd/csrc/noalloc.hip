// Allocation-free entry point: the counts-only symbolic factorisation of one matrix with every
// device buffer beyond the caller's inputs and outputs carved from one static device array per
// GPU. For callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc). It enqueues the same kernels as a plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "engine.h"

namespace gsofa {
namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;  // per GPU, held until the process ends
constexpr int32_t kMaxRows = 65536;
constexpr int64_t kMaxNnz = int64_t(1) << 20;
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

}  // namespace
}  // namespace gsofa

// The scratch memory. In host code this name is a stand-in for the symbol, not a device address:
// every host use goes through the per-device address hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[gsofa::kScratchBytes];

namespace gsofa {
namespace {

struct DeviceSlot {
  bool ready = false;
  unsigned char* base = nullptr;  // this device's copy of scratch
  DeviceInfo info;
};
DeviceSlot g_slots[kMaxDevices];
std::mutex g_slots_mu;

// Once per device, with that device current: its properties and its address of scratch.
int device_slot(int device, DeviceSlot** out) {
  if (device < 0 || device >= kMaxDevices) return fail(GSOFA_ERR_NO_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_slots_mu);
  DeviceSlot& d = g_slots[device];
  if (!d.ready) {
    int rc = device_info(device, d.info);
    if (rc != GSOFA_OK) return rc;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GSOFA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.base = static_cast<unsigned char*>(p);
    d.ready = true;
  }
  *out = &d;
  return GSOFA_OK;
}

// Consecutive 256-byte aligned pieces of the scratch; with base == nullptr it only measures.
struct Carve {
  unsigned char* base;
  size_t off = 0;
  explicit Carve(unsigned char* b) : base(b) {}
  template <typename T>
  T* take(int64_t count) {
    const size_t at = off;
    off = (off + (size_t)std::max<int64_t>(count, 1) * sizeof(T) + kAlign - 1) & ~(kAlign - 1);
    return base ? reinterpret_cast<T*>(base + at) : nullptr;
  }
};

// The buffers of one run, as the engine lays them out for a set of one matrix.
struct Layout {
  int32_t *rowptr_a, *adj_a, *ell_a, *col_cnt, *rowptr_t, *cursor, *adj_t, *ell_t, *block_sums, *total, *cnt_l;
  uint32_t *win, *blk_flag, *bad;
  unsigned long long* acc;
  MatDev* mats;
  int64_t* bases;  // item base [0, 1], row base [2, 3], supernode block base [4, 5], chain block base [6, 7]
  WorkState* ws;
  uint64_t* heavy_items;
  uint8_t* esc;
  int32_t* bucket;
  int32_t* spill;
  size_t fixed_bytes, spill_bytes;
};

void carve_layout(unsigned char* base, int32_t n, int64_t nnz, int64_t nblk, int32_t nb_max, int64_t esc_words,
                  int64_t heavy_grid, int64_t region_ints, Layout& L) {
  Carve c(base);
  L.bad = c.take<uint32_t>(1);
  L.rowptr_a = c.take<int32_t>(n + 1);
  L.adj_a = c.take<int32_t>(nnz);
  L.ell_a = c.take<int32_t>((int64_t)n * kEll);
  L.col_cnt = c.take<int32_t>(n);
  L.rowptr_t = c.take<int32_t>(n + 1);
  L.cursor = c.take<int32_t>(n + 1);
  L.adj_t = c.take<int32_t>(nnz);
  L.ell_t = c.take<int32_t>((int64_t)n * kEll);
  L.block_sums = c.take<int32_t>(n / 4096 + 2);
  L.total = c.take<int32_t>(1);
  L.cnt_l = c.take<int32_t>(n);
  L.win = c.take<uint32_t>((int64_t)n * 4);
  L.acc = c.take<unsigned long long>(kAccSlots);
  L.mats = c.take<MatDev>(1);
  L.bases = c.take<int64_t>(8);
  L.ws = c.take<WorkState>(1);
  L.heavy_items = c.take<uint64_t>(2 * nblk);
  L.esc = c.take<uint8_t>(esc_words * 4);
  L.blk_flag = c.take<uint32_t>(2 * nblk);
  L.bucket = c.take<int32_t>(nb_max);
  L.fixed_bytes = c.off;
  L.spill = c.take<int32_t>(heavy_grid * region_ints);
  L.spill_bytes = c.off - L.fixed_bytes;
}

// Sets *bad unless rowptr/adj is a CSR pattern of n rows whose rows are strictly increasing
// column ids in [0, n), with rowptr[0] == 0 and rowptr[n] == nnz. Reads adj only inside checked
// extents.
__global__ void k_noalloc_check(int32_t n, int64_t nnz, const int32_t* __restrict__ rowptr,
                                const int32_t* __restrict__ adj, uint32_t* __restrict__ bad) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = rowptr[v], e = rowptr[v + 1];
    bool ok = b >= 0 && b <= e && e <= nnz;
    if (v == 0) ok = ok && b == 0;
    if (v == n - 1) ok = ok && e == nnz;
    if (ok) {
      int32_t last = -1;
      for (int64_t p = b; p < e && ok; ++p) {
        const int32_t c = adj[p];
        ok = c > last && c < n;
        last = c;
      }
    }
    if (!ok) atomicOr(bad, 1u);
  }
}

// The pattern into the scratch, or, when the check failed, an empty pattern of n rows (every
// later kernel then runs on valid data and the totals carry the error).
__global__ void k_noalloc_copy(int32_t n, int64_t nnz, const int32_t* __restrict__ rowptr,
                               const int32_t* __restrict__ adj, const uint32_t* __restrict__ bad,
                               int32_t* __restrict__ rowptr_a, int32_t* __restrict__ adj_a) {
  const bool empty = *bad != 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride)
    rowptr_a[i] = empty ? 0 : rowptr[i];
  if (!empty)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride) adj_a[i] = adj[i];
}

// The matrix table and block bases, passed by value so that nothing is staged in host memory.
struct Tables {
  MatDev m;
  int64_t bases[8];
};
__global__ void k_noalloc_tables(Tables t, MatDev* __restrict__ mats, int64_t* __restrict__ bases) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *mats = t.m;
    for (int k = 0; k < 8; ++k) bases[k] = t.bases[k];
  }
}

// Totals in the order gsofa_run_noalloc documents: fill, nnz_F, nnz_L, nnz_U, W, |R|, reported
// supernodes, counted supernodes, input error.
__global__ void k_noalloc_totals(const unsigned long long* __restrict__ acc, const uint32_t* __restrict__ bad,
                                 int32_t n, int32_t n_processes, uint64_t* __restrict__ totals) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    totals[0] = acc[kAccFill];
    totals[1] = acc[kAccNnzL] + acc[kAccNnzU] + (uint64_t)n;
    totals[2] = acc[kAccNnzL];
    totals[3] = acc[kAccNnzU];
    totals[4] = acc[kAccW];
    totals[5] = acc[kAccR];
    totals[6] = acc[kAccNsup] + (uint64_t)(int64_t)n_processes;
    totals[7] = acc[kAccNsup];
    totals[8] = *bad;
  }
}

int hip_error(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? GSOFA_ERR_OUT_OF_MEMORY : GSOFA_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace
}  // namespace gsofa

using namespace gsofa;

#define NA_HIP(x, what)                           \
  do {                                            \
    hipError_t e_ = (x);                          \
    if (e_ != hipSuccess) return hip_error(e_, what); \
  } while (0)

extern "C" {

int gsofa_run_noalloc(int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* adj, const gsofa_options* opt,
                      int32_t* nzu, uint8_t* sup_flags, uint64_t* totals, void* hip_stream) {
  if (!rowptr || !nzu || !sup_flags || !totals || (nnz > 0 && !adj) || !opt)
    return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (n <= 0) return fail(GSOFA_ERR_INVALID_ARG, "n must be positive");
  if (nnz < 0) return fail(GSOFA_ERR_INVALID_ARG, "nnz must not be negative");
  if (n > kMaxRows || nnz > kMaxNnz)
    return fail(GSOFA_ERR_CAPACITY, "the allocation-free entry takes at most 65,536 rows and 1,048,576 entries "
                                    "(its device scratch is 128 MiB); use a plan for larger patterns");
  if (opt->structure) return fail(GSOFA_ERR_INVALID_ARG, "the allocation-free entry produces counts only (structure 0)");
  if (opt->autotune) return fail(GSOFA_ERR_INVALID_ARG, "the allocation-free entry does not autotune");
  if (opt->bitmap_in_hbm == 1) return fail(GSOFA_ERR_INVALID_ARG, "the allocation-free entry keeps bitmaps in LDS");
  if (opt->chunk_size <= 0) return fail(GSOFA_ERR_INVALID_ARG, "chunk size must be positive");
  if (opt->chain_block < 0) return fail(GSOFA_ERR_INVALID_ARG, "chain block must not be negative");
  Options o;
  o.chunk = opt->chunk_size;
  o.n_processes = opt->n_processes;
  o.structure = 0;
  o.device = opt->device;
  o.heavy_threads = opt->heavy_threads;
  o.heavy_wgs_per_cu = opt->heavy_wgs_per_cu;
  o.light_wgs_per_cu = opt->light_wgs_per_cu;
  o.qcap = opt->queue_capacity;
  o.chain_block = opt->chain_block;
  o.pair_directions = opt->pair_directions == 1 ? 1 : 0;  // the pattern is not read on the host
  o.bitmap_in_hbm = 0;
  const bool paired = o.pair_directions == 1;

  NA_HIP(hipSetDevice(o.device), "hipSetDevice");
  DeviceSlot* dev = nullptr;
  int rc = device_slot(o.device, &dev);
  if (rc != GSOFA_OK) return rc;

  // Launch configuration first without a grid cap, then with the bitmap-path grid cut to the
  // spill regions the scratch holds next to the fixed buffers (the block counts below depend on
  // the chain block, which depends on the grid, so the layout is measured for the smallest one).
  LaunchConfig cfg;
  rc = configure_launch(n, n, paired, o, dev->info, 0, cfg);
  if (rc != GSOFA_OK) return rc;
  if (cfg.hbm_bitmap) return fail(GSOFA_ERR_CAPACITY, "pattern needs bitmaps in HBM");
  const int64_t region = heavy_region_ints(n, paired);
  const int64_t esc_words = (2 * (int64_t)n + 3) / 4;
  Layout L;
  {
    // chain blocks: the option's size, or at least 4 rows when it is picked automatically
    const int64_t chain_min = o.chain_block ? o.chain_block : 4;
    const int64_t nblk_max = ((int64_t)n + chain_min - 1) / chain_min;
    carve_layout(nullptr, n, nnz, nblk_max, (int32_t)nblk_max, esc_words, 0, region, L);
    if (L.fixed_bytes >= kScratchBytes) return fail(GSOFA_ERR_CAPACITY, "pattern does not fit the device scratch");
    const int64_t fit = (int64_t)((kScratchBytes - L.fixed_bytes) / (size_t)(region * 4));
    if (fit < 1) return fail(GSOFA_ERR_CAPACITY, "pattern does not fit the device scratch");
    if (fit < cfg.heavy_grid) {
      rc = configure_launch(n, n, paired, o, dev->info, fit, cfg);
      if (rc != GSOFA_OK) return rc;
    }
  }
  const int32_t chain = cfg.chain_block;
  const int64_t nblk = ((int64_t)n + chain - 1) / chain;
  const int32_t nb_max = (int32_t)nblk;
  const int64_t nitems = 2 * (((int64_t)n + cfg.light_block - 1) / cfg.light_block);
  const int64_t nblocks_sup = ((int64_t)n + 255) / 256;
  carve_layout(dev->base, n, nnz, nblk, nb_max, esc_words, cfg.heavy_grid, region, L);
  if (L.fixed_bytes + L.spill_bytes > kScratchBytes)
    return fail(GSOFA_ERR_CAPACITY, "pattern does not fit the device scratch");

  Tables t{};
  MatDev& M = t.m;
  M.n = n;
  M.chunk = o.chunk;
  M.ell[0] = L.ell_a;
  M.ell[1] = L.ell_t;
  M.rowptr[0] = L.rowptr_a;
  M.rowptr[1] = L.rowptr_t;
  M.adj[0] = L.adj_a;
  M.adj[1] = L.adj_t;
  M.cnt[0] = nzu;  // rows are duplicate-free, so nzU is the strict U count
  M.cnt[1] = L.cnt_l;
  M.win = L.win;
  M.dup_u = nullptr;
  M.dup_all = nullptr;
  M.list_ptr[0] = M.list_ptr[1] = nullptr;
  M.pool[0] = M.pool[1] = nullptr;
  M.sup_flag = sup_flags;
  M.acc = L.acc;
  M.esc[0] = L.esc;
  M.esc[1] = L.esc + n;
  M.blk_flag[0] = L.blk_flag;
  M.blk_flag[1] = L.blk_flag + nblk;
  M.chain_block = chain;
  M.paired = paired ? 1 : 0;
  M.light_block = cfg.light_block;
  M.light_end = n;
  M.subtree = nullptr;  // no pre-escalation: the subtree sizes need the host's elimination tree
  M.pre_size = INT32_MAX;
  M.light_limit = n <= kLightSmallRows ? kLightLimitSmall
                  : cfg.light_bits == kLightBitsBig ? kLightLimitBig : kLightLimitMax;
  const int64_t bases[8] = {0, nitems, 0, n, 0, nblocks_sup, 0, nblk};
  for (int k = 0; k < 8; ++k) t.bases[k] = bases[k];
  int64_t* item_base = L.bases;
  int64_t* row_base = L.bases + 2;
  int64_t* blk_base = L.bases + 4;
  int64_t* blkq_base = L.bases + 6;

  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  const int g = (int)std::min<int64_t>(4096, ((int64_t)std::max<int64_t>(n + 1, nnz) + 255) / 256);
  NA_HIP(launch_fill_i32(reinterpret_cast<int32_t*>(L.bad), 1, 0, s), "clear input flag");
  hipLaunchKernelGGL(k_noalloc_check, dim3(std::min(4096, (n + 255) / 256)), dim3(256), 0, s, n, nnz, rowptr, adj,
                     L.bad);
  hipLaunchKernelGGL(k_noalloc_copy, dim3(g), dim3(256), 0, s, n, nnz, rowptr, adj, L.bad, L.rowptr_a, L.adj_a);
  hipLaunchKernelGGL(k_noalloc_tables, dim3(1), dim3(64), 0, s, t, L.mats, L.bases);
  NA_HIP(hipGetLastError(), "input kernels");
  // the launch sequence of a plan, without its allocation
  Sequence q;
  q.nrows = n;
  q.nitems = nitems;
  q.nblk_total = nblk;
  q.nblocks_sup = nblocks_sup;
  q.esc_words = esc_words;
  q.nmats = 1;
  q.nb_max = nb_max;
  q.n_max = n;
  q.paired_any = paired ? 1 : 0;
  q.cfg = cfg;
  q.rowptr_a = L.rowptr_a;
  q.adj_a = L.adj_a;
  q.ell_a = L.ell_a;
  q.col_cnt = L.col_cnt;
  q.rowptr_t = L.rowptr_t;
  q.cursor = L.cursor;
  q.adj_t = L.adj_t;
  q.ell_t = L.ell_t;
  q.block_sums = L.block_sums;
  q.total = L.total;
  q.bucket = L.bucket;
  q.spill = L.spill;
  q.acc = L.acc;
  q.ws = L.ws;
  q.heavy_items = L.heavy_items;
  q.esc = L.esc;
  q.blk_flag = L.blk_flag;
  q.hbm_bm = nullptr;
  q.mats = L.mats;
  q.item_base = item_base;
  q.row_base = row_base;
  q.blk_base = blk_base;
  q.blkq_base = blkq_base;
  rc = enqueue_sequence(q, s, nullptr);
  if (rc != GSOFA_OK) return rc;
  hipLaunchKernelGGL(k_noalloc_totals, dim3(1), dim3(64), 0, s, L.acc, L.bad, n, o.n_processes, totals);
  NA_HIP(hipGetLastError(), "k_noalloc_totals");
  return GSOFA_OK;
}

}  // extern "C"
