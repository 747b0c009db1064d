// Host engine: input preparation, device state and the launch sequence of one set of matrices.
#include "engine.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>

namespace gsofa {

int prepare(const gsofa_csr* a, Prepared& out) {
  int rc = validate_csr(a);
  if (rc != GSOFA_OK) return rc;
  if (a->nnz > (int64_t)INT32_MAX - 64) return fail(GSOFA_ERR_CAPACITY, "nnz must stay below 2^31 - 64");
  const int32_t n = a->n;
  out.n = n;
  out.rowptr.assign((size_t)n + 1, 0);
  out.adj.clear();
  out.adj.reserve((size_t)a->nnz);
  out.dup_u.clear();
  out.hold_prev = 0;
  std::vector<int32_t> row;
  bool any_dup = false;
  std::vector<int32_t> dup((size_t)n, 0), dup_all((size_t)n, 0);
  for (int32_t v = 0; v < n; ++v) {
    row.assign(a->adj + a->beg[v], a->adj + a->end[v]);
    std::sort(row.begin(), row.end());
    int32_t upper_all = 0;
    for (int32_t c : row) upper_all += c > v;
    const int32_t len_in = (int32_t)row.size();
    row.erase(std::unique(row.begin(), row.end()), row.end());
    dup_all[v] = len_in - (int32_t)row.size();
    int32_t upper_distinct = 0;
    for (int32_t c : row) upper_distinct += c > v;
    dup[v] = upper_all - upper_distinct;
    if (v > 0 && std::binary_search(row.begin(), row.end(), v - 1)) ++out.hold_prev;
    any_dup |= dup_all[v] != 0;
    out.adj.insert(out.adj.end(), row.begin(), row.end());
    out.rowptr[v + 1] = (int32_t)out.adj.size();
  }
  if (any_dup) {
    out.dup_u.swap(dup);
    out.dup_all.swap(dup_all);
  }
  // structural symmetry of the de-duplicated pattern (rows are sorted: one search per entry)
  out.symmetric = true;
  for (int32_t v = 0; v < n && out.symmetric; ++v)
    for (int32_t k = out.rowptr[v]; k < out.rowptr[v + 1]; ++k) {
      const int32_t c = out.adj[k];
      if (c != v && !std::binary_search(out.adj.begin() + out.rowptr[c], out.adj.begin() + out.rowptr[c + 1], v)) {
        out.symmetric = false;
        break;
      }
    }
  return GSOFA_OK;
}

bool mirrored(const Prepared& P, const Options& opt) { return opt.symmetric != -1 && P.symmetric; }

// Elimination tree of A + A^T by Liu's algorithm with path compression: every pair {k, i}, k < i,
// of A's off-diagonal entries is taken at step i. Then subtree sizes.
static void subtree_sizes(const Prepared& P, std::vector<int32_t>& size) {
  const int32_t n = P.n;
  std::vector<int32_t> cnt((size_t)n + 1, 0);
  for (int32_t r = 0; r < n; ++r)
    for (int32_t k = P.rowptr[r]; k < P.rowptr[r + 1]; ++k) {
      const int32_t c = P.adj[k];
      if (c != r) ++cnt[(size_t)std::max(c, r) + 1];
    }
  for (int32_t i = 0; i < n; ++i) cnt[(size_t)i + 1] += cnt[i];
  std::vector<int32_t> lower((size_t)cnt[n]);
  {
    std::vector<int32_t> cur(cnt.begin(), cnt.end() - 1);
    for (int32_t r = 0; r < n; ++r)
      for (int32_t k = P.rowptr[r]; k < P.rowptr[r + 1]; ++k) {
        const int32_t c = P.adj[k];
        if (c != r) lower[(size_t)cur[std::max(c, r)]++] = std::min(c, r);
      }
  }
  std::vector<int32_t> parent((size_t)n, -1), anc((size_t)n, -1);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t p = cnt[i]; p < cnt[(size_t)i + 1]; ++p) {
      int32_t r = lower[p];
      while (anc[r] != -1 && anc[r] != i) {
        const int32_t next = anc[r];
        anc[r] = i;
        r = next;
      }
      if (anc[r] == -1) {
        anc[r] = i;
        parent[r] = i;
      }
    }
  size.assign((size_t)n, 1);
  for (int32_t v = 0; v < n; ++v)
    if (parent[v] >= 0) size[parent[v]] += size[v];
}

int64_t block_cuts(const Prepared& P, int32_t chain_block, int32_t divisor, std::vector<int32_t>& cut) {
  const int32_t n = P.n;
  const int64_t nb = ((int64_t)n + chain_block - 1) / chain_block;
  cut.assign((size_t)nb, -1);
  const int64_t min_size = std::max<int64_t>(16 * (int64_t)chain_block, n / std::max(divisor, 1));
  if (chain_block < 2 || n < 2 * min_size) return 0;
  if (P.subtree.empty()) subtree_sizes(P, P.subtree);
  const std::vector<int32_t>& size = P.subtree;
  int64_t count = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int32_t r0 = (int32_t)(b * chain_block), r1 = (int32_t)std::min<int64_t>(n, r0 + (int64_t)chain_block);
    int32_t best = -1;
    for (int32_t s = r0 + 1; s < r1; ++s)
      if (2 * (int64_t)size[s] >= 3 * (int64_t)size[s - 1] && size[s] >= min_size && (best < 0 || size[s] > size[best]))
        best = s;
    if (best >= 0) {
      cut[(size_t)b] = best;
      ++count;
    }
  }
  return count;
}

// Share of off-diagonal entries (i, j) whose (j, i) is present, over every 8th row (rows are sorted).
static double symmetric_share(const Prepared& P) {
  int64_t seen = 0, matched = 0;
  for (int32_t v = 0; v < P.n; v += 8)
    for (int32_t k = P.rowptr[v]; k < P.rowptr[v + 1]; ++k) {
      const int32_t c = P.adj[k];
      if (c == v) continue;
      ++seen;
      matched += std::binary_search(P.adj.begin() + P.rowptr[c], P.adj.begin() + P.rowptr[c + 1], v) ? 1 : 0;
    }
  return seen ? (double)matched / (double)seen : 1.0;
}

int32_t light_limit_for(int32_t n, int light_bits) {
  return n <= kLightSmallRows ? kLightLimitSmall : light_bits == kLightBitsBig ? kLightLimitBig : kLightLimitMax;
}

int32_t top_separator_rows(const Prepared& P, const std::vector<int32_t>& cut) {
  if (P.subtree.empty()) return -1;  // no cut was computed
  int32_t best = -1;
  for (int32_t c : cut)
    if (c >= 0 && (best < 0 || P.subtree[c] > P.subtree[best])) best = c;
  return best >= 0 && 2 * (int64_t)P.subtree[best] >= P.n ? best : -1;
}

// Paired items pay a bitmap save per row and win where chaining by "row s holds s-1" is rare
// (unstructured patterns); structured ones (stencils in nested-dissection order) chain directly.
bool pair_directions(const Prepared& P, int32_t mode) {
  if (mode >= 0) return mode > 0;
  return P.n > 1 && 4 * P.hold_prev < (int64_t)P.n - 1;
}

int device_info(int device, DeviceInfo& info) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(GSOFA_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= count) return fail(GSOFA_ERR_NO_DEVICE, "device ordinal out of range");
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipGetDeviceProperties");
  info.cus = p.multiProcessorCount;
  info.lds_per_block = p.sharedMemPerBlock;
  info.lds_per_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : p.sharedMemPerBlock;
  info.arch = p.gcnArchName;
  if (info.arch.rfind("gfx950", 0) != 0)
    return fail(GSOFA_ERR_NO_DEVICE, "engine is built for gfx950, device is " + info.arch);
  return GSOFA_OK;
}

DeviceSet::~DeviceSet() {
  if (side_) (void)hipStreamSynchronize(side_);
  for (hipEvent_t e : {fork_, join_, top_ev_[0], top_ev_[1]})
    if (e) (void)hipEventDestroy(e);
  if (side_) (void)hipStreamDestroy(side_);
  for (void* p : owned_) {
    if (alloc_stream_) (void)hipFreeAsync(p, alloc_stream_);
    else (void)hipFree(p);
  }
}

int DeviceSet::download(void* dst, const void* src, size_t bytes, const char* what) {
  if (bytes == 0) return GSOFA_OK;
  hipError_t e;
  if (alloc_stream_) {
    e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, alloc_stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(alloc_stream_);
  } else {
    e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) return hip_fail(e, what);
  return GSOFA_OK;
}

int DeviceSet::upload(void* dst, const void* src, size_t bytes, const char* what) {
  if (bytes == 0) return GSOFA_OK;
  hipError_t e = alloc_stream_ ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, alloc_stream_)
                               : hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, what);
  return GSOFA_OK;
}

int DeviceSet::hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? GSOFA_ERR_OUT_OF_MEMORY : GSOFA_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
int DeviceSet::alloc(T*& p, size_t count, const char* what) {
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = alloc_stream_ ? hipMallocAsync(&q, bytes, alloc_stream_) : hipMalloc(&q, bytes);
  if (e != hipSuccess) return hip_fail(e, what);
  owned_.push_back(q);
  p = static_cast<T*>(q);
  return GSOFA_OK;
}

#define GSOFA_TRY(x)                 \
  do {                               \
    int rc_ = (x);                   \
    if (rc_ != GSOFA_OK) return rc_; \
  } while (0)
#define GSOFA_HIP(x, what)                      \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// Whether the bitmap path keeps its bitmaps in HBM: when asked for, or when one bitmap does not
// fit in a workgroup's LDS next to the static part and the smallest queues.
static bool place_bitmaps_in_hbm(int32_t n_max, const Options& opt, const DeviceInfo& info) {
  const size_t bitmap_bytes = (((size_t)n_max + 31) / 32) * 4;
  return opt.bitmap_in_hbm == 1 ||
         bitmap_bytes + kHeavyLdsReserve + heavy_stash_bytes(true) + 256 * 8 > info.lds_per_block;
}

int configure_launch(int32_t n_max, int64_t nrows, bool paired_any, const Options& opt, const DeviceInfo& info,
                     int64_t max_heavy_grid, LaunchConfig& out) {
  out = LaunchConfig();
  out.bw_max = (n_max + 31) / 32;
  // static LDS and the window's rows (per direction) besides the bitmap, queues and U lists
  const size_t reserve = kHeavyLdsReserve + heavy_stash_bytes(paired_any);
  // Bitmaps in HBM when asked for, or when one does not fit in a workgroup's LDS next to the
  // smallest queues (n above about 1.2 M rows): slower claims (L2 atomics), no size limit.
  out.hbm_bitmap = place_bitmaps_in_hbm(n_max, opt, info);
  const size_t bitmap_bytes = out.hbm_bitmap ? 0 : (size_t)out.bw_max * 4;
  // Workgroup size of the bitmap path. Registers allow 16 waves per CU whatever the size, so the
  // choice trades rows in flight per CU (more, smaller workgroups) against frontier width per
  // pass. Measured: unstructured sets (paired items) and small patterns run fastest as four
  // 256-thread workgroups per CU, structured ones as two of 512 threads up to 150,000 rows,
  // larger ones (whose bitmap leaves LDS for one) as one of 1024.
  auto lds_fits = [&](int wgs) { return (size_t)wgs * (bitmap_bytes + reserve + 256 * 8) <= info.lds_per_cu; };
  int threads;
  if (opt.heavy_threads) threads = opt.heavy_threads;
  else if (out.hbm_bitmap) threads = 1024;
  else if ((paired_any || n_max <= 16384) && lds_fits(4)) threads = 256;
  else if (n_max <= 150000 && lds_fits(2)) threads = 512;
  else threads = 1024;
  if (threads != 256 && threads != 512 && threads != 1024)
    return fail(GSOFA_ERR_INVALID_ARG, "heavy_threads must be 256, 512 or 1024");
  if (out.hbm_bitmap && threads != 1024)
    return fail(GSOFA_ERR_CAPACITY, "bitmaps in HBM run 1024-thread workgroups only");
  out.heavy_threads = threads;
  const int target_wgs = opt.heavy_wgs_per_cu ? opt.heavy_wgs_per_cu : 1024 / threads;
  const size_t budget = std::min(info.lds_per_block, info.lds_per_cu / (size_t)std::max(target_wgs, 1));
  // U lists (explicit structure only): their first 4,096 (or 1,024) entries per direction in LDS
  // when the frontier queues keep at least 2,048 entries per level; the rest stays in HBM.
  const int64_t rem = budget > bitmap_bytes + reserve ? (int64_t)(budget - bitmap_bytes - reserve) : 0;
  const int64_t udirs = paired_any ? 2 : 1;
  if (opt.structure) {
    const int64_t qmin = opt.qcap ? opt.qcap : 2048;
    for (int32_t c : {4096, 1024})
      if (!out.ucap && rem >= 8 * qmin + 4 * udirs * c) out.ucap = c;
  }
  if (opt.qcap) {
    out.qcap = opt.qcap;
  } else {
    int64_t q = rem > 4 * udirs * out.ucap ? (rem - 4 * udirs * out.ucap) / 8 : 256;
    q = std::max<int64_t>(256, std::min<int64_t>(16384, q));
    out.qcap = (int32_t)(q & ~int64_t(63));
  }
  const size_t lds = heavy_lds_bytes(out.bw_max, out.qcap, out.hbm_bitmap, out.ucap, paired_any);
  // lds holds the window's rows already; the static part is the rest
  if (lds + kHeavyLdsReserve > info.lds_per_block)
    return fail(GSOFA_ERR_CAPACITY, "heavy LDS request exceeds the device limit");
  const hipError_t e = heavy_set_lds_limit(threads, lds, out.hbm_bitmap);
  if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipFuncSetAttribute(k_heavy): ") + hipGetErrorString(e));
  const int occ = heavy_occupancy(threads, lds, out.hbm_bitmap);
  if (occ <= 0) return fail(GSOFA_ERR_CAPACITY, "k_heavy cannot be resident with this LDS size");
  const int cus = opt.cu_share > 0 ? std::min(opt.cu_share, info.cus) : info.cus;
  out.heavy_grid = cus * occ;
  if (max_heavy_grid > 0 && out.heavy_grid > max_heavy_grid) out.heavy_grid = (int)max_heavy_grid;
  out.heavy_wgs_per_cu = occ;
  out.heavy_lds = lds;
  // Rows per workgroup-path item: long blocks chain more, short blocks keep every workgroup
  // busy; aim for about two blocks per resident workgroup and direction on the largest matrix.
  // (Measured on the unstructured batch members: paired items, which carry both directions,
  // do best with the same rule; one block per workgroup left the tail to the costliest ones.)
  // Capped at 64 rows, or 128 with one workgroup per CU or paired items: measured best at 128 on
  // the 659,033-row stencil (1024 x 1), at 64 on the 110,592-row one (512 x 2, 128 rows left
  // too few blocks), at 80 = n / (2 grid) on a 165,532-row random member.
  out.chain_block = opt.chain_block;
  if (out.chain_block == 0) {
    const int64_t cap = (occ == 1 || paired_any) ? 128 : 64;
    int64_t k = (int64_t)n_max / std::max<int64_t>(1, 2 * (int64_t)out.heavy_grid);
    out.chain_block = (int32_t)std::max<int64_t>(4, std::min<int64_t>(cap, k));
  }
  out.light_bits = n_max > kLightBigRows ? kLightBitsBig : kLightBits;
  const int occ_l = light_occupancy(out.light_bits);
  if (occ_l <= 0) return fail(GSOFA_ERR_HIP, "k_light occupancy query failed");
  out.light_grid = cus * (opt.light_wgs_per_cu ? std::min(opt.light_wgs_per_cu, occ_l) : occ_l);
  const int64_t waves = (int64_t)out.light_grid * kLightWaves;
  out.light_block = kLightBlockMin;
  while (out.light_block < kLightBlockMax && 2 * nrows / (2 * out.light_block) >= 4 * waves) out.light_block *= 2;
  return GSOFA_OK;
}

int enqueue_preparation(const Sequence& q, hipStream_t s) {
  auto hf = [](hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? GSOFA_ERR_OUT_OF_MEMORY : GSOFA_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e = launch_ell_build(q.nrows, q.nmats, q.row_base, q.rowptr_a, q.adj_a, q.ell_a, s);
  if (e != hipSuccess) return hf(e, "ell A");
  if (!q.mirror_all) {  // A^T: traversed from every matrix that is not mirrored
    e = launch_transpose(q.nrows, q.nmats, q.row_base, q.rowptr_a, q.adj_a, q.col_cnt, q.rowptr_t, q.cursor, q.adj_t,
                         q.block_sums, q.total, s);
    if (e != hipSuccess) return hf(e, "transpose");
    e = launch_ell_build(q.nrows, q.nmats, q.row_base, q.rowptr_t, q.adj_t, q.ell_t, s);
    if (e != hipSuccess) return hf(e, "ell A^T");
  }
  return GSOFA_OK;
}

int enqueue_sequence(const Sequence& q, hipStream_t s, hipEvent_t* ev) {
  auto hf = [](hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? GSOFA_ERR_OUT_OF_MEMORY : GSOFA_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
  };
#define GSOFA_SEQ(x, what)                      \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) return hf(e_, what);  \
  } while (0)
  if (ev) GSOFA_SEQ(hipEventRecord(ev[0], s), "hipEventRecord");
  GSOFA_SEQ(launch_clear_run(q.acc, (int64_t)q.nmats * kAccSlots, reinterpret_cast<unsigned long long*>(q.ws), 4,
                             reinterpret_cast<int32_t*>(q.esc), q.esc_words, reinterpret_cast<int32_t*>(q.blk_flag),
                             2 * q.nblk_total, q.bucket, q.nb_max, s),
            "clear the run's state");
  if (q.prepare) GSOFA_TRY(enqueue_preparation(q, s));
  if (ev) GSOFA_SEQ(hipEventRecord(ev[1], s), "hipEventRecord");
  if (q.n_top > 0) {
    // the early launch: the top separators' items, beside the wave path
    GSOFA_SEQ(hipEventRecord(q.fork, s), "hipEventRecord(fork)");
    GSOFA_SEQ(hipStreamWaitEvent(q.side, q.fork, 0), "hipStreamWaitEvent(fork)");
    GSOFA_SEQ(launch_top_setup(q.mats, q.nmats, q.ws_top, q.n_top, q.side), "early launch set-up");
    if (q.top_ev) GSOFA_SEQ(hipEventRecord(q.top_ev[0], q.side), "hipEventRecord");
    GSOFA_SEQ(launch_heavy(q.cfg.heavy_threads, q.top_grid, q.mats, q.ws_top, q.top_items, q.cfg.bw_max, q.cfg.qcap,
                           q.n_max, q.paired_any, q.spill_top, nullptr, q.cfg.ucap, kAccTopExecEdges, q.side),
              "k_heavy (early launch)");
    if (q.top_ev) GSOFA_SEQ(hipEventRecord(q.top_ev[1], q.side), "hipEventRecord");
    GSOFA_SEQ(hipEventRecord(q.join, q.side), "hipEventRecord(join)");
  }
  GSOFA_SEQ(launch_light(q.mats, q.nmats, q.item_base, q.nitems, q.ws, q.cfg.light_grid, q.cfg.light_bits, s), "k_light");
  // cursor, block_sums and total are free again once the transpose is built
  GSOFA_SEQ(launch_heavy_order(q.mats, q.nblk_total, q.nmats, q.blkq_base, q.blk_flag, q.nb_max, q.bucket, q.cursor,
                               q.block_sums, q.total, q.ws, q.heavy_items, s),
            "workgroup-path queue");
  if (ev) GSOFA_SEQ(hipEventRecord(ev[2], s), "hipEventRecord");
  GSOFA_SEQ(launch_heavy(q.cfg.heavy_threads, q.cfg.heavy_grid, q.mats, q.ws, q.heavy_items, q.cfg.bw_max, q.cfg.qcap,
                         q.n_max, q.paired_any, q.spill, q.hbm_bm, q.cfg.ucap, kAccHeavyExecEdges, s),
            "k_heavy");
  if (ev) GSOFA_SEQ(hipEventRecord(ev[3], s), "hipEventRecord");
  if (q.n_top > 0) GSOFA_SEQ(hipStreamWaitEvent(s, q.join, 0), "hipStreamWaitEvent(join)");
  if (q.mirror_any) GSOFA_SEQ(launch_mirror(q.mats, q.nmats, q.blk_base, q.nblocks_sup, s), "k_mirror");
  GSOFA_SEQ(launch_supernodes(q.mats, q.nmats, q.blk_base, q.nblocks_sup, s), "k_supernodes");
  if (ev) GSOFA_SEQ(hipEventRecord(ev[4], s), "hipEventRecord");
#undef GSOFA_SEQ
  return GSOFA_OK;
}

int DeviceSet::create(const std::vector<const Prepared*>& mats, const Options& opt, hipStream_t alloc_stream) {
  opt_ = opt;
  alloc_stream_ = alloc_stream;
  nmats_ = (int)mats.size();
  if (nmats_ <= 0) return fail(GSOFA_ERR_INVALID_ARG, "empty matrix set");
  if (nmats_ > (1 << 20)) return fail(GSOFA_ERR_CAPACITY, "too many matrices in one set");
  if (opt.chunk <= 0) return fail(GSOFA_ERR_INVALID_ARG, "chunk size must be positive");
  if (opt.chain_block < 0) return fail(GSOFA_ERR_INVALID_ARG, "chain block must not be negative");
  DeviceInfo info;
  GSOFA_TRY(device_info(opt.device, info));
  GSOFA_HIP(hipSetDevice(opt.device), "hipSetDevice");

  ns_.resize(nmats_);
  row_base_h_.assign(nmats_ + 1, 0);
  nnz_base_h_.assign(nmats_ + 1, 0);
  item_base_h_.assign(nmats_ + 1, 0);
  blk_base_h_.assign(nmats_ + 1, 0);
  n_max_ = 0;
  bool any_dup = false;
  mirror_.assign((size_t)nmats_, 0);
  for (int b = 0; b < nmats_; ++b) {
    const Prepared& P = *mats[b];
    ns_[b] = P.n;
    n_max_ = std::max(n_max_, P.n);
    row_base_h_[b + 1] = row_base_h_[b] + P.n;
    nnz_base_h_[b + 1] = nnz_base_h_[b] + (int64_t)P.adj.size();
    blk_base_h_[b + 1] = blk_base_h_[b] + ((int64_t)P.n + 255) / 256;
    any_dup |= !P.dup_u.empty();
    mirror_[b] = mirrored(P, opt) ? 1 : 0;
    paired_any_ |= (!mirror_[b] && pair_directions(P, opt.pair_directions)) ? 1 : 0;
  }
  mirror_any_ = 0;
  mirror_all_ = 1;
  for (int b = 0; b < nmats_; ++b) {
    mirror_any_ |= mirror_[b];
    mirror_all_ &= mirror_[b];
  }
  nrows_ = row_base_h_[nmats_];
  nnz_ = nnz_base_h_[nmats_];
  nblocks_sup_ = blk_base_h_[nmats_];
  if (nnz_ > (int64_t)INT32_MAX - 64) return fail(GSOFA_ERR_CAPACITY, "total nnz of a set must stay below 2^31 - 64");

  // ---- launch configuration ----
  LaunchConfig cfg;
  GSOFA_TRY(configure_launch(n_max_, nrows_, paired_any_ != 0, opt, info, 0, cfg));
  cfg_ = cfg;
  bw_max_ = cfg.bw_max;
  hbm_bitmap_ = cfg.hbm_bitmap;
  heavy_threads_ = cfg.heavy_threads;
  heavy_grid_ = cfg.heavy_grid;
  heavy_wgs_per_cu_ = cfg.heavy_wgs_per_cu;
  light_grid_ = cfg.light_grid;
  light_block_ = cfg.light_block;
  qcap_ = cfg.qcap;
  ucap_ = cfg.ucap;
  opt_.chain_block = cfg.chain_block;
  blkq_base_h_.assign(nmats_ + 1, 0);
  for (int b = 0; b < nmats_; ++b)
    blkq_base_h_[b + 1] = blkq_base_h_[b] + ((int64_t)ns_[b] + opt_.chain_block - 1) / opt_.chain_block;
  nblk_total_ = blkq_base_h_[nmats_];
  nb_max_ = (int32_t)(((int64_t)n_max_ + opt_.chain_block - 1) / opt_.chain_block);
  // blocks cut at large separator starts (their parts are separate workgroup-path items)
  std::vector<int32_t> cut_all;
  std::vector<bool> has_cut((size_t)nmats_, false);
  int64_t ncut = 0;
  if (opt.split_blocks != -1) {
    std::vector<int32_t> cut;
    for (int b = 0; b < nmats_; ++b) {
      // automatic: only matrices of at least a quarter of the set's rows, whose longest item can
      // bound the launch; in a set of many small matrices the launch is bound by total work,
      // which a cut adds to (each part traverses its rows' subtrees afresh)
      if (opt.split_blocks == 0 && (int64_t)ns_[b] * kCutShare < nrows_) continue;
      const int64_t c = block_cuts(*mats[b], opt_.chain_block, opt.split_blocks > 0 ? opt.split_blocks : kCutDivisor, cut);
      if (c == 0) continue;
      if (cut_all.empty()) cut_all.assign((size_t)nblk_total_, -1);
      std::copy(cut.begin(), cut.end(), cut_all.begin() + blkq_base_h_[b]);
      has_cut[b] = true;
      ncut += c;
    }
  }
  // The early launch (on request): the top separator of large matrices (at least 1 / kCutShare
  // of the set's rows) when the wave path uses its 2,048-slot variant, whose items are taken on
  // demand (a few CUs held by the early launch then cost it little more than their share) and
  // the bitmaps live in LDS. Nothing orders its dispatch before the wave path's: when the wave
  // path fills the CUs first, the early launch runs after it, beside the main launch, and the
  // run takes longer than without it.
  light_end_.assign(ns_.begin(), ns_.end());
  std::vector<uint64_t> top_items;
  if (opt.early_top > 0 && cfg_.light_bits == kLightBitsBig && !hbm_bitmap_) {
    std::vector<int32_t> cut;
    for (int b = 0; b < nmats_; ++b) {
      if (!has_cut[b] || (int64_t)ns_[b] * kCutShare < nrows_) continue;
      cut.assign(cut_all.begin() + blkq_base_h_[b], cut_all.begin() + blkq_base_h_[b + 1]);
      const int32_t top = top_separator_rows(*mats[b], cut);
      if (top < 0) continue;
      const int32_t K = opt_.chain_block;
      // kTopItemsPerDir items per direction (a mirrored or paired matrix: one item for both),
      // each over e consecutive blocks from the cut on: every item begins with a traversal of
      // nearly the whole matrix and chains its other rows, so fewer, longer items execute fewer
      // traversals; the launch runs beside the wave path and the main launch, not before them
      const int ndir = (mirror_[b] || pair_directions(*mats[b], opt_.pair_directions)) ? 1 : 2;
      const int64_t b0 = top / K, nblk = ((int64_t)ns_[b] + K - 1) / K - b0;
      const int64_t e = std::min<int64_t>((nblk + kTopItemsPerDir - 1) / kTopItemsPerDir, kItemExtraMax + 1);
      const int64_t nit = (nblk + e - 1) / e;
      if ((int64_t)top_items.size() + ndir * nit > kTopMaxItems) continue;
      for (int64_t k = 0; k < nit; ++k) {
        const uint32_t row = k == 0 ? (uint32_t)top : (uint32_t)((b0 + k * e) * K);
        for (int d = 0; d < ndir; ++d) top_items.push_back(encode_item((uint32_t)b, (uint32_t)d, row, 0u, (uint32_t)(e - 1)));
      }
      light_end_[b] = top;
    }
  }
  n_top_ = (int64_t)top_items.size();
  top_grid_ = (int)std::min<int64_t>(n_top_, heavy_grid_);
  for (int b = 0; b < nmats_; ++b) {
    // wave-path items: two per block (one per direction), or for a mirrored matrix one per block
    // rounded up to even (the next matrix's items keep the parity the sibling exchange pairs by);
    // rows from light_end on are not the wave path's
    const int64_t nb = ((int64_t)light_end_[b] + light_block_ - 1) / light_block_;
    item_base_h_[b + 1] = item_base_h_[b] + (mirror_[b] ? nb + (nb & 1) : 2 * nb);
  }
  nitems_ = item_base_h_[nmats_];

  // ---- allocations ----
  GSOFA_TRY(alloc(rowptr_a_, nrows_ + 1, "rowptr A"));
  GSOFA_TRY(alloc(adj_a_, nnz_, "adj A"));
  GSOFA_TRY(alloc(ell_a_, nrows_ * kEll, "ell A"));
  GSOFA_TRY(alloc(col_cnt_, nrows_, "column counts"));
  GSOFA_TRY(alloc(rowptr_t_, nrows_ + 1, "rowptr A^T"));
  GSOFA_TRY(alloc(cursor_, nrows_ + 1, "transpose cursor"));
  GSOFA_TRY(alloc(adj_t_, nnz_, "adj A^T"));
  GSOFA_TRY(alloc(ell_t_, nrows_ * kEll, "ell A^T"));
  GSOFA_TRY(alloc(block_sums_, nrows_ / 4096 + 2, "scan block sums"));
  GSOFA_TRY(alloc(total_, 1, "scan total"));
  GSOFA_TRY(alloc(cnt_u_, nrows_, "U counts"));
  GSOFA_TRY(alloc(cnt_l_, nrows_, "L counts"));
  GSOFA_TRY(alloc(win_, nrows_ * 4, "window bits"));
  GSOFA_TRY(alloc(sup_flag_, nrows_, "supernode flags"));
  GSOFA_TRY(alloc(acc_, (size_t)nmats_ * kAccSlots, "accumulators"));
  GSOFA_TRY(alloc(mats_d_, nmats_, "matrix table"));
  GSOFA_TRY(alloc(item_base_d_, nmats_ + 1, "item base"));
  GSOFA_TRY(alloc(row_base_d_, nmats_ + 1, "row base"));
  GSOFA_TRY(alloc(blk_base_d_, nmats_ + 1, "block base"));
  GSOFA_TRY(alloc(ws_, n_top_ > 0 ? 2 : 1, "work state"));  // the second: the early launch's
  GSOFA_TRY(alloc(heavy_items_, 2 * (nblk_total_ + ncut), "escalation list"));
  if (ncut > 0) {
    GSOFA_TRY(alloc(cut_, (size_t)nblk_total_, "block cuts"));
    GSOFA_TRY(upload(cut_, cut_all.data(), cut_all.size() * 4, "upload block cuts"));
  }
  // per resident workgroup: two frontier spill queues, the U list (beyond its LDS part) and the
  // saved bitmap of the last row (with paired items: a U list per direction and a ring of the
  // last saved rows' bitmaps per direction)
  // (the early launch's workgroups have regions of their own after the main launch's)
  GSOFA_TRY(alloc(spill_, (size_t)(heavy_grid_ + top_grid_) * (size_t)heavy_region_ints(n_max_, paired_any_ != 0),
                  "heavy spill regions"));
  if (n_top_ > 0) {
    GSOFA_TRY(alloc(top_items_, (size_t)n_top_, "early items"));
    GSOFA_TRY(upload(top_items_, top_items.data(), top_items.size() * 8, "upload early items"));
    GSOFA_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    GSOFA_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreateWithFlags");
    GSOFA_HIP(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreateWithFlags");
    GSOFA_HIP(hipEventCreate(&top_ev_[0]), "hipEventCreate");
    GSOFA_HIP(hipEventCreate(&top_ev_[1]), "hipEventCreate");
  }
  if (hbm_bitmap_) GSOFA_TRY(alloc(hbm_bm_, (size_t)heavy_grid_ * (size_t)bw_max_, "heavy bitmaps"));
  esc_words_ = (2 * nrows_ + 3) / 4;
  GSOFA_TRY(alloc(esc_, (size_t)esc_words_ * 4, "escalation flags"));
  GSOFA_TRY(alloc(blk_flag_, (size_t)(2 * nblk_total_), "block flags"));
  GSOFA_TRY(alloc(blkq_base_d_, nmats_ + 1, "chain block base"));
  GSOFA_TRY(alloc(bucket_, nb_max_, "block index histogram"));
  // Pre-escalation: a row whose elimination subtree of A + A^T holds more vertices than the
  // escalation limit goes to the workgroup path without a wave-path attempt. A mirrored matrix's
  // reached set holds the whole subtree, so the rule is exact there; a directed reach can be
  // smaller, but on the unsymmetric stencils it sent 11 of 98,731 (659,033 rows) and 53 of
  // 76,125 (110,592 rows) workgroup-path traversals there needlessly. Patterns far from
  // symmetric (under kPreSymmetry of the sampled entries matched: one direction of a band
  // reaches almost nothing below its rows, whatever the subtree) and paired matrices
  // (unstructured: reaches far smaller than their subtrees) keep every attempt. Only matrices
  // with more rows than the limit carry the subtree sizes.
  std::vector<int32_t> pre_size((size_t)nmats_, INT32_MAX);
  bool any_pre = false;
  if (opt.pre_escalate != -1) {
    for (int b = 0; b < nmats_; ++b) {
      if (!mirror_[b] && pair_directions(*mats[b], opt_.pair_directions)) continue;
      const int32_t limit = light_limit_for(ns_[b], cfg_.light_bits);
      if (ns_[b] <= limit) continue;
      if (!mirror_[b] && symmetric_share(*mats[b]) < kPreSymmetry) continue;
      pre_size[b] = limit;
      any_pre = true;
    }
  }
  if (any_pre) {
    GSOFA_TRY(alloc(subtree_d_, (size_t)nrows_, "subtree sizes"));
    std::vector<int32_t> sub((size_t)nrows_, 0);
    for (int b = 0; b < nmats_; ++b) {
      if (pre_size[b] == INT32_MAX) continue;
      const Prepared& P = *mats[b];
      if (P.subtree.empty()) subtree_sizes(P, P.subtree);
      std::copy(P.subtree.begin(), P.subtree.end(), sub.begin() + row_base_h_[b]);
    }
    GSOFA_TRY(upload(subtree_d_, sub.data(), sub.size() * 4, "upload subtree sizes"));
    if (alloc_stream_) GSOFA_HIP(hipStreamSynchronize(alloc_stream_), "hipStreamSynchronize(subtree sizes)");
  }
  if (any_dup) {
    GSOFA_TRY(alloc(dup_u_, nrows_, "duplicate surplus (upper)"));
    GSOFA_TRY(alloc(dup_all_, nrows_, "duplicate surplus (row)"));
  }

  // ---- uploads ----
  std::vector<int32_t> rp((size_t)nrows_ + 1);
  std::vector<int32_t> adj_all((size_t)std::max<int64_t>(nnz_, 1));
  std::vector<int32_t> dup_all, dupr_all;
  if (any_dup) {
    dup_all.assign((size_t)nrows_, 0);
    dupr_all.assign((size_t)nrows_, 0);
  }
  for (int b = 0; b < nmats_; ++b) {
    const Prepared& P = *mats[b];
    const int64_t rb = row_base_h_[b], nb = nnz_base_h_[b];
    for (int32_t v = 0; v < P.n; ++v) rp[rb + v] = (int32_t)(nb + P.rowptr[v]);
    std::copy(P.adj.begin(), P.adj.end(), adj_all.begin() + nb);
    if (any_dup && !P.dup_u.empty()) {
      std::copy(P.dup_u.begin(), P.dup_u.end(), dup_all.begin() + rb);
      std::copy(P.dup_all.begin(), P.dup_all.end(), dupr_all.begin() + rb);
    }
  }
  rp[nrows_] = (int32_t)nnz_;
  GSOFA_TRY(upload(rowptr_a_, rp.data(), rp.size() * 4, "upload rowptr"));
  GSOFA_TRY(upload(adj_a_, adj_all.data(), (size_t)nnz_ * 4, "upload adj"));
  if (any_dup) {
    GSOFA_TRY(upload(dup_u_, dup_all.data(), dup_all.size() * 4, "upload dup"));
    GSOFA_TRY(upload(dup_all_, dupr_all.data(), dupr_all.size() * 4, "upload dup"));
  }
  GSOFA_TRY(upload(item_base_d_, item_base_h_.data(), item_base_h_.size() * 8, "upload"));
  GSOFA_TRY(upload(row_base_d_, row_base_h_.data(), row_base_h_.size() * 8, "upload"));
  GSOFA_TRY(upload(blk_base_d_, blk_base_h_.data(), blk_base_h_.size() * 8, "upload"));
  GSOFA_TRY(upload(blkq_base_d_, blkq_base_h_.data(), blkq_base_h_.size() * 8, "upload"));
  if (alloc_stream_) {
    // pageable sources: make sure the copies consumed them before the vectors go away
    GSOFA_HIP(hipStreamSynchronize(alloc_stream_), "hipStreamSynchronize(uploads)");
  }

  mats_h_.assign(nmats_, MatDev{});
  for (int b = 0; b < nmats_; ++b) {
    MatDev& M = mats_h_[b];
    const int64_t rb = row_base_h_[b];
    M.n = ns_[b];
    M.chunk = opt.chunk;
    M.ell[0] = ell_a_ + rb * kEll;
    M.ell[1] = ell_t_ + rb * kEll;
    M.rowptr[0] = rowptr_a_ + rb;
    M.rowptr[1] = rowptr_t_ + rb;
    M.adj[0] = adj_a_;
    M.adj[1] = adj_t_;
    M.cnt[0] = cnt_u_ + rb;
    M.cnt[1] = cnt_l_ + rb;
    M.win = win_ + rb * 4;
    M.dup_u = dup_u_ ? dup_u_ + rb : nullptr;
    M.dup_all = dup_all_ ? dup_all_ + rb : nullptr;
    M.list_ptr[0] = M.list_ptr[1] = nullptr;
    M.pool[0] = M.pool[1] = nullptr;
    M.sup_flag = sup_flag_ + rb;
    M.acc = acc_ + (size_t)b * kAccSlots;
    M.esc[0] = esc_ + rb;
    M.esc[1] = esc_ + nrows_ + rb;
    M.blk_flag[0] = blk_flag_ + blkq_base_h_[b];
    M.blk_flag[1] = blk_flag_ + nblk_total_ + blkq_base_h_[b];
    M.split = has_cut[b] ? cut_ + blkq_base_h_[b] : nullptr;
    M.chain_block = opt_.chain_block;
    M.sym = mirror_[b];
    M.paired = (!mirror_[b] && pair_directions(*mats[b], opt_.pair_directions)) ? 1 : 0;
    M.light_block = light_block_;
    M.light_end = light_end_[b];
    M.light_limit = light_limit_for(ns_[b], cfg_.light_bits);
    M.subtree = pre_size[b] != INT32_MAX ? subtree_d_ + rb : nullptr;
    M.pre_size = pre_size[b];
  }
  GSOFA_TRY(upload(mats_d_, mats_h_.data(), sizeof(MatDev) * nmats_, "upload table"));
  // The pattern is fixed for the set's lifetime, and so are its device forms (ELL rows, A^T): built
  // here once, like the upload, not on every run.
  GSOFA_TRY(enqueue_preparation(sequence(), alloc_stream_));
  GSOFA_HIP(hipStreamSynchronize(alloc_stream_), "hipStreamSynchronize(preparation)");
  return GSOFA_OK;
}

int64_t DeviceSet::footprint(const std::vector<const Prepared*>& mats, const Options& opt, int64_t* out_bytes) {
  // mirrors the allocations of create() (spill sized for the resident heavy grid)
  DeviceInfo info;
  int rc = device_info(opt.device, info);
  if (rc != GSOFA_OK) return rc;
  int64_t nrows = 0, nnz = 0, nblk = 0;
  int32_t nmax = 0;
  bool dup = false, paired = false;
  for (const Prepared* P : mats) {
    nrows += P->n;
    nnz += (int64_t)P->adj.size();
    nmax = std::max(nmax, P->n);
    nblk += ((int64_t)P->n + 3) / 4;  // smallest automatic chain block
    dup |= !P->dup_u.empty();
    paired |= pair_directions(*P, opt.pair_directions);
  }
  const int cus = opt.cu_share > 0 ? std::min(opt.cu_share, info.cus) : info.cus;
  const int64_t heavy_grid = (int64_t)cus * 4;  // upper bound on resident workgroups per CU
  int64_t b = 0;
  b += 4 * (nrows + 1) * 3;           // rowptr A, rowptr A^T, cursor
  b += 4 * nnz * 2;                   // adj, adj^T
  b += 4 * nrows * kEll * 2;          // ELL A, A^T
  b += 4 * nrows * 3 + 16 * nrows;    // column counts, U/L counts, window bits
  b += nrows + 2 * nrows + 4 * 2 * nblk + 8 * 2 * nrows;  // flags, escalations, block flags, items
  b += 4 * nblk + 8 * 2 * nblk;                           // block cuts and their extra items
  b += 4 * (int64_t)nmax + 8 * ((int64_t)mats.size() + 1);  // block index histogram, chain block bases
  b += 4 * (heavy_grid + kTopMaxItems) * heavy_region_ints(nmax, paired);  // spill (and the early launch's)
  if (place_bitmaps_in_hbm(nmax, opt, info)) b += 4 * (int64_t)cus * (((int64_t)nmax + 31) / 32);  // one per CU
  if (dup) b += 8 * nrows;
  if (opt.pre_escalate != -1) b += 4 * nrows;  // subtree sizes (at most one per row)
  if (opt.structure) b += 16 * (nrows + (int64_t)mats.size());  // offsets; pools are sized later
  b += 1 << 20;                       // tables, accumulators, scan scratch
  *out_bytes = b;
  return GSOFA_OK;
}

int DeviceSet::enqueue(hipStream_t s, hipEvent_t* ev) {
  GSOFA_HIP(hipSetDevice(opt_.device), "hipSetDevice");
  Sequence q = sequence();
  q.prepare = 0;  // built once by create()
  q.top_ev = (ev && n_top_ > 0) ? top_ev_ : nullptr;
  top_timed_ = q.top_ev != nullptr;
  return enqueue_sequence(q, s, ev);
}

int DeviceSet::top_ms(double* ms) {
  *ms = 0.0;
  if (!top_timed_) return GSOFA_OK;
  float t = 0.f;
  GSOFA_HIP(hipEventElapsedTime(&t, top_ev_[0], top_ev_[1]), "hipEventElapsedTime(early launch)");
  *ms = t;
  return GSOFA_OK;
}

Sequence DeviceSet::sequence() const {
  Sequence q;
  q.nrows = nrows_;
  q.nitems = nitems_;
  q.nblk_total = nblk_total_;
  q.nblocks_sup = nblocks_sup_;
  q.esc_words = esc_words_;
  q.nmats = nmats_;
  q.nb_max = nb_max_;
  q.n_max = n_max_;
  q.paired_any = paired_any_;
  q.mirror_any = mirror_any_;
  q.mirror_all = mirror_all_;
  q.cfg = cfg_;
  q.rowptr_a = rowptr_a_;
  q.adj_a = adj_a_;
  q.ell_a = ell_a_;
  q.col_cnt = col_cnt_;
  q.rowptr_t = rowptr_t_;
  q.cursor = cursor_;
  q.adj_t = adj_t_;
  q.ell_t = ell_t_;
  q.block_sums = block_sums_;
  q.total = total_;
  q.bucket = bucket_;
  q.spill = spill_;
  q.acc = acc_;
  q.ws = ws_;
  q.heavy_items = heavy_items_;
  q.esc = esc_;
  q.blk_flag = blk_flag_;
  q.hbm_bm = hbm_bm_;
  q.mats = mats_d_;
  q.item_base = item_base_d_;
  q.row_base = row_base_d_;
  q.blk_base = blk_base_d_;
  q.blkq_base = blkq_base_d_;
  q.n_top = n_top_;
  q.top_grid = top_grid_;
  q.top_items = top_items_;
  q.ws_top = n_top_ > 0 ? ws_ + 1 : nullptr;
  q.spill_top = spill_ + (int64_t)heavy_grid_ * heavy_region_ints(n_max_, paired_any_ != 0);
  q.side = side_;
  q.fork = fork_;
  q.join = join_;
  return q;
}

int DeviceSet::read_results(std::vector<MatResult>& out) {
  std::vector<unsigned long long> acc((size_t)nmats_ * kAccSlots);
  GSOFA_TRY(download(acc.data(), acc_, acc.size() * 8, "read accumulators"));
  out.assign(nmats_, MatResult{});
  for (int b = 0; b < nmats_; ++b) {
    const unsigned long long* a = acc.data() + (size_t)b * kAccSlots;
    if (a[kAccOverflow]) return fail(GSOFA_ERR_STATE, "structure offsets do not match this run's counts");
    if (a[kAccScheduleError]) return fail(GSOFA_ERR_STATE, "wave path: a wave stopped waiting for its sibling's item");
    MatResult& r = out[b];
    r.fill = a[kAccFill];
    r.W = a[kAccW];
    r.R = a[kAccR];
    r.nnz_u = a[kAccNnzU];
    r.nnz_l = a[kAccNnzL];
    r.heavy = a[kAccHeavy];
    r.heavy_w = a[kAccHeavyW];
    r.heavy_r = a[kAccHeavyR];
    r.heavy_nnz = a[kAccHeavyNnz];
    for (int k = 0; k < 6; ++k) r.exec[k] = a[kAccHeavyExecEdges + k];
    for (int k = 0; k < 3; ++k) {  // the workgroup path's work: both of its launches
      r.exec_top[k] = a[kAccTopExecEdges + k];
      r.exec[k] += r.exec_top[k];
    }
    r.nsup = (int64_t)a[kAccNsup];
  }
  return GSOFA_OK;
}

int DeviceSet::read_counts(int mat, int32_t* cnt_u, int32_t* cnt_l, uint8_t* sup_flag, int32_t* nzu) {
  if (mat < 0 || mat >= nmats_) return fail(GSOFA_ERR_INVALID_ARG, "matrix index");
  const int64_t rb = row_base_h_[mat];
  const size_t n = (size_t)ns_[mat];
  if (cnt_u || nzu) {
    std::vector<int32_t> tmp(n);
    GSOFA_TRY(download(tmp.data(), cnt_u_ + rb, n * 4, "read U counts"));
    if (cnt_u) std::copy(tmp.begin(), tmp.end(), cnt_u);
    if (nzu) {
      std::vector<int32_t> dup(n, 0);
      if (dup_u_) GSOFA_TRY(download(dup.data(), dup_u_ + rb, n * 4, "read dup"));
      for (size_t i = 0; i < n; ++i) nzu[i] = tmp[i] + dup[i];
    }
  }
  if (cnt_l) GSOFA_TRY(download(cnt_l, cnt_l_ + rb, n * 4, "read L counts"));
  if (sup_flag) GSOFA_TRY(download(sup_flag, sup_flag_ + rb, n, "read supernode flags"));
  return GSOFA_OK;
}

// Exact structure offsets from a completed run's counts: list_ptr = exclusive scan of counts.
int DeviceSet::size_pools_from_counts() {
  ptr_base_h_.assign(nmats_ + 1, 0);
  for (int b = 0; b < nmats_; ++b) ptr_base_h_[b + 1] = ptr_base_h_[b] + ns_[b] + 1;
  const int64_t nptr = ptr_base_h_[nmats_];
  std::vector<int64_t> up((size_t)nptr), lp((size_t)nptr);
  pool_base_u_.assign(nmats_ + 1, 0);
  pool_base_l_.assign(nmats_ + 1, 0);
  u_ptr_h_.assign(nmats_, {});
  l_ptr_h_.assign(nmats_, {});
  std::vector<int32_t> cu((size_t)nrows_), cl((size_t)nrows_);
  GSOFA_TRY(download(cu.data(), cnt_u_, cu.size() * 4, "read U counts"));
  GSOFA_TRY(download(cl.data(), cnt_l_, cl.size() * 4, "read L counts"));
  for (int b = 0; b < nmats_; ++b) {
    const int64_t rb = row_base_h_[b], pb = ptr_base_h_[b];
    int64_t su = 0, sl = 0;
    for (int32_t v = 0; v < ns_[b]; ++v) {
      up[pb + v] = su;
      lp[pb + v] = sl;
      su += cu[rb + v];
      if (!mirror_[b]) sl += cl[rb + v];  // a mirrored matrix's L lists are its U lists
    }
    up[pb + ns_[b]] = su;
    lp[pb + ns_[b]] = sl;
    pool_base_u_[b + 1] = pool_base_u_[b] + su;
    pool_base_l_[b + 1] = pool_base_l_[b] + sl;
    u_ptr_h_[b].assign(up.begin() + pb, up.begin() + pb + ns_[b] + 1);
    l_ptr_h_[b].assign(lp.begin() + pb, lp.begin() + pb + ns_[b] + 1);
    if (mirror_[b]) l_ptr_h_[b] = u_ptr_h_[b];
  }
  if (!list_ptr_u_) {
    GSOFA_TRY(alloc(list_ptr_u_, (size_t)nptr, "U offsets"));
    GSOFA_TRY(alloc(list_ptr_l_, (size_t)nptr, "L offsets"));
    GSOFA_TRY(alloc(pool_u_, (size_t)pool_base_u_[nmats_], "U structure"));
    GSOFA_TRY(alloc(pool_l_, (size_t)pool_base_l_[nmats_], "L structure"));
  }
  GSOFA_TRY(upload(list_ptr_u_, up.data(), up.size() * 8, "upload U offsets"));
  GSOFA_TRY(upload(list_ptr_l_, lp.data(), lp.size() * 8, "upload L offsets"));
  for (int b = 0; b < nmats_; ++b) {
    MatDev& M = mats_h_[b];
    M.list_ptr[0] = list_ptr_u_ + ptr_base_h_[b];
    M.list_ptr[1] = list_ptr_l_ + ptr_base_h_[b];
    M.pool[0] = pool_u_ + pool_base_u_[b];
    M.pool[1] = pool_l_ + pool_base_l_[b];
  }
  GSOFA_TRY(upload(mats_d_, mats_h_.data(), sizeof(MatDev) * nmats_, "upload table"));
  if (alloc_stream_) GSOFA_HIP(hipStreamSynchronize(alloc_stream_), "hipStreamSynchronize(offsets)");
  pools_sized_ = true;
  return GSOFA_OK;
}

int DeviceSet::read_structure(int mat, std::vector<int64_t>& u_ptr, std::vector<int32_t>& u_idx,
                              std::vector<int64_t>& l_ptr, std::vector<int32_t>& l_idx) {
  if (!pools_sized_) return fail(GSOFA_ERR_STATE, "structure was not produced (structure option off)");
  if (mat < 0 || mat >= nmats_) return fail(GSOFA_ERR_INVALID_ARG, "matrix index");
  u_ptr = u_ptr_h_[mat];
  l_ptr = l_ptr_h_[mat];
  const int64_t nu = u_ptr.back(), nl = l_ptr.back();
  u_idx.resize((size_t)nu);
  l_idx.resize((size_t)nl);
  if (nu) GSOFA_TRY(download(u_idx.data(), pool_u_ + pool_base_u_[mat], (size_t)nu * 4, "read U"));
  if (mirror_[mat]) {
    l_idx = u_idx;  // L's column v is U's row v
  } else if (nl) {
    GSOFA_TRY(download(l_idx.data(), pool_l_ + pool_base_l_[mat], (size_t)nl * 4, "read L"));
  }
  const int32_t n = ns_[mat];
  for (int32_t v = 0; v < n; ++v) {
    std::sort(u_idx.begin() + u_ptr[v], u_idx.begin() + u_ptr[v + 1]);
    std::sort(l_idx.begin() + l_ptr[v], l_idx.begin() + l_ptr[v + 1]);
  }
  return GSOFA_OK;
}

}  // namespace gsofa
