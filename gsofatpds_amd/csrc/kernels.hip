// HIP kernels of the symbolic-factorisation engine, written for gfx950 (CDNA4, wave64).
//
// Method. The reference decides row i of F = L+U by a per-source minimax fill-path traversal
// (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:777-1029): (i,m) is in F iff
// m is in row i of A or a path i -> x1 -> ... -> xk -> m exists in G(A) with every
// xj < min(i,m) (Rose-Tarjan). This engine evaluates the same predicate as two plain
// reachability traversals per index:
//   * direction 0, on A from source i through vertices < i: every m > i adjacent to the
//     reached set (or to i) is in U row i; the reached set below i is R_i;
//   * direction 1, on A^T from source j through vertices < j: every m > j adjacent to the
//     reached set (or to j) is in L column j, since the reversed path has all
//     intermediates below j = min(m, j).
// A reachability traversal needs one bit of state per vertex instead of a 32-bit minimax
// label, so a whole traversal of a 1.2M-vertex pattern fits in one CU's LDS (larger patterns
// keep the bitmap in HBM).
//
// Kernels, in launch order for one matrix or one batch:
//   k_fill_u64 / k_fill_i32     clear accumulators and counters (graph-safe, no memset nodes)
//   k_ell_build                 pack rows into 64-byte ELL-16 slots
//   k_col_hist, scan, k_transpose_scatter, k_ell_build   build A^T in the same layout
//   k_light                     one wave per block of rows, LDS hash set; escalates big sources
//   k_heavy_hist, scan, k_heavy_scatter   the workgroup path's queue, costliest blocks first
//   k_heavy                     one workgroup per block of escalated rows, n-bit bitmap
//   k_supernodes                the reference's supernode rule, one thread per row
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.h"

namespace gsofa {

// ------------------------------------------------------------------------------------------
// wave64 helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Pointers that reach a kernel through a struct are generic; these casts make the hot accesses
// global_load / global_store instead of flat ones.
#define GSOFA_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const GSOFA_GLOBAL T*)(p);
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, T v) {
  *(GSOFA_GLOBAL T*)(p) = v;
}
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4i32 gload_v4(const int32_t* p) { return *(const GSOFA_GLOBAL v4i32*)(p); }

__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, o, 64);
    hi = __shfl_xor(hi, o, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// A wave-uniform value moved to a scalar register (keeps per-wave counters out of VGPRs).
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// DPP lane moves: VALU operations, no LDS round trip. A lane whose source lies outside its row
// of 16 (row_shr) reads 0; rows outside kRowMask keep 0 (row_bcast).
template <int kCtrl, int kRowMask = 0xF, bool kBoundZero = true>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, kRowMask, 0xF, kBoundZero);
}

// Inclusive prefix sum over the 64 lanes: a Hillis-Steele scan inside each row of 16 lanes
// (row_shr 1, 2, 4, 8), then lane 15 of each row added into rows 1 and 3 and lane 31 into rows
// 2 and 3 (row_bcast 15 / 31). Every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  x += dpp_u32<0x111>(x);
  x += dpp_u32<0x112>(x);
  x += dpp_u32<0x114>(x);
  x += dpp_u32<0x118>(x);
  x += dpp_u32<0x142, 0xA, false>(x);
  x += dpp_u32<0x143, 0xC, false>(x);
  return x;
}

// Sum over the wave, the same in every lane.
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_u32(x), 63);
}

// OR over the wave, the same in every lane (the scan pattern above with OR).
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t x) {
  x |= dpp_u32<0x111>(x);
  x |= dpp_u32<0x112>(x);
  x |= dpp_u32<0x114>(x);
  x |= dpp_u32<0x118>(x);
  x |= dpp_u32<0x142, 0xA, false>(x);
  x |= dpp_u32<0x143, 0xC, false>(x);
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Lanes of one wave hand data to each other through LDS (one lane writes a count, others read
// it; lanes read queue entries other lanes wrote) between steps that no workgroup barrier
// separates. Release and acquire at wavefront scope around a wave barrier: every access of the
// step before is ordered before every access after, and no shared value is carried over in a
// register.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int find_segment(const int64_t* base, int nseg, int64_t k) {
  // base[0] = 0 <= ... <= base[nseg]; returns b with base[b] <= k < base[b + 1]
  int lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= k) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------------------------------------
// small utility kernels (graph-capturable replacements for memset)
// ------------------------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long* p, int64_t count, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int32_t* p, int64_t count, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// A run's clears in one launch: accumulators and work state (u64), escalation and block flags
// and the workgroup-path queue's buckets (i32 words), all to zero.
__global__ void k_clear_run(unsigned long long* __restrict__ acc, int64_t nacc, unsigned long long* __restrict__ ws,
                            int64_t nws, int32_t* __restrict__ esc, int64_t nesc, int32_t* __restrict__ blk_flag,
                            int64_t nblk, int32_t* __restrict__ bucket, int64_t nbucket) {
  const int64_t total = nacc + nws + nesc + nblk + nbucket;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < nacc) acc[i] = 0ull;
    else if (i < nacc + nws) ws[i - nacc] = 0ull;
    else if (i < nacc + nws + nesc) esc[i - nacc - nws] = 0;
    else if (i < nacc + nws + nesc + nblk) blk_flag[i - nacc - nws - nesc] = 0;
    else bucket[i - nacc - nws - nesc - nblk] = 0;
  }
}

// ------------------------------------------------------------------------------------------
// layout: ELL-16 packing and the transpose
// ------------------------------------------------------------------------------------------
// Row g of the concatenated pattern: entries adj[rowptr[g] .. rowptr[g+1]) (local column ids).
__global__ void k_ell_build(int64_t nrows, int nmats, const int64_t* __restrict__ row_base,
                            const int32_t* __restrict__ rowptr, const int32_t* __restrict__ adj,
                            int32_t* __restrict__ ell) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nrows; g += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = rowptr[g], e = rowptr[g + 1];
    const int32_t deg = e - b;
    int32_t slot[kEll];
#pragma unroll
    for (int k = 0; k < kEll; ++k) slot[k] = kPad;
    if (deg < kEll) {
      // off-diagonal entries first, then the diagonal's mark (at most slot 14)
      const int32_t v = (int32_t)(g - row_base[nmats == 1 ? 0 : find_segment(row_base, nmats, g)]);
      int k = 0;
      bool diag = false;
      for (int p = 0; p < deg; ++p) {
        const int32_t c = adj[b + p];
        if (c == v) diag = true;
        else slot[k++] = c;
      }
      if (diag) slot[k] = kDiag;
    } else if (deg == kEll) {
      for (int k = 0; k < kEll; ++k) slot[k] = adj[b + k];
    } else {
      for (int k = 0; k < kEll - 1; ++k) slot[k] = adj[b + k];
      slot[kEll - 1] = -(b + (kEll - 1) + 2);
    }
    int4* dst = reinterpret_cast<int4*>(ell + g * kEll);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = make_int4(slot[4 * q], slot[4 * q + 1], slot[4 * q + 2], slot[4 * q + 3]);
  }
}

// Column counts of every matrix of the set: one thread per global row.
__global__ void k_col_hist(int64_t nrows, int nmats, const int64_t* __restrict__ row_base,
                           const int32_t* __restrict__ rowptr, const int32_t* __restrict__ adj,
                           int32_t* __restrict__ col_cnt) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nrows; g += (int64_t)gridDim.x * blockDim.x) {
    const int mb = find_segment(row_base, nmats, g);
    const int64_t base = row_base[mb];
    for (int32_t p = rowptr[g]; p < rowptr[g + 1]; ++p) atomicAdd(&col_cnt[base + adj[p]], 1);
  }
}

// cursor[] starts as the exclusive scan of col_cnt (global positions); rows scatter themselves
// into their columns. The order inside a column is not deterministic and does not matter:
// every output is a set or a count.
__global__ void k_transpose_scatter(int64_t nrows, int nmats, const int64_t* __restrict__ row_base,
                                    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ adj,
                                    int32_t* __restrict__ cursor, int32_t* __restrict__ adj_t) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nrows; g += (int64_t)gridDim.x * blockDim.x) {
    const int mb = find_segment(row_base, nmats, g);
    const int64_t base = row_base[mb];
    const int32_t v = (int32_t)(g - base);
    for (int32_t p = rowptr[g]; p < rowptr[g + 1]; ++p) {
      const int32_t pos = atomicAdd(&cursor[base + adj[p]], 1);
      adj_t[pos] = v;
    }
  }
}

// Exclusive scan of int32 counts into int32 positions (three passes; counts fit in int32).
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int32_t* __restrict__ in, int64_t count,
                                                            int32_t* __restrict__ block_sums) {
  __shared__ int32_t s[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock * kScanItems;
  int32_t acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    if (i < count) acc += in[i];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane_id() == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) t += s[w];
    block_sums[blockIdx.x] = t;
  }
}

// Single block: exclusive scan of the block sums in place; writes the grand total to *total.
__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(int32_t* __restrict__ sums, int64_t nblocks,
                                                            int32_t* __restrict__ total) {
  __shared__ int32_t s_wave[kScanBlock / 64];
  __shared__ int32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nblocks; base += kScanBlock) {
    const int64_t i = base + threadIdx.x;
    const int32_t v = (i < nblocks) ? sums[i] : 0;
    int32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      if (lane_id() >= o) x += y;
    }
    if (lane_id() == 63) s_wave[threadIdx.x >> 6] = x;
    __syncthreads();
    int32_t wave_off = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) wave_off += s_wave[w];
    const int32_t carry = s_carry;
    if (i < nblocks) sums[i] = carry + wave_off + x - v;
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) s_carry = carry + wave_off + x;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int32_t* __restrict__ in, int64_t count,
                                                           const int32_t* __restrict__ block_off,
                                                           int32_t* __restrict__ out, int32_t* __restrict__ out2,
                                                           const int32_t* __restrict__ total) {
  __shared__ int32_t s_wave[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock * kScanItems;
  int32_t carry = block_off[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    const int32_t v = (i < count) ? in[i] : 0;
    int32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      if (lane_id() >= o) x += y;
    }
    if (lane_id() == 63) s_wave[threadIdx.x >> 6] = x;
    __syncthreads();
    int32_t wave_off = 0, all = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
      if (w < (int)(threadIdx.x >> 6)) wave_off += s_wave[w];
      all += s_wave[w];
    }
    if (i < count) {
      out[i] = carry + wave_off + x - v;
      if (out2) out2[i] = carry + wave_off + x - v;
    }
    carry += all;
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[count] = *total;
    if (out2) out2[count] = *total;
  }
}

// ------------------------------------------------------------------------------------------
// traversal: shared pieces
// ------------------------------------------------------------------------------------------
struct RowCursor {
  int32_t e[kEll];
  int32_t ov_pos, ov_end;  // overflow part of a long row, read from the CSR adjacency
};

__device__ __forceinline__ void load_row(const MatDev& M, int dir, int32_t v, RowCursor& rc) {
  rc.ov_pos = 0;
  rc.ov_end = 0;
  if (v < 0) {
#pragma unroll
    for (int k = 0; k < kEll; ++k) rc.e[k] = kPad;
    return;
  }
  const int32_t* src = M.ell[dir] + (int64_t)v * kEll;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const v4i32 t = gload_v4(src + 4 * q);
    rc.e[4 * q] = t.x;
    rc.e[4 * q + 1] = t.y;
    rc.e[4 * q + 2] = t.z;
    rc.e[4 * q + 3] = t.w;
  }
  if (rc.e[kEll - 1] <= -2) {
    rc.ov_pos = -(rc.e[kEll - 1] + 2);
    rc.ov_end = gload(M.rowptr[dir] + v + 1);
    rc.e[kEll - 1] = kPad;
  }
}

// 1 when the row's slots mark its diagonal (kDiag): one more entry for degree counts.
__device__ __forceinline__ uint32_t diag_mark(const RowCursor& rc) {
  bool d = false;
#pragma unroll
  for (int k = 0; k < kEll - 1; ++k) d = d || rc.e[k] == kDiag;
  return d ? 1u : 0u;
}

// Whether some entry u < src - 1 of row src (its ELL slots, one per lane 0..15 in sm) holds
// src - 1 in its own ELL slots: then src -> u -> src - 1 stays below src, so src - 1 is in R_src
// and X_src contains X_{src-1}. A sufficient test (CSR tails are not searched); 16 lanes per
// candidate row, four candidates per load round. Wave-uniform call.
__device__ __forceinline__ bool two_hop_prev(const int32_t* __restrict__ ell, int32_t src, int32_t sm) {
  const int lane = lane_id();
  uint64_t cand = __ballot(lane < kEll && sm >= 0 && sm < src - 1);
  while (cand != 0) {
    int32_t u[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      u[c] = -1;
      if (cand != 0) {
        const int p = __ffsll((unsigned long long)cand) - 1;
        cand &= cand - 1;
        u[c] = __builtin_amdgcn_readlane(sm, p);
      }
    }
    const int ci = lane >> 4;
    const int32_t mine = ci == 0 ? u[0] : ci == 1 ? u[1] : ci == 2 ? u[2] : u[3];
    const int32_t e = mine >= 0 ? gload(ell + (int64_t)mine * kEll + (lane & (kEll - 1))) : kPad;
    if (__ballot(e == src - 1) != 0) return true;
  }
  return false;
}

// Window bits of direction 1: column j = src, row m > j, same supernode chunk, m - j - 1 < 128.
// Selects instead of win[t >> 5]: a dynamically indexed register array would live in scratch.
__device__ __forceinline__ void window_bit(uint32_t win[4], int32_t src, int32_t m, int32_t chunk_end) {
  const int32_t t = m - src - 1;
  if (t < 128 && m < chunk_end) {
    const uint32_t b = 1u << (t & 31);
    const int w = t >> 5;
    win[0] |= (w == 0) ? b : 0u;
    win[1] |= (w == 1) ? b : 0u;
    win[2] |= (w == 2) ? b : 0u;
    win[3] |= (w == 3) ? b : 0u;
  }
}

// ------------------------------------------------------------------------------------------
// k_light: one wave per source, LDS hash set of the reached vertices
// ------------------------------------------------------------------------------------------
constexpr int kLightStep = 4;                // ELL slots claimed together between escalation checks
constexpr int kLightVpp = 64 / kEll;         // vertices of a narrow pass (one ELL slot per lane)

// The hash set of a wave path variant: 2^kBits slots, escalation limit (LDS sizing) and queue.
template <int kBits>
struct LightHash {
  static constexpr int kHash = 1 << kBits;
  static constexpr int kLimit = kBits == kLightBitsBig ? kLightLimitBig : kLightLimitMax;
  static constexpr int kQ = kLimit + 64 * kLightStep;  // R entries never exceed the limit plus one step
  static_assert(kLimit + 64 * kLightStep < kHash, "the hash set must never fill");
};

template <int kBits>
__device__ __forceinline__ uint32_t hash_home(int32_t m) { return ((uint32_t)m * 0x9E3779B1u) >> (32 - kBits); }

// Linear probing from slot h. The table holds at most kLimit + 64 * kLightStep < kHash
// entries, so an empty slot is always found.
template <int kBits>
__device__ __forceinline__ int hash_probe(int32_t* H, int32_t m, uint32_t h) {
  constexpr int kHash = 1 << kBits;
  while (true) {
    const int32_t old = atomicCAS(&H[h], -1, m);
    if (old == -1) return 1;
    if (old == m) return 0;
    h = (h + 1) & (kHash - 1);
  }
}
template <int kBits>
__device__ __forceinline__ int hash_insert(int32_t* H, int32_t m) { return hash_probe<kBits>(H, m, hash_home<kBits>(m)); }
// Membership without inserting (the table always keeps an empty slot).
template <int kBits>
__device__ __forceinline__ bool hash_has(const int32_t* H, int32_t m) {
  constexpr int kHash = 1 << kBits;
  uint32_t h = hash_home<kBits>(m);
  for (int k = 0; k < kHash; ++k) {
    const int32_t x = H[h];
    if (x == m) return true;
    if (x == -1) return false;
    h = (h + 1) & (kHash - 1);
  }
  return false;
}

// One wave per item = a block of M.light_block consecutive rows of one direction, in increasing
// order. Row s reuses the previous row's hash set when row s holds s-1 (then X_{s-1} is part of
// X_s). A row that holds an escalated u < s escalates without an attempt: X_u, already over the
// limit, is part of X_s.

template <int kBits>
__global__ __launch_bounds__(kLightWaves * 64) __attribute__((amdgpu_waves_per_eu(kBits == kLightBitsBig ? 3 : 5))) void k_light(
    const MatDev* __restrict__ mats, int nmats, const int64_t* __restrict__ item_base, int64_t nitems,
    WorkState* __restrict__ ws) {
  constexpr int kHash = LightHash<kBits>::kHash;
  constexpr int kLightQ = LightHash<kBits>::kQ;
  __shared__ int32_t s_hash[kLightWaves][kHash];
  __shared__ int32_t s_q[kLightWaves][kLightQ];
  const int wave = threadIdx.x >> 6;
  const int lane = lane_id();
  int32_t* H = s_hash[wave];
  int32_t* Q = s_q[wave];
  for (int i = lane; i < kHash; i += 64) H[i] = -1;
  // Sibling exchange. Waves 2i and 2i+1 run the two directions of one block (items 2j and
  // 2j+1). After each row t a wave publishes whether t+1 is in its X_t. For the other direction
  // that decides exactly whether t is in R_{t+1}: a path t+1 -> ... -> t through vertices below
  // t in one direction is the reverse of a path t -> ... -> t+1 in the other, i.e. t+1 in U_t.
  __shared__ unsigned long long s_pub[kLightWaves];  // (item << 32) | (row + 1) of the last row published
  __shared__ uint8_t s_nxt[kLightWaves][8];          // by row mod 8: 1 in, 2 not in, 3 unknown
  // Items. The 2,048-slot variant takes them in pairs (items 2p and 2p+1: the two directions of
  // a block, or two blocks of a mirrored matrix) from a device counter, lowest rows first, so that
  // waves that finish early take more (a fixed share per wave left the slowest wave 1.26x the mean
  // on the 659,033-row stencil). The even wave of a sibling pair takes the pair and hands it to
  // the odd wave: (fetch number << 32) | pair index; the odd wave acknowledges each fetch it has
  // read, and the even wave takes the next pair only after that acknowledgement. The 1,024-slot
  // variant strides over the items, item k to wave k mod waves (its shorter items measured faster
  // that way: the handshake per item costs more than the imbalance saves).
  __shared__ unsigned long long s_pair[kLightWaves / 2];
  __shared__ uint32_t s_ack[kLightWaves / 2];
  if (threadIdx.x < kLightWaves) s_pub[threadIdx.x] = 0ull;
  if (threadIdx.x < kLightWaves / 2) {
    s_pair[threadIdx.x] = 0ull;
    s_ack[threadIdx.x] = 0u;
  }
  __syncthreads();
  const int sib = wave ^ 1;

  // per-wave running totals, flushed once per matrix (one atomic per wave and slot)
  uint64_t t_fill = 0, t_w = 0, t_r = 0, t_nnz_u = 0, t_nnz_l = 0;
  uint64_t x_edges = 0, x_verts = 0, x_out = 0;  // wave-uniform
  int last_mat = -1;

  constexpr bool kDynamic = kBits == kLightBitsBig;
  const int64_t npairs = nitems / 2;  // nitems is even
  const bool even = (wave & 1) == 0;
  uint32_t fetch = 0;
  const int64_t gwave = (int64_t)blockIdx.x * kLightWaves + wave;
  const int64_t nwaves = (int64_t)gridDim.x * kLightWaves;

  auto flush = [&](int mb) {
    if (mb < 0) return;
    if (lane == 0) {
      unsigned long long* acc = mats[mb].acc;
      if (t_fill) atomicAdd(&acc[kAccFill], t_fill);
      if (t_w) atomicAdd(&acc[kAccW], t_w);
      if (t_r) atomicAdd(&acc[kAccR], t_r);
      if (t_nnz_u) atomicAdd(&acc[kAccNnzU], t_nnz_u);
      if (t_nnz_l) atomicAdd(&acc[kAccNnzL], t_nnz_l);
      if (x_edges) atomicAdd(&acc[kAccLightExecEdges], x_edges);
      if (x_verts) atomicAdd(&acc[kAccLightExecVerts], x_verts);
      if (x_out) atomicAdd(&acc[kAccLightExecOut], x_out);
    }
    t_fill = t_w = t_r = t_nnz_u = t_nnz_l = 0;
    x_edges = x_verts = x_out = 0;
  };

  for (int64_t round = 0;; ++round) {
    int64_t k;
    if constexpr (!kDynamic) {
      k = gwave + round * nwaves;
      if (k >= nitems) break;
    } else {
      int64_t pidx;
      ++fetch;
      // Both waits end once the other wave of the pair finishes the item it is on; they are bounded
      // all the same (a wave that stops waiting ends and the run reports an error).
      if (even) {
        bool ok = true;
        if (fetch > 1) {  // the odd wave has read the previous pair
          ok = false;
          for (int spin = 0; spin < (1 << 24); ++spin) {
            if (__hip_atomic_load(&s_ack[wave >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == fetch - 1) {
              ok = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          ok = __builtin_amdgcn_readfirstlane(ok ? 1 : 0) != 0;
        }
        if (!ok) {
          if (lane == 0 && nmats > 0) atomicOr(&mats[0].acc[kAccScheduleError], 1ull);
          break;
        }
        unsigned long long q = 0;
        if (lane == 0) q = atomicAdd(&ws->light_next, 1ull);
        q = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(q >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)q);
        pidx = q < (unsigned long long)npairs ? (int64_t)q : npairs;
        if (lane == 0) {  // always published, the last time with npairs (done)
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __hip_atomic_store(&s_pair[wave >> 1], ((unsigned long long)fetch << 32) | (uint32_t)pidx, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      } else {
        unsigned long long v = 0;
        bool got = false;
        for (int spin = 0; spin < (1 << 24); ++spin) {
          v = __hip_atomic_load(&s_pair[wave >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if ((uint32_t)(v >> 32) == fetch) {
            got = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        got = __builtin_amdgcn_readfirstlane(got ? 1 : 0) != 0;
        if (!got) {
          if (lane == 0 && nmats > 0) atomicOr(&mats[0].acc[kAccScheduleError], 1ull);
          break;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        pidx = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
        if (lane == 0) __hip_atomic_store(&s_ack[wave >> 1], fetch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      if (pidx >= npairs) break;
      k = 2 * pidx + (even ? 0 : 1);
    }
    const int mb = (nmats == 1) ? 0 : find_segment(item_base, nmats, k);
    if (mb != last_mat) {
      flush(last_mat);
      last_mat = mb;
    }
    const MatDev& M = mats[mb];
    const int64_t local = k - item_base[mb];
    // a mirrored matrix runs direction 0 only, one item per block (its last item may be padding)
    const bool sym = M.sym != 0;
    const int dir = sym ? 0 : (int)(local & 1);
    const int32_t lb = M.light_block;
    const int limit = M.light_limit;  // reached vertices before escalation
    // Lowest rows first, so that their escalations are seen by the rows above (the escalation
    // rules below; measured 3 % faster than highest first on the 659,033-row stencil).
    const int32_t blk = sym ? (int32_t)local : (int32_t)(local >> 1);
    // rows from light_end on belong to the early launch, which runs with the 2,048-slot variant
    // only (the 1,024-slot one keeps its registers: n)
    const int32_t s0 = blk * lb, s1 = min(kBits == kLightBitsBig ? M.light_end : M.n, s0 + lb);
    // the U list is kept where it is read: for the structure, and for the window bits of L's
    // columns (direction 1, or a mirrored matrix's direction 0: L's column s is U's row s)
    const bool track_u = M.list_ptr[dir] != nullptr || dir == 1 || sym;

    // state carried from row s-1 (all wave-uniform; hash H holds X_{s-1} when carry_ok)
    bool carry_ok = false;     // H holds X_{s-1}
    bool prev_esc = false;     // s-1 escalated
    int hcount = 0;            // entries of H
    // The U list of the row in H (entries above its source) sits at the back of Q: Q[kLightQ-1-i],
    // i < uq. New R entries fill Q from the front; both are disjoint subsets of H, whose size
    // never exceeds kLightQ, so they cannot meet.
    int uq = 0;
    uint32_t u_prev = 0;       // |U_{s-1}|
    uint64_t wr_prev = 0, deg_prev = 0, r_prev = 0;
    int self_st = 3;           // mirrored: is src in X_{src-1}? 1 in, 2 not in, 3 unknown
    for (int32_t src = s0; src < s1; ++src) {
      const int32_t chunk_end = min(M.n, (src / M.chunk + 1) * M.chunk);
      // A row whose elimination subtree exceeds pre_size escalates at once: its reached set holds
      // its whole subtree (mirrored matrices), or nearly (pre_size leaves a margin otherwise).
      const bool pre = M.subtree != nullptr && gload(M.subtree + src) > M.pre_size;
      RowCursor rc;
      // the source row, one ELL slot per lane (lanes 0..15); slot 15 <= -2 encodes a CSR tail
      const int32_t sm = lane < kEll ? gload(M.ell[dir] + (int64_t)src * kEll + lane) : kPad;
      // the row's structure offsets, loaded now and used when it finishes (no round trip there);
      // the 1,024-slot variant has no registers to spare and loads them at the finish
      int64_t lp_off = 0, lp_end = 0;
      if (kBits == kLightBitsBig && M.list_ptr[dir] != nullptr) {
        lp_off = gload(M.list_ptr[dir] + src);
        lp_end = gload(M.list_ptr[dir] + src + 1);
      }
      // does row src hold src-1?
      bool has_prev_in_row = false;
      if (src > s0) {
        has_prev_in_row = sm == src - 1;
        if (lane == kEll - 1 && sm <= -2)  // slot 15: a CSR tail
          for (int32_t p = -(sm + 2), pe = gload(M.rowptr[dir] + src + 1); p < pe; ++p)
            has_prev_in_row |= gload(M.adj[dir] + p) == src - 1;
      }
      bool chained = __ballot(has_prev_in_row) != 0;
      // otherwise (when X_{src-1} is at hand or src-1 escalated): from the sibling wave's row
      // src-1 of the other direction, or through an entry whose own row holds src-1
      if (!pre && !chained && src > s0 && (carry_ok || prev_esc)) {
        // Mirrored, the other direction is this one: its own row src-1 answers (src in U_{src-1}).
        int st = sym ? self_st : 3;
        const unsigned long long want = ((unsigned long long)(k ^ 1) << 32) | (uint32_t)src;
        for (int spin = 0; !sym && spin < (1 << 16); ++spin) {  // bounded: an unknown answer is also correct
          const unsigned long long p = __hip_atomic_load(&s_pub[sib], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (p >= want) {
            // Published row src-1 of this block (or went past it). Its ring slot is rewritten when
            // the sibling publishes row src+7, which it may be doing once row src+6 is out: read
            // the slot only while the sibling has published at most row src+5, before and after.
            if ((p >> 32) == (unsigned long long)(k ^ 1) && (uint32_t)p - (uint32_t)src < 7u) {
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
              st = s_nxt[sib][(src - 1) & 7];
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
              const unsigned long long p2 = __hip_atomic_load(&s_pub[sib], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              if ((p2 >> 32) != (unsigned long long)(k ^ 1) || (uint32_t)p2 - (uint32_t)src >= 7u) st = 3;
            }
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        st = __builtin_amdgcn_readfirstlane(st);
        if (st == 1) chained = true;
        else if (st != 2) chained = two_hop_prev(M.ell[dir], src, sm);
      }
      // an entry u < src (ELL slots) whose row escalated already: X_src contains X_u. A flag
      // written by another wave may not be seen yet; that only costs the attempt. Read here,
      // tested after the source row's pass (which hides the load).
      const bool held_esc_lane = !pre && sm >= 0 && sm < src && gload(M.esc[dir] + sm) != 0;
      bool esc = false;
      if (pre || (chained && prev_esc)) {
        esc = true;  // X_src contains the escalated X_{src-1}, or its subtree is over the limit
      } else {
        const bool carry = chained && carry_ok;
        if (!carry) {
          if (hcount > 0)
            for (int i = lane; i < kHash; i += 64) H[i] = -1;
          hcount = 0;
          uq = 0;
        }
        bool src_was_set = false;
        if (lane == 0) src_was_set = hash_insert<kBits>(H, src) == 0;
        src_was_set = __ballot(src_was_set) != 0;
        if (!src_was_set) ++hcount;
        if (track_u && carry && src_was_set) {
          // src was a U entry of src-1: drop it from the carried list (the last entry takes its place)
          for (int base = 0; base < uq; base += 64) {
            const int i = base + lane;
            const uint64_t hit = __ballot(i < uq && Q[kLightQ - 1 - i] == src);
            if (hit) {
              const int at = base + __ffsll((unsigned long long)hit) - 1;
              if (lane == 0) Q[kLightQ - 1 - at] = Q[kLightQ - uq];
              break;
            }
          }
          --uq;
        }
        int qn = 0;
        uint32_t u_new = 0, u_seed = 0;  // per lane, summed at the end of the row
        uint64_t w_new = 0, deg0 = 0;
        int head = 0;
        // Narrow pass for the source row and for batches of at most kLightVpp queued vertices:
        // lane = (vertex, ELL slot), one claim per lane, queue positions from the ballot rank.
        // (The wide path below would run all its slot groups for one or two active lanes.)
        auto claim_one = [&](int32_t m, bool seedpass) {
          const bool valid = m >= 0;
          // entries of the row, with the diagonal's mark
          const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid || m == kDiag));
          const uint32_t nup = (uint32_t)__popcll(__ballot(valid && m > src));
          if (seedpass) {
            deg0 += nvalid;
            if (lane == 0) u_seed += nup;  // entries of A's row above the diagonal
          } else if (lane == 0) {
            w_new += nvalid;
          }
          bool ins = false;
          if (valid) {
            const uint32_t home = hash_home<kBits>(m);
            const int32_t old = atomicCAS(&H[home], -1, m);
            ins = old == -1 || (old != m && hash_probe<kBits>(H, m, (home + 1) & (kHash - 1)) != 0);
          }
          const bool is_r = ins && m < src;
          const uint64_t mr = __ballot(is_r);
          if (is_r) {
            const int pos = qn + (int)rank_below(mr);
            if (pos < kLightQ) Q[pos] = m;
          }
          qn += __popcll(mr);
          hcount += __popcll(__ballot(ins));
          const bool is_u = ins && m > src;
          const uint64_t mu = __ballot(is_u);
          if (track_u) {
            if (is_u) Q[kLightQ - 1 - (uq + (int)rank_below(mu))] = m;
            uq += __popcll(mu);
          }
          if (lane == 0) u_new += (uint32_t)__popcll(mu);
          if (hcount > limit || qn >= kLightQ) esc = true;
        };
        // m: this lane's slot of vertex v (slot 15 <= -2: the row continues in a CSR tail)
        auto narrow = [&](int32_t v, int32_t m, bool seedpass) {
          const bool tail = (lane & (kEll - 1)) == kEll - 1 && m <= -2;
          claim_one(tail ? kPad : m, seedpass);
          uint64_t tm = __ballot(tail);
          while (tm != 0 && !esc) {
            const int t = __ffsll((unsigned long long)tm) - 1;
            tm &= tm - 1;
            const int32_t vt = __builtin_amdgcn_readlane(v, t);
            const int32_t a = -(__builtin_amdgcn_readlane(m, t) + 2);
            const int32_t b = gload(M.rowptr[dir] + vt + 1);
            for (int32_t p = a; p < b && !esc; p += 64) claim_one(p + lane < b ? gload(M.adj[dir] + p + lane) : kPad, seedpass);
          }
          if (dir == 0 && M.dup_all) {
            const uint32_t dv = (v >= 0 && (lane & (kEll - 1)) == 0) ? (uint32_t)gload(M.dup_all + v) : 0u;
            const uint32_t dsum = wave_sum_u32(dv);
            if (seedpass) deg0 += dsum;
            else if (lane == 0) w_new += dsum;
          }
        };
        narrow(lane < kEll ? src : -1, sm, true);
        if (__ballot(held_esc_lane) != 0) esc = true;  // X_src contains an escalated X_u
        bool seed = false;
        while (!esc && head < qn) {
          if (qn - head <= kLightVpp) {
            const int vi = head + (lane >> 4);
            const int32_t v = vi < qn ? Q[vi] : -1;
            const int32_t m = v >= 0 ? gload(M.ell[dir] + (int64_t)v * kEll + (lane & (kEll - 1))) : kPad;
            // a queued vertex is in R_src: if its own row escalated, X_src contains its X
            const bool vesc = v >= 0 && (lane & (kEll - 1)) == 0 && gload(M.esc[dir] + v) != 0;
            const int qend = qn;
            narrow(v, m, false);
            if (__ballot(vesc) != 0) esc = true;
            head = qend;
            continue;
          }
          int batch_end = 0;
          bool vesc_any = false;  // a vertex of this batch (in R_src) escalated: so does src
          {
            batch_end = min(head + 64, qn);
            const int32_t v = (head + lane < batch_end) ? Q[head + lane] : -1;
            load_row(M, dir, v, rc);
            vesc_any = v >= 0 && gload(M.esc[dir] + v) != 0;  // tested after the claims
          }
          uint32_t deg = diag_mark(rc);
          // kLightStep slots at a time: their home-slot CAS operations are in flight together;
          // a collision probes on. New entries below src are queued with one wave scan.
          auto claim = [&](int32_t m, int32_t old, uint32_t home) -> int {
            if (old == -1) return 1;
            if (old == m) return 0;
            return hash_probe<kBits>(H, m, (home + 1) & (kHash - 1));
          };
#pragma unroll
          for (int h = 0; h < kEll; h += kLightStep) {
            if (esc || __ballot(rc.e[h] >= 0) == 0) break;  // rows are packed: no later slot is used either
            int32_t old[kLightStep];
            uint32_t home[kLightStep];
#pragma unroll
            for (int j = 0; j < kLightStep; ++j) {
              const int32_t m = rc.e[h + j];
              home[j] = hash_home<kBits>(m);
              old[j] = m >= 0 ? atomicCAS(&H[home[j]], -1, m) : m;
            }
            uint32_t mr = 0, mu = 0, nnew = 0;
#pragma unroll
            for (int j = 0; j < kLightStep; ++j) {
              const int32_t m = rc.e[h + j];
              if (m >= 0) {
                ++deg;
                if (seed && m > src) ++u_seed;
                if (claim(m, old[j], home[j])) {
                  ++nnew;
                  if (m < src) mr |= 1u << j;
                  else if (m > src) mu |= 1u << j;
                }
              }
            }
            u_new += __popc(mu);
            const uint32_t c = __popc(mr);
            const uint32_t incl = wave_incl_scan_u32(c);
            int pos = qn + (int)(incl - c);
#pragma unroll
            for (int j = 0; j < kLightStep; ++j)
              if ((mr >> j) & 1u) {
                if (pos < kLightQ) Q[pos] = rc.e[h + j];
                ++pos;
              }
            qn += __builtin_amdgcn_readlane((int)incl, 63);
            if (track_u) {
              const uint32_t cu = __popc(mu);
              const uint32_t inclu = wave_incl_scan_u32(cu);
              int upos = uq + (int)(inclu - cu);
#pragma unroll
              for (int j = 0; j < kLightStep; ++j)
                if ((mu >> j) & 1u) {
                  if (upos < kLightQ) Q[kLightQ - 1 - upos] = rc.e[h + j];
                  ++upos;
                }
              uq += __builtin_amdgcn_readlane((int)inclu, 63);
            }
            hcount += (int)wave_sum_u32(nnew);
            if (hcount > limit || qn >= kLightQ) esc = true;
          }
          if (__ballot(vesc_any) != 0) esc = true;
          auto visit = [&](int32_t m) {
            const bool active = m >= 0;
            deg += active;
            if (seed && m > src) ++u_seed;
            const int ins = active ? hash_insert<kBits>(H, m) : 0;
            const bool is_r = ins == 1 && m < src;
            const bool is_u = ins == 1 && m > src;
            const uint64_t mr = __ballot(is_r);
            const uint64_t mn = __ballot(ins == 1);
            if (is_r) {
              const int pos = qn + (int)rank_below(mr);
              if (pos < kLightQ) Q[pos] = m;
            }
            qn += __popcll(mr);
            hcount += __popcll(mn);
            if (track_u) {
              const uint64_t mu = __ballot(is_u);
              if (is_u && uq + (int)rank_below(mu) < kLightQ) Q[kLightQ - 1 - (uq + (int)rank_below(mu))] = m;
              uq += __popcll(mu);
            }
            u_new += is_u;
            if (hcount > limit || qn >= kLightQ) esc = true;
          };
          while (!esc && __ballot(rc.ov_pos < rc.ov_end) != 0) {
            int32_t m = kPad;
            if (rc.ov_pos < rc.ov_end) m = gload(M.adj[dir] + rc.ov_pos++);
            visit(m);
          }
          const int32_t vexp = (head + lane < batch_end) ? Q[head + lane] : -1;
          if (vexp >= 0 && dir == 0 && M.dup_all) deg += gload(M.dup_all + vexp);
          w_new += deg;
          head = batch_end;
        }
        // per-row sums stay below the input nnz < 2^31
        const uint64_t w_new_all = wave_sum_u32((uint32_t)w_new);
        x_edges += w_new_all + deg0;
        x_verts += (uint64_t)min(qn, kLightQ) + 1;
        if (!esc) {
          const uint32_t U = (carry ? u_prev - (src_was_set ? 1u : 0u) : 0u) +
                             wave_sum_u32(u_new);
          const uint32_t Useed = wave_sum_u32(u_seed);
          const uint64_t wr = (carry ? wr_prev + deg_prev : 0) + w_new_all;
          const uint64_t rr = (carry ? r_prev + 1 : 0) + (uint64_t)qn;
          if (lane == 0) M.cnt[dir][src] = (int32_t)U;
          x_out += U;
          t_fill += U - Useed;
          if (dir == 0) {
            t_w += deg0 + wr;
            t_r += rr;
            t_nnz_u += U;
          } else {
            t_nnz_l += U;
          }
          if (dir == 1 || sym) {
            uint32_t win[4] = {0, 0, 0, 0};
            for (int i = lane; i < uq; i += 64) window_bit(win, src, Q[kLightQ - 1 - i], chunk_end);
            const uint32_t w0 = wave_or_u32(win[0]), w1 = wave_or_u32(win[1]);
            const uint32_t w2 = wave_or_u32(win[2]), w3 = wave_or_u32(win[3]);
            if (lane == 0) reinterpret_cast<uint4*>(M.win)[src] = make_uint4(w0, w1, w2, w3);
          }
          // structure: the U list at the back of Q
          if (M.list_ptr[dir] != nullptr) {
            const int64_t off = kBits == kLightBitsBig ? lp_off : gload(M.list_ptr[dir] + src);
            const int64_t end = kBits == kLightBitsBig ? lp_end : gload(M.list_ptr[dir] + src + 1);
            if (end - off != (int64_t)U || uq != (int)U || qn + uq > kLightQ) {
              if (lane == 0) atomicOr(&M.acc[kAccOverflow], 1ull);
            } else {
              for (int i = lane; i < uq; i += 64) gstore(M.pool[dir] + off + i, Q[kLightQ - 1 - i]);
            }
          }
          carry_ok = true;
          prev_esc = false;
          u_prev = U;
          wr_prev = wr;
          deg_prev = deg0;
          r_prev = rr;
        }
      }
      if (esc) {
        if (lane == 0) {
          gstore(M.esc[dir] + src, (uint8_t)1);
          gstore(M.blk_flag[dir] + src / M.chain_block, 1u);  // the block goes to the workgroup path
        }
        carry_ok = false;
        prev_esc = true;
      }
      // publish row src for the sibling wave: is src+1 in X_src?
      {
        const uint8_t st = esc ? 3 : (src + 1 < s1 && hash_has<kBits>(H, src + 1)) ? 1 : 2;
        self_st = st;
        if (!sym && lane == 0) {
          // the previous row's publication is visible before its ring slot's next use
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          s_nxt[wave][src & 7] = st;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __hip_atomic_store(&s_pub[wave], ((unsigned long long)k << 32) | (uint32_t)(src + 1), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
    if (hcount > 0) {
      for (int i = lane; i < kHash; i += 64) H[i] = -1;
      hcount = 0;
    }
  }
  flush(last_mat);
  if (lane == 0) __hip_atomic_store(&s_pub[wave], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);  // done
}

// ------------------------------------------------------------------------------------------
// The workgroup-path queue, in a fixed order: flagged blocks by descending block index (all
// matrices of a set share chain_block, and the highest rows reach the most vertices, so the
// costliest items start first and the tail is short), both directions and all matrices of
// one block index together. A counting sort over block indices: histogram, scan, scatter.
// Slot g < nblk_total is (direction 0, global block g); the rest are direction 1. A paired
// matrix gets one item per block (both directions), from its direction-0 slot.
__device__ __forceinline__ bool heavy_slot(const MatDev* mats, int64_t nblk_total, int nmats,
                                           const int64_t* blkq_base, const uint32_t* blk_flag, int64_t g,
                                           int& m, uint32_t& dir, int32_t& b) {
  const int64_t gb = g < nblk_total ? g : g - nblk_total;
  m = find_segment(blkq_base, nmats, gb);
  b = (int32_t)(gb - blkq_base[m]);
  dir = g < nblk_total ? 0u : 1u;
  if (mats[m].paired) return dir == 0 && (blk_flag[gb] | blk_flag[gb + nblk_total]) != 0;
  return blk_flag[g] != 0;
}

// The row where block b of matrix M is cut into two workgroup-path items, or -1.
__device__ __forceinline__ int32_t block_cut(const MatDev& M, int32_t b) {
  if (!M.split) return -1;
  const int32_t cut = M.split[b];
  const int64_t r0 = (int64_t)b * M.chain_block;
  return cut > r0 && cut < min((int64_t)M.n, r0 + M.chain_block) ? cut : -1;
}

// ------------------------------------------------------------------------------------------
__global__ void k_heavy_hist(const MatDev* __restrict__ mats, int64_t nblk_total, int nmats,
                             const int64_t* __restrict__ blkq_base, const uint32_t* __restrict__ blk_flag,
                             int32_t nb_max, int32_t* __restrict__ bucket, WorkState* __restrict__ ws) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < 2 * nblk_total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int m;
    uint32_t dir;
    int32_t b;
    if (!heavy_slot(mats, nblk_total, nmats, blkq_base, blk_flag, g, m, dir, b)) continue;
    // a block cut at light_end keeps only its part before the cut (the rest is an early item)
    const int32_t cut = block_cut(mats[m], b);
    const int parts = cut >= 0 && cut < mats[m].light_end ? 2 : 1;
    atomicAdd(&bucket[nb_max - 1 - b], parts);
    atomicAdd(&ws->heavy_count, (unsigned long long)parts);
  }
}

__global__ void k_heavy_scatter(const MatDev* __restrict__ mats, int64_t nblk_total, int nmats,
                                const int64_t* __restrict__ blkq_base, const uint32_t* __restrict__ blk_flag,
                                int32_t nb_max, int32_t* __restrict__ cursor, uint64_t* __restrict__ heavy_items) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < 2 * nblk_total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int m;
    uint32_t dir;
    int32_t b;
    if (!heavy_slot(mats, nblk_total, nmats, blkq_base, blk_flag, g, m, dir, b)) continue;
    const int32_t cut = block_cut(mats[m], b);
    if (cut >= 0 && cut < mats[m].light_end) {
      // a cut block: the part from the cut on (the larger subtree) goes first
      const int32_t pos = atomicAdd(&cursor[nb_max - 1 - b], 2);
      heavy_items[pos] = encode_item((uint32_t)m, dir, (uint32_t)cut, 0u);
      heavy_items[pos + 1] = encode_item((uint32_t)m, dir, (uint32_t)cut, 1u);
    } else if (cut >= 0) {
      // cut at light_end: the part from the cut on runs in the early launch
      const int32_t pos = atomicAdd(&cursor[nb_max - 1 - b], 1);
      heavy_items[pos] = encode_item((uint32_t)m, dir, (uint32_t)cut, 1u);
    } else {
      const int32_t pos = atomicAdd(&cursor[nb_max - 1 - b], 1);
      heavy_items[pos] = encode_item((uint32_t)m, dir, (uint32_t)(b * mats[m].chain_block), 0u);
    }
  }
}

// ------------------------------------------------------------------------------------------
// k_heavy: one workgroup per escalated source, n-bit bitmap in LDS, level-synchronous
// ------------------------------------------------------------------------------------------
// Dynamic LDS: bitmap[bw_max] | q0[qcap] | q1[qcap] | U list[ucap] per direction. Queue
// indices >= qcap and U-list indices >= ucap continue in the workgroup's global region:
// spill0 / spill1 (n_max entries each), ustage (n_max entries per direction).
__device__ __forceinline__ int32_t qget(const int32_t* l, const int32_t* g, int32_t qcap, int32_t e) {
  return e < qcap ? l[e] : gload(g + (e - qcap));
}
__device__ __forceinline__ void qput(int32_t* l, int32_t* g, int32_t qcap, int32_t e, int32_t v) {
  if (e < qcap) l[e] = v;
  else gstore(g + (e - qcap), v);
}

// Writes the entries of rc at the set bits of mask (one lane's new vertices) through put at
// pos, pos + 1, ... A slot no lane of the wave sets (a zero bit of any, the OR of mask over the
// wave) costs a scalar branch instead of a predicated write. Wave-uniform call.
template <typename Put>
__device__ __forceinline__ void put_row_entries(const RowCursor& rc, uint32_t mask, uint32_t any, int32_t pos, Put put) {
#pragma unroll
  for (int j = 0; j < kEll; ++j)
    if (((any >> j) & 1u) && ((mask >> j) & 1u)) put(pos++, rc.e[j]);
}

// A row traversed by the bitmap path: W_R, input degree, |R|, |U| and its staged U-list length.
struct RowFacts {
  uint64_t wr, deg, r;
  uint32_t u, ul;
};

struct Totals {  // cleared with Totals{}
  uint64_t fill, w, r, nnz_u, nnz_l, heavy;
  uint64_t x_edges, x_verts, x_out;
};
// Every k_heavy total is also the workgroup path's own share: W, R and F entries found there.

__device__ __forceinline__ void flush_totals(const MatDev* mats, int mb, Totals& t, int32_t exec_slot) {
  if (mb < 0) return;
  unsigned long long* acc = mats[mb].acc;
  if (t.fill) atomicAdd(&acc[kAccFill], t.fill);
  if (t.w) atomicAdd(&acc[kAccW], t.w);
  if (t.r) atomicAdd(&acc[kAccR], t.r);
  if (t.nnz_u) atomicAdd(&acc[kAccNnzU], t.nnz_u);
  if (t.nnz_l) atomicAdd(&acc[kAccNnzL], t.nnz_l);
  if (t.heavy) atomicAdd(&acc[kAccHeavy], t.heavy);
  if (t.w) atomicAdd(&acc[kAccHeavyW], t.w);
  if (t.r) atomicAdd(&acc[kAccHeavyR], t.r);
  if (t.nnz_u + t.nnz_l) atomicAdd(&acc[kAccHeavyNnz], t.nnz_u + t.nnz_l);
  if (t.x_edges) atomicAdd(&acc[exec_slot], t.x_edges);
  if (t.x_verts) atomicAdd(&acc[exec_slot + 1], t.x_verts);
  if (t.x_out) atomicAdd(&acc[exec_slot + 2], t.x_out);
  t = Totals{};
}

// One workgroup per item = a block of chain_block consecutive rows; it traverses the block's
// escalated rows in increasing order and reuses the previous row's reached set
// X_{s-1} = {s-1} + R_{s-1} + U_{s-1} whenever s-1 is in R_s, because then R_{s-1}, s-1 and
// all their neighbours belong to X_s:
//   * chaining: row s holds s-1 (or s-1 is known to be in R_s, below). The bitmap starts as
//     X_{s-1} (carried in LDS, or copied back from the saved copy) and only vertices new to s
//     are expanded.
//   * import: otherwise, if X_{s-1} was saved, s is traversed from scratch and, the moment s-1
//     is reached, X_{s-1} is OR-ed in at the next level boundary; queued vertices already in
//     X_{s-1} are then dropped (their neighbours are in).
// W_R(s) = sum of deg over R_s and |R_s| are carried exactly: the previous row's values, plus
// s-1, plus the vertices expanded here that are not in X_{s-1}.
//
// Paired items (M.paired) run both directions of the block, row by row (s dir 0, s dir 1,
// s+1 dir 0, ...), and save every row's bitmap. Then c in R_s is often known before any
// expansion, for the last rows c saved in this direction: if an entry u < c of row s lies in the
// other direction's R'_c, u reaches c through vertices < c, so s -> u -> ... -> c stays below s.
// That turns the expensive late imports of unstructured patterns (where row s rarely holds
// s-1) into chaining. Plain items run one direction and save a bitmap only for an import.
//
// kHbmBitmap: the bitmap lives in a per-workgroup HBM region instead of LDS (matrices whose
// n-bit bitmap does not fit in LDS). Its claims are then L2 atomics, its reads bypass L1, and a
// wave's plain stores to it (clears, copies) are acknowledged by L2 before each barrier.
//
// Work split inside the workgroup. Wave 0, the leader, runs the rows: it plans each row (chain
// base, early tests, import), runs the seed pass and every BFS level of at most kSoloMax
// vertices by itself, and finishes the row (counts, window bits, the U list into the
// structure). The other waves wait at one barrier and join only when the leader asks: to clear,
// copy in, OR in or save a bitmap, and for levels wider than kSoloMax vertices. Most escalated
// rows of structured patterns chain from their predecessor and reach nothing new below
// themselves (their traversal is the seed row alone): they cost the leader a few hundred
// cycles and the workgroup no barrier. A request is the command word in LDS, written by the
// leader before a barrier and read by every wave after it, so every wave takes the same
// barriers. The per-row facts of a window of up to kHeavyWin rows (escalated? holds its
// predecessor? structure offsets, and the rows themselves) are read in one round into LDS.
constexpr int kImportBatch = 8;
constexpr int kSoloNarrow = 4;  // levels of at most this many vertices: 16 lanes per vertex
constexpr int kSoloMax = 64;    // levels of at most this many vertices: the leader alone

enum : int32_t { kCmdDone = 0, kCmdClear = 1, kCmdCopyIn = 2, kCmdImport = 3, kCmdSave = 4, kCmdWide = 5 };

// The leader's request: written by its lane 0 before a barrier, read by every thread after it.
struct HeavyCtl {
  int32_t cmd, d, src, level;
  int32_t imported, slot;  // slot: the saved bitmap to copy in, OR in or write
  uint32_t ubase;          // carried U-list entries (the row's new entries follow them)
  int32_t has_prev;
};

// The leader's state between its turns, in LDS: a turn starts from it and writes it back before
// the barrier that ends the turn, so no register of the leader is live while the workgroup works.
struct LeaderState {
  int32_t prev[2];      // per direction: the last row traversed in this block
  int32_t sv[2][kSavedRing];  // per direction: the row each saved bitmap holds
  int32_t wp[2];        // per direction: the ring slot of the next save
  int32_t lds_row, lds_dir;   // what the LDS bitmap holds
  int32_t li_next, di_next, phase;  // cursor in the window; 0 start a row, 1 bitmap ready, 2 levels, 3 finish
  int32_t d, li, fi, slot1, tslot, level, src, t, chunk_end;  // the row in progress
  int32_t chained, has_prev, imported;
  uint32_t ubase, gap;
};

// kPairedAny: some matrix of the set runs paired items (otherwise that code is compiled out).
template <int kThreads, bool kPairedAny, bool kHbmBitmap>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_heavy(const MatDev* __restrict__ mats,
                                                    WorkState* __restrict__ ws,
                                                    const uint64_t* __restrict__ heavy_items,
                                                    int32_t bw_max, int32_t qcap, int64_t n_max,
                                                    int32_t* __restrict__ spill_all,
                                                    uint32_t* __restrict__ hbm_bitmaps, int32_t ucap,
                                                    int32_t exec_slot) {
  static_assert(kEll == 16, "the window scan maps one row to 16 lanes");
  extern __shared__ uint32_t dyn[];
  constexpr int kPfRounds = ((kHeavyWin + 1) * kEll + kThreads - 1) / kThreads;
  constexpr int kStash = (kHeavyWin + 1) * kEll;  // ELL slots of a window's rows, per direction
  __shared__ unsigned long long s_idx;
  __shared__ int32_t s_cnt[3];     // queue lengths: level L at index L % 3
  __shared__ uint32_t s_ust;       // U-list entries appended after the carried ones
  __shared__ uint32_t s_unew;      // new U entries found by cooperative phases
  __shared__ unsigned long long s_wall, s_wovl, s_xe;  // row sums of earlier turns and cooperative phases
  __shared__ uint32_t s_rall, s_rovl, s_useed, s_deg0;
  __shared__ int32_t s_import, s_src_was_set;
  __shared__ int32_t s_esc[2][kHeavyWin + 1];   // row w0 + i was escalated (per direction)
  __shared__ int32_t s_hold[2][kHeavyWin + 1];  // row w0 + i holds w0 + i - 1
  __shared__ int64_t s_lp[2][kHeavyWin + 2];    // structure offsets of rows w0 .. w0 + np
  __shared__ unsigned long long s_escm[2];      // rows w0 .. w0 + 63 escalated, as bits (per direction)
  __shared__ HeavyCtl s_ctl;
  __shared__ LeaderState s_ls;
  // The slots some lane of the wave sets (push loops skip the others). Sets with paired items
  // (unstructured rows, which fill most slots) keep every slot: measured faster there.
  auto row_any = [](uint32_t mask) -> uint32_t {
    if constexpr (kPairedAny) return ~0u;
    else return wave_or_u32(mask);
  };

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid >> 6;
  const bool leader = wave == 0;
  // the n-bit bitmap: LDS, or this workgroup's bw_max words of hbm_bitmaps
  uint32_t* const bm = kHbmBitmap ? hbm_bitmaps + (int64_t)blockIdx.x * bw_max : dyn;
  int32_t* const ql0 = reinterpret_cast<int32_t*>(kHbmBitmap ? dyn : dyn + bw_max);
  // the staged U list of each direction: its first ucap entries in LDS, the rest in ustage
  int32_t* const ul0 = ql0 + 2 * qcap;
  int32_t* const ul1 = kPairedAny ? ul0 + ucap : ul0;
  // the window's rows in ELL form, per direction (the seed rows of the leader)
  int32_t* const stash0 = ul0 + (kPairedAny ? 2 : 1) * ucap;
  int32_t* const stash1 = kPairedAny ? stash0 + kStash : stash0;
  // HBM bitmap: its claims are L2 atomics, so its reads bypass L1 (agent-scope loads, never a
  // stale line), and each wave's plain stores (clears, copies) are acknowledged by L2 before a
  // barrier lets another wave's atomics at the same words.
  auto bm_get = [&](int32_t w) -> uint32_t {
    if constexpr (kHbmBitmap) return __hip_atomic_load(bm + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return bm[w];
  };
  auto bsync = [&]() {
    if constexpr (kHbmBitmap) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  };
  int32_t* const ql1 = ql0 + qcap;
  // Global region of this workgroup: two frontier spill queues, the U list of the last row of
  // each direction, saved bitmaps: one, or with paired items a ring of the last kSavedRing rows
  // saved per direction (chain bases, and the other direction's early tests).
  constexpr int kRing = kPairedAny ? kSavedRing : 1;
  const int64_t nbw = (n_max + 31) / 32;
  const int64_t region_len = heavy_region_ints(n_max, kPairedAny);
  int32_t* const region = spill_all + (int64_t)blockIdx.x * region_len;
  int32_t* const qg0 = region;
  int32_t* const qg1 = region + n_max;
  int32_t* const ustage0 = region + 2 * n_max;
  int32_t* const ustage1 = kPairedAny ? region + 3 * n_max : ustage0;
  uint32_t* const gsave = reinterpret_cast<uint32_t*>(region + (kPairedAny ? 4 : 3) * n_max);
  // saved bitmap k of direction dd
  auto gslot = [&](int dd, int k) -> uint32_t* {
    return kPairedAny ? gsave + (int64_t)(kSavedRing * dd + k) * nbw : gsave;
  };
  // Per direction: the rows of the saved bitmaps (index k < kSavedRing) and the last row
  // traversed (index kSavedRing). Written by the leader when it finishes a row.
  __shared__ RowFacts s_facts[2][kSavedRing + 1];

  int last_mat = -1;
  __shared__ Totals tot;  // thread 0 only (in LDS: keeps the registers for the traversal)
  if (tid == 0) tot = Totals{};

  const unsigned long long total = ws->heavy_count;
  while (true) {
    if (tid == 0) s_idx = atomicAdd(&ws->heavy_next, 1ull);
    bsync();
    const unsigned long long idx = s_idx;
    if (idx >= total) break;
    const uint64_t item = heavy_items[idx];
    const int mb = (int)((item >> 33) & 0xfffffu);
    const int item_dir = (int)((item >> 32) & 1);
    const int32_t v = (int32_t)(item & 0xffffffffu);
    const int before_cut = (int)(item >> 63);
    const MatDev& M = mats[mb];
    const bool paired = kPairedAny && M.paired != 0;
    const int ndir = paired ? 2 : 1;
    const int32_t n = M.n;
    const int32_t bw = (n + 31) >> 5;
    // rows [r0, r1): from v to the end of its block, or, before a cut, from the block's start to
    // v (decoded from the item alone: a table read here costs the traversal registers)
    const int32_t vb = v - v % M.chain_block;
    const int32_t r0 = before_cut ? vb : v;
    const int32_t r1 = before_cut ? v : (int32_t)min((int64_t)n, vb + (int64_t)M.chain_block * (1 + (int32_t)((item >> 53) & 0x3ffu)));
    const bool want_list = M.list_ptr[0] != nullptr;  // both directions or neither

    // The leader's state of this item, per direction (index = direction; a plain item uses its
    // own direction's slot): the last row traversed in this block, which rows the saved bitmaps
    // hold, the ring slot of the next save, and what the LDS bitmap holds. Set before the first
    // window's barrier.
    if (tid == 0) {
      for (int dd = 0; dd < 2; ++dd) {
        s_ls.prev[dd] = -1;
        s_ls.wp[dd] = 0;
        for (int k = 0; k < kSavedRing; ++k) s_ls.sv[dd][k] = -1;
      }
      s_ls.lds_row = s_ls.lds_dir = -1;
    }
#define GSOFA_SV(dd, k) (s_ls.sv[dd][k])
#define GSOFA_SLOT(dd, r)                                                             \
  ((r) < 0 ? -1                                                                       \
   : s_ls.sv[dd][0] == (r) ? 0 : (kRing > 1 && s_ls.sv[dd][1] == (r)) ? 1 : (kRing > 2 && s_ls.sv[dd][2] == (r)) ? 2 : -1)
    for (int32_t w0 = r0; w0 < r1; w0 += kHeavyWin) {
      const int32_t we = min(r1, w0 + kHeavyWin);
      const int32_t np = min(r1, we + 1) - w0;  // rows read: the window plus one look-ahead row
      if (tid == 0) {
        s_ls.li_next = 0;
        s_ls.di_next = 0;
        s_ls.phase = 0;
      }
      // ---- window facts, one round of loads per direction ----
      for (int di = 0; di < ndir; ++di) {
        const int dd = paired ? di : item_dir;
        int32_t* const stash = (kPairedAny && dd) ? stash1 : stash0;
        int32_t esc_v = 0;
        int64_t lp_v = 0;
        if (tid < np) esc_v = gload(M.esc[dd] + w0 + tid);
        if (want_list && tid <= np) lp_v = gload(M.list_ptr[dd] + w0 + tid);
        int32_t ev[kPfRounds];
#pragma unroll
        for (int q = 0; q < kPfRounds; ++q) {
          const int32_t i = q * kThreads + tid;
          ev[q] = i < np * kEll ? gload(M.ell[dd] + (int64_t)w0 * kEll + i) : kPad;
        }
#pragma unroll
        for (int q = 0; q < kPfRounds; ++q) {
          const int32_t i = q * kThreads + tid;
          const int32_t row = w0 + (i >> 4);
          bool hit = row > 0 && ev[q] == row - 1;
          if ((i & (kEll - 1)) == kEll - 1 && ev[q] <= -2 && i < np * kEll) {  // long row: its CSR tail
            for (int32_t p = -(ev[q] + 2), pe = gload(M.rowptr[dd] + row + 1); p < pe; ++p)
              hit |= gload(M.adj[dd] + p) == row - 1;
          }
          const uint64_t b = __ballot(hit);
          if ((i & (kEll - 1)) == 0 && i < np * kEll) s_hold[dd][i >> 4] = ((b >> (lane & 48)) & 0xFFFFull) != 0;
          if (i < np * kEll) stash[i] = ev[q];
        }
        if (tid < np) s_esc[dd][tid] = esc_v;
        if (want_list && tid <= np) s_lp[dd][tid] = lp_v;
        // the window's escalated rows as bits (rows w0 .. we - 1, at most 64), for the leader
        const uint64_t em = __ballot(tid < we - w0 && esc_v != 0);
        if (tid == 0) s_escm[dd] = em;
      }
      bsync();
      // Rows that do not hold their predecessor directly but through an entry u < row - 1 whose
      // own row holds row - 1 (row -> u -> row - 1 stays below row): they chain too. Tested for
      // escalated rows whose predecessor is escalated, i.e. traversed here just before them.
      for (int di = 0; di < ndir; ++di) {
        const int dd = paired ? di : item_dir;
        const int32_t* const stash = (kPairedAny && dd) ? stash1 : stash0;
#pragma unroll 1
        for (int q = 0; q < kPfRounds; ++q) {
          const int32_t i = q * kThreads + tid;
          const int32_t li2 = i >> 4;
          const int32_t row = w0 + li2;
          bool hit = false;
          if (i < np * kEll && li2 >= 1 && s_esc[dd][li2] != 0 && s_esc[dd][li2 - 1] != 0 && s_hold[dd][li2] == 0) {
            const int32_t u = stash[i];
            if (u >= 0 && u < row - 1) {
              const int32_t* const ur = M.ell[dd] + (int64_t)u * kEll;
#pragma unroll
              for (int q4 = 0; q4 < 4; ++q4) {
                const v4i32 t4 = gload_v4(ur + 4 * q4);
                hit = hit || t4.x == row - 1 || t4.y == row - 1 || t4.z == row - 1 || t4.w == row - 1;
              }
            }
          }
          const uint64_t b = __ballot(hit);
          if ((i & (kEll - 1)) == 0 && i < np * kEll && ((b >> (lane & 48)) & 0xFFFFull) != 0) s_hold[dd][li2] = 1;
        }
      }
      bsync();

      // ---- the window's rows: the leader runs them, the workgroup helps on request ----
      while (true) {
        if (leader) {
          // a turn of the leader: its state from LDS (uniform within the wave)
          int li_next = s_ls.li_next, di_next = s_ls.di_next;
          int phase = s_ls.phase;
          int d = s_ls.d, li = s_ls.li, fi = s_ls.fi, slot1 = s_ls.slot1, tslot = s_ls.tslot, level = s_ls.level;
          int32_t src = s_ls.src, t = s_ls.t, chunk_end = s_ls.chunk_end;
          bool chained = s_ls.chained != 0, has_prev = s_ls.has_prev != 0, imported = s_ls.imported != 0;
          uint32_t ubase = s_ls.ubase, gap = s_ls.gap;
          // per-lane row sums of this turn's passes (earlier turns' are in the LDS sums)
          uint32_t w_all = 0, w_ovl = 0, x_e = 0, r_all = 0, r_ovl = 0, u_seed = 0, u_new = 0, deg0 = 0;
          int32_t cmd = -1;
          int32_t cslot = 0;
          while (cmd < 0) {
            // the staged U list of the row's direction (d changes when a row starts)
            auto uget = [&](uint32_t i) -> int32_t {
              return (int32_t)i < ucap ? (d ? ul1 : ul0)[i] : gload((d ? ustage1 : ustage0) + ((int32_t)i - ucap));
            };
            auto uput = [&](uint32_t i, int32_t v) {
              if ((int32_t)i < ucap) (d ? ul1 : ul0)[i] = v;
              else gstore((d ? ustage1 : ustage0) + ((int32_t)i - ucap), v);
            };
            if (phase == 0) {
              // ---- the next escalated (row, direction) of the window, from the window's bits ----
              const uint64_t em0 = s_escm[0], em1 = s_escm[1];
              const uint64_t rows_m = paired ? (em0 | em1) : (item_dir ? em1 : em0);
              const int nw = we - w0;
              bool found = false;
              while (!found && li_next < nw) {
                const int dd = paired ? di_next : item_dir;
                if ((((dd ? em1 : em0) >> li_next) & 1ull) != 0) {
                  found = true;
                  d = dd;
                  li = li_next;
                }
                if (++di_next >= ndir) {
                  di_next = 0;
                  ++li_next;
                  // skip the rows with no escalated direction
                  const uint64_t rest = li_next < nw ? rows_m >> li_next : 0ull;
                  li_next = rest ? li_next + __ffsll((unsigned long long)rest) - 1 : nw;
                }
              }
              if (!found) {
                cmd = kCmdDone;
                break;
              }
              src = w0 + li;
              chunk_end = min(n, (src / M.chunk + 1) * M.chunk);
              const int32_t prev_d = s_ls.prev[d];
              // t1: the last row traversed here in this direction, even with wave-path rows in
              // between (plain items: s-1 only)
              const int32_t t1 = (prev_d >= 0 && (paired || prev_d == src - 1)) ? prev_d : -1;
              slot1 = GSOFA_SLOT(d, t1);
              const bool lds_t1 = t1 >= 0 && s_ls.lds_row == t1 && s_ls.lds_dir == d;  // LDS still holds X_t1
              const bool hold = t1 == src - 1 && s_hold[d][li] != 0;           // row s holds t1 = s-1
              // The chain base t (some t in R_s: X_s contains X_t) and the index of its facts.
              t = t1;
              fi = kSavedRing;
              chained = hold && (lds_t1 || slot1 >= 0);
              if (!chained && M.sym && t1 == src - 1 && (lds_t1 || slot1 >= 0)) {
                // Mirrored (A^T = A): s-1 is in R_s exactly when s is in U_{s-1}, a bit of X_{s-1}
                const uint32_t xw = lds_t1 ? bm_get(src >> 5) : gload(gslot(d, slot1) + (src >> 5));
                chained = ((xw >> (src & 31)) & 1u) != 0;
              }
              if (kPairedAny && paired && !chained) {
                // ---- paired: early tests of the saved rows c of this direction whose other
                // direction is saved too. An entry u of row s with u == c, or u < c in R'_c (u
                // reaches c through vertices < c), puts c in R_s. With a staged U list only t1
                // qualifies (the list is t1's). ----
                const int o = 1 - d;
                int32_t c0 = GSOFA_SV(d, 0), c1 = GSOFA_SV(d, 1), c2 = GSOFA_SV(d, 2);
                if (want_list) {
                  c0 = c0 == t1 ? c0 : -1;
                  c1 = c1 == t1 ? c1 : -1;
                  c2 = c2 == t1 ? c2 : -1;
                }
                const int o0 = GSOFA_SLOT(o, c0), o1 = GSOFA_SLOT(o, c1), o2 = GSOFA_SLOT(o, c2);
                if (o0 >= 0 || o1 >= 0 || o2 >= 0) {
                  const uint32_t* const g0 = gslot(o, o0 < 0 ? 0 : o0);
                  const uint32_t* const g1 = gslot(o, o1 < 0 ? 0 : o1);
                  const uint32_t* const g2 = gslot(o, o2 < 0 ? 0 : o2);
                  auto in_other = [&](int32_t u, int32_t c, int oc, const uint32_t* g) -> bool {
                    return oc >= 0 && (u == c || (u >= 0 && u < c && ((gload(g + (u >> 5)) >> (u & 31)) & 1u) != 0));
                  };
                  const int32_t* const stash = d ? stash1 : stash0;
                  const int32_t e = lane < kEll ? stash[li * kEll + lane] : kPad;
                  bool h0 = in_other(e, c0, o0, g0), h1 = in_other(e, c1, o1, g1), h2 = in_other(e, c2, o2, g2);
                  if (lane == kEll - 1 && e <= -2) {
                    h0 = h1 = h2 = false;
                    for (int32_t p = -(e + 2), pe = gload(M.rowptr[d] + src + 1); p < pe && !(h0 && h1 && h2); ++p) {
                      const int32_t u = gload(M.adj[d] + p);
                      h0 = h0 || in_other(u, c0, o0, g0);
                      h1 = h1 || in_other(u, c1, o1, g1);
                      h2 = h2 || in_other(u, c2, o2, g2);
                    }
                  }
                  const int hits = (__ballot(h0) != 0 ? 1 : 0) | (__ballot(h1) != 0 ? 2 : 0) | (__ballot(h2) != 0 ? 4 : 0);
                  // of the rows found in R_s, the one with the largest R (fewest vertices left)
                  int best = -1;
                  uint64_t best_r = 0;
                  for (int k = 0; k < kRing; ++k)
                    if (((hits >> k) & 1) && (best < 0 || s_facts[d][k].r > best_r)) {
                      best = k;
                      best_r = s_facts[d][k].r;
                    }
                  if (best >= 0) {
                    chained = true;
                    t = GSOFA_SV(d, best);
                    fi = best;
                  }
                }
              }
              const bool chain_lds = chained && t == t1 && lds_t1;
              // the saved bitmap of the chain base, or of t1 for a late import
              tslot = chained ? GSOFA_SLOT(d, t) : slot1;
              // late import (of X_{s-1} only, when s-1 is reached)
              has_prev = !chained && t1 == src - 1 && slot1 >= 0;
              if (lane == 0) {
                s_cnt[0] = 1;
                s_cnt[1] = 0;
                s_cnt[2] = 0;
                s_ust = 0;
                s_unew = 0;
                s_wall = s_wovl = s_xe = 0;
                s_rall = s_rovl = s_useed = s_deg0 = 0;
                s_import = 0;
              }
              w_all = w_ovl = x_e = r_all = r_ovl = u_seed = u_new = deg0 = 0;
              gap = 0;
              phase = 1;
              if (!chained) {
                cmd = kCmdClear;  // a fresh bitmap holding src
                break;
              }
              if (!chain_lds) {
                cmd = kCmdCopyIn;  // X_t back from its saved copy, with the bit of src set
                cslot = tslot;
                break;
              }
              if (lane == 0) {
                const uint32_t sb = 1u << (src & 31);
                uint32_t old;
                if constexpr (kHbmBitmap) {
                  old = atomicOr(&bm[src >> 5], sb);
                } else {
                  old = bm[src >> 5];
                  bm[src >> 5] = old | sb;
                }
                s_src_was_set = (old & sb) != 0;
              }
            }
            wave_sync();  // row resets, the src bit
            if (phase == 1) {
              // ---- the bitmap holds X_t (or only src): queues, U list base, the gap ----
              // The carried U list keeps stale entries up to src; they are dropped when the list
              // is written out at the end of the row.
              ubase = (chained && want_list) ? s_facts[d][fi].ul : 0u;
              if (lane == 0) ql0[0] = src;
              if (chained && t < src - 1) {
                // X_t's vertices strictly between t and src are in R_s (reached through t's set)
                // but their own neighbours are not covered: they start the level after the seed.
                const int32_t wlo = (t + 1) >> 5, whi = (src - 1) >> 5;
                int32_t pos = 0;
                for (int32_t wb = wlo; wb <= whi; wb += 64) {
                  const int32_t w = wb + lane;
                  uint32_t x = 0;
                  if (w <= whi) {
                    x = bm_get(w);
                    const int32_t b0 = w * 32;
                    if (b0 <= t) x &= ~0u << (t - b0) << 1;                 // above t
                    if (b0 + 31 >= src) x &= (1u << (src - b0)) - 1u;      // below src (src > b0 here)
                  }
                  const uint32_t c = __popc(x);
                  const uint32_t incl = wave_incl_scan_u32(c);
                  int32_t p = pos + (int32_t)(incl - c);
                  while (x) {
                    const int b = __ffs(x) - 1;
                    x &= x - 1;
                    qput(ql1, qg1, qcap, p++, w * 32 + b);
                  }
                  pos += __builtin_amdgcn_readlane((int)incl, 63);
                }
                gap = (uint32_t)pos;
                if (lane == 0) s_cnt[1] = pos;
              }
              level = 0;
              imported = false;
              phase = 2;
              wave_sync();  // the seed, the gap queue and its count
            }
            if (phase == 2) {
              // ---- BFS levels: the leader alone while a level has at most kSoloMax vertices ----
              const uint32_t* const gprev = gslot(d, slot1 < 0 ? 0 : slot1);  // saved X_{s-1}
              const int32_t* const stash = (kPairedAny && d) ? stash1 : stash0;
              while (true) {
                const int32_t ncur = s_cnt[level % 3];
                if (ncur == 0) {
                  phase = 3;
                  break;
                }
                if (ncur > kSoloMax) {
                  cmd = kCmdWide;  // the workgroup takes the wide levels
                  cslot = slot1;
                  break;
                }
                const bool odd = (level & 1) != 0;
                const int32_t* const cur_l = odd ? ql1 : ql0;
                const int32_t* const cur_g = odd ? qg1 : qg0;
                int32_t* const nxt_l = odd ? ql0 : ql1;
                int32_t* const nxt_g = odd ? qg0 : qg1;
                const bool seed = level == 0;
                int32_t nxt_n = s_cnt[(level + 1) % 3];
                uint32_t ust = s_ust;
                bool imp = false;  // this lane reached src-1 with an import pending
                // One claim per lane: m < 0 claims nothing. New entries below src go to the next
                // level, new entries above src are U. Wave-uniform call.
                auto claim = [&](int32_t m, bool fresh_check) {
                  const bool valid = m >= 0;
                  bool isnew = false;
                  if (valid) {
                    const uint32_t bit = 1u << (m & 31);
                    if (!fresh_check || (bm_get(m >> 5) & bit) == 0) isnew = (atomicOr(&bm[m >> 5], bit) & bit) == 0;
                  }
                  if (isnew && has_prev && m == src - 1) {  // import X_{src-1} instead of expanding it
                    imp = true;
                    isnew = false;
                  }
                  const bool is_r = isnew && m < src;
                  const bool is_u = isnew && m > src;
                  const uint64_t mr = __ballot(is_r);
                  if (is_r) qput(nxt_l, nxt_g, qcap, nxt_n + (int32_t)rank_below(mr), m);
                  nxt_n += __popcll(mr);
                  u_new += is_u;
                  if (want_list) {
                    const uint64_t mu = __ballot(is_u);
                    if (is_u) uput(ubase + ust + rank_below(mu), m);
                    ust += (uint32_t)__popcll(mu);
                  }
                };
                if (ncur <= kSoloNarrow) {
                  // narrow pass: lane = (vertex, ELL slot)
                  const int vi = lane >> 4, sl = lane & (kEll - 1);
                  int32_t v = -1;
                  if (vi < ncur) v = seed ? src : qget(cur_l, cur_g, qcap, vi);
                  int32_t m = kPad;
                  if (v >= 0) m = seed ? stash[li * kEll + sl] : gload(M.ell[d] + (int64_t)v * kEll + sl);
                  const bool chk = v >= 0 && !seed && has_prev;
                  const uint32_t pword = chk ? gload(gprev + (v >> 5)) : 0u;
                  uint32_t dupv = (v >= 0 && sl == 0 && d == 0 && M.dup_all) ? (uint32_t)gload(M.dup_all + v) : 0u;
                  const bool in_prev = chk && ((pword >> (v & 31)) & 1u) != 0;
                  if (in_prev && imported) {  // covered by the imported X_{s-1}
                    v = -1;
                    m = kPad;
                    dupv = 0;
                  }
                  const bool tail = sl == kEll - 1 && m <= -2;
                  int32_t ov_pos = 0, ov_end = 0;
                  if (tail) {
                    ov_pos = -(m + 2);
                    ov_end = gload(M.rowptr[d] + v + 1);
                    m = kPad;
                  }
                  const uint32_t valid = (m >= 0 || m == kDiag) ? 1u : 0u;  // an entry, or the diagonal's mark
                  if (m == kDiag) m = kPad;
                  if (seed) {
                    deg0 += valid + dupv;
                    u_seed += (m > src) ? 1u : 0u;  // entries of A's row above the diagonal
                  } else {
                    w_all += valid + dupv;
                    if (in_prev) w_ovl += valid + dupv;
                    if (v >= 0 && sl == 0) {
                      r_all += 1;
                      if (in_prev) r_ovl += 1;
                    }
                  }
                  x_e += valid;
                  claim(m, false);
                  // CSR tails of long rows: the wave walks each one 64 entries at a time
                  uint64_t tm = __ballot(tail);
                  while (tm != 0) {
                    const int tl = __ffsll((unsigned long long)tm) - 1;
                    tm &= tm - 1;
                    const int32_t a = __builtin_amdgcn_readlane(ov_pos, tl);
                    const int32_t b = __builtin_amdgcn_readlane(ov_end, tl);
                    const bool tin = __builtin_amdgcn_readlane((int)in_prev, tl) != 0;
                    for (int32_t p = a; p < b; p += 64) {
                      const int32_t mm = p + lane < b ? gload(M.adj[d] + p + lane) : kPad;
                      const uint32_t vv = mm >= 0 ? 1u : 0u;
                      if (seed) {
                        deg0 += vv;
                        u_seed += (mm > src) ? 1u : 0u;
                      } else {
                        w_all += vv;
                        if (tin) w_ovl += vv;
                      }
                      x_e += vv;
                      claim(mm, true);
                    }
                  }
                } else {
                  // wide pass of the leader: one vertex per lane, its ELL slots 8 at a time
                  int32_t v = lane < ncur ? qget(cur_l, cur_g, qcap, lane) : -1;
                  const bool chk = v >= 0 && has_prev;
                  const uint32_t pword = chk ? gload(gprev + (v >> 5)) : 0u;
                  RowCursor rc;
                  load_row(M, d, v, rc);
                  uint32_t dupv = (v >= 0 && d == 0 && M.dup_all) ? (uint32_t)gload(M.dup_all + v) : 0u;
                  const bool in_prev = chk && ((pword >> (v & 31)) & 1u) != 0;
                  if (in_prev && imported) {
                    v = -1;
                    dupv = 0;
#pragma unroll
                    for (int k = 0; k < kEll; ++k) rc.e[k] = kPad;
                    rc.ov_pos = rc.ov_end = 0;
                  }
                  uint32_t deg = diag_mark(rc);
                  uint32_t mask_r = 0, mask_u = 0;
#pragma unroll
                  for (int h = 0; h < kEll; h += 8) {
                    if (h > 0 && __ballot(rc.e[h] >= 0) == 0) break;
                    uint32_t old[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                      const int32_t m = rc.e[h + j];
                      old[j] = m >= 0 ? atomicOr(&bm[m >> 5], 1u << (m & 31)) : ~0u;
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                      const int32_t m = rc.e[h + j];
                      if (m >= 0) {
                        ++deg;
                        if (((old[j] >> (m & 31)) & 1u) == 0) {
                          if (has_prev && m == src - 1) imp = true;
                          else if (m < src) mask_r |= 1u << (h + j);
                          else if (m > src) mask_u |= 1u << (h + j);
                        }
                      }
                    }
                  }
                  {
                    const uint32_t c = __popc(mask_r);
                    const uint32_t incl = wave_incl_scan_u32(c);
                    const int32_t pos = nxt_n + (int32_t)(incl - c);
                    const int32_t wend = nxt_n + __builtin_amdgcn_readlane((int)incl, 63);
                    const uint32_t any = row_any(mask_r);
                    if (!kPairedAny && wend <= qcap)  // the whole wave's range is in LDS
                      put_row_entries(rc, mask_r, any, pos, [&](int32_t i, int32_t x) { nxt_l[i] = x; });
                    else
                      put_row_entries(rc, mask_r, any, pos, [&](int32_t i, int32_t x) { qput(nxt_l, nxt_g, qcap, i, x); });
                    nxt_n = wend;
                  }
                  u_new += __popc(mask_u);
                  if (want_list && __ballot(mask_u != 0) != 0) {  // most levels find no U entry
                    const uint32_t c = __popc(mask_u);
                    const uint32_t incl = wave_incl_scan_u32(c);
                    const uint32_t pos = ubase + ust + (incl - c);
                    put_row_entries(rc, mask_u, row_any(mask_u), (int32_t)pos,
                                    [&](int32_t i, int32_t x) { uput((uint32_t)i, x); });
                    ust += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                  }
                  // CSR tails of long rows: one entry per lane per step
                  while (__ballot(rc.ov_pos < rc.ov_end) != 0) {
                    int32_t m = kPad;
                    if (rc.ov_pos < rc.ov_end) m = gload(M.adj[d] + rc.ov_pos++);
                    deg += m >= 0;
                    claim(m, true);
                  }
                  x_e += deg;
                  if (v >= 0) {
                    deg += dupv;
                    w_all += deg;
                    r_all += 1;
                    if (in_prev) {
                      w_ovl += deg;
                      r_ovl += 1;
                    }
                  }
                }
                if (lane == 0) {
                  s_cnt[(level + 2) % 3] = 0;
                  s_cnt[(level + 1) % 3] = nxt_n;
                  s_ust = ust;
                }
                wave_sync();  // the next level's queue, its count and the U list
                ++level;
                if (!imported && __ballot(imp) != 0) {
                  cmd = kCmdImport;  // OR in X_{src-1} before the next level
                  cslot = slot1;
                  break;
                }
              }
              if (cmd >= 0) break;
            }
            if (phase == 3) {
              // ---- this row's totals: the leader's own sums plus the cooperative phases' ----
              // (a sum no lane contributed to costs no reduction: a row that reaches nothing new
              // below itself has only its seed sums)
              auto wsum = [&](uint32_t x) -> uint32_t { return __ballot(x != 0) ? wave_sum_u32(x) : 0u; };
              const uint64_t a_w = (uint64_t)wsum(w_all) + s_wall;
              const uint64_t a_wo = (uint64_t)wsum(w_ovl) + s_wovl;
              const uint64_t a_x = (uint64_t)wsum(x_e) + s_xe;
              const uint64_t a_r = (uint64_t)wsum(r_all) + s_rall;
              const uint64_t a_ro = (uint64_t)wsum(r_ovl) + s_rovl;
              const uint32_t a_us = wsum(u_seed) + s_useed;
              const uint32_t a_un = wsum(u_new) + s_unew;
              const uint64_t dg0 = (uint64_t)wsum(deg0) + s_deg0;
              const RowFacts base = s_facts[d][fi];
              uint64_t wr, rr;
              uint32_t U;
              if (imported) {
                wr = base.wr + base.deg + (a_w - a_wo);
                rr = base.r + 1 + (a_r - a_ro);
                U = a_un;
              } else if (chained) {
                // U_t's entries up to src (the gap vertices and src itself) are not U entries of src
                wr = base.wr + base.deg + a_w;
                rr = base.r + 1 + a_r;
                U = base.u - gap - (uint32_t)s_src_was_set + a_un;
              } else {
                wr = a_w;
                rr = a_r;
                U = a_un;
              }
              // The U list: entries above src, compacted in place and written to the structure.
              uint32_t kept = 0;
              bool list_ok = false;
              if (want_list) {
                const uint32_t len = ubase + s_ust;
                const int64_t off = s_lp[d][li];
                const bool lp_ok = s_lp[d][li + 1] - off == (int64_t)U;
                int32_t* const dst = M.pool[d] + off;
                // 256 entries per round, all read before any is written: an entry moves to a
                // position at or below its own, so a round never overwrites what it still reads
                // and later rounds read only positions no round has written
                int32_t* const ul = d ? ul1 : ul0;
                for (uint32_t i0 = 0; i0 < len; i0 += 256) {
                  // a round wholly inside the list's LDS part (the usual case) reads and writes
                  // LDS without the per-entry test for the HBM part, so its reads overlap
                  const bool in_lds = (int64_t)i0 + 256 <= (int64_t)ucap;
                  int32_t e[4];
                  if (in_lds) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                      const uint32_t i = i0 + 64 * q + lane;
                      e[q] = i < len ? ul[i] : -1;
                    }
                  } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                      const uint32_t i = i0 + 64 * q + lane;
                      e[q] = i < len ? uget(i) : -1;
                    }
                  }
#pragma unroll
                  for (int q = 0; q < 4; ++q) {
                    const bool keep = e[q] > src;
                    const uint64_t mk = __ballot(keep);
                    const uint32_t pos = kept + rank_below(mk);
                    if (keep) {
                      if (lp_ok && pos < U) gstore(dst + pos, e[q]);
                      if (in_lds) ul[pos] = e[q];  // pos <= i0 + 255 < ucap
                      else uput(pos, e[q]);
                    }
                    kept += (uint32_t)__popcll(mk);
                  }
                }
                list_ok = lp_ok && kept == U;
              }
              const bool next_in = src + 1 < r1;
              // Save the bitmap for later rows: paired, whenever a row follows in the block (it
              // may chain from this one across wave-path rows, and the LDS copy will not survive
              // the other direction); plain, when the next row is escalated and does not hold
              // this one (only an import can use it).
              // Mirrored: never (s+1 chains from the LDS copy when it can, and otherwise s is not in
              // R_{s+1}, so no import takes X_s).
              const bool save = paired ? next_in
                                       : !M.sym && next_in && s_esc[d][li + 1] != 0 && s_hold[d][li + 1] == 0;
              const int wslot = s_ls.wp[d];  // the ring slot this row's save takes
              if (lane == 0) {
                const RowFacts f{wr, dg0, rr, U, want_list ? kept : 0u};
                s_facts[d][kSavedRing] = f;
                if (save) s_facts[d][wslot] = f;
                if (mb != last_mat) {
                  flush_totals(mats, last_mat, tot, exec_slot);
                  last_mat = mb;
                }
                gstore(M.cnt[d] + src, (int32_t)U);
                tot.fill += U - a_us;
                tot.heavy += 1;
                tot.x_edges += a_x;
                tot.x_verts += a_r + 1;
                tot.x_out += U;
                if (d == 0) {
                  tot.w += dg0 + wr;
                  tot.r += rr;
                  tot.nnz_u += U;
                } else {
                  tot.nnz_l += U;
                }
                if (d == 1 || M.sym) {
                  // window bits straight from the bitmap: rows src+1 .. src+128 of this column's
                  // chunk (a mirrored matrix: L's column src is U's row src)
                  uint32_t win[4];
                  for (int q = 0; q < 4; ++q) {
                    const int32_t lo = src + 1 + 32 * q;  // first row of this window word
                    uint32_t word = 0;
                    if (lo < chunk_end) {
                      const int32_t wi = lo >> 5, sh = lo & 31;
                      const uint32_t a = bm_get(wi);
                      const uint32_t b = (wi + 1 < bw) ? bm_get(wi + 1) : 0u;
                      word = sh ? ((a >> sh) | (b << (32 - sh))) : a;
                      const int32_t valid = chunk_end - lo;  // rows of the chunk left
                      if (valid < 32) word &= (1u << valid) - 1u;
                    }
                    win[q] = word;
                  }
                  reinterpret_cast<uint4*>(M.win)[src] = make_uint4(win[0], win[1], win[2], win[3]);
                }
                if (want_list && !list_ok) atomicOr(&M.acc[kAccOverflow], 1ull);
              }
              wave_sync();  // the compacted U list and the row's facts
              phase = 4;
              if (save) {
                cmd = kCmdSave;
                cslot = wslot;
                break;
              }
            }
            if (phase == 4) {
              // ---- bookkeeping after the row (and its save) ----
              const bool saved = paired ? src + 1 < r1
                                        : !M.sym && src + 1 < r1 && s_esc[d][li + 1] != 0 && s_hold[d][li + 1] == 0;
              if (lane == 0) {
                if (saved) {
                  const int wslot = s_ls.wp[d];
                  s_ls.sv[d][wslot] = src;
                  s_ls.wp[d] = wslot + 1 == kRing ? 0 : wslot + 1;
                }
                s_ls.prev[d] = src;
                s_ls.lds_row = src;
                s_ls.lds_dir = d;
              }
              wave_sync();  // the item state for the next row
              phase = 0;
            }
          }
          if (phase == 2) {
            // the row continues after the workgroup's part: keep this turn's sums in LDS
            const uint32_t a_w = wave_sum_u32(w_all), a_wo = wave_sum_u32(w_ovl), a_x = wave_sum_u32(x_e);
            const uint32_t a_r = wave_sum_u32(r_all), a_ro = wave_sum_u32(r_ovl), a_us = wave_sum_u32(u_seed);
            const uint32_t a_un = wave_sum_u32(u_new), a_d0 = wave_sum_u32(deg0);
            if (lane == 0) {
              s_wall += a_w;
              s_wovl += a_wo;
              s_xe += a_x;
              s_rall += a_r;
              s_rovl += a_ro;
              s_useed += a_us;
              s_unew += a_un;
              s_deg0 += a_d0;
            }
          }
          if (lane == 0) {
            s_ls.li_next = li_next;
            s_ls.di_next = di_next;
            s_ls.phase = phase;
            s_ls.d = d;
            s_ls.li = li;
            s_ls.fi = fi;
            s_ls.slot1 = slot1;
            s_ls.tslot = tslot;
            s_ls.level = level;
            s_ls.src = src;
            s_ls.t = t;
            s_ls.chunk_end = chunk_end;
            s_ls.chained = chained ? 1 : 0;
            s_ls.has_prev = has_prev ? 1 : 0;
            s_ls.imported = imported ? 1 : 0;
            s_ls.ubase = ubase;
            s_ls.gap = gap;
            s_ctl.cmd = cmd;
            s_ctl.d = d;
            s_ctl.src = src;
            s_ctl.level = level;
            s_ctl.imported = imported ? 1 : 0;
            s_ctl.slot = cslot;
            s_ctl.ubase = ubase;
            s_ctl.has_prev = has_prev ? 1 : 0;
          }
        }
        bsync();
        const HeavyCtl ctl = s_ctl;
        if (ctl.cmd == kCmdDone) break;
        const int cd = ctl.d;
        const int32_t csrc = ctl.src;
        if (ctl.cmd == kCmdClear) {
          const int32_t sw = csrc >> 5;
          const uint32_t sb = 1u << (csrc & 31);
          for (int32_t w = tid; w < bw; w += kThreads) bm[w] = w == sw ? sb : 0u;
          if (tid == 0) s_src_was_set = 0;
          bsync();
        } else if (ctl.cmd == kCmdCopyIn) {
          // X_t back from its saved copy, with the bit of src set
          const uint32_t* const g = gslot(cd, ctl.slot);
          for (int32_t wb = 0; wb < bw; wb += kThreads * kImportBatch) {
            uint32_t pw[kImportBatch];
#pragma unroll
            for (int q = 0; q < kImportBatch; ++q) {
              const int32_t w = wb + q * kThreads + tid;
              pw[q] = w < bw ? gload(g + w) : 0u;
            }
#pragma unroll
            for (int q = 0; q < kImportBatch; ++q) {
              const int32_t w = wb + q * kThreads + tid;
              if (w < bw) {
                uint32_t x = pw[q];
                if (w == (csrc >> 5)) {
                  s_src_was_set = (x >> (csrc & 31)) & 1u;
                  x |= 1u << (csrc & 31);
                }
                bm[w] = x;
              }
            }
          }
          bsync();
        } else if (ctl.cmd == kCmdSave) {
          uint32_t* const gout = gslot(cd, ctl.slot);
          for (int32_t w = tid; w < bw; w += kThreads) gstore(gout + w, bm_get(w));
          bsync();
        } else {
          // kCmdImport and kCmdWide share the level machinery of all waves.
          int32_t* const ustage = cd ? ustage1 : ustage0;
          int32_t* const ul = cd ? ul1 : ul0;
          auto uput = [&](uint32_t i, int32_t v) {
            if ((int32_t)i < ucap) ul[i] = v;
            else gstore(ustage + ((int32_t)i - ucap), v);
          };
          const uint32_t* const gprev = gslot(cd, ctl.slot < 0 ? 0 : ctl.slot);  // saved X_{s-1}
          const bool chas_prev = ctl.has_prev != 0;
          const uint32_t cubase = ctl.ubase;
          bool cimported = ctl.imported != 0;
          uint32_t p_w = 0, p_wo = 0, p_x = 0, p_r = 0, p_ro = 0, p_un = 0;
          // OR in X_{src-1}; its new bits above src are U entries of src, the rest is R_{src-1}
          auto import_prev = [&]() {
            for (int32_t wb = 0; wb < bw; wb += kThreads * kImportBatch) {
              uint32_t pw[kImportBatch];
#pragma unroll
              for (int q = 0; q < kImportBatch; ++q) {
                const int32_t w = wb + q * kThreads + tid;
                pw[q] = w < bw ? gload(gprev + w) : 0u;
              }
#pragma unroll
              for (int q = 0; q < kImportBatch; ++q) {
                const int32_t w = wb + q * kThreads + tid;
                uint32_t nb = 0;
                if (w < bw) {
                  uint32_t cw;
                  if constexpr (kHbmBitmap) {
                    cw = pw[q] ? atomicOr(&bm[w], pw[q]) : 0u;  // keeps the level's L2 claims
                  } else {
                    cw = bm[w];
                    bm[w] = cw | pw[q];
                  }
                  nb = pw[q] & ~cw;
                  const int32_t b0 = w * 32;
                  if (b0 + 31 <= csrc) nb = 0;
                  else if (b0 <= csrc) nb &= ~0u << ((csrc - b0) & 31) << 1;  // keep bits above src
                }
                p_un += __popc(nb);
                if (want_list) {
                  const uint32_t cnt = __popc(nb);
                  // wave-aggregated reservation (inclusive scan of the counts), then each lane writes
                  const uint32_t incl = wave_incl_scan_u32(cnt);
                  const uint32_t wave_total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                  uint32_t basep = 0;
                  if (lane == 0 && wave_total) basep = atomicAdd(&s_ust, wave_total);
                  basep = cubase + (uint32_t)__builtin_amdgcn_readlane((int)basep, 0);
                  uint32_t pos = basep + incl - cnt;
                  while (nb) {
                    const int b = __ffs(nb) - 1;
                    nb &= nb - 1;
                    uput(pos++, w * 32 + b);
                  }
                }
              }
            }
          };
          int clevel = ctl.level;
          if (ctl.cmd == kCmdImport) {
            import_prev();
            cimported = true;
            bsync();
          } else {
            // levels wider than kSoloMax, all waves (one vertex per thread), until one is narrow
            const MatDev& CM = M;
            for (;; ++clevel) {
              const int32_t ncur = s_cnt[clevel % 3];
              if (ncur <= kSoloMax) break;
              if (tid == 0) s_cnt[(clevel + 2) % 3] = 0;
              const bool odd = (clevel & 1) != 0;
              const int32_t* const cur_l = odd ? ql1 : ql0;
              const int32_t* const cur_g = odd ? qg1 : qg0;
              int32_t* const nxt_l = odd ? ql0 : ql1;
              int32_t* const nxt_g = odd ? qg0 : qg1;
              int32_t* nxt_cnt = &s_cnt[(clevel + 1) % 3];
              for (int32_t base = 0; base < ncur; base += kThreads) {
                const int32_t e = base + tid;
                int32_t v = e < ncur ? qget(cur_l, cur_g, qcap, e) : -1;
                // issued together: membership in the saved X_{s-1}, the row, its duplicate surplus
                const bool chk = v >= 0 && chas_prev;
                const uint32_t pword = chk ? gload(gprev + (v >> 5)) : 0u;
                RowCursor rc;
                load_row(CM, cd, v, rc);
                const uint32_t dupv = (v >= 0 && cd == 0 && CM.dup_all) ? (uint32_t)gload(CM.dup_all + v) : 0u;
                const bool in_prev = chk && ((pword >> (v & 31)) & 1u) != 0;
                if (in_prev && cimported) {  // covered by the imported X_{s-1}
                  v = -1;
#pragma unroll
                  for (int k = 0; k < kEll; ++k) rc.e[k] = kPad;
                  rc.ov_pos = rc.ov_end = 0;
                }
                uint32_t deg = diag_mark(rc);
                // Claim the bits of 8 slots at a time: the LDS atomics are independent and in
                // flight together. Rows are packed, so the second group is skipped when no lane
                // has a ninth entry. New entries below src go to the next level, new entries above
                // src are U.
                uint32_t mask_r = 0, mask_u = 0;
#pragma unroll
                for (int h = 0; h < kEll; h += 8) {
                  if (h > 0 && __ballot(rc.e[h] >= 0) == 0) break;
                  uint32_t old[8];
#pragma unroll
                  for (int j = 0; j < 8; ++j) {
                    const int32_t m = rc.e[h + j];
                    old[j] = m >= 0 ? atomicOr(&bm[m >> 5], 1u << (m & 31)) : ~0u;
                  }
#pragma unroll
                  for (int j = 0; j < 8; ++j) {
                    const int32_t m = rc.e[h + j];
                    if (m >= 0) {
                      ++deg;
                      if (((old[j] >> (m & 31)) & 1u) == 0) {
                        if (chas_prev && m == csrc - 1) s_import = 1;  // import X_{src-1} instead of expanding it
                        else if (m < csrc) mask_r |= 1u << (h + j);
                        else if (m > csrc) mask_u |= 1u << (h + j);
                      }
                    }
                  }
                }
                // One queue reservation per wave (exclusive scan of the per-lane counts).
                {
                  const uint32_t c = __popc(mask_r);
                  const uint32_t incl = wave_incl_scan_u32(c);
                  const uint32_t tot_w = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                  if (tot_w != 0) {
                    int32_t basepos = 0;
                    if (lane == 63) basepos = atomicAdd(nxt_cnt, (int32_t)tot_w);
                    const int32_t wbase = __builtin_amdgcn_readlane(basepos, 63);
                    const int32_t pos = wbase + (int32_t)(incl - c);
                    const uint32_t any = row_any(mask_r);
                    if (!kPairedAny && wbase + (int32_t)tot_w <= qcap)  // the whole wave's range is in LDS
                      put_row_entries(rc, mask_r, any, pos, [&](int32_t i, int32_t x) { nxt_l[i] = x; });
                    else
                      put_row_entries(rc, mask_r, any, pos, [&](int32_t i, int32_t x) { qput(nxt_l, nxt_g, qcap, i, x); });
                  }
                }
                p_un += __popc(mask_u);
                if (want_list && __ballot(mask_u != 0) != 0) {  // most passes find no U entry
                  const uint32_t c = __popc(mask_u);
                  const uint32_t incl = wave_incl_scan_u32(c);
                  const uint32_t tot_w = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                  if (tot_w != 0) {
                    uint32_t basepos = 0;
                    if (lane == 63) basepos = atomicAdd(&s_ust, tot_w);
                    const uint32_t pos = cubase + (uint32_t)__builtin_amdgcn_readlane((int)basepos, 63) + (incl - c);
                    put_row_entries(rc, mask_u, row_any(mask_u), (int32_t)pos,
                                    [&](int32_t i, int32_t x) { uput((uint32_t)i, x); });
                  }
                }
                // the CSR tail of a long row, one entry per lane per step
                auto visit = [&](int32_t m) {
                  const bool active = m >= 0;
                  deg += active;
                  bool isnew = false;
                  if (active) {
                    const uint32_t bit = 1u << (m & 31);
                    uint32_t* wp = &bm[m >> 5];
                    if ((bm_get(m >> 5) & bit) == 0) isnew = (atomicOr(wp, bit) & bit) == 0;
                  }
                  if (isnew && chas_prev && m == csrc - 1) {
                    s_import = 1;
                    isnew = false;
                  }
                  const bool is_r = isnew && m < csrc;
                  const bool is_u = isnew && m > csrc;
                  const uint64_t mr = __ballot(is_r);
                  if (mr) {
                    const int leader_lane = __ffsll((unsigned long long)mr) - 1;
                    int32_t basepos = 0;
                    if (lane == leader_lane) basepos = atomicAdd(nxt_cnt, (int32_t)__popcll(mr));
                    basepos = __shfl(basepos, leader_lane, 64);
                    if (is_r) qput(nxt_l, nxt_g, qcap, basepos + (int32_t)rank_below(mr), m);
                  }
                  p_un += is_u;
                  if (want_list) {
                    const uint64_t mu = __ballot(is_u);
                    if (mu) {
                      const int leader_lane = __ffsll((unsigned long long)mu) - 1;
                      uint32_t basepos = 0;
                      if (lane == leader_lane) basepos = atomicAdd(&s_ust, (uint32_t)__popcll(mu));
                      basepos = cubase + __shfl(basepos, leader_lane, 64);
                      if (is_u) uput(basepos + rank_below(mu), m);
                    }
                  }
                };
                while (__ballot(rc.ov_pos < rc.ov_end) != 0) {
                  int32_t m = kPad;
                  if (rc.ov_pos < rc.ov_end) m = gload(CM.adj[cd] + rc.ov_pos++);
                  visit(m);
                }
                p_x += deg;
                if (v >= 0) {
                  deg += dupv;
                  p_w += deg;
                  p_r += 1;
                  if (in_prev) {
                    p_wo += deg;
                    p_ro += 1;
                  }
                }
              }
              bsync();
              // s_import only ever goes 0 -> 1 during a row, and only in claims before this barrier
              if (!cimported && s_import != 0) {
                import_prev();
                cimported = true;
                bsync();
              }
            }
          }
          // this phase's sums into the row's LDS sums (per-row sums stay below the input nnz)
          {
            const uint32_t a_w = wave_sum_u32(p_w), a_wo = wave_sum_u32(p_wo), a_x = wave_sum_u32(p_x);
            const uint32_t a_r = wave_sum_u32(p_r), a_ro = wave_sum_u32(p_ro), a_un = wave_sum_u32(p_un);
            if (lane == 0) {
              if (a_w) atomicAdd(&s_wall, (unsigned long long)a_w);
              if (a_wo) atomicAdd(&s_wovl, (unsigned long long)a_wo);
              if (a_x) atomicAdd(&s_xe, (unsigned long long)a_x);
              if (a_r) atomicAdd(&s_rall, a_r);
              if (a_ro) atomicAdd(&s_rovl, a_ro);
              if (a_un) atomicAdd(&s_unew, a_un);
            }
          }
          if (tid == 0) {  // the leader resumes its levels where the workgroup stopped
            s_ls.level = clevel;
            s_ls.imported = cimported ? 1 : 0;
          }
          bsync();
        }
      }
    }
  }
  if (tid == 0) flush_totals(mats, last_mat, tot, exec_slot);
}
#undef GSOFA_SLOT
#undef GSOFA_SV

// ------------------------------------------------------------------------------------------
// k_supernodes: the reference's rule (warpcentric_local_detection_test,
// gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:1790-2123), one thread per row.
//   start(r): (r mod S) mod 128 == 0, or nzU[r-1] - nzU[r] != 1        (:1833, :1879, :1928)
//   other rows pass through; a pass-through row r extends the run of the last start s of its
//   chunk iff (r, s) is in F (:2034), else it counts one more supernode (:2047).
// Starts recur at least every 128 rows of a chunk, so s >= r - 127 and (r, s) is one of the
// 128 window bits of column s.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int32_t nzu_at(const MatDev& M, int32_t r) {
  return M.cnt[0][r] + (M.dup_u ? M.dup_u[r] : 0);
}

__global__ __launch_bounds__(256) void k_supernodes(const MatDev* __restrict__ mats, int nmats,
                                                    const int64_t* __restrict__ blk_base) {
  // blk_base[b] = first block of matrix b; each block covers 256 rows r0 .. r0 + 255 of one
  // matrix. A row's supernode start lies at most 127 rows before it (row r - ((r mod S) & 127)
  // starts one), so the block reads the counts of the 129 rows before r0 as well, marks the
  // starts of rows r0 - 128 .. r0 + 255 as LDS bits and finds each row's start with a few word
  // scans (a walk row by row through global counts chains up to 127 loads).
  constexpr int kBack = 128;
  __shared__ int32_t s_nzu[kBack + 1 + 256];      // rows r0 - 129 .. r0 + 255
  __shared__ uint32_t s_st[(kBack + 256) / 32];  // start bits of rows r0 - 128 .. r0 + 255
  __shared__ uint32_t s_sum;
  const int tid = threadIdx.x;
  if (tid == 0) s_sum = 0;
  if (tid < (kBack + 256) / 32) s_st[tid] = 0;
  const int mb = find_segment(blk_base, nmats, (int64_t)blockIdx.x);
  const MatDev& M = mats[mb];
  const int32_t n = M.n;
  const int32_t r0 = (int32_t)(((int64_t)blockIdx.x - blk_base[mb]) * 256);
  for (int i = tid; i < kBack + 1 + 256; i += 256) {
    const int32_t row = r0 - kBack - 1 + i;
    s_nzu[i] = (row >= 0 && row < n) ? nzu_at(M, row) : 0;
  }
  __syncthreads();
  for (int j = tid; j < kBack + 256; j += 256) {
    const int32_t row = r0 - kBack + j;  // its count at s_nzu[j + 1], row - 1's at s_nzu[j]
    if (row < 0 || row >= n) continue;
    if (((row % M.chunk) & 127) == 0 || s_nzu[j] - s_nzu[j + 1] != 1) atomicOr(&s_st[j >> 5], 1u << (j & 31));
  }
  __syncthreads();
  uint32_t flag = 0;
  const int32_t r = r0 + tid;
  if (r < n) {
    const int j = kBack + tid;
    if ((s_st[j >> 5] >> (j & 31)) & 1u) {
      flag = 1;
    } else {
      // the last start before r: the highest set bit below j (at least j - 127 >= 1)
      int w = j >> 5;
      uint32_t bits = s_st[w] & ((1u << (j & 31)) - 1u);
      while (bits == 0 && w > 0) bits = s_st[--w];
      if (bits == 0) {
        flag = 1;  // cannot happen (a start lies within 127 rows); keeps every read in bounds
      } else {
        const int32_t s = r0 - kBack + w * 32 + 31 - __clz(bits);
        const int32_t t = r - s - 1;
        const uint32_t word = M.win[(int64_t)s * 4 + (t >> 5)];
        flag = ((word >> (t & 31)) & 1u) ? 0u : 1u;
      }
    }
    M.sup_flag[r] = (uint8_t)flag;
  }
  const uint64_t c = wave_sum_u64(flag);
  if (lane_id() == 0 && c) atomicAdd(&s_sum, (uint32_t)c);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum) atomicAdd(&M.acc[kAccNsup], (unsigned long long)s_sum);
}

// k_mirror: a mirrored matrix was traversed from A only. From source s, A^T reaches exactly what
// A reaches (the pattern is its own transpose), so L's column s holds U's row s: its count is
// copied, the fill and the workgroup path's entries (both triangles) double, nnz_L = nnz_U.
// Runs after both traversal kernels, before k_supernodes.
__global__ __launch_bounds__(256) void k_mirror(const MatDev* __restrict__ mats, int nmats,
                                                const int64_t* __restrict__ blk_base) {
  const int mb = find_segment(blk_base, nmats, (int64_t)blockIdx.x);
  const MatDev& M = mats[mb];
  if (!M.sym) return;
  const int64_t r = ((int64_t)blockIdx.x - blk_base[mb]) * 256 + threadIdx.x;
  if (r < M.n) M.cnt[1][r] = M.cnt[0][r];
  if ((int64_t)blockIdx.x == blk_base[mb] && threadIdx.x == 0) {
    unsigned long long* acc = M.acc;
    acc[kAccFill] *= 2;
    acc[kAccNnzL] = acc[kAccNnzU];
    acc[kAccHeavyNnz] *= 2;
  }
}

// Before the early launch of the top separators' items: their rows escalated in both directions
// (k_heavy runs a window's escalated rows), and the launch's own work counter. One block per
// matrix; matrices without such rows (light_end == n) do nothing.
__global__ __launch_bounds__(256) void k_top_setup(const MatDev* __restrict__ mats, WorkState* __restrict__ ws,
                                                   int64_t count) {
  const MatDev& M = mats[blockIdx.x];
  for (int32_t r = M.light_end + (int32_t)threadIdx.x; r < M.n; r += (int32_t)blockDim.x) {
    M.esc[0][r] = 1;
    M.esc[1][r] = 1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ws->heavy_count = (unsigned long long)count;
    ws->heavy_next = 0ull;
  }
}

// ------------------------------------------------------------------------------------------
// host-side launchers (the engine calls these; kernels stay private to this file)
// ------------------------------------------------------------------------------------------
static int grid_for(int64_t work, int block, int cap) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

hipError_t launch_mirror(const MatDev* mats, int nmats, const int64_t* blk_base, int64_t nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mirror, dim3((unsigned)nblocks), dim3(256), 0, s, mats, nmats, blk_base);
  return hipGetLastError();
}

hipError_t launch_top_setup(const MatDev* mats, int nmats, WorkState* ws, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_top_setup, dim3((unsigned)nmats), dim3(256), 0, s, mats, ws, count);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, int64_t count, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(count, 256, 4096)), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

hipError_t launch_clear_run(unsigned long long* acc, int64_t nacc, unsigned long long* ws, int64_t nws, int32_t* esc,
                            int64_t nesc, int32_t* blk_flag, int64_t nblk, int32_t* bucket, int64_t nbucket,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_clear_run, dim3(grid_for(nacc + nws + nesc + nblk + nbucket, 256, 8192)), dim3(256), 0, s, acc,
                     nacc, ws, nws, esc, nesc, blk_flag, nblk, bucket, nbucket);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int64_t count, int32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(count, 256, 8192)), dim3(256), 0, s, p, count, v);
  return hipGetLastError();
}

hipError_t launch_ell_build(int64_t nrows, int nmats, const int64_t* row_base, const int32_t* rowptr,
                            const int32_t* adj, int32_t* ell, hipStream_t s) {
  hipLaunchKernelGGL(k_ell_build, dim3(grid_for(nrows, 256, 16384)), dim3(256), 0, s, nrows, nmats, row_base, rowptr,
                     adj, ell);
  return hipGetLastError();
}

hipError_t launch_transpose(int64_t nrows, int nmats, const int64_t* row_base, const int32_t* rowptr,
                            const int32_t* adj, int32_t* col_cnt, int32_t* rowptr_t, int32_t* cursor,
                            int32_t* adj_t, int32_t* block_sums, int32_t* total, hipStream_t s) {
  const int g = grid_for(nrows, 256, 16384);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nrows, 256, 8192)), dim3(256), 0, s, col_cnt, nrows, 0);
  hipLaunchKernelGGL(k_col_hist, dim3(g), dim3(256), 0, s, nrows, nmats, row_base, rowptr, adj, col_cnt);
  const int64_t per = (int64_t)kScanBlock * kScanItems;
  const int64_t nb = (nrows + per - 1) / per;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanBlock), 0, s, col_cnt, nrows, block_sums);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, s, block_sums, nb, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanBlock), 0, s, col_cnt, nrows, block_sums, rowptr_t,
                     cursor, total);
  hipLaunchKernelGGL(k_transpose_scatter, dim3(g), dim3(256), 0, s, nrows, nmats, row_base, rowptr, adj, cursor, adj_t);
  return hipGetLastError();
}

hipError_t launch_light(const MatDev* mats, int nmats, const int64_t* item_base, int64_t nitems, WorkState* ws, int grid,
                        int bits, hipStream_t s) {
  if (bits == kLightBitsBig)
    hipLaunchKernelGGL(k_light<kLightBitsBig>, dim3(grid), dim3(kLightWaves * 64), 0, s, mats, nmats, item_base, nitems,
                       ws);
  else
    hipLaunchKernelGGL(k_light<kLightBits>, dim3(grid), dim3(kLightWaves * 64), 0, s, mats, nmats, item_base, nitems, ws);
  return hipGetLastError();
}

hipError_t launch_heavy_order(const MatDev* mats, int64_t nblk_total, int nmats, const int64_t* blkq_base,
                              const uint32_t* blk_flag, int32_t nb_max, int32_t* bucket, int32_t* cursor,
                              int32_t* block_sums, int32_t* total, WorkState* ws, uint64_t* heavy_items, hipStream_t s) {
  const int64_t nslots = 2 * nblk_total;
  const int g = grid_for(nslots, 256, 8192);
  // bucket is zero already (the run's clears)
  hipLaunchKernelGGL(k_heavy_hist, dim3(g), dim3(256), 0, s, mats, nblk_total, nmats, blkq_base, blk_flag, nb_max, bucket,
                     ws);
  if (nb_max <= 16 * kScanBlock) {
    // one workgroup scans the buckets in place; they then serve as the scatter's cursors
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, s, bucket, (int64_t)nb_max, total);
    hipLaunchKernelGGL(k_heavy_scatter, dim3(g), dim3(256), 0, s, mats, nblk_total, nmats, blkq_base, blk_flag, nb_max,
                       bucket, heavy_items);
    return hipGetLastError();
  }
  const int64_t per = (int64_t)kScanBlock * kScanItems;
  const int64_t nb = (nb_max + per - 1) / per;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanBlock), 0, s, bucket, (int64_t)nb_max, block_sums);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, s, block_sums, nb, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanBlock), 0, s, bucket, (int64_t)nb_max, block_sums,
                     cursor, nullptr, total);
  hipLaunchKernelGGL(k_heavy_scatter, dim3(g), dim3(256), 0, s, mats, nblk_total, nmats, blkq_base, blk_flag, nb_max,
                     cursor, heavy_items);
  return hipGetLastError();
}

size_t heavy_lds_bytes(int32_t bw_max, int32_t qcap, bool hbm_bitmap, int32_t ucap, bool paired_any) {
  return (hbm_bitmap ? 0 : (size_t)bw_max * 4) + (size_t)qcap * 8 + (size_t)ucap * 4 * (paired_any ? 2 : 1) +
         heavy_stash_bytes(paired_any);
}

// The instantiation for (threads, paired, bitmap in HBM); the HBM bitmap runs 1024-thread
// workgroups only. nullptr: no such instantiation.
template <bool kP, bool kH>
const void* heavy_fn(int threads) {
  if constexpr (kH) {
    return threads == 1024 ? (const void*)k_heavy<1024, kP, true> : nullptr;
  } else {
    switch (threads) {
      case 1024: return (const void*)k_heavy<1024, kP, false>;
      case 512: return (const void*)k_heavy<512, kP, false>;
      case 256: return (const void*)k_heavy<256, kP, false>;
      default: return nullptr;
    }
  }
}

const void* heavy_fn(int threads, bool paired, bool hbm) {
  if (paired) return hbm ? heavy_fn<true, true>(threads) : heavy_fn<true, false>(threads);
  return hbm ? heavy_fn<false, true>(threads) : heavy_fn<false, false>(threads);
}

hipError_t heavy_set_lds_limit(int threads, size_t bytes, bool hbm_bitmap) {
  for (bool paired : {false, true}) {
    const void* f = heavy_fn(threads, paired, hbm_bitmap);
    if (!f) return hipErrorInvalidValue;
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_heavy(int threads, int grid, const MatDev* mats, WorkState* ws, const uint64_t* heavy_items,
                        int32_t bw_max, int32_t qcap, int64_t n_max, int32_t paired_any, int32_t* spill,
                        uint32_t* hbm_bitmaps, int32_t ucap, int32_t exec_slot, hipStream_t s) {
  const bool hbm = hbm_bitmaps != nullptr;
  const void* f = heavy_fn(threads, paired_any != 0, hbm);
  if (!f) return hipErrorInvalidValue;
  void* args[] = {&mats, &ws, &heavy_items, &bw_max, &qcap, &n_max, &spill, &hbm_bitmaps, &ucap, &exec_slot};
  return hipLaunchKernel(f, dim3(grid), dim3(threads), args,
                         heavy_lds_bytes(bw_max, qcap, hbm, ucap, paired_any != 0), s);
}

hipError_t launch_supernodes(const MatDev* mats, int nmats, const int64_t* blk_base, int64_t nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_supernodes, dim3((unsigned)nblocks), dim3(256), 0, s, mats, nmats, blk_base);
  return hipGetLastError();
}

}  // namespace gsofa

namespace gsofa {
// Streaming copy of 16-byte words: the calibration kernel for HBM byte counters and the
// achievable-bandwidth probe (gsofa_bandwidth_probe).
__global__ __launch_bounds__(256) void k_copy16(const v4i32* __restrict__ src, v4i32* __restrict__ dst, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy16(const void* src, void* dst, int64_t n16, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, static_cast<const v4i32*>(src), static_cast<v4i32*>(dst),
                     n16);
  return hipGetLastError();
}

// Resident workgroups per CU of both paired variants (the grid must fit either).
int heavy_occupancy(int threads, size_t lds_bytes, bool hbm_bitmap) {
  int occ = 1 << 30;
  for (bool paired : {false, true}) {
    const void* f = heavy_fn(threads, paired, hbm_bitmap);
    int nb = 0;
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, threads, lds_bytes) != hipSuccess) return 0;
    occ = std::min(occ, nb);
  }
  return occ;
}

int light_occupancy(int bits) {
  int nb = 0;
  const hipError_t e = bits == kLightBitsBig
                           ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_light<kLightBitsBig>, kLightWaves * 64, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_light<kLightBits>, kLightWaves * 64, 0);
  return e == hipSuccess ? nb : 0;
}
}  // namespace gsofa
