// Device-side data layout shared by the HIP kernels and the host engine.
#pragma once

#include <stdint.h>

namespace gsofa {

// Rows are packed into 64-byte slots of 16 int32 ("ELL-16"): the entries first, then -1
// padding. A row with more than 16 entries keeps its first 15 in the slot and stores
// -(p + 2) in entry 15, where p is the CSR position of its 16th entry; the rest of the row is
// read from the CSR adjacency up to rowptr[v + 1].
constexpr int kEll = 16;
constexpr int32_t kPad = -1;
// A row of at most 15 entries that holds its diagonal keeps the diagonal out of its slots (it is
// never new to a traversal) and marks it with kDiag after the off-diagonal entries: claims stop
// at the packed entries, degree counts add the mark. Slot 15 never holds kDiag.
constexpr int32_t kDiag = -2;

// Accumulator slots (u64) per matrix.
enum : int {
  kAccFill = 0,      // off-diagonal fill-ins, both triangles
  kAccW = 1,         // schedule-independent edge checks of the row traversals
  kAccR = 2,         // sum of |R_i| of the row traversals
  kAccOverflow = 5,  // nonzero: a row's count differs from the structure offsets
  kAccNsup = 6,      // supernodes counted by the reference rule
  kAccNnzU = 7,      // strict upper entries of F
  kAccNnzL = 8,      // strict lower entries of F
  kAccHeavy = 9,     // sources that took the workgroup path
  kAccHeavyW = 10,   // edge checks of row traversals done on the workgroup path
  kAccHeavyR = 11,   // |R_i| of row traversals done on the workgroup path
  kAccHeavyNnz = 12, // strict F entries found on the workgroup path (both directions)
  // work the kernels actually executed (both directions; chaining and the wave path's hash
  // make these differ from the schedule-independent W and |R|)
  kAccHeavyExecEdges = 13,  // neighbour inspections done by k_heavy
  kAccHeavyExecVerts = 14,  // rows expanded by k_heavy
  kAccHeavyExecOut = 15,    // F entries k_heavy wrote (U lists incl. carried entries)
  kAccLightExecEdges = 16,  // neighbour inspections done by k_light (escalated attempts included)
  kAccLightExecVerts = 17,  // rows expanded by k_light
  kAccLightExecOut = 18,    // F entries of the sources k_light finished
  kAccScheduleError = 19,   // nonzero: a wave-path wave stopped waiting for its sibling's next item
  // the same three for the early launch of the top separator's items (kept apart so that each
  // k_heavy launch is priced on its own work; kAccHeavyExec* counts the main launch only)
  kAccTopExecEdges = 20,
  kAccTopExecVerts = 21,
  kAccTopExecOut = 22,
  kAccSlots = 24
};

// One matrix as the kernels see it. Direction 0 traverses A (rows, gives U), direction 1
// traverses A^T (columns of A, gives L by columns).
struct MatDev {
  int32_t n;
  int32_t chunk;                  // supernode chunk size S
  const int32_t* ell[2];          // n * kEll
  const int32_t* rowptr[2];       // n + 1 (overflow rows only)
  const int32_t* adj[2];          // CSR adjacency (overflow rows only)
  int32_t* cnt[2];                // n: |U_i| (dir 0) / |L column j| (dir 1)
  uint32_t* win;                  // n * 4: bits t of column j: row j+1+t in L, same chunk
  const int32_t* dup_u;           // n: duplicate surplus of A's upper entries (may be null)
  const int32_t* dup_all;         // n: duplicate surplus of A's rows, for W (may be null)
  const int64_t* list_ptr[2];     // n + 1: row/column offsets into pool (null: no structure)
  int32_t* pool[2];               // strict U rows (dir 0) / strict L columns (dir 1)
  uint8_t* sup_flag;              // n: 1 where the supernode count increments
  unsigned long long* acc;        // kAccSlots
  uint8_t* esc[2];                // n: 1 where the wave path escalated the source
  uint32_t* blk_flag[2];          // ceil(n / chain_block): 1 where the block has an escalated row
  const int32_t* split;           // ceil(n / chain_block): a row where the block is cut into two
                                  // workgroup-path items, or -1 (null: no block is cut)
  int32_t chain_block;            // K: rows per workgroup-path item (1 disables chaining)
  int32_t paired;                 // 1: workgroup-path items run both directions of a block
  int32_t sym;                    // 1: mirrored (structurally symmetric: direction 0 only, L's
                                  // counts and totals copied from U's by k_mirror)
  int32_t light_block;            // rows per wave-path item
  int32_t light_end;              // the wave path runs rows [0, light_end); rows from light_end on
                                  // (a top separator's, n: none) are fixed workgroup-path items of
                                  // an early launch, escalated from the start
  int32_t light_limit;            // reached vertices before the wave path escalates a row
  const int32_t* subtree;         // n: elimination-subtree sizes of A + A^T (null: none)
  int32_t pre_size;               // a row whose subtree exceeds this escalates without a
                                  // wave-path attempt (its reached set would exceed the limit)
};

// Scheduler state of one launch set (plain matrix or batch).
struct WorkState {
  unsigned long long light_next;   // next pair of wave-path items
  unsigned long long heavy_count;  // items escalated to the workgroup path
  unsigned long long heavy_next;   // next escalated item
  unsigned long long pad;
};

// Item encoding for the escalation list: rows before a cut (bit 63), extra blocks e (bits
// 53..62), matrix index (bits 33..52: sets hold at most 2^20 matrices), direction (bit 32), a row
// v (bits 0..31). The item runs rows v to the end of the e-th block after v's block of
// chain_block rows (v is the block's first row, or a cut; e > 0 only for the early launch's
// items), or, with bit 63, the rows of v's block before v.
constexpr int kItemExtraMax = 1023;
__host__ __device__ inline uint64_t encode_item(uint32_t mat, uint32_t dir, uint32_t row, uint32_t before_cut,
                                                uint32_t extra = 0) {
  return ((uint64_t)before_cut << 63) | ((uint64_t)extra << 53) | ((uint64_t)mat << 33) | ((uint64_t)dir << 32) |
         (uint64_t)row;
}

}  // namespace gsofa

// ---- launchers (defined in kernels.hip) ---------------------------------------------------
#include <hip/hip_runtime_api.h>
#include <stddef.h>

namespace gsofa {
hipError_t launch_fill_u64(unsigned long long* p, int64_t count, unsigned long long v, hipStream_t s);
// A run's clears in one launch (all to zero): nacc u64 at acc, nws u64 at ws, nesc, nblk and
// nbucket i32 words at esc, blk_flag and bucket.
hipError_t launch_clear_run(unsigned long long* acc, int64_t nacc, unsigned long long* ws, int64_t nws, int32_t* esc,
                            int64_t nesc, int32_t* blk_flag, int64_t nblk, int32_t* bucket, int64_t nbucket,
                            hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int64_t count, int32_t v, hipStream_t s);
hipError_t launch_ell_build(int64_t nrows, int nmats, const int64_t* row_base, const int32_t* rowptr,
                            const int32_t* adj, int32_t* ell, hipStream_t s);
// Builds A^T of every matrix: col_cnt/cursor/rowptr_t have nrows (+1) entries, block_sums
// ceil(nrows / 4096), total one int.
hipError_t launch_transpose(int64_t nrows, int nmats, const int64_t* row_base, const int32_t* rowptr,
                            const int32_t* adj, int32_t* col_cnt, int32_t* rowptr_t, int32_t* cursor,
                            int32_t* adj_t, int32_t* block_sums, int32_t* total, hipStream_t s);
// bits: kLightBits, or kLightBitsBig (a hash set of 2,048 slots and an escalation limit of 1,024)
hipError_t launch_light(const MatDev* mats, int nmats, const int64_t* item_base, int64_t nitems, WorkState* ws, int grid, int bits,
                        hipStream_t s);
// The escalated blocks (blk_flag, 2 * nblk_total slots) into heavy_items in descending block
// order; sets ws->heavy_count. bucket has nb_max entries, cursor nb_max + 1, block_sums
// ceil(nb_max / 4096), total one int.
hipError_t launch_heavy_order(const MatDev* mats, int64_t nblk_total, int nmats, const int64_t* blkq_base,
                              const uint32_t* blk_flag, int32_t nb_max, int32_t* bucket, int32_t* cursor,
                              int32_t* block_sums, int32_t* total, WorkState* ws, uint64_t* heavy_items, hipStream_t s);
// LDS of k_heavy outside its dynamic part (static shared state, 2.6 KB) plus margin.
constexpr size_t kHeavyLdsReserve = 3072;
// Rows of a block whose facts (and ELL rows) k_heavy reads into LDS in one round.
constexpr int kHeavyWin = 64;
// The window's ELL rows (plus one look-ahead row) kept in LDS per direction.
constexpr size_t heavy_stash_bytes(bool paired_any) {
  return (size_t)(kHeavyWin + 1) * kEll * 4 * (paired_any ? 2 : 1);
}
// Dynamic LDS of k_heavy: the bitmap (unless it lives in HBM), two frontier queue levels, the
// first ucap entries of the staged U list of each direction (two with paired items), and the
// window's rows.
size_t heavy_lds_bytes(int32_t bw_max, int32_t qcap, bool hbm_bitmap, int32_t ucap, bool paired_any);
hipError_t heavy_set_lds_limit(int threads, size_t bytes, bool hbm_bitmap);
// Saved bitmaps per direction in a workgroup region when some matrix runs paired items: a ring
// of the last rows saved (chain bases for later rows, early tests for the other direction).
constexpr int kSavedRing = 3;
// spill: ints per resident workgroup, (3 n_max + ceil(n_max / 32)), or with paired items
// (4 n_max + 2 kSavedRing ceil(n_max / 32)).
constexpr int64_t heavy_region_ints(int64_t n_max, bool paired_any) {
  const int64_t nbw = (n_max + 31) / 32;
  return paired_any ? 4 * n_max + 2 * kSavedRing * nbw : 3 * n_max + nbw;
}
// hbm_bitmaps: null (bitmaps in LDS), or grid * bw_max words, one bitmap per workgroup
// (1024-thread workgroups only).
// exec_slot: the accumulator slot of the launch's first executed-work counter (kAccHeavyExecEdges,
// or kAccTopExecEdges for the early launch).
hipError_t launch_heavy(int threads, int grid, const MatDev* mats, WorkState* ws, const uint64_t* heavy_items,
                        int32_t bw_max, int32_t qcap, int64_t n_max, int32_t paired_any, int32_t* spill,
                        uint32_t* hbm_bitmaps, int32_t ucap, int32_t exec_slot, hipStream_t s);
// Before the early launch: every matrix's rows [light_end, n) escalated (both directions), and
// ws->heavy_count = count, ws->heavy_next = 0 (one block per matrix).
hipError_t launch_top_setup(const MatDev* mats, int nmats, WorkState* ws, int64_t count, hipStream_t s);
// Mirrored matrices: L counts = U counts, fill and workgroup-path entries doubled, nnz_L = nnz_U
// (one thread per row, blocks as for launch_supernodes).
hipError_t launch_mirror(const MatDev* mats, int nmats, const int64_t* blk_base, int64_t nblocks, hipStream_t s);
hipError_t launch_supernodes(const MatDev* mats, int nmats, const int64_t* blk_base, int64_t nblocks, hipStream_t s);
constexpr int kLightThreads = 256;
constexpr int kLightWaves = 4;  // wavefronts per wave-path workgroup
// Consecutive rows per wave-path item: the largest power of two in [kLightBlockMin,
// kLightBlockMax] that still leaves about four items per resident wavefront (longer blocks chain
// more rows inside the wave path; measured on the 659,033-row stencil: 8 rows 2.52 ms, 16 rows
// 2.35 ms, 64 rows 2.23 ms).
constexpr int kLightBlockMin = 8, kLightBlockMax = 64;
// Reached vertices before the wave path escalates a row: 256 for matrices up to 32,768 rows (a
// bitmap-path row is cheap there), else 512 (the LDS sizing). Measured: 4,096 rows 1.43 -> 1.17 ms
// with 256; 110,592 rows 2.89 ms with 512, 3.02 with 256; 659,033 rows 8.4 / 9.4 ms.
constexpr int kLightLimitMax = 512;
constexpr int kLightLimitSmall = 256, kLightSmallRows = 32768;
// Wave path hash sets: 1,024 slots (five workgroups per CU), or, for sets whose largest matrix
// has more than kLightBigRows rows, 2,048 slots with the limit at 1,024 reached vertices (three
// workgroups per CU). Measured: the 659,033-row stencil 7.46 -> 6.78 ms per step with the
// structure (rows of 512-1,024 reached vertices chain cheaply on the wave path instead of taking
// a workgroup row each), the 110,592-row one 2.58 -> 2.75 ms counting only (occupancy).
constexpr int kLightBits = 10, kLightBitsBig = 11, kLightLimitBig = 1024, kLightBigRows = 262144;
}  // namespace gsofa

namespace gsofa {
// Resident workgroups of k_heavy per CU for a block size and dynamic LDS size.
int heavy_occupancy(int threads, size_t lds_bytes, bool hbm_bitmap);
hipError_t launch_copy16(const void* src, void* dst, int64_t n16, int grid, hipStream_t s);
int light_occupancy(int bits);
}  // namespace gsofa
