/*
 * gsofa.h - C ABI of the MI355X symbolic-factorisation engine (libgsofa.so).
 *
 * Plain C types only: int32 CSR in the reference's layout (row v is adj[beg[v]..end[v]),
 * rows may be non-contiguous and unsorted), caller-owned inputs, library-owned plans.
 * Every entry point returns GSOFA_OK (0) or a GSOFA_ERR_* code and never exits the process;
 * gsofa_last_error() gives the text of the calling thread's last failure.
 */
#ifndef GSOFA_H
#define GSOFA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define GSOFA_OK 0
#define GSOFA_ERR_INVALID_ARG 1   /* bad argument count / value                                */
#define GSOFA_ERR_IO 2            /* a file could not be opened or parsed                      */
#define GSOFA_ERR_BAD_CSR 3       /* beg > end, extent outside adj, or column id outside [0,n) */
#define GSOFA_ERR_HIP 4           /* a HIP runtime call failed                                 */
#define GSOFA_ERR_NO_DEVICE 5     /* no usable gfx950 device                                   */
#define GSOFA_ERR_OUT_OF_MEMORY 6 /* device or host allocation failed                          */
#define GSOFA_ERR_CAPACITY 7      /* a request exceeds a fixed capacity                        */
#define GSOFA_ERR_STATE 8         /* call out of order (e.g. result before run)                */
#define GSOFA_ERR_REJECTED 9      /* serving: request can never fit the device budget          */
#define GSOFA_ERR_CANCELLED 10    /* serving: request cancelled before completion              */
#define GSOFA_ERR_NOT_READY 11    /* serving: result not ready yet                             */

/* ---- the pattern --------------------------------------------------------------------------- */
typedef struct gsofa_csr {
  int32_t n;          /* rows = columns                                                     */
  int64_t nnz;        /* length of adj                                                      */
  const int32_t* beg; /* n: first entry of row v                                            */
  const int32_t* end; /* n: one past the last entry of row v                                */
  const int32_t* adj; /* nnz column ids                                                     */
} gsofa_csr;

const char* gsofa_last_error(void);

/* ---- host patterns: generators and the reference's ASCII files -------------------------- */
typedef struct gsofa_host_matrix gsofa_host_matrix;

/* generator kinds / orderings / drop sides (see gsofa_generate) */
#define GSOFA_GEN_STENCIL 0
#define GSOFA_GEN_ERDOS_RENYI 1
#define GSOFA_ORDER_NATURAL 0
#define GSOFA_ORDER_ND 1
#define GSOFA_ORDER_RCM 2
#define GSOFA_DROP_ANY 0
#define GSOFA_DROP_LOWER 1
#define GSOFA_DROP_UPPER 2

/* the workloads: member selects a batch member / stream request (ignored otherwise) */
#define GSOFA_WL_GRID2D_4096 1   /* 64x64 5-point, natural order, 20% drop                 */
#define GSOFA_WL_GRID3D_110592 2 /* 48^3 7-point, ND, 15% upper drop                        */
#define GSOFA_WL_GRID2D_659033 3 /* 812-wide 9-point truncated to 659,033 rows, ND, extras */
#define GSOFA_WL_BATCH256 4      /* member k of the 256-matrix batch                        */
#define GSOFA_WL_STREAM 5        /* request k of the served stream                          */

int gsofa_generate_workload(int workload, int64_t member, gsofa_host_matrix** out);

/* Free-form generator. params (int64 each): kind, dims, points, nx, ny, nz, n_keep, order,
 * nd_leaf, drop_side, drop_ppm, extra_every, avg_degree_x100. */
#define GSOFA_GEN_NPARAMS 13
int gsofa_generate(const int64_t* params, uint64_t seed, gsofa_host_matrix** out);

int gsofa_host_matrix_csr(const gsofa_host_matrix* m, gsofa_csr* view);
void gsofa_host_matrix_free(gsofa_host_matrix* m);

/* Reference ASCII triplet (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:2567-2616):
 * csr file = "E n" then E column ids; beg file = n ints; end file = n ints. Rows may be
 * non-contiguous and unsorted. Unlike the reference, failures return codes instead of exit(1). */
int gsofa_read_ascii(const char* beg_file, const char* end_file, const char* csr_file, gsofa_host_matrix** out);
int gsofa_write_ascii(const gsofa_csr* a, const char* beg_file, const char* end_file, const char* csr_file);

/* Reference binary CSR (gSoFa_SOAComparison/graph.hpp:4-104, a template over its file integer
 * types): beg file = n+1 row offsets of index_bytes each, csr file = the column ids of rows
 * 0..n-1 back to back, vert_bytes each, native byte order; n = beg file size / index_bytes - 1,
 * the edge count is the last offset. Widths are 4 or 8. The reference's optional weight file is
 * not read (the symbolic path has no values). Errors return codes (the reference asserts). */
int gsofa_read_binary(const char* beg_file, const char* csr_file, int32_t index_bytes, int32_t vert_bytes,
                      gsofa_host_matrix** out);
int gsofa_write_binary(const gsofa_csr* a, const char* beg_file, const char* csr_file, int32_t index_bytes,
                       int32_t vert_bytes);

/* The reference's inputs are SuperLU_DIST's reordered matrices (README.md:27-30): SuperLU_DIST
 * factors Pc * Pr * A * Pc^T. gsofa_permute builds that pattern, B(perm_c[perm_r[i]], perm_c[j]) =
 * A(i, j); perm_r or perm_c may be null (identity) and must otherwise be permutations of [0, n).
 * gsofa_read_matrix_market reads a square coordinate Matrix Market file (pattern, real, integer or
 * complex; general, or symmetric / skew-symmetric / Hermitian expanded to both triangles; values
 * ignored). gsofa_read_permutation reads n integers (base 0 or 1) and checks the permutation.
 * The program gsofa_convert turns a .mtx file and permutation files into the ASCII triplet or
 * binary files. No counterpart in the reference's code (its dump step lives in SuperLU_DIST). */
int gsofa_read_matrix_market(const char* path, gsofa_host_matrix** out);
int gsofa_read_permutation(const char* path, int32_t n, int32_t base, int32_t* perm);
int gsofa_permute(const gsofa_csr* a, const int32_t* perm_r, const int32_t* perm_c, gsofa_host_matrix** out);

/* ---- the drop-in entry ---------------------------------------------------------------------
 * Replaces void symbfact_min_id(int args, char** argv, int myrank, ull_t& fill_count,
 * double& time, int& N_supernodes, ull_t& Process_TEPS)
 * (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:2529; caller
 * gSoFa_SOAComparison/main.cu:41). Same argv contract (argc == 11): [1] beg file, [2] end file,
 * [3] csr file, [4] num_process, [5] max_supernode_size (chunk S), [6] percent_cat0,
 * [7] percent_cat2, [8] N_blocks_source_cat2, [9] N_GPU_Node, [10] N_src_group.
 * All num_process x N_GPU_Node shares run on one device (device = rank % N_GPU_Node).
 * Outputs as in the reference: *fill_count += fill-ins, *time_ms += milliseconds from the first
 * kernel to the end of supernode detection, *n_supernodes = supernodes counted + num_process
 * (each reference process adds one, ...cuh:3549, summed by main.cu:44), *edge_checks =
 * schedule-independent edge checks W. */
int gsofa_symbfact_min_id(int argc, char** argv, int rank, uint64_t* fill_count, double* time_ms,
                          int64_t* n_supernodes, uint64_t* edge_checks);

/* Device facts the engine sizes itself by: compute units, LDS bytes per workgroup and per CU,
 * and the architecture name (e.g. "gfx950:sramecc+:xnack-"). No counterpart in the reference. */
int gsofa_device_query(int device, int32_t* cus, int64_t* lds_per_block, int64_t* lds_per_cu, char* arch,
                       int32_t arch_len);

/* Stream and synchronisation helpers for callers without a HIP runtime of their own (a process
 * must not load a second HIP runtime, e.g. a framework's bundled one, next to this library's).
 * No counterpart in the reference. */
int gsofa_device_synchronize(int device);
int gsofa_stream_create(int device, void** hip_stream);
int gsofa_stream_destroy(void* hip_stream);

/* Streaming device copy of `bytes` (read + write counted), `iters` times: the achievable HBM
 * bandwidth, and the reference kernel ("k_copy16") that calibrates byte counters in profiles.
 * No counterpart in the reference. */
int gsofa_bandwidth_probe(int device, int64_t bytes, int32_t iters, double* gb_per_s);

/* ---- structured API --------------------------------------------------------------------- */
typedef struct gsofa_options {
  int32_t chunk_size;       /* supernode chunk S (argv[5]); default 128                      */
  int32_t n_processes;      /* argv[4]: the reported supernode count adds one per process      */
  int32_t structure;        /* 1: also produce strict L (by columns) and strict U (by rows)    */
  int32_t device;           /* HIP device ordinal                                               */
  int32_t heavy_threads;    /* tuning: workgroup size of the bitmap path, 0 = automatic         */
  int32_t heavy_wgs_per_cu; /* tuning: resident bitmap workgroups per CU, 0 = automatic         */
  int32_t light_wgs_per_cu; /* tuning: resident hash-path workgroups per CU, 0 = automatic      */
  int32_t queue_capacity;   /* tuning: LDS frontier entries per level buffer, 0 = automatic     */
  int32_t chain_block;      /* tuning: rows per workgroup-path item (1 = no chaining), 0 = auto */
  int32_t pair_directions;  /* workgroup-path items run both directions of a block and learn
                               s-1 in R_s from the other direction: 1 on, 0 off, -1 (default)
                               on where fewer than a quarter of the rows hold their predecessor */
  int32_t autotune;         /* 1: at creation, time the bitmap-path workgroup shapes and larger
                               row blocks on this very set and keep the fastest (results are
                               identical under every candidate); 0 (default): built-in rules */
  int32_t bitmap_in_hbm;    /* 1: the bitmap path keeps its n-bit bitmaps in HBM (L2 atomics);
                               0 (default): only where one does not fit in LDS (n above about
                               1.2 M rows), so any n runs. The reference's labels live in global
                               memory at every n (SOA:470-479).                                */
  int32_t split_blocks;     /* a workgroup-path block that holds the start of a separator over
                               a large subtree (its elimination subtree at least 1.5x its
                               predecessor's and n / k rows) runs as two items, so two
                               traversals of whole subtrees do not queue in one workgroup:
                               0 (default) k = 64 for matrices of at least a quarter of the
                               set's rows (a plan's matrix; not the members of a set of many,
                               where the extra traversals cost more than they shorten), k > 0
                               that k for every matrix, -1 never                               */
  int32_t symmetric;        /* 0 (default): a structurally symmetric pattern (after duplicate
                               removal, every off-diagonal (i, j) has its (j, i)) is traversed
                               once, from A: the traversal from source j over A^T then reaches
                               exactly what the one over A reaches, so L's column j is U's row j
                               and the L counts, lists and totals are copied from U's.
                               -1: traverse A and A^T for every pattern                        */
  int32_t cu_share;         /* 0 (default): persistent grids sized for every CU; k > 0: for k
                               CUs, so that work the caller runs beside it on other CUs (a
                               stream with a CU mask) keeps its share of the device            */
  int32_t early_top;        /* 1: the rows from the start of a large matrix's top separator (a
                               block cut whose subtree holds half the rows; matrices of at least
                               a quarter of the set's rows, over 262,144 rows, bitmaps in LDS)
                               run as a few fixed workgroup-path items, three per direction over
                               consecutive blocks, in a launch of their own on an internal
                               stream, which starts with the wave path and joins the run before
                               supernode detection (counting only, about 3 % faster at 659,033
                               rows when it is dispatched before the wave path fills the device,
                               about 10 % slower when not: nothing orders the two). 0 (default)
                               or -1: not. Results do not depend on it                         */
  int32_t pre_escalate;     /* 0 (default): in matrices without paired items whose pattern is
                               symmetric or nearly (three quarters of the sampled off-diagonal
                               entries have their transpose), a row whose elimination subtree of
                               A + A^T holds more vertices than the hash path's escalation limit
                               goes to the bitmap path without a hash-path attempt (a symmetric
                               pattern's reached set holds the whole subtree; a nearly symmetric
                               one's nearly always does). -1:
                               every row is attempted first. The server does not use it (the
                               elimination tree costs as much host time as the preparation).
                               Results do not depend on it                                     */
} gsofa_options;

typedef struct gsofa_result {
  uint64_t fill_count;        /* off-diagonal fill-ins: |F \ A| ("Number of fill-ins detected") */
  uint64_t nnz_F;             /* |F| = |L+U| including the diagonal                           */
  uint64_t nnz_L, nnz_U;      /* strict lower / strict upper entries of F                     */
  uint64_t edge_checks;       /* W = sum_i deg(i) + sum_{v in R_i} deg(v)                      */
  uint64_t reach;             /* sum_i |R_i|                                                  */
  int64_t supernodes;         /* reported count: counted + n_processes                       */
  int64_t supernodes_counted; /* supernodes counted by the reference rule                     */
  uint64_t heavy_sources;     /* traversals that took the workgroup (bitmap) path; may vary
                                 between runs (escalation can follow other rows' outcomes)     */
  double time_ms;             /* device time of the last run                                   */
} gsofa_result;

typedef struct gsofa_plan gsofa_plan;

void gsofa_options_default(gsofa_options* opt);
/* Validates and copies the pattern (row sort + de-duplication, the reference's SortEdgeList),
 * allocates device state and uploads it. The caller keeps ownership of a. */
int gsofa_plan_create(const gsofa_csr* a, const gsofa_options* opt, gsofa_plan** out);
/* Enqueues one full symbolic factorisation on `hip_stream` (a hipStream_t, 0 = default
 * stream). With the structure option the first run also performs a sizing pass. */
int gsofa_run(gsofa_plan* plan, void* hip_stream);
int gsofa_sync(gsofa_plan* plan);
int gsofa_result_get(gsofa_plan* plan, gsofa_result* out);
/* nzU[i] as the reference counts it (U entries of row i, duplicates of A included); n ints. */
int gsofa_nzu_get(gsofa_plan* plan, int32_t* nzu);
/* flags[i] = 1 where the supernode count increments (a start or a failed pass-through); n bytes. */
int gsofa_supernodes_get(gsofa_plan* plan, uint8_t* flags);
/* Sizes of the strict parts; requires the structure option and a completed run. */
int gsofa_structure_size(gsofa_plan* plan, int64_t* nnz_u, int64_t* nnz_l);
/* Strict U by rows (u_rowptr n+1, u_col nnz_u) and strict L by columns (l_colptr n+1,
 * l_row nnz_l), each row/column sorted ascending. */
int gsofa_structure_get(gsofa_plan* plan, int64_t* u_rowptr, int32_t* u_col, int64_t* l_colptr, int32_t* l_row);
/* ---- supernodal export ------------------------------------------------------------------
 * The supernode partition the reference computes in struct Super {start, end}
 * (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:341-345, written at :1970-1971
 * and :2039, never returned): supernode k is the row range [xsup[k], xsup[k+1]), starting at a row
 * where the supernode count increments (flag 1) and ending before the next one; xsup[nsuper] = n.
 * The count equals gsofa_result.supernodes_counted. xsup (nsuper + 1 entries) and supno (n
 * entries: the supernode of each row) may be null; call once with both null to get nsuper. */
int gsofa_supernode_partition(gsofa_plan* plan, int64_t* nsuper, int32_t* xsup, int32_t* supno);
/* Per-supernode panel structure for a supernodal numeric factorisation (requires the structure
 * option and a completed run). For supernode k = [a, b): lsub[xlsub[k] .. xlsub[k+1]) = sorted
 * rows i >= b with (i, j) in F for some column j in [a, b) (the L panel below the diagonal
 * block); usub[xusub[k] .. xusub[k+1]) = sorted columns m >= b with (i, m) in F for some row i
 * in [a, b) (the U panel). The diagonal block is dense in that format and not listed. Sizes first
 * (nsuper, lsub and usub lengths), then the arrays (xsup nsuper + 1, xlsub and xusub nsuper + 1).
 * Built on the host from the structure. No counterpart in the reference beyond Super. */
int gsofa_supernodal_structure_size(gsofa_plan* plan, int64_t* nsuper, int64_t* nnz_lsub, int64_t* nnz_usub);
int gsofa_supernodal_structure_get(gsofa_plan* plan, int32_t* xsup, int64_t* xlsub, int32_t* lsub, int64_t* xusub,
                                   int32_t* usub);
/* The same from caller arrays (no device): supernode flags (n bytes) and the strict structure in
 * gsofa_structure_get's form. With every output array null only the sizes are written. */
int gsofa_supernodal_from_structure(int32_t n, const uint8_t* flags, const int64_t* u_rowptr, const int32_t* u_col,
                                    const int64_t* l_colptr, const int32_t* l_row, int64_t* nsuper, int64_t* nnz_lsub,
                                    int64_t* nnz_usub, int32_t* xsup, int64_t* xlsub, int32_t* lsub, int64_t* xusub,
                                    int32_t* usub);
/* Device time of the last run's stages, from events on the run's stream: ms[0] the run's clears
 * (the pattern's device forms, ELL rows and the transpose, are built once when the plan or batch
 * is created), ms[1] the wave (hash) path and the workgroup-path queue, ms[2] the workgroup
 * (bitmap) path's main launch, ms[3] the wait for the early launch (if it is still running),
 * supernode detection. No counterpart in the reference (its only timers are wtime()). */
int gsofa_stage_times(gsofa_plan* plan, double* ms);
/* Whether gsofa_run records the stage events above (on = 1, the default) or not (0: each event
 * between kernels costs a few microseconds of device time; gsofa_stage_times then reports
 * GSOFA_ERR_STATE and gsofa_early_top a time of 0). No counterpart in the reference. */
int gsofa_plan_stage_timing(gsofa_plan* plan, int32_t on);
/* The early launch of the top separator's items (gsofa_options.early_top) in the last run: its
 * item count (0: none), its device time (events on its own stream, which starts with the wave
 * path) and its executed work3 = {neighbour inspections, rows expanded, F entries written}
 * (included in gsofa_executed_work's workgroup-path figures). No counterpart in the reference. */
int gsofa_early_top(gsofa_plan* plan, int64_t* items, double* ms, uint64_t* work3);
/* The workgroup path's share of the last run: row-traversal edge checks, |R_i| sum, and strict
 * F entries (both directions). Used to price that kernel alone. */
int gsofa_work_split(gsofa_plan* plan, uint64_t* heavy_edge_checks, uint64_t* heavy_reach, uint64_t* heavy_entries);
/* Work the last run actually executed, both directions: work6 = {workgroup path: neighbour
 * inspections, rows expanded, F entries written; wave path: the same three}. Chaining and the
 * two directions make these differ from W and |R|; the roofline of each kernel is priced on
 * them. No counterpart in the reference. */
int gsofa_executed_work(gsofa_plan* plan, uint64_t* work6);
void gsofa_plan_destroy(gsofa_plan* plan);

/* ---- batched form: many matrices in one captured graph of persistent launches ----------- */
typedef struct gsofa_batch gsofa_batch;
/* The launch configuration in use: out4 = {heavy_threads, bitmap workgroups per CU,
 * chain_block, queue_capacity}. No counterpart in the reference. */
int gsofa_plan_config(gsofa_plan* plan, int32_t* out4);
int gsofa_batch_config(gsofa_batch* batch, int32_t* out4);
int gsofa_batch_create(const gsofa_csr* mats, int32_t count, const gsofa_options* opt, gsofa_batch** out);
/* Replays the batch's hipGraph on `hip_stream` (captured on the first call). */
int gsofa_batch_run(gsofa_batch* batch, void* hip_stream);
int gsofa_batch_sync(gsofa_batch* batch);
int gsofa_batch_result_get(gsofa_batch* batch, int32_t index, gsofa_result* out);
int gsofa_batch_nzu_get(gsofa_batch* batch, int32_t index, int32_t* nzu);
int gsofa_batch_supernodes_get(gsofa_batch* batch, int32_t index, uint8_t* flags);
/* gsofa_supernode_partition for batch member `index`. */
int gsofa_batch_supernode_partition(gsofa_batch* batch, int32_t index, int64_t* nsuper, int32_t* xsup, int32_t* supno);
void gsofa_batch_destroy(gsofa_batch* batch);

/* ---- serving: continuous batching of requests ------------------------------------------
 * Requests are queued FIFO and admitted onto one of `slots` device streams when a slot is free
 * and their device footprint fits the HBM budget; admitted requests run concurrently on the
 * device. A request that can never fit is rejected (GSOFA_ERR_REJECTED from
 * poll/wait). Results are the same as a plan's on the same pattern. No counterpart in the
 * reference. */
typedef struct gsofa_server gsofa_server;
typedef struct gsofa_server_options {
  int32_t device;
  int32_t slots;            /* concurrent requests on the device, default 4                      */
  int64_t hbm_budget_bytes; /* admission budget; 0 = 80% of the free device memory at start      */
  int32_t cu_share;         /* CUs each group's persistent grids are sized for; 0 = all          */
  int32_t max_group;        /* requests a free slot takes from the queue at once (same options,
                               FIFO order) and runs as one batch; default 16, 1 = one at a time */
} gsofa_server_options;
void gsofa_server_options_default(gsofa_server_options* opt);
int gsofa_server_start(const gsofa_server_options* opt, gsofa_server** out);
/* Copies and prepares the pattern on the calling thread, then queues it. opt: chunk_size,
 * n_processes and the tuning fields (heavy_threads, heavy_wgs_per_cu, light_wgs_per_cu,
 * queue_capacity, chain_block, pair_directions, bitmap_in_hbm) are used; requests with different
 * options are never grouped. structure = 1 and autotune = 1 are plan/batch features and return
 * GSOFA_ERR_INVALID_ARG. */
int gsofa_server_submit(gsofa_server* srv, const gsofa_csr* a, const gsofa_options* opt, int64_t* request_id);
/* GSOFA_OK with the result, or GSOFA_ERR_NOT_READY / _CANCELLED / _REJECTED / another error. */
int gsofa_server_poll(gsofa_server* srv, int64_t request_id, gsofa_result* out);
/* Blocks up to timeout_s seconds (< 0: no limit). */
int gsofa_server_wait(gsofa_server* srv, int64_t request_id, double timeout_s, gsofa_result* out);
/* Pending requests are dropped; running ones finish on the device and report GSOFA_ERR_CANCELLED. */
int gsofa_server_cancel(gsofa_server* srv, int64_t request_id);
/* Forgets a finished, failed, cancelled or rejected request (its id becomes unknown); a queued or
 * running one returns GSOFA_ERR_STATE. A finished request keeps only its totals and digests (its
 * pattern is released when it leaves the device); a long-running server releases ids it has
 * read to keep its host memory bounded. */
int gsofa_server_release(gsofa_server* srv, int64_t request_id);
/* Requests the server still holds (any state) and the host bytes they keep alive. */
int gsofa_server_retained(gsofa_server* srv, int64_t* requests, int64_t* host_bytes);
/* Submit-to-result latency, time queued before admission, and FNV-1a digests of nzU and the
 * supernode flags of a finished request. */
int gsofa_server_request_info(gsofa_server* srv, int64_t request_id, double* latency_ms, double* queue_ms,
                              uint64_t* nzu_digest, uint64_t* supernode_digest);
/* The FNV-1a digest (64-bit) gsofa_server_request_info reports, of any byte range. */
int gsofa_digest_fnv1a(const void* data, int64_t bytes, uint64_t* out);
int gsofa_server_stats(gsofa_server* srv, int64_t* completed, int32_t* running, int32_t* peak_running,
                       int64_t* pending, int64_t* hbm_in_use, int64_t* hbm_budget);
/* Drops pending requests, waits for running ones, releases everything. */
int gsofa_server_stop(gsofa_server* srv);

/* ---- allocation-free run ------------------------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The counts-only run of
 * gsofa_plan_create + gsofa_run + gsofa_result_get / gsofa_nzu_get / gsofa_supernodes_get, enqueued
 * on `hip_stream` with no allocation: every buffer beyond the inputs and outputs lives in one
 * static 128 MiB device array per GPU, held until the process ends. Limits: n <= 65,536 rows and
 * nnz <= 1,048,576 entries (GSOFA_ERR_CAPACITY beyond them); structure, autotune and
 * bitmap_in_hbm = 1 are rejected; pair_directions -1 means off. All pointers are device addresses:
 * rowptr (n + 1) and adj (nnz) with rowptr[0] = 0, rowptr[n] = nnz and rows of strictly increasing
 * column ids in [0, n); outputs nzu (n), sup_flags (n) and totals (GSOFA_NOALLOC_TOTALS: fill_count,
 * nnz_F, nnz_L, nnz_U, edge_checks, reach, supernodes, supernodes_counted, input_error). A pattern
 * that breaks the row rule is checked on the device: input_error is 1 and the counts are those of
 * an empty pattern. Calls on one GPU share its array: calls that overlap on one GPU (two streams,
 * two threads) race. */
#define GSOFA_NOALLOC_TOTALS 9
int gsofa_run_noalloc(int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* adj, const gsofa_options* opt,
                      int32_t* nzu, uint8_t* sup_flags, uint64_t* totals, void* hip_stream);

#ifdef __cplusplus
}
#endif
#endif /* GSOFA_H */
