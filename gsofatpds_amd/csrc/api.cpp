// C ABI of the engine: plans, batches, the reference's ASCII files and the drop-in entry.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>

#include "engine.h"

using namespace gsofa;

namespace {

Options to_options(const gsofa_options* o) {
  Options r;
  if (!o) return r;
  r.chunk = o->chunk_size;
  r.n_processes = o->n_processes;
  r.structure = o->structure;
  r.device = o->device;
  r.heavy_threads = o->heavy_threads;
  r.heavy_wgs_per_cu = o->heavy_wgs_per_cu;
  r.light_wgs_per_cu = o->light_wgs_per_cu;
  r.qcap = o->queue_capacity;
  r.chain_block = o->chain_block > 0 ? o->chain_block : 0;  // 0: chosen from n and the grid
  r.pair_directions = o->pair_directions < 0 ? -1 : (o->pair_directions ? 1 : 0);
  r.autotune = o->autotune ? 1 : 0;
  r.bitmap_in_hbm = o->bitmap_in_hbm ? 1 : 0;
  r.split_blocks = o->split_blocks < 0 ? -1 : o->split_blocks;
  r.symmetric = o->symmetric < 0 ? -1 : 0;
  r.cu_share = o->cu_share > 0 ? o->cu_share : 0;
  r.early_top = o->early_top < 0 ? -1 : (o->early_top > 0 ? 1 : 0);
  r.pre_escalate = o->pre_escalate < 0 ? -1 : 0;
  return r;
}

void fill_result(const MatResult& r, const Options& opt, int32_t n, double ms, gsofa_result* out) {
  out->fill_count = r.fill;
  out->nnz_L = r.nnz_l;
  out->nnz_U = r.nnz_u;
  out->nnz_F = r.nnz_l + r.nnz_u + (uint64_t)n;
  out->edge_checks = r.W;
  out->reach = r.R;
  out->supernodes_counted = r.nsup;
  out->supernodes = r.nsup + opt.n_processes;
  out->heavy_sources = r.heavy;
  out->time_ms = ms;
}

// A candidate configuration on its own device set: a warm-up run, then the faster of two timed
// runs (one run alone let noise pick the block size).
int time_candidate(const std::vector<const Prepared*>& mats, const Options& o, double* ms, Options* chosen) {
  std::unique_ptr<DeviceSet> set(new (std::nothrow) DeviceSet);
  if (!set) return fail(GSOFA_ERR_OUT_OF_MEMORY, "device set");
  int rc = set->create(mats, o);
  if (rc != GSOFA_OK) return rc;
  *chosen = set->options();
  chosen->heavy_threads = set->heavy_threads();
  chosen->heavy_wgs_per_cu = set->heavy_wgs_per_cu();
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = fail(GSOFA_ERR_HIP, "autotune stream/events");
  if (rc == GSOFA_OK) rc = set->enqueue(s);
  float t = 0.f;
  for (int rep = 0; rep < 2 && rc == GSOFA_OK; ++rep) {
    if (hipEventRecord(e0, s) != hipSuccess) rc = fail(GSOFA_ERR_HIP, "hipEventRecord");
    if (rc == GSOFA_OK) rc = set->enqueue(s);
    if (rc == GSOFA_OK && hipEventRecord(e1, s) != hipSuccess) rc = fail(GSOFA_ERR_HIP, "hipEventRecord");
    float tr = 0.f;
    if (rc == GSOFA_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&tr, e0, e1) != hipSuccess))
      rc = fail(GSOFA_ERR_HIP, "autotune timing");
    if (rc == GSOFA_OK) t = rep == 0 ? tr : std::min(t, tr);
  }
  if (s) (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  *ms = t;
  return rc;
}

// The autotuner: bitmap-path workgroup shapes (threads x resident per CU: rows in flight against
// frontier width per pass and LDS queue depth) at the automatic row-block size, then 2x, 4x, ...
// larger blocks for the winner (fewer chain starts, coarser balance) while they keep helping,
// up to 1024 rows. Candidates that do not fit the device are skipped. Fields the caller set
// explicitly are kept. The structure is not produced while timing.
int autotune(const std::vector<const Prepared*>& mats, const Options& base, Options& best) {
  Options b = base;
  b.structure = 0;
  double best_ms = 1e300;
  Options best_o = b;
  bool any = false;
  const int shapes[4][2] = {{256, 4}, {256, 3}, {512, 2}, {1024, 1}};
  for (const auto& sh : shapes) {
    if (base.heavy_threads && base.heavy_threads != sh[0]) continue;
    if (base.heavy_wgs_per_cu && sh[0] == 256 && sh[1] == 3) continue;  // same candidate twice
    Options o = b;
    o.heavy_threads = sh[0];
    if (!base.heavy_wgs_per_cu) o.heavy_wgs_per_cu = sh[1];
    double ms = 0.0;
    Options chosen;
    const int rc = time_candidate(mats, o, &ms, &chosen);
    if (rc == GSOFA_ERR_CAPACITY) continue;
    if (rc != GSOFA_OK) return rc;
    if (ms < best_ms) {
      best_ms = ms;
      best_o = chosen;
      any = true;
    }
  }
  if (!any) return fail(GSOFA_ERR_CAPACITY, "autotune: no candidate fits the device");
  if (base.chain_block == 0) {
    int32_t k = best_o.chain_block;
    while (k < 1024) {
      Options o = best_o;
      o.chain_block = std::min<int32_t>(1024, 2 * k);
      double ms = 0.0;
      Options chosen;
      const int rc = time_candidate(mats, o, &ms, &chosen);
      if (rc != GSOFA_OK && rc != GSOFA_ERR_CAPACITY) return rc;
      if (rc != GSOFA_OK || ms >= 0.98 * best_ms) break;  // a larger block must win clearly
      best_ms = ms;
      best_o = chosen;
      k = o.chain_block;
    }
  }
  best = best_o;
  best.structure = base.structure;
  best.autotune = 0;
  return GSOFA_OK;
}

// Events and stream bookkeeping shared by plans and batches.
struct Timing {
  hipEvent_t start = nullptr, stop = nullptr;
  hipStream_t stream = nullptr;
  bool pending = false;
  double last_ms = 0.0;
  int init() {
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess)
      return fail(GSOFA_ERR_HIP, "hipEventCreate");
    return GSOFA_OK;
  }
  ~Timing() {
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
  }
  int finish() {
    if (!pending) return GSOFA_OK;
    hipError_t e = hipEventSynchronize(stop);
    if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, start, stop) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipEventElapsedTime");
    last_ms = ms;
    pending = false;
    return GSOFA_OK;
  }
};

}  // namespace

struct gsofa_plan {
  Prepared prep;
  DeviceSet set;
  Timing timing;
  hipEvent_t stage[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  ~gsofa_plan() {
    for (hipEvent_t e : stage)
      if (e) (void)hipEventDestroy(e);
  }
  std::vector<MatResult> results;
  bool have_result = false;
  bool ran = false;  // a run was enqueued: results and timings exist only after one
  bool stage_timing = true;  // record the stage events in gsofa_run
  bool staged = false;       // the last run recorded them
  // host copy of the last run's structure and flags for the supernodal export (until the next run)
  bool sn_cached = false;
  std::vector<int64_t> sn_up, sn_lp;
  std::vector<int32_t> sn_ui, sn_li;
  std::vector<uint8_t> sn_flags;
};

struct gsofa_batch {
  std::vector<Prepared> preps;
  DeviceSet set;
  Timing timing;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<MatResult> results;
  bool have_result = false;
  bool ran = false;
  ~gsofa_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
};

extern "C" {

int gsofa_device_query(int device, int32_t* cus, int64_t* lds_per_block, int64_t* lds_per_cu, char* arch,
                       int32_t arch_len) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(GSOFA_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= count) return fail(GSOFA_ERR_NO_DEVICE, "device ordinal out of range");
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipGetDeviceProperties");
  if (cus) *cus = p.multiProcessorCount;
  if (lds_per_block) *lds_per_block = (int64_t)p.sharedMemPerBlock;
  if (lds_per_cu) *lds_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
  if (arch && arch_len > 0) {
    std::strncpy(arch, p.gcnArchName, (size_t)arch_len - 1);
    arch[arch_len - 1] = '\0';
  }
  return GSOFA_OK;
}

int gsofa_device_synchronize(int device) {
  if (hipSetDevice(device) != hipSuccess) return fail(GSOFA_ERR_NO_DEVICE, "hipSetDevice");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
  return GSOFA_OK;
}

int gsofa_stream_create(int device, void** out) {
  if (!out) return fail(GSOFA_ERR_INVALID_ARG, "out is null");
  if (hipSetDevice(device) != hipSuccess) return fail(GSOFA_ERR_NO_DEVICE, "hipSetDevice");
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  *out = s;
  return GSOFA_OK;
}

int gsofa_stream_destroy(void* s) {
  if (!s) return GSOFA_OK;
  hipError_t e = hipStreamDestroy(static_cast<hipStream_t>(s));
  if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipStreamDestroy: ") + hipGetErrorString(e));
  return GSOFA_OK;
}

int gsofa_bandwidth_probe(int device, int64_t bytes, int32_t iters, double* gb_per_s) {
  if (bytes < (1 << 20) || iters <= 0 || !gb_per_s) return fail(GSOFA_ERR_INVALID_ARG, "bad probe arguments");
  if (hipSetDevice(device) != hipSuccess) return fail(GSOFA_ERR_NO_DEVICE, "hipSetDevice");
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipGetDeviceProperties");
  const int64_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, (size_t)n16 * 16) != hipSuccess) return fail(GSOFA_ERR_OUT_OF_MEMORY, "probe buffer");
  if (hipMalloc(&b, (size_t)n16 * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(GSOFA_ERR_OUT_OF_MEMORY, "probe buffer");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int grid = p.multiProcessorCount * 8;
  int rc = GSOFA_OK;
  if (gsofa::launch_copy16(a, b, n16, grid, nullptr) != hipSuccess) rc = fail(GSOFA_ERR_HIP, "copy kernel");
  (void)hipEventRecord(e0, nullptr);
  for (int i = 0; i < iters && rc == GSOFA_OK; ++i)
    if (gsofa::launch_copy16(a, b, n16, grid, nullptr) != hipSuccess) rc = fail(GSOFA_ERR_HIP, "copy kernel");
  (void)hipEventRecord(e1, nullptr);
  float ms = 0.f;
  if (rc == GSOFA_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(GSOFA_ERR_HIP, "probe timing");
  if (rc == GSOFA_OK) *gb_per_s = 2.0 * (double)n16 * 16.0 * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

void gsofa_options_default(gsofa_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->chunk_size = 128;
  o->n_processes = 1;
  o->pair_directions = -1;
}

int gsofa_plan_create(const gsofa_csr* a, const gsofa_options* opt, gsofa_plan** out) {
  if (!out) return fail(GSOFA_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  gsofa_options def;
  gsofa_options_default(&def);
  Options o = to_options(opt ? opt : &def);
  std::unique_ptr<gsofa_plan> p(new (std::nothrow) gsofa_plan);
  if (!p) return fail(GSOFA_ERR_OUT_OF_MEMORY, "plan");
  int rc = prepare(a, p->prep);
  if (rc != GSOFA_OK) return rc;
  if (o.autotune) {
    Options tuned;
    rc = autotune({&p->prep}, o, tuned);
    if (rc != GSOFA_OK) return rc;
    o = tuned;
  }
  rc = p->set.create({&p->prep}, o);
  if (rc != GSOFA_OK) return rc;
  rc = p->timing.init();
  if (rc != GSOFA_OK) return rc;
  for (hipEvent_t& e : p->stage)
    if (hipEventCreate(&e) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipEventCreate");
  *out = p.release();
  return GSOFA_OK;
}

int gsofa_run(gsofa_plan* p, void* hip_stream) {
  if (!p) return fail(GSOFA_ERR_INVALID_ARG, "plan is null");
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  if (p->set.options().structure && !p->set.structure_ready()) {
    // sizing pass: counts give the exact row / column offsets of the structure
    int rc = p->set.enqueue(s);
    if (rc != GSOFA_OK) return rc;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("sizing pass: ") + hipGetErrorString(e));
    rc = p->set.size_pools_from_counts();
    if (rc != GSOFA_OK) return rc;
  }
  if (hipEventRecord(p->timing.start, s) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipEventRecord");
  int rc = p->set.enqueue(s, p->stage_timing ? p->stage : nullptr);
  if (rc != GSOFA_OK) return rc;
  p->staged = p->stage_timing;
  if (hipEventRecord(p->timing.stop, s) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipEventRecord");
  p->timing.stream = s;
  p->timing.pending = true;
  p->have_result = false;
  p->ran = true;
  p->sn_cached = false;
  return GSOFA_OK;
}

int gsofa_sync(gsofa_plan* p) {
  if (!p) return fail(GSOFA_ERR_INVALID_ARG, "plan is null");
  if (!p->ran) return fail(GSOFA_ERR_STATE, "no run was enqueued on this plan (call gsofa_run first)");
  int rc = p->timing.finish();
  if (rc != GSOFA_OK) return rc;
  rc = p->set.read_results(p->results);
  if (rc != GSOFA_OK) return rc;
  p->have_result = true;
  return GSOFA_OK;
}

int gsofa_result_get(gsofa_plan* p, gsofa_result* out) {
  if (!p || !out) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  fill_result(p->results[0], p->set.options(), p->prep.n, p->timing.last_ms, out);
  return GSOFA_OK;
}

int gsofa_nzu_get(gsofa_plan* p, int32_t* nzu) {
  if (!p || !nzu) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  return p->set.read_counts(0, nullptr, nullptr, nullptr, nzu);
}

int gsofa_supernodes_get(gsofa_plan* p, uint8_t* flags) {
  if (!p || !flags) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  return p->set.read_counts(0, nullptr, nullptr, flags, nullptr);
}

int gsofa_structure_size(gsofa_plan* p, int64_t* nnz_u, int64_t* nnz_l) {
  if (!p || !nnz_u || !nnz_l) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->set.structure_ready()) return fail(GSOFA_ERR_STATE, "structure not produced: enable it and run first");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  *nnz_u = (int64_t)p->results[0].nnz_u;
  *nnz_l = (int64_t)p->results[0].nnz_l;
  return GSOFA_OK;
}

int gsofa_structure_get(gsofa_plan* p, int64_t* u_rowptr, int32_t* u_col, int64_t* l_colptr, int32_t* l_row) {
  if (!p || !u_rowptr || !l_colptr) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  std::vector<int64_t> up, lp;
  std::vector<int32_t> ui, li;
  int rc = p->set.read_structure(0, up, ui, lp, li);
  if (rc != GSOFA_OK) return rc;
  std::memcpy(u_rowptr, up.data(), up.size() * 8);
  std::memcpy(l_colptr, lp.data(), lp.size() * 8);
  if (!ui.empty()) {
    if (!u_col) return fail(GSOFA_ERR_INVALID_ARG, "u_col is null");
    std::memcpy(u_col, ui.data(), ui.size() * 4);
  }
  if (!li.empty()) {
    if (!l_row) return fail(GSOFA_ERR_INVALID_ARG, "l_row is null");
    std::memcpy(l_row, li.data(), li.size() * 4);
  }
  return GSOFA_OK;
}

int gsofa_plan_stage_timing(gsofa_plan* p, int32_t on) {
  if (!p) return fail(GSOFA_ERR_INVALID_ARG, "plan is null");
  p->stage_timing = on != 0;
  return GSOFA_OK;
}

int gsofa_stage_times(gsofa_plan* p, double* ms) {
  if (!p || !ms) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  if (!p->staged) return fail(GSOFA_ERR_STATE, "the last run recorded no stage times (gsofa_plan_stage_timing)");
  for (int k = 0; k < 4; ++k) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, p->stage[k], p->stage[k + 1]) != hipSuccess)
      return fail(GSOFA_ERR_HIP, "hipEventElapsedTime");
    ms[k] = t;
  }
  return GSOFA_OK;
}

int gsofa_work_split(gsofa_plan* p, uint64_t* heavy_edge_checks, uint64_t* heavy_reach, uint64_t* heavy_entries) {
  if (!p || !heavy_edge_checks || !heavy_reach || !heavy_entries) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  *heavy_edge_checks = p->results[0].heavy_w;
  *heavy_reach = p->results[0].heavy_r;
  *heavy_entries = p->results[0].heavy_nnz;
  return GSOFA_OK;
}

int gsofa_early_top(gsofa_plan* p, int64_t* items, double* ms, uint64_t* work3) {
  if (!p || !items || !ms || !work3) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  *items = p->set.top_items();
  int rc = p->set.top_ms(ms);
  if (rc != GSOFA_OK) return rc;
  for (int k = 0; k < 3; ++k) work3[k] = p->results[0].exec_top[k];
  return GSOFA_OK;
}

int gsofa_executed_work(gsofa_plan* p, uint64_t* work6) {
  if (!p || !work6) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  for (int k = 0; k < 6; ++k) work6[k] = p->results[0].exec[k];
  return GSOFA_OK;
}

// ---- supernodal export -------------------------------------------------------------------

static int plan_flags(gsofa_plan* p, std::vector<uint8_t>& flags) {
  if (!p->have_result) {
    int rc = gsofa_sync(p);
    if (rc != GSOFA_OK) return rc;
  }
  flags.resize((size_t)p->prep.n);
  return p->set.read_counts(0, nullptr, nullptr, flags.data(), nullptr);
}

int gsofa_supernode_partition(gsofa_plan* p, int64_t* nsuper, int32_t* xsup, int32_t* supno) {
  if (!p || !nsuper) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  std::vector<uint8_t> flags;
  int rc = plan_flags(p, flags);
  if (rc != GSOFA_OK) return rc;
  *nsuper = supernode_partition(p->prep.n, flags.data(), xsup, supno);
  return GSOFA_OK;
}

int gsofa_batch_supernode_partition(gsofa_batch* b, int32_t index, int64_t* nsuper, int32_t* xsup, int32_t* supno) {
  if (!b || !nsuper) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (index < 0 || index >= b->set.nmats()) return fail(GSOFA_ERR_INVALID_ARG, "batch index out of range");
  if (!b->have_result) {
    int rc = gsofa_batch_sync(b);
    if (rc != GSOFA_OK) return rc;
  }
  std::vector<uint8_t> flags((size_t)b->preps[index].n);
  int rc = b->set.read_counts(index, nullptr, nullptr, flags.data(), nullptr);
  if (rc != GSOFA_OK) return rc;
  *nsuper = supernode_partition(b->preps[index].n, flags.data(), xsup, supno);
  return GSOFA_OK;
}

static int plan_structure_cache(gsofa_plan* p) {
  if (p->sn_cached) return GSOFA_OK;
  if (!p->set.structure_ready()) return fail(GSOFA_ERR_STATE, "structure not produced: enable it and run first");
  int rc = plan_flags(p, p->sn_flags);
  if (rc != GSOFA_OK) return rc;
  rc = p->set.read_structure(0, p->sn_up, p->sn_ui, p->sn_lp, p->sn_li);
  if (rc != GSOFA_OK) return rc;
  p->sn_cached = true;
  return GSOFA_OK;
}

int gsofa_supernodal_structure_size(gsofa_plan* p, int64_t* nsuper, int64_t* nnz_lsub, int64_t* nnz_usub) {
  if (!p || !nsuper || !nnz_lsub || !nnz_usub) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  int rc = plan_structure_cache(p);
  if (rc != GSOFA_OK) return rc;
  return supernodal_panels(p->prep.n, p->sn_flags.data(), p->sn_up.data(), p->sn_ui.data(), p->sn_lp.data(),
                           p->sn_li.data(), nsuper, nnz_lsub, nnz_usub, nullptr, nullptr, nullptr, nullptr, nullptr);
}

int gsofa_supernodal_structure_get(gsofa_plan* p, int32_t* xsup, int64_t* xlsub, int32_t* lsub, int64_t* xusub,
                                   int32_t* usub) {
  if (!p || !xsup || !xlsub || !xusub) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  int rc = plan_structure_cache(p);
  if (rc != GSOFA_OK) return rc;
  int64_t ns = 0, nl = 0, nu = 0;
  rc = supernodal_panels(p->prep.n, p->sn_flags.data(), p->sn_up.data(), p->sn_ui.data(), p->sn_lp.data(),
                         p->sn_li.data(), &ns, &nl, &nu, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc != GSOFA_OK) return rc;
  if ((nl > 0 && !lsub) || (nu > 0 && !usub)) return fail(GSOFA_ERR_INVALID_ARG, "lsub / usub is null");
  return supernodal_panels(p->prep.n, p->sn_flags.data(), p->sn_up.data(), p->sn_ui.data(), p->sn_lp.data(),
                           p->sn_li.data(), &ns, &nl, &nu, xsup, xlsub, lsub, xusub, usub);
}

void gsofa_plan_destroy(gsofa_plan* p) { delete p; }

// ---- batches ---------------------------------------------------------------------------------

static void config_of(const DeviceSet& s, int32_t* out4) {
  out4[0] = s.heavy_threads();
  out4[1] = s.heavy_wgs_per_cu();
  out4[2] = s.options().chain_block;
  out4[3] = s.qcap();
}

int gsofa_plan_config(gsofa_plan* p, int32_t* out4) {
  if (!p || !out4) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  config_of(p->set, out4);
  return GSOFA_OK;
}

int gsofa_batch_config(gsofa_batch* b, int32_t* out4) {
  if (!b || !out4) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  config_of(b->set, out4);
  return GSOFA_OK;
}

int gsofa_batch_create(const gsofa_csr* mats, int32_t count, const gsofa_options* opt, gsofa_batch** out) {
  if (!out || !mats || count <= 0) return fail(GSOFA_ERR_INVALID_ARG, "bad batch arguments");
  *out = nullptr;
  gsofa_options def;
  gsofa_options_default(&def);
  Options o = to_options(opt ? opt : &def);
  if (o.structure) return fail(GSOFA_ERR_INVALID_ARG, "batches report counts and supernodes; structure is per plan");
  std::unique_ptr<gsofa_batch> b(new (std::nothrow) gsofa_batch);
  if (!b) return fail(GSOFA_ERR_OUT_OF_MEMORY, "batch");
  b->preps.resize(count);
  std::vector<const Prepared*> ptrs(count);
  for (int32_t i = 0; i < count; ++i) {
    int rc = prepare(&mats[i], b->preps[i]);
    if (rc != GSOFA_OK) return fail(rc, "batch member " + std::to_string(i) + ": " + gsofa_last_error());
    ptrs[i] = &b->preps[i];
  }
  int rc = GSOFA_OK;
  if (o.autotune) {
    Options tuned;
    rc = autotune(ptrs, o, tuned);
    if (rc != GSOFA_OK) return rc;
    o = tuned;
  }
  rc = b->set.create(ptrs, o);
  if (rc != GSOFA_OK) return rc;
  rc = b->timing.init();
  if (rc != GSOFA_OK) return rc;
  if (hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(GSOFA_ERR_HIP, "hipStreamCreate");
  *out = b.release();
  return GSOFA_OK;
}

int gsofa_batch_run(gsofa_batch* b, void* hip_stream) {
  if (!b) return fail(GSOFA_ERR_INVALID_ARG, "batch is null");
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  if (!b->exec) {
    if (hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail(GSOFA_ERR_HIP, "hipStreamBeginCapture");
    int rc = b->set.enqueue(b->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc != GSOFA_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  if (hipEventRecord(b->timing.start, s) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipEventRecord");
  hipError_t e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) return fail(GSOFA_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
  if (hipEventRecord(b->timing.stop, s) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipEventRecord");
  b->timing.stream = s;
  b->timing.pending = true;
  b->have_result = false;
  b->ran = true;
  return GSOFA_OK;
}

int gsofa_batch_sync(gsofa_batch* b) {
  if (!b) return fail(GSOFA_ERR_INVALID_ARG, "batch is null");
  if (!b->ran) return fail(GSOFA_ERR_STATE, "no run was enqueued on this batch (call gsofa_batch_run first)");
  int rc = b->timing.finish();
  if (rc != GSOFA_OK) return rc;
  rc = b->set.read_results(b->results);
  if (rc != GSOFA_OK) return rc;
  b->have_result = true;
  return GSOFA_OK;
}

int gsofa_batch_result_get(gsofa_batch* b, int32_t index, gsofa_result* out) {
  if (!b || !out) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (index < 0 || index >= b->set.nmats()) return fail(GSOFA_ERR_INVALID_ARG, "batch index out of range");
  if (!b->have_result) {
    int rc = gsofa_batch_sync(b);
    if (rc != GSOFA_OK) return rc;
  }
  fill_result(b->results[index], b->set.options(), b->preps[index].n, b->timing.last_ms, out);
  return GSOFA_OK;
}

int gsofa_batch_nzu_get(gsofa_batch* b, int32_t index, int32_t* nzu) {
  if (!b || !nzu) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (index < 0 || index >= b->set.nmats()) return fail(GSOFA_ERR_INVALID_ARG, "batch index out of range");
  if (!b->have_result) {
    int rc = gsofa_batch_sync(b);
    if (rc != GSOFA_OK) return rc;
  }
  return b->set.read_counts(index, nullptr, nullptr, nullptr, nzu);
}

int gsofa_batch_supernodes_get(gsofa_batch* b, int32_t index, uint8_t* flags) {
  if (!b || !flags) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if (index < 0 || index >= b->set.nmats()) return fail(GSOFA_ERR_INVALID_ARG, "batch index out of range");
  if (!b->have_result) {
    int rc = gsofa_batch_sync(b);
    if (rc != GSOFA_OK) return rc;
  }
  return b->set.read_counts(index, nullptr, nullptr, flags, nullptr);
}

void gsofa_batch_destroy(gsofa_batch* b) { delete b; }

// ---- the reference's ASCII files ----------------------------------------------------------

static int read_ints(const char* path, std::vector<int64_t>& out, int64_t expect_min) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::vector<char> buf;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  std::fclose(f);
  buf.push_back('\0');
  out.clear();
  char* p = buf.data();
  while (true) {
    while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') ++p;
    if (!*p) break;
    char* e = nullptr;
    errno = 0;
    const long long v = std::strtoll(p, &e, 10);
    if (e == p || errno) return fail(GSOFA_ERR_IO, std::string("malformed integer in ") + path);
    out.push_back(v);
    p = e;
  }
  if ((int64_t)out.size() < expect_min) return fail(GSOFA_ERR_IO, std::string("too few integers in ") + path);
  return GSOFA_OK;
}

int gsofa_read_ascii(const char* beg_file, const char* end_file, const char* csr_file, gsofa_host_matrix** out) {
  if (!beg_file || !end_file || !csr_file || !out) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  std::vector<int64_t> c, b, e;
  int rc = read_ints(csr_file, c, 2);
  if (rc != GSOFA_OK) return rc;
  const int64_t E = c[0], n = c[1];  // header order: edge count, then vertex count (...cuh:2577-2580)
  if (E < 0 || n <= 0 || n > INT32_MAX || E > INT32_MAX) return fail(GSOFA_ERR_IO, "bad csr header");
  if ((int64_t)c.size() < 2 + E) return fail(GSOFA_ERR_IO, "csr file holds fewer than E column ids");
  rc = read_ints(beg_file, b, n);
  if (rc != GSOFA_OK) return rc;
  rc = read_ints(end_file, e, n);
  if (rc != GSOFA_OK) return rc;
  std::unique_ptr<gsofa_host_matrix> h(new (std::nothrow) gsofa_host_matrix);
  if (!h) return fail(GSOFA_ERR_OUT_OF_MEMORY, "host matrix");
  HostMatrix& M = h->m;
  M.n = (int32_t)n;
  M.adj.resize((size_t)E);
  M.beg.resize((size_t)n);
  M.end.resize((size_t)n);
  for (int64_t k = 0; k < E; ++k) M.adj[k] = (int32_t)c[2 + k];
  for (int64_t v = 0; v < n; ++v) {
    M.beg[v] = (int32_t)b[v];
    M.end[v] = (int32_t)e[v];
  }
  gsofa_csr view = M.view();
  rc = validate_csr(&view);
  if (rc != GSOFA_OK) return rc;
  *out = h.release();
  return GSOFA_OK;
}

int gsofa_write_ascii(const gsofa_csr* a, const char* beg_file, const char* end_file, const char* csr_file) {
  int rc = validate_csr(a);
  if (rc != GSOFA_OK) return rc;
  FILE* f = std::fopen(csr_file, "wb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot create ") + csr_file);
  std::fprintf(f, "%lld\n%d\n", (long long)a->nnz, a->n);
  for (int64_t k = 0; k < a->nnz; ++k) std::fprintf(f, "%d ", a->adj[k]);
  std::fclose(f);
  f = std::fopen(beg_file, "wb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot create ") + beg_file);
  for (int32_t v = 0; v < a->n; ++v) std::fprintf(f, "%d ", a->beg[v]);
  std::fclose(f);
  f = std::fopen(end_file, "wb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot create ") + end_file);
  for (int32_t v = 0; v < a->n; ++v) std::fprintf(f, "%d ", a->end[v]);
  std::fclose(f);
  return GSOFA_OK;
}

// ---- binary CSR -------------------------------------------------------------------------------

static int read_file(const char* path, std::vector<unsigned char>& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  buf.clear();
  unsigned char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  if (err) return fail(GSOFA_ERR_IO, std::string("read error on ") + path);
  return GSOFA_OK;
}

static int64_t word_at(const unsigned char* p, int32_t bytes) {
  if (bytes == 4) {
    int32_t v;
    std::memcpy(&v, p, 4);
    return v;
  }
  int64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

static bool write_words(FILE* f, const int64_t* v, size_t count, int32_t bytes) {
  for (size_t i = 0; i < count; ++i) {
    if (bytes == 4) {
      const int32_t x = (int32_t)v[i];
      if (std::fwrite(&x, 4, 1, f) != 1) return false;
    } else if (std::fwrite(&v[i], 8, 1, f) != 1) {
      return false;
    }
  }
  return true;
}

int gsofa_read_binary(const char* beg_file, const char* csr_file, int32_t index_bytes, int32_t vert_bytes,
                      gsofa_host_matrix** out) {
  if (!beg_file || !csr_file || !out) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if ((index_bytes != 4 && index_bytes != 8) || (vert_bytes != 4 && vert_bytes != 8))
    return fail(GSOFA_ERR_INVALID_ARG, "integer widths must be 4 or 8 bytes");
  std::vector<unsigned char> b, c;
  int rc = read_file(beg_file, b);
  if (rc != GSOFA_OK) return rc;
  if (b.size() % (size_t)index_bytes != 0 || b.size() < 2 * (size_t)index_bytes)
    return fail(GSOFA_ERR_IO, "beg file must hold n+1 offsets (n >= 1)");
  const int64_t n = (int64_t)(b.size() / index_bytes) - 1;  // vert_count = fsize / sizeof - 1 (graph.hpp:18)
  if (n > INT32_MAX) return fail(GSOFA_ERR_CAPACITY, "n must fit in int32");
  std::vector<int64_t> off((size_t)n + 1);
  for (int64_t v = 0; v <= n; ++v) off[v] = word_at(b.data() + v * index_bytes, index_bytes);
  for (int64_t v = 0; v < n; ++v)
    if (off[v] < 0 || off[v + 1] < off[v]) return fail(GSOFA_ERR_IO, "row offsets must be non-negative and non-decreasing");
  const int64_t E = off[n];  // edge_count = beg_pos[vert_count] (graph.hpp:34)
  if (E > INT32_MAX) return fail(GSOFA_ERR_CAPACITY, "edge count must fit in int32");
  rc = read_file(csr_file, c);
  if (rc != GSOFA_OK) return rc;
  if ((int64_t)(c.size() / vert_bytes) < E) return fail(GSOFA_ERR_IO, "csr file holds fewer column ids than the last offset");
  std::unique_ptr<gsofa_host_matrix> h(new (std::nothrow) gsofa_host_matrix);
  if (!h) return fail(GSOFA_ERR_OUT_OF_MEMORY, "host matrix");
  HostMatrix& M = h->m;
  M.n = (int32_t)n;
  M.adj.resize((size_t)E);
  M.beg.resize((size_t)n);
  M.end.resize((size_t)n);
  for (int64_t k = 0; k < E; ++k) {
    const int64_t x = word_at(c.data() + k * vert_bytes, vert_bytes);
    if (x < 0 || x >= n) return fail(GSOFA_ERR_IO, "column id out of range in " + std::string(csr_file));
    M.adj[k] = (int32_t)x;
  }
  for (int64_t v = 0; v < n; ++v) {
    M.beg[v] = (int32_t)off[v];
    M.end[v] = (int32_t)off[v + 1];
  }
  gsofa_csr view = M.view();
  rc = validate_csr(&view);
  if (rc != GSOFA_OK) return rc;
  *out = h.release();
  return GSOFA_OK;
}

int gsofa_write_binary(const gsofa_csr* a, const char* beg_file, const char* csr_file, int32_t index_bytes,
                       int32_t vert_bytes) {
  int rc = validate_csr(a);
  if (rc != GSOFA_OK) return rc;
  if (!beg_file || !csr_file) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  if ((index_bytes != 4 && index_bytes != 8) || (vert_bytes != 4 && vert_bytes != 8))
    return fail(GSOFA_ERR_INVALID_ARG, "integer widths must be 4 or 8 bytes");
  // rows back to back in row order (the input may hold them anywhere in adj)
  std::vector<int64_t> off((size_t)a->n + 1, 0), col;
  col.reserve((size_t)a->nnz);
  for (int32_t v = 0; v < a->n; ++v) {
    for (int32_t k = a->beg[v]; k < a->end[v]; ++k) col.push_back(a->adj[k]);
    off[v + 1] = (int64_t)col.size();
  }
  FILE* f = std::fopen(beg_file, "wb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot create ") + beg_file);
  bool ok = write_words(f, off.data(), off.size(), index_bytes);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(GSOFA_ERR_IO, std::string("write error on ") + beg_file);
  f = std::fopen(csr_file, "wb");
  if (!f) return fail(GSOFA_ERR_IO, std::string("cannot create ") + csr_file);
  ok = write_words(f, col.data(), col.size(), vert_bytes);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(GSOFA_ERR_IO, std::string("write error on ") + csr_file);
  return GSOFA_OK;
}

// ---- the drop-in entry ------------------------------------------------------------------------

int gsofa_symbfact_min_id(int argc, char** argv, int rank, uint64_t* fill_count, double* time_ms,
                          int64_t* n_supernodes, uint64_t* edge_checks) {
  if (argc != 11 || !argv) return fail(GSOFA_ERR_INVALID_ARG, "expected 10 arguments: beg end csr num_process "
                                                              "chunk_size percent_cat0 percent_cat2 "
                                                              "N_blocks_source_cat2 N_GPU_Node N_src_group");
  if (!fill_count || !time_ms || !n_supernodes || !edge_checks) return fail(GSOFA_ERR_INVALID_ARG, "null output");
  const int num_process = std::atoi(argv[4]);
  const int chunk = std::atoi(argv[5]);
  const int n_gpu_node = std::atoi(argv[9]);
  if (num_process <= 0 || chunk <= 0 || n_gpu_node <= 0)
    return fail(GSOFA_ERR_INVALID_ARG, "num_process, chunk size and N_GPU_Node must be positive");
  gsofa_host_matrix* h = nullptr;
  int rc = gsofa_read_ascii(argv[1], argv[2], argv[3], &h);
  if (rc != GSOFA_OK) return rc;
  std::unique_ptr<gsofa_host_matrix> hold(h);
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) return fail(GSOFA_ERR_NO_DEVICE, "no HIP device");
  gsofa_options o;
  gsofa_options_default(&o);
  o.chunk_size = chunk;
  o.n_processes = num_process;
  o.device = (rank % n_gpu_node) % devices;
  gsofa_csr view = h->m.view();
  gsofa_plan* plan = nullptr;
  rc = gsofa_plan_create(&view, &o, &plan);
  if (rc != GSOFA_OK) return rc;
  std::unique_ptr<gsofa_plan, void (*)(gsofa_plan*)> hold_plan(plan, gsofa_plan_destroy);
  rc = gsofa_run(plan, nullptr);
  if (rc != GSOFA_OK) return rc;
  gsofa_result r;
  rc = gsofa_result_get(plan, &r);
  if (rc != GSOFA_OK) return rc;
  *fill_count += r.fill_count;
  *time_ms += r.time_ms;
  *n_supernodes = r.supernodes;
  *edge_checks = r.edge_checks;
  return GSOFA_OK;
}

}  // extern "C"
