// batch.cpp -- several independent compressed bases stepped in one persistent
// launch, with a HIP-graph capture of the whole step and an autotuner for the
// launch and tile shape (waves per workgroup, workgroups per CU, the work
// split between bases).
//
// The reference steps one basis pair per call (stepPlume,
// SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:453-501).  A batch holds bases of
// different grids and column counts; every launch splits the persistent grid
// into one contiguous range of workgroups per basis, sized by the basis's
// share of the work, and a second launch reduces every basis's q.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"
#include "cfs_device.h"

namespace cfs {
hipError_t launch_batch_step(const BatchEntry* d_entries, int n, int mode, int grid, int qrows, size_t lds, int wpw,
                             hipStream_t s);
hipError_t batch_occupancy(int mode, size_t lds, int wpw, int* wg_per_cu);
size_t step_lds_bytes_for(int r, int nk, int wpw);
}  // namespace cfs

using namespace cfs;

struct cfs_batch {
  int device = 0;
  std::vector<cfs_basis*> bases;
  std::vector<BatchEntry> entries;  // host copy
  BatchEntry* d_entries = nullptr;
  double* d_partials = nullptr;
  size_t partial_doubles = 0;
  // per tile shape (index 0: 2 waves per workgroup, 1: 4): LDS bytes and
  // workgroups per CU that fit
  size_t lds_of[2] = {0, 0};
  int occupancy_of[2] = {1, 1};
  size_t lds = 0;
  int cus = 0, occupancy = 1;
  int wpw = 4;           // tuned: waves per workgroup (2 or 4)
  int wg_per_cu = 0;     // tuned: workgroups per CU of the persistent grid
  float item_weight = 0; // tuned: per-item work (in tokens) added to a basis's token count when splitting
  int grid = 0, qrows = 0;
  bool bound = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~cfs_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (d_entries) (void)hipFree(d_entries);
    if (d_partials) (void)hipFree(d_partials);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

// Splits `grid` workgroups over the entries in proportion to
// tokens + item_weight * items (each entry gets at least one) and assigns the
// partial-q rows.  Returns false if the grid is smaller than the entry count.
bool plan(cfs_batch* b, int grid) {
  const int n = (int)b->bases.size();
  if (grid < n) return false;
  std::vector<double> work(n);
  double total = 0.0;
  for (int i = 0; i < n; i++) {
    const cfs_basis* B = b->bases[i];
    const double items = 3.0 * basis_num_blocks(B);
    work[i] = (double)basis_total_tokens(B) + b->item_weight * items;
    total += work[i];
  }
  // largest remainder apportionment of grid - n extra workgroups
  std::vector<int> cnt(n, 1);
  std::vector<std::pair<double, int>> rem;
  int left = grid - n;
  for (int i = 0; i < n; i++) {
    const double ideal = total > 0 ? (grid - n) * work[i] / total : 0.0;
    // never more workgroups than items
    const int maxw = std::max(1, 3 * basis_num_blocks(b->bases[i]));
    const int add = std::min((int)std::floor(ideal), maxw - 1);
    cnt[i] += add;
    left -= add;
    rem.push_back({ideal - std::floor(ideal), i});
  }
  std::sort(rem.begin(), rem.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  for (size_t j = 0; left > 0 && j < rem.size(); j++) {
    const int i = rem[j].second;
    if (cnt[i] < 3 * basis_num_blocks(b->bases[i])) {
      cnt[i]++;
      left--;
    }
  }
  int wg = 0, q = 0;
  size_t part = 0;
  for (int i = 0; i < n; i++) {
    BatchEntry& e = b->entries[i];
    e.wg_begin = wg;
    e.wg_count = cnt[i];
    e.q_begin = q;
    e.partials = b->d_partials + part;
    wg += cnt[i];
    q += e.B.r;
    part += (size_t)cnt[i] * e.B.r;
  }
  if (part > b->partial_doubles) return false;
  b->grid = wg;
  b->qrows = q;
  return true;
}

int upload(cfs_batch* b) {
  hipError_t e = hipMemcpy(b->d_entries, b->entries.data(), b->entries.size() * sizeof(BatchEntry),
                           hipMemcpyHostToDevice);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "hipMemcpy batch entries");
}

int apply(cfs_batch* b, int wpw, int wg_per_cu, float item_weight) {
  const int k = wpw == 2 ? 0 : 1;
  b->wpw = wpw == 2 ? 2 : 4;
  b->lds = b->lds_of[k];
  b->occupancy = b->occupancy_of[k];
  b->wg_per_cu = std::max(1, std::min(wg_per_cu, b->occupancy));
  b->item_weight = item_weight;
  if (!plan(b, std::max((int)b->bases.size(), b->cus * b->wg_per_cu)))
    return fail(CFS_ERR_ARG, "batch too large for one launch");
  b->drop_graph();
  return upload(b);
}

}  // namespace

extern "C" {

int cfs_batch_create(cfs_basis* const* bases, int n, cfs_batch** out) {
  if (!out || !bases || n < 1) return fail(CFS_ERR_ARG, "need at least one basis");
  *out = nullptr;
  std::unique_ptr<cfs_batch> b(new cfs_batch);
  b->device = basis_device(bases[0]);
  int max_r = 0;
  for (int i = 0; i < n; i++) {
    if (!bases[i]) return fail(CFS_ERR_ARG, "null basis in batch");
    if (basis_device(bases[i]) != b->device) return fail(CFS_ERR_ARG, "batch bases must share one device");
    b->bases.push_back(bases[i]);
    const DevBasis& D = basis_dev(bases[i]);
    max_r = std::max(max_r, D.r);
    for (int k = 0; k < 2; k++) b->lds_of[k] = std::max(b->lds_of[k], step_lds_bytes_for(D.r, D.nk, k ? 4 : 2));
  }
  cfs::DeviceGuard g(b->device);
  hipError_t e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
  for (int k = 0; k < 2; k++) {
    e = batch_occupancy(3, b->lds_of[k], k ? 4 : 2, &b->occupancy_of[k]);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (b->occupancy_of[k] < 1) return fail(CFS_ERR_ARG, "batch LDS needs exceed one workgroup per CU");
  }
  b->entries.resize(n);
  for (int i = 0; i < n; i++) {
    BatchEntry& be = b->entries[i];
    std::memset(&be, 0, sizeof(be));
    be.B = basis_dev(bases[i]);
    be.damp = basis_damp(bases[i]);
    be.tab = reinterpret_cast<const Tables*>(basis_tables(bases[i]));
  }
  const int max_grid = std::max(n, b->cus * std::max(b->occupancy_of[0], b->occupancy_of[1]));
  b->partial_doubles = (size_t)(max_grid + n) * max_r;
  e = hipMalloc(reinterpret_cast<void**>(&b->d_partials), b->partial_doubles * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc batch partials");
  e = hipMalloc(reinterpret_cast<void**>(&b->d_entries), (size_t)n * sizeof(BatchEntry));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc batch entries");
  int st = apply(b.get(), 4, b->occupancy_of[1], 0.0f);
  if (st) return st;
  *out = b.release();
  return CFS_OK;
}

int cfs_batch_bind(cfs_batch* b, const float* const* v_in, double* const* q_out, const double* const* q_in,
                   float* const* field) {
  if (!b || !v_in || !q_out || !q_in || !field) return fail(CFS_ERR_ARG, "null argument");
  for (size_t i = 0; i < b->entries.size(); i++) {
    if (!v_in[i] || !q_out[i] || !q_in[i] || !field[i]) return fail(CFS_ERR_ARG, "null field or q in batch");
    b->entries[i].v_in = v_in[i];
    b->entries[i].q_out = q_out[i];
    b->entries[i].q_in = q_in[i];
    b->entries[i].field = field[i];
  }
  cfs::DeviceGuard g(b->device);
  b->bound = true;
  b->drop_graph();
  return upload(b);
}

int cfs_batch_step(cfs_batch* b, void* stream) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (!b->bound) return fail(CFS_ERR_ARG, "cfs_batch_bind must precede cfs_batch_step");
  cfs::DeviceGuard g(b->device);
  hipError_t e = launch_batch_step(b->d_entries, (int)b->entries.size(), 3, b->grid, b->qrows, b->lds,
                                   b->wpw, (hipStream_t)stream);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "batch step launch");
}

int cfs_batch_tune(cfs_batch* b, int reps, void* stream) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (!b->bound) return fail(CFS_ERR_ARG, "cfs_batch_bind must precede cfs_batch_tune");
  if (reps < 1) reps = 3;
  cfs::DeviceGuard g(b->device);
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
  const float weights[] = {0.0f, 256.0f, 1024.0f, 4096.0f};
  float best_ms = 1e30f;
  int best_wpw = 4, best_w = b->occupancy_of[1];
  float best_iw = 0.0f;
  int st = CFS_OK;
  for (int wpw : {2, 4})
  for (int w = 1; w <= b->occupancy_of[wpw == 2 ? 0 : 1] && st == CFS_OK; w++) {
    for (float iw : weights) {
      st = apply(b, wpw, w, iw);
      if (st) break;
      // one untimed launch, then the timed ones
      e = launch_batch_step(b->d_entries, (int)b->entries.size(), 3, b->grid, b->qrows, b->lds, b->wpw, s);
      if (e == hipSuccess) e = hipEventRecord(e0, s);
      for (int i = 0; e == hipSuccess && i < reps; i++)
        e = launch_batch_step(b->d_entries, (int)b->entries.size(), 3, b->grid, b->qrows, b->lds, b->wpw, s);
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        st = hip_fail(e, "batch tuning launch");
        break;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best_wpw = wpw;
        best_w = w;
        best_iw = iw;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  return apply(b, best_wpw, best_w, best_iw);
}

int cfs_batch_get_config(const cfs_batch* b, int* grid, int* wg_per_cu, float* item_weight) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (grid) *grid = b->grid;
  if (wg_per_cu) *wg_per_cu = b->wg_per_cu;
  if (item_weight) *item_weight = b->item_weight;
  return CFS_OK;
}

int cfs_batch_set_config(cfs_batch* b, int wg_per_cu, float item_weight) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (wg_per_cu < 1 || item_weight < 0) return fail(CFS_ERR_ARG, "bad batch configuration");
  cfs::DeviceGuard g(b->device);
  return apply(b, b->wpw, wg_per_cu, item_weight);
}

int cfs_batch_get_shape(const cfs_batch* b, int* waves_per_wg) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (waves_per_wg) *waves_per_wg = b->wpw;
  return CFS_OK;
}

int cfs_batch_set_shape(cfs_batch* b, int waves_per_wg) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (waves_per_wg != 2 && waves_per_wg != 4) return fail(CFS_ERR_ARG, "waves per workgroup must be 2 or 4");
  cfs::DeviceGuard g(b->device);
  return apply(b, waves_per_wg, b->wg_per_cu, b->item_weight);
}

int cfs_batch_graph_capture(cfs_batch* b, void* stream) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (!b->bound) return fail(CFS_ERR_ARG, "cfs_batch_bind must precede cfs_batch_graph_capture");
  if (!stream) return fail(CFS_ERR_ARG, "graph capture needs a non-default stream");
  cfs::DeviceGuard g(b->device);
  b->drop_graph();
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
  hipError_t le = launch_batch_step(b->d_entries, (int)b->entries.size(), 3, b->grid, b->qrows, b->lds, b->wpw, s);
  e = hipStreamEndCapture(s, &b->graph);
  if (le != hipSuccess) return hip_fail(le, "batch step launch during capture");
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  return CFS_OK;
}

int cfs_batch_graph_launch(cfs_batch* b, void* stream) {
  if (!b) return fail(CFS_ERR_ARG, "null batch");
  if (!b->exec) return fail(CFS_ERR_ARG, "no captured graph (cfs_batch_graph_capture)");
  cfs::DeviceGuard g(b->device);
  hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "hipGraphLaunch");
}

int cfs_batch_destroy(cfs_batch* b) {
  delete b;
  return CFS_OK;
}

}  // extern "C"
