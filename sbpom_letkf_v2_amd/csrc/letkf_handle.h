// letkf_handle.h -- host-side handle, workspaces and error plumbing shared by
// the C ABI translation units.  Not part of the public ABI.
#pragma once
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "letkf_internal.h"

int set_err(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      return set_err(LETKF_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),     \
                     __FILE__, __LINE__);                                                      \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    want = want < 256 ? 256 : want;
    hipError_t e = hipMalloc(&ptr, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <typename X>
  X* as() const {
    return (X*)ptr;
  }
};

// The geometry-dependent part of an analysis (local-obs counts, chunk plan,
// size buckets), made by letkf_da_main_plan or inside letkf_da_main_device
// and reused by letkf_da_main_enqueue.
struct AnalysisPlan {
  bool valid = false;
  letkf_params_t prm{};     // parameters it was made with
  const void* key[12] = {};  // grid and obs position arrays it was made from
  int nobs = 0;
  long long P = 0, PC = 0, nch = 0, yp = 0;
  bool reg_eig = false;
  int dual = 0;
  bool rat = false;  // f(T) by rational functions (batched form, K <= 128)
  int pm_ld = 0;     // register update through point-major chunk copies: their row pitch (0: none)
  bool pm_stats = false;  // the copies also form each point's forecast mean and spread (pst)
  bool ysplit = false;  // the dual-form BF16x3 Gram reads the obs rows pre-split (ysp)
  bool fastloc = false; // rloc from the FP32 exponential (precision bf16x3t)
  double kmax = 1e308;  // spectrum width past which a slot takes the FP64 fallback
  std::vector<long long> bpts, bv;  // chunk point bounds; per-chunk (local obs, active slot) offsets + totals
  std::vector<int32_t> hist, boff;  // per (chunk, size bucket) slot counts and list offsets
  letkf_stats_t st{};               // counters known from the plan
  uint64_t epoch = 0;               // distinct for every plan made on the handle
};

struct letkf_handle {
  letkf_params_t prm;
  AnalysisPlan plan;
  // Graph guard: plans made so far, the plan of the last public enqueue, and
  // the entry point that last overwrote the plan's workspaces (a graph
  // captured from an enqueue refers to them).
  uint64_t plans_made = 0, enq_epoch = 0;
  const char* invalidated_by = nullptr;
  // points whose local-obs counts nloc holds (count_nobsl of the last
  // analysis, plan or adjoint on the handle; -1 none); nloc_zero: no obs
  long long nloc_P = -1;
  bool nloc_zero = false;
  void invalidate_plan(const char* by) {
    plan.valid = false;
    invalidated_by = by;
  }
  int vpath[5] = {0, 0, 0, 0, 0};  // vector path per size bucket (16, 32, 48, 64, 128): letkf_autotune
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf nloc, loff, actidx, act_p, scan_tmp, yT, lobs, lrloc, G, lam, V, bvec, status, bounds;
  DevBuf nrow, bin_tmp;
  DevBuf eD, eE, logc, logs, logi, logn, logx, logr, yvec;  // eigensolve workspaces
  DevBuf rotsum;  // stage_timing 2: logged QL rotations, summed over slots
  DevBuf otrig;   // (sin, cos) of every obs latitude
  DevBuf ratt;    // coefficient table of the rational path (uploaded once: ratt_ok)
  bool ratt_ok = false;
  DevBuf adj_part;  // adjoint: per-point (dL/dalpha, dL/drho, dL/dsigma_obs, dL/dsigma_obsv), block sums
  DevBuf adj_pair, adj_sort;  // adjoint: per-pair terms; sort keys, values and scratch
  DevBuf bhist, bcur, blist;              // per-chunk size buckets of active slots
  DevBuf bblk;                            // per-256-point-block bucket counts / offsets
  DevBuf xpm, apm;                        // point-major forecast / analysis of a chunk (register update)
  DevBuf pst;                             // their forecast (mean, spread) per point and variable
  DevBuf ysp;                             // obs rows as BF16 hi / lo groups (dual-form BF16x3 Gram)
  // staging for the host-pointer entry point
  DevBuf h_grid, h_obs, h_f3, h_f2, h_a3, h_a2;
  letkf_stats_t stats{};
  std::vector<hipEvent_t> kevents;  // eigensolve kernel marks (stage_timing 2)
  std::vector<int> ktags;
  hipError_t ensure_kevents(size_t n) {
    while (kevents.size() < n) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      kevents.push_back(e);
    }
    return hipSuccess;
  }
  std::vector<hipEvent_t> events;  // stage-timing marks, grown on demand
  hipError_t ensure_events(size_t n) {
    while (events.size() < n) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      events.push_back(e);
    }
    return hipSuccess;
  }
  void release_all() {
    DevBuf* all[] = {&nloc, &loff, &actidx, &act_p, &scan_tmp, &yT, &lobs, &lrloc, &G, &lam, &V, &bvec,
                     &status, &bounds, &nrow, &bin_tmp, &eD, &eE, &logc, &logs, &logi, &logn, &logx, &logr, &yvec, &rotsum, &otrig, &ratt, &adj_part, &adj_pair, &adj_sort, &bhist, &bcur, &blist, &bblk, &xpm, &apm, &pst, &ysp, &h_grid, &h_obs, &h_f3, &h_f2, &h_a3, &h_a2};
    for (DevBuf* b : all) b->release();
    ratt_ok = false;
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    events.clear();
    for (hipEvent_t e : kevents) (void)hipEventDestroy(e);
    kevents.clear();
  }
};

// Kernel status bits -> return code (with the error text set).
int status_to_rc(int status);

namespace letkf {
// Resampling draws (sorted for MR, the SR formula) and the running sums of
// 1/nmonte for the LPF kernels, uploaded into buf (lpfgm.hip).
hipError_t lpf_prepare(int K, int nmonte, int rm_sr, bool f64, const double* rand, DevBuf& buf, hipStream_t s,
                       const double** d_rt, const double** d_tt);
}  // namespace letkf
