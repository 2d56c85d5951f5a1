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

struct letkf_handle {
  letkf_params_t prm;
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf nloc, loff, actidx, act_p, scan_tmp, yT, lobs, lrloc, G, lam, V, bvec, status, bounds;
  DevBuf nrow, bin_tmp;
  DevBuf eD, eE, logc, logs, logi, logn, logx, logr, yvec;  // eigensolve workspaces
  DevBuf rotsum;  // stage_timing 2: logged QL rotations, summed over slots
  DevBuf otrig;   // (sin, cos) of every obs latitude
  DevBuf adj_part;  // adjoint: per-point (dL/dalpha, dL/drho)
  DevBuf bhist, bcur, blist;              // per-chunk size buckets of active slots
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
                     &status, &bounds, &nrow, &bin_tmp, &eD, &eE, &logc, &logs, &logi, &logn, &logx, &logr, &yvec, &rotsum, &otrig, &adj_part, &bhist, &bcur, &blist, &h_grid, &h_obs, &h_f3, &h_f2, &h_a3, &h_a2};
    for (DevBuf* b : all) b->release();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    events.clear();
    for (hipEvent_t e : kevents) (void)hipEventDestroy(e);
    kevents.clear();
  }
};

// Kernel status bits -> return code (with the error text set).
int status_to_rc(int status);
