// noalloc.hip -- letkf_core_batched_noalloc (include/letkf_engine.h): the
// batched letkf_core of letkf_core_batched on the same kernels (Gram,
// eigensolve, weights), its scratch in one file-scope device array instead
// of a handle's workspaces, so that the call allocates nothing.
//
// Scratch layout (one per device, kScratchBytes in all), for nb problems of
// nmem K:  status [4 int, padded to 256 B] | nrow [nb int32] | lam [nb K] |
// bvec [nb K] | G [nb K K] | V [nb K K]   (doubles; each part 256-B aligned).
// Calls on one device share the array: calls that overlap in time on one
// device (different streams, or threads) race on it.
#include <mutex>

#include "letkf_handle.h"

using namespace letkf;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this device's scratch, looked up once per device with
// that device current (each device holds its own copy of the array).
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

int scratch_ptr(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_err(LETKF_ERR_ARG, "device %d beyond the supported %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(LETKF_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch array");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return LETKF_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of scratch for nb problems of nmem K (the layout above).
size_t scratch_need(long long nb, int K) {
  return 256 + align256(sizeof(int32_t) * nb) + 2 * align256(sizeof(double) * nb * K) +
         2 * align256(sizeof(double) * nb * K * K);
}

__global__ void k_clear_status(int* st) {
  if (threadIdx.x < 4) st[threadIdx.x] = 0;
}

// The kernels' status bits as the letkf_status_t the synchronous entry
// points return (same priority as their status_to_rc).
__global__ void k_copy_status(const int* st, int32_t* out) {
  if (threadIdx.x != 0) return;
  const int b = st[0];
  *out = (b & ST_NONFINITE) ? LETKF_ERR_NONFINITE
         : (b & (ST_EIG_NOCONV | ST_EIG_NEG)) ? LETKF_ERR_EIG
         : (b & ST_CAPACITY) ? LETKF_ERR_CAPACITY : LETKF_OK;
}

// Largest number of problems of nmem K whose layout fits the scratch.
int capacity(int K) {
  long long lo = 0, hi = (long long)(kScratchBytes / (16.0 * K * K)) + 1;
  while (lo < hi) {
    const long long mid = (lo + hi + 1) / 2;
    if (scratch_need(mid, K) <= kScratchBytes) lo = mid; else hi = mid - 1;
  }
  return (int)lo;
}

}  // namespace

extern "C" int letkf_core_batched_noalloc(const letkf_params_t* p, int32_t nb, const int64_t* offs,
                                          const void* hdxf, const void* rdiag, const void* rloc, const void* dep,
                                          void* wvec, void* wmat, int32_t* status, void* stream) {
  if (!p) return set_err(LETKF_ERR_ARG, "params is NULL");
  const int K = p->nmem;
  if (K < 2 || K > kMaxK) return set_err(LETKF_ERR_ARG, "nmem must be in [2,%d] (got %d)", kMaxK, K);
  if (p->real_kind != 4 && p->real_kind != 8)
    return set_err(LETKF_ERR_ARG, "real_kind must be 4 or 8 (got %d)", p->real_kind);
  if (!(p->cov_infl_mul > 0.0)) return set_err(LETKF_ERR_ARG, "cov_infl_mul must be > 0");
  if (nb < 0 || !status || (nb > 0 && (!offs || !wvec || !wmat)))
    return set_err(LETKF_ERR_ARG, "bad core_batched_noalloc arguments");
  const int cap = capacity(K);
  if (nb > cap)
    return set_err(LETKF_ERR_CAPACITY,
                   "letkf_core_batched_noalloc: %d problems of nmem %d need %zu bytes of scratch; the "
                   "scratch holds %zu (at most %d problems of this nmem per call)",
                   nb, K, scratch_need(nb, K), kScratchBytes, cap);
  void* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  char* q = (char*)base;
  int* st = (int*)q;
  q += 256;
  int32_t* nrow = (int32_t*)q;
  q += align256(sizeof(int32_t) * (size_t)nb);
  double* lam = (double*)q;
  q += align256(sizeof(double) * (size_t)nb * K);
  double* bvec = (double*)q;
  q += align256(sizeof(double) * (size_t)nb * K);
  double* G = (double*)q;
  q += align256(sizeof(double) * (size_t)nb * K * K);
  double* V = (double*)q;
  // the status word keeps the previous call's bits: cleared first (a fill
  // kernel, so that a captured call clears it on every replay)
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, s, st);
  HIP_TRY(hipGetLastError());
  if (nb > 0) {
    const double rho = p->cov_infl_mul;
    if (p->real_kind == 8)
      HIP_TRY(launch_gram_csr<double>(K, rho, nb, offs, (const double*)hdxf, (const double*)rdiag,
                                      (const double*)rloc, (const double*)dep, G, bvec, nrow, s));
    else
      HIP_TRY(launch_gram_csr<float>(K, rho, nb, offs, (const float*)hdxf, (const float*)rdiag, (const float*)rloc,
                                     (const float*)dep, G, bvec, nrow, s));
    HIP_TRY(launch_eig_fp64(K, nb, G, lam, V, st, 0, nullptr, nullptr, 0, s));
    if (p->real_kind == 8)
      HIP_TRY(launch_form_weights<double>(K, rho, nb, nrow, lam, V, bvec, (double*)wvec, (double*)wmat, s));
    else
      HIP_TRY(launch_form_weights<float>(K, rho, nb, nrow, lam, V, bvec, (float*)wvec, (float*)wmat, s));
  }
  hipLaunchKernelGGL(k_copy_status, dim3(1), dim3(64), 0, s, st, status);
  HIP_TRY(hipGetLastError());
  return LETKF_OK;
}
