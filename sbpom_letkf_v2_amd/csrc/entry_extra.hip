// entry_extra.hip -- C ABI entry points beside da_main: letkf_core_batched,
// letkf_bin_obs and letkf_aoei (include/letkf_engine.h).
#include <hipcub/hipcub.hpp>

#include "letkf_handle.h"

using namespace letkf;

namespace {

__global__ void k_key_fix(int n, int32_t* key, int32_t* vals) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  if (key[o] < 0) key[o] = INT32_MAX;  // outside the grid: sorted last, then dropped
  vals[o] = o;
}

// nobsgrd holds the inclusive cumulative count of every cell; the reference
// never fills row nlat (common_obs.f90:271-345), which stays 0.
__global__ void k_last_row(int nlon, int nlat, int32_t* grd, int32_t* nsorted) {
  if (threadIdx.x == 0) *nsorted = grd[(long long)nlon * (nlat - 1) - 1];
  for (int i = threadIdx.x; i < nlon; i += blockDim.x) grd[(long long)nlon * (nlat - 1) + i] = 0;
}

}  // namespace

extern "C" int letkf_aoei(letkf_handle_t* h, int32_t nobs, void* err, const void* dep, const void* sprd) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  if (nobs < 0 || (nobs > 0 && (!err || !dep || !sprd))) return set_err(LETKF_ERR_ARG, "bad AOEI arguments");
  HIP_TRY(hipSetDevice(h->device));
  if (h->prm.real_kind == 8)
    HIP_TRY(launch_aoei<double>(nobs, (double*)err, (const double*)dep, (const double*)sprd, h->stream));
  else
    HIP_TRY(launch_aoei<float>(nobs, (float*)err, (const float*)dep, (const float*)sprd, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return LETKF_OK;
}

extern "C" int letkf_bin_obs(letkf_handle_t* h, int32_t nlon, int32_t nlat, const void* lon, const void* lat,
                             int32_t nobs, const void* olon, const void* olat, int32_t* idx, int32_t* idy,
                             int32_t* perm, int32_t* nobsgrd, int32_t* nsorted) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  if (nlon < 2 || nlat < 2 || nobs < 0 || !lon || !lat || !nobsgrd || !nsorted)
    return set_err(LETKF_ERR_ARG, "bad bin_obs arguments");
  if (nobs > 0 && (!olon || !olat || !idx || !idy || !perm)) return set_err(LETKF_ERR_ARG, "obs array is NULL");
  if ((long long)nlon * nlat >= INT32_MAX) return set_err(LETKF_ERR_ARG, "grid too large");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const int ncell = nlon * nlat;
  // scratch: key, key_sorted, vals, counts, cub temp
  size_t cub_sort = 0, cub_scan = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_sort, (int32_t*)nullptr, (int32_t*)nullptr,
                                             (int32_t*)nullptr, (int32_t*)nullptr, nobs > 0 ? nobs : 1, 0, 32,
                                             s));
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, cub_scan, (int32_t*)nullptr, (int32_t*)nullptr, ncell, s));
  const size_t n1 = (size_t)(nobs > 0 ? nobs : 1);
  const size_t bytes = sizeof(int32_t) * (3 * n1 + ncell + 1) + std::max(cub_sort, cub_scan) + 256;
  HIP_TRY(h->bin_tmp.ensure(bytes));
  int32_t* key = h->bin_tmp.as<int32_t>();
  int32_t* key2 = key + n1;
  int32_t* vals = key2 + n1;
  int32_t* counts = vals + n1;
  void* cub_tmp = (void*)(((uintptr_t)(counts + ncell + 1) + 255) & ~(uintptr_t)255);
  HIP_TRY(hipMemsetAsync(counts, 0, sizeof(int32_t) * ncell, s));
  if (h->prm.real_kind == 8)
    HIP_TRY(launch_obs_cells<double>(nlon, nlat, (const double*)lon, (const double*)lat, nobs,
                                     (const double*)olon, (const double*)olat, idx, idy, key, counts, s));
  else
    HIP_TRY(launch_obs_cells<float>(nlon, nlat, (const float*)lon, (const float*)lat, nobs, (const float*)olon,
                                    (const float*)olat, idx, idy, key, counts, s));
  if (nobs > 0) {
    hipLaunchKernelGGL(k_key_fix, dim3((nobs + 255) / 256), dim3(256), 0, s, nobs, key, vals);
    HIP_TRY(hipGetLastError());
    // LSD radix sort is stable: equal cells keep the input order, as the
    // reference's row-then-column gathering does (common_obs.f90:266-345)
    size_t tb = cub_sort;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(cub_tmp, tb, key, key2, vals, perm, nobs, 0, 32, s));
  }
  size_t tb = cub_scan;
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(cub_tmp, tb, counts, nobsgrd, ncell, s));
  hipLaunchKernelGGL(k_last_row, dim3(1), dim3(256), 0, s, nlon, nlat, nobsgrd, counts + ncell);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(nsorted, counts + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LETKF_OK;
}

extern "C" int letkf_core_batched(letkf_handle_t* h, int32_t nb, const int64_t* offs, const void* hdxf,
                                  const void* rdiag, const void* rloc, const void* dep, void* wvec, void* wmat) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  if (nb < 0 || (nb > 0 && (!offs || !wvec || !wmat))) return set_err(LETKF_ERR_ARG, "bad core_batched arguments");
  if (nb == 0) return LETKF_OK;
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const int K = h->prm.nmem;
  const double rho = h->prm.cov_infl_mul;
  h->invalidate_plan("invalidated by letkf_core_batched, which may reallocate the plan's workspaces");
  HIP_TRY(h->G.ensure(sizeof(double) * (size_t)nb * K * K));
  HIP_TRY(h->V.ensure(sizeof(double) * (size_t)nb * K * K));
  HIP_TRY(h->lam.ensure(sizeof(double) * (size_t)nb * K));
  HIP_TRY(h->bvec.ensure(sizeof(double) * (size_t)nb * K));
  HIP_TRY(h->nrow.ensure(sizeof(int32_t) * (size_t)nb));
  HIP_TRY(hipMemsetAsync(h->status.ptr, 0, sizeof(int) * 4, s));
  if (h->prm.real_kind == 8)
    HIP_TRY(launch_gram_csr<double>(K, rho, nb, offs, (const double*)hdxf, (const double*)rdiag,
                                    (const double*)rloc, (const double*)dep, h->G.as<double>(),
                                    h->bvec.as<double>(), h->nrow.as<int32_t>(), s));
  else
    HIP_TRY(launch_gram_csr<float>(K, rho, nb, offs, (const float*)hdxf, (const float*)rdiag, (const float*)rloc,
                                   (const float*)dep, h->G.as<double>(), h->bvec.as<double>(),
                                   h->nrow.as<int32_t>(), s));
  HIP_TRY(launch_eig_fp64(K, nb, h->G.as<double>(), h->lam.as<double>(), h->V.as<double>(), h->status.as<int>(),
                          0, nullptr, nullptr, 0, s));
  if (h->prm.real_kind == 8)
    HIP_TRY(launch_form_weights<double>(K, rho, nb, h->nrow.as<int32_t>(), h->lam.as<double>(), h->V.as<double>(),
                                        h->bvec.as<double>(), (double*)wvec, (double*)wmat, s));
  else
    HIP_TRY(launch_form_weights<float>(K, rho, nb, h->nrow.as<int32_t>(), h->lam.as<double>(), h->V.as<double>(),
                                       h->bvec.as<double>(), (float*)wvec, (float*)wmat, s));
  int status = 0;
  HIP_TRY(hipMemcpyAsync(&status, h->status.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return status_to_rc(status);
}
