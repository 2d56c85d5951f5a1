// prepass.hip -- the steps on either side of the analysis kernels that the
// C ABI also exposes:
//  * AOEI (common_obs.f90:389-400): sigma_o <- sqrt(max(sigma_o^2, d^2 - s^2)),
//  * obs_id_uniform (common.f90:112-168): first lon cell / lat cell with
//    min <= x < max of the two bounding grid lines, 999 when none,
//  * weight formation of letkf_core (common_letkf.f90:37-57, 139-208) for
//    letkf_core_batched: wvec = V lam^-1 V^T b, Wmat = V sqrt((K-1)/lam) V^T.
#include "letkf_internal.h"

namespace letkf {

template <typename T> __device__ __forceinline__ T sqrt_r(T x);
template <> __device__ __forceinline__ double sqrt_r<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_r<float>(float x) { return sqrtf(x); }

template <typename T>
__global__ void k_aoei(int nobs, T* err, const T* dep, const T* sprd) {
#pragma clang fp contract(off)
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nobs) return;
  const T e = err[o];
  const T d2 = dep[o] * dep[o] - sprd[o] * sprd[o];
  const T e2 = e * e;
  err[o] = sqrt_r<T>(e2 > d2 ? e2 : d2);
}

template <typename T>
hipError_t launch_aoei(int nobs, T* err, const T* dep, const T* sprd, hipStream_t s) {
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aoei<T>, dim3((nobs + 255) / 256), dim3(256), 0, s, nobs, err, dep, sprd);
  return hipGetLastError();
}
template hipError_t launch_aoei<float>(int, float*, const float*, const float*, hipStream_t);
template hipError_t launch_aoei<double>(int, double*, const double*, const double*, hipStream_t);

// One thread per obs: the reference's first-match cell search over the
// lon cells of row 1 and the lat cells of column 1 of the global grid.
// key = (idy-1)*nlon + (idx-1) for obs inside the grid, -1 otherwise.
template <typename T>
__global__ void k_obs_cells(int nlon, int nlat, const T* lon, const T* lat, int nobs, const T* olon,
                            const T* olat, int32_t* idx, int32_t* idy, int32_t* key, int32_t* counts) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nobs) return;
  const T x = olon[o], y = olat[o];
  int iy = 999, ix = 999;
  for (int c = 0; c < nlat - 1; ++c) {
    const T a = lat[(long long)nlon * c], b = lat[(long long)nlon * (c + 1)];
    const T lo = a < b ? a : b, hi = a < b ? b : a;
    if (y < lo || hi <= y) continue;
    iy = c + 1;
    break;
  }
  for (int c = 0; c < nlon - 1; ++c) {
    const T a = lon[c], b = lon[c + 1];
    const T lo = a < b ? a : b, hi = a < b ? b : a;
    if (x < lo || hi <= x) continue;
    ix = c + 1;
    break;
  }
  idx[o] = ix;
  idy[o] = iy;
  int k = -1;
  if (ix != 999 && iy != 999) {
    k = (iy - 1) * nlon + (ix - 1);
    atomicAdd(&counts[k], 1);
  }
  key[o] = k;
}

template <typename T>
hipError_t launch_obs_cells(int nlon, int nlat, const T* lon, const T* lat, int nobs, const T* olon,
                            const T* olat, int32_t* idx, int32_t* idy, int32_t* key, int32_t* counts,
                            hipStream_t s) {
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_obs_cells<T>, dim3((nobs + 255) / 256), dim3(256), 0, s, nlon, nlat, lon, lat, nobs, olon,
                     olat, idx, idy, key, counts);
  return hipGetLastError();
}
template hipError_t launch_obs_cells<float>(int, int, const float*, const float*, int, const float*,
                                            const float*, int32_t*, int32_t*, int32_t*, int32_t*, hipStream_t);
template hipError_t launch_obs_cells<double>(int, int, const double*, const double*, int, const double*,
                                             const double*, int32_t*, int32_t*, int32_t*, int32_t*, hipStream_t);

// Explicit w and W of letkf_core for one problem per workgroup.
template <typename T>
__global__ void __launch_bounds__(256) k_form_weights(int K, double sqrt_rho, double km1, const int32_t* nrow,
                                                      const double* lam, const double* V, const double* bvec,
                                                      T* wvec, T* wmat) {
  __shared__ double sq[kMaxK], ss[kMaxK], sw[kMaxK];
  const int b = blockIdx.x, t = threadIdx.x;
  T* Wb = wmat + (long long)b * K * K;
  T* wb = wvec + (long long)b * K;
  if (nrow[b] == 0) {
    for (int q = t; q < K * K; q += 256) Wb[q] = (q / K == q % K) ? (T)sqrt_rho : (T)0;
    for (int q = t; q < K; q += 256) wb[q] = (T)0;
    return;
  }
  const double* Vb = V + (long long)b * K * K;
  const double* lb = lam + (long long)b * K;
  const double* bb = bvec + (long long)b * K;
  for (int c = t; c < K; c += 256) {
    double acc = 0.0;
    for (int r = 0; r < K; ++r) acc = fma(Vb[r * K + c], bb[r], acc);
    sq[c] = acc / lb[c];
    ss[c] = sqrt(km1 / lb[c]);
  }
  __syncthreads();
  for (int r = t; r < K; r += 256) {
    double acc = 0.0;
    for (int c = 0; c < K; ++c) acc = fma(Vb[r * K + c], sq[c], acc);
    sw[r] = acc;
    wb[r] = (T)acc;
  }
  for (int q = t; q < K * K; q += 256) {
    const int i = q / K, j = q - i * K;
    double acc = 0.0;
    for (int c = 0; c < K; ++c) acc = fma(Vb[i * K + c] * ss[c], Vb[j * K + c], acc);
    Wb[q] = (T)acc;
  }
}

template <typename T>
hipError_t launch_form_weights(int K, double rho, int nb, const int32_t* nrow, const double* lam, const double* V,
                               const double* bvec, T* wvec, T* wmat, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const double sqrt_rho = (double)sqrt((T)rho);
  const double km1 = (double)(T)(K - 1);
  hipLaunchKernelGGL(k_form_weights<T>, dim3(nb), dim3(256), 0, s, K, sqrt_rho, km1, nrow, lam, V, bvec, wvec,
                     wmat);
  return hipGetLastError();
}
template hipError_t launch_form_weights<float>(int, double, int, const int32_t*, const double*, const double*,
                                               const double*, float*, float*, hipStream_t);
template hipError_t launch_form_weights<double>(int, double, int, const int32_t*, const double*, const double*,
                                                const double*, double*, double*, hipStream_t);

}  // namespace letkf
