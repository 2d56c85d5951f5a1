// output.hip -- the device half of the output step after da_main
// (write_ens_mpi, letkf/run/src/common_io.f90:863-1074):
//  * ensemble_mesp (common.f90:505-596): grid ensemble mean and spread of
//    the forecast and of the analysis;
//  * the single-precision global grid of every written field: gather_grd_mpi
//    on one rank (common_mpi.f90:133-206), the quasi-global wrap
//    (common_io.f90:960-963) and, for the file image, the vertical order of
//    write_state_vector_all (common_io.f90:655-661).  With the wrap and in
//    model order the same grids are what Trans_XtoY reads for H(x^a)
//    (common_io.f90:966-989), so letkf_trans_xtoy_device takes them as is.
// Both are streaming kernels with consecutive threads on consecutive grid
// values (coalesced): k_mesp_reg holds a grid value's members in registers
// (up to 128; k_mesp reads them twice beyond), its sums run in member order
// as in the reference so that results are bit-identical; k_grid_pack
// strides along longitude rows.
#include "letkf_handle.h"

using namespace letkf;

namespace {

template <typename T> __device__ __forceinline__ T sqrt_t(T x);
template <> __device__ __forceinline__ double sqrt_t<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_t<float>(float x) { return sqrtf(x); }

// Row r < nv3d nlev is level r % nlev of 3-D variable r / nlev, the rest are
// the 2-D variables.  Loads go out eight members at a time; the sums take
// them in member order.
template <typename T>
__global__ void __launch_bounds__(256) k_mesp(int nij1, int nlev, int nmem, int nv3d, int nv2d, T undef,
                                              const T* __restrict__ v3d, const T* __restrict__ v2d,
                                              T* __restrict__ v3dm, T* __restrict__ v3ds, T* __restrict__ v2dm,
                                              T* __restrict__ v2ds) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nij1) return;
  const long long ni = nij1;
  const int rows = nv3d * nlev + nv2d;
  for (int r = blockIdx.y; r < rows; r += gridDim.y) {
    const T* src;
    long long ms;
    T *om, *os;
    if (r < nv3d * nlev) {
      const int v = r / nlev, k = r - v * nlev;
      src = v3d + ni * (k + (long long)nlev * nmem * v) + i;
      ms = ni * nlev;
      om = v3dm + ni * (k + (long long)nlev * v) + i;
      os = v3ds + ni * (k + (long long)nlev * v) + i;
    } else {
      const int v = r - nv3d * nlev;
      src = v2d + ni * nmem * v + i;
      ms = ni;
      om = v2dm + ni * v + i;
      os = v2ds + ni * v + i;
    }
    T s = (T)0;
    int m = 0;
    for (; m + 8 <= nmem; m += 8) {
      T x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = src[ms * (m + u)];
#pragma unroll
      for (int u = 0; u < 8; ++u) s = s + x[u];
    }
    for (; m < nmem; ++m) s = s + src[ms * m];
    s = s / (T)nmem;
    *om = s;
    T q = (T)0;
    m = 0;
    for (; m + 8 <= nmem; m += 8) {
      T x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = src[ms * (m + u)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const T d = x[u] - s;
        q = q + d * d;
      }
    }
    for (; m < nmem; ++m) {
      const T d = src[ms * m] - s;
      q = q + d * d;
    }
    *os = nmem <= 1 ? undef : sqrt_t<T>(q / (T)(nmem - 1));
  }
}

// nmem <= NM: the members are loaded once into registers (all loads in
// flight together) and both sums run over the registers; same operations in
// the same order as k_mesp.
template <typename T, int NM>
__global__ void __launch_bounds__(256) k_mesp_reg(int nij1, int nlev, int nmem, int nv3d, int nv2d, T undef,
                                                  const T* __restrict__ v3d, const T* __restrict__ v2d,
                                                  T* __restrict__ v3dm, T* __restrict__ v3ds, T* __restrict__ v2dm,
                                                  T* __restrict__ v2ds) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nij1) return;
  const long long ni = nij1;
  const int rows = nv3d * nlev + nv2d;
  for (int r = blockIdx.y; r < rows; r += gridDim.y) {
    const T* src;
    long long ms;
    T *om, *os;
    if (r < nv3d * nlev) {
      const int v = r / nlev, k = r - v * nlev;
      src = v3d + ni * (k + (long long)nlev * nmem * v) + i;
      ms = ni * nlev;
      om = v3dm + ni * (k + (long long)nlev * v) + i;
      os = v3ds + ni * (k + (long long)nlev * v) + i;
    } else {
      const int v = r - nv3d * nlev;
      src = v2d + ni * nmem * v + i;
      ms = ni;
      om = v2dm + ni * v + i;
      os = v2ds + ni * v + i;
    }
    T x[NM];
#pragma unroll
    for (int u = 0; u < NM; ++u) x[u] = u < nmem ? src[ms * u] : (T)0;
    T s = (T)0;
#pragma unroll
    for (int u = 0; u < NM; ++u)
      if (u < nmem) s = s + x[u];
    s = s / (T)nmem;
    *om = s;
    T q = (T)0;
#pragma unroll
    for (int u = 0; u < NM; ++u)
      if (u < nmem) {
        const T d = x[u] - s;
        q = q + d * d;
      }
    *os = nmem <= 1 ? undef : sqrt_t<T>(q / (T)(nmem - 1));
  }
}

template <typename T>
void launch_mesp(dim3 grid, hipStream_t s, int nij1, int nlev, int nmem, int nv3d, int nv2d, T undef, const T* v3d,
                 const T* v2d, T* v3dm, T* v3ds, T* v2dm, T* v2ds) {
#define MESP_ARGS s, nij1, nlev, nmem, nv3d, nv2d, undef, v3d, v2d, v3dm, v3ds, v2dm, v2ds
  if (nmem <= 16)
    hipLaunchKernelGGL((k_mesp_reg<T, 16>), grid, dim3(256), 0, MESP_ARGS);
  else if (nmem <= 32)
    hipLaunchKernelGGL((k_mesp_reg<T, 32>), grid, dim3(256), 0, MESP_ARGS);
  else if (nmem <= 64)
    hipLaunchKernelGGL((k_mesp_reg<T, 64>), grid, dim3(256), 0, MESP_ARGS);
  else if (nmem <= 104)  // two wavefronts per SIMD in FP64 (the 128 form holds one)
    hipLaunchKernelGGL((k_mesp_reg<T, 104>), grid, dim3(256), 0, MESP_ARGS);
  else if (nmem <= 128)
    hipLaunchKernelGGL((k_mesp_reg<T, 128>), grid, dim3(256), 0, MESP_ARGS);
  else
    hipLaunchKernelGGL(k_mesp<T>, grid, dim3(256), 0, MESP_ARGS);
#undef MESP_ARGS
}

struct PackArgs {
  int nlon, nlat, nlev, nv3d, nv2d, nfield, wrap, file_order;
  long long f3, s3v, s3k, f2, s2v;
};

// Output row r of field f: rows < nv3d nlev are (variable, output level),
// the rest 2-D variables.  A workgroup covers kPackRows latitude rows of
// one output row, its threads striding along longitude (no division per
// value); value (ilon, ilat) reads local value (ilon, ilat), or (ilon -
// (nlon - 2), ilat) in the last two columns under the wrap.
constexpr int kPackRows = 4;

template <typename T>
__global__ void __launch_bounds__(256) k_grid_pack(PackArgs a, const T* __restrict__ v3d, const T* __restrict__ v2d,
                                                   float* __restrict__ g3, float* __restrict__ g2) {
  const long long nij = (long long)a.nlon * a.nlat;
  const int per = a.nv3d * a.nlev + a.nv2d;
  const long long rows = (long long)per * a.nfield;
  const int j0 = blockIdx.x * kPackRows;
  const int j1 = min(j0 + kPackRows, a.nlat);
  const int wcol = a.wrap ? a.nlon - 2 : a.nlon;  // first wrapped column
  for (long long rr = blockIdx.y; rr < rows; rr += gridDim.y) {
    const int f = (int)(rr / per), r = (int)(rr - (long long)f * per);
    const T* src;
    float* dst;
    if (r < a.nv3d * a.nlev) {
      const int v = r / a.nlev, ko = r - v * a.nlev;
      const int k = a.file_order ? a.nlev - 1 - ko : ko;
      src = v3d + a.f3 * f + a.s3v * v + a.s3k * k;
      dst = g3 + nij * ((long long)a.nlev * a.nv3d * f + (long long)a.nlev * v + ko);
    } else {
      const int v = r - a.nv3d * a.nlev;
      src = v2d + a.f2 * f + a.s2v * v;
      dst = g2 + nij * ((long long)a.nv2d * f + v);
    }
    for (int j = j0; j < j1; ++j) {
      const long long row = (long long)a.nlon * j;
      for (int ilon = threadIdx.x; ilon < a.nlon; ilon += 256) {
        const int sc = ilon >= wcol ? ilon - wcol : ilon;
        dst[row + ilon] = (float)src[row + sc];
      }
    }
  }
}

hipError_t finish(hipStream_t s) {
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(s);
  return e != hipSuccess ? e : e2;
}

}  // namespace

extern "C" int letkf_ensemble_mesp_device(int32_t real_kind, int32_t nij1, int32_t nlev, int32_t nmem,
                                          int32_t nv3d, int32_t nv2d, double undef, const void* v3d,
                                          const void* v2d, void* v3dm, void* v3ds, void* v2dm, void* v2ds,
                                          void* stream) {
  if (real_kind != 4 && real_kind != 8)
    return set_err(LETKF_ERR_ARG, "letkf_ensemble_mesp_device: real_kind must be 4 or 8 (got %d)", real_kind);
  if (nij1 < 0 || nlev < 0 || nmem < 1 || nv3d < 0 || nv2d < 0)
    return set_err(LETKF_ERR_ARG, "letkf_ensemble_mesp_device: bad sizes nij1 %d nlev %d nmem %d nv3d %d nv2d %d",
                   nij1, nlev, nmem, nv3d, nv2d);
  const int rows = nv3d * nlev + nv2d;
  if (nij1 == 0 || rows == 0) return LETKF_OK;
  if ((nv3d * nlev > 0 && (!v3d || !v3dm || !v3ds)) || (nv2d > 0 && (!v2d || !v2dm || !v2ds)))
    return set_err(LETKF_ERR_ARG, "letkf_ensemble_mesp_device: NULL array");
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)((nij1 + 255) / 256), (unsigned)std::min(rows, 65535));
  if (real_kind == 8)
    launch_mesp<double>(grid, s, nij1, nlev, nmem, nv3d, nv2d, undef, (const double*)v3d, (const double*)v2d,
                        (double*)v3dm, (double*)v3ds, (double*)v2dm, (double*)v2ds);
  else
    launch_mesp<float>(grid, s, nij1, nlev, nmem, nv3d, nv2d, (float)undef, (const float*)v3d, (const float*)v2d,
                       (float*)v3dm, (float*)v3ds, (float*)v2dm, (float*)v2ds);
  const hipError_t e = finish(s);
  if (e != hipSuccess) return set_err(LETKF_ERR_HIP, "letkf_ensemble_mesp_device: %s", hipGetErrorString(e));
  return LETKF_OK;
}

extern "C" int letkf_grid_pack_device(int32_t real_kind, int32_t nlon, int32_t nlat, int32_t nlev, int32_t nv3d,
                                      int32_t nv2d, int32_t nfield, int32_t wrap, int32_t file_order,
                                      const void* v3d, const int64_t* s3, const void* v2d, const int64_t* s2,
                                      float* g3, float* g2, void* stream) {
  if (real_kind != 4 && real_kind != 8)
    return set_err(LETKF_ERR_ARG, "letkf_grid_pack_device: real_kind must be 4 or 8 (got %d)", real_kind);
  if (nlon < 1 || nlat < 1 || nlev < 0 || nv3d < 0 || nv2d < 0 || nfield < 0 || (wrap && nlon < 4))
    return set_err(LETKF_ERR_ARG,
                   "letkf_grid_pack_device: bad sizes nlon %d nlat %d nlev %d nv3d %d nv2d %d nfield %d "
                   "(the wrap needs nlon >= 4)",
                   nlon, nlat, nlev, nv3d, nv2d, nfield);
  const long long rows = (long long)(nv3d * nlev + nv2d) * nfield;
  if (rows == 0) return LETKF_OK;
  if ((nv3d * nlev > 0 && (!v3d || !s3 || !g3)) || (nv2d > 0 && (!v2d || !s2 || !g2)))
    return set_err(LETKF_ERR_ARG, "letkf_grid_pack_device: NULL array or strides");
  PackArgs a{nlon, nlat, nlev, nv3d, nv2d, nfield, wrap ? 1 : 0, file_order ? 1 : 0,
             s3 ? s3[0] : 0, s3 ? s3[1] : 0, s3 ? s3[2] : 0, s2 ? s2[0] : 0, s2 ? s2[1] : 0};
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)((nlat + kPackRows - 1) / kPackRows), (unsigned)std::min<long long>(rows, 65535));
  if (real_kind == 8)
    hipLaunchKernelGGL(k_grid_pack<double>, grid, dim3(256), 0, s, a, (const double*)v3d, (const double*)v2d, g3,
                       g2);
  else
    hipLaunchKernelGGL(k_grid_pack<float>, grid, dim3(256), 0, s, a, (const float*)v3d, (const float*)v2d, g3, g2);
  const hipError_t e = finish(s);
  if (e != hipSuccess) return set_err(LETKF_ERR_HIP, "letkf_grid_pack_device: %s", hipGetErrorString(e));
  return LETKF_OK;
}
