// gram.hip -- the ensemble-space Gram products of letkf_core
// (letkf/run/src/common_letkf.f90:93-126) for a batch of points, in FP64:
//   Rlinv_dYf = hdxf * rloc / rdiag                 (:93-97)
//   G = Rlinv_dYf^T hdxf + (K-1)/rho I              (:112-126)
//   b = Rlinv_dYf^T dep                              (the right-hand side of
//       wvec = Pa (Rlinv_dYf^T dep), an exact re-association of :162-184)
// One workgroup per point; the local obs rows are gathered from the
// obs-major FP64 copy of obshdxf through LDS in panels of 2048/KT obs.
#include "letkf_internal.h"

namespace letkf {


// KT: padded ensemble tile (64 or 128); each of the 16x16 threads owns a
// TT x TT block of G with TT = KT/16.
template <typename T, int KT>
__global__ void __launch_bounds__(256) k_gram_fp64(KParams kp, Tables<T> tb, const double* __restrict__ yT,
                                                   const int32_t* __restrict__ act_p,
                                                   const int64_t* __restrict__ loff, long long lbase,
                                                   const int32_t* __restrict__ lobs,
                                                   const double* __restrict__ lrloc, int a0,
                                                   double* __restrict__ G, double* __restrict__ bvec) {
  constexpr int TT = KT / 16;
  constexpr int kPanel = 2048 / KT;  // 32 obs at KT=64, 16 at KT=128 (33 KiB LDS)
  __shared__ double As[kPanel][KT + 1];  // Rlinv_dYf rows
  __shared__ double Bs[kPanel][KT + 1];  // hdxf rows
  __shared__ double ds[kPanel];
  const int K = kp.K;
  const int a = a0 + blockIdx.x;
  const int p = act_p[a];
  const long long o0 = loff[p] - lbase;
  const int n = (int)(loff[p + 1] - loff[p]);
  const int t = threadIdx.x;
  const int tr = t >> 4, tc = t & 15;
  double acc[TT][TT];
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y) acc[x][y] = 0.0;
  double bacc = 0.0;
  for (int l0 = 0; l0 < n; l0 += kPanel) {
    const int nb = min(kPanel, n - l0);
    // stage panel: thread t loads member m = t % KT of obs rows t / KT, ...
    for (int q = t; q < kPanel * KT; q += 256) {
      const int l = q / KT, m = q - l * KT;
      double av = 0.0, bv = 0.0;
      if (l < nb && m < K) {
        const int io = lobs[o0 + l0 + l];
        const double rl = lrloc[o0 + l0 + l];
        const T er = tb.oerr[io];
        const double rd = (double)(T)(er * er);  // rdiag = REAL(obserr**2, r_size)
        bv = yT[(long long)io * K + m];
        av = bv * rl / rd;
      }
      As[l][m] = av;
      Bs[l][m] = bv;
    }
    if (t < kPanel) ds[t] = t < nb ? (double)tb.odep[lobs[o0 + l0 + t]] : 0.0;
    __syncthreads();
    for (int l = 0; l < nb; ++l) {
      double ar[TT], bc[TT];
#pragma unroll
      for (int x = 0; x < TT; ++x) ar[x] = As[l][tr * TT + x];
#pragma unroll
      for (int y = 0; y < TT; ++y) bc[y] = Bs[l][tc * TT + y];
#pragma unroll
      for (int x = 0; x < TT; ++x)
#pragma unroll
        for (int y = 0; y < TT; ++y) acc[x][y] = fma(ar[x], bc[y], acc[x][y]);
    }
    if (t < K)
      for (int l = 0; l < nb; ++l) bacc = fma(As[l][t], ds[l], bacc);
    __syncthreads();
  }
  const double diag = (double)((T)(K - 1) / (T)kp.rho);
  double* Gp = G + (long long)blockIdx.x * K * K;
#pragma unroll
  for (int x = 0; x < TT; ++x) {
    const int r = tr * TT + x;
    if (r >= K) continue;
#pragma unroll
    for (int y = 0; y < TT; ++y) {
      const int c = tc * TT + y;
      if (c < K) Gp[r * K + c] = acc[x][y] + (r == c ? diag : 0.0);
    }
  }
  if (t < K) bvec[(long long)blockIdx.x * K + t] = bacc;
}

template <typename T>
hipError_t launch_gram_fp64(const KParams& kp, const Tables<T>& tb, const double* yT, const int32_t* act_p,
                            const int64_t* loff, long long lbase, const int32_t* lobs, const double* lrloc,
                            int a0, int na, double* G, double* bvec, hipStream_t s) {
  if (na <= 0) return hipSuccess;
  if (kp.K <= 64)
    hipLaunchKernelGGL((k_gram_fp64<T, 64>), dim3(na), dim3(256), 0, s, kp, tb, yT, act_p, loff, lbase,
                       lobs, lrloc, a0, G, bvec);
  else
    hipLaunchKernelGGL((k_gram_fp64<T, 128>), dim3(na), dim3(256), 0, s, kp, tb, yT, act_p, loff, lbase,
                       lobs, lrloc, a0, G, bvec);
  return hipGetLastError();
}

template hipError_t launch_gram_fp64<float>(const KParams&, const Tables<float>&, const double*,
                                            const int32_t*, const int64_t*, long long, const int32_t*,
                                            const double*, int, int, double*, double*, hipStream_t);
template hipError_t launch_gram_fp64<double>(const KParams&, const Tables<double>&, const double*,
                                             const int32_t*, const int64_t*, long long, const int32_t*,
                                             const double*, int, int, double*, double*, hipStream_t);

}  // namespace letkf
