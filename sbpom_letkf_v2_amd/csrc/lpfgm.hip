// lpfgm.hip -- the local particle filter and its Gaussian-mixture extension
// (letkf/run/src/common_lpfgm.f90), batched over local problems like
// letkf_core_batched: letkf_lpfgm_core_batched.
//
// One workgroup per problem:
//  * particle weights (pf_weight_gauss_likelihood, :38-77): lane j sums
//    -0.5 (d_i - y_ij)^2 rloc_i / rdiag_i over the problem's rows in row
//    order in FP64; max, exp, the normalising sum and the cumulative
//    weights in member order (one lane);
//  * the nmonte resamplings (resampling_matrix, :89-197): lane m replays
//    draw set m; the draws arrive sorted (MR) or in the reference's SR form,
//    so the particle picked per column (first i with draw <= cw_i) comes
//    from one forward walk over cw, and diagonal priority needs only a
//    128-bit column mask: a particle's first pick takes its own column, the
//    repeats take the free columns in ascending order.  Picks are counted
//    per (particle, column) in LDS bytes; the average of the nmonte 0/1
//    matrices is a table lookup of the count (the running sum of 1/nmonte in
//    the real kind, as the reference accumulates it);
//  * LPFGM (lpfgm_core, :444-494): W_GM (gm_core, :279-405) in its closed
//    form alpha Pa + (Pa b) 1^T, Pa = C^-1, C = Y^T R^-1 Y + alpha I
//    (alpha = (m-1)/cov_infl_gm; the Gram of letkf_core_batched).  The
//    reference forms Pa from mtx_eigen; C >= alpha I is well conditioned, so
//    the workgroup inverts it through its Cholesky factor in LDS instead (the
//    same Pa to FP64 rounding, a tenth of the eigensolve's time).  Then
//    W = W_GM W_LPF, W_GM rows staged through the same LDS.
#include <algorithm>
#include <vector>

#include "letkf_handle.h"

using namespace letkf;

namespace {

constexpr int kLpfMaxK = 128;
constexpr int kLpfRows = 16;  // W_GM rows staged per pass

template <typename T>
__global__ void __launch_bounds__(256) k_lpfgm(int K, int mode, int nmonte, int dp, double alpha,
                                               const int64_t* __restrict__ offs, const T* __restrict__ hdxf,
                                               const T* __restrict__ rdiag, const T* __restrict__ rloc,
                                               const T* __restrict__ dep, const double* __restrict__ rtab,
                                               const double* __restrict__ ttab, const double* __restrict__ G,
                                               const double* __restrict__ bvec, T* __restrict__ wvec,
                                               T* __restrict__ wmat) {
#pragma clang fp contract(off)
  __shared__ double ll[kLpfMaxK], cw[kLpfMaxK], sq[kLpfMaxK], wb[kLpfMaxK], il[kLpfMaxK];
  __shared__ uint32_t cnt[kLpfMaxK * kLpfMaxK / 4];  // byte count per (particle, column)
  __shared__ double tt[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const long long o0 = offs[b], o1 = offs[b + 1];
  T* Wb = wmat + (long long)b * K * K;
  if (t < K) wvec[(long long)b * K + t] = (T)0;
  if (o1 <= o0) {  // lpfgm_core_noobs: W = I
    for (int q = t; q < K * K; q += 256) Wb[q] = (q / K == q % K) ? (T)1 : (T)0;
    return;
  }
  for (int q = t; q < K * K / 4 + 1 && q < kLpfMaxK * kLpfMaxK / 4; q += 256) cnt[q] = 0u;
  for (int q = t; q <= nmonte; q += 256) tt[q] = ttab[q];
  // ---- log-likelihood per particle, rows in order ----
  if (t < K) {
    double s = 0.0;
    for (long long r = o0; r < o1; ++r) {
      const T dx = dep[r] - hdxf[r * K + t];
      double x = 0.5 * (double)dx;
      x = x * (double)dx;
      x = x * (double)rloc[r];
      x = x / (double)rdiag[r];
      s = s - x;
    }
    ll[t] = s;
  }
  __syncthreads();
  if (t == 0) {
    double mx = ll[0];
    for (int j = 1; j < K; ++j) mx = ll[j] > mx ? ll[j] : mx;
    double se = 0.0;
    for (int j = 0; j < K; ++j) {
      const double e = exp(ll[j] - mx);
      cw[j] = e;
      se = se + e;
    }
    double c = cw[0] / se;
    cw[0] = c;
    for (int j = 1; j < K; ++j) {
      c = c + cw[j] / se;
      cw[j] = c;
    }
  }
  __syncthreads();
  // ---- resamplings: lane m replays draw set m ----
  if (t < nmonte) {
    const double* r = rtab + (long long)t * K;
    if (dp) {
      uint64_t u0 = 0, u1 = 0;  // columns taken
      int p = 0, prev = -1;
      for (int i = 0; i < K; ++i) {  // first picks: own column
        const double x = r[i];
        while (p < K - 1 && !(x <= cw[p])) ++p;
        if (p != prev) {
          if (p < 64) u0 |= 1ull << p;
          else u1 |= 1ull << (p - 64);
          const int q = p * K + p;
          atomicAdd(&cnt[q >> 2], 1u << (8 * (q & 3)));
          prev = p;
        }
      }
      int col = 0;
      p = 0;
      prev = -1;
      for (int i = 0; i < K; ++i) {  // repeats: free columns in ascending order
        const double x = r[i];
        while (p < K - 1 && !(x <= cw[p])) ++p;
        if (p == prev) {
          while (col < K && (col < 64 ? (u0 >> col) & 1 : (u1 >> (col - 64)) & 1)) ++col;
          if (col < K) {
            if (col < 64) u0 |= 1ull << col;
            else u1 |= 1ull << (col - 64);
            const int q = p * K + col;
            atomicAdd(&cnt[q >> 2], 1u << (8 * (q & 3)));
          }
        }
        prev = p;
      }
    } else {
      for (int j = 0; j < K; ++j) {  // column j takes the first i with draw <= cw_i
        const double x = r[j];
        int p = 0;
        while (p < K - 1 && !(x <= cw[p])) ++p;
        const int q = p * K + j;
        atomicAdd(&cnt[q >> 2], 1u << (8 * (q & 3)));
      }
    }
  }
  __syncthreads();
  auto wlpf = [&](int i, int k) -> double {
    const int q = i * K + k;
    return tt[(cnt[q >> 2] >> (8 * (q & 3))) & 0xffu];
  };
  if (mode == 2) {
    for (int q = t; q < K * K; q += 256) Wb[q] = (T)wlpf(q / K, q % K);
    return;
  }
  // ---- W_GM = alpha Pa + (Pa b) 1^T with Pa = C^-1 = L^-T L^-1 ----
  // C (upper triangle in G) -> LDS, lower Cholesky factor L in place
  extern __shared__ double cm[];  // K x K, row pitch K
  const double* Gb = G + (long long)b * K * K;
  for (int q = t; q < K * K; q += 256) {
    const int i = q / K, j = q - i * K;
    if (j <= i) cm[q] = Gb[j * K + i];
  }
  __syncthreads();
  for (int j = 0; j < K; ++j) {
    if (t == 0) {
      const double d = cm[j * K + j];
      cm[j * K + j] = d > 0.0 ? sqrt(d) : NAN;
    }
    __syncthreads();
    const double ljj = cm[j * K + j];
    for (int i = j + 1 + t; i < K; i += 256) cm[i * K + j] /= ljj;
    __syncthreads();
    const int m = K - j - 1;  // trailing block rows/cols j+1..K-1, lower part
    for (int q = t; q < m * m; q += 256) {
      const int i = j + 1 + q / m, k = j + 1 + q % m;
      if (k <= i) cm[i * K + k] -= cm[i * K + j] * cm[k * K + j];
    }
    __syncthreads();
  }
  // L^-1: column j by lane j, stored transposed in the upper triangle
  // (Linv(i,j) at cm[j*K + i], i > j); its diagonal in il
  if (t < K) il[t] = 1.0 / cm[t * K + t];
  __syncthreads();
  if (t < K) {
    const int j = t;
    for (int i = j + 1; i < K; ++i) {
      double acc = cm[i * K + j] * il[j];
      for (int k = j + 1; k < i; ++k) acc = fma(cm[i * K + k], cm[j * K + k], acc);
      cm[j * K + i] = -acc * il[i];
    }
  }
  __syncthreads();
  auto linv = [&](int i, int j) -> double { return i == j ? il[i] : cm[j * K + i]; };  // i >= j
  // w = Pa b = L^-T (L^-1 b)
  const double* bb = bvec + (long long)b * K;
  if (t < K) {
    double acc = 0.0;
    for (int m2 = 0; m2 <= t; ++m2) acc = fma(linv(t, m2), bb[m2], acc);
    sq[t] = acc;
  }
  __syncthreads();
  if (t < K) {
    double acc = 0.0;
    for (int k = t; k < K; ++k) acc = fma(linv(k, t), sq[k], acc);
    wb[t] = acc;
  }
  __syncthreads();
  for (int q = t; q < K * K; q += 256) {
    const int i = q / K, j = q - i * K;
    const int k0 = i > j ? i : j;
    double acc = 0.0;
    for (int k = k0; k < K; ++k) acc = fma(linv(k, i), linv(k, j), acc);
    Wb[q] = (T)(alpha * acc + wb[i]);
  }
  __syncthreads();
  // ---- W = W_GM W_LPF, kLpfRows rows of W_GM at a time (staged in cm) ----
  for (int i0 = 0; i0 < K; i0 += kLpfRows) {
    const int nr = min(kLpfRows, K - i0);
    for (int q = t; q < nr * K; q += 256) cm[q] = (double)Wb[(long long)i0 * K + q];
    __syncthreads();
    for (int q = t; q < nr * K; q += 256) {
      const int i = q / K, k = q - i * K;
      double acc = 0.0;
      for (int j = 0; j < K; ++j) acc = fma(cm[i * K + j], wlpf(j, k), acc);
      Wb[(long long)(i0 + i) * K + k] = (T)acc;
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" int letkf_lpfgm_core_batched(letkf_handle_t* h, int32_t mode, double cov_infl_gm, int32_t nmonte,
                                        int32_t rm_sr, int32_t dp, const double* rand, int32_t nb,
                                        const int64_t* offs, const void* hdxf, const void* rdiag, const void* rloc,
                                        const void* dep, void* wvec, void* wmat) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  const int K = h->prm.nmem;
  if (mode != 2 && mode != 3) return set_err(LETKF_ERR_ARG, "lpfgm mode must be 2 (LPF) or 3 (LPFGM), got %d", mode);
  if (K < 1 || K > kLpfMaxK) return set_err(LETKF_ERR_ARG, "lpfgm_core_batched needs 1 <= nmem <= %d", kLpfMaxK);
  if (nmonte < 1 || nmonte > 255) return set_err(LETKF_ERR_ARG, "nmonte must be 1..255 (got %d)", nmonte);
  if (mode == 3 && !(cov_infl_gm > 0.0)) return set_err(LETKF_ERR_ARG, "cov_infl_gm must be > 0");
  if (nb < 0 || (nb > 0 && (!offs || !wvec || !wmat || !rand)))
    return set_err(LETKF_ERR_ARG, "bad lpfgm_core_batched arguments");
  if (nb == 0) return LETKF_OK;
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const bool f64 = h->prm.real_kind == 8;
  // the draws of each resampling as resampling_matrix uses them (:111-128)
  std::vector<double> rt((size_t)nmonte * K);
  for (int m = 0; m < nmonte; ++m) {
    double* r = rt.data() + (size_t)m * K;
    std::copy(rand + (size_t)m * K, rand + (size_t)(m + 1) * K, r);
    if (rm_sr) {
      for (int i = 0; i < K; ++i) r[i] = ((r[0] + (double)(i + 1)) - 1.0) / (double)K;
    } else {
      std::sort(r, r + K);
    }
  }
  // the running average Wmat + RSmat / dble(nmonte) in the real kind (:259)
  std::vector<double> tt((size_t)nmonte + 1);
  tt[0] = 0.0;
  for (int c = 1; c <= nmonte; ++c)
    tt[c] = f64 ? tt[c - 1] + 1.0 / (double)nmonte : (double)(float)(tt[c - 1] + 1.0 / (double)nmonte);
  HIP_TRY(h->bin_tmp.ensure(sizeof(double) * (rt.size() + tt.size())));
  double* d_rt = h->bin_tmp.as<double>();
  double* d_tt = d_rt + rt.size();
  HIP_TRY(hipMemcpyAsync(d_rt, rt.data(), sizeof(double) * rt.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d_tt, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice, s));
  const double* Gp = nullptr;
  const double* bv = nullptr;
  double alpha = 0.0;
  if (mode == 3) {
    HIP_TRY(h->G.ensure(sizeof(double) * (size_t)nb * K * K));
    HIP_TRY(h->bvec.ensure(sizeof(double) * (size_t)nb * K));
    HIP_TRY(h->nrow.ensure(sizeof(int32_t) * (size_t)nb));
    if (f64)
      HIP_TRY(launch_gram_csr<double>(K, cov_infl_gm, nb, offs, (const double*)hdxf, (const double*)rdiag,
                                      (const double*)rloc, (const double*)dep, h->G.as<double>(),
                                      h->bvec.as<double>(), h->nrow.as<int32_t>(), s));
    else
      HIP_TRY(launch_gram_csr<float>(K, cov_infl_gm, nb, offs, (const float*)hdxf, (const float*)rdiag,
                                     (const float*)rloc, (const float*)dep, h->G.as<double>(), h->bvec.as<double>(),
                                     h->nrow.as<int32_t>(), s));
    Gp = h->G.as<double>();
    bv = h->bvec.as<double>();
    alpha = f64 ? (double)(K - 1) / cov_infl_gm : (double)((float)(K - 1) / (float)cov_infl_gm);
  }
  const size_t lds = mode == 3 ? sizeof(double) * (size_t)K * K : 0;
  if (lds > 65536) {
    if (f64)
      HIP_TRY(hipFuncSetAttribute((const void*)k_lpfgm<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    else
      HIP_TRY(hipFuncSetAttribute((const void*)k_lpfgm<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  if (f64)
    hipLaunchKernelGGL(k_lpfgm<double>, dim3(nb), dim3(256), lds, s, K, mode, nmonte, dp ? 1 : 0, alpha, offs,
                       (const double*)hdxf, (const double*)rdiag, (const double*)rloc, (const double*)dep, d_rt, d_tt,
                       Gp, bv, (double*)wvec, (double*)wmat);
  else
    hipLaunchKernelGGL(k_lpfgm<float>, dim3(nb), dim3(256), lds, s, K, mode, nmonte, dp ? 1 : 0, alpha, offs,
                       (const float*)hdxf, (const float*)rdiag, (const float*)rloc, (const float*)dep, d_rt, d_tt,
                       Gp, bv, (float*)wvec, (float*)wmat);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return LETKF_OK;
}
