// lpfgm.hip -- the local particle filter and its Gaussian-mixture extension
// (letkf/run/src/common_lpfgm.f90): letkf_lpfgm_core_batched over local
// problems given in CSR form (k_lpfgm), and the per-point kernel of da_main
// with iswitch_da 2/3 (k_lpfgm_points), which reads a point's local obs
// through the search's lists and then updates its ensemble.
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

struct LpfShared {
  double ll[kLpfMaxK], cw[kLpfMaxK], sq[kLpfMaxK], wb[kLpfMaxK], il[kLpfMaxK];
  double tt[256];
};

// Dynamic LDS of the LPF kernels: [K*K doubles for the Cholesky work (mode
// 3)] then the pick counts, one byte per (particle, column), so that small
// ensembles fit more workgroups per CU.
__host__ __device__ inline size_t lpf_dyn_bytes(int K, int mode) {
  return (mode == 3 ? sizeof(double) * (size_t)K * K : 0) + sizeof(uint32_t) * ((size_t)K * K / 4 + 1);
}

// Particle weights and the pick counts of the nmonte resamplings of one
// problem with n rows: rows.dx(r, j) = dep - hdxf of row r and member j (in
// the real kind), rows.rloc(r), rows.rdiag(r).  Ends with a barrier.
template <typename Rows>
__device__ void lpf_counts(LpfShared& sh, uint32_t* cnt, int K, int n, int nmonte, int dp, const Rows& rows,
                           const double* __restrict__ rtab, const double* __restrict__ ttab) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  for (int q = t; q < K * K / 4 + 1; q += 256) cnt[q] = 0u;
  for (int q = t; q <= nmonte; q += 256) sh.tt[q] = ttab[q];
  if (t < K) {  // log-likelihood per particle, rows in order, 8 rows of loads in flight
    double s = 0.0;
    for (int r0 = 0; r0 < n; r0 += 8) {
      decltype(rows.dx(0, 0)) dx[8];
      double rl[8], rd[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = r0 + u < n ? r0 + u : r0;
        dx[u] = rows.dx(r, t);
        rl[u] = rows.rloc(r);
        rd[u] = rows.rdiag(r);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (r0 + u >= n) break;
        double x = 0.5 * (double)dx[u];
        x = x * (double)dx[u];
        x = x * rl[u];
        x = x / rd[u];
        s = s - x;
      }
    }
    sh.ll[t] = s;
  }
  __syncthreads();
  if (t == 0) {
    double mx = sh.ll[0];
    for (int j = 1; j < K; ++j) mx = sh.ll[j] > mx ? sh.ll[j] : mx;
    double se = 0.0;
    for (int j = 0; j < K; ++j) {
      const double e = exp(sh.ll[j] - mx);
      sh.cw[j] = e;
      se = se + e;
    }
    double c = sh.cw[0] / se;
    sh.cw[0] = c;
    for (int j = 1; j < K; ++j) {
      c = c + sh.cw[j] / se;
      sh.cw[j] = c;
    }
  }
  __syncthreads();
  if (t < nmonte) {  // lane m replays draw set m
    const double* r = rtab + (long long)t * K;
    if (dp) {
      uint64_t u0 = 0, u1 = 0;  // columns taken
      int p = 0, prev = -1;
      for (int i = 0; i < K; ++i) {  // first picks: own column
        const double x = r[i];
        while (p < K - 1 && !(x <= sh.cw[p])) ++p;
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
        while (p < K - 1 && !(x <= sh.cw[p])) ++p;
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
        while (p < K - 1 && !(x <= sh.cw[p])) ++p;
        const int q = p * K + j;
        atomicAdd(&cnt[q >> 2], 1u << (8 * (q & 3)));
      }
    }
  }
  __syncthreads();
}

// Entry (i, k) of the averaged resampling matrix W_LPF.
__device__ __forceinline__ double lpf_w(const LpfShared& sh, const uint32_t* cnt, int K, int i, int k) {
  const int q = i * K + k;
  return sh.tt[(cnt[q >> 2] >> (8 * (q & 3))) & 0xffu];
}

// W = W_GM W_LPF into Wb (K x K, row-major), W_GM = alpha Pa + (Pa b) 1^T
// with Pa = C^-1 = L^-T L^-1 from the Cholesky factor of C (upper triangle
// in Gb, pitch K) in LDS cm.  Wb may alias Gb.  Ends with a barrier.
template <typename WT>
__device__ void gm_times_lpf(LpfShared& sh, const uint32_t* cnt, double* cm, int K, double alpha, const double* Gb,
                             const double* __restrict__ bb, WT* Wb) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
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
  if (t < K) sh.il[t] = 1.0 / cm[t * K + t];
  __syncthreads();
  if (t < K) {  // four partial sums per dot product: LDS latency, not the chain, sets the pace
    const int j = t;
    for (int i = j + 1; i < K; ++i) {
      double a0 = cm[i * K + j] * sh.il[j], a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int k = j + 1;
      for (; k + 4 <= i; k += 4) {
        a0 = fma(cm[i * K + k], cm[j * K + k], a0);
        a1 = fma(cm[i * K + k + 1], cm[j * K + k + 1], a1);
        a2 = fma(cm[i * K + k + 2], cm[j * K + k + 2], a2);
        a3 = fma(cm[i * K + k + 3], cm[j * K + k + 3], a3);
      }
      for (; k < i; ++k) a0 = fma(cm[i * K + k], cm[j * K + k], a0);
      cm[j * K + i] = -((a0 + a1) + (a2 + a3)) * sh.il[i];
    }
  }
  __syncthreads();
  auto linv = [&](int i, int j) -> double { return i == j ? sh.il[i] : cm[j * K + i]; };  // i >= j
  if (t < K) {  // w = Pa b = L^-T (L^-1 b)
    double acc = 0.0;
    for (int m2 = 0; m2 <= t; ++m2) acc = fma(linv(t, m2), bb[m2], acc);
    sh.sq[t] = acc;
  }
  __syncthreads();
  if (t < K) {
    double acc = 0.0;
    for (int k = t; k < K; ++k) acc = fma(linv(k, t), sh.sq[k], acc);
    sh.wb[t] = acc;
  }
  __syncthreads();
  for (int q = t; q < K * K; q += 256) {  // Pa(i,j) = sum_k Linv(k,i) Linv(k,j), k >= max(i,j)
    const int i = q / K, j = q - i * K;
    const int k0 = i > j ? i : j;
    double a0 = linv(k0, i) * linv(k0, j), a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = k0 + 1;
    for (; k + 4 <= K; k += 4) {
      a0 = fma(cm[i * K + k], cm[j * K + k], a0);
      a1 = fma(cm[i * K + k + 1], cm[j * K + k + 1], a1);
      a2 = fma(cm[i * K + k + 2], cm[j * K + k + 2], a2);
      a3 = fma(cm[i * K + k + 3], cm[j * K + k + 3], a3);
    }
    for (; k < K; ++k) a0 = fma(cm[i * K + k], cm[j * K + k], a0);
    Wb[q] = (WT)(alpha * ((a0 + a1) + (a2 + a3)) + sh.wb[i]);
  }
  __syncthreads();
  for (int i0 = 0; i0 < K; i0 += kLpfRows) {  // W = W_GM W_LPF, rows staged in cm
    const int nr = min(kLpfRows, K - i0);
    for (int q = t; q < nr * K; q += 256) cm[q] = (double)Wb[(long long)i0 * K + q];
    __syncthreads();
    for (int q = t; q < nr * K; q += 256) {
      const int i = q / K, k = q - i * K;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int j = 0;
      for (; j + 4 <= K; j += 4) {
        a0 = fma(cm[i * K + j], lpf_w(sh, cnt, K, j, k), a0);
        a1 = fma(cm[i * K + j + 1], lpf_w(sh, cnt, K, j + 1, k), a1);
        a2 = fma(cm[i * K + j + 2], lpf_w(sh, cnt, K, j + 2, k), a2);
        a3 = fma(cm[i * K + j + 3], lpf_w(sh, cnt, K, j + 3, k), a3);
      }
      for (; j < K; ++j) a0 = fma(cm[i * K + j], lpf_w(sh, cnt, K, j, k), a0);
      Wb[(long long)(i0 + i) * K + k] = (WT)((a0 + a1) + (a2 + a3));
    }
    __syncthreads();
  }
}

template <typename T>
struct CsrRows {
  const T *hd, *rd, *rl, *dp;
  long long o0;
  int K;
  __device__ T dx(int r, int j) const { return dp[o0 + r] - hd[(o0 + r) * K + j]; }
  __device__ double rloc(int r) const { return (double)rl[o0 + r]; }
  __device__ double rdiag(int r) const { return (double)rd[o0 + r]; }
};

template <typename T>
__global__ void __launch_bounds__(256) k_lpfgm(int K, int mode, int nmonte, int dp, double alpha,
                                               const int64_t* __restrict__ offs, const T* __restrict__ hdxf,
                                               const T* __restrict__ rdiag, const T* __restrict__ rloc,
                                               const T* __restrict__ dep, const double* __restrict__ rtab,
                                               const double* __restrict__ ttab, const double* __restrict__ G,
                                               const double* __restrict__ bvec, T* __restrict__ wvec,
                                               T* __restrict__ wmat) {
  __shared__ LpfShared sh;
  extern __shared__ double cm[];  // K x K (mode 3), then the pick counts
  const int b = blockIdx.x, t = threadIdx.x;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cm + (mode == 3 ? K * K : 0));
  const long long o0 = offs[b], o1 = offs[b + 1];
  T* Wb = wmat + (long long)b * K * K;
  if (t < K) wvec[(long long)b * K + t] = (T)0;
  if (o1 <= o0) {  // lpfgm_core_noobs: W = I
    for (int q = t; q < K * K; q += 256) Wb[q] = (q / K == q % K) ? (T)1 : (T)0;
    return;
  }
  const CsrRows<T> rows{hdxf, rdiag, rloc, dep, o0, K};
  lpf_counts(sh, cnt, K, (int)(o1 - o0), nmonte, dp, rows, rtab, ttab);
  if (mode == 2) {
    for (int q = t; q < K * K; q += 256) Wb[q] = (T)lpf_w(sh, cnt, K, q / K, q % K);
    return;
  }
  gm_times_lpf<T>(sh, cnt, cm, K, alpha, G + (long long)b * K * K, bvec + (long long)b * K, Wb);
}

// Local obs of one analysis point, as obs_local gathers them
// (common_da.f90:276-287): row r is obs lobs[l0 + r].
template <typename T>
struct LocalRows {
  const int32_t* lobs;
  const double* lrloc;
  const double* yT;
  const T *oerr, *odep;
  long long l0;
  int K;
  __device__ T dx(int r, int j) const {
    const int o = lobs[l0 + r];
    return odep[o] - (T)yT[(long long)o * K + j];
  }
  __device__ double rloc(int r) const { return lrloc[l0 + r]; }
  __device__ double rdiag(int r) const {
    const T e = oerr[lobs[l0 + r]];
    return (double)(T)(e * e);
  }
};

template <typename T> __device__ __forceinline__ T sqrt_t(T x);
template <> __device__ __forceinline__ double sqrt_t<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_t<float>(float x) { return sqrtf(x); }

// One point of da_main with iswitch_da 2/3 (common_da.f90:67-114 with
// lpfgm_core / lpfgm_core_noobs) and its share of round_off (:506-533):
// W from the core (identity without local obs), then per variable the
// forecast mean/spread/perturbations, Analysis_Ensemble with wvec = 0,
// the analysis statistics, RTPP, RTPS and the clamps, each a serial loop in
// member order in the real kind as the reference runs it.
template <typename T>
__global__ void __launch_bounds__(256) k_lpfgm_points(KParams kp, Tables<T> tb, const T* __restrict__ f3,
                                                      const T* __restrict__ f2, T* __restrict__ a3,
                                                      T* __restrict__ a2, const int32_t* __restrict__ nloc,
                                                      const int64_t* __restrict__ loff, long long lbase,
                                                      const int32_t* __restrict__ lobs,
                                                      const double* __restrict__ lrloc,
                                                      const double* __restrict__ yT,
                                                      const int32_t* __restrict__ actidx, int a0, double* G,
                                                      const double* __restrict__ bvec, long long p0,
                                                      LpfPointArgs la) {
#pragma clang fp contract(off)
  __shared__ LpfShared sh;
  __shared__ T xs[kLpfMaxK], dxf[kLpfMaxK], xa[kLpfMaxK];
  __shared__ T stat[4];  // forecast mean, spread; analysis mean, spread
  extern __shared__ double cm[];  // K x K (mode 3), then the pick counts
  const int K = kp.K, t = threadIdx.x;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(cm + (la.mode == 3 ? K * K : 0));
  const long long p = p0 + blockIdx.x;
  const int n = nloc[p];
  int wsrc = 0;  // 0 identity, 2 W_LPF, 3 W in the G slot
  double* Wg = nullptr;
  if (n > 0) {
    const LocalRows<T> rows{lobs, lrloc, yT, tb.oerr, tb.odep, loff[p] - lbase, K};
    lpf_counts(sh, cnt, K, n, la.nmonte, la.dp, rows, la.rtab, la.ttab);
    wsrc = 2;
    if (la.mode == 3) {
      const long long a = actidx[p] - a0;
      Wg = G + a * K * K;
      gm_times_lpf<double>(sh, cnt, cm, K, la.alpha, Wg, bvec + a * K, Wg);
      wsrc = 3;
    }
  }
  auto wij = [&](int i, int j) -> T {
    if (wsrc == 0) return i == j ? (T)1 : (T)0;
    if (wsrc == 2) return (T)lpf_w(sh, cnt, K, i, j);
    return (T)Wg[i * K + j];
  };
  const long long P = kp.P;
  const int i = (int)(p % kp.nij1), k = (int)(p / kp.nij1);
  const bool surf = k == kp.nlev - 1;
  const T art = (T)kp.alpha_rtpp, ars = (T)kp.alpha_rtps;
  const int nvt = kp.nv3d + (surf ? kp.nv2d : 0);
  for (int vv = 0; vv < nvt; ++vv) {
    const bool is3 = vv < kp.nv3d;
    const int v = is3 ? vv : vv - kp.nv3d;
    const long long base = is3 ? p + P * ((long long)K * v) : i + (long long)kp.nij1 * ((long long)K * v);
    const long long ms = is3 ? P : (long long)kp.nij1;
    const T* src = (is3 ? f3 : f2) + base;
    T* dst = (is3 ? a3 : a2) + base;
    if (t < K) xs[t] = src[ms * t];
    __syncthreads();
    if (t == 0) {
      T sm = (T)0;
      for (int m = 0; m < K; ++m) sm += xs[m];
      const T fm = sm / (T)K;
      T q = (T)0;
      for (int m = 0; m < K; ++m) {
        const T d = xs[m] - fm;
        q += d * d;
      }
      stat[0] = fm;
      stat[1] = sqrt_t<T>(q / (T)(K - 1));
    }
    __syncthreads();
    if (t < K) dxf[t] = xs[t] - stat[0];
    __syncthreads();
    if (t < K) {
      T acc = stat[0];
      for (int ii = 0; ii < K; ++ii) acc += dxf[ii] * ((T)0 + wij(ii, t));
      xa[t] = acc;
    }
    __syncthreads();
    if (t == 0) {
      T sm = (T)0;
      for (int m = 0; m < K; ++m) sm += xa[m];
      const T am = sm / (T)K;
      T q = (T)0;
      for (int m = 0; m < K; ++m) {
        const T d = xa[m] - am;
        q += d * d;
      }
      stat[2] = am;
      stat[3] = sqrt_t<T>(q / (T)(K - 1));
    }
    __syncthreads();
    if (t < K) {
      const T am = stat[2], as = stat[3];
      const T dxa = xa[t] - am;
      T out = xa[t];
      if ((double)art > 0.0) {
        const T head = am + art * dxf[t];
        out = (T)((double)head + (1.0 - (double)art) * (double)dxa);
      }
      if ((double)ars > 0.0) {
        const T f = as == (T)0 ? (T)0 : (T)((double)(ars * stat[1] / as) + (1.0 - (double)ars));
        out = am + dxa * f;
      }
      if (is3 && kp.roff && k >= 1 && tb.phi1[i] != (T)1) {
        if (v == kp.iv3d_t - 1) {
          out = out < (T)kp.tmin ? (T)kp.tmin : out;
          out = out > (T)kp.tmax ? (T)kp.tmax : out;
        }
        if (v == kp.iv3d_s - 1) {
          out = out < (T)kp.smin ? (T)kp.smin : out;
          out = out > (T)kp.smax ? (T)kp.smax : out;
        }
      }
      dst[ms * t] = out;
      if (!isfinite((double)out)) atomicOr(la.status, ST_NONFINITE);
    }
    __syncthreads();
  }
}

}  // namespace

namespace letkf {

// The draws of each resampling as resampling_matrix uses them
// (common_lpfgm.f90:111-128: sorted for MR, the SR formula in place) and
// the running sums of 1/nmonte in the real kind (:259), uploaded to buf.
hipError_t lpf_prepare(int K, int nmonte, int rm_sr, bool f64, const double* rand, DevBuf& buf, hipStream_t s,
                       const double** d_rt, const double** d_tt) {
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
  std::vector<double> tt((size_t)nmonte + 1);
  tt[0] = 0.0;
  for (int c = 1; c <= nmonte; ++c)
    tt[c] = f64 ? tt[c - 1] + 1.0 / (double)nmonte : (double)(float)(tt[c - 1] + 1.0 / (double)nmonte);
  hipError_t e = buf.ensure(sizeof(double) * (rt.size() + tt.size()));
  if (e != hipSuccess) return e;
  double* p = buf.as<double>();
  e = hipMemcpyAsync(p, rt.data(), sizeof(double) * rt.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p + rt.size(), tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host vectors go out of scope
  *d_rt = p;
  *d_tt = p + rt.size();
  return e;
}

template <typename T>
hipError_t launch_lpfgm_points(const KParams& kp, const Tables<T>& tb, const T* f3, const T* f2, T* a3, T* a2,
                               const int32_t* nloc, const int64_t* loff, long long lbase, const int32_t* lobs,
                               const double* lrloc, const double* yT, const int32_t* actidx, int a0, double* G,
                               const double* bvec, long long p0, long long p1, const LpfPointArgs& la,
                               hipStream_t s) {
  if (p1 <= p0) return hipSuccess;
  const size_t lds = lpf_dyn_bytes(kp.K, la.mode);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_lpfgm_points<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_lpfgm_points<T>, dim3((unsigned)(p1 - p0)), dim3(256), lds, s, kp, tb, f3, f2, a3, a2, nloc,
                     loff, lbase, lobs, lrloc, yT, actidx, a0, G, bvec, p0, la);
  return hipGetLastError();
}
template hipError_t launch_lpfgm_points<float>(const KParams&, const Tables<float>&, const float*, const float*,
                                               float*, float*, const int32_t*, const int64_t*, long long,
                                               const int32_t*, const double*, const double*, const int32_t*, int,
                                               double*, const double*, long long, long long, const LpfPointArgs&,
                                               hipStream_t);
template hipError_t launch_lpfgm_points<double>(const KParams&, const Tables<double>&, const double*,
                                                const double*, double*, double*, const int32_t*, const int64_t*,
                                                long long, const int32_t*, const double*, const double*,
                                                const int32_t*, int, double*, const double*, long long, long long,
                                                const LpfPointArgs&, hipStream_t);

}  // namespace letkf

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
  h->invalidate_plan("invalidated by letkf_lpfgm_core_batched, which may reallocate the plan's workspaces");
  const double* d_rt = nullptr;
  const double* d_tt = nullptr;
  HIP_TRY(lpf_prepare(K, nmonte, rm_sr, f64, rand, h->bin_tmp, s, &d_rt, &d_tt));
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
  const size_t lds = lpf_dyn_bytes(K, mode);
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
