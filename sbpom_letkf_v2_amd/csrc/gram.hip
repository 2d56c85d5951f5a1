// gram.hip -- the ensemble-space Gram products of letkf_core
// (letkf/run/src/common_letkf.f90:93-126) for a batch of points, in FP64:
//   Rlinv_dYf = hdxf * rloc / rdiag                 (:93-97)
//   G = Rlinv_dYf^T hdxf + (K-1)/rho I              (:112-126)
//   b = Rlinv_dYf^T dep                              (the right-hand side of
//       wvec = Pa (Rlinv_dYf^T dep), an exact re-association of :162-184)
// One workgroup per local problem; its rows are staged through LDS in panels
// of 2048/KT rows.
#include <algorithm>
#include <cstdlib>

#include "letkf_internal.h"

namespace letkf {

typedef float mf_v16 __attribute__((ext_vector_type(16)));
typedef __bf16 mf_bf8 __attribute__((ext_vector_type(8)));

// Where the rows of one local problem come from.
//  ObsSource: the da_main path -- row l of point p is obs lobs[o0+l] of the
//    sorted tables (obs-major FP64 copy yT), rdiag = REAL(obserr**2, r_size),
//    rloc from the search.
//  CsrSource: letkf_core_batched -- rows given directly.
template <typename T>
struct ObsSource {
  Tables<T> tb;
  const double* yT;
  const int32_t* act_p;
  const int64_t* loff;
  long long lbase;
  const int32_t* lobs;
  const double* lrloc;
  int a0;
  long long o0;
  const mf_bf8* ysp = nullptr;  // the rows as BF16 hi / lo groups (launch_split_obs), or none
  __device__ int begin(int b) {
    const int p = act_p[a0 + b];
    o0 = loff[p] - lbase;
    return (int)(loff[p + 1] - loff[p]);
  }
  __device__ void row(int l, int K, int m, double& y, double& rl, double& rd) const {
    const int io = lobs[o0 + l];
    rl = lrloc[o0 + l];
    const T er = tb.oerr[io];
    rd = (double)(T)(er * er);
    y = yT[(long long)io * K + m];
  }
  __device__ double dep(int l) const { return (double)tb.odep[lobs[o0 + l]]; }
  __device__ void row_of(int io, int K, int m, double& y) const { y = yT[(long long)io * K + m]; }
  // row l of diag(rloc/rdiag) Y: obs index, weight and innovation
  __device__ void wrow(int l, int& io, double& w, double& d) const {
    io = lobs[o0 + l];
    const T er = tb.oerr[io];
    w = lrloc[o0 + l] / (double)(T)(er * er);
    d = (double)tb.odep[io];
  }
  // row l of B = diag(sqrt(rloc/rdiag)) Y: obs index and sqrt weight
  __device__ void brow(int l, int& io, double& sw) const {
    io = lobs[o0 + l];
    const T er = tb.oerr[io];
    sw = sqrt(lrloc[o0 + l] / (double)(T)(er * er));
  }
};

template <typename T>
struct CsrSource {
  const int64_t* offs;
  const T *hdxf, *rdiag, *rloc, *depv;
  long long o0;
  __device__ int begin(int b) {
    o0 = offs[b];
    return (int)(offs[b + 1] - offs[b]);
  }
  __device__ void row(int l, int K, int m, double& y, double& rl, double& rd) const {
    const long long q = o0 + l;
    y = (double)hdxf[q * K + m];
    rl = (double)rloc[q];
    rd = (double)rdiag[q];
  }
  __device__ double dep(int l) const { return (double)depv[o0 + l]; }
  __device__ void brow(int l, int& io, double& sw) const {
    io = (int)(o0 + l);
    sw = sqrt((double)rloc[o0 + l] / (double)rdiag[o0 + l]);
  }
  __device__ void row_of(int io, int K, int m, double& y) const { y = (double)hdxf[(long long)io * K + m]; }
  __device__ void wrow(int l, int& io, double& w, double& d) const {
    io = (int)(o0 + l);
    w = (double)rloc[o0 + l] / (double)rdiag[o0 + l];
    d = (double)depv[o0 + l];
  }
};

// KT: padded ensemble tile (64 or 128); each of the 16x16 threads owns a
// TT x TT block of G with TT = KT/16.  diag = (K-1)/rho in r_size.
//
// dual == 2: only the points with n < K (the others are left to the MFMA
// Gram, which then skips these).
// dual != 0 (batched form): a point with n < K local obs instead gets the
// n x n matrix M = B B^T, B = diag(sqrt(rloc/rdiag)) Y, summed over the K
// members, and bvec = diag(sqrt(rloc/rdiag)) dep; both are stored with
// row pitch n.  C = (K-1)/rho I + B^T B shares its nonzero spectrum with M.
template <int KT, typename Src>
__global__ void __launch_bounds__(256) k_gram_fp64(int K, double diag, Src src, double* __restrict__ G,
                                                   double* __restrict__ bvec, int32_t* __restrict__ nrow,
                                                   int dual) {
  constexpr int TT = KT / 16;
  constexpr int kPanel = 2048 / KT;  // 32 obs at KT=64, 16 at KT=128 (33 KiB LDS)
  __shared__ double As[kPanel][KT + 1];  // Rlinv_dYf rows
  __shared__ double Bs[kPanel][KT + 1];  // hdxf rows
  __shared__ double ds[kPanel];
  __shared__ double sws[KT];
  __shared__ int ios[KT];
  const int n = src.begin(blockIdx.x);
  if (dual == 2 && n >= K) return;  // dual slots only (the MFMA kernel takes the others)
  const int t = threadIdx.x;
  const int tr = t >> 4, tc = t & 15;
  double acc[TT][TT];
#pragma unroll
  for (int x = 0; x < TT; ++x)
#pragma unroll
    for (int y = 0; y < TT; ++y) acc[x][y] = 0.0;
  double bacc = 0.0;
  if (dual && n < K) {
    // ---- M = B B^T over members, panels of kPanel members ----
    for (int l = t; l < n; l += 256) {
      int io;
      double sw;
      src.brow(l, io, sw);
      ios[l] = io;
      sws[l] = sw;
      bvec[(long long)blockIdx.x * K + l] = sw * src.dep(l);
    }
    __syncthreads();
    for (int m0 = 0; m0 < K; m0 += kPanel) {
      const int pm = min(kPanel, K - m0);
      for (int q = t; q < kPanel * KT; q += 256) {
        const int l = q / kPanel, mm = q - l * kPanel;
        double v = 0.0;
        if (l < n && mm < pm) {
          double y;
          src.row_of(ios[l], K, m0 + mm, y);
          v = sws[l] * y;
        }
        Bs[mm][l] = v;
      }
      __syncthreads();
      for (int mm = 0; mm < pm; ++mm) {
        double ar[TT], bc[TT];
#pragma unroll
        for (int x = 0; x < TT; ++x) ar[x] = Bs[mm][tr * TT + x];
#pragma unroll
        for (int y = 0; y < TT; ++y) bc[y] = Bs[mm][tc * TT + y];
#pragma unroll
        for (int x = 0; x < TT; ++x)
#pragma unroll
          for (int y = 0; y < TT; ++y) acc[x][y] = fma(ar[x], bc[y], acc[x][y]);
      }
      __syncthreads();
    }
    double* Mp = G + (long long)blockIdx.x * K * K;
#pragma unroll
    for (int x = 0; x < TT; ++x) {
      const int r = tr * TT + x;
      if (r >= n) continue;
#pragma unroll
      for (int y = 0; y < TT; ++y) {
        const int c = tc * TT + y;
        if (c < n && r <= c) Mp[r * n + c] = acc[x][y];  // upper triangle: all its readers use
      }
    }
    if (t == 0 && nrow) nrow[blockIdx.x] = n;
    return;
  }
  for (int l0 = 0; l0 < n; l0 += kPanel) {
    const int nb = min(kPanel, n - l0);
    for (int q = t; q < kPanel * KT; q += 256) {
      const int l = q / KT, m = q - l * KT;
      double av = 0.0, bv = 0.0;
      if (l < nb && m < K) {
        double rl, rd;
        src.row(l0 + l, K, m, bv, rl, rd);
        av = bv * rl / rd;  // Rlinv_dYf = hdxf * rloc / rdiag (common_letkf.f90:95)
      }
      As[l][m] = av;
      Bs[l][m] = bv;
    }
    if (t < kPanel) ds[t] = t < nb ? src.dep(l0 + t) : 0.0;
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
  double* Gp = G + (long long)blockIdx.x * K * K;
#pragma unroll
  for (int x = 0; x < TT; ++x) {
    const int r = tr * TT + x;
    if (r >= K) continue;
#pragma unroll
    for (int y = 0; y < TT; ++y) {
      const int c = tc * TT + y;
      if (c < K && r <= c) Gp[r * K + c] = acc[x][y] + (r == c ? diag : 0.0);
    }
  }
  if (t < K) bvec[(long long)blockIdx.x * K + t] = bacc;
  if (t == 0 && nrow) nrow[blockIdx.x] = n;
}

// ------------------------------------------------------------------------
// The same FP64 products on the FP64 matrix cores (v_mfma_f64_16x16x4f64):
// the output (K x K, or n x n in the dual form) in 16 x 16 tiles, only the
// T (T + 1) / 2 tiles that hold part of the upper triangle (T = OUT / 16),
// dealt round-robin to the four wavefronts (3 / 3 / 2 / 2 of 10 at OUT 64,
// 9 each of 36 at OUT 128); the summation index (obs, or members in the
// dual form) staged through LDS in panels of 16 rows, four rows per MFMA.
// Lane layout (checked on gfx950): a = A[i][k], b = B[k][j] with
// i, j = lane % 16, k = lane / 16; D[i][j] in lane j + 16 (i % 4), element
// i / 4.  Operands and accumulation are FP64, as in k_gram_fp64.
typedef double mf_d4 __attribute__((ext_vector_type(4)));
constexpr int kP64 = 16;  // summation rows per LDS panel

// One slot sb of the FP64 products on the whole workgroup (every thread
// calls it, with the same sb).
template <int OUT, typename Src>
__device__ __forceinline__ void gram64_slot(int sb, int K, double diag, Src src, double* __restrict__ G,
                                            double* __restrict__ bvec, int32_t* __restrict__ nrow, int dual) {
  constexpr int T = OUT / 16;
  constexpr int NU = T * (T + 1) / 2;
  constexpr int NT = (NU + 3) / 4;
  constexpr int LP = OUT + 1;            // odd pitch: row and column walks hit distinct banks
  constexpr int kStage = kP64 * OUT / 256;  // panel elements staged per thread (4 or 8)
  __shared__ double As[kP64][LP];        // primal: Rlinv_dYf rows; dual: B^T panel (A = B)
  __shared__ double Bs[kP64][LP];        // primal: hdxf rows
  __shared__ double ds[kP64];
  __shared__ double sws[OUT];
  __shared__ int ios[OUT];
  const int n = src.begin(sb);
  const bool dm = dual && n < K;
  if (dual == 2 && !dm) return;  // dual slots only (the FP32 MFMA kernel takes the others)
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int N = dm ? n : K;  // output dimension
  const int S = dm ? K : n;  // summation length
  // this wavefront's upper-triangle tiles u = wv, wv + 4, ... (row-major over ti <= tj)
  int tr[NT], tc[NT];
  bool live[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    const int u = wv + 4 * q;
    int ti = 0, rem = u;
    while (ti < T && rem >= T - ti) {
      rem -= T - ti;
      ++ti;
    }
    tr[q] = 16 * ti;
    tc[q] = 16 * (ti + rem);
    live[q] = u < NU && tr[q] < N && tc[q] < N;
  }
  mf_d4 acc[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) acc[q] = mf_d4{0.0, 0.0, 0.0, 0.0};
  if (dm) {
    for (int l = t; l < n; l += 256) {
      int io;
      double sw;
      src.brow(l, io, sw);
      ios[l] = io;
      sws[l] = sw;
      bvec[(long long)sb * K + l] = sw * src.dep(l);
    }
    __syncthreads();
  }
  double bacc = 0.0;  // primal: b = Rlinv_dYf^T dep, thread t = member t
  for (int s0 = 0; s0 < S; s0 += kP64) {
    const int ps = min(kP64, S - s0);
    if (dm) {
      // B^T panel: lanes walk 16 consecutive members of one obs (contiguous in yT)
      double yv[kStage];
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int q = t + 256 * u, mm = q & (kP64 - 1), l = q / kP64;
        yv[u] = 0.0;
        if (mm < ps && l < N) src.row_of(ios[l], K, s0 + mm, yv[u]);
      }
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int q = t + 256 * u, mm = q & (kP64 - 1), l = q / kP64;
        As[mm][l] = l < N ? sws[l] * yv[u] : 0.0;
      }
    } else {
      // Y panel: lanes walk the members of one obs, four loads in flight per batch
      constexpr int kB = 4;
#pragma unroll
      for (int u0 = 0; u0 < kStage; u0 += kB) {
        double yv[kB], rl[kB], rd[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          const int q = t + 256 * (u0 + u), l = q / OUT, m = q % OUT;
          yv[u] = 0.0;
          rl[u] = 0.0;
          rd[u] = 1.0;
          if (l < ps && m < K) src.row(s0 + l, K, m, yv[u], rl[u], rd[u]);
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          const int q = t + 256 * (u0 + u), l = q / OUT, m = q % OUT;
          As[l][m] = yv[u] * rl[u] / rd[u];  // Rlinv_dYf = hdxf * rloc / rdiag (common_letkf.f90:95)
          Bs[l][m] = yv[u];
        }
      }
      if (t < kP64) ds[t] = t < ps ? src.dep(s0 + t) : 0.0;
    }
    __syncthreads();
    const double(*Xs)[LP] = As;
    const double(*Ys)[LP] = dm ? As : Bs;
    for (int k0 = 0; k0 < ps; k0 += 4) {
      const int kk = k0 + (lane >> 4);
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        if (!live[q]) continue;
        const double x = Xs[kk][tr[q] + (lane & 15)];
        const double y = Ys[kk][tc[q] + (lane & 15)];
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[q], 0, 0, 0);
      }
    }
    if (!dm && t < K)
      for (int l = 0; l < ps; ++l) bacc = fma(As[l][t], ds[l], bacc);
    __syncthreads();
  }
  double* Gp = G + (long long)sb * K * K;
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    if (!live[q]) continue;
    const int col = tc[q] + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = tr[q] + 4 * r + (lane >> 4);
      if (row < N && col < N && row <= col)  // upper triangle: all its readers use
        Gp[row * N + col] = acc[q][r] + (!dm && row == col ? diag : 0.0);
    }
  }
  if (!dm && t < K) bvec[(long long)sb * K + t] = bacc;
  if (t == 0 && nrow) nrow[sb] = n;
}

template <int OUT, typename Src>
__global__ void __launch_bounds__(256) k_gram_mfma64(int K, double diag, Src src, double* __restrict__ G,
                                                     double* __restrict__ bvec, int32_t* __restrict__ nrow,
                                                     int dual) {
  gram64_slot<OUT, Src>(blockIdx.x, K, diag, src, G, bvec, nrow, dual);
}

// The slots list[0..cnt) with flags[slot] == -1 (those the mixed-precision
// update left to the fallback): each workgroup tests 256 positions a pass
// and runs the products of the flagged ones.
template <int OUT, typename Src>
__global__ void __launch_bounds__(256) k_gram_mfma64_flagged(int K, double diag, Src src, double* __restrict__ G,
                                                             double* __restrict__ bvec, int dual,
                                                             const int32_t* __restrict__ list, int cnt,
                                                             const int32_t* __restrict__ flags) {
  __shared__ unsigned long long fm[4];
  const int t = threadIdx.x;
  for (int base = blockIdx.x * 256; base < cnt; base += gridDim.x * 256) {
    const int pos = base + t;
    const bool f = pos < cnt && flags[list[pos]] == -1;
    const unsigned long long m = __ballot(f);
    if ((t & 63) == 0) fm[t >> 6] = m;
    __syncthreads();
    for (int wq = 0; wq < 4; ++wq) {
      unsigned long long mm = fm[wq];
      while (mm) {
        const int bit = __ffsll((long long)mm) - 1;
        mm &= mm - 1;
        gram64_slot<OUT, Src>(list[base + 64 * wq + bit], K, diag, src, G, bvec, nullptr, dual);
      }
    }
    __syncthreads();  // the next pass rewrites fm
  }
}

template <typename Src>
static hipError_t launch_gram_src(int K, double diag, const Src& src, int na, double* G, double* bvec,
                                  int32_t* nrow, int dual, hipStream_t s) {
  if (na <= 0) return hipSuccess;
  if (K <= 64)
    hipLaunchKernelGGL((k_gram_mfma64<64, Src>), dim3(na), dim3(256), 0, s, K, diag, src, G, bvec, nrow, dual);
  else if (K <= 128)
    hipLaunchKernelGGL((k_gram_mfma64<128, Src>), dim3(na), dim3(256), 0, s, K, diag, src, G, bvec, nrow, dual);
  else
    hipLaunchKernelGGL((k_gram_fp64<128, Src>), dim3(na), dim3(256), 0, s, K, diag, src, G, bvec, nrow, dual);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gram_fp64(const KParams& kp, const Tables<T>& tb, const double* yT, const int32_t* act_p,
                            const int64_t* loff, long long lbase, const int32_t* lobs, const double* lrloc,
                            int a0, int na, double* G, double* bvec, int dual, hipStream_t s) {
  ObsSource<T> src{tb, yT, act_p, loff, lbase, lobs, lrloc, a0, 0};
  const double diag = (double)((T)(kp.K - 1) / (T)kp.rho);
  return launch_gram_src(kp.K, diag, src, na, G, bvec, nullptr, dual, s);
}

template <typename T>
hipError_t launch_gram_fp64_flagged(const KParams& kp, const Tables<T>& tb, const double* yT, const int32_t* act_p,
                                    const int64_t* loff, long long lbase, const int32_t* lobs, const double* lrloc,
                                    int a0, const int32_t* list, int cnt, const int32_t* flags, double* G,
                                    double* bvec, int dual, hipStream_t s) {
  if (cnt <= 0 || kp.K > 128) return hipSuccess;
  ObsSource<T> src{tb, yT, act_p, loff, lbase, lobs, lrloc, a0, 0};
  const double diag = (double)((T)(kp.K - 1) / (T)kp.rho);
  const unsigned grid = (unsigned)std::min((cnt + 255) / 256, 1024);
  if (kp.K <= 64)
    hipLaunchKernelGGL((k_gram_mfma64_flagged<64, ObsSource<T>>), dim3(grid), dim3(256), 0, s, kp.K, diag, src, G,
                       bvec, dual ? 1 : 0, list, cnt, flags);
  else
    hipLaunchKernelGGL((k_gram_mfma64_flagged<128, ObsSource<T>>), dim3(grid), dim3(256), 0, s, kp.K, diag, src, G,
                       bvec, dual ? 1 : 0, list, cnt, flags);
  return hipGetLastError();
}
template hipError_t launch_gram_fp64_flagged<float>(const KParams&, const Tables<float>&, const double*,
                                                    const int32_t*, const int64_t*, long long, const int32_t*,
                                                    const double*, int, const int32_t*, int, const int32_t*,
                                                    double*, double*, int, hipStream_t);
template hipError_t launch_gram_fp64_flagged<double>(const KParams&, const Tables<double>&, const double*,
                                                     const int32_t*, const int64_t*, long long, const int32_t*,
                                                     const double*, int, const int32_t*, int, const int32_t*,
                                                     double*, double*, int, hipStream_t);

template <typename T>
hipError_t launch_gram_csr(int K, double rho, int nb, const int64_t* offs, const T* hdxf, const T* rdiag,
                           const T* rloc, const T* dep, double* G, double* bvec, int32_t* nrow, hipStream_t s) {
  CsrSource<T> src{offs, hdxf, rdiag, rloc, dep, 0};
  const double diag = (double)((T)(K - 1) / (T)rho);
  return launch_gram_src(K, diag, src, nb, G, bvec, nrow, 0, s);
}
template hipError_t launch_gram_csr<float>(int, double, int, const int64_t*, const float*, const float*,
                                           const float*, const float*, double*, double*, int32_t*, hipStream_t);
template hipError_t launch_gram_csr<double>(int, double, int, const int64_t*, const double*, const double*,
                                            const double*, const double*, double*, double*, int32_t*,
                                            hipStream_t);

template hipError_t launch_gram_fp64<float>(const KParams&, const Tables<float>&, const double*,
                                            const int32_t*, const int64_t*, long long, const int32_t*,
                                            const double*, int, int, double*, double*, int, hipStream_t);
template hipError_t launch_gram_fp64<double>(const KParams&, const Tables<double>&, const double*,
                                             const int32_t*, const int64_t*, long long, const int32_t*,
                                             const double*, int, int, double*, double*, int, hipStream_t);

// ------------------------------------------------------------------------
// Mixed-precision Gram on MFMA (K <= 128): the K x K (or dual n x n)
// product runs as 32x32 output tiles with FP32 accumulation.  Only the
// tiles that hold part of the upper triangle are computed (the eigensolvers
// read no other entry): 3 of the 2x2 tiles of a 64-wide output, one per
// wavefront; 10 of the 4x4 tiles of a 128-wide output, 3/3/2/2 over the
// four wavefronts (kTile), tiles past the problem skipped:
//   prec FP32   : v_mfma_f32_32x32x2f32 on FP32-rounded operands;
//   prec BF16X3 : operands split x = hi + lo (bf16 each), three products
//                 hi*hi + hi*lo + lo*hi with v_mfma_f32_32x32x16_bf16; the
//                 split is done once as a panel is staged, and the primal
//                 rows are scaled by sqrt(rloc / rdiag) so that one panel
//                 serves as both operands (as in the dual form).
// The innovation term b (and dt in the dual form) stays FP64; the result is
// widened to FP64 for the eigensolve.  prec FP32_PANEL accumulates in two
// levels: the MFMA accumulators sum one panel of 32 obs or members and are
// then added into FP32 running totals, so a rounding error is relative to a
// panel sum or to the total rather than to a running sum over hundreds of
// terms (2-6x smaller Gram errors at 400-800 local obs), at the price of
// 16 more registers per tile (occupancy 8 -> 6 at OUT 64, 4 -> 2 at 128).
// prec FP32X3 removes the operands' FP32 rounding instead: each operand is
// staged as hi = fp32(x) and lo = fp32(x - hi) (panels of 16 rows, so that
// the four panels take the LDS of two) and every tile sums hi*hi + hi*lo +
// lo*hi on v_mfma_f32_32x32x2f32.
// Lane layouts (checked on gfx950):
//   32x32x2 f32 : a = A[i][k], i = lane%32, k = lane/32;  b = B[k][j], j = lane%32
//   32x32x16 bf16: a = A[i][8*(lane/32)+q], b = B[8*(lane/32)+q][j], q = 0..7
//   C (16 floats): C[(r/4)*8 + (lane/32)*4 + r%4][lane%32]

constexpr int kMfPanel = 32;  // obs (primal) or members (dual) per LDS panel

// Upper-triangle tiles (row tile * 4 + column tile, -1 none) per wavefront.
__constant__ int kTile64[4][1] = {{0 * 4 + 0}, {0 * 4 + 1}, {1 * 4 + 1}, {-1}};
__constant__ int kTile128[4][3] = {{0 * 4 + 0, 0 * 4 + 1, 0 * 4 + 2},
                                   {1 * 4 + 1, 1 * 4 + 2, 0 * 4 + 3},
                                   {2 * 4 + 2, 1 * 4 + 3, -1},
                                   {3 * 4 + 3, 2 * 4 + 3, -1}};

template <int PREC, int OUT, typename Src>
__global__ void __launch_bounds__(256) k_gram_mfma(int K, double diag, Src src, double* __restrict__ G,
                                                   double* __restrict__ bvec, int dual,
                                                   const int32_t* __restrict__ list) {
  constexpr int NT = OUT == 128 ? 3 : 1;          // tiles per wavefront
  constexpr bool kTwoLevel = PREC == LETKF_PREC_FP32_PANEL;
  constexpr bool kSplit = PREC == LETKF_PREC_FP32X3;  // operands as FP32 hi + lo
  // the split keeps a lo panel per operand in the LDS of half-height panels
  constexpr int kPanel = kSplit ? kMfPanel / 2 : kMfPanel;
  constexpr int kLoRows = kSplit ? kPanel : 1;
  constexpr int LP = OUT + 1;    // odd pitch: row-wise and column-wise panel writes hit distinct banks
  constexpr int RS = 256 / OUT;  // primal: threads sharing one output column
  constexpr int kStage = kPanel * OUT / 256;  // panel elements staged per thread (4 .. 16)
  // loads issued together: four at OUT 64 (one at a time left the K x K
  // Gram waiting a memory latency per element; 2 / 4 / 8 measured 121 / 103
  // / 129 ms at c2, the last at occupancy 3)
  constexpr int kBatch = OUT == 128 ? (kSplit ? 4 : 8) : 4;
  // BF16X3: one operand panel for both sides (the primal rows scaled by the
  // square root of their weight, as the dual rows are), split into BF16 hi
  // and lo once at staging and stored transposed -- rows = output index,
  // kPanel summation values + 8 of padding (80 bytes: 16-byte aligned rows)
  // -- so each lane loads its eight operand values with one 16-byte read
  constexpr bool kBf = PREC == LETKF_PREC_BF16X3;
  constexpr int PB = kPanel + 8;
  // panel: rows = summation index (obs or member), cols = output index
  __shared__ float As[kBf ? 1 : kPanel][LP];
  __shared__ float Bs[kBf ? 1 : kPanel][LP];
  __shared__ float Al[kLoRows][LP];  // FP32X3: x - (double)hi rounded to FP32
  __shared__ float Bl[kLoRows][LP];
  __shared__ __align__(16) __bf16 Sh[kBf ? OUT : 1][PB];
  __shared__ __align__(16) __bf16 Sl[kBf ? OUT : 1][PB];
  __shared__ double sqw[kBf ? kPanel : 1];  // BF16X3 primal: sqrt weight per panel row
  __shared__ double sws[OUT];  // dual: sqrt weight per obs; primal: weight per panel row
  __shared__ double dps[kPanel];
  __shared__ int ios[OUT];
  __shared__ double bred[256];
  // list (optional): the slots this launch covers, one workgroup each
  const int b = list ? list[blockIdx.x] : (int)blockIdx.x;
  const int n = src.begin(b);
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);  // wave-uniform: tile tables in scalar registers
  // BF16X3 staging: x -> FP32 -> BF16 hi + BF16 (x32 - hi), each step rounded
  // (the empty asm keeps the two conversions from being fused into one)
  auto put_bf = [&](int row, int col, double x) {
    float v = (float)x;
    asm volatile("" : "+v"(v));
    const __bf16 h = (__bf16)v;
    Sh[col][row] = h;
    Sl[col][row] = (__bf16)(v - (float)h);
  };
  const bool dm = dual && n < K;
  if (dual == 2 && dm) return;  // K x K slots only (the FP64 kernel takes the dual ones)
  if (dual == 3 && dm && n <= 64) return;  // dual slots up to 64 rows: k_gram_wave
  const int N = dm ? n : K;  // output dimension
  const int S = dm ? K : n;  // summation length
  // this wavefront's tiles: row / column offsets and liveness (wave-uniform)
  int tr[NT], tc[NT];
  bool live[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    const int code = OUT == 128 ? kTile128[wv][q] : kTile64[wv][q];
    tr[q] = code >= 0 ? 32 * (code >> 2) : 0;
    tc[q] = code >= 0 ? 32 * (code & 3) : 0;
    live[q] = code >= 0 && tr[q] < N && tc[q] < N;
  }
  mf_v16 acc[NT], tot[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    acc[q] = mf_v16{};
    tot[q] = mf_v16{};
  }
  if (dm) {
    for (int l = t; l < n; l += 256) {
      int io;
      double sw;
      src.brow(l, io, sw);
      ios[l] = io;
      sws[l] = sw;
      bvec[(long long)b * K + l] = sw * src.dep(l);
    }
    __syncthreads();
  }
  // primal: b = Y^T diag(rloc/rdiag) dep in FP64, thread t summing column
  // t % OUT over the panel rows it stages
  double bpart = 0.0;
  for (int s0 = 0; s0 < S; s0 += kPanel) {
    const int ps = min(kPanel, S - s0);
    if (dm) {
      // B^T panel: lanes walk the members of one obs (contiguous in yT);
      // a batch of kBatch loads is issued before its first store
#pragma unroll 1
      for (int u0 = 0; u0 < kStage; u0 += kBatch) {
        double yv[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = t + 256 * (u0 + u), row = q & (kPanel - 1), col = q / kPanel;
          yv[u] = 0.0;
          if (row < ps && col < N) src.row_of(ios[col], K, s0 + row, yv[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = t + 256 * (u0 + u), row = q & (kPanel - 1), col = q / kPanel;
          const double x = col < N ? sws[col] * yv[u] : 0.0;
          if constexpr (kBf) {
            put_bf(row, col, x);
            continue;
          }
          const float v = (float)x;
          As[row][col] = v;
          Bs[row][col] = v;
          if (kSplit) {
            const float lo = (float)(x - (double)v);
            Al[row][col] = lo;
            Bl[row][col] = lo;
          }
        }
      }
    } else {
      if (t < ps) {
        int io;
        double w, d;
        src.wrow(s0 + t, io, w, d);
        ios[t] = io;
        sws[t] = w;
        dps[t] = d;
        if (kBf) sqw[t] = sqrt(w);
      }
      __syncthreads();
      // Y panel: lanes walk the members of one obs, batches of kBatch loads
#pragma unroll 1
      for (int u0 = 0; u0 < kStage; u0 += kBatch) {
        double yv[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = t + 256 * (u0 + u), row = q / OUT, col = q % OUT;
          yv[u] = 0.0;
          if (row < ps && col < N) src.row_of(ios[row], K, col, yv[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = t + 256 * (u0 + u), row = q / OUT, col = q % OUT;
          float av = 0.f, bv = 0.f, al = 0.f, bl = 0.f;
          if constexpr (kBf) {
            double xs = 0.0;
            if (row < ps && col < N) {
              bpart = fma(yv[u] * sws[row], dps[row], bpart);
              xs = yv[u] * sqw[row];
            }
            put_bf(row, col, xs);
            continue;
          }
          if (row < ps && col < N) {
            const double wy = yv[u] * sws[row];
            bpart = fma(wy, dps[row], bpart);
            av = (float)wy;
            bv = (float)yv[u];
            if (kSplit) {
              al = (float)(wy - (double)av);
              bl = (float)(yv[u] - (double)bv);
            }
          }
          As[row][col] = av;
          Bs[row][col] = bv;
          if (kSplit) {
            Al[row][col] = al;
            Bl[row][col] = bl;
          }
        }
      }
    }
    __syncthreads();
    if constexpr (PREC != LETKF_PREC_BF16X3) {
#pragma unroll 2
      for (int k = 0; k < ps; k += 2) {
        const int kk = k + (lane >> 5);
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          if (!live[q]) continue;
          const float av = kk < ps ? As[kk][tr[q] + (lane & 31)] : 0.f;
          const float bv = kk < ps ? Bs[kk][tc[q] + (lane & 31)] : 0.f;
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[q], 0, 0, 0);
          if (kSplit) {  // hi*lo + lo*hi: the operands to ~2^-48
            const float al = kk < ps ? Al[kk][tr[q] + (lane & 31)] : 0.f;
            const float bl = kk < ps ? Bl[kk][tc[q] + (lane & 31)] : 0.f;
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bl, acc[q], 0, 0, 0);
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(al, bv, acc[q], 0, 0, 0);
          }
        }
      }
    } else {
      // summation values k .. k + 15 (zero past ps: the staging fills the panel)
      for (int k = 0; k < ps; k += 16) {
        const int ko = k + 8 * (lane >> 5);
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          if (!live[q]) continue;
          const int ra = tr[q] + (lane & 31), rb = tc[q] + (lane & 31);
          const mf_bf8 ah = *reinterpret_cast<const mf_bf8*>(&Sh[ra][ko]);
          const mf_bf8 al = *reinterpret_cast<const mf_bf8*>(&Sl[ra][ko]);
          const mf_bf8 bh = *reinterpret_cast<const mf_bf8*>(&Sh[rb][ko]);
          const mf_bf8 bl = *reinterpret_cast<const mf_bf8*>(&Sl[rb][ko]);
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[q], 0, 0, 0);
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[q], 0, 0, 0);
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[q], 0, 0, 0);
        }
      }
    }
    if (kTwoLevel) {  // panel sum -> running total
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        tot[q] += acc[q];
        acc[q] = mf_v16{};
      }
    }
    __syncthreads();
  }
  if (!dm) {
    bred[t] = bpart;
    __syncthreads();
    if (t < K) {
      double bs = 0.0;
#pragma unroll
      for (int r = 0; r < RS; ++r) bs += bred[t + r * OUT];
      bvec[(long long)b * K + t] = bs;
    }
  }
  double* Gp = G + (long long)b * K * K;
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    if (!live[q]) continue;
    const mf_v16 v = kTwoLevel ? tot[q] : acc[q];
    const int col = tc[q] + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = tr[q] + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
      if (row < N && col < N && row <= col)  // upper triangle: all its readers use
        Gp[row * N + col] = (double)v[r] + (!dm && row == col ? diag : 0.0);
    }
  }
}

// BF16x3 Gram of the dual-form points (n < K, nmem <= 64) with one
// wavefront per point, its MFMA operands loaded straight from the local obs
// rows -- no LDS panel, no block barrier, no idle wavefront while the (at
// most three) 32 x 32 tiles of M = B B^T run.  M sums over the K members in
// chunks of 16; lane (h, c) = (lane / 32, lane % 32) loads members
// 16 k + 8 h .. + 7 of rows c and 32 + c (eight consecutive doubles of an
// obs row), which are both the A operand of its row tile and the B operand
// of its column tile (A = B^T), so each value is split into BF16 hi + lo
// once.  Operand values, their splits and the MFMA sequence are those of
// k_gram_mfma<LETKF_PREC_BF16X3, 64> (upper-triangle tiles only); the K x K
// points (and, for nmem > 64, the dual ones of more than 64 rows) are left
// to k_gram_wave_kk or k_gram_mfma.
// SPLIT: the operands are the unweighted rows, pre-split once per analysis
// (launch_split_obs: one 32-byte load gives a lane its hi and lo vectors,
// half the bytes of the FP64 row and no conversion), and the weights come
// in at the store, M(i, j) = sw_i sw_j (Y Y^T)(i, j) -- the same matrix,
// the BF16 splits taken of y instead of sw y.
template <typename T, bool SPLIT>
__global__ void __launch_bounds__(256) k_gram_wave(int K, ObsSource<T> src, int na, double* __restrict__ G,
                                                   double* __restrict__ bvec) {
  __shared__ double swl[4][64];  // SPLIT: the rows' sqrt weights
  const int lane = threadIdx.x & 63, c = lane & 31, h = lane >> 5;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x * 4 + w;
  if (b >= na) return;
  const int n = src.begin(b);
  if (n >= K || n > 64) return;  // K x K form, or more than two row tiles: k_gram_mfma
  const bool two = n > 32;  // tiles (0,0), (0,1), (1,1); else (0,0) only
  mf_v16 acc00 = mf_v16{}, acc01 = mf_v16{}, acc11 = mf_v16{};
  // x -> (hi, lo) in element e of two operand vectors
  // x -> FP32 -> BF16 hi + lo, each step rounded to nearest (the empty asm
  // keeps the compiler from fusing the two conversions of hi into a
  // correctly rounded FP64 -> BF16 one, a dozen instructions more)
  auto split = [](double x, mf_bf8& hi, mf_bf8& lo, int e) {
    float v = (float)x;
    asm volatile("" : "+v"(v));
    const __bf16 vh = (__bf16)v;
    hi[e] = vh;
    lo[e] = (__bf16)(v - (float)vh);
  };
  // rows c and 32 + c (obs), their sqrt weights; bvec = sw dep
  int io0, io1 = 0;
  double sw0, sw1 = 0.0;
  src.brow(c < n ? c : n - 1, io0, sw0);
  if (two) src.brow(32 + c < n ? 32 + c : n - 1, io1, sw1);
  if (c >= n) sw0 = 0.0;
  if (32 + c >= n) sw1 = 0.0;
  if (h == 0) {
    if (c < n) bvec[(long long)b * K + c] = sw0 * src.dep(c);
    if (32 + c < n) bvec[(long long)b * K + 32 + c] = sw1 * src.dep(32 + c);
  }
  const double* y0 = src.yT + (long long)io0 * K;
  const double* y1 = src.yT + (long long)io1 * K;
  const int ng = split_obs_groups(K);
  const mf_bf8* s0 = SPLIT ? src.ysp + ((long long)io0 * ng + h) * 2 : nullptr;
  const mf_bf8* s1 = SPLIT ? src.ysp + ((long long)io1 * ng + h) * 2 : nullptr;
  if (SPLIT) {
    swl[w][h * 32 + c] = h == 0 ? sw0 : sw1;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  for (int k0 = 0; k0 < K; k0 += 16) {
    mf_bf8 x0h, x0l, x1h, x1l;
    if (SPLIT) {
      x0h = s0[k0 / 4];
      x0l = s0[k0 / 4 + 1];
      if (two) {
        x1h = s1[k0 / 4];
        x1l = s1[k0 / 4 + 1];
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int m = k0 + 8 * h + e;
        const double v0 = m < K ? sw0 * y0[m < K ? m : K - 1] : 0.0;
        split(v0, x0h, x0l, e);
      }
    }
    acc00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x0h, acc00, 0, 0, 0);
    acc00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x0l, acc00, 0, 0, 0);
    acc00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0l, x0h, acc00, 0, 0, 0);
    if (two) {
      if (!SPLIT) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int m = k0 + 8 * h + e;
          const double v1 = m < K ? sw1 * y1[m < K ? m : K - 1] : 0.0;
          split(v1, x1h, x1l, e);
        }
      }
      acc01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x1h, acc01, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x1l, acc01, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0l, x1h, acc01, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1h, x1h, acc11, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1h, x1l, acc11, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1l, x1h, acc11, 0, 0, 0);
    }
  }
  // upper-triangle entries of the live tiles, row pitch n
  double* Gp = G + (long long)b * K * K;
  auto store = [&](const mf_v16& v, int tr, int tc) {
    const int col = tc + c;
    const double swc = SPLIT ? swl[w][col < 64 ? col : 63] : 1.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = tr + (r >> 2) * 8 + h * 4 + (r & 3);
      if (row < n && col < n && row <= col)
        Gp[row * n + col] = SPLIT ? (double)v[r] * (swl[w][row] * swc) : (double)v[r];
    }
  };
  store(acc00, 0, 0);
  if (two) {
    store(acc01, 0, 32);
    store(acc11, 32, 32);
  }
}

// yT rows -> BF16 hi / lo groups of 8 members, one thread a group
__global__ void __launch_bounds__(256) k_split_obs(int nobs, int K, int ng, const double* __restrict__ yT,
                                                   mf_bf8* __restrict__ ysp) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nobs * ng) return;
  const long long io = t / ng;
  const int g = (int)(t - io * ng);
  mf_bf8 hi, lo;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int m = 8 * g + e;
    const float v = m < K ? (float)yT[io * K + m] : 0.f;
    const __bf16 vh = (__bf16)v;
    hi[e] = vh;
    lo[e] = (__bf16)(v - (float)vh);
  }
  ysp[2 * t] = hi;
  ysp[2 * t + 1] = lo;
}

hipError_t launch_split_obs(int nobs, int K, const double* yT, void* ysp, hipStream_t s) {
  const long long nt = (long long)nobs * split_obs_groups(K);
  if (nt <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_obs, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, nobs, K,
                     split_obs_groups(K), yT, static_cast<mf_bf8*>(ysp));
  return hipGetLastError();
}

// The K x K form (n >= K in the batched form, every point in the plain one)
// of the BF16x3 Gram, one wavefront per point (nmem <= 64): C = B^T B with
// B = diag(sqrt(w)) Y, w = rloc / rdiag, summed over the n obs in chunks of
// 16.  Lane (h, c) holds members c and 32 + c of obs rows 16 k + 8 h .. + 7
// (the 32 lanes of a half read a row's members contiguously): as the A
// operand of its row tile (A[member][obs]) and the B operand of its column
// tile (B[obs][member]) these are the same registers.  The row data (obs
// index, sqrt weight, sqrt weight x innovation) of 64 rows at a time is
// fetched one row per lane into a per-wavefront LDS table, so a chunk waits
// on one memory latency (its rows' members) instead of three; b =
// Y^T diag(w) dep sums in FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_gram_wave_kk(int K, double diag, ObsSource<T> src, int na,
                                                      double* __restrict__ G, double* __restrict__ bvec, int dual,
                                                      const int32_t* __restrict__ list) {
  struct RowMeta {
    double sw, swd;
  };
  __shared__ RowMeta meta[4][64];
  __shared__ int mio[4][64];
  const int lane = threadIdx.x & 63, c = lane & 31, h = lane >> 5;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // list (optional): the na slots to visit, else slots 0 .. na - 1
  const int i = blockIdx.x * 4 + w;
  if (i >= na) return;
  const int b = list ? list[i] : i;
  const int n = src.begin(b);
  if (dual && n < K) return;  // dual form: k_gram_wave
  const bool two = K > 32;
  const int m0 = c < K ? c : K - 1, m1 = 32 + c < K ? 32 + c : K - 1;
  const bool ok0 = c < K, ok1 = 32 + c < K;
  mf_v16 acc00 = mf_v16{}, acc01 = mf_v16{}, acc11 = mf_v16{};
  double bp0 = 0.0, bp1 = 0.0;
  // x -> FP32 -> BF16 hi + lo, each step rounded to nearest (the empty asm
  // keeps the compiler from fusing the two conversions of hi into a
  // correctly rounded FP64 -> BF16 one, a dozen instructions more)
  auto split = [](double x, mf_bf8& hi, mf_bf8& lo, int e) {
    float v = (float)x;
    asm volatile("" : "+v"(v));
    const __bf16 vh = (__bf16)v;
    hi[e] = vh;
    lo[e] = (__bf16)(v - (float)vh);
  };
  for (int g0 = 0; g0 < n; g0 += 64) {
    {
      const int l = g0 + lane;
      int io;
      double wt, d;
      src.wrow(l < n ? l : n - 1, io, wt, d);
      const double sw = l < n ? sqrt(wt) : 0.0;
      meta[w][lane] = RowMeta{sw, sw * d};
      mio[w][lane] = io;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int gend = n - g0 < 64 ? n - g0 : 64;
    for (int l0 = 0; l0 < gend; l0 += 16) {
      mf_bf8 x0h, x0l, x1h, x1l;
      double y0[8], y1[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const double* yr = src.yT + (long long)mio[w][l0 + 8 * h + e] * K;
        y0[e] = yr[m0];
        y1[e] = two ? yr[m1] : 0.0;
      }
      // (a lane past K holds a clamped member: it feeds only rows and
      // columns >= K of the tiles and its own b, none of which is stored)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const RowMeta rm = meta[w][l0 + 8 * h + e];
        const double v0 = rm.sw * y0[e];
        bp0 = fma(v0, rm.swd, bp0);
        split(v0, x0h, x0l, e);
        if (two) {
          const double v1 = rm.sw * y1[e];
          bp1 = fma(v1, rm.swd, bp1);
          split(v1, x1h, x1l, e);
        }
      }
      acc00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x0h, acc00, 0, 0, 0);
      acc00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x0l, acc00, 0, 0, 0);
      acc00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0l, x0h, acc00, 0, 0, 0);
      if (two) {
        acc01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x1h, acc01, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0h, x1l, acc01, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x0l, x1h, acc01, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1h, x1h, acc11, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1h, x1l, acc11, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x1l, x1h, acc11, 0, 0, 0);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the next 64 rows overwrite the table
  }
  // b for members c and 32 + c: the two row halves of the obs chunks
  bp0 += __shfl_xor(bp0, 32);
  bp1 += __shfl_xor(bp1, 32);
  if (h == 0) {
    if (ok0) bvec[(long long)b * K + c] = bp0;
    if (two && ok1) bvec[(long long)b * K + 32 + c] = bp1;
  }
  double* Gp = G + (long long)b * K * K;
  auto store = [&](const mf_v16& v, int tr, int tc) {
    const int col = tc + c;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = tr + (r >> 2) * 8 + h * 4 + (r & 3);
      if (row < K && col < K && row <= col) Gp[row * K + col] = (double)v[r] + (row == col ? diag : 0.0);
    }
  };
  store(acc00, 0, 0);
  if (two) {
    store(acc01, 0, 32);
    store(acc11, 32, 32);
  }
}

template <int PREC, typename T>
static void launch_gram_mixed_p(int K, double diag, const ObsSource<T>& src, int na, double* G, double* bvec,
                                int dual, hipStream_t s, const int32_t* list = nullptr) {
  if (K <= 64)
    hipLaunchKernelGGL((k_gram_mfma<PREC, 64, ObsSource<T>>), dim3(na), dim3(256), 0, s, K, diag, src, G, bvec,
                       dual, list);
  else
    hipLaunchKernelGGL((k_gram_mfma<PREC, 128, ObsSource<T>>), dim3(na), dim3(256), 0, s, K, diag, src, G, bvec,
                       dual, list);
}

template <typename T>
hipError_t launch_gram_mixed(int prec, const KParams& kp, const Tables<T>& tb, const double* yT,
                             const int32_t* act_p, const int64_t* loff, long long lbase, const int32_t* lobs,
                             const double* lrloc, int a0, int na, double* G, double* bvec, int dual,
                             const void* ysp, hipStream_t s, const int32_t* klist, int kcnt) {
  if (na <= 0) return hipSuccess;
  if (kp.K > 128) return hipErrorInvalidValue;
  ObsSource<T> src{tb, yT, act_p, loff, lbase, lobs, lrloc, a0, 0, static_cast<const mf_bf8*>(ysp)};
  const double diag = (double)((T)(kp.K - 1) / (T)kp.rho);
  if (prec == LETKF_PREC_BF16X3 && dual != 2 && !getenv("LETKF_GRAM_BLOCK")) {
    // one wavefront per point: the dual-form points of up to 64 rows, then
    // (nmem <= 64) the K x K ones; for nmem > 64 the block kernel takes the
    // rest (dual == 3: it skips the points done here)
    if (dual && ysp)
      hipLaunchKernelGGL((k_gram_wave<T, true>), dim3((na + 3) / 4), dim3(256), 0, s, kp.K, src, na, G, bvec);
    else if (dual)
      hipLaunchKernelGGL((k_gram_wave<T, false>), dim3((na + 3) / 4), dim3(256), 0, s, kp.K, src, na, G, bvec);
    // the K x K slots lie in the size bucket of K: over its list when given
    if (kp.K <= 64) {
      if (!klist)
        hipLaunchKernelGGL(k_gram_wave_kk<T>, dim3((na + 3) / 4), dim3(256), 0, s, kp.K, diag, src, na, G, bvec,
                           dual, nullptr);
      else if (kcnt > 0)
        hipLaunchKernelGGL(k_gram_wave_kk<T>, dim3((kcnt + 3) / 4), dim3(256), 0, s, kp.K, diag, src, kcnt, G,
                           bvec, dual, klist);
      return hipGetLastError();
    }
    if (dual) dual = 3;
    // the rest are the slots of the 128 bucket (dual 65 .. K - 1 and K x K)
    if (dual && klist) {
      if (kcnt > 0) launch_gram_mixed_p<LETKF_PREC_BF16X3, T>(kp.K, diag, src, kcnt, G, bvec, dual, s, klist);
      return hipGetLastError();
    }
  }
  if (prec == LETKF_PREC_FP32)
    launch_gram_mixed_p<LETKF_PREC_FP32, T>(kp.K, diag, src, na, G, bvec, dual, s);
  else if (prec == LETKF_PREC_FP32_PANEL)
    launch_gram_mixed_p<LETKF_PREC_FP32_PANEL, T>(kp.K, diag, src, na, G, bvec, dual, s);
  else if (prec == LETKF_PREC_FP32X3)
    launch_gram_mixed_p<LETKF_PREC_FP32X3, T>(kp.K, diag, src, na, G, bvec, dual, s);
  else
    launch_gram_mixed_p<LETKF_PREC_BF16X3, T>(kp.K, diag, src, na, G, bvec, dual, s);
  return hipGetLastError();
}
template hipError_t launch_gram_mixed<float>(int, const KParams&, const Tables<float>&, const double*,
                                             const int32_t*, const int64_t*, long long, const int32_t*,
                                             const double*, int, int, double*, double*, int, const void*,
                                             hipStream_t, const int32_t*, int);
template hipError_t launch_gram_mixed<double>(int, const KParams&, const Tables<double>&, const double*,
                                              const int32_t*, const int64_t*, long long, const int32_t*,
                                              const double*, int, int, double*, double*, int, const void*,
                                              hipStream_t, const int32_t*, int);

}  // namespace letkf
