// search.hip -- local-observation search of da_main (count_nobsl /
// obs_local / obs_local_sub, letkf/run/src/common_da.f90:133-348), the
// point scans that turn local counts into list offsets, and the obs-major
// transpose of obshdxf.
//
// The search keeps the reference's r_size semantics: the box edges use
// NINT (round half away from zero) of r_size values, distance_2p
// (common.f90:79-101) stores cos(d) in r_size before ACOS, and rloc is
// rounded to r_size.  T is the caller's REAL.
#include <cstdlib>

#include "letkf_internal.h"

namespace letkf {

template <typename T> __device__ __forceinline__ int nint_r(T x);
template <> __device__ __forceinline__ int nint_r<double>(double x) { return (int)round(x); }
template <> __device__ __forceinline__ int nint_r<float>(float x) { return (int)roundf(x); }

template <typename T> __device__ __forceinline__ T acos_r(T x);
template <> __device__ __forceinline__ double acos_r<double>(double x) { return acos(x); }
template <> __device__ __forceinline__ float acos_r<float>(float x) { return acosf(x); }

template <typename T> __device__ __forceinline__ T abs_r(T x) { return x < (T)0 ? -x : x; }

// Walks the search box of point (i,k) in the reference's order and calls
// visit(io, dist, dlev) for each accepted obs io (0-based).
// NEED_DIST false (count only): an obs whose cos d is clearly inside the
// cut-off is accepted without acos; dist is then not formed.
template <typename T, bool NEED_DIST, typename Visit>
__device__ __forceinline__ int search_point(const KParams& kp, const Tables<T>& tb,
                                            int i, int k, Visit&& visit) {
#pragma clang fp contract(off)
  const T dz = (T)kp.dist_zero, dzv = (T)kp.dist_zerov;
  const T pi = (T)kp.pi, re = (T)kp.re;
  const T rdep = tb.depth1[i + (long long)kp.nij1 * k];
  const T hx = dz / tb.dlon1[i], hy = dz / tb.dlat1[i];
  int imin = nint_r<T>(tb.ri1[i] - hx);
  int imax = nint_r<T>(tb.ri1[i] + hx);
  int jmin = nint_r<T>(tb.rj1[i] - hy);
  int jmax = nint_r<T>(tb.rj1[i] + hy);
  imin = imin < 1 ? 1 : imin;
  imax = imax > kp.nlon ? kp.nlon : imax;
  jmin = jmin < 1 ? 1 : jmin;
  jmax = jmax > kp.nlat ? kp.nlat : jmax;
  const T lonb = tb.lon1[i];
  const double bb = (double)(T)(pi * tb.lat1[i]) / 180.0;
  const double sinb = sin(bb), cosb = cos(bb);
  const bool below_surface = k < kp.nlev - 1;
  int count = 0;
  for (int j = jmin; j <= jmax; ++j) {
    long long ib;
    if (imin > 1)
      ib = (long long)tb.nobsgrd[(imin - 2) + (long long)kp.nlon * (j - 1)] + 1;
    else if (j > 1)
      ib = (long long)tb.nobsgrd[(kp.nlon - 1) + (long long)kp.nlon * (j - 2)] + 1;
    else
      ib = 1;
    const long long ie = tb.nobsgrd[(imax - 1) + (long long)kp.nlon * (j - 1)];
    for (long long ip = ib; ip <= ie; ++ip) {
      const int io = (int)(ip - 1);
      const int el = tb.elm[io];
      T dlev;
      if (el == kp.id_z_obs && below_surface)
        dlev = abs_r<T>(rdep);
      else if (el != kp.id_z_obs)
        dlev = abs_r<T>(tb.olev[io] - rdep);
      else
        dlev = (T)0;
      if (dzv < dlev) continue;
      const double2 sca = tb.osc[io];  // sin(aa), cos(aa) of this obs (k_obs_trig)
      const double cc = (double)(T)(pi * (T)(lonb - tb.olon[io])) / 180.0;
      const double prod1 = sca.x * sinb;
      const double prod2 = sca.y * cosb;
      T cosd = (T)(prod1 + prod2 * cos(cc));
      if ((double)cosd > 1.0) cosd = (T)1;
      if ((double)cosd < -1.0) cosd = (T)-1;
      if ((double)cosd <= kp.cos_out) continue;  // beyond dist_zero whatever acos rounds to
      T dist = (T)0;
      if (NEED_DIST || (double)cosd < kp.cos_in) {
        dist = re * acos_r<T>(cosd);
        if (dz < dist) continue;
      }
      visit(count, io, dist, dlev);
      ++count;
    }
  }
  return count;
}

template <typename T>
__global__ void __launch_bounds__(256) k_search_count(KParams kp, Tables<T> tb, int32_t* nloc) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < kp.P; p += stride) {
    const int k = (int)(p / kp.nij1), i = (int)(p - (long long)k * kp.nij1);
    nloc[p] = search_point<T, false>(kp, tb, i, k, [](int, int, T, T) {});
  }
}

// obs_local's gather half: the obs index and rloc of every accepted obs,
// rloc = exp(-0.5*((dist/sigma_obs)^2 + (dlev/sigma_obsv)^2)) rounded to
// r_size (common_da.f90:282-287).
template <typename T>
__global__ void __launch_bounds__(256) k_search_fill(KParams kp, Tables<T> tb, long long p0, long long p1,
                                                     const int64_t* loff, int32_t* lobs, double* lrloc) {
  const long long base = loff[p0];
  const long long stride = (long long)gridDim.x * blockDim.x;
  const T so = (T)kp.sigma_obs, sv = (T)kp.sigma_obsv;
  const bool vloc = !(kp.sigma_obsv <= 0.0);
  for (long long p = p0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; p < p1; p += stride) {
    const long long o = loff[p] - base;
    if (loff[p + 1] == loff[p]) continue;
    const int k = (int)(p / kp.nij1), i = (int)(p - (long long)k * kp.nij1);
    search_point<T, true>(kp, tb, i, k, [&](int l, int io, T dist, T dlev) {
#pragma clang fp contract(off)
      T q = (dist / so) * (dist / so);
      if (vloc) q = q + (dlev / sv) * (dlev / sv);
      lobs[o + l] = io;
      lrloc[o + l] = (double)(T)exp(-0.5 * (double)q);
    });
  }
}

// ---------------------------------------------------------------- column search
// The same acceptance tests as search_point for every level of one grid
// column at once: one wavefront per column i, its lanes over the column's
// candidate obs -- the CSR ranges of the box's rows, flattened in the
// reference's walk order (row j ascending, then ip) -- so the horizontal
// test (cos d, acos: the costly part, the same at every level) runs once per
// (column, obs) instead of once per (point, obs).  A level loop then applies
// the vertical test of each level to the chunk's 64 candidates; a ballot and
// its prefix count give each accepted obs its place in the point's list, so
// a point's obs keep the reference's order and a level's entries are stored
// contiguously.  The running count of level k lives in lane k % 64 of cnt0
// (k < 64) or cnt1.  FILL: the obs index and rloc of every accepted obs of the
// levels whose points lie in [p0, p1) (search_point with dist, as
// k_search_fill); otherwise the counts of every level (as k_search_count).
constexpr int kColMaxLev = 128;

// one wavefront's LDS writes visible to its later reads (program order)
__device__ __forceinline__ void col_wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int lane_mbcnt(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
}

__device__ __forceinline__ long long readlane_ll(long long x, int l) {
  const int lo = __builtin_amdgcn_readlane((int)x, l);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
  return ((long long)hi << 32) | (unsigned)lo;
}

template <typename T, bool FILL, bool FASTLOC = false>
__global__ void __launch_bounds__(256) k_search_col(KParams kp, Tables<T> tb, long long p0, long long p1,
                                                    int32_t* __restrict__ nloc, const int64_t* __restrict__ loff,
                                                    int32_t* __restrict__ lobs, double* __restrict__ lrloc) {
#pragma clang fp contract(off)
  __shared__ int rstart[4][64];        // exclusive scan of the row lengths of a group of 64 rows
  __shared__ long long rib[4][64];     // first ip of each row
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * 4 + w;
  if (i >= kp.nij1) return;
  const long long nij1 = kp.nij1;
  int kbeg = 0, kend = kp.nlev;
  if (FILL) {  // levels k with p = i + nij1 k in [p0, p1)
    kbeg = p0 > i ? (int)((p0 - i + nij1 - 1) / nij1) : 0;
    kend = p1 > i ? (int)((p1 - i + nij1 - 1) / nij1) : 0;
    kend = kend < kp.nlev ? kend : kp.nlev;
    if (kbeg >= kend) return;
  }
  const T dz = (T)kp.dist_zero, dzv = (T)kp.dist_zerov;
  const T pi = (T)kp.pi, re = (T)kp.re;
  const T so = (T)kp.sigma_obs, sv = (T)kp.sigma_obsv;
  const bool vloc = !(kp.sigma_obsv <= 0.0);
  const T hx = dz / tb.dlon1[i], hy = dz / tb.dlat1[i];
  int imin = nint_r<T>(tb.ri1[i] - hx);
  int imax = nint_r<T>(tb.ri1[i] + hx);
  int jmin = nint_r<T>(tb.rj1[i] - hy);
  int jmax = nint_r<T>(tb.rj1[i] + hy);
  imin = imin < 1 ? 1 : imin;
  imax = imax > kp.nlon ? kp.nlon : imax;
  jmin = jmin < 1 ? 1 : jmin;
  jmax = jmax > kp.nlat ? kp.nlat : jmax;
  const T lonb = tb.lon1[i];
  const double bb = (double)(T)(pi * tb.lat1[i]) / 180.0;
  const double sinb = sin(bb), cosb = cos(bb);
  // per level: running count (lane k % 64) and, FILL, the chunk-relative
  // offset of the point's list
  int cnt0 = 0, cnt1 = 0;
  long long base0 = 0, base1 = 0;
  if (FILL) {
    const long long lb = loff[p0];
    if (lane >= kbeg && lane < kend) base0 = loff[i + nij1 * lane] - lb;
    if (lane + 64 >= kbeg && lane + 64 < kend) base1 = loff[i + nij1 * (lane + 64)] - lb;
  }
  for (int jg = jmin; jg <= jmax; jg += 64) {
    // the rows of this group: first ip and length, exclusive scan
    const int j = jg + lane;
    long long ib = 1;
    int len = 0;
    if (j <= jmax) {
      if (imin > 1)
        ib = (long long)tb.nobsgrd[(imin - 2) + (long long)kp.nlon * (j - 1)] + 1;
      else if (j > 1)
        ib = (long long)tb.nobsgrd[(kp.nlon - 1) + (long long)kp.nlon * (j - 2)] + 1;
      const long long ie = tb.nobsgrd[(imax - 1) + (long long)kp.nlon * (j - 1)];
      len = ie >= ib ? (int)(ie - ib + 1) : 0;
    }
    int incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int up = __shfl_up(incl, o);
      if (lane >= o) incl += up;
    }
    const int total = __builtin_amdgcn_readlane(incl, 63);
    rstart[w][lane] = incl - len;
    rib[w][lane] = ib;
    col_wsync();
    for (int v0 = 0; v0 < total; v0 += 64) {
      const int v = v0 + lane;
      const bool valid = v < total;
      // the row holding candidate v: the last row whose start is <= v
      int r = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (rstart[w][r + step] <= v) r += step;
      const int io = valid ? (int)(rib[w][r] + (v - rstart[w][r]) - 1) : 0;
      // horizontal test (search_point's arithmetic)
      bool hok = false;
      T dist = (T)0;
      int el = 0;
      T ol = (T)0;
      if (valid) {
        el = tb.elm[io];
        ol = tb.olev[io];
        const double2 sca = tb.osc[io];
        const double cc = (double)(T)(pi * (T)(lonb - tb.olon[io])) / 180.0;
        const double prod1 = sca.x * sinb;
        const double prod2 = sca.y * cosb;
        T cosd = (T)(prod1 + prod2 * cos(cc));
        if ((double)cosd > 1.0) cosd = (T)1;
        if ((double)cosd < -1.0) cosd = (T)-1;
        hok = !((double)cosd <= kp.cos_out);
        if (hok && (FILL || (double)cosd < kp.cos_in)) {
          dist = re * acos_r<T>(cosd);
          hok = !(dz < dist);
        }
      }
      if (__ballot(hok) == 0) continue;
      for (int k = kbeg; k < kend; ++k) {
        const T rdep = tb.depth1[i + nij1 * k];
        const bool below_surface = k < kp.nlev - 1;
        T dlev;
        if (el == kp.id_z_obs && below_surface)
          dlev = abs_r<T>(rdep);
        else if (el != kp.id_z_obs)
          dlev = abs_r<T>(ol - rdep);
        else
          dlev = (T)0;
        const bool acc = hok && !(dzv < dlev);
        const unsigned long long m = __ballot(acc);
        if (m == 0) continue;
        const int kl = k & 63;
        if (FILL) {
          const int ck = __builtin_amdgcn_readlane(k < 64 ? cnt0 : cnt1, kl);
          const long long bk = readlane_ll(k < 64 ? base0 : base1, kl);
          if (acc) {
            const long long pos = bk + ck + lane_mbcnt(m);
            T q = (dist / so) * (dist / so);
            if (vloc) q = q + (dlev / sv) * (dlev / sv);
            lobs[pos] = io;
            if (FASTLOC)
              lrloc[pos] = (double)__expf(-0.5f * (float)q);
            else
              lrloc[pos] = (double)(T)exp(-0.5 * (double)q);
          }
        }
        const int nacc = __popcll(m);
        if (lane == kl) {
          if (k < 64)
            cnt0 += nacc;
          else
            cnt1 += nacc;
        }
      }
    }
    col_wsync();  // the next row group overwrites rstart / rib
  }
  if (!FILL) {
    if (lane < kp.nlev) nloc[i + nij1 * lane] = cnt0;
    if (lane + 64 < kp.nlev) nloc[i + nij1 * (lane + 64)] = cnt1;
  }
}

// (sin, cos) of every obs latitude in radians, formed as distance_2p forms
// them (common.f90:79-101): aa = real(pi*lat, r_size)/180 in FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_obs_trig(int nobs, double pi_d, const T* __restrict__ olat,
                                                  double2* __restrict__ osc) {
#pragma clang fp contract(off)
  const int io = blockIdx.x * 256 + threadIdx.x;
  if (io >= nobs) return;
  const T pi = (T)pi_d;
  const double aa = (double)(T)(pi * olat[io]) / 180.0;
  osc[io] = make_double2(sin(aa), cos(aa));
}

template <typename T>
hipError_t launch_obs_trig(int nobs, double pi, const T* olat, double2* osc, hipStream_t s) {
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_obs_trig<T>, dim3((unsigned)((nobs + 255) / 256)), dim3(256), 0, s, nobs, pi, olat, osc);
  return hipGetLastError();
}
template hipError_t launch_obs_trig<float>(int, double, const float*, double2*, hipStream_t);
template hipError_t launch_obs_trig<double>(int, double, const double*, double2*, hipStream_t);

template <typename T>
hipError_t launch_search_count(const KParams& kp, const Tables<T>& tb, int32_t* nloc, hipStream_t s) {
  if (kp.nlev <= kColMaxLev && !getenv("LETKF_SEARCH_POINT")) {
    hipLaunchKernelGGL((k_search_col<T, false>), dim3((unsigned)((kp.nij1 + 3) / 4)), dim3(256), 0, s, kp, tb, 0LL,
                       kp.P, nloc, nullptr, nullptr, nullptr);
    return hipGetLastError();
  }
  long long blocks = (kp.P + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_search_count<T>, dim3((unsigned)blocks), dim3(256), 0, s, kp, tb, nloc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_search_fill(const KParams& kp, const Tables<T>& tb, long long p0, long long p1,
                              const int64_t* loff, int32_t* lobs, double* lrloc, hipStream_t s, bool fast_loc) {
  if (p1 <= p0) return hipSuccess;
  if (kp.nlev <= kColMaxLev && !getenv("LETKF_SEARCH_POINT")) {
    if (fast_loc)
      hipLaunchKernelGGL((k_search_col<T, true, true>), dim3((unsigned)((kp.nij1 + 3) / 4)), dim3(256), 0, s, kp, tb,
                         p0, p1, nullptr, loff, lobs, lrloc);
    else
      hipLaunchKernelGGL((k_search_col<T, true, false>), dim3((unsigned)((kp.nij1 + 3) / 4)), dim3(256), 0, s, kp,
                         tb, p0, p1, nullptr, loff, lobs, lrloc);
    return hipGetLastError();
  }
  long long blocks = (p1 - p0 + 255) / 256;
  if (blocks < 1) return hipSuccess;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_search_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, kp, tb, p0, p1, loff,
                     lobs, lrloc);
  return hipGetLastError();
}

template hipError_t launch_search_count<float>(const KParams&, const Tables<float>&, int32_t*, hipStream_t);
template hipError_t launch_search_count<double>(const KParams&, const Tables<double>&, int32_t*, hipStream_t);
template hipError_t launch_search_fill<float>(const KParams&, const Tables<float>&, long long, long long,
                                              const int64_t*, int32_t*, double*, hipStream_t, bool);
template hipError_t launch_search_fill<double>(const KParams&, const Tables<double>&, long long, long long,
                                               const int64_t*, int32_t*, double*, hipStream_t, bool);

// ------------------------------------------------------------------ scans
// Exclusive scans over points: loff (sum of nloc, int64) and actidx (count
// of points with nloc > 0), plus the compacted list act_p of active points.
// Three phases: per-tile totals, one-block scan of tile totals, tile rescans.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ void block_excl_scan2(long long& a, int& b, long long* sa, int* sb,
                                                 long long& tot_a, int& tot_b) {
  const int t = threadIdx.x;
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    long long va = t >= off ? sa[t - off] : 0;
    int vb = t >= off ? sb[t - off] : 0;
    __syncthreads();
    sa[t] += va;
    sb[t] += vb;
    __syncthreads();
  }
  tot_a = sa[kScanThreads - 1];
  tot_b = sb[kScanThreads - 1];
  a = sa[t] - a;
  b = sb[t] - b;
  __syncthreads();
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(long long P, const int32_t* nloc,
                                                             long long* tsum, int* tcnt) {
  __shared__ long long sa[kScanThreads];
  __shared__ int sb[kScanThreads];
  const long long base = (long long)blockIdx.x * kScanTile;
  long long a = 0;
  int b = 0;
  for (int q = 0; q < kScanItems; ++q) {
    long long p = base + (long long)q * kScanThreads + threadIdx.x;
    if (p < P) {
      int n = nloc[p];
      a += n;
      b += n > 0;
    }
  }
  long long ta;
  int tb;
  block_excl_scan2(a, b, sa, sb, ta, tb);
  if (threadIdx.x == 0) {
    tsum[blockIdx.x] = ta;
    tcnt[blockIdx.x] = tb;
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_totals(int ntiles, long long* tsum, int* tcnt) {
  __shared__ long long sa[kScanThreads];
  __shared__ int sb[kScanThreads];
  long long carry_a = 0;
  int carry_b = 0;
  for (int base = 0; base < ntiles; base += kScanThreads) {
    int q = base + threadIdx.x;
    long long a = q < ntiles ? tsum[q] : 0;
    int b = q < ntiles ? tcnt[q] : 0;
    long long ta;
    int tb;
    block_excl_scan2(a, b, sa, sb, ta, tb);
    if (q < ntiles) {
      tsum[q] = a + carry_a;
      tcnt[q] = b + carry_b;
    }
    carry_a += ta;
    carry_b += tb;
  }
  if (threadIdx.x == 0) {
    tsum[ntiles] = carry_a;
    tcnt[ntiles] = carry_b;
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(long long P, const int32_t* nloc,
                                                             const long long* tsum, const int* tcnt,
                                                             int64_t* loff, int32_t* actidx, int32_t* act_p) {
  __shared__ long long sa[kScanThreads];
  __shared__ int sb[kScanThreads];
  const long long base = (long long)blockIdx.x * kScanTile;
  // each thread owns kScanItems consecutive points
  const long long mine = base + (long long)threadIdx.x * kScanItems;
  int n[kScanItems];
  long long a = 0;
  int b = 0;
  for (int q = 0; q < kScanItems; ++q) {
    long long p = mine + q;
    n[q] = p < P ? nloc[p] : 0;
    a += n[q];
    b += n[q] > 0;
  }
  long long ta;
  int tb;
  block_excl_scan2(a, b, sa, sb, ta, tb);
  a += tsum[blockIdx.x];
  b += tcnt[blockIdx.x];
  for (int q = 0; q < kScanItems; ++q) {
    long long p = mine + q;
    if (p < P) {
      loff[p] = a;
      actidx[p] = b;
      if (n[q] > 0) act_p[b] = (int32_t)p;
      a += n[q];
      b += n[q] > 0;
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    loff[P] = tsum[gridDim.x];
    actidx[P] = tcnt[gridDim.x];
  }
}

size_t scan_tmp_bytes(long long P) {
  long long nt = (P + kScanTile - 1) / kScanTile + 1;
  return (size_t)nt * (sizeof(long long) + sizeof(int)) + 64;
}

hipError_t launch_scan_points(long long P, const int32_t* nloc, int64_t* loff, int32_t* actidx,
                              int32_t* act_p, void* tmp, size_t tmp_bytes, hipStream_t s) {
  long long nt = (P + kScanTile - 1) / kScanTile;
  if (nt < 1) nt = 1;
  if (tmp_bytes < scan_tmp_bytes(P)) return hipErrorInvalidValue;
  long long* tsum = (long long*)tmp;
  int* tcnt = (int*)(tsum + nt + 1);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(kScanThreads), 0, s, P, nloc, tsum, tcnt);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, s, (int)nt, tsum, tcnt);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kScanThreads), 0, s, P, nloc, tsum, tcnt,
                     loff, actidx, act_p);
  return hipGetLastError();
}

// ---------------------------------------------------------- obs transpose
// obshdxf(nobs,K) column-major (member-major) -> yT[o*K + m] in FP64, so that
// the K values of one obs are contiguous for the per-point gathers.
template <typename T>
__global__ void __launch_bounds__(256) k_transpose_obs(int nobs, int K, const T* hdxf, double* yT) {
  __shared__ double tile[32][33];
  const int o0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    int o = o0 + tx, m = m0 + r;
    tile[r][tx] = (o < nobs && m < K) ? (double)hdxf[o + (long long)nobs * m] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int o = o0 + r, m = m0 + tx;
    if (o < nobs && m < K) yT[(long long)o * K + m] = tile[tx][r];
  }
}

template <typename T>
hipError_t launch_transpose_obs(int nobs, int K, const T* hdxf, double* yT, hipStream_t s) {
  if (nobs <= 0) return hipSuccess;
  dim3 grid((nobs + 31) / 32, (K + 31) / 32);
  hipLaunchKernelGGL(k_transpose_obs<T>, grid, dim3(256), 0, s, nobs, K, hdxf, yT);
  return hipGetLastError();
}
template hipError_t launch_transpose_obs<float>(int, int, const float*, double*, hipStream_t);
template hipError_t launch_transpose_obs<double>(int, int, const double*, double*, hipStream_t);

template <typename T>
__global__ void k_copy(const T* src, T* dst, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = src[q];
}
template <typename T>
hipError_t launch_copy(const T* src, T* dst, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy<T>, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}
template hipError_t launch_copy<float>(const float*, float*, size_t, hipStream_t);
template hipError_t launch_copy<double>(const double*, double*, size_t, hipStream_t);

}  // namespace letkf
