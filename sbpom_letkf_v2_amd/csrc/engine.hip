// engine.hip -- host side of the C ABI in include/letkf_engine.h: argument
// checks, device workspaces, and the per-analysis pipeline
//   search count -> point scans -> per chunk of points:
//   search fill -> Gram -> eigensolve -> transform (+RTPP/RTPS/round_off)
// that replaces da_main (letkf/run/src/common_da.f90:22-124).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "letkf_handle.h"

using namespace letkf;

static thread_local char g_err[512] = "";

int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int status_to_rc(int status) {
  if (status & ST_NONFINITE) return set_err(LETKF_ERR_NONFINITE, "non-finite ensemble-space matrix");
  if (status & ST_EIG_NOCONV) return set_err(LETKF_ERR_EIG, "eigensolve did not converge");
  if (status & ST_EIG_NEG) return set_err(LETKF_ERR_EIG, "all eigenvalues negative");
  if (status & ST_CAPACITY) return set_err(LETKF_ERR_CAPACITY, "local problem beyond engine capacity");
  return LETKF_OK;
}

static int check_params(const letkf_params_t* p) {
  if (!p) return set_err(LETKF_ERR_ARG, "params is NULL");
  if (p->real_kind != 4 && p->real_kind != 8)
    return set_err(LETKF_ERR_ARG, "real_kind must be 4 or 8 (got %d)", p->real_kind);
  if (p->nmem < 2 || p->nmem > kMaxK)
    return set_err(LETKF_ERR_ARG, "nmem must be in [2,%d] (got %d)", kMaxK, p->nmem);
  if (p->nlon < 2 || p->nlat < 2 || p->nlev < 1 || p->nij1 < 1)
    return set_err(LETKF_ERR_ARG, "bad grid sizes nlon=%d nlat=%d nlev=%d nij1=%d", p->nlon, p->nlat,
                   p->nlev, p->nij1);
  if (p->nv3d < 0 || p->nv2d < 0 || p->nv3d + p->nv2d > 8 || p->nv3d + p->nv2d < 1)
    return set_err(LETKF_ERR_ARG, "nv3d+nv2d must be in [1,8] (got %d+%d)", p->nv3d, p->nv2d);
  if (!(p->cov_infl_mul > 0.0)) return set_err(LETKF_ERR_ARG, "cov_infl_mul must be > 0");
  if (p->workspace_bytes < 0) return set_err(LETKF_ERR_ARG, "workspace_bytes must be >= 0");
  if (p->precision != LETKF_PREC_FP64 && p->precision != LETKF_PREC_FP32 && p->precision != LETKF_PREC_BF16X3 &&
      p->precision != LETKF_PREC_FP32_PANEL && p->precision != LETKF_PREC_FP32X3 &&
      p->precision != LETKF_PREC_FP32_DUAL64 && p->precision != LETKF_PREC_BF16X3_TRI32)
    return set_err(LETKF_ERR_ARG, "unknown precision %d", p->precision);
  if (p->form != LETKF_FORM_PLAIN && p->form != LETKF_FORM_BATCHED && p->form != LETKF_FORM_BATCHED_QL)
    return set_err(LETKF_ERR_ARG, "unknown form %d", p->form);
  if ((long long)p->nij1 * p->nlev > (1LL << 31) - 2)
    return set_err(LETKF_ERR_ARG, "nij1*nlev beyond int32 range");
  return LETKF_OK;
}

extern "C" int letkf_params_default(letkf_params_t* p, int32_t real_kind) {
  if (!p) return set_err(LETKF_ERR_ARG, "params is NULL");
  if (real_kind != 4 && real_kind != 8) return set_err(LETKF_ERR_ARG, "real_kind must be 4 or 8");
  memset(p, 0, sizeof *p);
  auto R = [&](double x) { return real_kind == 4 ? (double)(float)x : x; };
  p->real_kind = real_kind;
  p->iv3d_t = 3;
  p->iv3d_s = 4;
  p->id_z_obs = 2567;
  p->roff = 1;
  p->precision = LETKF_PREC_FP64;
  p->form = LETKF_FORM_PLAIN;
  p->pi = R(4.0 * atan(1.0));
  p->re = R(6371.3e3);
  p->cov_infl_mul = R(1.0);
  p->alpha_rtpp = R(0.9);
  p->alpha_rtps = R(0.0);
  p->sigma_obs = R(300.0e3);
  p->sigma_obsv = R(100.0);
  // REAL(2*sqrt(10/3), r_size) * sigma, product in r_size
  if (real_kind == 4) {
    p->dist_zero = (double)((float)(2.0 * sqrt(10.0 / 3.0)) * (float)300.0e3);
    p->dist_zerov = (double)((float)(2.0 * sqrt(10.0 / 3.0)) * (float)100.0);
  } else {
    p->dist_zero = 2.0 * sqrt(10.0 / 3.0) * 300.0e3;
    p->dist_zerov = 2.0 * sqrt(10.0 / 3.0) * 100.0;
  }
  p->tmin = R(-1.8);
  p->tmax = R(50.0);
  p->smin = R(25.0);
  p->smax = R(50.0);
  return LETKF_OK;
}

extern "C" int letkf_create(const letkf_params_t* p, letkf_handle_t** out) {
  if (!out) return set_err(LETKF_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  letkf_handle* h = new letkf_handle();
  h->prm = *p;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = h->status.ensure(sizeof(int) * 4);
  if (e != hipSuccess) {
    delete h;
    return set_err(LETKF_ERR_HIP, "letkf_create: %s", hipGetErrorString(e));
  }
  *out = h;
  return LETKF_OK;
}

extern "C" int letkf_destroy(letkf_handle_t* h) {
  if (!h) return LETKF_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->release_all();
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return LETKF_OK;
}

extern "C" const char* letkf_last_error(void) { return g_err; }

extern "C" int letkf_get_nobsl(letkf_handle_t* h, int32_t* nobsl, void* stream) {
  if (!h || !nobsl) return set_err(LETKF_ERR_ARG, "NULL argument");
  if (h->nloc_P < 0) return set_err(LETKF_ERR_ARG, "letkf_get_nobsl: no analysis or plan on this handle yet");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  const size_t bytes = sizeof(int32_t) * (size_t)h->nloc_P;
  if (h->nloc_zero)
    HIP_TRY(hipMemsetAsync(nobsl, 0, bytes, s));
  else
    HIP_TRY(hipMemcpyAsync(nobsl, h->nloc.ptr, bytes, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LETKF_OK;
}

extern "C" int letkf_rational_table_entry(int32_t kind, int32_t cls, double* d, double* a, double* s) {
  if (!d || !a || !s) return set_err(LETKF_ERR_ARG, "NULL argument");
  if (kind < 0 || kind >= kRatTableKinds || cls < 0 || cls >= kRatTableClasses)
    return set_err(LETKF_ERR_ARG, "kind must be in [0,%d) and cls in [0,%d)", kRatTableKinds, kRatTableClasses);
  const double* e = rat_table_data() + ((size_t)kind * kRatTableClasses + cls) * (2 + 2 * kRatTablePoles);
  const int n = (int)e[0];
  *d = e[1];
  for (int l = 0; l < n; ++l) {
    a[l] = e[2 + l];
    s[l] = e[2 + kRatTablePoles + l];
  }
  return n;
}

extern "C" int letkf_get_stats(const letkf_handle_t* h, letkf_stats_t* out) {
  if (!h || !out) return set_err(LETKF_ERR_ARG, "NULL argument");
  *out = h->stats;
  return LETKF_OK;
}

static KParams make_kparams(const letkf_params_t& p, int nobs) {
  KParams kp;
  kp.K = p.nmem;
  kp.nlon = p.nlon;
  kp.nlat = p.nlat;
  kp.nlev = p.nlev;
  kp.nij1 = p.nij1;
  kp.nv3d = p.nv3d;
  kp.nv2d = p.nv2d;
  kp.iv3d_t = p.iv3d_t;
  kp.iv3d_s = p.iv3d_s;
  kp.id_z_obs = p.id_z_obs;
  kp.roff = p.roff;
  kp.nobs = nobs;
  kp.P = (long long)p.nij1 * p.nlev;
  kp.pi = p.pi;
  kp.re = p.re;
  kp.rho = p.cov_infl_mul;
  kp.alpha_rtpp = p.alpha_rtpp;
  kp.alpha_rtps = p.alpha_rtps;
  kp.sigma_obs = p.sigma_obs;
  kp.sigma_obsv = p.sigma_obsv;
  kp.dist_zero = p.dist_zero;
  kp.dist_zerov = p.dist_zerov;
  kp.tmin = p.tmin;
  kp.tmax = p.tmax;
  kp.smin = p.smin;
  kp.smax = p.smax;
  kp.cos_in = 2.0;  // no shortcut unless the cut-off angle is well inside (0, pi)
  kp.cos_out = -2.0;
  if (p.dist_zero > 0.0 && p.re > 0.0) {
    const double th = p.dist_zero / p.re, m = p.real_kind == 4 ? 1e-4 : 1e-9;
    if (th * (1.0 + m) < 3.0) {
      kp.cos_in = cos(th * (1.0 - m));
      kp.cos_out = cos(th * (1.0 + m));
    }
  }
  return kp;
}

// Collect loff/actidx at chunk boundaries: out[2c] = loff[b[c]], out[2c+1] = actidx[b[c]].
__global__ void k_gather_bounds(int nb, const long long* pts, const int64_t* loff, const int32_t* actidx,
                                long long* out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nb) {
    out[2 * c] = loff[pts[c]];
    out[2 * c + 1] = actidx[pts[c]];
  }
}

// Per-call point statistics: out[0] max nobsl, out[1] dual-form points,
// out[2] sum of n^2 over dual points, out[3] sum of n over the others
// (Gram flops 2 n^2 K and 2 n K^2), out[4] sum of m^3 over active points,
// m the eigenproblem size.
__global__ void k_point_stats(long long P, long long surf0, const int32_t* nloc, int K, int dual,
                              unsigned long long* out) {
  __shared__ unsigned long long red[6][256];
  unsigned long long a[6] = {0, 0, 0, 0, 0, 0};
  for (long long p = blockIdx.x * 256LL + threadIdx.x; p < P; p += 256LL * gridDim.x) {
    const unsigned long long n = (unsigned long long)nloc[p];
    a[0] = n > a[0] ? n : a[0];
    if (n == 0) {
      a[5] += p >= surf0;  // surface points without local obs
      continue;
    }
    const bool d = dual && n < (unsigned long long)K;
    a[1] += d;
    a[2] += d ? n * n : 0;
    a[3] += d ? 0 : n;
    const unsigned long long m = d ? n : (unsigned long long)K;
    a[4] += m * m * m;
  }
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[0][threadIdx.x] = max(red[0][threadIdx.x], red[0][threadIdx.x + o]);
      for (int q = 1; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(out, red[0][0]);
    for (int q = 1; q < 6; ++q) atomicAdd(out + q, red[q][0]);
  }
}

// Size buckets of the register eigensolve: problem size s (K, or nobsl < K
// in the dual form) goes to bucket min((s-1)/16, 4), i.e. nb = 16, 32, 48,
// 64, 128.
constexpr int kNQ = 5;
static int bucket_nb(int q) { return q < 4 ? 16 * (q + 1) : 128; }

__device__ __forceinline__ int slot_size(int n, int K, int dual) { return dual ? (n < K ? n : K) : K; }
__device__ __forceinline__ int size_bucket(int m) { return (m - 1) / 16 < 4 ? (m - 1) / 16 : 4; }

// Per-(chunk, bucket) slot counts: each block counts 256 consecutive points
// (one chunk: PC is a multiple of 256) with wave ballots into LDS and adds
// one total per bucket (a global atomic per point contended on the few
// bins: 45 ms at 10^7 points); bblk[t * kNQ + q] keeps the count of
// 256-point block t in bucket q for the ordered scatter.
__global__ void __launch_bounds__(256) k_bucket_count(long long P, const int32_t* nloc, int K, int dual,
                                                      long long PC, int32_t* hist, int32_t* bblk) {
  __shared__ int cnt[kNQ];
  const int lane = threadIdx.x & 63;
  for (long long p0 = blockIdx.x * 256LL; p0 < P; p0 += 256LL * gridDim.x) {
    if (threadIdx.x < kNQ) cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long p = p0 + threadIdx.x;
    int q = -1;
    if (p < P) {
      const int n = nloc[p];
      if (n > 0) q = size_bucket(slot_size(n, K, dual));
    }
#pragma unroll
    for (int b = 0; b < kNQ; ++b) {
      const unsigned long long m = __ballot(q == b);
      if (lane == 0 && m) atomicAdd(&cnt[b], __popcll(m));
    }
    __syncthreads();
    if (threadIdx.x < kNQ) {
      bblk[(p0 / 256) * kNQ + threadIdx.x] = cnt[threadIdx.x];
      if (cnt[threadIdx.x]) atomicAdd(&hist[(p0 / PC) * kNQ + threadIdx.x], cnt[threadIdx.x]);
    }
    __syncthreads();
  }
}

// The block counts of each (chunk, bucket) turned into exclusive prefix
// sums in point order, one workgroup per (chunk, bucket): the offset of
// block t's slots within the chunk's list of that bucket.
__global__ void __launch_bounds__(256) k_bucket_blockscan(long long nblk, long long PC, int32_t* bblk) {
  __shared__ int wsum[4];
  const int c = blockIdx.x / kNQ, q = blockIdx.x % kNQ;
  const long long t0 = (long long)c * PC / 256;
  const long long t1 = min(((long long)(c + 1) * PC + 255) / 256, nblk);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int carry = 0;
  for (long long b0 = t0; b0 < t1; b0 += 256) {
    const long long t = b0 + threadIdx.x;
    const int v = t < t1 ? bblk[t * kNQ + q] : 0;
    int x = v;  // inclusive scan over the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int pre = carry;
    for (int u = 0; u < w; ++u) pre += wsum[u];
    if (t < t1) bblk[t * kNQ + q] = pre + x - v;
    carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

// Scatter of active points into per-(chunk, bucket) lists in point order:
// each block takes 256 consecutive points, ranks them per bucket with wave
// ballots and places them at the list start of (chunk, bucket) plus the
// block's prefix offset (bblk after k_bucket_blockscan), so neighbouring
// points stay together in every bucket's kernels (shared obs rows in L2).
__global__ void __launch_bounds__(256) k_bucket_scatter(long long P, const int32_t* nloc, const int32_t* actidx,
                                                        int K, int dual, long long PC, const int32_t* cursor,
                                                        const int32_t* bblk, int32_t* list) {
  __shared__ int wcnt[4][kNQ];
  __shared__ int base[kNQ];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (long long p0 = blockIdx.x * 256LL; p0 < P; p0 += 256LL * gridDim.x) {
    const long long p = p0 + threadIdx.x;
    int q = -1;
    if (p < P) {
      const int n = nloc[p];
      if (n > 0) q = size_bucket(slot_size(n, K, dual));
    }
    // all points of a 256-block lie in one chunk when PC is a multiple of 256
    const long long c = p0 / PC;
    int rank[kNQ];
#pragma unroll
    for (int b = 0; b < kNQ; ++b) {
      const unsigned long long m = __ballot(q == b);
      rank[b] = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
      if (lane == 0) wcnt[wv][b] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kNQ) {
      const int b = threadIdx.x;
      base[b] = cursor[c * kNQ + b] + bblk[(p0 / 256) * kNQ + b];
    }
    __syncthreads();
    if (q >= 0) {
      int pos = base[q] + rank[q];
      for (int w = 0; w < wv; ++w) pos += wcnt[w][q];
      list[pos] = actidx[p] - actidx[c * PC];
    }
    __syncthreads();
  }
}

// The Gram of active slots [a0, a0+na) in the handle's precision: FP64, the
// MFMA kernel (K <= 128), or the MFMA kernel for the K x K slots and the FP64
// kernel for the dual n x n ones.
template <typename T>
static hipError_t launch_gram_any(int prec, const KParams& kp, const Tables<T>& tb, const double* yT,
                                  const int32_t* act_p, const int64_t* loff, long long lbase, const int32_t* lobs,
                                  const double* lrloc, int a0, int na, double* G, double* bvec, int dual,
                                  const void* ysp, hipStream_t s, const int32_t* klist = nullptr,
                                  int kcnt = 0) {
  if (prec == LETKF_PREC_BF16X3_TRI32) prec = LETKF_PREC_BF16X3;  // the Gram of the mode
  if (prec == LETKF_PREC_FP64 || kp.K > 128)
    return launch_gram_fp64<T>(kp, tb, yT, act_p, loff, lbase, lobs, lrloc, a0, na, G, bvec, dual, s);
  if (prec != LETKF_PREC_FP32_DUAL64 || !dual)
    return launch_gram_mixed<T>(prec == LETKF_PREC_FP32_DUAL64 ? LETKF_PREC_FP32 : prec, kp, tb, yT, act_p, loff,
                                lbase, lobs, lrloc, a0, na, G, bvec, dual,
                                prec == LETKF_PREC_BF16X3 ? ysp : nullptr, s,
                                prec == LETKF_PREC_BF16X3 ? klist : nullptr, kcnt);
  hipError_t e = launch_gram_mixed<T>(LETKF_PREC_FP32, kp, tb, yT, act_p, loff, lbase, lobs, lrloc, a0, na, G, bvec,
                                      2, nullptr, s);
  if (e != hipSuccess) return e;
  return launch_gram_fp64<T>(kp, tb, yT, act_p, loff, lbase, lobs, lrloc, a0, na, G, bvec, 2, s);
}

template <typename T>
static Tables<T> make_tables(const letkf_grid_t* g, const letkf_obs_t* o, const double2* osc) {
  Tables<T> tb;
  tb.lon1 = (const T*)g->lon1;
  tb.lat1 = (const T*)g->lat1;
  tb.ri1 = (const T*)g->ri1;
  tb.rj1 = (const T*)g->rj1;
  tb.dlon1 = (const T*)g->dlon1;
  tb.dlat1 = (const T*)g->dlat1;
  tb.phi1 = (const T*)g->phi1;
  tb.depth1 = (const T*)g->depth1;
  tb.nobsgrd = o->nobsgrd;
  tb.elm = o->elm;
  tb.olon = (const T*)o->lon;
  tb.olat = (const T*)o->lat;
  tb.olev = (const T*)o->lev;
  tb.oerr = (const T*)o->err;
  tb.odep = (const T*)o->dep;
  tb.osc = osc;
  return tb;
}

__global__ void k_zero_status(int* st) {
  if (threadIdx.x < 4) st[threadIdx.x] = 0;
}

// Parameters an analysis plan depends on: sizes, form and localisation.
static bool same_geometry_params(const letkf_params_t& a, const letkf_params_t& b) {
  return a.nmem == b.nmem && a.nlon == b.nlon && a.nlat == b.nlat && a.nlev == b.nlev && a.nij1 == b.nij1 &&
         a.nv3d == b.nv3d && a.nv2d == b.nv2d && a.real_kind == b.real_kind && a.form == b.form &&
         a.id_z_obs == b.id_z_obs && a.pi == b.pi && a.re == b.re && a.sigma_obs == b.sigma_obs &&
         a.sigma_obsv == b.sigma_obsv && a.dist_zero == b.dist_zero && a.dist_zerov == b.dist_zerov;
}

static void plan_key(const letkf_grid_t* g, const letkf_obs_t* o, const void* key[12]) {
  const void* k[12] = {g->lon1, g->lat1, g->ri1, g->rj1, g->dlon1, g->dlat1, g->phi1, g->depth1,
                       o->nobsgrd, o->lon, o->lat, o->lev};
  for (int q = 0; q < 12; ++q) key[q] = k[q];
}

// The part of an analysis fixed by the geometry (grid, obs positions and
// levels, localisation scales): local-obs counts and offsets of every
// point, the chunk plan, the size buckets of the active slots and the
// workspaces.  Host read-backs and allocations happen here only.
template <typename T>
static int plan_device(letkf_handle* h, const letkf_grid_t* g, const letkf_obs_t* o, hipStream_t s, bool timing) {
  AnalysisPlan& pl = h->plan;
  pl.valid = false;
  const letkf_params_t& p = h->prm;
  const int K = p.nmem;
  KParams kp = make_kparams(p, o->nobs);
  const long long P = kp.P;
  letkf_stats_t st{};
  st.points = P;
  pl.prm = p;
  pl.nobs = o->nobs;
  plan_key(g, o, pl.key);
  pl.P = P;
  if (o->nobs == 0) {
    h->nloc_P = P;
    h->nloc_zero = true;
    pl.nch = 0;
    pl.st = st;
    pl.valid = true;
    pl.epoch = ++h->plans_made;
    h->invalidated_by = nullptr;
    return LETKF_OK;
  }
  HIP_TRY(h->otrig.ensure(sizeof(double2) * (size_t)(o->nobs + 1)));
  Tables<T> tb = make_tables<T>(g, o, h->otrig.as<double2>());
  HIP_TRY(launch_obs_trig<T>(o->nobs, p.pi, tb.olat, h->otrig.as<double2>(), s));
  HIP_TRY(h->nloc.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->loff.ensure(sizeof(int64_t) * (P + 1)));
  HIP_TRY(h->actidx.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->act_p.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->scan_tmp.ensure(scan_tmp_bytes(P)));
  HIP_TRY(h->yT.ensure(sizeof(double) * (size_t)o->nobs * K));
  const int dual = p.form != LETKF_FORM_PLAIN;
  // batched form, register path: f(T) by rational functions (no QL log)
  const bool rat = p.form == LETKF_FORM_BATCHED && K <= 128;
  // count_nobsl for every point, then the scans
  if (timing) {
    HIP_TRY(h->ensure_events(2));
    HIP_TRY(hipEventRecord(h->events[0], s));
  }
  HIP_TRY(launch_search_count<T>(kp, tb, h->nloc.as<int32_t>(), s));
  h->nloc_P = P;
  h->nloc_zero = false;
  HIP_TRY(launch_scan_points(P, h->nloc.as<int32_t>(), h->loff.as<int64_t>(), h->actidx.as<int32_t>(),
                             h->act_p.as<int32_t>(), h->scan_tmp.ptr, h->scan_tmp.bytes, s));
  if (timing) HIP_TRY(hipEventRecord(h->events[1], s));

  // Chunk of consecutive points: its active slots own a K x K Gram and
  // eigenvector block and the eigensolve's rotation log.  The register
  // eigensolve (K <= 128) launches once per (chunk, size bucket) and needs
  // tens of thousands of matrices per launch to fill the GPU, so its chunks
  // are sized by a workspace budget of up to 96 GB: 45% of the device memory
  // free now plus what this handle's per-slot buffers already hold.  The
  // LDS path keeps 2 GiB per block.
  long long PC;
  if (K <= 128) {
    const double ncol = 1.0 + p.nv3d + p.nv2d;
    const double log_bytes =
        rat ? 0.0 : 16.0 * eigr_rot_capacity(K) + (double)eigr_ix_pitch(K) + 8.0 * eigr_sweep_capacity(K);
    // point-major forecast + analysis copies: per point of the chunk, counted
    // per slot (a chunk has at least as many points as active slots)
    const double pm_bytes = 2.0 * sizeof(T) * K * (ncol - 1);
    const double per_slot = 16.0 * K * K + 8.0 * K * ncol + log_bytes + 64.0 * K + 8.0 + pm_bytes;
    size_t mfree = 0, mtotal = 0;
    HIP_TRY(hipMemGetInfo(&mfree, &mtotal));
    const double held = (double)(h->G.bytes + h->V.bytes + h->yvec.bytes + h->lam.bytes + h->bvec.bytes +
                                 h->eD.bytes + h->eE.bytes + h->logc.bytes + h->logs.bytes + h->logi.bytes +
                                 h->logn.bytes + h->logx.bytes + h->logr.bytes + h->xpm.bytes + h->apm.bytes);
    double budget = std::min(96e9, std::max(1e9, 0.45 * ((double)mfree + held)));
    if (p.workspace_bytes > 0)  // the caller's cap (room for several handles, or for the caller's own data)
      budget = std::min(std::max(1e9, (double)p.workspace_bytes), 0.9 * ((double)mfree + held));
    PC = (long long)(budget / per_slot) / 256 * 256;  // multiple of 256 (k_bucket_scatter)
  } else {
    PC = (long long)((1ULL << 31) / (sizeof(double) * (size_t)K * K));
    PC = std::min(PC, 65536LL);
  }
  PC = std::max(1024LL, std::min(PC, P));
  const long long nch = (P + PC - 1) / PC;
  std::vector<long long>& bpts = pl.bpts;
  bpts.assign(nch + 1, 0);
  for (long long c = 0; c <= nch; ++c) bpts[c] = std::min(P, c * PC);
  HIP_TRY(h->bounds.ensure(sizeof(long long) * (3 * (nch + 1) + 6)));
  long long* d_pts = h->bounds.as<long long>();
  long long* d_out = d_pts + (nch + 1);
  unsigned long long* d_pst = (unsigned long long*)(d_out + 2 * (nch + 1));
  HIP_TRY(hipMemsetAsync(d_pst, 0, 6 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_point_stats, dim3(1024), dim3(256), 0, s, P, (long long)p.nij1 * (p.nlev - 1),
                     h->nloc.as<int32_t>(), K, dual, d_pst);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(d_pts, bpts.data(), sizeof(long long) * (nch + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_bounds, dim3((unsigned)((nch + 1 + 255) / 256)), dim3(256), 0, s, (int)(nch + 1),
                     d_pts, h->loff.as<int64_t>(), h->actidx.as<int32_t>(), d_out);
  HIP_TRY(hipGetLastError());
  std::vector<long long>& bv = pl.bv;
  bv.assign(2 * (nch + 1) + 6, 0);
  HIP_TRY(hipMemcpyAsync(bv.data(), d_out, sizeof(long long) * (2 * (nch + 1) + 6), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  st.max_nobsl = bv[2 * (nch + 1)];
  st.dual_points = dual ? bv[2 * (nch + 1) + 1] : 0;
  long long maxL = 0, maxA = 0;
  for (long long c = 0; c < nch; ++c) {
    maxL = std::max(maxL, bv[2 * (c + 1)] - bv[2 * c]);
    maxA = std::max(maxA, bv[2 * (c + 1) + 1] - bv[2 * c + 1]);
  }
  st.local_obs = bv[2 * nch];
  st.active_points = bv[2 * nch + 1];
  // register eigensolve for K <= 128: bucket the active slots of every chunk
  // by problem size; bucket (c, q) occupies list[boff[c*kNQ+q] ...)
  const bool reg_eig = K <= 128;
  std::vector<int32_t>& hist = pl.hist;
  std::vector<int32_t>& boff = pl.boff;
  hist.clear();
  boff.clear();
  if (reg_eig) {
    const size_t nh = (size_t)nch * kNQ;
    HIP_TRY(h->bhist.ensure(sizeof(int32_t) * nh));
    HIP_TRY(h->bcur.ensure(sizeof(int32_t) * nh));
    HIP_TRY(h->blist.ensure(sizeof(int32_t) * (size_t)(st.active_points + 1)));
    const long long nblk = (P + 255) / 256;
    HIP_TRY(h->bblk.ensure(sizeof(int32_t) * (size_t)nblk * kNQ));
    HIP_TRY(hipMemsetAsync(h->bhist.ptr, 0, sizeof(int32_t) * nh, s));
    hipLaunchKernelGGL(k_bucket_count, dim3(2048), dim3(256), 0, s, P, h->nloc.as<int32_t>(), K, dual, PC,
                       h->bhist.as<int32_t>(), h->bblk.as<int32_t>());
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_bucket_blockscan, dim3((unsigned)nh), dim3(256), 0, s, nblk, PC, h->bblk.as<int32_t>());
    HIP_TRY(hipGetLastError());
    hist.resize(nh);
    boff.resize(nh);
    HIP_TRY(hipMemcpyAsync(hist.data(), h->bhist.ptr, sizeof(int32_t) * nh, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int32_t run = 0;
    for (size_t q = 0; q < nh; ++q) {
      boff[q] = run;
      run += hist[q];
      st.bucket_slots[q % kNQ] += hist[q];
    }
    HIP_TRY(hipMemcpyAsync(h->bcur.ptr, boff.data(), sizeof(int32_t) * nh, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bucket_scatter, dim3(2048), dim3(256), 0, s, P, h->nloc.as<int32_t>(),
                       h->actidx.as<int32_t>(), K, dual, PC, h->bcur.as<int32_t>(), h->bblk.as<int32_t>(),
                       h->blist.as<int32_t>());
    HIP_TRY(hipGetLastError());
    // the host copy of boff must stay alive until the copy has run
    HIP_TRY(hipStreamSynchronize(s));
  }
  HIP_TRY(h->lobs.ensure(sizeof(int32_t) * (maxL + 1)));
  HIP_TRY(h->lrloc.ensure(sizeof(double) * (maxL + 1)));
  HIP_TRY(h->G.ensure(sizeof(double) * (size_t)(maxA + 1) * K * K));
  HIP_TRY(h->V.ensure(sizeof(double) * (size_t)(maxA + 1) * K * K));
  // vector blocks of the register path (rows x (1 + nv3d + nv2d) per slot)
  const long long yp = (long long)K * (1 + p.nv3d + p.nv2d);
  if (reg_eig) HIP_TRY(h->yvec.ensure(sizeof(double) * (size_t)(maxA + 1) * yp));
  HIP_TRY(h->lam.ensure(sizeof(double) * (size_t)(maxA + 1) * K));
  HIP_TRY(h->bvec.ensure(sizeof(double) * (size_t)(maxA + 1) * K));
  const long long cap = reg_eig ? 2 * eigr_rot_capacity(K) : eig_log_capacity(K);   // doubles (c,s)
  const long long hcap = reg_eig ? 2 * eigr_sweep_capacity(K) : eig_log_capacity(K); // int32 (col,len)
  HIP_TRY(h->eD.ensure(sizeof(double) * (size_t)(maxA + 1) * K));
  HIP_TRY(h->eE.ensure(sizeof(double) * (size_t)(maxA + 1) * K));
  HIP_TRY(h->logn.ensure(sizeof(int32_t) * (size_t)(maxA + 1)));  // per slot: sweeps logged, or -1 fallback
  if (!rat) {  // the QL rotation log
    HIP_TRY(h->logc.ensure(sizeof(double) * (size_t)(maxA + 1) * cap));
    if (!reg_eig) HIP_TRY(h->logs.ensure(sizeof(double) * (size_t)(maxA + 1) * cap));
    HIP_TRY(h->logi.ensure(sizeof(int32_t) * (size_t)(maxA + 1) * hcap));
    if (reg_eig) {
      HIP_TRY(h->logx.ensure((size_t)(maxA + 1) * eigr_ix_pitch(K)));
      HIP_TRY(h->logr.ensure(sizeof(int32_t) * (size_t)(maxA + 1)));
    }
  }
  // point-major copies of each chunk's ensemble for the register update
  // (test hook LETKF_NO_POINT_MAJOR: the update on the caller's arrays)
  pl.pm_ld = 0;
  if (reg_eig && !getenv("LETKF_NO_POINT_MAJOR")) {
    pl.pm_ld = K * (p.nv3d + p.nv2d);
    HIP_TRY(h->xpm.ensure(sizeof(T) * (size_t)PC * pl.pm_ld));
    HIP_TRY(h->apm.ensure(sizeof(T) * (size_t)PC * pl.pm_ld));
    // and the forecast statistics the copy forms (nmem <= 64; test hook
    // LETKF_NO_PM_STATS: the vector kernels form them)
    pl.pm_stats = K <= 64 && !getenv("LETKF_NO_PM_STATS");
    if (pl.pm_stats) HIP_TRY(h->pst.ensure(sizeof(double2) * (size_t)PC * (p.nv3d + p.nv2d)));
  }
  // a reduced-precision Gram resolves the weights only while the spectrum
  // bound of the slot's tridiagonal matrix stays within kmax of beta (the
  // analysis error grows with it, ~1e-8 of the field scale per unit with the
  // BF16x3 Gram, ~1e-9 with the FP32 ones); wider slots go to the fallback
  // on an FP64 Gram.  LETKF_KAPPA_MAX overrides.
  pl.kmax = p.precision == LETKF_PREC_FP64 ? 1e308
            : (p.precision == LETKF_PREC_BF16X3 || p.precision == LETKF_PREC_BF16X3_TRI32) ? 1024.0
                                                                                              : 8192.0;
  if (const char* env = getenv("LETKF_KAPPA_MAX")) {
    const double v = atof(env);
    if (v > 1.0) pl.kmax = v;
  }
  // rloc from the FP32 exponential in precision bf16x3t (test hook
  // LETKF_EXACT_RLOC: the FP64 one)
  pl.fastloc = p.precision == LETKF_PREC_BF16X3_TRI32 && !getenv("LETKF_EXACT_RLOC");
  // the dual-form BF16x3 Gram's operands split once per analysis (test hook
  // LETKF_GRAM_NOSPLIT: split per gather from the FP64 rows)
  pl.ysplit = dual && (p.precision == LETKF_PREC_BF16X3 || p.precision == LETKF_PREC_BF16X3_TRI32) && K <= 128 &&
              !getenv("LETKF_GRAM_BLOCK") && !getenv("LETKF_GRAM_NOSPLIT");
  if (pl.ysplit) HIP_TRY(h->ysp.ensure((size_t)32 * split_obs_groups(K) * (size_t)o->nobs));
  if (rat && !h->ratt_ok) {  // the coefficient table, once per handle (host copy built once per process)
    HIP_TRY(h->ratt.ensure(sizeof(double) * rat_table_size()));
    HIP_TRY(hipMemcpyAsync(h->ratt.ptr, rat_table_data(), sizeof(double) * rat_table_size(),
                           hipMemcpyHostToDevice, s));
    h->ratt_ok = true;
  }
  st.chunks = nch;
  st.gram_flops = 2.0 * K * ((double)bv[2 * (nch + 1) + 2] + (double)K * (double)bv[2 * (nch + 1) + 3]);
  st.eig_flops = 4.0 / 3.0 * (double)bv[2 * (nch + 1) + 4];
  // the tile transform reads and writes the ensemble of the points it
  // updates: every point on the LDS path, only points without local obs
  // (and the rare overflowed eigensolves, not counted) on the register path
  if (reg_eig) {
    const double noobs = (double)(P - st.active_points), noobs_surf = (double)bv[2 * (nch + 1) + 5];
    st.transform_bytes = 2.0 * sizeof(T) * K * (noobs * p.nv3d + noobs_surf * p.nv2d);
  } else {
    const double n3 = (double)P * K * p.nv3d, n2 = (double)p.nij1 * K * p.nv2d;
    st.transform_bytes = 2.0 * sizeof(T) * (n3 + n2);
  }
  pl.PC = PC;
  pl.nch = nch;
  pl.yp = yp;
  pl.reg_eig = reg_eig;
  pl.dual = dual;
  pl.rat = rat;
  pl.st = st;
  pl.valid = true;
  pl.epoch = ++h->plans_made;
  h->invalidated_by = nullptr;
  return LETKF_OK;
}

// The part of an analysis that depends on the values (forecast ensemble,
// innovations, obs errors and perturbations, inflation): only kernel
// launches on `s` -- no allocation, no host synchronisation, no read-back,
// no memset -- so it can be captured in a hipGraph.  Status bits stay in
// the handle's device word.  timing (not under capture): stage events, and
// per-kernel events of the eigensolve when ktm is given.
template <typename T>
static int enqueue_device(letkf_handle* h, const letkf_grid_t* g, const letkf_obs_t* o, const T* f3, const T* f2,
                          T* a3, T* a2, hipStream_t s, bool timing, EigTimer* ktm) {
  const AnalysisPlan& pl = h->plan;
  const letkf_params_t& p = h->prm;
  const int K = p.nmem;
  KParams kp = make_kparams(p, o->nobs);
  const long long P = kp.P;
  hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, h->status.as<int>());
  HIP_TRY(hipGetLastError());
  if (o->nobs == 0) {
    // common_da.f90:57-61: analysis = forecast, no inflation, no round_off
    HIP_TRY(launch_copy<T>(f3, a3, (size_t)P * K * p.nv3d, s));
    HIP_TRY(launch_copy<T>(f2, a2, (size_t)p.nij1 * K * p.nv2d, s));
    return LETKF_OK;
  }
  Tables<T> tb = make_tables<T>(g, o, h->otrig.as<double2>());
  const int dual = pl.dual;
  const bool reg_eig = pl.reg_eig;
  const long long nch = pl.nch, yp = pl.yp;
  const std::vector<long long>& bpts = pl.bpts;
  const std::vector<long long>& bv = pl.bv;
  auto mark = [&](size_t q) -> hipError_t { return timing ? hipEventRecord(h->events[q], s) : hipSuccess; };
  // obs-major FP64 copy of obshdxf (and its BF16 split for the dual Gram)
  HIP_TRY(launch_transpose_obs<T>(o->nobs, K, (const T*)o->hdxf, h->yT.as<double>(), s));
  const void* ysp = pl.ysplit ? h->ysp.ptr : nullptr;
  if (ysp) HIP_TRY(launch_split_obs(o->nobs, K, h->yT.as<double>(), h->ysp.ptr, s));
  for (long long c = 0; c < nch; ++c) {
    const size_t e0 = 2 + 5 * (size_t)c;
    const long long p0 = bpts[c], p1 = bpts[c + 1];
    const long long lbase = bv[2 * c];
    const int a0 = (int)bv[2 * c + 1];
    const int na = (int)(bv[2 * (c + 1) + 1] - a0);
    HIP_TRY(mark(e0));
    if (na > 0) {
      HIP_TRY(launch_search_fill<T>(kp, tb, p0, p1, h->loff.as<int64_t>(), h->lobs.as<int32_t>(),
                                    h->lrloc.as<double>(), s, pl.fastloc));
    }
    HIP_TRY(mark(e0 + 1));
    if (na > 0) {
      // with the size buckets planned, the K x K Gram kernels visit the
      // slots of the bucket that holds size K only
      const size_t qk = (size_t)c * kNQ + (K <= 64 ? (K - 1) / 16 : 4);
      const bool bl = reg_eig && qk < pl.hist.size();
      HIP_TRY(launch_gram_any<T>(p.precision, kp, tb, h->yT.as<double>(), h->act_p.as<int32_t>(),
                                 h->loff.as<int64_t>(), lbase, h->lobs.as<int32_t>(), h->lrloc.as<double>(), a0, na,
                                 h->G.as<double>(), h->bvec.as<double>(), dual, ysp, s,
                                 bl ? h->blist.as<int32_t>() + pl.boff[qk] : nullptr, bl ? pl.hist[qk] : 0));
    }
    HIP_TRY(mark(e0 + 2));
    PointMajor<T> pm{nullptr, nullptr, p0, pl.pm_ld};
    if (na > 0 && reg_eig && pl.pm_ld > 0) {
      pm.X = h->xpm.as<T>();
      pm.A = h->apm.as<T>();
      double2* st = pl.pm_stats ? h->pst.as<double2>() : nullptr;
      pm.S = st;
      HIP_TRY(launch_to_point_major<T>(kp, p0, p1, f3, f2, h->xpm.as<T>(), pl.pm_ld, s, st));
    }
    if (na > 0) {
      if (reg_eig) {
        for (int q = 0; q < kNQ; ++q) {
          const int cnt = pl.hist[(size_t)c * kNQ + q];
          if (cnt == 0) continue;
          HIP_TRY(launch_eigr_update<T>(bucket_nb(q), K, dual, h->nloc.as<int32_t>(), h->act_p.as<int32_t>(), a0,
                                        h->blist.as<int32_t>() + pl.boff[(size_t)c * kNQ + q], cnt,
                                        h->G.as<double>(), h->V.as<double>(), h->eD.as<double>(),
                                        h->eE.as<double>(), h->logc.as<double>(), h->logx.as<uint8_t>(),
                                        h->logi.as<int32_t>(), h->logn.as<int32_t>(), h->logr.as<int32_t>(),
                                        h->lam.as<double>(), h->V.as<double>(), h->yvec.as<double>(), yp, kp, tb, f3,
                                        f2, a3, a2, h->loff.as<int64_t>(), lbase, h->lobs.as<int32_t>(),
                                        h->lrloc.as<double>(), h->yT.as<double>(), h->bvec.as<double>(),
                                        h->status.as<int>(), h->vpath[q],
                                        pl.rat ? h->ratt.as<double>() : nullptr,
                                        p.precision == LETKF_PREC_BF16X3_TRI32, ktm, pm, pl.kmax, s));
        }
      } else {
        HIP_TRY(launch_eig3(K, na, h->G.as<double>(), h->V.as<double>(), h->eD.as<double>(), h->eE.as<double>(),
                            h->logc.as<double>(), h->logs.as<double>(), h->logi.as<int32_t>(),
                            h->logn.as<int32_t>(), h->lam.as<double>(), h->V.as<double>(), h->status.as<int>(),
                            dual, h->nloc.as<int32_t>(), h->act_p.as<int32_t>(), a0, s));
      }
    }
    HIP_TRY(mark(e0 + 3));
    if (pm.A) HIP_TRY(launch_from_point_major<T>(kp, p0, p1, pm.A, pl.pm_ld, h->nloc.as<int32_t>(), a3, a2, s));
    HIP_TRY(launch_transform<T>(kp, tb, p0, p1, h->nloc.as<int32_t>(), h->actidx.as<int32_t>(), a0,
                                h->lam.as<double>(), h->V.as<double>(), h->bvec.as<double>(), f3, f2, a3, a2, dual,
                                h->loff.as<int64_t>(), lbase, h->lobs.as<int32_t>(), h->lrloc.as<double>(),
                                h->yT.as<double>(), reg_eig ? h->logn.as<int32_t>() : nullptr,
                                h->status.as<int>() + 1, s));
    HIP_TRY(mark(e0 + 4));
  }
  return LETKF_OK;
}

// One analysis: plan, enqueue, wait, status and stats.
template <typename T>
static int run_device(letkf_handle* h, const letkf_grid_t* g, const letkf_obs_t* o, const T* f3, const T* f2, T* a3,
                      T* a2, hipStream_t s) {
  const letkf_params_t& p = h->prm;
  const bool timing = p.stage_timing != 0;
  // the two events live until every return below
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } evs;
  HIP_TRY(hipEventCreate(&evs.e[0]));
  HIP_TRY(hipEventCreate(&evs.e[1]));
  hipEvent_t ev0 = evs.e[0], ev1 = evs.e[1];
  HIP_TRY(hipEventRecord(ev0, s));
  int rc = plan_device<T>(h, g, o, s, timing && o->nobs > 0);
  if (rc) return rc;
  const AnalysisPlan& pl = h->plan;
  const long long nch = pl.nch;
  if (timing && o->nobs > 0) HIP_TRY(h->ensure_events(2 + 5 * (size_t)nch));
  // per-kernel marks of the register eigensolve: 5 kernels per bucket launch
  EigTimer ktm{nullptr, nullptr, 0, 0, nullptr};
  if (p.stage_timing >= 2 && pl.reg_eig && o->nobs > 0) {
    const int cap = (int)(5 * kNQ * nch);
    HIP_TRY(h->ensure_kevents(2 * (size_t)cap));
    h->ktags.assign(cap, 0);
    HIP_TRY(h->rotsum.ensure(sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(h->rotsum.ptr, 0, sizeof(unsigned long long), s));
    ktm = EigTimer{h->kevents.data(), h->ktags.data(), cap, 0, h->rotsum.as<unsigned long long>()};
  }
  rc = enqueue_device<T>(h, g, o, f3, f2, a3, a2, s, timing && o->nobs > 0, ktm.ev ? &ktm : nullptr);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(ev1, s));
  int stat4[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(stat4, h->status.ptr, sizeof(stat4), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  letkf_stats_t st = pl.st;
  const int status = stat4[0];
  st.fallback_slots = stat4[2];
  if (ktm.rot) {
    unsigned long long nrot_total = 0;
    HIP_TRY(hipMemcpy(&nrot_total, ktm.rot, sizeof(nrot_total), hipMemcpyDeviceToHost));
    st.rotations = (int64_t)nrot_total;
  }
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  st.ms_total = ms;
  if (timing && o->nobs > 0) {
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, h->events[0], h->events[1]));
    st.ms_search = t;
    for (long long c = 0; c < nch; ++c) {
      const size_t e0 = 2 + 5 * (size_t)c;
      double* dst[4] = {&st.ms_search, &st.ms_gram, &st.ms_eig, &st.ms_transform};
      for (int q = 0; q < 4; ++q) {
        HIP_TRY(hipEventElapsedTime(&t, h->events[e0 + q], h->events[e0 + q + 1]));
        *dst[q] += t;
      }
    }
  }
  for (int u = 0; u < ktm.used; ++u) {
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, ktm.ev[2 * u], ktm.ev[2 * u + 1]));
    st.ms_eigk[ktm.tag[u]] += t;
    st.n_eigk[ktm.tag[u]] += 1;
  }
  h->stats = st;
  return status_to_rc(status);
}

// Adjoint of run_device<double> (see adjoint.inc): search and FP64 Gram in
// the K x K form per chunk, then the per-point backward kernels.
static int run_adjoint(letkf_handle* h, const letkf_grid_t* g, const letkf_obs_t* o, const double* f3,
                       const double* f2, const double* b3, const double* b2, double* gdep, double* glogvar,
                       double* gsc, hipStream_t s) {
  const letkf_params_t& p = h->prm;
  const int K = p.nmem;
  KParams kp = make_kparams(p, o->nobs);
  const long long P = kp.P;
  HIP_TRY(hipMemsetAsync(gsc, 0, 4 * sizeof(double), s));
  if (o->nobs > 0) {
    HIP_TRY(hipMemsetAsync(gdep, 0, sizeof(double) * o->nobs, s));
    HIP_TRY(hipMemsetAsync(glogvar, 0, sizeof(double) * o->nobs, s));
  }
  if (o->nobs == 0) {  // analysis = forecast (common_da.f90:57-61): no dependence
    h->nloc_P = P;
    h->nloc_zero = true;
    HIP_TRY(hipStreamSynchronize(s));
    return LETKF_OK;
  }
  Tables<double> tb;
  tb.lon1 = (const double*)g->lon1;
  tb.lat1 = (const double*)g->lat1;
  tb.ri1 = (const double*)g->ri1;
  tb.rj1 = (const double*)g->rj1;
  tb.dlon1 = (const double*)g->dlon1;
  tb.dlat1 = (const double*)g->dlat1;
  tb.phi1 = (const double*)g->phi1;
  tb.depth1 = (const double*)g->depth1;
  tb.nobsgrd = o->nobsgrd;
  tb.elm = o->elm;
  tb.olon = (const double*)o->lon;
  tb.olat = (const double*)o->lat;
  tb.olev = (const double*)o->lev;
  tb.oerr = (const double*)o->err;
  tb.odep = (const double*)o->dep;
  HIP_TRY(h->otrig.ensure(sizeof(double2) * (size_t)(o->nobs + 1)));
  tb.osc = h->otrig.as<double2>();
  HIP_TRY(launch_obs_trig<double>(o->nobs, p.pi, tb.olat, h->otrig.as<double2>(), s));
  HIP_TRY(h->nloc.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->loff.ensure(sizeof(int64_t) * (P + 1)));
  HIP_TRY(h->actidx.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->act_p.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->scan_tmp.ensure(scan_tmp_bytes(P)));
  HIP_TRY(h->yT.ensure(sizeof(double) * (size_t)o->nobs * K));
  HIP_TRY(h->adj_part.ensure(sizeof(double) * (4 * (size_t)P + 1024)));
  HIP_TRY(hipMemsetAsync(h->status.ptr, 0, sizeof(int) * 4, s));
  HIP_TRY(hipMemsetAsync(h->adj_part.ptr, 0, sizeof(double) * 4 * (size_t)P, s));
  HIP_TRY(launch_transpose_obs<double>(o->nobs, K, (const double*)o->hdxf, h->yT.as<double>(), s));
  HIP_TRY(launch_search_count<double>(kp, tb, h->nloc.as<int32_t>(), s));
  h->nloc_P = P;
  h->nloc_zero = false;
  HIP_TRY(launch_scan_points(P, h->nloc.as<int32_t>(), h->loff.as<int64_t>(), h->actidx.as<int32_t>(),
                             h->act_p.as<int32_t>(), h->scan_tmp.ptr, h->scan_tmp.bytes, s));
  // chunks with a K x K Gram block per active slot (~2 GiB)
  long long PC = (long long)((1ULL << 31) / (sizeof(double) * (size_t)K * K));
  PC = std::max(1024LL, std::min(PC, P));
  const long long nch = (P + PC - 1) / PC;
  std::vector<long long> bpts(nch + 1);
  for (long long c = 0; c <= nch; ++c) bpts[c] = std::min(P, c * PC);
  HIP_TRY(h->bounds.ensure(sizeof(long long) * (3 * (nch + 1) + 5)));
  long long* d_pts = h->bounds.as<long long>();
  long long* d_out = d_pts + (nch + 1);
  HIP_TRY(hipMemcpyAsync(d_pts, bpts.data(), sizeof(long long) * (nch + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_bounds, dim3((unsigned)((nch + 1 + 255) / 256)), dim3(256), 0, s, (int)(nch + 1),
                     d_pts, h->loff.as<int64_t>(), h->actidx.as<int32_t>(), d_out);
  HIP_TRY(hipGetLastError());
  std::vector<long long> bv(2 * (nch + 1));
  HIP_TRY(hipMemcpyAsync(bv.data(), d_out, sizeof(long long) * 2 * (nch + 1), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  long long maxL = 0, maxA = 0;
  for (long long c = 0; c < nch; ++c) {
    maxL = std::max(maxL, bv[2 * (c + 1)] - bv[2 * c]);
    maxA = std::max(maxA, bv[2 * (c + 1) + 1] - bv[2 * c + 1]);
  }
  HIP_TRY(h->lobs.ensure(sizeof(int32_t) * (maxL + 1)));
  HIP_TRY(h->lrloc.ensure(sizeof(double) * (maxL + 1)));
  HIP_TRY(h->G.ensure(sizeof(double) * (size_t)(maxA + 1) * K * K));
  HIP_TRY(h->bvec.ensure(sizeof(double) * (size_t)(maxA + 1) * K));
  // deterministic per-obs sums: pair terms, (obs, position) sort keys and values, sort scratch
  int nobs_bits = 1;
  while (nobs_bits < 31 && (1LL << nobs_bits) <= o->nobs) ++nobs_bits;
  const size_t sort_tmp = adjoint_sort_tmp_bytes(maxL, nobs_bits);
  HIP_TRY(h->adj_pair.ensure(sizeof(double) * 2 * (size_t)(maxL + 1)));
  HIP_TRY(h->adj_sort.ensure(sizeof(int32_t) * 3 * (size_t)(maxL + 1) + sort_tmp + 256));
  int32_t* kv = h->adj_sort.as<int32_t>();
  void* sort_buf = (void*)(((uintptr_t)(kv + 3 * (maxL + 1)) + 255) & ~(uintptr_t)255);
  for (long long c = 0; c < nch; ++c) {
    const long long p0 = bpts[c], p1 = bpts[c + 1];
    const long long lbase = bv[2 * c];
    const long long nl = bv[2 * (c + 1)] - lbase;
    const int a0 = (int)bv[2 * c + 1];
    const int na = (int)(bv[2 * (c + 1) + 1] - a0);
    if (na > 0) {
      HIP_TRY(launch_search_fill<double>(kp, tb, p0, p1, h->loff.as<int64_t>(), h->lobs.as<int32_t>(),
                                         h->lrloc.as<double>(), s));
      HIP_TRY(launch_gram_fp64<double>(kp, tb, h->yT.as<double>(), h->act_p.as<int32_t>(), h->loff.as<int64_t>(),
                                       lbase, h->lobs.as<int32_t>(), h->lrloc.as<double>(), a0, na,
                                       h->G.as<double>(), h->bvec.as<double>(), 0, s));
    }
    HIP_TRY(launch_adjoint_points(kp, tb, f3, f2, b3, b2, h->nloc.as<int32_t>(), h->loff.as<int64_t>(), lbase,
                                  h->lobs.as<int32_t>(), h->lrloc.as<double>(), h->yT.as<double>(),
                                  h->G.as<double>(), h->bvec.as<double>(), h->act_p.as<int32_t>(), a0, na, p0, p1,
                                  h->adj_pair.as<double>(), h->adj_part.as<double>(), h->status.as<int>(), s));
    if (na > 0)
      HIP_TRY(launch_adjoint_pairs(nl, nobs_bits, h->lobs.as<int32_t>(), h->adj_pair.as<double>(), kv, sort_buf,
                                   sort_tmp, gdep, glogvar, s));
  }
  HIP_TRY(launch_adjoint_reduce(h->adj_part.as<double>(), P, h->adj_part.as<double>() + 4 * (size_t)P, gsc, s));
  int status = 0;
  HIP_TRY(hipMemcpyAsync(&status, h->status.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return status_to_rc(status);
}

// da_main with the LPF / LPFGM core (iswitch_da 2/3): search count, point
// scans, then per chunk the search fill, for LPFGM the FP64 Gram with alpha
// = (K-1)/cov_infl_gm, and the per-point core + ensemble update kernel.
template <typename T>
static int run_lpfgm(letkf_handle* h, const letkf_grid_t* g, const letkf_obs_t* o, const letkf_lpfgm_cfg_t* cfg,
                     const T* f3, const T* f2, T* a3, T* a2, hipStream_t s) {
  const letkf_params_t& p = h->prm;
  const int K = p.nmem;
  KParams kp = make_kparams(p, o->nobs);
  const long long P = kp.P;
  if (o->nobs == 0) {  // analysis = forecast (common_da.f90:57-61)
    h->nloc_P = P;
    h->nloc_zero = true;
    if (p.nv3d) HIP_TRY(hipMemcpyAsync(a3, f3, sizeof(T) * P * K * p.nv3d, hipMemcpyDeviceToDevice, s));
    if (p.nv2d)
      HIP_TRY(hipMemcpyAsync(a2, f2, sizeof(T) * (size_t)p.nij1 * K * p.nv2d, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return LETKF_OK;
  }
  LpfPointArgs la{cfg->mode, cfg->nmonte, cfg->dp ? 1 : 0, 0.0, nullptr, nullptr, h->status.as<int>()};
  hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, h->status.as<int>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(lpf_prepare(K, cfg->nmonte, cfg->rm_sr, sizeof(T) == 8, cfg->rand, h->bin_tmp, s, &la.rtab, &la.ttab));
  la.alpha = (double)((T)(K - 1) / (T)cfg->cov_infl_gm);
  KParams kg = kp;
  kg.rho = cfg->cov_infl_gm;  // the Gram's diagonal (K-1)/rho
  HIP_TRY(h->otrig.ensure(sizeof(double2) * (size_t)(o->nobs + 1)));
  HIP_TRY(launch_obs_trig<T>(o->nobs, p.pi, (const T*)o->lat, h->otrig.as<double2>(), s));
  const Tables<T> tb = make_tables<T>(g, o, h->otrig.as<double2>());
  HIP_TRY(h->nloc.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->loff.ensure(sizeof(int64_t) * (P + 1)));
  HIP_TRY(h->actidx.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->act_p.ensure(sizeof(int32_t) * (P + 1)));
  HIP_TRY(h->scan_tmp.ensure(scan_tmp_bytes(P)));
  HIP_TRY(h->yT.ensure(sizeof(double) * (size_t)o->nobs * K));
  HIP_TRY(launch_transpose_obs<T>(o->nobs, K, (const T*)o->hdxf, h->yT.as<double>(), s));
  HIP_TRY(launch_search_count<T>(kp, tb, h->nloc.as<int32_t>(), s));
  h->nloc_P = P;
  h->nloc_zero = false;
  HIP_TRY(launch_scan_points(P, h->nloc.as<int32_t>(), h->loff.as<int64_t>(), h->actidx.as<int32_t>(),
                             h->act_p.as<int32_t>(), h->scan_tmp.ptr, h->scan_tmp.bytes, s));
  // chunks with a K x K slot per active point (~2 GiB)
  long long PC = (long long)((1ULL << 31) / (sizeof(double) * (size_t)K * K));
  PC = std::max(1024LL, std::min(PC, P));
  const long long nch = (P + PC - 1) / PC;
  std::vector<long long> bpts(nch + 1);
  for (long long c = 0; c <= nch; ++c) bpts[c] = std::min(P, c * PC);
  HIP_TRY(h->bounds.ensure(sizeof(long long) * (3 * (nch + 1) + 5)));
  long long* d_pts = h->bounds.as<long long>();
  long long* d_out = d_pts + (nch + 1);
  HIP_TRY(hipMemcpyAsync(d_pts, bpts.data(), sizeof(long long) * (nch + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_bounds, dim3((unsigned)((nch + 1 + 255) / 256)), dim3(256), 0, s, (int)(nch + 1),
                     d_pts, h->loff.as<int64_t>(), h->actidx.as<int32_t>(), d_out);
  HIP_TRY(hipGetLastError());
  std::vector<long long> bv(2 * (nch + 1));
  HIP_TRY(hipMemcpyAsync(bv.data(), d_out, sizeof(long long) * 2 * (nch + 1), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  long long maxL = 0, maxA = 0;
  for (long long c = 0; c < nch; ++c) {
    maxL = std::max(maxL, bv[2 * (c + 1)] - bv[2 * c]);
    maxA = std::max(maxA, bv[2 * (c + 1) + 1] - bv[2 * c + 1]);
  }
  HIP_TRY(h->lobs.ensure(sizeof(int32_t) * (maxL + 1)));
  HIP_TRY(h->lrloc.ensure(sizeof(double) * (maxL + 1)));
  if (cfg->mode == 3) {
    HIP_TRY(h->G.ensure(sizeof(double) * (size_t)(maxA + 1) * K * K));
    HIP_TRY(h->bvec.ensure(sizeof(double) * (size_t)(maxA + 1) * K));
  }
  for (long long c = 0; c < nch; ++c) {
    const long long p0 = bpts[c], p1 = bpts[c + 1];
    const long long lbase = bv[2 * c];
    const int a0 = (int)bv[2 * c + 1];
    const int na = (int)(bv[2 * (c + 1) + 1] - a0);
    if (na > 0) {
      HIP_TRY(launch_search_fill<T>(kp, tb, p0, p1, h->loff.as<int64_t>(), h->lobs.as<int32_t>(),
                                    h->lrloc.as<double>(), s));
      if (cfg->mode == 3)
        HIP_TRY(launch_gram_fp64<T>(kg, tb, h->yT.as<double>(), h->act_p.as<int32_t>(), h->loff.as<int64_t>(),
                                    lbase, h->lobs.as<int32_t>(), h->lrloc.as<double>(), a0, na, h->G.as<double>(),
                                    h->bvec.as<double>(), 0, s));
    }
    HIP_TRY(launch_lpfgm_points<T>(kp, tb, f3, f2, a3, a2, h->nloc.as<int32_t>(), h->loff.as<int64_t>(), lbase,
                                   h->lobs.as<int32_t>(), h->lrloc.as<double>(), h->yT.as<double>(),
                                   h->actidx.as<int32_t>(), a0, cfg->mode == 3 ? h->G.as<double>() : nullptr,
                                   cfg->mode == 3 ? h->bvec.as<double>() : nullptr, p0, p1, la, s));
  }
  int stat4[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(stat4, h->status.ptr, sizeof(stat4), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return status_to_rc(stat4[0]);
}

static int check_call(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o, const void* f3,
                      const void* f2, void* a3, void* a2) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  if (!g || !o) return set_err(LETKF_ERR_ARG, "grid/obs is NULL");
  if (o->nobs < 0) return set_err(LETKF_ERR_ARG, "nobs < 0");
  const letkf_params_t& p = h->prm;
  if ((p.nv3d && (!f3 || !a3)) || (p.nv2d && (!f2 || !a2)))
    return set_err(LETKF_ERR_ARG, "fcst/anal array is NULL");
  if (o->nobs > 0) {
    const void* need[] = {g->lon1, g->lat1, g->ri1, g->rj1, g->dlon1, g->dlat1, g->phi1, g->depth1,
                          o->nobsgrd, o->elm, o->lon, o->lat, o->lev, o->err, o->dep, o->hdxf};
    for (const void* q : need)
      if (!q) return set_err(LETKF_ERR_ARG, "a grid/obs array is NULL");
  }
  return LETKF_OK;
}

extern "C" int letkf_da_main_device(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o,
                                    const void* f3, const void* f2, void* a3, void* a2, void* stream) {
  int rc = check_call(h, g, o, f3, f2, a3, a2);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  if (h->prm.real_kind == 8)
    return run_device<double>(h, g, o, (const double*)f3, (const double*)f2, (double*)a3, (double*)a2, s);
  return run_device<float>(h, g, o, (const float*)f3, (const float*)f2, (float*)a3, (float*)a2, s);
}

extern "C" int letkf_da_main_lpfgm_device(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o,
                                          const letkf_lpfgm_cfg_t* cfg, const void* f3, const void* f2, void* a3,
                                          void* a2, void* stream) {
  int rc = check_call(h, g, o, f3, f2, a3, a2);
  if (rc) return rc;
  if (!cfg) return set_err(LETKF_ERR_ARG, "lpfgm settings are NULL");
  if (cfg->mode != 2 && cfg->mode != 3) return set_err(LETKF_ERR_ARG, "lpfgm mode must be 2 or 3 (got %d)", cfg->mode);
  if (cfg->nmonte < 1 || cfg->nmonte > 255 || !cfg->rand)
    return set_err(LETKF_ERR_ARG, "lpfgm needs 1 <= nmonte <= 255 and the draws");
  if (h->prm.nmem > 128 || h->prm.nmem < 2) return set_err(LETKF_ERR_ARG, "lpfgm analysis needs 2 <= nmem <= 128");
  if (cfg->mode == 3 && !(cfg->cov_infl_gm > 0.0)) return set_err(LETKF_ERR_ARG, "cov_infl_gm must be > 0");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  h->invalidate_plan("invalidated by letkf_da_main_lpfgm(_device), which reuses the plan's workspaces");
  if (h->prm.real_kind == 8)
    return run_lpfgm<double>(h, g, o, cfg, (const double*)f3, (const double*)f2, (double*)a3, (double*)a2, s);
  return run_lpfgm<float>(h, g, o, cfg, (const float*)f3, (const float*)f2, (float*)a3, (float*)a2, s);
}

// In-project autotuner of the register eigensolve's vector path: on the
// first chunk of this analysis, every size bucket runs its eigensolve +
// update under each vector path it can take (1: Q accumulated, K <= 64;
// 2: reflectors, one wavefront sum per column; 3: reflectors in the
// blocked layout, 1 + nv <= 8), three timed runs each after a warm-up, and
// keeps the fastest; then the whole analysis runs with the chosen paths.
// The fixed rule's path is kept unless another is faster by more than 3%
// (timing noise on a tie must not flip it).  ms[3q + v - 1] receives the
// best time of bucket q under path v (-1: not measured).
template <typename T>
static int autotune_device(letkf_handle* h, const letkf_grid_t* g, const letkf_obs_t* o, const T* f3, const T* f2,
                           T* a3, T* a2, hipStream_t s, double* ms) {
  for (int q = 0; q < 3 * kNQ; ++q) ms[q] = -1.0;
  const letkf_params_t& p = h->prm;
  const int K = p.nmem;
  const int ncol = 1 + p.nv3d + p.nv2d;
  if (K <= 128 && o->nobs > 0) {
    int rc = plan_device<T>(h, g, o, s, false);
    if (rc) return rc;
    const AnalysisPlan& pl = h->plan;
    const int a0 = pl.nch > 0 ? (int)pl.bv[1] : 0;
    const int na = pl.nch > 0 ? (int)(pl.bv[3] - a0) : 0;
    if (pl.reg_eig && na > 0) {
      KParams kp = make_kparams(p, o->nobs);
      Tables<T> tb = make_tables<T>(g, o, h->otrig.as<double2>());
      const int dual = pl.dual;
      const long long lbase = pl.bv[0];
      hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, h->status.as<int>());
      HIP_TRY(hipGetLastError());
      HIP_TRY(launch_transpose_obs<T>(o->nobs, K, (const T*)o->hdxf, h->yT.as<double>(), s));
      const void* ysp = pl.ysplit ? h->ysp.ptr : nullptr;
      if (ysp) HIP_TRY(launch_split_obs(o->nobs, K, h->yT.as<double>(), h->ysp.ptr, s));
      HIP_TRY(launch_search_fill<T>(kp, tb, pl.bpts[0], pl.bpts[1], h->loff.as<int64_t>(), h->lobs.as<int32_t>(),
                                    h->lrloc.as<double>(), s, pl.fastloc));
      HIP_TRY(launch_gram_any<T>(p.precision, kp, tb, h->yT.as<double>(), h->act_p.as<int32_t>(),
                                 h->loff.as<int64_t>(), lbase, h->lobs.as<int32_t>(), h->lrloc.as<double>(), a0, na,
                                 h->G.as<double>(), h->bvec.as<double>(), dual, ysp, s));
      // the first chunk's point-major forecast, as the analysis uses it
      PointMajor<T> pm{nullptr, nullptr, pl.bpts[0], pl.pm_ld};
      if (pl.pm_ld > 0) {
        pm.X = h->xpm.as<T>();
        pm.A = h->apm.as<T>();
        double2* st = pl.pm_stats ? h->pst.as<double2>() : nullptr;
        pm.S = st;
        HIP_TRY(launch_to_point_major<T>(kp, pl.bpts[0], pl.bpts[1], f3, f2, h->xpm.as<T>(), pl.pm_ld, s, st));
      }
      hipEvent_t e0, e1;
      HIP_TRY(hipEventCreate(&e0));
      HIP_TRY(hipEventCreate(&e1));
      for (int q = 0; q < kNQ; ++q) {
        const int cnt = pl.hist[q];
        if (cnt == 0) continue;
        for (int v = 1; v <= 3; ++v) {
          if (eigr_vpath(v, bucket_nb(q), K, ncol) != v) continue;  // a path this bucket cannot take
          float best = 1e30f;
          for (int r = 0; r < 4; ++r) {  // run 0 warms up
            HIP_TRY(hipEventRecord(e0, s));
            HIP_TRY(launch_eigr_update<T>(bucket_nb(q), K, dual, h->nloc.as<int32_t>(), h->act_p.as<int32_t>(), a0,
                                          h->blist.as<int32_t>() + pl.boff[q], cnt, h->G.as<double>(),
                                          h->V.as<double>(), h->eD.as<double>(), h->eE.as<double>(),
                                          h->logc.as<double>(), h->logx.as<uint8_t>(), h->logi.as<int32_t>(),
                                          h->logn.as<int32_t>(), h->logr.as<int32_t>(), h->lam.as<double>(),
                                          h->V.as<double>(), h->yvec.as<double>(), pl.yp, kp, tb, f3, f2, a3, a2,
                                          h->loff.as<int64_t>(), lbase, h->lobs.as<int32_t>(),
                                          h->lrloc.as<double>(), h->yT.as<double>(), h->bvec.as<double>(),
                                          h->status.as<int>(), v, pl.rat ? h->ratt.as<double>() : nullptr,
                                          p.precision == LETKF_PREC_BF16X3_TRI32, nullptr, pm, pl.kmax, s));
            HIP_TRY(hipEventRecord(e1, s));
            HIP_TRY(hipEventSynchronize(e1));
            float t = 0.f;
            HIP_TRY(hipEventElapsedTime(&t, e0, e1));
            if (r > 0 && t < best) best = t;
          }
          ms[3 * q + v - 1] = best;
        }
        // the fixed rule's path unless another is faster by more than 3%
        const int dflt = eigr_vpath(0, bucket_nb(q), K, ncol);
        int pick = dflt;
        for (int v = 1; v <= 3; ++v)
          if (ms[3 * q + v - 1] >= 0.0 && ms[3 * q + v - 1] < 0.97 * ms[3 * q + dflt - 1] &&
              (pick == dflt || ms[3 * q + v - 1] < ms[3 * q + pick - 1]))
            pick = v;
        h->vpath[q] = pick;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  }
  return run_device<T>(h, g, o, f3, f2, a3, a2, s);
}

extern "C" int letkf_autotune(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o, const void* f3,
                              const void* f2, void* a3, void* a2, double* ms, void* stream) {
  int rc = check_call(h, g, o, f3, f2, a3, a2);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  double loc[3 * kNQ];
  double* m = ms ? ms : loc;
  if (h->prm.real_kind == 8)
    return autotune_device<double>(h, g, o, (const double*)f3, (const double*)f2, (double*)a3, (double*)a2, s, m);
  return autotune_device<float>(h, g, o, (const float*)f3, (const float*)f2, (float*)a3, (float*)a2, s, m);
}

extern "C" int letkf_set_tuning(letkf_handle_t* h, const int32_t* vpath) {
  if (!h || !vpath) return set_err(LETKF_ERR_ARG, "NULL argument");
  for (int q = 0; q < kNQ; ++q)
    if (vpath[q] < 0 || vpath[q] > 3) return set_err(LETKF_ERR_ARG, "vector path %d of bucket %d not in 0..3", vpath[q], q);
  for (int q = 0; q < kNQ; ++q) h->vpath[q] = vpath[q];
  return LETKF_OK;
}

extern "C" int letkf_get_tuning(const letkf_handle_t* h, int32_t* vpath) {
  if (!h || !vpath) return set_err(LETKF_ERR_ARG, "NULL argument");
  for (int q = 0; q < kNQ; ++q) vpath[q] = h->vpath[q];
  return LETKF_OK;
}

extern "C" int letkf_da_main_plan(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o, void* stream) {
  int rc = check_call(h, g, o, (const void*)1, (const void*)1, (void*)1, (void*)1);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  rc = h->prm.real_kind == 8 ? plan_device<double>(h, g, o, s, false) : plan_device<float>(h, g, o, s, false);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  h->stats = h->plan.st;
  return LETKF_OK;
}

extern "C" int letkf_da_main_enqueue(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o,
                                     const void* f3, const void* f2, void* a3, void* a2, void* stream) {
  int rc = check_call(h, g, o, f3, f2, a3, a2);
  if (rc) return rc;
  const AnalysisPlan& pl = h->plan;
  if (!pl.valid) return set_err(LETKF_ERR_ARG, "letkf_da_main_enqueue: no plan (call letkf_da_main_plan first)");
  const void* key[12];
  plan_key(g, o, key);
  bool same = o->nobs == pl.nobs && same_geometry_params(h->prm, pl.prm);
  for (int q = 0; q < 12 && same; ++q) same = key[q] == pl.key[q];
  if (!same)
    return set_err(LETKF_ERR_ARG, "letkf_da_main_enqueue: grid, obs positions or localisation differ from the plan");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  h->enq_epoch = pl.epoch;
  if (h->prm.real_kind == 8)
    return enqueue_device<double>(h, g, o, (const double*)f3, (const double*)f2, (double*)a3, (double*)a2, s, false,
                                  nullptr);
  return enqueue_device<float>(h, g, o, (const float*)f3, (const float*)f2, (float*)a3, (float*)a2, s, false,
                               nullptr);
}

extern "C" int letkf_da_main_status(letkf_handle_t* h, void* stream) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  int stat4[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(stat4, h->status.ptr, sizeof(stat4), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const AnalysisPlan& pl = h->plan;
  if (h->enq_epoch != 0 && (!pl.valid || pl.epoch != h->enq_epoch))
    return set_err(LETKF_ERR_ARG,
                   "letkf_da_main_status: the plan of the last letkf_da_main_enqueue was %s since; a graph "
                   "captured from that enqueue refers to workspaces that have been reused and must be captured "
                   "again after a new letkf_da_main_plan",
                   pl.valid ? "replaced by a new plan (letkf_da_main_plan, letkf_da_main_device, letkf_autotune)"
                            : (h->invalidated_by ? h->invalidated_by : "invalidated"));
  return status_to_rc(stat4[0]);
}

extern "C" int letkf_da_adjoint_device(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o,
                                       const void* f3, const void* f2, const void* ga3, const void* ga2,
                                       double* g_dep, double* g_logvar, double* g_scalars, void* stream) {
  int rc = check_call(h, g, o, f3, f2, (void*)ga3, (void*)ga2);
  if (rc) return rc;
  const letkf_params_t& p = h->prm;
  if (p.real_kind != 8) return set_err(LETKF_ERR_ARG, "the adjoint needs real_kind 8");
  if (p.nmem > 64) return set_err(LETKF_ERR_ARG, "the adjoint supports nmem <= 64 (got %d)", p.nmem);
  if (p.alpha_rtps != 0.0) return set_err(LETKF_ERR_ARG, "the adjoint needs alpha_rtps = 0 (RTPS is not differentiated)");
  if (!g_scalars || (o->nobs > 0 && (!g_dep || !g_logvar))) return set_err(LETKF_ERR_ARG, "gradient array is NULL");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  h->invalidate_plan("invalidated by letkf_da_adjoint_device, which reuses the plan's workspaces");
  return run_adjoint(h, g, o, (const double*)f3, (const double*)f2, (const double*)ga3, (const double*)ga2, g_dep,
                     g_logvar, g_scalars, s);
}

extern "C" int letkf_set_params(letkf_handle_t* h, const letkf_params_t* p) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  int rc = check_params(p);
  if (rc) return rc;
  const letkf_params_t& q = h->prm;
  if (p->nmem != q.nmem || p->nlon != q.nlon || p->nlat != q.nlat || p->nlev != q.nlev || p->nij1 != q.nij1 ||
      p->nv3d != q.nv3d || p->nv2d != q.nv2d || p->real_kind != q.real_kind)
    return set_err(LETKF_ERR_ARG, "letkf_set_params: sizes and real_kind must not change");
  h->prm = *p;
  return LETKF_OK;
}

// Host arrays: staged into the handle's device copies, analysed by the
// device entry (the LETKF, or the LPF / LPFGM core when cfg is given), the
// analysis copied back.
static int da_main_host(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o,
                        const letkf_lpfgm_cfg_t* cfg, const void* f3, const void* f2, void* a3, void* a2) {
  int rc = check_call(h, g, o, f3, f2, a3, a2);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(h->device));
  const letkf_params_t& p = h->prm;
  const size_t R = p.real_kind;
  const size_t nij = p.nij1, nlev = p.nlev, K = p.nmem, nobs = o->nobs;
  const size_t n3 = nij * nlev * K * p.nv3d, n2 = nij * K * p.nv2d;
  hipStream_t s = h->stream;
  // grid arrays: 7 x nij1 + depth1 (nij1 x nlev)
  const size_t gbytes = R * (7 * nij + nij * nlev);
  const size_t obytes = sizeof(int32_t) * ((size_t)p.nlon * p.nlat + nobs) + R * (5 * nobs + nobs * K);
  HIP_TRY(h->h_grid.ensure(gbytes));
  HIP_TRY(h->h_obs.ensure(obytes + 64));
  HIP_TRY(h->h_f3.ensure(R * n3));
  HIP_TRY(h->h_f2.ensure(R * n2));
  HIP_TRY(h->h_a3.ensure(R * n3));
  HIP_TRY(h->h_a2.ensure(R * n2));
  char* gp = h->h_grid.as<char>();
  letkf_grid_t dg;
  const void* gsrc[8] = {g->lon1, g->lat1, g->ri1, g->rj1, g->dlon1, g->dlat1, g->phi1, g->depth1};
  const void** gdst[8] = {&dg.lon1, &dg.lat1, &dg.ri1, &dg.rj1, &dg.dlon1, &dg.dlat1, &dg.phi1, &dg.depth1};
  for (int q = 0; q < 8; ++q) {
    const size_t b = R * (q == 7 ? nij * nlev : nij);
    *gdst[q] = gp;
    if (gsrc[q]) HIP_TRY(hipMemcpyAsync(gp, gsrc[q], b, hipMemcpyHostToDevice, s));
    gp += b;
  }
  char* op = h->h_obs.as<char>();
  letkf_obs_t dobs;
  dobs.nobs = o->nobs;
  auto put = [&](const void* src, size_t b) -> const void* {
    const void* dst = op;
    if (src && b) {
      hipError_t e = hipMemcpyAsync(op, src, b, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return nullptr;
    }
    op += (b + 7) & ~(size_t)7;
    return dst;
  };
  dobs.nobsgrd = (const int32_t*)put(o->nobsgrd, sizeof(int32_t) * p.nlon * p.nlat);
  dobs.elm = (const int32_t*)put(o->elm, sizeof(int32_t) * nobs);
  dobs.lon = put(o->lon, R * nobs);
  dobs.lat = put(o->lat, R * nobs);
  dobs.lev = put(o->lev, R * nobs);
  dobs.err = put(o->err, R * nobs);
  dobs.dep = put(o->dep, R * nobs);
  dobs.hdxf = put(o->hdxf, R * nobs * K);
  if (n3) HIP_TRY(hipMemcpyAsync(h->h_f3.ptr, f3, R * n3, hipMemcpyHostToDevice, s));
  if (n2) HIP_TRY(hipMemcpyAsync(h->h_f2.ptr, f2, R * n2, hipMemcpyHostToDevice, s));
  rc = cfg ? letkf_da_main_lpfgm_device(h, &dg, &dobs, cfg, h->h_f3.ptr, h->h_f2.ptr, h->h_a3.ptr, h->h_a2.ptr, s)
           : letkf_da_main_device(h, &dg, &dobs, h->h_f3.ptr, h->h_f2.ptr, h->h_a3.ptr, h->h_a2.ptr, s);
  if (rc) return rc;
  if (n3) HIP_TRY(hipMemcpyAsync(a3, h->h_a3.ptr, R * n3, hipMemcpyDeviceToHost, s));
  if (n2) HIP_TRY(hipMemcpyAsync(a2, h->h_a2.ptr, R * n2, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return LETKF_OK;
}

extern "C" int letkf_da_main(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o, const void* f3,
                             const void* f2, void* a3, void* a2) {
  return da_main_host(h, g, o, nullptr, f3, f2, a3, a2);
}

extern "C" int letkf_da_main_lpfgm(letkf_handle_t* h, const letkf_grid_t* g, const letkf_obs_t* o,
                                   const letkf_lpfgm_cfg_t* cfg, const void* f3, const void* f2, void* a3,
                                   void* a2) {
  if (!cfg) return set_err(LETKF_ERR_ARG, "lpfgm settings are NULL");
  return da_main_host(h, g, o, cfg, f3, f2, a3, a2);
}
