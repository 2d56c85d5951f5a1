// setobs.hip -- the body of set_obs after the observation operator
// (common_obs.f90:139-402) on device arrays: per-obs ensemble statistics,
// gross-error check, temporal localisation, removal of rejected obs, the
// stable (lat cell, lon cell) sort with nobsgrd, and AOEI.
//
// Three launches and one radix sort, no host round trip before the end:
//  k_setobs_stats   one thread per input obs: mean/spread over the members
//                   (column reads of hx (nobs,nmem) are coalesced across the
//                   wave), innovation, QC, localised error, sort key, cell
//                   count, and the kept/rejected counters;
//  radix sort       (key, input index) pairs, stable, so obs of one cell keep
//                   their input order as the reference's two gathering
//                   passes (common_obs.f90:266-345) do, and the qc 0 obs
//                   (key INT32_MAX) go last: the compaction of
//                   common_obs.f90:236-256 and the sort in one step;
//  inclusive scan   of the per-cell counts -> nobsgrd;
//  k_setobs_gather  one thread per sorted slot (slots past the kept count
//                   exit): scalars, hdxf columns (written coalesced), AOEI.
#include <cmath>
#include <cstring>
#include <hipcub/hipcub.hpp>

#include "letkf_handle.h"

using namespace letkf;

namespace {

template <typename T> __device__ __forceinline__ T sqrt_t(T x);
template <> __device__ __forceinline__ double sqrt_t<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_t<float>(float x) { return sqrtf(x); }

constexpr int kMaxSlots = 1024;

template <typename T>
struct StatsArgs {
  int nobs, nmem, nlon, nlat, nslots;
  int lge, id_u, id_v, id_t, id_s, id_z;
  T undef, gross, qmin[9], qmax[9];
  const int32_t *elm, *idx, *idy, *qc;
  const T *lev, *dat, *err, *hx;
  const int32_t* slot_end;  // [nslots] exclusive end of each slot's obs
  const double* slot_fac;   // [nslots] exp(0.25 ((islot-nbslot)/sigma_obst)^2)
  T *mean, *sprd, *dep, *erl;
  int32_t *key, *vals, *counts, *ctr;  // ctr[0]: qc 1 after projection, ctr[1]: after gross check
};

// Gross-error class of common_obs.f90:191-213 (-1: the |dep| > 10 err rule).
template <typename T>
__device__ __forceinline__ int gross_class(const StatsArgs<T>& a, int el, T lev) {
  const bool sfc = lev == T(0);
  if (el == a.id_z) return 0;
  if (el == a.id_t) return sfc ? 1 : 2;
  if (el == a.id_s) return sfc ? 3 : 4;
  if (el == a.id_u) return sfc ? 5 : 6;
  if (el == a.id_v) return sfc ? 7 : 8;
  return -1;
}

template <typename T>
__global__ void __launch_bounds__(256) k_setobs_stats(StatsArgs<T> a) {
#pragma clang fp contract(off)
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= a.nobs) return;
  int q = a.qc ? a.qc[o] : 1;
  const int ix = a.idx[o], iy = a.idy[o];
  // H(x) is undef outside cells 1..nlon-1 x 1..nlat-1 (common.f90:428-432),
  // so such obs never reach the sort with qc 1 in the reference
  if (ix < 1 || a.nlon - 1 < ix || iy < 1 || a.nlat - 1 < iy) q = 0;
  T mean = a.undef, sprd = a.undef, dep = a.undef;
  if (q == 1) {
    const T* h = a.hx + o;
    const long long ld = a.nobs;
    T s = T(0);
    for (int m = 0; m < a.nmem; ++m) s += h[ld * m];
    mean = s / T(a.nmem);
    T v = T(0);
    for (int m = 0; m < a.nmem; ++m) {
      const T d = h[ld * m] - mean;
      v += d * d;
    }
    sprd = a.nmem > 1 ? sqrt_t<T>(v / T(a.nmem - 1)) : a.undef;
    dep = a.dat[o] - mean;
  } else {
    q = 0;
  }
  const int q1 = q;
  const T e = a.err[o];
  if (q == 1 && a.lge) {
    const int c = gross_class(a, a.elm[o], a.lev[o]);
    if (c >= 0) {
      if (dep < a.qmin[c] || a.qmax[c] < dep) q = 0;
    } else {
      const T ad = dep < T(0) ? -dep : dep;
      if (ad > a.gross * e) q = 0;
    }
  }
  // slot of obs o: the first s with o < slot_end[s]
  int sl = 0;
  while (sl < a.nslots - 1 && a.slot_end[sl] <= o) ++sl;
  a.mean[o] = mean;
  a.sprd[o] = sprd;
  a.dep[o] = dep;
  a.erl[o] = (T)((double)e * a.slot_fac[sl]);
  int k = INT32_MAX;
  if (q == 1) {
    k = (iy - 1) * a.nlon + (ix - 1);
    atomicAdd(&a.counts[k], 1);
  }
  a.key[o] = k;
  a.vals[o] = o;
  // wave-aggregated counters
  const unsigned long long b1 = __ballot(q1 == 1), b2 = __ballot(q == 1);
  if ((threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1) {
    if (b1) atomicAdd(&a.ctr[0], __popcll(b1));
    if (b2) atomicAdd(&a.ctr[1], __popcll(b2));
  }
}

template <typename T>
struct GatherArgs {
  int nobs, nmem, aoei;
  const int32_t *perm, *ctr;
  const int32_t *elm, *ins, *idx, *idy;
  const T *lon, *lat, *lev, *dat, *hx, *mean, *sprd, *dep, *erl;
  int32_t *o_elm, *o_ins, *o_idx, *o_idy;
  T *o_lon, *o_lat, *o_lev, *o_dat, *o_err, *o_dep, *o_mean, *o_sprd, *o_hdxf;
};

template <typename T>
__global__ void __launch_bounds__(256) k_setobs_gather(GatherArgs<T> a) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int nout = a.ctr[1];
  if (r >= nout) return;
  const int o = a.perm[r];
  const T mean = a.mean[o], sprd = a.sprd[o], dep = a.dep[o];
  T e = a.erl[o];
  if (a.aoei) {  // common_obs.f90:389-400
    const T d2 = dep * dep - sprd * sprd;
    const T e2 = e * e;
    e = sqrt_t<T>(e2 > d2 ? e2 : d2);
  }
  a.o_elm[r] = a.elm[o];
  if (a.o_ins) a.o_ins[r] = a.ins ? a.ins[o] : 0;
  if (a.o_idx) a.o_idx[r] = a.idx[o];
  if (a.o_idy) a.o_idy[r] = a.idy[o];
  a.o_lon[r] = a.lon[o];
  a.o_lat[r] = a.lat[o];
  a.o_lev[r] = a.lev[o];
  if (a.o_dat) a.o_dat[r] = a.dat[o];
  a.o_err[r] = e;
  a.o_dep[r] = dep;
  if (a.o_mean) a.o_mean[r] = mean;
  if (a.o_sprd) a.o_sprd[r] = sprd;
  const long long ldi = a.nobs, ldo = nout;
  for (int m = 0; m < a.nmem; ++m) a.o_hdxf[r + ldo * m] = a.hx[o + ldi * m] - mean;
}

// nobsgrd row nlat is never filled by the reference (common_obs.f90:271-345).
__global__ void k_zero_last_row(int nlon, int nlat, int32_t* grd) {
  for (int i = threadIdx.x; i < nlon; i += blockDim.x) grd[(long long)nlon * (nlat - 1) + i] = 0;
}

template <typename T>
int run_set_obs(letkf_handle_t* h, const letkf_setobs_cfg_t& c, int nlon, int nlat, int nobs,
                const int32_t* nobslots, const int32_t* elm, const int32_t* ins, const T* lon, const T* lat,
                const T* lev, const T* dat, const T* err, const int32_t* idx, const int32_t* idy, const T* hx,
                const int32_t* qc, letkf_obs_set_t* out) {
  hipStream_t s = h->stream;
  const int nmem = h->prm.nmem;
  const int ncell = nlon * nlat;
  const int ns = c.nslots;
  // per-slot error factors on the host (common_obs.f90:224-232): the slot
  // offset and its square in the arrays' REAL, then 0.25d0 and exp in double
  std::vector<int32_t> hend(ns);
  std::vector<double> hfac(ns);
  long long acc = 0;
  for (int k = 0; k < ns; ++k) {
    acc += nobslots ? nobslots[k] : nobs;
    hend[k] = (int32_t)acc;
    const T x = (T)(k + 1 - c.nbslot) / (T)c.sigma_obst;
    const T x2 = x * x;
    hfac[k] = std::exp(0.25 * (double)x2);
  }
  if (acc != nobs) return set_err(LETKF_ERR_ARG, "nobslots sum to %lld, not nobs = %d", acc, nobs);

  const size_t n1 = (size_t)(nobs > 0 ? nobs : 1);
  size_t cub_sort = 0, cub_scan = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_sort, (int32_t*)nullptr, (int32_t*)nullptr,
                                             (int32_t*)nullptr, (int32_t*)nullptr, (int)n1, 0, 32, s));
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, cub_scan, (int32_t*)nullptr, (int32_t*)nullptr, ncell, s));
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_real = up(sizeof(T) * n1), b_int = up(sizeof(int32_t) * n1);
  const size_t b_cnt = up(sizeof(int32_t) * (ncell + 2)), b_slot = up((sizeof(int32_t) + sizeof(double)) * ns + 8);
  const size_t bytes = 4 * b_real + 4 * b_int + b_cnt + b_slot + up(std::max(cub_sort, cub_scan)) + 256;
  HIP_TRY(h->bin_tmp.ensure(bytes));
  char* p = (char*)(((uintptr_t)h->bin_tmp.ptr + 255) & ~(uintptr_t)255);
  auto take = [&](size_t b) { char* q = p; p += b; return q; };
  T* mean = (T*)take(b_real);
  T* sprd = (T*)take(b_real);
  T* dep = (T*)take(b_real);
  T* erl = (T*)take(b_real);
  int32_t* key = (int32_t*)take(b_int);
  int32_t* key2 = (int32_t*)take(b_int);
  int32_t* vals = (int32_t*)take(b_int);
  int32_t* perm = (int32_t*)take(b_int);
  int32_t* counts = (int32_t*)take(b_cnt);  // [ncell] cells, then 2 counters
  int32_t* ctr = counts + ncell;
  char* slot = take(b_slot);
  double* sfac = (double*)slot;
  int32_t* send = (int32_t*)(slot + sizeof(double) * ns);
  void* cub_tmp = take(up(std::max(cub_sort, cub_scan)));

  HIP_TRY(hipMemsetAsync(counts, 0, sizeof(int32_t) * (ncell + 2), s));
  HIP_TRY(hipMemcpyAsync(sfac, hfac.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(send, hend.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
  if (nobs > 0) {
    StatsArgs<T> a{};
    a.nobs = nobs;
    a.nmem = nmem;
    a.nlon = nlon;
    a.nlat = nlat;
    a.nslots = ns;
    a.lge = c.lge_io && !c.aoei;
    a.id_u = c.id_u_obs;
    a.id_v = c.id_v_obs;
    a.id_t = c.id_t_obs;
    a.id_s = c.id_s_obs;
    a.id_z = c.id_z_obs;
    a.undef = (T)c.undef;
    a.gross = (T)c.gross_error;
    for (int k = 0; k < 9; ++k) {
      a.qmin[k] = (T)c.qc_min[k];
      a.qmax[k] = (T)c.qc_max[k];
    }
    a.elm = elm;
    a.idx = idx;
    a.idy = idy;
    a.qc = qc;
    a.lev = lev;
    a.dat = dat;
    a.err = err;
    a.hx = hx;
    a.slot_end = send;
    a.slot_fac = sfac;
    a.mean = mean;
    a.sprd = sprd;
    a.dep = dep;
    a.erl = erl;
    a.key = key;
    a.vals = vals;
    a.counts = counts;
    a.ctr = ctr;
    hipLaunchKernelGGL(k_setobs_stats<T>, dim3((nobs + 255) / 256), dim3(256), 0, s, a);
    HIP_TRY(hipGetLastError());
    size_t tb = cub_sort;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(cub_tmp, tb, key, key2, vals, perm, nobs, 0, 32, s));
  }
  size_t tb = cub_scan;
  HIP_TRY(hipcub::DeviceScan::InclusiveSum(cub_tmp, tb, counts, out->nobsgrd, ncell, s));
  hipLaunchKernelGGL(k_zero_last_row, dim3(1), dim3(256), 0, s, nlon, nlat, out->nobsgrd);
  HIP_TRY(hipGetLastError());
  if (nobs > 0) {
    GatherArgs<T> g{};
    g.nobs = nobs;
    g.nmem = nmem;
    g.aoei = c.aoei;
    g.perm = perm;
    g.ctr = ctr;
    g.elm = elm;
    g.ins = ins;
    g.idx = idx;
    g.idy = idy;
    g.lon = lon;
    g.lat = lat;
    g.lev = lev;
    g.dat = dat;
    g.hx = hx;
    g.mean = mean;
    g.sprd = sprd;
    g.dep = dep;
    g.erl = erl;
    g.o_elm = out->elm;
    g.o_ins = out->ins;
    g.o_idx = out->idx;
    g.o_idy = out->idy;
    g.o_lon = (T*)out->lon;
    g.o_lat = (T*)out->lat;
    g.o_lev = (T*)out->lev;
    g.o_dat = (T*)out->dat;
    g.o_err = (T*)out->err;
    g.o_dep = (T*)out->dep;
    g.o_mean = (T*)out->hxfmean;
    g.o_sprd = (T*)out->hxfsprd;
    g.o_hdxf = (T*)out->hdxf;
    hipLaunchKernelGGL(k_setobs_gather<T>, dim3((nobs + 255) / 256), dim3(256), 0, s, g);
    HIP_TRY(hipGetLastError());
  }
  int32_t hc[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(hc, ctr, sizeof hc, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  out->nobs_out = hc[1];
  out->ndel_proj = nobs - hc[0];
  out->ndel_gross = hc[0] - hc[1];
  return LETKF_OK;
}

}  // namespace

extern "C" void letkf_setobs_default(letkf_setobs_cfg_t* c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->id_u_obs = 2819;
  c->id_v_obs = 2820;
  c->id_t_obs = 3073;
  c->id_s_obs = 3332;
  c->id_z_obs = 2567;
  c->lge_io = 0;
  c->aoei = 1;
  c->nslots = 1;
  c->nbslot = 1;
  c->sigma_obst = 1.0;
  c->undef = 999.0;
  c->gross_error = 10.0;
  // SSH, SST, T, SSS, S, SSU, U, SSV, V (common_setting.f90:109-118)
  const double lo[9] = {-1, -5, -5, -2, -2, -2, -2, -2, -2};
  for (int k = 0; k < 9; ++k) {
    c->qc_min[k] = lo[k];
    c->qc_max[k] = -lo[k];
  }
}

extern "C" int letkf_set_obs_device(letkf_handle_t* h, const letkf_setobs_cfg_t* cfg, int32_t nlon, int32_t nlat,
                                    int32_t nobs, const int32_t* nobslots, const int32_t* elm, const int32_t* ins,
                                    const void* lon, const void* lat, const void* lev, const void* dat,
                                    const void* err, const int32_t* idx, const int32_t* idy, const void* hx,
                                    const int32_t* qc, letkf_obs_set_t* out) {
  if (!h) return set_err(LETKF_ERR_ARG, "handle is NULL");
  if (!cfg || !out) return set_err(LETKF_ERR_ARG, "cfg or out is NULL");
  if (nlon < 2 || nlat < 2 || nobs < 0) return set_err(LETKF_ERR_ARG, "bad set_obs sizes");
  if ((long long)nlon * nlat >= INT32_MAX) return set_err(LETKF_ERR_ARG, "grid too large");
  if (cfg->nslots < 1 || cfg->nslots > kMaxSlots) return set_err(LETKF_ERR_ARG, "nslots must be in 1..%d", kMaxSlots);
  if (cfg->nslots > 1 && !nobslots) return set_err(LETKF_ERR_ARG, "nobslots is NULL with nslots > 1");
  if (!(cfg->sigma_obst > 0)) return set_err(LETKF_ERR_ARG, "sigma_obst must be > 0");
  if (nobslots)
    for (int k = 0; k < cfg->nslots; ++k)
      if (nobslots[k] < 0) return set_err(LETKF_ERR_ARG, "nobslots[%d] < 0", k);
  if (!out->nobsgrd) return set_err(LETKF_ERR_ARG, "out->nobsgrd is NULL");
  if (nobs > 0 && (!elm || !lon || !lat || !lev || !dat || !err || !idx || !idy || !hx))
    return set_err(LETKF_ERR_ARG, "input obs array is NULL");
  if (nobs > 0 && (!out->elm || !out->lon || !out->lat || !out->lev || !out->err || !out->dep || !out->hdxf))
    return set_err(LETKF_ERR_ARG, "output obs array is NULL");
  const int nmem = h->prm.nmem;
  if (nmem < 1) return set_err(LETKF_ERR_ARG, "nmem < 1");
  HIP_TRY(hipSetDevice(h->device));
  if (h->prm.real_kind == 8)
    return run_set_obs<double>(h, *cfg, nlon, nlat, nobs, nobslots, elm, ins, (const double*)lon,
                               (const double*)lat, (const double*)lev, (const double*)dat, (const double*)err, idx,
                               idy, (const double*)hx, qc, out);
  return run_set_obs<float>(h, *cfg, nlon, nlat, nobs, nobslots, elm, ins, (const float*)lon, (const float*)lat,
                            (const float*)lev, (const float*)dat, (const float*)err, idx, idy, (const float*)hx, qc,
                            out);
}
