// xtoy.hip -- the observation operator H(x) of the reference,
// Trans_XtoY (letkf/run/src/common.f90:391-500) with its bilinear_interpolation
// (:246-269) and akima_spline (:276-385), batched over ensemble members: the
// step that produces hdxf (common_io.f90:829-831, one call per member there).
//
// A pass per obs finds the member-independent part -- the path through the
// operator's tests and the bracket of the target in each corner's depth
// column -- then one thread per (obs, member), obs fastest, so that a
// wavefront reads the neighbouring obs of one member, evaluates it.  Obs
// arrive in file order, scattered over the grid; above a few thousand obs
// the threads visit them in grid-cell order (a radix sort of the cell
// index, stable), so that neighbouring threads read the same or adjacent
// model columns.  Every value follows the reference's operation order in r_size, with the double
// literals of akima_spline promoted to FP64 as the Fortran promotes them;
// contraction into FMA is off, so results are bit-identical to the CPU
// restatement of the same routines.
#include <hipcub/hipcub.hpp>

#include "letkf_handle.h"

namespace letkf {

namespace {

template <typename T>
struct XtoyArgs {
  int nlon, nlat, nlev, nv3d, nv2d;
  int id_u, id_v, id_t, id_s, id_z;
  int iv_u, iv_v, iv_t, iv_s, iv_z;
  T hnosea, undef;
  const T *lon, *lat, *depth, *fsm;
  int nobs, nmem;
  const int32_t *elm, *idx, *idy;
  const T *lono, *lato, *levo;
  const T *v3d, *v2d;  // element (i,j,k,v,m) at i + nlon j + nlon nlat k + sv3 v + sm3 m (v2d: no k)
  long long sv3, sm3, sv2, sm2;
  const int32_t* order;  // visit order of the obs (NULL: input order)
  int32_t* pk;           // per obs: path + 4 * variable slot (k_xtoy_prep)
  int32_t* br;           // per obs: the four corner columns' bracket codes
  T* hx;               // (nobs, nmem)
  int32_t* qc;         // 0 where any member's H(x) is undef
};

template <typename T>
__device__ __forceinline__ T bilinear(T x0, T x1, T y0, T y1, T x, T y, T f00, T f10, T f01, T f11, T undef) {
#pragma clang fp contract(off)
  if (x0 == x1 || y0 == y1) return undef;
  return (y1 - y) / (y1 - y0) * ((x1 - x) / (x1 - x0) * f00 + (x - x0) / (x1 - x0) * f10) +
         (y - y0) / (y1 - y0) * ((x1 - x) / (x1 - x0) * f01 + (x - x0) / (x1 - x0) * f11);
}

// akima_spline (common.f90:276-385) at one target on x(1:ndim), y(1:ndim)
// read with stride st (a model column, levels 2..nlev), in two parts: the
// bracket search depends on the depth column only and is shared by all
// members (akima_bracket); the interpolation reads the member's column
// (akima_eval).  A target equal to x(ndim) ends the reference's bracket
// search past the column (it then reads beyond x and y with dx = 0); y(ndim)
// is its value wherever those reads are finite and is returned here.
//
// Bracket code: -1 target outside [x(1), x(ndim)] (undef); otherwise
// i | h << 16 with i the first level above the target (ndim + 1: none, the
// value is y(h)) and h the last level equal to it (0: none).
template <typename T>
__device__ int akima_bracket(int ndim, const T* __restrict__ x, long long st, T x5) {
#define AX(k) x[((long long)(k) - 1) * st]
  if (x5 < AX(1) || AX(ndim) < x5) return -1;
  int i, h = 0;
  for (i = 1; i <= ndim; ++i) {
    const T xi = AX(i);
    if (x5 == xi) {
      h = i;
      continue;
    }
    if (x5 < xi) break;
  }
  return i | (h << 16);
#undef AX
}

template <typename T>
__device__ T akima_eval(int ndim, const T* __restrict__ x, const T* __restrict__ y, long long st, T x5, int code,
                        T undef) {
#pragma clang fp contract(off)
#define AX(k) x[((long long)(k) - 1) * st]
#define AY(k) y[((long long)(k) - 1) * st]
#define DYDX(m) ((AY(i - 3 + (m)) - AY(i - 4 + (m))) / (AX(i - 3 + (m)) - AX(i - 4 + (m))))
  if (code < 0) return undef;
  const int i = code & 0xffff, h = code >> 16;
  if (i > ndim) return h > 0 ? AY(h) : undef;
  T dydx[6], ddydx[5], t[3];
  if (i == 2) {
    for (int m = 3; m <= 5; ++m) dydx[m] = DYDX(m);
    dydx[2] = (T)(2.0 * (double)dydx[3] - (double)dydx[4]);
    dydx[1] = (T)(2.0 * (double)dydx[2] - (double)dydx[3]);
  } else if (i == 3) {
    for (int m = 2; m <= 5; ++m) dydx[m] = DYDX(m);
    dydx[1] = (T)(2.0 * (double)dydx[2] - (double)dydx[3]);
  } else if (i == ndim) {
    for (int m = 1; m <= 3; ++m) dydx[m] = DYDX(m);
    dydx[4] = (T)(2.0 * (double)dydx[3] - (double)dydx[2]);
    dydx[5] = (T)(2.0 * (double)dydx[4] - (double)dydx[3]);
  } else if (i == ndim - 1) {
    for (int m = 1; m <= 4; ++m) dydx[m] = DYDX(m);
    dydx[5] = (T)(2.0 * (double)dydx[4] - (double)dydx[3]);
  } else {
    for (int m = 1; m <= 5; ++m) dydx[m] = DYDX(m);
  }
  for (int m = 1; m <= 4; ++m) {
    const T dd = dydx[m + 1] - dydx[m];
    ddydx[m] = dd < (T)0 ? -dd : dd;
  }
  for (int m = 1; m <= 2; ++m) {
    const T wk = ddydx[m + 2] + ddydx[m];
    t[m] = wk == (T)0 ? (T)0 : (ddydx[m + 2] * dydx[m + 1] + ddydx[m] * dydx[m + 2]) / wk;
  }
  const T dx21 = AX(i) - AX(i - 1), dx = x5 - AX(i - 1);
  double v = (double)(T)(AY(i - 1) + dx * t[1]);
  v = v + (double)(dx * dx) * (3.0 * (double)dydx[3] - 2.0 * (double)t[1] - (double)t[2]) / (double)dx21;
  v = v + (double)(dx * dx * dx) * ((double)(t[1] + t[2]) - 2.0 * (double)dydx[3]) / (double)dx21 / (double)dx21;
  return (T)v;
#undef DYDX
#undef AX
#undef AY
}

template <typename T>
__global__ void __launch_bounds__(256) k_qc_fill(int n, int32_t* qc) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o < n) qc[o] = 1;
}

// The member-independent part of Trans_XtoY for one obs, in the
// reference's order of tests (common.f90:425-497): the path -- 0 undef,
// 1 SSH (bilinear of the 2D field), 2 above the top level of a corner (the
// surface values), 3 Akima in the four corner columns -- plus 4 x the
// variable slot, and for path 3 the bracket code of each corner column.
template <typename T>
__global__ void __launch_bounds__(256) k_xtoy_prep(XtoyArgs<T> a) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= a.nobs) return;
  const long long nij = (long long)a.nlon * a.nlat;
  const int i = a.idx[o], j = a.idy[o];
  int path = 0, iv = 0;
  int br[4] = {-1, -1, -1, -1};
#define G2(p, ii, jj) (p)[((ii) - 1) + (long long)a.nlon * ((jj) - 1)]
#define G3(p, ii, jj, kk) (p)[((ii) - 1) + (long long)a.nlon * ((jj) - 1) + nij * ((kk) - 1)]
  do {
    if (i < 1 || a.nlon - 1 < i || j < 1 || a.nlat - 1 < j) break;
    if (G2(a.fsm, i, j) == a.hnosea || G2(a.fsm, i + 1, j) == a.hnosea || G2(a.fsm, i, j + 1) == a.hnosea ||
        G2(a.fsm, i + 1, j + 1) == a.hnosea)
      break;
    const int el = a.elm[o];
    if (el == a.id_z) {
      path = 1;
      break;
    }
    iv = el == a.id_u ? a.iv_u : el == a.id_v ? a.iv_v : el == a.id_t ? a.iv_t : el == a.id_s ? a.iv_s : 0;
    if (iv == 0) break;  // an element Trans_XtoY does not know
    const T lv = a.levo[o];
    if ((T)0 < lv) break;  // above the sea surface
    if (lv < G3(a.depth, i, j, 2) || lv < G3(a.depth, i + 1, j, 2) || lv < G3(a.depth, i, j + 1, 2) ||
        lv < G3(a.depth, i + 1, j + 1, 2))
      break;  // below the sea bottom
    if (G3(a.depth, i, j, a.nlev) < lv || G3(a.depth, i + 1, j, a.nlev) < lv ||
        G3(a.depth, i, j + 1, a.nlev) < lv || G3(a.depth, i + 1, j + 1, a.nlev) < lv) {
      path = 2;
      break;
    }
    path = 3;
    const int ci[4] = {i, i + 1, i, i + 1}, cj[4] = {j, j, j + 1, j + 1};
    for (int q = 0; q < 4; ++q) br[q] = akima_bracket<T>(a.nlev - 1, &G3(a.depth, ci[q], cj[q], 2), nij, lv);
  } while (false);
#undef G2
#undef G3
  a.pk[o] = path + 4 * iv;
  *reinterpret_cast<int4*>(a.br + 4LL * o) = make_int4(br[0], br[1], br[2], br[3]);
}

// One thread per (obs, member): the member's values along the path and
// brackets k_xtoy_prep found.
template <typename T>
__global__ void __launch_bounds__(256) k_trans_xtoy(XtoyArgs<T> a) {
#pragma clang fp contract(off)
  const long long tid = blockIdx.x * 256LL + threadIdx.x;
  if (tid >= (long long)a.nobs * a.nmem) return;
  const int t = (int)(tid % a.nobs), mem = (int)(tid / a.nobs);
  const int o = a.order ? a.order[t] : t;
  const long long nij = (long long)a.nlon * a.nlat;
  const T undef = a.undef;
  T h = undef;
  const int pk = a.pk[o], path = pk & 3, iv = pk >> 2;
#define G2(p, ii, jj) (p)[((ii) - 1) + (long long)a.nlon * ((jj) - 1)]
#define G3(p, ii, jj, kk) (p)[((ii) - 1) + (long long)a.nlon * ((jj) - 1) + nij * ((kk) - 1)]
  if (path != 0) {
    const int i = a.idx[o], j = a.idy[o];
    const T x0 = G2(a.lon, i, j), x1 = G2(a.lon, i + 1, j), y0 = G2(a.lat, i, j), y1 = G2(a.lat, i, j + 1);
    if (path == 1) {
      const T* f = a.v2d + a.sm2 * mem + a.sv2 * (a.iv_z - 1);
      h = bilinear<T>(x0, x1, y0, y1, a.lono[o], a.lato[o], G2(f, i, j), G2(f, i + 1, j), G2(f, i, j + 1),
                      G2(f, i + 1, j + 1), undef);
    } else {
      const T* f = a.v3d + a.sm3 * mem + a.sv3 * (iv - 1);
      const int ci[4] = {i, i + 1, i, i + 1}, cj[4] = {j, j, j + 1, j + 1};
      T c[4];
      if (path == 2) {
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = G3(f, ci[q], cj[q], a.nlev);
      } else {
        const int4 b4 = *reinterpret_cast<const int4*>(a.br + 4LL * o);
        const int br[4] = {b4.x, b4.y, b4.z, b4.w};
        const T lv = a.levo[o];
        for (int q = 0; q < 4; ++q)
          c[q] = akima_eval<T>(a.nlev - 1, &G3(a.depth, ci[q], cj[q], 2), &G3(f, ci[q], cj[q], 2), nij, lv, br[q],
                               undef);
      }
      h = bilinear<T>(x0, x1, y0, y1, a.lono[o], a.lato[o], c[0], c[1], c[2], c[3], undef);
    }
  }
#undef G2
#undef G3
  a.hx[o + (long long)a.nobs * mem] = h;
  if (a.qc && h == undef) a.qc[o] = 0;
}

// Cell index of each obs (cells outside 1..nlon-1 x 1..nlat-1 share the
// key ncell) and the identity permutation to sort along.
__global__ void __launch_bounds__(256) k_xtoy_key(int n, int nlon, int nlat, const int32_t* idx, const int32_t* idy,
                                                  int32_t* key, int32_t* vals) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= n) return;
  const int i = idx[o], j = idy[o];
  const bool in = i >= 1 && i <= nlon - 1 && j >= 1 && j <= nlat - 1;
  key[o] = in ? (j - 1) * nlon + (i - 1) : nlon * nlat;
  vals[o] = o;
}

constexpr int kXtoySortMin = 4096;  // below this the input order is kept

template <typename T>
int run_xtoy(const letkf_xtoy_grid_t* g, int nobs, const int32_t* elm, const int32_t* idx, const int32_t* idy,
             const void* lono, const void* lato, const void* levo, int nmem, const void* v3d, const void* v2d,
             const int64_t* strides, void* hx, int32_t* qc, hipStream_t s) {
  XtoyArgs<T> a;
  a.nlon = g->nlon; a.nlat = g->nlat; a.nlev = g->nlev; a.nv3d = g->nv3d; a.nv2d = g->nv2d;
  a.id_u = g->id_u_obs; a.id_v = g->id_v_obs; a.id_t = g->id_t_obs; a.id_s = g->id_s_obs; a.id_z = g->id_z_obs;
  a.iv_u = g->iv3d_u; a.iv_v = g->iv3d_v; a.iv_t = g->iv3d_t; a.iv_s = g->iv3d_s; a.iv_z = g->iv2d_z;
  a.hnosea = (T)g->hnosea;
  a.undef = (T)g->undef;
  a.lon = (const T*)g->lon; a.lat = (const T*)g->lat; a.depth = (const T*)g->depth; a.fsm = (const T*)g->fsm;
  a.nobs = nobs; a.nmem = nmem;
  a.elm = elm; a.idx = idx; a.idy = idy;
  a.lono = (const T*)lono; a.lato = (const T*)lato; a.levo = (const T*)levo;
  a.v3d = (const T*)v3d; a.v2d = (const T*)v2d;
  const long long nij = (long long)g->nlon * g->nlat;
  // default: one packed (nlon,nlat,nlev,nv3d) / (nlon,nlat,nv2d) block per member
  a.sv3 = strides && strides[0] ? strides[0] : nij * g->nlev;
  a.sm3 = strides && strides[1] ? strides[1] : nij * g->nlev * g->nv3d;
  a.sv2 = strides && strides[2] ? strides[2] : nij;
  a.sm2 = strides && strides[3] ? strides[3] : nij * g->nv2d;
  a.hx = (T*)hx; a.qc = qc;
  a.order = nullptr;
  // scratch (freed after the final sync): per-obs path and brackets, and
  // above kXtoySortMin obs the cell sort giving the visit order
  const bool sorted = nobs >= kXtoySortMin;
  const int ncell = g->nlon * g->nlat;
  int bits = 1;
  while (bits < 31 && (1LL << bits) <= ncell) ++bits;
  size_t cub_bytes = 0;
  if (sorted && hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, (int32_t*)nullptr, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int32_t*)nullptr, nobs, 0, bits,
                                                   s) != hipSuccess)
    return set_err(LETKF_ERR_HIP, "letkf_trans_xtoy_device: sort sizing failed");
  const size_t nb = ((sizeof(int32_t) * (size_t)nobs + 255) / 256) * 256;
  void* tmp = nullptr;
  hipError_t e = hipMalloc(&tmp, 5 * nb + (sorted ? 4 * nb + cub_bytes : 0));
  if (e != hipSuccess) return set_err(LETKF_ERR_HIP, "letkf_trans_xtoy_device: %s", hipGetErrorString(e));
  a.pk = (int32_t*)tmp;
  a.br = (int32_t*)((char*)tmp + nb);  // 4 per obs
  auto fail = [&](hipError_t err) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(tmp);
    return set_err(LETKF_ERR_HIP, "letkf_trans_xtoy_device: %s", hipGetErrorString(err));
  };
  if (sorted) {
    char* sb = (char*)tmp + 5 * nb;
    int32_t* key = (int32_t*)sb;
    int32_t* key2 = (int32_t*)(sb + nb);
    int32_t* vals = (int32_t*)(sb + 2 * nb);
    int32_t* order = (int32_t*)(sb + 3 * nb);
    hipLaunchKernelGGL(k_xtoy_key, dim3((unsigned)((nobs + 255) / 256)), dim3(256), 0, s, nobs, g->nlon, g->nlat,
                       idx, idy, key, vals);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortPairs(sb + 4 * nb, cub_bytes, key, key2, vals, order, nobs, 0, bits, s);
    if (e != hipSuccess) return fail(e);
    a.order = order;
  }
  hipLaunchKernelGGL(k_xtoy_prep<T>, dim3((unsigned)((nobs + 255) / 256)), dim3(256), 0, s, a);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e);
  if (qc) {
    hipLaunchKernelGGL(k_qc_fill<T>, dim3((unsigned)((nobs + 255) / 256)), dim3(256), 0, s, nobs, qc);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
  }
  const long long n = (long long)nobs * nmem;
  hipLaunchKernelGGL(k_trans_xtoy<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(s);
  if (e == hipSuccess) e = e2;
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess) return set_err(LETKF_ERR_HIP, "letkf_trans_xtoy_device: %s", hipGetErrorString(e));
  return LETKF_OK;
}

}  // namespace

}  // namespace letkf

extern "C" int letkf_trans_xtoy_device(const letkf_xtoy_grid_t* g, int32_t nobs, const int32_t* elm,
                                       const int32_t* idx, const int32_t* idy, const void* lono, const void* lato,
                                       const void* levo, int32_t nmem, const void* v3d, const void* v2d,
                                       const int64_t* strides, void* hx, int32_t* qc, void* stream) {
  using namespace letkf;
  if (!g) return set_err(LETKF_ERR_ARG, "letkf_trans_xtoy_device: grid is NULL");
  if (g->real_kind != 4 && g->real_kind != 8)
    return set_err(LETKF_ERR_ARG, "real_kind must be 4 or 8 (got %d)", g->real_kind);
  if (g->nlon < 2 || g->nlat < 2 || g->nlev < 6 || g->nv3d < 1 || g->nv2d < 1)
    return set_err(LETKF_ERR_ARG, "grid (%d,%d,%d) nv3d %d nv2d %d: need nlon, nlat >= 2, nlev >= 6 (the Akima "
                                  "stencil), nv3d, nv2d >= 1", g->nlon, g->nlat, g->nlev, g->nv3d, g->nv2d);
  const int iv[4] = {g->iv3d_u, g->iv3d_v, g->iv3d_t, g->iv3d_s};
  for (int q = 0; q < 4; ++q)
    if (iv[q] < 1 || iv[q] > g->nv3d) return set_err(LETKF_ERR_ARG, "iv3d slot %d out of 1..%d", iv[q], g->nv3d);
  if (g->iv2d_z < 1 || g->iv2d_z > g->nv2d)
    return set_err(LETKF_ERR_ARG, "iv2d_z %d out of 1..%d", g->iv2d_z, g->nv2d);
  if (nobs < 0 || nmem < 1) return set_err(LETKF_ERR_ARG, "nobs %d, nmem %d", nobs, nmem);
  if ((long long)g->nlon * g->nlat >= INT32_MAX) return set_err(LETKF_ERR_ARG, "letkf_trans_xtoy_device: grid too large");
  if (nobs == 0) return LETKF_OK;
  if (!elm || !idx || !idy || !lono || !lato || !levo || !v3d || !v2d || !hx || !g->lon || !g->lat || !g->depth ||
      !g->fsm)
    return set_err(LETKF_ERR_ARG, "letkf_trans_xtoy_device: NULL array");
  hipStream_t s = (hipStream_t)stream;
  if (strides)
    for (int q = 0; q < 4; ++q)
      if (strides[q] < 0) return set_err(LETKF_ERR_ARG, "letkf_trans_xtoy_device: negative stride");
  return g->real_kind == 8
             ? run_xtoy<double>(g, nobs, elm, idx, idy, lono, lato, levo, nmem, v3d, v2d, strides, hx, qc, s)
             : run_xtoy<float>(g, nobs, elm, idx, idy, lono, lato, levo, nmem, v3d, v2d, strides, hx, qc, s);
}
