// transform.hip -- weights, ensemble transform and inflation of da_main
// for a range of points (letkf/run/src/common_da.f90:74-113 and the
// round_off of :506-533), given the eigen-decomposition C = V diag(lam) V^T
// of each active point's ensemble-space matrix and b = Rlinv_dYf^T dep.
//
// With s_c = sqrt((K-1)/lam_c) the reference's
//   Pa = V lam^-1 V^T, wvec = Pa b, Wmat = V diag(s) V^T      (common_letkf.f90:139-208)
//   anal(j) = xf_mean + sum_i dxf(i) (wvec(i) + Wmat(i,j))    (common_da.f90:423-430)
// are applied without forming Pa or Wmat:
//   q = lam^-1 (V^T b), w = V q, z = diag(s) V^T dxf, anal = mean + dxf.w + V z
// which costs 4K^2 per point instead of 4K^3.  Points without local obs take
// letkf_core_noobs (W = sqrt(rho) I, w = 0; common_letkf.f90:37-57).
//
// Layout: fcst3d(nij1,nlev,K,nv3d) is read as [v][m][p] with p = i + nij1*k,
// so a tile of consecutive points is one contiguous run per (m,v): each
// workgroup stages TI consecutive points x all members x all variables in
// LDS with coalesced loads, each wavefront then owns TI/4 points, lane j
// owning members j and j+64.
#include "letkf_internal.h"
#include "wave.h"

namespace letkf {

// Wavefront sum (DPP butterfly; every lane receives the total).
__device__ __forceinline__ double wsum(double v) { return wave_sum(v); }

// Orders this wavefront's LDS writes before its later LDS reads; waves of
// the workgroup work on different points, so no workgroup barrier here.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

constexpr int kMaxVars = 8;

// What the dual branch needs to rebuild B = diag(sqrt(rloc/rdiag)) Y of a
// point: its local list (lobs/lrloc from loff[p]-lbase) and the obs-major
// FP64 copy of obshdxf.  enabled == 0 for the plain form.
struct DualRows {
  int enabled;
  // non-null: update only the points whose slot count is -1 (the others
  // were updated by the fused eigensolve or by k_noobs)
  const int32_t* fallback_count;
  const int64_t* loff;
  long long lbase;
  const int32_t* lobs;
  const double* lrloc;
  const double* yT;
  double alpha;     // (K-1)/rho in r_size
  const int* ovf;   // register path: nonzero once any eigensolve log overflowed
  int* status;      // status bits (ST_NONFINITE: a non-finite analysis at a point with local obs)
};

// What one launch of k_transform updates.
enum TransformMode {
  TM_ALL = 0,    // every point (the LDS eigensolve path)
  TM_OVF = 2     // points whose eigensolve log overflowed (E4 decomposition)
};

// Points without local obs on the register path (letkf_core_noobs: W =
// sqrt(rho) I, w = 0, common_letkf.f90:37-57), then RTPP / RTPS /
// round_off: one thread per (point, variable), the members walked in order
// (consecutive threads hold consecutive points, so every member row is one
// coalesced run); the passes after the first re-read the point's members
// from the caches.  Points with local obs are left to the eigensolve.
template <typename T>
__global__ void __launch_bounds__(256) k_noobs(KParams kp, Tables<T> tb, long long p0, long long p1,
                                               const int32_t* __restrict__ nloc, const T* __restrict__ X3,
                                               const T* __restrict__ X2, T* __restrict__ A3, T* __restrict__ A2) {
  const long long p = p0 + blockIdx.x * 256LL + threadIdx.x;
  const int v = blockIdx.y;
  if (p >= p1 || nloc[p] != 0) return;
  const int K = kp.K, nv3 = kp.nv3d;
  const long long P = kp.P, surf0 = (long long)kp.nij1 * (kp.nlev - 1);
  const T* x;
  T* y;
  long long st;  // member stride
  if (v < nv3) {
    x = X3 + p + P * (long long)K * v;
    y = A3 + p + P * (long long)K * v;
    st = P;
  } else {
    if (p < surf0) return;
    x = X2 + (p - surf0) + (long long)kp.nij1 * K * (v - nv3);
    y = A2 + (p - surf0) + (long long)kp.nij1 * K * (v - nv3);
    st = kp.nij1;
  }
  const double rK = 1.0 / (double)K;
  const double sqrt_rho = (double)sqrt((T)kp.rho);
  const double ap = (double)(T)kp.alpha_rtpp, as_ = (double)(T)kp.alpha_rtps;
  // member loops unrolled by 8: eight loads in flight per thread
  double s = 0.0;
#pragma unroll 8
  for (int m = 0; m < K; ++m) s += (double)x[m * st];
  const double fm = s * rK;
  double sd = 0.0, sa = 0.0;
#pragma unroll 8
  for (int m = 0; m < K; ++m) {
    const double d = (double)x[m * st] - fm;
    sd = fma(d, d, sd);
    sa += fm + d * sqrt_rho;
  }
  const double fs = sqrt(sd / (double)(K - 1));
  const double am = sa * rK;
  double sda = 0.0;
  if (as_ > 0.0)
#pragma unroll 8
    for (int m = 0; m < K; ++m) {
      const double da = fm + ((double)x[m * st] - fm) * sqrt_rho - am;
      sda = fma(da, da, sda);
    }
  const double asd = sqrt(sda / (double)(K - 1));
  const int k = (int)(p / kp.nij1), i = (int)(p - (long long)k * kp.nij1);
  const bool clampable = kp.roff && k >= 1 && !((T)tb.phi1[i] == (T)1) && v < nv3;
  const bool is_t = v == kp.iv3d_t - 1, is_s = v == kp.iv3d_s - 1;
  const double lo = is_t ? (double)(T)kp.tmin : (double)(T)kp.smin;
  const double hi = is_t ? (double)(T)kp.tmax : (double)(T)kp.smax;
  const double f = asd == 0.0 ? 0.0 : as_ * fs / asd + (1.0 - as_);
#pragma unroll 8
  for (int m = 0; m < K; ++m) {
    const double d = (double)x[m * st] - fm;
    const double a = fm + d * sqrt_rho;
    const double da = a - am;
    double out = a;
    if (ap > 0.0) out = am + ap * d + (1.0 - ap) * da;
    if (as_ > 0.0) out = am + da * f;
    if (clampable && (is_t || is_s)) out = fmin(fmax(out, lo), hi);
    y[m * st] = (T)out;
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_transform(KParams kp, Tables<T> tb, long long p0, long long p1, int TI,
                                                   const int32_t* __restrict__ nloc,
                                                   const int32_t* __restrict__ actidx, int a0,
                                                   const double* __restrict__ lam, const double* __restrict__ V,
                                                   const double* __restrict__ bvec, const T* __restrict__ X3,
                                                   const T* __restrict__ X2, T* __restrict__ A3,
                                                   T* __restrict__ A2, DualRows dr) {
  extern __shared__ double lds[];
  const int K = kp.K, nv3 = kp.nv3d, nv2 = kp.nv2d, nvt = nv3 + nv2;
  const long long P = kp.P;
  const long long surf0 = (long long)kp.nij1 * (kp.nlev - 1);
  const long long pt = p0 + (long long)blockIdx.x * TI;
  const int ntile = (int)min((long long)TI, p1 - pt);
  double* xs = lds;                                  // [nvt][K][TI]
  double* scr = xs + (size_t)nvt * K * TI;           // per wave: 2 banks of (1+nvt)*K + K ints
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double* q_s = scr + (size_t)wave * (2 * (1 + nvt) * K + K / 2 + 1);
  double* z_s = q_s + K;                             // bank A: q[K], z[nvt][K]
  double* y_s = q_s + (size_t)(1 + nvt) * K;         // bank B: [nvt+1][K]
  int* io_s = (int*)(y_s + (size_t)(1 + nvt) * K);   // K obs indices
  // a tile with no point for this launch is done before anything is
  // staged (and only its points are staged)
  if (MODE == TM_OVF && *dr.ovf == 0) return;
  __shared__ int need_s[16];
  if (t < TI) {
    bool need = t < ntile;
    if (need && MODE != TM_ALL) {
      const long long p = pt + t;
      need = nloc[p] > 0 && dr.fallback_count[actidx[p] - a0] == -1;
    }
    need_s[t] = need;
  }
  if (!__syncthreads_or(t < TI && need_s[t])) return;
  // ---- stage the forecast tile ----
  // Thread t keeps point tp = t % TI (TI is 8 or 16) and walks the (variable,
  // member) rows vm = t / TI, t / TI + 256 / TI, ...; element e = vm TI + tp.
  const int tpt = t & (TI - 1), vstep = 256 / TI;
  const bool tneed = need_s[tpt];
  {
    int m = (t / TI) % K, v = (t / TI) / K;
#pragma unroll 4
    for (int vm = t / TI; vm < nvt * K; vm += vstep) {
      double x = 0.0;
      if (tneed) {
        const long long p = pt + tpt;
        if (v < nv3)
          x = (double)X3[p + P * (m + (long long)K * v)];
        else if (p >= surf0)
          x = (double)X2[(p - surf0) + (long long)kp.nij1 * (m + (long long)K * (v - nv3))];
      }
      xs[vm * TI + tpt] = x;
      m += vstep;
      while (m >= K) {
        m -= K;
        ++v;
      }
    }
  }
  __syncthreads();
  const T rho_r = (T)kp.rho;
  const double sqrt_rho = (double)sqrt(rho_r);
  const double ap = (double)(T)kp.alpha_rtpp, as_ = (double)(T)kp.alpha_rtps;
  const int m0 = lane, m1 = lane + 64;
  const bool h0 = m0 < K, h1 = m1 < K;
  const double rK = 1.0 / (double)K;
  const int ppw = (TI + 3) / 4;
  const double tmin = (double)(T)kp.tmin, tmax = (double)(T)kp.tmax;
  const double smin = (double)(T)kp.smin, smax = (double)(T)kp.smax;
  for (int jj = 0; jj < ppw; ++jj) {
    const int tp = wave * ppw + jj;
    if (tp >= ntile) break;
    const long long p = pt + tp;
    const int k = (int)(p / kp.nij1), i = (int)(p - (long long)k * kp.nij1);
    const bool surf = p >= surf0;
    const int nvp = surf ? nvt : nv3;
    if (!need_s[tp]) continue;  // updated by the fused eigensolve
    const bool clampable = kp.roff && k >= 1 && !((T)tb.phi1[i] == (T)1);
    const int n = nloc[p];
    // per-variable forecast mean and perturbations (members m0, m1)
    double dx0[kMaxVars], dx1[kMaxVars], fm[kMaxVars], fs[kMaxVars], ta0[kMaxVars], ta1[kMaxVars];
    _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
      const double x0 = h0 ? xs[((size_t)v * K + m0) * TI + tp] : 0.0;
      const double x1 = h1 ? xs[((size_t)v * K + m1) * TI + tp] : 0.0;
      fm[v] = wsum(x0 + x1) * rK;
      dx0[v] = h0 ? x0 - fm[v] : 0.0;
      dx1[v] = h1 ? x1 - fm[v] : 0.0;
      fs[v] = sqrt(wsum(dx0[v] * dx0[v] + dx1[v] * dx1[v]) / (double)(K - 1));
    }
    if (n == 0) {
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        ta0[v] = fm[v] + dx0[v] * sqrt_rho;
        ta1[v] = fm[v] + dx1[v] * sqrt_rho;
      }
    } else if (dr.enabled && n < K) {
      // ---- dual form: n x n problem M = B B^T = U diag(s) U^T ----
      //   w      = B^T U diag(1/(alpha+s)) U^T dt,   dt = sqrt(rloc/rdiag) dep
      //   W dx   = sqrt(rho) dx + sqrt(K-1) B^T U diag(g(s)) U^T B dx
      //   g(s)   = ((alpha+s)^-1/2 - alpha^-1/2)/s, evaluated without cancellation
      const long long sl = (long long)(actidx[p] - a0);
      const double* Ub = V + sl * K * K;   // pitch n
      const double* sb = lam + sl * K;
      const double* db = bvec + sl * K;
      const long long o0 = dr.loff[p] - dr.lbase;
      const bool g0 = lane < n, g1 = lane + 64 < n;
      int io0 = 0, io1 = 0;
      double sw0 = 0.0, sw1 = 0.0;
      if (g0) {
        io0 = dr.lobs[o0 + lane];
        const T er = tb.oerr[io0];
        sw0 = sqrt(dr.lrloc[o0 + lane] / (double)(T)(er * er));
        io_s[lane] = io0;
      }
      if (g1) {
        io1 = dr.lobs[o0 + lane + 64];
        const T er = tb.oerr[io1];
        sw1 = sqrt(dr.lrloc[o0 + lane + 64] / (double)(T)(er * er));
        io_s[lane + 64] = io1;
      }
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        if (h0) z_s[v * K + m0] = dx0[v];
        if (h1) z_s[v * K + m1] = dx1[v];
      }
      wave_sync();
      // (1) lanes over obs: t[v]_l = sw_l (y_l . dx[v])
      double t0[kMaxVars], t1[kMaxVars];
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) t0[v] = t1[v] = 0.0;
      for (int m = 0; m < K; ++m) {
        const double y0 = g0 ? dr.yT[(long long)io0 * K + m] : 0.0;
        const double y1 = g1 ? dr.yT[(long long)io1 * K + m] : 0.0;
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
          const double dm = z_s[v * K + m];
          t0[v] = fma(y0, dm, t0[v]);
          t1[v] = fma(y1, dm, t1[v]);
        }
      }
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        if (g0) y_s[v * K + lane] = sw0 * t0[v];
        if (g1) y_s[v * K + lane + 64] = sw1 * t1[v];
      }
      if (g0) y_s[nvp * K + lane] = db[lane];
      if (g1) y_s[nvp * K + lane + 64] = db[lane + 64];
      wave_sync();
      // (2) lanes over eigenpairs c: u[v]_c = g(s_c) (U^T t[v])_c, q_c = (U^T dt)_c / (alpha + s_c)
      double u0[kMaxVars], u1[kMaxVars], qd0 = 0.0, qd1 = 0.0;
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) u0[v] = u1[v] = 0.0;
      for (int l = 0; l < n; ++l) {
        const double U0 = g0 ? Ub[(long long)l * n + lane] : 0.0;
        const double U1 = g1 ? Ub[(long long)l * n + lane + 64] : 0.0;
        const double dl = y_s[nvp * K + l];
        qd0 = fma(U0, dl, qd0);
        qd1 = fma(U1, dl, qd1);
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
          const double tl = y_s[v * K + l];
          u0[v] = fma(U0, tl, u0[v]);
          u1[v] = fma(U1, tl, u1[v]);
        }
      }
      wave_sync();
      const double al = dr.alpha, sal = sqrt(al);
      if (g0) {
        const double sc = sb[lane], ap_ = al + sc, sap = sqrt(ap_);
        const bool ok = ap_ > 0.0;  // s of M = B B^T above -alpha (rounding keeps it there)
        const double gc = ok ? -1.0 / (sal * sap * (sal + sap)) : 0.0;
        q_s[lane] = ok ? qd0 / ap_ : 0.0;
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) z_s[v * K + lane] = gc * u0[v];
      }
      if (g1) {
        const double sc = sb[lane + 64], ap_ = al + sc, sap = sqrt(ap_);
        const bool ok = ap_ > 0.0;
        const double gc = ok ? -1.0 / (sal * sap * (sal + sap)) : 0.0;
        q_s[lane + 64] = ok ? qd1 / ap_ : 0.0;
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) z_s[v * K + lane + 64] = gc * u1[v];
      }
      wave_sync();
      // (3) lanes over obs: vv[v]_l = sw_l (U u[v])_l, zq_l = sw_l (U q)_l
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) t0[v] = t1[v] = 0.0;
      double zq0 = 0.0, zq1 = 0.0;
      for (int c = 0; c < n; ++c) {
        const double U0 = g0 ? Ub[(long long)lane * n + c] : 0.0;
        const double U1 = g1 ? Ub[(long long)(lane + 64) * n + c] : 0.0;
        const double qc = q_s[c];
        zq0 = fma(U0, qc, zq0);
        zq1 = fma(U1, qc, zq1);
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
          const double uc = z_s[v * K + c];
          t0[v] = fma(U0, uc, t0[v]);
          t1[v] = fma(U1, uc, t1[v]);
        }
      }
      wave_sync();
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        if (g0) y_s[v * K + lane] = sw0 * t0[v];
        if (g1) y_s[v * K + lane + 64] = sw1 * t1[v];
      }
      if (g0) y_s[nvp * K + lane] = sw0 * zq0;
      if (g1) y_s[nvp * K + lane + 64] = sw1 * zq1;
      wave_sync();
      // (4) lanes over members: r[v]_m = (Y^T vv[v])_m, w_m = (Y^T zq)_m
      double r0[kMaxVars], r1[kMaxVars], w0 = 0.0, w1 = 0.0;
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) r0[v] = r1[v] = 0.0;
      for (int l = 0; l < n; ++l) {
        const long long row = (long long)io_s[l] * K;
        const double y0 = h0 ? dr.yT[row + m0] : 0.0;
        const double y1 = h1 ? dr.yT[row + m1] : 0.0;
        const double zl = y_s[nvp * K + l];
        w0 = fma(y0, zl, w0);
        w1 = fma(y1, zl, w1);
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
          const double vl = y_s[v * K + l];
          r0[v] = fma(y0, vl, r0[v]);
          r1[v] = fma(y1, vl, r1[v]);
        }
      }
      const double skm1 = sqrt((double)(T)(K - 1)), srho = sqrt((double)(T)(K - 1) / al);
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        const double mu = wsum(dx0[v] * w0 + dx1[v] * w1);
        ta0[v] = fm[v] + mu + (srho * dx0[v] + skm1 * r0[v]);
        ta1[v] = fm[v] + mu + (srho * dx1[v] + skm1 * r1[v]);
      }
      wave_sync();
    } else {
      const long long sl = (long long)(actidx[p] - a0);
      const double* Vb = V + sl * K * K;
      const double* lb = lam + sl * K;
      const double* bb = bvec + sl * K;
      // pass A over V by columns: q_c = (V^T b)_c / lam_c, z_c[v] = s_c (V^T dx[v])_c
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        if (h0) z_s[v * K + m0] = dx0[v];
        if (h1) z_s[v * K + m1] = dx1[v];
      }
      wave_sync();
      double qa0 = 0.0, qa1 = 0.0, za0[kMaxVars], za1[kMaxVars];
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) za0[v] = za1[v] = 0.0;
      for (int r = 0; r < K; ++r) {
        const double v0 = h0 ? Vb[(long long)r * K + m0] : 0.0;
        const double v1 = h1 ? Vb[(long long)r * K + m1] : 0.0;
        const double br = bb[r];
        qa0 = fma(v0, br, qa0);
        qa1 = fma(v1, br, qa1);
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
          const double dr = z_s[v * K + r];
          za0[v] = fma(v0, dr, za0[v]);
          za1[v] = fma(v1, dr, za1[v]);
        }
      }
      wave_sync();
      const double km1 = (double)(T)(K - 1);
      if (h0) {
        const double l0 = lb[m0], s0 = sqrt(km1 / l0);
        q_s[m0] = qa0 / l0;
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) z_s[v * K + m0] = s0 * za0[v];
      }
      if (h1) {
        const double l1 = lb[m1], s1 = sqrt(km1 / l1);
        q_s[m1] = qa1 / l1;
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) z_s[v * K + m1] = s1 * za1[v];
      }
      wave_sync();
      // pass B over V by rows: w_j = (V q)_j, t_j[v] = (V z[v])_j
      double w0 = 0.0, w1 = 0.0, tb0[kMaxVars], tb1[kMaxVars];
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) tb0[v] = tb1[v] = 0.0;
      for (int c = 0; c < K; ++c) {
        const double v0 = h0 ? Vb[(long long)m0 * K + c] : 0.0;
        const double v1 = h1 ? Vb[(long long)m1 * K + c] : 0.0;
        const double qc = q_s[c];
        w0 = fma(v0, qc, w0);
        w1 = fma(v1, qc, w1);
        _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
          const double zc = z_s[v * K + c];
          tb0[v] = fma(v0, zc, tb0[v]);
          tb1[v] = fma(v1, zc, tb1[v]);
        }
      }
      _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
        const double mu = wsum(dx0[v] * w0 + dx1[v] * w1);
        ta0[v] = fm[v] + mu + tb0[v];
        ta1[v] = fm[v] + mu + tb1[v];
      }
      wave_sync();
    }
    // analysis mean / spread / perturbation, RTPP, RTPS, round_off
    _Pragma("unroll") for (int v = 0; v < kMaxVars; ++v) if (v < nvp) {
      const double am = wsum(h0 * ta0[v] + h1 * ta1[v]) * rK;
      const double da0 = h0 ? ta0[v] - am : 0.0, da1 = h1 ? ta1[v] - am : 0.0;
      const double asd = sqrt(wsum(da0 * da0 + da1 * da1) / (double)(K - 1));
      double y0 = ta0[v], y1 = ta1[v];
      if (ap > 0.0) {
        y0 = am + ap * dx0[v] + (1.0 - ap) * da0;
        y1 = am + ap * dx1[v] + (1.0 - ap) * da1;
      }
      if (as_ > 0.0) {
        const double f = asd == 0.0 ? 0.0 : as_ * fs[v] / asd + (1.0 - as_);
        y0 = am + da0 * f;
        y1 = am + da1 * f;
      }
      if (clampable && v < nv3) {
        if (v == kp.iv3d_t - 1) {
          y0 = fmin(fmax(y0, tmin), tmax);
          y1 = fmin(fmax(y1, tmin), tmax);
        } else if (v == kp.iv3d_s - 1) {
          y0 = fmin(fmax(y0, smin), smax);
          y1 = fmin(fmax(y1, smin), smax);
        }
      }
      if (h0) xs[((size_t)v * K + m0) * TI + tp] = y0;
      if (h1) xs[((size_t)v * K + m1) * TI + tp] = y1;
      // NaN / Inf in obsdep, obshdxf or the forecast of a point with local obs
      if (!isfinite(wsum((h0 ? y0 : 0.0) + (h1 ? y1 : 0.0))) && lane == 0)
        atomicOr(dr.status, ST_NONFINITE);
    }
  }
  __syncthreads();
  // ---- write the analysis tile ----
  // points past the tile, or updated by the fused eigensolve, are skipped
  if (!tneed) return;
  int m = (t / TI) % K, v = (t / TI) / K;
  const long long p = pt + tpt;
  for (int vm = t / TI; vm < nvt * K; vm += vstep) {
    const double y = xs[vm * TI + tpt];
    if (v < nv3)
      A3[p + P * (m + (long long)K * v)] = (T)y;
    else if (p >= surf0)
      A2[(p - surf0) + (long long)kp.nij1 * (m + (long long)K * (v - nv3))] = (T)y;
    m += vstep;
    while (m >= K) {
      m -= K;
      ++v;
    }
  }
}

// ---------------------------------------------------------------- point-major copies
// The register update reads every member of every variable of one point and
// writes them back; in the caller's layout (point index fastest) those are
// K (nv3d + nv2d) separate cache lines per point, in a point-major copy a few
// contiguous rows.  The copies run per chunk as 64 x 64 (points x members)
// tiles through LDS, both sides coalesced: forecast in before the update,
// analysis out after it (only points with local obs: the others are written
// by k_noobs, and the fallback slots by the tile transform after this).
template <typename T>
__global__ void __launch_bounds__(256) k_to_point_major(KParams kp, long long p0, long long p1, const T* __restrict__ X3,
                                                        const T* __restrict__ X2, T* __restrict__ Xp, int ld,
                                                        double2* __restrict__ S) {
  __shared__ T tile[64][65];
  const int v = blockIdx.z, K = kp.K, nv3 = kp.nv3d;
  const long long pb = p0 + (long long)blockIdx.x * 64;
  const int mb = blockIdx.y * 64;
  const long long surf0 = (long long)kp.nij1 * (kp.nlev - 1);
  if (v >= nv3 && pb + 63 < surf0) return;  // 2-D variables: surface points only
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // the thread's 16 loads issued together, then the LDS stores
  T x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int m = mb + ty + 4 * k;
    const long long p = pb + tx;
    x[k] = T(0);
    if (m < K && p < p1) {
      if (v < nv3)
        x[k] = X3[p + kp.P * (m + (long long)K * v)];
      else if (p >= surf0)
        x[k] = X2[(p - surf0) + (long long)kp.nij1 * (m + (long long)K * (v - nv3))];
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) tile[ty + 4 * k][tx] = x[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int r = ty + 4 * k;
    const long long p = pb + r;
    const int m = mb + tx;
    if (p < p1 && m < K && (v < nv3 || p >= surf0)) Xp[(p - p0) * ld + (long long)K * v + m] = tile[tx][r];
  }
  if (S != nullptr && K <= 64) {
    // the forecast mean and spread of point tx: partial sums over the
    // members ty + 4 k, then the four partials
    __shared__ double red[4][64];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (ty + 4 * k < K) s += (double)tile[ty + 4 * k][tx];
    red[ty][tx] = s;
    __syncthreads();
    const double fm = ((red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx])) * (1.0 / (double)K);
    __syncthreads();
    double s2 = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (ty + 4 * k < K) {
        const double d = (double)tile[ty + 4 * k][tx] - fm;
        s2 = fma(d, d, s2);
      }
    red[ty][tx] = s2;
    __syncthreads();
    const long long p = pb + tx;
    if (ty == 0 && p < p1 && (v < nv3 || p >= surf0))
      S[(p - p0) * (nv3 + kp.nv2d) + v] =
          make_double2(fm, sqrt(((red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx])) / (double)(K - 1)));
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_from_point_major(KParams kp, long long p0, long long p1,
                                                          const T* __restrict__ Ap, int ld,
                                                          const int32_t* __restrict__ nloc, T* __restrict__ A3,
                                                          T* __restrict__ A2) {
  __shared__ T tile[64][65];
  const int v = blockIdx.z, K = kp.K, nv3 = kp.nv3d;
  const long long pb = p0 + (long long)blockIdx.x * 64;
  const int mb = blockIdx.y * 64;
  const long long surf0 = (long long)kp.nij1 * (kp.nlev - 1);
  if (v >= nv3 && pb + 63 < surf0) return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // the thread's 16 loads issued together, then the LDS stores
  T x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const long long p = pb + ty + 4 * k;
    const int m = mb + tx;
    x[k] = p < p1 && m < K && (v < nv3 || p >= surf0) ? Ap[(p - p0) * ld + (long long)K * v + m] : T(0);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) tile[tx][ty + 4 * k] = x[k];
  __syncthreads();
  const long long p = pb + tx;
  const bool act = p < p1 && nloc[p] > 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int r = ty + 4 * k;
    const int m = mb + r;
    if (!act || m >= K) continue;
    if (v < nv3)
      A3[p + kp.P * (m + (long long)K * v)] = tile[r][tx];
    else if (p >= surf0)
      A2[(p - surf0) + (long long)kp.nij1 * (m + (long long)K * (v - nv3))] = tile[r][tx];
  }
}

template <typename T>
hipError_t launch_to_point_major(const KParams& kp, long long p0, long long p1, const T* X3, const T* X2, T* Xp,
                                 int ld, hipStream_t s, double2* S) {
  if (p1 <= p0) return hipSuccess;
  const dim3 grid((unsigned)((p1 - p0 + 63) / 64), (unsigned)((kp.K + 63) / 64), (unsigned)(kp.nv3d + kp.nv2d));
  hipLaunchKernelGGL(k_to_point_major<T>, grid, dim3(256), 0, s, kp, p0, p1, X3, X2, Xp, ld, S);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_from_point_major(const KParams& kp, long long p0, long long p1, const T* Ap, int ld,
                                   const int32_t* nloc, T* A3, T* A2, hipStream_t s) {
  if (p1 <= p0) return hipSuccess;
  const dim3 grid((unsigned)((p1 - p0 + 63) / 64), (unsigned)((kp.K + 63) / 64), (unsigned)(kp.nv3d + kp.nv2d));
  hipLaunchKernelGGL(k_from_point_major<T>, grid, dim3(256), 0, s, kp, p0, p1, Ap, ld, nloc, A3, A2);
  return hipGetLastError();
}
template hipError_t launch_to_point_major<float>(const KParams&, long long, long long, const float*, const float*,
                                                 float*, int, hipStream_t, double2*);
template hipError_t launch_to_point_major<double>(const KParams&, long long, long long, const double*,
                                                  const double*, double*, int, hipStream_t, double2*);
template hipError_t launch_from_point_major<float>(const KParams&, long long, long long, const float*, int,
                                                   const int32_t*, float*, float*, hipStream_t);
template hipError_t launch_from_point_major<double>(const KParams&, long long, long long, const double*, int,
                                                    const int32_t*, double*, double*, hipStream_t);

template <typename T>
hipError_t launch_transform(const KParams& kp, const Tables<T>& tb, long long p0, long long p1,
                            const int32_t* nloc, const int32_t* actidx, int a0, const double* lam,
                            const double* V, const double* bvec, const T* fcst3d, const T* fcst2d, T* anal3d,
                            T* anal2d, int dual, const int64_t* loff, long long lbase, const int32_t* lobs,
                            const double* lrloc, const double* yT, const int32_t* fallback_count,
                            const int* ovf, hipStream_t s) {
  if (p1 <= p0) return hipSuccess;
  const int nvt = kp.nv3d + kp.nv2d;
  if (nvt > kMaxVars) return hipErrorInvalidValue;
  const int TI = kp.K <= 64 ? 16 : 8;
  const size_t tile = sizeof(double) * (size_t)nvt * kp.K * TI;
  const size_t lds = tile + sizeof(double) * 4 * (2 * (size_t)(1 + nvt) * kp.K + kp.K / 2 + 1);
  // ovf is word 1 of the handle's status block, whose word 0 holds the status bits
  DualRows dr{dual, fallback_count, loff, lbase, lobs, lrloc, yT, (double)((T)(kp.K - 1) / (T)kp.rho), ovf,
              const_cast<int*>(ovf) - 1};
  const long long blocks = (p1 - p0 + TI - 1) / TI;
  auto big = [&](auto kern) -> hipError_t {
    if (lds > 65536) {
      hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), lds, s, kp, tb, p0, p1, TI, nloc, actidx, a0, lam,
                       V, bvec, fcst3d, fcst2d, anal3d, anal2d, dr);
    return hipGetLastError();
  };
  if (!fallback_count) return big(k_transform<T, TM_ALL>);
  // register path: the points without local obs, then the overflowed
  // eigensolves, if any
  hipLaunchKernelGGL(k_noobs<T>, dim3((unsigned)((p1 - p0 + 255) / 256), (unsigned)nvt), dim3(256), 0, s, kp, tb, p0,
                     p1, nloc, fcst3d, fcst2d, anal3d, anal2d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return big(k_transform<T, TM_OVF>);
}

template hipError_t launch_transform<float>(const KParams&, const Tables<float>&, long long, long long,
                                            const int32_t*, const int32_t*, int, const double*, const double*,
                                            const double*, const float*, const float*, float*, float*, int,
                                            const int64_t*, long long, const int32_t*, const double*,
                                            const double*, const int32_t*, const int*, hipStream_t);
template hipError_t launch_transform<double>(const KParams&, const Tables<double>&, long long, long long,
                                             const int32_t*, const int32_t*, int, const double*, const double*,
                                             const double*, const double*, const double*, double*, double*,
                                             int, const int64_t*, long long, const int32_t*, const double*,
                                             const double*, const int32_t*, const int*, hipStream_t);

}  // namespace letkf
