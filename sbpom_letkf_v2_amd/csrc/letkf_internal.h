// letkf_internal.h -- device-side types and launcher declarations shared by
// the engine's HIP translation units.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/letkf_engine.h"

namespace letkf {

// Scalars of one analysis, by value into every kernel.
struct KParams {
  int K, nlon, nlat, nlev, nij1, nv3d, nv2d, iv3d_t, iv3d_s, id_z_obs, roff;
  int nobs;
  long long P;  // nij1 * nlev
  double pi, re, rho, alpha_rtpp, alpha_rtps, sigma_obs, sigma_obsv;
  double dist_zero, dist_zerov, tmin, tmax, smin, smax;
  // cos of the great-circle angle dist_zero/re widened by a relative margin
  // far above the rounding of re*acos in r_size: cos d >= cos_in is inside
  // the cut-off and cos d <= cos_out beyond it without evaluating acos
  double cos_in, cos_out;
};

// Local grid + sorted obs tables, typed by the caller's REAL.
template <typename T>
struct Tables {
  const T *lon1, *lat1, *ri1, *rj1, *dlon1, *dlat1, *phi1, *depth1;
  const int32_t *nobsgrd, *elm;
  const T *olon, *olat, *olev, *oerr, *odep;
  const double2* osc;  // per obs (sin, cos) of its latitude in radians, as distance_2p forms them
};

// Per-analysis device arrays over all P points.
struct PointArrays {
  int32_t* nloc;    // [P]   local obs count (nobsl)
  int64_t* loff;    // [P+1] exclusive scan of nloc
  int32_t* actidx;  // [P+1] exclusive scan of (nloc > 0)
  int32_t* act_p;   // [nactive] point of each active slot
};

// Status bits raised by kernels (atomicOr into a device word).
enum : int {
  ST_EIG_NOCONV = 1,
  ST_EIG_NEG = 2,
  ST_CAPACITY = 4,
  ST_NONFINITE = 8
};

constexpr int kMaxK = 128;  // largest ensemble handled on chip

// ---- launchers (each defined in the .hip that holds its kernel) ----
template <typename T>
hipError_t launch_obs_trig(int nobs, double pi, const T* olat, double2* osc, hipStream_t s);
template <typename T>
hipError_t launch_search_count(const KParams& kp, const Tables<T>& tb,
                               int32_t* nloc, hipStream_t s);
// fast_loc: rloc from the FP32 exponential (v_exp_f32, ~1e-7 relative) in
// place of the FP64 one -- the mixed-precision analysis (column kernel only)
template <typename T>
hipError_t launch_search_fill(const KParams& kp, const Tables<T>& tb,
                              long long p0, long long p1, const int64_t* loff,
                              int32_t* lobs, double* lrloc, hipStream_t s, bool fast_loc = false);

hipError_t launch_scan_points(long long P, const int32_t* nloc, int64_t* loff,
                              int32_t* actidx, int32_t* act_p, void* tmp,
                              size_t tmp_bytes, hipStream_t s);
size_t scan_tmp_bytes(long long P);

template <typename T>
hipError_t launch_transpose_obs(int nobs, int K, const T* hdxf, double* yT,
                                hipStream_t s);

// Gram G = Y^T diag(rloc/rdiag) Y + (K-1)/rho I and b = Y^T diag(rloc/rdiag) d
// for active slots [a0, a0+na) -> G point-major [na][K][K], bvec [na][K].
// Only the upper triangle (with the diagonal) of each G is written: the
// eigensolvers read no other entry.
template <typename T>
hipError_t launch_gram_fp64(const KParams& kp, const Tables<T>& tb,
                            const double* yT, const int32_t* act_p,
                            const int64_t* loff, long long lbase,
                            const int32_t* lobs, const double* lrloc, int a0,
                            int na, double* G, double* bvec, int dual,
                            hipStream_t s);

// The same products with the Gram on MFMA in FP32 (prec LETKF_PREC_FP32)
// or from BF16 hi/lo splits (LETKF_PREC_BF16X3), FP32 accumulation (per
// panel into FP32 totals for LETKF_PREC_FP32_PANEL), b in
// FP64; K <= 128.
// ysp (optional, BF16X3 only): the obs rows pre-split by launch_split_obs,
// which the dual-form points' Gram then reads instead of yT.  klist / kcnt
// (optional, BF16X3): the slots of the size bucket that holds the K x K
// problems, the only ones the K x K kernel (nmem <= 64) or, in the dual
// form, the block kernel (nmem > 64) then has to visit.
template <typename T>
hipError_t launch_gram_mixed(int prec, const KParams& kp, const Tables<T>& tb,
                             const double* yT, const int32_t* act_p,
                             const int64_t* loff, long long lbase,
                             const int32_t* lobs, const double* lrloc, int a0,
                             int na, double* G, double* bvec, int dual,
                             const void* ysp, hipStream_t s,
                             const int32_t* klist = nullptr, int kcnt = 0);

// Obs rows of yT (nobs x K) as BF16 hi / lo pairs, 8 members a group
// (zero past K), split_obs_groups(K) groups a row: 32 bytes a group, hi
// then lo.
__host__ __device__ inline int split_obs_groups(int K) { return 2 * ((K + 15) / 16); }
hipError_t launch_split_obs(int nobs, int K, const double* yT, void* ysp, hipStream_t s);

// The FP64 products again for the slots list[0..cnt) with flags[slot] == -1
// (a size bucket's slots that the mixed-precision update left to the
// fallback), so that the fallback solves them from an FP64 Gram.
template <typename T>
hipError_t launch_gram_fp64_flagged(const KParams& kp, const Tables<T>& tb, const double* yT, const int32_t* act_p,
                                    const int64_t* loff, long long lbase, const int32_t* lobs, const double* lrloc,
                                    int a0, const int32_t* list, int cnt, const int32_t* flags, double* G,
                                    double* bvec, int dual, hipStream_t s);

// Same products for letkf_core_batched: problem b owns rows offs[b]..offs[b+1)
// of hdxf (K values per row); nrow[b] receives the row count.
template <typename T>
hipError_t launch_gram_csr(int K, double rho, int nb, const int64_t* offs,
                           const T* hdxf, const T* rdiag, const T* rloc,
                           const T* dep, double* G, double* bvec,
                           int32_t* nrow, hipStream_t s);

// From the eigen-decomposition: wvec = V lam^-1 V^T b and
// Wmat = V diag(sqrt((K-1)/lam)) V^T, or letkf_core_noobs where nrow == 0.
template <typename T>
hipError_t launch_form_weights(int K, double rho, int nb, const int32_t* nrow,
                               const double* lam, const double* V,
                               const double* bvec, T* wvec, T* wmat,
                               hipStream_t s);

// LPF / LPFGM inside da_main (iswitch_da 2/3): settings, draws and the
// running-sum table on the device, and the per-point kernel over points
// [p0, p1) of a chunk (G/bvec: the chunk's K x K Gram slots, mode 3 only;
// the kernel overwrites a point's slot with its W).
struct LpfPointArgs {
  int mode, nmonte, dp;
  double alpha;  // (K-1)/cov_infl_gm in the real kind
  const double* rtab;
  const double* ttab;
  int* status;   // status bits: ST_NONFINITE when an analysis value is not finite
};
template <typename T>
hipError_t launch_lpfgm_points(const KParams& kp, const Tables<T>& tb, const T* f3, const T* f2, T* a3, T* a2,
                               const int32_t* nloc, const int64_t* loff, long long lbase, const int32_t* lobs,
                               const double* lrloc, const double* yT, const int32_t* actidx, int a0, double* G,
                               const double* bvec, long long p0, long long p1, const LpfPointArgs& la,
                               hipStream_t s);

// AOEI and obs binning pre-passes.
template <typename T>
hipError_t launch_aoei(int nobs, T* err, const T* dep, const T* sprd,
                       hipStream_t s);
template <typename T>
hipError_t launch_obs_cells(int nlon, int nlat, const T* lon, const T* lat,
                            int nobs, const T* olon, const T* olat,
                            int32_t* idx, int32_t* idy, int32_t* key,
                            int32_t* counts, hipStream_t s);

// Batched symmetric eigensolve: slot b holds an n_b x n_b matrix at
// G + b*K*K with row pitch n_b (upper triangle read) -> lam[b*K + j]
// descending, V + b*K*K row pitch n_b (V[r][j] = component r of vector j).
// n_b = K, or (dual != 0) n_b = min(nloc[act_p[a0+b]], K).
hipError_t launch_eig_fp64(int K, int na, const double* G, double* lam,
                           double* V, int* status, int dual,
                           const int32_t* nloc, const int32_t* act_p, int a0,
                           hipStream_t s);

// Three-phase variant (tridiagonalise / QL with rotation log / apply);
// Q, D, E, lc, ls, li, cnt are per-slot workspaces (K*K, K, K, cap, cap,
// cap, 1 per slot, cap = eig_log_capacity(K)).
long long eig_log_capacity(int K);
hipError_t launch_eig3(int K, int na, const double* G, double* Q, double* D,
                       double* E, double* lc, double* ls, int32_t* li,
                       int32_t* cnt, double* lam, double* V, int* status,
                       int dual, const int32_t* nloc, const int32_t* act_p,
                       int a0, hipStream_t s);

// Register-resident eigensolve (eigreg.inc, K <= 64) fused with the
// ensemble update: writes the analysis of the point of every listed slot
// (all of size <= nb, nb in {16, 32, 48, 64}).  Per-slot workspaces: Q K*K,
// D/E K, cs 2*eigr_rot_capacity, ix eigr_ix_pitch bytes, hdr
// 2*eigr_sweep_capacity, count and nrot 1, Y yp >= K*(1+nv3d+nv2d); Q may
// alias V.  Slots whose rotation log overflowed get lam/V instead
// (count[slot] == -1, and status[1] set) for launch_transform's fallback
// mode.  vpath selects the vector path (0 default rule, 1 Q accumulated
// from the reflectors, K <= 64 only; 2 reflectors applied to the vectors
// one wavefront sum per column; 3 reflectors applied in the (row group,
// column) layout, 1 + nv3d + nv2d <= 8); eigr_vpath resolves it for a
// bucket (a path the bucket cannot take falls back to the rule).
// tri32: E1 (the Householder reduction) in FP32 arithmetic.
// rat (the table of rat_table_data on the device): f(T) on the vectors by
// rational functions and shifted tridiagonal solves instead of E2 + V2;
// slots it cannot take (not SPD, spectrum past the table) go to the same
// fallback.
// Events around every eigensolve kernel launch (stage_timing 2): launch u
// is bracketed by ev[2u], ev[2u+1] and tagged 4*kind + bucket.
struct EigTimer {
  hipEvent_t* ev;
  int* tag;
  int cap, used;
  unsigned long long* rot;  // device counter: logged rotations of every slot (may be null)
};

// Point-major copy of a chunk's ensemble for the register update: member m
// of variable v at point p in X[(p - p0) ld + v K + m], ld = (nv3d + nv2d) K
// (the 2-D variables at the surface points only).  X null: the update reads
// the caller's arrays; A null: it writes them.
template <typename T>
struct PointMajor {
  const T* X;
  T* A;
  long long p0;
  int ld;
  const double2* S = nullptr;  // forecast (mean, spread) of each (chunk point, variable), or none
};
// S (nmem <= 64, optional): the forecast mean and spread of every (point,
// variable) of the chunk, S[(p - p0) (nv3d + nv2d) + v], formed from the
// tile the copy holds (the variables a point does not have are left).
template <typename T>
hipError_t launch_to_point_major(const KParams& kp, long long p0, long long p1, const T* X3, const T* X2,
                                 T* Xp, int ld, hipStream_t s, double2* S = nullptr);
template <typename T>
hipError_t launch_from_point_major(const KParams& kp, long long p0, long long p1, const T* Ap, int ld,
                                   const int32_t* nloc, T* A3, T* A2, hipStream_t s);

long long eigr_rot_capacity(int K);
long long eigr_sweep_capacity(int K);
int eigr_vpath(int vpath, int nb, int K, int ncol);
long long eigr_ix_pitch(int K);
template <typename T>
hipError_t launch_eigr_update(int nb, int K, int dual, const int32_t* nloc,
                              const int32_t* act_p, int a0, const int32_t* list,
                              int cnt, const double* G, double* Q, double* D,
                              double* E, double* cs, uint8_t* ix, int32_t* hdr,
                              int32_t* count, int32_t* nrot, double* lam,
                              double* V, double* Y, long long yp,
                              const KParams& kp, const Tables<T>& tb,
                              const T* X3, const T* X2, T* A3, T* A2,
                              const int64_t* loff, long long lbase,
                              const int32_t* lobs, const double* lrloc,
                              const double* yT, const double* bvec,
                              int* status, int vpath, const double* rat, bool tri32, EigTimer* tm,
                              const PointMajor<T>& pm, double kmax, hipStream_t s);
// The rational path's coefficient table (ratfun.inc), rat_table_size()
// doubles, built once per process; rat == nullptr above selects E2 + V2.
// Layout: [kind][class][n, d, a[kRatTablePoles], s[kRatTablePoles]], kinds
// primal, dual, and the two after an FP32 reduction (looser targets).
constexpr int kRatTableClasses = 24, kRatTablePoles = 40, kRatTableKinds = 4;
const double* rat_table_data();
size_t rat_table_size();
// Zolotarev's approximation of x^-1/2 on [1, kappa] of order n: d, a[n],
// s[n]; returns its max relative error.
long double rat_invsqrt(long double kappa, int n, long double* d, long double* a, long double* s);

// Adjoint of the analysis (FP64, K <= 64): the active slots [0,na) of a
// chunk (C and b from launch_gram_fp64 in the K x K form) and the points
// [p0,p1) without local obs: each (point, local obs) pair writes its
// (dL/d dep, dL/d log r) terms into gpair (2 per local obs of the chunk),
// each point (dL/dalpha, dL/drho, dL/dsigma_obs, dL/dsigma_obsv) into part.
hipError_t launch_adjoint_points(const KParams& kp, const Tables<double>& tb, const double* X3, const double* X2,
                                 const double* B3, const double* B2, const int32_t* nloc, const int64_t* loff,
                                 long long lbase, const int32_t* lobs, const double* lrloc, const double* yT,
                                 const double* G, const double* bvec, const int32_t* act_p, int a0, int na,
                                 long long p0, long long p1, double* gpair, double* part, int* status,
                                 hipStream_t s);
// gdep / glogvar += the chunk's pair terms summed per obs in point order
// (kv: 3 nl int32 of scratch, tmp: adjoint_sort_tmp_bytes).
hipError_t launch_adjoint_pairs(long long nl, int nobs_bits, const int32_t* lobs, const double* gpair, int32_t* kv,
                                void* tmp, size_t tmp_bytes, double* gdep, double* glogvar, hipStream_t s);
size_t adjoint_sort_tmp_bytes(long long nl, int nobs_bits);
// out[0..3] = sums over the P points of part (blk: 1024 doubles of scratch).
hipError_t launch_adjoint_reduce(const double* part, long long P, double* blk, double* out, hipStream_t s);

// Weights + ensemble transform + RTPP/RTPS + round_off for points [p0,p1).
template <typename T>
hipError_t launch_transform(const KParams& kp, const Tables<T>& tb,
                            long long p0, long long p1, const int32_t* nloc,
                            const int32_t* actidx, int a0, const double* lam,
                            const double* V, const double* bvec,
                            const T* fcst3d, const T* fcst2d, T* anal3d,
                            T* anal2d, int dual, const int64_t* loff,
                            long long lbase, const int32_t* lobs,
                            const double* lrloc, const double* yT,
                            const int32_t* fallback_count, const int* ovf,
                            hipStream_t s);

template <typename T>
hipError_t launch_copy(const T* src, T* dst, size_t n, hipStream_t s);

}  // namespace letkf
