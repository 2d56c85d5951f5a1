// eig.hip -- batched FP64 symmetric eigensolve of the ensemble-space
// matrices, replacing mtx_eigen / LAPACK DSYEV
// (letkf/run/src/common.f90:992-1050): upper triangle read, eigenvalues
// returned in descending order, negative eigenvalues set to 0 with their
// vectors ("reconditioning"), status raised when every eigenvalue is
// negative or an eigenvalue does not converge.
//
// Method: Householder reduction to tridiagonal form with accumulated
// transformations, then implicit-shift QL iterations (Wilkinson & Reinsch,
// Handbook for Automatic Computation II, 1971).  One wavefront per matrix;
// the matrix lives in LDS with an odd row pitch (n+1 doubles) so that
// row-per-lane and column-per-lane sweeps both spread over the banks.  The
// O(n^2) work of every Householder step and every QL rotation is spread
// over the 64 lanes (lane r owns rows r, r+64); the scalar QL recurrence is
// evaluated in every lane.
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "letkf_internal.h"
#include "wave.h"

namespace letkf {

// The eigensolver workgroups are single wavefronts: a wavefront's LDS
// accesses execute in program order, so ordering the compiler is enough.
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Eigen-decomposition of one n x n symmetric matrix held in LDS Z (pitch
// ld, row-major), d/e scratch of n doubles each.  On return d holds the
// eigenvalues (unordered) and column j of Z the eigenvector of d[j].
// Returns false if the QL iteration did not converge.
// Householder reduction of the symmetric n x n matrix in LDS Z (lower
// triangle used) to tridiagonal form, accumulating Q in Z: on return d is
// the diagonal, e[1..n-1] the sub-diagonal (e[0] = 0), Z = Q.
__device__ void tridiag_wave(int n, double* Z, int ld, double* d, double* e) {
  const int lane = threadIdx.x & 63;
#define Z_(r, c) Z[(r) * ld + (c)]
  // ---- Householder tridiagonalisation (lower triangle) ----
  for (int i = n - 1; i >= 1; --i) {
    const int l = i - 1;
    if (l == 0) {
      if (lane == 0) {
        e[i] = Z_(i, 0);
        d[i] = 0.0;
      }
      wsync();
      continue;
    }
    double part = 0.0;
    for (int c = lane; c <= l; c += 64) part += fabs(Z_(i, c));
    const double scale = wave_sum(part);
    if (scale == 0.0) {
      wsync();
      if (lane == 0) {
        e[i] = Z_(i, l);
        d[i] = 0.0;
      }
      wsync();
      continue;
    }
    part = 0.0;
    for (int c = lane; c <= l; c += 64) {
      const double v = Z_(i, c) / scale;
      Z_(i, c) = v;
      part += v * v;
    }
    double h = wave_sum(part);
    wsync();
    const double f = Z_(i, l);
    const double g = f >= 0.0 ? -sqrt(h) : sqrt(h);
    h -= f * g;
    wsync();
    if (lane == 0) {
      e[i] = scale * g;
      Z_(i, l) = f - g;
    }
    wsync();
    // p = A u / h into e[0..l]
    part = 0.0;
    for (int r = lane; r <= l; r += 64) {
      Z_(r, i) = Z_(i, r) / h;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int c = 0;
      for (; c + 3 <= l; c += 4) {
        a0 += (c <= r ? Z_(r, c) : Z_(c, r)) * Z_(i, c);
        a1 += (c + 1 <= r ? Z_(r, c + 1) : Z_(c + 1, r)) * Z_(i, c + 1);
        a2 += (c + 2 <= r ? Z_(r, c + 2) : Z_(c + 2, r)) * Z_(i, c + 2);
        a3 += (c + 3 <= r ? Z_(r, c + 3) : Z_(c + 3, r)) * Z_(i, c + 3);
      }
      for (; c <= l; ++c) a0 += (c <= r ? Z_(r, c) : Z_(c, r)) * Z_(i, c);
      const double acc = (a0 + a1) + (a2 + a3);
      const double pr = acc / h;
      e[r] = pr;
      part += pr * Z_(i, r);
    }
    const double kk = wave_sum(part) / (h + h);
    wsync();
    for (int r = lane; r <= l; r += 64) e[r] = e[r] - kk * Z_(i, r);
    wsync();
    for (int r = lane; r <= l; r += 64) {
      const double ur = Z_(i, r), qr = e[r];
#pragma unroll 4
      for (int c = 0; c <= r; ++c) Z_(r, c) -= ur * e[c] + qr * Z_(i, c);
    }
    if (lane == 0) d[i] = h;
    wsync();
  }
  if (lane == 0) {
    d[0] = 0.0;
    e[0] = 0.0;
  }
  wsync();
  // ---- accumulate Q ----
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int c = lane; c <= l; c += 64) {
        double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
        int r = 0;
        for (; r + 3 <= l; r += 4) {
          g0 += Z_(i, r) * Z_(r, c);
          g1 += Z_(i, r + 1) * Z_(r + 1, c);
          g2 += Z_(i, r + 2) * Z_(r + 2, c);
          g3 += Z_(i, r + 3) * Z_(r + 3, c);
        }
        for (; r <= l; ++r) g0 += Z_(i, r) * Z_(r, c);
        const double g = (g0 + g1) + (g2 + g3);
#pragma unroll 4
        for (int rr = 0; rr <= l; ++rr) Z_(rr, c) -= g * Z_(rr, i);
      }
    }
    wsync();
    if (lane == 0) d[i] = Z_(i, i);
    wsync();
    for (int c = lane; c <= l; c += 64) {
      Z_(c, i) = 0.0;
      Z_(i, c) = 0.0;
    }
    if (lane == 0) Z_(i, i) = 1.0;
    wsync();
  }
#undef Z_
}

// Implicit-shift QL on (d, e) from tridiag_wave, rotations applied to the
// rows of Z by the 64 lanes; the scalar recurrence runs in every lane.
__device__ bool ql_wave(int n, double* Z, int ld, double* d, double* e) {
  const int lane = threadIdx.x & 63;
#define Z_(r, c) Z[(r) * ld + (c)]
  // ---- implicit QL ----
  for (int q = lane + 1; q < n; q += 64) e[q - 1] = e[q];
  wsync();
  if (lane == 0) e[n - 1] = 0.0;
  wsync();
  bool ok = true;
  for (int l = 0; l < n; ++l) {
    int sweeps = 0;
    for (;;) {
      int m;
      for (m = l; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++sweeps > 60) {
        ok = false;
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      // column i+1 of the row carried in a register across the sweep
      double carry0 = lane < n ? Z_(lane, m) : 0.0;
      double carry1 = lane + 64 < n ? Z_(lane + 64, m) : 0.0;
      for (int i = m - 1; i >= l; --i) {
        const double ei = e[i], f = s * ei, b = c * ei;
        r = hypot(f, g);
        wsync();
        if (lane == 0) e[i + 1] = r;
        if (r == 0.0) {
          if (lane == 0) {
            d[i + 1] -= p;
            e[m] = 0.0;
          }
          if (lane < n) Z_(lane, i + 1) = carry0;
          if (lane + 64 < n) Z_(lane + 64, i + 1) = carry1;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        wsync();
        if (lane == 0) d[i + 1] = g + p;
        g = c * r - b;
        if (lane < n) {
          const double zi = Z_(lane, i);
          Z_(lane, i + 1) = s * zi + c * carry0;
          carry0 = c * zi - s * carry0;
        }
        if (lane + 64 < n) {
          const double zi = Z_(lane + 64, i);
          Z_(lane + 64, i + 1) = s * zi + c * carry1;
          carry1 = c * zi - s * carry1;
        }
      }
      if (!ok) break;
      if (deflated) {
        wsync();
        continue;
      }
      if (lane < n) Z_(lane, l) = carry0;
      if (lane + 64 < n) Z_(lane + 64, l) = carry1;
      wsync();
      if (lane == 0) {
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
      wsync();
    }
    if (!ok) break;
  }
  wsync();
#undef Z_
  return ok;
}

__device__ bool eig_wave(int n, double* Z, int ld, double* d, double* e) {
  tridiag_wave(n, Z, ld, d, e);
  return ql_wave(n, Z, ld, d, e);
}

// Descending order (ties keep the reverse of DSYEV's ascending order),
// negative eigenvalues set to 0 with their vectors (common.f90:1030-1043),
// status bits for no convergence / all negative / non-finite.  lb gets the
// n eigenvalues, Vb the eigenvectors with row pitch n.  keep_neg (the dual
// n x n matrix M = B B^T, semi-definite): negative eigenvalues are rounding
// of M and are kept with their vectors (the update uses alpha + s), and an
// all-negative spectrum is no error.
__device__ void finish_eig(int n, const double* Z, int ld, const double* d, double* lb, double* Vb, bool ok,
                           int* status, bool keep_neg) {
  const int lane = threadIdx.x & 63;
  double dmax = -INFINITY;
  bool finite = true;
  for (int j = 0; j < n; ++j) {
    dmax = fmax(dmax, d[j]);
    finite = finite && isfinite(d[j]);
  }
  for (int j = lane; j < n && finite; j += 64) {
    const double dj = d[j];
    int rank = 0;
    for (int c = 0; c < n; ++c) rank += (d[c] > dj) || (d[c] == dj && c > j);
    const bool neg = !keep_neg && dj < 0.0;
    lb[rank] = neg ? 0.0 : dj;
    for (int r = 0; r < n; ++r) Vb[r * n + rank] = neg ? 0.0 : Z[r * ld + j];
  }
  if (lane == 0) {
    int st = 0;
    if (!ok) st |= ST_EIG_NOCONV;
    if (!keep_neg && dmax < 0.0) st |= ST_EIG_NEG;
    if (!finite) st |= ST_NONFINITE;
    if (st) atomicOr(status, st);
  }
}

// One matrix per 64-lane workgroup.  Slot b starts at G + b*K*K; its size
// is K, or in the dual form the point's local-obs count when below K.
// Dynamic LDS: K*(K+1) + 2K doubles.
__global__ void __launch_bounds__(64) k_eig_fp64(int K, const double* __restrict__ G, double* __restrict__ lam,
                                                 double* __restrict__ V, int* __restrict__ status, int dual,
                                                 const int32_t* __restrict__ nloc,
                                                 const int32_t* __restrict__ act_p, int a0) {
  extern __shared__ double lds[];
  int n = K;
  if (dual) {
    const int nl = nloc[act_p[a0 + blockIdx.x]];
    n = nl < K ? nl : K;
  }
  const int ld = n + 1;
  double* Z = lds;
  double* d = Z + n * ld;
  double* e = d + n;
  const int lane = threadIdx.x;
  const double* Gb = G + (long long)blockIdx.x * K * K;
  // symmetric copy from the upper triangle (DSYEV('V','U') semantics)
  for (int q = lane; q < n * n; q += 64) {
    const int r = q / n, c = q - r * n;
    Z[r * ld + c] = r <= c ? Gb[r * n + c] : Gb[c * n + r];
  }
  wsync();
  bool ok;
  if (n == 1) {
    if (lane == 0) {
      d[0] = Z[0];
      Z[0] = 1.0;
    }
    wsync();
    ok = true;
  } else {
    ok = eig_wave(n, Z, ld, d, e);
  }
  finish_eig(n, Z, ld, d, lam + (long long)blockIdx.x * K, V + (long long)blockIdx.x * K * K, ok, status,
             dual && n < K);
}

hipError_t launch_eig_fp64(int K, int na, const double* G, double* lam, double* V, int* status, int dual,
                           const int32_t* nloc, const int32_t* act_p, int a0, hipStream_t s) {
  if (na <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * ((size_t)K * (K + 1) + 2 * (size_t)K);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_eig_fp64, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_eig_fp64, dim3(na), dim3(64), lds, s, K, G, lam, V, status, dual, nloc, act_p, a0);
  return hipGetLastError();
}


// ------------------------------------------------------------------------
// Three-phase form of the same solve, used by the analysis pipeline:
//   1. k_eig_tridiag  one wavefront per matrix: tridiag_wave in LDS, Q out;
//   2. k_eig_ql       one lane per matrix: the implicit-QL recurrence on
//                     (d, e) of 64 matrices per wavefront at once, every
//                     Givens rotation (c, s, column) appended to a log;
//   3. k_eig_apply    one wavefront per matrix: the logged rotations applied
//                     to the rows of Q in LDS, then ordering as finish_eig.
// The scalar recurrence, which is a long dependent chain, thus runs once
// per matrix instead of once per lane.  A matrix whose log would overflow
// is solved whole by eig_wave in phase 3.

constexpr int kSweepStart = 1 << 20;   // flag on the first rotation of a sweep

struct EigSlots {
  int K;  // slot pitch: K*K doubles per matrix in G/Q, K in D/E/lam
  int dual;
  const int32_t* nloc;
  const int32_t* act_p;
  int a0;
  __device__ int size(int b) const {
    if (!dual) return K;
    const int nl = nloc[act_p[a0 + b]];
    return nl < K ? nl : K;
  }
};

struct RotLog {
  double* c;
  double* s;
  int32_t* col;
  int32_t* count;  // per matrix; -1 overflow (phase 3 falls back), -2 no convergence
  long long cap;   // rotations per matrix
};

__device__ __forceinline__ void load_sym(int n, const double* Gb, double* Z, int ld) {
  for (int q = threadIdx.x; q < n * n; q += 64) {
    const int r = q / n, c = q - r * n;
    Z[r * ld + c] = r <= c ? Gb[r * n + c] : Gb[c * n + r];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(64) k_eig_tridiag(EigSlots es, const double* __restrict__ G,
                                                    double* __restrict__ Q, double* __restrict__ D,
                                                    double* __restrict__ E) {
  extern __shared__ double lds[];
  const int b = blockIdx.x, n = es.size(b), ld = n + 1, K = es.K;
  double* Z = lds;
  double* d = Z + n * ld;
  double* e = d + n;
  load_sym(n, G + (long long)b * K * K, Z, ld);
  if (n == 1) {
    if (threadIdx.x == 0) {
      d[0] = Z[0];
      e[0] = 0.0;
      Z[0] = 1.0;
    }
    __syncthreads();
  } else {
    tridiag_wave(n, Z, ld, d, e);
  }
  double* Qb = Q + (long long)b * K * K;
  for (int q = threadIdx.x; q < n * n; q += 64) {
    const int r = q / n, c = q - r * n;
    Qb[q] = Z[r * ld + c];
  }
  for (int q = threadIdx.x; q < n; q += 64) {
    D[(long long)b * K + q] = d[q];
    E[(long long)b * K + q] = e[q];
  }
}

__global__ void __launch_bounds__(256) k_eig_ql(int na, EigSlots es, double* __restrict__ D, double* __restrict__ E,
                                                RotLog lg) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= na) return;
  const int n = es.size(b);
  double* d = D + (long long)b * es.K;
  double* e = E + (long long)b * es.K;
  double* lc = lg.c + (long long)b * lg.cap;
  double* ls = lg.s + (long long)b * lg.cap;
  int32_t* li = lg.col + (long long)b * lg.cap;
  long long nr = 0;
  bool overflow = false, ok = true;
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  for (int l = 0; l < n && ok; ++l) {
    int sweeps = 0;
    for (;;) {
      int m;
      for (m = l; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++sweeps > 60) {
        ok = false;
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], bb = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * bb;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - bb;
        if (nr < lg.cap) {
          lc[nr] = c;
          ls[nr] = s;
          li[nr] = i | (i == m - 1 ? kSweepStart : 0);
        } else {
          overflow = true;
        }
        ++nr;
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  lg.count[b] = !ok ? -2 : (overflow ? -1 : (int)nr);
}

__global__ void __launch_bounds__(64) k_eig_apply(EigSlots es, const double* __restrict__ G,
                                                  const double* __restrict__ Q, const double* __restrict__ D,
                                                  const RotLog lg, double* __restrict__ lam, double* __restrict__ V,
                                                  int* __restrict__ status) {
  extern __shared__ double lds[];
  const int b = blockIdx.x, n = es.size(b), ld = n + 1, K = es.K;
  const int lane = threadIdx.x;
  double* Z = lds;
  double* d = Z + n * ld;
  double* e = d + n;
  const int cnt = lg.count[b];
  bool ok = true;
  if (cnt == -1) {
    // rotation log overflowed: solve this matrix whole
    load_sym(n, G + (long long)b * K * K, Z, ld);
    ok = eig_wave(n, Z, ld, d, e);
  } else {
    ok = cnt >= 0;
    const double* Qb = Q + (long long)b * K * K;
    for (int q = lane; q < n * n; q += 64) {
      const int r = q / n, c = q - r * n;
      Z[r * ld + c] = Qb[q];
    }
    for (int q = lane; q < n; q += 64) d[q] = D[(long long)b * K + q];
    __syncthreads();
    const double* __restrict__ lc = lg.c + (long long)b * lg.cap;
    const double* __restrict__ ls = lg.s + (long long)b * lg.cap;
    const int32_t* __restrict__ li = lg.col + (long long)b * lg.cap;
    const bool r0 = lane < n, r1 = lane + 64 < n;
    double* z0 = Z + (r0 ? lane : 0) * ld;
    double* z1 = Z + (r1 ? lane + 64 : 0) * ld;
    // Rotations arrive 64 at a time (one coalesced load per array, lane t
    // holding rotation j0+t) and are broadcast with readlane, so the chain
    // per rotation is LDS + FMA only.  Column i+1 of each row is carried in
    // registers along a sweep; the next column is read before this one is
    // stored (i-1 is never written by rotation i).
    int i = -1;                      // column of the last applied rotation
    double carry0 = 0.0, carry1 = 0.0, zp0 = 0.0, zp1 = 0.0;
    bool have_pref = false;
    for (int j0 = 0; j0 < cnt; j0 += 64) {
      const int nb = min(64, cnt - j0);
      const double cv = lane < nb ? lc[j0 + lane] : 0.0;
      const double sv = lane < nb ? ls[j0 + lane] : 0.0;
      const int iv = lane < nb ? li[j0 + lane] : 0;
      have_pref = false;
      for (int t = 0; t < nb; ++t) {
        const int code = __builtin_amdgcn_readlane(iv, t);
        const double c = __longlong_as_double(
            ((long long)__builtin_amdgcn_readlane((int)(__double_as_longlong(cv) >> 32), t) << 32) |
            (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(cv), t));
        const double s = __longlong_as_double(
            ((long long)__builtin_amdgcn_readlane((int)(__double_as_longlong(sv) >> 32), t) << 32) |
            (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(sv), t));
        const int ic = code & (kSweepStart - 1);
        if (code & kSweepStart) {
          if (i >= 0) {              // close the previous sweep
            if (r0) z0[i] = carry0;
            if (r1) z1[i] = carry1;
          }
          carry0 = r0 ? z0[ic + 1] : 0.0;
          carry1 = r1 ? z1[ic + 1] : 0.0;
        }
        i = ic;
        double zi0, zi1;
        if (have_pref && !(code & kSweepStart)) {
          zi0 = zp0;
          zi1 = zp1;
        } else {
          zi0 = r0 ? z0[i] : 0.0;
          zi1 = r1 ? z1[i] : 0.0;
        }
        if (r0) z0[i + 1] = s * zi0 + c * carry0;
        if (r1) z1[i + 1] = s * zi1 + c * carry1;
        carry0 = c * zi0 - s * carry0;
        carry1 = c * zi1 - s * carry1;
        // read the next rotation's column now when it continues this sweep
        // (column i-1, untouched by rotation i)
        have_pref = false;
        if (t + 1 < nb) {
          const int code2 = __builtin_amdgcn_readlane(iv, t + 1);
          if (!(code2 & kSweepStart)) {
            const int i2 = code2 & (kSweepStart - 1);
            zp0 = r0 ? z0[i2] : 0.0;
            zp1 = r1 ? z1[i2] : 0.0;
            have_pref = true;
          }
        }
      }
    }
    if (i >= 0) {
      if (r0) z0[i] = carry0;
      if (r1) z1[i] = carry1;
    }
    __syncthreads();
  }
  finish_eig(n, Z, ld, d, lam + (long long)b * K, V + (long long)b * K * K, ok, status, es.dual && n < es.K);
}

// Rotations logged per matrix (capacity); a typical implicit-QL solve needs
// about 1.3 n^2, the rest is the fallback's business.
long long eig_log_capacity(int K) { return 2LL * K * K + 64; }

hipError_t launch_eig3(int K, int na, const double* G, double* Q, double* D, double* E, double* lc, double* ls,
                       int32_t* li, int32_t* cnt, double* lam, double* V, int* status, int dual,
                       const int32_t* nloc, const int32_t* act_p, int a0, hipStream_t st) {
  if (na <= 0) return hipSuccess;
  EigSlots es{K, dual, nloc, act_p, a0};
  RotLog lg{lc, ls, li, cnt, eig_log_capacity(K)};
  const size_t lds = sizeof(double) * ((size_t)K * (K + 1) + 2 * (size_t)K);
  if (lds > 65536) {
    hipError_t e1 = hipFuncSetAttribute((const void*)k_eig_tridiag, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds);
    hipError_t e2 = hipFuncSetAttribute((const void*)k_eig_apply, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds);
    if (e1 != hipSuccess) return e1;
    if (e2 != hipSuccess) return e2;
  }
  hipLaunchKernelGGL(k_eig_tridiag, dim3(na), dim3(64), lds, st, es, G, Q, D, E);
  hipLaunchKernelGGL(k_eig_ql, dim3((na + 255) / 256), dim3(256), 0, st, na, es, D, E, lg);
  hipLaunchKernelGGL(k_eig_apply, dim3(na), dim3(64), lds, st, es, G, Q, D, lg, lam, V, status);
  return hipGetLastError();
}

#include "eigreg.inc"
#include "adjoint.inc"

}  // namespace letkf
