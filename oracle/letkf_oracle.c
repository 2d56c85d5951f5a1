/*
 * letkf_oracle.c -- CPU restatement of the sbPOM-LETKF local analysis
 * (letkf/run/src/common_da.f90, common_letkf.f90, common.f90,
 * common_obs.f90).  TEST INFRASTRUCTURE ONLY: the checker for the HIP
 * engine, never linked by it.  See letkf_oracle.h.
 *
 * The symmetric eigensolver stands in for LAPACK DSYEV (which the reference
 * links from Fujitsu SSL2, no pinned version).  It is Householder reduction
 * to tridiagonal form with accumulated transformations followed by the
 * implicit-shift QL iteration (Wilkinson & Reinsch, Handbook for Automatic
 * Computation II, 1971, procedures tred2/tql2), written here for row-major
 * storage.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "letkf_oracle.h"

/* Householder tridiagonalisation of the symmetric n x n row-major matrix a.
 * On return d = diagonal, e[1..n-1] = sub-diagonal (e[0] = 0) and a holds
 * the orthogonal Q with A = Q T Q^T. */
static void oracle_tridiagonalize(int n, double* a, double* d, double* e) {
#define A_(r, c) a[(size_t)(r) * n + (c)]
  for (int i = n - 1; i >= 1; --i) {
    const int l = i - 1;
    double h = 0.0;
    if (l == 0) {
      e[i] = A_(i, 0);
      d[i] = 0.0;
      continue;
    }
    double scale = 0.0;
    for (int c = 0; c <= l; ++c) scale += fabs(A_(i, c));
    if (scale == 0.0) {
      e[i] = A_(i, l);
      d[i] = 0.0;
      continue;
    }
    for (int c = 0; c <= l; ++c) {
      A_(i, c) /= scale;
      h += A_(i, c) * A_(i, c);
    }
    double f = A_(i, l);
    double g = f >= 0.0 ? -sqrt(h) : sqrt(h);
    e[i] = scale * g;
    h -= f * g;
    A_(i, l) = f - g;
    /* p = A u / h (stored in e[0..l]); u is row i */
    double upu = 0.0;
    for (int r = 0; r <= l; ++r) {
      A_(r, i) = A_(i, r) / h;
      double acc = 0.0;
      for (int c = 0; c <= r; ++c) acc += A_(r, c) * A_(i, c);
      for (int c = r + 1; c <= l; ++c) acc += A_(c, r) * A_(i, c);
      e[r] = acc / h;
      upu += e[r] * A_(i, r);
    }
    const double kk = upu / (h + h);
    /* q = p - kk u ; A <- A - u q^T - q u^T on the lower triangle */
    for (int r = 0; r <= l; ++r) {
      const double ur = A_(i, r);
      const double qr = e[r] - kk * ur;
      e[r] = qr;
      for (int c = 0; c <= r; ++c) A_(r, c) -= ur * e[c] + qr * A_(i, c);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  /* accumulate Q */
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int c = 0; c <= l; ++c) {
        double g = 0.0;
        for (int r = 0; r <= l; ++r) g += A_(i, r) * A_(r, c);
        for (int r = 0; r <= l; ++r) A_(r, c) -= g * A_(r, i);
      }
    }
    d[i] = A_(i, i);
    A_(i, i) = 1.0;
    for (int c = 0; c <= l; ++c) {
      A_(c, i) = 0.0;
      A_(i, c) = 0.0;
    }
  }
#undef A_
}

/* Implicit-shift QL on the tridiagonal (d, e) with Wilkinson-type shift,
 * rotations accumulated into the row-major z (columns = eigenvectors).
 * Returns 0, or -1 if an eigenvalue needs more than 60 sweeps. */
static int oracle_tridiag_ql(int n, double* d, double* e, double* z) {
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int sweeps = 0;
    for (;;) {
      int m;
      for (m = l; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++sweeps > 60) return -1;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int deflated = 0;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = 1;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          double* zr = z + (size_t)k * n;
          const double zf = zr[i + 1];
          zr[i + 1] = s * zr[i] + c * zf;
          zr[i] = c * zr[i] - s * zf;
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return 0;
}

#define REAL double
#define SUF _f64
#define ROUND lround
#define RSQRT sqrt
#define RACOS acos
#include "letkf_oracle_body.inc"
#undef REAL
#undef SUF
#undef ROUND
#undef RSQRT
#undef RACOS

#define REAL float
#define SUF _f32
#define ROUND lroundf
#define RSQRT sqrtf
#define RACOS acosf
#include "letkf_oracle_body.inc"
#undef REAL
#undef SUF
#undef ROUND
#undef RSQRT
#undef RACOS
