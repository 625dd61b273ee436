/* TEST INFRASTRUCTURE ONLY -- CPU oracle (FP64) for the batched complex
 * eigendecomposition zgehrd -> zunghr -> zlahqr -> ztrevc.
 *
 * A restatement, in C, of the reference algorithm as the reference computes
 * it (each function cites the reference lines it follows).  Floating-point
 * operations are kept in the reference's order so that, with FMA contraction
 * off on both sides, this file reproduces the reference bit for bit; the
 * pinning tests check exactly that against oracle/_ref/libref_exact.so and
 * against closed-form invariants (Q^H Q = I, A = Q H Q^H, A = Z T Z^H,
 * A V = V diag(E)) and numpy's LAPACK.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library, and only as the checker.  Never shipped, never measured
 * as the product.
 *
 * Indices inside the per-matrix routines are 1-based (Fortran style) like the
 * reference's, over a matrix slot whose consecutive elements are T = thread_cnt
 * doubles apart.
 */
#include "eig_oracle.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double *re, *im; /* slot pointer: element (1,1) of this matrix */
  int n;           /* order (leading dimension) */
  size_t T;        /* element stride (thread_cnt) */
} mat_t;

#define IX(M, r, c) ((((size_t)(r)-1) + ((size_t)(c)-1) * (size_t)(M).n) * (M).T)
#define VX(T_, r) (((size_t)(r)-1) * (T_))

/* ---- scalar complex arithmetic, ComplexDouble.hpp (default, non-MKL) ---- */

/* cinv, ComplexDouble.hpp:283-317 (Smith-style reciprocal). */
static inline void c_inv(double *r, double *i) {
  double rr = *r, ri = *i, ratio, den, cr, ci;
  double ar = fabs(rr), ai = fabs(ri);
  if (ar <= ai) { cr = rr; ci = ri; } else { ci = rr; cr = ri; }
  ratio = cr / ci;
  den = ci * (1.0 + ratio * ratio);
  cr = ratio / den;
  ci = 1.0 / den;
  if (ar <= ai) { *r = cr; *i = -ci; } else { *r = ci; *i = -cr; }
}

/* cdiv (complex divisor), ComplexDouble.hpp:319-353 (Smith's method). */
static inline void c_div(double *r, double *i, double rr, double ri) {
  double ratio, den, cr, ci, ar = fabs(rr), ai = fabs(ri);
  if (ar <= ai) { cr = rr; ci = ri; } else { ci = rr; cr = ri; }
  ratio = cr / ci;
  den = ci * (1.0 + ratio * ratio);
  if (ar <= ai) {
    cr = *r * ratio + *i;
    ci = *i * ratio - *r;
  } else {
    cr = *r + *i * ratio;
    ci = *i - *r * ratio;
  }
  *r = cr / den;
  *i = ci / den;
}

/* cmul, ComplexDouble.hpp:367-372: (r,i) *= (rr,ri). */
static inline void c_mul(double *r, double *i, double rr, double ri) {
  double lr = *r, li = *i;
  *r = lr * rr - li * ri;
  *i = li * rr + lr * ri;
}

/* cmulc, ComplexDouble.hpp:374-379: (r,i) *= conj(rr,ri). */
static inline void c_mulc(double *r, double *i, double rr, double ri) {
  double lr = *r, li = *i;
  *r = lr * rr + li * ri;
  *i = li * rr - lr * ri;
}

static inline double cabs1_(double r, double i) { return fabs(r) + fabs(i); } /* :432 */
static inline double cabs_(double r, double i) { return sqrt(r * r + i * i); } /* :446 */

/* csqrt (principal branch), ComplexDouble.hpp:451-488. */
static inline void c_sqrt(double *r, double *i) {
  double lr = *r, li = *i, zr, zi, cr, ci;
  zr = sqrt(0.5 * (cabs_(lr, li) + fabs(lr)));
  zi = (zr == 0.0) ? 0.0 : fabs(li) / (2.0 * zr);
  if (lr < 0.0) { cr = zi; ci = zr; } else { cr = zr; ci = zi; }
  if (li < 0.0) ci = -ci;
  *r = cr;
  *i = ci;
}

/* ---- zgemv_simple, Toolbox.hpp:1033-1131 ----
 * y := beta*y + A*x (hermitian=0) or beta*y + A^H*x (hermitian=1), A an m x n
 * block of a matrix with leading dimension lda, elements T apart.  beta*y is
 * formed by multiplication even for beta == 0 (Toolbox.hpp:1067-1080).  The
 * Hermitian branch adds to the REAL part of y for the imaginary part
 * (Toolbox.hpp:1125) unless ORACLE_ZGEMV_FIX. */
static void zgemv_ref(int hermitian, int m, int n, const double *rA, const double *iA, int lda,
                      size_t T, const double *rx, const double *ix, double rbeta, double ibeta,
                      double *ry, double *iy, int flags) {
  const int lenx = hermitian ? m : n;
  const int leny = hermitian ? n : m;
  (void)lenx;
  for (int i = 1; i <= leny; ++i) {
    double vr = ry[VX(T, i)], vi = iy[VX(T, i)];
    c_mul(&vr, &vi, rbeta, ibeta);
    ry[VX(T, i)] = vr;
    iy[VX(T, i)] = vi;
  }
  if (!hermitian) {
    for (int j = 1; j <= n; ++j) {
      const double tr = rx[VX(T, j)], ti = ix[VX(T, j)];
      for (int i = 1; i <= m; ++i) {
        const size_t a = (((size_t)i - 1) + ((size_t)j - 1) * lda) * T;
        double mr = rA[a], mi = iA[a];
        c_mul(&mr, &mi, tr, ti);
        ry[VX(T, i)] = ry[VX(T, i)] + mr;
        iy[VX(T, i)] = iy[VX(T, i)] + mi;
      }
    }
  } else {
    for (int j = 1; j <= n; ++j) {
      double tr = 0.0, ti = 0.0;
      for (int i = 1; i <= m; ++i) {
        const size_t a = (((size_t)i - 1) + ((size_t)j - 1) * lda) * T;
        double vr = rx[VX(T, i)], vi = ix[VX(T, i)];
        c_mulc(&vr, &vi, rA[a], iA[a]);
        tr = tr + vr;
        ti = ti + vi;
      }
      const double yr = ry[VX(T, j)];
      const double yi = (flags & ORACLE_ZGEMV_FIX) ? iy[VX(T, j)] : yr;
      ry[VX(T, j)] = yr + tr;
      iy[VX(T, j)] = yi + ti;
    }
  }
}

/* ---- zgehrd, Toolbox.hpp:1315-1490 (unblocked zgehd2 with inline zlarfg) ----
 * work: n complex entries, stride T. */
static void zgehrd_one(mat_t A, double *rtau, double *itau, double *rwork, double *iwork,
                       int flags) {
  const int n = A.n;
  const size_t T = A.T;
  for (int i = 1; i <= n - 1; ++i) {
    double ar = A.re[IX(A, i + 1, i)], ai = A.im[IX(A, i + 1, i)];
    const int len = n - i; /* length of (alpha, x) */
    /* ||x||^2 over rows i+2..n, Toolbox.hpp:1354-1363 */
    double norm = 0.0;
    for (int r = i + 2; r <= n; ++r) {
      const double xr = A.re[IX(A, r, i)], xi = A.im[IX(A, r, i)];
      norm += xr * xr + xi * xi;
    }
    double tr, ti, beta;
    if ((flags & ORACLE_LAPACK_SIGN) && norm == 0.0 && ai == 0.0) {
      /* LAPACK zlarfg: H = I, tau = 0 */
      tr = 0.0; ti = 0.0; beta = ar;
    } else {
      double temp = sqrt(ar * ar + ai * ai + norm);
      beta = copysign(temp, ar); /* Toolbox.hpp:1368: same sign as Re alpha */
      if (flags & ORACLE_LAPACK_SIGN) beta = -beta;
      temp = 1.0 / beta;
      tr = (beta - ar) * temp;
      ti = -ai * temp;
      double sr = ar - beta, si = ai;
      c_inv(&sr, &si);
      for (int r = i + 2; r <= n; ++r) {
        double xr = A.re[IX(A, r, i)], xi = A.im[IX(A, r, i)];
        c_mul(&xr, &xi, sr, si);
        A.re[IX(A, r, i)] = xr;
        A.im[IX(A, r, i)] = xi;
      }
    }
    (void)len;
    rtau[VX(T, i)] = tr;
    itau[VX(T, i)] = ti;
    A.re[IX(A, i + 1, i)] = 1.0;
    A.im[IX(A, i + 1, i)] = 0.0;

    /* H from the right on A(1:n, i+1:n): w = A*v, A -= tau*w*v^H
     * (Toolbox.hpp:1405-1441) */
    {
      const int lastv = n - i, lastc = n;
      double *rC = A.re + IX(A, 1, i + 1), *iC = A.im + IX(A, 1, i + 1);
      const double *rv = A.re + IX(A, i + 1, i), *iv = A.im + IX(A, i + 1, i);
      zgemv_ref(0, lastc, lastv, rC, iC, n, T, rv, iv, 0.0, 0.0, rwork, iwork, flags);
      for (int j = 1; j <= lastv; ++j) {
        double cr = -rtau[VX(T, i)], ci = -itau[VX(T, i)];
        c_mulc(&cr, &ci, rv[VX(T, j)], iv[VX(T, j)]);
        for (int k = 1; k <= lastc; ++k) {
          const size_t a = (((size_t)k - 1) + ((size_t)j - 1) * n) * T;
          double wr = rwork[VX(T, k)], wi = iwork[VX(T, k)];
          c_mul(&wr, &wi, cr, ci);
          rC[a] = wr + rC[a];
          iC[a] = wi + iC[a];
        }
      }
    }
    /* H^H from the left on A(i+1:n, i+1:n): w = A^H v, A -= conj(tau)*v*w^H
     * (Toolbox.hpp:1448-1481) */
    {
      const int lastv = n - i, lastc = n - i;
      double *rC = A.re + IX(A, i + 1, i + 1), *iC = A.im + IX(A, i + 1, i + 1);
      const double *rv = A.re + IX(A, i + 1, i), *iv = A.im + IX(A, i + 1, i);
      zgemv_ref(1, lastv, lastc, rC, iC, n, T, rv, iv, 0.0, 0.0, rwork, iwork, flags);
      for (int j = 1; j <= lastc; ++j) {
        double cr = -rtau[VX(T, i)], ci = itau[VX(T, i)];
        c_mulc(&cr, &ci, rwork[VX(T, j)], iwork[VX(T, j)]);
        for (int k = 1; k <= lastv; ++k) {
          const size_t a = (((size_t)k - 1) + ((size_t)j - 1) * n) * T;
          double wr = rv[VX(T, k)], wi = iv[VX(T, k)];
          c_mul(&wr, &wi, cr, ci);
          rC[a] = wr + rC[a];
          iC[a] = wi + iC[a];
        }
      }
    }
    A.re[IX(A, i + 1, i)] = beta;
    A.im[IX(A, i + 1, i)] = 0.0;
  }
}

/* ---- zunghr, Toolbox.hpp:1492-1659 (shift reflectors, then zung2r) ---- */
static void zunghr_one(mat_t A, const double *rtau, const double *itau, double *rwork,
                       double *iwork, int flags) {
  const int n = A.n;
  const size_t T = A.T;
  /* shift the reflector vectors one column to the right (1524-1544) */
  for (int j = n; j >= 2; --j) {
    for (int i = 1; i <= j - 1; ++i) {
      A.re[IX(A, i, j)] = 0.0;
      A.im[IX(A, i, j)] = 0.0;
    }
    for (int i = j + 1; i <= n; ++i) {
      A.re[IX(A, i, j)] = A.re[IX(A, i, j - 1)];
      A.im[IX(A, i, j)] = A.im[IX(A, i, j - 1)];
    }
  }
  /* first row/column of the identity (1546-1558) */
  for (int i = 1; i <= n; ++i) {
    A.re[IX(A, i, 1)] = 0.0;
    A.im[IX(A, i, 1)] = 0.0;
  }
  A.re[IX(A, 1, 1)] = 1.0;
  A.im[IX(A, 1, 1)] = 0.0;
  if (n < 2) return;
  /* zung2r on the trailing (n-1) x (n-1) block (1563-1654) */
  mat_t S = A;
  S.re = A.re + IX(A, 2, 2);
  S.im = A.im + IX(A, 2, 2);
  const int nh = n - 1;
  for (int i = nh; i >= 1; --i) {
    S.re[IX(S, i, i)] = 1.0;
    S.im[IX(S, i, i)] = 0.0;
    const int lastv = nh - i + 1, lastc = nh - i;
    double *rC = S.re + IX(S, i, i + 1), *iC = S.im + IX(S, i, i + 1);
    const double *rv = S.re + IX(S, i, i), *iv = S.im + IX(S, i, i);
    zgemv_ref(1, lastv, lastc, rC, iC, n, T, rv, iv, 0.0, 0.0, rwork, iwork, flags);
    for (int j = 1; j <= lastc; ++j) {
      double cr = -rtau[VX(T, i)], ci = -itau[VX(T, i)];
      c_mulc(&cr, &ci, rwork[VX(T, j)], iwork[VX(T, j)]);
      for (int k = 1; k <= lastv; ++k) {
        const size_t a = (((size_t)k - 1) + ((size_t)j - 1) * n) * T;
        double wr = rv[VX(T, k)], wi = iv[VX(T, k)];
        c_mul(&wr, &wi, cr, ci);
        rC[a] = wr + rC[a];
        iC[a] = wi + iC[a];
      }
    }
    {
      const double sr = -rtau[VX(T, i)], si = -itau[VX(T, i)];
      for (int r = i + 1; r <= nh; ++r) {
        double xr = S.re[IX(S, r, i)], xi = S.im[IX(S, r, i)];
        c_mul(&xr, &xi, sr, si);
        S.re[IX(S, r, i)] = xr;
        S.im[IX(S, r, i)] = xi;
      }
    }
    S.re[IX(S, i, i)] = 1.0 - rtau[VX(T, i)];
    S.im[IX(S, i, i)] = 0.0 - itau[VX(T, i)];
    for (int r = 1; r <= i - 1; ++r) {
      S.re[IX(S, r, i)] = 0.0;
      S.im[IX(S, r, i)] = 0.0;
    }
  }
}

/* ---- one single-shift QR sweep, Toolbox.hpp:138-369 (wantT = wantZ = 1) ----
 * Bulge chase over K = M..I-1 with 2-element reflectors (zlarfg, LAPACK sign),
 * rows/cols updated over [I1, I2]; then H(I,I-1) made real (scale_rc 14-29). */
static void qr_sweep(mat_t H, mat_t Z, int I, int M, int I1, int I2, double v0r, double v0i,
                     double v1r, double v1i) {
  const int n = H.n;
  for (int K = M; K <= I - 1; ++K) {
    if (K > M) {
      v0r = H.re[IX(H, K, K - 1)];
      v0i = H.im[IX(H, K, K - 1)];
      v1r = H.re[IX(H, K + 1, K - 1)];
      v1i = H.im[IX(H, K + 1, K - 1)];
    }
    double t1r, t1i;
    if (v1r == 0.0 && v1i == 0.0 && v0i == 0.0) {
      t1r = 0.0;
      t1i = 0.0;
    } else {
      const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
      double norm = sqrt(v0r * v0r + v0i * v0i + v1r * v1r + v1i * v1i);
      int knt = 0;
      if (norm < safmin) {
        const double rsafmn = 1.0 / safmin;
        do {
          ++knt;
          v1r *= rsafmn; v1i *= rsafmn;
          v0r *= rsafmn; v0i *= rsafmn;
          norm *= rsafmn;
        } while (norm < safmin);
        norm = sqrt(v0r * v0r + v0i * v0i + v1r * v1r + v1i * v1i);
      }
      double beta = copysign(norm, v0r);
      double zr = v0r + beta, zi = v0i;
      t1r = zr;
      t1i = zi;
      c_inv(&zr, &zi);
      c_mul(&v1r, &v1i, zr, zi);
      t1r /= beta;
      t1i /= beta;
      for (int k = 0; k < knt; ++k) beta *= safmin;
      v0r = -beta;
      v0i = 0.0;
    }
    if (K > M) {
      H.re[IX(H, K, K - 1)] = v0r;
      H.im[IX(H, K, K - 1)] = v0i;
      H.re[IX(H, K + 1, K - 1)] = 0.0;
      H.im[IX(H, K + 1, K - 1)] = 0.0;
    }
    const double t2 = t1r * v1r - t1i * v1i;
    /* rows K, K+1, columns K..I2 (loop 80) */
    for (int j = K; j <= I2; ++j) {
      double h1r = H.re[IX(H, K, j)], h1i = H.im[IX(H, K, j)];
      double h2r = H.re[IX(H, K + 1, j)], h2i = H.im[IX(H, K + 1, j)];
      double zr = h1r, zi = h1i;
      c_mulc(&zr, &zi, t1r, t1i);
      zr = zr + h2r * t2;
      zi = zi + h2i * t2;
      h1r -= zr; h1i -= zi;
      c_mul(&zr, &zi, v1r, v1i);
      h2r -= zr; h2i -= zi;
      H.re[IX(H, K, j)] = h1r; H.im[IX(H, K, j)] = h1i;
      H.re[IX(H, K + 1, j)] = h2r; H.im[IX(H, K + 1, j)] = h2i;
    }
    /* columns K, K+1, rows I1..min(K+2, I) (loop 90) */
    const int i3 = (K + 2) < I ? (K + 2) : I;
    for (int j = I1; j <= i3; ++j) {
      double h1r = H.re[IX(H, j, K)], h1i = H.im[IX(H, j, K)];
      double h2r = H.re[IX(H, j, K + 1)], h2i = H.im[IX(H, j, K + 1)];
      double zr = h1r, zi = h1i;
      c_mul(&zr, &zi, t1r, t1i);
      zr = zr + h2r * t2;
      zi = zi + h2i * t2;
      h1r -= zr; h1i -= zi;
      c_mulc(&zr, &zi, v1r, v1i);
      h2r -= zr; h2i -= zi;
      H.re[IX(H, j, K)] = h1r; H.im[IX(H, j, K)] = h1i;
      H.re[IX(H, j, K + 1)] = h2r; H.im[IX(H, j, K + 1)] = h2i;
    }
    /* Schur vectors, columns K, K+1, all rows (loop 100; absent in the
     * eigenvalues-only single_shift_qr, Toolbox.hpp:372-522) */
    for (int j = 1; Z.re && j <= n; ++j) {
      double z1r = Z.re[IX(Z, j, K)], z1i = Z.im[IX(Z, j, K)];
      double z2r = Z.re[IX(Z, j, K + 1)], z2i = Z.im[IX(Z, j, K + 1)];
      double zr = z1r, zi = z1i;
      c_mul(&zr, &zi, t1r, t1i);
      zr = zr + t2 * z2r;
      zi = zi + t2 * z2i;
      z1r -= zr; z1i -= zi;
      c_mulc(&zr, &zi, v1r, v1i);
      z2r -= zr; z2i -= zi;
      Z.re[IX(Z, j, K)] = z1r; Z.im[IX(Z, j, K)] = z1i;
      Z.re[IX(Z, j, K + 1)] = z2r; Z.im[IX(Z, j, K + 1)] = z2i;
    }
  }
  /* make H(I, I-1) real (337-367) */
  const double hi = H.im[IX(H, I, I - 1)];
  if (hi != 0.0) {
    double zr = H.re[IX(H, I, I - 1)], zi = hi;
    const double d = cabs_(zr, zi);
    zr /= d;
    zi /= d;
    zi = -zi;
    H.re[IX(H, I, I - 1)] = d;
    H.im[IX(H, I, I - 1)] = 0.0;
    for (int j = I + 1; j <= I2; ++j) c_mul(&H.re[IX(H, I, j)], &H.im[IX(H, I, j)], zr, zi);
    for (int j = I1; j <= I - 1; ++j) c_mulc(&H.re[IX(H, j, I)], &H.im[IX(H, j, I)], zr, zi);
    for (int j = 1; Z.re && j <= n; ++j) c_mulc(&Z.re[IX(Z, j, I)], &Z.im[IX(Z, j, I)], zr, zi);
  }
}

/* ---- zlahqr_schur(wantT=1, wantZ=1), Toolbox.hpp:690-1031; with
 * ORACLE_EIG_ONLY the eigenvalues-only zlahqr_ (Toolbox.hpp:1663-1890): no
 * loop 20, no Z (Z.re == NULL), and each sweep updates H only inside the
 * active block (I1 = L, I2 = I, 1761-1762; single_shift_qr 372-522). ---- */
static void zlahqr_one(mat_t H, mat_t Z, double *rE, double *iE, int flags, int32_t *info,
                       int32_t *sweeps, int32_t *steps) {
  const int n = H.n;
  const size_t T = H.T;
  const int ILO = 1, IHI = n;
  const double safemin = DBL_MIN, ulp = DBL_EPSILON;
  const double smlnum = safemin * (((double)(IHI - ILO + 1)) / ulp);
  const int eig_only = (flags & ORACLE_EIG_ONLY) != 0;
  int nfail = 0, nsweep = 0, nstep = 0;

  /* loop 20: make the subdiagonal real (735-793) */
  for (int i = ILO + 1; !eig_only && i <= IHI; ++i) {
    const double v0i = H.im[IX(H, i, i - 1)];
    if (v0i != 0.0) {
      const double v0r = H.re[IX(H, i, i - 1)];
      double d = cabs1_(v0r, v0i);
      double wr = v0r / d, wi = v0i / d;
      d = cabs_(wr, wi);
      wi = -wi;
      wr /= d;
      wi /= d;
      d = cabs_(v0r, v0i);
      H.re[IX(H, i, i - 1)] = d;
      H.im[IX(H, i, i - 1)] = 0.0;
      for (int j = i; j <= n; ++j) c_mul(&H.re[IX(H, i, j)], &H.im[IX(H, i, j)], wr, wi);
      const int hi = n < i + 1 ? n : i + 1;
      for (int j = 1; j <= hi; ++j) c_mulc(&H.re[IX(H, j, i)], &H.im[IX(H, j, i)], wr, wi);
      for (int j = 1; j <= n; ++j) c_mulc(&Z.re[IX(Z, j, i)], &Z.im[IX(Z, j, i)], wr, wi);
    }
  }

  int I1 = 1, I2 = n;
  for (int I = IHI; I >= ILO; --I) {
    int L = ILO, K = I, converged = 0;
    for (int its = 0; its <= 30; ++its) {
      /* look for a single small subdiagonal element (loop 40, 810-867) */
      for (K = I; K >= L + 1; --K) {
        const double wr4 = H.re[IX(H, K, K - 1)], wi4 = H.im[IX(H, K, K - 1)];
        if (cabs1_(wr4, wi4) <= smlnum) break;
        const double wr1 = H.re[IX(H, K - 1, K - 1)], wi1 = H.im[IX(H, K - 1, K - 1)];
        const double wr2 = H.re[IX(H, K, K)], wi2 = H.im[IX(H, K, K)];
        double tst = cabs1_(wr1, wi1) + cabs1_(wr2, wi2);
        if (tst == 0.0) {
          if (K - 2 >= ILO) tst += fabs(H.re[IX(H, K - 1, K - 2)]);
          if (K + 1 <= IHI)
            tst += fabs((flags & ORACLE_LAPACK_TST) ? H.re[IX(H, K + 1, K)] : H.im[IX(H, K + 1, K)]);
        }
        if (fabs(wr4) <= ulp * tst) {
          const double wr5 = H.re[IX(H, K - 1, K)], wi5 = H.im[IX(H, K - 1, K)];
          double d2 = cabs1_(wr4, wi4), d3 = cabs1_(wr5, wi5);
          const double ab = fmax(d2, d3), ba = fmin(d2, d3);
          const double zr = wr1 - wr2, zi = wi1 - wi2;
          d2 = cabs1_(wr2, wi2);
          d3 = cabs1_(zr, zi);
          const double aa = fmax(d2, d3), bb = fmin(d2, d3);
          const double s = aa + ab;
          if (ba * (ab / s) <= fmax(smlnum, ulp * (bb * (aa / s)))) break;
        }
      }
      L = K;
      if (L > ILO) {
        H.re[IX(H, L, L - 1)] = 0.0;
        H.im[IX(H, L, L - 1)] = 0.0;
      }
      if (L >= I) {
        converged = 1;
        break;
      }
      if (eig_only) {
        I1 = L;
        I2 = I;
      }
      /* shift (896-990) */
      double tr, ti;
      if (its == 10) {
        const double d = fabs(H.re[IX(H, L + 1, L)]) * 0.75;
        tr = d + H.re[IX(H, L, L)];
        ti = H.im[IX(H, L, L)];
      } else if (its == 20) {
        const double d = fabs(H.re[IX(H, I, I - 1)]) * 0.75;
        tr = d + H.re[IX(H, I, I)];
        ti = H.im[IX(H, I, I)];
      } else {
        tr = H.re[IX(H, I, I)];
        ti = H.im[IX(H, I, I)];
        double h2r = H.re[IX(H, I - 1, I)], h2i = H.im[IX(H, I - 1, I)];
        double h3r = H.re[IX(H, I, I - 1)], h3i = H.im[IX(H, I, I - 1)];
        c_sqrt(&h2r, &h2i);
        c_sqrt(&h3r, &h3i);
        c_mul(&h2r, &h2i, h3r, h3i);
        double ur = h2r, ui = h2i;
        double s = cabs1_(ur, ui);
        if (s != 0.0) {
          double xr = H.re[IX(H, I - 1, I - 1)], xi = H.im[IX(H, I - 1, I - 1)];
          xr -= tr;
          xi -= ti;
          xr *= 0.5;
          xi *= 0.5;
          const double sx = cabs1_(xr, xi);
          s = fmax(s, sx);
          double z1r = xr / s, z1i = xi / s;
          double z2r = ur / s, z2i = ui / s;
          c_mul(&z1r, &z1i, z1r, z1i);
          c_mul(&z2r, &z2i, z2r, z2i);
          z1r += z2r;
          z1i += z2i;
          c_sqrt(&z1r, &z1i);
          z1r *= s;
          z1i *= s;
          double yr = z1r, yi = z1i;
          if (sx > 0.0) {
            const double zr = xr / sx, zi = xi / sx;
            if (zr * yr + zi * yi < 0.0) {
              yr = -yr;
              yi = -yi;
            }
          }
          const double vr = ur, vi = ui;
          xr += yr;
          xi += yi;
          c_div(&ur, &ui, xr, xi);
          c_mul(&ur, &ui, vr, vi);
          tr -= ur;
          ti -= ui;
        }
      }
      /* first column of the shifted matrix, M = L (loop 60 omitted, 997-1015) */
      const int M = L;
      double zr = H.re[IX(H, L, L)] - tr, zi = H.im[IX(H, L, L)] - ti;
      double d1r = H.re[IX(H, L + 1, L)];
      const double s = cabs1_(zr, zi) + fabs(d1r);
      zr /= s;
      zi /= s;
      d1r /= s;
      qr_sweep(H, Z, I, M, I1, I2, zr, zi, d1r, 0.0);
      ++nsweep;
      nstep += I - M;
    }
    if (!converged) ++nfail;
  }
  for (int I = IHI; I >= ILO; --I) {
    rE[VX(T, I)] = H.re[IX(H, I, I)];
    iE[VX(T, I)] = H.im[IX(H, I, I)];
  }
  if (info) *info = nfail;
  if (sweeps) *sweeps = nsweep;
  if (steps) *steps = nstep;
}

/* ---- backsubstitution, Toolbox.hpp:31-136 (live loop 104-135) ----
 * Solves the leading m x m upper-triangular system in place in x. */
static void backsub_one(int m, const double *rA, const double *iA, int lda, size_t T, double *rx,
                        double *ix) {
  for (int j = m; j >= 1; --j) {
    const size_t a = (((size_t)j - 1) + ((size_t)j - 1) * lda) * T;
    double tr = rx[VX(T, j)], ti = ix[VX(T, j)];
    c_div(&tr, &ti, rA[a], iA[a]);
    rx[VX(T, j)] = tr;
    ix[VX(T, j)] = ti;
    for (int i = j - 1; i >= 1; --i) {
      const size_t b = (((size_t)i - 1) + ((size_t)j - 1) * lda) * T;
      double ar = rA[b], ai = iA[b];
      c_mul(&ar, &ai, tr, ti);
      rx[VX(T, i)] = rx[VX(T, i)] - ar;
      ix[VX(T, i)] = ix[VX(T, i)] - ai;
    }
  }
}

/* ---- ztrevc (right eigenvectors, back-transformed), Toolbox.hpp:1133-1313 ----
 * work: 2n complex entries, stride T (diagonal copy, then right-hand side). */
static void ztrevc_one(mat_t Tm, mat_t R, double *rwork, double *iwork, int flags) {
  const int n = Tm.n;
  const size_t T = Tm.T;
  const double unfl = DBL_MIN, ulp = DBL_EPSILON;
  const double smlnum = unfl * ((double)n / ulp);
  double *rdiag = rwork, *idiag = iwork;
  double *rrhs = rwork + (size_t)n * T, *irhs = iwork + (size_t)n * T;
  for (int i = 1; i <= n; ++i) {
    rdiag[VX(T, i)] = Tm.re[IX(Tm, i, i)];
    idiag[VX(T, i)] = Tm.im[IX(Tm, i, i)];
  }
  for (int ki = n; ki >= 1; --ki) {
    const double lr = Tm.re[IX(Tm, ki, ki)], li = Tm.im[IX(Tm, ki, ki)];
    double smin = cabs1_(lr, li);
    smin *= ulp;
    smin = fmax(smin, smlnum);
    for (int k = 1; k <= ki - 1; ++k) { /* loop 40 */
      rrhs[VX(T, k)] = -Tm.re[IX(Tm, k, ki)];
      irhs[VX(T, k)] = -Tm.im[IX(Tm, k, ki)];
    }
    for (int k = 1; k <= ki - 1; ++k) { /* loop 50 */
      double tr = Tm.re[IX(Tm, k, k)] - lr, ti = Tm.im[IX(Tm, k, k)] - li;
      if (cabs1_(tr, ti) < smin) {
        tr = smin;
        ti = 0.0;
      }
      Tm.re[IX(Tm, k, k)] = tr;
      Tm.im[IX(Tm, k, k)] = ti;
    }
    if (ki > 1) {
      backsub_one(ki - 1, Tm.re, Tm.im, n, T, rrhs, irhs);
      zgemv_ref(0, n, ki - 1, R.re, R.im, n, T, rrhs, irhs, 1.0, 0.0, R.re + IX(R, 1, ki),
                R.im + IX(R, 1, ki), flags);
      if (!(flags & ORACLE_NO_T_RESCALE)) {
        /* max-norm "normalisation" applied to column ki of T (1278-1295) */
        double mx = 0.0;
        for (int k = 1; k <= n; ++k) {
          const double a = cabs1_(Tm.re[IX(Tm, k, ki)], Tm.im[IX(Tm, k, ki)]);
          mx = a > mx ? a : mx;
        }
        const double s = 1.0 / mx;
        for (int k = 1; k <= n; ++k) {
          Tm.re[IX(Tm, k, ki)] *= s;
          Tm.im[IX(Tm, k, ki)] *= s;
        }
      }
    }
    for (int k = 1; k <= ki - 1; ++k) {
      Tm.re[IX(Tm, k, k)] = rdiag[VX(T, k)];
      Tm.im[IX(Tm, k, k)] = idiag[VX(T, k)];
    }
  }
}

/* ---- batch dispatch (the reference's Toolbox.hpp:1984-2206 wrappers) ---- */

static int g_threads = 0;
void oracle_set_threads(int nthreads) { g_threads = nthreads; }
int oracle_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

static mat_t slot(double *re, double *im, int n, int T, int b, int t) {
  mat_t M;
  const size_t off = (size_t)b * (size_t)T * (size_t)n * (size_t)n + (size_t)t;
  M.re = re + off;
  M.im = im + off;
  M.n = n;
  M.T = (size_t)T;
  return M;
}
static size_t vslot(int n, int T, int b, int t, int mult) {
  return (size_t)mult * (size_t)b * (size_t)T * (size_t)n + (size_t)t;
}
static double *zalloc(size_t count) {
  double *p = (double *)calloc(count ? count : 1, sizeof(double));
  if (!p) abort();
  return p;
}

#ifdef _OPENMP
#define OMP_THREADS() (g_threads > 0 ? g_threads : omp_get_max_threads())
#endif

void oracle_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                   double *itau, int flags) {
  const size_t sv = (size_t)dim * block_cnt * thread_cnt;
  double *rw = zalloc(sv), *iw = zalloc(sv);
#pragma omp parallel for schedule(static, 8) num_threads(OMP_THREADS())
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t v = vslot(dim, thread_cnt, b, t, 1);
      zgehrd_one(slot(rA, iA, dim, thread_cnt, b, t), rtau + v, itau + v, rw + v, iw + v, flags);
    }
  free(rw);
  free(iw);
}

void oracle_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                   const double *rtau, const double *itau, int flags) {
  const size_t sv = (size_t)dim * block_cnt * thread_cnt;
  double *rw = zalloc(sv), *iw = zalloc(sv);
#pragma omp parallel for schedule(static, 8) num_threads(OMP_THREADS())
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t v = vslot(dim, thread_cnt, b, t, 1);
      zunghr_one(slot(rA, iA, dim, thread_cnt, b, t), rtau + v, itau + v, rw + v, iw + v, flags);
    }
  free(rw);
  free(iw);
}

/* hessenberg(), utils.cpp:417-441: zgehrd, copy, zunghr. */
void oracle_hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rQ,
                       double *iQ, int flags) {
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  double *rt = zalloc(sv), *it = zalloc(sv);
  oracle_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rt, it, flags);
  memcpy(rQ, rA, sm * sizeof(double));
  memcpy(iQ, iA, sm * sizeof(double));
  oracle_zunghr(dim, block_cnt, thread_cnt, rQ, iQ, rt, it, flags);
  free(rt);
  free(it);
}

void oracle_zlahqr(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rH,
                   double *iH, double *rZ, double *iZ, int flags, oracle_stats *stats) {
#pragma omp parallel for schedule(static, 8) num_threads(OMP_THREADS())
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t v = vslot(dim, thread_cnt, b, t, 1);
      const size_t m = (size_t)b * thread_cnt + t;
      zlahqr_one(slot(rH, iH, dim, thread_cnt, b, t), slot(rZ, iZ, dim, thread_cnt, b, t),
                 rE + v, iE + v, flags, stats && stats->info ? stats->info + m : NULL,
                 stats && stats->sweeps ? stats->sweeps + m : NULL,
                 stats && stats->steps ? stats->steps + m : NULL);
    }
}

/* call_zlahqr_new, utils.cpp:159-171: Z is set to zero and test_diag is
 * called with the planes swapped (utils.cpp:169), so Z starts as i*I. */
void oracle_call_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE,
                            double *rM, double *iM, double *rS, double *iS, int flags) {
  const size_t sm = (size_t)dim * dim * block_cnt * thread_cnt;
  memset(rS, 0, sm * sizeof(double));
  memset(iS, 0, sm * sizeof(double));
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      mat_t S = slot(iS, rS, dim, thread_cnt, b, t); /* swapped, as the reference */
      for (int j = 1; j <= dim; ++j) {
        S.re[IX(S, j, j)] = 1.0;
        S.im[IX(S, j, j)] = 0.0;
      }
    }
  oracle_zlahqr(dim, block_cnt, thread_cnt, rE, iE, rM, iM, rS, iS, flags, NULL);
}

/* zlahqr_ (eigenvalues only), Toolbox.hpp:1663-1890, over the batch: H is
 * upper Hessenberg with a real subdiagonal and is updated inside the active
 * blocks only; E receives its diagonal. */
void oracle_zlahqr_eig(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rH,
                       double *iH, int flags, oracle_stats *stats) {
  const mat_t none = {NULL, NULL, dim, (size_t)thread_cnt};
#pragma omp parallel for schedule(static, 8) num_threads(OMP_THREADS())
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t v = vslot(dim, thread_cnt, b, t, 1);
      const size_t m = (size_t)b * thread_cnt + t;
      zlahqr_one(slot(rH, iH, dim, thread_cnt, b, t), none, rE + v, iE + v, flags | ORACLE_EIG_ONLY,
                 stats && stats->info ? stats->info + m : NULL,
                 stats && stats->sweeps ? stats->sweeps + m : NULL,
                 stats && stats->steps ? stats->steps + m : NULL);
    }
}

/* create_companion_matrices, Toolbox.hpp:605-643: the coefficients
 * c_0..c_dim of each polynomial (vector slot of length dim+1) give the
 * companion matrix with ones on the subdiagonal and last column
 * -c_i / c_dim (Smith division, ComplexDouble.hpp:319-353). */
void oracle_companion(int dim, int block_cnt, int thread_cnt, const double *rC, const double *iC,
                      double *rM, double *iM) {
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t v = vslot(dim + 1, thread_cnt, b, t, 1);
      const size_t T = (size_t)thread_cnt;
      mat_t M = slot(rM, iM, dim, thread_cnt, b, t);
      for (int c = 1; c <= dim; ++c)
        for (int r = 1; r <= dim; ++r) {
          M.re[IX(M, r, c)] = 0.0;
          M.im[IX(M, r, c)] = 0.0;
        }
      const double cr = rC[v + (size_t)dim * T], ci = iC[v + (size_t)dim * T];
      for (int i = 0; i < dim; ++i) {
        double a = rC[v + (size_t)i * T], bb = iC[v + (size_t)i * T];
        c_div(&a, &bb, cr, ci);
        M.re[IX(M, i + 1, dim)] = -a;
        M.im[IX(M, i + 1, dim)] = -bb;
      }
      for (int i = 1; i < dim; ++i) M.re[IX(M, i + 1, i)] = 1.0;
    }
}

/* zlahqr kernel (polynomial roots), Toolbox.hpp:1894-1957: companion matrix
 * into V, Z = I, zlahqr_schur(wantT=1, wantZ=1); E receives the roots, V the
 * Schur form T and Z the Schur vectors. */
void oracle_zlahqr_roots(int dim, int block_cnt, int thread_cnt, const double *rC, const double *iC,
                         double *rE, double *iE, double *rV, double *iV, double *rZ, double *iZ,
                         int flags, oracle_stats *stats) {
  oracle_companion(dim, block_cnt, thread_cnt, rC, iC, rV, iV);
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      mat_t Z = slot(rZ, iZ, dim, thread_cnt, b, t);
      for (int c = 1; c <= dim; ++c)
        for (int r = 1; r <= dim; ++r) {
          Z.re[IX(Z, r, c)] = r == c ? 1.0 : 0.0;
          Z.im[IX(Z, r, c)] = 0.0;
        }
    }
  oracle_zlahqr(dim, block_cnt, thread_cnt, rE, iE, rV, iV, rZ, iZ, flags, stats);
}

void oracle_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                   double *iR, int flags) {
  const size_t sv = (size_t)dim * block_cnt * thread_cnt;
  double *rw = zalloc(2 * sv), *iw = zalloc(2 * sv);
#pragma omp parallel for schedule(static, 8) num_threads(OMP_THREADS())
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t v = vslot(dim, thread_cnt, b, t, 2);
      ztrevc_one(slot(rT, iT, dim, thread_cnt, b, t), slot(rR, iR, dim, thread_cnt, b, t), rw + v,
                 iw + v, flags);
    }
  free(rw);
  free(iw);
}

/* eigensolver(), utils.cpp:578-615: zgehrd -> copy A to V -> zunghr ->
 * zlahqr(H=A, Z=V) -> ztrevc(T=A, R=V).  Unless ORACLE_KEEP_TRASH, A(i+2:n,i)
 * is zeroed before the QR stage. */
void oracle_eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rE,
                        double *iE, double *rV, double *iV, int flags, oracle_stats *stats) {
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  double *rt = zalloc(sv), *it = zalloc(sv);
  oracle_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rt, it, flags);
  memcpy(rV, rA, sm * sizeof(double));
  memcpy(iV, iA, sm * sizeof(double));
  oracle_zunghr(dim, block_cnt, thread_cnt, rV, iV, rt, it, flags);
  if (!(flags & ORACLE_KEEP_TRASH)) {
    for (int b = 0; b < block_cnt; ++b)
      for (int t = 0; t < thread_cnt; ++t) {
        mat_t A = slot(rA, iA, dim, thread_cnt, b, t);
        for (int c = 1; c <= dim; ++c)
          for (int r = c + 2; r <= dim; ++r) {
            A.re[IX(A, r, c)] = 0.0;
            A.im[IX(A, r, c)] = 0.0;
          }
      }
  }
  oracle_zlahqr(dim, block_cnt, thread_cnt, rE, iE, rA, iA, rV, iV, flags, stats);
  oracle_ztrevc(dim, block_cnt, thread_cnt, rA, iA, rV, iV, flags);
  free(rt);
  free(it);
}

/* test_zgemv batch entry, Toolbox.hpp:2118-2148 (A is dim x dim). */
void oracle_zgemv(int hermitian, int dim, int block_cnt, int thread_cnt, const double *rM,
                  const double *iM, const double *rx, const double *ix, double *ry, double *iy,
                  double rbeta, double ibeta, int flags) {
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t mo = (size_t)b * thread_cnt * dim * dim + t;
      const size_t v = vslot(dim, thread_cnt, b, t, 1);
      zgemv_ref(hermitian, dim, dim, rM + mo, iM + mo, dim, (size_t)thread_cnt, rx + v, ix + v,
                rbeta, ibeta, ry + v, iy + v, flags);
    }
}

/* test_backsubstitution batch entry, Toolbox.hpp:2090-2116. */
void oracle_backsubstitution(int dim, int block_cnt, int thread_cnt, const double *rT,
                             const double *iT, double *rb, double *ib) {
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t mo = (size_t)b * thread_cnt * dim * dim + t;
      const size_t v = vslot(dim, thread_cnt, b, t, 1);
      backsub_one(dim, rT + mo, iT + mo, dim, (size_t)thread_cnt, rb + v, ib + v);
    }
}
