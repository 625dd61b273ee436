// veig_complex.hpp -- the scalar complex layer of the reference header
// library (ComplexDouble.hpp:20-488, default non-MKL branch) with the same
// names and semantics, usable on the host and in HIP device code.
//
// Split (re, im) doubles passed by reference, as the reference's free
// functions take them:
//   cinv     r+ii <- 1/(r+ii), Smith-style (ComplexDouble.hpp:283-317)
//   cdiv     r+ii <- (r+ii)/(rr+i ri), Smith's method (319-353); real divisor (387-391)
//   cadd, csub, cmul, cmulc (a*conj(b)), real cmul                    (355-385)
//   dabs2, dabs3  scaled 2- and 3-norms                               (393-430)
//   cabs1 = |re| + |im|; cabs = sqrt(re^2 + im^2) unscaled            (432-449)
//   csqrt    principal square root                                  (451-488)
// and the header library's row/column scaling scale_rc (Toolbox.hpp:14-29).
#pragma once

#include <cmath>
#include <ostream>

#if defined(__HIPCC__)
#define VEIG_HD __host__ __device__ inline
#else
#define VEIG_HD inline
#endif

VEIG_HD double dabs(double x) { return std::fabs(x); }

VEIG_HD void cinv(double &r, double &i) {
  const double rr = r, ri = i;
  const bool a = std::fabs(rr) <= std::fabs(ri);
  const double num = a ? rr : ri, den0 = a ? ri : rr;
  const double ratio = num / den0;
  const double den = den0 * (1.0 + ratio * ratio);
  const double p = ratio / den, q = 1.0 / den;
  if (a) {
    r = p;
    i = -q;
  } else {
    r = q;
    i = -p;
  }
}

VEIG_HD void cdiv(double &r, double &i, const double rr, const double ri) {
  const bool a = std::fabs(rr) <= std::fabs(ri);
  const double ratio = a ? rr / ri : ri / rr;
  const double den = (a ? ri : rr) * (1.0 + ratio * ratio);
  double cr, ci;
  if (a) {
    cr = r * ratio + i;
    ci = i * ratio - r;
  } else {
    cr = r + i * ratio;
    ci = i - r * ratio;
  }
  r = cr / den;
  i = ci / den;
}

VEIG_HD void cdiv(double &r, double &i, const double rr) {
  r /= rr;
  i /= rr;
}

VEIG_HD void cadd(double &r, double &i, const double rr, const double ri) {
  r += rr;
  i += ri;
}

VEIG_HD void csub(double &r, double &i, const double rr, const double ri) {
  r -= rr;
  i -= ri;
}

VEIG_HD void cmul(double &r, double &i, const double rr, const double ri) {
  const double lr = r, li = i;
  r = lr * rr - li * ri;
  i = li * rr + lr * ri;
}

VEIG_HD void cmulc(double &r, double &i, const double rr, const double ri) {
  const double lr = r, li = i;
  r = lr * rr + li * ri;
  i = li * rr - lr * ri;
}

VEIG_HD void cmul(double &r, double &i, const double rr) {
  r *= rr;
  i *= rr;
}

VEIG_HD double dabs2(double x, double y) {
  x = std::fabs(x);
  y = std::fabs(y);
  const double big = x > y ? x : y, small = x > y ? y : x;
  const double q = big == 0.0 ? 0.0 : small / big;
  return big * std::sqrt(1.0 + q * q);
}

VEIG_HD double dabs3(double x, double y, double z) {
  x = std::fabs(x);
  y = std::fabs(y);
  z = std::fabs(z);
  const double w = std::fmax(std::fmax(x, y), z);
  if (w == 0.0) return x + y + z;
  x /= w;
  y /= w;
  z /= w;
  return w * std::sqrt(x * x + y * y + z * z);
}

VEIG_HD double cabs1(double r, double i) { return std::fabs(r) + std::fabs(i); }

VEIG_HD double cabs(double r, double i) { return std::sqrt(r * r + i * i); }

VEIG_HD void csqrt(double &r, double &i) {
  const double lr = r, li = i;
  const double zr = std::sqrt(0.5 * (cabs(lr, li) + std::fabs(lr)));
  const double zi = zr == 0.0 ? 0.0 : std::fabs(li) / (2.0 * zr);
  double cr = lr < 0.0 ? zi : zr;
  double ci = lr < 0.0 ? zr : zi;
  if (li < 0.0) ci = -ci;
  r = cr;
  i = ci;
}

// The value type of the reference (ComplexDouble.hpp:110-281): a pair of
// doubles with the usual compound operators built on the functions above.
// scale_rc (Toolbox.hpp:14-29), the step that makes a subdiagonal entry
// real: in the dim x dim column-major matrix whose elements lie n apart
// from (rV, iV), row I times z over columns I+1..I2, and column I times
// conj(z) over rows I1..I-1 (1-based indices, as in the reference)
VEIG_HD void scale_rc(double *rV, double *iV, double zr, double zi, int I, int I1, int I2, int dim, int n) {
  const long long ld = (long long)dim * n;
  for (int c = I + 1; c <= I2; ++c) {
    const long long k = (long long)(I - 1) * n + (long long)(c - 1) * ld;
    cmul(rV[k], iV[k], zr, zi);
  }
  for (int r = I1; r <= I - 1; ++r) {
    const long long k = (long long)(r - 1) * n + (long long)(I - 1) * ld;
    cmulc(rV[k], iV[k], zr, zi);
  }
}

struct ComplexDouble {
  double r, i;
  VEIG_HD ComplexDouble(double real, double imag) : r(real), i(imag) {}
  VEIG_HD ComplexDouble() : r(0.0), i(0.0) {}
  VEIG_HD ComplexDouble &operator+=(const ComplexDouble &o) {
    cadd(r, i, o.r, o.i);
    return *this;
  }
  VEIG_HD ComplexDouble &operator-=(const ComplexDouble &o) {
    csub(r, i, o.r, o.i);
    return *this;
  }
  VEIG_HD ComplexDouble &operator*=(const ComplexDouble &o) {
    cmul(r, i, o.r, o.i);
    return *this;
  }
  VEIG_HD ComplexDouble &operator/=(const ComplexDouble &o) {
    cdiv(r, i, o.r, o.i);
    return *this;
  }
  VEIG_HD ComplexDouble &operator*=(const double &s) {
    cmul(r, i, s);
    return *this;
  }
  VEIG_HD ComplexDouble &operator/=(const double &s) {
    cdiv(r, i, s);
    return *this;
  }
  VEIG_HD ComplexDouble &cinv() {
    ::cinv(r, i);
    return *this;
  }
  VEIG_HD ComplexDouble &csqrt() {
    ::csqrt(r, i);
    return *this;
  }
  friend std::ostream &operator<<(std::ostream &os, ComplexDouble z) {
    return os << z.r << (z.i >= 0.0 ? "+" : "") << z.i << "i";
  }
};
