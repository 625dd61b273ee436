// Device-side building blocks of the batched complex eigensolver for gfx950.
//
// Execution model ("lane groups"): a wavefront of 64 lanes holds 64/NB
// matrices of order n <= NB; lane r of a group owns index r of its matrix:
//   * row r of the unitary factor (Q, then the Schur vectors Z, then V) in
//     registers, as NB complex values z[0..NB-1];
//   * in the reflector updates, row r (updates from the right) and column r
//     (updates from the left) of the Hessenberg / Schur matrix, which lives in
//     LDS as an NB x (NB+1) complex column-major tile (the +1 pad keeps both
//     row-wise and column-wise walks free of bank conflicts).
// Lanes of one group cooperate through LDS within a single wavefront, so the
// only synchronisation needed is a wavefront-scope fence (LDS operations of a
// wavefront complete in order).
//
// The arithmetic follows the reference chain zgehrd -> zunghr -> zlahqr ->
// ztrevc (Toolbox.hpp:1315, 1492, 690, 1133) with these MI355X-first changes:
//   * Q is accumulated forward (Q <- Q*H_i) in registers while the
//     Hessenberg reduction runs, instead of a second pass over the stored
//     reflectors (zunghr); the product is the same unitary matrix.
//   * zgehrd's reflectors use LAPACK zlarfg's sign and its x == 0 guard
//     (the reference's opposite sign, Toolbox.hpp:1367-1373, cancels and its
//     missing guard divides by zero on triangular columns).
//   * the entries below the subdiagonal are zeroed as the reduction goes, so
//     the QR stage sees a clean Hessenberg matrix (the reference leaves the
//     reflectors there, Toolbox.hpp:721-722, which corrupts its first sweep).
//   * the deflation scan of every subdiagonal entry runs lane-parallel and is
//     combined with one ballot.
//   * the triangular solves for all eigenvectors run lane-parallel (lane k
//     solves for eigenvector k) and the back-transform V = Z*X runs
//     lane-parallel over rows of Z.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <type_traits>
#include <utility>

namespace veig {

// --------------------------------------------------------------------------
// scalar complex arithmetic (split re/im, as ComplexDouble.hpp:283-488)
// --------------------------------------------------------------------------
template <typename R>
struct cx {
  R re, im;
};

template <typename R> struct lds2;
template <> struct lds2<double> { using type = double2; };
template <> struct lds2<float> { using type = float2; };

template <typename R> struct limits;
template <> struct limits<double> {
  static constexpr double eps = DBL_EPSILON;
  static constexpr double tiny = DBL_MIN;
};
template <> struct limits<float> {
  static constexpr float eps = FLT_EPSILON;
  static constexpr float tiny = FLT_MIN;
};

// --------------------------------------------------------------------------
// reciprocal and square root from the hardware approximations (v_rcp_f64,
// v_rsq_f64; the seeds are good to 2^-24.2, measured over 2^26 arguments
// across the exponent range) refined by one third-order step, in place of
// the IEEE division / sqrt expansions.  With e = 1 - x y^2 (rsqrt) or
// e = 1 - x y (reciprocal), |e| < 2^-23:
//   1/sqrt(x) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + O(e^3)),
//   1/x       = y (1 - e)^(-1)   = y (1 + e + e^2 + O(e^3)),
// and the dropped terms are below 2^-68 relative: the result is within one
// ulp, in fewer dependent steps than two Newton iterations.  Arguments are
// finite and nonzero where rcp/rsq are used (guarded by the callers);
// rsqrt-based sqrt maps 0 to 0.
// --------------------------------------------------------------------------
__device__ __forceinline__ double frcp(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, y, 1.0);
  return fma(y, fma(e, e, e), y);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ double frsq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// One Newton step from the seed: relative error ~1.5 (2^-24.2)^2 = 2^-47.8,
// for quantities that steer the iteration without entering the similarity
// transform (the shift), where full precision buys nothing measurable
__device__ __forceinline__ double frcp1(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return fma(y, fma(-x, y, 1.0), y);
}
__device__ __forceinline__ float frcp1(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frsq1(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return fma(y, fma(-0.5 * x * y, y, 0.5), y);
}
__device__ __forceinline__ float frsq1(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ double fsqrt(double x) {
  // s = x*rsq(x), then one Newton correction on s itself
  if (!(x > 0.0)) return x == 0.0 ? 0.0 : __builtin_nan("");
  const double y = frsq(x);
  const double s = x * y;
  const double d = fma(-s, s, x);
  return fma(0.5 * y, d, s);
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_sqrtf(x); }

template <typename R> __device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <typename R> __device__ __forceinline__ cx<R> add(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __device__ __forceinline__ cx<R> sub(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __device__ __forceinline__ cx<R> mul(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename R> __device__ __forceinline__ cx<R> mulc(cx<R> a, cx<R> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// conj(a) * b
template <typename R> __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <typename R> __device__ __forceinline__ cx<R> scale(cx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> __device__ __forceinline__ cx<R> conj(cx<R> a) { return {a.re, -a.im}; }
template <typename R> __device__ __forceinline__ R cabs1(cx<R> a) { return fabs(a.re) + fabs(a.im); }
template <typename R> __device__ __forceinline__ R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
// 1/a as conj(a)/|a|^2 (one real division)
template <typename R> __device__ __forceinline__ cx<R> inv(cx<R> a) {
  const R s = frcp(abs2(a));
  return {a.re * s, -a.im * s};
}
// a/b by Smith's method (ComplexDouble.hpp:319-353), robust to |b| range
template <typename R> __device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) {
  if (fabs(b.re) <= fabs(b.im)) {
    const R ratio = b.re / b.im, den = b.im * (R(1) + ratio * ratio);
    return {(a.re * ratio + a.im) / den, (a.im * ratio - a.re) / den};
  } else {
    const R ratio = b.im / b.re, den = b.re * (R(1) + ratio * ratio);
    return {(a.re + a.im * ratio) / den, (a.im - a.re * ratio) / den};
  }
}
// principal square root (ComplexDouble.hpp:451-488)
template <typename R> __device__ __forceinline__ cx<R> csqrt(cx<R> a) {
  const R zr = fsqrt(R(0.5) * (fsqrt(abs2(a)) + fabs(a.re)));
  const R zi = (zr == R(0)) ? R(0) : fabs(a.im) * (R(0.5) * frcp(zr));
  cx<R> o = (a.re < R(0)) ? mk(zi, zr) : mk(zr, zi);
  if (a.im < R(0)) o.im = -o.im;
  return o;
}

// frexp's exponent (x = m 2^e, 1/2 <= |m| < 1; 0 for x = 0) and ldexp: exact
// power-of-two scaling
__device__ __forceinline__ int fexp(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ __forceinline__ int fexp(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ __forceinline__ double fldexp(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
__device__ __forceinline__ float fldexp(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
template <typename R> __device__ __forceinline__ cx<R> cldexp(cx<R> a, int e) {
  return {fldexp(a.re, e), fldexp(a.im, e)};
}

// principal square root of an a with |a| of order one (the scaled Wilkinson
// discriminant): t = sqrt((|a| + |Re a|) / 2) = q rsq(q) and |Im a| / (2 t) =
// |Im a| rsq(q) / 2, two reciprocal square roots in place of two square roots
// and a reciprocal (one Newton step each: it feeds the shift only); 0 for a = 0
template <typename R> __device__ __forceinline__ cx<R> csqrt_unit(cx<R> a) {
  const R m2 = abs2(a);
  const R m = m2 > R(0) ? m2 * frsq1(m2) : R(0);
  const R q = R(0.5) * (m + fabs(a.re));
  const R rq = q > R(0) ? frsq1(q) : R(0);
  const R zr = q * rq, zi = fabs(a.im) * (R(0.5) * rq);
  cx<R> o = (a.re < R(0)) ? mk(zi, zr) : mk(zr, zi);
  if (a.im < R(0)) o.im = -o.im;
  return o;
}

template <typename R>
__device__ __forceinline__ cx<R> ld(const typename lds2<R>::type* p) {
  const typename lds2<R>::type v = *p;
  return {v.x, v.y};
}
template <typename R>
__device__ __forceinline__ void st(typename lds2<R>::type* p, cx<R> a) {
  typename lds2<R>::type v;
  v.x = a.re;
  v.y = a.im;
  *p = v;
}

// Wavefront-scope ordering of LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x from the lane that DPP control CTRL names within this lane's row of 16
// (every lane of the row active)
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xf, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
}

// Sum over the NB lanes of a lane group (butterfly, all lanes get the total;
// every lane of the wavefront active).  Groups of 8 or 16 lanes sit in one
// DPP row: partners lane^1 and lane^2 (quad permutations), then the mirrored
// lane within the half row and within the row, in registers -- the same sum
// in every lane, as each step adds two values that are swapped between the
// partners.  Wider groups exchange through the LDS crossbar.
template <int NB, typename R>
__device__ __forceinline__ R group_sum(R x) {
  if constexpr (NB == 8 || NB == 16) {
    x += dpp_mov<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dpp_mov<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp_mov<0x141>(x);  // row_half_mirror
    if constexpr (NB == 16) x += dpp_mov<0x140>(x);  // row_mirror
  } else {
#pragma unroll
    for (int o = NB / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, NB);
  }
  return x;
}

// Call f(std::integral_constant<int, k>) for a wavefront-uniform runtime k in
// [LO, HI): a binary tree of scalar branches, so that f can index register
// arrays with a compile-time constant.
template <int LO, int HI, typename F>
__device__ __forceinline__ void uniform_dispatch(int k, F&& f) {
  if constexpr (HI - LO <= 1) {
    f(std::integral_constant<int, LO>{});
  } else {
    constexpr int MID = (LO + HI) / 2;
    if (k < MID) uniform_dispatch<LO, MID>(k, f);
    else uniform_dispatch<MID, HI>(k, f);
  }
}

// f(integral_constant<int, I>) for I = 0 .. N-1 in order, unrolled
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// --------------------------------------------------------------------------
// per-group context
// --------------------------------------------------------------------------
// A group of LG = NB/P lanes holds one matrix of order n <= NB; lane q of the
// group owns the P indices q, q + LG, ..., q + (P-1) LG ("slots").  P = 1 is
// one lane per index; P = 2 and 4 put 2 or 4 times as many matrices in a
// wavefront, which spreads the per-step scalar work (reflector generation,
// shifts, deflation bookkeeping) over more matrices per instruction.
// LAY = LAY_DENSE: the NB x NB matrix in an NB x (NB+1) column-major tile.
// LAY = LAY_BAND: only the band row <= col + 2 (the Hessenberg matrix and the
// bulge of the QR sweep), column c holding rows 0..min(c+2, NB-1) from
// offset c(c+5)/2: 166 instead of 272 entries at NB = 16.  Enough for the
// eigenvalues-only QR iteration, which never leaves that band.
enum { LAY_DENSE = 0, LAY_BAND = 1 };
// offset of column c in the banded tile (see Group)
template <int NB>
__host__ __device__ constexpr int band_off(int c) {
  return (c * (c + 5)) >> 1;
}

// LW: lanes per matrix when wider than NB / P (a slot of capacity NB on a
// lane group of LW lanes, the lanes past NB / P idle in row operations)
template <typename R, int NB, int P = 1, int LAY = LAY_DENSE, int LW = 0>
struct Group {
  static constexpr int LG = LW ? LW : NB / P;  // lanes per matrix
  static_assert(NB <= LG * P && 64 % LG == 0, "lane group must divide the wavefront and hold NB indices");
  static constexpr int LD = NB + 1;
  static constexpr int MPW = 64 / LG;  // matrices per wavefront
  // complex entries of one matrix's tile
  static constexpr int kSize = LAY == LAY_BAND ? band_off<NB>(NB - 1) + NB : NB * LD;
  using L2 = typename lds2<R>::type;
  L2* H;     // the tile: H(r,c) at H[c*LD + r], or banded (see above)
  L2* vsh;   // NB complex scratch (reflector vector)
  int r;     // lane index within the group (slot 0 index)
  int mi;    // group index within the wavefront
  int n;     // order of this group's matrix (0 if the slot is empty)
  __device__ __forceinline__ int idx(int p) const { return r + p * LG; }
  __device__ __forceinline__ L2* at(int row, int col) const {
    if constexpr (LAY == LAY_BAND) return H + band_off<NB>(col) + row;
    else return H + col * LD + row;
  }
};

// --------------------------------------------------------------------------
// Hessenberg reduction with forward accumulation of Q (zgehrd + zunghr,
// Toolbox.hpp:1315-1659).  On exit H holds the upper Hessenberg matrix with
// a real subdiagonal and exact zeros below it, and z[p][] holds row idx(p) of
// Q with A = Q H Q^H.  nmax is the wavefront-uniform loop bound (max n in the
// wave).  If keep_reflectors, v(i+2:n) is left in H(i+2:n, i) and tau(i) is
// returned in tau_out[p] of the lane owning index i (zgehrd's convention).
// --------------------------------------------------------------------------
template <typename R, int NB, int P>
__device__ __forceinline__ void gehrd_q(const Group<R, NB, P>& g, cx<R> (&z)[P][NB], int nmax,
                                        bool accumulate_q, bool keep_reflectors, cx<R>* tau_out) {
  constexpr int LG = Group<R, NB, P>::LG;
  const int n = g.n;
  for (int i = 0; i + 1 < nmax; ++i) {
    // ---- reflector annihilating H(i+2:n, i) (zlarfg, LAPACK sign) ----
    cx<R> hri[P];
    R part = R(0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int row = g.idx(p);
      hri[p] = row < n ? ld<R>(g.at(row, i)) : mk<R>(0, 0);
      if (row >= i + 2 && row < n) part += abs2(hri[p]);
    }
    const R xnorm2 = group_sum<LG>(part);
    const bool live = i + 1 < n;  // this matrix still has a column to reduce
    const cx<R> alpha = live ? ld<R>(g.at(i + 1, i)) : mk<R>(0, 0);
    cx<R> tau = mk<R>(0, 0), scal = mk<R>(0, 0);
    R beta = alpha.re;
    if (live && !(xnorm2 == R(0) && alpha.im == R(0))) {
      const R nrm = fsqrt(abs2(alpha) + xnorm2);
      beta = alpha.re >= R(0) ? -nrm : nrm;
      const R ib = frcp(beta);
      tau = mk((beta - alpha.re) * ib, -alpha.im * ib);
      scal = inv(mk(alpha.re - beta, alpha.im));
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int row = g.idx(p);
      const bool in = row < n;
      cx<R> vr = mk<R>(0, 0);
      if (live) {
        if (row == i + 1) vr = mk<R>(1, 0);
        else if (row >= i + 2 && in) vr = mul(hri[p], scal);
      }
      st<R>(g.vsh + row, vr);
      if (live && in && row >= i + 1) {
        if (row == i + 1) st<R>(g.at(row, i), mk<R>(beta, 0));
        else st<R>(g.at(row, i), keep_reflectors ? vr : mk<R>(0, 0));
      }
      if (keep_reflectors && row == i) tau_out[p] = tau;
    }
    wave_sync();
    // ---- H <- H * (I - tau v v^H) on rows 0..n-1, columns i+1..n-1 ----
    if (live) {
      cx<R> w[P];
#pragma unroll
      for (int p = 0; p < P; ++p) w[p] = mk<R>(0, 0);
      for (int c = i + 1; c < nmax; ++c) {
        const cx<R> v = ld<R>(g.vsh + c);  // zero beyond n
#pragma unroll
        for (int p = 0; p < P; ++p) w[p] = add(w[p], mul(ld<R>(g.at(g.idx(p), c)), v));
      }
#pragma unroll
      for (int p = 0; p < P; ++p) w[p] = mul(tau, w[p]);
      for (int c = i + 1; c < nmax; ++c) {
        const cx<R> v = ld<R>(g.vsh + c);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int row = g.idx(p);
          if (row < n && c < n) st<R>(g.at(row, c), sub(ld<R>(g.at(row, c)), mulc(w[p], v)));
        }
      }
    }
    // ---- Q <- Q * (I - tau v v^H) (rows idx(p) of Q in registers) ----
    if (accumulate_q) {
      cx<R> w[P];
#pragma unroll
      for (int p = 0; p < P; ++p) w[p] = mk<R>(0, 0);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const cx<R> v = ld<R>(g.vsh + c);
#pragma unroll
        for (int p = 0; p < P; ++p) w[p] = add(w[p], mul(z[p][c], v));
      }
#pragma unroll
      for (int p = 0; p < P; ++p) w[p] = mul(tau, w[p]);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const cx<R> v = ld<R>(g.vsh + c);
#pragma unroll
        for (int p = 0; p < P; ++p) z[p][c] = sub(z[p][c], mulc(w[p], v));
      }
    }
    wave_sync();
    // ---- H <- (I - conj(tau) v v^H) H on rows i+1..n-1, columns i+1..n-1 ----
    if (live) {
      cx<R> y[P];
#pragma unroll
      for (int p = 0; p < P; ++p) y[p] = mk<R>(0, 0);
      for (int k = i + 1; k < nmax; ++k) {
        const cx<R> v = ld<R>(g.vsh + k);
#pragma unroll
        for (int p = 0; p < P; ++p)
          if ((p + 1) * LG > i + 1) y[p] = add(y[p], cmul(v, ld<R>(g.at(k, g.idx(p)))));
      }
#pragma unroll
      for (int p = 0; p < P; ++p) y[p] = cmul(tau, y[p]);  // conj(tau) * y
      for (int k = i + 1; k < nmax; ++k) {
        const cx<R> v = ld<R>(g.vsh + k);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int col = g.idx(p);
          if ((p + 1) * LG > i + 1 && col < n && col >= i + 1 && k < n)
            st<R>(g.at(k, col), sub(ld<R>(g.at(k, col)), mul(v, y[p])));
        }
      }
    }
    wave_sync();
  }
}

// The Hessenberg reduction of gehrd_q (zgehrd with zlarfg's LAPACK sign,
// Toolbox.hpp:1315-1469) for one index per lane (NB <= 16), leaving the
// compact form: v(i+2:n) below the subdiagonal, tau(i) in tau_out of lane i.
// The reflectors are unrolled, so each update touches only the columns
// (right update) or rows (left update) i+1..NB-1 that reflector i changes --
// half the loads and flops of full-width updates -- and takes them into
// registers at once, loads the hardware overlaps.  The tile is zero outside
// the n x n matrix and v is zero past n, so entries past n are rewritten
// unchanged rather than masked.
template <int NB, int LW = 0>
__device__ __forceinline__ void gehrd_compact(const Group<double, NB, 1, LAY_DENSE, LW> &g, int nmax,
                                              cx<double> &tau_out) {
  using R = double;
  constexpr int LG = Group<R, NB, 1, LAY_DENSE, LW>::LG;
  static_assert(NB <= 16 && LG == NB, "one index per lane, at most 16");
  const int n = g.n, r = g.r;
  const bool in = r < n;
  static_for<NB - 1>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    if (i + 1 >= nmax) return;
    // ---- reflector annihilating H(i+2:n, i) ----
    const cx<R> hri = in ? ld<R>(g.at(r, i)) : mk<R>(0, 0);
    const R xnorm2 = group_sum<LG, R>(r >= i + 2 && in ? abs2(hri) : R(0));
    const bool live = i + 1 < n;
    const cx<R> alpha = live ? ld<R>(g.at(i + 1, i)) : mk<R>(0, 0);
    cx<R> tau = mk<R>(0, 0), scal = mk<R>(0, 0);
    R beta = alpha.re;
    if (live && !(xnorm2 == R(0) && alpha.im == R(0))) {
      // |beta| and 1/beta from one reciprocal square root
      const R nrm2 = abs2(alpha) + xnorm2, ir = frsq(nrm2), nrm = nrm2 * ir;
      beta = alpha.re >= R(0) ? -nrm : nrm;
      const R ib = alpha.re >= R(0) ? -ir : ir;
      tau = mk((beta - alpha.re) * ib, -alpha.im * ib);
      scal = inv(mk(alpha.re - beta, alpha.im));
    }
    cx<R> vr = mk<R>(0, 0);
    if (live) {
      if (r == i + 1) vr = mk<R>(1, 0);
      else if (r >= i + 2 && in) vr = mul(hri, scal);
    }
    st<R>(g.vsh + r, vr);
    if (live && in && r >= i + 1) st<R>(g.at(r, i), r == i + 1 ? mk<R>(beta, 0) : vr);
    if (r == i) tau_out = tau;
    wave_sync();
    constexpr int M = NB - 1 - i;  // columns / rows i+1..NB-1
    cx<R> h[M], v[M];
    // ---- H <- H (I - tau v v^H): row r over columns i+1..NB-1 ----
    if (live) {
#pragma unroll
      for (int j = 0; j < M; ++j) {
        h[j] = ld<R>(g.at(r, i + 1 + j));
        v[j] = ld<R>(g.vsh + i + 1 + j);
      }
      cx<R> w = mk<R>(0, 0);
#pragma unroll
      for (int j = 0; j < M; ++j) w = add(w, mul(h[j], v[j]));
      w = mul(tau, w);
#pragma unroll
      for (int j = 0; j < M; ++j) st<R>(g.at(r, i + 1 + j), sub(h[j], mulc(w, v[j])));
    }
    wave_sync();
    // ---- H <- (I - conj(tau) v v^H) H: column r > i over rows i+1..NB-1 ----
    if (live && r > i) {
#pragma unroll
      for (int j = 0; j < M; ++j) h[j] = ld<R>(g.at(i + 1 + j, r));
      cx<R> y = mk<R>(0, 0);
#pragma unroll
      for (int j = 0; j < M; ++j) y = add(y, cmul(v[j], h[j]));
      y = cmul(tau, y);  // conj(tau) * y
#pragma unroll
      for (int j = 0; j < M; ++j) st<R>(g.at(i + 1 + j, r), sub(h[j], mul(v[j], y)));
    }
    wave_sync();
  });
}

// --------------------------------------------------------------------------
// 2-element Householder reflector of the bulge chase (zlarfg with n = 2,
// LAPACK sign, Toolbox.hpp:183-242): given (a0, a1) returns t1, v2 and the
// new a0 = beta (real) such that (I - t1 [1;v2][1;v2]^H)^H [a0;a1] = [beta;0].
// --------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void larfg2(cx<R> a0, cx<R> a1, cx<R>& t1, cx<R>& v2, cx<R>& beta_out) {
  if (a1.re == R(0) && a1.im == R(0) && a0.im == R(0)) {
    t1 = mk<R>(0, 0);
    v2 = a1;
    beta_out = a0;
    return;
  }
  const R nrm2 = abs2(a0) + abs2(a1);
  const R rn = frsq(nrm2);                    // 1/norm
  const R nrm = nrm2 * rn;
  const R beta = a0.re >= R(0) ? nrm : -nrm;  // copysign(norm, Re a0)
  const cx<R> zz = mk(a0.re + beta, a0.im);
  const R ib = a0.re >= R(0) ? rn : -rn;      // 1/beta
  t1 = scale(zz, ib);
  v2 = mul(a1, inv(zz));
  beta_out = mk<R>(-beta, 0);
}

// The unitary of one bulge-chase step.  The reference chases the bulge
// with zlarfg's 2x2 reflectors (Toolbox.hpp:138-335); a plane rotation
// G = [c, -s; s, conj(c)] with c = a0 / r, s = a1 / r, r = |[a0; a1]|, maps
// [a0; a1] to G^H [a0; a1] = [r; 0] as the reflector maps it to [beta; 0],
// beta = -copysign(r, Re a0).  Both start every step from [a0; a1] with a1
// real -- the first step's a1 is the scaled subdiagonal entry, every later one
// the bulge s H(K+2, K+1) of a real subdiagonal entry (Toolbox.hpp:172-174)
// -- and both leave H(K, K-1) real, so the two chases differ by a similarity
// with a diagonal of signs: the same diagonal, subdiagonal and entry moduli,
// the same deflations and shifts, Schur vectors and eigenvectors equal up to
// the sign of each column.  The rotation needs one reciprocal square root
// (the reflector one more reciprocal) and applies with the same 12 flops.
template <typename R>
struct Rot {
  cx<R> c;
  R s;
};

// The rotation for [a0; a1], a1 real, r (the new H(K, K-1)) and r^2; where
// r^2 = 0 (the zero vector, or one below the underflow threshold) 1/r is
// infinite and the caller takes G = I, r = 0 instead
template <typename R>
__device__ __forceinline__ Rot<R> rot2(cx<R> a0, R a1, R& r_out, R& nrm2) {
  nrm2 = fma(a0.re, a0.re, fma(a0.im, a0.im, a1 * a1));
  const R ir = frsq(nrm2);
  r_out = nrm2 * ir;
  return Rot<R>{mk(a0.re * ir, a0.im * ir), a1 * ir};
}

// [h1; h2] <- G^H [h1; h2] (rows K, K+1): h1 = conj(c) h1 + s h2,
// h2 = c h2 - s h1.  Each part is a product and two fused multiply-adds;
// the minus signs sit on a product or on the addend of the last one, where
// the instruction carries them for free (an accumulating multiply-add would
// need the negated factor in a register of its own).
template <typename R>
__device__ __forceinline__ void rot_left(const Rot<R>& q, cx<R>& h1, cx<R>& h2) {
  const R r1 = fma(q.c.re, h1.re, fma(q.c.im, h1.im, q.s * h2.re));
  const R i1 = fma(q.c.re, h1.im, fma(q.s, h2.im, -(q.c.im * h1.re)));
  const R r2 = fma(q.c.re, h2.re, -fma(q.c.im, h2.im, q.s * h1.re));
  const R i2 = fma(q.c.re, h2.im, fma(q.c.im, h2.re, -(q.s * h1.im)));
  h1 = mk(r1, i1);
  h2 = mk(r2, i2);
}

// [h1, h2] <- [h1, h2] G (columns K, K+1 of H and of the Schur vectors):
// h1 = c h1 + s h2, h2 = conj(c) h2 - s h1
template <typename R>
__device__ __forceinline__ void rot_right(const Rot<R>& q, cx<R>& h1, cx<R>& h2) {
  const R r1 = fma(q.c.re, h1.re, fma(q.s, h2.re, -(q.c.im * h1.im)));
  const R i1 = fma(q.c.re, h1.im, fma(q.c.im, h1.re, q.s * h2.im));
  const R r2 = fma(q.c.re, h2.re, fma(q.c.im, h2.im, -(q.s * h1.re)));
  const R i2 = fma(q.c.re, h2.im, -fma(q.c.im, h2.re, q.s * h1.im));
  h1 = mk(r1, i1);
  h2 = mk(r2, i2);
}

// The Schur-vector update of the unrolled banded sweep at order 16 with
// row entry z[k] in fixed registers (re in v[104+4k : 105+4k], im in
// v[106+4k : 107+4k]): each step rotates columns K, K+1 in place, and
// making H(I, I-1) real scales z[I] in place, so within a sweep the row
// stays in one set of registers instead of migrating through fresh ones
// (measured 0.45% of the order-16 step; the allocator still keeps the row
// elsewhere between sweeps, about 60 moves per sweep).  The arithmetic is
// rot_right's and mulc's, operation for operation, so the results are
// bit-identical.
// Requires the 168-register budget of three wavefronts per SIMD (a smaller
// budget fails to compile, it cannot miscompile).
template <int K> struct ZPin;
template <int C> struct ZPinMulc;
// where the fixed registers apply: FP64, order 16 on the banded tile, several
// matrices per wavefront (the kernels held to three wavefronts per SIMD)
template <typename R, int NB, int P, int LAY, int MPW>
inline constexpr bool kPinnedRow = std::is_same_v<R, double> && NB == 16 && P == 1 && LAY == LAY_BAND && MPW > 1;
// ZPin<K>: [z[K], z[K+1]] <- [z[K], z[K+1]] G, G = [c, -s; s, conj(c)]:
// each output is c.re times its own entry plus a two-product partial sum
#define VEIG_ZPIN(K, A, B, C_, D, E, F, G_, H)                                                      \
  template <> struct ZPin<K> {                                                                    \
    static __device__ __forceinline__ void rot(double cr, double ci, double sn, double& ar,        \
                                               double& ai, double& br, double& bi) {              \
      double t0, t1, t2, t3;                                                                      \
      asm("v_mul_f64 %0, %9, %5\n\tv_mul_f64 %1, %10, %7\n\t"                                   \
          "v_mul_f64 %2, %10, %4\n\tv_mul_f64 %3, %10, %5\n\t"                                  \
          "v_fma_f64 %0, %10, %6, -%0\n\tv_fma_f64 %1, %9, %4, %1\n\t"                          \
          "v_fma_f64 %2, %9, %7, -%2\n\tv_fma_f64 %3, %9, %6, %3\n\t"                           \
          "v_fma_f64 %4, %8, %4, %0\n\tv_fma_f64 %5, %8, %5, %1\n\t"                            \
          "v_fma_f64 %6, %8, %6, %2\n\tv_fma_f64 %7, %8, %7, -%3"                                \
          : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3), "+{v[" #A ":" #B "]}"(ar),                \
            "+{v[" #C_ ":" #D "]}"(ai), "+{v[" #E ":" #F "]}"(br), "+{v[" #G_ ":" #H "]}"(bi)      \
          : "v"(cr), "v"(ci), "v"(sn));                                                          \
    }                                                                                             \
  };
// ZPinMulc<C>: z[C] <- z[C] conj(zz)
#define VEIG_ZPIN_MULC(C, A, B, C_, D)                                                           \
  template <> struct ZPinMulc<C> {                                                                \
    static __device__ __forceinline__ void mul(double zr, double zi, double& ar, double& ai) {     \
      double t0, t1;                                                                              \
      asm("v_mul_f64 %0, %3, %5\n\tv_mul_f64 %1, %2, %5\n\t"                                    \
          "v_fma_f64 %2, %2, %4, %0\n\tv_fma_f64 %3, %3, %4, -%1"                                 \
          : "=&v"(t0), "=&v"(t1), "+{v[" #A ":" #B "]}"(ar), "+{v[" #C_ ":" #D "]}"(ai)          \
          : "v"(zr), "v"(zi));                                                                    \
    }                                                                                             \
  };
VEIG_ZPIN(0, 104, 105, 106, 107, 108, 109, 110, 111)
VEIG_ZPIN(1, 108, 109, 110, 111, 112, 113, 114, 115)
VEIG_ZPIN(2, 112, 113, 114, 115, 116, 117, 118, 119)
VEIG_ZPIN(3, 116, 117, 118, 119, 120, 121, 122, 123)
VEIG_ZPIN(4, 120, 121, 122, 123, 124, 125, 126, 127)
VEIG_ZPIN(5, 124, 125, 126, 127, 128, 129, 130, 131)
VEIG_ZPIN(6, 128, 129, 130, 131, 132, 133, 134, 135)
VEIG_ZPIN(7, 132, 133, 134, 135, 136, 137, 138, 139)
VEIG_ZPIN(8, 136, 137, 138, 139, 140, 141, 142, 143)
VEIG_ZPIN(9, 140, 141, 142, 143, 144, 145, 146, 147)
VEIG_ZPIN(10, 144, 145, 146, 147, 148, 149, 150, 151)
VEIG_ZPIN(11, 148, 149, 150, 151, 152, 153, 154, 155)
VEIG_ZPIN(12, 152, 153, 154, 155, 156, 157, 158, 159)
VEIG_ZPIN(13, 156, 157, 158, 159, 160, 161, 162, 163)
VEIG_ZPIN(14, 160, 161, 162, 163, 164, 165, 166, 167)
VEIG_ZPIN_MULC(0, 104, 105, 106, 107)
VEIG_ZPIN_MULC(1, 108, 109, 110, 111)
VEIG_ZPIN_MULC(2, 112, 113, 114, 115)
VEIG_ZPIN_MULC(3, 116, 117, 118, 119)
VEIG_ZPIN_MULC(4, 120, 121, 122, 123)
VEIG_ZPIN_MULC(5, 124, 125, 126, 127)
VEIG_ZPIN_MULC(6, 128, 129, 130, 131)
VEIG_ZPIN_MULC(7, 132, 133, 134, 135)
VEIG_ZPIN_MULC(8, 136, 137, 138, 139)
VEIG_ZPIN_MULC(9, 140, 141, 142, 143)
VEIG_ZPIN_MULC(10, 144, 145, 146, 147)
VEIG_ZPIN_MULC(11, 148, 149, 150, 151)
VEIG_ZPIN_MULC(12, 152, 153, 154, 155)
VEIG_ZPIN_MULC(13, 156, 157, 158, 159)
VEIG_ZPIN_MULC(14, 160, 161, 162, 163)
VEIG_ZPIN_MULC(15, 164, 165, 166, 167)
#undef VEIG_ZPIN
#undef VEIG_ZPIN_MULC

// Apply one bulge-chase rotation at row K to this lane's slots: rows K, K+1
// of H over the lane's columns >= K, columns K, K+1 of H over the lane's rows
// <= i3, and columns K, K+1 of the Schur vectors.  KC >= 0 is K as a
// compile-time constant (unrolled sweeps); KC < 0 takes the wavefront-uniform
// runtime Krt and selects the register columns by a uniform branch tree.
// The rotation's input is column K-1, rows K and K+1; at the first step of
// a matrix's sweep (K = L) the caller has staged the shifted first column
// (a0; a1 real) there -- in vsh[0..1] when L = 0 -- so that every step reads
// its input from a fixed place, without a select or a branch.
// TZ = false is the eigenvalues-only sweep (single_shift_qr, Toolbox.hpp:
// 372-522): H is updated inside the active block L..I only (columns <= I,
// rows >= L) and there are no Schur vectors.
template <int KC, bool TZ = true, typename R, int NB, int P, int LAY, int LW>
__device__ __forceinline__ void sweep_step(const Group<R, NB, P, LAY, LW>& g, cx<R> (&z)[P][NB], bool act,
                                           int L, int I, int Krt = 0) {
  using L2 = typename Group<R, NB, P, LAY, LW>::L2;
  constexpr int LG = Group<R, NB, P, LAY, LW>::LG;
  const int n = g.n;
  const int K = KC >= 0 ? KC : Krt;
  constexpr int MPW = Group<R, NB, P, LAY, LW>::MPW;
  // with several matrices per wavefront, groups outside their sweep at this
  // K read the input (1, 0; 0) that lahqr keeps in vsh[2..3], whose rotation
  // is the identity exactly: the Schur-vector update then needs no branch
  constexpr bool kIdle = TZ && MPW > 1;
  const L2* pb;
  if constexpr (KC == 0) pb = g.vsh;
  else if constexpr (KC > 0) pb = g.at(KC, KC - 1);
  else pb = K > 0 ? g.at(K, K - 1) : g.vsh;
  if constexpr (kIdle) pb = act ? pb : g.vsh + 2;
  // the row update's operands, loaded ahead of the rotation they wait for
  // (lanes outside the update read an entry of the tile and use none)
  cx<R> rh1[P], rh2[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if ((p + 1) * LG - 1 >= K) {
      const int col = g.idx(p) < NB ? g.idx(p) : NB - 1;
      rh1[p] = ld<R>(g.at(K, col));
      rh2[p] = ld<R>(g.at(K + 1, col));
    }
  }
  const cx<R> b0 = ld<R>(pb);
  const R b1 = pb[1].x;
  R nb0, nrm2;
  Rot<R> q = rot2(b0, b1, nb0, nrm2);
  const bool zero_in = nrm2 == R(0);
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(zero_in) != 0, 0)) {
    if (zero_in) {
      q.c = mk<R>(1, 0);
      q.s = R(0);
      nb0 = R(0);
    }
  }
  // H(K, K-1) = r, H(K+1, K-1) = 0; at K = L this overwrites the staged
  // first column, which lahqr clears after the sweep.  With several matrices
  // per wavefront every lane writes its group's pb, without a branch: a
  // group outside its sweep rewrites its identity input unchanged (r = 1
  // exactly), and at K = 0 pb is the staged column in vsh, read only above
  if constexpr (kIdle) {
    L2* pk = const_cast<L2*>(pb);
    st<R>(pk, mk<R>(nb0, R(0)));
    st<R>(pk + 1, mk<R>(0, 0));
  } else if (act && K > 0 && g.r == 0) {
    L2* pk = g.at(K, K > 0 ? K - 1 : 0);
    pk[0].x = nb0;
    pk[0].y = R(0);
    st<R>(pk + 1, mk<R>(0, 0));
  }
  // rows K, K+1 over columns K..n-1 (slot = column); groups outside their
  // sweep apply their identity rotation too, where it rewrites what they read
  const bool upd = kIdle || act;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if ((p + 1) * LG - 1 >= K) {
      const int col = g.idx(p);
      if (upd && col >= K && col < n && (TZ || col <= I)) {
        cx<R> h1 = rh1[p], h2 = rh2[p];
        rot_left(q, h1, h2);
        st<R>(g.at(K, col), h1);
        st<R>(g.at(K + 1, col), h2);
      }
    }
  }
  wave_sync();
  // columns K, K+1 over rows 0..min(K+2, I) (slot = row)
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if (p * LG <= K + 2) {
      const int row = g.idx(p);
      if (upd && row <= K + 2 && row <= I && (TZ || row >= L)) {
        L2* p1 = g.at(row, K);
        L2* p2 = g.at(row, K + 1);
        cx<R> h1 = ld<R>(p1), h2 = ld<R>(p2);
        rot_right(q, h1, h2);
        st<R>(p1, h1);
        st<R>(p2, h2);
      }
    }
  }
  // Schur vectors: columns K, K+1 of each owned row
  if (TZ && upd) {
    auto zupd = [&](auto kc) {
      constexpr int KK = decltype(kc)::value;
      if constexpr (kPinnedRow<R, NB, P, LAY, MPW>) {
        ZPin<KK>::rot(q.c.re, q.c.im, q.s, z[0][KK].re, z[0][KK].im, z[0][KK + 1].re, z[0][KK + 1].im);
      } else {
#pragma unroll
        for (int p = 0; p < P; ++p) rot_right(q, z[p][KK], z[p][KK + 1]);
      }
    };
    if constexpr (!TZ) {
    } else if constexpr (KC >= 0) {
      zupd(std::integral_constant<int, KC>{});
    } else {
      uniform_dispatch<0, NB - 1>(K, zupd);
    }
  }
  wave_sync();
}

// The unrolled sweep over K = 0 ... KEND-1 by bit masks: bit K of the
// wavefront-uniform kmask says some group's sweep passes step K (one scalar
// bit test per step), bit K of the lane's amask that its own group's does
// (one vector bit test): the range compares they replace cost 2% of the
// order-16 step in scalar instructions alone.
template <int K, int KEND, bool TZ = true, typename R, int NB, int P, int LAY, int LW>
__device__ __forceinline__ void sweep_range(const Group<R, NB, P, LAY, LW>& g, cx<R> (&z)[P][NB], unsigned amask,
                                            int L, int I, unsigned kmask) {
  if constexpr (K < KEND) {
    if ((kmask >> K) & 1u) sweep_step<K, TZ>(g, z, (amask >> K) & 1u, L, I);
    sweep_range<K + 1, KEND, TZ>(g, z, amask, L, I, kmask);
  }
}

// --------------------------------------------------------------------------
// Single-shift complex QR to Schur form with Schur vectors (zlahqr_schur with
// wantT = wantZ = 1, Toolbox.hpp:690-1031): the reference's deflation test
// (Ahues & Tisseur), its exceptional shifts at iterations 10 and 20 and its
// Wilkinson shift otherwise.  Every matrix of the wavefront iterates
// independently; a sweep is one pass of the bulge chase over the union of
// the active row ranges.  An eigenvalue not converged after itmax + 1 sweeps
// is accepted as it stands (the reference's silent case, Toolbox.hpp:808);
// the count of such eigenvalues is returned.
// If make_real_subdiag, the subdiagonal is first made real (loop 20,
// Toolbox.hpp:735-793); after gehrd_q it already is.
// TZ = false: the eigenvalues-only iteration (zlahqr_, Toolbox.hpp:1663-1890):
// no Schur vectors and H updated inside the active block only; the active
// block evolves exactly as with TZ = true, so the eigenvalues are the same.
// Every group needs its own g.vsh[0..1] (TZ with several matrices per
// wavefront: g.vsh[0..3]): a sweep that starts at row 0 stages its first
// column there.
// --------------------------------------------------------------------------
template <typename R, int NB, int P, bool TZ = true, int LAY = LAY_DENSE, int LW = 0>
__device__ __forceinline__ int lahqr(const Group<R, NB, P, LAY, LW>& g, cx<R> (&z)[P][NB], int itmax,
                                     bool make_real_subdiag) {
  constexpr int LG = Group<R, NB, P, LAY, LW>::LG;
  constexpr int MPW = Group<R, NB, P, LAY, LW>::MPW;
  const int n = g.n;
  const R ulp = limits<R>::eps;
  const R smlnum = limits<R>::tiny * (R(n > 0 ? n : 1) / ulp);

  if (make_real_subdiag) {
#pragma unroll
    for (int i = 1; i < NB; ++i) {
      cx<R> h = (i < n) ? ld<R>(g.at(i, i - 1)) : mk<R>(0, 0);
      const bool doit = i < n && h.im != R(0);
      cx<R> sc = mk<R>(1, 0);
      if (doit) {
        const R d1 = cabs1(h);
        cx<R> w = scale(h, frcp(d1));
        sc = scale(conj(w), frsq(abs2(w)));  // conj(h/|h|)
      }
      wave_sync();
      if (doit) {
        if (g.r == 0) st<R>(g.at(i, i - 1), mk<R>(fsqrt(abs2(h)), 0));
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int x = g.idx(p);
          if (x >= i && x < n) st<R>(g.at(i, x), mul(ld<R>(g.at(i, x)), sc));
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int x = g.idx(p);
          if (x <= i + 1 && x < n) st<R>(g.at(x, i), mulc(ld<R>(g.at(x, i)), sc));
        }
        if constexpr (TZ) {
#pragma unroll
          for (int p = 0; p < P; ++p) z[p][i] = mulc(z[p][i], sc);
        }
      }
      wave_sync();
    }
  }

  // the identity input of the groups a sweep step passes by (sweep_step)
  if (TZ && MPW > 1 && g.r == 0) {
    st<R>(g.vsh + 2, mk<R>(1, 0));
    st<R>(g.vsh + 3, mk<R>(0, 0));
  }
  int I = n - 1, L = 0, its = 0, nfail = 0;
  bool done = (n <= 0);
  const int base = g.mi * LG;
  constexpr uint64_t kMask = (LG == 64) ? ~0ull : ((1ull << LG) - 1);
  for (;;) {
    // ---- bookkeeping until every matrix either needs a sweep or is done ----
    // Every subdiagonal entry k = 1..I is tested at once.  The entries at or
    // below L are frozen and H(L, L-1) is an exact zero, so the highest
    // negligible k is what the reference's scan from I down to L+1 finds
    // (Toolbox.hpp:810-867), and one ballot serves every deflation it shows,
    // from I downwards, without testing again.
    bool ready = false;
    for (;;) {
      const bool need = !done && !ready;
      if (!__any(need)) break;
      uint64_t bal[P];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        bool neg = false;
        const int k = g.idx(p);
        if (need && its <= itmax && k > 0 && k <= I) {
          const cx<R> h = ld<R>(g.at(k, k - 1));
          const cx<R> hkk1 = ld<R>(g.at(k - 1, k - 1)), hkk = ld<R>(g.at(k, k));
          R tst = cabs1(hkk1) + cabs1(hkk);
          // the rare zero diagonal pair behind a wavefront-uniform test, and
          // a subdiagonal entry below smlnum (negligible outright) folded
          // into the result: fewer scalar instructions than the reference's
          // nested branches, the same outcome
          if (__builtin_expect(__any(tst == R(0)), 0)) {
            if (tst == R(0)) {
              if (k - 2 >= 0) tst += fabs(ld<R>(g.at(k - 1, k - 2)).re);
              if (k + 1 <= n - 1) tst += fabs(ld<R>(g.at(k + 1, k)).re);
            }
          }
          const bool tiny = cabs1(h) <= smlnum;
          if (!tiny && fabs(h.re) <= ulp * tst) {
            const cx<R> hup = ld<R>(g.at(k - 1, k));
            const R ab = fmax(cabs1(h), cabs1(hup)), ba = fmin(cabs1(h), cabs1(hup));
            const R d3 = cabs1(sub(hkk1, hkk));
            const R aa = fmax(cabs1(hkk), d3), bb = fmin(cabs1(hkk), d3);
            const R is = frcp(aa + ab);  // aa + ab >= cabs1(h) > 0
            neg = ba * (ab * is) <= fmax(smlnum, ulp * (bb * (aa * is)));
          }
          neg = neg || tiny;
        }
        bal[p] = __ballot(neg);
      }
      if (need) {
        if (its > itmax) {
          // accepted without convergence (the reference's silent case)
          ++nfail;
          --I;
          L = 0;
          its = 0;
          if (I < 0) done = true;
        } else {
          uint64_t neg = 0;  // bit k: H(k, k-1) negligible
#pragma unroll
          for (int p = 0; p < P; ++p) neg |= ((bal[p] >> base) & kMask) << (p * LG);
          for (;;) {
            const uint64_t cand = I >= 63 ? neg : (neg & ((2ull << I) - 1));
            L = cand ? 63 - __clzll(cand) : 0;
            if (L > 0 && g.r == 0) st<R>(g.at(L, L - 1), mk<R>(0, 0));
            if (L < I) {
              ready = true;
              break;
            }
            --I;  // H(I, I) has deflated
            its = 0;
            if (I < 0) {
              done = true;
              break;
            }
          }
        }
      }
      wave_sync();
    }
    if (!__any(ready)) break;

    // ---- shift and first column (Toolbox.hpp:896-1015) ----
    // The Wilkinson shift t = H(I,I) - u^2 / (x + y), x = (H(I-1,I-1) -
    // H(I,I)) / 2, y = +-sqrt(x^2 + u^2), needs u only through u^2 =
    // H(I-1,I) H(I,I-1) (the reference forms u = sqrt(H(I-1,I)) sqrt(H(I,I-1))
    // and squares it again): one complex square root.  The reference scales
    // x and u by s = max(|u|, |x|) (cabs1); the power of two 2^E >= s used
    // here scales exactly and needs no square root or reciprocal, and so
    // does 2^E1 for the first column (the reference divides it by its cabs1
    // norm; the rotation normalises it anyway).
    cx<R> a0 = mk<R>(0, 0);
    R a1 = R(0);
    if (ready) {
      cx<R> t;
      // the Wilkinson shift t = H(I,I) - u^2 / (x + y)
      auto wilkinson = [&]() {
        cx<R> tw = ld<R>(g.at(I, I));
        const cx<R> u2 = mul(ld<R>(g.at(I - 1, I)), ld<R>(g.at(I, I - 1)));
        if (u2.re != R(0) || u2.im != R(0)) {
          const cx<R> x = scale(sub(ld<R>(g.at(I - 1, I - 1)), tw), R(0.5));
          const R sx = cabs1(x);
          // 2^E > max(|u|, |x|): |x / 2^E| < 1 and |u^2 / 2^2E| <= 1
          const int eu = fexp(cabs1(u2)), ex = sx > R(0) ? fexp(sx) : -4096;
          const int E = max((eu + 1) >> 1, ex);
          const cx<R> xs = cldexp(x, -E);
          cx<R> ys = csqrt_unit(add(mul(xs, xs), cldexp(u2, -2 * E)));  // y / 2^E
          if (sx > R(0) && xs.re * ys.re + xs.im * ys.im < R(0)) ys = mk(-ys.re, -ys.im);
          const cx<R> d = add(xs, ys);
          const R id = frcp1(abs2(d));
          tw = sub(tw, mul(cldexp(u2, -E), mk(d.re * id, -d.im * id)));
        }
        return tw;
      };
      // the exceptional shifts of iterations 10 and 20 behind a
      // wavefront-uniform test, so that the usual sweep pays for no branch
      // around them
      if (!__builtin_expect(__any(its == 10 || its == 20), 0)) {
        t = wilkinson();
      } else if (its == 10) {
        const R s = R(0.75) * fabs(ld<R>(g.at(L + 1, L)).re);
        t = ld<R>(g.at(L, L));
        t.re = s + t.re;
      } else if (its == 20) {
        const R s = R(0.75) * fabs(ld<R>(g.at(I, I - 1)).re);
        t = ld<R>(g.at(I, I));
        t.re = s + t.re;
      } else {
        t = wilkinson();
      }
      const cx<R> h11s = sub(ld<R>(g.at(L, L)), t);
      const R h21 = ld<R>(g.at(L + 1, L)).re;
      const int E1 = fexp(cabs1(h11s) + fabs(h21));
      a0 = cldexp(h11s, -E1);
      a1 = fldexp(h21, -E1);
      // the first step's reflector input, where sweep_step reads it: column
      // L-1, rows L and L+1 (H(L, L-1) is a deflated zero; both are restored
      // after the sweep), or vsh when L = 0
      if (g.r == 0) {
        typename Group<R, NB, P, LAY, LW>::L2* sp = L > 0 ? g.at(L, L > 0 ? L - 1 : 0) : g.vsh;
        st<R>(sp, a0);
        st<R>(sp + 1, mk<R>(a1, 0));
      }
    }

    // ---- wavefront-uniform range of bulge positions ----
    int kmin = NB, kmax = 0;
    // one readlane per group: (L, I) packed, -1 when the group is not ready
    const int packed = ready ? ((L << 8) | I) : -1;
#pragma unroll
    for (int q = 0; q < MPW; ++q) {
      const int pq = __builtin_amdgcn_readlane(packed, q * LG);
      if (pq >= 0) {
        const int lq = pq >> 8, iq = pq & 255;
        kmin = lq < kmin ? lq : kmin;
        kmax = iq > kmax ? iq : kmax;
      }
    }
    wave_sync();

    // ---- the sweep (Toolbox.hpp:138-335), K = bulge row ----
    if constexpr (MPW == 1) {
      // one matrix per wavefront: a runtime loop over K, with the Schur
      // vector columns selected by a uniform branch tree
      static_assert(P == 1, "one matrix per wavefront uses one index per lane");
      for (int K = kmin; K < kmax; ++K) sweep_step<-1, TZ>(g, z, ready && K >= L && K < I, L, I, K);
    } else {
      // steps kmin..kmax-1 of the wavefront, L..I-1 of this group
      static_assert(NB <= 32, "step masks are 32 bits");
      const unsigned kmask = kmax > kmin ? ((1u << kmax) - 1u) & ~((1u << kmin) - 1u) : 0u;
      const unsigned amask = ready ? ((1u << I) - 1u) & ~((1u << L) - 1u) : 0u;
      sweep_range<0, NB - 1, TZ>(g, z, amask, L, I, kmask);
    }
    if (ready && L > 0 && g.r == 0) {  // the staged first column
      st<R>(g.at(L, L - 1), mk<R>(0, 0));
      st<R>(g.at(L + 1, L - 1), mk<R>(0, 0));
    }

    // ---- make H(I, I-1) real (Toolbox.hpp:337-367) ----
    if (ready) {
      const cx<R> h = ld<R>(g.at(I, I - 1));
      if (h.im != R(0) && abs2(h) > R(0)) {  // |h|^2 may underflow: leave such h alone
        const R id = frsq(abs2(h));
        const R d = abs2(h) * id;
        const cx<R> zz = mk(h.re * id, -h.im * id);  // conj(h/|h|)
        if (g.r == 0) st<R>(g.at(I, I - 1), mk<R>(d, 0));
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int x = g.idx(p);
          if (TZ && x > I && x < n) st<R>(g.at(I, x), mul(ld<R>(g.at(I, x)), zz));
          if (x < I && (TZ || x >= L)) st<R>(g.at(x, I), mulc(ld<R>(g.at(x, I)), zz));
        }
        if constexpr (!TZ) {
        } else if constexpr (MPW == 1) {
          uniform_dispatch<0, NB>(I, [&](auto kc) {
            constexpr int KK = decltype(kc)::value;
            z[0][KK] = mulc(z[0][KK], zz);
          });
        } else {
#pragma unroll
          for (int c = 1; c < NB; ++c)
            if (c == I) {
              if constexpr (kPinnedRow<R, NB, P, LAY, MPW>) {
                static_for<NB>([&](auto cc) {
                  constexpr int C = decltype(cc)::value;
                  if (C == c) ZPinMulc<C>::mul(zz.re, zz.im, z[0][C].re, z[0][C].im);
                });
              } else {
#pragma unroll
                for (int p = 0; p < P; ++p) z[p][c] = mulc(z[p][c], zz);
              }
            }
        }
      }
      ++its;
    }
    wave_sync();
  }
  return nfail;
}

// --------------------------------------------------------------------------
// Right eigenvectors of the upper-triangular T in LDS (ztrevc +
// backsubstitution, Toolbox.hpp:1133-1313, 31-136).  The lane owning index k
// solves (T(0:k-1,0:k-1) - T(k,k) I) x = -T(0:k-1,k) with the reference's
// small-pivot clamp (Toolbox.hpp:1235-1258); the unit upper triangular X
// (column k = [x; 1; 0]) is written over the strict upper triangle of T.
// Slots are solved from the last to the first: slot p only reads columns
// below (p+1) LG, which later slots have not yet overwritten.
// Returns lam[p] = T(k,k) for k = idx(p).
// --------------------------------------------------------------------------
template <typename R, int NB, int P>
__device__ __forceinline__ void trevc_solve(const Group<R, NB, P>& g, int nmax, cx<R> (&lam)[P]) {
  constexpr int LG = Group<R, NB, P>::LG;
  const int n = g.n;
  const R ulp = limits<R>::eps;
  const R smlnum = limits<R>::tiny * (R(n > 0 ? n : 1) / ulp);
#pragma unroll
  for (int p = P - 1; p >= 0; --p) {
    const int k = g.idx(p);
    const int kcap = (p + 1) * LG;  // k < kcap in this slot
    const bool in = k < n;
    lam[p] = in ? ld<R>(g.at(k, k)) : mk<R>(0, 0);
    const R smin = fmax(ulp * cabs1(lam[p]), smlnum);
    cx<R> x[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const bool a = j < kcap && in && j < k;
      x[j] = a ? ld<R>(g.at(j, k)) : mk<R>(0, 0);
      x[j] = mk(-x[j].re, -x[j].im);
    }
#pragma unroll
    for (int j = NB - 2; j >= 0; --j) {
      if (j + 1 < kcap && j + 1 < nmax) {
        const bool act = in && j < k;
        cx<R> d = sub(ld<R>(g.at(j, j)), lam[p]);
        if (cabs1(d) < smin) d = mk<R>(smin, 0);
        const cx<R> xj = mul(x[j], inv(d));
        if (act) x[j] = xj;
#pragma unroll
        for (int i = 0; i < j; ++i) {
          const cx<R> tij = ld<R>(g.at(i, j));
          if (act) x[i] = sub(x[i], mul(tij, xj));
        }
      }
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < kcap && in && j < k) st<R>(g.at(j, k), x[j]);
    wave_sync();
  }
}

// --------------------------------------------------------------------------
// Y = X^{-1} for the unit upper triangular X left in LDS by trevc_solve: the
// lane owning index k forms column k of Y by back substitution and stores its
// strict upper part transposed into the (unused) strict lower triangle of the
// tile (Y(j,k), j < k, at tile position (k, j)).
// --------------------------------------------------------------------------
template <typename R, int NB, int P>
__device__ __forceinline__ void inverse_unit_upper(const Group<R, NB, P>& g, int nmax) {
  constexpr int LG = Group<R, NB, P>::LG;
  const int n = g.n;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = g.idx(p);
    const int kcap = (p + 1) * LG;
    const bool in = k < n;
    cx<R> y[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) y[j] = mk<R>(j == k ? R(1) : R(0), 0);
    // y_j = -sum_{i=j+1..k} X(j,i) y_i, j = k-1 .. 0
#pragma unroll
    for (int j = NB - 2; j >= 0; --j) {
      if (j + 1 < kcap && j + 1 < nmax) {
        cx<R> acc = mk<R>(0, 0);
#pragma unroll
        for (int i = j + 1; i < NB; ++i)
          if (i < kcap) acc = add(acc, mul(ld<R>(g.at(j, i)), y[i]));
        if (in && j < k) y[j] = mk(-acc.re, -acc.im);
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < kcap && in && j < k) st<R>(g.at(k, j), y[j]);
  }
  wave_sync();
}

// --------------------------------------------------------------------------
// Approximate inverse of the eigenvector matrix, W = X^{-1} Z^H = Y Z^H,
// from Y (inverse_unit_upper) and the Schur-vector rows z[p][] of each lane:
// for each owned row r of Z it produces column r of W, passed as
// out(p, k, W(k, r)).
// --------------------------------------------------------------------------
template <typename R, int NB, int P, typename Out>
__device__ __forceinline__ void inverse_vectors(const Group<R, NB, P>& g, const cx<R> (&z)[P][NB],
                                                int nmax, Out&& out) {
  inverse_unit_upper<R, NB, P>(g, nmax);
  // W(k, r) = conj(z[k]) + sum_{j>k} Y(k,j) conj(z[j])
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    cx<R> acc[P];
#pragma unroll
    for (int p = 0; p < P; ++p) acc[p] = conj(z[p][k]);
    if (k < nmax) {
#pragma unroll
      for (int j = k + 1; j < NB; ++j) {
        const cx<R> yk = ld<R>(g.at(j, k));
#pragma unroll
        for (int p = 0; p < P; ++p) acc[p] = add(acc[p], mulc(yk, z[p][j]));
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) out(p, k, acc[p]);
  }
}

// --------------------------------------------------------------------------
// Back-transform (the zgemv of ztrevc, Toolbox.hpp:1272-1275): rows idx(p)
// of V = Z * X, in place in z[][] (column k of V = Z [x; 1; 0], the
// reference's scaling, not normalised).
// --------------------------------------------------------------------------
template <typename R, int NB, int P>
__device__ __forceinline__ void backtransform(const Group<R, NB, P>& g, cx<R> (&z)[P][NB], int nmax) {
#pragma unroll
  for (int k = NB - 1; k >= 1; --k) {
    if (k < nmax) {
      cx<R> acc[P];
#pragma unroll
      for (int p = 0; p < P; ++p) acc[p] = z[p][k];
#pragma unroll
      for (int j = 0; j < k; ++j) {
        const cx<R> xjk = ld<R>(g.at(j, k));
#pragma unroll
        for (int p = 0; p < P; ++p) acc[p] = add(acc[p], mul(z[p][j], xjk));
      }
#pragma unroll
      for (int p = 0; p < P; ++p) z[p][k] = acc[p];
    }
  }
}

// --------------------------------------------------------------------------
// The same eigenvector solve for one matrix per wavefront (n up to 64, one
// index per lane), with the right-hand sides kept in LDS instead of
// registers: lane k keeps y = -x of its eigenvector in column k of the tile
// (rows < k).  At step j lanes k > j read column j of T, which lane j only
// starts to overwrite at step j-1, so T and the solutions share the tile.
// --------------------------------------------------------------------------
template <typename R, int NB, int LAY = LAY_DENSE, int LW = 0>
__device__ __forceinline__ cx<R> trevc_solve_lds(const Group<R, NB, 1, LAY, LW>& g, int nmax) {
  const int r = g.r, n = g.n;
  const R ulp = limits<R>::eps;
  const R smlnum = limits<R>::tiny * (R(n > 0 ? n : 1) / ulp);
  const bool in = r < n;
  const cx<R> lam = in ? ld<R>(g.at(r, r)) : mk<R>(0, 0);
  const R smin = fmax(ulp * cabs1(lam), smlnum);
  for (int j = nmax - 2; j >= 0; --j) {
    const bool act = in && j < r;
    cx<R> d = sub(ld<R>(g.at(j, j)), lam);
    if (cabs1(d) < smin) d = mk<R>(smin, 0);
    cx<R> yj = mk<R>(0, 0);
    if (act) {
      yj = mul(ld<R>(g.at(j, r)), inv(d));
      st<R>(g.at(j, r), yj);
    }
    for (int i = 0; i < j; ++i) {
      const cx<R> tij = ld<R>(g.at(i, j));
      if (act) st<R>(g.at(i, r), sub(ld<R>(g.at(i, r)), mul(tij, yj)));
    }
    wave_sync();
  }
  return lam;
}

// Row r of V = Z X from the tile left by trevc_solve_lds, one column at a
// time, stored straight to the output (column k of V = Z [x; 1; 0]).
template <typename R, int NB, typename Store>
__device__ __forceinline__ void backtransform_lds(const Group<R, NB, 1>& g, const cx<R> (&z)[1][NB],
                                                  int nmax, Store&& store) {
  const int r = g.r;
  // tile column r becomes -X(:, r): y above the diagonal, -1 on it, 0 below
  if (r < g.n) {
    st<R>(g.at(r, r), mk<R>(-1, 0));
    for (int i = r + 1; i < g.n; ++i) st<R>(g.at(i, r), mk<R>(0, 0));
  }
  wave_sync();
  for (int k = 0; k < nmax; ++k) {
    cx<R> acc = mk<R>(0, 0);
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < nmax) acc = add(acc, mul(z[0][j], ld<R>(g.at(j, k))));
    store(k, mk(-acc.re, -acc.im));
  }
}

template <typename R, int NB, int P>
__device__ __forceinline__ void trevc_backtransform(const Group<R, NB, P>& g, cx<R> (&z)[P][NB],
                                                    cx<R> (&lam)[P], int nmax) {
  trevc_solve<R, NB, P>(g, nmax, lam);
  backtransform<R, NB, P>(g, z, nmax);
}

}  // namespace veig
