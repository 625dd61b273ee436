// FP64 Newton refinement of a batch of eigendecompositions on the FP64
// matrix cores of gfx950 (v_mfma_f64_16x16x4f64).
//
// One wavefront refines one matrix of order n <= 16*NT, held as NT x NT
// blocks of 16 x 16.  With R = A V - V diag(lam) and an approximate inverse
// W ~ V^{-1}, one Newton step is
//     E = W R,   dlam_k = E_kk,   C_ik = E_ik / (lam_k - lam_i) (i != k),
//     V <- V + V C,   lam <- lam + dlam,
// (Dongarra, Moler & Wilkinson, "Improving the accuracy of computed
// eigenvalues and eigenvectors", SIAM J. Numer. Anal. 20 (1983)), i.e. three
// complex matrix products, each 16 MFMA per 16 x 16 block product (four real
// products of four k-steps).  The backward error ||R||_F / (||A||_F ||V||_F)
// is both the stopping test and the certificate returned to the caller.
//
// Register layouts of a 16 x 16 block on the 64 lanes (lane l, c = l % 16,
// g = l / 16, slot s = 0..3 of a 4-double vector):
//   A operand       X(c,          4s + g)
//   B operand       X(4s + g,     c)
//   result (C, D)   X(4s + g,     c)      -- the B layout
// (checked on gfx950 by a one-hot probe of every A lane).  A product can
// therefore feed the next one as its B operand directly; only V as the A
// operand of V C needs a transpose, which goes through an LDS tile.
#pragma once
#include "veig_common.hpp"

namespace veig {

typedef double d4 __attribute__((ext_vector_type(4)));

enum { LAY_A = 0, LAY_B = 1 };

template <int NT>
struct CMat {
  d4 re[NT][NT], im[NT][NT];
};

// (row, col) held by lane (c, g) in slot s of block (I, J) of layout LAY
template <int LAY>
__device__ __forceinline__ void lay_rc(int I, int J, int s, int c, int g, int &row, int &col) {
  if constexpr (LAY == LAY_A) {
    row = 16 * I + c;
    col = 16 * J + 4 * s + g;
  } else {
    row = 16 * I + 4 * s + g;
    col = 16 * J + c;
  }
}

template <int NT>
__device__ __forceinline__ void cm_zero(CMat<NT> &x) {
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J) {
      x.re[I][J] = d4{0, 0, 0, 0};
      x.im[I][J] = d4{0, 0, 0, 0};
    }
}

template <int NT>
struct RefTile {
  static constexpr int N = 16 * NT;
  static constexpr int LD = N + 1;
  double2 t[N * LD];  // X(row, col) at t[col * LD + row]
  double2 dl[N];      // eigenvalue corrections
};

template <int NT, int LAY>
__device__ __forceinline__ void cm_store(RefTile<NT> &sh, const CMat<NT> &x, int c, int g) {
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        int row, col;
        lay_rc<LAY>(I, J, s, c, g, row, col);
        sh.t[col * RefTile<NT>::LD + row] = double2{x.re[I][J][s], x.im[I][J][s]};
      }
}

template <int NT, int LAY>
__device__ __forceinline__ void cm_load(const RefTile<NT> &sh, CMat<NT> &x, int c, int g) {
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        int row, col;
        lay_rc<LAY>(I, J, s, c, g, row, col);
        const double2 v = sh.t[col * RefTile<NT>::LD + row];
        x.re[I][J][s] = v.x;
        x.im[I][J][s] = v.y;
      }
}

// d += a * b  (a in A layout, b and d in B layout).  One block: the
// three-product form of the complex product, P1 = ar br, P2 = ai bi,
// P3 = (ar + ai)(br + bi), re += P1 - P2, im += P3 - P1 - P2 (12 MFMAs
// instead of 16; the error bound grows by at most a factor of about two),
// the three products in independent chains.
template <int NT>
__device__ __forceinline__ void cm_gemm(const CMat<NT> &a, const CMat<NT> &b, CMat<NT> &d) {
  if constexpr (NT > 1) {
    // 2 x 2 blocks: registers are short; four products in two chains
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int K = 0; K < NT; ++K)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double ar = a.re[I][K][s], ai = a.im[I][K][s];
            const double br = b.re[K][J][s], bi = b.im[K][J][s];
            d.re[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, d.re[I][J], 0, 0, 0);
            d.im[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, d.im[I][J], 0, 0, 0);
            d.re[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, d.re[I][J], 0, 0, 0);
            d.im[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, d.im[I][J], 0, 0, 0);
          }
    return;
  }
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J) {
      d4 p1 = d4{0, 0, 0, 0}, p2 = d4{0, 0, 0, 0}, p3 = d4{0, 0, 0, 0};
#pragma unroll
      for (int K = 0; K < NT; ++K)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const double ar = a.re[I][K][s], ai = a.im[I][K][s];
          const double br = b.re[K][J][s], bi = b.im[K][J][s];
          p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, p1, 0, 0, 0);
          p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, p2, 0, 0, 0);
          p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, p3, 0, 0, 0);
        }
      d.re[I][J] += p1 - p2;
      d.im[I][J] += p3 - p1 - p2;
    }
}

// DPP move of a double (both halves with the same control)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)__double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)__double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_f64(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane((int)__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane((int)__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront, returned in every lane: DPP sums
// within each row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror), then the four row sums are read as scalars.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return (lane_f64(x, 0) + lane_f64(x, 16)) + (lane_f64(x, 32) + lane_f64(x, 48));
}

// load a matrix of the tiled batch layout (re/im planes of type T) into
// layout LAY, zero outside the n x n matrix
template <int NT, int LAY, typename T>
__device__ __forceinline__ void cm_load_global(CMat<NT> &x, const Geom &G, int64_t base,
                                               const T *re, const T *im, int c, int g) {
  const int n = G.n;
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        int row, col;
        lay_rc<LAY>(I, J, s, c, g, row, col);
        double vr = 0, vi = 0;
        if (row < n && col < n) {
          const int64_t idx = base + G.el(row, col);
          vr = (double)re[idx];
          vi = (double)im[idx];
        }
        x.re[I][J][s] = vr;
        x.im[I][J][s] = vi;
      }
}

struct RefineArgs {
  const double *rA, *iA;
  double *rE, *iE, *rV, *iV;
  const float *rW, *iW;
  int32_t *flags;           // out: 1 = needs the FP64 solve
  double *berr;             // out (may be NULL): final backward error
  const int32_t *info32;    // in (may be NULL): complex64 QR failures
  const int32_t *mask;      // if set, only matrices with mask[m] != 0
  int32_t *info;            // out (may be NULL): 0 for matrices refined here
  Geom G;
  double tol;
  int maxit;                // 0 = measure the backward error only
  // if set, only the matrices list[0 .. *list_count) (persistent grid)
  const int32_t *list = nullptr;
  const int32_t *list_count = nullptr;
  // k_refine_zxy: the complex64 factors (Z in rW/iW; X above, X^{-1} below
  // the diagonal in rX/iX) from which V and W are formed
  const float *rX = nullptr, *iX = nullptr;
};

// Newton iterations on one matrix held by the whole wavefront: V (B layout)
// and the eigenvalues of its columns (lc) and rows (lr) are updated in place;
// returns the final backward error ||A V - V Lam||_F / (||A||_F ||V||_F).
// load_a / load_w fill A and W in the A layout where they are needed;
// to_a(V, X) transposes V into the A layout (through LDS); dl is LDS scratch
// for the 16*NT eigenvalue corrections.
template <int NT, typename LoadA, typename LoadW, typename ToA>
__device__ __forceinline__ double newton_refine(CMat<NT> &V, cx<double> (&lc)[NT], cx<double> (&lr)[NT][4],
                                                int n, double nrmA, double tol, int maxit, bool qr_ok,
                                                double2 *dl, LoadA &&load_a, LoadW &&load_w, ToA &&to_a) {
  const int l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  double be = INFINITY, be_prev = INFINITY;
  for (int it = 0;; ++it) {
    // ---- R = A V - V diag(lam), backward error ----
    CMat<NT> Rm;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const double vr = V.re[I][J][s], vi = V.im[I][J][s];
          Rm.re[I][J][s] = -(vr * lc[J].re - vi * lc[J].im);
          Rm.im[I][J][s] = -(vr * lc[J].im + vi * lc[J].re);
        }
    {
      CMat<NT> A;
      load_a(A);
      cm_gemm<NT>(A, V, Rm);
    }
    double rn2 = 0, vn2 = 0;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          rn2 += Rm.re[I][J][s] * Rm.re[I][J][s] + Rm.im[I][J][s] * Rm.im[I][J][s];
          vn2 += V.re[I][J][s] * V.re[I][J][s] + V.im[I][J][s] * V.im[I][J][s];
        }
    rn2 = wave_sum(rn2);
    vn2 = wave_sum(vn2);
    be = sqrt(rn2) / (nrmA * sqrt(vn2));
    if (n == 0) be = 0;
    // wavefront-uniform stop: converged, out of iterations, or contracting
    // too slowly to reach tol in the iterations left at the observed rate
    // (the complex64 start is outside Newton's fast basin: the matrix goes
    // to the FP64 solve instead)
    if (!qr_ok || be <= tol || it >= maxit || !(be < be_prev)) break;
    if (it > 0) {
      const double rate = be / be_prev;
      double reach = be;
      for (int k = it; k < maxit; ++k) reach *= rate;
      if (reach > tol) break;
    }
    be_prev = be;

    // ---- E = W R ----
    CMat<NT> Em;
    cm_zero<NT>(Em);
    {
      CMat<NT> W;
      load_w(W);
      cm_gemm<NT>(W, Rm, Em);
    }
    // ---- dlam = diag(E); C = E / (lam_k - lam_i) off the diagonal ----
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (4 * s + g == c) dl[16 * I + c] = double2{Em.re[I][I][s], Em.im[I][I][s]};
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = 16 * I + 4 * s + g, col = 16 * J + c;
          const cx<double> d = sub(lc[J], lr[I][s]);
          const double d2 = abs2(d);
          const bool zero = row == col || row >= n || col >= n || !(d2 > 0.0);
          // E / d = E conj(d) / |d|^2 (|d|^2 underflowing: no correction)
          const cx<double> q = zero ? mk<double>(0, 0)
                                    : scale(mulc(mk<double>(Em.re[I][J][s], Em.im[I][J][s]), d), frcp(d2));
          Em.re[I][J][s] = q.re;
          Em.im[I][J][s] = q.im;
        }
    // ---- V <- V + V C ----
    {
      CMat<NT> X;
      to_a(V, X);  // includes the wavefront syncs around the LDS round trip
      cm_gemm<NT>(X, Em, V);
    }
#pragma unroll
    for (int J = 0; J < NT; ++J) {
      const double2 d = dl[16 * J + c];
      lc[J] = add(lc[J], mk<double>(d.x, d.y));
    }
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double2 d = dl[16 * I + 4 * s + g];
        lr[I][s] = add(lr[I][s], mk<double>(d.x, d.y));
      }
    wave_sync();
  }
  return be;
}

template <int NT>
__device__ __forceinline__ double cm_norm2(const CMat<NT> &x) {
  double a = 0;
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J)
#pragma unroll
      for (int s = 0; s < 4; ++s) a += x.re[I][J][s] * x.re[I][J][s] + x.im[I][J][s] * x.im[I][J][s];
  return wave_sum(a);
}

// V (B layout) and eigenvalues of one refined matrix to the batch layout
template <int NT>
__device__ __forceinline__ void store_refined(const Geom &G, int64_t base, int64_t vbase, double *rV, double *iV,
                                              double *rE, double *iE, const CMat<NT> &V,
                                              const cx<double> (&lc)[NT]) {
  const int l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  const int n = G.n;
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int row = 16 * I + 4 * s + g, col = 16 * J + c;
        if (row < n && col < n) {
          const int64_t idx = base + G.el(row, col);
          rV[idx] = V.re[I][J][s];
          iV[idx] = V.im[I][J][s];
        }
      }
  if (g == 0) {
#pragma unroll
    for (int J = 0; J < NT; ++J) {
      const int k = 16 * J + c;
      if (k < n) {
        rE[vbase + G.ve(k)] = lc[J].re;
        iE[vbase + G.ve(k)] = lc[J].im;
      }
    }
  }
}

// one matrix of the batch: V, E (and W) from global memory
template <int NT>
__device__ __forceinline__ void refine_one(const RefineArgs &a, RefTile<NT> &sh, int64_t m) {
  const int l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  const int n = a.G.n;
  const int64_t base = a.G.mat_base(m), vbase = a.G.vec_base(m);
  // A and W stay in registers for one block; for larger orders they are
  // re-read (from L2) where they are used, to stay within 256 VGPRs.
  constexpr bool kHold = NT == 1;
  CMat<NT> A, V, W;
  cm_load_global<NT, LAY_A>(A, a.G, base, a.rA, a.iA, c, g);
  cm_load_global<NT, LAY_B>(V, a.G, base, a.rV, a.iV, c, g);
  if (kHold && a.maxit > 0) cm_load_global<NT, LAY_A>(W, a.G, base, a.rW, a.iW, c, g);
  const double nrmA = sqrt(cm_norm2<NT>(A));
  // eigenvalues of the B-layout columns (lc) and rows (lr)
  cx<double> lc[NT], lr[NT][4];
#pragma unroll
  for (int J = 0; J < NT; ++J) {
    const int k = 16 * J + c;
    lc[J] = k < n ? mk<double>(a.rE[vbase + a.G.ve(k)], a.iE[vbase + a.G.ve(k)]) : mk<double>(0, 0);
  }
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 16 * I + 4 * s + g;
      lr[I][s] = k < n ? mk<double>(a.rE[vbase + a.G.ve(k)], a.iE[vbase + a.G.ve(k)]) : mk<double>(0, 0);
    }
  const bool qr_ok = !(a.info32 && a.info32[m] != 0);
  const double be = newton_refine<NT>(
      V, lc, lr, n, nrmA, a.tol, a.maxit, qr_ok, sh.dl,
      [&](CMat<NT> &x) {
        if constexpr (kHold) {
          x = A;
        } else {
          int64_t b = base;
          asm volatile("" : "+v"(b));  // opaque: keep the reload inside the loop
          cm_load_global<NT, LAY_A>(x, a.G, b, a.rA, a.iA, c, g);
        }
      },
      [&](CMat<NT> &x) {
        if constexpr (kHold) {
          x = W;
        } else {
          int64_t b = base;
          asm volatile("" : "+v"(b));
          cm_load_global<NT, LAY_A>(x, a.G, b, a.rW, a.iW, c, g);
        }
      },
      [&](const CMat<NT> &v, CMat<NT> &x) {
        cm_store<NT, LAY_B>(sh, v, c, g);
        wave_sync();
        cm_load<NT, LAY_A>(sh, x, c, g);
      });
  const bool ok = (be <= a.tol) && qr_ok;
  if (a.maxit > 0) {
    store_refined<NT>(a.G, base, vbase, a.rV, a.iV, a.rE, a.iE, V, lc);
    if (l == 0) {
      a.flags[m] = ok ? 0 : 1;
      if (a.info && ok) a.info[m] = 0;
    }
  }
  if (l == 0 && a.berr) a.berr[m] = be;
}

// Workgroups of kRefWaves wavefronts refine kRefWaves consecutive matrices:
// in the tiled layout the same element of consecutive matrices shares a
// cache line, so the wavefronts of one workgroup (one CU) reuse each line
// from L1 instead of each issuing its own L2 request per element.
constexpr int kRefWaves = 4;

template <int NT>
__global__ __launch_bounds__(64 * kRefWaves) void k_refine_mfma(RefineArgs a) {
  __shared__ RefTile<NT> sh[kRefWaves];
  const int w = threadIdx.x / 64;
  if (a.list) {
    const int64_t cnt = *a.list_count;
    for (int64_t i = (int64_t)blockIdx.x * kRefWaves + w; i < cnt; i += (int64_t)gridDim.x * kRefWaves) {
      refine_one<NT>(a, sh[w], a.list[i]);
      wave_sync();
    }
    return;
  }
  const int64_t m = xcd_block() * kRefWaves + w;
  if (m >= a.G.nmat) return;
  if (a.mask && a.mask[m] == 0) return;
  refine_one<NT>(a, sh[w], m);
}

// ---------------------------------------------------------------------------
// Orders up to 16, whole batch: workgroups of kStageWaves wavefronts refine
// kStageWaves consecutive matrices and move A, V, W in and V out through LDS
// with a coalesced pattern (consecutive threads take the same element of
// consecutive matrices, which the tiled layout keeps in one cache segment),
// instead of one scattered 8-byte access per lane and element.
// ---------------------------------------------------------------------------
constexpr int kStageWaves = 4;
constexpr int kStageLd = 16 * 16 + 1;  // per-matrix slice (+1: spreads the slices over the LDS banks)

struct StagedShared {
  double2 st[kStageWaves * kStageLd];
  double2 dl[kStageWaves][16];
};

// element e (column-major) of matrix j of the workgroup -> slice j
template <typename T>
__device__ __forceinline__ void stage_in(StagedShared &sh, const Geom &G, int64_t mbase, bool have, const T *re,
                                         const T *im) {
  const int n = G.n, nn = n * n;
  const int j = threadIdx.x % kStageWaves;
  for (int e = threadIdx.x / kStageWaves; e < nn; e += blockDim.x / kStageWaves) {
    const int row = e % n, col = e / n;
    double2 v = double2{0, 0};
    if (have) {
      const int64_t idx = mbase + G.el(row, col);
      v = double2{(double)re[idx], (double)im[idx]};
    }
    sh.st[j * kStageLd + col * 16 + row] = v;
  }
}

__device__ __forceinline__ void stage_out(const StagedShared &sh, const Geom &G, int64_t mbase, bool have, double *re,
                                          double *im) {
  const int n = G.n, nn = n * n;
  const int j = threadIdx.x % kStageWaves;
  for (int e = threadIdx.x / kStageWaves; e < nn; e += blockDim.x / kStageWaves) {
    const int row = e % n, col = e / n;
    if (have) {
      const int64_t idx = mbase + G.el(row, col);
      const double2 v = sh.st[j * kStageLd + col * 16 + row];
      re[idx] = v.x;
      im[idx] = v.y;
    }
  }
}

// a wavefront's 16 x 16 operand from / to its slice (zero outside n on load)
template <int LAY>
__device__ __forceinline__ void slice_load(const double2 *sl, CMat<1> &x, int n, int c, int g) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    int row, col;
    lay_rc<LAY>(0, 0, s, c, g, row, col);
    double2 v = double2{0, 0};
    if (row < n && col < n) v = sl[col * 16 + row];
    x.re[0][0][s] = v.x;
    x.im[0][0][s] = v.y;
  }
}

template <int LAY>
__device__ __forceinline__ void slice_store(double2 *sl, const CMat<1> &x, int c, int g) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    int row, col;
    lay_rc<LAY>(0, 0, s, c, g, row, col);
    sl[col * 16 + row] = double2{x.re[0][0][s], x.im[0][0][s]};
  }
}

__global__ __launch_bounds__(64 * kStageWaves) void k_refine_staged(RefineArgs a) {
  __shared__ StagedShared sh;
  const int w = threadIdx.x / 64, l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  const int n = a.G.n;
  const int64_t m0 = xcd_block() * kStageWaves;
  // this thread's staging matrix (threadIdx % kStageWaves) and this wavefront's matrix (w)
  const int js = threadIdx.x % kStageWaves;
  const bool have_s = m0 + js < a.G.nmat;
  const int64_t base_s = have_s ? a.G.mat_base(m0 + js) : 0;
  const int64_t m = m0 + w;
  const bool have = m < a.G.nmat;
  double2 *sl = sh.st + w * kStageLd;

  CMat<1> A, V, W;
  stage_in(sh, a.G, base_s, have_s, a.rA, a.iA);
  __syncthreads();
  slice_load<LAY_A>(sl, A, n, c, g);
  __syncthreads();
  stage_in(sh, a.G, base_s, have_s, a.rV, a.iV);
  __syncthreads();
  slice_load<LAY_B>(sl, V, n, c, g);
  __syncthreads();
  if (a.maxit > 0) {
    stage_in(sh, a.G, base_s, have_s, a.rW, a.iW);
    __syncthreads();
    slice_load<LAY_A>(sl, W, n, c, g);
    __syncthreads();
  }
  double be = 0;
  bool ok = false;
  cx<double> lc[1], lr[1][4];
  if (have) {
    const int64_t vbase = a.G.vec_base(m);
    const double nrmA = sqrt(cm_norm2<1>(A));
    lc[0] = c < n ? mk<double>(a.rE[vbase + a.G.ve(c)], a.iE[vbase + a.G.ve(c)]) : mk<double>(0, 0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * s + g;
      lr[0][s] = k < n ? mk<double>(a.rE[vbase + a.G.ve(k)], a.iE[vbase + a.G.ve(k)]) : mk<double>(0, 0);
    }
    const bool qr_ok = !(a.info32 && a.info32[m] != 0);
    be = newton_refine<1>(
        V, lc, lr, n, nrmA, a.tol, a.maxit, qr_ok, sh.dl[w], [&](CMat<1> &x) { x = A; },
        [&](CMat<1> &x) { x = W; },
        [&](const CMat<1> &v, CMat<1> &x) {
          slice_store<LAY_B>(sl, v, c, g);
          wave_sync();
          slice_load<LAY_A>(sl, x, 16, c, g);
        });
    ok = (be <= a.tol) && qr_ok;
    wave_sync();
    slice_store<LAY_B>(sl, V, c, g);
  }
  __syncthreads();
  if (a.maxit > 0) {
    stage_out(sh, a.G, base_s, have_s, a.rV, a.iV);
    if (have) {
      const int64_t vbase = a.G.vec_base(m);
      if (g == 0 && c < n) {
        a.rE[vbase + a.G.ve(c)] = lc[0].re;
        a.iE[vbase + a.G.ve(c)] = lc[0].im;
      }
      if (l == 0) {
        a.flags[m] = ok ? 0 : 1;
        if (a.info && ok) a.info[m] = 0;
      }
    }
  }
  if (have && l == 0 && a.berr) a.berr[m] = be;
}

// ---------------------------------------------------------------------------
// Orders up to 16 from the complex64 factors: V = Z X and W = X^{-1} Z^H are
// formed on the matrix cores (FP64 products of complex64 data), then refined.
// A workgroup of kStageWaves wavefronts takes kStageWaves consecutive
// matrices; Z, X | X^{-1}, A and the eigenvalues come in through LDS in one
// coalesced phase (one memory latency per workgroup), V goes out the same way.
// ---------------------------------------------------------------------------
struct ZxyShared {
  double2 z[kStageWaves * kStageLd];   // Z; then per-wavefront transpose scratch
  double2 xy[kStageWaves * kStageLd];  // X | X^{-1}; then the refined V
  double2 a[kStageWaves * kStageLd];   // A
  double2 e[kStageWaves][16];          // eigenvalues
  double2 dl[kStageWaves][16];
};

__global__ __launch_bounds__(64 * kStageWaves) __attribute__((amdgpu_waves_per_eu(3))) void k_refine_zxy(RefineArgs a) {
  __shared__ ZxyShared sh;
  const int w = threadIdx.x / 64, l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  const int n = a.G.n, nn = n * n;
  const int64_t m0 = xcd_block() * kStageWaves;
  // ---- one staging phase: consecutive threads take consecutive matrices ----
  {
    const int js = threadIdx.x % kStageWaves;
    const bool have_s = m0 + js < a.G.nmat;
    const int64_t base_s = have_s ? a.G.mat_base(m0 + js) : 0;
    constexpr int kStep = 64;  // elements per pass: blockDim / kStageWaves
    constexpr int kPass = 256 / kStep;
    double2 vz[kPass], vx[kPass], va[kPass];
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const int e = threadIdx.x / kStageWaves + k * kStep;
      vz[k] = vx[k] = va[k] = double2{0, 0};
      if (have_s && e < nn) {
        const int64_t idx = base_s + a.G.el(e % n, e / n);
        vz[k] = double2{(double)a.rW[idx], (double)a.iW[idx]};
        vx[k] = double2{(double)a.rX[idx], (double)a.iX[idx]};
        va[k] = double2{a.rA[idx], a.iA[idx]};
      }
    }
    if (threadIdx.x < kStageWaves * 16) {
      const int r = threadIdx.x / kStageWaves;
      double2 ev = double2{0, 0};
      if (have_s && r < n) {
        const int64_t vi = a.G.vec_base(m0 + js) + a.G.ve(r);
        ev = double2{a.rE[vi], a.iE[vi]};
      }
      sh.e[js][r] = ev;
    }
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const int e = threadIdx.x / kStageWaves + k * kStep;
      if (e < nn) {
        const int o = js * kStageLd + (e / n) * 16 + e % n;
        sh.z[o] = vz[k];
        sh.xy[o] = vx[k];
        sh.a[o] = va[k];
      }
    }
  }
  __syncthreads();
  const int64_t m = m0 + w;
  const bool have = m < a.G.nmat;
  double2 *zs = sh.z + w * kStageLd, *xs = sh.xy + w * kStageLd, *as = sh.a + w * kStageLd;
  double be = 0;
  bool ok = false;
  cx<double> lc[1];
  if (have) {
    // ---- V = Z X, W = Y Z^H (Y = X^{-1}) ----
    CMat<1> V, W;
    {
      CMat<1> Za, Xb, Ya, Zh;
      slice_load<LAY_A>(zs, Za, n, c, g);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        {  // X in the B layout: unit upper triangular
          const int row = 4 * s + g, col = c;
          double2 v = double2{0, 0};
          if (row < n && col < n) v = row < col ? xs[col * 16 + row] : (row == col ? double2{1, 0} : v);
          Xb.re[0][0][s] = v.x;
          Xb.im[0][0][s] = v.y;
        }
        {  // Y in the A layout: Y(row, col), row < col, is stored at (col, row)
          const int row = c, col = 4 * s + g;
          double2 v = double2{0, 0};
          if (row < n && col < n) v = row < col ? xs[row * 16 + col] : (row == col ? double2{1, 0} : v);
          Ya.re[0][0][s] = v.x;
          Ya.im[0][0][s] = v.y;
        }
        Zh.re[0][0][s] = Za.re[0][0][s];  // Z^H in the B layout = conj(Z in the A layout)
        Zh.im[0][0][s] = -Za.im[0][0][s];
      }
      cm_zero<1>(V);
      cm_gemm<1>(Za, Xb, V);
      cm_zero<1>(W);
      cm_gemm<1>(Ya, Zh, W);
    }
    wave_sync();  // Z read: its slice becomes this wavefront's transpose scratch
    CMat<1> Wa;
    slice_store<LAY_B>(zs, W, c, g);
    wave_sync();
    slice_load<LAY_A>(zs, Wa, 16, c, g);
    // ---- Newton refinement ----
    CMat<1> A;
    slice_load<LAY_A>(as, A, n, c, g);
    const double nrmA = sqrt(cm_norm2<1>(A));
    cx<double> lr[1][4];
    {
      const double2 v = sh.e[w][c];
      lc[0] = mk<double>(v.x, v.y);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double2 v = sh.e[w][4 * s + g];
      lr[0][s] = mk<double>(v.x, v.y);
    }
    const bool qr_ok = !(a.info32 && a.info32[m] != 0);
    wave_sync();
    be = newton_refine<1>(
        V, lc, lr, n, nrmA, a.tol, a.maxit, qr_ok, sh.dl[w], [&](CMat<1> &x) { x = A; },
        [&](CMat<1> &x) { x = Wa; },
        [&](const CMat<1> &v, CMat<1> &x) {
          slice_store<LAY_B>(zs, v, c, g);
          wave_sync();
          slice_load<LAY_A>(zs, x, 16, c, g);
        });
    ok = (be <= a.tol) && qr_ok;
    slice_store<LAY_B>(xs, V, c, g);  // X | Y were read before the refinement
  }
  __syncthreads();
  {
    const int js = threadIdx.x % kStageWaves;
    const bool have_s = m0 + js < a.G.nmat;
    if (have_s) {
      const int64_t base_s = a.G.mat_base(m0 + js);
      for (int e = threadIdx.x / kStageWaves; e < nn; e += 64) {
        const double2 v = sh.xy[js * kStageLd + (e / n) * 16 + e % n];
        const int64_t idx = base_s + a.G.el(e % n, e / n);
        a.rV[idx] = v.x;
        a.iV[idx] = v.y;
      }
    }
  }
  if (have) {
    const int64_t vbase = a.G.vec_base(m);
    if (g == 0 && c < n) {
      a.rE[vbase + a.G.ve(c)] = lc[0].re;
      a.iE[vbase + a.G.ve(c)] = lc[0].im;
    }
    if (l == 0) {
      a.flags[m] = ok ? 0 : 1;
      if (a.info && ok) a.info[m] = 0;
      if (a.berr) a.berr[m] = be;
    }
  }
}

// ---------------------------------------------------------------------------
// Orders up to 64: one matrix per workgroup of four wavefronts.  Wavefront j
// owns block column j (16 columns) of V, R = A V - V Lam, E = W R and C in
// the B layout, one 16 x 16 block per row block I (d4 per lane and block).
// A and W (row blocks, A layout) are read from global memory where each
// product needs them; V as the A operand of V C is shared through an LDS
// tile.  Norms and eigenvalue corrections are combined through LDS at
// workgroup barriers, in the same order in every wavefront, so the stopping
// decision is workgroup-uniform.
// ---------------------------------------------------------------------------
constexpr int kColBlocks = 4;           // 64 = 4 x 16
constexpr int kColLd = 16 * kColBlocks + 1;

struct ColShared {
  double2 t[16 * kColBlocks * kColLd];  // V(row, col) at t[col * kColLd + row]
  double2 dl[16 * kColBlocks];          // eigenvalue corrections
  double red[3][kColBlocks];            // per-wavefront partial norms (R, V, A)
};

struct ColBlock {
  d4 re[kColBlocks], im[kColBlocks];
};

// d(I) += M(I, :) x b for I = 0..3: M in the A layout from `ld_a(I, K, s,
// ar, ai)`, b and d in the B layout (this wavefront's block column).  The
// 16 k-steps (K, s) are software-pipelined one deep: the operands of step
// t + 1 are requested before the products of step t, and a scheduling
// barrier keeps the compiler from hoisting all 64 operand loads at once
// (which would not fit in registers next to the three block columns).
template <typename LdA>
__device__ __forceinline__ void col_gemm(LdA &&ld_a, const ColBlock &b, ColBlock &d) {
  constexpr int kSteps = 4 * kColBlocks;
  double ar[2][kColBlocks], ai[2][kColBlocks];
#pragma unroll
  for (int I = 0; I < kColBlocks; ++I) ld_a(I, 0, 0, ar[0][I], ai[0][I]);
#pragma unroll
  for (int t = 0; t < kSteps; ++t) {
    const int K = t / 4, s = t % 4, cur = t & 1;
    if (t + 1 < kSteps) {
#pragma unroll
      for (int I = 0; I < kColBlocks; ++I) ld_a(I, (t + 1) / 4, (t + 1) % 4, ar[cur ^ 1][I], ai[cur ^ 1][I]);
    }
    const double br = b.re[K][s], bi = b.im[K][s];
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I) {
      d.re[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[cur][I], br, d.re[I], 0, 0, 0);
      d.im[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[cur][I], bi, d.im[I], 0, 0, 0);
      d.re[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[cur][I], bi, d.re[I], 0, 0, 0);
      d.im[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[cur][I], br, d.im[I], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// the sum of the four wavefronts' partials, in a fixed order
__device__ __forceinline__ double col_total(const double (&red)[kColBlocks]) {
  return (red[0] + red[1]) + (red[2] + red[3]);
}

template <int NB>
__device__ __forceinline__ void refine_cols_one(const RefineArgs &a, ColShared &sh, int64_t m) {
  static_assert(NB == 16 * kColBlocks, "four block columns");
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  const int n = a.G.n;
  const int64_t base = a.G.mat_base(m), vbase = a.G.vec_base(m);
  const int col = 16 * w + c;  // this lane's column in the B layout
  // V (B layout), eigenvalue of this lane's column (lc) and of its rows (lr)
  ColBlock V;
#pragma unroll
  for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int row = 16 * I + 4 * s + g;
      double vr = 0, vi = 0;
      if (row < n && col < n) {
        const int64_t idx = base + a.G.el(row, col);
        vr = a.rV[idx];
        vi = a.iV[idx];
      }
      V.re[I][s] = vr;
      V.im[I][s] = vi;
    }
  cx<double> lc = col < n ? mk<double>(a.rE[vbase + a.G.ve(col)], a.iE[vbase + a.G.ve(col)]) : mk<double>(0, 0);
  cx<double> lr[kColBlocks][4];
#pragma unroll
  for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 16 * I + 4 * s + g;
      lr[I][s] = k < n ? mk<double>(a.rE[vbase + a.G.ve(k)], a.iE[vbase + a.G.ve(k)]) : mk<double>(0, 0);
    }
  // ||A||_F: this wavefront's block column, then the workgroup total
  double an2 = 0;
#pragma unroll
  for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int row = 16 * I + 4 * s + g;
      if (row < n && col < n) {
        const int64_t idx = base + a.G.el(row, col);
        an2 += a.rA[idx] * a.rA[idx] + a.iA[idx] * a.iA[idx];
      }
    }
  an2 = wave_sum(an2);
  if (l == 0) sh.red[2][w] = an2;
  __syncthreads();
  const double nrmA = sqrt(col_total(sh.red[2]));
  const bool qr_ok = !(a.info32 && a.info32[m] != 0);

  // A operand of a matrix in the batch layout (zero outside n x n); the base
  // is made opaque so that the loads stay where the products use them
  auto ld_global = [&](const auto *re, const auto *im) {
    int64_t b = base, tt = a.G.T;
    int nn = n;
    asm volatile("" : "+s"(b), "+s"(tt), "+s"(nn));
    const int64_t lane_off = b + (int64_t)(c + g * nn) * tt;  // row c, column g
    const auto *pr = re + lane_off;
    const auto *pi = im + lane_off;
    return [=](int I, int K, int s, double &ar, double &ai) {
      const int row = 16 * I + c, cc = 16 * K + 4 * s + g;
      const int64_t off = (int64_t)(16 * I + (16 * K + 4 * s) * nn) * tt;  // wavefront-uniform
      ar = 0;
      ai = 0;
      if (row < n && cc < n) {
        ar = (double)pr[off];
        ai = (double)pi[off];
      }
    };
  };

  double be = INFINITY, be_prev = INFINITY;
  for (int it = 0;; ++it) {
    // ---- R = A V - V diag(lam) and the backward error ----
    ColBlock Rm;
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double vr = V.re[I][s], vi = V.im[I][s];
        Rm.re[I][s] = -(vr * lc.re - vi * lc.im);
        Rm.im[I][s] = -(vr * lc.im + vi * lc.re);
      }
    col_gemm(ld_global(a.rA, a.iA), V, Rm);
    double rn2 = 0, vn2 = 0;
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        rn2 += Rm.re[I][s] * Rm.re[I][s] + Rm.im[I][s] * Rm.im[I][s];
        vn2 += V.re[I][s] * V.re[I][s] + V.im[I][s] * V.im[I][s];
      }
    rn2 = wave_sum(rn2);
    vn2 = wave_sum(vn2);
    if (l == 0) {
      sh.red[0][w] = rn2;
      sh.red[1][w] = vn2;
    }
    __syncthreads();
    be = sqrt(col_total(sh.red[0])) / (nrmA * sqrt(col_total(sh.red[1])));
    if (n == 0) be = 0;
    // workgroup-uniform stop (the same rule as newton_refine); red is next
    // written after two more barriers of this iteration

    if (!qr_ok || be <= a.tol || it >= a.maxit || !(be < be_prev)) break;
    if (it > 0) {
      const double rate = be / be_prev;
      double reach = be;
      for (int k = it; k < a.maxit; ++k) reach *= rate;
      if (reach > a.tol) break;
    }
    be_prev = be;

    // ---- E = W R ----
    ColBlock Em;
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I) {
      Em.re[I] = d4{0, 0, 0, 0};
      Em.im[I] = d4{0, 0, 0, 0};
    }
    col_gemm(ld_global(a.rW, a.iW), Rm, Em);
    // ---- dlam = diag(E); C = E / (lam_k - lam_i) off the diagonal ----
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (4 * s + g == c) {
#pragma unroll
        for (int I = 0; I < kColBlocks; ++I)
          if (I == w) sh.dl[col] = double2{Em.re[I][s], Em.im[I][s]};
      }
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int row = 16 * I + 4 * s + g;
        const cx<double> d = sub(lc, lr[I][s]);
        const double d2 = abs2(d);
        const bool zero = row == col || row >= n || col >= n || !(d2 > 0.0);
        const cx<double> q = zero ? mk<double>(0, 0)
                                  : scale(mulc(mk<double>(Em.re[I][s], Em.im[I][s]), d), frcp(d2));
        Em.re[I][s] = q.re;
        Em.im[I][s] = q.im;
      }
    // ---- V <- V + V C, with V shared through the tile ----
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        sh.t[col * kColLd + 16 * I + 4 * s + g] = double2{V.re[I][s], V.im[I][s]};
    __syncthreads();
    col_gemm(
        [&](int I, int K, int s, double &ar, double &ai) {
          const double2 v = sh.t[(16 * K + 4 * s + g) * kColLd + 16 * I + c];
          ar = v.x;
          ai = v.y;
        },
        Em, V);
    {
      const double2 d = sh.dl[col];
      lc = add(lc, mk<double>(d.x, d.y));
    }
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double2 d = sh.dl[16 * I + 4 * s + g];
        lr[I][s] = add(lr[I][s], mk<double>(d.x, d.y));
      }
    __syncthreads();  // the tile and dl are rewritten by the next iteration
  }
  const bool ok = (be <= a.tol) && qr_ok;
  if (a.maxit > 0) {
#pragma unroll
    for (int I = 0; I < kColBlocks; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int row = 16 * I + 4 * s + g;
        if (row < n && col < n) {
          const int64_t idx = base + a.G.el(row, col);
          a.rV[idx] = V.re[I][s];
          a.iV[idx] = V.im[I][s];
        }
      }
    if (g == 0 && col < n) {
      a.rE[vbase + a.G.ve(col)] = lc.re;
      a.iE[vbase + a.G.ve(col)] = lc.im;
    }
    if (threadIdx.x == 0) {
      a.flags[m] = ok ? 0 : 1;
      if (a.info && ok) a.info[m] = 0;
    }
  }
  if (threadIdx.x == 0 && a.berr) a.berr[m] = be;
}

// one matrix per workgroup of four wavefronts; the batch, or a.list on a
// persistent grid
template <int NB>
__global__ __launch_bounds__(64 * kColBlocks) void k_refine_cols(RefineArgs a) {
  __shared__ ColShared sh;
  // one call site (the body is large): entries i = first, first + step, ...
  int64_t first, end, step;
  if (a.list) {
    first = blockIdx.x;
    end = *a.list_count;
    step = gridDim.x;
  } else {
    first = xcd_block();
    end = first < a.G.nmat ? first + 1 : 0;
    step = 1;
  }
  for (int64_t i = first; i < end; i += step) {  // workgroup-uniform
    const int64_t m = a.list ? a.list[i] : i;
    if (!a.mask || a.mask[m] != 0) refine_cols_one<NB>(a, sh, m);
    __syncthreads();
  }
}

}  // namespace veig
