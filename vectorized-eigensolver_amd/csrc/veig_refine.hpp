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

// d += a * b  (a in A layout, b and d in B layout)
template <int NT>
__device__ __forceinline__ void cm_gemm(const CMat<NT> &a, const CMat<NT> &b, CMat<NT> &d) {
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
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
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
};

template <int NT>
__global__ __launch_bounds__(64) void k_refine_mfma(RefineArgs a) {
  __shared__ RefTile<NT> sh;
  const int64_t m = xcd_block();  // one matrix per wavefront
  if (m >= a.G.nmat) return;
  if (a.mask && a.mask[m] == 0) return;
  const int l = threadIdx.x, c = l & 15, g = l >> 4;
  const int n = a.G.n;
  const int64_t base = a.G.mat_base(m), vbase = a.G.vec_base(m);

  // A and W stay in registers for one block; for larger orders they are
  // re-read (from L2) where they are used, to stay within 256 VGPRs.
  constexpr bool kHold = NT == 1;
  CMat<NT> A, V, W;
  cm_load_global<NT, LAY_A>(A, a.G, base, a.rA, a.iA, c, g);
  cm_load_global<NT, LAY_B>(V, a.G, base, a.rV, a.iV, c, g);
  if (kHold && a.maxit > 0) cm_load_global<NT, LAY_A>(W, a.G, base, a.rW, a.iW, c, g);
  double an2 = 0;
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int J = 0; J < NT; ++J)
#pragma unroll
      for (int s = 0; s < 4; ++s) an2 += A.re[I][J][s] * A.re[I][J][s] + A.im[I][J][s] * A.im[I][J][s];
  const double nrmA = sqrt(wave_sum(an2));
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
  double be = INFINITY;
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
    if constexpr (!kHold) {
      int64_t b = base;
      asm volatile("" : "+v"(b));  // opaque: keep the reload inside the loop
      cm_load_global<NT, LAY_A>(A, a.G, b, a.rA, a.iA, c, g);
    }
    cm_gemm<NT>(A, V, Rm);
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
    if (!qr_ok || be <= a.tol || it >= a.maxit) break;  // wavefront-uniform

    // ---- E = W R ----
    CMat<NT> Em;
    cm_zero<NT>(Em);
    if constexpr (!kHold) {
      int64_t b = base;
      asm volatile("" : "+v"(b));
      cm_load_global<NT, LAY_A>(W, a.G, b, a.rW, a.iW, c, g);
    }
    cm_gemm<NT>(W, Rm, Em);
    // ---- dlam = diag(E); C = E / (lam_k - lam_i) off the diagonal ----
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (4 * s + g == c) sh.dl[16 * I + c] = double2{Em.re[I][I][s], Em.im[I][I][s]};
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = 16 * I + 4 * s + g, col = 16 * J + c;
          const cx<double> d = sub(lc[J], lr[I][s]);
          const bool zero = row == col || row >= n || col >= n || (d.re == 0.0 && d.im == 0.0);
          const cx<double> q = zero ? mk<double>(0, 0)
                                    : cdiv(mk<double>(Em.re[I][J][s], Em.im[I][J][s]), d);
          Em.re[I][J][s] = q.re;
          Em.im[I][J][s] = q.im;
        }
    // ---- V <- V + V C (V transposed to the A layout through LDS) ----
    CMat<NT> X;
    cm_store<NT, LAY_B>(sh, V, c, g);
    wave_sync();
    cm_load<NT, LAY_A>(sh, X, c, g);
    cm_gemm<NT>(X, Em, V);
#pragma unroll
    for (int J = 0; J < NT; ++J) {
      const double2 d = sh.dl[16 * J + c];
      lc[J] = add(lc[J], mk<double>(d.x, d.y));
    }
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double2 d = sh.dl[16 * I + 4 * s + g];
        lr[I][s] = add(lr[I][s], mk<double>(d.x, d.y));
      }
    wave_sync();
  }
  const bool ok = (be <= a.tol) && qr_ok;
  if (a.maxit > 0) {
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = 16 * I + 4 * s + g, col = 16 * J + c;
          if (row < n && col < n) {
            const int64_t idx = base + a.G.el(row, col);
            a.rV[idx] = V.re[I][J][s];
            a.iV[idx] = V.im[I][J][s];
          }
        }
    if (g == 0) {
#pragma unroll
      for (int J = 0; J < NT; ++J) {
        const int k = 16 * J + c;
        if (k < n) {
          a.rE[vbase + a.G.ve(k)] = lc[J].re;
          a.iE[vbase + a.G.ve(k)] = lc[J].im;
        }
      }
    }
    if (l == 0) {
      a.flags[m] = ok ? 0 : 1;
      if (a.info && ok) a.info[m] = 0;
    }
  }
  if (l == 0 && a.berr) a.berr[m] = be;
}

}  // namespace veig
