// Mixed-precision batched eigensolver for MI355X (gfx950).
//
//   1. k_zgeev<float, NB, true>: Hessenberg reduction, QR and eigenvectors in
//      complex64 (the same lane-group kernel as the FP64 path), which also
//      emits W = X^{-1} Z^H, an approximate inverse of the eigenvector matrix.
//   2. k_refine<NB> (FP64): Newton refinement of all eigenpairs at once.  With
//      R = A V - V diag(lam) and E = W R, the first-order correction of
//      A (V + dV) = (V + dV)(Lam + dLam) with dV = V C is
//          dlam_k = E_kk,    C_ik = E_ik / (lam_k - lam_i)  (i != k),  C_kk = 0,
//      so each iteration costs three n x n complex products.  The backward
//      error ||R||_F / (||A||_F ||V||_F) is the convergence test; iteration
//      stops at `tol`.  (Newton's method for the eigenproblem in the form of
//      Dongarra, Moler & Wilkinson, "Improving the accuracy of computed
//      eigenvalues and eigenvectors", SIAM J. Numer. Anal. 20 (1983).)
//   3. Matrices that do not reach `tol` within the iteration cap (or whose
//      complex64 QR did not converge) are re-solved by the FP64 kernel in a
//      masked launch, and their backward error is measured again.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "veig_common.hpp"

using namespace veig;

namespace {

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

template <int NB>
__global__ __launch_bounds__(64) void k_refine(RefineArgs a) {
  using R = double;
  __shared__ Tiles<R, NB> sm;
  constexpr int MPW = Tiles<R, NB>::MPW;
  const int64_t m = xcd_block() * MPW + threadIdx.x / NB;
  bool valid = m < a.G.nmat;
  if (valid && a.mask) valid = a.mask[m] != 0;
  if (!__any(valid)) return;
  const int n = valid ? a.G.n : 0;
  Group<R, NB> g = make_group<R, NB>(sm, n);
  const int r = g.r;
  const bool in = r < n;
  const int64_t base = valid ? a.G.mat_base(m) : 0;
  const int64_t vbase = valid ? a.G.vec_base(m) : 0;

  cx<R> av[NB], v[NB], w[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    av[c] = mk<R>(0, 0);
    v[c] = mk<R>(0, 0);
    w[c] = mk<R>(0, 0);
    if (in && c < n) {
      const int64_t idx = base + a.G.el(r, c);
      av[c] = mk<R>(a.rA[idx], a.iA[idx]);
      v[c] = mk<R>(a.rV[idx], a.iV[idx]);
      if (a.maxit > 0) w[c] = mk<R>((R)a.rW[idx], (R)a.iW[idx]);
    }
  }
  cx<R> lam = in ? mk<R>(a.rE[vbase + a.G.ve(r)], a.iE[vbase + a.G.ve(r)]) : mk<R>(0, 0);
  st<R>(g.vsh + r, lam);
  R an2 = 0;
#pragma unroll
  for (int c = 0; c < NB; ++c) an2 += abs2(av[c]);
  const R nrmA = fsqrt(group_sum<NB>(an2));
  const bool qr_ok = !(a.info32 && valid && a.info32[m] != 0);
  R be = R(INFINITY);
  bool active = valid && qr_ok;
  for (int it = 0;; ++it) {
    // tile <- V
#pragma unroll
    for (int c = 0; c < NB; ++c) st<R>(g.at(r, c), v[c]);
    wave_sync();
    // row r of R = A V - V diag(lam)
    cx<R> acc[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const cx<R> lk = ld<R>(g.vsh + k);
      acc[k] = mk(-(v[k].re * lk.re - v[k].im * lk.im), -(v[k].re * lk.im + v[k].im * lk.re));
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < a.G.n) {
#pragma unroll
        for (int k = 0; k < NB; ++k) acc[k] = add(acc[k], mul(av[j], ld<R>(g.at(j, k))));
      }
    }
    R rn2 = 0, vn2 = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      rn2 += in ? abs2(acc[k]) : R(0);
      vn2 += in ? abs2(v[k]) : R(0);
    }
    rn2 = group_sum<NB>(rn2);
    vn2 = group_sum<NB>(vn2);
    if (active || (valid && a.maxit == 0)) be = fsqrt(rn2) / (nrmA * fsqrt(vn2));
    if (!(be <= a.tol) && it < a.maxit && active) {
      // keep iterating
    } else {
      active = false;
    }
    if (!__any(active)) break;
    wave_sync();
    // tile <- R
#pragma unroll
    for (int c = 0; c < NB; ++c) st<R>(g.at(r, c), acc[c]);
    wave_sync();
    // row r of E = W R
    cx<R> e[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) e[k] = mk<R>(0, 0);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < a.G.n) {
#pragma unroll
        for (int k = 0; k < NB; ++k) e[k] = add(e[k], mul(w[j], ld<R>(g.at(j, k))));
      }
    }
    // dlam_r = E_rr; C_rk = E_rk / (lam_k - lam_r)
    cx<R> dl = mk<R>(0, 0);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (k == r) {
        dl = e[k];
        e[k] = mk<R>(0, 0);
      } else if (k < n) {
        const cx<R> d = sub(ld<R>(g.vsh + k), lam);
        e[k] = (d.re == R(0) && d.im == R(0)) ? mk<R>(0, 0) : mul(e[k], inv(d));
      }
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < NB; ++c) st<R>(g.at(r, c), e[c]);
    wave_sync();
    // V <- V + V C, lam <- lam + dlam (matrices still iterating)
    cx<R> nv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) nv[k] = v[k];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < a.G.n) {
#pragma unroll
        for (int k = 0; k < NB; ++k) nv[k] = add(nv[k], mul(v[j], ld<R>(g.at(j, k))));
      }
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < NB; ++k) v[k] = nv[k];
      lam = add(lam, dl);
    }
    wave_sync();
    st<R>(g.vsh + r, lam);
    wave_sync();
  }
  if (!valid) return;
  const bool ok = (be <= a.tol) && qr_ok;
  if (a.maxit > 0) {
    store_row_regs<R, NB>(g, a.G, base, a.rV, a.iV, v);
    if (in) {
      a.rE[vbase + a.G.ve(r)] = lam.re;
      a.iE[vbase + a.G.ve(r)] = lam.im;
    }
    if (r == 0) {
      a.flags[m] = ok ? 0 : 1;
      if (a.info && ok) a.info[m] = 0;
    }
  }
  if (r == 0 && a.berr) a.berr[m] = be;
}

int launch_refine(const RefineArgs &a, hipStream_t s) {
  const int nb = nb_for(a.G.n);
  const int mpw = 64 / nb;
  const int64_t grid = (a.G.nmat + mpw - 1) / mpw;
  if (grid > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  switch (nb) {
    case 8: hipLaunchKernelGGL((k_refine<8>), dim3((unsigned)grid), dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_refine<16>), dim3((unsigned)grid), dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL((k_refine<32>), dim3((unsigned)grid), dim3(64), 0, s, a); break;
  }
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct MixedWs {
  float *rW, *iW;
  int32_t *flags, *info32;
};

size_t mixed_ws_bytes(int64_t dim, int64_t nmat) {
  const size_t nm2 = (size_t)nmat * dim * dim;
  return 2 * align_up(nm2 * sizeof(float)) + 2 * align_up((size_t)nmat * sizeof(int32_t));
}

MixedWs carve(void *ws, int64_t dim, int64_t nmat) {
  char *p = static_cast<char *>(ws);
  const size_t nm2 = (size_t)nmat * dim * dim;
  MixedWs w;
  w.rW = reinterpret_cast<float *>(p);
  p += align_up(nm2 * sizeof(float));
  w.iW = reinterpret_cast<float *>(p);
  p += align_up(nm2 * sizeof(float));
  w.flags = reinterpret_cast<int32_t *>(p);
  p += align_up((size_t)nmat * sizeof(int32_t));
  w.info32 = reinterpret_cast<int32_t *>(p);
  return w;
}

}  // namespace

extern "C" {

size_t veig_mixed_workspace_size(int64_t dim, int64_t block_cnt, int64_t thread_cnt) {
  if (dim < 0 || block_cnt < 0 || thread_cnt < 1) return 0;
  return mixed_ws_bytes(dim, block_cnt * thread_cnt);
}

int veig_zgeev_mixed(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rA,
                     const double *iA, double *rE, double *iE, double *rV, double *iV,
                     int32_t *info, int32_t *fallback, double *berr, double tol, int32_t itmax,
                     void *workspace, size_t workspace_bytes, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rV);
  VEIG_REQUIRE(iV); VEIG_REQUIRE(workspace);
  if (!(tol > 0.0)) return fail(VEIG_ERR_INVALID, "tol must be positive");
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  const size_t need = mixed_ws_bytes(dim, G.nmat);
  if (workspace_bytes < need)
    return fail(VEIG_ERR_INVALID, "workspace too small: need " + std::to_string(need) + " bytes");
  if (reinterpret_cast<uintptr_t>(workspace) % kAlign != 0)
    return fail(VEIG_ERR_INVALID, "workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int its = itmax > 0 ? itmax : 30;
  MixedWs w = carve(workspace, dim, G.nmat);
  double *rAm = const_cast<double *>(rA), *iAm = const_cast<double *>(iA);
  // 1. complex64 solve, W = V^{-1}
  FusedArgs f32{rAm, iAm, rE, iE, rV, iV, w.info32, G, its, 0u, nullptr, w.rW, w.iW};
  VEIG_TRY((launch_zgeev<float, true>(f32, s)));
  // 2. FP64 refinement
  RefineArgs ra{rA, iA, rE, iE, rV, iV, w.rW, w.iW, w.flags, berr, w.info32, nullptr, info, G, tol, 3};
  VEIG_TRY(launch_refine(ra, s));
  // 3. FP64 solve of the matrices that did not reach tol, then their backward error
  FusedArgs f64{rAm, iAm, rE, iE, rV, iV, info, G, its, 0u, w.flags, nullptr, nullptr};
  VEIG_TRY((launch_zgeev<double, false>(f64, s)));
  if (berr) {
    RefineArgs chk{rA, iA, rE, iE, rV, iV, nullptr, nullptr, w.flags, berr, nullptr, w.flags, nullptr, G, tol, 0};
    VEIG_TRY(launch_refine(chk, s));
  }
  if (fallback)
    VEIG_HIP(hipMemcpyAsync(fallback, w.flags, (size_t)G.nmat * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  return VEIG_OK;
}

}  // extern "C"
