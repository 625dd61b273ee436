// Mixed-precision batched eigensolver for MI355X (gfx950).
//
//   1. k_zgeev<float, NB, true>: Hessenberg reduction, QR and eigenvectors in
//      complex64 (the same lane-group kernel as the FP64 path), which also
//      emits W = X^{-1} Z^H, an approximate inverse of the eigenvector matrix.
//   2. k_refine_mfma<NT> (FP64 matrix cores, one wavefront per matrix):
//      Newton refinement of all eigenpairs at once.  With R = A V - V diag(lam)
//      and E = W R, the first-order correction of
//      A (V + dV) = (V + dV)(Lam + dLam) with dV = V C is
//          dlam_k = E_kk,    C_ik = E_ik / (lam_k - lam_i)  (i != k),  C_kk = 0,
//      so each iteration costs three n x n complex products.  The backward
//      error ||R||_F / (||A||_F ||V||_F) is the convergence test; iteration
//      stops at `tol`.
//   3. Matrices that do not reach `tol` within the iteration cap (or whose
//      complex64 QR did not converge) are re-solved by the FP64 kernel in a
//      masked launch, and their backward error is measured again.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "veig_common.hpp"
#include "veig_refine.hpp"

using namespace veig;

namespace {

int launch_refine(const RefineArgs &a, hipStream_t s) {
  // one wavefront per matrix
  if (a.G.nmat > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  const dim3 grid((unsigned)a.G.nmat);
  if (a.G.n <= 16) hipLaunchKernelGGL((k_refine_mfma<1>), grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL((k_refine_mfma<2>), grid, dim3(64), 0, s, a);
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
