// Mixed-precision batched eigensolver for MI355X (gfx950).
//
//   1. k_zgeev<float, NB, P, OUT>: Hessenberg reduction, QR and the
//      eigenvector factors in complex64 (the same lane-group kernel as the FP64
//      path).  For n <= 16 it writes the Schur vectors Z, the unit upper
//      triangular X (V = Z X) and Y = X^{-1}; k_refine_zxy then forms
//      V = Z X and W = Y Z^H ~ V^{-1} on the FP64 matrix cores and refines,
//      so the VALU kernel does no O(n^3) eigenvector work and keeps its
//      occupancy.  For 16 < n <= 64 the VALU kernel writes V and W itself
//      (n > 32: the two-wavefront kernel, complex64).
//   2. k_refine_mfma<NT> (FP64 matrix cores, one wavefront per matrix; for
//      n > 32 k_refine_cols, one workgroup of four wavefronts per matrix):
//      Newton refinement of all eigenpairs at once.  With R = A V - V diag(lam)
//      and E = W R, the first-order correction of
//      A (V + dV) = (V + dV)(Lam + dLam) with dV = V C is
//          dlam_k = E_kk,    C_ik = E_ik / (lam_k - lam_i)  (i != k),  C_kk = 0,
//      so each iteration costs three n x n complex products.  The backward
//      error ||R||_F / (||A||_F ||V||_F) is the convergence test; iteration
//      stops at `tol`.
//   3. Matrices that do not reach `tol` within the iteration cap (or whose
//      complex64 QR did not converge) are collected in a list and re-solved
//      by the FP64 kernel on a persistent grid walking that list, so the
//      fallback costs in proportion to its size; their backward error is
//      then measured again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "veig_common.hpp"
#include "veig_refine.hpp"

using namespace veig;

namespace {

int launch_refine(const RefineArgs &a, hipStream_t s) {
  // one wavefront per matrix (a persistent grid for list launches)
  if (a.G.nmat > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  if (a.G.n <= 16 && !a.list && !a.mask) {
    // the whole batch: coalesced staging through LDS
    const int64_t groups = (a.G.nmat + kStageWaves - 1) / kStageWaves;
    hipLaunchKernelGGL(k_refine_staged, dim3((unsigned)groups), dim3(64 * kStageWaves), 0, s, a);
    VEIG_HIP(hipGetLastError());
    return VEIG_OK;
  }
  if (a.G.n > 32) {
    // one matrix per workgroup of four wavefronts
    const dim3 grid((unsigned)(a.list ? std::min<int64_t>(a.G.nmat, 4096) : a.G.nmat));
    hipLaunchKernelGGL((k_refine_cols<64>), grid, dim3(64 * kColBlocks), 0, s, a);
    VEIG_HIP(hipGetLastError());
    return VEIG_OK;
  }
  const int64_t groups = (a.G.nmat + kRefWaves - 1) / kRefWaves;
  const dim3 grid((unsigned)(a.list ? std::min<int64_t>(groups, 4096) : groups));
  const dim3 block(64 * kRefWaves);
  if (a.G.n <= 16) hipLaunchKernelGGL((k_refine_mfma<1>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_refine_mfma<2>), grid, block, 0, s, a);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct MixedWs {
  float *rW, *iW;  // W = V^{-1}, or the Schur vectors Z (n <= 16)
  float *rX, *iX;  // n <= 16: X above and X^{-1} below the diagonal
  int32_t *flags, *info32;
  int32_t *list, *count;  // matrices left to the FP64 solve
};

size_t mixed_ws_bytes(int64_t dim, int64_t nmat) {
  const size_t nm2 = (size_t)nmat * dim * dim;
  return 4 * align_up(nm2 * sizeof(float)) + 3 * align_up((size_t)nmat * sizeof(int32_t)) + kAlign;
}

MixedWs carve(void *ws, int64_t dim, int64_t nmat) {
  char *p = static_cast<char *>(ws);
  const size_t nm2 = (size_t)nmat * dim * dim;
  MixedWs w;
  w.rW = reinterpret_cast<float *>(p);
  p += align_up(nm2 * sizeof(float));
  w.iW = reinterpret_cast<float *>(p);
  p += align_up(nm2 * sizeof(float));
  w.rX = reinterpret_cast<float *>(p);
  p += align_up(nm2 * sizeof(float));
  w.iX = reinterpret_cast<float *>(p);
  p += align_up(nm2 * sizeof(float));
  w.flags = reinterpret_cast<int32_t *>(p);
  p += align_up((size_t)nmat * sizeof(int32_t));
  w.info32 = reinterpret_cast<int32_t *>(p);
  p += align_up((size_t)nmat * sizeof(int32_t));
  w.list = reinterpret_cast<int32_t *>(p);
  p += align_up((size_t)nmat * sizeof(int32_t));
  w.count = reinterpret_cast<int32_t *>(p);
  return w;
}

__global__ void k_zero_count(int32_t *count) {
  if (threadIdx.x == 0) count[0] = 0;
}

// list of the matrices flagged for the FP64 solve.  Any order is correct
// (every matrix is solved independently); each chunk of 256 consecutive
// matrices is written as one ascending run (one atomic per chunk, ranks from
// ballots), so that list neighbours are mostly batch neighbours and share
// cache lines of the tiled layout when the list is walked.
constexpr int kCompactThreads = 256;

__global__ __launch_bounds__(kCompactThreads) void k_compact(const int32_t *flags, int64_t nmat, int32_t *list,
                                                             int32_t *count) {
  __shared__ int32_t wave_cnt[kCompactThreads / 64];
  __shared__ int32_t chunk_base;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t c0 = (int64_t)blockIdx.x * kCompactThreads; c0 < nmat; c0 += (int64_t)gridDim.x * kCompactThreads) {
    const int64_t m = c0 + threadIdx.x;
    const bool f = m < nmat && flags[m] != 0;
    const uint64_t bits = __ballot(f);
    const int rank = __popcll(bits & ((1ull << lane) - 1));
    if (lane == 0) wave_cnt[w] = __popcll(bits);
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int j = 0; j < kCompactThreads / 64; ++j) {
      before += j < w ? wave_cnt[j] : 0;
      total += wave_cnt[j];
    }
    if (threadIdx.x == 0) chunk_base = total ? atomicAdd(count, total) : 0;
    __syncthreads();
    if (f) list[chunk_base + before + rank] = (int32_t)m;
    __syncthreads();  // wave_cnt and chunk_base are rewritten for the next chunk
  }
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
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM_FAST));
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
  if (dim <= 16) {
    // 1. complex64 reduction, QR and eigenvector factors Z, X, X^{-1}
    FusedArgs f32{rAm, iAm, rE, iE, rV, iV, w.info32, G, its, 0u, nullptr, w.rW, w.iW};
    f32.rX = w.rX;
    f32.iX = w.iX;
    VEIG_TRY((launch_zgeev<float, OUT_ZXY>(f32, s)));
    // 2. V = Z X, W = X^{-1} Z^H and the refinement on the FP64 matrix cores
    RefineArgs ra{rA, iA, rE, iE, rV, iV, w.rW, w.iW, w.flags, berr, w.info32, nullptr, info, G, tol, 3};
    ra.rX = w.rX;
    ra.iX = w.iX;
    const int64_t groups = (G.nmat + kStageWaves - 1) / kStageWaves;
    if (groups > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
    hipLaunchKernelGGL(k_refine_zxy, dim3((unsigned)groups), dim3(64 * kStageWaves), 0, s, ra);
    VEIG_HIP(hipGetLastError());
  } else {
    // 1. complex64 solve, V and W = V^{-1}
    FusedArgs f32{rAm, iAm, rE, iE, rV, iV, w.info32, G, its, 0u, nullptr, w.rW, w.iW};
    VEIG_TRY((launch_zgeev<float, OUT_VW>(f32, s)));
    // 2. FP64 refinement
    RefineArgs ra{rA, iA, rE, iE, rV, iV, w.rW, w.iW, w.flags, berr, w.info32, nullptr, info, G, tol, 3};
    VEIG_TRY(launch_refine(ra, s));
  }
  // the list of matrices that did not reach tol
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, w.count);
  {
    const int64_t blocks = std::min<int64_t>((G.nmat + kCompactThreads - 1) / kCompactThreads, 4096);
    hipLaunchKernelGGL(k_compact, dim3((unsigned)blocks), dim3(kCompactThreads), 0, s, w.flags, G.nmat, w.list,
                       w.count);
  }
  VEIG_HIP(hipGetLastError());
  // 3. FP64 solve of the listed matrices (persistent grid), then their backward error
  FusedArgs f64{rAm, iAm, rE, iE, rV, iV, info, G, its, 0u, nullptr, nullptr, nullptr};
  f64.list = w.list;
  f64.list_count = w.count;
  VEIG_TRY((launch_zgeev<double, OUT_V>(f64, s)));
  if (berr) {
    RefineArgs chk{rA, iA, rE, iE, rV, iV, nullptr, nullptr, w.flags, berr, nullptr, nullptr, nullptr, G, tol, 0,
                   w.list, w.count};
    VEIG_TRY(launch_refine(chk, s));
  }
  if (fallback)
    VEIG_HIP(hipMemcpyAsync(fallback, w.flags, (size_t)G.nmat * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  return VEIG_OK;
}

}  // extern "C"
