// Eigenvalues only, and polynomial roots, on the lane-group kernels.
//
//   veig_zlahqr_eig    the reference's eigenvalues-only zlahqr_
//                      (Toolbox.hpp:1663-1890) over the batch: H upper
//                      Hessenberg with a real subdiagonal, updated in place
//                      inside the active blocks; E = its diagonal.
//   veig_eigvals       eigenvalues of general matrices: the Hessenberg
//                      reduction (no Q) followed by the same iteration.
//   veig_zlahqr_roots  the reference's polynomial-roots kernel zlahqr
//                      (Toolbox.hpp:1894-1957): the companion matrix of each
//                      polynomial (create_companion_matrices, 605-643) is
//                      built in LDS from its coefficients and iterated to
//                      Schur form; E = roots, and on request T and Z.
//
// Without Schur vectors a lane holds no row of Z, so these kernels need a
// fraction of the registers of the fused eigensolver (and about a quarter of
// its flops: no T outside the active block, no Z, no eigenvector solves).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "veig_common.hpp"

using namespace veig;

namespace {

enum { SRC_MATRIX = 0, SRC_HESSENBERG = 1, SRC_POLY = 2 };

struct EigArgs {
  const double *rA, *iA;  // matrices, Hessenberg matrices or polynomial coefficients
  double *rH, *iH;        // out: H after the iteration (SRC_HESSENBERG) or T (SRC_POLY, Schur)
  double *rZ, *iZ;        // out: Schur vectors (SRC_POLY, Schur)
  double *rE, *iE;
  int32_t *info;
  Geom G;
  int itmax;
};

// Row idx(p) of the companion matrix of c_0 + c_1 x + ... + c_n x^n into
// the tile: 1 on the subdiagonal, -c_r / c_n in the last column (Smith's
// division, as create_companion_matrices).  The tile is zero elsewhere.
template <typename G>
__device__ __forceinline__ void companion_rows(const G &g, const EigArgs &a, int64_t m) {
  const int n = g.n;
  if (n <= 0) return;
  const int64_t T = a.G.T, b = m / T, t = m - b * T;
  const int64_t cbase = b * T * (int64_t)(n + 1) + t;  // vector slot of length n + 1
  const cx<double> lead = mk<double>(a.rA[cbase + (int64_t)n * T], a.iA[cbase + (int64_t)n * T]);
  constexpr int kSlots = (G::LD - 1) / G::LG;  // indices per lane
#pragma unroll
  for (int p = 0; p < kSlots; ++p) {
    const int r = g.idx(p);
    if (r < n) {
      const cx<double> c = mk<double>(a.rA[cbase + (int64_t)r * T], a.iA[cbase + (int64_t)r * T]);
      const cx<double> q = cdiv(c, lead);
      st<double>(g.at(r, n - 1), mk(-q.re, -q.im));
      if (r >= 1 && r - 1 != n - 1) st<double>(g.at(r, r - 1), mk<double>(1, 0));
    }
  }
}

// LDS of a wavefront's MPW matrices in layout LAY, and each group's vector:
// the reflector of the Hessenberg reduction (dense layout), or just the two
// entries where the QR iteration stages a sweep's first column when it starts
// at row 0 (every group needs its own)
template <int NB, int P, int LAY>
struct EigTiles {
  using G = Group<double, NB, P, LAY>;
  static constexpr int kTile = G::kSize + 1;  // +1: each group starts on another bank
  static constexpr int kVec = LAY == LAY_DENSE ? NB + 1 : 2;
  double2 H[G::MPW * kTile];
  double2 v[G::MPW * kVec];
};

template <int NB, int P, int SRC, bool TZ>
__global__ __launch_bounds__(64) void k_eig(EigArgs a) {
  using R = double;
  // the banded tile serves the iterations that never leave the band: no
  // Hessenberg reduction (dense trailing matrix) and no Schur form output
  constexpr int LAY = (SRC != SRC_MATRIX && !TZ) ? LAY_BAND : LAY_DENSE;
  using G = Group<R, NB, P, LAY>;
  using Tl = EigTiles<NB, P, LAY>;
  __shared__ Tl sm;
  constexpr int MPW = G::MPW;
  const int64_t m = xcd_block() * MPW + threadIdx.x / G::LG;
  const bool valid = m < a.G.nmat;
  if (!__any(valid)) return;
  const int n = valid ? a.G.n : 0, nmax = a.G.n;
  G g;
  g.mi = threadIdx.x / G::LG;
  g.r = threadIdx.x % G::LG;
  g.n = n;
  g.H = sm.H + g.mi * Tl::kTile;
  g.vsh = sm.v + g.mi * Tl::kVec;
  for (int i = g.r; i < G::kSize; i += G::LG) st<R>(g.H + i, mk<R>(0, 0));
  if constexpr (LAY == LAY_DENSE) {
#pragma unroll
    for (int p = 0; p < P; ++p) st<R>(g.vsh + g.idx(p), mk<R>(0, 0));
  }
  const int64_t base = valid ? a.G.mat_base(m) : 0, vbase = valid ? a.G.vec_base(m) : 0;
  wave_sync();
  if constexpr (SRC == SRC_POLY) {
    companion_rows(g, a, m);
  } else {
    // rows idx(p): entries with row <= col + band (the whole row for a
    // general matrix, the Hessenberg part otherwise)
    constexpr int band = SRC == SRC_HESSENBERG ? 1 : NB;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int row = g.idx(p);
      if (row < n)
        for (int c = row > band ? row - band : 0; c < n; ++c) {
          const int64_t idx = base + a.G.el(row, c);
          st<R>(g.at(row, c), mk<R>(a.rA[idx], a.iA[idx]));
        }
    }
  }
  cx<R> z[P][NB];
  if constexpr (TZ) identity_rows<R, NB, P>(g, z);
  wave_sync();
  if constexpr (SRC == SRC_MATRIX) gehrd_q<R, NB, P>(g, z, nmax, false, false, nullptr);
  const int nfail = lahqr<R, NB, P, TZ>(g, z, a.itmax, false);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = g.idx(p);
    if (k < n) {
      const cx<R> lam = ld<R>(g.at(k, k));
      a.rE[vbase + a.G.ve(k)] = lam.re;
      a.iE[vbase + a.G.ve(k)] = lam.im;
    }
  }
  if constexpr (SRC == SRC_HESSENBERG) {
    // the Hessenberg part back (entries below the subdiagonal untouched)
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int row = g.idx(p);
      if (row < n)
        for (int c = row > 1 ? row - 1 : 0; c < n; ++c) {
          const int64_t idx = base + a.G.el(row, c);
          const cx<R> h = ld<R>(g.at(row, c));
          a.rH[idx] = h.re;
          a.iH[idx] = h.im;
        }
    }
  }
  if constexpr (SRC == SRC_POLY && TZ) {
    store_rows<R, NB, P>(g, a.G, base, a.rH, a.iH);
    store_row_regs<R, NB, P>(g, a.G, base, a.rZ, a.iZ, z);
  }
  if (a.info && valid && g.r == 0) a.info[m] = nfail;
}

template <int SRC, bool TZ>
int launch_eig(const EigArgs &a, hipStream_t s) {
  if (a.G.nmat > (int64_t)INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  const int nb = nb_for(a.G.n);
  const int mpw = 64 / nb;
  const unsigned grid = (unsigned)((a.G.nmat + mpw - 1) / mpw);
  switch (nb) {
    case 8: hipLaunchKernelGGL((k_eig<8, 1, SRC, TZ>), dim3(grid), dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_eig<16, 1, SRC, TZ>), dim3(grid), dim3(64), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_eig<32, 1, SRC, TZ>), dim3(grid), dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL((k_eig<64, 1, SRC, TZ>), dim3(grid), dim3(64), 0, s, a); break;
  }
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

}  // namespace

extern "C" {

int veig_zlahqr_eig(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rE, double *iE,
                    double *rH, double *iH, int32_t *info, int32_t itmax, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM_FAST));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rH); VEIG_REQUIRE(iH);
  EigArgs a{rH, iH, rH, iH, nullptr, nullptr, rE, iE, info, make_geom(dim, block_cnt, thread_cnt),
            itmax > 0 ? itmax : 30};
  return launch_eig<SRC_HESSENBERG, false>(a, (hipStream_t)stream);
}

int veig_eigvals(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rA, const double *iA,
                 double *rE, double *iE, int32_t *info, int32_t itmax, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM_FAST));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rE); VEIG_REQUIRE(iE);
  EigArgs a{rA, iA, nullptr, nullptr, nullptr, nullptr, rE, iE, info, make_geom(dim, block_cnt, thread_cnt),
            itmax > 0 ? itmax : 30};
  return launch_eig<SRC_MATRIX, false>(a, (hipStream_t)stream);
}

int veig_zlahqr_roots(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rC,
                      const double *iC, double *rE, double *iE, double *rV, double *iV, double *rZ,
                      double *iZ, int32_t *info, int32_t itmax, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM_FAST));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rC); VEIG_REQUIRE(iC); VEIG_REQUIRE(rE); VEIG_REQUIRE(iE);
  const bool schur = rV || iV || rZ || iZ;
  if (schur && !(rV && iV && rZ && iZ))
    return fail(VEIG_ERR_INVALID, "rV, iV, rZ, iZ: all four (Schur form and vectors) or none (roots only)");
  EigArgs a{rC, iC, rV, iV, rZ, iZ, rE, iE, info, make_geom(dim, block_cnt, thread_cnt), itmax > 0 ? itmax : 30};
  if (schur) return launch_eig<SRC_POLY, true>(a, (hipStream_t)stream);
  return launch_eig<SRC_POLY, false>(a, (hipStream_t)stream);
}

}  // extern "C"
