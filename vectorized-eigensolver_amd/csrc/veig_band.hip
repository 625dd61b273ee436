// FP64 batches (and lists of matrices from a batch) at orders 9..32 in two
// kernels, so that the QR iteration runs on a banded tile and more matrices
// stay resident per CU.
//
//   k_gehrd_c<NB, LW>  the Hessenberg reduction alone (zgehrd, Toolbox.hpp:
//                      1315-1469) on the dense tile, by one wavefront (NB = 16)
//                      or two sharing each update (NB = 24..32), written in
//                      zgehrd's compact form -- H on and above the
//                      subdiagonal, the reflector vectors v(i+2:n) below it,
//                      tau(i) -- to the V and E planes, which serve as its
//                      workspace.
//   k_qr_band<NB, LW>  Q from those reflectors (zunghr, Toolbox.hpp:
//                      1470-1659), the QR iteration with Schur vectors
//                      (zlahqr_schur, Toolbox.hpp:690-1031) and the
//                      eigenvectors (ztrevc, Toolbox.hpp:1143-1311) on the
//                      band row <= col + 2 of the tile: every entry a bulge
//                      chase, the deflation test and the triangular solve
//                      touch lies in it.
//
// NB is the slot's capacity, LW the lanes per matrix (0: NB): orders 9..12
// take <12, 16>, 13..16 <16, 0>, 17..24 <24, 32>, 25..28 <28, 32>, 29..32
// <32, 0>.  At NB = 16
// the band holds 166 of the dense tile's 272 entries: 11.8 KB of LDS per
// wavefront of four matrices instead of 18.6 KB, which with at most 168
// registers lets three wavefronts share a SIMD instead of two; with two
// matrices per wavefront, 11.3 to 19.5 KB instead of 19.2 to 33.8 KB: two
// wavefronts per SIMD instead of one (the 32-entry Schur-vector rows of
// NB = 32 partly spilled at 256 registers, the 24-entry ones not).  The
// reduction's round trip through HBM (the compact form written and read
// back) costs 2 x 16 n^2 bytes per matrix.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "veig_band.hpp"

using namespace veig;

namespace {

// the matrix of this lane's group (of MPW per wavefront): from the batch,
// or from a.list, whose *a.list_count entries the grid's first blocks take
// -- blocks are dispatched in order, round-robin over the XCDs, so the work
// spreads over all of them -- while the blocks past them find no matrix
template <int MPW>
__device__ __forceinline__ bool slot_matrix(const FusedArgs &a, int group, int64_t &m) {
  if (a.list) {
    const int64_t cnt = *a.list_count;
    const int64_t used = (cnt + MPW - 1) / MPW;
    const int64_t b = blockIdx.x;
    const int64_t slot = b < used ? xcd_remap(b, used) * MPW + group : cnt;
    const bool valid = slot < cnt;
    m = valid ? a.list[slot] : 0;
    return valid;
  }
  m = xcd_block() * MPW + group;
  return m < a.G.nmat;
}

template <int NB, int LW>
__global__ __launch_bounds__(64 * reduce_waves<NB>()) void k_gehrd_c(FusedArgs a) {
  using G = Group<double, NB, 1, LAY_DENSE, LW>;
  __shared__ ReduceTiles<NB, LW> sm;
  int64_t m;
  // every wavefront of the workgroup holds the same lane groups
  const bool valid = slot_matrix<G::MPW>(a, (threadIdx.x & 63) / G::LG, m);
  if (!__syncthreads_or(valid)) return;  // workgroup-uniform
  const int n = valid ? a.G.n : 0;
  const int64_t base = valid ? a.G.mat_base(m) : 0, vbase = valid ? a.G.vec_base(m) : 0;
  band_reduce_one<NB, LW>(sm, n, a.G, base, vbase, a.G.n, a.rA, a.iA, a.rV, a.iV, a.rE, a.iE);
}

template <int NB, int LW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(qr_band_waves<NB>()))) void k_qr_band(FusedArgs a) {
  using G = Group<double, NB, 1, LAY_BAND, LW>;
  __shared__ BandTiles<NB, LW> sm;
  const int lane = threadIdx.x;
  int64_t m;
  const bool valid = slot_matrix<G::MPW>(a, lane / G::LG, m);
  if (!__any(valid)) return;
  const int n = valid ? a.G.n : 0;
  const G g = band_group<NB, LW>(sm, n, lane);
  const int64_t base = valid ? a.G.mat_base(m) : 0, vbase = valid ? a.G.vec_base(m) : 0;
  const bool schur = a.flags & VEIG_WRITE_SCHUR;
  const int nfail = band_qr_one<NB, LW>(g, a.G, base, vbase, a.G.n, a.itmax, a.rV, a.iV, a.rE, a.iE,
                                    schur ? a.rA : nullptr, schur ? a.iA : nullptr);
  if (a.info && valid && g.r == 0) a.info[m] = nfail;
}

}  // namespace

namespace veig {

template <int NB, int LW>
int launch_banded(const FusedArgs &a, hipStream_t s) {
  constexpr int mpw = Group<double, NB, 1, LAY_DENSE, LW>::MPW;
  // a list launch gets the whole batch's grid: the list's length is only
  // known on the device, and it never exceeds the batch
  const unsigned grid = (unsigned)((a.G.nmat + mpw - 1) / mpw);
  hipLaunchKernelGGL((k_gehrd_c<NB, LW>), dim3(grid), dim3(64 * reduce_waves<NB>()), 0, s, a);
  hipLaunchKernelGGL((k_qr_band<NB, LW>), dim3(grid), dim3(64), 0, s, a);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}
template int launch_banded<12, 16>(const FusedArgs &, hipStream_t);
template int launch_banded<16, 0>(const FusedArgs &, hipStream_t);
template int launch_banded<24, 32>(const FusedArgs &, hipStream_t);
template int launch_banded<28, 32>(const FusedArgs &, hipStream_t);
template int launch_banded<32, 0>(const FusedArgs &, hipStream_t);

}  // namespace veig
