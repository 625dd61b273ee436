// FP64 batches (and lists of matrices from a batch) at orders 9..32 in two
// kernels, so that the QR iteration runs on a banded tile and more matrices
// stay resident per CU.
//
//   k_gehrd_c<NB>   the Hessenberg reduction alone (zgehrd, Toolbox.hpp:
//                   1315-1469) on the dense tile, written in zgehrd's compact
//                   form -- H on and above the subdiagonal, the reflector
//                   vectors v(i+2:n) below it, tau(i) -- to the V and E
//                   planes, which serve as its workspace.
//   k_qr_band<NB>   Q from those reflectors (zunghr, Toolbox.hpp:1470-1659),
//                   the QR iteration with Schur vectors (zlahqr_schur,
//                   Toolbox.hpp:690-1031) and the eigenvectors (ztrevc,
//                   Toolbox.hpp:1143-1311) on the band row <= col + 2 of the
//                   tile: every entry a bulge chase, the deflation test and
//                   the triangular solve touch lies in it.
//
// At NB = 16 the band holds 166 of the dense tile's 272 entries: 11.8 KB of
// LDS per wavefront of four matrices instead of 18.6 KB, which with at most
// 168 registers lets three wavefronts share a SIMD instead of two; at
// NB = 32, 19.5 KB per wavefront of two instead of 33.8 KB: two wavefronts
// per SIMD instead of one (at 256 registers, part of the Schur-vector rows
// spilled).  The
// reduction round trip through HBM (the compact form written and read back)
// costs 2 x 16 n^2 bytes per matrix.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "veig_common.hpp"

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

template <int NB>
__global__ __launch_bounds__(64) void k_gehrd_c(FusedArgs a) {
  using G = Group<double, NB, 1>;
  __shared__ Tiles<double, NB, 1> sm;
  int64_t m;
  const bool valid = slot_matrix<G::MPW>(a, threadIdx.x / G::LG, m);
  if (!__any(valid)) return;
  const int n = valid ? a.G.n : 0, nmax = a.G.n;
  G g = make_group<double, NB, 1>(sm, n, threadIdx.x);
  const int64_t base = valid ? a.G.mat_base(m) : 0, vbase = valid ? a.G.vec_base(m) : 0;
  wave_sync();
  load_rows<double, NB, 1>(g, a.G, base, a.rA, a.iA, NB);
  wave_sync();
  cx<double> tau = mk<double>(0, 0);
  gehrd_compact<NB>(g, nmax, tau);
  store_rows<double, NB, 1>(g, a.G, base, a.rV, a.iV);
  if (g.r + 1 < n) {
    a.rE[vbase + a.G.ve(g.r)] = tau.re;
    a.iE[vbase + a.G.ve(g.r)] = tau.im;
  }
}

template <int NB>
struct BandTiles {
  using G = Group<double, NB, 1, LAY_BAND>;
  static constexpr int kTile = G::kSize + 1;  // +1: each group starts on another bank
  static constexpr int kVec = NB + 1;
  double2 H[G::MPW * kTile];
  double2 v[G::MPW * kVec];  // the reflector being applied while Q is formed
};

// wavefronts per SIMD the register allocator is held to: what the LDS
// allows (11.8 KB per wavefront at NB = 16, 19.5 KB at NB = 32)
template <int NB>
constexpr int qr_band_waves() { return NB == 16 ? 3 : 2; }

template <int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(qr_band_waves<NB>()))) void k_qr_band(FusedArgs a) {
  using G = Group<double, NB, 1, LAY_BAND>;
  using Tl = BandTiles<NB>;
  __shared__ Tl sm;
  const int lane = threadIdx.x;
  int64_t m;
  const bool valid = slot_matrix<G::MPW>(a, lane / G::LG, m);
  if (!__any(valid)) return;
  const int n = valid ? a.G.n : 0, nmax = a.G.n;
  G g;
  g.mi = lane / G::LG;
  g.r = lane % G::LG;
  g.n = n;
  g.H = sm.H + g.mi * Tl::kTile;
  g.vsh = sm.v + g.mi * Tl::kVec;
  const int r = g.r;
  for (int i = r; i < G::kSize; i += G::LG) st<double>(g.H + i, mk<double>(0, 0));
  const int64_t base = valid ? a.G.mat_base(m) : 0, vbase = valid ? a.G.vec_base(m) : 0;
  wave_sync();
  // the Hessenberg part of the reduction's output, row r
  if (r < n)
    for (int c = r > 0 ? r - 1 : 0; c < n; ++c) {
      const int64_t idx = base + a.G.el(r, c);
      st<double>(g.at(r, c), mk<double>(a.rV[idx], a.iV[idx]));
    }
  cx<double> z[1][NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) z[0][c] = mk<double>(c == r ? 1.0 : 0.0, 0);
  // Q = H_0 H_1 ... H_{n-2}: lane r stages v(r) of reflector i in the
  // group's vector, then every lane applies Z <- Z - tau (Z v) v^H to its row
  for (int i = 0; i + 1 < nmax; ++i) {
    cx<double> vl = mk<double>(r == i + 1 ? 1.0 : 0.0, 0);
    if (r >= i + 2 && r < n) {
      const int64_t idx = base + a.G.el(r, i);
      vl = mk<double>(a.rV[idx], a.iV[idx]);
    }
    const cx<double> tau = mk<double>(a.rE[vbase + a.G.ve(i)], a.iE[vbase + a.G.ve(i)]);
    st<double>(g.vsh + r, vl);
    wave_sync();
    cx<double> w = mk<double>(0, 0);
#pragma unroll
    for (int c = 0; c < NB; ++c) w = add(w, mul(z[0][c], ld<double>(g.vsh + c)));
    const cx<double> tw = mul(tau, w);
    int off = 0;
    asm volatile("" : "+v"(off));  // read v again rather than keep NB values live
    const double2 *vs = g.vsh + off;
#pragma unroll
    for (int c = 0; c < NB; ++c) z[0][c] = sub(z[0][c], mulc(tw, ld<double>(vs + c)));
    wave_sync();  // every lane has read v before the next one is staged
  }
  wave_sync();
  const int nfail = lahqr<double, NB, 1, true, LAY_BAND>(g, z, a.itmax, false);
  if (a.flags & VEIG_WRITE_SCHUR) {
    if (r < n)
      for (int c = 0; c < n; ++c) {
        const int64_t idx = base + a.G.el(r, c);
        const cx<double> t = c >= r ? ld<double>(g.at(r, c)) : mk<double>(0, 0);
        a.rA[idx] = t.re;
        a.iA[idx] = t.im;
      }
  }
  const cx<double> lam = trevc_solve_lds<double, NB, LAY_BAND>(g, nmax);  // leaves y = -x above the diagonal
  // V(:, k) = Z [x; 1; 0] = z_k - sum_{j<k} z_j y(j, k), in place, from the last column
#pragma unroll
  for (int k = NB - 1; k >= 1; --k) {
    if (k < nmax) {
      cx<double> acc = z[0][k];
#pragma unroll
      for (int j = 0; j < k; ++j) acc = sub(acc, mul(z[0][j], ld<double>(g.at(j, k))));
      z[0][k] = acc;
    }
  }
  if (r < n) {
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (c < n) {
        const int64_t idx = base + a.G.el(r, c);
        a.rV[idx] = z[0][c].re;
        a.iV[idx] = z[0][c].im;
      }
    }
    a.rE[vbase + a.G.ve(r)] = lam.re;
    a.iE[vbase + a.G.ve(r)] = lam.im;
  }
  if (a.info && valid && r == 0) a.info[m] = nfail;
}

}  // namespace

namespace veig {

template <int NB>
int launch_banded(const FusedArgs &a, hipStream_t s) {
  constexpr int mpw = Group<double, NB, 1>::MPW;
  // a list launch gets the whole batch's grid: the list's length is only
  // known on the device, and it never exceeds the batch
  const unsigned grid = (unsigned)((a.G.nmat + mpw - 1) / mpw);
  hipLaunchKernelGGL(k_gehrd_c<NB>, dim3(grid), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_qr_band<NB>, dim3(grid), dim3(64), 0, s, a);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}
template int launch_banded<16>(const FusedArgs &, hipStream_t);
template int launch_banded<32>(const FusedArgs &, hipStream_t);

}  // namespace veig
