// Per-matrix bodies of the two-kernel FP64 eigensolver (orders up to 32):
// the Hessenberg reduction written in zgehrd's compact form, then Q, the QR
// iteration and the eigenvectors on a banded tile.  Shared by the batch and
// list kernels (veig_band.hip) and the size-binned kernels of uneven batches
// (veig_varsize.hip), which differ only in how a lane group finds its matrix
// and its order.
//
// Addressing: element (r, c) of the group's matrix at base + G.el(r, c),
// vector entry r at vbase + G.ve(r), with G.n the matrix's own order.  The V
// and E planes carry the compact form between the two kernels: H on and above
// the subdiagonal with v(i+2:n) below it in V, tau(i) in E(i).
#pragma once
#include "veig_common.hpp"

namespace veig {

// LDS of a wavefront of the QR kernel: the band row <= col + 2 of each
// group's tile and the reflector vector used while Q is formed
template <int NB>
struct BandTiles {
  using G = Group<double, NB, 1, LAY_BAND>;
  static constexpr int kTile = G::kSize + 1;  // +1: each group starts on another bank
  static constexpr int kVec = NB + 1;
  double2 H[G::MPW * kTile];
  double2 v[G::MPW * kVec];
};

// wavefronts per SIMD the QR kernel's register allocation is held to: what
// the LDS allows (11.8 KB per wavefront at NB = 16, 19.5 KB at NB = 32)
template <int NB>
constexpr int qr_band_waves() { return NB == 16 ? 3 : 2; }

// Reduction of the group's matrix (order g.n; nmax = the largest order in
// the wavefront) into its compact form in the V and E planes.  g from
// make_group (tile zeroed); every lane of the wavefront calls it.
template <int NB>
__device__ __forceinline__ void band_reduce_one(const Group<double, NB, 1> &g, const Geom &G, int64_t base,
                                                int64_t vbase, int nmax, const double *rA, const double *iA,
                                                double *rV, double *iV, double *rE, double *iE) {
  wave_sync();
  load_rows<double, NB, 1>(g, G, base, rA, iA, NB);
  wave_sync();
  cx<double> tau = mk<double>(0, 0);
  gehrd_compact<NB>(g, nmax, tau);
  store_rows<double, NB, 1>(g, G, base, rV, iV);
  if (g.r + 1 < g.n) {
    rE[vbase + G.ve(g.r)] = tau.re;
    iE[vbase + G.ve(g.r)] = tau.im;
  }
}

// From the compact form: Q (zunghr, Toolbox.hpp:1470-1659), the QR iteration
// with Schur vectors (zlahqr_schur, Toolbox.hpp:690-1031) and the eigenvector
// solve (ztrevc, Toolbox.hpp:1143-1311) on the banded tile; V and E
// overwrite the compact form; T to rT/iT if given.  Returns the count of
// eigenvalues accepted at the iteration cap.  g: the group's banded tile and
// vector (g.H, g.vsh, g.r, g.n set; the tile is zeroed here).
template <int NB>
__device__ __forceinline__ int band_qr_one(const Group<double, NB, 1, LAY_BAND> &g, const Geom &G, int64_t base,
                                           int64_t vbase, int nmax, int itmax, double *rV, double *iV,
                                           double *rE, double *iE, double *rT, double *iT) {
  using Gb = Group<double, NB, 1, LAY_BAND>;
  const int r = g.r, n = g.n;
  for (int i = r; i < Gb::kSize; i += Gb::LG) st<double>(g.H + i, mk<double>(0, 0));
  wave_sync();
  // the Hessenberg part of the compact form, row r
  if (r < n)
    for (int c = r > 0 ? r - 1 : 0; c < n; ++c) {
      const int64_t idx = base + G.el(r, c);
      st<double>(g.at(r, c), mk<double>(rV[idx], iV[idx]));
    }
  cx<double> z[1][NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) z[0][c] = mk<double>(c == r ? 1.0 : 0.0, 0);
  // Q = H_0 H_1 ... H_{n-2}: lane r stages v(r) of reflector i in the
  // group's vector, then every lane applies Z <- Z - tau (Z v) v^H to its row
  for (int i = 0; i + 1 < nmax; ++i) {
    cx<double> vl = mk<double>(r == i + 1 ? 1.0 : 0.0, 0);
    cx<double> tau = mk<double>(0, 0);
    if (i + 1 < n) {
      if (r >= i + 2 && r < n) {
        const int64_t idx = base + G.el(r, i);
        vl = mk<double>(rV[idx], iV[idx]);
      }
      tau = mk<double>(rE[vbase + G.ve(i)], iE[vbase + G.ve(i)]);
    }
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
  const int nfail = lahqr<double, NB, 1, true, LAY_BAND>(g, z, itmax, false);
  if (rT) {
    if (r < n)
      for (int c = 0; c < n; ++c) {
        const int64_t idx = base + G.el(r, c);
        const cx<double> t = c >= r ? ld<double>(g.at(r, c)) : mk<double>(0, 0);
        rT[idx] = t.re;
        iT[idx] = t.im;
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
        const int64_t idx = base + G.el(r, c);
        rV[idx] = z[0][c].re;
        iV[idx] = z[0][c].im;
      }
    }
    rE[vbase + G.ve(r)] = lam.re;
    iE[vbase + G.ve(r)] = lam.im;
  }
  return nfail;
}

// the banded group of lane `lane` in a wavefront's BandTiles
template <int NB>
__device__ __forceinline__ Group<double, NB, 1, LAY_BAND> band_group(BandTiles<NB> &sm, int n, int lane) {
  using Gb = Group<double, NB, 1, LAY_BAND>;
  Gb g;
  g.mi = lane / Gb::LG;
  g.r = lane % Gb::LG;
  g.n = n;
  g.H = sm.H + g.mi * BandTiles<NB>::kTile;
  g.vsh = sm.v + g.mi * BandTiles<NB>::kVec;
  return g;
}

}  // namespace veig
