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
template <int NB, int LW = 0>
struct BandTiles {
  using G = Group<double, NB, 1, LAY_BAND, LW>;
  static constexpr int kTile = G::kSize + 1;  // +1: each group starts on another bank
  static constexpr int kVec = NB + 1;
  double2 H[G::MPW * kTile];
  double2 v[G::MPW * kVec];
};

// wavefronts per SIMD the QR kernel's register allocation is held to: what
// the LDS allows (11.8 KB per wavefront at NB = 16, 11.3 KB for two order-24
// slots, 19.5 KB at NB = 32) where the registers can follow (three up to
// NB = 16; two otherwise, the Schur-vector rows being 24 to 32 entries)
template <int NB>
constexpr int qr_band_waves() { return NB <= 16 ? 3 : 2; }

// Wavefronts of a workgroup that share one reduction, and the LDS it needs
// beside the dense tiles: one partial product per wavefront and lane.  At
// NB = 16 one wavefront (gehrd_compact) measured faster than two sharing
// the work (7.6 vs 6.2 ms for 2^20 matrices): the tiles already allow two
// wavefronts per SIMD there, against one at NB = 32 and half at NB = 64.
// Slots narrower than their lane group (NB = 12 on 16 lanes) take the
// shared form too, whose loads and stores skip the lanes past the slot.
template <int NB>
constexpr int reduce_waves() { return NB == 64 ? 4 : (NB == 16 ? 1 : 2); }

template <int NB, int LW = 0>
struct ReduceTiles {
  Tiles<double, NB, 1, LW> t;
  double2 part[reduce_waves<NB>() > 1 ? reduce_waves<NB>() * 64 : 1];
};

// The Hessenberg reduction of gehrd_q (zgehrd with zlarfg's LAPACK sign,
// Toolbox.hpp:1315-1469), leaving the compact form: v(i+2:n) below the
// subdiagonal, tau(i) in tau_out of lane i.  The W = reduce_waves<NB>()
// wavefronts of the workgroup hold the same lane groups and split each update
// by chunks of CH = NB / W columns (right update) or rows (left update): a
// wavefront takes its chunk of the lane's row or column and the reflector into
// registers, puts its partial product in LDS, and after a barrier adds the W
// partials in a fixed order and updates its chunk.  Every wavefront forms the
// reflector from column i (the same arithmetic on the same data); wavefront 0
// writes it.  Five workgroup barriers per reflector; the loop bound nmax must
// be the same in every wavefront of the workgroup.
template <int NB, int LW = 0>
__device__ __forceinline__ void gehrd_split(const Group<double, NB, 1, LAY_DENSE, LW> &g, int nmax,
                                            cx<double> &tau_out, int wq, double2 *part) {
  using R = double;
  constexpr int LG = Group<R, NB, 1, LAY_DENSE, LW>::LG;
  constexpr int W = reduce_waves<NB>();
  constexpr int CH = NB / W;
  const int n = g.n, r = g.r, lane = threadIdx.x & 63;
  const bool in = r < n;
  const int c0 = wq * CH;
  const int rr = r < NB ? r : NB - 1;  // lanes past the slot's capacity read a valid entry and use none
  for (int i = 0; i + 1 < nmax; ++i) {
    // ---- reflector annihilating H(i+2:n, i) ----
    const cx<R> hri = in ? ld<R>(g.at(r, i)) : mk<R>(0, 0);
    const R xnorm2 = group_sum<LG, R>(r >= i + 2 && in ? abs2(hri) : R(0));
    const bool live = i + 1 < n;
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
    cx<R> vr = mk<R>(0, 0);
    if (live) {
      if (r == i + 1) vr = mk<R>(1, 0);
      else if (r >= i + 2 && in) vr = mul(hri, scal);
    }
    if (r == i) tau_out = tau;
    __syncthreads();  // every wavefront has read column i
    if (wq == 0 && r < NB) {
      st<R>(g.vsh + r, vr);
      if (live && in && r >= i + 1) st<R>(g.at(r, i), r == i + 1 ? mk<R>(beta, 0) : vr);
    }
    __syncthreads();
    const bool mine = c0 + CH - 1 > i;  // this wavefront's chunk reaches past column i
    // ---- H <- H (I - tau v v^H): row r, columns of the chunk ----
    cx<R> h[CH], v[CH];
    if (live && mine) {
      cx<R> w = mk<R>(0, 0);
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        h[j] = ld<R>(g.at(rr, c0 + j));
        v[j] = ld<R>(g.vsh + c0 + j);  // zero at c <= i and beyond n
      }
#pragma unroll
      for (int j = 0; j < CH; ++j)
        if (c0 + j > i) w = add(w, mul(h[j], v[j]));
      st<R>(part + wq * 64 + lane, w);
    } else {
      st<R>(part + wq * 64 + lane, mk<R>(0, 0));
    }
    __syncthreads();
    if (live && mine) {
      cx<R> w = mk<R>(0, 0);
#pragma unroll
      for (int q = 0; q < W; ++q) w = add(w, ld<R>(part + q * 64 + lane));
      w = mul(tau, w);
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int c = c0 + j;
        if (c > i && c < n && in) st<R>(g.at(r, c), sub(h[j], mulc(w, v[j])));
      }
    }
    __syncthreads();
    // ---- H <- (I - conj(tau) v v^H) H: column r, rows of the chunk ----
    if (live && mine) {
      cx<R> y = mk<R>(0, 0);
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        h[j] = ld<R>(g.at(c0 + j, rr));
        v[j] = ld<R>(g.vsh + c0 + j);
      }
#pragma unroll
      for (int j = 0; j < CH; ++j)
        if (c0 + j > i) y = add(y, cmul(v[j], h[j]));
      st<R>(part + wq * 64 + lane, y);
    } else {
      st<R>(part + wq * 64 + lane, mk<R>(0, 0));
    }
    __syncthreads();
    if (live && mine) {
      cx<R> y = mk<R>(0, 0);
#pragma unroll
      for (int q = 0; q < W; ++q) y = add(y, ld<R>(part + q * 64 + lane));
      y = cmul(tau, y);  // conj(tau) * y
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int k = c0 + j;
        if (k > i && k < n && r > i && in) st<R>(g.at(k, r), sub(h[j], mul(v[j], y)));
      }
    }
    __syncthreads();
  }
}

// Reduction of the lane group's matrix (order n; nmax = the largest order in
// the workgroup) into its compact form in the V and E planes, by the W
// wavefronts of the workgroup (wavefront index wq).  Every thread of the
// workgroup calls it; the tile needs no zeroing beforehand.
template <int NB, int LW = 0>
__device__ __forceinline__ void band_reduce_one(ReduceTiles<NB, LW> &sm, int n, const Geom &G, int64_t base,
                                                int64_t vbase, int nmax, const double *rA, const double *iA,
                                                double *rV, double *iV, double *rE, double *iE) {
  using Gd = Group<double, NB, 1, LAY_DENSE, LW>;
  constexpr int CH = NB / reduce_waves<NB>();
  const int lane = threadIdx.x & 63, wq = threadIdx.x >> 6;
  Gd g;
  g.mi = lane / Gd::LG;
  g.r = lane % Gd::LG;
  g.n = n;
  g.H = sm.t.H + g.mi * Tiles<double, NB, 1, LW>::kTile;
  g.vsh = sm.t.v + g.mi * Tiles<double, NB, 1, LW>::kVec;
  const int r = g.r, c0 = wq * CH;
  cx<double> tau = mk<double>(0, 0);
  if constexpr (reduce_waves<NB>() == 1) {
    // one wavefront: the whole row in flight at once (this kernel's LDS,
    // not its registers, limits its occupancy; a runtime loop over the row
    // waited for each pair of loads in turn: 6% of its time), entries
    // outside the matrix as zeros
    cx<double> row[1][NB];
    load_row_regs<double, NB, 1>(g, G, base, rA, iA, row);
#pragma unroll
    for (int c = 0; c < NB; ++c) st<double>(g.at(r, c), row[0][c]);
    st<double>(g.vsh + r, mk<double>(0, 0));
    wave_sync();
    gehrd_compact<NB>(g, nmax, tau);
    store_rows<double, NB, 1>(g, G, base, rV, iV);
  } else {
    // row r, the wavefront's chunk of columns: the matrix, zeros outside it
    if (r < NB) {
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int c = c0 + j;
        cx<double> x = mk<double>(0, 0);
        if (r < n && c < n) {
          const int64_t idx = base + G.el(r, c);
          x = mk<double>(rA[idx], iA[idx]);
        }
        st<double>(g.at(r, c), x);
      }
      if (wq == 0) st<double>(g.vsh + r, mk<double>(0, 0));
    }
    __syncthreads();
    gehrd_split<NB>(g, nmax, tau, wq, sm.part);
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int c = c0 + j;
      if (r < n && c < n) {
        const int64_t idx = base + G.el(r, c);
        const cx<double> x = ld<double>(g.at(r, c));
        rV[idx] = x.re;
        iV[idx] = x.im;
      }
    }
  }
  if (wq == 0 && r + 1 < n) {
    rE[vbase + G.ve(r)] = tau.re;
    iE[vbase + G.ve(r)] = tau.im;
  }
  __syncthreads();  // the tile is free for the workgroup's next matrices
}

// From the compact form: Q (zunghr, Toolbox.hpp:1470-1659), the QR iteration
// with Schur vectors (zlahqr_schur, Toolbox.hpp:690-1031) and the eigenvector
// solve (ztrevc, Toolbox.hpp:1143-1311) on the banded tile; V and E
// overwrite the compact form; T to rT/iT if given.  Returns the count of
// eigenvalues accepted at the iteration cap.  g: the group's banded tile and
// vector (g.H, g.vsh, g.r, g.n set; the tile is zeroed here).
template <int NB, int LW = 0>
__device__ __forceinline__ int band_qr_one(const Group<double, NB, 1, LAY_BAND, LW> &g, const Geom &G, int64_t base,
                                           int64_t vbase, int nmax, int itmax, double *rV, double *iV,
                                           double *rE, double *iE, double *rT, double *iT) {
  using Gb = Group<double, NB, 1, LAY_BAND, LW>;
  const int r = g.r, n = g.n;
  for (int i = r; i < Gb::kSize; i += Gb::LG) st<double>(g.H + i, mk<double>(0, 0));
  wave_sync();
  // the Hessenberg part of the compact form, row r: at NB <= 16 the loads
  // are unrolled so that they are all in flight at once
  if constexpr (NB <= 16) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (r < n && c + 1 >= r && c < n) {
        const int64_t idx = base + G.el(r, c);
        st<double>(g.at(r, c), mk<double>(rV[idx], iV[idx]));
      }
  } else if (r < n) {
    for (int c = r > 0 ? r - 1 : 0; c < n; ++c) {
      const int64_t idx = base + G.el(r, c);
      st<double>(g.at(r, c), mk<double>(rV[idx], iV[idx]));
    }
  }
  cx<double> z[1][NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) z[0][c] = mk<double>(c == r ? 1.0 : 0.0, 0);
  // Q = H_0 H_1 ... H_{n-2}: lane r stages v(r) of reflector i in the
  // group's vector, then every lane applies Z <- Z - tau (Z v) v^H to its
  // row.  The next reflector's v(r) and tau are loaded during this one.
  auto fetch = [&](int i, cx<double> &vl, cx<double> &tau) {
    vl = mk<double>(r == i + 1 ? 1.0 : 0.0, 0);
    tau = mk<double>(0, 0);
    if (i + 1 < n) {
      if (r >= i + 2 && r < n) {
        const int64_t idx = base + G.el(r, i);
        vl = mk<double>(rV[idx], iV[idx]);
      }
      tau = mk<double>(rE[vbase + G.ve(i)], iE[vbase + G.ve(i)]);
    }
  };
  cx<double> vnext, tnext;
  fetch(0, vnext, tnext);
  for (int i = 0; i + 1 < nmax; ++i) {
    const cx<double> vl = vnext, tau = tnext;
    if (i + 2 < nmax) fetch(i + 1, vnext, tnext);
    if (r < NB) st<double>(g.vsh + r, vl);
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
template <int NB, int LW = 0>
__device__ __forceinline__ Group<double, NB, 1, LAY_BAND, LW> band_group(BandTiles<NB, LW> &sm, int n, int lane) {
  using Gb = Group<double, NB, 1, LAY_BAND, LW>;
  Gb g;
  g.mi = lane / Gb::LG;
  g.r = lane % Gb::LG;
  g.n = n;
  g.H = sm.H + g.mi * BandTiles<NB, LW>::kTile;
  g.vsh = sm.v + g.mi * BandTiles<NB, LW>::kVec;
  return g;
}

}  // namespace veig
