// Orders 33..64: one matrix per workgroup of two wavefronts.
//
// At n = 64 a lane owning a row of the unitary factor needs 64 complex
// doubles (256 registers) next to everything else, which does not fit.  So
// the work is split by data:
//   * wavefront 0 (the Hessenberg engine) owns the Hessenberg / Schur matrix
//     in LDS: it generates every reflector, applies it to H and runs the QR
//     iteration (the same arithmetic and order as the lane-group kernels);
//   * wavefront 1 (the Schur-vector engine) owns the unitary factor, one row
//     per lane in registers: it applies each Hessenberg reflector (Q <- Q H_i)
//     and then, one sweep behind, replays the rotations that wavefront 0
//     records for every bulge step (and the phase that makes H(I, I-1) real),
//     and finally forms V = Z X.
// The engines meet at one workgroup barrier per Hessenberg step and one per
// QR sweep; reflectors and rotation records are double-buffered in LDS so
// that neither engine waits for the other inside a sweep.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "veig_band.hpp"

using namespace veig;

namespace {

constexpr int kN = 64;

// One QR sweep as the Schur-vector engine needs it: the bulge-chase
// reflectors for rows K = L..I-1 (t1, v2 of each) and, if has_phase, the
// phase that made H(I, I-1) real, applied to column I.
template <typename R>
struct SweepLog {
  using L2 = typename lds2<R>::type;
  L2 t1[kN];
  L2 v2[kN];
  L2 phase;
  int32_t L, I, has_phase, last;
};

template <typename R>
struct LargeShared {
  using L2 = typename lds2<R>::type;
  L2 H[kN * (kN + 1)];
  L2 v[2][kN];         // Hessenberg reflectors (double-buffered)
  L2 tau[2];
  SweepLog<R> log[2];  // sweeps (double-buffered)
};

// The Hessenberg engine's QR iteration (lahqr with one matrix per wavefront),
// recording the Schur-vector updates instead of applying them.  Ends every
// sweep (and the iteration) at a workgroup barrier; the last one is flagged.
template <typename R, int LAY, typename Shared>
__device__ int lahqr_recorded(const Group<R, kN, 1, LAY> &g, Shared &sh, int itmax) {
  using L2 = typename lds2<R>::type;
  const int r = g.r, n = g.n;
  const R ulp = limits<R>::eps;
  const R smlnum = limits<R>::tiny * (R(n > 0 ? n : 1) / ulp);
  int I = n - 1, L = 0, its = 0, nfail = 0;
  bool done = (n <= 0);
  int buf = 0;
  for (;;) {
    // ---- bookkeeping until the matrix needs a sweep or is done ----
    bool ready = false;
    while (!done && !ready) {
      bool neg = false;
      const int k = r;
      if (its <= itmax && k > L && k <= I) {
        const cx<R> h = ld<R>(g.at(k, k - 1));
        if (cabs1(h) <= smlnum) {
          neg = true;
        } else {
          const cx<R> hkk1 = ld<R>(g.at(k - 1, k - 1)), hkk = ld<R>(g.at(k, k));
          R tst = cabs1(hkk1) + cabs1(hkk);
          if (tst == R(0)) {
            if (k - 2 >= 0) tst += fabs(ld<R>(g.at(k - 1, k - 2)).re);
            if (k + 1 <= n - 1) tst += fabs(ld<R>(g.at(k + 1, k)).re);
          }
          if (fabs(h.re) <= ulp * tst) {
            const cx<R> hup = ld<R>(g.at(k - 1, k));
            const R ab = fmax(cabs1(h), cabs1(hup)), ba = fmin(cabs1(h), cabs1(hup));
            const R d3 = cabs1(sub(hkk1, hkk));
            const R aa = fmax(cabs1(hkk), d3), bb = fmin(cabs1(hkk), d3);
            const R s = aa + ab;
            neg = ba * (ab / s) <= fmax(smlnum, ulp * (bb * (aa / s)));
          }
        }
      }
      const uint64_t bits = __ballot(neg);
      if (its > itmax) {
        ++nfail;  // accepted without convergence (the reference's silent case)
        --I;
        L = 0;
        its = 0;
        if (I < 0) done = true;
      } else {
        if (bits) L = 63 - __clzll(bits);
        if (L > 0 && r == 0) st<R>(g.at(L, L - 1), mk<R>(0, 0));
        if (L >= I) {
          --I;
          L = 0;
          its = 0;
          if (I < 0) done = true;
        } else {
          ready = true;
        }
      }
      wave_sync();
    }
    if (!ready) break;

    // ---- shift and first column (Toolbox.hpp:896-1015) ----
    cx<R> t;
    if (its == 10) {
      const R s = R(0.75) * fabs(ld<R>(g.at(L + 1, L)).re);
      t = ld<R>(g.at(L, L));
      t.re = s + t.re;
    } else if (its == 20) {
      const R s = R(0.75) * fabs(ld<R>(g.at(I, I - 1)).re);
      t = ld<R>(g.at(I, I));
      t.re = s + t.re;
    } else {
      t = ld<R>(g.at(I, I));
      cx<R> u = mul(csqrt(ld<R>(g.at(I - 1, I))), csqrt(ld<R>(g.at(I, I - 1))));
      R s = cabs1(u);
      if (s != R(0)) {
        cx<R> x = scale(sub(ld<R>(g.at(I - 1, I - 1)), t), R(0.5));
        const R sx = cabs1(x);
        s = fmax(s, sx);
        const R is = frcp(s);
        const cx<R> xs = scale(x, is), us = scale(u, is);
        cx<R> y = scale(csqrt(add(mul(xs, xs), mul(us, us))), s);
        if (sx > R(0) && x.re * y.re + x.im * y.im < R(0)) y = mk(-y.re, -y.im);
        t = sub(t, mul(cdiv(u, add(x, y)), u));
      }
    }
    const cx<R> h11s = sub(ld<R>(g.at(L, L)), t);
    const R h21 = ld<R>(g.at(L + 1, L)).re;
    const R s0 = cabs1(h11s) + fabs(h21);
    const R is0 = frcp(s0);
    const cx<R> a0 = scale(h11s, is0), a1 = mk<R>(h21 * is0, 0);

    // ---- the sweep (Toolbox.hpp:138-335), rotations recorded ----
    SweepLog<R> &lg = sh.log[buf];
    for (int K = L; K < I; ++K) {
      cx<R> b0 = a0, b1 = a1;
      if (K > L) {
        b0 = ld<R>(g.at(K, K - 1));
        b1 = ld<R>(g.at(K + 1, K - 1));
      }
      cx<R> t1, v2, nb0;
      larfg2(b0, b1, t1, v2, nb0);
      const R t2 = t1.re * v2.re - t1.im * v2.im;  // Re(t1 * v2)
      if (r == 0) {
        if (K > L) {
          st<R>(g.at(K, K - 1), nb0);
          st<R>(g.at(K + 1, K - 1), mk<R>(0, 0));
        }
        st<R>(&lg.t1[K], t1);
        st<R>(&lg.v2[K], v2);
      }
      if (r >= K && r < n) {  // rows K, K+1 over columns K..n-1
        L2 *p1 = g.at(K, r);
        L2 *p2 = g.at(K + 1, r);
        cx<R> h1 = ld<R>(p1), h2 = ld<R>(p2);
        const cx<R> sum = add(cmul(t1, h1), scale(h2, t2));
        h1 = sub(h1, sum);
        h2 = sub(h2, mul(sum, v2));
        st<R>(p1, h1);
        st<R>(p2, h2);
      }
      wave_sync();
      const int i3 = (K + 2) < I ? (K + 2) : I;
      if (r <= i3) {  // columns K, K+1 over rows 0..min(K+2, I)
        L2 *p1 = g.at(r, K);
        L2 *p2 = g.at(r, K + 1);
        cx<R> h1 = ld<R>(p1), h2 = ld<R>(p2);
        const cx<R> sum = add(mul(t1, h1), scale(h2, t2));
        h1 = sub(h1, sum);
        h2 = sub(h2, mulc(sum, v2));
        st<R>(p1, h1);
        st<R>(p2, h2);
      }
      wave_sync();
    }
    // ---- make H(I, I-1) real (Toolbox.hpp:337-367) ----
    bool has_phase = false;
    {
      const cx<R> h = ld<R>(g.at(I, I - 1));
      if (h.im != R(0) && abs2(h) > R(0)) {
        has_phase = true;
        const R id = frsq(abs2(h));
        const R d = abs2(h) * id;
        const cx<R> zz = mk(h.re * id, -h.im * id);  // conj(h/|h|)
        if (r == 0) {
          st<R>(g.at(I, I - 1), mk<R>(d, 0));
          st<R>(&lg.phase, zz);
        }
        if (r > I && r < n) st<R>(g.at(I, r), mul(ld<R>(g.at(I, r)), zz));
        if (r < I) st<R>(g.at(r, I), mulc(ld<R>(g.at(r, I)), zz));
      }
    }
    ++its;
    if (r == 0) {
      lg.L = L;
      lg.I = I;
      lg.has_phase = has_phase;
      lg.last = 0;
    }
    __syncthreads();  // hand the sweep's record to the Schur-vector engine
    buf ^= 1;
  }
  if (r == 0) {
    sh.log[buf].L = 0;
    sh.log[buf].I = 0;
    sh.log[buf].has_phase = 0;
    sh.log[buf].last = 1;
  }
  __syncthreads();
  return nfail;
}

// The Schur-vector engine's side of the QR iteration: replay each recorded
// sweep on the row of Z held by this lane, until the record flagged last.
// A sweep's reflectors act on consecutive column pairs (K, K+1), K = L..I-1,
// so the replay is a fully unrolled pass over K with a wavefront-uniform
// guard: every z[] index is a compile-time constant and Z stays in registers.
template <typename R>
__device__ void replay_sweeps(const LargeShared<R> &sh, cx<R> (&z)[kN]) {
  int buf = 0;
  for (;;) {
    __syncthreads();
    const SweepLog<R> &lg = sh.log[buf];
    const int L = lg.L, I = lg.I;
#pragma unroll
    for (int K = 0; K + 1 < kN; ++K) {
      if (K >= L && K < I) {
        const cx<R> t1 = ld<R>(&lg.t1[K]), v2 = ld<R>(&lg.v2[K]);
        const R t2 = t1.re * v2.re - t1.im * v2.im;
        const cx<R> sum = add(mul(t1, z[K]), scale(z[K + 1], t2));
        z[K] = sub(z[K], sum);
        z[K + 1] = sub(z[K + 1], mulc(sum, v2));
      }
    }
    if (lg.has_phase) {
      const cx<R> ph = ld<R>(&lg.phase);
#pragma unroll
      for (int c = 1; c < kN; ++c)
        if (c == I) z[c] = mulc(z[c], ph);
    }
    const int last = lg.last;
    buf ^= 1;
    if (last) break;
  }
}

// Hessenberg engine, reduction step i: the reflector annihilating
// H(i+2:n, i) (zlarfg, LAPACK sign) is published to the Schur-vector engine
// at a workgroup barrier, then applied to H from both sides.
template <typename R>
__device__ void gehrd_step(const Group<R, kN, 1> &g, LargeShared<R> &sh, int i) {
  using L2 = typename lds2<R>::type;
  L2 *vs = sh.v[i & 1];
  const int r = g.r, n = g.n;
  const bool in = r < n;
  const cx<R> hri = in ? ld<R>(g.at(r, i)) : mk<R>(0, 0);
  const R xnorm2 = group_sum<kN>((r >= i + 2 && in) ? abs2(hri) : R(0));
  const cx<R> alpha = ld<R>(g.at(i + 1, i));
  cx<R> tau = mk<R>(0, 0), scal = mk<R>(0, 0);
  R beta = alpha.re;
  if (!(xnorm2 == R(0) && alpha.im == R(0))) {
    const R nrm = fsqrt(abs2(alpha) + xnorm2);
    beta = alpha.re >= R(0) ? -nrm : nrm;
    const R ib = frcp(beta);
    tau = mk((beta - alpha.re) * ib, -alpha.im * ib);
    scal = inv(mk(alpha.re - beta, alpha.im));
  }
  cx<R> vr = mk<R>(0, 0);
  if (r == i + 1) vr = mk<R>(1, 0);
  else if (r >= i + 2 && in) vr = mul(hri, scal);
  st<R>(vs + r, vr);
  if (in && r >= i + 1) st<R>(g.at(r, i), r == i + 1 ? mk<R>(beta, 0) : mk<R>(0, 0));
  if (r == 0) st<R>(&sh.tau[i & 1], tau);
  __syncthreads();  // the reflector goes to the Schur-vector engine
  // H <- H (I - tau v v^H), rows 0..n-1, columns i+1..n-1
  if (in) {
    cx<R> wv = mk<R>(0, 0);
    for (int c = i + 1; c < n; ++c) wv = add(wv, mul(ld<R>(g.at(r, c)), ld<R>(vs + c)));
    const cx<R> tw = mul(tau, wv);
    for (int c = i + 1; c < n; ++c) st<R>(g.at(r, c), sub(ld<R>(g.at(r, c)), mulc(tw, ld<R>(vs + c))));
  }
  wave_sync();
  // H <- (I - conj(tau) v v^H) H, rows i+1..n-1, columns i+1..n-1
  if (in && r >= i + 1) {
    cx<R> y = mk<R>(0, 0);
    for (int k = i + 1; k < n; ++k) y = add(y, cmul(ld<R>(vs + k), ld<R>(g.at(k, r))));
    const cx<R> ty = cmul(tau, y);
    for (int k = i + 1; k < n; ++k) st<R>(g.at(k, r), sub(ld<R>(g.at(k, r)), mul(ld<R>(vs + k), ty)));
  }
  wave_sync();
}

// Schur-vector engine, reduction step i: Q <- Q (I - tau v v^H) on the row
// of Q held by this lane.
template <typename R>
__device__ void q_step(const LargeShared<R> &sh, cx<R> (&z)[kN], int i) {
  using L2 = typename lds2<R>::type;
  __syncthreads();
  const L2 *vs = sh.v[i & 1];
  const cx<R> tau = ld<R>(&sh.tau[i & 1]);
  cx<R> wv = mk<R>(0, 0);
#pragma unroll
  for (int c = 0; c < kN; ++c) wv = add(wv, mul(z[c], ld<R>(vs + c)));
  const cx<R> tw = mul(tau, wv);
  // re-read v rather than keep 64 loaded values live across the sum
  int zoff = 0;
  asm volatile("" : "+s"(zoff));
  const L2 *vs2 = vs + zoff;
#pragma unroll
  for (int c = 0; c < kN; ++c) z[c] = sub(z[c], mulc(tw, ld<R>(vs2 + c)));
}

// One matrix on the workgroup (w = wavefront index, 0 or 1).  Every
// workgroup barrier below is reached by both wavefronts the same number of
// times: once per reduction step, once per QR sweep plus once at the end of
// the iteration, once when the eigenvector factors are ready.
//   OUT_V:  E, V = Z X (and T over A with VEIG_WRITE_SCHUR);
//   OUT_VW: E, V = Z X and W = X^{-1} Z^H ~ V^{-1} in the complex64 planes
//           rW/iW (the input of the FP64 refinement).
template <typename R, int OUT>
__device__ __forceinline__ void solve_one(const FusedArgs &a, LargeShared<R> &sh, int64_t m, int w, int lane) {
  using G = Group<R, kN, 1>;
  const int n = a.G.n, nmax = a.G.n;
  const int64_t base = a.G.mat_base(m), vbase = a.G.vec_base(m);
  G g;
  g.mi = 0;
  g.r = lane;
  g.n = n;
  g.H = sh.H;
  g.vsh = sh.v[0];
  if (w == 0) {
#pragma unroll
    for (int c = 0; c < kN; ++c) st<R>(g.at(lane, c), mk<R>(0, 0));
    wave_sync();
    load_rows<R, kN, 1>(g, a.G, base, a.rA, a.iA, kN);
  }
  __syncthreads();

  if (w == 0) {
    // ---------------- Hessenberg engine ----------------
    for (int i = 0; i + 1 < nmax; ++i) gehrd_step<R>(g, sh, i);
    const int nfail = lahqr_recorded(g, sh, a.itmax);
    if (OUT == OUT_V && (a.flags & VEIG_WRITE_SCHUR)) store_rows<R, kN, 1>(g, a.G, base, a.rA, a.iA);
    const cx<R> lam = trevc_solve_lds<R, kN>(g, nmax);
    if (lane < n) {
      if constexpr (OUT == OUT_V) {
        // tile column r becomes -X(:, r): y above the diagonal, -1 on it, 0 below
        st<R>(g.at(lane, lane), mk<R>(-1, 0));
        for (int i = lane + 1; i < n; ++i) st<R>(g.at(i, lane), mk<R>(0, 0));
      } else {
        // X above the diagonal (x = -y)
        for (int i = 0; i < lane; ++i) {
          const cx<R> y = ld<R>(g.at(i, lane));
          st<R>(g.at(i, lane), mk(-y.re, -y.im));
        }
      }
      a.rE[vbase + a.G.ve(lane)] = (double)lam.re;
      a.iE[vbase + a.G.ve(lane)] = (double)lam.im;
    }
    if constexpr (OUT == OUT_VW) {
      wave_sync();
      inverse_unit_upper<R, kN, 1>(g, nmax);  // Y = X^{-1} below the diagonal
    }
    if (a.info && lane == 0) a.info[m] = nfail;
    __syncthreads();  // the eigenvector factors are ready
  } else {
    // ---------------- Schur-vector engine ----------------
    cx<R> z[1][kN];
#pragma unroll
    for (int c = 0; c < kN; ++c) z[0][c] = mk<R>(c == lane ? R(1) : R(0), 0);
    for (int i = 0; i + 1 < nmax; ++i) q_step<R>(sh, z[0], i);
    replay_sweeps<R>(sh, z[0]);
    __syncthreads();  // the eigenvector factors are ready
    if constexpr (OUT == OUT_V) {
      // row `lane` of V = Z X = -Z (-X), column by column, straight to V
      if (lane < n) {
        for (int k = 0; k < n; ++k) {
          cx<R> acc = mk<R>(0, 0);
#pragma unroll
          for (int j = 0; j < kN; ++j)
            if (j < nmax) acc = add(acc, mul(z[0][j], ld<R>(g.at(j, k))));
          const int64_t idx = base + a.G.el(lane, k);
          a.rV[idx] = -(double)acc.re;
          a.iV[idx] = -(double)acc.im;
        }
      }
    } else {
      // column `lane` of W: W(k, r) = conj(z[k]) + sum_{j>k} Y(k,j) conj(z[j]),
      // Y(k, j) at tile position (j, k)
#pragma unroll
      for (int k = 0; k < kN; ++k) {
        if (k < nmax) {
          cx<R> acc = conj(z[0][k]);
#pragma unroll
          for (int j = k + 1; j < kN; ++j)
            if (j < nmax) acc = add(acc, mulc(ld<R>(g.at(j, k)), z[0][j]));
          if (lane < n) {
            const int64_t idx = base + a.G.el(k, lane);
            a.rW[idx] = (float)acc.re;
            a.iW[idx] = (float)acc.im;
          }
        }
      }
      backtransform<R, kN, 1>(g, z, nmax);  // row `lane` of V = Z X, in place
      if (lane < n) {
#pragma unroll
        for (int k = 0; k < kN; ++k) {
          if (k < n) {
            const int64_t idx = base + a.G.el(lane, k);
            a.rV[idx] = (double)z[0][k].re;
            a.iV[idx] = (double)z[0][k].im;
          }
        }
      }
    }
  }
}

// the batch, one matrix per workgroup
template <typename R, int OUT>
__global__ __launch_bounds__(128) void k_zgeev64(FusedArgs a) {
  __shared__ LargeShared<R> sh;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t m = xcd_block();
  if (m >= a.G.nmat) return;  // workgroup-uniform
  solve_one<R, OUT>(a, sh, m, w, lane);
}

// the matrices of a.list, on a persistent grid
template <typename R, int OUT>
__global__ __launch_bounds__(128) void k_zgeev64_list(FusedArgs a) {
  __shared__ LargeShared<R> sh;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t cnt = *a.list_count;
  for (int64_t i = blockIdx.x; i < cnt; i += gridDim.x) {
    // an opaque copy of the lane: addresses derived from it are recomputed
    // for every matrix instead of being hoisted out of the loop
    int ln = lane;
    asm volatile("" : "+v"(ln));
    solve_one<R, OUT>(a, sh, a.list[i], w, ln);
    __syncthreads();  // the tile and logs are reused by the next matrix
  }
}

// ---------------------------------------------------------------------------
// Orders 33..64 in FP64, whole batches: two kernels.
//
// k_gehrd64 (four wavefronts per matrix on its dense tile, each owning 16
// columns / rows of every update): the Hessenberg reduction alone, in
// zgehrd's compact form -- H on and above the subdiagonal, the reflector
// vectors v(i+2:n) below it, tau(i) -- written to the V and E planes, which
// serve as its workspace.
//
// k_qr64 (one workgroup of three wavefronts per matrix): the QR iteration and
// the eigenvectors with only the band row <= col + 2 of H in LDS (35 KB
// instead of 66 KB) and the Schur-vector rows split over two wavefronts, two
// lanes per row (each lane holds half a row, 32 complex values): wavefront 0
// is the Hessenberg engine of k_zgeev64 on the banded tile; wavefronts 1 and
// 2 (rows 0..31, 32..63) first form Q from the reflectors (Q <- Q H_i, the
// row's two halves combined by a lane swap), then replay each recorded sweep;
// the reflector on columns (31, 32) pairs the two halves of a row through the
// same swap.  About 40 KB of LDS and ~170 registers per wavefront let four
// matrices share a CU, against two for k_zgeev64.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * reduce_waves<kN>()) void k_gehrd64(FusedArgs a) {
  __shared__ ReduceTiles<kN> sm;
  const int64_t m = xcd_block();
  if (m >= a.G.nmat) return;  // workgroup-uniform
  band_reduce_one<kN>(sm, a.G.n, a.G, a.G.mat_base(m), a.G.vec_base(m), a.G.n, a.rA, a.iA, a.rV, a.iV,
                      a.rE, a.iE);
}

using GBand = Group<double, kN, 1, LAY_BAND>;

struct QrShared64 {
  double2 H[GBand::kSize + 1];
  SweepLog<double> log[2];
};

// the other half of this lane's row (lane ^ 32)
__device__ __forceinline__ cx<double> half_swap(cx<double> x) {
  return mk<double>(__shfl_xor(x.re, 32), __shfl_xor(x.im, 32));
}

// Schur-vector engine of k_qr64: this lane's half row zh[c] = Z(row, 32 h + c).
__device__ __forceinline__ void replay_sweeps_half(const QrShared64 &sh, int h, cx<double> (&zh)[32]) {
  int buf = 0;
  for (;;) {
    __syncthreads();
    const SweepLog<double> &lg = sh.log[buf];
    const int L = lg.L, I = lg.I;
#pragma unroll
    for (int K = 0; K + 1 < kN; ++K) {
      if (K >= L && K < I) {
        const cx<double> t1 = ld<double>(&lg.t1[K]), v2 = ld<double>(&lg.v2[K]);
        const double t2 = t1.re * v2.re - t1.im * v2.im;
        if (K + 1 < 32) {
          if (h == 0) {
            const cx<double> sum = add(mul(t1, zh[K]), scale(zh[K + 1], t2));
            zh[K] = sub(zh[K], sum);
            zh[K + 1] = sub(zh[K + 1], mulc(sum, v2));
          }
        } else if (K >= 32) {
          if (h == 1) {
            const cx<double> sum = add(mul(t1, zh[K - 32]), scale(zh[K - 31], t2));
            zh[K - 32] = sub(zh[K - 32], sum);
            zh[K - 31] = sub(zh[K - 31], mulc(sum, v2));
          }
        } else {
          // columns 31 and 32: one in each half of the row
          const cx<double> mine = h == 0 ? zh[31] : zh[0];
          const cx<double> other = half_swap(mine);
          const cx<double> z31 = h == 0 ? mine : other, z32 = h == 0 ? other : mine;
          const cx<double> sum = add(mul(t1, z31), scale(z32, t2));
          if (h == 0) zh[31] = sub(z31, sum);
          else zh[0] = sub(z32, mulc(sum, v2));
        }
      }
    }
    if (lg.has_phase) {
      const cx<double> ph = ld<double>(&lg.phase);
#pragma unroll
      for (int c = 1; c < kN; ++c)
        if (c == I && h == c / 32) zh[c % 32] = mulc(zh[c % 32], ph);
    }
    const int last = lg.last;
    buf ^= 1;
    if (last) break;
  }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// The back-transform V = Z X of ztrevc (Toolbox.hpp:1268-1275) for the 32
// rows row0..row0+31 of Z this wavefront holds as half rows (lane l: row
// row0 + l % 32, columns 32 (l / 32) .. +31 in zh), on the FP64 matrix
// cores: v_mfma_f64_16x16x4f64 multiplies a 16 x 4 slice of Z (A operand:
// lane (c, q) = (l % 16, l / 16) supplies Z(16 I + c, k0 + q)) by a 4 x 16
// slice of X (B operand: X(k0 + q, 16 J + c), read from the tile, where the
// solve left y = -x above the diagonal) into a 16 x 16 block of V (lane
// (c, q), slot s: V(16 I + 4 s + q, 16 J + c)).  The complex product is four
// real products.  The A operand's entry lives in lane 16 I + c + 32 (k0 / 32)
// of the wavefront, register k0 % 32 + q: four lane permutes per slice (one
// per q) and a select.  X is unit upper triangular: column block J needs the
// row blocks K <= J only.  The result goes straight to the V planes.
__device__ __forceinline__ void backtransform_mfma(const GBand &g, const cx<double> (&zh)[32], int row0, int lane,
                                                   int64_t base, const FusedArgs &a) {
  const int n = g.n;
  const int c = lane & 15, q = lane >> 4;
  // one 16 x 16 block of V at a time (runtime loops: Z's 32 half-row entries
  // take most of the registers, so only one block's accumulators stay live)
  for (int blk = 0; blk < 8; ++blk) {
    const int I = blk & 1, J = blk >> 1;
    if (16 * J >= n) break;  // wavefront-uniform
    d4 vre = d4{0, 0, 0, 0}, vim = d4{0, 0, 0, 0};
#pragma unroll
    for (int K = 0; K < 4; ++K) {
      if (K > J) break;  // X is upper triangular
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = 16 * K + 4 * s + q, j = 16 * J + c;
        cx<double> x = mk<double>(k == j ? 1.0 : 0.0, 0);
        if (k < j) {
          const cx<double> y = ld<double>(g.at(k, j));
          x = mk<double>(-y.re, -y.im);
        }
        const int src = 16 * I + c + 32 * (K >> 1);
        cx<double> za = mk<double>(0, 0);
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const cx<double> zr = zh[16 * (K & 1) + 4 * s + qq];
          const double re = __shfl(zr.re, src), im = __shfl(zr.im, src);
          if (qq == q) za = mk<double>(re, im);
        }
        vre = __builtin_amdgcn_mfma_f64_16x16x4f64(za.re, x.re, vre, 0, 0, 0);
        vre = __builtin_amdgcn_mfma_f64_16x16x4f64(-za.im, x.im, vre, 0, 0, 0);
        vim = __builtin_amdgcn_mfma_f64_16x16x4f64(za.re, x.im, vim, 0, 0, 0);
        vim = __builtin_amdgcn_mfma_f64_16x16x4f64(za.im, x.re, vim, 0, 0, 0);
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int row = row0 + 16 * I + 4 * s + q, col = 16 * J + c;
      if (row < n && col < n) {
        const int64_t idx = base + a.G.el(row, col);
        a.rV[idx] = vre[s];
        a.iV[idx] = vim[s];
      }
    }
  }
}

__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3))) void k_qr64(FusedArgs a) {
  __shared__ QrShared64 sh;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t m = xcd_block();
  if (m >= a.G.nmat) return;  // workgroup-uniform
  const int n = a.G.n, nmax = a.G.n;
  const int64_t base = a.G.mat_base(m), vbase = a.G.vec_base(m);
  GBand g;
  g.mi = 0;
  g.r = lane;
  g.n = n;
  g.H = sh.H;
  g.vsh = nullptr;
  if (w == 0) {
    // the Hessenberg part of the reduction's output, into the band
    for (int i = lane; i < GBand::kSize; i += 64) st<double>(sh.H + i, mk<double>(0, 0));
    wave_sync();
    if (lane < n)
      for (int c = lane > 0 ? lane - 1 : 0; c < n; ++c) {
        const int64_t idx = base + a.G.el(lane, c);
        st<double>(g.at(lane, c), mk<double>(a.rV[idx], a.iV[idx]));
      }
  }
  __syncthreads();
  if (w == 0) {
    // ---------------- Hessenberg engine ----------------
    const int nfail = lahqr_recorded(g, sh, a.itmax);
    if (a.flags & VEIG_WRITE_SCHUR) {
      if (lane < n)
        for (int c = 0; c < n; ++c) {
          const int64_t idx = base + a.G.el(lane, c);
          const cx<double> t = c >= lane ? ld<double>(g.at(lane, c)) : mk<double>(0, 0);
          a.rA[idx] = t.re;
          a.iA[idx] = t.im;
        }
    }
    const cx<double> lam = trevc_solve_lds(g, nmax);  // leaves y = -x above the diagonal
    if (lane < n) {
      a.rE[vbase + a.G.ve(lane)] = lam.re;
      a.iE[vbase + a.G.ve(lane)] = lam.im;
    }
    if (a.info && lane == 0) a.info[m] = nfail;
    __syncthreads();  // the eigenvector factor is ready
  } else {
    // ---------------- Schur-vector engines: half a row per lane ----------------
    const int h = lane >> 5;
    const int row = (w - 1) * 32 + (lane & 31);
    cx<double> zh[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) zh[c] = mk<double>(32 * h + c == row ? 1.0 : 0.0, 0);
    // Q = H_0 H_1 ... H_{n-2} from the reflectors the reduction left in V and
    // E: lane l loads v(l) of reflector i, and each lane fetches the 32
    // entries of its half row from the other lanes (a lane permute, no LDS),
    // once for w = Z v and again for the update Z -= tau w v^H
    for (int i = 0; i + 1 < nmax; ++i) {
      cx<double> vl = mk<double>(lane == i + 1 ? 1.0 : 0.0, 0);
      if (lane >= i + 2 && lane < n) {
        const int64_t idx = base + a.G.el(lane, i);
        vl = mk<double>(a.rV[idx], a.iV[idx]);
      }
      const cx<double> tau = mk<double>(a.rE[vbase + a.G.ve(i)], a.iE[vbase + a.G.ve(i)]);
      int src = 32 * h;
      asm volatile("" : "+v"(src));
      cx<double> wv = mk<double>(0, 0);
#pragma unroll
      for (int c = 0; c < 32; ++c)
        wv = add(wv, mul(zh[c], mk<double>(__shfl(vl.re, src + c), __shfl(vl.im, src + c))));
      wv = add(wv, half_swap(wv));
      const cx<double> tw = mul(tau, wv);
      asm volatile("" : "+v"(src));  // permute again rather than keep 32 values live
#pragma unroll
      for (int c = 0; c < 32; ++c)
        zh[c] = sub(zh[c], mulc(tw, mk<double>(__shfl(vl.re, src + c), __shfl(vl.im, src + c))));
    }
    replay_sweeps_half(sh, h, zh);
    __syncthreads();  // the eigenvector factor is ready
    backtransform_mfma(g, zh, (w - 1) * 32, lane, base, a);
  }
}

}  // namespace

namespace veig {
template <typename R, int OUT>
int launch_zgeev64(const FusedArgs &a, hipStream_t s) {
  if (a.G.nmat > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  if (a.mask) return fail(VEIG_ERR_UNSUPPORTED, "orders above 32: no masked launches");
  if (a.list) {
    // the FP64 solve of the matrices the mixed path hands back
    if constexpr (std::is_same<R, double>::value && OUT == OUT_V) {
      const int64_t g = a.G.nmat < kListGrid ? a.G.nmat : kListGrid;
      hipLaunchKernelGGL((k_zgeev64_list<R, OUT>), dim3((unsigned)g), dim3(128), 0, s, a);
    } else {
      return fail(VEIG_ERR_UNSUPPORTED, "orders above 32: list launches are FP64 only");
    }
  } else if constexpr (std::is_same<R, double>::value && OUT == OUT_V) {
    hipLaunchKernelGGL(k_gehrd64, dim3((unsigned)a.G.nmat), dim3(64 * reduce_waves<kN>()), 0, s, a);
    hipLaunchKernelGGL(k_qr64, dim3((unsigned)a.G.nmat), dim3(192), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_zgeev64<R, OUT>), dim3((unsigned)a.G.nmat), dim3(128), 0, s, a);
  }
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}
template int launch_zgeev64<double, OUT_V>(const FusedArgs &, hipStream_t);
template int launch_zgeev64<float, OUT_V>(const FusedArgs &, hipStream_t);
template int launch_zgeev64<float, OUT_VW>(const FusedArgs &, hipStream_t);
}  // namespace veig
