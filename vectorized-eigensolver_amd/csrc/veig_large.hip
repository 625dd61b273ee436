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

#include "veig_common.hpp"

using namespace veig;

namespace {

constexpr int kN = 64;

// One QR sweep as the Schur-vector engine needs it: the bulge-chase
// reflectors for rows K = L..I-1 (t1, v2 of each) and, if has_phase, the
// phase that made H(I, I-1) real, applied to column I.
struct SweepLog {
  double2 t1[kN];
  double2 v2[kN];
  double2 phase;
  int32_t L, I, has_phase, last;
};

struct LargeShared {
  double2 H[kN * (kN + 1)];
  double2 v[2][kN];   // Hessenberg reflectors (double-buffered)
  double2 tau[2];
  SweepLog log[2];    // sweeps (double-buffered)
};

using Gr = Group<double, kN, 1>;

// The Hessenberg engine's QR iteration (lahqr with one matrix per wavefront),
// recording the Schur-vector updates instead of applying them.  Ends every
// sweep (and the iteration) at a workgroup barrier; the last one is flagged.
__device__ int lahqr_recorded(const Gr &g, LargeShared &sh, int itmax) {
  const int r = g.r, n = g.n;
  const double ulp = limits<double>::eps;
  const double smlnum = limits<double>::tiny * (double(n > 0 ? n : 1) / ulp);
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
        const cx<double> h = ld<double>(g.at(k, k - 1));
        if (cabs1(h) <= smlnum) {
          neg = true;
        } else {
          const cx<double> hkk1 = ld<double>(g.at(k - 1, k - 1)), hkk = ld<double>(g.at(k, k));
          double tst = cabs1(hkk1) + cabs1(hkk);
          if (tst == 0.0) {
            if (k - 2 >= 0) tst += fabs(ld<double>(g.at(k - 1, k - 2)).re);
            if (k + 1 <= n - 1) tst += fabs(ld<double>(g.at(k + 1, k)).re);
          }
          if (fabs(h.re) <= ulp * tst) {
            const cx<double> hup = ld<double>(g.at(k - 1, k));
            const double ab = fmax(cabs1(h), cabs1(hup)), ba = fmin(cabs1(h), cabs1(hup));
            const double d3 = cabs1(sub(hkk1, hkk));
            const double aa = fmax(cabs1(hkk), d3), bb = fmin(cabs1(hkk), d3);
            const double s = aa + ab;
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
        if (L > 0 && r == 0) st<double>(g.at(L, L - 1), mk<double>(0, 0));
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
    cx<double> t;
    if (its == 10) {
      const double s = 0.75 * fabs(ld<double>(g.at(L + 1, L)).re);
      t = ld<double>(g.at(L, L));
      t.re = s + t.re;
    } else if (its == 20) {
      const double s = 0.75 * fabs(ld<double>(g.at(I, I - 1)).re);
      t = ld<double>(g.at(I, I));
      t.re = s + t.re;
    } else {
      t = ld<double>(g.at(I, I));
      cx<double> u = mul(csqrt(ld<double>(g.at(I - 1, I))), csqrt(ld<double>(g.at(I, I - 1))));
      double s = cabs1(u);
      if (s != 0.0) {
        cx<double> x = scale(sub(ld<double>(g.at(I - 1, I - 1)), t), 0.5);
        const double sx = cabs1(x);
        s = fmax(s, sx);
        const double is = frcp(s);
        const cx<double> xs = scale(x, is), us = scale(u, is);
        cx<double> y = scale(csqrt(add(mul(xs, xs), mul(us, us))), s);
        if (sx > 0.0 && x.re * y.re + x.im * y.im < 0.0) y = mk(-y.re, -y.im);
        t = sub(t, mul(cdiv(u, add(x, y)), u));
      }
    }
    const cx<double> h11s = sub(ld<double>(g.at(L, L)), t);
    const double h21 = ld<double>(g.at(L + 1, L)).re;
    const double s0 = cabs1(h11s) + fabs(h21);
    const double is0 = frcp(s0);
    const cx<double> a0 = scale(h11s, is0), a1 = mk<double>(h21 * is0, 0);

    // ---- the sweep (Toolbox.hpp:138-335), rotations recorded ----
    SweepLog &lg = sh.log[buf];
    for (int K = L; K < I; ++K) {
      cx<double> b0 = a0, b1 = a1;
      if (K > L) {
        b0 = ld<double>(g.at(K, K - 1));
        b1 = ld<double>(g.at(K + 1, K - 1));
      }
      cx<double> t1, v2, nb0;
      larfg2(b0, b1, t1, v2, nb0);
      const double t2 = t1.re * v2.re - t1.im * v2.im;  // Re(t1 * v2)
      if (r == 0) {
        if (K > L) {
          st<double>(g.at(K, K - 1), nb0);
          st<double>(g.at(K + 1, K - 1), mk<double>(0, 0));
        }
        lg.t1[K] = double2{t1.re, t1.im};
        lg.v2[K] = double2{v2.re, v2.im};
      }
      if (r >= K && r < n) {  // rows K, K+1 over columns K..n-1
        double2 *p1 = g.at(K, r);
        double2 *p2 = g.at(K + 1, r);
        cx<double> h1 = ld<double>(p1), h2 = ld<double>(p2);
        const cx<double> sum = add(cmul(t1, h1), scale(h2, t2));
        h1 = sub(h1, sum);
        h2 = sub(h2, mul(sum, v2));
        st<double>(p1, h1);
        st<double>(p2, h2);
      }
      wave_sync();
      const int i3 = (K + 2) < I ? (K + 2) : I;
      if (r <= i3) {  // columns K, K+1 over rows 0..min(K+2, I)
        double2 *p1 = g.at(r, K);
        double2 *p2 = g.at(r, K + 1);
        cx<double> h1 = ld<double>(p1), h2 = ld<double>(p2);
        const cx<double> sum = add(mul(t1, h1), scale(h2, t2));
        h1 = sub(h1, sum);
        h2 = sub(h2, mulc(sum, v2));
        st<double>(p1, h1);
        st<double>(p2, h2);
      }
      wave_sync();
    }
    // ---- make H(I, I-1) real (Toolbox.hpp:337-367) ----
    bool has_phase = false;
    {
      const cx<double> h = ld<double>(g.at(I, I - 1));
      if (h.im != 0.0 && abs2(h) > 0.0) {
        has_phase = true;
        const double id = frsq(abs2(h));
        const double d = abs2(h) * id;
        const cx<double> zz = mk(h.re * id, -h.im * id);  // conj(h/|h|)
        if (r == 0) {
          st<double>(g.at(I, I - 1), mk<double>(d, 0));
          lg.phase = double2{zz.re, zz.im};
        }
        if (r > I && r < n) st<double>(g.at(I, r), mul(ld<double>(g.at(I, r)), zz));
        if (r < I) st<double>(g.at(r, I), mulc(ld<double>(g.at(r, I)), zz));
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
__device__ void replay_sweeps(const LargeShared &sh, cx<double> (&z)[kN]) {
  int buf = 0;
  for (;;) {
    __syncthreads();
    const SweepLog &lg = sh.log[buf];
    const int L = lg.L, I = lg.I;
#pragma unroll
    for (int K = 0; K + 1 < kN; ++K) {
      if (K >= L && K < I) {
        const cx<double> t1 = ld<double>(&lg.t1[K]), v2 = ld<double>(&lg.v2[K]);
        const double t2 = t1.re * v2.re - t1.im * v2.im;
        const cx<double> sum = add(mul(t1, z[K]), scale(z[K + 1], t2));
        z[K] = sub(z[K], sum);
        z[K + 1] = sub(z[K + 1], mulc(sum, v2));
      }
    }
    if (lg.has_phase) {
      const cx<double> ph = ld<double>(&lg.phase);
#pragma unroll
      for (int c = 1; c < kN; ++c)
        if (c == I) z[c] = mulc(z[c], ph);
    }
    const int last = lg.last;
    buf ^= 1;
    if (last) break;
  }
}

__global__ __launch_bounds__(128) void k_zgeev64(FusedArgs a) {
  __shared__ LargeShared sh;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t m = xcd_block();  // one matrix per workgroup
  if (m >= a.G.nmat) return;      // workgroup-uniform
  const int n = a.G.n, nmax = a.G.n;
  const int64_t base = a.G.mat_base(m), vbase = a.G.vec_base(m);
  Gr g;
  g.mi = 0;
  g.r = lane;
  g.n = n;
  g.H = sh.H;
  g.vsh = sh.v[0];
  if (w == 0) {
#pragma unroll
    for (int c = 0; c < kN; ++c) st<double>(g.at(lane, c), mk<double>(0, 0));
    wave_sync();
    load_rows<double, kN, 1>(g, a.G, base, a.rA, a.iA, kN);
  }
  __syncthreads();

  if (w == 0) {
    // ---------------- Hessenberg engine ----------------
    for (int i = 0; i + 1 < nmax; ++i) {
      double2 *vs = sh.v[i & 1];
      const int r = lane;
      const bool in = r < n;
      const cx<double> hri = in ? ld<double>(g.at(r, i)) : mk<double>(0, 0);
      const double xnorm2 = group_sum<kN>((r >= i + 2 && in) ? abs2(hri) : 0.0);
      const cx<double> alpha = ld<double>(g.at(i + 1, i));
      cx<double> tau = mk<double>(0, 0), scal = mk<double>(0, 0);
      double beta = alpha.re;
      if (!(xnorm2 == 0.0 && alpha.im == 0.0)) {
        const double nrm = fsqrt(abs2(alpha) + xnorm2);
        beta = alpha.re >= 0.0 ? -nrm : nrm;
        const double ib = frcp(beta);
        tau = mk((beta - alpha.re) * ib, -alpha.im * ib);
        scal = inv(mk(alpha.re - beta, alpha.im));
      }
      cx<double> vr = mk<double>(0, 0);
      if (r == i + 1) vr = mk<double>(1, 0);
      else if (r >= i + 2 && in) vr = mul(hri, scal);
      st<double>(vs + r, vr);
      if (in && r >= i + 1) st<double>(g.at(r, i), r == i + 1 ? mk<double>(beta, 0) : mk<double>(0, 0));
      if (r == 0) st<double>(&sh.tau[i & 1], tau);
      __syncthreads();  // the reflector goes to the Schur-vector engine
      // H <- H (I - tau v v^H), rows 0..n-1, columns i+1..n-1
      if (in) {
        cx<double> wv = mk<double>(0, 0);
        for (int c = i + 1; c < n; ++c) wv = add(wv, mul(ld<double>(g.at(r, c)), ld<double>(vs + c)));
        const cx<double> tw = mul(tau, wv);
        for (int c = i + 1; c < n; ++c)
          st<double>(g.at(r, c), sub(ld<double>(g.at(r, c)), mulc(tw, ld<double>(vs + c))));
      }
      wave_sync();
      // H <- (I - conj(tau) v v^H) H, rows i+1..n-1, columns i+1..n-1
      if (in && r >= i + 1) {
        cx<double> y = mk<double>(0, 0);
        for (int k = i + 1; k < n; ++k) y = add(y, cmul(ld<double>(vs + k), ld<double>(g.at(k, r))));
        const cx<double> ty = cmul(tau, y);
        for (int k = i + 1; k < n; ++k) st<double>(g.at(k, r), sub(ld<double>(g.at(k, r)), mul(ld<double>(vs + k), ty)));
      }
      wave_sync();
    }
    const int nfail = lahqr_recorded(g, sh, a.itmax);
    if (a.flags & VEIG_WRITE_SCHUR) store_rows<double, kN, 1>(g, a.G, base, a.rA, a.iA);
    const cx<double> lam = trevc_solve_lds<double, kN>(g, nmax);
    // tile column r becomes -X(:, r): y above the diagonal, -1 on it, 0 below
    if (lane < n) {
      st<double>(g.at(lane, lane), mk<double>(-1, 0));
      for (int i = lane + 1; i < n; ++i) st<double>(g.at(i, lane), mk<double>(0, 0));
      a.rE[vbase + a.G.ve(lane)] = lam.re;
      a.iE[vbase + a.G.ve(lane)] = lam.im;
    }
    if (a.info && lane == 0) a.info[m] = nfail;
    __syncthreads();  // X is ready
  } else {
    // ---------------- Schur-vector engine ----------------
    cx<double> z[kN];
#pragma unroll
    for (int c = 0; c < kN; ++c) z[c] = mk<double>(c == lane ? 1.0 : 0.0, 0);
    for (int i = 0; i + 1 < nmax; ++i) {
      __syncthreads();
      const double2 *vs = sh.v[i & 1];
      const cx<double> tau = ld<double>(&sh.tau[i & 1]);
      // Q <- Q (I - tau v v^H): row `lane` of Q in registers
      cx<double> wv = mk<double>(0, 0);
#pragma unroll
      for (int c = 0; c < kN; ++c) wv = add(wv, mul(z[c], ld<double>(vs + c)));
      const cx<double> tw = mul(tau, wv);
      // re-read v rather than keep 64 loaded values live across the sum
      int zoff = 0;
      asm volatile("" : "+s"(zoff));
      const double2 *vs2 = vs + zoff;
#pragma unroll
      for (int c = 0; c < kN; ++c) z[c] = sub(z[c], mulc(tw, ld<double>(vs2 + c)));
    }
    replay_sweeps(sh, z);
    __syncthreads();  // X is ready
    // row `lane` of V = Z X = -Z (-X), column by column, straight to V
    if (lane < n) {
      for (int k = 0; k < n; ++k) {
        cx<double> acc = mk<double>(0, 0);
#pragma unroll
        for (int j = 0; j < kN; ++j)
          if (j < nmax) acc = add(acc, mul(z[j], ld<double>(g.at(j, k))));
        const int64_t idx = base + a.G.el(lane, k);
        a.rV[idx] = -acc.re;
        a.iV[idx] = -acc.im;
      }
    }
  }
}

}  // namespace

namespace veig {
int launch_zgeev64(const FusedArgs &a, hipStream_t s) {
  if (a.G.nmat > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  if (a.list || a.mask) return fail(VEIG_ERR_UNSUPPORTED, "orders above 32: whole batches only");
  hipLaunchKernelGGL(k_zgeev64, dim3((unsigned)a.G.nmat), dim3(128), 0, s, a);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}
}  // namespace veig
