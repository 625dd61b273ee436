// Orders beyond the lane-group kernels: the fused eigendecomposition at
// 64 < n <= VEIG_MAX_DIM and the stage entries at 32 < n <= VEIG_MAX_DIM.
//
// One workgroup of 256 threads per matrix (a persistent grid walks the
// batch).  The matrix being reduced or iterated lives in LDS as a dense
// n x (n+1) column-major tile (at n = 96: 146 KB, most of the LDS of a CU);
// the unitary factor (Q, the Schur vectors, V) stays in global memory, in the
// caller's output planes, each thread owning one of its rows.  The
// arithmetic is that of the lane-group kernels (veig_device.hpp) spread over
// threads instead of lanes of a group, with workgroup barriers where those
// use wavefront fences:
//   reduction   zgehrd (Toolbox.hpp:1315-1490), LAPACK zlarfg sign, Q
//               accumulated forward (Q <- Q H_i) when wanted;
//   Q alone     zunghr (Toolbox.hpp:1492-1659) as the same forward product;
//   QR          zlahqr_schur (Toolbox.hpp:690-1031): the deflation test,
//               shifts and 30-sweep cap of lahqr, the bulge chase of
//               sweep_step (rows from the left, barrier, columns and Schur
//               vectors from the right, barrier);
//   eigenvectors ztrevc + backsubstitution (Toolbox.hpp:1133-1313, 31-136):
//               thread k solves for eigenvector k in column k of the tile,
//               then thread r forms row r of V = Z X in place.
// These orders are for coverage of the reference's interface (its own
// driver sweeps n = 10..79, zlahqr.cpp:206-249); the bulge chase is a chain
// of barriers and runs far below the lane-group kernels' rate.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "veig_common.hpp"

using namespace veig;

namespace {

constexpr int kThreads = 256;

// element (r, c) of a matrix held in global split planes
struct GMat {
  double *re, *im;
  int64_t base, sr, sc;
  __device__ __forceinline__ int64_t idx(int r, int c) const { return base + r * sr + c * sc; }
  __device__ __forceinline__ cx<double> ld(int r, int c) const {
    const int64_t i = idx(r, c);
    return mk<double>(re[i], im[i]);
  }
  __device__ __forceinline__ void st(int r, int c, cx<double> v) const {
    const int64_t i = idx(r, c);
    re[i] = v.re;
    im[i] = v.im;
  }
};

// the LDS tile: H(r, c) at t[c * ld + r], and a vector of n entries after it
struct Tile {
  double2 *t;
  double2 *v;
  int n, ld;
  __device__ __forceinline__ double2 *at(int r, int c) const { return t + c * ld + r; }
};

// workgroup-wide helpers (every thread calls them)
struct Shared {
  double part[kThreads / 64];
  int flag;
};

__device__ __forceinline__ double block_sum(double x, Shared &sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh.part[w] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int q = 0; q < kThreads / 64; ++q) s += sh.part[q];
  __syncthreads();
  return s;
}

// largest value of `k` over the threads with `pred` (0 if none)
__device__ __forceinline__ int block_max_index(bool pred, int k, Shared &sh) {
  if (threadIdx.x == 0) sh.flag = 0;
  __syncthreads();
  if (pred) atomicMax(&sh.flag, k);
  __syncthreads();
  const int r = sh.flag;
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Hessenberg reduction of the tile (zgehrd with zlarfg's LAPACK sign and its
// x = 0 guard).  keep: leave v(i+2:n) below the subdiagonal and tau(i) in
// tau_out (zgehrd's compact form); else zeros there.  Q (if non-null rows):
// Q <- Q H_i for each reflector, Q in global memory, thread r owning row r.
// ---------------------------------------------------------------------------
__device__ void reduce_tile(const Tile &T, Shared &sh, bool keep, const GMat *Q, double2 *tau_out) {
  const int n = T.n, tid = threadIdx.x;
  for (int i = 0; i + 1 < n; ++i) {
    double part = 0;
    for (int r = i + 2 + tid; r < n; r += kThreads) {
      const double2 h = *T.at(r, i);
      part += h.x * h.x + h.y * h.y;
    }
    const double xnorm2 = block_sum(part, sh);
    const double2 al = *T.at(i + 1, i);
    const cx<double> alpha = mk(al.x, al.y);
    cx<double> tau = mk<double>(0, 0), scal = mk<double>(0, 0);
    double beta = alpha.re;
    if (!(xnorm2 == 0.0 && alpha.im == 0.0)) {
      const double nrm = fsqrt(abs2(alpha) + xnorm2);
      beta = alpha.re >= 0.0 ? -nrm : nrm;
      const double ib = frcp(beta);
      tau = mk((beta - alpha.re) * ib, -alpha.im * ib);
      scal = inv(mk(alpha.re - beta, alpha.im));
    }
    // v(i+1) = 1, v(r) = H(r, i) scal below; v zero elsewhere
    for (int r = tid; r < n; r += kThreads) {
      cx<double> vr = mk<double>(0, 0);
      if (r == i + 1) vr = mk<double>(1, 0);
      else if (r >= i + 2) vr = mul(ld<double>(T.at(r, i)), scal);
      st<double>(T.v + r, vr);
    }
    __syncthreads();
    for (int r = i + 1 + tid; r < n; r += kThreads)
      st<double>(T.at(r, i), r == i + 1 ? mk(beta, 0.0) : (keep ? ld<double>(T.v + r) : mk<double>(0, 0)));
    if (tid == 0 && tau_out) tau_out[i] = double2{tau.re, tau.im};
    // H <- H (I - tau v v^H): row r, columns i+1..n-1
    for (int r = tid; r < n; r += kThreads) {
      cx<double> w = mk<double>(0, 0);
      for (int c = i + 1; c < n; ++c) w = add(w, mul(ld<double>(T.at(r, c)), ld<double>(T.v + c)));
      const cx<double> tw = mul(tau, w);
      for (int c = i + 1; c < n; ++c) st<double>(T.at(r, c), sub(ld<double>(T.at(r, c)), mulc(tw, ld<double>(T.v + c))));
      if (Q) {
        cx<double> wq = mk<double>(0, 0);
        for (int c = i + 1; c < n; ++c) wq = add(wq, mul(Q->ld(r, c), ld<double>(T.v + c)));
        const cx<double> twq = mul(tau, wq);
        for (int c = i + 1; c < n; ++c) Q->st(r, c, sub(Q->ld(r, c), mulc(twq, ld<double>(T.v + c))));
      }
    }
    __syncthreads();
    // H <- (I - conj(tau) v v^H) H: column c, rows i+1..n-1
    for (int c = i + 1 + tid; c < n; c += kThreads) {
      cx<double> y = mk<double>(0, 0);
      for (int k = i + 1; k < n; ++k) y = add(y, cmul(ld<double>(T.v + k), ld<double>(T.at(k, c))));
      const cx<double> ty = cmul(tau, y);
      for (int k = i + 1; k < n; ++k) st<double>(T.at(k, c), sub(ld<double>(T.at(k, c)), mul(ld<double>(T.v + k), ty)));
    }
    __syncthreads();
  }
}

// Q <- Q H_0 H_1 ... H_{n-2} from the compact form in the tile (v(i+2:n)
// below the subdiagonal, v(i+1) = 1) and tau (zunghr's product, formed
// forward), Q in global memory.
__device__ void unghr_tile(const Tile &T, Shared &sh, const GMat &Q, const double2 *tau) {
  const int n = T.n, tid = threadIdx.x;
  (void)sh;
  for (int i = 0; i + 1 < n; ++i) {
    const cx<double> ta = mk(tau[i].x, tau[i].y);
    for (int r = tid; r < n; r += kThreads) {
      auto v = [&](int c) { return c == i + 1 ? mk<double>(1, 0) : ld<double>(T.at(c, i)); };
      cx<double> w = mk<double>(0, 0);
      for (int c = i + 1; c < n; ++c) w = add(w, mul(Q.ld(r, c), v(c)));
      const cx<double> tw = mul(ta, w);
      for (int c = i + 1; c < n; ++c) Q.st(r, c, sub(Q.ld(r, c), mulc(tw, v(c))));
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// QR iteration on the tile (upper Hessenberg, real subdiagonal after
// make_real) with the Schur vectors Z in global memory (thread r: row r,
// n <= kThreads; each thread carries Z(r, K+1) in a register from one bulge
// step to the next and loads the following column one step ahead).  Returns
// the count of eigenvalues accepted at the iteration cap.
// ---------------------------------------------------------------------------
__device__ int lahqr_tile(const Tile &T, Shared &sh, const GMat &Z, int itmax, bool make_real) {
  const int n = T.n, tid = threadIdx.x;
  const double ulp = DBL_EPSILON;
  const double smlnum = DBL_MIN * (double(n > 0 ? n : 1) / ulp);
  auto H = [&](int r, int c) { return ld<double>(T.at(r, c)); };
  if (make_real) {
    for (int i = 1; i < n; ++i) {
      const cx<double> h = H(i, i - 1);
      if (h.im != 0.0) {  // uniform: every thread reads the same entry
        const double d1 = cabs1(h);
        const cx<double> w = scale(h, frcp(d1));
        const cx<double> sc = scale(conj(w), frsq(abs2(w)));  // conj(h/|h|)
        __syncthreads();
        if (tid == 0) st<double>(T.at(i, i - 1), mk(fsqrt(abs2(h)), 0.0));
        for (int x = i + tid; x < n; x += kThreads) st<double>(T.at(i, x), mul(H(i, x), sc));
        for (int x = tid; x <= i + 1 && x < n; x += kThreads)
          if (x != i) st<double>(T.at(x, i), mulc(H(x, i), sc));
        __syncthreads();
        if (tid == 0) st<double>(T.at(i, i), mulc(H(i, i), sc));
        for (int r = tid; r < n; r += kThreads) Z.st(r, i, mulc(Z.ld(r, i), sc));
        __syncthreads();
      }
    }
  }
  int I = n - 1, L = 0, its = 0, nfail = 0;
  while (I >= 0) {
    // ---- deflation: the largest negligible k in L+1..I, as the reference's
    // scan from I down to L+1 finds it (each thread tests its k's) ----
    int best = 0;
    for (int k = L + 1 + tid; k <= I; k += kThreads) {
      const cx<double> h = H(k, k - 1);
      bool nk = cabs1(h) <= smlnum;
      if (!nk) {
        const cx<double> hkk1 = H(k - 1, k - 1), hkk = H(k, k);
        double tst = cabs1(hkk1) + cabs1(hkk);
        if (tst == 0.0) {
          if (k - 2 >= 0) tst += fabs(H(k - 1, k - 2).re);
          if (k + 1 <= n - 1) tst += fabs(H(k + 1, k).re);
        }
        if (fabs(h.re) <= ulp * tst) {
          const cx<double> hup = H(k - 1, k);
          const double ab = fmax(cabs1(h), cabs1(hup)), ba = fmin(cabs1(h), cabs1(hup));
          const double d3 = cabs1(sub(hkk1, hkk));
          const double aa = fmax(cabs1(hkk), d3), bb = fmin(cabs1(hkk), d3);
          const double s = aa + ab;
          nk = ba * (ab / s) <= fmax(smlnum, ulp * (bb * (aa / s)));
        }
      }
      if (nk) best = k;
    }
    const int Lf = block_max_index(best > 0, best, sh);
    if (its > itmax) {
      ++nfail;  // accepted without convergence (the reference's silent case)
      --I;
      L = 0;
      its = 0;
      continue;
    }
    if (Lf > L) {
      L = Lf;
      if (tid == 0) st<double>(T.at(L, L - 1), mk<double>(0, 0));
      __syncthreads();
    }
    if (L >= I) {
      --I;
      L = 0;
      its = 0;
      continue;
    }
    // ---- shift and first column (every thread computes the same) ----
    cx<double> t;
    if (its == 10) {
      t = H(L, L);
      t.re += 0.75 * fabs(H(L + 1, L).re);
    } else if (its == 20) {
      t = H(I, I);
      t.re += 0.75 * fabs(H(I, I - 1).re);
    } else {
      t = H(I, I);
      const cx<double> u2 = mul(H(I - 1, I), H(I, I - 1));
      if (u2.re != 0.0 || u2.im != 0.0) {
        const cx<double> x = scale(sub(H(I - 1, I - 1), t), 0.5);
        const double sx = cabs1(x);
        const double s = fmax(fsqrt(cabs1(u2)), sx);
        const double is = frcp(s);
        const cx<double> xs = scale(x, is);
        cx<double> ys = csqrt(add(mul(xs, xs), scale(u2, is * is)));
        if (sx > 0.0 && xs.re * ys.re + xs.im * ys.im < 0.0) ys = mk(-ys.re, -ys.im);
        t = sub(t, mul(scale(u2, is), inv(add(xs, ys))));
      }
    }
    const cx<double> h11s = sub(H(L, L), t);
    const double h21 = H(L + 1, L).re;
    const double is0 = frcp(cabs1(h11s) + fabs(h21));
    const cx<double> a0 = scale(h11s, is0);
    const double a1 = h21 * is0;
    // ---- the sweep; each thread carries Z(r, K+1) of its row between steps ----
    cx<double> zk1 = tid < n ? Z.ld(tid, L) : mk<double>(0, 0);
    for (int K = L; K < I; ++K) {
      cx<double> b0 = mk(a0.re, a0.im);
      double b1 = a1;
      if (K > L) {
        b0 = H(K, K - 1);
        b1 = H(K + 1, K - 1).re;
      }
      double nb0, nrm2;
      Rot<double> q = rot2(b0, b1, nb0, nrm2);
      if (nrm2 == 0.0) {  // G = I
        q.c = mk<double>(1, 0);
        q.s = 0.0;
        nb0 = 0.0;
      }
      const cx<double> znext = tid < n ? Z.ld(tid, K + 1) : mk<double>(0, 0);
      // rows K, K+1 over columns K..n-1
      for (int c = K + tid; c < n; c += kThreads) {
        cx<double> h1 = H(K, c), h2 = H(K + 1, c);
        rot_left(q, h1, h2);
        st<double>(T.at(K, c), h1);
        st<double>(T.at(K + 1, c), h2);
      }
      __syncthreads();
      if (K > L && tid == kThreads - 1) {
        st<double>(T.at(K, K - 1), mk(nb0, 0.0));
        st<double>(T.at(K + 1, K - 1), mk<double>(0, 0));
      }
      // columns K, K+1 over rows 0..min(K+2, I)
      const int i3 = (K + 2) < I ? (K + 2) : I;
      for (int r = tid; r <= i3; r += kThreads) {
        cx<double> h1 = H(r, K), h2 = H(r, K + 1);
        rot_right(q, h1, h2);
        st<double>(T.at(r, K), h1);
        st<double>(T.at(r, K + 1), h2);
      }
      // Schur vectors: columns K, K+1 of row tid
      if (tid < n) {
        cx<double> z0 = zk1, z1 = znext;
        rot_right(q, z0, z1);
        Z.st(tid, K, z0);
        zk1 = z1;
      }
      __syncthreads();
    }
    if (tid < n) Z.st(tid, I, zk1);
    // ---- make H(I, I-1) real (Toolbox.hpp:337-367) ----
    const cx<double> h = H(I, I - 1);
    __syncthreads();
    if (h.im != 0.0 && abs2(h) > 0.0) {
      const double id = frsq(abs2(h));
      const double d = abs2(h) * id;
      const cx<double> zz = mk(h.re * id, -h.im * id);  // conj(h/|h|)
      if (tid == 0) st<double>(T.at(I, I - 1), mk(d, 0.0));
      for (int x = I + 1 + tid; x < n; x += kThreads) st<double>(T.at(I, x), mul(H(I, x), zz));
      for (int x = tid; x < I; x += kThreads) st<double>(T.at(x, I), mulc(H(x, I), zz));
      for (int r = tid; r < n; r += kThreads) Z.st(r, I, mulc(Z.ld(r, I), zz));
    }
    ++its;
    __syncthreads();
  }
  return nfail;
}

// ---------------------------------------------------------------------------
// Eigenvectors of the triangular tile (ztrevc + backsubstitution): thread k
// solves for eigenvector k with the reference's small-pivot clamp and leaves
// y = -x in column k above the diagonal; then R <- R X in global memory,
// thread r owning row r (column k of R X = R [x; 1; 0], the reference's
// scaling).  Returns nothing; lam(k) = T(k, k) stays on the diagonal.
// ---------------------------------------------------------------------------
__device__ void trevc_tile(const Tile &T, const GMat &R) {
  const int n = T.n, tid = threadIdx.x;
  const double ulp = DBL_EPSILON;
  const double smlnum = DBL_MIN * (double(n > 0 ? n : 1) / ulp);
  for (int j = n - 2; j >= 0; --j) {
    const cx<double> tjj = ld<double>(T.at(j, j));
    for (int r = j + 1 + tid; r < n; r += kThreads) {
      const cx<double> lam = ld<double>(T.at(r, r));
      const double smin = fmax(ulp * cabs1(lam), smlnum);
      cx<double> d = sub(tjj, lam);
      if (cabs1(d) < smin) d = mk(smin, 0.0);
      const cx<double> yj = mul(ld<double>(T.at(j, r)), inv(d));
      st<double>(T.at(j, r), yj);
      for (int i = 0; i < j; ++i) st<double>(T.at(i, r), sub(ld<double>(T.at(i, r)), mul(ld<double>(T.at(i, j)), yj)));
    }
    __syncthreads();
  }
  // R(r, k) <- R(r, k) - sum_{j<k} R(r, j) y(j, k), from the last column
  for (int r = tid; r < n; r += kThreads)
    for (int k = n - 1; k >= 1; --k) {
      cx<double> acc = R.ld(r, k);
      for (int j = 0; j < k; ++j) acc = sub(acc, mul(R.ld(r, j), ld<double>(T.at(j, k))));
      R.st(r, k, acc);
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct BigArgs {
  double *rA, *iA;  // A / H / T (in place for the stages)
  double *rB, *iB;  // V / Q / Z / R
  double *rE, *iE;  // E / tau
  int32_t *info;
  Geom G;
  int itmax;
  int mode;
  uint32_t flags;
};

__device__ GMat gmat(double *re, double *im, const Geom &G, int64_t m) {
  return GMat{re, im, G.mat_base(m), G.T, (int64_t)G.n * G.T};
}

__global__ __launch_bounds__(kThreads) void k_big(BigArgs a) {
  extern __shared__ double2 lds[];
  __shared__ Shared sh;
  const int n = a.G.n, tid = threadIdx.x;
  Tile T;
  T.n = n;
  T.ld = n + 1;
  T.t = lds;
  T.v = lds + n * (n + 1);
  double2 *tau = T.v + n;  // n entries
  for (int64_t m = blockIdx.x; m < a.G.nmat; m += gridDim.x) {
    const GMat A = gmat(a.rA, a.iA, a.G, m);
    const GMat B = gmat(a.rB, a.iB, a.G, m);
    const int64_t vb = a.G.vec_base(m);
    // the matrix into the tile (full; the QR and eigenvector stages read the
    // Hessenberg part or the upper triangle only)
    for (int c = 0; c < n; ++c)
      for (int r = tid; r < n; r += kThreads) {
        cx<double> x = A.ld(r, c);
        if ((a.mode == BIG_LAHQR && r > c + 1) || (a.mode == BIG_TREVC && r > c)) x = mk<double>(0, 0);
        st<double>(T.at(r, c), x);
      }
    if (a.mode == BIG_UNGHR)
      for (int i = tid; i < n; i += kThreads)
        tau[i] = i + 1 < n ? double2{a.rE[vb + a.G.ve(i)], a.iE[vb + a.G.ve(i)]} : double2{0, 0};
    // Q or Z starts as the identity where this kernel forms it
    if (a.mode == BIG_FUSED || a.mode == BIG_HESS || a.mode == BIG_UNGHR) {
      const GMat &Q = a.mode == BIG_UNGHR ? A : B;
      for (int c = 0; c < n; ++c)
        for (int r = tid; r < n; r += kThreads) Q.st(r, c, mk<double>(r == c ? 1.0 : 0.0, 0));
    }
    __syncthreads();
    if (a.mode == BIG_GEHRD || a.mode == BIG_HESS) {
      reduce_tile(T, sh, true, a.mode == BIG_HESS ? &B : nullptr, tau);
      for (int c = 0; c < n; ++c)
        for (int r = tid; r < n; r += kThreads) A.st(r, c, ld<double>(T.at(r, c)));
      if (a.rE)
        for (int i = tid; i + 1 < n; i += kThreads) {
          a.rE[vb + a.G.ve(i)] = tau[i].x;
          a.iE[vb + a.G.ve(i)] = tau[i].y;
        }
    } else if (a.mode == BIG_UNGHR) {
      unghr_tile(T, sh, A, tau);
    } else if (a.mode == BIG_LAHQR) {
      const int nfail = lahqr_tile(T, sh, B, a.itmax, true);
      for (int c = 0; c < n; ++c)
        for (int r = tid; r < n; r += kThreads) A.st(r, c, r <= c ? ld<double>(T.at(r, c)) : mk<double>(0, 0));
      for (int r = tid; r < n; r += kThreads) {
        const double2 d = *T.at(r, r);
        a.rE[vb + a.G.ve(r)] = d.x;
        a.iE[vb + a.G.ve(r)] = d.y;
      }
      if (a.info && tid == 0) a.info[m] = nfail;
    } else if (a.mode == BIG_TREVC) {
      trevc_tile(T, B);
    } else {  // BIG_FUSED: A read only (unless the Schur form is asked for)
      reduce_tile(T, sh, false, &B, nullptr);
      const int nfail = lahqr_tile(T, sh, B, a.itmax, false);
      if (a.flags & VEIG_WRITE_SCHUR)
        for (int c = 0; c < n; ++c)
          for (int r = tid; r < n; r += kThreads) A.st(r, c, r <= c ? ld<double>(T.at(r, c)) : mk<double>(0, 0));
      for (int r = tid; r < n; r += kThreads) {
        const double2 d = *T.at(r, r);
        a.rE[vb + a.G.ve(r)] = d.x;
        a.iE[vb + a.G.ve(r)] = d.y;
      }
      __syncthreads();
      trevc_tile(T, B);
      if (a.info && tid == 0) a.info[m] = nfail;
    }
    __syncthreads();  // the tile is reused by the next matrix
  }
}

size_t big_lds_bytes(int n) { return (size_t)(n * (n + 1) + 2 * n) * sizeof(double2); }

}  // namespace

namespace veig {

// The kernel above for one batch (mode: what it computes, see BigArgs).
int launch_big(int mode, const Geom &G, double *rA, double *iA, double *rB, double *iB, double *rE, double *iE,
               int32_t *info, int itmax, uint32_t flags, hipStream_t s) {
  if (G.n > VEIG_MAX_DIM) return fail(VEIG_ERR_UNSUPPORTED, "order beyond VEIG_MAX_DIM");
  const size_t bytes = big_lds_bytes(G.n);
  VEIG_HIP(hipFuncSetAttribute((const void *)k_big, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  int dev = 0, cus = 256;
  VEIG_HIP(hipGetDevice(&dev));
  VEIG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int per_cu = (int)(160 * 1024 / (bytes + 1024)) > 0 ? (int)(160 * 1024 / (bytes + 1024)) : 1;
  const int64_t want = (int64_t)cus * per_cu;
  const unsigned grid = (unsigned)(G.nmat < want ? G.nmat : want);
  BigArgs a{rA, iA, rB, iB, rE, iE, info, G, itmax > 0 ? itmax : 30, mode, flags};
  hipLaunchKernelGGL(k_big, dim3(grid), dim3(kThreads), bytes, s, a);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

}  // namespace veig
