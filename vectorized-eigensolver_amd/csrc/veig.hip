// Batched complex eigensolver for MI355X (gfx950): FP64 kernels and C ABI.
// See include/veig.h for the interface and veig_device.hpp for the
// lane-group execution model.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "veig_common.hpp"

namespace veig {
std::string &last_error() {
  static thread_local std::string s;
  return s;
}

}  // namespace veig

using namespace veig;

namespace {

// --------------------------------------------------------------------------
// stage kernels
// --------------------------------------------------------------------------
struct StageArgs {
  double *rA, *iA;      // matrix in/out (A, H, T)
  double *rB, *iB;      // second matrix (Q, Z, R)
  double *rE, *iE;      // vector (tau or eigenvalues)
  int32_t *info;
  Geom G;
  int itmax;
  int mode;
};

enum { ST_GEHRD = 0, ST_HESS = 1, ST_UNGHR = 2, ST_LAHQR = 3, ST_TREVC = 4 };

template <int NB>
__global__ __launch_bounds__(64) void k_stage(StageArgs a) {
  using R = double;
  __shared__ Tiles<R, NB> sm;
  constexpr int MPW = Tiles<R, NB>::MPW;
  const int64_t m = xcd_block() * MPW + threadIdx.x / NB;
  const bool valid = m < a.G.nmat;
  const int n = valid ? a.G.n : 0;
  Group<R, NB> g = make_group<R, NB>(sm, n);
  const int64_t base = valid ? a.G.mat_base(m) : 0;
  const int64_t vbase = valid ? a.G.vec_base(m) : 0;
  const int nmax = a.G.n;
  wave_sync();
  cx<R> z[1][NB];
  if (a.mode == ST_GEHRD || a.mode == ST_HESS) {
    load_rows<R, NB>(g, a.G, base, a.rA, a.iA, NB);
    identity_rows<R, NB>(g, z);
    wave_sync();
    cx<R> tau = mk<R>(0, 0);
    gehrd_q<R, NB>(g, z, nmax, a.mode == ST_HESS, true, &tau);
    store_rows<R, NB>(g, a.G, base, a.rA, a.iA);
    if (a.rE && g.r + 1 < g.n) {
      a.rE[vbase + a.G.ve(g.r)] = tau.re;
      a.iE[vbase + a.G.ve(g.r)] = tau.im;
    }
    if (a.mode == ST_HESS) store_row_regs<R, NB>(g, a.G, base, a.rB, a.iB, z);
  } else if (a.mode == ST_UNGHR) {
    // reflector i: v = e_{i+1} + A(i+2:n, i), tau(i); Q = H_0 H_1 ... H_{n-2}
    load_rows<R, NB>(g, a.G, base, a.rA, a.iA, NB);
    identity_rows<R, NB>(g, z);
    wave_sync();
    for (int i = 0; i + 1 < nmax; ++i) {
      const bool in = g.r < g.n;
      cx<R> vr = mk<R>(0, 0);
      if (g.r == i + 1) vr = mk<R>(1, 0);
      else if (g.r >= i + 2 && in) vr = ld<R>(g.at(g.r, i));
      cx<R> tau = mk<R>(0, 0);
      if (i + 1 < g.n) {
        const int64_t v = vbase + a.G.ve(i);
        tau = mk<R>(a.rE[v], a.iE[v]);
      }
      wave_sync();
      st<R>(g.vsh + g.r, vr);
      wave_sync();
      cx<R> w = mk<R>(0, 0);
#pragma unroll
      for (int c = 0; c < NB; ++c) w = add(w, mul(z[0][c], ld<R>(g.vsh + c)));
      const cx<R> tw = mul(tau, w);
#pragma unroll
      for (int c = 0; c < NB; ++c) z[0][c] = sub(z[0][c], mulc(tw, ld<R>(g.vsh + c)));
    }
    store_row_regs<R, NB>(g, a.G, base, a.rA, a.iA, z);
  } else if (a.mode == ST_LAHQR) {
    load_rows<R, NB>(g, a.G, base, a.rA, a.iA, 1);  // Hessenberg part only
    load_row_regs<R, NB>(g, a.G, base, a.rB, a.iB, z);
    wave_sync();
    const int nfail = lahqr<R, NB>(g, z, a.itmax, true);
    store_rows<R, NB>(g, a.G, base, a.rA, a.iA);
    store_row_regs<R, NB>(g, a.G, base, a.rB, a.iB, z);
    if (g.r < g.n) {
      const cx<R> d = ld<R>(g.at(g.r, g.r));
      a.rE[vbase + a.G.ve(g.r)] = d.re;
      a.iE[vbase + a.G.ve(g.r)] = d.im;
    }
    if (a.info && valid && g.r == 0) a.info[m] = nfail;
  } else {  // ST_TREVC
    load_rows<R, NB>(g, a.G, base, a.rA, a.iA, 0);  // upper triangle only
    load_row_regs<R, NB>(g, a.G, base, a.rB, a.iB, z);
    wave_sync();
    cx<R> lam[1];
    trevc_backtransform<R, NB>(g, z, lam, nmax);
    store_row_regs<R, NB>(g, a.G, base, a.rB, a.iB, z);
  }
}

// --------------------------------------------------------------------------
// one-lane-per-matrix helpers (test entries of the reference)
// --------------------------------------------------------------------------
__global__ void k_zgemv(int herm, Geom G, const double *rM, const double *iM, const double *rx,
                        const double *ix, double *ry, double *iy, double br, double bi) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= G.nmat) return;
  const int n = G.n;
  const int64_t mb = G.mat_base(m), vb = G.vec_base(m);
  for (int i = 0; i < n; ++i) {
    const int64_t yi = vb + G.ve(i);
    const cx<double> y = mul(mk(ry[yi], iy[yi]), mk(br, bi));
    cx<double> acc = y;
    for (int j = 0; j < n; ++j) {
      const int64_t xj = vb + G.ve(j);
      const cx<double> x = mk(rx[xj], ix[xj]);
      if (!herm) {
        const int64_t e = mb + G.el(i, j);
        acc = add(acc, mul(mk(rM[e], iM[e]), x));
      } else {
        const int64_t e = mb + G.el(j, i);
        acc = add(acc, cmul(mk(rM[e], iM[e]), x));
      }
    }
    ry[yi] = acc.re;
    iy[yi] = acc.im;
  }
}

__global__ void k_backsub(Geom G, const double *rT, const double *iT, double *rb, double *ib) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= G.nmat) return;
  const int n = G.n;
  const int64_t mb = G.mat_base(m), vb = G.vec_base(m);
  for (int j = n - 1; j >= 0; --j) {
    const int64_t d = mb + G.el(j, j), bj = vb + G.ve(j);
    const cx<double> x = cdiv(mk(rb[bj], ib[bj]), mk(rT[d], iT[d]));
    rb[bj] = x.re;
    ib[bj] = x.im;
    for (int i = j - 1; i >= 0; --i) {
      const int64_t e = mb + G.el(i, j), bi = vb + G.ve(i);
      const cx<double> v = sub(mk(rb[bi], ib[bi]), mul(mk(rT[e], iT[e]), x));
      rb[bi] = v.re;
      ib[bi] = v.im;
    }
  }
}

// AoS <-> tiled SoA.  Within a block of T matrices the AoS plane is a
// T x n^2 row-major array ([t][e]) and the SoA plane its transpose ([e][t]),
// so each conversion is a batch of transposes of rows x cols arrays (AoS ->
// SoA: T x n^2; SoA -> AoS: n^2 x T), every block at offset b T n^2 in both.
// One workgroup moves a 64 x 64 tile through LDS: reads along the input's
// rows, writes along the output's rows, both coalesced.
constexpr int kTr = 64;
__global__ __launch_bounds__(256) void k_transpose(int64_t rows, int64_t cols, const double *in, double *out) {
  __shared__ double tile[kTr][kTr + 1];
  const int64_t ntr = (rows + kTr - 1) / kTr, ntc = (cols + kTr - 1) / kTr;
  const int64_t id = blockIdx.x, per = ntr * ntc;
  const int64_t b = id / per, rem = id - b * per;
  const int64_t r0 = rem / ntc * kTr, c0 = rem % ntc * kTr;
  const int64_t off = b * rows * cols;
  const int tx = threadIdx.x % kTr, ty = threadIdx.x / kTr;
  for (int i = ty; i < kTr; i += 256 / kTr) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) tile[i][tx] = in[off + r * cols + c];
  }
  __syncthreads();
  for (int i = ty; i < kTr; i += 256 / kTr) {
    const int64_t c = c0 + i, r = r0 + tx;  // output row c, column r
    if (c < cols && r < rows) out[off + c * rows + r] = tile[tx][i];
  }
}

// replication (and the T = 1 conversions, plain copies): one thread per
// element of the output
__global__ void k_layout(int dir, Geom G, const double *in, double *out, int64_t total) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int64_t n2 = (int64_t)G.n * G.n;
  if (dir == 0) {  // AoS -> SoA; k indexes the SoA output
    const int64_t b = k / (G.T * n2), rem = k - b * G.T * n2;
    const int64_t e = rem / G.T, t = rem - e * G.T;
    out[k] = in[(b * G.T + t) * n2 + e];
  } else if (dir == 1) {  // SoA -> AoS; k indexes the AoS output
    const int64_t m = k / n2, e = k - m * n2;
    const int64_t b = m / G.T, t = m - b * G.T;
    out[k] = in[b * G.T * n2 + e * G.T + t];
  } else if (dir == 2) {  // replicate matrix
    const int64_t b = k / (G.T * n2), rem = k - b * G.T * n2;
    out[k] = in[rem / G.T];
  } else {  // replicate vector
    const int64_t b = k / (G.T * G.n), rem = k - b * G.T * G.n;
    out[k] = in[rem / G.T];
  }
}

int launch_stage(const StageArgs &a, hipStream_t s) {
  const int nb = nb_for(a.G.n);
  const int mpw = 64 / nb;
  const int64_t grid = (a.G.nmat + mpw - 1) / mpw;
  if (grid > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  switch (nb) {
    case 8: hipLaunchKernelGGL((k_stage<8>), dim3((unsigned)grid), dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_stage<16>), dim3((unsigned)grid), dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL((k_stage<32>), dim3((unsigned)grid), dim3(64), 0, s, a); break;
  }
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}



// RAII device buffer for the host mirrors
struct DevBuf {
  double *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int dev_alloc(DevBuf &b, size_t count) {
  if (count == 0) return VEIG_OK;
  hipError_t e = hipMalloc(&b.p, count * sizeof(double));
  if (e != hipSuccess) return fail(VEIG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return VEIG_OK;
}
int h2d(DevBuf &b, const double *h, size_t count) {
  if (count) VEIG_HIP(hipMemcpy(b.p, h, count * sizeof(double), hipMemcpyHostToDevice));
  return VEIG_OK;
}
int d2h(double *h, const DevBuf &b, size_t count) {
  if (count) VEIG_HIP(hipMemcpy(h, b.p, count * sizeof(double), hipMemcpyDeviceToHost));
  return VEIG_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char *veig_status_string(int status) {
  switch (status) {
    case VEIG_OK: return "ok";
    case VEIG_ERR_INVALID: return "invalid argument";
    case VEIG_ERR_UNSUPPORTED: return "unsupported order";
    case VEIG_ERR_HIP: return "HIP runtime error";
    case VEIG_ERR_NOMEM: return "out of device memory";
    default: return "unknown status";
  }
}

const char *veig_last_error(void) { return last_error().c_str(); }

const char *veig_version(void) { return "veig 0.3 gfx950 lane-group kernels (fused and stages n <= 96)"; }

int veig_zgeev(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
               double *rE, double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
               uint32_t flags, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rV); VEIG_REQUIRE(iV);
  FusedArgs a{rA, iA, rE, iE, rV, iV, info, make_geom(dim, block_cnt, thread_cnt),
              itmax > 0 ? itmax : 30, flags, nullptr, nullptr, nullptr};
  return launch_zgeev<double>(a, (hipStream_t)stream);
}

int veig_zgeev_fp32(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rA,
                    const double *iA, double *rE, double *iE, double *rV, double *iV,
                    int32_t *info, int32_t itmax, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM_FAST));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rV); VEIG_REQUIRE(iV);
  FusedArgs a{const_cast<double *>(rA), const_cast<double *>(iA), rE, iE, rV, iV, info,
              make_geom(dim, block_cnt, thread_cnt), itmax > 0 ? itmax : 30, 0u, nullptr, nullptr, nullptr};
  return launch_zgeev<float>(a, (hipStream_t)stream);
}

int veig_zgehrd(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
                double *rtau, double *itau, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rtau); VEIG_REQUIRE(itau);
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  if (dim > 32) return launch_big(BIG_GEHRD, G, rA, iA, nullptr, nullptr, rtau, itau, nullptr, 30, 0, (hipStream_t)stream);
  StageArgs a{rA, iA, nullptr, nullptr, rtau, itau, nullptr, G, 30, ST_GEHRD};
  return launch_stage(a, (hipStream_t)stream);
}

int veig_hessenberg(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
                    double *rQ, double *iQ, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rQ); VEIG_REQUIRE(iQ);
  // tau is produced and consumed on chip (Q is accumulated forward), so no
  // tau planes are needed: the kernels skip writing it
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  if (dim > 32) return launch_big(BIG_HESS, G, rA, iA, rQ, iQ, nullptr, nullptr, nullptr, 30, 0, (hipStream_t)stream);
  StageArgs a{rA, iA, rQ, iQ, nullptr, nullptr, nullptr, G, 30, ST_HESS};
  return launch_stage(a, (hipStream_t)stream);
}

int veig_zunghr(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
                const double *rtau, const double *itau, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rA); VEIG_REQUIRE(iA); VEIG_REQUIRE(rtau); VEIG_REQUIRE(itau);
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  double *rt = const_cast<double *>(rtau), *it = const_cast<double *>(itau);
  if (dim > 32) return launch_big(BIG_UNGHR, G, rA, iA, nullptr, nullptr, rt, it, nullptr, 30, 0, (hipStream_t)stream);
  StageArgs a{rA, iA, nullptr, nullptr, rt, it, nullptr, G, 30, ST_UNGHR};
  return launch_stage(a, (hipStream_t)stream);
}

int veig_zlahqr(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rE, double *iE,
                double *rH, double *iH, double *rZ, double *iZ, int32_t *info, int32_t itmax,
                void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rH); VEIG_REQUIRE(iH); VEIG_REQUIRE(rZ); VEIG_REQUIRE(iZ);
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  const int its = itmax > 0 ? itmax : 30;
  if (dim > 32) return launch_big(BIG_LAHQR, G, rH, iH, rZ, iZ, rE, iE, info, its, 0, (hipStream_t)stream);
  StageArgs a{rH, iH, rZ, iZ, rE, iE, info, G, its, ST_LAHQR};
  return launch_stage(a, (hipStream_t)stream);
}

int veig_ztrevc(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rT,
                const double *iT, double *rR, double *iR, void *stream) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rT); VEIG_REQUIRE(iT); VEIG_REQUIRE(rR); VEIG_REQUIRE(iR);
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  double *rt = const_cast<double *>(rT), *it = const_cast<double *>(iT);
  if (dim > 32) return launch_big(BIG_TREVC, G, rt, it, rR, iR, nullptr, nullptr, nullptr, 30, 0, (hipStream_t)stream);
  StageArgs a{rt, it, rR, iR, nullptr, nullptr, nullptr, G, 30, ST_TREVC};
  return launch_stage(a, (hipStream_t)stream);
}

int veig_zgemv(int hermitian, int64_t dim, int64_t block_cnt, int64_t thread_cnt,
               const double *rM, const double *iM, const double *rx, const double *ix,
               double *ry, double *iy, double rbeta, double ibeta, void *stream) {
  if (dim < 0 || block_cnt < 0 || thread_cnt < 1) return fail(VEIG_ERR_INVALID, "invalid batch geometry");
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rM); VEIG_REQUIRE(iM); VEIG_REQUIRE(rx); VEIG_REQUIRE(ix); VEIG_REQUIRE(ry); VEIG_REQUIRE(iy);
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  const int64_t grid = (G.nmat + 255) / 256;
  hipLaunchKernelGGL(k_zgemv, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, hermitian, G, rM, iM,
                     rx, ix, ry, iy, rbeta, ibeta);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

int veig_backsubstitution(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rT,
                          const double *iT, double *rb, double *ib, void *stream) {
  if (dim < 0 || block_cnt < 0 || thread_cnt < 1) return fail(VEIG_ERR_INVALID, "invalid batch geometry");
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  VEIG_REQUIRE(rT); VEIG_REQUIRE(iT); VEIG_REQUIRE(rb); VEIG_REQUIRE(ib);
  const Geom G = make_geom(dim, block_cnt, thread_cnt);
  const int64_t grid = (G.nmat + 255) / 256;
  hipLaunchKernelGGL(k_backsub, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, G, rT, iT, rb, ib);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

static int layout(int dir, int64_t dim, int64_t bc, int64_t tc, const double *in, double *out,
                  void *stream) {
  if (dim < 0 || bc < 0 || tc < 1) return fail(VEIG_ERR_INVALID, "invalid batch geometry");
  if (empty_batch(dim, bc, tc)) return VEIG_OK;
  VEIG_REQUIRE(in); VEIG_REQUIRE(out);
  const Geom G = make_geom(dim, bc, tc);
  if (bc > 0 && tc > INT64_MAX / bc / (dim * dim + 1)) return fail(VEIG_ERR_INVALID, "batch size overflows");
  if (dir <= 1 && tc > 1) {
    const int64_t n2 = dim * dim, rows = dir == 0 ? tc : n2, cols = dir == 0 ? n2 : tc;
    const int64_t tiles = bc * ((rows + kTr - 1) / kTr) * ((cols + kTr - 1) / kTr);
    if (tiles > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)tiles), dim3(256), 0, (hipStream_t)stream, rows, cols, in, out);
    VEIG_HIP(hipGetLastError());
    return VEIG_OK;
  }
  const int64_t total = (dir <= 2) ? G.nmat * dim * dim : G.nmat * dim;
  const int64_t grid = (total + 255) / 256;
  if (grid > INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  hipLaunchKernelGGL(k_layout, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, dir, G, in, out, total);
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

int veig_rearrange_matrices(int64_t dim, int64_t bc, int64_t tc, const double *in, double *out, void *s) {
  return layout(0, dim, bc, tc, in, out, s);
}
int veig_arrange_matrices_aos(int64_t dim, int64_t bc, int64_t tc, const double *in, double *out, void *s) {
  return layout(1, dim, bc, tc, in, out, s);
}
int veig_rep_matrix_soa(int64_t dim, int64_t bc, int64_t tc, const double *m, double *out, void *s) {
  return layout(2, dim, bc, tc, m, out, s);
}
int veig_rep_vector_soa(int64_t dim, int64_t bc, int64_t tc, const double *v, double *out, void *s) {
  return layout(3, dim, bc, tc, v, out, s);
}

// ---------------------------------------------------------------------------
// host-pointer mirrors
// ---------------------------------------------------------------------------
int veig_host_eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                          double *rE, double *iE, double *rV, double *iV) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt, VEIG_MAX_DIM));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  DevBuf dra, dia, dre, die, drv, div;
  VEIG_TRY(dev_alloc(dra, sm)); VEIG_TRY(dev_alloc(dia, sm));
  VEIG_TRY(dev_alloc(dre, sv)); VEIG_TRY(dev_alloc(die, sv));
  VEIG_TRY(dev_alloc(drv, sm)); VEIG_TRY(dev_alloc(div, sm));
  VEIG_TRY(h2d(dra, rA, sm)); VEIG_TRY(h2d(dia, iA, sm));
  VEIG_TRY(veig_zgeev(dim, block_cnt, thread_cnt, dra.p, dia.p, dre.p, die.p, drv.p, div.p, nullptr, 30,
                 VEIG_WRITE_SCHUR, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rE, dre, sv)); VEIG_TRY(d2h(iE, die, sv));
  VEIG_TRY(d2h(rV, drv, sm)); VEIG_TRY(d2h(iV, div, sm));
  VEIG_TRY(d2h(rA, dra, sm)); VEIG_TRY(d2h(iA, dia, sm));
  return VEIG_OK;
}

int veig_host_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                     double *itau) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  DevBuf a, b, t, u;
  VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm)); VEIG_TRY(dev_alloc(t, sv)); VEIG_TRY(dev_alloc(u, sv));
  VEIG_TRY(h2d(a, rA, sm)); VEIG_TRY(h2d(b, iA, sm)); VEIG_TRY(h2d(t, rtau, sv)); VEIG_TRY(h2d(u, itau, sv));
  VEIG_TRY(veig_zgehrd(dim, block_cnt, thread_cnt, a.p, b.p, t.p, u.p, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rA, a, sm)); VEIG_TRY(d2h(iA, b, sm)); VEIG_TRY(d2h(rtau, t, sv)); VEIG_TRY(d2h(itau, u, sv));
  return VEIG_OK;
}

int veig_host_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                     double *itau) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  DevBuf a, b, t, u;
  VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm)); VEIG_TRY(dev_alloc(t, sv)); VEIG_TRY(dev_alloc(u, sv));
  VEIG_TRY(h2d(a, rA, sm)); VEIG_TRY(h2d(b, iA, sm)); VEIG_TRY(h2d(t, rtau, sv)); VEIG_TRY(h2d(u, itau, sv));
  VEIG_TRY(veig_zunghr(dim, block_cnt, thread_cnt, a.p, b.p, t.p, u.p, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rA, a, sm)); VEIG_TRY(d2h(iA, b, sm));
  return VEIG_OK;
}

int veig_host_hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                         double *rQ, double *iQ) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sm = (size_t)dim * dim * block_cnt * thread_cnt;
  DevBuf a, b, q, p;
  VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm)); VEIG_TRY(dev_alloc(q, sm)); VEIG_TRY(dev_alloc(p, sm));
  VEIG_TRY(h2d(a, rA, sm)); VEIG_TRY(h2d(b, iA, sm));
  VEIG_TRY(veig_hessenberg(dim, block_cnt, thread_cnt, a.p, b.p, q.p, p.p, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rA, a, sm)); VEIG_TRY(d2h(iA, b, sm)); VEIG_TRY(d2h(rQ, q, sm)); VEIG_TRY(d2h(iQ, p, sm));
  return VEIG_OK;
}

int veig_host_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE,
                         double *rM, double *iM, double *rS, double *iS) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  // S := 0 with the diagonal written through swapped planes (utils.cpp:167-169):
  // the imaginary plane gets 1.0 and the real plane 0.0, so S = i*I.
  std::memset(rS, 0, sm * sizeof(double));
  std::memset(iS, 0, sm * sizeof(double));
  for (int64_t b = 0; b < block_cnt; ++b)
    for (int64_t t = 0; t < thread_cnt; ++t)
      for (int64_t j = 0; j < dim; ++j) iS[b * thread_cnt * dim * dim + (j + j * dim) * thread_cnt + t] = 1.0;
  DevBuf e, f, a, b, s, u;
  VEIG_TRY(dev_alloc(e, sv)); VEIG_TRY(dev_alloc(f, sv)); VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm));
  VEIG_TRY(dev_alloc(s, sm)); VEIG_TRY(dev_alloc(u, sm));
  VEIG_TRY(h2d(a, rM, sm)); VEIG_TRY(h2d(b, iM, sm)); VEIG_TRY(h2d(s, rS, sm)); VEIG_TRY(h2d(u, iS, sm));
  VEIG_TRY(veig_zlahqr(dim, block_cnt, thread_cnt, e.p, f.p, a.p, b.p, s.p, u.p, nullptr, 30, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rE, e, sv)); VEIG_TRY(d2h(iE, f, sv)); VEIG_TRY(d2h(rM, a, sm)); VEIG_TRY(d2h(iM, b, sm));
  VEIG_TRY(d2h(rS, s, sm)); VEIG_TRY(d2h(iS, u, sm));
  return VEIG_OK;
}

int veig_host_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                     double *iR) {
  VEIG_TRY(check_geom(dim, block_cnt, thread_cnt));
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sm = (size_t)dim * dim * block_cnt * thread_cnt;
  DevBuf a, b, c, d;
  VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm)); VEIG_TRY(dev_alloc(c, sm)); VEIG_TRY(dev_alloc(d, sm));
  VEIG_TRY(h2d(a, rT, sm)); VEIG_TRY(h2d(b, iT, sm)); VEIG_TRY(h2d(c, rR, sm)); VEIG_TRY(h2d(d, iR, sm));
  VEIG_TRY(veig_ztrevc(dim, block_cnt, thread_cnt, a.p, b.p, c.p, d.p, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rR, c, sm)); VEIG_TRY(d2h(iR, d, sm));
  return VEIG_OK;
}

int veig_host_zgemv(int hermitian, int dim, int block_cnt, int thread_cnt, double *rM, double *iM,
                    double *rx, double *ix, double *ry, double *iy, double rbeta, double ibeta) {
  if (dim < 0 || block_cnt < 0 || thread_cnt < 1) return fail(VEIG_ERR_INVALID, "invalid batch geometry");
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  DevBuf a, b, x, y, u, v;
  VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm)); VEIG_TRY(dev_alloc(x, sv)); VEIG_TRY(dev_alloc(y, sv));
  VEIG_TRY(dev_alloc(u, sv)); VEIG_TRY(dev_alloc(v, sv));
  VEIG_TRY(h2d(a, rM, sm)); VEIG_TRY(h2d(b, iM, sm)); VEIG_TRY(h2d(x, rx, sv)); VEIG_TRY(h2d(y, ix, sv));
  VEIG_TRY(h2d(u, ry, sv)); VEIG_TRY(h2d(v, iy, sv));
  VEIG_TRY(veig_zgemv(hermitian, dim, block_cnt, thread_cnt, a.p, b.p, x.p, y.p, u.p, v.p, rbeta, ibeta, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(ry, u, sv)); VEIG_TRY(d2h(iy, v, sv));
  return VEIG_OK;
}

int veig_host_backsubstitution(int dim, int block_cnt, int thread_cnt, double *rT, double *iT,
                               double *rb, double *ib) {
  if (dim < 0 || block_cnt < 0 || thread_cnt < 1) return fail(VEIG_ERR_INVALID, "invalid batch geometry");
  if (empty_batch(dim, block_cnt, thread_cnt)) return VEIG_OK;
  const size_t sv = (size_t)dim * block_cnt * thread_cnt, sm = sv * dim;
  DevBuf a, b, x, y;
  VEIG_TRY(dev_alloc(a, sm)); VEIG_TRY(dev_alloc(b, sm)); VEIG_TRY(dev_alloc(x, sv)); VEIG_TRY(dev_alloc(y, sv));
  VEIG_TRY(h2d(a, rT, sm)); VEIG_TRY(h2d(b, iT, sm)); VEIG_TRY(h2d(x, rb, sv)); VEIG_TRY(h2d(y, ib, sv));
  VEIG_TRY(veig_backsubstitution(dim, block_cnt, thread_cnt, a.p, b.p, x.p, y.p, nullptr));
  VEIG_HIP(hipDeviceSynchronize());
  VEIG_TRY(d2h(rb, x, sv)); VEIG_TRY(d2h(ib, y, sv));
  return VEIG_OK;
}

}  // extern "C"
