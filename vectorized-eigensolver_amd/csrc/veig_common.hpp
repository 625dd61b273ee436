// Shared host/device plumbing of the gfx950 eigensolver translation units:
// batch geometry of the tiled split-plane layout, the per-wavefront LDS
// tiles, row loads/stores, error reporting and the fused kernel template.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "veig_device.hpp"
#include "../../include/veig.h"

namespace veig {

// last error message of the calling thread (veig_last_error)
std::string &last_error();



inline int fail(int status, const std::string &msg) {
  last_error() = msg;
  return status;
}

inline int hip_fail(hipError_t e, const char *what) {
  return fail(VEIG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define VEIG_HIP(call)                                  \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #call);   \
  } while (0)

// Batch geometry shared by all kernels.
struct Geom {
  int64_t nmat;  // block_cnt * thread_cnt
  int64_t T;     // thread_cnt (tile)
  int n;         // order
  __device__ __forceinline__ int64_t mat_base(int64_t m) const {
    const int64_t b = m / T, t = m - b * T;
    return b * T * (int64_t)n * n + t;
  }
  __device__ __forceinline__ int64_t vec_base(int64_t m, int mult = 1) const {
    const int64_t b = m / T, t = m - b * T;
    return (int64_t)mult * b * T * n + t;
  }
  __device__ __forceinline__ int64_t el(int r, int c) const { return ((int64_t)r + (int64_t)c * n) * T; }
  __device__ __forceinline__ int64_t ve(int r) const { return (int64_t)r * T; }
};

inline int check_geom(int64_t dim, int64_t block_cnt, int64_t thread_cnt, int64_t max_dim = VEIG_MAX_DIM_STAGES) {
  if (dim < 0 || block_cnt < 0 || thread_cnt < 1)
    return fail(VEIG_ERR_INVALID, "invalid batch geometry (dim < 0, block_cnt < 0 or thread_cnt < 1)");
  if (dim > max_dim)
    return fail(VEIG_ERR_UNSUPPORTED, "order " + std::to_string(dim) + " exceeds " +
                                          std::to_string(max_dim) + " for this entry point");
  if (block_cnt > 0 && thread_cnt > INT64_MAX / block_cnt) return fail(VEIG_ERR_INVALID, "batch size overflows");
  return VEIG_OK;
}

inline int nb_for(int64_t n) { return n <= 8 ? 8 : (n <= 16 ? 16 : (n <= 32 ? 32 : 64)); }


// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Adjacent matrices share cache lines in the tiled layout, so remap the
// launch order so that each XCD walks one contiguous range of matrices.
// (xcd_remap: the same over the first g blocks of a larger grid)
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t g) {
  constexpr int kXcd = 8;
  const int64_t g8 = g / kXcd * kXcd;
  if (b >= g8) return b;
  return (b % kXcd) * (g8 / kXcd) + b / kXcd;
}
__device__ __forceinline__ int64_t xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }

// --------------------------------------------------------------------------
// common kernel prologue: lane group setup and LDS tile
// --------------------------------------------------------------------------
template <typename R, int NB, int P = 1, int LW = 0>
struct Tiles {
  using L2 = typename lds2<R>::type;
  static constexpr int MPW = Group<R, NB, P, LAY_DENSE, LW>::MPW;
  static constexpr int LD = NB + 1;
  // one element of padding between the groups' tiles (and vectors) starts
  // each group on a different LDS bank: the groups' broadcast reads of their
  // own matrices (reflector inputs, shifts, T(i,j) in the solves) would
  // otherwise all hit the same bank with different addresses
  static constexpr int kTile = NB * LD + 1;
  static constexpr int kVec = NB + 1;
  L2 H[MPW * kTile];
  L2 v[MPW * kVec];
};

template <typename R, int NB, int P = 1>
__device__ __forceinline__ Group<R, NB, P> make_group(Tiles<R, NB, P> &sm, int n_of_matrix,
                                                      int lane = threadIdx.x) {
  using G = Group<R, NB, P>;
  G g;
  g.mi = lane / G::LG;
  g.r = lane % G::LG;
  g.n = n_of_matrix;
  g.H = sm.H + g.mi * Tiles<R, NB, P>::kTile;
  g.vsh = sm.v + g.mi * Tiles<R, NB, P>::kVec;
  // zero the whole tile: entries outside the matrix (and below the
  // subdiagonal when only a Hessenberg part is loaded) read as exact zeros
#pragma unroll
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int c = 0; c < NB; ++c) st<R>(g.at(g.idx(p), c), mk<R>(0, 0));
    st<R>(g.vsh + g.idx(p), mk<R>(0, 0));
  }
  return g;
}

// each lane loads its rows of a full matrix into the tile (entries with
// row <= col + lower_band)
template <typename R, int NB, int P>
__device__ __forceinline__ void load_rows(const Group<R, NB, P> &g, const Geom &G, int64_t base,
                                          const double *re, const double *im, int lower_band) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int row = g.idx(p);
    if (row < g.n) {
      for (int c = 0; c < g.n; ++c) {
        if (row <= c + lower_band) {
          const int64_t idx = base + G.el(row, c);
          st<R>(g.at(row, c), mk<R>((R)re[idx], (R)im[idx]));
        }
      }
    }
  }
}

// each lane loads its rows of a matrix into registers
template <typename R, int NB, int P>
__device__ __forceinline__ void load_row_regs(const Group<R, NB, P> &g, const Geom &G, int64_t base,
                                              const double *re, const double *im, cx<R> (&z)[P][NB]) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int row = g.idx(p);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      z[p][c] = mk<R>(0, 0);
      if (row < g.n && c < g.n) {
        const int64_t idx = base + G.el(row, c);
        z[p][c] = mk<R>((R)re[idx], (R)im[idx]);
      }
    }
  }
}

template <typename R, int NB, int P>
__device__ __forceinline__ void store_row_regs(const Group<R, NB, P> &g, const Geom &G, int64_t base,
                                               double *re, double *im, const cx<R> (&z)[P][NB]) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int row = g.idx(p);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (row < g.n && c < g.n) {
        const int64_t idx = base + G.el(row, c);
        re[idx] = (double)z[p][c].re;
        im[idx] = (double)z[p][c].im;
      }
    }
  }
}

// each lane stores its rows of the tile (full matrix)
template <typename R, int NB, int P>
__device__ __forceinline__ void store_rows(const Group<R, NB, P> &g, const Geom &G, int64_t base,
                                           double *re, double *im) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int row = g.idx(p);
    if (row < g.n) {
      for (int c = 0; c < g.n; ++c) {
        const int64_t idx = base + G.el(row, c);
        const cx<R> v = ld<R>(g.at(row, c));
        re[idx] = (double)v.re;
        im[idx] = (double)v.im;
      }
    }
  }
}

// each lane's eigenvalues lam[p] = T(idx(p), idx(p)) to the E planes
template <typename R, int NB, int P>
__device__ __forceinline__ void store_eigenvalues(const Group<R, NB, P> &g, const Geom &G, int64_t vbase,
                                                  double *re, double *im, const cx<R> (&lam)[P]) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int k = g.idx(p);
    if (k < g.n) {
      re[vbase + G.ve(k)] = (double)lam[p].re;
      im[vbase + G.ve(k)] = (double)lam[p].im;
    }
  }
}

template <typename R, int NB, int P>
__device__ __forceinline__ void identity_rows(const Group<R, NB, P> &g, cx<R> (&z)[P][NB]) {
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < NB; ++c) z[p][c] = mk<R>(c == g.idx(p) ? R(1) : R(0), 0);
}

// --------------------------------------------------------------------------
// fused eigendecomposition: A -> (E, V)
// --------------------------------------------------------------------------
struct FusedArgs {
  double *rA, *iA;
  double *rE, *iE, *rV, *iV;
  int32_t *info;
  Geom G;
  int itmax;
  uint32_t flags;
  const int32_t *mask;  // if set, only matrices with mask[m] != 0 are solved
  float *rW, *iW;       // OUT_VW: W = V^{-1}; OUT_ZXY: the Schur vectors Z (complex64 planes)
  float *rX = nullptr, *iX = nullptr;  // OUT_ZXY: X strictly above, X^{-1} strictly below the diagonal
  // if set, solve only the matrices list[0 .. *list_count) (device memory),
  // with a persistent grid walking the list
  const int32_t *list = nullptr;
  const int32_t *list_count = nullptr;
};

// What the fused kernel writes: the eigenvectors V; V and an approximate
// Eigenvector solves of the fused kernels (X above the diagonal of the tile,
// lam[p] = T(idx(p), idx(p))).  At NB = 32 with one index per lane they run
// in LDS: the register form would hold NB more complex values per lane next
// to the Schur vectors.
template <typename R, int NB, int P>
__device__ __forceinline__ void solve_eigvecs(const Group<R, NB, P> &g, int nmax, cx<R> (&lam)[P]) {
  if constexpr (P == 1 && NB == 32) {
    lam[0] = trevc_solve_lds<R, NB>(g, nmax);  // leaves y = -x
    if (g.r < g.n)
      for (int i = 0; i < g.r; ++i) {
        const cx<R> y = ld<R>(g.at(i, g.r));
        st<R>(g.at(i, g.r), mk(-y.re, -y.im));
      }
    wave_sync();
  } else {
    trevc_solve<R, NB, P>(g, nmax, lam);
  }
}

// inverse W = V^{-1} (complex64); or, for refinement on the matrix cores, the
// Schur vectors Z and the eigenvector factor X of T (V = Z X) with
// Y = X^{-1} (complex64), instead of V.
enum { OUT_V = 0, OUT_VW = 1, OUT_ZXY = 2 };

// one wavefront of the fused kernel: MPW matrices, from the batch (wave
// index `wave`) or from a.list (list entries wave*MPW ..)
// (lane: threadIdx.x; the list kernel passes it through an opaque copy so
// that the LDS addresses derived from it are recomputed for every matrix
// instead of being hoisted out of its loop and held in registers)
template <typename R, int NB, int P, int OUT>
__device__ __forceinline__ void zgeev_wave(const FusedArgs &a, Tiles<R, NB, P> &sm, int64_t wave,
                                           int64_t list_count, int lane) {
  using G = Group<R, NB, P>;
  constexpr int MPW = G::MPW;
  const int64_t slot = wave * MPW + lane / G::LG;
  int64_t m;
  bool valid;
  if (a.list) {
    valid = slot < list_count;
    m = valid ? a.list[slot] : 0;
  } else {
    m = slot;
    valid = m < a.G.nmat;
    if (valid && a.mask) valid = a.mask[m] != 0;
  }
  if (!__any(valid)) return;  // wave-uniform early exit (masked launches)
  const int n = valid ? a.G.n : 0;
  G g = make_group<R, NB, P>(sm, n, lane);
  const int64_t base = valid ? a.G.mat_base(m) : 0;
  const int64_t vbase = valid ? a.G.vec_base(m) : 0;
  wave_sync();
  load_rows<R, NB, P>(g, a.G, base, a.rA, a.iA, NB);
  cx<R> z[P][NB];
  identity_rows<R, NB, P>(g, z);
  wave_sync();
  const int nmax = a.G.n;  // every matrix of a launch has the same order
  gehrd_q<R, NB, P>(g, z, nmax, true, false, nullptr);
  const int nfail = lahqr<R, NB, P>(g, z, a.itmax, false);
  if (a.flags & VEIG_WRITE_SCHUR) store_rows<R, NB, P>(g, a.G, base, a.rA, a.iA);
  if constexpr (NB == 64) {
    // one matrix per wavefront: eigenvector solves in LDS, V streamed out
    static_assert(OUT == OUT_V && P == 1, "one matrix per wavefront: FP64, one index per lane");
    const cx<R> lam = trevc_solve_lds<R, NB>(g, nmax);
    backtransform_lds<R, NB>(g, z, nmax, [&](int k, cx<R> v) {
      if (g.r < g.n) {
        const int64_t idx = base + a.G.el(g.r, k);
        a.rV[idx] = (double)v.re;
        a.iV[idx] = (double)v.im;
      }
    });
    if (g.r < g.n) {
      a.rE[vbase + a.G.ve(g.r)] = (double)lam.re;
      a.iE[vbase + a.G.ve(g.r)] = (double)lam.im;
    }
    if (a.info && valid && g.r == 0) a.info[m] = nfail;
    return;
  }
  cx<R> lam[P];
  solve_eigvecs<R, NB, P>(g, nmax, lam);
  if constexpr (OUT == OUT_ZXY) {
    // Z first, so that its registers are free for the triangular inverse
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int row = g.idx(p);
      if (row < g.n) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          if (c < g.n) {
            const int64_t idx = base + a.G.el(row, c);
            a.rW[idx] = (float)z[p][c].re;
            a.iW[idx] = (float)z[p][c].im;
          }
        }
      }
    }
    inverse_unit_upper<R, NB, P>(g, nmax);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int row = g.idx(p);
      if (row < g.n) {
        for (int c = 0; c < g.n; ++c) {
          const int64_t idx = base + a.G.el(row, c);
          const cx<R> t = ld<R>(g.at(row, c));
          a.rX[idx] = (float)t.re;
          a.iX[idx] = (float)t.im;
        }
      }
    }
    store_eigenvalues<R, NB, P>(g, a.G, vbase, a.rE, a.iE, lam);
    if (a.info && valid && g.r == 0) a.info[m] = nfail;
    return;
  }
  if constexpr (OUT == OUT_VW) {
    // column idx(p) of W
    inverse_vectors<R, NB, P>(g, z, nmax, [&](int p, int k, cx<R> w) {
      const int col = g.idx(p);
      if (col < g.n && k < g.n) {
        const int64_t idx = base + a.G.el(k, col);
        a.rW[idx] = (float)w.re;
        a.iW[idx] = (float)w.im;
      }
    });
  }
  backtransform<R, NB, P>(g, z, nmax);
  store_row_regs<R, NB, P>(g, a.G, base, a.rV, a.iV, z);
  store_eigenvalues<R, NB, P>(g, a.G, vbase, a.rE, a.iE, lam);
  if (a.info && valid && g.r == 0) a.info[m] = nfail;
}

// the batch, one wavefront per MPW matrices (kept free of the list loop:
// a second inlined copy of the solver would double the code footprint)
// occupancy the register allocator is asked to keep (1: no constraint); the
// complex64 kernel that feeds the matrix-core refinement fits 4 wavefronts
template <typename R, int NB, int P, int OUT>
constexpr int zgeev_waves_per_eu() {
  return (std::is_same<R, float>::value && OUT == OUT_ZXY && NB == 16 && P == 1) ? 4 : 1;
}

template <typename R, int NB, int P, int OUT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(zgeev_waves_per_eu<R, NB, P, OUT>())))
void k_zgeev(FusedArgs a) {
  __shared__ Tiles<R, NB, P> sm;
  zgeev_wave<R, NB, P, OUT>(a, sm, xcd_block(), 0, threadIdx.x);
}

// the matrices of a.list, on a persistent grid
template <typename R, int NB, int P, int OUT>
__global__ __launch_bounds__(64) void k_zgeev_list(FusedArgs a) {
  __shared__ Tiles<R, NB, P> sm;
  constexpr int MPW = Group<R, NB, P>::MPW;
  const int64_t cnt = *a.list_count;
  const int64_t nw = (cnt + MPW - 1) / MPW;
  for (int64_t w = blockIdx.x; w < nw; w += gridDim.x) {
    int lane = threadIdx.x;
    asm volatile("" : "+v"(lane));
    zgeev_wave<R, NB, P, OUT>(a, sm, w, cnt, lane);
    wave_sync();
  }
}

// persistent grid of list launches (the list length is only known on the device)
constexpr int64_t kListGrid = 4096;

template <typename R, int NB, int P, int OUT>
inline void launch_one(const FusedArgs &a, hipStream_t s) {
  constexpr int mpw = Group<R, NB, P>::MPW;
  const int64_t grid = (a.G.nmat + mpw - 1) / mpw;
  if constexpr (std::is_same<R, double>::value && OUT == OUT_V) {
    if (a.list) {
      const int64_t g = grid < kListGrid ? grid : kListGrid;
      hipLaunchKernelGGL((k_zgeev_list<R, NB, P, OUT>), dim3((unsigned)g), dim3(64), 0, s, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_zgeev<R, NB, P, OUT>), dim3((unsigned)grid), dim3(64), 0, s, a);
}

// orders 33..64, one matrix per two-wavefront workgroup (instantiated for
// <double, OUT_V>, <float, OUT_V> and <float, OUT_VW>)
template <typename R, int OUT>
int launch_zgeev64(const FusedArgs &a, hipStream_t s);

// Orders beyond the lane-group kernels (veig_big.hip): one workgroup per
// matrix, the matrix in LDS, the unitary factor in global memory.  Modes:
// the fused eigendecomposition (A read only; Q, Z and then V in B; E), and
// the stage entries in place: zgehrd (A -> compact form, tau in E),
// hessenberg (also Q in B), zunghr (reflectors in A, tau in E -> Q in A),
// zlahqr (H in A -> T, Z in B updated, E), ztrevc (T in A, B <- B X).
enum { BIG_FUSED = 0, BIG_GEHRD = 1, BIG_HESS = 2, BIG_UNGHR = 3, BIG_LAHQR = 4, BIG_TREVC = 5 };
int launch_big(int mode, const Geom &G, double *rA, double *iA, double *rB, double *iB, double *rE, double *iE,
               int32_t *info, int itmax, uint32_t flags, hipStream_t s);

// FP64 batches and lists of orders up to NB on lane groups of LW lanes
// (0: NB): Hessenberg reduction, then the QR iteration and eigenvectors on
// a banded tile (two kernels)
template <int NB, int LW = 0>
int launch_banded(const FusedArgs &a, hipStream_t s);

// One index per lane (P = 1) everywhere: two or four indices per lane put
// more matrices in a wavefront but measured slower at every order (at n = 16
// FP64 the slot registers drop the occupancy to one wavefront per SIMD).
template <typename R, int OUT = OUT_V>
int launch_zgeev(const FusedArgs &a, hipStream_t s) {
  constexpr bool f32 = std::is_same<R, float>::value;
  if (a.G.nmat > (int64_t)INT32_MAX) return fail(VEIG_ERR_INVALID, "batch too large for one launch");
  if (a.G.n > 64) {
    if constexpr (!f32 && OUT == OUT_V) {
      if (!a.mask && !a.list)
        return launch_big(BIG_FUSED, a.G, a.rA, a.iA, a.rV, a.iV, a.rE, a.iE, a.info, a.itmax, a.flags, s);
    }
    return fail(VEIG_ERR_UNSUPPORTED, "orders above 64: FP64 whole batches only");
  }
  const int nb = nb_for(a.G.n);
  switch (nb) {
    case 8:
      launch_one<R, 8, 1, OUT>(a, s);
      break;
    case 16:
      if constexpr (!f32 && OUT == OUT_V) {
        if (!a.mask) return a.G.n <= 12 ? launch_banded<12, 16>(a, s) : launch_banded<16>(a, s);
      }
      launch_one<R, 16, 1, OUT>(a, s);
      break;
    case 32:
      if constexpr (!f32 && OUT == OUT_V) {
        // slots of capacity 24 or 28 where they fit: shorter Schur-vector rows, fewer spills
        if (!a.mask)
          return a.G.n <= 24 ? launch_banded<24, 32>(a, s)
                             : (a.G.n <= 28 ? launch_banded<28, 32>(a, s) : launch_banded<32>(a, s));
      }
      launch_one<R, 32, 1, OUT>(a, s);
      break;
    default:
      if constexpr (OUT == OUT_V || (f32 && OUT == OUT_VW)) {
        return launch_zgeev64<R, OUT>(a, s);
      } else {
        return fail(VEIG_ERR_UNSUPPORTED, "this output form handles orders up to 32");
      }
  }
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

#define VEIG_TRY(x)                   \
  do {                                \
    int st_ = (x);                    \
    if (st_ != VEIG_OK) return st_;   \
  } while (0)

#define VEIG_REQUIRE(p)                                                                     \
  do {                                                                                      \
    if (!(p)) return fail(VEIG_ERR_INVALID, std::string("NULL pointer argument: ") + #p);   \
  } while (0)

inline Geom make_geom(int64_t dim, int64_t bc, int64_t tc) {
  Geom G;
  G.nmat = bc * tc;
  G.T = tc;
  G.n = (int)dim;
  return G;
}

inline bool empty_batch(int64_t dim, int64_t bc, int64_t tc) { return dim == 0 || bc == 0 || tc == 0; }

}  // namespace veig
