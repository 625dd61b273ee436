// Batches of matrices of different orders (1 <= n <= 32) in one enqueue,
// ready for capture into a hipGraph.
//
// Layout: matrix m is column-major with leading dimension dims[m], element
// (0,0) at a_off[m] of the rA/iA planes; V uses the same offsets in rV/iV;
// eigenvalue 0 of matrix m sits at e_off[m] of rE/iE (the reference has no
// uneven-size layout; this is its one-matrix-per-tile case, thread_cnt = 1,
// with explicit offsets).
//
// Launch structure (no host synchronisation, no allocation, no memset, so the
// whole sequence can be captured and replayed):
//   k_bin_init     clears the size histogram (a fill kernel, not a memset)
//   k_bin_count    histogram of the orders
//   k_bin_scan     orders grouped by kernel variant (the per-order choice of
//                  `variant_of_n`: lane-group width NB and lanes per matrix),
//                  ordered by n inside a variant: start offsets and the bin
//                  ranges
//   k_bin_scatter  permutation of matrix indices in that order
//   k_zgeev_var<NB,P> one persistent kernel per bin: a fixed grid walks the bin's
//                  range, read from device memory, wavefront by wavefront; the
//                  matrices of a wavefront then have (nearly) the same order.
//                  Bins (16, 16) and (32, 32) run two such kernels instead,
//                  k_reduce_var and k_qr_var (the banded path); the bins'
//                  orders <= 12 and <= 24 in slots of capacity 12 and 24.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "veig_band.hpp"

using namespace veig;

namespace {

constexpr int kMaxN = 32;
// kernel variants (bins): lane-group width NB and slots P, i.e. NB/P lanes
// per matrix -- the threads-per-matrix choice of the autotuner
constexpr int kBins = 5;
__host__ __device__ constexpr int var_nb(int v) { return v < 2 ? 8 : (v < 4 ? 16 : 32); }
__host__ __device__ constexpr int var_p(int v) { return (v == 1 || v == 3) ? 2 : 1; }
__host__ __device__ constexpr int default_variant(int n) { return n <= 8 ? 0 : (n <= 16 ? 2 : 4); }
// orders of a width-16 / width-32 bin that run in slots of capacity 12 / 24
__host__ __device__ constexpr int small_cut(int v) { return var_nb(v) == 16 ? 12 : (var_nb(v) == 32 ? 24 : 0); }

struct BinState {
  int32_t hist[kMaxN + 1];     // count per order
  int32_t cursor[kMaxN + 1];   // scatter cursor per order
  int64_t start[kBins];        // bin range in the permutation
  int64_t count[kBins];
  int64_t small[kBins];        // leading part of a bin with orders <= small_cut
  int32_t var_of_n[kMaxN + 1]; // kernel variant per order
};

struct VarArgs {
  int64_t nmat;
  const int32_t *dims;
  const int64_t *a_off, *e_off;
  const double *rA, *iA;
  double *rE, *iE, *rV, *iV;
  int32_t *info;
  int32_t *perm;
  BinState *bins;
  int itmax;
};

__global__ void k_bin_init(BinState *b, const int8_t *var_table) {
  const int i = threadIdx.x;
  if (i <= kMaxN) {
    b->hist[i] = 0;
    b->cursor[i] = 0;
    int v = var_table ? var_table[i] : -1;
    // unknown codes, and groups too narrow to hold the matrix: the default
    if (v < 0 || v >= kBins || var_nb(v) < i) v = default_variant(i);
    b->var_of_n[i] = v;
  }
  if (i < kBins) {
    b->start[i] = 0;
    b->count[i] = 0;
    b->small[i] = 0;
  }
}

__global__ void k_bin_count(VarArgs a) {
  __shared__ int32_t h[kMaxN + 1];
  if (threadIdx.x <= kMaxN) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.nmat;
       m += (int64_t)gridDim.x * blockDim.x) {
    const int n = a.dims[m];
    if (n >= 1 && n <= kMaxN) {
      atomicAdd(&h[n], 1);
    } else if (a.info) {
      a.info[m] = (n == 0) ? 0 : -1;  // n == 0: nothing to do; invalid order: -1
    }
  }
  __syncthreads();
  if (threadIdx.x <= kMaxN && h[threadIdx.x]) atomicAdd(&a.bins->hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_bin_scan(BinState *b) {
  if (threadIdx.x != 0) return;
  int64_t pos = 0;
  for (int bi = 0; bi < kBins; ++bi) {
    b->start[bi] = pos;
    for (int n = 1; n <= kMaxN; ++n) {
      if (b->var_of_n[n] != bi) continue;
      b->cursor[n] = (int32_t)pos;
      pos += b->hist[n];
      if (n <= small_cut(bi)) b->small[bi] = pos - b->start[bi];
    }
    b->count[bi] = pos - b->start[bi];
  }
}

__global__ void k_bin_scatter(VarArgs a) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.nmat;
       m += (int64_t)gridDim.x * blockDim.x) {
    const int n = a.dims[m];
    if (n >= 1 && n <= kMaxN) a.perm[atomicAdd(&a.bins->cursor[n], 1)] = (int32_t)m;
  }
}

template <int NB, int P>
__global__ __launch_bounds__(64) void k_zgeev_var(VarArgs a, int bin) {
  using R = double;
  using Gr = Group<R, NB, P>;
  __shared__ Tiles<R, NB, P> sm;
  constexpr int MPW = Gr::MPW;
  const int64_t start = a.bins->start[bin], count = a.bins->count[bin];
  const int64_t nwaves = (count + MPW - 1) / MPW;
  for (int64_t w = blockIdx.x; w < nwaves; w += gridDim.x) {
    // an opaque copy of the lane: addresses derived from it are recomputed
    // for every wavefront of matrices instead of being hoisted out of the loop
    int lane = threadIdx.x;
    asm volatile("" : "+v"(lane));
    const int64_t idx = w * MPW + lane / Gr::LG;
    const bool valid = idx < count;
    const int64_t m = valid ? a.perm[start + idx] : 0;
    const int n = valid ? a.dims[m] : 0;
    int nmax = 0;
#pragma unroll
    for (int q = 0; q < MPW; ++q) {
      const int nq = __builtin_amdgcn_readlane(n, q * Gr::LG);
      nmax = nq > nmax ? nq : nmax;
    }
    Gr g = make_group<R, NB, P>(sm, n, lane);
    Geom G;
    G.nmat = 0;
    G.T = 1;
    G.n = n;
    const int64_t base = valid ? a.a_off[m] : 0;
    const int64_t vbase = valid ? a.e_off[m] : 0;
    wave_sync();
    load_rows<R, NB, P>(g, G, base, a.rA, a.iA, NB);
    cx<R> z[P][NB];
    identity_rows<R, NB, P>(g, z);
    wave_sync();
    gehrd_q<R, NB, P>(g, z, nmax, true, false, nullptr);
    const int nfail = lahqr<R, NB, P>(g, z, a.itmax, false);
    cx<R> lam[P];
    solve_eigvecs<R, NB, P>(g, nmax, lam);
    backtransform<R, NB, P>(g, z, nmax);
    store_row_regs<R, NB, P>(g, G, base, a.rV, a.iV, z);
    store_eigenvalues<R, NB, P>(g, G, vbase, a.rE, a.iE, lam);
    if (a.info && valid && g.r == 0) a.info[m] = nfail;
    wave_sync();
  }
}

// Bins (16, 16) and (32, 32) on the two-kernel banded path (one index per
// lane): k_reduce_var<NB> writes each matrix's compact Hessenberg form into
// its V and E slots, then k_qr_var<NB> forms Q, iterates and solves on the
// banded tile -- the LDS of a wavefront drops from 18.6 to 11.8 KB (NB = 16)
// and from 33.8 to 19.5 KB (NB = 32), one more resident wavefront per SIMD.
// Both walk the bin's range like k_zgeev_var.
// the range of a bin, or of its part 0 (orders <= small_cut) or part 1 (the rest)
__device__ __forceinline__ void bin_part(const VarArgs &a, int bin, int part, int64_t &start, int64_t &count) {
  start = a.bins->start[bin];
  count = a.bins->count[bin];
  const int64_t small = a.bins->small[bin];
  if (part == 0) {
    count = small;
  } else if (part == 1) {
    start += small;
    count -= small;
  }
}

template <int LG, int MPW>
__device__ __forceinline__ int wave_max_order(int n) {
  int nmax = 0;
#pragma unroll
  for (int q = 0; q < MPW; ++q) {
    const int nq = __builtin_amdgcn_readlane(n, q * LG);
    nmax = nq > nmax ? nq : nmax;
  }
  return nmax;
}

template <int NB, int LW>
__global__ __launch_bounds__(64 * reduce_waves<NB>()) __attribute__((amdgpu_waves_per_eu(2))) void k_reduce_var(
    VarArgs a, int bin, int part) {
  using G = Group<double, NB, 1, LAY_DENSE, LW>;
  __shared__ ReduceTiles<NB, LW> sm;
  constexpr int MPW = G::MPW;
  int64_t start, count;
  bin_part(a, bin, part, start, count);
  const int64_t nwaves = (count + MPW - 1) / MPW;
  // the workgroup's wavefronts share each wavefront-sized set of matrices
  for (int64_t w = blockIdx.x; w < nwaves; w += gridDim.x) {
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int64_t idx = w * MPW + lane / G::LG;
    const bool valid = idx < count;
    const int64_t m = valid ? a.perm[start + idx] : 0;
    const int n = valid ? a.dims[m] : 0;
    const int nmax = wave_max_order<G::LG, MPW>(n);
    Geom Gm;
    Gm.nmat = 0;
    Gm.T = 1;
    Gm.n = n;
    const int64_t base = valid ? a.a_off[m] : 0, vbase = valid ? a.e_off[m] : 0;
    band_reduce_one<NB, LW>(sm, n, Gm, base, vbase, nmax, a.rA, a.iA, a.rV, a.iV, a.rE, a.iE);
  }
}

template <int NB, int LW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(qr_band_waves<NB>()))) void k_qr_var(VarArgs a,
                                                                                                       int bin,
                                                                                                       int part) {
  using G = Group<double, NB, 1, LAY_BAND, LW>;
  __shared__ BandTiles<NB, LW> sm;
  constexpr int MPW = G::MPW;
  int64_t start, count;
  bin_part(a, bin, part, start, count);
  const int64_t nwaves = (count + MPW - 1) / MPW;
  for (int64_t w = blockIdx.x; w < nwaves; w += gridDim.x) {
    int lane = threadIdx.x;
    asm volatile("" : "+v"(lane));
    const int64_t idx = w * MPW + lane / G::LG;
    const bool valid = idx < count;
    const int64_t m = valid ? a.perm[start + idx] : 0;
    const int n = valid ? a.dims[m] : 0;
    const int nmax = wave_max_order<G::LG, MPW>(n);
    const G g = band_group<NB, LW>(sm, n, lane);
    Geom Gm;
    Gm.nmat = 0;
    Gm.T = 1;
    Gm.n = n;
    const int64_t base = valid ? a.a_off[m] : 0, vbase = valid ? a.e_off[m] : 0;
    const int nfail =
        band_qr_one<NB, LW>(g, Gm, base, vbase, nmax, a.itmax, a.rV, a.iV, a.rE, a.iE, nullptr, nullptr);
    if (a.info && valid && g.r == 0) a.info[m] = nfail;
    wave_sync();
  }
}

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

size_t varsize_ws_bytes(int64_t nmat) {
  return align_up(sizeof(BinState)) + align_up((size_t)nmat * sizeof(int32_t));
}

int enqueue_varsize(int64_t nmat, const int32_t *dims, const int64_t *a_off, const int64_t *e_off,
                    const double *rA, const double *iA, double *rE, double *iE, double *rV, double *iV,
                    int32_t *info, int32_t itmax, const int8_t *var_table, const int32_t *grid_per_bin,
                    void *workspace, hipStream_t s) {
  char *p = static_cast<char *>(workspace);
  VarArgs a;
  a.nmat = nmat;
  a.dims = dims;
  a.a_off = a_off;
  a.e_off = e_off;
  a.rA = rA;
  a.iA = iA;
  a.rE = rE;
  a.iE = iE;
  a.rV = rV;
  a.iV = iV;
  a.info = info;
  a.bins = reinterpret_cast<BinState *>(p);
  a.perm = reinterpret_cast<int32_t *>(p + align_up(sizeof(BinState)));
  a.itmax = itmax > 0 ? itmax : 30;
  int dev = 0, cus = 256;
  VEIG_HIP(hipGetDevice(&dev));
  VEIG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t cgrid64 = (nmat + 255) / 256;
  const unsigned cgrid = (unsigned)(cgrid64 < 4 * (int64_t)cus ? (cgrid64 > 0 ? cgrid64 : 1) : 4 * cus);
  hipLaunchKernelGGL(k_bin_init, dim3(1), dim3(64), 0, s, a.bins, var_table);
  hipLaunchKernelGGL(k_bin_count, dim3(cgrid), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(64), 0, s, a.bins);
  hipLaunchKernelGGL(k_bin_scatter, dim3(cgrid), dim3(256), 0, s, a);
  // persistent grids: a few resident waves per CU and bin (the bins run back
  // to back; every wave of a bin loops until the bin is drained)
  const int defaults[kBins] = {8 * cus, 8 * cus, 12 * cus, 4 * cus, 8 * cus};
  for (int b = 0; b < kBins; ++b) {
    const int g = (grid_per_bin && grid_per_bin[b] > 0) ? grid_per_bin[b] : defaults[b];
    switch (b) {
      case 0: hipLaunchKernelGGL((k_zgeev_var<8, 1>), dim3(g), dim3(64), 0, s, a, 0); break;
      case 1: hipLaunchKernelGGL((k_zgeev_var<8, 2>), dim3(g), dim3(64), 0, s, a, 1); break;
      case 2:
        // orders <= 12 of the bin in slots of 12, the rest in slots of 16
        hipLaunchKernelGGL((k_reduce_var<12, 16>), dim3(g), dim3(64 * reduce_waves<12>()), 0, s, a, 2, 0);
        hipLaunchKernelGGL((k_qr_var<12, 16>), dim3(g), dim3(64), 0, s, a, 2, 0);
        hipLaunchKernelGGL((k_reduce_var<16, 0>), dim3(g), dim3(64 * reduce_waves<16>()), 0, s, a, 2, 1);
        hipLaunchKernelGGL((k_qr_var<16, 0>), dim3(g), dim3(64), 0, s, a, 2, 1);
        break;
      case 3: hipLaunchKernelGGL((k_zgeev_var<16, 2>), dim3(g), dim3(64), 0, s, a, 3); break;
      default:
        // orders <= 24 of the bin in slots of 24, the rest in slots of 32
        hipLaunchKernelGGL((k_reduce_var<24, 32>), dim3(g), dim3(64 * reduce_waves<24>()), 0, s, a, 4, 0);
        hipLaunchKernelGGL((k_qr_var<24, 32>), dim3(g), dim3(64), 0, s, a, 4, 0);
        hipLaunchKernelGGL((k_reduce_var<32, 0>), dim3(g), dim3(64 * reduce_waves<32>()), 0, s, a, 4, 1);
        hipLaunchKernelGGL((k_qr_var<32, 0>), dim3(g), dim3(64), 0, s, a, 4, 1);
        break;
    }
  }
  VEIG_HIP(hipGetLastError());
  return VEIG_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the binning state and the permutation live in a
// file-scope device array instead of a caller workspace, for callers that
// cannot allocate (graph capture, serving loops).  Its size caps the batch at
// kNoallocMaxMat matrices (16 MiB); every GPU the library runs on holds one
// copy until the process ends.  Host code never uses the array's name as an
// address: each device's copy is looked up once (hipGetSymbolAddress, with
// that device current) and that address is what the kernels get.
constexpr int64_t kNoallocMaxMat = int64_t(1) << 22;
constexpr size_t kNoallocBytes = align_up(sizeof(BinState)) + align_up(size_t(kNoallocMaxMat) * sizeof(int32_t));
__device__ __align__(256) char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_addr[kMaxDevices];

int scratch_for_current_device(void **out) {
  int dev = 0;
  VEIG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(VEIG_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_addr[dev]) {
    void *p = nullptr;
    VEIG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p || reinterpret_cast<uintptr_t>(p) % kAlign != 0)
      return fail(VEIG_ERR_HIP, "scratch array address missing or misaligned");
    scratch_addr[dev] = p;
  }
  *out = scratch_addr[dev];
  return VEIG_OK;
}

}  // namespace

struct veig_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

size_t veig_varsize_workspace_size(int64_t nmat) {
  if (nmat < 0) return 0;
  return varsize_ws_bytes(nmat);
}

int veig_zgeev_varsize(int64_t nmat, const int32_t *dims, const int64_t *a_off, const int64_t *e_off,
                       const double *rA, const double *iA, double *rE, double *iE, double *rV,
                       double *iV, int32_t *info, int32_t itmax, const int8_t *variant_of_n,
                       const int32_t *grid_per_bin, void *workspace, size_t workspace_bytes,
                       void *stream) {
  if (nmat < 0) return fail(VEIG_ERR_INVALID, "nmat < 0");
  if (nmat == 0) return VEIG_OK;
  if (nmat > INT32_MAX) return fail(VEIG_ERR_INVALID, "nmat exceeds 2^31-1");
  VEIG_REQUIRE(dims); VEIG_REQUIRE(a_off); VEIG_REQUIRE(e_off); VEIG_REQUIRE(rA); VEIG_REQUIRE(iA);
  VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rV); VEIG_REQUIRE(iV); VEIG_REQUIRE(workspace);
  if (workspace_bytes < varsize_ws_bytes(nmat))
    return fail(VEIG_ERR_INVALID, "workspace too small: need " + std::to_string(varsize_ws_bytes(nmat)) + " bytes");
  if (reinterpret_cast<uintptr_t>(workspace) % kAlign != 0)
    return fail(VEIG_ERR_INVALID, "workspace must be 256-byte aligned");
  return enqueue_varsize(nmat, dims, a_off, e_off, rA, iA, rE, iE, rV, iV, info, itmax, variant_of_n,
                         grid_per_bin, workspace, (hipStream_t)stream);
}

int veig_zgeev_varsize_noalloc(int64_t nmat, const int32_t *dims, const int64_t *a_off,
                               const int64_t *e_off, const double *rA, const double *iA, double *rE,
                               double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
                               const int8_t *variant_of_n, const int32_t *grid_per_bin, void *stream) {
  if (nmat < 0) return fail(VEIG_ERR_INVALID, "nmat < 0");
  if (nmat == 0) return VEIG_OK;
  if (nmat > kNoallocMaxMat)
    return fail(VEIG_ERR_INVALID, "nmat " + std::to_string(nmat) + " exceeds the allocation-free entry's " +
                                      std::to_string(kNoallocMaxMat) + " matrices");
  VEIG_REQUIRE(dims); VEIG_REQUIRE(a_off); VEIG_REQUIRE(e_off); VEIG_REQUIRE(rA); VEIG_REQUIRE(iA);
  VEIG_REQUIRE(rE); VEIG_REQUIRE(iE); VEIG_REQUIRE(rV); VEIG_REQUIRE(iV);
  void *ws = nullptr;
  VEIG_TRY(scratch_for_current_device(&ws));
  // k_bin_init clears the binning state (a kernel, not a memset); the
  // permutation is written in full before it is read
  return enqueue_varsize(nmat, dims, a_off, e_off, rA, iA, rE, iE, rV, iV, info, itmax, variant_of_n,
                         grid_per_bin, ws, (hipStream_t)stream);
}

int veig_varsize_graph_create(int64_t nmat, const int32_t *dims, const int64_t *a_off,
                              const int64_t *e_off, const double *rA, const double *iA, double *rE,
                              double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
                              const int8_t *variant_of_n, const int32_t *grid_per_bin, void *workspace,
                              size_t workspace_bytes, veig_graph **out) {
  VEIG_REQUIRE(out);
  *out = nullptr;
  hipStream_t cs;
  VEIG_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return hip_fail(e, "hipStreamBeginCapture");
  }
  const int st = veig_zgeev_varsize(nmat, dims, a_off, e_off, rA, iA, rE, iE, rV, iV, info, itmax,
                                    variant_of_n, grid_per_bin, workspace, workspace_bytes, cs);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(cs, &graph);
  (void)hipStreamDestroy(cs);
  if (st != VEIG_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  veig_graph *g = new veig_graph;
  g->graph = graph;
  e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    delete g;
    return hip_fail(e, "hipGraphInstantiate");
  }
  *out = g;
  return VEIG_OK;
}

int veig_varsize_autotune(int64_t nmat, const int32_t *dims, const int64_t *a_off, const int64_t *e_off,
                          const double *rA, const double *iA, double *rE, double *iE, double *rV,
                          double *iV, int32_t *info, int32_t itmax, int8_t *variant_of_n_out,
                          int32_t *grid_per_bin_out, void *workspace, size_t workspace_bytes,
                          void *stream) {
  VEIG_REQUIRE(variant_of_n_out);
  VEIG_REQUIRE(grid_per_bin_out);
  for (int n = 0; n <= kMaxN; ++n) variant_of_n_out[n] = (int8_t)default_variant(n);
  for (int b = 0; b < kBins; ++b) grid_per_bin_out[b] = 0;
  if (nmat <= 0) return nmat < 0 ? fail(VEIG_ERR_INVALID, "nmat < 0") : VEIG_OK;
  VEIG_REQUIRE(dims);
  if (workspace_bytes < varsize_ws_bytes(nmat))
    return fail(VEIG_ERR_INVALID, "workspace too small: need " + std::to_string(varsize_ws_bytes(nmat)) + " bytes");
  hipStream_t s = (hipStream_t)stream;
  // orders present (host copy of dims: the tuner is an offline step)
  std::vector<int32_t> hd((size_t)nmat);
  VEIG_HIP(hipMemcpyAsync(hd.data(), dims, (size_t)nmat * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  VEIG_HIP(hipStreamSynchronize(s));
  int64_t hist[kMaxN + 1] = {0};
  for (int32_t n : hd)
    if (n >= 1 && n <= kMaxN) ++hist[n];
  // device scratch: a dims copy restricted to one order, and a variant table
  int32_t *dsub = nullptr;
  int8_t *dtab = nullptr;
  VEIG_HIP(hipMalloc(&dsub, (size_t)nmat * sizeof(int32_t)));
  if (hipMalloc(&dtab, kMaxN + 1) != hipSuccess) {
    (void)hipFree(dsub);
    return fail(VEIG_ERR_NOMEM, "hipMalloc (tuner table)");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int st = VEIG_OK;
  auto time_run = [&](const int32_t *d, const int8_t *tab, const int32_t *grid, float &ms) -> int {
    // one untimed run, then the best of two timed runs
    ms = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      if (hipEventRecord(e0, s) != hipSuccess) return fail(VEIG_ERR_HIP, "hipEventRecord");
      const int r = enqueue_varsize(nmat, d, a_off, e_off, rA, iA, rE, iE, rV, iV, info, itmax, tab, grid,
                                    workspace, s);
      if (r != VEIG_OK) return r;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        return fail(VEIG_ERR_HIP, "hipEventSynchronize");
      float t = 0;
      (void)hipEventElapsedTime(&t, e0, e1);
      if (rep > 0 && t < ms) ms = t;
    }
    return VEIG_OK;
  };
  // 1. per order, the fastest variant (matrices of the other orders masked to n = 0)
  int8_t tab[kMaxN + 1];
  std::vector<int32_t> sub((size_t)nmat);
  for (int n = 1; n <= kMaxN && st == VEIG_OK; ++n) {
    if (!hist[n] || n > 16) continue;  // orders above 16 have a single variant
    for (size_t m = 0; m < (size_t)nmat; ++m) sub[m] = hd[m] == n ? n : 0;
    if (hipMemcpy(dsub, sub.data(), (size_t)nmat * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(VEIG_ERR_HIP, "hipMemcpy (tuner dims)");
      break;
    }
    float best = 1e30f;
    for (int v = 0; v < kBins && st == VEIG_OK; ++v) {
      if (var_nb(v) < n) continue;
      for (int k = 0; k <= kMaxN; ++k) tab[k] = (int8_t)default_variant(k);
      tab[n] = (int8_t)v;
      if (hipMemcpy(dtab, tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess) {
        st = fail(VEIG_ERR_HIP, "hipMemcpy (tuner table)");
        break;
      }
      float ms = 0;
      st = time_run(dsub, dtab, nullptr, ms);
      if (st == VEIG_OK && ms < best) {
        best = ms;
        variant_of_n_out[n] = (int8_t)v;
      }
    }
  }
  // 2. per bin, the persistent grid (bins run back to back: their times add)
  if (st == VEIG_OK && hipMemcpy(dtab, variant_of_n_out, kMaxN + 1, hipMemcpyHostToDevice) != hipSuccess)
    st = fail(VEIG_ERR_HIP, "hipMemcpy (tuner table)");
  int dev = 0, cus = 256;
  if (st == VEIG_OK && (hipGetDevice(&dev) != hipSuccess ||
                        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess))
    st = fail(VEIG_ERR_HIP, "hipDeviceGetAttribute");
  for (int b = 0; b < kBins && st == VEIG_OK; ++b) {
    bool used = false;
    for (int n = 1; n <= kMaxN; ++n) used |= hist[n] && variant_of_n_out[n] == b;
    if (!used) continue;
    float best = 1e30f;
    const int cand[4] = {2, 4, 8, 16};
    for (int c : cand) {
      int32_t grid[kBins] = {0};
      for (int k = 0; k < kBins; ++k) grid[k] = grid_per_bin_out[k];
      grid[b] = c * cus;
      float ms = 0;
      st = time_run(dims, dtab, grid, ms);
      if (st != VEIG_OK) break;
      if (ms < best) {
        best = ms;
        grid_per_bin_out[b] = c * cus;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dsub);
  (void)hipFree(dtab);
  return st;
}

int veig_graph_launch(veig_graph *g, void *stream) {
  VEIG_REQUIRE(g);
  VEIG_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return VEIG_OK;
}

int veig_graph_destroy(veig_graph *g) {
  if (!g) return VEIG_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return VEIG_OK;
}

}  // extern "C"
