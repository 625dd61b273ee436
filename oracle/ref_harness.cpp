// TEST INFRASTRUCTURE ONLY -- never linked into the product library.
//
// C-ABI harness around the reference's own compiled code (utils.cpp and the
// headers it pulls in, compiled from where they lie; see oracle/Makefile).
// Nothing here restates the algorithm: every entry point calls the
// reference's functions so that tests can compare the oracle restatement
// (oracle/eig_oracle.c) with the reference on identical inputs.
//
// Reference interfaces called (file:line in the reference tree):
//   eigensolver(dim,bc,tc,rA,iA,rE,iE,rV,iV)   utils.cpp:578
//   test_zgehrd / test_zunghr                    Toolbox.hpp:2150 / 2179
//   zlahqr_new                                   Toolbox.hpp:1984
//   test_ztrevc                                  Toolbox.hpp:2060
//   call_zlahqr_new                              utils.cpp:159
//   call_ztrevc                                  utils.cpp:279
//   hessenberg                                   utils.cpp:417
//   call_zgemv / call_backsubstitution           utils.cpp:218 / 269
//   zlahqr_ (eigenvalues only)                   Toolbox.hpp:1663
//   zlahqr (polynomial roots kernel)             Toolbox.hpp:1894
#include <omp.h>

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

#include "definitions.hpp"
struct ComplexDouble;
#include "utils.h"
#include "Toolbox-def.hpp"

// Not declared in the reference's headers (defined in Toolbox.hpp:2035, 2090).
void test_diag(double *rM, double *iM, double real, double imag, int dim, int block_cnt, int thread_cnt);
void test_backsubstitution(double *rT, double *iT, double *rx, double *ix, int dim, int block_cnt,
                           int thread_cnt);
// Defined in Toolbox.hpp (1663, 1894) and not declared in its headers: the
// eigenvalues-only QR of one matrix slot and the polynomial-roots kernel
// (a host loop over the batch in this CPU build).
void zlahqr_(double *rE, double *iE, double *rV, double *iV, int dim, int block_cnt, int thread_cnt);
void zlahqr(double *rC, double *iC, double *rE, double *iE, double *rV, double *iV, double *rZ, double *iZ,
            int dim, int dim1, int dim2, int block_cnt, int thread_cnt);
// Declared only in the Cython binding (python_interface.pyx:24-25).
void call_backsubstitution(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rb,
                           double *ib);

namespace {
// eigensolver() prints its elapsed time on std::cout (utils.cpp:619); keep
// that out of the callers' output.
struct MuteCout {
  std::streambuf *old;
  std::ostringstream sink;
  MuteCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
  ~MuteCout() { std::cout.rdbuf(old); }
};

double *alloc(size_t n) {
  void *p = nullptr;
  if (posix_memalign(&p, 4096, n * sizeof(double)) != 0) abort();
  std::memset(p, 0, n * sizeof(double));
  return static_cast<double *>(p);
}

// Zero A(i+2:n, i) for every column i: the entries where zgehrd leaves its
// reflector vectors (Toolbox.hpp:1374-1384).  LAPACK's zlahqr clears them
// itself; the reference's zlahqr_schur does not (Toolbox.hpp:721-722).
void clear_below_subdiag(int dim, int block_cnt, int thread_cnt, double *rA, double *iA) {
  const size_t T = thread_cnt;
  const size_t n2 = (size_t)dim * dim;
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t)
      for (int c = 0; c < dim; ++c)
        for (int r = c + 2; r < dim; ++r) {
          size_t idx = (size_t)b * T * n2 + ((size_t)r + (size_t)c * dim) * T + t;
          rA[idx] = 0.0;
          iA[idx] = 0.0;
        }
}
}  // namespace

extern "C" {

// Threads of the reference's OpenMP loops for calls from this thread
// (<= 0: the number of processors); returns the previous setting.  The
// bit-exact comparisons run the reference single-threaded so that they
// compare arithmetic, not the OpenMP schedule.
int ref_set_threads(int nthreads) {
  const int prev = omp_get_max_threads();
  omp_set_num_threads(nthreads > 0 ? nthreads : omp_get_num_procs());
  return prev;
}

// Threads the reference's OpenMP loops use for calls from this thread.
int ref_get_threads(void) { return omp_get_max_threads(); }

// The reference's fused host call, exactly as shipped.
void ref_eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rE,
                     double *iE, double *rV, double *iV) {
  MuteCout mute;
  eigensolver(dim, block_cnt, thread_cnt, rA, iA, rE, iE, rV, iV);
}

// The stage sequence of eigensolver() exactly as shipped (utils.cpp:598-615,
// the reflector entries left below the subdiagonal during QR), but with tau
// and the work vectors zeroed: the shipped call takes them from
// posix_memalign uninitialised, and the reference's zgemv forms beta*y
// before adding A*x even for beta = 0 (Toolbox.hpp:1070-1080), so its result
// depends on whatever the heap held (NaN bytes give NaN eigenvalues, a
// negative value a -0 that can flip a later sign test).  Zeroed buffers make
// the as-shipped path deterministic, as the oracle assumes.
void ref_eigensolver_shipped(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                             double *rE, double *iE, double *rV, double *iV) {
  const size_t nm = (size_t)block_cnt * thread_cnt;
  const size_t sizem = (size_t)dim * dim * nm, sizev = (size_t)dim * nm;
  double *rtau = alloc(sizev), *itau = alloc(sizev);
  double *rwork = alloc(2 * sizev), *iwork = alloc(2 * sizev);
  test_zgehrd(rA, iA, rtau, itau, rwork, iwork, dim, block_cnt, thread_cnt);
  std::memcpy(rV, rA, sizem * sizeof(double));
  std::memcpy(iV, iA, sizem * sizeof(double));
  test_zunghr(rV, iV, rtau, itau, rwork, iwork, dim, block_cnt, thread_cnt);
  zlahqr_new(rE, iE, rA, iA, rV, iV, dim, dim * dim, block_cnt, thread_cnt);
  test_ztrevc(rA, iA, rV, iV, rwork, iwork, dim, dim * dim, block_cnt, thread_cnt);
  free(rtau); free(itau); free(rwork); free(iwork);
}

// The same four reference stages in the same order and with the same work
// sizes as utils.cpp:598-615, with one difference: the reflector entries below
// the subdiagonal are zeroed before the QR stage, so zlahqr_schur sees a clean
// Hessenberg matrix (what its own Python tests feed it, python_interface.pyx:174,201).
void ref_eigensolver_clean(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                           double *rE, double *iE, double *rV, double *iV) {
  const size_t nm = (size_t)block_cnt * thread_cnt;
  const size_t sizem = (size_t)dim * dim * nm, sizev = (size_t)dim * nm;
  double *rtau = alloc(sizev), *itau = alloc(sizev);
  double *rwork = alloc(2 * sizev), *iwork = alloc(2 * sizev);
  test_zgehrd(rA, iA, rtau, itau, rwork, iwork, dim, block_cnt, thread_cnt);
  std::memcpy(rV, rA, sizem * sizeof(double));
  std::memcpy(iV, iA, sizem * sizeof(double));
  test_zunghr(rV, iV, rtau, itau, rwork, iwork, dim, block_cnt, thread_cnt);
  clear_below_subdiag(dim, block_cnt, thread_cnt, rA, iA);
  zlahqr_new(rE, iE, rA, iA, rV, iV, dim, dim * dim, block_cnt, thread_cnt);
  test_ztrevc(rA, iA, rV, iV, rwork, iwork, dim, dim * dim, block_cnt, thread_cnt);
  free(rtau); free(itau); free(rwork); free(iwork);
}

// Stage entries.
void ref_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                double *itau) {
  call_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rtau, itau);
}
void ref_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                double *itau) {
  call_zunghr(dim, block_cnt, thread_cnt, rA, iA, rtau, itau);
}
void ref_hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rQ,
                    double *iQ) {
  hessenberg(dim, block_cnt, thread_cnt, rA, iA, rQ, iQ);
}
// zlahqr_new with a caller-supplied Z (Toolbox.hpp:1984).
void ref_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rH,
                    double *iH, double *rZ, double *iZ) {
  zlahqr_new(rE, iE, rH, iH, rZ, iZ, dim, dim * dim, block_cnt, thread_cnt);
}
// call_zlahqr_new as shipped: initialises Z itself (utils.cpp:167-169).
void ref_call_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE,
                         double *rM, double *iM, double *rS, double *iS) {
  call_zlahqr_new(dim, block_cnt, thread_cnt, rE, iE, rM, iM, rS, iS);
}
void ref_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                double *iR) {
  call_ztrevc(dim, block_cnt, thread_cnt, rT, iT, rR, iR);
}
void ref_zgemv(int hermitian, int dim, int block_cnt, int thread_cnt, double *rM, double *iM,
               double *rx, double *ix, double *ry, double *iy, double rbeta, double ibeta) {
  call_zgemv(hermitian != 0, dim, block_cnt, thread_cnt, rM, iM, rx, ix, ry, iy, rbeta, ibeta);
}
void ref_backsubstitution(int dim, int block_cnt, int thread_cnt, double *rT, double *iT,
                          double *rb, double *ib) {
  call_backsubstitution(dim, block_cnt, thread_cnt, rT, iT, rb, ib);
}
// zlahqr_ on every matrix slot of the batch (H Hessenberg, updated in place).
void ref_zlahqr_eig(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rH,
                    double *iH) {
  const size_t T = thread_cnt, n2 = (size_t)dim * dim;
  for (int b = 0; b < block_cnt; ++b)
    for (int t = 0; t < thread_cnt; ++t) {
      const size_t m = (size_t)b * T * n2 + t, v = (size_t)b * T * dim + t;
      zlahqr_(rE + v, iE + v, rH + m, iH + m, dim, block_cnt, thread_cnt);
    }
}
// The roots kernel as shipped: coefficients c_0..c_dim per polynomial.
void ref_zlahqr_roots(int dim, int block_cnt, int thread_cnt, double *rC, double *iC, double *rE,
                      double *iE, double *rV, double *iV, double *rZ, double *iZ) {
  zlahqr(rC, iC, rE, iE, rV, iV, rZ, iZ, dim, dim + 1, dim * dim, block_cnt, thread_cnt);
}

}  // extern "C"
