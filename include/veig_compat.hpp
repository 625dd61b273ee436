// veig_compat.hpp -- drop-in replacement for the reference's host API
// header (utils.h:37-65) and batch entry points (Toolbox-def.hpp:5-22):
// the same C++ names and argument lists on host pointers in the tiled
// split-plane layout, backed by the gfx950 kernels of libveig.so.
//
// A program written against the reference switches by including this header
// instead of "utils.h" / "Toolbox-def.hpp" and linking libveig.so; nothing
// else changes.  Each call copies its inputs to the GPU, runs, and copies the
// results back before returning (as the reference's CUDA branch does,
// utils.cpp:631-692).  Failures raise veig::Error (the reference returns void
// and reports nothing).
//
// Semantics that differ from the reference, because the reference's are
// defects (see the oracle's switches and the tests that pin them):
//   * eigensolver(): the QR stage runs on a clean Hessenberg matrix (the
//     reference leaves zgehrd's reflectors below the subdiagonal,
//     Toolbox.hpp:721-722, which corrupts its result); A receives the Schur
//     form T unscaled (the reference rescales T's columns in ztrevc,
//     Toolbox.hpp:1278-1295).
//   * zgehrd's reflectors use LAPACK's sign and guard (Toolbox.hpp:1366-1373
//     divides by zero on triangular columns); H, tau differ accordingly, Q,
//     eigenvalues and eigenvectors do not.
//   * call_zgemv with hermitian = true adds to Im(y) (Toolbox.hpp:1125 adds
//     Re(y)).
#pragma once

#include <stdexcept>
#include <string>

#include "veig.h"
#include "veig_complex.hpp"

namespace veig {
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};
}  // namespace veig

// layout helpers (utils.cpp:85-157)
void rearrangeMatrices(int dim, int block_cnt, int thread_cnt, double *in, double *out);
void arrangeMatricesAoS(int dim, int block_cnt, int thread_cnt, double *in, double *out);
void repMatrixSoA(int dim, int block_cnt, int thread_cnt, double *matrix, double *out);
void repVectorSoA(int dim, int block_cnt, int thread_cnt, double *vec, double *out);

// host API (utils.h:42-65)
void call_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rM,
                     double *iM, double *rS, double *iS);
void call_zgemv(bool hermitian, int dim, int block_cnt, int thread_cnt, double *rM, double *iM,
                double *rx, double *ix, double *ry, double *iy, double rbeta, double ibeta);
void call_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                 double *iR);
void call_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                 double *itau);
void call_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                 double *itau);
void call_backsubstitution(int dim, int block_cnt, int thread_cnt, double *rT, double *iT,
                           double *rb, double *ib);
void hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rQ,
                double *iQ);
void eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rE,
                 double *iE, double *rV, double *iV);
// loadPackages (utils.cpp:17-50): `packages` files testcases/f_coeffs_<k>.csv
// (k = 1..packages, relative to the working directory), each with `matrices`
// lines of 2*n_coeffs comma-separated numbers (re, im of c_0 .. c_{n-1});
// C[coeff + line*n_coeffs + package*matrices*n_coeffs].  Throws veig::Error
// if a file cannot be opened (the reference reads nothing and goes on).
void loadPackages(int packages, int matrices, int n_coeffs, ComplexDouble *C);
// raw binary readers (utils.cpp:52-83): dim*dim (readMatrix), dim
// (readVector) or dim*dim*n_matrices (readMatricesFromFile) native doubles
// from `file` into `array`.  Sizes are 64-bit (the reference's int byte counts
// overflow at 2^31); a file that cannot be opened or holds fewer values
// throws veig::Error (the reference leaves the array partly unread).
void readMatrix(int dim, std::string file, double *array);
void readVector(int dim, std::string file, double *array);
void readMatricesFromFile(int dim, int n_matrices, double *array, std::string file);

// batch entry points (Toolbox-def.hpp:8-22).  The work arrays are accepted
// and unused: the kernels keep their work on chip.
void test_zgemv(bool hermitian, double *rA, double *iA, double *rx, double *ix, double rbeta,
                double ibeta, double *ry, double *iy, int dim, int block_cnt, int thread_cnt);
void test_ztrevc(double *rV, double *iV, double *rZ, double *iZ, double *rwork, double *iwork,
                 int dim, int dim2, int block_cnt, int thread_cnt);
void zlahqr_new(double *rE, double *iE, double *rV, double *iV, double *rZ, double *iZ, int dim,
                int dim2, int block_cnt, int thread_cnt);
void test_ident(double *rZ, double *iZ, int dim, int block_cnt, int thread_cnt);
// the polynomial-roots kernel zlahqr (Toolbox.hpp:1894): dim1 = dim + 1
// coefficients per polynomial, dim2 = dim * dim; E = roots, V = Schur form of
// the companion matrix, Z = Schur vectors
void zlahqr(double *rC, double *iC, double *rE, double *iE, double *rV, double *iV, double *rZ,
            double *iZ, int dim, int dim1, int dim2, int block_cnt, int thread_cnt);
void test_diag(double *rM, double *iM, double real, double imag, int dim, int block_cnt,
               int thread_cnt);
void test_zgehrd(double *rA, double *iA, double *rtau, double *itau, double *rwork, double *iwork,
                 int dim, int block_cnt, int thread_cnt);
void test_zunghr(double *rA, double *iA, double *rtau, double *itau, double *rwork, double *iwork,
                 int dim, int block_cnt, int thread_cnt);
void test_backsubstitution(double *rT, double *iT, double *rx, double *ix, int dim, int block_cnt,
                           int thread_cnt);

// Per-matrix entries of the header library (Toolbox.hpp:1315, 1492, 690,
// 1133): ONE matrix of the tiled layout, the pointers pre-offset to its slot
// and its elements thread_cnt apart (element (r, c) of the dim x dim matrix
// at p[(r + c*dim) * thread_cnt], vector entry i at p[i * thread_cnt]), as
// the batch entries above pass them.  Each call gathers the matrix, runs the
// batch entry on it as a batch of one on the GPU and scatters the results
// back, so its values are the batch entry's for that slot; a caller that
// loops over slots should call the batch entry instead (one launch instead
// of one per matrix).  The work arrays are accepted and unused.
void zgehrd(double *rA, double *iA, double *rtau, double *itau, double *rwork, double *iwork, int dim,
            int block_cnt, int thread_cnt);
void zunghr(double *rA, double *iA, double *rtau, double *itau, double *rwork, double *iwork, int dim,
            int block_cnt, int thread_cnt);
// wantT = false computes T in full as well (the reference leaves the entries
// outside the active block unupdated; every reference caller passes true);
// wantZ = false leaves Z untouched.
void zlahqr_schur(bool wantT, bool wantZ, int dim, double *rV, double *iV, double *rE, double *iE,
                  double *rZ, double *iZ, int block_cnt, int thread_cnt);
// right eigenvectors into R (which holds the Schur vectors on entry); T is
// left as it is; L is not touched (the reference's left eigenvectors are
// unimplemented, its batch entry passes none, Toolbox.hpp:2079)
void ztrevc(double *rT, double *iT, double *rL, double *iL, double *rR, double *iR, double *rwork,
            double *iwork, int dim, int block_cnt, int thread_cnt);
