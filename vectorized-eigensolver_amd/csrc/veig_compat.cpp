// C++ drop-in shim: the reference's host API names (utils.h:37-65,
// Toolbox-def.hpp:5-22) over the C ABI of libveig.so.  See
// include/veig_compat.hpp for the semantics.
#include "veig_compat.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace {

void check(int st, const char *what) {
  if (st != VEIG_OK) throw veig::Error(st, std::string(what) + ": " + veig_last_error());
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw veig::Error(VEIG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// device copy of a host plane; copied back on request
struct Dev {
  double *d = nullptr;
  size_t n = 0;
  Dev(const double *h, size_t count) : n(count) {
    if (!n) return;
    hip_check(hipMalloc(&d, n * sizeof(double)), "hipMalloc");
    if (h) hip_check(hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  void back(double *h) const {
    if (n) hip_check(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
  ~Dev() {
    if (d) (void)hipFree(d);
  }
  Dev(const Dev &) = delete;
  Dev &operator=(const Dev &) = delete;
};

size_t nmat(int bc, int tc) { return (size_t)(bc > 0 ? bc : 0) * (size_t)(tc > 0 ? tc : 0); }
size_t msize(int dim, int bc, int tc) { return nmat(bc, tc) * (size_t)(dim > 0 ? dim : 0) * (dim > 0 ? dim : 0); }
size_t vsize(int dim, int bc, int tc) { return nmat(bc, tc) * (size_t)(dim > 0 ? dim : 0); }

void sync() { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

}  // namespace

// ---------------------------------------------------------------------------
// layout helpers (host loops, utils.cpp:85-157)
// ---------------------------------------------------------------------------
void rearrangeMatrices(int dim, int block_cnt, int thread_cnt, double *in, double *out) {
  const size_t n2 = (size_t)dim * dim;
  for (size_t b = 0; b < (size_t)block_cnt; ++b)
    for (size_t t = 0; t < (size_t)thread_cnt; ++t)
      for (size_t e = 0; e < n2; ++e)
        out[b * thread_cnt * n2 + e * thread_cnt + t] = in[(b * thread_cnt + t) * n2 + e];
}

void arrangeMatricesAoS(int dim, int block_cnt, int thread_cnt, double *in, double *out) {
  const size_t n2 = (size_t)dim * dim;
  for (size_t b = 0; b < (size_t)block_cnt; ++b)
    for (size_t t = 0; t < (size_t)thread_cnt; ++t)
      for (size_t e = 0; e < n2; ++e)
        out[(b * thread_cnt + t) * n2 + e] = in[b * thread_cnt * n2 + e * thread_cnt + t];
}

void repMatrixSoA(int dim, int block_cnt, int thread_cnt, double *matrix, double *out) {
  const size_t n2 = (size_t)dim * dim;
  for (size_t e = 0; e < n2; ++e)
    for (size_t b = 0; b < (size_t)block_cnt; ++b)
      for (size_t t = 0; t < (size_t)thread_cnt; ++t) out[b * thread_cnt * n2 + t + e * thread_cnt] = matrix[e];
}

void repVectorSoA(int dim, int block_cnt, int thread_cnt, double *vec, double *out) {
  for (size_t e = 0; e < (size_t)dim; ++e)
    for (size_t b = 0; b < (size_t)block_cnt; ++b)
      for (size_t t = 0; t < (size_t)thread_cnt; ++t) out[b * thread_cnt * dim + t + e * thread_cnt] = vec[e];
}

// ---------------------------------------------------------------------------
// host API
// ---------------------------------------------------------------------------
void call_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rM,
                     double *iM, double *rS, double *iS) {
  check(veig_host_zlahqr_new(dim, block_cnt, thread_cnt, rE, iE, rM, iM, rS, iS), "call_zlahqr_new");
}

void call_zgemv(bool hermitian, int dim, int block_cnt, int thread_cnt, double *rM, double *iM,
                double *rx, double *ix, double *ry, double *iy, double rbeta, double ibeta) {
  check(veig_host_zgemv(hermitian ? 1 : 0, dim, block_cnt, thread_cnt, rM, iM, rx, ix, ry, iy, rbeta,
                        ibeta),
        "call_zgemv");
}

void call_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                 double *iR) {
  check(veig_host_ztrevc(dim, block_cnt, thread_cnt, rT, iT, rR, iR), "call_ztrevc");
}

void call_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                 double *itau) {
  check(veig_host_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rtau, itau), "call_zgehrd");
}

void call_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                 double *itau) {
  check(veig_host_zunghr(dim, block_cnt, thread_cnt, rA, iA, rtau, itau), "call_zunghr");
}

void call_backsubstitution(int dim, int block_cnt, int thread_cnt, double *rT, double *iT,
                           double *rb, double *ib) {
  check(veig_host_backsubstitution(dim, block_cnt, thread_cnt, rT, iT, rb, ib), "call_backsubstitution");
}

void hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rQ,
                double *iQ) {
  check(veig_host_hessenberg(dim, block_cnt, thread_cnt, rA, iA, rQ, iQ), "hessenberg");
}

void loadPackages(int packages, int matrices, int n_coeffs, ComplexDouble *C) {
  for (int package = 0; package < packages; ++package) {
    const std::string path = "testcases/f_coeffs_" + std::to_string(package + 1) + ".csv";
    std::ifstream in(path);
    if (!in) throw veig::Error(VEIG_ERR_INVALID, "loadPackages: cannot open " + path);
    std::string line;
    for (int row = 0; row < matrices; ++row) {
      std::getline(in, line);
      std::stringstream ls(line);
      std::vector<double> v(2 * (size_t)n_coeffs, 0.0);
      std::string cell;
      // as the reference: missing or malformed cells read as 0 (atof)
      for (size_t k = 0; k < v.size() && std::getline(ls, cell, ','); ++k) v[k] = std::atof(cell.c_str());
      for (int c = 0; c < n_coeffs; ++c)
        C[c + (size_t)row * n_coeffs + (size_t)package * matrices * n_coeffs] = ComplexDouble(v[2 * c], v[2 * c + 1]);
    }
  }
}

namespace {
void read_doubles(const std::string &file, size_t count, double *array, const char *what) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw veig::Error(VEIG_ERR_INVALID, std::string(what) + ": cannot open " + file);
  in.read(reinterpret_cast<char *>(array), (std::streamsize)(count * sizeof(double)));
  if ((size_t)in.gcount() != count * sizeof(double))
    throw veig::Error(VEIG_ERR_INVALID, std::string(what) + ": " + file + " holds fewer than " +
                                           std::to_string(count) + " doubles");
}
}  // namespace

void readMatrix(int dim, std::string file, double *array) {
  read_doubles(file, (size_t)(dim > 0 ? dim : 0) * (dim > 0 ? dim : 0), array, "readMatrix");
}

void readVector(int dim, std::string file, double *array) {
  read_doubles(file, (size_t)(dim > 0 ? dim : 0), array, "readVector");
}

void readMatricesFromFile(int dim, int n_matrices, double *array, std::string file) {
  read_doubles(file, msize(dim, n_matrices, 1), array, "readMatricesFromFile");
}

void eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rE,
                 double *iE, double *rV, double *iV) {
  check(veig_host_eigensolver(dim, block_cnt, thread_cnt, rA, iA, rE, iE, rV, iV), "eigensolver");
}

// ---------------------------------------------------------------------------
// batch entry points
// ---------------------------------------------------------------------------
void test_zgemv(bool hermitian, double *rA, double *iA, double *rx, double *ix, double rbeta,
                double ibeta, double *ry, double *iy, int dim, int block_cnt, int thread_cnt) {
  call_zgemv(hermitian, dim, block_cnt, thread_cnt, rA, iA, rx, ix, ry, iy, rbeta, ibeta);
}

void test_ztrevc(double *rV, double *iV, double *rZ, double *iZ, double *, double *, int dim, int,
                 int block_cnt, int thread_cnt) {
  call_ztrevc(dim, block_cnt, thread_cnt, rV, iV, rZ, iZ);
}

void zlahqr_new(double *rE, double *iE, double *rV, double *iV, double *rZ, double *iZ, int dim,
                int, int block_cnt, int thread_cnt) {
  if (dim <= 0 || block_cnt <= 0 || thread_cnt <= 0) {
    check(veig_zlahqr(dim, block_cnt, thread_cnt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr, 30, nullptr),
          "zlahqr_new");
    return;
  }
  const size_t sm = msize(dim, block_cnt, thread_cnt), sv = vsize(dim, block_cnt, thread_cnt);
  Dev h1(rV, sm), h2(iV, sm), z1(rZ, sm), z2(iZ, sm), e1(nullptr, sv), e2(nullptr, sv);
  check(veig_zlahqr(dim, block_cnt, thread_cnt, e1.d, e2.d, h1.d, h2.d, z1.d, z2.d, nullptr, 30, nullptr),
        "zlahqr_new");
  sync();
  h1.back(rV); h2.back(iV); z1.back(rZ); z2.back(iZ); e1.back(rE); e2.back(iE);
}

void zlahqr(double *rC, double *iC, double *rE, double *iE, double *rV, double *iV, double *rZ,
            double *iZ, int dim, int dim1, int dim2, int block_cnt, int thread_cnt) {
  if (dim > 0 && (dim1 != dim + 1 || dim2 != dim * dim))
    throw veig::Error(VEIG_ERR_INVALID, "zlahqr: dim1 must be dim + 1 and dim2 dim * dim");
  if (dim <= 0 || block_cnt <= 0 || thread_cnt <= 0) {
    check(veig_zlahqr_roots(dim, block_cnt, thread_cnt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, nullptr, nullptr, 30, nullptr),
          "zlahqr");
    return;
  }
  const size_t sm = msize(dim, block_cnt, thread_cnt), sv = vsize(dim, block_cnt, thread_cnt);
  const size_t sc = nmat(block_cnt, thread_cnt) * (size_t)(dim + 1);
  Dev c1(rC, sc), c2(iC, sc), v1(nullptr, sm), v2(nullptr, sm), z1(nullptr, sm), z2(nullptr, sm);
  Dev e1(nullptr, sv), e2(nullptr, sv);
  check(veig_zlahqr_roots(dim, block_cnt, thread_cnt, c1.d, c2.d, e1.d, e2.d, v1.d, v2.d, z1.d, z2.d, nullptr,
                          30, nullptr),
        "zlahqr");
  sync();
  e1.back(rE); e2.back(iE); v1.back(rV); v2.back(iV); z1.back(rZ); z2.back(iZ);
}

void test_ident(double *rZ, double *iZ, int dim, int block_cnt, int thread_cnt) {
  const size_t n2 = (size_t)dim * dim;
  for (size_t b = 0; b < (size_t)block_cnt; ++b)
    for (size_t t = 0; t < (size_t)thread_cnt; ++t)
      for (size_t c = 0; c < (size_t)dim; ++c)
        for (size_t r = 0; r < (size_t)dim; ++r) {
          const size_t idx = b * thread_cnt * n2 + (r + c * dim) * thread_cnt + t;
          rZ[idx] = r == c ? 1.0 : 0.0;
          iZ[idx] = 0.0;
        }
}

void test_diag(double *rM, double *iM, double real, double imag, int dim, int block_cnt,
               int thread_cnt) {
  const size_t n2 = (size_t)dim * dim;
  for (size_t b = 0; b < (size_t)block_cnt; ++b)
    for (size_t t = 0; t < (size_t)thread_cnt; ++t)
      for (size_t j = 0; j < (size_t)dim; ++j) {
        const size_t idx = b * thread_cnt * n2 + (j + j * dim) * thread_cnt + t;
        rM[idx] = real;
        iM[idx] = imag;
      }
}

void test_zgehrd(double *rA, double *iA, double *rtau, double *itau, double *, double *, int dim,
                 int block_cnt, int thread_cnt) {
  call_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rtau, itau);
}

void test_zunghr(double *rA, double *iA, double *rtau, double *itau, double *, double *, int dim,
                 int block_cnt, int thread_cnt) {
  call_zunghr(dim, block_cnt, thread_cnt, rA, iA, rtau, itau);
}

void test_backsubstitution(double *rT, double *iT, double *rx, double *ix, int dim, int block_cnt,
                           int thread_cnt) {
  call_backsubstitution(dim, block_cnt, thread_cnt, rT, iT, rx, ix);
}

// ---------------------------------------------------------------------------
// per-matrix entries (Toolbox.hpp:1315, 1492, 690, 1133)
// ---------------------------------------------------------------------------
namespace {
// the count elements of one slot, stride apart from (re, im), gathered into
// contiguous planes (the tiled layout with thread_cnt = 1) and scattered back
struct Slot {
  std::vector<double> re, im;
  double *pr, *pi;
  int64_t stride;
  Slot(double *r, double *i, size_t count, int s, const char *what) : re(count), im(count), pr(r), pi(i), stride(s) {
    if (count && (!r || !i)) throw veig::Error(VEIG_ERR_INVALID, std::string(what) + ": null pointer");
    for (size_t k = 0; k < count; ++k) {
      re[k] = pr[(int64_t)k * stride];
      im[k] = pi[(int64_t)k * stride];
    }
  }
  void back() {
    for (size_t k = 0; k < re.size(); ++k) {
      pr[(int64_t)k * stride] = re[k];
      pi[(int64_t)k * stride] = im[k];
    }
  }
};

bool slot_args(int dim, int thread_cnt, const char *what) {
  if (thread_cnt <= 0) throw veig::Error(VEIG_ERR_INVALID, std::string(what) + ": thread_cnt must be positive");
  return dim > 0;  // an empty matrix: nothing to do, as in the reference
}
}  // namespace

void zgehrd(double *rA, double *iA, double *rtau, double *itau, double *, double *, int dim, int,
            int thread_cnt) {
  if (!slot_args(dim, thread_cnt, "zgehrd")) return;
  Slot a(rA, iA, (size_t)dim * dim, thread_cnt, "zgehrd"), t(rtau, itau, (size_t)dim, thread_cnt, "zgehrd");
  call_zgehrd(dim, 1, 1, a.re.data(), a.im.data(), t.re.data(), t.im.data());
  a.back();
  t.back();
}

void zunghr(double *rA, double *iA, double *rtau, double *itau, double *, double *, int dim, int,
            int thread_cnt) {
  if (!slot_args(dim, thread_cnt, "zunghr")) return;
  Slot a(rA, iA, (size_t)dim * dim, thread_cnt, "zunghr"), t(rtau, itau, (size_t)dim, thread_cnt, "zunghr");
  call_zunghr(dim, 1, 1, a.re.data(), a.im.data(), t.re.data(), t.im.data());
  a.back();
}

void zlahqr_schur(bool, bool wantZ, int dim, double *rV, double *iV, double *rE, double *iE, double *rZ,
                  double *iZ, int, int thread_cnt) {
  if (!slot_args(dim, thread_cnt, "zlahqr_schur")) return;
  const size_t nn = (size_t)dim * dim;
  Slot v(rV, iV, nn, thread_cnt, "zlahqr_schur"), e(rE, iE, (size_t)dim, thread_cnt, "zlahqr_schur");
  std::vector<double> zr(nn, 0.0), zi(nn, 0.0);
  Slot *z = nullptr;
  Slot zs = wantZ ? Slot(rZ, iZ, nn, thread_cnt, "zlahqr_schur") : Slot(nullptr, nullptr, 0, 1, "zlahqr_schur");
  if (wantZ) z = &zs;
  double *pzr = z ? z->re.data() : zr.data(), *pzi = z ? z->im.data() : zi.data();
  zlahqr_new(e.re.data(), e.im.data(), v.re.data(), v.im.data(), pzr, pzi, dim, dim * dim, 1, 1);
  v.back();
  e.back();
  if (z) z->back();
}

void ztrevc(double *rT, double *iT, double *, double *, double *rR, double *iR, double *, double *, int dim, int,
            int thread_cnt) {
  if (!slot_args(dim, thread_cnt, "ztrevc")) return;
  const size_t nn = (size_t)dim * dim;
  Slot t(rT, iT, nn, thread_cnt, "ztrevc"), r(rR, iR, nn, thread_cnt, "ztrevc");
  call_ztrevc(dim, 1, 1, t.re.data(), t.im.data(), r.re.data(), r.im.data());
  r.back();
}
