/* TEST INFRASTRUCTURE ONLY -- the CPU oracle for the batched complex
 * eigensolver.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, and only as a checker; the product library never
 * links or calls it.
 *
 * Layout (tiled split-plane SoA, the reference's "vectorized data structure"):
 *   matrix m = b*T + t, element (r,c) 0-based, column-major:
 *       idx = b*T*n*n + (r + c*n)*T + t                 (utils.cpp:96)
 *   vector element r:  idx = b*T*n + r*T + t            (utils.cpp:153)
 * with T = thread_cnt, b < block_cnt.
 */
#ifndef VEIG_ORACLE_H_
#define VEIG_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Behaviour switches.  0 = the reference's own behaviour where it is well
 * defined, i.e. what the pinning tests compare bit for bit. */
enum {
  /* Keep zgehrd's reflector vectors below the subdiagonal when the QR stage
   * starts (what eigensolver() does, utils.cpp:606-613).  Off = zero them
   * first, as LAPACK's zlahqr does and as the reference's own QR tests feed. */
  ORACLE_KEEP_TRASH = 1,
  /* zgehrd with LAPACK zlarfg's sign (beta = -sign(Re alpha)*norm) and its
   * norm==0 guard instead of the reference's +sign (Toolbox.hpp:1367-1373). */
  ORACLE_LAPACK_SIGN = 2,
  /* Deflation test's tst==0 branch with Re H(K+1,K) (LAPACK) instead of the
   * reference's Im H(K+1,K) (Toolbox.hpp:840). */
  ORACLE_LAPACK_TST = 4,
  /* ztrevc without rescaling column ki of T afterwards (Toolbox.hpp:1278-1295). */
  ORACLE_NO_T_RESCALE = 8,
  /* Hermitian zgemv reading the imaginary part of y (LAPACK) instead of the
   * reference's iv = ry[...] (Toolbox.hpp:1125). */
  ORACLE_ZGEMV_FIX = 16,
  /* QR stage as the eigenvalues-only zlahqr_ (Toolbox.hpp:1663-1890): no Z,
   * no initial real-subdiagonal pass, updates inside the active block only.
   * Set by oracle_zlahqr_eig. */
  ORACLE_EIG_ONLY = 32
};

/* Per-matrix statistics from the QR stage (any pointer may be NULL).
 * info[m]  : number of eigenvalues accepted after the iteration cap without a
 *            negligible subdiagonal (the reference reports nothing, Toolbox.hpp:806-808)
 * sweeps[m]: QR sweeps performed;  steps[m]: bulge-chase steps performed. */
typedef struct {
  int32_t *info;
  int32_t *sweeps;
  int32_t *steps;
} oracle_stats;

void oracle_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                   double *itau, int flags);
void oracle_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                   const double *rtau, const double *itau, int flags);
void oracle_hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rQ,
                       double *iQ, int flags);
void oracle_zlahqr(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rH,
                   double *iH, double *rZ, double *iZ, int flags, oracle_stats *stats);
void oracle_call_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE,
                            double *rM, double *iM, double *rS, double *iS, int flags);
/* Eigenvalues only (zlahqr_, Toolbox.hpp:1663): H upper Hessenberg with a
 * real subdiagonal, updated in place inside the active blocks. */
void oracle_zlahqr_eig(int dim, int block_cnt, int thread_cnt, double *rE, double *iE, double *rH,
                       double *iH, int flags, oracle_stats *stats);
/* Companion matrices of polynomials c_0 + c_1 x + ... + c_dim x^dim, the
 * coefficients a vector slot of length dim+1 (Toolbox.hpp:605-643). */
void oracle_companion(int dim, int block_cnt, int thread_cnt, const double *rC, const double *iC,
                      double *rM, double *iM);
/* Polynomial roots as the reference's zlahqr kernel (Toolbox.hpp:1894):
 * E = roots, V = Schur form of the companion matrix, Z = Schur vectors. */
void oracle_zlahqr_roots(int dim, int block_cnt, int thread_cnt, const double *rC, const double *iC,
                         double *rE, double *iE, double *rV, double *iV, double *rZ, double *iZ,
                         int flags, oracle_stats *stats);
void oracle_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                   double *iR, int flags);
void oracle_eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rE,
                        double *iE, double *rV, double *iV, int flags, oracle_stats *stats);
void oracle_zgemv(int hermitian, int dim, int block_cnt, int thread_cnt, const double *rM,
                  const double *iM, const double *rx, const double *ix, double *ry, double *iy,
                  double rbeta, double ibeta, int flags);
void oracle_backsubstitution(int dim, int block_cnt, int thread_cnt, const double *rT,
                             const double *iT, double *rb, double *ib);
/* Threads the OpenMP loops use (0 = library default). */
void oracle_set_threads(int nthreads);
int oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
