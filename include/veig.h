/* veig.h -- C ABI of the MI355X (gfx950) batched complex eigensolver.
 *
 * Drop-in for the hot path of the reference header library: the batched
 * eigendecomposition zgehrd -> zunghr -> zlahqr -> ztrevc of equal-size complex
 * n x n matrices held in the reference's tiled split-plane layout.
 *
 * Layout (the reference's "vectorized data structure", utils.cpp:85-157):
 *   T = thread_cnt, matrix m = b*T + t (b < block_cnt, t < T)
 *   matrix element (r,c), 0-based, column-major:  b*T*n*n + (r + c*n)*T + t
 *   vector element r (eigenvalues, tau):           b*T*n   + r*T       + t
 * Real and imaginary parts live in separate double planes (rX, iX).
 *
 * Two families of entry points:
 *   veig_*       device pointers, enqueued on `stream` (a hipStream_t; NULL =
 *                the default stream), asynchronous, no allocation except where
 *                stated.
 *   veig_host_*  host pointers with the reference's exact argument lists
 *                (utils.h:37-65); they allocate device memory, copy in, run the
 *                device entry and copy back, like utils.cpp:631-692, and return
 *                after the results are on the host.
 * Every function returns a veig_status; veig_last_error() describes the last
 * failure of the calling thread.  Sizes are 64-bit (the reference's 32-bit
 * int byte counts overflow at 2^31 bytes per plane).
 */
#ifndef VEIG_H_
#define VEIG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  VEIG_OK = 0,
  VEIG_ERR_INVALID = 1,     /* bad argument (negative size, NULL pointer, ...) */
  VEIG_ERR_UNSUPPORTED = 2, /* order beyond what the entry point handles */
  VEIG_ERR_HIP = 3,         /* HIP runtime error (launch, copy, allocation) */
  VEIG_ERR_NOMEM = 4        /* device allocation failed */
} veig_status;

/* flags for veig_zgeev */
#define VEIG_WRITE_SCHUR 1u /* also store the Schur form T over A (zeros below the diagonal) */

/* Largest orders.  VEIG_MAX_DIM: veig_zgeev (FP64), veig_host_eigensolver and
 * the stage entries (orders above 64, and above 32 at the stage entries, run
 * on one workgroup per matrix with the matrix in LDS: 96 x 97 complex doubles
 * fill the LDS of a CU).  VEIG_MAX_DIM_FAST: veig_zgeev_fp32,
 * veig_zgeev_mixed, the eigenvalues-only and roots entries.
 * VEIG_MAX_DIM_VARSIZE: veig_zgeev_varsize (larger orders get info = -1). */
#define VEIG_MAX_DIM 96
#define VEIG_MAX_DIM_STAGES VEIG_MAX_DIM
#define VEIG_MAX_DIM_FAST 64
#define VEIG_MAX_DIM_VARSIZE 32

const char *veig_status_string(int status);
const char *veig_last_error(void);
/* Build/version string and the device the library runs on. */
const char *veig_version(void);

/* ---------------------------------------------------------------------------
 * Fused eigendecomposition, device pointers.
 * Replaces eigensolver(dim, block_cnt, thread_cnt, rA, iA, rE, iE, rV, iV)
 * (utils.cpp:578-698): E receives the eigenvalues in Schur (deflation) order,
 * column k of V the right eigenvector Z*[x;1;0] of E(k) (the reference's
 * scaling: the k-th entry in the Schur basis is 1, not normalised,
 * Toolbox.hpp:1272-1275).  A is read only, unless VEIG_WRITE_SCHUR.
 * info (may be NULL): per matrix, the number of eigenvalues accepted after
 * `itmax` sweeps without a negligible subdiagonal (the reference caps at 30,
 * Toolbox.hpp:808, and reports nothing; itmax <= 0 selects 30).
 * A, E and V are separate arrays: at orders 9..64 the call passes the
 * Hessenberg reduction's compact form between its kernels through V and E,
 * which then receive the results; above 64, V holds Q and then the Schur
 * vectors on the way to the eigenvectors.
 * ------------------------------------------------------------------------- */
int veig_zgeev(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
               double *rE, double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
               uint32_t flags, void *stream);

/* Same with the reduction, QR and eigenvector stages in complex64 (inputs
 * and outputs stay complex128), orders up to VEIG_MAX_DIM_FAST.  No
 * refinement: the building block of the mixed-precision path, also usable on
 * its own. */
int veig_zgeev_fp32(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rA,
                    const double *iA, double *rE, double *iE, double *rV, double *iV,
                    int32_t *info, int32_t itmax, void *stream);

/* ---------------------------------------------------------------------------
 * Mixed precision.  No counterpart in the reference (new in this design, for
 * the same eigendecomposition as veig_zgeev, orders up to VEIG_MAX_DIM_FAST): the
 * reduction, QR sweeps and eigenvector solves run in complex64; every eigenpair is then refined in
 * FP64 by Newton's method on R = A V - V diag(E) until the backward error
 * ||R||_F / (||A||_F ||V||_F) <= tol (at most 3 corrections); matrices that do
 * not get there, or whose complex64 QR did not converge, are solved again by
 * the FP64 path.  A is read only.
 *   info     (may be NULL): 0, or for matrices re-solved in FP64 that path's info
 *   fallback (may be NULL): 1 for matrices re-solved in FP64, else 0
 *   berr     (may be NULL): final backward error per matrix
 *   workspace: device memory of at least veig_mixed_workspace_size() bytes,
 *              256-byte aligned; contents need no initialisation.
 * ------------------------------------------------------------------------- */
size_t veig_mixed_workspace_size(int64_t dim, int64_t block_cnt, int64_t thread_cnt);
int veig_zgeev_mixed(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rA,
                     const double *iA, double *rE, double *iE, double *rV, double *iV,
                     int32_t *info, int32_t *fallback, double *berr, double tol, int32_t itmax,
                     void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Batches of different orders.  No counterpart in the reference (its batches
 * are equal-size, utils.cpp:578); the per-matrix arithmetic is veig_zgeev's.
 * Matrix m (order dims[m], 0 <= dims[m] <= 32) is column-major with leading
 * dimension dims[m] at offset a_off[m] of the rA/iA planes; its eigenvectors
 * go to the same offsets of rV/iV and its eigenvalues to e_off[m] of rE/iE.
 * The matrices are sorted on the device by the kernel variant chosen for
 * their order, variant_of_n[n] (33 int8 in device memory, may be NULL for the
 * defaults): 0 = lane groups of 8 (one lane per index), 1 = groups of 8 with
 * 2 indices per lane (4 lanes per matrix), 2 = groups of 16, 3 = groups of 16
 * with 2 indices per lane (8 lanes per matrix), 4 = groups of 32; a variant
 * too narrow for n is replaced by the default (0 for n <= 8, 2 for n <= 16,
 * 4 above).  One persistent kernel per variant (grid_per_bin[5] workgroups,
 * a host array read when the launches are enqueued or captured; NULL or 0 =
 * default).
 * info[m] (may be NULL): as veig_zgeev, -1 for an order outside [0, 32].
 * Enqueues only kernels (no allocation, copy or memset), so it can be
 * captured into a graph; veig_varsize_graph_create does that.
 * workspace: >= veig_varsize_workspace_size(nmat) bytes, 256-byte aligned.
 * ------------------------------------------------------------------------- */
typedef struct veig_graph veig_graph;
size_t veig_varsize_workspace_size(int64_t nmat);
int veig_zgeev_varsize(int64_t nmat, const int32_t *dims, const int64_t *a_off,
                       const int64_t *e_off, const double *rA, const double *iA, double *rE,
                       double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
                       const int8_t *variant_of_n, const int32_t *grid_per_bin, void *workspace,
                       size_t workspace_bytes, void *stream);
/* veig_zgeev_varsize without a workspace: no counterpart in the reference.
 * For callers that cannot allocate (code under graph capture, serving loops,
 * harnesses that forbid hipMalloc): the binning state and permutation live in
 * a static device array of the library (16 MiB per GPU, held until the
 * process ends), so the call allocates nothing and enqueues only kernels.
 * nmat is limited to 4,194,304 (VEIG_ERR_INVALID beyond).  Calls share their
 * GPU's array: calls that overlap on one GPU (other streams, other threads)
 * race; calls in order on one stream do not.  The array is looked up on the
 * first call on each device.  Other arguments as veig_zgeev_varsize. */
int veig_zgeev_varsize_noalloc(int64_t nmat, const int32_t *dims, const int64_t *a_off,
                               const int64_t *e_off, const double *rA, const double *iA, double *rE,
                               double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
                               const int8_t *variant_of_n, const int32_t *grid_per_bin, void *stream);
/* Capture the same launch sequence into an instantiated hipGraph (arguments as
 * above; pointers are baked in, array contents are read at each launch). */
int veig_varsize_graph_create(int64_t nmat, const int32_t *dims, const int64_t *a_off,
                              const int64_t *e_off, const double *rA, const double *iA,
                              double *rE, double *iE, double *rV, double *iV, int32_t *info,
                              int32_t itmax, const int8_t *variant_of_n, const int32_t *grid_per_bin,
                              void *workspace, size_t workspace_bytes, veig_graph **out);
/* Autotuner for batches like this one (no counterpart in the reference):
 * times every kernel variant on the matrices of each order n <= 16 (lanes per
 * matrix 4/8/16/32 as the widths allow), keeps the fastest per order, then
 * times grids of 2/4/8/16 workgroups per CU for each variant in use.  Writes
 * variant_of_n_out[33] and grid_per_bin_out[5] (host arrays) for
 * veig_zgeev_varsize / veig_varsize_graph_create.  Runs the batch several
 * times (its outputs are overwritten), synchronises the stream and allocates
 * a little device scratch: an offline step, not for capture. */
int veig_varsize_autotune(int64_t nmat, const int32_t *dims, const int64_t *a_off,
                          const int64_t *e_off, const double *rA, const double *iA, double *rE,
                          double *iE, double *rV, double *iV, int32_t *info, int32_t itmax,
                          int8_t *variant_of_n_out, int32_t *grid_per_bin_out, void *workspace,
                          size_t workspace_bytes, void *stream);
int veig_graph_launch(veig_graph *graph, void *stream);
int veig_graph_destroy(veig_graph *graph);

/* ---------------------------------------------------------------------------
 * Stage entries, device pointers (the batch kernels of Toolbox-def.hpp:5-22).
 * ------------------------------------------------------------------------- */
/* test_zgehrd (Toolbox.hpp:2150) / call_zgehrd (utils.cpp:330): A -> upper
 * Hessenberg H on and above the subdiagonal, reflector vectors v(i+2:n) below
 * it, tau(1:n-1) (LAPACK zgehd2 convention; LAPACK zlarfg sign). */
int veig_zgehrd(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
                double *rtau, double *itau, void *stream);
/* test_zunghr (Toolbox.hpp:2179) / call_zunghr (utils.cpp:376): reflectors in
 * A (as left by veig_zgehrd) and tau -> the unitary Q, in place in A. */
int veig_zunghr(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
                const double *rtau, const double *itau, void *stream);
/* hessenberg (utils.cpp:417): A -> H (with reflectors, as zgehrd) and Q. */
int veig_hessenberg(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rA, double *iA,
                    double *rQ, double *iQ, void *stream);
/* zlahqr_new (Toolbox.hpp:1984): upper Hessenberg H (entries below the
 * subdiagonal are ignored) -> Schur form T in place (zeros below the
 * diagonal), Z <- Z * (Schur vectors of H), E = diag(T). */
int veig_zlahqr(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rE, double *iE,
                double *rH, double *iH, double *rZ, double *iZ, int32_t *info, int32_t itmax,
                void *stream);
/* test_ztrevc (Toolbox.hpp:2060) / call_ztrevc (utils.cpp:279): upper
 * triangular T (read only), R <- R * X with X the unit upper triangular
 * eigenvector matrix of T (column k of R*X is the eigenvector of T(k,k)). */
int veig_ztrevc(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rT,
                const double *iT, double *rR, double *iR, void *stream);
/* test_zgemv (Toolbox.hpp:2118): y <- beta*y + A*x, or beta*y + A^H*x. */
int veig_zgemv(int hermitian, int64_t dim, int64_t block_cnt, int64_t thread_cnt,
               const double *rM, const double *iM, const double *rx, const double *ix,
               double *ry, double *iy, double rbeta, double ibeta, void *stream);
/* test_backsubstitution (Toolbox.hpp:2090): solve T x = b (T upper
 * triangular), x over b in place. */
int veig_backsubstitution(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rT,
                          const double *iT, double *rb, double *ib, void *stream);

/* ---------------------------------------------------------------------------
 * Eigenvalues only, and polynomial roots, device pointers.
 * ------------------------------------------------------------------------- */
/* zlahqr_ (Toolbox.hpp:1663-1890, the eigenvalues-only QR) on every matrix
 * of the batch: H upper Hessenberg with a real subdiagonal (entries below the
 * subdiagonal are ignored), updated in place inside the active blocks only
 * (as the reference; T is not formed), E = the eigenvalues in deflation
 * order.  info as veig_zgeev. */
int veig_zlahqr_eig(int64_t dim, int64_t block_cnt, int64_t thread_cnt, double *rE, double *iE,
                    double *rH, double *iH, int32_t *info, int32_t itmax, void *stream);
/* Eigenvalues of general matrices: the Hessenberg reduction without Q, then
 * the iteration of veig_zlahqr_eig (A read only).  No single counterpart in
 * the reference (zgehrd followed by its zlahqr_); E is the E of veig_zgeev on
 * the same input, to rounding. */
int veig_eigvals(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rA, const double *iA,
                 double *rE, double *iE, int32_t *info, int32_t itmax, void *stream);
/* The polynomial-roots kernel zlahqr (Toolbox.hpp:1894-1957): for each
 * polynomial c_0 + c_1 x + ... + c_dim x^dim (coefficients rC/iC in the
 * vector layout with dim+1 entries per polynomial) the companion matrix of
 * create_companion_matrices (Toolbox.hpp:605-643) is iterated to Schur form;
 * E = the roots.  With rV, iV, rZ, iZ (all four) V receives the Schur form T
 * and Z the Schur vectors, as in the reference; with all four NULL only the
 * roots are computed (the eigenvalues-only iteration). */
int veig_zlahqr_roots(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *rC,
                      const double *iC, double *rE, double *iE, double *rV, double *iV, double *rZ,
                      double *iZ, int32_t *info, int32_t itmax, void *stream);

/* ---------------------------------------------------------------------------
 * Layout helpers on the device (utils.cpp:85-157).  `in` and `out` must not
 * overlap.  AoS = one column-major n x n matrix after the other.
 * ------------------------------------------------------------------------- */
/* rearrangeMatrices (utils.cpp:85): AoS -> tiled SoA, one plane. */
int veig_rearrange_matrices(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *in,
                            double *out, void *stream);
/* arrangeMatricesAoS (utils.cpp:105): tiled SoA -> AoS, one plane. */
int veig_arrange_matrices_aos(int64_t dim, int64_t block_cnt, int64_t thread_cnt,
                              const double *in, double *out, void *stream);
/* repMatrixSoA (utils.cpp:125): one column-major n x n plane (device) copied
 * into every batch slot. */
int veig_rep_matrix_soa(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *matrix,
                        double *out, void *stream);
/* repVectorSoA (utils.cpp:145). */
int veig_rep_vector_soa(int64_t dim, int64_t block_cnt, int64_t thread_cnt, const double *vec,
                        double *out, void *stream);

/* ---------------------------------------------------------------------------
 * Host-pointer mirrors of the reference host API (utils.h:37-65; the first
 * two are the Cython binding's externs, python_interface.pyx:12-28).
 * ------------------------------------------------------------------------- */
/* eigensolver(dim, block_cnt, thread_cnt, rA, iA, rE, iE, rV, iV), utils.cpp:578 */
int veig_host_eigensolver(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                          double *rE, double *iE, double *rV, double *iV);
/* call_zgehrd, utils.cpp:330 */
int veig_host_zgehrd(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                     double *itau);
/* call_zunghr, utils.cpp:376 */
int veig_host_zunghr(int dim, int block_cnt, int thread_cnt, double *rA, double *iA, double *rtau,
                     double *itau);
/* hessenberg, utils.cpp:417 */
int veig_host_hessenberg(int dim, int block_cnt, int thread_cnt, double *rA, double *iA,
                         double *rQ, double *iQ);
/* call_zlahqr_new, utils.cpp:159.  As in the reference, S is overwritten with
 * the initial Schur-vector matrix before the QR stage; the reference's
 * argument swap at utils.cpp:169 makes that i*I, reproduced here so results
 * match (pass the fixed-Z variant veig_zlahqr for identity). */
int veig_host_zlahqr_new(int dim, int block_cnt, int thread_cnt, double *rE, double *iE,
                         double *rM, double *iM, double *rS, double *iS);
/* call_ztrevc, utils.cpp:279 */
int veig_host_ztrevc(int dim, int block_cnt, int thread_cnt, double *rT, double *iT, double *rR,
                     double *iR);
/* call_zgemv, utils.cpp:218 */
int veig_host_zgemv(int hermitian, int dim, int block_cnt, int thread_cnt, double *rM, double *iM,
                    double *rx, double *ix, double *ry, double *iy, double rbeta, double ibeta);
/* call_backsubstitution, utils.cpp:269 (declared in python_interface.pyx:24) */
int veig_host_backsubstitution(int dim, int block_cnt, int thread_cnt, double *rT, double *iT,
                               double *rb, double *ib);

#ifdef __cplusplus
}
#endif
#endif /* VEIG_H_ */
