/*
 * rnp_oracle.h -- CPU restatement of RNP2's LA::QR path (test infrastructure).
 *
 * THIS IS THE ORACLE, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker (or as
 * the timed CPU baseline).  The product path (rnp2_amd/, include/) never links
 * or calls anything here.
 *
 * What it restates (reference file:line):
 *   QR::Tuning            RNP/LA/QR.hpp:41-51       (orc_tuning, defaults 32/2/128)
 *   QR::Factor_unblocked  RNP/LA/QR.hpp:78-102
 *   QR::Factor            RNP/LA/QR.hpp:129-197
 *   QR::MultQ_unblocked   RNP/LA/QR.hpp:234-298
 *   QR::MultQ             RNP/LA/QR.hpp:337-429
 *   QR::GenerateQ_unblocked RNP/LA/QR.hpp:451-485
 *   QR::GenerateQ         RNP/LA/QR.hpp:509-570
 *   Reflector::LastNonzeroColumnLimit/RowLimit  RNP/LA/Reflector.hpp:36-75
 *   Reflector::Generate   RNP/LA/Reflector.hpp:117-158
 *   Reflector::Apply      RNP/LA/Reflector.hpp:191-360 (vone in {0,1}, vconj=false:
 *                         the branches QR reaches; vconj/bottom-one serve LQ/QL/RQ)
 *   Reflector::GenerateBlockTr RNP/LA/Reflector.hpp:573-626 (dir "F", storev "C")
 *   Reflector::ApplyBlock RNP/LA/Reflector.hpp:725-815 (dir "F", storev "C", side L/R)
 *   BLAS::Norm2           RNP/BLAS/BLAS_mix.hpp:725-758 (one-pass scaled sum of squares)
 *   Traits<real>::hypot3  RNP/Types.hpp:73-92
 *   Level-2/3 BLAS (gemv, gerc, trmv, gemm, trmm, trsm): the external Fortran
 *     BLAS that src/BLAS_mix.cpp:104-693 forwards to; restated from the BLAS
 *     reference semantics (beta == 0 clears y/C).
 *   Random::UniformRealVector / GenerateVector  RNP/Random.hpp:45-320
 *   MatrixGen::GenerateDiagonal / RandomMatrix (Nonsymmetric)  RNP/MatrixGen.hpp:31-198
 *
 * Complex numbers are C99 double complex, laid out like std::complex<double>
 * (interleaved re, im).  All sizes are size_t, matrices column-major, 0-based.
 * Functions return void like the reference; bad arguments print an
 * "Assertion failed" line to stderr and continue (RNP/Debug.hpp:9-17).
 */
#ifndef RNP_ORACLE_H
#define RNP_ORACLE_H

#include <stddef.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double complex orc_z;

/* Block-size tuning, QR::Tuning<T> (QR.hpp:41-51).  A NULL pointer means the
 * reference defaults: nb = 32, nbmin = 2, crossover nx = 128. */
typedef struct orc_tuning {
	size_t nb;
	size_t nbmin;
	size_t nx;
} orc_tuning;

/* ---- level-1 / reflector pieces ---- */
double orc_d_norm2(size_t n, const double *x, size_t incx);
double orc_z_norm2(size_t n, const orc_z *x, size_t incx);
double orc_hypot3(double x, double y, double z);

void orc_d_generate(size_t n, double *alpha, double *x, size_t incx, double *tau);
void orc_z_generate(size_t n, orc_z *alpha, orc_z *x, size_t incx, orc_z *tau);

size_t orc_d_last_nonzero_column_limit(size_t m, size_t n, const double *a, size_t lda);
size_t orc_d_last_nonzero_row_limit(size_t m, size_t n, const double *a, size_t lda);
size_t orc_z_last_nonzero_column_limit(size_t m, size_t n, const orc_z *a, size_t lda);
size_t orc_z_last_nonzero_row_limit(size_t m, size_t n, const orc_z *a, size_t lda);

void orc_d_apply(const char *side, int vone, size_t m, size_t n, const double *v, size_t incv,
                 double tau, double *c, size_t ldc, double *work);
void orc_z_apply(const char *side, int vone, size_t m, size_t n, const orc_z *v, size_t incv,
                 orc_z tau, orc_z *c, size_t ldc, orc_z *work);

void orc_d_generate_block_tr(size_t n, size_t k, const double *v, size_t ldv, const double *tau,
                             double *t, size_t ldt);
void orc_z_generate_block_tr(size_t n, size_t k, const orc_z *v, size_t ldv, const orc_z *tau,
                             orc_z *t, size_t ldt);

void orc_d_apply_block(const char *side, const char *trans, size_t m, size_t n, size_t k,
                       const double *v, size_t ldv, const double *t, size_t ldt,
                       double *c, size_t ldc, double *work, size_t ldwork);
void orc_z_apply_block(const char *side, const char *trans, size_t m, size_t n, size_t k,
                       const orc_z *v, size_t ldv, const orc_z *t, size_t ldt,
                       orc_z *c, size_t ldc, orc_z *work, size_t ldwork);

/* ---- QR ---- */
void orc_d_factor_unblocked(size_t m, size_t n, double *a, size_t lda, double *tau, double *work);
void orc_z_factor_unblocked(size_t m, size_t n, orc_z *a, size_t lda, orc_z *tau, orc_z *work);
void orc_d_factor(size_t m, size_t n, double *a, size_t lda, double *tau, size_t *lwork,
                  double *work, const orc_tuning *tune);
void orc_z_factor(size_t m, size_t n, orc_z *a, size_t lda, orc_z *tau, size_t *lwork,
                  orc_z *work, const orc_tuning *tune);
/* Factor stopped after the first `npanels` blocked panels (the CPU-baseline
 * sample: the same loop as orc_d_factor, QR.hpp:173-190, cut short). */
void orc_d_factor_panels(size_t m, size_t n, double *a, size_t lda, double *tau,
                         double *work, size_t nb, size_t npanels);

void orc_d_multq_unblocked(const char *side, const char *trans, size_t m, size_t n, size_t k,
                           const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                           double *work);
void orc_z_multq_unblocked(const char *side, const char *trans, size_t m, size_t n, size_t k,
                           const orc_z *a, size_t lda, const orc_z *tau, orc_z *c, size_t ldc,
                           orc_z *work);
void orc_d_multq(const char *side, const char *trans, size_t m, size_t n, size_t k,
                 const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                 size_t *lwork, double *work, const orc_tuning *tune);
void orc_z_multq(const char *side, const char *trans, size_t m, size_t n, size_t k,
                 const orc_z *a, size_t lda, const orc_z *tau, orc_z *c, size_t ldc,
                 size_t *lwork, orc_z *work, const orc_tuning *tune);

void orc_d_generateq_unblocked(size_t m, size_t n, size_t k, double *a, size_t lda,
                               const double *tau, double *work);
void orc_z_generateq_unblocked(size_t m, size_t n, size_t k, orc_z *a, size_t lda,
                               const orc_z *tau, orc_z *work);
void orc_d_generateq(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                     size_t *lwork, double *work, const orc_tuning *tune);
void orc_z_generateq(size_t m, size_t n, size_t k, orc_z *a, size_t lda, const orc_z *tau,
                     size_t *lwork, orc_z *work, const orc_tuning *tune);

/* ---- LQ (RNP/LA/LQ.hpp) ----
 *   LQ::Factor_unblocked  LQ.hpp:78-102,   LQ::Factor    LQ.hpp:126-190
 *   LQ::MultQ_unblocked   LQ.hpp:220-270,  LQ::MultQ     LQ.hpp:309-405
 *   LQ::GenerateQ_unblocked LQ.hpp:429-469, LQ::GenerateQ LQ.hpp:491-559
 *   Reflector::Apply vconj branches, Reflector.hpp:238-291 */
void orc_d_lq_factor_unblocked(size_t m, size_t n, double *a, size_t lda, double *tau, double *work);
void orc_z_lq_factor_unblocked(size_t m, size_t n, orc_z *a, size_t lda, orc_z *tau, orc_z *work);
void orc_d_lq_factor(size_t m, size_t n, double *a, size_t lda, double *tau, size_t *lwork, double *work);
void orc_z_lq_factor(size_t m, size_t n, orc_z *a, size_t lda, orc_z *tau, size_t *lwork, orc_z *work);
void orc_d_lq_multq(const char *side, const char *trans, size_t m, size_t n, size_t k, const double *a,
                    size_t lda, const double *tau, double *c, size_t ldc, size_t *lwork, double *work);
void orc_z_lq_multq(const char *side, const char *trans, size_t m, size_t n, size_t k, const orc_z *a,
                    size_t lda, const orc_z *tau, orc_z *c, size_t ldc, size_t *lwork, orc_z *work);
void orc_d_lq_generateq(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                        size_t *lwork, double *work);
void orc_z_lq_generateq(size_t m, size_t n, size_t k, orc_z *a, size_t lda, const orc_z *tau,
                        size_t *lwork, orc_z *work);

/* ---- QL (RNP/LA/QL.hpp) and RQ (RNP/LA/RQ.hpp) ----
 *   QL::Factor_unblocked  QL.hpp:79-100,   QL::Factor    QL.hpp:126-205
 *   QL::MultQ_unblocked   QL.hpp:245-300,  QL::MultQ     QL.hpp:335-430
 *   QL::GenerateQ_unblocked QL.hpp:445-471, QL::GenerateQ QL.hpp:501-565
 *   RQ::Factor_unblocked  RQ.hpp:80-102,   RQ::Factor    RQ.hpp:131-205
 *   RQ::MultQ_unblocked   RQ.hpp:245-303,  RQ::MultQ     RQ.hpp:344-440
 *   RQ::GenerateQ_unblocked RQ.hpp:457-491, RQ::GenerateQ RQ.hpp:519-580
 *   Reflector::Apply vone == -1 branches, Reflector.hpp:297-357 */
void orc_d_ql_factor(size_t m, size_t n, double *a, size_t lda, double *tau, size_t *lwork, double *work);
void orc_d_ql_multq(const char *side, const char *trans, size_t m, size_t n, size_t k, const double *a,
                    size_t lda, const double *tau, double *c, size_t ldc, size_t *lwork, double *work);
void orc_d_ql_generateq(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                        size_t *lwork, double *work);
void orc_d_rq_factor(size_t m, size_t n, double *a, size_t lda, double *tau, size_t *lwork, double *work);
void orc_d_rq_multq(const char *side, const char *trans, size_t m, size_t n, size_t k, const double *a,
                    size_t lda, const double *tau, double *c, size_t ldc, size_t *lwork, double *work);
void orc_d_rq_generateq(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                        size_t *lwork, double *work);
void orc_z_ql_factor(size_t m, size_t n, orc_z *a, size_t lda, orc_z *tau, size_t *lwork, orc_z *work);
void orc_z_ql_multq(const char *side, const char *trans, size_t m, size_t n, size_t k, const orc_z *a,
                    size_t lda, const orc_z *tau, orc_z *c, size_t ldc, size_t *lwork, orc_z *work);
void orc_z_ql_generateq(size_t m, size_t n, size_t k, orc_z *a, size_t lda, const orc_z *tau,
                        size_t *lwork, orc_z *work);
void orc_z_rq_factor(size_t m, size_t n, orc_z *a, size_t lda, orc_z *tau, size_t *lwork, orc_z *work);
void orc_z_rq_multq(const char *side, const char *trans, size_t m, size_t n, size_t k, const orc_z *a,
                    size_t lda, const orc_z *tau, orc_z *c, size_t ldc, size_t *lwork, orc_z *work);
void orc_z_rq_generateq(size_t m, size_t n, size_t k, orc_z *a, size_t lda, const orc_z *tau,
                        size_t *lwork, orc_z *work);

/* LU::RefineSolution LU.hpp:216-366 (MatrixNorm1Estimate_unblocked,
 * MatrixNormEstimates.hpp:17-107, for the forward error). */
void orc_d_lu_refine(const char *trans, size_t n, size_t nrhs, const double *a, size_t lda, const double *af,
                     size_t ldaf, const size_t *ipiv, const double *b, size_t ldb, double *x, size_t ldx,
                     double *ferr, double *berr);
void orc_z_lu_refine(const char *trans, size_t n, size_t nrhs, const orc_z *a, size_t lda, const orc_z *af,
                     size_t ldaf, const size_t *ipiv, const orc_z *b, size_t ldb, orc_z *x, size_t ldx,
                     double *ferr, double *berr);

/* ---- Hessenberg (RNP/LA/Hessenberg.hpp) ----
 *   Hessenberg::MultQ  Hessenberg.hpp:478-514, Hessenberg::GenerateQ Hessenberg.hpp:540-588 */
void orc_d_hess_multq(const char *side, const char *trans, size_t m, size_t n, size_t ilo, size_t ihi,
                      const double *a, size_t lda, const double *tau, double *c, size_t ldc, size_t *lwork, double *work);
void orc_d_hess_generateq(size_t n, size_t ilo, size_t ihi, double *a, size_t lda, const double *tau,
                          size_t *lwork, double *work);
void orc_z_hess_multq(const char *side, const char *trans, size_t m, size_t n, size_t ilo, size_t ihi,
                      const orc_z *a, size_t lda, const orc_z *tau, orc_z *c, size_t ldc, size_t *lwork, orc_z *work);
void orc_z_hess_generateq(size_t n, size_t ilo, size_t ihi, orc_z *a, size_t lda, const orc_z *tau,
                          size_t *lwork, orc_z *work);

/* ---- LU (RNP/LA/LU.hpp) ----
 *   LU::Factor_unblocked LU.hpp:52-75, LU::Factor LU.hpp:77-115, LU::Solve LU.hpp:117-144
 *   BLAS::MaximumIndex BLAS_mix.hpp:798-810, ApplyPermutations Permutation.hpp:33-52 */
int orc_d_lu_factor_unblocked(size_t m, size_t n, double *a, size_t lda, size_t *piv);
int orc_d_lu_factor(size_t m, size_t n, double *a, size_t lda, size_t *piv);
int orc_d_lu_invert(size_t n, double *a, size_t lda, const size_t *ipiv, double *work);
void orc_d_lu_solve(const char *trans, size_t n, size_t nrhs, const double *a, size_t lda, const size_t *ipiv,
                    double *b, size_t ldb);
int orc_z_lu_factor_unblocked(size_t m, size_t n, orc_z *a, size_t lda, size_t *piv);
int orc_z_lu_factor(size_t m, size_t n, orc_z *a, size_t lda, size_t *piv);
int orc_z_lu_invert(size_t n, orc_z *a, size_t lda, const size_t *ipiv, orc_z *work);
void orc_z_lu_solve(const char *trans, size_t n, size_t nrhs, const orc_z *a, size_t lda, const size_t *ipiv,
                    orc_z *b, size_t ldb);

/* Least squares min ||A x - B|| for m >= n by Factor, MultQ("L","C") and an
 * upper-triangular back substitution with R (BLAS dtrsm "L","U","N","N";
 * the composition of QR.hpp:129, QR.hpp:337 and BLAS_mix.hpp:2096).
 * A is overwritten by its factorization, B (m-by-nrhs) by Q^H B with the
 * solution in its first n rows. */
void orc_d_ls_solve(size_t m, size_t n, size_t nrhs, double *a, size_t lda, double *b, size_t ldb);
void orc_z_ls_solve(size_t m, size_t n, size_t nrhs, orc_z *a, size_t lda, orc_z *b, size_t ldb);

/* ---- BLAS restatement (column-major, Fortran BLAS semantics) ---- */
void orc_d_gemm(const char *transa, const char *transb, size_t m, size_t n, size_t k,
                double alpha, const double *a, size_t lda, const double *b, size_t ldb,
                double beta, double *c, size_t ldc);
void orc_z_gemm(const char *transa, const char *transb, size_t m, size_t n, size_t k,
                orc_z alpha, const orc_z *a, size_t lda, const orc_z *b, size_t ldb,
                orc_z beta, orc_z *c, size_t ldc);
void orc_set_threads(int nthreads);
int orc_get_threads(void);

/* ---- harness: RNP's generators ---- */
enum { ORC_UNIFORM01 = 0, ORC_UNIFORM_11 = 1, ORC_NORMAL01 = 2, ORC_UNITDISC = 3, ORC_UNITCIRCLE = 4 };
/* iseed == NULL uses the internal static seed {2398,691,782,721} (Random.hpp:126). */
void orc_uniform_real_vector(size_t n, double *x, int *iseed);
void orc_reset_internal_seed(void);
void orc_d_generate_vector(int dist, size_t n, double *x, int *iseed);
void orc_z_generate_vector(int dist, size_t n, orc_z *x, int *iseed);

enum { ORC_DIAG_GIVEN = 0, ORC_DIAG_FIRST1_REMAINING_COND = 1, ORC_DIAG_LASTCOND_REMAINING1 = 2,
       ORC_DIAG_GRADED_EXPONENTIALLY = 3, ORC_DIAG_GRADED_LINEARLY = 4,
       ORC_DIAG_RANDOM_EXPONENTIAL = 5, ORC_DIAG_RANDOM_DIST = 6 };
void orc_generate_diagonal(int dist, int mode, int reverse, int randsign, double cond,
                           double dscale, size_t n, double *d, int *iseed);
void orc_d_random_matrix(int dist, int mode, int reverse, double cond, double dmax,
                         size_t m, size_t n, double *a, size_t lda, int *iseed);
void orc_z_random_matrix(int dist, int mode, int reverse, double cond, double dmax,
                         size_t m, size_t n, orc_z *a, size_t lda, int *iseed);

#ifdef __cplusplus
}
#endif
#endif
