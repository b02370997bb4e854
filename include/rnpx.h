/*
 * rnpx.h -- C ABI of the MI355X (gfx950) blocked Householder QR that drops
 * in for RNP2's LA::QR path.
 *
 * Conventions are RNP's (RNP/LA/QR.hpp): column-major, 0-based, size_t
 * sizes, complex numbers interleaved like std::complex<double>, side/trans
 * read from the first character ("L"/"R", "N"/"C").  Every function returns
 * an int status: 0 on success, -i when argument i (1-based) is invalid
 * (LAPACK style), or one of the RNPX_ERR_* codes below.
 *
 * Two kinds of entry point:
 *  - drop-in ones (rnpx_dgeqrf, ...) take RNP's exact argument list.  The
 *    matrix pointers may be host memory (staged to the GPU and back inside
 *    the call) or device memory (used in place; the call still returns only
 *    when the result is in place).  The workspace query (work == NULL or
 *    *lwork == 0) returns the same size the reference returns.  work itself
 *    is not touched: the GPU keeps its own workspace.
 *  - handle ones (rnpx_*_dev, rnpx_ls_*, batched, serving) take device
 *    pointers, run on the handle's HIP stream and return without waiting.
 */
#ifndef RNPX_H
#define RNPX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rnpx_zcomplex {
	double re, im;
} rnpx_zcomplex;

typedef struct rnpx_handle_s *rnpx_handle;

enum {
	RNPX_OK = 0,
	RNPX_ERR_HIP = -1000,       /* a HIP runtime call failed */
	RNPX_ERR_NOGPU = -1001,     /* no usable gfx950 device */
	RNPX_ERR_NOMEM = -1002,     /* device allocation failed / over budget */
	RNPX_ERR_SYNC = -1003,      /* a grid barrier timed out (results invalid) */
	RNPX_ERR_UNSUPPORTED = -1004, /* shape/option outside what is built */
	RNPX_ERR_NOTCONV = -1005,   /* iterative refinement did not reach its target */
	RNPX_ERR_ARG = -1006        /* an invalid argument where -i would read as a result (rnpx_*getri) */
};

/* One problem of a batched call: A is m-by-n at A + a_off (leading dimension
 * lda), tau has min(m,n) entries at tau + tau_off; with nrhs > 0 (and
 * m >= n) the least-squares solution of min ||A x - b|| for the m-by-nrhs
 * right-hand sides at B + b_off (leading dimension ldb) overwrites their
 * first n rows (Q^H b below them).  Offsets count elements. */
typedef struct rnpx_qr_desc {
	int m, n, lda, ldb, nrhs, pad;
	long long a_off, tau_off, b_off;
} rnpx_qr_desc;

typedef struct rnpx_batched_plan_s *rnpx_batched_plan;

/* Tuning hook: the counterpart of specialising QR::Tuning<T>
 * (RNP/LA/QR.hpp:41-51).  Zero fields keep the built-in choice. */
typedef struct rnpx_tuning {
	int nb;           /* outer block width: 32, 64, 128, 256 (0 = the autotuned
	                     table's entry for the shape, else 128 with 64-column panels) */
	int panel_rpt;    /* panel rows per thread: 1, 2, 4 */
	int ksplit;       /* split-K slices for W = V^H C (0 = automatic) */
	int lookahead;    /* 1 = overlap the next panel with the trailing update */
	int panel;        /* real nb = 64 panels: 0 automatic, 1 column by column only,
	                     2 CholeskyQR2 + Householder reconstruction where it applies */
	int panel_cus;    /* look-ahead: CUs reserved for the panel stream, the trailing
	                     update runs on the others (0 automatic, -1 no partition) */
} rnpx_tuning;

/* ---------------------------------------------------------------------
 * Drop-in entry points (host or device pointers).
 * ------------------------------------------------------------------- */

/* QR::Factor<double> -- RNP/LA/QR.hpp:129-197 (dgeqrf). */
int rnpx_dgeqrf(size_t m, size_t n, double *a, size_t lda, double *tau,
                size_t *lwork, double *work);
/* QR::Factor<std::complex<double>> -- RNP/LA/QR.hpp:129-197 (zgeqrf). */
int rnpx_zgeqrf(size_t m, size_t n, rnpx_zcomplex *a, size_t lda, rnpx_zcomplex *tau,
                size_t *lwork, rnpx_zcomplex *work);

/* Reflector::LastNonzeroColumnLimit / LastNonzeroRowLimit -- RNP/LA/Reflector.hpp:
 * 36-75: one plus the index of the last column (row) of the m x n matrix a (host or
 * device memory, lda >= m) that holds an entry other than zero (NaN counts, -0 does
 * not), 0 when there is none.  The limits ApplyBlock and Apply trim to. */
int rnpx_dlast_nonzero_column_limit(size_t m, size_t n, const double *a, size_t lda,
                                    size_t *limit);
int rnpx_zlast_nonzero_column_limit(size_t m, size_t n, const rnpx_zcomplex *a, size_t lda,
                                    size_t *limit);
int rnpx_dlast_nonzero_row_limit(size_t m, size_t n, const double *a, size_t lda,
                                 size_t *limit);
int rnpx_zlast_nonzero_row_limit(size_t m, size_t n, const rnpx_zcomplex *a, size_t lda,
                                 size_t *limit);

/* QR::MultQ<double> -- RNP/LA/QR.hpp:337-429 (dormqr). */
int rnpx_dormqr(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                size_t *lwork, double *work);
/* QR::MultQ<std::complex<double>> -- RNP/LA/QR.hpp:337-429 (zunmqr). */
int rnpx_zunmqr(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const rnpx_zcomplex *a, size_t lda, const rnpx_zcomplex *tau,
                rnpx_zcomplex *c, size_t ldc, size_t *lwork, rnpx_zcomplex *work);

/* QR::GenerateQ<double> -- RNP/LA/QR.hpp:509-570 (dorgqr). */
int rnpx_dorgqr(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                size_t *lwork, double *work);
/* QR::GenerateQ<std::complex<double>> -- RNP/LA/QR.hpp:509-570 (zungqr). */
int rnpx_zungqr(size_t m, size_t n, size_t k, rnpx_zcomplex *a, size_t lda,
                const rnpx_zcomplex *tau, size_t *lwork, rnpx_zcomplex *work);

/* LQ (RNP/LA/LQ.hpp): A = L Q with the reflectors conjugated in the rows
 * above L.  Same argument lists, conventions and workspace queries as the
 * reference: LQ::Factor nb*n (LQ.hpp:135-138), LQ::MultQ nb*nw + nb*nb
 * (LQ.hpp:329-333), LQ::GenerateQ n*nb (LQ.hpp:500-504), nb = 32. */
/* LQ::Factor<double> -- RNP/LA/LQ.hpp:126-190 (dgelqf). */
int rnpx_dgelqf(size_t m, size_t n, double *a, size_t lda, double *tau,
                size_t *lwork, double *work);
/* LQ::Factor<std::complex<double>> -- RNP/LA/LQ.hpp:126-190 (zgelqf). */
int rnpx_zgelqf(size_t m, size_t n, rnpx_zcomplex *a, size_t lda, rnpx_zcomplex *tau,
                size_t *lwork, rnpx_zcomplex *work);
/* LQ::MultQ<double> -- RNP/LA/LQ.hpp:309-405 (dormlq): c <- op(Q) c or c op(Q),
 * the product LQ.hpp:195-196 documents. */
int rnpx_dormlq(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                size_t *lwork, double *work);
/* LQ::MultQ<std::complex<double>> -- RNP/LA/LQ.hpp:309-405 (zunmlq). */
int rnpx_zunmlq(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const rnpx_zcomplex *a, size_t lda, const rnpx_zcomplex *tau,
                rnpx_zcomplex *c, size_t ldc, size_t *lwork, rnpx_zcomplex *work);
/* LQ::GenerateQ<double> -- RNP/LA/LQ.hpp:491-559 (dorglq): the first m rows of Q. */
int rnpx_dorglq(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                size_t *lwork, double *work);
/* LQ::GenerateQ<std::complex<double>> -- RNP/LA/LQ.hpp:491-559 (zunglq). */
int rnpx_zunglq(size_t m, size_t n, size_t k, rnpx_zcomplex *a, size_t lda,
                const rnpx_zcomplex *tau, size_t *lwork, rnpx_zcomplex *work);

/* QL (RNP/LA/QL.hpp): A = Q L, L ending at the bottom-right corner, the
 * reflectors in the columns above it, Q = H[k-1] ... H[0].  RQ
 * (RNP/LA/RQ.hpp): A = R Q, R ending at the bottom-right corner, the
 * reflectors conjugated in the rows left of it.  Same argument lists,
 * conventions and workspace queries as the reference: QL::Factor nb*n
 * (QL.hpp:139-141), RQ::Factor nb*m (RQ.hpp:142-144; lwork 0 or >= m),
 * MultQ nb*nw + nb*nb (QL.hpp:351-354, RQ.hpp:359-362), QL::GenerateQ n*nb
 * (QL.hpp:509-512), RQ::GenerateQ m*nb (RQ.hpp:528-531), nb = 32. */
/* QL::Factor<double> -- RNP/LA/QL.hpp:126-205 (dgeqlf). */
int rnpx_dgeqlf(size_t m, size_t n, double *a, size_t lda, double *tau,
                size_t *lwork, double *work);
/* QL::Factor<std::complex<double>> -- RNP/LA/QL.hpp:126-205 (zgeqlf). */
int rnpx_zgeqlf(size_t m, size_t n, rnpx_zcomplex *a, size_t lda, rnpx_zcomplex *tau,
                size_t *lwork, rnpx_zcomplex *work);
/* QL::MultQ<double> -- RNP/LA/QL.hpp:335-430 (dormql): c <- op(Q) c or c op(Q);
 * a holds the k reflector columns (lda >= the order of Q). */
int rnpx_dormql(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                size_t *lwork, double *work);
/* QL::MultQ<std::complex<double>> -- RNP/LA/QL.hpp:335-430 (zunmql). */
int rnpx_zunmql(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const rnpx_zcomplex *a, size_t lda, const rnpx_zcomplex *tau,
                rnpx_zcomplex *c, size_t ldc, size_t *lwork, rnpx_zcomplex *work);
/* QL::GenerateQ<double> -- RNP/LA/QL.hpp:501-565 (dorgql): the last n columns of Q. */
int rnpx_dorgql(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                size_t *lwork, double *work);
/* QL::GenerateQ<std::complex<double>> -- RNP/LA/QL.hpp:501-565 (zungql). */
int rnpx_zungql(size_t m, size_t n, size_t k, rnpx_zcomplex *a, size_t lda,
                const rnpx_zcomplex *tau, size_t *lwork, rnpx_zcomplex *work);
/* RQ::Factor<double> -- RNP/LA/RQ.hpp:131-205 (dgerqf). */
int rnpx_dgerqf(size_t m, size_t n, double *a, size_t lda, double *tau,
                size_t *lwork, double *work);
/* RQ::Factor<std::complex<double>> -- RNP/LA/RQ.hpp:131-205 (zgerqf). */
int rnpx_zgerqf(size_t m, size_t n, rnpx_zcomplex *a, size_t lda, rnpx_zcomplex *tau,
                size_t *lwork, rnpx_zcomplex *work);
/* RQ::MultQ<double> -- RNP/LA/RQ.hpp:344-440 (dormrq): c <- op(Q) c or c op(Q);
 * a holds the k reflector rows. */
int rnpx_dormrq(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                size_t *lwork, double *work);
/* RQ::MultQ<std::complex<double>> -- RNP/LA/RQ.hpp:344-440 (zunmrq). */
int rnpx_zunmrq(const char *side, const char *trans, size_t m, size_t n, size_t k,
                const rnpx_zcomplex *a, size_t lda, const rnpx_zcomplex *tau,
                rnpx_zcomplex *c, size_t ldc, size_t *lwork, rnpx_zcomplex *work);
/* RQ::GenerateQ<double> -- RNP/LA/RQ.hpp:519-580 (dorgrq): the last m rows of Q. */
int rnpx_dorgrq(size_t m, size_t n, size_t k, double *a, size_t lda, const double *tau,
                size_t *lwork, double *work);
/* RQ::GenerateQ<std::complex<double>> -- RNP/LA/RQ.hpp:519-580 (zungrq). */
int rnpx_zungrq(size_t m, size_t n, size_t k, rnpx_zcomplex *a, size_t lda,
                const rnpx_zcomplex *tau, size_t *lwork, rnpx_zcomplex *work);

/* Hessenberg (RNP/LA/Hessenberg.hpp): the Q of a Hessenberg reduction
 * (Reduce's reflectors one row below the diagonal, 0-based ilo, exclusive
 * ihi) through the QR routines.  Checks and workspace queries as the
 * reference: MultQ passes QR::MultQ's (nb*nw + nb*nb, none when
 * ihi-ilo-1 = 0), GenerateQ nh*nb with nh = ihi-ilo-1 (Hessenberg.hpp:553-556). */
/* Hessenberg::MultQ<double> -- RNP/LA/Hessenberg.hpp:478-514 (dormhr). */
int rnpx_dormhr(const char *side, const char *trans, size_t m, size_t n, size_t ilo, size_t ihi,
                const double *a, size_t lda, const double *tau, double *c, size_t ldc,
                size_t *lwork, double *work);
/* Hessenberg::MultQ<std::complex<double>> -- RNP/LA/Hessenberg.hpp:478-514 (zunmhr). */
int rnpx_zunmhr(const char *side, const char *trans, size_t m, size_t n, size_t ilo, size_t ihi,
                const rnpx_zcomplex *a, size_t lda, const rnpx_zcomplex *tau, rnpx_zcomplex *c,
                size_t ldc, size_t *lwork, rnpx_zcomplex *work);
/* Hessenberg::GenerateQ<double> -- RNP/LA/Hessenberg.hpp:540-588 (dorghr): the n x n Q. */
int rnpx_dorghr(size_t n, size_t ilo, size_t ihi, double *a, size_t lda, const double *tau,
                size_t *lwork, double *work);
/* Hessenberg::GenerateQ<std::complex<double>> -- RNP/LA/Hessenberg.hpp:540-588 (zunghr). */
int rnpx_zunghr(size_t n, size_t ilo, size_t ihi, rnpx_zcomplex *a, size_t lda,
                const rnpx_zcomplex *tau, size_t *lwork, rnpx_zcomplex *work);

/* LU (RNP/LA/LU.hpp): partial pivoting, pivots size_t 0-based absolute rows
 * (row i was interchanged with row pivots[i]).  Factor returns the
 * reference's value (>= 0; LU.hpp:63-74 and 112-113: the last exactly-zero
 * pivot column of the first panel that has one, with the reference's
 * offsets) or a negative status. */
/* LU::Factor<double> -- RNP/LA/LU.hpp:77-115 (dgetrf). */
int rnpx_dgetrf(size_t m, size_t n, double *a, size_t lda, size_t *pivots);
/* LU::Factor<std::complex<double>> -- RNP/LA/LU.hpp:77-115 (zgetrf). */
int rnpx_zgetrf(size_t m, size_t n, rnpx_zcomplex *a, size_t lda, size_t *pivots);
/* LU::Solve<double> -- RNP/LA/LU.hpp:117-144 (dgetrs): op(A) X = B from the factors. */
int rnpx_dgetrs(const char *trans, size_t n, size_t nrhs, const double *a, size_t lda,
                const size_t *ipiv, double *b, size_t ldb);
/* LU::Solve<std::complex<double>> -- RNP/LA/LU.hpp:117-144 (zgetrs). */
int rnpx_zgetrs(const char *trans, size_t n, size_t nrhs, const rnpx_zcomplex *a, size_t lda,
                const size_t *ipiv, rnpx_zcomplex *b, size_t ldb);
/* LU::Invert<double> -- RNP/LA/LU.hpp:146-214 (dgetri): inv(A) in place from
 * its LU factors and pivots.  Returns the reference's value: 0, or -(i+1)
 * for the first exactly-zero U_ii (a left unchanged), or RNPX_ERR_* (an
 * invalid argument is RNPX_ERR_ARG, not -i, which would read as a zero U_ii).
 * Workspace query as the reference's (work == NULL or *lwork == 0 writes
 * 64 n); work is not used.  The reference's blocked branch (n > 64 with a
 * full workspace) mis-indexes L and returns a wrong matrix; this computes the
 * inverse both of its branches intend (its unblocked branch's). */
int rnpx_dgetri(size_t n, double *a, size_t lda, const size_t *ipiv, size_t *lwork, double *work);
/* LU::Invert<std::complex<double>> -- RNP/LA/LU.hpp:146-214 (zgetri). */
int rnpx_zgetri(size_t n, rnpx_zcomplex *a, size_t lda, const size_t *ipiv, size_t *lwork,
                rnpx_zcomplex *work);
/* LU::RefineSolution<double> -- RNP/LA/LU.hpp:216-366 (dgerfs): refines x
 * (n x nrhs) for op(A) x = b from the factors af/ipiv, and returns the
 * componentwise backward error berr and the estimated forward error ferr
 * per right-hand side.  work and rwork are accepted and not used. */
int rnpx_dgerfs(const char *trans, size_t n, size_t nrhs, const double *a, size_t lda,
                const double *af, size_t ldaf, const size_t *ipiv, const double *b, size_t ldb,
                double *x, size_t ldx, double *ferr, double *berr, double *work, double *rwork);
/* LU::RefineSolution<std::complex<double>> -- RNP/LA/LU.hpp:216-366 (zgerfs). */
int rnpx_zgerfs(const char *trans, size_t n, size_t nrhs, const rnpx_zcomplex *a, size_t lda,
                const rnpx_zcomplex *af, size_t ldaf, const size_t *ipiv, const rnpx_zcomplex *b,
                size_t ldb, rnpx_zcomplex *x, size_t ldx, double *ferr, double *berr,
                rnpx_zcomplex *work, double *rwork);

/* ---------------------------------------------------------------------
 * Handles (device-resident work on one HIP stream).
 * ------------------------------------------------------------------- */

/* Create a handle on the current device with its own stream.
 * hbm_budget_bytes bounds the handle's own workspace (0 = unbounded). */
int rnpx_create(rnpx_handle *h, size_t hbm_budget_bytes);
int rnpx_destroy(rnpx_handle h);
/* Use an existing hipStream_t (NULL = the handle's own stream). */
int rnpx_set_stream(rnpx_handle h, void *hip_stream);
void *rnpx_get_stream(rnpx_handle h);
int rnpx_set_tuning(rnpx_handle h, const rnpx_tuning *t);
int rnpx_get_tuning(rnpx_handle h, rnpx_tuning *t);

/* The autotuning table behind the automatic tuning (the device counterpart
 * of RNP's QR::Tuning hooks, RNP/LA/QR.hpp:41-51, which fix nb = 32 for
 * every shape).  rnp2_amd/autotune.py times candidate configurations on the
 * device and writes it as JSON; it is read once per process on first use,
 * from RNPX_TUNING_FILE or rnpx_tuning.json beside librnpx.so.  A
 * factorization whose handle tuning has nb = 0 takes the entry nearest its
 * shape (element type equal, m and n within a factor of two): the outer
 * block width (32, 64 or 128) and the W = V^H C split.  Batched calls take
 * the small tier's LDS per team and the blocked-path column threshold. */
/* Replace the table from path (NULL: the default location, where no file
 * means an empty table); returns the entry count, RNPX_ERR_UNSUPPORTED for an
 * unreadable path, RNPX_ERR_ARG for a malformed file (the previous table is
 * kept in both cases). */
int rnpx_tuning_load(const char *path);
/* 1 and the entry's nb and W split when the table has an entry for an
 * m x n factorization of element type dtype ('d' or 'z'), else 0. */
int rnpx_tuning_lookup_factor(char dtype, size_t m, size_t n, int *nb, int *ksplit);
/* 1 and the entry's panel_cus (CUs reserved for the look-ahead panel stream;
 * 0 when the entry gives none) for an m x n factorization of element type
 * dtype, else 0.  No counterpart in the reference (its Tuning hooks,
 * RNP/LA/QR.hpp:41-51, have no device to partition). */
int rnpx_tuning_lookup_panel_cus(char dtype, size_t m, size_t n, int *panel_cus);
/* 1 and the batched parameters for element type dtype, else 0. */
int rnpx_tuning_lookup_batched(char dtype, int *lds_bytes, int *blocked_min);
/* Wait for the handle's stream; reports RNPX_ERR_SYNC if a panel barrier
 * timed out in any call since the last check. */
int rnpx_synchronize(rnpx_handle h);
/* Bytes of device workspace the handle currently holds. */
size_t rnpx_workspace_bytes(rnpx_handle h);

/* Live kernel timing on the handle's streams (HIP events around each launch
 * of a kernel family; the FP64 trailing update's launch records its own
 * start and stop events, so its time is the kernel's alone).
 * kernel: 0 trailing update C -= V Z, 1 W = V^H C,
 * 2 panel, 3 T factor.  on: 0 off, 1 every family, otherwise bit (k + 1)
 * of on selects family k (fewer events perturb the timed work less).
 * Enabling resets the records.  flops/bytes are the algorithmic counts of
 * the recorded launches. */
int rnpx_profile_enable(rnpx_handle h, int on);
int rnpx_profile_read(rnpx_handle h, int kernel, long *launches, double *ms, double *flops,
                      double *bytes);

/* Device-pointer forms of Factor / MultQ / GenerateQ (same semantics). */
/* Factorizations on h (NULL: the drop-in calls' handle) whose matrix held Inf or NaN:
 * those follow the reference's own blocking (nb 32, crossover 128, RNP/LA/QR.hpp:
 * 141-196) and every LastNonzero* trim it makes, since which 0 * Inf products the
 * reference forms depends on them; finite matrices take the blocked fast path. */
int rnpx_exact_path_count(rnpx_handle h, long *count);

int rnpx_dgeqrf_dev(rnpx_handle h, size_t m, size_t n, double *a, size_t lda, double *tau);
int rnpx_zgeqrf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex *a, size_t lda,
                    rnpx_zcomplex *tau);
/* LQ::Factor on device pointers, asynchronous (RNP/LA/LQ.hpp:126-190). */
int rnpx_dgelqf_dev(rnpx_handle h, size_t m, size_t n, double *a, size_t lda, double *tau);
int rnpx_zgelqf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex *a, size_t lda,
                    rnpx_zcomplex *tau);
/* QL::Factor and RQ::Factor on device pointers, asynchronous
 * (RNP/LA/QL.hpp:126-205, RNP/LA/RQ.hpp:131-205). */
int rnpx_dgeqlf_dev(rnpx_handle h, size_t m, size_t n, double *a, size_t lda, double *tau);
int rnpx_zgeqlf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex *a, size_t lda,
                    rnpx_zcomplex *tau);
int rnpx_dgerqf_dev(rnpx_handle h, size_t m, size_t n, double *a, size_t lda, double *tau);
int rnpx_zgerqf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex *a, size_t lda,
                    rnpx_zcomplex *tau);
/* Hessenberg::GenerateQ on device pointers, asynchronous (RNP/LA/Hessenberg.hpp:540-588). */
int rnpx_dorghr_dev(rnpx_handle h, size_t n, size_t ilo, size_t ihi, double *a, size_t lda,
                    const double *tau);
int rnpx_zunghr_dev(rnpx_handle h, size_t n, size_t ilo, size_t ihi, rnpx_zcomplex *a, size_t lda,
                    const rnpx_zcomplex *tau);
int rnpx_dormqr_dev(rnpx_handle h, const char *side, const char *trans, size_t m, size_t n,
                    size_t k, const double *a, size_t lda, const double *tau, double *c,
                    size_t ldc);
int rnpx_zunmqr_dev(rnpx_handle h, const char *side, const char *trans, size_t m, size_t n,
                    size_t k, const rnpx_zcomplex *a, size_t lda, const rnpx_zcomplex *tau,
                    rnpx_zcomplex *c, size_t ldc);
int rnpx_dorgqr_dev(rnpx_handle h, size_t m, size_t n, size_t k, double *a, size_t lda,
                    const double *tau);
int rnpx_zungqr_dev(rnpx_handle h, size_t m, size_t n, size_t k, rnpx_zcomplex *a, size_t lda,
                    const rnpx_zcomplex *tau);

/* Batched QR / least squares (no counterpart in the reference: RNP factors
 * one matrix per call).  desc is a host array of count descriptors; A, tau,
 * B are device buffers.  Problems are cost-sorted and run by persistent
 * teams in one launch per tier; asynchronous on the handle's stream. */
int rnpx_dgeqrf_batched(rnpx_handle h, int count, const rnpx_qr_desc *desc, double *A, double *tau,
                        double *B);
int rnpx_zgeqrf_batched(rnpx_handle h, int count, const rnpx_qr_desc *desc, rnpx_zcomplex *A,
                        rnpx_zcomplex *tau, rnpx_zcomplex *B);
/* The same with the descriptor upload done once and the launches captured
 * in a hipGraph bound to A/tau/B; run replays it on the handle's stream. */
int rnpx_dgeqrf_batched_plan(rnpx_handle h, int count, const rnpx_qr_desc *desc, double *A,
                             double *tau, double *B, rnpx_batched_plan *plan);
int rnpx_zgeqrf_batched_plan(rnpx_handle h, int count, const rnpx_qr_desc *desc,
                             rnpx_zcomplex *A, rnpx_zcomplex *tau, rnpx_zcomplex *B,
                             rnpx_batched_plan *plan);
int rnpx_batched_plan_run(rnpx_batched_plan plan);
int rnpx_batched_plan_destroy(rnpx_batched_plan plan);
/* rnpx_dgeqrf_batched without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as rnpx_dgeqrf_batched; both tiers run on the handle's stream.  The
 * descriptor table, problem order and queue counters live in a static device
 * array sized for at most 16384 problems: a larger count returns
 * RNPX_ERR_UNSUPPORTED before anything is touched.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (different streams or
 * threads) race; calls on one stream are ordered by it. */
int rnpx_dgeqrf_batched_noalloc(rnpx_handle h, int count, const rnpx_qr_desc *desc, double *A,
                                double *tau, double *B);

/* Least squares min ||A X - B||, m >= n, device pointers: Factor, then
 * MultQ("L","C") on B, then back substitution with R (the composition of
 * RNP/LA/QR.hpp:129, QR.hpp:337 and SolveTrM, RNP/BLAS/BLAS_mix.hpp:2096).
 * A is overwritten by its factorization (tau, length n), B (m-by-nrhs) by
 * Q^H B with X in its first n rows. */
int rnpx_dgels_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, double *a, size_t lda,
                   double *tau, double *b, size_t ldb);
int rnpx_zgels_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, rnpx_zcomplex *a, size_t lda,
                   rnpx_zcomplex *tau, rnpx_zcomplex *b, size_t ldb);

/* Mixed-precision least squares (no counterpart in the reference): QR with
 * the trailing-update products on FP32 MFMA (reflector generation, T,
 * residuals, triangular solves and the update of x in FP64), refined by
 * corrected semi-normal equations in FP64 until the correction reaches
 * working precision (||d|| <= 2^-50 ||x||), stops halving, or max_iter
 * corrections.  berr receives the normwise backward error estimate
 * eta = ||A^T r|| / (||A||_F (||A||_F ||x|| + ||b||)) of the returned x;
 * RNPX_ERR_NOTCONV when eta > tol.  A and B are left intact; X is
 * n-by-nrhs.  Real only. */
int rnpx_dgels_mixed_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, const double *a,
                         size_t lda, const double *b, size_t ldb, double *x, size_t ldx,
                         int max_iter, double tol, int *iters, double *berr);
/* eta (as above) of a solution x of min ||A x - b||, worst over the nrhs. */
int rnpx_dls_berr_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, const double *a, size_t lda,
                      const double *b, size_t ldb, const double *x, size_t ldx, double *berr);

/* ---------------------------------------------------------------------
 * Least-squares request server (no counterpart in the reference).
 * Requests are submitted from the host and solved on the device by worker
 * threads (one HIP stream each): small requests (m <= 256) are batched into
 * one persistent launch, larger ones run concurrently; admission is gated by
 * the HBM budget.  mode 0: FP64; mode 1: mixed precision with FP64
 * refinement and FP64 fallback when refinement does not reach tol.
 * ------------------------------------------------------------------- */
typedef struct rnpx_server_s *rnpx_server;
typedef struct rnpx_server_config {
	size_t hbm_budget;  /* bytes for in-flight requests (0 = unbounded) */
	int workers;        /* concurrent requests / batches on the device, 1..16 */
	int mode;           /* 0 FP64, 1 mixed + refinement (FP64 fallback) */
	int max_iter;       /* refinement corrections (0 = 30) */
	double tol;         /* backward-error target of the mixed path (0 = 1e-13) */
} rnpx_server_config;
typedef struct rnpx_request_info {
	int status;   /* RNPX_OK or an error code */
	int path;     /* 0 FP64 dense, 1 mixed, 2 mixed fell back to FP64, 3 batched */
	int iters;    /* refinement corrections (mixed) */
	int worker;   /* worker that ran it */
	double berr;  /* eta of the solution (dense paths) */
	/* steady-clock seconds: submit called (before its copy of A and b),
	 * admitted to a worker, result ready */
	double t_submit, t_admit, t_done;
} rnpx_request_info;
typedef struct rnpx_server_stats {
	long long submitted, admitted, completed, rejected, failed, batches, queued;
	size_t bytes_in_use;
} rnpx_server_stats;
int rnpx_server_create(const rnpx_server_config *cfg, rnpx_server *server);
/* Copies A (m-by-n, lda) and B (m-by-nrhs, ldb); m >= n. */
int rnpx_server_submit(rnpx_server s, size_t m, size_t n, size_t nrhs, const double *a, size_t lda,
                       const double *b, size_t ldb, long long *id);
/* 0: done (X n-by-nrhs written when status is OK), 1: pending (wait == 0). */
int rnpx_server_poll(rnpx_server s, long long id, int wait, double *x, size_t ldx,
                     rnpx_request_info *info);
int rnpx_server_release(rnpx_server s, long long id);
int rnpx_server_stats_get(rnpx_server s, rnpx_server_stats *stats);
/* Waits for every submitted request, then stops the workers. */
int rnpx_server_destroy(rnpx_server s);

/* Building blocks, device pointers (the reference's Reflector functions):
 * T factor of k <= nb forward columnwise reflectors stored in a (n-by-k,
 * unit diagonal implicit) -- Reflector::GenerateBlockTr("F","C"),
 * RNP/LA/Reflector.hpp:573-626.  t is k-by-k with leading dimension ldt. */
int rnpx_dlarft_dev(rnpx_handle h, size_t n, size_t k, const double *v, size_t ldv,
                    const double *tau, double *t, size_t ldt);
int rnpx_zlarft_dev(rnpx_handle h, size_t n, size_t k, const rnpx_zcomplex *v, size_t ldv,
                    const rnpx_zcomplex *tau, rnpx_zcomplex *t, size_t ldt);
/* Reflector::ApplyBlock(side, trans, "F", "C", ...) -- Reflector.hpp:725-815. */
int rnpx_dlarfb_dev(rnpx_handle h, const char *side, const char *trans, size_t m, size_t n,
                    size_t k, const double *v, size_t ldv, const double *t, size_t ldt,
                    double *c, size_t ldc);
int rnpx_zlarfb_dev(rnpx_handle h, const char *side, const char *trans, size_t m, size_t n,
                    size_t k, const rnpx_zcomplex *v, size_t ldv, const rnpx_zcomplex *t,
                    size_t ldt, rnpx_zcomplex *c, size_t ldc);

/* C = alpha op(A) op(B) + beta C on FP64 MFMA, op in "N","T","C" (the
 * MultMM the reference forwards to xgemm, src/BLAS_mix.cpp:498-517).
 * With op(B) = "N" and 1 <= n <= 4 (matrix-vector shapes) the product runs
 * on an HBM-bound kernel that reads A once; the reduction is split and
 * added in a fixed order either way. */
int rnpx_dgemm_dev(rnpx_handle h, const char *transa, const char *transb, size_t m, size_t n,
                   size_t k, double alpha, const double *a, size_t lda, const double *b,
                   size_t ldb, double beta, double *c, size_t ldc);
int rnpx_zgemm_dev(rnpx_handle h, const char *transa, const char *transb, size_t m, size_t n,
                   size_t k, rnpx_zcomplex alpha, const rnpx_zcomplex *a, size_t lda,
                   const rnpx_zcomplex *b, size_t ldb, rnpx_zcomplex beta, rnpx_zcomplex *c,
                   size_t ldc);

/* RNP's random generator on the device (RNP/Random.hpp:45-320): fills
 * count values of the stream that starts at iseed (column-major fill of a
 * matrix = one stream).  dist: 0 Uniform01, 1 Uniform_11, 2 Normal01. */
int rnpx_dlarnv_dev(rnpx_handle h, int dist, const int iseed[4], size_t count, double *x);
int rnpx_zlarnv_dev(rnpx_handle h, int dist, const int iseed[4], size_t count, rnpx_zcomplex *x);

/* Library version string and the device it runs on. */
const char *rnpx_version(void);
int rnpx_device_name(char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
