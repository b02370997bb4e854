// rnpx_ls.cpp -- least-squares drivers and the batched form.
//
// Dense least squares min ||A X - B|| (m >= n) is the composition the
// reference leaves to its callers: QR::Factor (RNP/LA/QR.hpp:129),
// QR::MultQ("L","C") on B (QR.hpp:337) and back substitution with R
// (SolveTrM "L","U","N","N", RNP/BLAS/BLAS_mix.hpp:2096).
//
// Batched: descriptors are validated and cost-sorted on the host (largest
// first, ties by index), split into the LDS tier and the in-place tier, and
// run by persistent teams pulling from an atomic counter.  A plan uploads the
// table once and captures both tier launches (fork/join over the handle's
// two streams) in a hipGraph that replays with one call.
#include "rnpx_internal.hpp"

#include <algorithm>
#include <functional>
#include <numeric>
#include <vector>

namespace rnpx {

// Triangular solve op(M) X = B in place on B, M the upper ('U') or lower
// ('L') triangle of R (n x n), op 'N', 'T' or 'C', unit diagonal for diag
// 'U'.  Each call takes a new flag epoch, so a solve is not replayable from
// a captured graph.
template <typename T>
int trsm_tri_dev(Handle* h, char uplo, char op, char diag, int n, int nrhs, const T* R, long ldr, T* B, long ldb) {
  if (n == 0 || nrhs == 0) return RNPX_OK;
  const size_t words = (size_t)trsm_flag_words(n, nrhs);
  const size_t before = h->trsm_flags.bytes;
  RNPX_TRY(h->trsm_flags.ensure(sizeof(unsigned) * words, h));
  if (h->trsm_flags.bytes != before || h->trsm_epoch == 0xFFFFFFFFu) {
    // new (uninitialised) or exhausted flags: clear them with a kernel
    RNPX_TRY(launch_zero_u32((unsigned*)h->trsm_flags.p, (int)(h->trsm_flags.bytes / sizeof(unsigned)), h->stream));
    h->trsm_epoch = 0;
  }
  ++h->trsm_epoch;
  RNPX_TRY(launch_trsm_tri<T>(uplo, op, diag, n, nrhs, R, ldr, B, ldb, (unsigned*)h->trsm_flags.p, h->trsm_epoch,
                              (unsigned*)h->flag.p, h->stream));
  return RNPX_OK;
}
template int trsm_tri_dev<double>(Handle*, char, char, char, int, int, const double*, long, double*, long);
template int trsm_tri_dev<zd>(Handle*, char, char, char, int, int, const zd*, long, zd*, long);

// With the upper triangle R of A: trans 'N' R X = B, 'C' R^H X = B.
template <typename T>
static int trsm_dev(Handle* h, char trans, int n, int nrhs, const T* R, long ldr, T* B, long ldb) {
  return trsm_tri_dev<T>(h, 'U', trans, 'N', n, nrhs, R, ldr, B, ldb);
}

template <typename T>
int gels_dev(Handle* h, int m, int n, int nrhs, T* A, long lda, T* tau, T* B, long ldb) {
  if (m < 0) return -2;
  if (n < 0 || n > m) return -3;
  if (nrhs < 0) return -4;
  if (lda < std::max(1, m)) return -6;
  if (ldb < std::max(1, m)) return -9;
  if (n == 0 || nrhs == 0) return RNPX_OK;
  // Q^H B is formed inside the factorization (each block reflector applied
  // to B as it is produced), then R X = (Q^H B)(0:n, :).
  int s = geqrf_dev<T>(h, m, n, A, lda, tau, B, nrhs, ldb);
  if (s) return s;
  return trsm_dev<T>(h, 'N', n, nrhs, A, lda, B, ldb);
}

template int gels_dev<double>(Handle*, int, int, int, double*, long, double*, double*, long);

// Normwise backward error of a least-squares solution, per right-hand side,
// worst over the columns:
//   eta = ||A^T r|| / (||A||_F (||A||_F ||x|| + ||b||)),  r = b - A x,
// all in FP64 on the device (scratch: the handle's lsbuf/lsnrm buffers).
struct LsScratch {
  double *R, *S, *nrm;  // m x nrhs residual, n x nrhs normal residual, norms
};

static int ls_scratch(Handle* h, int m, int n, int nrhs, LsScratch* sc) {
  RNPX_TRY(h->lsbuf.ensure(sizeof(double) * ((size_t)m * nrhs + (size_t)n * nrhs), h));
  RNPX_TRY(h->lsnrm.ensure(sizeof(double) * (size_t)(4 * nrhs + n + 64), h));
  sc->R = (double*)h->lsbuf.p;
  sc->S = sc->R + (size_t)m * nrhs;
  sc->nrm = (double*)h->lsnrm.p;
  return RNPX_OK;
}

// Y = alpha op(A) X + beta Y for the refinement's products: the few-column
// kernel (HBM-bound, every entry of A read once) up to four right-hand
// sides, the MFMA GEMM (split over the reduction for A^T) beyond.
static int ls_prod(Handle* h, char ta, int M, int N, int K, double alpha, const double* A, long lda, const double* X,
                   long ldx, double beta, double* Y, long ldy) {
  hipStream_t st = h->stream;
  if (N <= 4) return gemv_dev<double>(h, ta, M, N, K, alpha, A, lda, X, ldx, beta, Y, ldy, st);
  if (ta == 'N') {
    RNPX_TRY(gemm<double>('N', 'N', M, N, K, alpha, A, lda, X, ldx, beta, Y, ldy, 1, nullptr, st));
    return RNPX_OK;
  }
  const int ks = std::max(1, std::min(32, K / 512));
  RNPX_TRY(h->part.ensure(sizeof(double) * (size_t)ks * M * N, h));
  RNPX_TRY(gemm<double>('T', 'N', M, N, K, alpha, A, lda, X, ldx, beta, Y, ldy, ks, (double*)h->part.p, st));
  return RNPX_OK;
}

// Enqueues r = b - A x, s = A^T r and the squared column norms
//   nrm[q] = ||s_q||^2, nrm[nrhs + q] = ||x_q||^2,
//   nrm[3 nrhs + q] = ||b_q||^2 (with_b), nrm[4 nrhs + j] = ||A(:, j)||^2 (with_a);
// nrm[2 nrhs + q] is the refinement's ||d_q||^2, so that one read-back of
// nrm[0:3 nrhs] serves a step.
static int ls_residual_enqueue(Handle* h, int m, int n, int nrhs, const double* A, long lda, const double* B,
                               long ldb, const double* X, long ldx, const LsScratch& sc, bool with_b, bool with_a) {
  hipStream_t st = h->stream;
  RNPX_TRY(launch_copy<double>(m, nrhs, B, ldb, sc.R, m, st));
  RNPX_CHECK(ls_prod(h, 'N', m, nrhs, n, -1.0, A, lda, X, ldx, 1.0, sc.R, m));
  RNPX_CHECK(ls_prod(h, 'T', n, nrhs, m, 1.0, A, lda, sc.R, m, 0.0, sc.S, n));
  double* nr = sc.nrm;
  RNPX_TRY(launch_colnorm2(n, nrhs, sc.S, n, nr, st));
  RNPX_TRY(launch_colnorm2(n, nrhs, X, ldx, nr + nrhs, st));
  if (with_b) RNPX_TRY(launch_colnorm2(m, nrhs, B, ldb, nr + 3 * nrhs, st));
  if (with_a) RNPX_TRY(launch_colnorm2(m, n, A, lda, nr + 4 * nrhs, st));
  return RNPX_OK;
}

// eta (worst column) from ||s||^2, ||x||^2 (hn[0:2 nrhs] as above) and
// ||b||^2 (b2[0:nrhs]).
static double ls_eta(const double* hn, const double* b2, int nrhs, double anorm2) {
  const double an = sqrt(anorm2);
  double worst = 0.0;
  for (int q = 0; q < nrhs; ++q) {
    const double den = an * (an * sqrt(hn[nrhs + q]) + sqrt(b2[q]));
    const double e = den > 0 ? sqrt(hn[q]) / den : sqrt(hn[q]);
    worst = std::max(worst, e);
  }
  return worst;
}

// r = b - A x; s = A^T r; eta of x, synchronously (||A||_F^2 computed when
// *anorm2 < 0); ||b||^2 per column left in b2 when given.
static int ls_residual(Handle* h, int m, int n, int nrhs, const double* A, long lda, const double* B, long ldb,
                       const double* X, long ldx, const LsScratch& sc, double* anorm2, double* eta,
                       std::vector<double>* b2 = nullptr) {
  hipStream_t st = h->stream;
  const bool need_a = (*anorm2 < 0);
  RNPX_CHECK(ls_residual_enqueue(h, m, n, nrhs, A, lda, B, ldb, X, ldx, sc, true, need_a));
  const size_t cnt = 4 * (size_t)nrhs + (need_a ? n : 0);
  RNPX_TRY(h->hpin.ensure(sizeof(double) * cnt, h));
  double* hn = (double*)h->hpin.p;
  RNPX_TRY(hipMemcpyAsync(hn, sc.nrm, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
  RNPX_TRY(hipStreamSynchronize(st));
  if (need_a) {
    double a2 = 0.0;
    for (int j = 0; j < n; ++j) a2 += hn[4 * nrhs + j];
    *anorm2 = a2;
  }
  *eta = ls_eta(hn, hn + 3 * nrhs, nrhs, *anorm2);
  if (b2) b2->assign(hn + 3 * nrhs, hn + 4 * nrhs);
  return RNPX_OK;
}

int ls_berr_dev(Handle* h, int m, int n, int nrhs, const double* A, long lda, const double* B, long ldb,
                const double* X, long ldx, double* berr) {
  if (m < 0 || n < 0 || n > m || nrhs < 0) return -2;
  if (nrhs == 0 || n == 0) {
    *berr = 0.0;
    return RNPX_OK;
  }
  LsScratch sc;
  int s = ls_scratch(h, m, n, nrhs, &sc);
  if (s) return s;
  double a2 = -1.0;
  return ls_residual(h, m, n, nrhs, A, lda, B, ldb, X, ldx, sc, &a2, berr);
}

// Mixed-precision least squares: QR with the trailing-update products on
// FP32 MFMA (panel: norms, beta, tau in FP64; T factor in FP64), then
// refinement by corrected semi-normal equations in FP64:
//   r = b - A x,  s = A^T r,  R^T R d = s,  x += d
// until the correction is at working precision (||d|| <= 2^-50 ||x|| in
// every column), or ||d|| stops halving, or max_iter corrections; then eta
// of the final x is checked against tol.  A and B are not modified; X
// (n x nrhs) receives the solution.
// X = R^{-1} for the upper-triangular n x n R (X upper triangular, zero
// below): the 64 x 64 diagonal blocks inverted in LDS, then bottom-up: at
// width w, every pair of neighbouring inverted w-blocks becomes
// [X11 X12; 0 X22] with X12 = -X11 (R12 X22), all the level's full pairs in
// one strided batch of MFMA products (and a last, narrower pair alone).
// tmp holds n^2 / 4 + 64 n doubles.
static int trtri_upper(Handle* h, int n, const double* R, long ldr, double* X, long ldx, double* tmp) {
  hipStream_t st = h->stream;
  RNPX_TRY(launch_set<double>(n, n, 0.0, 0.0, X, ldx, st));
  RNPX_TRY(launch_trtri_diag(n, R, ldr, X, ldx, st));
  for (int w = 64; w < n; w *= 2) {
    const int full = n / (2 * w), rest = n - full * 2 * w;
    if (full > 0) {
      const long sr = 2L * w * (1 + ldr), sx = 2L * w * (1 + ldx), st2 = (long)w * w;
      // tmp_b = R12_b X22_b; X12_b = -X11_b tmp_b
      RNPX_TRY(dgemm_mfma_batched(w, w, w, 1.0, R + (size_t)w * ldr, ldr, sr, X + (size_t)w * (1 + ldx), ldx, sx, 0.0,
                                  tmp, w, st2, full, st));
      RNPX_TRY(dgemm_mfma_batched(w, w, w, -1.0, X, ldx, sx, tmp, w, st2, 0.0, X + (size_t)w * ldx, ldx, sx, full,
                                  st));
    }
    if (rest > w) {
      const int i0 = full * 2 * w, s1 = w, s2 = rest - w;
      const double* R12 = R + (size_t)i0 + (size_t)(i0 + s1) * ldr;
      const double* X11 = X + (size_t)i0 + (size_t)i0 * ldx;
      const double* X22 = X + (size_t)(i0 + s1) + (size_t)(i0 + s1) * ldx;
      double* X12 = X + (size_t)i0 + (size_t)(i0 + s1) * ldx;
      RNPX_TRY(gemm<double>('N', 'N', s1, s2, s2, 1.0, R12, ldr, X22, ldx, 0.0, tmp, s1, 1, nullptr, st));
      RNPX_TRY(gemm<double>('N', 'N', s1, s2, s1, -1.0, X11, ldx, tmp, s1, 0.0, X12, ldx, 1, nullptr, st));
    }
  }
  return RNPX_OK;
}

int gels_mixed_dev(Handle* h, int m, int n, int nrhs, const double* A, long lda, const double* B, long ldb, double* X,
                   long ldx, int max_iter, double tol, int* iters, double* berr) {
  if (m < 0) return -2;
  if (n < 0 || n > m) return -3;
  if (nrhs < 0) return -4;
  if (lda < std::max(1, m)) return -6;
  if (ldb < std::max(1, m)) return -8;
  if (ldx < std::max(1, n)) return -10;
  *iters = 0;
  *berr = 0.0;
  if (n == 0 || nrhs == 0) return RNPX_OK;
  hipStream_t st = h->stream;
  RNPX_TRY(h->vfull.ensure(sizeof(double) * ((size_t)m * n + n + (size_t)m * nrhs), h));
  double* Af = (double*)h->vfull.p;
  double* tau = Af + (size_t)m * n;
  double* Bq = tau + n;
  RNPX_TRY(launch_copy<double>(m, n, A, lda, Af, m, st));
  // x0 = R^{-1} (Q^T b)(0:n), Q^T b formed inside the factorization
  RNPX_TRY(launch_copy<double>(m, nrhs, B, ldb, Bq, m, st));
  h->update_prec = 1;
  int s = geqrf_dev<double>(h, m, n, Af, m, tau, Bq, nrhs, m);
  h->update_prec = 0;
  if (s) return s;
  // R^{-1} once: the refinement's two triangular solves per step become two
  // products (its error, kappa(R) eps, is far below the FP32 updates' that
  // the refinement corrects)
  const size_t tmpn = (size_t)n * n / 4 + 64 * (size_t)n;
  RNPX_TRY(h->rinv.ensure(sizeof(double) * ((size_t)n * n + tmpn + (size_t)n * nrhs), h));
  double* Ri = (double*)h->rinv.p;
  double* tmp = Ri + (size_t)n * n;
  double* Tn = tmp + tmpn;
  RNPX_CHECK(trtri_upper(h, n, Af, m, Ri, n, tmp));
  // x0 = R^{-1} (Q^T b)(0:n)
  RNPX_CHECK(ls_prod(h, 'N', n, nrhs, n, 1.0, Ri, n, Bq, m, 0.0, X, ldx));
  LsScratch sc;
  if ((s = ls_scratch(h, m, n, nrhs, &sc))) return s;
  double a2 = -1.0, eta = 0.0, prev = 0.0;
  std::vector<double> b2;
  if ((s = ls_residual(h, m, n, nrhs, A, lda, B, ldb, X, ldx, sc, &a2, &eta, &b2))) return s;
  // one read-back per step (pinned): the correction's norms and the eta of
  // the corrected x; ||b|| is kept from the first residual
  RNPX_TRY(h->hpin.ensure(sizeof(double) * 3 * nrhs, h));
  const double* hn = (const double*)h->hpin.p;
  double* dnrm = sc.nrm + 2 * nrhs;
  int it = 0;
  while (it < max_iter) {
    // d = R^{-1} R^{-T} s  (s is in sc.S, n x nrhs), x += d
    RNPX_CHECK(ls_prod(h, 'T', n, nrhs, n, 1.0, Ri, n, sc.S, n, 0.0, Tn, n));
    RNPX_CHECK(ls_prod(h, 'N', n, nrhs, n, 1.0, Ri, n, Tn, n, 0.0, sc.S, n));
    RNPX_TRY(launch_axpy_mat(n, nrhs, sc.S, n, X, ldx, st));
    RNPX_TRY(launch_colnorm2(n, nrhs, sc.S, n, dnrm, st));
    RNPX_CHECK(ls_residual_enqueue(h, m, n, nrhs, A, lda, B, ldb, X, ldx, sc, false, false));
    RNPX_TRY(hipMemcpyAsync(h->hpin.p, sc.nrm, sizeof(double) * 3 * nrhs, hipMemcpyDeviceToHost, st));
    RNPX_TRY(hipStreamSynchronize(st));
    ++it;
    eta = ls_eta(hn, b2.data(), nrhs, a2);
    double rel = 0.0;
    for (int q = 0; q < nrhs; ++q) {
      const double d2 = hn[2 * nrhs + q], x2 = hn[nrhs + q];
      rel = std::max(rel, x2 > 0 ? sqrt(d2 / x2) : sqrt(d2));
    }
    // stop when the correction is at working precision or stops halving
    const bool small = rel <= 0x1p-50;
    const bool stalled = it > 1 && !(rel < 0.5 * prev);
    prev = rel;
    if (small || stalled) break;
  }
  *iters = it;
  *berr = eta;
  return (eta <= tol) ? RNPX_OK : RNPX_ERR_NOTCONV;
}
template int gels_dev<zd>(Handle*, int, int, int, zd*, long, zd*, zd*, long);

// ---------------------------------------------------------------------------
// Batched
// ---------------------------------------------------------------------------
template <typename T>
int batch_setup(int count, const rnpx_qr_desc* desc, BatchSetup* out) {
  std::vector<double> cost(count);
  for (int i = 0; i < count; ++i) {
    const rnpx_qr_desc& d = desc[i];
    if (d.m < 0 || d.n < 0 || d.lda < std::max(1, d.m) || d.nrhs < 0) return -3;
    if (d.nrhs > 0 && (d.ldb < std::max(1, d.m) || d.m < d.n)) return -3;
    if (d.a_off < 0 || d.tau_off < 0 || d.b_off < 0) return -3;
    const double k = std::min(d.m, d.n);
    cost[i] = (double)d.m * d.n * k + (double)d.nrhs * d.m * d.n;
    const long need = (long)(d.m | 1) * d.n * (long)sizeof(T);
    (need <= out->lds_bytes ? out->small : out->large).push_back(i);
  }
  auto by_cost = [&](int a, int b) { return cost[a] != cost[b] ? cost[a] > cost[b] : a < b; };
  std::sort(out->small.begin(), out->small.end(), by_cost);
  std::sort(out->large.begin(), out->large.end(), by_cost);
  return RNPX_OK;
}

struct BatchBuffers {
  Buffer desc, order, counters;
};

template <typename T>
static int batch_upload(Handle* h, int count, const rnpx_qr_desc* desc, const BatchSetup& bs, BatchBuffers* bb) {
  RNPX_TRY(bb->desc.ensure(sizeof(rnpx_qr_desc) * (size_t)std::max(count, 1), h));
  RNPX_TRY(bb->order.ensure(sizeof(int) * (size_t)std::max(count, 1), h));
  RNPX_TRY(bb->counters.ensure(64, h));
  std::vector<int> order(bs.large);
  order.insert(order.end(), bs.small.begin(), bs.small.end());
  RNPX_TRY(hipMemcpyAsync(bb->desc.p, desc, sizeof(rnpx_qr_desc) * count, hipMemcpyHostToDevice, h->stream));
  RNPX_TRY(hipMemcpyAsync(bb->order.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, h->stream));
  // the host vectors die with this call: make the copies complete first
  RNPX_TRY(hipStreamSynchronize(h->stream));
  return RNPX_OK;
}

// Enqueue both tiers: the in-place tier on the panel stream, the LDS tier on
// the main stream, joined back into the main stream.
template <typename T>
static int batch_launch(Handle* h, const BatchSetup& bs, BatchBuffers* bb, T* A, T* tau, T* B) {
  const int* order = (const int*)bb->order.p;
  const rnpx_qr_desc* d = (const rnpx_qr_desc*)bb->desc.p;
  unsigned* ctr = (unsigned*)bb->counters.p;
  const int nl = (int)bs.large.size(), ns = (int)bs.small.size();
  int s = h->ensure_pipeline();
  if (s) return s;
  if (nl > 0) {
    RNPX_TRY(hipEventRecord(h->ev_start, h->stream));
    RNPX_TRY(hipStreamWaitEvent(h->pstream, h->ev_start, 0));
    RNPX_TRY(launch_batched<T>(1, nl, order, d, A, tau, B, ctr, batched_large_teams<T>(nl, h->num_cus), bs.blocked_min, h->pstream));
  }
  if (ns > 0) {
    const int per_cu = std::max(1, (160 * 1024) / bs.lds_bytes);
    const int teams = std::min(ns, per_cu * h->num_cus);
    RNPX_TRY(launch_batched<T>(0, ns, order + nl, d, A, tau, B, ctr + 1, teams, bs.lds_bytes, h->stream));
  }
  if (nl > 0) {
    RNPX_TRY(hipEventRecord(h->ev_next, h->pstream));
    RNPX_TRY(hipStreamWaitEvent(h->stream, h->ev_next, 0));
  }
  return RNPX_OK;
}

template <typename T>
int geqrf_batched(Handle* h, int count, const rnpx_qr_desc* desc, T* A, T* tau, T* B) {
  if (count < 0) return -2;
  if (count == 0) return RNPX_OK;
  if (!desc) return -3;
  if (!A) return -4;
  if (!tau) return -5;
  BatchSetup bs;
  batch_params<T>(&bs);
  int s = batch_setup<T>(count, desc, &bs);
  if (s) return s;
  // one set of descriptor buffers per handle
  BatchBuffers* bb = reinterpret_cast<BatchBuffers*>(h->batch);
  if (!bb) {
    bb = new BatchBuffers();
    h->batch = bb;
    h->batch_free = [](void* p) {
      BatchBuffers* b = reinterpret_cast<BatchBuffers*>(p);
      b->desc.release();
      b->order.release();
      b->counters.release();
      delete b;
    };
  }
  if ((s = batch_upload<T>(h, count, desc, bs, bb))) return s;
  return batch_launch<T>(h, bs, bb, A, tau, B);
}

struct BatchedPlan {
  Handle* h = nullptr;
  BatchBuffers bb;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

template <typename T>
int geqrf_batched_plan(Handle* h, int count, const rnpx_qr_desc* desc, T* A, T* tau, T* B, BatchedPlan** out) {
  if (!out) return -7;
  if (count <= 0) return -2;
  if (!desc) return -3;
  BatchSetup bs;
  batch_params<T>(&bs);
  int s = batch_setup<T>(count, desc, &bs);
  if (s) return s;
  BatchedPlan* p = new BatchedPlan();
  p->h = h;
  if ((s = batch_upload<T>(h, count, desc, bs, &p->bb)) || (s = h->ensure_pipeline())) {
    delete p;
    return s;
  }
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    s = batch_launch<T>(h, bs, &p->bb, A, tau, B);
    hipError_t e2 = hipStreamEndCapture(h->stream, &p->graph);
    if (!s && e2 == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    else e = (e2 != hipSuccess) ? e2 : hipErrorUnknown;
  }
  if (e != hipSuccess || s) {
    if (p->graph) hipGraphDestroy(p->graph);
    p->bb.desc.release();
    p->bb.order.release();
    p->bb.counters.release();
    delete p;
    return s ? s : hip_status(e);
  }
  *out = p;
  return RNPX_OK;
}

int batched_plan_run(BatchedPlan* p) {
  if (!p) return -1;
  return hip_status(hipGraphLaunch(p->exec, p->h->stream));
}

int batched_plan_destroy(BatchedPlan* p) {
  if (!p) return -1;
  hipStreamSynchronize(p->h->stream);
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  p->bb.desc.release();
  p->bb.order.release();
  p->bb.counters.release();
  delete p;
  return RNPX_OK;
}

template <typename T>
void batch_params(BatchSetup* out) {
  int lds = 0, bmin = 0;
  if (tuning_batched(S<T>::cplx ? 'z' : 'd', &lds, &bmin)) {
    out->lds_bytes = lds;
    out->blocked_min = bmin;
  }
}
template void batch_params<double>(BatchSetup*);
template void batch_params<zd>(BatchSetup*);
template int batch_setup<double>(int, const rnpx_qr_desc*, BatchSetup*);
template int batch_setup<zd>(int, const rnpx_qr_desc*, BatchSetup*);
template int geqrf_batched<double>(Handle*, int, const rnpx_qr_desc*, double*, double*, double*);
template int geqrf_batched<zd>(Handle*, int, const rnpx_qr_desc*, zd*, zd*, zd*);
template int geqrf_batched_plan<double>(Handle*, int, const rnpx_qr_desc*, double*, double*, double*, BatchedPlan**);
template int geqrf_batched_plan<zd>(Handle*, int, const rnpx_qr_desc*, zd*, zd*, zd*, BatchedPlan**);

}  // namespace rnpx
