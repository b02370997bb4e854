// rnpx_lu_driver.cpp -- LU::Factor and LU::Solve (RNP/LA/LU.hpp:77-144) on the
// device: the blocked loop of LU::Factor with each 64-column panel factored
// by one cooperative launch, the panel's row interchanges applied to the
// other columns, U12 by substitution and the trailing update as a GEMM
// (rocBLAS DGEMM for real matrices, the library's MFMA GEMM otherwise).
//
// Return value: the reference's arithmetic (LU.hpp:63-74, 94-96, 112-113):
// a panel whose last exact-zero pivot is its column z > 0 reports
// iinfo = z + 1; the first such panel gives info = iinfo + j, and Factor
// returns info + 1 (info = 0: success).  With min(m, n) <= 64 the whole
// factorization is Factor_unblocked's, which returns z + 1 for z > 0.
#include "rnpx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

namespace rnpx {

template <typename T>
int getrf_dev(Handle* h, int m, int n, T* A, long lda, int* piv_host, int* info) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  *info = 0;
  const int k = std::min(m, n);
  if (k == 0) return RNPX_OK;
  const int nb = kLuPanelNB;
  const int npan = (k + nb - 1) / nb;
  const int maxg = std::min(256, h->num_cus);
  const int rw = lu_panel_rows(m, maxg, S<T>::cplx ? 1 : 0);
  if (rw == 0) return RNPX_ERR_UNSUPPORTED;
  // scratch: candidates, row c, results (doubles); sync words; zero-pivot
  // columns and relative pivots (ints)
  const size_t nd = (size_t)2 * 256 * kLuPartPerGroup + 4 * nb + 2 * nb;
  const size_t bytes = sizeof(double) * nd + sizeof(unsigned) * 2 * nb + sizeof(int) * ((size_t)npan + k);
  RNPX_TRY(h->lubuf.ensure(bytes, h));
  double* part = (double*)h->lubuf.p;
  double* rowc = part + (size_t)2 * 256 * kLuPartPerGroup;
  unsigned* sync = (unsigned*)(rowc + 6 * nb);
  int* zl = (int*)(sync + 2 * nb);
  int* prel = zl + npan;
  // Panel pi factors columns [j, j+jb).  Its update of the columns to the
  // right is split: the next panel's columns are brought up to date on the
  // panel stream right behind it (so the chain of panels crosses no stream),
  // the rest on the handle's stream, overlapped with the next panel.  The
  // panel stream waits only for the previous panel's rest-of-update, which
  // covered the next panel's columns; the interchanges of columns left of
  // the panel go to the handle's stream in panel order.
  hipStream_t st = h->stream, ps = st;
  const bool look = npan > 1;
  if (look) {
    const int e = h->ensure_pipeline();
    if (e) return e;
    ps = h->pstream;
    RNPX_TRY(hipEventRecord(h->ev_start, st));
    RNPX_TRY(hipStreamWaitEvent(ps, h->ev_start, 0));
  }
  auto panel = [&](int pi) -> int {
    const int j = pi * nb, jb = std::min(nb, k - j), mp = m - j;
    RNPX_TRY(launch_zero_u32(sync, 2 * nb, ps));
    LuPanelArgs pa;
    pa.mp = mp;
    pa.jb = jb;
    pa.A = A + (size_t)j + (size_t)j * lda;
    pa.lda = lda;
    pa.piv = prel + j;
    pa.zero_last = zl + pi;
    pa.part = part;
    pa.rowc = rowc;
    pa.sync = sync;
    pa.abortf = (unsigned*)h->flag.p;
    RNPX_TRY(launch_lu_panel<T>(pa, lu_panel_rows(mp, maxg, S<T>::cplx ? 1 : 0), ps));
    if (look) RNPX_TRY(hipEventRecord(h->ev_pan[pi & 1], ps));
    return RNPX_OK;
  };
  // Panel pi's interchanges, U12 and A22 update for columns [c0, c0+nc).
  auto right = [&](int pi, int c0, int nc, hipStream_t s) -> int {
    if (nc <= 0) return RNPX_OK;
    const int j = pi * nb, jb = std::min(nb, k - j);
    RNPX_TRY(launch_lu_laswp<T>(j, jb, prel + j, A, lda, c0, nc, s));
    RNPX_TRY(launch_lu_trsm<T>(jb, A + (size_t)j + (size_t)j * lda, lda, A + (size_t)j + (size_t)c0 * lda, lda, nc, s));
    if (j + jb < m) {
      T* L21 = A + (size_t)(j + jb) + (size_t)j * lda;
      T* U12 = A + (size_t)j + (size_t)c0 * lda;
      T* A22 = A + (size_t)(j + jb) + (size_t)c0 * lda;
      const T one = S<T>::one(), mone = S<T>::zero() - S<T>::one();
      RNPX_TRY(gemm<T>('N', 'N', m - j - jb, nc, jb, mone, L21, lda, U12, lda, one, A22, lda, 1, nullptr, s));
    }
    return RNPX_OK;
  };
  int e = panel(0);
  if (e) return e;
  for (int pi = 0; pi < npan; ++pi) {
    const int j = pi * nb, jb = std::min(nb, k - j);
    const int c1 = j + jb;  // first column right of the panel
    if (look) RNPX_TRY(hipStreamWaitEvent(st, h->ev_pan[pi & 1], 0));
    int cr = c1;  // first column of the handle stream's part
    if (pi + 1 < npan) {
      const int jb2 = std::min(nb, k - c1);
      if (pi > 0) RNPX_TRY(hipStreamWaitEvent(ps, h->ev_rest, 0));
      if ((e = right(pi, c1, jb2, ps))) return e;
      if ((e = panel(pi + 1))) return e;
      cr = c1 + jb2;
    }
    if (j > 0) RNPX_TRY(launch_lu_laswp<T>(j, jb, prel + j, A, lda, 0, j, st));
    if ((e = right(pi, cr, n - cr, st))) return e;
    if (look) RNPX_TRY(hipEventRecord(h->ev_rest, st));
  }
  std::vector<int> z(npan), pr(k);
  RNPX_TRY(hipMemcpyAsync(z.data(), zl, sizeof(int) * npan, hipMemcpyDeviceToHost, h->stream));
  RNPX_TRY(hipMemcpyAsync(pr.data(), prel, sizeof(int) * k, hipMemcpyDeviceToHost, h->stream));
  int s = sync_handle(h);
  if (s) return s;
  for (int i = 0; i < k; ++i) piv_host[i] = pr[i] + (i / nb) * nb;
  if (k <= nb) {
    *info = z[0] > 0 ? z[0] + 1 : 0;
  } else {
    int inf = 0;
    for (int pi = 0; pi < npan && inf == 0; ++pi) {
      const int iinfo = z[pi] > 0 ? z[pi] + 1 : 0;
      if (iinfo > 0) inf = iinfo + pi * nb;
    }
    *info = inf ? inf + 1 : 0;
  }
  return RNPX_OK;
}

// LU::Solve (LU.hpp:117-144); ipiv on the host (0-based absolute rows).
template <typename T>
int getrs_dev(Handle* h, char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb) {
  if (trans != 'N' && trans != 'T' && trans != 'C') return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldb < std::max(1, n)) return -8;
  if (n == 0 || nrhs == 0) return RNPX_OK;
  // the interchanges composed into one gather: row i <- row perm[i]
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  if (trans == 'N') {
    for (int i = 0; i < n; ++i) std::swap(perm[i], perm[ipiv[i]]);
  } else {
    for (int i = n - 1; i >= 0; --i) std::swap(perm[i], perm[ipiv[i]]);
  }
  bool ident = true;
  for (int i = 0; i < n && ident; ++i) ident = (perm[i] == i);
  const size_t pbytes = sizeof(int) * (size_t)n;
  RNPX_TRY(h->lubuf.ensure(pbytes + sizeof(T) * (size_t)n * nrhs + 256, h));
  int* dperm = (int*)h->lubuf.p;
  T* tmp = (T*)((char*)h->lubuf.p + ((pbytes + 255) / 256) * 256);
  auto gather = [&]() -> int {
    if (ident) return RNPX_OK;
    RNPX_TRY(hipMemcpyAsync(dperm, perm.data(), pbytes, hipMemcpyHostToDevice, h->stream));
    RNPX_TRY(launch_lu_gather<T>(n, nrhs, dperm, B, ldb, tmp, n, h->stream));
    RNPX_TRY(launch_copy<T>(n, nrhs, tmp, n, B, ldb, h->stream));
    return RNPX_OK;
  };
  int s = 0;
  if (trans == 'N') {
    if ((s = gather())) return s;
    if ((s = trsm_tri_dev<T>(h, 'L', 'N', 'U', n, nrhs, A, lda, B, ldb))) return s;
    if ((s = trsm_tri_dev<T>(h, 'U', 'N', 'N', n, nrhs, A, lda, B, ldb))) return s;
  } else {
    if ((s = trsm_tri_dev<T>(h, 'U', trans, 'N', n, nrhs, A, lda, B, ldb))) return s;
    if ((s = trsm_tri_dev<T>(h, 'L', trans, 'U', n, nrhs, A, lda, B, ldb))) return s;
    if ((s = gather())) return s;
  }
  // the host permutation must outlive the copy
  RNPX_TRY(hipStreamSynchronize(h->stream));
  return RNPX_OK;
}

namespace {

template <typename T> struct HostT;
template <> struct HostT<double> {
  using type = double;
  static double norm1(double v) { return std::fabs(v); }
  static double abs(double v) { return std::fabs(v); }
  static double re(double v) { return v; }
};
template <> struct HostT<zd> {
  using type = zd;
  static double norm1(zd v) { return std::fabs(v.x) + std::fabs(v.y); }
  static double abs(zd v) { return std::hypot(v.x, v.y); }
  static double re(zd v) { return v.x; }
};
inline double scal(double v, double s) { return v * s; }
inline zd scal(zd v, double s) { return mkz(v.x * s, v.y * s); }
inline double setv(double, double r) { return r; }
inline zd setv(zd, double r) { return mkz(r, 0.0); }

}  // namespace

// LU::RefineSolution (LU.hpp:216-366) with its documented intent where the
// reference does not compile (the LU solve it calls PLUSolve; the estimator
// it calls MatrixNorm1Estimate with MatrixNorm1Estimate_unblocked's
// argument list, MatrixNormEstimates.hpp:17-107; the safe1 term added, as
// the comment at LU.hpp:322-323 says).  The residuals, |op(A)||x| + |b| and
// the corrections run on the device; the per-column scalars (berr, the
// stopping test, the estimator's vectors of length n) on the host.
template <typename T>
int gerfs_dev(Handle* h, char trans, int n, int nrhs, const T* A, long lda, const T* AF, long ldaf, const int* ipiv,
              const T* B, long ldb, T* X, long ldx, double* ferr, double* berr) {
  using H = HostT<T>;
  if (trans != 'N' && trans != 'T' && trans != 'C') return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) {
      if (ferr) ferr[j] = 0.0;
      if (berr) berr[j] = 0.0;
    }
    return RNPX_OK;
  }
  const double eps = std::numeric_limits<double>::epsilon(), dmin = std::numeric_limits<double>::min();
  const double nz = (double)(n + 1), safe1 = nz * dmin, safe2 = safe1 / eps;
  RNPX_TRY(h->rfbuf.ensure(sizeof(T) * (size_t)n + sizeof(double) * (size_t)n + 256, h));
  T* r = (T*)h->rfbuf.p;
  double* rwd = (double*)((char*)h->rfbuf.p + ((sizeof(T) * (size_t)n + 255) / 256) * 256);
  std::vector<T> rh(n), xh(n);
  std::vector<double> rw(n);
  const T one = S<T>::one(), mone = S<T>::zero() - S<T>::one();
  const char tn = (trans == 'N') ? 'N' : 'C', tt = (trans == 'N') ? 'C' : 'N';
  for (int j = 0; j < nrhs; ++j) {
    T* xj = X + (size_t)j * ldx;
    const T* bj = B + (size_t)j * ldb;
    int count = 1;
    double lstres = 3.0, s = 0.0;
    for (;;) {
      RNPX_TRY(launch_copy<T>(n, 1, bj, std::max(1L, ldb), r, n, h->stream));
      RNPX_TRY(gemm<T>(trans, 'N', n, 1, n, mone, A, lda, xj, std::max(1L, ldx), one, r, n, 1, nullptr, h->stream));
      RNPX_TRY(launch_lu_absmv<T>(n, trans != 'N', A, lda, xj, bj, rwd, h->stream));
      RNPX_TRY(hipMemcpyAsync(rh.data(), r, sizeof(T) * n, hipMemcpyDeviceToHost, h->stream));
      RNPX_TRY(hipMemcpyAsync(rw.data(), rwd, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
      RNPX_TRY(hipStreamSynchronize(h->stream));
      s = 0.0;
      for (int i = 0; i < n; ++i) {
        const double q = (rw[i] > safe2) ? H::norm1(rh[i]) / rw[i] : (H::norm1(rh[i]) + safe1) / (rw[i] + safe1);
        if (q > s) s = q;
      }
      if (berr) berr[j] = s;
      if (s > eps && 2.0 * s <= lstres && count < 5) {
        int st = getrs_dev<T>(h, trans, n, 1, AF, ldaf, ipiv, r, n);
        if (st) return st;
        RNPX_TRY(launch_lu_axpy<T>(n, r, xj, h->stream));
        lstres = s;
        ++count;
        continue;
      }
      break;
    }
    // w = |r| + nz eps (|op(A)||x| + |b|), + safe1 where the latter <= safe2
    for (int i = 0; i < n; ++i)
      rw[i] = (rw[i] > safe2) ? H::norm1(rh[i]) + nz * eps * rw[i] : H::norm1(rh[i]) + nz * eps * rw[i] + safe1;
    // the estimator's operator on host vectors: "N" solves with op_t then
    // scales by w; "C" scales then solves with op_n (LU.hpp:34-50)
    int est_status = 0;
    auto aop = [&](char op, std::vector<T>& v) {
      if (est_status) return;
      if (op == 'C')
        for (int i = 0; i < n; ++i) v[i] = scal(v[i], rw[i]);
      hipError_t e = hipMemcpyAsync(r, v.data(), sizeof(T) * n, hipMemcpyHostToDevice, h->stream);
      if (e != hipSuccess) { est_status = hip_status(e); return; }
      est_status = getrs_dev<T>(h, op == 'N' ? tt : tn, n, 1, AF, ldaf, ipiv, r, n);
      if (est_status) return;
      e = hipMemcpyAsync(v.data(), r, sizeof(T) * n, hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) { est_status = hip_status(e); return; }
      if (op == 'N')
        for (int i = 0; i < n; ++i) v[i] = scal(v[i], rw[i]);
    };
    // MatrixNorm1Estimate_unblocked (MatrixNormEstimates.hpp:17-107)
    std::vector<T> v(n), x(n);
    double est = 0.0;
    for (int i = 0; i < n; ++i) x[i] = setv(x[i], 1.0 / (double)n);
    aop('N', x);
    if (n == 1) {
      est = H::abs(x[0]);
    } else {
      for (int i = 0; i < n; ++i) est += H::abs(x[i]);
      auto sgn = [&]() {
        for (int i = 0; i < n; ++i) {
          const double ax = H::abs(x[i]);
          x[i] = (ax > dmin) ? scal(x[i], 1.0 / ax) : setv(x[i], 1.0);
        }
      };
      auto argmax_re = [&]() {
        int im = 0;
        double am = 0.0;
        for (int i = 0; i < n; ++i) {
          const double ar = std::fabs(H::re(x[i]));
          if (ar > am) { am = ar; im = i; }
        }
        return im;
      };
      sgn();
      aop('C', x);
      int imax = argmax_re(), jlast = 0, iter = 1;
      do {
        ++iter;
        for (int i = 0; i < n; ++i) x[i] = setv(x[i], 0.0);
        x[imax] = setv(x[imax], 1.0);
        aop('N', x);
        v = x;
        const double estold = est;
        est = 0.0;
        for (int i = 0; i < n; ++i) est += H::abs(v[i]);
        if (est <= estold) break;
        sgn();
        aop('C', x);
        jlast = imax;
        imax = argmax_re();
      } while (H::abs(x[jlast]) != H::abs(x[imax]) && iter < 5);
      double alt = 1.0;
      for (int i = 0; i < n; ++i) {
        x[i] = setv(x[i], alt * (1.0 + (double)i / (double)(n - 1)));
        alt = -alt;
      }
      aop('N', x);
      double t = 0.0;
      for (int i = 0; i < n; ++i) t += H::abs(x[i]);
      t = 2.0 * t / (double)(3 * n);
      if (t >= est) est = t;
    }
    if (est_status) return est_status;
    RNPX_TRY(hipMemcpyAsync(xh.data(), xj, sizeof(T) * n, hipMemcpyDeviceToHost, h->stream));
    RNPX_TRY(hipStreamSynchronize(h->stream));
    double xm = 0.0;
    for (int i = 0; i < n; ++i) xm = std::max(xm, H::norm1(xh[i]));
    if (ferr) ferr[j] = (xm != 0.0) ? est / xm : est;
  }
  return RNPX_OK;
}

template int gerfs_dev<double>(Handle*, char, int, int, const double*, long, const double*, long, const int*,
                               const double*, long, double*, long, double*, double*);
template int gerfs_dev<zd>(Handle*, char, int, int, const zd*, long, const zd*, long, const int*, const zd*, long,
                           zd*, long, double*, double*);

template int getrf_dev<double>(Handle*, int, int, double*, long, int*, int*);
template int getrf_dev<zd>(Handle*, int, int, zd*, long, int*, int*);
// LU::Invert (RNP/LA/LU.hpp:146-214).  The reference first inverts U in
// place (Triangular::Invert, which returns -(i+1) for the first zero U_ii
// before touching anything), then solves inv(A) L = inv(U) and undoes the
// column interchanges.  Here the same inverse comes from the factors through
// the device LU solve: X = A^{-1} I (the interchanges gathered into I, L then
// U solved by the blocked triangular kernels), copied into A.  The zero test
// reads U's diagonal to the host first, so A is untouched when it fails.
template <typename T>
int getri_dev(Handle* h, int n, T* A, long lda, const int* ipiv, int* info) {
  *info = 0;
  if (n < 0) return -1;
  if (lda < std::max(1, n)) return -3;
  if (n == 0) return RNPX_OK;
  RNPX_TRY(h->invbuf.ensure(sizeof(T) * (size_t)n * n, h));
  T* X = (T*)h->invbuf.p;
  std::vector<T> diag(n);
  RNPX_TRY(hipMemcpy2DAsync(diag.data(), sizeof(T), A, sizeof(T) * (size_t)(lda + 1), sizeof(T), n,
                            hipMemcpyDeviceToHost, h->stream));
  RNPX_TRY(hipStreamSynchronize(h->stream));
  for (int i = 0; i < n; ++i) {
    if (S<T>::is_zero(diag[i])) {
      *info = -(i + 1);
      return RNPX_OK;
    }
  }
  RNPX_TRY(launch_set<T>(n, n, S<T>::zero(), S<T>::one(), X, n, h->stream));
  int s = getrs_dev<T>(h, 'N', n, n, A, lda, ipiv, X, n);
  if (s) return s;
  RNPX_TRY(launch_copy<T>(n, n, X, n, A, lda, h->stream));
  return RNPX_OK;
}
template int getri_dev<double>(Handle*, int, double*, long, const int*, int*);
template int getri_dev<zd>(Handle*, int, zd*, long, const int*, int*);

template int getrs_dev<double>(Handle*, char, int, int, const double*, long, const int*, double*, long);
template int getrs_dev<zd>(Handle*, char, int, int, const zd*, long, const int*, zd*, long);

}  // namespace rnpx
