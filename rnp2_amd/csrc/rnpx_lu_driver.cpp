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
#include <numeric>
#include <vector>

namespace rnpx {

#define RNPX_TRY(x)                               \
  do {                                            \
    hipError_t _e = (x);                          \
    if (_e != hipSuccess) return hip_status(_e);  \
  } while (0)

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
  for (int pi = 0; pi < npan; ++pi) {
    const int j = pi * nb, jb = std::min(nb, k - j), mp = m - j;
    RNPX_TRY(launch_zero_u32(sync, 2 * nb, h->stream));
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
    RNPX_TRY(launch_lu_panel<T>(pa, lu_panel_rows(mp, maxg, S<T>::cplx ? 1 : 0), h->stream));
    if (j > 0) RNPX_TRY(launch_lu_laswp<T>(j, jb, prel + j, A, lda, 0, j, h->stream));
    if (j + jb < n) {
      RNPX_TRY(launch_lu_laswp<T>(j, jb, prel + j, A, lda, j + jb, n - j - jb, h->stream));
      RNPX_TRY(launch_lu_trsm<T>(jb, A + (size_t)j + (size_t)j * lda, lda, A + (size_t)j + (size_t)(j + jb) * lda, lda,
                                 n - j - jb, h->stream));
      if (j + jb < m) {
        T* L21 = A + (size_t)(j + jb) + (size_t)j * lda;
        T* U12 = A + (size_t)j + (size_t)(j + jb) * lda;
        T* A22 = A + (size_t)(j + jb) + (size_t)(j + jb) * lda;
        bool done = false;
        if constexpr (!S<T>::cplx) {
          if (vendor_update_enabled()) {
            const int s = vendor_dgemm_nn(h, h->stream, m - j - jb, n - j - jb, jb, -1.0, L21, lda, U12, lda, 1.0,
                                          A22, lda);
            if (s) return s;
            done = true;
          }
        }
        if (!done) {
          const T one = S<T>::one(), mone = S<T>::zero() - S<T>::one();
          RNPX_TRY(gemm<T>('N', 'N', m - j - jb, n - j - jb, jb, mone, L21, lda, U12, lda, one, A22, lda, 1, nullptr,
                           h->stream));
        }
      }
    }
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

template int getrf_dev<double>(Handle*, int, int, double*, long, int*, int*);
template int getrf_dev<zd>(Handle*, int, int, zd*, long, int*, int*);
template int getrs_dev<double>(Handle*, char, int, int, const double*, long, const int*, double*, long);
template int getrs_dev<zd>(Handle*, char, int, int, const zd*, long, const int*, zd*, long);

}  // namespace rnpx
