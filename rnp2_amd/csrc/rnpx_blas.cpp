// rnpx_blas.cpp -- the vendor DGEMM (rocBLAS) for the plain trailing-update
// product C -= V Z of real factorizations, where it beats the library's own
// MFMA GEMM at the update's shapes (K = 128).  One rocBLAS handle per
// library handle, created on first use, bound to the calling stream per call.
#include "rnpx_internal.hpp"

#include <rocblas/rocblas.h>

#include <cstdlib>
#include <cstring>

namespace rnpx {

// RNPX_UPDATE_GEMM: "vendor" or "own" (read once); unset = vendor.
bool vendor_update_enabled() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("RNPX_UPDATE_GEMM");
    v = (e && strcmp(e, "own") == 0) ? 0 : 1;
  }
  return v == 1;
}

int vendor_dgemm_nn(Handle* h, hipStream_t st, int m, int n, int k, double alpha, const double* A, long lda,
                    const double* B, long ldb, double beta, double* C, long ldc) {
  return vendor_dgemm(h, st, 'N', 'N', m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int vendor_dgemm(Handle* h, hipStream_t st, char ta, char tb, int m, int n, int k, double alpha, const double* A,
                 long lda, const double* B, long ldb, double beta, double* C, long ldc) {
  if (!h->blas) {
    rocblas_handle rh = nullptr;
    if (rocblas_create_handle(&rh) != rocblas_status_success) return RNPX_ERR_HIP;
    h->blas = rh;
  }
  rocblas_handle rh = (rocblas_handle)h->blas;
  if (rocblas_set_stream(rh, st) != rocblas_status_success) return RNPX_ERR_HIP;
  auto op = [](char t) { return t == 'N' ? rocblas_operation_none : rocblas_operation_transpose; };
  const rocblas_status s = rocblas_dgemm(rh, op(ta), op(tb), m, n, k, &alpha, A, (rocblas_int)lda, B,
                                         (rocblas_int)ldb, &beta, C, (rocblas_int)ldc);
  return s == rocblas_status_success ? RNPX_OK : RNPX_ERR_HIP;
}

void vendor_release(Handle* h) {
  if (h->blas) rocblas_destroy_handle((rocblas_handle)h->blas);
  h->blas = nullptr;
}

}  // namespace rnpx
