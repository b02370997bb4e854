// rnpx_blas.cpp -- the vendor DGEMM (rocBLAS) for the plain trailing-update
// product C -= V Z of real factorizations, where it beats the library's own
// MFMA GEMM at the update's shapes (K = 128).  Two rocBLAS handles per
// library handle, created on first use: one for the look-ahead's panel stream
// and one for every other stream, so that products running concurrently on
// the panel and update streams never share a handle's workspace.  Each
// is bound to the calling stream per call.
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
  void*& slot = (st && (st == h->pstream || st == h->part_ps)) ? h->blas_side : h->blas;
  if (!slot) {
    rocblas_handle rh = nullptr;
    if (rocblas_create_handle(&rh) != rocblas_status_success) return RNPX_ERR_HIP;
    // Atomics would let rocBLAS sum split products in arrival order; without
    // them a factorization is reproducible bit for bit (RNPX_BLAS_ATOMICS=1
    // allows them).
    static const bool atomics = [] {
      const char* e = getenv("RNPX_BLAS_ATOMICS");
      return e && atoi(e) == 1;
    }();
    if (rocblas_set_atomics_mode(rh, atomics ? rocblas_atomics_allowed : rocblas_atomics_not_allowed) !=
        rocblas_status_success) {
      rocblas_destroy_handle(rh);
      return RNPX_ERR_HIP;
    }
    slot = rh;
  }
  rocblas_handle rh = (rocblas_handle)slot;
  if (rocblas_set_stream(rh, st) != rocblas_status_success) return RNPX_ERR_HIP;
  auto op = [](char t) { return t == 'N' ? rocblas_operation_none : rocblas_operation_transpose; };
  const rocblas_status s = rocblas_dgemm(rh, op(ta), op(tb), m, n, k, &alpha, A, (rocblas_int)lda, B,
                                         (rocblas_int)ldb, &beta, C, (rocblas_int)ldc);
  return s == rocblas_status_success ? RNPX_OK : RNPX_ERR_HIP;
}

void vendor_release(Handle* h) {
  for (void** p : {&h->blas, &h->blas_side}) {
    if (*p) rocblas_destroy_handle((rocblas_handle)*p);
    *p = nullptr;
  }
}

}  // namespace rnpx
