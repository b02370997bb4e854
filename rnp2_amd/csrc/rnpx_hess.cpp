// rnpx_hess.cpp -- Hessenberg::MultQ and Hessenberg::GenerateQ
// (RNP/LA/Hessenberg.hpp:478-588) on the device: the Q of a Hessenberg
// reduction is the QR product of its ihi-ilo-1 reflectors, which sit one
// row below the diagonal of A's columns ilo.., so both routines are the QR
// ones on shifted submatrices.
//
//   MultQ (Hessenberg.hpp:478-514): QR::MultQ(side, trans, mi, ni, nh,
//     A + (ilo+1) + ilo lda, tau + ilo, C + i1 + i2 ldc), with nh = ihi-ilo-1
//     and (mi, ni, i1, i2) = (nh, n, ilo+1, 0) for side "L", (m, nh, 0,
//     ilo+1) for side "R".
//   GenerateQ (Hessenberg.hpp:540-588): the reflector block shifted one
//     column right, identity outside rows and columns [ilo+1, ihi), then
//     QR::GenerateQ(nh, nh, nh) on the block at (ilo+1, ilo+1).
#include "rnpx_internal.hpp"

#include <algorithm>

namespace rnpx {

template <typename T>
int ormhr_dev(Handle* h, char side, char trans, int m, int n, int ilo, int ihi, const T* A, long lda, const T* tau,
              T* C, long ldc) {
  const bool left = (side == 'L');
  if (side != 'L' && side != 'R') return -1;
  if (trans != 'N' && trans != 'C' && trans != 'T') return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  const int nq = left ? m : n;
  if (ilo < 0 || ilo > nq) return -5;
  if (ihi <= ilo || ihi > nq) return -6;
  if (lda < std::max(1, nq)) return -8;
  if (ldc < std::max(1, m)) return -11;
  const int nh = ihi - ilo - 1;
  if (nh <= 0 || m == 0 || n == 0) return RNPX_OK;
  const int mi = left ? nh : m, ni = left ? n : nh, i1 = left ? ilo + 1 : 0, i2 = left ? 0 : ilo + 1;
  return ormqr_dev<T>(h, side, trans, mi, ni, nh, A + (size_t)(ilo + 1) + (size_t)ilo * lda, lda, tau + ilo,
                      C + (size_t)i1 + (size_t)i2 * ldc, ldc);
}

template <typename T>
int orghr_dev(Handle* h, int n, int ilo, int ihi, T* A, long lda, const T* tau) {
  if (n < 0) return -1;
  if (ilo < 0 || ilo > n) return -2;
  if (ihi <= ilo || ihi > n) return -3;
  if (lda < std::max(1, n)) return -5;
  if (n == 0) return RNPX_OK;
  const int nh = ihi - ilo - 1;
  const int ls = std::max(1, nh);
  RNPX_TRY(h->qlbuf.ensure(sizeof(T) * (size_t)ls * std::max(1, nh), h));
  T* S = (T*)h->qlbuf.p;
  if (nh > 0) RNPX_TRY(launch_copy<T>(nh, nh, A + (size_t)(ilo + 1) + (size_t)ilo * lda, lda, S, ls, h->stream));
  RNPX_TRY(launch_hessq_setup<T>(n, ilo, ihi, A, lda, S, ls, h->stream));
  if (nh > 0) {
    int s = orgqr_dev<T>(h, nh, nh, nh, A + (size_t)(ilo + 1) + (size_t)(ilo + 1) * lda, lda, tau + ilo);
    if (s) return s;
  }
  return RNPX_OK;
}

#define RNPX_INST(T)                                                                                              \
  template int ormhr_dev<T>(Handle*, char, char, int, int, int, int, const T*, long, const T*, T*, long);         \
  template int orghr_dev<T>(Handle*, int, int, int, T*, long, const T*);
RNPX_INST(double)
RNPX_INST(zd)
#undef RNPX_INST

}  // namespace rnpx
