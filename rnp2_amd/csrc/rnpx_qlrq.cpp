// rnpx_qlrq.cpp -- QL (RNP/LA/QL.hpp) and RQ (RNP/LA/RQ.hpp) on the device
// through the QR and LQ paths.
//
// QL::Factor_unblocked (QL.hpp:79-100) generates reflector i from column
// n-k+i (alpha at row m-k+i, x the rows above it) for i = k-1 down to 0 and
// applies it to the columns on its left.  With J the reversal permutation,
// B = J_m A J_n turns that loop into QR::Factor_unblocked (QR.hpp:78-102) on
// B step for step: column n-k+i of A is column k-1-i of B read bottom-up,
// alpha and x are the same numbers (x in reverse order, which only changes
// the rounding of its norm), so beta, tau and v agree and the trailing
// updates correspond.  Hence QL's factors are J_m QR(B) J_n with tau
// reversed, and QL's Q = H[k-1] ... H[0] (the order QL::MultQ_unblocked
// applies, QL.hpp:265-298) is J Q_qr J: MultQ reverses C's rows (side "L")
// or columns (side "R") around QR's MultQ of the reversed reflectors, and
// GenerateQ reverses around QR's GenerateQ (the last n columns of Q map to
// the first n of Q_qr).
//
// RQ::Factor_unblocked (RQ.hpp:80-102) is, the same way, LQ::Factor_unblocked
// (LQ.hpp:78-102) on J_m A J_n: both conjugate the row, generate, apply from
// the right and conjugate back, so RQ's factors are J LQ(J A J) J with tau
// reversed, and RQ's MultQ and GenerateQ reverse around LQ's.
//
// Each reversal is one read and one write of the matrix (flip_kernel); the
// reference's blocking (QL.hpp:175-196, RQ.hpp:175-200) fixes only the
// rounding order, so results agree to the QR tolerances.
#include "rnpx_internal.hpp"

#include <algorithm>

namespace rnpx {

namespace {

// The k reflectors of a MultQ call reversed into the handle's qlbuf (rows x
// cols, leading dimension rows) with tau reversed after them.
template <typename T>
int reversed_reflectors(Handle* h, int rows, int cols, const T* A, long lda, const T* tau, int k, T** V, T** tv) {
  const size_t nv = (size_t)rows * cols;
  RNPX_TRY(h->qlbuf.ensure(sizeof(T) * (nv + (size_t)k + 32), h));
  *V = (T*)h->qlbuf.p;
  *tv = *V + ((nv + 31) / 32) * 32;
  RNPX_TRY(launch_flip_copy<T>(rows, cols, A, lda, *V, rows, true, true, h->stream));
  RNPX_TRY(launch_flip_copy<T>(k, 1, tau, std::max(1, k), *tv, std::max(1, k), true, false, h->stream));
  return RNPX_OK;
}

template <typename T>
int reversed_tau(Handle* h, const T* tau, int k, T** tv) {
  RNPX_TRY(h->qlbuf.ensure(sizeof(T) * (size_t)std::max(k, 1), h));
  *tv = (T*)h->qlbuf.p;
  RNPX_TRY(launch_flip_copy<T>(k, 1, tau, std::max(1, k), *tv, std::max(1, k), true, false, h->stream));
  return RNPX_OK;
}

}  // namespace

// QL::Factor (QL.hpp:126-205).
template <typename T>
int geqlf_dev(Handle* h, int m, int n, T* A, long lda, T* tau) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  const int k = std::min(m, n);
  if (k == 0) return RNPX_OK;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  int s = geqrf_dev<T>(h, m, n, A, lda, tau);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  RNPX_TRY(launch_flip<T>(k, 1, tau, k, true, false, h->stream));
  return RNPX_OK;
}

// QL::MultQ (QL.hpp:335-430): the k reflectors are the columns of A (nq x k).
template <typename T>
int ormql_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc) {
  const bool left = (side == 'L');
  if (side != 'L' && side != 'R') return -1;
  if (trans != 'N' && trans != 'C' && trans != 'T') return -2;
  if (S<T>::cplx && trans == 'T') return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  const int nq = left ? m : n;
  if (k < 0 || k > nq) return -5;
  if (lda < std::max(1, nq)) return -7;
  if (ldc < std::max(1, m)) return -10;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  T *V = nullptr, *tv = nullptr;
  int s = reversed_reflectors<T>(h, nq, k, A, lda, tau, k, &V, &tv);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, C, ldc, left, !left, h->stream));
  s = ormqr_dev<T>(h, side, trans, m, n, k, V, nq, tv, C, ldc);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, C, ldc, left, !left, h->stream));
  return RNPX_OK;
}

// QL::GenerateQ (QL.hpp:501-565): a (m x n, m >= n, the reflectors in its
// last k columns) <- the last n columns of Q.
template <typename T>
int orgql_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau) {
  if (m < 0) return -1;
  if (n < 0 || n > m) return -2;
  if (k < 0 || k > n) return -3;
  if (lda < std::max(1, m)) return -5;
  if (n == 0) return RNPX_OK;
  T* tv = nullptr;
  int s = reversed_tau<T>(h, tau, k, &tv);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  s = orgqr_dev<T>(h, m, n, k, A, lda, tv);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  return RNPX_OK;
}

// RQ::Factor (RQ.hpp:131-205).
template <typename T>
int gerqf_dev(Handle* h, int m, int n, T* A, long lda, T* tau) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  const int k = std::min(m, n);
  if (k == 0) return RNPX_OK;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  int s = gelqf_dev<T>(h, m, n, A, lda, tau);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  RNPX_TRY(launch_flip<T>(k, 1, tau, k, true, false, h->stream));
  return RNPX_OK;
}

// RQ::MultQ (RQ.hpp:344-440): the k reflectors are the rows of A (k x nq).
template <typename T>
int ormrq_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc) {
  const bool left = (side == 'L');
  if (side != 'L' && side != 'R') return -1;
  if (trans != 'N' && trans != 'C' && trans != 'T') return -2;
  if (S<T>::cplx && trans == 'T') return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  const int nq = left ? m : n;
  if (k < 0 || k > nq) return -5;
  if (lda < std::max(1, k)) return -7;
  if (ldc < std::max(1, m)) return -10;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  T *V = nullptr, *tv = nullptr;
  int s = reversed_reflectors<T>(h, k, nq, A, lda, tau, k, &V, &tv);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, C, ldc, left, !left, h->stream));
  s = ormlq_dev<T>(h, side, trans, m, n, k, V, k, tv, C, ldc);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, C, ldc, left, !left, h->stream));
  return RNPX_OK;
}

// RQ::GenerateQ (RQ.hpp:519-580): a (m x n, m <= n, the reflectors in its
// last k rows) <- the last m rows of Q.
template <typename T>
int orgrq_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau) {
  if (m < 0) return -1;
  if (n < m) return -2;
  if (k < 0 || k > m) return -3;
  if (lda < std::max(1, m)) return -5;
  if (m == 0) return RNPX_OK;
  T* tv = nullptr;
  int s = reversed_tau<T>(h, tau, k, &tv);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  s = orglq_dev<T>(h, m, n, k, A, lda, tv);
  if (s) return s;
  RNPX_TRY(launch_flip<T>(m, n, A, lda, true, true, h->stream));
  return RNPX_OK;
}

#define RNPX_INST(T)                                                                                           \
  template int geqlf_dev<T>(Handle*, int, int, T*, long, T*);                                                  \
  template int ormql_dev<T>(Handle*, char, char, int, int, int, const T*, long, const T*, T*, long);           \
  template int orgql_dev<T>(Handle*, int, int, int, T*, long, const T*);                                       \
  template int gerqf_dev<T>(Handle*, int, int, T*, long, T*);                                                  \
  template int ormrq_dev<T>(Handle*, char, char, int, int, int, const T*, long, const T*, T*, long);           \
  template int orgrq_dev<T>(Handle*, int, int, int, T*, long, const T*);
RNPX_INST(double)
RNPX_INST(zd)
#undef RNPX_INST

}  // namespace rnpx
