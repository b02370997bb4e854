// rnpx_lq.cpp -- LQ (RNP/LA/LQ.hpp) on the device through the QR kernels.
//
// LQ::Factor_unblocked (LQ.hpp:78-102) conjugates row i, generates its
// reflector with Reflector::Generate and applies it from the right: row by
// row that is QR::Factor_unblocked (QR.hpp:78-102) on A^H, column i of A^H
// being the conjugated row i of A, with the same alpha, x, beta and tau.  So
// the LQ factors are the QR factors of A^H conjugate-transposed: L = R^H on
// and below the diagonal, the reflectors conjugated in the rows above it,
// tau unchanged.  With Q_qr = H_0 H_1 ... H_{k-1} the QR product of those
// reflectors, LQ's Q is Q_qr^H (LQ.hpp:195-196 and MultQ_unblocked,
// LQ.hpp:245-270: "N" applies the H_i with conj(tau) in the order that
// forms Q_qr^H), so MultQ is QR's MultQ with trans flipped, and GenerateQ's
// rows are the conjugate transpose of QR's GenerateQ columns.
//
// Each routine conjugate-transposes into the handle's lqbuf (a tiled LDS
// kernel, one read and one write of the matrix), runs the QR path, and
// writes back.  The reference's blocking (row blocks of nb = 32, LQ.hpp:
// 163-183) only fixes rounding order; results agree to the QR tolerances.
#include "rnpx_internal.hpp"

#include <algorithm>

namespace rnpx {

// LQ::Factor (LQ.hpp:126-190).
template <typename T>
int gelqf_dev(Handle* h, int m, int n, T* A, long lda, T* tau) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (lda < std::max(1, m)) return -5;
  if (std::min(m, n) == 0) return RNPX_OK;
  RNPX_TRY(h->lqbuf.ensure(sizeof(T) * (size_t)m * n, h));
  T* B = (T*)h->lqbuf.p;  // n x m, ld n
  RNPX_TRY(launch_conj_transpose<T>(m, n, A, lda, B, n, h->stream));
  int s = geqrf_dev<T>(h, n, m, B, n, tau);
  if (s) return s;
  RNPX_TRY(launch_conj_transpose<T>(n, m, B, n, A, lda, h->stream));
  return RNPX_OK;
}

// LQ::MultQ (LQ.hpp:309-405): c <- op(Q) c or c op(Q); the k reflectors are
// the rows of A (k x nq).
template <typename T>
int ormlq_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
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
  RNPX_TRY(h->lqbuf.ensure(sizeof(T) * (size_t)nq * k, h));
  T* V = (T*)h->lqbuf.p;  // nq x k, ld nq: the reflectors as columns
  RNPX_TRY(launch_conj_transpose<T>(k, nq, A, lda, V, nq, h->stream));
  return ormqr_dev<T>(h, side, trans == 'N' ? 'C' : 'N', m, n, k, V, nq, tau, C, ldc);
}

// LQ::GenerateQ (LQ.hpp:491-559): a (m x n, m <= n) <- the first m rows of Q.
template <typename T>
int orglq_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau) {
  if (m < 0) return -1;
  if (n < m) return -2;
  if (k < 0 || k > m) return -3;
  if (lda < std::max(1, m)) return -5;
  if (m == 0) return RNPX_OK;
  RNPX_TRY(h->lqbuf.ensure(sizeof(T) * (size_t)m * n, h));
  T* B = (T*)h->lqbuf.p;  // n x m, ld n
  RNPX_TRY(launch_conj_transpose<T>(m, n, A, lda, B, n, h->stream));
  int s = orgqr_dev<T>(h, n, m, k, B, n, tau);
  if (s) return s;
  RNPX_TRY(launch_conj_transpose<T>(n, m, B, n, A, lda, h->stream));
  return RNPX_OK;
}

template int gelqf_dev<double>(Handle*, int, int, double*, long, double*);
template int gelqf_dev<zd>(Handle*, int, int, zd*, long, zd*);
template int ormlq_dev<double>(Handle*, char, char, int, int, int, const double*, long, const double*, double*, long);
template int ormlq_dev<zd>(Handle*, char, char, int, int, int, const zd*, long, const zd*, zd*, long);
template int orglq_dev<double>(Handle*, int, int, int, double*, long, const double*);
template int orglq_dev<zd>(Handle*, int, int, int, zd*, long, const zd*);

}  // namespace rnpx
