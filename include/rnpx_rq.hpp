// rnpx_rq.hpp -- RNP's templated RQ call surface (namespace RNP::LA::RQ),
// forwarding to the gfx950 library through include/rnpx.h.
//
// A caller of RNP/LA/RQ.hpp switches by including this header instead; the
// template signatures are the reference's:
//   Factor    RNP/LA/RQ.hpp:131-133   (dgerqf / zgerqf)
//   MultQ     RNP/LA/RQ.hpp:344-348   (dormrq / zunmrq)
//   GenerateQ RNP/LA/RQ.hpp:519-522   (dorgrq / zungrq)
//   Tuning    RNP/LA/RQ.hpp:41-51     (block-size hooks; they size the
//                                      workspace query as in the reference;
//                                      GenerateQ's query uses QR::Tuning as
//                                      RQ.hpp:528 does)
// Conventions, error behaviour and element types are those of rnpx_qr.hpp.
#ifndef RNPX_RQ_HPP
#define RNPX_RQ_HPP

#include "rnpx_qr.hpp"

namespace RNP {
namespace LA {
namespace RQ {

template <typename T>
struct Tuning {
  static size_t factor_block_size_opt(size_t, size_t) { return 32; }
  static size_t factor_block_size_min(size_t, size_t) { return 2; }
  static size_t factor_crossover_size(size_t, size_t) { return 128; }
  static size_t multQ_block_size_opt(const char*, const char*, size_t, size_t, size_t) { return 32; }
  static size_t multQ_block_size_min(const char*, const char*, size_t, size_t, size_t) { return 2; }
};

namespace detail {

inline int gerqf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dgerqf(m, n, a, lda, tau, lw, &w);
}
inline int gerqf(size_t m, size_t n, std::complex<double>* a, size_t lda, std::complex<double>* tau, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zgerqf(m, n, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<rnpx_zcomplex*>(tau), lw, &w);
}
inline int ormrq(const char* s, const char* t, size_t m, size_t n, size_t k, const double* a, size_t lda,
                 const double* tau, double* c, size_t ldc, size_t* lw) {
  double w = 0;
  return rnpx_dormrq(s, t, m, n, k, a, lda, tau, c, ldc, lw, &w);
}
inline int ormrq(const char* s, const char* t, size_t m, size_t n, size_t k, const std::complex<double>* a,
                 size_t lda, const std::complex<double>* tau, std::complex<double>* c, size_t ldc, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zunmrq(s, t, m, n, k, reinterpret_cast<const rnpx_zcomplex*>(a), lda,
                     reinterpret_cast<const rnpx_zcomplex*>(tau), reinterpret_cast<rnpx_zcomplex*>(c), ldc, lw, &w);
}
inline int orgrq(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dorgrq(m, n, k, a, lda, tau, lw, &w);
}
inline int orgrq(size_t m, size_t n, size_t k, std::complex<double>* a, size_t lda, const std::complex<double>* tau,
                 size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zungrq(m, n, k, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<const rnpx_zcomplex*>(tau),
                     lw, &w);
}

}  // namespace detail

// RQ::Factor -- RNP/LA/RQ.hpp:131-205.  On exit a holds R in the upper
// triangle ending at its bottom-right corner and the conjugated reflectors
// in the rows left of it.
template <typename T>
void Factor(size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork, T* work) {
  namespace qd = ::RNP::LA::QR::detail;
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (!RNPX_QR_CHECK(NULL != lwork)) return;
  // The reference asserts this (print and continue) although its own query
  // returns nb n: a short workspace still factors, as there.
  RNPX_QR_CHECK(0 == *lwork || *lwork >= m);
  const size_t k = (m < n ? m : n);
  if (0 == k) return;
  if (NULL == work || 0 == *lwork) {
    *lwork = RQ::Tuning<T>::factor_block_size_opt(m, n) * m;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("RQ::Factor", detail::gerqf(m, n, reinterpret_cast<W*>(a), lda, reinterpret_cast<W*>(tau), &lw));
    return;
  }
  qd::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::gerqf(m, n, A.data(), m, TAU.data(), &lw);
  qd::report("RQ::Factor", st);
  if (st == RNPX_OK) {
    A.back(a, m, n, lda);
    TAU.back(tau, k, 1, k);
  }
}

// RQ::MultQ -- RNP/LA/RQ.hpp:344-440.  c <- op(Q) c (side "L") or c op(Q)
// (side "R"), op(Q) = Q for trans "N" and Q^H for trans "C"; a holds the k
// reflector rows.
template <typename T>
void MultQ(const char* side, const char* trans, size_t m, size_t n, size_t k, const T* a, size_t lda, const T* tau,
           T* c, size_t ldc, size_t* lwork, T* work) {
  namespace qd = ::RNP::LA::QR::detail;
  if (!RNPX_QR_CHECK(ldc >= m)) return;
  if (0 == m || 0 == n || 0 == k) return;
  const bool left = ('L' == side[0]);
  if (!RNPX_QR_CHECK((left && k <= m) || (!left && k <= n))) return;
  if (!RNPX_QR_CHECK(lda >= k)) return;
  const size_t nq = (left ? m : n), nw = (left ? n : m);
  const size_t nb = RQ::Tuning<T>::multQ_block_size_opt(side, trans, m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = nb * nw + nb * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("RQ::MultQ", detail::ormrq(side, trans, m, n, k, reinterpret_cast<const W*>(a), lda,
                                          reinterpret_cast<const W*>(tau), reinterpret_cast<W*>(c), ldc, &lw));
    return;
  }
  qd::Widened<T> A(a, k, nq, lda), TAU(tau, k, 1, k), C(c, m, n, ldc);
  int st = detail::ormrq(side, trans, m, n, k, A.data(), k, TAU.data(), C.data(), m, &lw);
  qd::report("RQ::MultQ", st);
  if (st == RNPX_OK) C.back(c, m, n, ldc);
}

// RQ::GenerateQ -- RNP/LA/RQ.hpp:519-580.  Overwrites the m-by-n a (m <= n,
// the reflectors in its last k rows) with the last m rows of Q.
template <typename T>
void GenerateQ(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, T* work) {
  namespace qd = ::RNP::LA::QR::detail;
  if (!RNPX_QR_CHECK(n >= m)) return;
  if (!RNPX_QR_CHECK(k <= m)) return;
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (0 == m) return;
  const size_t nb = QR::Tuning<T>::genQ_block_size_opt(m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = m * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("RQ::GenerateQ", detail::orgrq(m, n, k, reinterpret_cast<W*>(a), lda,
                                              reinterpret_cast<const W*>(tau), &lw));
    return;
  }
  qd::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::orgrq(m, n, k, A.data(), m, TAU.data(), &lw);
  qd::report("RQ::GenerateQ", st);
  if (st == RNPX_OK) A.back(a, m, n, lda);
}

// Status of the last call on this thread (shared with RNP::LA::QR).
inline int LastStatus() { return ::RNP::LA::QR::detail::last_status(); }

}  // namespace RQ
}  // namespace LA
}  // namespace RNP

#endif  // RNPX_RQ_HPP
