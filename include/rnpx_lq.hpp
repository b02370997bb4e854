// rnpx_lq.hpp -- RNP's templated LQ call surface (namespace RNP::LA::LQ),
// forwarding to the gfx950 library through include/rnpx.h.
//
// A caller of RNP/LA/LQ.hpp switches by including this header instead; the
// template signatures are the reference's:
//   Factor    RNP/LA/LQ.hpp:126-127   (dgelqf / zgelqf)
//   MultQ     RNP/LA/LQ.hpp:309-317   (dormlq / zunmlq)
//   GenerateQ RNP/LA/LQ.hpp:491-495   (dorglq / zunglq)
//   Tuning    RNP/LA/LQ.hpp:41-51     (block-size hooks; they size the
//                                      workspace query as in the reference)
// Conventions, error behaviour and element types are those of rnpx_qr.hpp.
//
// One documented difference: for k <= nb the reference's LQ::MultQ with
// side "L" runs MultQ_unblocked, whose Reflector::Apply branch
// (Reflector.hpp:258-263) updates only column 0 of C; this surface returns
// the product LQ.hpp:195-196 documents (Q C or Q^H C) for every k.
#ifndef RNPX_LQ_HPP
#define RNPX_LQ_HPP

#include "rnpx_qr.hpp"

namespace RNP {
namespace LA {
namespace LQ {

template <typename T>
struct Tuning {
  static size_t factor_block_size_opt(size_t, size_t) { return 32; }
  static size_t factor_block_size_min(size_t, size_t) { return 2; }
  static size_t factor_crossover_size(size_t, size_t) { return 128; }
  static size_t multQ_block_size_opt(const char*, const char*, size_t, size_t, size_t) { return 32; }
  static size_t multQ_block_size_min(const char*, const char*, size_t, size_t, size_t) { return 2; }
  static size_t genQ_block_size_opt(size_t, size_t, size_t) { return 32; }
  static size_t genQ_block_size_min(size_t, size_t, size_t) { return 2; }
  static size_t genQ_crossover_size(size_t, size_t, size_t) { return 128; }
};

namespace detail {

inline int gelqf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dgelqf(m, n, a, lda, tau, lw, &w);
}
inline int gelqf(size_t m, size_t n, std::complex<double>* a, size_t lda, std::complex<double>* tau, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zgelqf(m, n, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<rnpx_zcomplex*>(tau), lw, &w);
}
inline int ormlq(const char* s, const char* t, size_t m, size_t n, size_t k, const double* a, size_t lda,
                 const double* tau, double* c, size_t ldc, size_t* lw) {
  double w = 0;
  return rnpx_dormlq(s, t, m, n, k, a, lda, tau, c, ldc, lw, &w);
}
inline int ormlq(const char* s, const char* t, size_t m, size_t n, size_t k, const std::complex<double>* a,
                 size_t lda, const std::complex<double>* tau, std::complex<double>* c, size_t ldc, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zunmlq(s, t, m, n, k, reinterpret_cast<const rnpx_zcomplex*>(a), lda,
                     reinterpret_cast<const rnpx_zcomplex*>(tau), reinterpret_cast<rnpx_zcomplex*>(c), ldc, lw, &w);
}
inline int orglq(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dorglq(m, n, k, a, lda, tau, lw, &w);
}
inline int orglq(size_t m, size_t n, size_t k, std::complex<double>* a, size_t lda, const std::complex<double>* tau,
                 size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zunglq(m, n, k, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<const rnpx_zcomplex*>(tau),
                     lw, &w);
}

}  // namespace detail

// LQ::Factor -- RNP/LA/LQ.hpp:126-190.  On exit a holds L on and below the
// diagonal and the conjugated reflectors in the rows above it.
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
    *lwork = LQ::Tuning<T>::factor_block_size_opt(m, n) * n;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("LQ::Factor", detail::gelqf(m, n, reinterpret_cast<W*>(a), lda, reinterpret_cast<W*>(tau), &lw));
    return;
  }
  qd::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::gelqf(m, n, A.data(), m, TAU.data(), &lw);
  qd::report("LQ::Factor", st);
  if (st == RNPX_OK) {
    A.back(a, m, n, lda);
    TAU.back(tau, k, 1, k);
  }
}

// LQ::MultQ -- RNP/LA/LQ.hpp:309-405.  c <- op(Q) c (side "L") or c op(Q)
// (side "R"), op(Q) = Q for trans "N" and Q^H for trans "C".
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
  const size_t nb = LQ::Tuning<T>::multQ_block_size_opt(side, trans, m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = nb * nw + nb * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("LQ::MultQ", detail::ormlq(side, trans, m, n, k, reinterpret_cast<const W*>(a), lda,
                                          reinterpret_cast<const W*>(tau), reinterpret_cast<W*>(c), ldc, &lw));
    return;
  }
  qd::Widened<T> A(a, k, nq, lda), TAU(tau, k, 1, k), C(c, m, n, ldc);
  int st = detail::ormlq(side, trans, m, n, k, A.data(), k, TAU.data(), C.data(), m, &lw);
  qd::report("LQ::MultQ", st);
  if (st == RNPX_OK) C.back(c, m, n, ldc);
}

// LQ::GenerateQ -- RNP/LA/LQ.hpp:491-559.  Overwrites the m-by-n a (m <= n)
// with the first m rows of Q.
template <typename T>
void GenerateQ(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, T* work) {
  namespace qd = ::RNP::LA::QR::detail;
  if (!RNPX_QR_CHECK(k <= m)) return;
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (0 == m) return;
  const size_t nb = LQ::Tuning<T>::genQ_block_size_opt(m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = n * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("LQ::GenerateQ", detail::orglq(m, n, k, reinterpret_cast<W*>(a), lda,
                                              reinterpret_cast<const W*>(tau), &lw));
    return;
  }
  qd::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::orglq(m, n, k, A.data(), m, TAU.data(), &lw);
  qd::report("LQ::GenerateQ", st);
  if (st == RNPX_OK) A.back(a, m, n, lda);
}

// Status of the last call on this thread (shared with RNP::LA::QR).
inline int LastStatus() { return ::RNP::LA::QR::detail::last_status(); }

}  // namespace LQ
}  // namespace LA
}  // namespace RNP

#endif  // RNPX_LQ_HPP
