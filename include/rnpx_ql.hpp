// rnpx_ql.hpp -- RNP's templated QL call surface (namespace RNP::LA::QL),
// forwarding to the gfx950 library through include/rnpx.h.
//
// A caller of RNP/LA/QL.hpp switches by including this header instead; the
// template signatures are the reference's:
//   Factor    RNP/LA/QL.hpp:126-128   (dgeqlf / zgeqlf)
//   MultQ     RNP/LA/QL.hpp:335-339   (dormql / zunmql)
//   GenerateQ RNP/LA/QL.hpp:501-504   (dorgql / zungql)
//   Tuning    RNP/LA/QL.hpp:41-51     (block-size hooks; they size the
//                                      workspace query as in the reference;
//                                      GenerateQ's query uses QR::Tuning as
//                                      QL.hpp:508 does)
// Conventions, error behaviour and element types are those of rnpx_qr.hpp.
#ifndef RNPX_QL_HPP
#define RNPX_QL_HPP

#include "rnpx_qr.hpp"

namespace RNP {
namespace LA {
namespace QL {

template <typename T>
struct Tuning {
  static size_t factor_block_size_opt(size_t, size_t) { return 32; }
  static size_t factor_block_size_min(size_t, size_t) { return 2; }
  static size_t factor_crossover_size(size_t, size_t) { return 128; }
  static size_t multQ_block_size_opt(const char*, const char*, size_t, size_t, size_t) { return 32; }
  static size_t multQ_block_size_min(const char*, const char*, size_t, size_t, size_t) { return 2; }
};

namespace detail {

inline int geqlf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dgeqlf(m, n, a, lda, tau, lw, &w);
}
inline int geqlf(size_t m, size_t n, std::complex<double>* a, size_t lda, std::complex<double>* tau, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zgeqlf(m, n, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<rnpx_zcomplex*>(tau), lw, &w);
}
inline int ormql(const char* s, const char* t, size_t m, size_t n, size_t k, const double* a, size_t lda,
                 const double* tau, double* c, size_t ldc, size_t* lw) {
  double w = 0;
  return rnpx_dormql(s, t, m, n, k, a, lda, tau, c, ldc, lw, &w);
}
inline int ormql(const char* s, const char* t, size_t m, size_t n, size_t k, const std::complex<double>* a,
                 size_t lda, const std::complex<double>* tau, std::complex<double>* c, size_t ldc, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zunmql(s, t, m, n, k, reinterpret_cast<const rnpx_zcomplex*>(a), lda,
                     reinterpret_cast<const rnpx_zcomplex*>(tau), reinterpret_cast<rnpx_zcomplex*>(c), ldc, lw, &w);
}
inline int orgql(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dorgql(m, n, k, a, lda, tau, lw, &w);
}
inline int orgql(size_t m, size_t n, size_t k, std::complex<double>* a, size_t lda, const std::complex<double>* tau,
                 size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zungql(m, n, k, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<const rnpx_zcomplex*>(tau),
                     lw, &w);
}

}  // namespace detail

// QL::Factor -- RNP/LA/QL.hpp:126-205.  On exit a holds L in the lower
// triangle ending at its bottom-right corner and the reflectors above it.
template <typename T>
void Factor(size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork, T* work) {
  namespace qd = ::RNP::LA::QR::detail;
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (!RNPX_QR_CHECK(NULL != lwork)) return;
  const size_t k = (m < n ? m : n);
  if (0 == k) return;
  if (NULL == work || 0 == *lwork) {
    *lwork = QL::Tuning<T>::factor_block_size_opt(m, n) * n;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("QL::Factor", detail::geqlf(m, n, reinterpret_cast<W*>(a), lda, reinterpret_cast<W*>(tau), &lw));
    return;
  }
  qd::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::geqlf(m, n, A.data(), m, TAU.data(), &lw);
  qd::report("QL::Factor", st);
  if (st == RNPX_OK) {
    A.back(a, m, n, lda);
    TAU.back(tau, k, 1, k);
  }
}

// QL::MultQ -- RNP/LA/QL.hpp:335-430.  c <- op(Q) c (side "L") or c op(Q)
// (side "R"), op(Q) = Q for trans "N" and Q^H for trans "C"; a holds the k
// reflector columns.
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
  const size_t nb = QL::Tuning<T>::multQ_block_size_opt(side, trans, m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = nb * nw + nb * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("QL::MultQ", detail::ormql(side, trans, m, n, k, reinterpret_cast<const W*>(a), lda,
                                          reinterpret_cast<const W*>(tau), reinterpret_cast<W*>(c), ldc, &lw));
    return;
  }
  qd::Widened<T> A(a, nq, k, lda), TAU(tau, k, 1, k), C(c, m, n, ldc);
  int st = detail::ormql(side, trans, m, n, k, A.data(), nq, TAU.data(), C.data(), m, &lw);
  qd::report("QL::MultQ", st);
  if (st == RNPX_OK) C.back(c, m, n, ldc);
}

// QL::GenerateQ -- RNP/LA/QL.hpp:501-565.  Overwrites the m-by-n a (m >= n,
// the reflectors in its last k columns) with the last n columns of Q.
template <typename T>
void GenerateQ(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, T* work) {
  namespace qd = ::RNP::LA::QR::detail;
  if (!RNPX_QR_CHECK(k <= n)) return;
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (0 == n) return;
  const size_t nb = QR::Tuning<T>::genQ_block_size_opt(m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = n * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename qd::Wide<T>::type W;
  if (qd::IsWide<T>::value) {
    qd::report("QL::GenerateQ", detail::orgql(m, n, k, reinterpret_cast<W*>(a), lda,
                                              reinterpret_cast<const W*>(tau), &lw));
    return;
  }
  qd::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::orgql(m, n, k, A.data(), m, TAU.data(), &lw);
  qd::report("QL::GenerateQ", st);
  if (st == RNPX_OK) A.back(a, m, n, lda);
}

// Status of the last call on this thread (shared with RNP::LA::QR).
inline int LastStatus() { return ::RNP::LA::QR::detail::last_status(); }

}  // namespace QL
}  // namespace LA
}  // namespace RNP

#endif  // RNPX_QL_HPP
