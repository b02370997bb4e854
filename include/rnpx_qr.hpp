// rnpx_qr.hpp -- RNP's templated QR call surface (namespace RNP::LA::QR),
// forwarding to the gfx950 library through include/rnpx.h.
//
// A caller of RNP/LA/QR.hpp switches by including this header instead; the
// template signatures are the reference's:
//   Factor    RNP/LA/QR.hpp:129-131   (dgeqrf / zgeqrf)
//   MultQ     RNP/LA/QR.hpp:337-341   (dormqr / zunmqr)
//   GenerateQ RNP/LA/QR.hpp:509-512   (dorgqr / zungqr)
//   Tuning    RNP/LA/QR.hpp:41-51     (block-size hooks; they size the
//                                      workspace query exactly as in the
//                                      reference, the GPU picks its own blocking)
// Conventions are unchanged: column-major, 0-based, size_t, side/trans read
// from their first character, workspace query when work == NULL or
// *lwork == 0.  Pointers may be host or device memory.
//
// Error behaviour follows the reference's RNPAssert (RNP/Debug.hpp:9-17):
// a failed argument check prints "Assertion failed in file:line: cond" to
// stderr and the function returns without computing (the reference would
// carry on into undefined behaviour).  A GPU-side failure (no device, out of
// memory, ...) prints "rnpx: <function>: status <code>" and returns; the code
// is kept in RNP::LA::QR::LastStatus() for callers that want to check.
//
// float and std::complex<float> are computed in double on the GPU (inputs
// widened, results rounded back).
#ifndef RNPX_QR_HPP
#define RNPX_QR_HPP

#include <complex>
#include <cstddef>
#include <cstdio>
#include <vector>

#include "rnpx.h"

namespace RNP {
namespace LA {
namespace QR {

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

inline int& last_status() {
  static thread_local int s = 0;
  return s;
}

inline bool check(bool cond, const char* file, int line, const char* str) {
  if (!cond) std::fprintf(stderr, "Assertion failed in %s:%d: %s\n", file, line, str);
  return cond;
}

inline void report(const char* fn, int st) {
  last_status() = st;
  if (st != RNPX_OK) std::fprintf(stderr, "rnpx: %s: status %d\n", fn, st);
}

#define RNPX_QR_CHECK(COND) ::RNP::LA::QR::detail::check((COND), __FILE__, __LINE__, #COND)

// The two types the GPU computes in, and how other element types map to them.
template <typename T> struct Wide;
template <> struct Wide<double> { typedef double type; };
template <> struct Wide<float> { typedef double type; };
template <> struct Wide<std::complex<double> > { typedef std::complex<double> type; };
template <> struct Wide<std::complex<float> > { typedef std::complex<double> type; };

inline int geqrf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dgeqrf(m, n, a, lda, tau, lw, &w);
}
inline int geqrf(size_t m, size_t n, std::complex<double>* a, size_t lda, std::complex<double>* tau, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zgeqrf(m, n, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<rnpx_zcomplex*>(tau), lw, &w);
}
inline int ormqr(const char* s, const char* t, size_t m, size_t n, size_t k, const double* a, size_t lda,
                 const double* tau, double* c, size_t ldc, size_t* lw) {
  double w = 0;
  return rnpx_dormqr(s, t, m, n, k, a, lda, tau, c, ldc, lw, &w);
}
inline int ormqr(const char* s, const char* t, size_t m, size_t n, size_t k, const std::complex<double>* a,
                 size_t lda, const std::complex<double>* tau, std::complex<double>* c, size_t ldc, size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zunmqr(s, t, m, n, k, reinterpret_cast<const rnpx_zcomplex*>(a), lda,
                     reinterpret_cast<const rnpx_zcomplex*>(tau), reinterpret_cast<rnpx_zcomplex*>(c), ldc, lw, &w);
}
inline int orgqr(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lw) {
  double w = 0;
  return rnpx_dorgqr(m, n, k, a, lda, tau, lw, &w);
}
inline int orgqr(size_t m, size_t n, size_t k, std::complex<double>* a, size_t lda, const std::complex<double>* tau,
                 size_t* lw) {
  rnpx_zcomplex w = {0, 0};
  return rnpx_zungqr(m, n, k, reinterpret_cast<rnpx_zcomplex*>(a), lda, reinterpret_cast<const rnpx_zcomplex*>(tau),
                     lw, &w);
}

// Host staging of a narrower element type into the wide one and back.
template <typename T>
struct Widened {
  typedef typename Wide<T>::type W;
  std::vector<W> buf;
  Widened(const T* p, size_t rows, size_t cols, size_t ld) : buf(rows * cols) {
    for (size_t j = 0; j < cols; ++j)
      for (size_t i = 0; i < rows; ++i) buf[i + j * rows] = W(p[i + j * ld]);
  }
  void back(T* p, size_t rows, size_t cols, size_t ld) const {
    for (size_t j = 0; j < cols; ++j)
      for (size_t i = 0; i < rows; ++i) p[i + j * ld] = T(buf[i + j * rows]);
  }
  W* data() { return buf.empty() ? 0 : &buf[0]; }
};

template <typename T>
struct IsWide {
  static const bool value = false;
};
template <> struct IsWide<double> { static const bool value = true; };
template <> struct IsWide<std::complex<double> > { static const bool value = true; };

}  // namespace detail

// QR::Factor -- RNP/LA/QR.hpp:129-197.  On exit a holds R on and above the
// diagonal and the Householder vectors below it; tau holds min(m,n) scalars.
template <typename T>
void Factor(size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork, T* work) {
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (!RNPX_QR_CHECK(NULL != lwork)) return;
  // The reference asserts this (print and continue) although its own query
  // returns nb n: a short workspace still factors, as there.
  RNPX_QR_CHECK(0 == *lwork || *lwork >= m);
  const size_t k = (m < n ? m : n);
  if (0 == k) return;
  if (NULL == work || 0 == *lwork) {
    *lwork = QR::Tuning<T>::factor_block_size_opt(m, n) * n;
    return;
  }
  size_t lw = *lwork;
  if (detail::IsWide<T>::value) {
    detail::report("Factor", detail::geqrf(m, n, reinterpret_cast<typename detail::Wide<T>::type*>(a), lda,
                                           reinterpret_cast<typename detail::Wide<T>::type*>(tau), &lw));
    return;
  }
  detail::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::geqrf(m, n, A.data(), m, TAU.data(), &lw);
  detail::report("Factor", st);
  if (st == RNPX_OK) {
    A.back(a, m, n, lda);
    TAU.back(tau, k, 1, k);
  }
}

// QR::MultQ -- RNP/LA/QR.hpp:337-429.  c <- op(Q) c (side "L") or c op(Q)
// (side "R"), op(Q) = Q for trans "N" and Q^H for trans "C".
template <typename T>
void MultQ(const char* side, const char* trans, size_t m, size_t n, size_t k, const T* a, size_t lda, const T* tau,
           T* c, size_t ldc, size_t* lwork, T* work) {
  if (!RNPX_QR_CHECK(ldc >= m)) return;
  if (0 == m || 0 == n || 0 == k) return;
  const bool left = ('L' == side[0]);
  if (!RNPX_QR_CHECK((left && k <= m) || (!left && k <= n))) return;
  if (!RNPX_QR_CHECK(lda >= k)) return;
  const size_t nq = (left ? m : n), nw = (left ? n : m);
  const size_t nb = QR::Tuning<T>::multQ_block_size_opt(side, trans, m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = nb * nw + nb * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename detail::Wide<T>::type W;
  if (detail::IsWide<T>::value) {
    detail::report("MultQ", detail::ormqr(side, trans, m, n, k, reinterpret_cast<const W*>(a), lda,
                                          reinterpret_cast<const W*>(tau), reinterpret_cast<W*>(c), ldc, &lw));
    return;
  }
  detail::Widened<T> A(a, nq, k, lda), TAU(tau, k, 1, k), C(c, m, n, ldc);
  int st = detail::ormqr(side, trans, m, n, k, A.data(), nq, TAU.data(), C.data(), m, &lw);
  detail::report("MultQ", st);
  if (st == RNPX_OK) C.back(c, m, n, ldc);
}

// QR::GenerateQ -- RNP/LA/QR.hpp:509-570.  Overwrites the m-by-n a with the
// first n columns of Q = H_0 H_1 ... H_{k-1}.
template <typename T>
void GenerateQ(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, T* work) {
  if (!RNPX_QR_CHECK(k <= n)) return;
  if (!RNPX_QR_CHECK(lda >= m)) return;
  if (0 == n) return;
  const size_t nb = QR::Tuning<T>::genQ_block_size_opt(m, n, k);
  if (0 == *lwork || NULL == work) {
    *lwork = n * nb;
    return;
  }
  size_t lw = *lwork;
  typedef typename detail::Wide<T>::type W;
  if (detail::IsWide<T>::value) {
    detail::report("GenerateQ", detail::orgqr(m, n, k, reinterpret_cast<W*>(a), lda,
                                              reinterpret_cast<const W*>(tau), &lw));
    return;
  }
  detail::Widened<T> A(a, m, n, lda), TAU(tau, k, 1, k);
  int st = detail::orgqr(m, n, k, A.data(), m, TAU.data(), &lw);
  detail::report("GenerateQ", st);
  if (st == RNPX_OK) A.back(a, m, n, lda);
}

// Status of the last Factor / MultQ / GenerateQ call on this thread.
inline int LastStatus() { return detail::last_status(); }

}  // namespace QR
}  // namespace LA
}  // namespace RNP

#endif  // RNPX_QR_HPP
