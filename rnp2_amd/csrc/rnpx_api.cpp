// rnpx_api.cpp -- the extern "C" surface of include/rnpx.h.
//
// Drop-in entry points keep RNP's argument lists and conventions
// (RNP/LA/QR.hpp:129-131, 337-341, 509-512): the workspace query returns
// exactly what the reference returns (its Tuning default nb = 32), zero
// sizes return before the query as in the reference, and each matrix
// argument may be host memory (staged through the handle's device buffers)
// or device memory (used in place).  Invalid arguments return -i instead of
// the reference's print-and-continue assert; the C++ surface in
// include/rnpx_qr.hpp prints the reference's message for them.
#include "rnpx_internal.hpp"

#include <climits>
#include <cstdio>
#include <cstring>

using namespace rnpx;

namespace {

constexpr size_t kRefNb = 32;  // QR::Tuning<T> defaults, RNP/LA/QR.hpp:43-50

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

bool fits_int(size_t v) { return v <= (size_t)INT_MAX; }

// One matrix argument of a drop-in call: device pointer used in place, or a
// host array staged into buf (rows x cols, leading dimension ld on the host,
// rows on the device).
template <typename T>
struct Arg {
  T* host = nullptr;
  T* dev = nullptr;
  long ld = 0;
  size_t rows = 0, cols = 0;
  bool staged = false;
  int in(Handle* h, Buffer& buf, T* p, size_t r, size_t c, size_t ldp, bool copy_in) {
    rows = r;
    cols = c;
    if (is_device_ptr(p)) {
      dev = p;
      ld = (long)ldp;
      return RNPX_OK;
    }
    host = p;
    staged = true;
    ld = (long)(r ? r : 1);
    hipError_t e = buf.ensure(sizeof(T) * (size_t)ld * (c ? c : 1), h);
    if (e != hipSuccess) return hip_status(e);
    dev = (T*)buf.p;
    if (copy_in && r && c) {
      e = hipMemcpy2DAsync(dev, sizeof(T) * ld, host, sizeof(T) * ldp, sizeof(T) * r, c, hipMemcpyHostToDevice,
                           h->stream);
      if (e != hipSuccess) return hip_status(e);
    }
    return RNPX_OK;
  }
  int out(Handle* h, size_t ldp) {
    if (!staged || !rows || !cols) return RNPX_OK;
    hipError_t e = hipMemcpy2DAsync(host, sizeof(T) * ldp, dev, sizeof(T) * ld, sizeof(T) * rows, cols,
                                    hipMemcpyDeviceToHost, h->stream);
    return hip_status(e);
  }
};

template <typename T>
int dropin_geqrf(size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork) {
  if (lda < m || lda == 0) return -4;
  if (!lwork) return -6;
  // 0 < *lwork < m is accepted: the reference asserts *lwork >= m (print and
  // continue) while its own workspace query returns nb n, and factors anyway.
  const size_t k = m < n ? m : n;
  if (k == 0) return RNPX_OK;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, tau, k, 1, k, false))) return st;
  // the workspace decides the reference's block width, which the path for
  // matrices holding Inf or NaN follows
  h->exact_lwork = *lwork;
  st = geqrf_dev<T>(h, (int)m, (int)n, A.dev, A.ld, TAU.dev);
  h->exact_lwork = 0;
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  if ((st = TAU.out(h, k))) return st;
  return sync_handle(h);
}

// Reflector::LastNonzeroColumnLimit (which 0) / LastNonzeroRowLimit (which 1),
// RNP/LA/Reflector.hpp:36-75.
template <typename T>
int dropin_last_nonzero(int which, size_t m, size_t n, const T* a, size_t lda, size_t* limit) {
  if (!limit) return -5;
  *limit = 0;
  if (m == 0 || n == 0) return RNPX_OK;
  if (lda < m) return -4;
  if (!a) return -3;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), m, n, lda, true))) return st;
  int out = 0;
  if ((st = last_nonzero_limit<T>(h, which, (int)m, (int)n, A.dev, A.ld, &out))) return st;
  *limit = (size_t)out;
  return RNPX_OK;
}

template <typename T>
int dropin_ormqr(const char* side, const char* trans, size_t m, size_t n, size_t k, const T* a, size_t lda,
                 const T* tau, T* c, size_t ldc, size_t* lwork, bool query) {
  if (!side || !trans) return !side ? -1 : -2;
  if (ldc < m || ldc == 0) return -10;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  const bool left = ('L' == side[0]);
  if (side[0] != 'L' && side[0] != 'R') return -1;
  if (trans[0] != 'N' && trans[0] != 'C' && trans[0] != 'T') return -2;
  if ((left && k > m) || (!left && k > n)) return -5;
  const size_t nq = left ? m : n, nw = left ? n : m;
  if (lda < k || lda < nq) return -7;
  if (!lwork) return -11;
  if (query) {
    *lwork = kRefNb * nw + kRefNb * kRefNb;
    return RNPX_OK;
  }
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU, C;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), nq, k, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), k, 1, k, true))) return st;
  if ((st = C.in(h, h->stage3, c, m, n, ldc, true))) return st;
  st = ormqr_dev<T>(h, side[0], trans[0], (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev, C.dev, C.ld);
  if (st) return st;
  if ((st = C.out(h, ldc))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_orgqr(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, bool query) {
  if (k > n) return -3;
  if (lda < m || lda == 0) return -5;
  if (n == 0) return RNPX_OK;
  if (!lwork) return -7;
  if (query) {
    *lwork = n * kRefNb;
    return RNPX_OK;
  }
  if (n > m) return -2;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), k ? k : 1, 1, k ? k : 1, k > 0))) return st;
  st = orgqr_dev<T>(h, (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  return sync_handle(h);
}

// LQ drop-ins (RNP/LA/LQ.hpp): argument checks and workspace queries as the
// reference's LQ::Factor (LQ.hpp:126-138: nb n), LQ::MultQ (LQ.hpp:318-333:
// nb nw + nb^2) and LQ::GenerateQ (LQ.hpp:496-504: n nb).
template <typename T>
int dropin_gelqf(size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork) {
  if (lda < m || lda == 0) return -4;
  if (!lwork) return -6;
  // 0 < *lwork < m is accepted: the reference asserts *lwork >= m (print and
  // continue) while its own workspace query returns nb n, and factors anyway.
  const size_t k = m < n ? m : n;
  if (k == 0) return RNPX_OK;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, tau, k, 1, k, false))) return st;
  st = gelqf_dev<T>(h, (int)m, (int)n, A.dev, A.ld, TAU.dev);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  if ((st = TAU.out(h, k))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_ormlq(const char* side, const char* trans, size_t m, size_t n, size_t k, const T* a, size_t lda,
                 const T* tau, T* c, size_t ldc, size_t* lwork, bool query) {
  if (!side || !trans) return !side ? -1 : -2;
  if (ldc < m || ldc == 0) return -10;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  const bool left = ('L' == side[0]);
  if (side[0] != 'L' && side[0] != 'R') return -1;
  if (trans[0] != 'N' && trans[0] != 'C' && trans[0] != 'T') return -2;
  if ((left && k > m) || (!left && k > n)) return -5;
  const size_t nq = left ? m : n, nw = left ? n : m;
  if (lda < k) return -7;
  if (!lwork) return -11;
  if (query) {
    *lwork = kRefNb * nw + kRefNb * kRefNb;
    return RNPX_OK;
  }
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU, C;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), k, nq, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), k, 1, k, true))) return st;
  if ((st = C.in(h, h->stage3, c, m, n, ldc, true))) return st;
  st = ormlq_dev<T>(h, side[0], trans[0], (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev, C.dev, C.ld);
  if (st) return st;
  if ((st = C.out(h, ldc))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_orglq(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, bool query) {
  if (k > m) return -3;
  if (lda < m || lda == 0) return -5;
  if (m == 0) return RNPX_OK;
  if (!lwork) return -7;
  if (query) {
    *lwork = n * kRefNb;
    return RNPX_OK;
  }
  if (n < m) return -2;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), k ? k : 1, 1, k ? k : 1, k > 0))) return st;
  st = orglq_dev<T>(h, (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  return sync_handle(h);
}

// QL and RQ drop-ins (RNP/LA/QL.hpp, RNP/LA/RQ.hpp): argument checks and
// workspace queries as the reference's QL::Factor (QL.hpp:132-141: nb n),
// RQ::Factor (RQ.hpp:134-144: nb m, lwork 0 or >= m), MultQ (QL.hpp:342-354,
// RQ.hpp:349-362: nb nw + nb^2) and GenerateQ (QL.hpp:505-512: n nb;
// RQ.hpp:523-531: m nb).
enum class Fam { QL, RQ };

template <typename T>
int dropin_gexxf(Fam f, size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork) {
  if (lda < m || lda == 0) return -4;
  if (!lwork) return -6;
  // RQ: 0 < *lwork < m accepted as in QR (the reference prints and continues)
  const size_t k = m < n ? m : n;
  if (k == 0) return RNPX_OK;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, tau, k, 1, k, false))) return st;
  st = f == Fam::QL ? geqlf_dev<T>(h, (int)m, (int)n, A.dev, A.ld, TAU.dev)
                    : gerqf_dev<T>(h, (int)m, (int)n, A.dev, A.ld, TAU.dev);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  if ((st = TAU.out(h, k))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_ormxx(Fam f, const char* side, const char* trans, size_t m, size_t n, size_t k, const T* a, size_t lda,
                 const T* tau, T* c, size_t ldc, size_t* lwork, bool query) {
  if (!side || !trans) return !side ? -1 : -2;
  if (ldc < m || ldc == 0) return -10;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  const bool left = ('L' == side[0]);
  if (side[0] != 'L' && side[0] != 'R') return -1;
  if (trans[0] != 'N' && trans[0] != 'C' && trans[0] != 'T') return -2;
  if ((left && k > m) || (!left && k > n)) return -5;
  const size_t nq = left ? m : n, nw = left ? n : m;
  if (lda < k || (f == Fam::QL && lda < nq)) return -7;
  if (!lwork) return -11;
  if (query) {
    *lwork = kRefNb * nw + kRefNb * kRefNb;
    return RNPX_OK;
  }
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU, C;
  const size_t ar = f == Fam::QL ? nq : k, ac = f == Fam::QL ? k : nq;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), ar, ac, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), k, 1, k, true))) return st;
  if ((st = C.in(h, h->stage3, c, m, n, ldc, true))) return st;
  st = f == Fam::QL
           ? ormql_dev<T>(h, side[0], trans[0], (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev, C.dev, C.ld)
           : ormrq_dev<T>(h, side[0], trans[0], (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev, C.dev, C.ld);
  if (st) return st;
  if ((st = C.out(h, ldc))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_orgxx(Fam f, size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau, size_t* lwork, bool query) {
  const bool ql = f == Fam::QL;
  if (ql ? k > n : k > m) return -3;
  if (lda < m || lda == 0) return -5;
  if (ql ? n == 0 : m == 0) return RNPX_OK;
  if (!lwork) return -7;
  if (query) {
    *lwork = (ql ? n : m) * kRefNb;
    return RNPX_OK;
  }
  if (ql ? n > m : n < m) return -2;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), k ? k : 1, 1, k ? k : 1, k > 0))) return st;
  st = ql ? orgql_dev<T>(h, (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev)
          : orgrq_dev<T>(h, (int)m, (int)n, (int)k, A.dev, A.ld, TAU.dev);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  return sync_handle(h);
}

template <typename T>
int gexxf_entry(Fam f, size_t m, size_t n, T* a, size_t lda, T* tau, size_t* lwork, T* work) {
  if (lwork && (work == nullptr || *lwork == 0)) {
    if (lda < m || lda == 0) return -4;
    if ((m < n ? m : n) == 0) return RNPX_OK;
    *lwork = kRefNb * (f == Fam::QL ? n : m);
    return RNPX_OK;
  }
  return dropin_gexxf<T>(f, m, n, a, lda, tau, lwork);
}

// Hessenberg drop-ins (RNP/LA/Hessenberg.hpp): the checks of MultQ
// (Hessenberg.hpp:488-494; the query is QR::MultQ's, which returns without
// one when nh = 0) and GenerateQ (Hessenberg.hpp:545-556: query nh nb).
template <typename T>
int dropin_ormhr(const char* side, const char* trans, size_t m, size_t n, size_t ilo, size_t ihi, const T* a,
                 size_t lda, const T* tau, T* c, size_t ldc, size_t* lwork, bool query) {
  if (!side || (side[0] != 'L' && side[0] != 'R')) return -1;
  if (!trans || (trans[0] != 'N' && trans[0] != 'C' && trans[0] != 'T')) return -2;
  const bool left = ('L' == side[0]);
  const size_t nq = left ? m : n, nw = left ? n : m;
  if (ilo > nq) return -5;
  if (ihi <= ilo || ihi > nq) return -6;
  if (!lwork) return -12;
  if (!(*lwork == 0 || *lwork >= nw)) return -12;
  const size_t nh = ihi - ilo - 1;
  if (ldc < m || ldc == 0) return -11;
  if (m == 0 || n == 0 || nh == 0) return RNPX_OK;
  if (lda < nq) return -8;
  if (query) {
    *lwork = kRefNb * nw + kRefNb * kRefNb;
    return RNPX_OK;
  }
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU, C;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), nq, nq, lda, true))) return st;
  const size_t tl = nq > 1 ? nq - 1 : 1;  // tau holds nq-1 scalars (Hessenberg.hpp:466-467)
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), tl, 1, tl, nq > 1))) return st;
  if ((st = C.in(h, h->stage3, c, m, n, ldc, true))) return st;
  st = ormhr_dev<T>(h, side[0], trans[0], (int)m, (int)n, (int)ilo, (int)ihi, A.dev, A.ld, TAU.dev, C.dev, C.ld);
  if (st) return st;
  if ((st = C.out(h, ldc))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_orghr(size_t n, size_t ilo, size_t ihi, T* a, size_t lda, const T* tau, size_t* lwork, bool query) {
  if (ilo > n) return -2;
  if (ihi <= ilo || ihi > n) return -3;
  if (lda < n || lda == 0) return -5;
  if (!lwork) return -7;
  const size_t nh = ihi - ilo - 1;
  if (!(*lwork == 0 || *lwork >= nh)) return -7;
  if (query) {
    *lwork = nh * kRefNb;
    return RNPX_OK;
  }
  if (n == 0) return RNPX_OK;
  if (!fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, TAU;
  if ((st = A.in(h, h->stage1, a, n, n, lda, true))) return st;
  const size_t tl = n > 1 ? n - 1 : 1;  // tau holds n-1 scalars (Hessenberg.hpp:535)
  if ((st = TAU.in(h, h->stage2, const_cast<T*>(tau), tl, 1, tl, n > 1))) return st;
  st = orghr_dev<T>(h, (int)n, (int)ilo, (int)ihi, A.dev, A.ld, TAU.dev);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  return sync_handle(h);
}

// LU drop-ins (RNP/LA/LU.hpp): Factor returns the reference's value
// (LU.hpp:63-74, 112-113; >= 0) or a negative status; pivots are size_t,
// 0-based absolute rows, as there.
template <typename T>
int dropin_getrf(size_t m, size_t n, T* a, size_t lda, size_t* pivots) {
  if (lda < m || lda == 0) return -4;
  const size_t k = m < n ? m : n;
  if (k == 0) return RNPX_OK;
  if (!pivots) return -5;
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A;
  if ((st = A.in(h, h->stage1, a, m, n, lda, true))) return st;
  std::vector<int> piv(k);
  int info = 0;
  st = getrf_dev<T>(h, (int)m, (int)n, A.dev, A.ld, piv.data(), &info);
  if (st) return st;
  if ((st = A.out(h, lda))) return st;
  if ((st = sync_handle(h))) return st;
  for (size_t i = 0; i < k; ++i) pivots[i] = (size_t)piv[i];
  return info;
}

template <typename T>
int dropin_getrs(const char* trans, size_t n, size_t nrhs, const T* a, size_t lda, const size_t* ipiv, T* b,
                 size_t ldb) {
  if (!trans || (trans[0] != 'N' && trans[0] != 'T' && trans[0] != 'C')) return -1;
  if (n == 0 || nrhs == 0) return RNPX_OK;
  if (lda < n) return -5;
  if (!ipiv) return -6;
  if (ldb < n) return -8;
  if (!fits_int(n) || !fits_int(nrhs)) return RNPX_ERR_UNSUPPORTED;
  std::vector<int> piv(n);
  for (size_t i = 0; i < n; ++i) {
    if (ipiv[i] >= n) return -6;
    piv[i] = (int)ipiv[i];
  }
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, B;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), n, n, lda, true))) return st;
  if ((st = B.in(h, h->stage3, b, n, nrhs, ldb, true))) return st;
  st = getrs_dev<T>(h, trans[0], (int)n, (int)nrhs, A.dev, A.ld, piv.data(), B.dev, B.ld);
  if (st) return st;
  if ((st = B.out(h, ldb))) return st;
  return sync_handle(h);
}

template <typename T>
int dropin_getri(size_t n, T* a, size_t lda, const size_t* ipiv, size_t* lwork, bool query) {
  if (!lwork) return RNPX_ERR_ARG;
  if (n == 0) return RNPX_OK;        // LU.hpp:149, before the query
  if (query) {
    *lwork = 64 * n;                 // LU::Tuning<T>::invert_block_size_opt(n) * n, LU.hpp:151-154
    return RNPX_OK;
  }
  if (lda < n || !a || !ipiv) return RNPX_ERR_ARG;
  if (!fits_int(n)) return RNPX_ERR_UNSUPPORTED;
  std::vector<int> piv(n);
  for (size_t i = 0; i < n; ++i) {
    if (ipiv[i] >= n) return RNPX_ERR_ARG;
    piv[i] = (int)ipiv[i];
  }
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A;
  if ((st = A.in(h, h->stage1, a, n, n, lda, true))) return st;
  int info = 0;
  if ((st = getri_dev<T>(h, (int)n, A.dev, A.ld, piv.data(), &info))) return st > -1000 ? RNPX_ERR_ARG : st;
  if (info != 0) return info;        // a left as it was
  if ((st = A.out(h, lda))) return st;
  if ((st = sync_handle(h))) return st;
  return info;
}

template <typename T>
int dropin_gerfs(const char* trans, size_t n, size_t nrhs, const T* a, size_t lda, const T* af, size_t ldaf,
                 const size_t* ipiv, const T* b, size_t ldb, T* x, size_t ldx, double* ferr, double* berr) {
  if (!trans || (trans[0] != 'N' && trans[0] != 'T' && trans[0] != 'C')) return -1;
  if (n == 0 || nrhs == 0) {
    for (size_t j = 0; j < nrhs; ++j) {
      if (ferr) ferr[j] = 0.0;
      if (berr) berr[j] = 0.0;
    }
    return RNPX_OK;
  }
  if (lda < n) return -5;
  if (ldaf < n) return -7;
  if (!ipiv) return -8;
  if (ldb < n) return -10;
  if (ldx < n) return -12;
  if (!fits_int(n) || !fits_int(nrhs)) return RNPX_ERR_UNSUPPORTED;
  std::vector<int> piv(n);
  for (size_t i = 0; i < n; ++i) {
    if (ipiv[i] >= n) return -8;
    piv[i] = (int)ipiv[i];
  }
  int st = 0;
  Handle* h = default_handle(&st);
  if (!h) return st;
  Arg<T> A, AF, B, X;
  if ((st = A.in(h, h->stage1, const_cast<T*>(a), n, n, lda, true))) return st;
  if ((st = AF.in(h, h->stage2, const_cast<T*>(af), n, n, ldaf, true))) return st;
  if ((st = B.in(h, h->stage3, const_cast<T*>(b), n, nrhs, ldb, true))) return st;
  if ((st = X.in(h, h->lsbuf, x, n, nrhs, ldx, true))) return st;
  st = gerfs_dev<T>(h, trans[0], (int)n, (int)nrhs, A.dev, A.ld, AF.dev, AF.ld, piv.data(), B.dev, B.ld, X.dev, X.ld,
                    ferr, berr);
  if (st) return st;
  if ((st = X.out(h, ldx))) return st;
  return sync_handle(h);
}

Handle* H(rnpx_handle h) { return reinterpret_cast<Handle*>(h); }

}  // namespace

struct rnpx_handle_s {};

extern "C" {

const char* rnpx_version(void) { return "rnpx 0.1 (gfx950, FP64 MFMA)"; }

int rnpx_device_name(char* buf, size_t len) {
  int st = Handle::check_gpu();
  if (st) return st;
  int dev = 0;
  hipGetDevice(&dev);
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_status(e);
  if (buf && len) {
    snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  }
  return RNPX_OK;
}

int rnpx_dgeqrf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lwork, double* work) {
  if (lwork && (work == nullptr || *lwork == 0)) {
    if (lda < m || lda == 0) return -4;
    if ((m < n ? m : n) == 0) return RNPX_OK;
    *lwork = kRefNb * n;
    return RNPX_OK;
  }
  return dropin_geqrf<double>(m, n, a, lda, tau, lwork);
}
int rnpx_zgeqrf(size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  if (lwork && (work == nullptr || *lwork == 0)) {
    if (lda < m || lda == 0) return -4;
    if ((m < n ? m : n) == 0) return RNPX_OK;
    *lwork = kRefNb * n;
    return RNPX_OK;
  }
  return dropin_geqrf<zd>(m, n, reinterpret_cast<zd*>(a), lda, reinterpret_cast<zd*>(tau), lwork);
}

int rnpx_dlast_nonzero_column_limit(size_t m, size_t n, const double* a, size_t lda, size_t* limit) {
  return dropin_last_nonzero<double>(0, m, n, a, lda, limit);
}
int rnpx_zlast_nonzero_column_limit(size_t m, size_t n, const rnpx_zcomplex* a, size_t lda, size_t* limit) {
  return dropin_last_nonzero<zd>(0, m, n, reinterpret_cast<const zd*>(a), lda, limit);
}
int rnpx_dlast_nonzero_row_limit(size_t m, size_t n, const double* a, size_t lda, size_t* limit) {
  return dropin_last_nonzero<double>(1, m, n, a, lda, limit);
}
int rnpx_zlast_nonzero_row_limit(size_t m, size_t n, const rnpx_zcomplex* a, size_t lda, size_t* limit) {
  return dropin_last_nonzero<zd>(1, m, n, reinterpret_cast<const zd*>(a), lda, limit);
}

int rnpx_dormqr(const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a, size_t lda,
                const double* tau, double* c, size_t ldc, size_t* lwork, double* work) {
  return dropin_ormqr<double>(side, trans, m, n, k, a, lda, tau, c, ldc, lwork,
                              lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunmqr(const char* side, const char* trans, size_t m, size_t n, size_t k, const rnpx_zcomplex* a, size_t lda,
                const rnpx_zcomplex* tau, rnpx_zcomplex* c, size_t ldc, size_t* lwork, rnpx_zcomplex* work) {
  return dropin_ormqr<zd>(side, trans, m, n, k, reinterpret_cast<const zd*>(a), lda,
                          reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(c), ldc, lwork,
                          lwork && (work == nullptr || *lwork == 0));
}

int rnpx_dorgqr(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lwork,
                double* work) {
  return dropin_orgqr<double>(m, n, k, a, lda, tau, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zungqr(size_t m, size_t n, size_t k, rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  return dropin_orgqr<zd>(m, n, k, reinterpret_cast<zd*>(a), lda, reinterpret_cast<const zd*>(tau), lwork,
                          lwork && (work == nullptr || *lwork == 0));
}

int rnpx_dgeqlf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lwork, double* work) {
  return gexxf_entry<double>(Fam::QL, m, n, a, lda, tau, lwork, work);
}
int rnpx_zgeqlf(size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  return gexxf_entry<zd>(Fam::QL, m, n, reinterpret_cast<zd*>(a), lda, reinterpret_cast<zd*>(tau), lwork,
                         reinterpret_cast<zd*>(work));
}
int rnpx_dormql(const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a, size_t lda,
                const double* tau, double* c, size_t ldc, size_t* lwork, double* work) {
  return dropin_ormxx<double>(Fam::QL, side, trans, m, n, k, a, lda, tau, c, ldc, lwork,
                              lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunmql(const char* side, const char* trans, size_t m, size_t n, size_t k, const rnpx_zcomplex* a, size_t lda,
                const rnpx_zcomplex* tau, rnpx_zcomplex* c, size_t ldc, size_t* lwork, rnpx_zcomplex* work) {
  return dropin_ormxx<zd>(Fam::QL, side, trans, m, n, k, reinterpret_cast<const zd*>(a), lda,
                          reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(c), ldc, lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dorgql(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lwork,
                double* work) {
  return dropin_orgxx<double>(Fam::QL, m, n, k, a, lda, tau, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zungql(size_t m, size_t n, size_t k, rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  return dropin_orgxx<zd>(Fam::QL, m, n, k, reinterpret_cast<zd*>(a), lda, reinterpret_cast<const zd*>(tau), lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dgerqf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lwork, double* work) {
  return gexxf_entry<double>(Fam::RQ, m, n, a, lda, tau, lwork, work);
}
int rnpx_zgerqf(size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  return gexxf_entry<zd>(Fam::RQ, m, n, reinterpret_cast<zd*>(a), lda, reinterpret_cast<zd*>(tau), lwork,
                         reinterpret_cast<zd*>(work));
}
int rnpx_dormrq(const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a, size_t lda,
                const double* tau, double* c, size_t ldc, size_t* lwork, double* work) {
  return dropin_ormxx<double>(Fam::RQ, side, trans, m, n, k, a, lda, tau, c, ldc, lwork,
                              lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunmrq(const char* side, const char* trans, size_t m, size_t n, size_t k, const rnpx_zcomplex* a, size_t lda,
                const rnpx_zcomplex* tau, rnpx_zcomplex* c, size_t ldc, size_t* lwork, rnpx_zcomplex* work) {
  return dropin_ormxx<zd>(Fam::RQ, side, trans, m, n, k, reinterpret_cast<const zd*>(a), lda,
                          reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(c), ldc, lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dorgrq(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lwork,
                double* work) {
  return dropin_orgxx<double>(Fam::RQ, m, n, k, a, lda, tau, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zungrq(size_t m, size_t n, size_t k, rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  return dropin_orgxx<zd>(Fam::RQ, m, n, k, reinterpret_cast<zd*>(a), lda, reinterpret_cast<const zd*>(tau), lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dormhr(const char* side, const char* trans, size_t m, size_t n, size_t ilo, size_t ihi, const double* a,
                size_t lda, const double* tau, double* c, size_t ldc, size_t* lwork, double* work) {
  return dropin_ormhr<double>(side, trans, m, n, ilo, ihi, a, lda, tau, c, ldc, lwork,
                              lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunmhr(const char* side, const char* trans, size_t m, size_t n, size_t ilo, size_t ihi,
                const rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau, rnpx_zcomplex* c, size_t ldc,
                size_t* lwork, rnpx_zcomplex* work) {
  return dropin_ormhr<zd>(side, trans, m, n, ilo, ihi, reinterpret_cast<const zd*>(a), lda,
                          reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(c), ldc, lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dorghr(size_t n, size_t ilo, size_t ihi, double* a, size_t lda, const double* tau, size_t* lwork,
                double* work) {
  return dropin_orghr<double>(n, ilo, ihi, a, lda, tau, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunghr(size_t n, size_t ilo, size_t ihi, rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau,
                size_t* lwork, rnpx_zcomplex* work) {
  return dropin_orghr<zd>(n, ilo, ihi, reinterpret_cast<zd*>(a), lda, reinterpret_cast<const zd*>(tau), lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dgetrf(size_t m, size_t n, double* a, size_t lda, size_t* pivots) {
  return dropin_getrf<double>(m, n, a, lda, pivots);
}
int rnpx_zgetrf(size_t m, size_t n, rnpx_zcomplex* a, size_t lda, size_t* pivots) {
  return dropin_getrf<zd>(m, n, reinterpret_cast<zd*>(a), lda, pivots);
}
int rnpx_dgetri(size_t n, double* a, size_t lda, const size_t* ipiv, size_t* lwork, double* work) {
  return dropin_getri<double>(n, a, lda, ipiv, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zgetri(size_t n, rnpx_zcomplex* a, size_t lda, const size_t* ipiv, size_t* lwork, rnpx_zcomplex* work) {
  return dropin_getri<zd>(n, reinterpret_cast<zd*>(a), lda, ipiv, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dgetrs(const char* trans, size_t n, size_t nrhs, const double* a, size_t lda, const size_t* ipiv, double* b,
                size_t ldb) {
  return dropin_getrs<double>(trans, n, nrhs, a, lda, ipiv, b, ldb);
}
int rnpx_zgetrs(const char* trans, size_t n, size_t nrhs, const rnpx_zcomplex* a, size_t lda, const size_t* ipiv,
                rnpx_zcomplex* b, size_t ldb) {
  return dropin_getrs<zd>(trans, n, nrhs, reinterpret_cast<const zd*>(a), lda, ipiv, reinterpret_cast<zd*>(b), ldb);
}
int rnpx_dgerfs(const char* trans, size_t n, size_t nrhs, const double* a, size_t lda, const double* af, size_t ldaf,
                const size_t* ipiv, const double* b, size_t ldb, double* x, size_t ldx, double* ferr, double* berr,
                double* work, double* rwork) {
  (void)work;
  (void)rwork;
  return dropin_gerfs<double>(trans, n, nrhs, a, lda, af, ldaf, ipiv, b, ldb, x, ldx, ferr, berr);
}
int rnpx_zgerfs(const char* trans, size_t n, size_t nrhs, const rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* af,
                size_t ldaf, const size_t* ipiv, const rnpx_zcomplex* b, size_t ldb, rnpx_zcomplex* x, size_t ldx,
                double* ferr, double* berr, rnpx_zcomplex* work, double* rwork) {
  (void)work;
  (void)rwork;
  return dropin_gerfs<zd>(trans, n, nrhs, reinterpret_cast<const zd*>(a), lda, reinterpret_cast<const zd*>(af), ldaf,
                          ipiv, reinterpret_cast<const zd*>(b), ldb, reinterpret_cast<zd*>(x), ldx, ferr, berr);
}
int rnpx_dgelqf(size_t m, size_t n, double* a, size_t lda, double* tau, size_t* lwork, double* work) {
  if (lwork && (work == nullptr || *lwork == 0)) {
    if (lda < m || lda == 0) return -4;
    if ((m < n ? m : n) == 0) return RNPX_OK;
    *lwork = kRefNb * n;
    return RNPX_OK;
  }
  return dropin_gelqf<double>(m, n, a, lda, tau, lwork);
}
int rnpx_zgelqf(size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  if (lwork && (work == nullptr || *lwork == 0)) {
    if (lda < m || lda == 0) return -4;
    if ((m < n ? m : n) == 0) return RNPX_OK;
    *lwork = kRefNb * n;
    return RNPX_OK;
  }
  return dropin_gelqf<zd>(m, n, reinterpret_cast<zd*>(a), lda, reinterpret_cast<zd*>(tau), lwork);
}
int rnpx_dormlq(const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a, size_t lda,
                const double* tau, double* c, size_t ldc, size_t* lwork, double* work) {
  return dropin_ormlq<double>(side, trans, m, n, k, a, lda, tau, c, ldc, lwork,
                              lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunmlq(const char* side, const char* trans, size_t m, size_t n, size_t k, const rnpx_zcomplex* a, size_t lda,
                const rnpx_zcomplex* tau, rnpx_zcomplex* c, size_t ldc, size_t* lwork, rnpx_zcomplex* work) {
  return dropin_ormlq<zd>(side, trans, m, n, k, reinterpret_cast<const zd*>(a), lda,
                          reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(c), ldc, lwork,
                          lwork && (work == nullptr || *lwork == 0));
}
int rnpx_dorglq(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau, size_t* lwork,
                double* work) {
  return dropin_orglq<double>(m, n, k, a, lda, tau, lwork, lwork && (work == nullptr || *lwork == 0));
}
int rnpx_zunglq(size_t m, size_t n, size_t k, rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau, size_t* lwork,
                rnpx_zcomplex* work) {
  return dropin_orglq<zd>(m, n, k, reinterpret_cast<zd*>(a), lda, reinterpret_cast<const zd*>(tau), lwork,
                          lwork && (work == nullptr || *lwork == 0));
}

int rnpx_create(rnpx_handle* h, size_t hbm_budget_bytes) {
  if (!h) return -1;
  Handle* p = nullptr;
  int st = create_handle(&p, hbm_budget_bytes);
  if (st) return st;
  *h = reinterpret_cast<rnpx_handle>(p);
  return RNPX_OK;
}
int rnpx_destroy(rnpx_handle h) {
  delete H(h);
  return RNPX_OK;
}
int rnpx_set_stream(rnpx_handle h, void* s) {
  if (!h) return -1;
  H(h)->stream = s ? (hipStream_t)s : H(h)->own;
  return RNPX_OK;
}
void* rnpx_get_stream(rnpx_handle h) { return h ? (void*)H(h)->stream : nullptr; }
int rnpx_set_tuning(rnpx_handle h, const rnpx_tuning* t) {
  if (!h) return -1;
  if (!t) return -2;
  H(h)->tune = *t;
  return RNPX_OK;
}
int rnpx_get_tuning(rnpx_handle h, rnpx_tuning* t) {
  if (!h) return -1;
  if (!t) return -2;
  *t = H(h)->tune;
  return RNPX_OK;
}
int rnpx_synchronize(rnpx_handle h) {
  if (!h) return -1;
  return sync_handle(H(h));
}
size_t rnpx_workspace_bytes(rnpx_handle h) { return h ? H(h)->held_bytes() : 0; }
int rnpx_exact_path_count(rnpx_handle h, long* count) {
  if (!count) return -2;
  int st = 0;
  Handle* hh = h ? H(h) : default_handle(&st);
  if (!hh) return st;
  *count = hh->exact_calls;
  return RNPX_OK;
}
int rnpx_profile_enable(rnpx_handle h, int on) {
  if (!h) return -1;
  prof_reset(H(h));
  H(h)->prof = (on != 0);
  H(h)->prof_mask = (on == 1) ? ~0u : ((unsigned)on >> 1);
  return RNPX_OK;
}
int rnpx_profile_read(rnpx_handle h, int kernel, long* launches, double* ms, double* flops, double* bytes) {
  if (!h) return -1;
  if (kernel < 0 || kernel >= K_COUNT) return -2;
  if (!launches || !ms || !flops || !bytes) return -3;
  return prof_collect(H(h), kernel, launches, ms, flops, bytes);
}

#define RNPX_DEV_CHECK(h, m, n)                            \
  if (!h) return -1;                                        \
  if (!fits_int(m) || !fits_int(n)) return RNPX_ERR_UNSUPPORTED;

int rnpx_dgeqrf_dev(rnpx_handle h, size_t m, size_t n, double* a, size_t lda, double* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return geqrf_dev<double>(H(h), (int)m, (int)n, a, (long)lda, tau);
}
int rnpx_zgeqrf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return geqrf_dev<zd>(H(h), (int)m, (int)n, reinterpret_cast<zd*>(a), (long)lda, reinterpret_cast<zd*>(tau));
}
int rnpx_dgelqf_dev(rnpx_handle h, size_t m, size_t n, double* a, size_t lda, double* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return gelqf_dev<double>(H(h), (int)m, (int)n, a, (long)lda, tau);
}
int rnpx_zgelqf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return gelqf_dev<zd>(H(h), (int)m, (int)n, reinterpret_cast<zd*>(a), (long)lda, reinterpret_cast<zd*>(tau));
}
int rnpx_dgeqlf_dev(rnpx_handle h, size_t m, size_t n, double* a, size_t lda, double* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return geqlf_dev<double>(H(h), (int)m, (int)n, a, (long)lda, tau);
}
int rnpx_zgeqlf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return geqlf_dev<zd>(H(h), (int)m, (int)n, reinterpret_cast<zd*>(a), (long)lda, reinterpret_cast<zd*>(tau));
}
int rnpx_dgerqf_dev(rnpx_handle h, size_t m, size_t n, double* a, size_t lda, double* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return gerqf_dev<double>(H(h), (int)m, (int)n, a, (long)lda, tau);
}
int rnpx_zgerqf_dev(rnpx_handle h, size_t m, size_t n, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau) {
  RNPX_DEV_CHECK(h, m, n);
  return gerqf_dev<zd>(H(h), (int)m, (int)n, reinterpret_cast<zd*>(a), (long)lda, reinterpret_cast<zd*>(tau));
}
int rnpx_dorghr_dev(rnpx_handle h, size_t n, size_t ilo, size_t ihi, double* a, size_t lda, const double* tau) {
  RNPX_DEV_CHECK(h, n, n);
  return orghr_dev<double>(H(h), (int)n, (int)ilo, (int)ihi, a, (long)lda, tau);
}
int rnpx_zunghr_dev(rnpx_handle h, size_t n, size_t ilo, size_t ihi, rnpx_zcomplex* a, size_t lda,
                    const rnpx_zcomplex* tau) {
  RNPX_DEV_CHECK(h, n, n);
  return orghr_dev<zd>(H(h), (int)n, (int)ilo, (int)ihi, reinterpret_cast<zd*>(a), (long)lda,
                       reinterpret_cast<const zd*>(tau));
}
int rnpx_dormqr_dev(rnpx_handle h, const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a,
                    size_t lda, const double* tau, double* c, size_t ldc) {
  RNPX_DEV_CHECK(h, m, n);
  if (!side) return -2;
  if (!trans) return -3;
  int st = ormqr_dev<double>(H(h), side[0], trans[0], (int)m, (int)n, (int)k, a, (long)lda, tau, c, (long)ldc);
  return st < 0 && st > -100 ? st - 1 : st;
}
int rnpx_zunmqr_dev(rnpx_handle h, const char* side, const char* trans, size_t m, size_t n, size_t k,
                    const rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* tau, rnpx_zcomplex* c, size_t ldc) {
  RNPX_DEV_CHECK(h, m, n);
  if (!side) return -2;
  if (!trans) return -3;
  int st = ormqr_dev<zd>(H(h), side[0], trans[0], (int)m, (int)n, (int)k, reinterpret_cast<const zd*>(a), (long)lda,
                         reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(c), (long)ldc);
  return st < 0 && st > -100 ? st - 1 : st;
}
int rnpx_dorgqr_dev(rnpx_handle h, size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau) {
  RNPX_DEV_CHECK(h, m, n);
  int st = orgqr_dev<double>(H(h), (int)m, (int)n, (int)k, a, (long)lda, tau);
  return st < 0 && st > -100 ? st - 1 : st;
}
int rnpx_zungqr_dev(rnpx_handle h, size_t m, size_t n, size_t k, rnpx_zcomplex* a, size_t lda,
                    const rnpx_zcomplex* tau) {
  RNPX_DEV_CHECK(h, m, n);
  int st = orgqr_dev<zd>(H(h), (int)m, (int)n, (int)k, reinterpret_cast<zd*>(a), (long)lda,
                         reinterpret_cast<const zd*>(tau));
  return st < 0 && st > -100 ? st - 1 : st;
}
int rnpx_dgels_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, double* a, size_t lda, double* tau, double* b,
                   size_t ldb) {
  RNPX_DEV_CHECK(h, m, n);
  if (!fits_int(nrhs)) return RNPX_ERR_UNSUPPORTED;
  return gels_dev<double>(H(h), (int)m, (int)n, (int)nrhs, a, (long)lda, tau, b, (long)ldb);
}
int rnpx_zgels_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, rnpx_zcomplex* a, size_t lda, rnpx_zcomplex* tau,
                   rnpx_zcomplex* b, size_t ldb) {
  RNPX_DEV_CHECK(h, m, n);
  if (!fits_int(nrhs)) return RNPX_ERR_UNSUPPORTED;
  return gels_dev<zd>(H(h), (int)m, (int)n, (int)nrhs, reinterpret_cast<zd*>(a), (long)lda,
                      reinterpret_cast<zd*>(tau), reinterpret_cast<zd*>(b), (long)ldb);
}
int rnpx_dgels_mixed_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, const double* a, size_t lda, const double* b,
                         size_t ldb, double* x, size_t ldx, int max_iter, double tol, int* iters, double* berr) {
  RNPX_DEV_CHECK(h, m, n);
  if (!fits_int(nrhs)) return RNPX_ERR_UNSUPPORTED;
  if (!iters || !berr) return -13;
  return gels_mixed_dev(H(h), (int)m, (int)n, (int)nrhs, a, (long)lda, b, (long)ldb, x, (long)ldx, max_iter, tol,
                        iters, berr);
}
int rnpx_dls_berr_dev(rnpx_handle h, size_t m, size_t n, size_t nrhs, const double* a, size_t lda, const double* b,
                      size_t ldb, const double* x, size_t ldx, double* berr) {
  RNPX_DEV_CHECK(h, m, n);
  if (!berr) return -11;
  return ls_berr_dev(H(h), (int)m, (int)n, (int)nrhs, a, (long)lda, b, (long)ldb, x, (long)ldx, berr);
}
int rnpx_dgeqrf_batched(rnpx_handle h, int count, const rnpx_qr_desc* desc, double* A, double* tau, double* B) {
  if (!h) return -1;
  return geqrf_batched<double>(H(h), count, desc, A, tau, B);
}
int rnpx_zgeqrf_batched(rnpx_handle h, int count, const rnpx_qr_desc* desc, rnpx_zcomplex* A, rnpx_zcomplex* tau,
                        rnpx_zcomplex* B) {
  if (!h) return -1;
  return geqrf_batched<zd>(H(h), count, desc, reinterpret_cast<zd*>(A), reinterpret_cast<zd*>(tau),
                           reinterpret_cast<zd*>(B));
}
int rnpx_dgeqrf_batched_plan(rnpx_handle h, int count, const rnpx_qr_desc* desc, double* A, double* tau, double* B,
                             rnpx_batched_plan* plan) {
  if (!h) return -1;
  return geqrf_batched_plan<double>(H(h), count, desc, A, tau, B, reinterpret_cast<BatchedPlan**>(plan));
}
int rnpx_zgeqrf_batched_plan(rnpx_handle h, int count, const rnpx_qr_desc* desc, rnpx_zcomplex* A,
                             rnpx_zcomplex* tau, rnpx_zcomplex* B, rnpx_batched_plan* plan) {
  if (!h) return -1;
  return geqrf_batched_plan<zd>(H(h), count, desc, reinterpret_cast<zd*>(A), reinterpret_cast<zd*>(tau),
                                reinterpret_cast<zd*>(B), reinterpret_cast<BatchedPlan**>(plan));
}
int rnpx_batched_plan_run(rnpx_batched_plan plan) { return batched_plan_run(reinterpret_cast<BatchedPlan*>(plan)); }
int rnpx_batched_plan_destroy(rnpx_batched_plan plan) {
  return batched_plan_destroy(reinterpret_cast<BatchedPlan*>(plan));
}
int rnpx_dlarft_dev(rnpx_handle h, size_t n, size_t k, const double* v, size_t ldv, const double* tau, double* t,
                    size_t ldt) {
  RNPX_DEV_CHECK(h, n, k);
  return larft_dev<double>(H(h), (int)n, (int)k, v, (long)ldv, tau, t, (long)ldt);
}
int rnpx_zlarft_dev(rnpx_handle h, size_t n, size_t k, const rnpx_zcomplex* v, size_t ldv, const rnpx_zcomplex* tau,
                    rnpx_zcomplex* t, size_t ldt) {
  RNPX_DEV_CHECK(h, n, k);
  return larft_dev<zd>(H(h), (int)n, (int)k, reinterpret_cast<const zd*>(v), (long)ldv,
                       reinterpret_cast<const zd*>(tau), reinterpret_cast<zd*>(t), (long)ldt);
}
int rnpx_dlarfb_dev(rnpx_handle h, const char* side, const char* trans, size_t m, size_t n, size_t k,
                    const double* v, size_t ldv, const double* t, size_t ldt, double* c, size_t ldc) {
  RNPX_DEV_CHECK(h, m, n);
  if (!side) return -2;
  if (!trans) return -3;
  return larfb_dev<double>(H(h), side[0], trans[0], (int)m, (int)n, (int)k, v, (long)ldv, t, (long)ldt, c,
                           (long)ldc);
}
int rnpx_zlarfb_dev(rnpx_handle h, const char* side, const char* trans, size_t m, size_t n, size_t k,
                    const rnpx_zcomplex* v, size_t ldv, const rnpx_zcomplex* t, size_t ldt, rnpx_zcomplex* c,
                    size_t ldc) {
  RNPX_DEV_CHECK(h, m, n);
  if (!side) return -2;
  if (!trans) return -3;
  return larfb_dev<zd>(H(h), side[0], trans[0], (int)m, (int)n, (int)k, reinterpret_cast<const zd*>(v), (long)ldv,
                       reinterpret_cast<const zd*>(t), (long)ldt, reinterpret_cast<zd*>(c), (long)ldc);
}
int rnpx_dgemm_dev(rnpx_handle h, const char* transa, const char* transb, size_t m, size_t n, size_t k, double alpha,
                   const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc) {
  RNPX_DEV_CHECK(h, m, n);
  if (!transa) return -2;
  if (!transb) return -3;
  if (!fits_int(k)) return RNPX_ERR_UNSUPPORTED;
  if (n >= 1 && n <= 4 && transb[0] == 'N' && (transa[0] == 'N' || transa[0] == 'T' || transa[0] == 'C'))
    return gemv_dev<double>(H(h), transa[0], (int)m, (int)n, (int)k, alpha, a, (long)lda, b, (long)ldb, beta, c,
                            (long)ldc, H(h)->stream);
  return hip_status(gemm<double>(transa[0], transb[0], (int)m, (int)n, (int)k, alpha, a, (long)lda, b, (long)ldb,
                                 beta, c, (long)ldc, 1, nullptr, H(h)->stream));
}
int rnpx_zgemm_dev(rnpx_handle h, const char* transa, const char* transb, size_t m, size_t n, size_t k,
                   rnpx_zcomplex alpha, const rnpx_zcomplex* a, size_t lda, const rnpx_zcomplex* b, size_t ldb,
                   rnpx_zcomplex beta, rnpx_zcomplex* c, size_t ldc) {
  RNPX_DEV_CHECK(h, m, n);
  if (!transa) return -2;
  if (!transb) return -3;
  if (!fits_int(k)) return RNPX_ERR_UNSUPPORTED;
  if (n >= 1 && n <= 4 && transb[0] == 'N' && (transa[0] == 'N' || transa[0] == 'T' || transa[0] == 'C'))
    return gemv_dev<zd>(H(h), transa[0], (int)m, (int)n, (int)k, mkz(alpha.re, alpha.im),
                        reinterpret_cast<const zd*>(a), (long)lda, reinterpret_cast<const zd*>(b), (long)ldb,
                        mkz(beta.re, beta.im), reinterpret_cast<zd*>(c), (long)ldc, H(h)->stream);
  return hip_status(gemm<zd>(transa[0], transb[0], (int)m, (int)n, (int)k, mkz(alpha.re, alpha.im),
                             reinterpret_cast<const zd*>(a), (long)lda, reinterpret_cast<const zd*>(b), (long)ldb,
                             mkz(beta.re, beta.im), reinterpret_cast<zd*>(c), (long)ldc, 1, nullptr, H(h)->stream));
}
int rnpx_dlarnv_dev(rnpx_handle h, int dist, const int iseed[4], size_t count, double* x) {
  if (!h) return -1;
  if (dist < 0 || dist > 2) return -2;
  if (!iseed) return -3;
  return hip_status(launch_larnv<double>(dist, iseed, (long long)count, x, H(h)->stream));
}
int rnpx_zlarnv_dev(rnpx_handle h, int dist, const int iseed[4], size_t count, rnpx_zcomplex* x) {
  if (!h) return -1;
  if (dist < 0 || dist > 2) return -2;
  if (!iseed) return -3;
  return hip_status(launch_larnv<zd>(dist, iseed, (long long)count, reinterpret_cast<zd*>(x), H(h)->stream));
}

}  // extern "C"
