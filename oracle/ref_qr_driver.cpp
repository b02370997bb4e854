// ref_qr_driver.cpp -- C entry points over the real RNP reference's blocked
// QR, compiled from the reference's own sources where they lie (Makefile
// target `ref`, output oracle/_ref/librnpref_qr.so) and linked to the image's
// Fortran-ABI LP64 BLAS (conda MKL, libmkl_rt.so.1, by full path).  Test and
// CPU-baseline infrastructure only: it lets the tests check the oracle's
// restatement against the reference itself and lets the benchmark time the
// reference's own Factor on the host cores.
//   RNP::LA::QR::Factor        (RNP/LA/QR.hpp:129-197)
//   RNP::LA::QR::MultQ         (RNP/LA/QR.hpp:337-429)
//   RNP::LA::QR::GenerateQ     (RNP/LA/QR.hpp:509-570)
//   RNP::MatrixGen::RandomMatrix (RNP/MatrixGen.hpp:120-267)
//   RNP::LA::LU::Factor / Invert (RNP/LA/LU.hpp:77-214)
// Workspace is sized by the reference's own query (work == NULL).
#include <complex>
#include <cstddef>
#include <vector>
#include <RNP/LA/QR.hpp>
#include <RNP/LA/Triangular.hpp>  // LU::Invert calls Triangular::Invert; LU.hpp does not include it
#include <RNP/LA/LU.hpp>
#include <RNP/MatrixGen.hpp>

typedef std::complex<double> zc;

namespace {
template <typename T>
void factor(size_t m, size_t n, T* a, size_t lda, T* tau, size_t lwork_cap) {
  size_t lwork = 0;
  RNP::LA::QR::Factor(m, n, a, lda, tau, &lwork, (T*)NULL);
  if (lwork_cap && lwork_cap < lwork) lwork = lwork_cap;  // a short workspace: the reference degrades nb
  std::vector<T> work(lwork ? lwork : 1);
  RNP::LA::QR::Factor(m, n, a, lda, tau, &lwork, &work[0]);
}
template <typename T>
void multq(const char* side, const char* trans, size_t m, size_t n, size_t k, const T* a, size_t lda, const T* tau,
           T* c, size_t ldc) {
  size_t lwork = 0;
  RNP::LA::QR::MultQ(side, trans, m, n, k, a, lda, tau, c, ldc, &lwork, (T*)NULL);
  std::vector<T> work(lwork ? lwork : 1);
  RNP::LA::QR::MultQ(side, trans, m, n, k, a, lda, tau, c, ldc, &lwork, &work[0]);
}
template <typename T>
void generateq(size_t m, size_t n, size_t k, T* a, size_t lda, const T* tau) {
  size_t lwork = 0;
  RNP::LA::QR::GenerateQ(m, n, k, a, lda, tau, &lwork, (T*)NULL);
  std::vector<T> work(lwork ? lwork : 1);
  RNP::LA::QR::GenerateQ(m, n, k, a, lda, tau, &lwork, &work[0]);
}
}  // namespace

extern "C" {

void refqr_d_factor(size_t m, size_t n, double* a, size_t lda, double* tau, size_t lwork_cap) {
  factor(m, n, a, lda, tau, lwork_cap);
}
void refqr_z_factor(size_t m, size_t n, double* a, size_t lda, double* tau, size_t lwork_cap) {
  factor(m, n, reinterpret_cast<zc*>(a), lda, reinterpret_cast<zc*>(tau), lwork_cap);
}
void refqr_d_multq(const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a, size_t lda,
                   const double* tau, double* c, size_t ldc) {
  multq(side, trans, m, n, k, a, lda, tau, c, ldc);
}
void refqr_z_multq(const char* side, const char* trans, size_t m, size_t n, size_t k, const double* a, size_t lda,
                   const double* tau, double* c, size_t ldc) {
  multq(side, trans, m, n, k, reinterpret_cast<const zc*>(a), lda, reinterpret_cast<const zc*>(tau),
        reinterpret_cast<zc*>(c), ldc);
}
void refqr_d_generateq(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau) {
  generateq(m, n, k, a, lda, tau);
}
void refqr_z_generateq(size_t m, size_t n, size_t k, double* a, size_t lda, const double* tau) {
  generateq(m, n, k, reinterpret_cast<zc*>(a), lda, reinterpret_cast<const zc*>(tau));
}
// Nonsymmetric RandomMatrix; d receives the singular values used.
void refqr_d_random_matrix(int dist, int mode, int reverse, double cond, double dmax, size_t m, size_t n, double* a,
                           size_t lda, double* d, int* iseed) {
  std::vector<double> work(m + n + 1);
  RNP::MatrixGen::RandomMatrix(RNP::MatrixGen::Symmetry::Nonsymmetric, (RNP::Random::Distribution::Distribution)dist,
                               (RNP::MatrixGen::DiagonalMode::Mode)mode, reverse != 0, cond, dmax, d, m, n, a, lda,
                               iseed, &work[0]);
}
int refqr_d_lu_factor(size_t m, size_t n, double* a, size_t lda, size_t* ipiv) {
  return RNP::LA::LU::Factor(m, n, a, lda, ipiv);
}
int refqr_z_lu_factor(size_t m, size_t n, double* a, size_t lda, size_t* ipiv) {
  return RNP::LA::LU::Factor(m, n, reinterpret_cast<zc*>(a), lda, ipiv);
}
// lwork_cap > 0 shortens the workspace (lwork = n takes the unblocked branch).
int refqr_d_lu_invert(size_t n, double* a, size_t lda, const size_t* ipiv, size_t lwork_cap) {
  size_t lwork = 0;
  RNP::LA::LU::Invert(n, a, lda, ipiv, &lwork, (double*)NULL);
  if (lwork_cap && lwork_cap < lwork) lwork = lwork_cap;
  std::vector<double> work(lwork ? lwork : 1);
  return RNP::LA::LU::Invert(n, a, lda, ipiv, &lwork, &work[0]);
}
int refqr_z_lu_invert(size_t n, double* a, size_t lda, const size_t* ipiv, size_t lwork_cap) {
  size_t lwork = 0;
  RNP::LA::LU::Invert(n, reinterpret_cast<zc*>(a), lda, ipiv, &lwork, (zc*)NULL);
  if (lwork_cap && lwork_cap < lwork) lwork = lwork_cap;
  std::vector<zc> work(lwork ? lwork : 1);
  return RNP::LA::LU::Invert(n, reinterpret_cast<zc*>(a), lda, ipiv, &lwork, &work[0]);
}

}  // extern "C"
