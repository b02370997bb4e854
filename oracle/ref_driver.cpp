// ref_driver.cpp -- C entry points over the header-only part of the real RNP
// reference, compiled from the reference's own headers where they lie
// (Makefile target `ref`, output oracle/_ref/librnpref.so).  Test
// infrastructure only: it lets the tests check the oracle's restatement
// against the reference itself, bit for bit, for the functions that need no
// external BLAS:
//   RNP::LA::Reflector::Generate      (RNP/LA/Reflector.hpp:117-158)
//   RNP::BLAS::Norm2                  (RNP/BLAS/BLAS_mix.hpp:725-758)
//   RNP::Traits<double>::hypot3       (RNP/Types.hpp:73-92)
//   RNP::Random::UniformRealVector    (RNP/Random.hpp:45-183)
//   RNP::Random::GenerateVector       (RNP/Random.hpp:262-320)
//   RNP::MatrixGen::GenerateDiagonal  (RNP/MatrixGen.hpp:31-110)
// The blocked QR (Factor, MultQ, GenerateQ, Apply, ApplyBlock,
// GenerateBlockTr) calls a Fortran BLAS through src/BLAS_mix.cpp, which this
// image does not provide, so it is not built here.
#include <complex>
#include <cstddef>
#include <RNP/LA/Reflector.hpp>
#include <RNP/Random.hpp>
#include <RNP/MatrixGen.hpp>

typedef std::complex<double> zc;

extern "C" {

void ref_d_generate(size_t n, double* alpha, double* x, size_t incx, double* tau) {
  RNP::LA::Reflector::Generate(n, alpha, x, incx, tau);
}
void ref_z_generate(size_t n, double* alpha, double* x, size_t incx, double* tau) {
  RNP::LA::Reflector::Generate(n, reinterpret_cast<zc*>(alpha), reinterpret_cast<zc*>(x), incx,
                               reinterpret_cast<zc*>(tau));
}
double ref_d_norm2(size_t n, const double* x, size_t incx) { return RNP::BLAS::Norm2(n, x, incx); }
double ref_z_norm2(size_t n, const double* x, size_t incx) {
  return RNP::BLAS::Norm2(n, reinterpret_cast<const zc*>(x), incx);
}
double ref_hypot3(double a, double b, double c) { return RNP::Traits<double>::hypot3(a, b, c); }

// iseed == NULL draws from the reference's internal seed (state kept in the library).
void ref_uniform_real_vector(size_t n, double* x, int* iseed) { RNP::Random::UniformRealVector(n, x, iseed); }
void ref_d_generate_vector(int dist, size_t n, double* x, int* iseed) {
  RNP::Random::GenerateVector((RNP::Random::Distribution::Distribution)dist, n, x, iseed);
}
void ref_z_generate_vector(int dist, size_t n, double* x, int* iseed) {
  RNP::Random::GenerateVector((RNP::Random::Distribution::Distribution)dist, n, reinterpret_cast<zc*>(x), iseed);
}
void ref_generate_diagonal(int dist, int mode, int reverse, int randsign, double cond, double dscale, size_t n,
                           double* d, int* iseed) {
  RNP::MatrixGen::GenerateDiagonal((RNP::Random::Distribution::Distribution)dist,
                                   (RNP::MatrixGen::DiagonalMode::Mode)mode, reverse != 0, randsign != 0, cond,
                                   dscale, n, d, iseed);
}

}  // extern "C"
