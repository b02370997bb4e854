// rnpx_aux.hip -- small elementwise kernels: matrix set/copy, clean-V
// extraction from a packed factor, zeroing, and RNP's random generator.
#include "rnpx_kernels.hpp"

#include <algorithm>

namespace rnpx {

template <typename T>
__global__ void set_kernel(int m, int n, T off, T diag, T* A, long lda) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  A[(size_t)i + (size_t)j * lda] = (i == j) ? diag : off;
}

template <typename T>
__global__ void copy_kernel(int m, int n, const T* A, long lda, T* B, long ldb) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  B[(size_t)i + (size_t)j * ldb] = A[(size_t)i + (size_t)j * lda];
}

template <typename T>
__global__ void extract_v_kernel(int n, int k, const T* A, long lda, T* V, long ldv) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * k) return;
  const int i = (int)(e % n), j = (int)(e / n);
  V[(size_t)i + (size_t)j * ldv] = (i > j) ? A[(size_t)i + (size_t)j * lda] : (i == j ? S<T>::one() : S<T>::zero());
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0u;
}

static inline unsigned blocks_for(long total) { return (unsigned)((total + 255) / 256); }

template <typename T>
hipError_t launch_set(int m, int n, T offdiag, T diag, T* A, long lda, hipStream_t st) {
  if (m < 0 || n < 0 || lda < m) return hipErrorInvalidValue;
  if ((long)m * n == 0) return hipSuccess;
  hipLaunchKernelGGL((set_kernel<T>), dim3(blocks_for((long)m * n)), dim3(256), 0, st, m, n, offdiag, diag, A, lda);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy(int m, int n, const T* A, long lda, T* B, long ldb, hipStream_t st) {
  if (m < 0 || n < 0 || lda < m || ldb < m) return hipErrorInvalidValue;
  if ((long)m * n == 0) return hipSuccess;
  hipLaunchKernelGGL((copy_kernel<T>), dim3(blocks_for((long)m * n)), dim3(256), 0, st, m, n, A, lda, B, ldb);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_extract_v(int n, int k, const T* A, long lda, T* V, long ldv, hipStream_t st) {
  if (n < 0 || k < 0 || lda < n || ldv < n) return hipErrorInvalidValue;
  if ((long)n * k == 0) return hipSuccess;
  hipLaunchKernelGGL((extract_v_kernel<T>), dim3(blocks_for((long)n * k)), dim3(256), 0, st, n, k, A, lda, V, ldv);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_zero_rows(int rows, int n, T* A, long lda, hipStream_t st) {
  return launch_set<T>(rows, n, S<T>::zero(), S<T>::zero(), A, lda, st);
}

hipError_t launch_zero_u32(unsigned* p, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

// ---- RNP's 48-bit MCG (RNP/Random.hpp:45-183) ----
// Stream value k (0-based) = 1 - (seed * a^(k+1) mod 2^48) / 2^48; the low 48
// bits of a 64-bit product are the product mod 2^48.
__device__ inline uint64_t mcg_pow(uint64_t e) {
  const uint64_t mask = (1ull << 48) - 1;
  uint64_t r = 1, b = 33952834046453ull;
  while (e) {
    if (e & 1) r = (r * b) & mask;
    b = (b * b) & mask;
    e >>= 1;
  }
  return r;
}

template <typename T>
__global__ void larnv_kernel(int dist, uint64_t seed, long long count, T* x, int per_thread) {
  const uint64_t mask = (1ull << 48) - 1;
  const long long e0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * per_thread;
  if (e0 >= count) return;
  constexpr int U = S<T>::cplx ? 2 : 1;       // uniforms per element (Uniform01/_11)
  const int upe = (dist == 2) ? 2 : U;           // Normal01 always draws pairs
  uint64_t st = (seed * mcg_pow((uint64_t)(e0 * upe) + 1)) & mask;
  const uint64_t a = 33952834046453ull;
  const double twopi = 6.28318530717958647692528676656;
  for (int t = 0; t < per_thread && e0 + t < count; ++t) {
    double u[2];
    for (int q = 0; q < upe; ++q) {
      u[q] = 1.0 - (double)st * 0x1p-48;
      st = (st * a) & mask;
    }
    T v;
    if (dist == 2) {
      const double r = sqrt(-2.0 * log1p(-u[0]));
      const double ang = twopi * u[1];
      if constexpr (S<T>::cplx) v = S<T>::make(r * cos(ang), r * sin(ang));
      else v = S<T>::make(r * cos(ang), 0.0);
    } else if (dist == 1) {
      v = S<T>::make(2.0 * u[0] - 1.0, U == 2 ? 2.0 * u[1] - 1.0 : 0.0);
    } else {
      v = S<T>::make(u[0], U == 2 ? u[1] : 0.0);
    }
    x[e0 + t] = v;
  }
}

template <typename T>
hipError_t launch_larnv(int dist, const int iseed[4], long long count, T* x, hipStream_t st) {
  if (dist < 0 || dist > 2 || count < 0) return hipErrorInvalidValue;
  if (count == 0) return hipSuccess;
  const uint64_t seed = ((uint64_t)(iseed[0] & 4095) << 36) | ((uint64_t)(iseed[1] & 4095) << 24) |
                        ((uint64_t)(iseed[2] & 4095) << 12) | (uint64_t)((iseed[3] | 1) & 4095);
  const int per = 64;
  const long long threads = (count + per - 1) / per;
  hipLaunchKernelGGL((larnv_kernel<T>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, dist, seed, count,
                     x, per);
  return hipGetLastError();
}

// B (cols x rows) = conj(A)^T for A (rows x cols): 32x32 tiles through LDS,
// both sides read and written down their columns.
template <typename T>
__global__ __launch_bounds__(256) void conj_transpose_kernel(int rows, int cols, const T* __restrict__ A, long lda,
                                                             T* __restrict__ B, long ldb) {
  __shared__ T tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
  for (int q = 0; q < 32; q += 8) {
    const int r = r0 + tx, c = c0 + ty + q;
    if (r < rows && c < cols) tile[ty + q][tx] = A[(size_t)r + (size_t)c * lda];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 32; q += 8) {
    const int c = c0 + tx, r = r0 + ty + q;  // B(c, r) = conj(A(r, c))
    if (r < rows && c < cols) B[(size_t)c + (size_t)r * ldb] = S<T>::conj(tile[tx][ty + q]);
  }
}

template <typename T>
hipError_t launch_conj_transpose(int rows, int cols, const T* A, long lda, T* B, long ldb, hipStream_t st) {
  if (rows < 0 || cols < 0 || lda < std::max(1, rows) || ldb < std::max(1, cols)) return hipErrorInvalidValue;
  if ((long)rows * cols == 0) return hipSuccess;
  if ((rows + 31) / 32 > 2147483647 || (cols + 31) / 32 > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL((conj_transpose_kernel<T>), dim3((rows + 31) / 32, (cols + 31) / 32), dim3(256), 0, st, rows, cols,
                     A, lda, B, ldb);
  return hipGetLastError();
}

// Reversal of the rows (fr) and/or columns (fc) of A (rows x cols).  The
// element (r, c) trades places with (fr ? rows-1-r : r, fc ? cols-1-c : c);
// in place, the thread whose element comes first in column-major order
// swaps the pair (self-pairs stay).  Out of place (B != A), B(r, c) takes
// A's partner element.  One thread per element, rows fastest.
template <typename T>
__global__ __launch_bounds__(256) void flip_kernel(int rows, int cols, const T* A, long lda, T* B, long ldb, int fr,
                                                   int fc, int inplace) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)rows * cols) return;
  const int r = (int)(e % rows), c = (int)(e / rows);
  const int pr = fr ? rows - 1 - r : r, pc = fc ? cols - 1 - c : c;
  if (inplace) {
    const long long pe = (long long)pr + (long long)pc * rows;
    if (pe <= e) return;
    T* a = const_cast<T*>(A);
    T* x = a + (size_t)r + (size_t)c * lda;
    T* y = a + (size_t)pr + (size_t)pc * lda;
    const T t = *x;
    *x = *y;
    *y = t;
  } else {
    B[(size_t)r + (size_t)c * ldb] = A[(size_t)pr + (size_t)pc * lda];
  }
}

template <typename T>
hipError_t launch_flip(int rows, int cols, T* A, long lda, bool fr, bool fc, hipStream_t st) {
  if (rows < 0 || cols < 0 || lda < std::max(1, rows)) return hipErrorInvalidValue;
  const long long total = (long long)rows * cols;
  if (total == 0 || (!fr && !fc)) return hipSuccess;
  const long long blocks = (total + 255) / 256;
  if (blocks > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((flip_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, rows, cols, (const T*)A, lda, A, lda,
                     (int)fr, (int)fc, 1);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_flip_copy(int rows, int cols, const T* A, long lda, T* B, long ldb, bool fr, bool fc, hipStream_t st) {
  if (rows < 0 || cols < 0 || lda < std::max(1, rows) || ldb < std::max(1, rows)) return hipErrorInvalidValue;
  const long long total = (long long)rows * cols;
  if (total == 0) return hipSuccess;
  const long long blocks = (total + 255) / 256;
  if (blocks > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((flip_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, rows, cols, A, lda, B, ldb, (int)fr,
                     (int)fc, 0);
  return hipGetLastError();
}

// Hessenberg::GenerateQ's set-up (RNP/LA/Hessenberg.hpp:555-580) on the
// n x n A: identity outside rows/columns [ilo+1, ihi), and inside that
// block the reflectors shifted one column right (V(i, j) = old A(i, j-1),
// from the copy S of the old block A[ilo+1.., ilo..], below the diagonal;
// zero on and above it).
template <typename T>
__global__ __launch_bounds__(256) void hessq_setup_kernel(int n, int ilo, int ihi, T* A, long lda, const T* Sv,
                                                          long lds) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n);
  const bool in_i = (i > ilo && i < ihi), in_j = (j > ilo && j < ihi);
  T v = S<T>::zero();
  if (in_i && in_j) {
    if (i > j) v = Sv[(size_t)(i - ilo - 1) + (size_t)(j - 1 - ilo) * lds];
  } else if (i == j) {
    v = S<T>::one();
  }
  A[(size_t)i + (size_t)j * lda] = v;
}

template <typename T>
hipError_t launch_hessq_setup(int n, int ilo, int ihi, T* A, long lda, const T* Sv, long lds, hipStream_t st) {
  if (n < 0 || ilo < 0 || ilo >= ihi || ihi > n || lda < std::max(1, n) || lds < std::max(1, ihi - ilo - 1))
    return hipErrorInvalidValue;
  const long long total = (long long)n * n;
  if (total == 0) return hipSuccess;
  const long long blocks = (total + 255) / 256;
  if (blocks > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((hessq_setup_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, n, ilo, ihi, A, lda, Sv, lds);
  return hipGetLastError();
}

#define RNPX_INST(T)                                                                  \
  template hipError_t launch_set<T>(int, int, T, T, T*, long, hipStream_t);           \
  template hipError_t launch_copy<T>(int, int, const T*, long, T*, long, hipStream_t); \
  template hipError_t launch_conj_transpose<T>(int, int, const T*, long, T*, long, hipStream_t); \
  template hipError_t launch_flip<T>(int, int, T*, long, bool, bool, hipStream_t);    \
  template hipError_t launch_hessq_setup<T>(int, int, int, T*, long, const T*, long, hipStream_t); \
  template hipError_t launch_flip_copy<T>(int, int, const T*, long, T*, long, bool, bool, hipStream_t); \
  template hipError_t launch_extract_v<T>(int, int, const T*, long, T*, long, hipStream_t); \
  template hipError_t launch_zero_rows<T>(int, int, T*, long, hipStream_t);           \
  template hipError_t launch_larnv<T>(int, const int*, long long, T*, hipStream_t);
RNPX_INST(double)
RNPX_INST(zd)
#undef RNPX_INST

}  // namespace rnpx

namespace rnpx {

// Back substitution R X = B for an upper-triangular n x n R (BLAS xtrsm
// "L","U","N","N", the SolveTrM of RNP/BLAS/BLAS_mix.hpp:2096, in place on the
// first n rows of B).  One workgroup per right-hand side; 64-row blocks from
// the bottom: wave 0 solves the diagonal block with the row values in its
// lanes (each x_i broadcast by v_readlane), then all threads update the rows
// above with the block's contribution.
template <typename T>
__global__ __launch_bounds__(256) void trsm_un_kernel(int n, const T* __restrict__ R, long ldr, T* __restrict__ B,
                                                      long ldb) {
  T* b = B + (size_t)blockIdx.x * ldb;
  __shared__ T xs[64];
  const int t = threadIdx.x, lane = t & 63;
  for (int i1 = n; i1 > 0; i1 -= 64) {
    const int i0 = i1 > 64 ? i1 - 64 : 0;
    const int bs = i1 - i0;
    if (t < 64) {
      T y = S<T>::zero();
      if (lane < bs) y = b[i0 + lane];
      for (int i = bs - 1; i >= 0; --i) {
        const T rii = R[(size_t)(i0 + i) + (size_t)(i0 + i) * ldr];
        T yi = readlane_t(y, i);
        yi = S<T>::cplx ? yi * recip_t(rii) : S<T>::make(S<T>::re(yi) / S<T>::re(rii), 0.0);
        if (lane == i) y = yi;
        if (lane < i) y = y - R[(size_t)(i0 + lane) + (size_t)(i0 + i) * ldr] * yi;
      }
      if (lane < bs) {
        xs[lane] = y;
        b[i0 + lane] = y;
      }
    }
    __syncthreads();
    for (int r = t; r < i0; r += 256) {
      T s = S<T>::zero();
      for (int j = 0; j < bs; ++j) s = s + R[(size_t)r + (size_t)(i0 + j) * ldr] * xs[j];
      b[r] = b[r] - s;
    }
    __syncthreads();
  }
}

template <typename T>
hipError_t launch_trsm_un(int n, int nrhs, const T* R, long ldr, T* B, long ldb, hipStream_t st) {
  if (n < 0 || nrhs < 0 || ldr < n || ldb < n) return hipErrorInvalidValue;
  if (n == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL((trsm_un_kernel<T>), dim3(nrhs), dim3(256), 0, st, n, R, ldr, B, ldb);
  return hipGetLastError();
}
template hipError_t launch_trsm_un<double>(int, int, const double*, long, double*, long, hipStream_t);
template hipError_t launch_trsm_un<zd>(int, int, const zd*, long, zd*, long, hipStream_t);

}  // namespace rnpx
