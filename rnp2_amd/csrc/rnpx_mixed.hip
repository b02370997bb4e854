// rnpx_mixed.hip -- the mixed-precision pieces of the least-squares form:
//
//  * gemm32: C = alpha op(A) op(B) + beta C with FP64 operands in memory but
//    the product on FP32 MFMA (v_mfma_f32_16x16x4f32, FP32 accumulation):
//    slices are rounded to FP32 when staged into LDS, the epilogue adds the
//    FP32 product to C in FP64.  Used for the two trailing-update products
//    of the mixed factorization (W = V^H C and C -= V Z); split-K slices are
//    reduced in FP64.
//  * triangular solves with R and R^H, residuals and the solution update in
//    FP64 (rnpx_ls.cpp drives the refinement).
#include "rnpx_kernels.hpp"

namespace rnpx {

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int BM, int BN, int BK, char TA, char TB>
struct Tile32 {
  static constexpr int PAD = 4;
  static constexpr int NA = BM * BK / 256;
  static constexpr int NBv = BN * BK / 256;
  __device__ static inline void fetchA(float (&r)[NA], const double* A, long lda, int M, int K, int m0, int k0,
                                       int tid) {
    if constexpr (TA == 'N') {
      constexpr int KS = 256 / BM;
      const int i = tid % BM, kb = tid / BM;
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        const int gi = m0 + i, gk = k0 + kb + KS * s;
        r[s] = (gi < M && gk < K) ? (float)A[(size_t)gi + (size_t)gk * lda] : 0.0f;
      }
    } else {
      constexpr int IS = 256 / BK;
      const int k = tid % BK, ib = tid / BK;
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        const int gi = m0 + ib + IS * s, gk = k0 + k;
        r[s] = (gi < M && gk < K) ? (float)A[(size_t)gk + (size_t)gi * lda] : 0.0f;
      }
    }
  }
  __device__ static inline void storeA(const float (&r)[NA], float (*As)[BM + PAD], int tid) {
    if constexpr (TA == 'N') {
      constexpr int KS = 256 / BM;
      const int i = tid % BM, kb = tid / BM;
#pragma unroll
      for (int s = 0; s < NA; ++s) As[kb + KS * s][i] = r[s];
    } else {
      constexpr int IS = 256 / BK;
      const int k = tid % BK, ib = tid / BK;
#pragma unroll
      for (int s = 0; s < NA; ++s) As[k][ib + IS * s] = r[s];
    }
  }
  __device__ static inline void fetchB(float (&r)[NBv], const double* B, long ldb, int N, int K, int n0, int k0,
                                       int tid) {
    if constexpr (TB == 'N') {
      constexpr int JS = 256 / BK;
      const int k = tid % BK, jb = tid / BK;
#pragma unroll
      for (int s = 0; s < NBv; ++s) {
        const int gj = n0 + jb + JS * s, gk = k0 + k;
        r[s] = (gj < N && gk < K) ? (float)B[(size_t)gk + (size_t)gj * ldb] : 0.0f;
      }
    } else {
      constexpr int KS = 256 / BN;
      const int j = tid % BN, kb = tid / BN;
#pragma unroll
      for (int s = 0; s < NBv; ++s) {
        const int gj = n0 + j, gk = k0 + kb + KS * s;
        r[s] = (gj < N && gk < K) ? (float)B[(size_t)gj + (size_t)gk * ldb] : 0.0f;
      }
    }
  }
  __device__ static inline void storeB(const float (&r)[NBv], float (*Bs)[BN + PAD], int tid) {
    if constexpr (TB == 'N') {
      constexpr int JS = 256 / BK;
      const int k = tid % BK, jb = tid / BK;
#pragma unroll
      for (int s = 0; s < NBv; ++s) Bs[k][jb + JS * s] = r[s];
    } else {
      constexpr int KS = 256 / BN;
      const int j = tid % BN, kb = tid / BN;
#pragma unroll
      for (int s = 0; s < NBv; ++s) Bs[kb + KS * s][j] = r[s];
    }
  }
};

// FP32 MFMA layout (measured on gfx950): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator item r of lane l holds D[4*(l/16)+r][l%16].
template <char TA, char TB, int BM, int BN>
__global__ __launch_bounds__(256) void gemm32_kernel(int M, int N, int K, double alpha, const double* __restrict__ A,
                                                     long lda, const double* __restrict__ B, long ldb, double beta,
                                                     double* __restrict__ C, long ldc, int kslice,
                                                     double* __restrict__ part) {
  constexpr int BK = 32;
  using Tl = Tile32<BM, BN, BK, TA, TB>;
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ float As[BK][BM + Tl::PAD];
  __shared__ float Bs[BK][BN + Tl::PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kslice;
  const int kend = min(K, kbeg + kslice);
  f32x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float ra[Tl::NA], rb[Tl::NBv];
  if (kbeg < kend) {
    Tl::fetchA(ra, A, lda, M, kend, m0, kbeg, tid);
    Tl::fetchB(rb, B, ldb, N, kend, n0, kbeg, tid);
  }
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    Tl::storeA(ra, As, tid);
    Tl::storeB(rb, Bs, tid);
    __syncthreads();
    if (k0 + BK < kend) {
      Tl::fetchA(ra, A, lda, M, kend, m0, k0 + BK, tid);
      Tl::fetchB(rb, B, ldb, N, kend, n0, k0 + BK, tid);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[kk + (lane >> 4)][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + (lane >> 4)][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int gi = m0 + wm * (BM / 2) + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * (BN / 2) + j * 16 + 4 * (lane >> 4) + r;
        if (gi < M && gj < N) {
          const double v = (double)acc[i][j][r];
          if (part) {
            part[(size_t)blockIdx.z * M * N + (size_t)gi + (size_t)gj * M] = v;
          } else {
            double* cp = C + (size_t)gi + (size_t)gj * ldc;
            *cp = (beta == 0.0) ? alpha * v : alpha * v + beta * (*cp);
          }
        }
      }
    }
  }
}

__global__ void reduce64_kernel(int M, int N, int ns, const double* __restrict__ part, double alpha, double beta,
                                double* __restrict__ C, long ldc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)M * N) return;
  const int i = (int)(e % M), j = (int)(e / M);
  const double s = sum_strided(part + e, ns, (size_t)M * N);
  double* cp = C + (size_t)i + (size_t)j * ldc;
  *cp = (beta == 0.0) ? alpha * s : alpha * s + beta * (*cp);
}

template <char TA, char TB>
static void launch32(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                     double beta, double* C, long ldc, int kslice, int ns, double* part, hipStream_t st) {
  dim3 grid((M + 63) / 64, (N + 63) / 64, ns);
  hipLaunchKernelGGL((gemm32_kernel<TA, TB, 64, 64>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C,
                     ldc, kslice, part);
}

hipError_t gemm32(char ta, char tb, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                  long ldb, double beta, double* C, long ldc, int ksplit, double* part, hipStream_t st) {
  if (M < 0 || N < 0 || K < 0) return hipErrorInvalidValue;
  if (M == 0 || N == 0) return hipSuccess;
  if (ldc < M) return hipErrorInvalidValue;
  if (ta == 'C') ta = 'T';
  if (tb == 'C') tb = 'T';
  if ((ta != 'N' && ta != 'T') || (tb != 'N' && tb != 'T')) return hipErrorInvalidValue;
  if (ta == 'N' ? lda < M : lda < K) return hipErrorInvalidValue;
  if (tb == 'N' ? ldb < K : ldb < N) return hipErrorInvalidValue;
  if (ksplit < 1) ksplit = 1;
  if (ksplit > 1 && !part) return hipErrorInvalidValue;
  int kslice = K, ns = 1;
  if (ksplit > 1 && K > 0) {
    kslice = (K + ksplit - 1) / ksplit;
    kslice = ((kslice + 31) / 32) * 32;
    ns = (K + kslice - 1) / kslice;
  }
  double* pp = (ns > 1) ? part : nullptr;
  // N/T times N: the double-buffered FP32-MFMA kernel (rnpx_dgemm.hip)
  {
    const hipError_t e = sgemm_mfma(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);
    if (e != hipErrorNotSupported) {
      if (e != hipSuccess) return e;
      if (ns > 1) {
        const long tot = (long)M * N;
        hipLaunchKernelGGL(reduce64_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M, N, ns, part,
                           alpha, beta, C, ldc);
      }
      return hipGetLastError();
    }
  }
  if (ta == 'N' && tb == 'N') launch32<'N', 'N'>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);
  else if (ta == 'N') launch32<'N', 'T'>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);
  else if (tb == 'N') launch32<'T', 'N'>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);
  else launch32<'T', 'T'>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);
  if (ns > 1) {
    const long tot = (long)M * N;
    hipLaunchKernelGGL(reduce64_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M, N, ns, part, alpha,
                       beta, C, ldc);
  }
  return hipGetLastError();
}

// Forward substitution R^H Y = S for upper-triangular R (xtrsm "L","U","C","N"),
// in place on the first n rows of S; one workgroup per right-hand side, the
// mirror image of trsm_un_kernel.
template <typename T>
__global__ __launch_bounds__(256) void trsm_uc_kernel(int n, const T* __restrict__ R, long ldr, T* __restrict__ B,
                                                      long ldb) {
  T* b = B + (size_t)blockIdx.x * ldb;
  __shared__ T xs[64];
  const int t = threadIdx.x, lane = t & 63;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int bs = min(64, n - i0);
    if (t < 64) {
      T y = S<T>::zero();
      if (lane < bs) y = b[i0 + lane];
      for (int i = 0; i < bs; ++i) {
        const T rii = S<T>::conj(R[(size_t)(i0 + i) + (size_t)(i0 + i) * ldr]);
        T yi = readlane_t(y, i);
        yi = S<T>::cplx ? yi * recip_t(rii) : S<T>::make(S<T>::re(yi) / S<T>::re(rii), 0.0);
        if (lane == i) y = yi;
        // R^H(i0+lane, i0+i) = conj(R(i0+i, i0+lane)) for lane > i
        if (lane > i && lane < bs) y = y - S<T>::conj(R[(size_t)(i0 + i) + (size_t)(i0 + lane) * ldr]) * yi;
      }
      if (lane < bs) {
        xs[lane] = y;
        b[i0 + lane] = y;
      }
    }
    __syncthreads();
    // rows below the block: y_r -= sum_j conj(R(j, r)) x_j, j in the block
    for (int r = i0 + bs + t; r < n; r += 256) {
      T s = S<T>::zero();
      for (int j = 0; j < bs; ++j) s = s + S<T>::conj(R[(size_t)(i0 + j) + (size_t)r * ldr]) * xs[j];
      b[r] = b[r] - s;
    }
    __syncthreads();
  }
}

template <typename T>
hipError_t launch_trsm_uc(int n, int nrhs, const T* R, long ldr, T* B, long ldb, hipStream_t st) {
  if (n < 0 || nrhs < 0 || ldr < n || ldb < n) return hipErrorInvalidValue;
  if (n == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL((trsm_uc_kernel<T>), dim3(nrhs), dim3(256), 0, st, n, R, ldr, B, ldb);
  return hipGetLastError();
}
template hipError_t launch_trsm_uc<double>(int, int, const double*, long, double*, long, hipStream_t);
template hipError_t launch_trsm_uc<zd>(int, int, const zd*, long, zd*, long, hipStream_t);

// Column norms / sums for the refinement's stopping test and the reported
// normwise backward error: out[j] = sum_i |X(i,j)|^2 for j < n (one block
// per column, fixed-order tree reduction).
__global__ __launch_bounds__(256) void colnorm2_kernel(int m, const double* __restrict__ X, long ldx,
                                                       double* __restrict__ out) {
  __shared__ double sh[4];
  const double* x = X + (size_t)blockIdx.x * ldx;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) s += x[i] * x[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

hipError_t launch_colnorm2(int m, int n, const double* X, long ldx, double* out, hipStream_t st) {
  if (m < 0 || n < 0 || ldx < m) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(colnorm2_kernel, dim3(n), dim3(256), 0, st, m, X, ldx, out);
  return hipGetLastError();
}

// y = y + x (elementwise, m x n blocks)
__global__ void axpy_mat_kernel(int m, int n, const double* __restrict__ X, long ldx, double* __restrict__ Y,
                                long ldy) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  Y[(size_t)i + (size_t)j * ldy] += X[(size_t)i + (size_t)j * ldx];
}
hipError_t launch_axpy_mat(int m, int n, const double* X, long ldx, double* Y, long ldy, hipStream_t st) {
  if (m < 0 || n < 0 || ldx < m || ldy < m) return hipErrorInvalidValue;
  if ((long)m * n == 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_mat_kernel, dim3((unsigned)(((long)m * n + 255) / 256)), dim3(256), 0, st, m, n, X, ldx, Y,
                     ldy);
  return hipGetLastError();
}

}  // namespace rnpx
