// rnpx_gemv.hip -- products with one to four right-hand sides,
//   Y = alpha op(A) X + beta Y,  op(A) in {N, T, C}, X with N <= 4 columns,
// for the least-squares paths (Q^H b inside the factorization, the residuals
// r = b - A x and s = A^H r of the refinement, the products with R^{-1}).
// These are HBM-bound: every entry of A is read once, coalesced, by enough
// workgroups to keep the memory system busy (the MFMA tiles would spend 63 of
// every 64 lanes on padding and run a long chain of K slices on few
// workgroups).  Each workgroup writes partial sums over its slice of the
// reduction index; a second kernel adds the slices in a fixed order and
// applies alpha and beta, so results do not depend on timing.
//   N: workgroup (row block of 256, K chunk): thread i walks its row over the
//      chunk, X(k, :) broadcast (uniform loads).
//   T/C: workgroup (column j of A, row chunk): the threads stride down the
//      column, then a workgroup reduction.
#include "rnpx_internal.hpp"

#include <algorithm>

namespace rnpx {

namespace gv {

template <typename T, int NR>
__global__ __launch_bounds__(256) void gemv_n_kernel(int M, int K, int N, const T* __restrict__ A, long lda,
                                                     const T* __restrict__ X, long ldx, T* __restrict__ part, int kc) {
  const int i = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  const int k0 = c * kc, k1 = min(K, k0 + kc);
  T acc[NR];
#pragma unroll
  for (int n = 0; n < NR; ++n) acc[n] = S<T>::zero();
  if (i < M) {
#pragma unroll 8
    for (int k = k0; k < k1; ++k) {
      const T a = A[(size_t)i + (size_t)k * lda];
#pragma unroll
      for (int n = 0; n < NR; ++n)
        if (n < N) acc[n] = acc[n] + a * X[(size_t)k + (size_t)n * ldx];
    }
#pragma unroll
    for (int n = 0; n < NR; ++n)
      if (n < N) part[(size_t)c * M * N + (size_t)i + (size_t)n * M] = acc[n];
  }
}

template <typename T, int NR, bool CONJ>
__global__ __launch_bounds__(256) void gemv_t_kernel(int M, int K, int N, const T* __restrict__ A, long lda,
                                                     const T* __restrict__ X, long ldx, T* __restrict__ part, int rc) {
  __shared__ T red[4][NR];
  const int j = blockIdx.x, c = blockIdx.y;
  const int r0 = c * rc, r1 = min(K, r0 + rc);
  T acc[NR];
#pragma unroll
  for (int n = 0; n < NR; ++n) acc[n] = S<T>::zero();
  for (int r = r0 + (int)threadIdx.x; r < r1; r += 256) {
    T a = A[(size_t)r + (size_t)j * lda];
    if constexpr (CONJ) a = S<T>::conj(a);
#pragma unroll
    for (int n = 0; n < NR; ++n)
      if (n < N) acc[n] = acc[n] + a * X[(size_t)r + (size_t)n * ldx];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int n = 0; n < NR; ++n) {
    double re = wave_sum(S<T>::re(acc[n])), im = 0.0;
    if constexpr (S<T>::cplx) im = wave_sum(S<T>::im(acc[n]));
    if (lane == 0) red[wave][n] = S<T>::make(re, im);
  }
  __syncthreads();
  if (threadIdx.x < N)
    part[(size_t)c * M * N + (size_t)j + (size_t)threadIdx.x * M] =
        (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// Y = alpha * (sum of the ns slices) + beta * Y, slices added in order
// (eight loads in flight); 64-thread workgroups, one output each thread, so
// that a vector's few thousand outputs still spread over the CUs.
template <typename T>
__global__ __launch_bounds__(64) void gemv_reduce_kernel(int M, int N, int ns, const T* __restrict__ part, T alpha,
                                                         T beta, T* __restrict__ Y, long ldy) {
  const long e = (long)blockIdx.x * 64 + threadIdx.x;
  if (e >= (long)M * N) return;
  const int i = (int)(e % M), n = (int)(e / M);
  const T s = sum_strided(part + e, ns, (size_t)M * N);
  T* yp = Y + (size_t)i + (size_t)n * ldy;
  *yp = S<T>::is_zero(beta) ? alpha * s : alpha * s + beta * (*yp);
}

template <typename T, int NR>
static hipError_t launch(char ta, int M, int N, int K, const T* A, long lda, const T* X, long ldx, T* part, int ns,
                         int chunk, hipStream_t st) {
  if (ta == 'N') {
    hipLaunchKernelGGL((gemv_n_kernel<T, NR>), dim3((unsigned)((M + 255) / 256), (unsigned)ns), dim3(256), 0, st, M,
                       K, N, A, lda, X, ldx, part, chunk);
  } else if (ta == 'C' && S<T>::cplx) {
    hipLaunchKernelGGL((gemv_t_kernel<T, NR, true>), dim3((unsigned)M, (unsigned)ns), dim3(256), 0, st, M, K, N, A,
                       lda, X, ldx, part, chunk);
  } else {
    hipLaunchKernelGGL((gemv_t_kernel<T, NR, false>), dim3((unsigned)M, (unsigned)ns), dim3(256), 0, st, M, K, N, A,
                       lda, X, ldx, part, chunk);
  }
  return hipGetLastError();
}

}  // namespace gv

// Y (M x N, N <= 4) = alpha op(A) X + beta Y on stream st; op(A) is M x K.
template <typename T>
int gemv_dev(Handle* h, char ta, int M, int N, int K, T alpha, const T* A, long lda, const T* X, long ldx, T beta,
             T* Y, long ldy, hipStream_t st) {
  if (N < 1 || N > 4 || M < 0 || K < 0) return RNPX_ERR_UNSUPPORTED;
  if (ta != 'N' && ta != 'T' && ta != 'C') return RNPX_ERR_UNSUPPORTED;
  if (M == 0) return RNPX_OK;
  // the leading dimensions the MFMA GEMM requires, with its status
  if (ldy < M || (ta == 'N' ? lda < M : lda < K) || ldx < K) return RNPX_ERR_UNSUPPORTED;
  int ns = 0, chunk = 0;
  if (K > 0) {
    if (ta == 'N') {
      // row blocks of 256 x K chunks of >= 32: about 4096 workgroups
      const int rb = (M + 255) / 256;
      ns = std::max(1, std::min((K + 31) / 32, (4096 + rb - 1) / rb));
    } else {
      // columns x row chunks of >= 2048 rows: about 4096 workgroups
      ns = std::max(1, std::min((K + 2047) / 2048, (4096 + M - 1) / M));
    }
    chunk = (K + ns - 1) / ns;
    ns = (K + chunk - 1) / chunk;  // no empty slice
    RNPX_TRY(h->gemvbuf.ensure(sizeof(T) * (size_t)ns * M * N, h));
    T* part = (T*)h->gemvbuf.p;
    hipError_t e;
    if (N == 1)
      e = gv::launch<T, 1>(ta, M, N, K, A, lda, X, ldx, part, ns, chunk, st);
    else if (N == 2)
      e = gv::launch<T, 2>(ta, M, N, K, A, lda, X, ldx, part, ns, chunk, st);
    else
      e = gv::launch<T, 4>(ta, M, N, K, A, lda, X, ldx, part, ns, chunk, st);
    RNPX_TRY(e);
  }
  // K = 0: Y = beta Y (no slices; beta = 0 clears without reading Y)
  const long tot = (long)M * N;
  hipLaunchKernelGGL((gv::gemv_reduce_kernel<T>), dim3((unsigned)((tot + 63) / 64)), dim3(64), 0, st, M, N, ns,
                     (const T*)h->gemvbuf.p, alpha, beta, Y, ldy);
  RNPX_TRY(hipGetLastError());
  return RNPX_OK;
}

template int gemv_dev<double>(Handle*, char, int, int, int, double, const double*, long, const double*, long, double,
                              double*, long, hipStream_t);
template int gemv_dev<zd>(Handle*, char, int, int, int, zd, const zd*, long, const zd*, long, zd, zd*, long,
                          hipStream_t);

}  // namespace rnpx
