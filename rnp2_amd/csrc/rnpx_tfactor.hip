// rnpx_tfactor.hip -- the T-factor kernel (the larft analogue,
// Reflector::GenerateBlockTr "F","C", RNP/LA/Reflector.hpp:573-626) and the
// Gram kernel that feeds it when V comes from a stored factorization
// (QR::MultQ / QR::GenerateQ recompute T per block, QR.hpp:395, 557).
//
// RNP builds T column by column: T(0:c,c) = -tau_c T(0:c,0:c) V^H v_c.
// Row j of that recurrence only touches row j of T, so here it runs as:
//   1. leaf blocks of 16 on the diagonal, one thread per row, each thread
//      walking its own row of the leaf with the recurrence above;
//   2. recursive doubling: for adjacent blocks, T12 = -T11 (V1^H V2) T22,
//      which is the same product H1 H2 written block-wise.
// Both use the Gram matrix G = V^H V (strict upper part) and tau, all in LDS.
// A tau_c == 0 column comes out zero, as at Reflector.hpp:584-587.
#include "rnpx_kernels.hpp"

namespace rnpx {

template <typename T, int NB>
__global__ __launch_bounds__(256) void tfactor_kernel(const T* __restrict__ gram, int ldg, int k,
                                                       T* __restrict__ Tout, int ldt, const unsigned* run_if) {
  if (run_if && *run_if == 0u) return;
  constexpr int LEAF = 16;
  __shared__ T Gs[NB][NB + 1];  // Gs[l][c] = G(l,c), l < c; diagonal = tau
  __shared__ T Ts[NB][NB + 1];
  __shared__ T Xs[NB / 2][NB / 2 + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += 256) {
    const int l = e % NB, c = e / NB;
    Gs[l][c] = (l <= c && c < k) ? gram[(size_t)l + (size_t)c * ldg] : S<T>::zero();
    Ts[l][c] = S<T>::zero();
  }
  __syncthreads();
  // 1. leaves
  if (tid < k) {
    const int j = tid;
    const int b1 = min((j / LEAF + 1) * LEAF, k);
    Ts[j][j] = Gs[j][j];
    for (int c = j + 1; c < b1; ++c) {
      T s = S<T>::zero();
      for (int l = j; l < c; ++l) s = s + Ts[j][l] * Gs[l][c];
      Ts[j][c] = (S<T>::zero() - Gs[c][c]) * s;
    }
  }
  __syncthreads();
  // 2. doubling: blocks of h -> 2h
  for (int h = LEAF; h < k; h *= 2) {
    const int npairs = (k + 2 * h - 1) / (2 * h);
    // X = G12 * T22 for every pair: X(r, c) = sum_{l in blk2, l <= c} G(i1+r, l) T(l, c)
    for (int e = tid; e < npairs * h * h; e += 256) {
      const int pr = e / (h * h), rc = e % (h * h);
      const int r = rc % h, cc = rc / h;
      const int i1 = pr * 2 * h, i2 = i1 + h;
      const int c = i2 + cc;
      T s = S<T>::zero();
      if (i2 < k && c < k) {
        for (int l = i2; l <= c; ++l) s = s + Gs[i1 + r][l] * Ts[l][c];
      }
      // Xs is (NB/2) x (NB/2); pairs are packed along the rows
      if (pr * h + r < NB / 2 && cc < NB / 2) Xs[pr * h + r][cc] = s;
    }
    __syncthreads();
    // T12 = -T11 * X
    for (int e = tid; e < npairs * h * h; e += 256) {
      const int pr = e / (h * h), rc = e % (h * h);
      const int r = rc % h, cc = rc / h;
      const int i1 = pr * 2 * h, i2 = i1 + h;
      const int c = i2 + cc;
      if (i2 < k && c < k) {
        T s = S<T>::zero();
        for (int l = r; l < h; ++l) s = s + Ts[i1 + r][i1 + l] * Xs[pr * h + l][cc];
        Ts[i1 + r][c] = S<T>::zero() - s;
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < NB * NB; e += 256) {
    const int l = e % NB, c = e / NB;
    Tout[(size_t)l + (size_t)c * ldt] = Ts[l][c];
  }
}

template <typename T>
hipError_t launch_tfactor(const T* gram, int ldg, int k, int nb, T* Tout, int ldt, hipStream_t st,
                          const unsigned* run_if) {
  if (k < 0 || k > nb || ldg < nb || ldt < nb) return hipErrorInvalidValue;
  if (nb == 16)
    hipLaunchKernelGGL((tfactor_kernel<T, 16>), dim3(1), dim3(256), 0, st, gram, ldg, k, Tout, ldt, run_if);
  else if (nb == 32)
    hipLaunchKernelGGL((tfactor_kernel<T, 32>), dim3(1), dim3(256), 0, st, gram, ldg, k, Tout, ldt, run_if);
  else if (nb == 64)
    hipLaunchKernelGGL((tfactor_kernel<T, 64>), dim3(1), dim3(256), 0, st, gram, ldg, k, Tout, ldt, run_if);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// Gram stage 1: chunk s of the rows, one workgroup per (chunk, column c):
// part[s](j,c) = sum_{r in chunk} conj(V(r,j)) V(r,c) for j <= c.
template <typename T>
__global__ __launch_bounds__(256) void gram_partial_kernel(int n, int k, const T* __restrict__ V, long ldv,
                                                           int rows_per_chunk, int nb, T* __restrict__ part) {
  const int s = blockIdx.x, c = blockIdx.y;
  const int r0 = s * rows_per_chunk, r1 = min(n, r0 + rows_per_chunk);
  __shared__ T red[4][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* out = part + (size_t)s * nb * nb + (size_t)c * nb;
  for (int j0 = 0; j0 <= c; j0 += 64) {
    // each wave takes every 4th row; lanes own column j0 + lane
    const int j = j0 + lane;
    T acc = S<T>::zero();
    if (j <= c) {
      for (int r = r0 + wave; r < r1; r += 4)
        acc = acc + S<T>::cmul(V[(size_t)r + (size_t)j * ldv], V[(size_t)r + (size_t)c * ldv]);
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && j <= c) out[j] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    __syncthreads();
  }
}

template <typename T>
__global__ void gram_final_kernel(int k, int nb, int nchunks, const T* __restrict__ part,
                                  const T* __restrict__ tau, T* __restrict__ gram) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb * nb) return;
  const int j = e % nb, c = e / nb;
  T s = S<T>::zero();
  if (c < k && j < c) {
    s = sum_strided(part + (size_t)c * nb + j, nchunks, (size_t)nb * nb);
  } else if (c < k && j == c) {
    s = tau[c];
  }
  gram[(size_t)j + (size_t)c * nb] = s;
}

template <typename T>
hipError_t launch_gram(int n, int k, const T* V, long ldv, const T* tau, T* gram, int nb,
                       T* scratch, int nchunks, hipStream_t st) {
  if (k < 1 || k > nb || n < k || nchunks < 1) return hipErrorInvalidValue;
  const int rpc = (n + nchunks - 1) / nchunks;
  hipLaunchKernelGGL((gram_partial_kernel<T>), dim3(nchunks, k), dim3(256), 0, st, n, k, V, ldv, rpc, nb, scratch);
  hipLaunchKernelGGL((gram_final_kernel<T>), dim3((nb * nb + 255) / 256), dim3(256), 0, st, k, nb, nchunks,
                     scratch, tau, gram);
  return hipGetLastError();
}

template hipError_t launch_tfactor<double>(const double*, int, int, int, double*, int, hipStream_t,
                                           const unsigned*);
template hipError_t launch_tfactor<zd>(const zd*, int, int, int, zd*, int, hipStream_t, const unsigned*);
template hipError_t launch_gram<double>(int, int, const double*, long, const double*, double*, int, double*, int,
                                        hipStream_t);
template hipError_t launch_gram<zd>(int, int, const zd*, long, const zd*, zd*, int, zd*, int, hipStream_t);

}  // namespace rnpx
