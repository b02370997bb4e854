// rnpx_trsm.hip -- triangular solves with the QR factor R for the least-squares
// driver and its refinement, in place on B (n x nrhs):
//
//   trans 'N':  R X = B     (back substitution; SolveTrM "L","U","N","N",
//                            RNP/BLAS/BLAS_mix.hpp:2096)
//   trans 'C':  R^H X = B   (forward substitution; the semi-normal
//                            equations' first solve)
//
// R is upper triangular (n x n, leading dimension ldr) -- the upper triangle
// of the factored A.
//
// Blocked and sync-free: one workgroup per (64-row block p, right-hand side
// c).  Workgroup w solves block p = nblk-1-w ('N') or p = w ('C'), so it
// depends only on lower-numbered workgroups of the same right-hand side,
// which the dispatcher started before it: every wait ends.  A workgroup
// stages its diagonal block in LDS while it waits, adds the off-diagonal
// products R_pq x_q in a fixed order as the blocks q it needs are published
// (per-block flags holding the call's epoch), then one wave solves the
// diagonal block by substitution (lane = row, the pivot broadcast by
// v_readlane) and publishes x_p.  Quotients are y * (1/r) with one
// correction step, y/r to within an ulp.
#include "rnpx_kernels.hpp"

namespace rnpx {

namespace {

constexpr int TB = 64;  // block rows per workgroup

__device__ __forceinline__ void wait_vm_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <typename T, bool CONJ>
__global__ __launch_bounds__(256) void trsm_blocked_kernel(int n, const T* __restrict__ R, long ldr, T* B, long ldb,
                                                           unsigned* flags, unsigned epoch, unsigned* abortf) {
  __shared__ T Rs[TB][TB + 1];  // N: R_pp;  C: R_pp^H (lower)
  __shared__ T rinv[TB];
  __shared__ T red[4][TB];
  __shared__ int bad;
  const int nblk = (n + TB - 1) / TB;
  const int w = blockIdx.x, c = blockIdx.y;
  const int p = CONJ ? w : nblk - 1 - w;
  const int i0 = p * TB, bs = min(TB, n - i0);
  T* b = B + (size_t)c * ldb;
  unsigned* fl = flags + (size_t)c * nblk;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const T zero = S<T>::zero(), one = S<T>::one();
  if (t == 0) bad = 0;
  // Diagonal block, identity beyond bs (coalesced down R's columns).
  for (int e = t; e < TB * TB; e += 256) {
    const int r = e % TB, q = e / TB;  // R(i0 + r, i0 + q)
    T v = (r == q) ? one : zero;
    if (r < bs && q < bs && r <= q) v = R[(size_t)(i0 + r) + (size_t)(i0 + q) * ldr];
    if (CONJ)
      Rs[q][r] = S<T>::conj(v);
    else
      Rs[r][q] = v;
  }
  __syncthreads();
  if (t < TB) rinv[t] = recip_t(Rs[t][t]);

  // Off-diagonal blocks, in the order they are finished.
  T acc = zero;
  for (int d = 0; d < w; ++d) {
    const int q = CONJ ? d : nblk - 1 - d;
    if (t == 0) {
      const uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(&fl[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
          bad = 1;
          break;
        }
        if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
          __hip_atomic_store(abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          bad = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (bad) return;
    const int j0 = q * TB, bq = min(TB, n - j0);
    if (lane < bs) {
#pragma unroll 4
      for (int jj = 0; jj < TB / 4; ++jj) {
        const int j = wave * (TB / 4) + jj;
        if (j < bq) {
          const T xj = b[j0 + j];
          const T rij = CONJ ? S<T>::conj(R[(size_t)(j0 + j) + (size_t)(i0 + lane) * ldr])
                             : R[(size_t)(i0 + lane) + (size_t)(j0 + j) * ldr];
          acc = acc + rij * xj;
        }
      }
    }
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  T y = zero;
  if (lane < bs) y = b[i0 + lane] - ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
  if (!CONJ) {
#pragma unroll
    for (int i = TB - 1; i >= 0; --i) {
      const T ri = Rs[lane][i];
      const T yi = readlane_t(y, i);
      T xi = yi * rinv[i];
      xi = xi + (yi - xi * Rs[i][i]) * rinv[i];
      if (lane == i) y = xi;
      else if (lane < i) y = y - ri * xi;
    }
  } else {
#pragma unroll
    for (int i = 0; i < TB; ++i) {
      const T ri = Rs[lane][i];
      const T yi = readlane_t(y, i);
      T xi = yi * rinv[i];
      xi = xi + (yi - xi * Rs[i][i]) * rinv[i];
      if (lane == i) y = xi;
      else if (lane > i) y = y - ri * xi;
    }
  }
  if (lane < bs) b[i0 + lane] = y;
  wait_vm_stores();
  if (lane == 0) __hip_atomic_store(&fl[p], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

int trsm_flag_words(int n, int nrhs) { return ((n + TB - 1) / TB) * nrhs; }

template <typename T>
hipError_t launch_trsm_blocked(char trans, int n, int nrhs, const T* R, long ldr, T* B, long ldb, unsigned* flags,
                               unsigned epoch, unsigned* abortf, hipStream_t st) {
  if (n < 0 || nrhs < 0 || ldr < std::max(1, n) || ldb < std::max(1, n) || epoch == 0u || !flags || !abortf)
    return hipErrorInvalidValue;
  if (trans != 'N' && trans != 'C') return hipErrorInvalidValue;
  if (n == 0 || nrhs == 0) return hipSuccess;
  if (nrhs > 65535) return hipErrorInvalidValue;
  const dim3 grid((n + TB - 1) / TB, nrhs);
  if (trans == 'N')
    hipLaunchKernelGGL((trsm_blocked_kernel<T, false>), grid, dim3(256), 0, st, n, R, ldr, B, ldb, flags, epoch, abortf);
  else
    hipLaunchKernelGGL((trsm_blocked_kernel<T, true>), grid, dim3(256), 0, st, n, R, ldr, B, ldb, flags, epoch, abortf);
  return hipGetLastError();
}

template hipError_t launch_trsm_blocked<double>(char, int, int, const double*, long, double*, long, unsigned*, unsigned,
                                                unsigned*, hipStream_t);
template hipError_t launch_trsm_blocked<zd>(char, int, int, const zd*, long, zd*, long, unsigned*, unsigned, unsigned*,
                                            hipStream_t);

}  // namespace rnpx
