// rnpx_trsm.hip -- triangular solves with the QR factor R for the least-squares
// driver and its refinement, in place on B (n x nrhs):
//
//   trans 'N':  R X = B     (back substitution; SolveTrM "L","U","N","N",
//                            RNP/BLAS/BLAS_mix.hpp:2096)
//   trans 'C':  R^H X = B   (forward substitution; the semi-normal
//                            equations' first solve)
//
// R is upper triangular (n x n, leading dimension ldr) -- the upper triangle
// of the factored A.  The same kernel, templated on the triangle, the
// operation (N, T, C) and a unit diagonal, does LU::Solve's four solves
// (SolveTrM "L","L"/"U", LU.hpp:117-144).
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

// op(M) X = B with M the upper (UPPER) or lower triangle of R, op = N, T
// (transpose, OP = 1) or C (conjugate transpose, OP = 2), unit diagonal
// when UNIT.  E = op(M) is upper when UPPER == (OP == 0): backward
// substitution, blocks from the last; forward otherwise.
template <typename T, bool UPPER, int OP, bool UNIT>
__global__ __launch_bounds__(256) void trsm_tri_kernel(int n, const T* __restrict__ R, long ldr, T* B, long ldb,
                                                       unsigned* flags, unsigned epoch, unsigned* abortf) {
  constexpr bool EUP = (UPPER == (OP == 0));
  __shared__ T Rs[TB][TB + 1];  // E's diagonal block, row-major
  __shared__ T rinv[TB];
  __shared__ T red[4][TB];
  __shared__ int bad;
  const int nblk = (n + TB - 1) / TB;
  const int w = blockIdx.x, c = blockIdx.y;
  const int p = EUP ? nblk - 1 - w : w;
  const int i0 = p * TB, bs = min(TB, n - i0);
  T* b = B + (size_t)c * ldb;
  unsigned* fl = flags + (size_t)c * nblk;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const T zero = S<T>::zero(), one = S<T>::one();
  if (t == 0) bad = 0;
  // Diagonal block E(i0 + r, i0 + q), identity beyond bs (read down R's columns).
  for (int e = t; e < TB * TB; e += 256) {
    const int r = e % TB, q = e / TB;  // R(i0 + r, i0 + q)
    T v = (r == q) ? one : zero;
    const bool in_tri = UPPER ? (r <= q) : (r >= q);
    if (r < bs && q < bs && in_tri && !(UNIT && r == q)) v = R[(size_t)(i0 + r) + (size_t)(i0 + q) * ldr];
    if (OP == 0)
      Rs[r][q] = v;
    else
      Rs[q][r] = (OP == 2) ? S<T>::conj(v) : v;
  }
  __syncthreads();
  if (t < TB) rinv[t] = UNIT ? one : recip_t(Rs[t][t]);

  // Off-diagonal blocks, in the order they are finished.
  T acc = zero;
  for (int d = 0; d < w; ++d) {
    const int q = EUP ? nblk - 1 - d : d;
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
          T rij;
          if (OP == 0) {
            rij = R[(size_t)(i0 + lane) + (size_t)(j0 + j) * ldr];
          } else {
            rij = R[(size_t)(j0 + j) + (size_t)(i0 + lane) * ldr];
            if (OP == 2) rij = S<T>::conj(rij);
          }
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
  if (EUP) {
#pragma unroll
    for (int i = TB - 1; i >= 0; --i) {
      const T ri = Rs[lane][i];
      const T yi = readlane_t(y, i);
      T xi = yi * rinv[i];
      if (!UNIT) xi = xi + (yi - xi * Rs[i][i]) * rinv[i];
      if (lane == i) y = xi;
      else if (lane < i) y = y - ri * xi;
    }
  } else {
#pragma unroll
    for (int i = 0; i < TB; ++i) {
      const T ri = Rs[lane][i];
      const T yi = readlane_t(y, i);
      T xi = yi * rinv[i];
      if (!UNIT) xi = xi + (yi - xi * Rs[i][i]) * rinv[i];
      if (lane == i) y = xi;
      else if (lane > i) y = y - ri * xi;
    }
  }
  if (lane < bs) b[i0 + lane] = y;
  wait_vm_stores();
  if (lane == 0) {
    release_agent();
    __hip_atomic_store(&fl[p], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

int trsm_flag_words(int n, int nrhs) { return ((n + TB - 1) / TB) * nrhs; }

template <typename T, bool UPPER, int OP, bool UNIT>
static void launch_tri(const dim3& grid, hipStream_t st, int n, const T* R, long ldr, T* B, long ldb,
                       unsigned* flags, unsigned epoch, unsigned* abortf) {
  hipLaunchKernelGGL((trsm_tri_kernel<T, UPPER, OP, UNIT>), grid, dim3(256), 0, st, n, R, ldr, B, ldb, flags, epoch,
                     abortf);
}

template <typename T>
hipError_t launch_trsm_tri(char uplo, char op, char diag, int n, int nrhs, const T* R, long ldr, T* B, long ldb,
                           unsigned* flags, unsigned epoch, unsigned* abortf, hipStream_t st) {
  if (n < 0 || nrhs < 0 || ldr < std::max(1, n) || ldb < std::max(1, n) || epoch == 0u || !flags || !abortf)
    return hipErrorInvalidValue;
  if ((uplo != 'U' && uplo != 'L') || (op != 'N' && op != 'T' && op != 'C') || (diag != 'N' && diag != 'U'))
    return hipErrorInvalidValue;
  if (n == 0 || nrhs == 0) return hipSuccess;
  if (nrhs > 65535) return hipErrorInvalidValue;
  const dim3 grid((n + TB - 1) / TB, nrhs);
  const bool up = (uplo == 'U'), unit = (diag == 'U');
  const int o = (op == 'N') ? 0 : (op == 'T' ? 1 : 2);
#define RNPX_TRI(U, O, D) \
  if (up == U && o == O && unit == D) { launch_tri<T, U, O, D>(grid, st, n, R, ldr, B, ldb, flags, epoch, abortf); return hipGetLastError(); }
  RNPX_TRI(true, 0, false) RNPX_TRI(true, 1, false) RNPX_TRI(true, 2, false)
  RNPX_TRI(true, 0, true) RNPX_TRI(true, 1, true) RNPX_TRI(true, 2, true)
  RNPX_TRI(false, 0, false) RNPX_TRI(false, 1, false) RNPX_TRI(false, 2, false)
  RNPX_TRI(false, 0, true) RNPX_TRI(false, 1, true) RNPX_TRI(false, 2, true)
#undef RNPX_TRI
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_trsm_blocked(char trans, int n, int nrhs, const T* R, long ldr, T* B, long ldb, unsigned* flags,
                               unsigned epoch, unsigned* abortf, hipStream_t st) {
  if (trans != 'N' && trans != 'C') return hipErrorInvalidValue;
  return launch_trsm_tri<T>('U', trans, 'N', n, nrhs, R, ldr, B, ldb, flags, epoch, abortf, st);
}

// The inverses of the 64 x 64 diagonal blocks of an upper-triangular R
// (the last one may be smaller), written over the same blocks of X (the
// rest of X untouched).  One wave per block, the block staged in LDS; lane j
// back-substitutes column j of the inverse: x_j = 1 / r_jj, then for
// i = j-1 .. 0, x_i = -(sum_{k=i+1..j} r_ik x_k) / r_ii.
__global__ __launch_bounds__(64) void trtri_diag_kernel(int n, const double* __restrict__ R, long ldr,
                                                       double* __restrict__ X, long ldx) {
  __shared__ double Rs[64][65];
  __shared__ double Xs[64][65];
  const int lane = threadIdx.x;
  const int b0 = blockIdx.x * 64, nb = min(64, n - b0);
  for (int j = 0; j < nb; ++j) {
    Rs[lane][j] = lane < nb ? R[(size_t)(b0 + lane) + (size_t)(b0 + j) * ldr] : 0.0;
    Xs[lane][j] = 0.0;
  }
  __syncthreads();
  if (lane < nb) {
    const int j = lane;
    Xs[j][j] = 1.0 / Rs[j][j];
    for (int i = j - 1; i >= 0; --i) {
      double sacc = 0.0;
      for (int k = i + 1; k <= j; ++k) sacc += Rs[i][k] * Xs[k][j];
      Xs[i][j] = -sacc / Rs[i][i];
    }
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j)
    if (lane < nb) X[(size_t)(b0 + lane) + (size_t)(b0 + j) * ldx] = Xs[lane][j];
}

hipError_t launch_trtri_diag(int n, const double* R, long ldr, double* X, long ldx, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(trtri_diag_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, n, R, ldr, X, ldx);
  return hipGetLastError();
}

template hipError_t launch_trsm_blocked<double>(char, int, int, const double*, long, double*, long, unsigned*, unsigned,
                                                unsigned*, hipStream_t);
template hipError_t launch_trsm_blocked<zd>(char, int, int, const zd*, long, zd*, long, unsigned*, unsigned, unsigned*,
                                            hipStream_t);
template hipError_t launch_trsm_tri<double>(char, char, char, int, int, const double*, long, double*, long, unsigned*,
                                            unsigned, unsigned*, hipStream_t);
template hipError_t launch_trsm_tri<zd>(char, char, char, int, int, const zd*, long, zd*, long, unsigned*, unsigned,
                                        unsigned*, hipStream_t);

}  // namespace rnpx
