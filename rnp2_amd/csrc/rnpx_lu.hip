// rnpx_lu.hip -- LU with partial pivoting (RNP/LA/LU.hpp) on gfx950.
//
// LU::Factor (LU.hpp:77-115) works in panels of 64 columns.  Each panel is
// factored by one cooperative launch (lu_panel_kernel) that keeps the panel
// resident: workgroup g holds rows [g RW, g RW + RW) of it in LDS.  For each
// column c (Factor_unblocked, LU.hpp:52-75):
//   1. every workgroup finds its candidate pivot (BLAS::MaximumIndex,
//      BLAS_mix.hpp:798-810: the first row of largest |re|+|im|) and writes
//      it, with that row, to a per-column-parity slot; the owner of row c
//      also writes row c;
//   2. every workgroup waits until all have arrived and reduces the
//      candidates itself, in workgroup order (ties to the lower row, so the
//      result is the sequential scan's);
//   3. every workgroup swaps (the owner of row c takes the pivot row, the
//      owner of the pivot row takes old row c), scales its part of column c
//      by 1/pivot and applies the rank-1 update to its rows.
// One grid hand-off per column; waits are bounded (2 s, abort flag).  The
// row interchanges are then applied to the columns outside the panel
// (lu_laswp_kernel: the panel's swaps composed into one permutation of the
// touched rows, gathered through LDS), U12 = L11^-1 A12 by substitution
// (lu_trsm_kernel, L11 in LDS, a column per thread), and the trailing
// update A22 -= L21 U12 is a GEMM.
#include "rnpx_kernels.hpp"

#include <cstdlib>

namespace rnpx {

namespace {

constexpr int kLuNB = 64;  // panel width (LU::Tuning::factor_block_size)

__device__ __forceinline__ double norm1_of(double x) { return fabs(x); }
__device__ __forceinline__ double norm1_of(zd x) { return fabs(x.x) + fabs(x.y); }

// (value, row) candidate order: the larger value, on ties the lower row.
// NaN values never win (the sequential scan's "cv > mv" never takes one).
__device__ __forceinline__ bool better(double va, int ia, double vb, int ib) {
  return (va > vb) || (va == vb && ia < ib);
}

template <typename T>
__device__ __forceinline__ T ld_part(const double* p) {
  if constexpr (S<T>::cplx) return mkz(p[0], p[1]);
  else return p[0];
}
template <typename T>
__device__ __forceinline__ void st_part(double* p, T v) {
  if constexpr (S<T>::cplx) {
    p[0] = v.x;
    p[1] = v.y;
  } else {
    p[0] = v;
  }
}

}  // namespace

// Scratch layout (doubles): part[2][G][kLuSlot] candidate slots (value,
// row, then the row's kLuNB entries), rowc[2][2*kLuNB] (row c).
constexpr int kLuSlot = 2 + 2 * kLuNB;

template <typename T, int RW>
__global__ __launch_bounds__(256) void lu_panel_kernel(LuPanelArgs a) {
  constexpr int NB = kLuNB, LDS_S = NB + 1, TE = S<T>::cplx ? 2 : 1;
  __shared__ T Sl[RW][LDS_S];
  __shared__ T ubuf[NB], ocbuf[NB];
  __shared__ double redv[4];
  __shared__ int redi[4], redw[4];
  __shared__ int pw[2];
  RNPX_LU_GLOBAL(a);
  T* A = reinterpret_cast<T*>(a.A);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x, G = gridDim.x;
  const int gr0 = g * RW, mp = a.mp, jb = a.jb;
  const T zero = S<T>::zero();

  // Stage this workgroup's rows (zero padding beyond mp and jb).
  for (int e = tid; e < RW * NB; e += 256) {
    const int r = e % RW, c = e / RW;
    T v = zero;
    if (gr0 + r < mp && c < jb) v = A[(size_t)(gr0 + r) + (size_t)c * a.lda];
    Sl[r][c] = v;
  }
  __syncthreads();

  int zero_last = -1;
  for (int c = 0; c < jb; ++c) {
    const int par = c & 1;
    double* part = a.part + (size_t)par * G * kLuSlot;
    double* rowc = a.rowc + (size_t)par * 2 * NB;
    // 1. local candidate
    double bv = -2.0;
    int bi = 0x7fffffff;
    for (int r = tid; r < RW; r += 256) {
      const int gr = gr0 + r;
      if (gr >= c && gr < mp) {
        double v = norm1_of(Sl[r][c]);
        if (v != v) v = -1.0;  // NaN: never preferred
        if (better(v, gr, bv, bi)) {
          bv = v;
          bi = gr;
        }
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (better(ov, oi, bv, bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      redv[wave] = bv;
      redi[wave] = bi;
    }
    __syncthreads();
    bv = redv[0];
    bi = redi[0];
    for (int w = 1; w < 4; ++w)
      if (better(redv[w], redi[w], bv, bi)) {
        bv = redv[w];
        bi = redi[w];
      }
    // 2. candidate slot (and row c from its owner), then arrive
    double* my = part + (size_t)g * kLuSlot;
    if (tid == 0) {
      my[0] = bv;
      my[1] = (double)bi;
    }
    if (bi != 0x7fffffff)
      for (int q = tid; q < NB; q += 256) st_part<T>(my + 2 + TE * q, Sl[bi - gr0][q]);
    if (c >= gr0 && c < gr0 + RW)
      for (int q = tid; q < NB; q += 256) st_part<T>(rowc + TE * q, Sl[c - gr0][q]);
    // every workgroup waits for all the candidates, then reduces them in
    // workgroup order itself (no second hand-off)
    unsigned* cnt = a.sync + 2 * c;
    wait_stores();
    __syncthreads();
    if (tid == 0) {
      release_agent();
      __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < (unsigned)G) {
        if (__hip_atomic_load(a.abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
        if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
          __hip_atomic_store(a.abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    // Old row c does not depend on the pivot: its loads go out with the
    // candidates' (one round trip fewer per column).
    static_assert(NB <= 256, "a thread per entry of row c");
    T ocv = zero;
    if (tid < NB) ocv = ld_part<T>(rowc + TE * tid);
    {
      double v = -2.0;
      int i = 0x7fffffff, w = -1;
      for (int q = tid; q < G; q += 256) {
        const double* o = part + (size_t)q * kLuSlot;
        const double ov = o[0];
        const int oi = (int)o[1];
        if (better(ov, oi, v, i)) {
          v = ov;
          i = oi;
          w = q;
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(i, off, 64);
        const int ow = __shfl_xor(w, off, 64);
        if (better(ov, oi, v, i)) {
          v = ov;
          i = oi;
          w = ow;
        }
      }
      if (lane == 0) {
        redv[wave] = v;
        redi[wave] = i;
        redw[wave] = w;
      }
      if (tid < NB) ocbuf[tid] = ocv;
      __syncthreads();
      if (tid == 0) {
        double bv2 = redv[0];
        int bi2 = redi[0], bw2 = redw[0];
        for (int q = 1; q < 4; ++q)
          if (better(redv[q], redi[q], bv2, bi2)) {
            bv2 = redv[q];
            bi2 = redi[q];
            bw2 = redw[q];
          }
        // The sequential scan starts from row c: a NaN there is kept.
        const T xc = ocbuf[c];
        const double nc = norm1_of(xc);
        if (nc != nc || bi2 == 0x7fffffff) {
          bi2 = c;
          bw2 = -1;
        }
        pw[0] = bi2;
        pw[1] = bw2;
      }
      __syncthreads();
    }
    const int p = pw[0], w = pw[1];
    // 3. pivot row u: the winner's row, or row c itself (already in ocbuf)
    if (tid < NB) ubuf[tid] = (w >= 0) ? ld_part<T>(part + (size_t)w * kLuSlot + 2 + TE * tid) : ocbuf[tid];
    __syncthreads();
    if (c >= gr0 && c < gr0 + RW)
      for (int q = tid; q < NB; q += 256) Sl[c - gr0][q] = ubuf[q];
    if (p != c && p >= gr0 && p < gr0 + RW)
      for (int q = tid; q < NB; q += 256) Sl[p - gr0][q] = ocbuf[q];
    __syncthreads();
    const T piv = ubuf[c];
    const bool nz = !S<T>::is_zero(piv);
    if (!nz) zero_last = c;
    const T rinv = nz ? recip_t(piv) : S<T>::one();
    // scale column c of the rows below c, then the rank-1 update
    for (int r = tid; r < RW; r += 256) {
      const int gr = gr0 + r;
      if (gr > c && gr < mp && nz) Sl[r][c] = Sl[r][c] * rinv;
    }
    __syncthreads();
    const int ncols = jb - c - 1;
    for (int e = tid; e < RW * ncols; e += 256) {
      const int r = e % RW, cc = c + 1 + e / RW;
      const int gr = gr0 + r;
      if (gr > c && gr < mp) Sl[r][cc] = Sl[r][cc] + Sl[r][c] * (zero - ubuf[cc]);
    }
    if (g == 0 && tid == 0) a.piv[c] = p;
    __syncthreads();
  }
  // Write the panel back.
  for (int e = tid; e < RW * NB; e += 256) {
    const int r = e % RW, c = e / RW;
    if (gr0 + r < mp && c < jb) A[(size_t)(gr0 + r) + (size_t)c * a.lda] = Sl[r][c];
  }
  if (g == 0 && tid == 0) *a.zero_last = zero_last;
}

// Row interchanges of one panel (rows j.. j+jb-1 swapped with piv[i], in
// order) applied to ncols columns starting at column c0 (stride lda).
// Each workgroup takes 64 columns: the swaps are composed into one
// permutation of the touched rows (at most 2 jb), then the rows are
// gathered through LDS.
template <typename T>
__global__ __launch_bounds__(256) void lu_laswp_kernel(int j, int jb, const int* __restrict__ piv, T* A, long lda,
                                                       int c0, int ncols) {
  constexpr int MAXT = 2 * kLuNB;
  __shared__ int rows[MAXT], src[MAXT];
  __shared__ int nt;
  __shared__ T buf[MAXT][65];
  __shared__ int tp[kLuNB], pv[kLuNB], fq[kLuNB];
  const int tid = threadIdx.x;
  // Touched rows: the panel rows j..j+jb-1 (slots 0..jb-1), then each pivot
  // row below them at its first appearance (slots jb.., in that order).
  if (tid < jb) pv[tid] = piv[tid] + j;
  __syncthreads();
  if (tid < jb) {  // first appearance of this pivot's row
    const int p = pv[tid];
    int f = tid;
    if (p >= j + jb)
      for (int q = 0; q < tid; ++q)
        if (pv[q] == p) {
          f = q;
          break;
        }
    fq[tid] = f;
  }
  __syncthreads();
  if (tid < jb) {  // its slot
    const int p = pv[tid];
    int t;
    if (p < j + jb) {
      t = p - j;
    } else {
      int cnt = 0;
      for (int q = 0; q < fq[tid]; ++q) cnt += (pv[q] >= j + jb && fq[q] == q) ? 1 : 0;
      t = jb + cnt;
      if (fq[tid] == tid) rows[t] = p;
    }
    tp[tid] = t;
    rows[tid] = j + tid;
  }
  __syncthreads();
  if (tid == 0) {
    int n = jb;
    for (int i = 0; i < jb; ++i) n = max(n, tp[i] + 1);
    for (int t = 0; t < n; ++t) src[t] = rows[t];
    // the swaps in order, on the slot contents
    for (int i = 0; i < jb; ++i) {
      const int t2 = tp[i];
      if (t2 == i) continue;
      const int x = src[i];
      src[i] = src[t2];
      src[t2] = x;
    }
    nt = n;
  }
  __syncthreads();
  const int cb = c0 + blockIdx.x * 64;
  const int nc = min(64, c0 + ncols - cb);
  const int n = nt;
  for (int e = tid; e < n * 64; e += 256) {
    const int t = e % n, q = e / n;
    if (q < nc) buf[t][q] = A[(size_t)src[t] + (size_t)(cb + q) * lda];
  }
  __syncthreads();
  for (int e = tid; e < n * 64; e += 256) {
    const int t = e % n, q = e / n;
    if (q < nc) A[(size_t)rows[t] + (size_t)(cb + q) * lda] = buf[t][q];
  }
}

// U12 = L11^-1 A12 (L11 the jb x jb unit lower triangle at L, A12 jb x ncols
// at B): a wave per column, lane = row, forward substitution with x_k
// broadcast by v_readlane and L11 read from LDS (each lane its own row).
// A workgroup takes 16 columns.
template <typename T>
__global__ __launch_bounds__(256) void lu_trsm_kernel(int jb, const T* __restrict__ L, long ldl, T* B, long ldb,
                                                      int ncols) {
  constexpr int NB = kLuNB;
  __shared__ T Ls[NB][NB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e % NB, c = e / NB;
    Ls[r][c] = (r < jb && c < jb && r > c) ? L[(size_t)r + (size_t)c * ldl] : S<T>::zero();
  }
  __syncthreads();
  const int cb = blockIdx.x * 16;
  for (int q = wave; q < 16; q += 4) {
    const int col = cb + q;
    if (col >= ncols) break;
    T* b = B + (size_t)col * ldb;
    T x = (lane < jb) ? b[lane] : S<T>::zero();
    for (int k = 0; k < jb - 1; ++k) {
      const T xk = readlane_t(x, k);
      if (lane > k) x = x - Ls[lane][k] * xk;
    }
    if (lane < jb) b[lane] = x;
  }
}

// Row gather b(i, :) = a(src[i], :) for an n x nrhs block (LU::Solve's
// interchanges composed into one permutation on the host).
template <typename T>
__global__ __launch_bounds__(256) void lu_gather_kernel(int n, int nrhs, const int* __restrict__ src,
                                                        const T* __restrict__ X, long ldx, T* __restrict__ Y,
                                                        long ldy) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)n * nrhs) return;
  const int i = (int)(e % n), c = (int)(e / n);
  Y[(size_t)i + (size_t)c * ldy] = X[(size_t)src[i] + (size_t)c * ldx];
}

// RefineSolution's denominators: rw_i = |b_i| + sum_k |A(i,k)| |x_k| for
// op "N" (a thread per row, coalesced down the columns) and
// rw_k = |b_k| + sum_i |A(i,k)| |x_i| for "T"/"C" (a wave per column);
// |.| is |re| + |im| (Traits<T>::norm1).
template <typename T>
__global__ __launch_bounds__(256) void lu_absmv_kernel(int n, int transposed, const T* __restrict__ A, long lda,
                                                       const T* __restrict__ x, const T* __restrict__ b,
                                                       double* __restrict__ rw) {
  if (!transposed) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += norm1_of(A[(size_t)i + (size_t)k * lda]) * norm1_of(x[k]);
    rw[i] = norm1_of(b[i]) + s;
  } else {
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (k >= n) return;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += norm1_of(A[(size_t)i + (size_t)k * lda]) * norm1_of(x[i]);
    s = wave_sum(s);
    if (lane == 0) rw[k] = norm1_of(b[k]) + s;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void lu_axpy_kernel(int n, const T* __restrict__ x, T* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = y[i] + x[i];
}

template <typename T>
hipError_t launch_lu_absmv(int n, bool transposed, const T* A, long lda, const T* x, const T* b, double* rw,
                           hipStream_t st) {
  if (n < 0 || lda < std::max(1, n)) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (!transposed)
    hipLaunchKernelGGL((lu_absmv_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, st, n, 0, A, lda, x, b, rw);
  else
    hipLaunchKernelGGL((lu_absmv_kernel<T>), dim3((n + 3) / 4), dim3(256), 0, st, n, 1, A, lda, x, b, rw);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lu_axpy(int n, const T* x, T* y, hipStream_t st) {
  if (n < 0) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((lu_axpy_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

int lu_panel_rows(int mp, int max_groups, int cplx) {
  // 128-row blocks, or 256 when 128 would need too many workgroups (measured
  // at 8192^2 with the panels overlapped by the trailing update: 64 rows
  // 72.4 ms, 128 rows 69.9 ms; fewer workgroups contend less with the
  // update's GEMM for CUs); RNPX_LU_PANEL_RW overrides.
  static const int pref = [] {
    const char* e = getenv("RNPX_LU_PANEL_RW");
    return e ? atoi(e) : 0;
  }();
  if (pref == 64 || pref == 128 || (pref == 256 && !cplx)) {
    const int G = (mp + pref - 1) / pref;
    if (G <= max_groups) return pref;
  }
  for (int rw : {128, 256}) {
    if (cplx && rw > 128) break;
    const int G = (mp + rw - 1) / rw;
    if (G <= max_groups) return rw;
  }
  return 0;
}

template <typename T>
hipError_t launch_lu_panel(const LuPanelArgs& a, int rw, hipStream_t st) {
  if (a.mp <= 0 || a.jb <= 0 || a.jb > kLuNB || a.lda < a.mp || !a.sync || !a.abortf || !a.piv || !a.zero_last)
    return hipErrorInvalidValue;
  const int G = (a.mp + rw - 1) / rw;
  if (G < 1 || G > 256) return hipErrorInvalidValue;
  switch (rw) {
    case 64: hipLaunchKernelGGL((lu_panel_kernel<T, 64>), dim3(G), dim3(256), 0, st, a); break;
    case 128: hipLaunchKernelGGL((lu_panel_kernel<T, 128>), dim3(G), dim3(256), 0, st, a); break;
    case 256:
      if constexpr (S<T>::cplx) return hipErrorInvalidValue;
      else hipLaunchKernelGGL((lu_panel_kernel<T, 256>), dim3(G), dim3(256), 0, st, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lu_laswp(int j, int jb, const int* piv, T* A, long lda, int c0, int ncols, hipStream_t st) {
  if (jb <= 0 || jb > kLuNB || j < 0 || c0 < 0 || ncols < 0) return hipErrorInvalidValue;
  if (ncols == 0) return hipSuccess;
  hipLaunchKernelGGL((lu_laswp_kernel<T>), dim3((ncols + 63) / 64), dim3(256), 0, st, j, jb, piv, A, lda, c0, ncols);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lu_trsm(int jb, const T* L, long ldl, T* B, long ldb, int ncols, hipStream_t st) {
  if (jb <= 0 || jb > kLuNB || ncols < 0 || ldl < jb || ldb < jb) return hipErrorInvalidValue;
  if (ncols == 0) return hipSuccess;
  hipLaunchKernelGGL((lu_trsm_kernel<T>), dim3((ncols + 15) / 16), dim3(256), 0, st, jb, L, ldl, B, ldb, ncols);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lu_gather(int n, int nrhs, const int* src, const T* X, long ldx, T* Y, long ldy, hipStream_t st) {
  if (n < 0 || nrhs < 0 || ldx < std::max(1, n) || ldy < std::max(1, n)) return hipErrorInvalidValue;
  const long long total = (long long)n * nrhs;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL((lu_gather_kernel<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, n, nrhs, src, X,
                     ldx, Y, ldy);
  return hipGetLastError();
}

#define RNPX_INST(T)                                                                                    \
  template hipError_t launch_lu_panel<T>(const LuPanelArgs&, int, hipStream_t);                         \
  template hipError_t launch_lu_laswp<T>(int, int, const int*, T*, long, int, int, hipStream_t);        \
  template hipError_t launch_lu_trsm<T>(int, const T*, long, T*, long, int, hipStream_t);               \
  template hipError_t launch_lu_gather<T>(int, int, const int*, const T*, long, T*, long, hipStream_t);      \
  template hipError_t launch_lu_absmv<T>(int, bool, const T*, long, const T*, const T*, double*, hipStream_t); \
  template hipError_t launch_lu_axpy<T>(int, const T*, T*, hipStream_t);
RNPX_INST(double)
RNPX_INST(zd)
#undef RNPX_INST

}  // namespace rnpx
