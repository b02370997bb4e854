// rnpx_kernels.hpp -- launch interfaces of the gfx950 kernels used by the
// host orchestration (rnpx_qr.cpp).  Every launcher checks the shapes it is
// given against what its kernel and grid assume and returns
// hipErrorInvalidValue instead of launching when they do not match.
#pragma once
#include "rnpx_device.hpp"
#include "../../include/rnpx.h"

#define RNPX_LAUNCH_CHECK(x)            \
  do {                                  \
    hipError_t _le = (x);               \
    if (_le != hipSuccess) return _le;  \
  } while (0)

namespace rnpx {

// Panel kernel: at most this many workgroups (each holds 256*RPT rows; one
// 256-thread workgroup per CU at these register counts, all co-resident).
constexpr int kMaxPanelGroups = 256;
// Grid-barrier timeout: 2 s of the 100 MHz wall clock.
constexpr uint64_t kBarrierTimeoutTicks = 200000000ull;

// Agent-scope release of the global stores made before it (this wave's, and
// after a workgroup barrier every wave's): the fence writes back this XCD's
// L2, then a hard wait for that write-back, before the caller's relaxed flag
// or counter update.  The wait must be explicit: with an earlier
// s_waitcnt vmcnt(0) in the source, hipcc (gfx950) merges the release's own
// wait into that earlier one, so the write-back was still in flight when the
// flag became visible and a reader on another XCD could acquire stale data.
__device__ __forceinline__ void release_agent() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}

// Bounded grid barrier: arrive with release, spin on a relaxed load, then an
// acquire fence.  Gives up (and raises the abort flag, which every spinner
// also watches) after ~2 s of wall clock so that no wave can spin forever.
__device__ inline void grid_barrier(unsigned* counter, unsigned* abortf, unsigned target) {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // this wave's stores done
  __syncthreads();
  if (threadIdx.x == 0) {
    release_agent();
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t0 = wall_clock64();
    while (true) {
      unsigned v = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (v >= target) break;
      if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
      if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
        __hip_atomic_store(abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Cross-workgroup hand-offs of the cooperative kernels (CholeskyQR2, LU panel).
// Wait until this wave's global stores have completed (vmcnt(0)).
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The workgroup arrives on an agent-scope counter after all its global
// stores are done (release by thread 0, acquire after); returns the count
// before this arrival.
__device__ __forceinline__ unsigned arrive(unsigned* counter, unsigned* slot) {
  wait_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    release_agent();
    *slot = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return *slot;
}

// Set a flag after all the workgroup's global stores are done.
__device__ __forceinline__ void publish(unsigned* flag) {
  wait_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    release_agent();
    __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Bounded wait for a flag (abort flag and ~2 s wall-clock limit), then
// acquire for the whole workgroup.
__device__ __forceinline__ void wait_flag(unsigned* flag, unsigned* abortf) {
  if (threadIdx.x == 0) {
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
      if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
      if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
        __hip_atomic_store(abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Sum of p[q * stride] over q < n, added in the order of q (the result is
// that of the plain loop), with eight loads in flight instead of one.
template <typename T>
__device__ __forceinline__ T sum_strided(const T* p, int n, size_t stride) {
  T s = S<T>::zero();
  int q = 0;
  for (; q + 8 <= n; q += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(q + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + v[u];
  }
  for (; q < n; ++q) s = s + p[(size_t)q * stride];
  return s;
}

// Bounded wait until an agent-scope counter reaches target (abort flag and
// ~2 s wall-clock limit), then acquire for the whole workgroup.
__device__ __forceinline__ void wait_count(unsigned* counter, unsigned target, unsigned* abortf) {
  if (threadIdx.x == 0) {
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
      if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
        __hip_atomic_store(abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// LU panel (rnpx_lu.hip): one cooperative launch factors an mp x jb panel
// (jb <= 64) with partial pivoting.
struct LuPanelArgs {
  int mp, jb;           // panel rows and columns
  void* A; long lda;    // panel top-left (T = double or zd)
  int* piv;             // out: jb pivot rows, relative to the panel top
  int* zero_last;       // out: the last column with an exact zero pivot, -1 if none
  double* part;         // scratch: 2 * G * (2 + 2*64) doubles (candidates)
  double* rowc;         // scratch: 2 * 2*64 doubles (row c)
  unsigned* sync;       // 2 words per column (an arrival counter used), zero at launch
  unsigned* abortf;     // sticky error flag
};
#define RNPX_LU_GLOBAL(a)                                                                              \
  do {                                                                                                 \
    (a).A = as_global((a).A); (a).piv = as_global((a).piv); (a).zero_last = as_global((a).zero_last); \
    (a).part = as_global((a).part); (a).rowc = as_global((a).rowc);                                    \
    (a).sync = as_global((a).sync); (a).abortf = as_global((a).abortf);                               \
  } while (0)
constexpr int kLuPanelNB = 64;
constexpr int kLuPartPerGroup = 2 + 2 * kLuPanelNB;
// Rows per workgroup for an mp-row panel within max_groups workgroups (64,
// 128 or 256; 0 when the panel is too tall).
int lu_panel_rows(int mp, int max_groups, int cplx);
template <typename T>
hipError_t launch_lu_panel(const LuPanelArgs& a, int rw, hipStream_t st);
template <typename T>
hipError_t launch_lu_laswp(int j, int jb, const int* piv, T* A, long lda, int c0, int ncols, hipStream_t st);
template <typename T>
hipError_t launch_lu_trsm(int jb, const T* L, long ldl, T* B, long ldb, int ncols, hipStream_t st);
template <typename T>
hipError_t launch_lu_gather(int n, int nrhs, const int* src, const T* X, long ldx, T* Y, long ldy, hipStream_t st);
// RefineSolution's |op(A)| |x| + |b| (rw, doubles) and y += x.
template <typename T>
hipError_t launch_lu_absmv(int n, bool transposed, const T* A, long lda, const T* x, const T* b, double* rw,
                           hipStream_t st);
template <typename T>
hipError_t launch_lu_axpy(int n, const T* x, T* y, hipStream_t st);

template <typename T>
struct PanelArgs {
  int mp, nbp;         // panel rows, panel columns (nbp <= NB)
  T* A; long lda;      // panel top-left (in: panel, out: R and V packed)
  T* tau;              // out: tau[0 .. min(mp,nbp))
  T* V; long ldv;      // out: clean V (unit diagonal, zeros above), or null
  T* gram;             // out: NB x NB, gram(j,c) for j < c, tau_c on the diagonal
  T* part;             // scratch: 2 * G * NB elements
  T* rowbuf;           // scratch: 2 * NB elements
  unsigned* counter;   // scratch: arrival counter, zero at launch
  unsigned* abortf;    // sticky error flag: set when a barrier times out
  const unsigned* run_if = nullptr;  // when set, the kernel runs only if *run_if != 0
};

template <typename T>
hipError_t launch_panel(const PanelArgs<T>& a, int nb, int rpt, hipStream_t st);

// Real 64-column panel by CholeskyQR2 + Householder reconstruction
// (rnpx_cqr.hip).  Outputs as the column-wise panel, plus T directly; a
// panel it rejects is refactored column by column inside the same launch
// (one workgroup), with the same outputs and *fail = 1.
// A generic pointer known to address global memory, rewritten through the
// global address space so that the compiler issues global instead of flat
// loads and stores (flat ones also count against the LDS counter, so every
// LDS wait would wait for them too).
template <typename T>
__device__ __forceinline__ T* as_global(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)p;
}
#define RNPX_CQR_GLOBAL(a)                                                                        \
  do {                                                                                            \
    (a).A = as_global((a).A); (a).tau = as_global((a).tau); (a).V = as_global((a).V);              \
    (a).T = as_global((a).T); (a).part = as_global((a).part);                                      \
    (a).gsum = as_global((a).gsum); (a).scratch = as_global((a).scratch);                         \
    (a).sync = as_global((a).sync); (a).abortf = as_global((a).abortf); (a).fail = as_global((a).fail); \
  } while (0)

struct CqrArgs {
  int mp;               // panel rows (>= 128)
  double* A; long lda;  // panel top-left
  double* tau;          // out: 64 scalars
  double* V; long ldv;  // out: clean V (ldv >= mp)
  double* T; long ldt;  // out: 64 x 64 T (ldt >= 64)
  double* part;         // scratch: G * 2560 (partial Gram tiles)
  double* gpart;        // scratch: 16 * 2560 (group sums)
  double* gsum;         // scratch: 2 * 2560 (reduced Gram of each pass)
  double* scratch;      // scratch: kCqrScratch doubles
  unsigned* sync;       // kCqrSyncWords words, zero at launch
  unsigned* abortf;     // sticky error flag
  unsigned* fail;       // zero at launch; 1 = redo this panel column by column
};
constexpr int kCqrSyncWords = 128;
constexpr int kCqrScratch = 3 * 4096 + 64;
constexpr int kCqrMaxGroups = 256;
// Workgroups (and rows per workgroup in *rw) for an mp-row panel, 0 if the
// panel is outside what the kernel takes (mp < 128 or too many rows).
int cqr_panel_groups(int mp, int max_groups, int* rw);
hipError_t launch_cqr_panel(const CqrArgs& a, int G, int rw, hipStream_t st);

// A real outer block of 128 columns (two 64-column panels) in one launch
// (rnpx_cqr.hip): both panels by CholeskyQR2 + reconstruction, panel 1
// applied to panel 2 and the compact-WY forms merged in LDS.  Outputs: R and
// V packed in A (128 columns), tau (128), the block's clean V (ldv >= mp,
// zeros above the second panel) and T (128 x 128, ldt >= 128).  If either
// panel is rejected, A is untouched until workgroup 0 refactors the block
// column by column in the same launch (*fail = 1).  G = ceil(mp/64)
// workgroups, all resident at once (they meet at grid-wide reductions).
struct CqrBlockArgs {
  int mp;               // block rows (>= 192)
  double* A; long lda;  // block top-left
  double* tau;          // out: 128 scalars
  double* V; long ldv;  // out: clean V, mp x 128
  double* T; long ldt;  // out: 128 x 128 T
  double* part;         // scratch: G * 4096 (each workgroup's partial products)
  double* gsum;         // scratch: 6 * 4096 (reduced products of each pass)
  double* scratch;      // scratch: kCqrBlockScratch doubles
  unsigned* sync;       // kCqrBlockSyncWords words, zero at launch
  unsigned* abortf;     // sticky error flag
  unsigned* fail;       // zero at launch; 1 = refactored column by column
};
constexpr int kCqrBlockSyncWords = 256;
constexpr int kCqrBlockScratch = 65728;
// Workgroups for an mp-row block, 0 when the block kernel does not apply.
int cqr_block_groups(int mp, int max_groups);
hipError_t launch_cqr_block(const CqrBlockArgs& a, int G, hipStream_t st);
// Waits (bounded) until the G workgroups of the block kernel with these sync
// words have all started.
hipError_t launch_cqr_gate(const unsigned* block_sync, int G, unsigned* off, hipStream_t st);

// T factor of a forward, columnwise block reflector from its Gram matrix
// (the larft analogue, RNP/LA/Reflector.hpp:573-626): k <= NB reflectors,
// gram(j,c) = V(:,j)^H V(:,c) for j < c and tau_c on the diagonal.
// Writes the NB x NB block of T (zeros below the diagonal and beyond k).
// With run_if set the kernel returns at once unless *run_if != 0.
template <typename T>
hipError_t launch_tfactor(const T* gram, int ldg, int k, int nb, T* Tout, int ldt, hipStream_t st,
                          const unsigned* run_if = nullptr);

// Gram of a clean V and tau into the layout launch_tfactor reads:
// gram(j,c) = sum_r conj(V(r,j)) V(r,c) for j < c, tau_c on the diagonal.
// Deterministic two-stage reduction over row chunks (scratch holds
// nchunks * nb * nb elements).
template <typename T>
hipError_t launch_gram(int n, int k, const T* V, long ldv, const T* tau, T* gram, int nb,
                       T* scratch, int nchunks, hipStream_t st);

// General matrix multiply on FP64 MFMA:
//   C = alpha * op(A) * op(B) + beta * C,  op in {'N','T','C'}, column-major.
// With ksplit > 1 the K range is cut into ksplit slices whose raw products go
// to part[s] (M x N, ld M) and a second kernel sums them in slice order and
// applies alpha/beta (deterministic).  part must hold ksplit*M*N elements.
// ev0/ev1 (optional timing events) are recorded at the start and the end of
// the product: carried by the kernel launch itself where the product is one
// FP64 MFMA launch, as marker packets around its kernels otherwise.
template <typename T>
hipError_t gemm(char ta, char tb, int M, int N, int K, T alpha, const T* A, long lda,
                const T* B, long ldb, T beta, T* C, long ldc, int ksplit, T* part,
                hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// The real trailing-update product on FP64 MFMA with LDS double-buffered
// 128 x 128 tiles (rnpx_dgemm.hip): op(A) 'N', 'T' or 'C', op(B) 'N'; ns > 1
// writes the raw products of K slices of kslice (a multiple of 16) to part
// (M x N each, leading dimension M); cfg 0 picks the tile (1: 128 x 128,
// 2: 128 x 64, 3: 64 x 64).  hipErrorNotSupported for other operand forms
// (nothing launched).  ev0/ev1: start and stop events the launch records.
hipError_t dgemm_mfma(char ta, char tb, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                      long ldb, double beta, double* C, long ldc, int kslice, int ns, double* part, int cfg,
                      hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// The complex counterpart (rnpx_zgemm.hip): op(A) N/T/C, op(B) N.
hipError_t zgemm_mfma(char ta, char tb, int M, int N, int K, zd alpha, const zd* A, long lda, const zd* B, long ldb,
                      zd beta, zd* C, long ldc, int kslice, int ns, zd* part, hipStream_t st);
int dgemm_tile_cfg(int M, int N, int K, int ksplit, int num_cus);

// Small helpers.
template <typename T>
hipError_t launch_set(int m, int n, T offdiag, T diag, T* A, long lda, hipStream_t st);
// B (cols x rows, leading dimension ldb) = conj(A)^T, A rows x cols.
template <typename T>
hipError_t launch_conj_transpose(int rows, int cols, const T* A, long lda, T* B, long ldb, hipStream_t st);
template <typename T>
hipError_t launch_copy(int m, int n, const T* A, long lda, T* B, long ldb, hipStream_t st);
// Rows (fr) and/or columns (fc) of A (rows x cols) reversed, in place or
// into B (rows x cols, leading dimension ldb).
template <typename T>
hipError_t launch_flip(int rows, int cols, T* A, long lda, bool fr, bool fc, hipStream_t st);
// Hessenberg::GenerateQ's set-up of the n x n A from S, a copy of the old
// reflector block A[ilo+1..ihi-1, ilo..ihi-2] (ld lds).
template <typename T>
hipError_t launch_hessq_setup(int n, int ilo, int ihi, T* A, long lda, const T* S, long lds, hipStream_t st);
template <typename T>
hipError_t launch_flip_copy(int rows, int cols, const T* A, long lda, T* B, long ldb, bool fr, bool fc, hipStream_t st);

// Clean V (unit lower trapezoidal, n x k) out of a packed factor.
template <typename T>
hipError_t launch_extract_v(int n, int k, const T* A, long lda, T* V, long ldv, hipStream_t st);
// Zero the strictly upper triangle of columns [0,k) above row j (rows < j)
// of an m x n block: A(l, j) = 0 for l < row_limit, j in [0, n).
template <typename T>
hipError_t launch_zero_rows(int rows, int n, T* A, long lda, hipStream_t st);
hipError_t launch_zero_u32(unsigned* p, int n, hipStream_t st);

// Random fill with RNP's generator (RNP/Random.hpp:45-320): the 48-bit MCG
// stream of one seed, value k = 1 - (seed * a^(k+1) mod 2^48) / 2^48 with the
// stream laid out column by column; dist 0 Uniform01, 1 Uniform_11,
// 2 Normal01 (pairs of uniforms through sqrt(-2 log1p(-a)) cos/sin 2 pi b).
template <typename T>
hipError_t launch_larnv(int dist, const int iseed[4], long long count, T* x, hipStream_t st);

// R X = B, R upper triangular n x n, B's first n rows (n x nrhs) in place.
// Blocked triangular solve with upper R, in place on B (rnpx_trsm.hip):
// trans 'N' R X = B, 'C' R^H X = B.  flags: trsm_flag_words(n, nrhs) words
// that never hold `epoch` at launch (epoch != 0, a new value per launch).
int trsm_flag_words(int n, int nrhs);
// op(M) X = B, M the upper ('U') or lower ('L') triangle of R, op 'N', 'T'
// or 'C', unit diagonal for diag 'U' (same flags protocol as below).
template <typename T>
hipError_t launch_trsm_tri(char uplo, char op, char diag, int n, int nrhs, const T* R, long ldr, T* B, long ldb,
                           unsigned* flags, unsigned epoch, unsigned* abortf, hipStream_t st);
template <typename T>
hipError_t launch_trsm_blocked(char trans, int n, int nrhs, const T* R, long ldr, T* B, long ldb, unsigned* flags,
                               unsigned epoch, unsigned* abortf, hipStream_t st);

template <typename T>
hipError_t launch_trsm_un(int n, int nrhs, const T* R, long ldr, T* B, long ldb, hipStream_t st);
// Inverses of the 64 x 64 diagonal blocks of upper-triangular R into X.
hipError_t launch_trtri_diag(int n, const double* R, long ldr, double* X, long ldx, hipStream_t st);

// Strided batch of FP64 MFMA products C_b = alpha A_b B_b + beta C_b (op N,
// N; X_b = X + b sX), one launch.
hipError_t dgemm_mfma_batched(int M, int N, int K, double alpha, const double* A, long lda, long sa, const double* B,
                              long ldb, long sb, double beta, double* C, long ldc, long sc, int batch,
                              hipStream_t st);

// Mixed precision: FP64 operands, FP32 MFMA product (FP32 accumulation),
// FP64 epilogue; split-K slices reduced in FP64.
// FP32-MFMA product of FP64 operands (rnpx_dgemm.hip), used by gemm32.
hipError_t sgemm_mfma(char ta, char tb, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                      long ldb, double beta, double* C, long ldc, int kslice, int ns, double* part, hipStream_t st);
hipError_t gemm32(char ta, char tb, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                  long ldb, double beta, double* C, long ldc, int ksplit, double* part, hipStream_t st);
// R^H Y = S (R upper), S's first n rows in place.
template <typename T>
hipError_t launch_trsm_uc(int n, int nrhs, const T* R, long ldr, T* B, long ldb, hipStream_t st);
// out[j] = ||X(:,j)||_2^2
hipError_t launch_colnorm2(int m, int n, const double* X, long ldx, double* out, hipStream_t st);
// Y += X
hipError_t launch_axpy_mat(int m, int n, const double* X, long ldx, double* Y, long ldy, hipStream_t st);

// Batched QR / LS: tier 0 = one wave per problem in LDS (param: bytes of
// dynamic LDS per team), tier 1 = 512 threads per problem in place (param:
// the column count from which real problems take the blocked path, 0 = the
// default).
// Workgroups of the large tier for count problems (a few per CU).
template <typename T> int batched_large_teams(int count, int num_cus);
template <typename T>
hipError_t launch_batched(int tier, int count, const int* order, const rnpx_qr_desc* desc, T* A, T* tau, T* B,
                          unsigned* counter, int teams, int param, hipStream_t st);

}  // namespace rnpx
