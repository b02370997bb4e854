// rnpx_gemm.hip -- FP64 MFMA matrix multiply for the trailing update
// (the larfb analogue's contractions, Reflector::ApplyBlock at
// RNP/LA/Reflector.hpp:739-815: W = C^H V, W T, C -= V W^H and the side-R
// forms), and for everything else here that is a plain product.
//
//   C = alpha * op(A) * op(B) + beta * C,   op in {N, T, C}, column-major.
//
// Tiles: a 256-thread workgroup computes a BM x BN tile of C as 2 x 2 waves
// (BM, BN in {64, 128}),
// each wave (BM/2) x (BN/2) in 16 x 16 MFMA blocks (v_mfma_f64_16x16x4_f64).
// op(A) and op(B) K-slices of depth BK are staged through LDS with the
// global loads coalesced along whichever index is contiguous in memory, and
// the next slice is fetched into registers while the current one feeds the
// MFMAs.  The MFMA is issued as (op(B)^T)(op(A)^T) so that each lane's four
// accumulators lie along a column of C: stores of C are 16 consecutive rows
// per column, i.e. full 128-byte lines.
// complex<double>: four real MFMAs per block (re*re - im*im, re*im + im*re),
// conjugation applied when the slice is staged.
// Split-K: blockIdx.z picks a K slice; raw sums go to part[z] and
// gemm_reduce sums the slices in order (results do not depend on timing).
#include "rnpx_kernels.hpp"

#include <cstdlib>

namespace rnpx {

// Tile shapes: (BM, BN) per workgroup, BK the K depth staged per LDS slice.
// Real: 128 x 128 (or 64 x 128 when M <= 64), BK 32 -- one slice is 8K
// cycles of MFMA per wave, longer than a global-load round trip, so the
// register prefetch of the next slice hides its latency.  Complex: 64 x 64.
template <typename T> struct Cfg;
template <> struct Cfg<double> { static constexpr int BK = 32; };
template <> struct Cfg<zd> { static constexpr int BK = 16; };

template <typename T, char OP>
__device__ inline T ld_op(const T* p) {
  T v = *p;
  if constexpr (OP == 'C') return S<T>::conj(v);
  return v;
}

template <typename T, int BM, int BN, int BK, char TA, char TB>
struct Tile {
  static constexpr int PADA = 2, PADB = 2;
  static constexpr int NA = BM * BK / 256;  // A elements per thread per slice
  static constexpr int NBv = BN * BK / 256;
  // Global -> register fetch of one K slice of op(A) (rows m0.., k0..).
  __device__ static inline void fetchA(T (&r)[NA], const T* A, long lda, int M, int K, int m0, int k0, int tid) {
    if constexpr (TA == 'N') {
      constexpr int KS = 256 / BM;  // k rows covered per pass
      const int i = tid % BM, kb = tid / BM;
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        const int k = kb + KS * s;
        const int gi = m0 + i, gk = k0 + k;
        r[s] = (gi < M && gk < K) ? A[(size_t)gi + (size_t)gk * lda] : S<T>::zero();
      }
    } else {
      constexpr int IS = 256 / BK;
      const int k = tid % BK, ib = tid / BK;
#pragma unroll
      for (int s = 0; s < NA; ++s) {
        const int i = ib + IS * s;
        const int gi = m0 + i, gk = k0 + k;
        r[s] = (gi < M && gk < K) ? ld_op<T, TA>(A + (size_t)gk + (size_t)gi * lda) : S<T>::zero();
      }
    }
  }
  __device__ static inline void storeA(const T (&r)[NA], T (*As)[BM + PADA], int tid) {
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
  __device__ static inline void fetchB(T (&r)[NBv], const T* B, long ldb, int N, int K, int n0, int k0, int tid) {
    if constexpr (TB == 'N') {
      constexpr int JS = 256 / BK;
      const int k = tid % BK, jb = tid / BK;
#pragma unroll
      for (int s = 0; s < NBv; ++s) {
        const int j = jb + JS * s;
        const int gj = n0 + j, gk = k0 + k;
        r[s] = (gj < N && gk < K) ? B[(size_t)gk + (size_t)gj * ldb] : S<T>::zero();
      }
    } else {
      constexpr int KS = 256 / BN;
      const int j = tid % BN, kb = tid / BN;
#pragma unroll
      for (int s = 0; s < NBv; ++s) {
        const int k = kb + KS * s;
        const int gj = n0 + j, gk = k0 + k;
        r[s] = (gj < N && gk < K) ? ld_op<T, TB>(B + (size_t)gj + (size_t)gk * ldb) : S<T>::zero();
      }
    }
  }
  __device__ static inline void storeB(const T (&r)[NBv], T (*Bs)[BN + PADB], int tid) {
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

template <typename T> struct Acc;
template <> struct Acc<double> {
  f64x4 v;
  __device__ inline void zero() { v = f64x4{0.0, 0.0, 0.0, 0.0}; }
  __device__ inline void mma(double b, double a) { v = mfma16(b, a, v); }
  __device__ inline double get(int r) const { return v[r]; }
  __device__ inline void set(int r, double x) { v[r] = x; }
};
template <> struct Acc<zd> {
  f64x4 re, im;
  __device__ inline void zero() { re = f64x4{0.0, 0.0, 0.0, 0.0}; im = re; }
  __device__ inline void mma(zd b, zd a) {
    re = mfma16(b.x, a.x, re);
    re = mfma16(-b.y, a.y, re);
    im = mfma16(b.x, a.y, im);
    im = mfma16(b.y, a.x, im);
  }
  __device__ inline zd get(int r) const { return mkz(re[r], im[r]); }
  __device__ inline void set(int r, zd x) {
    re[r] = x.x;
    im[r] = x.y;
  }
};

// CINIT: C is loaded into the accumulators before the K loop (beta = 1,
// alpha = +-1, the sign folded into the op(A) fragments), so the C read
// overlaps the first slice instead of stalling the epilogue.
template <typename T, char TA, char TB, int BM, int BN, bool CINIT>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, T alpha, const T* __restrict__ A, long lda,
                                                   const T* __restrict__ B, long ldb, T beta, T* __restrict__ C,
                                                   long ldc, int kslice, T* __restrict__ part) {
  constexpr int BK = Cfg<T>::BK;
  using Tl = Tile<T, BM, BN, BK, TA, TB>;
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ T As[BK][BM + Tl::PADA];
  __shared__ T Bs[BK][BN + Tl::PADB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kslice;
  const int kend = min(K, kbeg + kslice);

  T ra[Tl::NA], rb[Tl::NBv];
  if (kbeg < kend) {
    Tl::fetchA(ra, A, lda, M, kend, m0, kbeg, tid);
    Tl::fetchB(rb, B, ldb, N, kend, n0, kbeg, tid);
  }
  Acc<T> acc[FM][FN];
  // accumulator item r of lane l is C(row0 + l%16, col0 + l/16 + 4r)
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc[i][j].zero();
      if constexpr (CINIT) {
        const int gi = m0 + wm * (BM / 2) + i * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = n0 + wn * (BN / 2) + j * 16 + (lane >> 4) + 4 * r;
          if (gi < M && gj < N) acc[i][j].set(r, C[(size_t)gi + (size_t)gj * ldc]);
        }
      }
    }
  const bool neg = CINIT && (S<T>::re(alpha) < 0.0);
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
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        af[i] = As[kk + (lane >> 4)][wm * (BM / 2) + i * 16 + (lane & 15)];
        if (neg) af[i] = S<T>::zero() - af[i];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + (lane >> 4)][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j].mma(bf[j], af[i]);
    }
  }

  // Epilogue: lane holds C(i, j + 4r), r = 0..3, with i = row base + lane%16
  // and j = column base + lane/16.
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int gi = m0 + wm * (BM / 2) + i * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wn * (BN / 2) + j * 16 + (lane >> 4) + 4 * r;
        if (gi < M && gj < N) {
          const T v = acc[i][j].get(r);
          if (CINIT) {
            C[(size_t)gi + (size_t)gj * ldc] = v;
          } else if (part) {
            part[(size_t)blockIdx.z * M * N + (size_t)gi + (size_t)gj * M] = v;
          } else {
            T* cp = C + (size_t)gi + (size_t)gj * ldc;
            *cp = S<T>::is_zero(beta) ? alpha * v : alpha * v + beta * (*cp);
          }
        }
      }
    }
  }
}

template <typename T>
__global__ void gemm_reduce_kernel(int M, int N, int ns, const T* __restrict__ part, T alpha, T beta,
                                   T* __restrict__ C, long ldc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)M * N) return;
  const int i = (int)(e % M), j = (int)(e / M);
  const T s = sum_strided(part + e, ns, (size_t)M * N);
  T* cp = C + (size_t)i + (size_t)j * ldc;
  *cp = S<T>::is_zero(beta) ? alpha * s : alpha * s + beta * (*cp);
}

template <typename T, char TA, char TB, int BM, int BN, bool CINIT>
static void launch_gemm_tile(int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta, T* C,
                             long ldc, int kslice, int ns, T* part, hipStream_t st) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, ns);
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB, BM, BN, CINIT>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, ldb,
                     beta, C, ldc, kslice, part);
}

// Real tile choice: RNPX_GEMM_TILE (read once) overrides the default
// (0 auto, 1 128x128, 2 128x64, 3 64x128, 4 64x64) -- a tuning knob.
static int gemm_tile_override() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("RNPX_GEMM_TILE");
    v = e ? atoi(e) : 0;
  }
  return v;
}

template <typename T, char TA, char TB, bool CI>
static void launch_real_tiles(int cfg, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb,
                              T beta, T* C, long ldc, int kslice, int ns, T* part, hipStream_t st) {
  switch (cfg) {
    case 1: launch_gemm_tile<T, TA, TB, 128, 128, CI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st); break;
    case 2: launch_gemm_tile<T, TA, TB, 128, 64, CI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st); break;
    case 3: launch_gemm_tile<T, TA, TB, 64, 128, CI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st); break;
    default: launch_gemm_tile<T, TA, TB, 64, 64, CI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st); break;
  }
}

template <typename T, char TA, char TB>
static void launch_gemm_t(int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta, T* C,
                          long ldc, int kslice, int ns, T* part, hipStream_t st) {
  const bool cinit = (ns == 1) && !part && S<T>::re(beta) == 1.0 && S<T>::im(beta) == 0.0 &&
                     (S<T>::re(alpha) == 1.0 || S<T>::re(alpha) == -1.0) && S<T>::im(alpha) == 0.0;
  if constexpr (S<T>::cplx) {
    if (cinit) launch_gemm_tile<T, TA, TB, 64, 64, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st);
    else launch_gemm_tile<T, TA, TB, 64, 64, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st);
  } else {
    // Measured on MI355X: 64 x 64 tiles (2-3 workgroups per CU) are fastest
    // for the trailing-update shapes (K = 64..128 or M = 64..128); 64 x 128
    // for large products in all three dimensions.
    int cfg = gemm_tile_override();
    if (cfg == 0) cfg = (M >= 2048 && N >= 2048 && K >= 1024) ? 3 : 4;
    if (cinit) launch_real_tiles<T, TA, TB, true>(cfg, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st);
    else launch_real_tiles<T, TA, TB, false>(cfg, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, part, st);
  }
}

// RNPX_DGEMM (read once): -1 keeps every product on gemm_kernel (a tuning
// and comparison knob); 0 (default) lets dgemm_mfma pick its tile; 1..3 force
// one of its tiles.
static int dgemm_cfg_override() {
  static const int v = [] {
    const char* e = getenv("RNPX_DGEMM");
    return e ? atoi(e) : 0;
  }();
  return v;
}

template <typename T>
hipError_t gemm(char ta, char tb, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta,
                T* C, long ldc, int ksplit, T* part, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  if (M < 0 || N < 0 || K < 0) return hipErrorInvalidValue;
  if (M == 0 || N == 0) return hipSuccess;
  if (ldc < M) return hipErrorInvalidValue;
  const bool ca = (ta == 'N' || ta == 'T' || ta == 'C'), cb = (tb == 'N' || tb == 'T' || tb == 'C');
  if (!ca || !cb) return hipErrorInvalidValue;
  if (ta == 'N' ? lda < M : lda < K) return hipErrorInvalidValue;
  if (tb == 'N' ? ldb < K : ldb < N) return hipErrorInvalidValue;
  if (ksplit < 1) ksplit = 1;
  if (ksplit > 1 && !part) return hipErrorInvalidValue;
  int kslice = K;
  int ns = 1;
  if (ksplit > 1 && K > 0) {
    kslice = (K + ksplit - 1) / ksplit;
    kslice = ((kslice + Cfg<T>::BK - 1) / Cfg<T>::BK) * Cfg<T>::BK;
    ns = (K + kslice - 1) / kslice;
  }
  if (K == 0) kslice = 0;
  T* pp = (ns > 1) ? part : nullptr;
  bool done = false;
  // The timing events ride on the launch when the product is one FP64 MFMA
  // launch; otherwise they are recorded around the product's kernels.
  bool ride = false;
  if constexpr (!S<T>::cplx) ride = ev0 && ev1 && ns == 1 && tb == 'N' && dgemm_cfg_override() >= 0;
  if (ev0 && ev1 && !ride) {
    const hipError_t e = hipEventRecord(ev0, st);
    if (e != hipSuccess) return e;
  }
  if constexpr (!S<T>::cplx) {
    // Real op(A) N/T/C times op(B) N: the double-buffered 128 x 128 kernel.
    const int cfg = dgemm_cfg_override();
    if (cfg >= 0 && tb == 'N') {
      const hipError_t e = dgemm_mfma(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, cfg, st,
                                      ride ? ev0 : nullptr, ride ? ev1 : nullptr);
      if (e != hipErrorNotSupported) {
        if (e != hipSuccess) return e;
        done = true;
      }
    }
    if (!done && ride) {
      // not launched there: the events go around the general kernel
      const hipError_t e = hipEventRecord(ev0, st);
      if (e != hipSuccess) return e;
      ride = false;
    }
  } else {
    // Complex op(A) N/T/C times op(B) N: the double-buffered complex kernel
    // (RNPX_ZGEMM=-1 keeps the general one).
    static const bool zoff = [] {
      const char* e = getenv("RNPX_ZGEMM");
      return e && atoi(e) < 0;
    }();
    if (!zoff && tb == 'N') {
      const hipError_t e = zgemm_mfma(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);
      if (e != hipErrorNotSupported) {
        if (e != hipSuccess) return e;
        done = true;
      }
    }
  }
  if (!done) {
#define RNPX_GEMM_CASE(A_, B_)                                                                          \
  if (ta == A_ && tb == B_) {                                                                            \
    launch_gemm_t<T, A_, B_>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kslice, ns, pp, st);          \
  }
  RNPX_GEMM_CASE('N', 'N') else RNPX_GEMM_CASE('N', 'T') else RNPX_GEMM_CASE('N', 'C')
  else RNPX_GEMM_CASE('T', 'N') else RNPX_GEMM_CASE('T', 'T') else RNPX_GEMM_CASE('T', 'C')
  else RNPX_GEMM_CASE('C', 'N') else RNPX_GEMM_CASE('C', 'T') else RNPX_GEMM_CASE('C', 'C')
#undef RNPX_GEMM_CASE
  }
  if (ns > 1) {
    const long tot = (long)M * N;
    hipLaunchKernelGGL((gemm_reduce_kernel<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M, N, ns,
                       part, alpha, beta, C, ldc);
  }
  if (ev0 && ev1 && !ride) {
    const hipError_t e = hipEventRecord(ev1, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

template hipError_t gemm<double>(char, char, int, int, int, double, const double*, long, const double*, long, double,
                                 double*, long, int, double*, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t gemm<zd>(char, char, int, int, int, zd, const zd*, long, const zd*, long, zd, zd*, long, int, zd*,
                             hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace rnpx
