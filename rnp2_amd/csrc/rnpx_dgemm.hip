// rnpx_dgemm.hip -- the FP64 MFMA product of the trailing update (the larfb
// analogue's two large contractions, Reflector::ApplyBlock at
// RNP/LA/Reflector.hpp:755-768: W = C^H V and C -= V W^H):
//
//   C = alpha op(A) op(B) + beta C,  op(A) M x K, op(B) K x N, column-major,
//
// for real operands whose op leaves them either contiguous along the output
// index (A 'N', B 'T') or along the reduction index (A 'T'/'C', B 'N').
//
// Design (gfx950: 64-wide waves, FP64 MFMA 16x16x4, 160 KB of LDS per CU):
//  * a 256-thread workgroup computes a BM x BN tile as 2 x 2 waves, each wave
//    (BM/2) x (BN/2) in 16 x 16 MFMA blocks (v_mfma_f64_16x16x4_f64), all
//    accumulators in registers for the whole K range;
//  * K advances in slices of BK staged through LDS, double-buffered: the
//    global loads of slice s+1 are in flight (in registers) while the MFMAs
//    consume slice s from one LDS buffer, then land in the other buffer; one
//    barrier per slice;
//  * global loads are 16 bytes per thread, coalesced along whichever index is
//    contiguous in memory; the LDS copy keeps that layout (no transpose while
//    staging): rows of R + 2 doubles (16-byte aligned, 128-bit writes) for
//    operands contiguous along the output index, rows of BK + 1 doubles (odd
//    stride, conflict-free MFMA reads) for operands contiguous along K;
//  * modes: CINIT (beta = 1, alpha = +-1): C is loaded into the accumulators
//    before the K loop and the sign is applied while staging op(A), so the
//    tile is read and written once; SPLIT: raw products of one K slice go to
//    part[z] (M x N, leading dimension M) for a fixed-order reduction;
//    otherwise the alpha / beta epilogue;
//  * tiles are dealt to the XCDs in contiguous chunks (workgroup w runs on XCD
//    w mod 8) and, inside a chunk, in groups of 8 row tiles, so that the
//    operand blocks a chunk shares stay in that XCD's L2.
// The MFMA is issued as op(B)^T op(A)^T, so each lane's four accumulators lie
// along one column of C: lane l holds C(i0 + l%16, j0 + l/16 + 4r), r = 0..3,
// and a store instruction writes 128 contiguous bytes per column.
#include "rnpx_kernels.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdint>

namespace rnpx {

namespace dg {

constexpr int kThreads = 256;

// One operand tile: R rows (the output index: i of op(A), j of op(B)) by BK.
// KC: the operand is contiguous along K in memory (element (r, k) at
// p[k + r ld]); otherwise along r (p[r + k ld]).
template <bool KC, int R, int BK>
struct Op {
  static constexpr int LD = KC ? BK + 1 : R + 2;
  static constexpr int ELEMS = (KC ? R : BK) * LD;
  static constexpr int NV = R * BK / (2 * kThreads);  // 16-byte loads per thread per slice
  static_assert(NV >= 1 && (R * BK) % (2 * kThreads) == 0, "tile / thread mismatch");

  __device__ static inline void coords(int tid, int s, int& r, int& k) {
    if constexpr (KC) {
      r = tid / (BK / 2) + s * (2 * kThreads / BK);
      k = 2 * (tid % (BK / 2));
    } else {
      r = 2 * (tid % (R / 2));
      k = tid / (R / 2) + s * (2 * kThreads / R);
    }
  }
  // Global -> registers for the slice at (r0, k0), without branches or
  // masking, so that every load stays in flight while the MFMAs of the
  // previous slice run: coordinates are clamped into the matrix, so an element
  // outside it is some other entry.  Outside rows (rows of op(A) or columns of
  // op(B)) only feed outputs that are never stored; outside K is zeroed by
  // store().  VEC: 16-byte loads (the operand is 16-byte aligned with an even
  // leading dimension and an even extent along its pairs, so a clamped pair
  // start stays even).
  template <bool VEC>
  __device__ static inline void fetch(double2 (&v)[NV], const double* __restrict__ p, long ld, int r0, int k0,
                                      int rlim, int klim, int tid) {
#pragma unroll
    for (int s = 0; s < NV; ++s) {
      int r, k;
      coords(tid, s, r, k);
      const int cr = min(r0 + r, rlim - 1), ck = min(k0 + k, klim - 1);
      if constexpr (VEC) {
        const int cr2 = KC ? cr : (cr & ~1), ck2 = KC ? (ck & ~1) : ck;
        v[s] = *reinterpret_cast<const double2*>(KC ? p + (size_t)ck2 + (size_t)cr2 * ld
                                                    : p + (size_t)cr2 + (size_t)ck2 * ld);
      } else {
        const int cr1 = KC ? cr : min(r0 + r + 1, rlim - 1), ck1 = KC ? min(k0 + k + 1, klim - 1) : ck;
        v[s].x = KC ? p[(size_t)ck + (size_t)cr * ld] : p[(size_t)cr + (size_t)ck * ld];
        v[s].y = KC ? p[(size_t)ck1 + (size_t)cr1 * ld] : p[(size_t)cr1 + (size_t)ck1 * ld];
      }
    }
  }
  // Registers -> LDS; K indices >= kvalid (only in a last, partial slice)
  // are zeroed, and op(A) is negated for alpha = -1.
  __device__ static inline void store(const double2 (&v)[NV], double* __restrict__ sm, bool neg, int kvalid,
                                      int tid) {
#pragma unroll
    for (int s = 0; s < NV; ++s) {
      int r, k;
      coords(tid, s, r, k);
      double2 x = v[s];
      if (kvalid < BK) {
        const int k1 = KC ? k + 1 : k;
        x.x = (k < kvalid) ? x.x : 0.0;
        x.y = (k1 < kvalid) ? x.y : 0.0;
      }
      if (neg) {
        x.x = -x.x;
        x.y = -x.y;
      }
      if constexpr (KC) {
        sm[r * LD + k] = x.x;
        sm[r * LD + k + 1] = x.y;
      } else {
        *reinterpret_cast<double2*>(sm + k * LD + r) = x;
      }
    }
  }
  __device__ static inline double read(const double* __restrict__ sm, int r, int k) {
    return KC ? sm[r * LD + k] : sm[k * LD + r];
  }
};

// kGeneral: C = alpha P + beta C; kBeta0: C = alpha P; kCInit: C = C + alpha P
// with alpha = +-1 (C in the accumulators); kSplit: the raw P of K slice z at
// C + z zstride (split-K partials, or C itself when alpha = 1, beta = 0).
enum Mode { kGeneral = 0, kCInit = 1, kSplit = 2, kBeta0 = 3 };

struct Args {
  int M, N, K;
  int mode;
  double alpha, beta;
  const double* A;
  long lda;
  const double* B;
  long ldb;
  double* C;  // C, or the split partials
  long ldc;
  size_t zstride;  // kSplit: elements between the slices' outputs
  int kslice;  // K per split slice (multiple of BK)
  int tilesM, tilesN, ntiles;
  int avec, bvec;  // operand base and leading dimension allow 16-byte loads
  long sa, sb, sc;  // strided batch (grid y): elements between the problems' A, B, C
};

// Workgroup -> (tile m, tile n, split z): contiguous chunks per XCD, then
// groups of 8 row tiles inside a chunk.
__device__ inline void tile_of(const Args& a, int w, int nwg, int& tm, int& tn, int& z) {
  int t = w;
  const int full = (nwg / 8) * 8;
  if (w < full) t = (w % 8) * (full / 8) + w / 8;
  const int per = a.tilesM * a.tilesN;
  z = t / per;
  int rem = t - z * per;
  constexpr int G = 8;
  const int group = rem / (G * a.tilesN);
  const int first = group * G;
  const int gsz = min(a.tilesM - first, G);
  const int in = rem - group * G * a.tilesN;
  tm = first + in % gsz;
  tn = in / gsz;
}

// The K loop over [kbeg, kend) into acc, LDS double-buffered (one barrier
// per slice).
template <bool VEC, bool AKC, bool BKC, int BM, int BN, int BK, int FM, int FN>
__device__ inline void k_loop(const Args& a, f64x4 (&acc)[FM][FN], double* __restrict__ As0,
                              double* __restrict__ As1, double* __restrict__ Bs0, double* __restrict__ Bs1, int m0,
                              int n0, int kbeg, int kend, bool neg, int tid, int lane, int wm, int wn) {
  using OA = Op<AKC, BM, BK>;
  using OB = Op<BKC, BN, BK>;
  double2 ra[OA::NV], rb[OB::NV];
  const int nsl = (kend - kbeg + BK - 1) / BK;
  if (nsl > 0) {
    OA::template fetch<VEC>(ra, a.A, a.lda, m0, kbeg, a.M, kend, tid);
    OB::template fetch<VEC>(rb, a.B, a.ldb, n0, kbeg, a.N, kend, tid);
    OA::store(ra, As0, neg, kend - kbeg, tid);
    OB::store(rb, Bs0, false, kend - kbeg, tid);
  }
  __syncthreads();
  for (int s = 0; s < nsl; ++s) {
    const bool odd = s & 1;
    if (s + 1 < nsl) {
      OA::template fetch<VEC>(ra, a.A, a.lda, m0, kbeg + (s + 1) * BK, a.M, kend, tid);
      OB::template fetch<VEC>(rb, a.B, a.ldb, n0, kbeg + (s + 1) * BK, a.N, kend, tid);
    }
    const double* as = odd ? As1 : As0;
    const double* bs = odd ? Bs1 : Bs0;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = OA::read(as, wm * (BM / 2) + 16 * i + (lane & 15), kk + (lane >> 4));
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = OB::read(bs, wn * (BN / 2) + 16 * j + (lane & 15), kk + (lane >> 4));
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = mfma16(bf[j], af[i], acc[i][j]);
    }
    if (s + 1 < nsl) {
      const int kvalid = kend - (kbeg + (s + 1) * BK);
      OA::store(ra, odd ? As0 : As1, neg, kvalid, tid);
      OB::store(rb, odd ? Bs0 : Bs1, false, kvalid, tid);
    }
    __syncthreads();
  }
}

template <int MODE, bool VEC, bool AKC, bool BKC, int BM, int BN, int BK>
__global__ __launch_bounds__(kThreads) void dgemm_kernel(Args a0) {
  using OA = Op<AKC, BM, BK>;
  using OB = Op<BKC, BN, BK>;
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ __align__(16) double As[2][OA::ELEMS];
  __shared__ __align__(16) double Bs[2][OB::ELEMS];
  Args a = a0;  // problem blockIdx.y of a strided batch (0 for one product)
  a.A += (size_t)blockIdx.y * a0.sa;
  a.B += (size_t)blockIdx.y * a0.sb;
  a.C += (size_t)blockIdx.y * a0.sc;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tm, tn, z;
  tile_of(a, blockIdx.x, gridDim.x, tm, tn, z);
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = z * a.kslice;
  const int kend = min(a.K, kbeg + a.kslice);
  const bool inside = (m0 + BM <= a.M) && (n0 + BN <= a.N);
  const bool neg = (MODE == kCInit) && a.alpha < 0.0;

  f64x4 acc[FM][FN];
  const int ib = m0 + wm * (BM / 2) + (lane & 15);  // C row of this lane, block i adds 16 i
  const int jb = n0 + wn * (BN / 2) + (lane >> 4);  // C column, block j adds 16 j, item r adds 4 r
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double v = 0.0;
        if constexpr (MODE == kCInit) {
          // from the clamped, always valid address: an accumulator outside C
          // starts from some other entry of C and is never stored
          const int gi = ib + 16 * i, gj = jb + 16 * j + 4 * r;
          v = a.C[(size_t)min(gi, a.M - 1) + (size_t)min(gj, a.N - 1) * a.ldc];
        }
        acc[i][j][r] = v;
      }
    }

  k_loop<VEC, AKC, BKC, BM, BN, BK, FM, FN>(a, acc, As[0], As[1], Bs[0], Bs[1], m0, n0, kbeg, kend, neg, tid, lane, wm,
                                           wn);

  // Epilogue.
  double* out = a.C;
  long ldo = a.ldc;
  if constexpr (MODE == kSplit) out = a.C + (size_t)z * a.zstride;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = ib + 16 * i, gj = jb + 16 * j + 4 * r;
        if (inside || (gi < a.M && gj < a.N)) {
          double* cp = out + (size_t)gi + (size_t)gj * ldo;
          const double v = acc[i][j][r];
          if constexpr (MODE == kGeneral)
            *cp = a.alpha * v + a.beta * (*cp);
          else if constexpr (MODE == kBeta0)
            *cp = a.alpha * v;
          else
            *cp = v;
        }
      }
}

// ev0/ev1 (optional): start and stop events recorded by the launch itself.
template <bool AKC, bool BKC, int BM, int BN, int BK>
static hipError_t launch(Args a, hipStream_t st, int batch = 1, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  // 16-byte loads: both operands aligned with even leading dimensions, and an
  // even extent along each operand's pairs (rows of op(A) or op(B) for an
  // operand contiguous along the output index, K for one contiguous along K)
  const bool vec = a.avec && a.bvec && (AKC ? a.K % 2 == 0 : a.M % 2 == 0) && (BKC ? a.K % 2 == 0 : a.N % 2 == 0) &&
                   (a.kslice % 2 == 0);
  a.tilesM = (a.M + BM - 1) / BM;
  a.tilesN = (a.N + BN - 1) / BN;
  const int ns = (a.mode == kSplit) ? std::max(1, (a.K + a.kslice - 1) / a.kslice) : 1;  // K = 0: one empty slice
  a.ntiles = a.tilesM * a.tilesN * ns;
  const dim3 grid((unsigned)a.ntiles, (unsigned)std::max(1, batch)), block(kThreads);
#define RNPX_DG_LAUNCH(MODE_)                                                                         \
  do {                                                                                                \
    if (ev0 && vec)                                                                                   \
      hipExtLaunchKernelGGL((dgemm_kernel<MODE_, true, AKC, BKC, BM, BN, BK>), grid, block, 0, st,    \
                            ev0, ev1, 0, a);                                                          \
    else if (ev0)                                                                                     \
      hipExtLaunchKernelGGL((dgemm_kernel<MODE_, false, AKC, BKC, BM, BN, BK>), grid, block, 0, st,   \
                            ev0, ev1, 0, a);                                                          \
    else if (vec)                                                                                     \
      hipLaunchKernelGGL((dgemm_kernel<MODE_, true, AKC, BKC, BM, BN, BK>), grid, block, 0, st, a);   \
    else                                                                                              \
      hipLaunchKernelGGL((dgemm_kernel<MODE_, false, AKC, BKC, BM, BN, BK>), grid, block, 0, st, a);  \
  } while (0)
  if (a.mode == kCInit)
    RNPX_DG_LAUNCH(kCInit);
  else if (a.mode == kSplit)
    RNPX_DG_LAUNCH(kSplit);
  else if (a.mode == kBeta0)
    RNPX_DG_LAUNCH(kBeta0);
  else
    RNPX_DG_LAUNCH(kGeneral);
#undef RNPX_DG_LAUNCH
  return hipGetLastError();
}

static bool aligned16(const double* p, long ld) {
  return ((reinterpret_cast<uintptr_t>(p) & 15u) == 0) && (ld % 2 == 0);
}


// ---------------------------------------------------------------------------
// The FP32 product of the mixed form: FP64 operands in memory, rounded to
// FP32 when a slice is staged into LDS, v_mfma_f32_16x16x4f32 with FP32
// accumulation, the result added to C in FP64 (or written raw for the split-K
// reduction, which sums in FP64).  Same tiling, staging, clamped loads and XCD
// mapping as the FP64 kernel; lane l's four accumulators of a 16 x 16 block
// lie along one column of C: C(i0 + l%16, j0 + 4 (l/16) + r).
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <bool KC, int R, int BK>
struct OpF {
  using O = Op<KC, R, BK>;
  static constexpr int LD = KC ? BK + 1 : R + 1;
  static constexpr int ELEMS = (KC ? R : BK) * LD;
  static constexpr int NV = O::NV;
  __device__ static inline void store(const double2 (&v)[NV], float* __restrict__ sm, bool neg, int kvalid, int tid) {
#pragma unroll
    for (int s = 0; s < NV; ++s) {
      int r, k;
      O::coords(tid, s, r, k);
      double2 x = v[s];
      if (kvalid < BK) {
        const int k1 = KC ? k + 1 : k;
        x.x = (k < kvalid) ? x.x : 0.0;
        x.y = (k1 < kvalid) ? x.y : 0.0;
      }
      const float a = (float)(neg ? -x.x : x.x), b = (float)(neg ? -x.y : x.y);
      if constexpr (KC) {
        sm[r * LD + k] = a;
        sm[r * LD + k + 1] = b;
      } else {
        sm[k * LD + r] = a;
        sm[k * LD + r + 1] = b;
      }
    }
  }
  __device__ static inline float read(const float* __restrict__ sm, int r, int k) {
    return KC ? sm[r * LD + k] : sm[k * LD + r];
  }
};

template <int MODE, bool VEC, bool AKC, bool BKC, int BM, int BN, int BK>
__global__ __launch_bounds__(kThreads) void sgemm_kernel(Args a) {
  using OA = OpF<AKC, BM, BK>;
  using OB = OpF<BKC, BN, BK>;
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ float As[2][OA::ELEMS];
  __shared__ float Bs[2][OB::ELEMS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tm, tn, z;
  tile_of(a, blockIdx.x, gridDim.x, tm, tn, z);
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = z * a.kslice;
  const int kend = min(a.K, kbeg + a.kslice);
  const bool inside = (m0 + BM <= a.M) && (n0 + BN <= a.N);
  f32x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  double2 ra[OA::NV], rb[OB::NV];
  const int nsl = (kend - kbeg + BK - 1) / BK;
  if (nsl > 0) {
    OA::O::template fetch<VEC>(ra, a.A, a.lda, m0, kbeg, a.M, kend, tid);
    OB::O::template fetch<VEC>(rb, a.B, a.ldb, n0, kbeg, a.N, kend, tid);
    OA::store(ra, As[0], false, kend - kbeg, tid);
    OB::store(rb, Bs[0], false, kend - kbeg, tid);
  }
  __syncthreads();
  for (int s = 0; s < nsl; ++s) {
    const int cur = s & 1;
    if (s + 1 < nsl) {
      OA::O::template fetch<VEC>(ra, a.A, a.lda, m0, kbeg + (s + 1) * BK, a.M, kend, tid);
      OB::O::template fetch<VEC>(rb, a.B, a.ldb, n0, kbeg + (s + 1) * BK, a.N, kend, tid);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = OA::read(As[cur], wm * (BM / 2) + 16 * i + (lane & 15), kk + (lane >> 4));
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = OB::read(Bs[cur], wn * (BN / 2) + 16 * j + (lane & 15), kk + (lane >> 4));
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < nsl) {
      const int kvalid = kend - (kbeg + (s + 1) * BK);
      OA::store(ra, As[cur ^ 1], false, kvalid, tid);
      OB::store(rb, Bs[cur ^ 1], false, kvalid, tid);
    }
    __syncthreads();
  }
  double* out = a.C;
  const long ldo = a.ldc;
  if constexpr (MODE == kSplit) out = a.C + (size_t)z * a.zstride;
  const int ib = m0 + wm * (BM / 2) + (lane & 15);
  const int jb = n0 + wn * (BN / 2) + 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = ib + 16 * i, gj = jb + 16 * j + r;
        if (inside || (gi < a.M && gj < a.N)) {
          double* cp = out + (size_t)gi + (size_t)gj * ldo;
          const double v = (double)acc[i][j][r];
          if constexpr (MODE == kCInit)
            *cp = *cp + a.alpha * v;
          else if constexpr (MODE == kGeneral)
            *cp = a.alpha * v + a.beta * (*cp);
          else if constexpr (MODE == kBeta0)
            *cp = a.alpha * v;
          else
            *cp = v;
        }
      }
}

template <bool AKC, bool BKC, int BM, int BN, int BK>
static hipError_t launch32(Args a, hipStream_t st) {
  const bool vec = a.avec && a.bvec && (AKC ? a.K % 2 == 0 : a.M % 2 == 0) && (BKC ? a.K % 2 == 0 : a.N % 2 == 0) &&
                   (a.kslice % 2 == 0);
  a.tilesM = (a.M + BM - 1) / BM;
  a.tilesN = (a.N + BN - 1) / BN;
  const int ns = (a.mode == kSplit) ? std::max(1, (a.K + a.kslice - 1) / a.kslice) : 1;
  a.ntiles = a.tilesM * a.tilesN * ns;
  const dim3 grid((unsigned)a.ntiles), block(kThreads);
#define RNPX_SG_LAUNCH(MODE_)                                                                         \
  do {                                                                                                \
    if (vec)                                                                                          \
      hipLaunchKernelGGL((sgemm_kernel<MODE_, true, AKC, BKC, BM, BN, BK>), grid, block, 0, st, a);   \
    else                                                                                              \
      hipLaunchKernelGGL((sgemm_kernel<MODE_, false, AKC, BKC, BM, BN, BK>), grid, block, 0, st, a);  \
  } while (0)
  if (a.mode == kCInit)
    RNPX_SG_LAUNCH(kCInit);
  else if (a.mode == kSplit)
    RNPX_SG_LAUNCH(kSplit);
  else if (a.mode == kBeta0)
    RNPX_SG_LAUNCH(kBeta0);
  else
    RNPX_SG_LAUNCH(kGeneral);
#undef RNPX_SG_LAUNCH
  return hipGetLastError();
}

}  // namespace dg

// The tile configurations built (see the header comment): 128 x 128 for
// products reduced over a long K (W = V^H C, split-K), 128 x 64 for the
// K <= 256 updates (measured 2-7% faster than 128 x 128 there), and smaller
// tiles where the grid would leave CUs idle.
int dgemm_tile_cfg(int M, int N, int K, int ksplit, int num_cus) {
  const long t128 = (long)((M + 127) / 128) * ((N + 127) / 128) * ksplit;
  const long t64 = (long)((M + 127) / 128) * ((N + 63) / 64) * ksplit;
  if (K > 256 && 2 * t128 >= 3L * num_cus) return 1;
  if (t64 >= 2L * num_cus || (K > 256 && t64 >= num_cus)) return 2;
  if (t128 >= 2L * num_cus) return 1;
  return 3;
}

// The mixed form's FP32-MFMA product (FP64 operands and result, FP32
// accumulation): op(A) N/T/C, op(B) N; ns > 1: raw slices of kslice (a
// multiple of 32) to part.  Other forms return hipErrorNotSupported.
hipError_t sgemm_mfma(char ta, char tb, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                      long ldb, double beta, double* C, long ldc, int kslice, int ns, double* part, hipStream_t st) {
  const bool akc = (ta == 'T' || ta == 'C');
  if (tb != 'N' || !(akc || ta == 'N')) return hipErrorNotSupported;
  constexpr int BK = 32;
  if (ns > 1 && (kslice % BK != 0 || !part)) return hipErrorNotSupported;
  static const int cus = [] {
    int dev = 0, n = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    return n > 0 ? n : 256;
  }();
  const int cfg = dgemm_tile_cfg(M, N, ns > 1 ? kslice : K, ns, cus);
  dg::Args a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.alpha = alpha;
  a.beta = beta;
  a.A = A;
  a.lda = lda;
  a.B = B;
  a.ldb = ldb;
  a.avec = dg::aligned16(A, lda);
  a.bvec = dg::aligned16(B, ldb);
  if (ns > 1) {
    a.mode = dg::kSplit;
    a.kslice = kslice;
    a.C = part;
    a.ldc = M;
    a.zstride = (size_t)M * N;
  } else {
    if (beta == 1.0)
      a.mode = dg::kCInit;  // C += alpha P in FP64
    else if (beta == 0.0)
      a.mode = (alpha == 1.0) ? dg::kSplit : dg::kBeta0;
    else
      a.mode = dg::kGeneral;
    a.kslice = std::max(BK, ((K + BK - 1) / BK) * BK);
    a.C = C;
    a.ldc = ldc;
    a.zstride = 0;
  }
  if (akc) {
    if (cfg == 1) return dg::launch32<true, true, 128, 128, BK>(a, st);
    if (cfg == 2) return dg::launch32<true, true, 128, 64, BK>(a, st);
    return dg::launch32<true, true, 64, 64, BK>(a, st);
  }
  if (cfg == 1) return dg::launch32<false, true, 128, 128, BK>(a, st);
  if (cfg == 2) return dg::launch32<false, true, 128, 64, BK>(a, st);
  return dg::launch32<false, true, 64, 64, BK>(a, st);
}

// op(A) in {'N', 'T', 'C'} (real: 'C' = 'T'), op(B) 'N'; ns > 1: split-K
// into slices of kslice (a multiple of 16) written to part; returns
// hipErrorNotSupported for combinations not built here (the caller then uses
// the general kernel).
hipError_t dgemm_mfma(char ta, char tb, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                      long ldb, double beta, double* C, long ldc, int kslice, int ns, double* part, int cfg,
                      hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  const bool akc = (ta == 'T' || ta == 'C');
  if (tb != 'N' || !(akc || ta == 'N')) return hipErrorNotSupported;
  constexpr int BK = 16;
  if (ns > 1 && (kslice % BK != 0 || !part)) return hipErrorNotSupported;
  if (cfg <= 0) {
    static const int cus = [] {
      int dev = 0, n = 256;
      if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
      return n > 0 ? n : 256;
    }();
    cfg = dgemm_tile_cfg(M, N, ns > 1 ? kslice : K, ns, cus);
  }
  dg::Args a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.alpha = alpha;
  a.beta = beta;
  a.A = A;
  a.lda = lda;
  a.B = B;
  a.ldb = ldb;
  a.avec = dg::aligned16(A, lda);
  a.bvec = dg::aligned16(B, ldb);
  if (ns > 1) {
    a.mode = dg::kSplit;
    a.kslice = kslice;
    a.C = part;
    a.ldc = M;
    a.zstride = (size_t)M * N;
  } else {
    if (beta == 1.0 && (alpha == 1.0 || alpha == -1.0))
      a.mode = dg::kCInit;
    else if (beta == 0.0)
      a.mode = (alpha == 1.0) ? dg::kSplit : dg::kBeta0;  // alpha = 1: the raw product, one slice
    else
      a.mode = dg::kGeneral;
    a.kslice = std::max(BK, ((K + BK - 1) / BK) * BK);
    a.C = C;
    a.ldc = ldc;
    a.zstride = 0;
  }
  if (akc) {
    switch (cfg) {
      case 1: return dg::launch<true, true, 128, 128, BK>(a, st, 1, ev0, ev1);
      case 2: return dg::launch<true, true, 128, 64, BK>(a, st, 1, ev0, ev1);
      default: return dg::launch<true, true, 64, 64, BK>(a, st, 1, ev0, ev1);
    }
  }
  switch (cfg) {
    case 1: return dg::launch<false, true, 128, 128, BK>(a, st, 1, ev0, ev1);
    case 2: return dg::launch<false, true, 128, 64, BK>(a, st, 1, ev0, ev1);
    default: return dg::launch<false, true, 64, 64, BK>(a, st, 1, ev0, ev1);
  }
}

// batch products C_b = alpha A_b B_b + beta C_b (op N, N), X_b = X + b sX:
// one launch, problem b in grid row b.
hipError_t dgemm_mfma_batched(int M, int N, int K, double alpha, const double* A, long lda, long sa, const double* B,
                              long ldb, long sb, double beta, double* C, long ldc, long sc, int batch,
                              hipStream_t st) {
  if (M < 0 || N < 0 || K < 0 || batch < 0 || lda < std::max(1, M) || ldb < std::max(1, K) || ldc < std::max(1, M))
    return hipErrorInvalidValue;
  if (M == 0 || N == 0 || batch == 0) return hipSuccess;
  if (batch > 65535) return hipErrorInvalidValue;
  constexpr int BK = 16;
  static const int cus = [] {
    int dev = 0, n = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    return n > 0 ? n : 256;
  }();
  // the tile choice for the whole launch: the batch multiplies the tile count
  const int cfg = dgemm_tile_cfg(M, N, K, batch, cus);
  dg::Args a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.alpha = alpha;
  a.beta = beta;
  a.A = A;
  a.lda = lda;
  a.B = B;
  a.ldb = ldb;
  a.sa = sa;
  a.sb = sb;
  a.sc = sc;
  a.avec = dg::aligned16(A, lda) && sa % 2 == 0;
  a.bvec = dg::aligned16(B, ldb) && sb % 2 == 0;
  if (beta == 1.0 && (alpha == 1.0 || alpha == -1.0))
    a.mode = dg::kCInit;
  else if (beta == 0.0)
    a.mode = (alpha == 1.0) ? dg::kSplit : dg::kBeta0;
  else
    a.mode = dg::kGeneral;
  a.kslice = std::max(BK, ((K + BK - 1) / BK) * BK);
  a.C = C;
  a.ldc = ldc;
  a.zstride = 0;
  switch (cfg) {
    case 1: return dg::launch<false, true, 128, 128, BK>(a, st, batch);
    case 2: return dg::launch<false, true, 128, 64, BK>(a, st, batch);
    default: return dg::launch<false, true, 64, 64, BK>(a, st, batch);
  }
}

}  // namespace rnpx
