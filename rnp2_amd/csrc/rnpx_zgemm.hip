// rnpx_zgemm.hip -- the complex<double> FP64 MFMA product of the trailing
// update (Reflector::ApplyBlock at RNP/LA/Reflector.hpp:755-768 for complex
// matrices: W = V^H C and C -= V W^H), the complex counterpart of
// rnpx_dgemm.hip:
//
//   C = alpha op(A) op(B) + beta C,  op(A) in {N, T, C}, op(B) = N,
//
// interleaved (re, im) column-major operands.
//
// Design (gfx950, FP64 MFMA 16x16x4): a 256-thread workgroup computes a
// BM x BN complex tile as 2 x 2 waves; K advances in slices of BK staged
// through LDS, double-buffered, with the next slice's global loads in flight
// in registers while the MFMAs consume the current one (one barrier per
// slice).  Each operand slice is split into a real and an imaginary plane
// when staged (conjugation and the sign of alpha = -1 applied there), and
// every 16 x 16 x 4 complex block product is four real MFMAs:
//   Cr += Ar Br - Ai Bi,   Ci += Ar Bi + Ai Br,
// so the complex "flops" (8 per complex multiply-add) run at the real MFMA
// rate.  Loads are branch-free (coordinates clamped into the matrix; only
// the K tail is zeroed when staged); tiles are dealt to the XCDs in
// contiguous chunks.  Modes as in the real kernel: CINIT (alpha = +-1,
// beta = 1: C loaded into the accumulators, read and written once), SPLIT
// (raw slice products for the fixed-order reduction), BETA0, GENERAL.
#include "rnpx_kernels.hpp"

#include <algorithm>
#include <cstdint>

namespace rnpx {

namespace zg {

constexpr int kThreads = 256;

// One operand tile: R rows (the output index) by BK (the reduction index).
// KC: contiguous along K in memory (element (r, k) at p[k + r ld]); else
// along r (p[r + k ld]).  LDS planes: KC -> [r][k] rows of BK + 1, else
// [k][r] rows of R + 1 (odd strides: conflict-free MFMA reads).
template <bool KC, int R, int BK>
struct ZOp {
  static constexpr int LD = KC ? BK + 1 : R + 1;
  static constexpr int ELEMS = (KC ? R : BK) * LD;  // per plane
  static constexpr int NV = R * BK / kThreads;     // complex loads per thread per slice
  static_assert(NV >= 1 && (R * BK) % kThreads == 0, "tile / thread mismatch");

  __device__ static inline void coords(int tid, int s, int& r, int& k) {
    if constexpr (KC) {
      k = tid % BK;
      r = tid / BK + s * (kThreads / BK);
    } else {
      r = tid % R;
      k = tid / R + s * (kThreads / R);
    }
  }
  __device__ static inline void fetch(zd (&v)[NV], const zd* __restrict__ p, long ld, int r0, int k0, int rlim,
                                      int klim, int tid) {
#pragma unroll
    for (int s = 0; s < NV; ++s) {
      int r, k;
      coords(tid, s, r, k);
      const int cr = min(r0 + r, rlim - 1), ck = min(k0 + k, klim - 1);
      v[s] = KC ? p[(size_t)ck + (size_t)cr * ld] : p[(size_t)cr + (size_t)ck * ld];
    }
  }
  // conj: negate the imaginary part; neg: negate both (alpha = -1).
  __device__ static inline void store(const zd (&v)[NV], double* __restrict__ re, double* __restrict__ im, bool conj,
                                      bool neg, int kvalid, int tid) {
#pragma unroll
    for (int s = 0; s < NV; ++s) {
      int r, k;
      coords(tid, s, r, k);
      double x = v[s].x, y = v[s].y;
      if (conj) y = -y;
      if (neg) {
        x = -x;
        y = -y;
      }
      if (k >= kvalid) {
        x = 0.0;
        y = 0.0;
      }
      const int o = KC ? r * LD + k : k * LD + r;
      re[o] = x;
      im[o] = y;
    }
  }
  __device__ static inline int at(int r, int k) { return KC ? r * LD + k : k * LD + r; }
};

enum Mode { kGeneral = 0, kCInit = 1, kSplit = 2, kBeta0 = 3 };

struct Args {
  int M, N, K;
  int mode;
  zd alpha, beta;
  const zd* A;
  long lda;
  const zd* B;
  long ldb;
  zd* C;
  long ldc;
  size_t zstride;
  int kslice;
  int tilesM, tilesN;
  int conja;  // op(A) = 'C'
};

// Workgroup -> (tile m, tile n, split z): contiguous chunks per XCD, then
// groups of 8 row tiles inside a chunk (as the real kernel).
__device__ inline void tile_of(const Args& a, int w, int nwg, int& tm, int& tn, int& z) {
  int t = w;
  const int full = (nwg / 8) * 8;
  if (w < full) t = (w % 8) * (full / 8) + w / 8;
  const int per = a.tilesM * a.tilesN;
  z = t / per;
  const int rem = t - z * per;
  constexpr int G = 8;
  const int group = rem / (G * a.tilesN);
  const int first = group * G;
  const int gsz = min(a.tilesM - first, G);
  const int in = rem - group * G * a.tilesN;
  tm = first + in % gsz;
  tn = in / gsz;
}

template <int MODE, bool AKC, int BM, int BN, int BK>
__global__ __launch_bounds__(kThreads) void zgemm_kernel(Args a) {
  using OA = ZOp<AKC, BM, BK>;
  using OB = ZOp<true, BN, BK>;  // op(B) = N: contiguous along K
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ double Ar[2][OA::ELEMS], Ai[2][OA::ELEMS];
  __shared__ double Br[2][OB::ELEMS], Bi[2][OB::ELEMS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tm, tn, z;
  tile_of(a, blockIdx.x, gridDim.x, tm, tn, z);
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = z * a.kslice;
  const int kend = min(a.K, kbeg + a.kslice);
  const bool inside = (m0 + BM <= a.M) && (n0 + BN <= a.N);
  const bool neg = (MODE == kCInit) && a.alpha.x < 0.0;
  const bool conja = a.conja != 0;

  f64x4 cr[FM][FN], ci[FM][FN];
  const int ib = m0 + wm * (BM / 2) + (lane & 15);
  const int jb = n0 + wn * (BN / 2) + (lane >> 4);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double x = 0.0, y = 0.0;
        if constexpr (MODE == kCInit) {
          const int gi = ib + 16 * i, gj = jb + 16 * j + 4 * r;
          const zd c = a.C[(size_t)min(gi, a.M - 1) + (size_t)min(gj, a.N - 1) * a.ldc];
          x = c.x;
          y = c.y;
        }
        cr[i][j][r] = x;
        ci[i][j][r] = y;
      }

  zd ra[OA::NV], rb[OB::NV];
  const int nsl = (kend - kbeg + BK - 1) / BK;
  if (nsl > 0) {
    OA::fetch(ra, a.A, a.lda, m0, kbeg, a.M, kend, tid);
    OB::fetch(rb, a.B, a.ldb, n0, kbeg, a.N, kend, tid);
    OA::store(ra, Ar[0], Ai[0], conja, neg, kend - kbeg, tid);
    OB::store(rb, Br[0], Bi[0], false, false, kend - kbeg, tid);
  }
  __syncthreads();
  for (int s = 0; s < nsl; ++s) {
    const int cur = s & 1;
    if (s + 1 < nsl) {
      OA::fetch(ra, a.A, a.lda, m0, kbeg + (s + 1) * BK, a.M, kend, tid);
      OB::fetch(rb, a.B, a.ldb, n0, kbeg + (s + 1) * BK, a.N, kend, tid);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double ar[FM], ai[FM], br[FN], bi[FN];
      const int kq = kk + (lane >> 4);
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int o = OA::at(wm * (BM / 2) + 16 * i + (lane & 15), kq);
        ar[i] = Ar[cur][o];
        ai[i] = Ai[cur][o];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int o = OB::at(wn * (BN / 2) + 16 * j + (lane & 15), kq);
        br[j] = Br[cur][o];
        bi[j] = Bi[cur][o];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const double nai = -ai[i];
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          cr[i][j] = mfma16(br[j], ar[i], cr[i][j]);
          cr[i][j] = mfma16(bi[j], nai, cr[i][j]);
          ci[i][j] = mfma16(bi[j], ar[i], ci[i][j]);
          ci[i][j] = mfma16(br[j], ai[i], ci[i][j]);
        }
      }
    }
    if (s + 1 < nsl) {
      const int kvalid = kend - (kbeg + (s + 1) * BK);
      OA::store(ra, Ar[cur ^ 1], Ai[cur ^ 1], conja, neg, kvalid, tid);
      OB::store(rb, Br[cur ^ 1], Bi[cur ^ 1], false, false, kvalid, tid);
    }
    __syncthreads();
  }

  zd* out = a.C;
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
          zd* cp = out + (size_t)gi + (size_t)gj * ldo;
          const zd v = mkz(cr[i][j][r], ci[i][j][r]);
          if constexpr (MODE == kGeneral)
            *cp = a.alpha * v + a.beta * (*cp);
          else if constexpr (MODE == kBeta0)
            *cp = a.alpha * v;
          else
            *cp = v;
        }
      }
}

template <bool AKC, int BM, int BN, int BK>
static hipError_t launch(Args a, int ns, hipStream_t st) {
  a.tilesM = (a.M + BM - 1) / BM;
  a.tilesN = (a.N + BN - 1) / BN;
  const dim3 grid((unsigned)(a.tilesM * a.tilesN * ns)), block(kThreads);
  switch (a.mode) {
    case kCInit: hipLaunchKernelGGL((zgemm_kernel<kCInit, AKC, BM, BN, BK>), grid, block, 0, st, a); break;
    case kSplit: hipLaunchKernelGGL((zgemm_kernel<kSplit, AKC, BM, BN, BK>), grid, block, 0, st, a); break;
    case kBeta0: hipLaunchKernelGGL((zgemm_kernel<kBeta0, AKC, BM, BN, BK>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((zgemm_kernel<kGeneral, AKC, BM, BN, BK>), grid, block, 0, st, a); break;
  }
  return hipGetLastError();
}

}  // namespace zg

// op(A) in {'N', 'T', 'C'}, op(B) 'N'; ns > 1: split-K into slices of kslice
// (a multiple of 16) written to part (M x N per slice).  Other forms return
// hipErrorNotSupported (the caller then uses the general kernel).
hipError_t zgemm_mfma(char ta, char tb, int M, int N, int K, zd alpha, const zd* A, long lda, const zd* B, long ldb,
                      zd beta, zd* C, long ldc, int kslice, int ns, zd* part, hipStream_t st) {
  const bool akc = (ta == 'T' || ta == 'C');
  if (tb != 'N' || !(akc || ta == 'N')) return hipErrorNotSupported;
  constexpr int BK = 16;
  if (ns > 1 && (kslice % BK != 0 || !part)) return hipErrorNotSupported;
  zg::Args a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.alpha = alpha;
  a.beta = beta;
  a.A = A;
  a.lda = lda;
  a.B = B;
  a.ldb = ldb;
  a.conja = (ta == 'C');
  const bool one_im0 = alpha.y == 0.0 && (alpha.x == 1.0 || alpha.x == -1.0);
  if (ns > 1) {
    a.mode = zg::kSplit;
    a.kslice = kslice;
    a.C = part;
    a.ldc = M;
    a.zstride = (size_t)M * N;
  } else {
    if (beta.x == 1.0 && beta.y == 0.0 && one_im0)
      a.mode = zg::kCInit;
    else if (beta.x == 0.0 && beta.y == 0.0)
      a.mode = (alpha.x == 1.0 && alpha.y == 0.0) ? zg::kSplit : zg::kBeta0;
    else
      a.mode = zg::kGeneral;
    a.kslice = std::max(BK, ((K + BK - 1) / BK) * BK);
    a.C = C;
    a.ldc = ldc;
    a.zstride = 0;
  }
  // 64 x 64 complex tiles (the accumulators of a 128 x 128 real tile);
  // 32-row tiles when M is a block of reflectors (W = V^H C, Z = T^H W).
  if (akc) {
    if (M <= 32) return zg::launch<true, 32, 64, BK>(a, ns, st);
    return zg::launch<true, 64, 64, BK>(a, ns, st);
  }
  if (M <= 32) return zg::launch<false, 32, 64, BK>(a, ns, st);
  return zg::launch<false, 64, 64, BK>(a, ns, st);
}

}  // namespace rnpx
