// rnpx_cqr.hip -- panel factorization by CholeskyQR2 with Householder
// reconstruction, for real panels of 64 columns.
//
// The column-by-column panel (rnpx_panel.hip) needs one grid-wide reduction
// per column.  This kernel needs two grid-wide reductions per panel:
//
//   1. Q1 R1 = A     (CholeskyQR: Gram A^T A over the grid, Cholesky, A R1^-1)
//   2. Q  R2 = Q1    (the second pass restores orthogonality to working
//                     precision when the first pass was accurate enough)
//   3. Householder reconstruction: Q - [D; 0] = Y U (LU without pivoting,
//      D = diag(+-1) chosen column by column), so that
//        Y = the Householder vectors V (unit lower trapezoidal),
//        tau_j = -U_jj D_j,  T = -U D Y1^-T,  R_householder = D R2 R1.
//
// In exact arithmetic this is the factorization the column-by-column
// reflectors produce (RNP/LA/QR.hpp:78-102 with Reflector::Generate,
// RNP/LA/Reflector.hpp:117-158): with the panel's Householder form
// H = I - Y T Y^T, H [I; 0] = Q D, hence Q D - [I; 0] = -Y T Y1^T, an LU
// factorization with a unit lower factor.  The sign D_j = -sign of the
// running pivot makes |U_jj| = tau_j in [1, 2] and matches Generate's
// choice beta_j = -sign(alpha_j) |.|; a zero pivot takes D_j = +1, the
// reference's beta > 0 when real(alpha) == 0 (Reflector.hpp:131).
//
// The second pass's Gram G2 = Q1^T Q1 = I + E is within 2^-27 of I or the
// panel is rejected.  When |E| <= 2^-40 its Cholesky factor is taken to
// first order, R2 = I + E_U with E_U = triu(E) - diag(E)/2, and
// R2^-1 = I - E_U (both exact to O(|E|^2) <= 2^-68); otherwise it is
// factored.
//
// The panel is refactored column by column instead (by workgroup 0, inside
// the same launch; the fail flag records it) when the first Cholesky meets a
// pivot below 2^-40 of its diagonal entry, a column whose squared norm is
// outside [2^-600, 2^600] (so the Gram entries cannot under/overflow), a
// non-finite value, or the second pass's test above fails.  Every workgroup
// takes the same decision from the same reduced Gram.
//
// Layout: G workgroups of 256 threads; workgroup g stages rows
// [g*RW, (g+1)*RW) of the panel in LDS (row-major, padded), so the panel is
// read from HBM once and written once.  Gram partials use FP64 MFMA over the
// 10 upper 16x16 tiles; they are summed in groups of 16 workgroups by the
// last workgroup of each group to arrive, then across groups by the last
// group, always in workgroup order (deterministic); everybody else waits on
// a done flag.  The 64x64 Cholesky, triangular inverses and the LU run in
// 16x16 diagonal blocks (one wave each step, v_readlane broadcasts) with the
// off-diagonal blocks as MFMA block products over the four waves
// (rnpx_small64.hpp); workgroup 0 does the LU and publishes U^-1, workgroup 1
// forms R, one more workgroup forms T.
#include "rnpx_kernels.hpp"
#include "rnpx_small64.hpp"
#include "rnpx_team.hpp"

#include <cstdlib>

namespace rnpx {

namespace {

constexpr int NB = 64;
constexpr int LDA_S = NB + 1;  // padded LDS row stride (doubles)
constexpr int kTiles = 10;     // upper 16x16 tiles of a 64x64 symmetric matrix
constexpr int kGramElems = kTiles * 256;
constexpr int kGroup = 16;     // workgroups per first-level reduction group
constexpr double kFirstOrder = 0x1p-40;
constexpr double kSecondPass = 0x1p-27;
// sync words: [pass*32 + grp] group arrivals, [pass*32 + 16] group-leader
// arrivals, [pass*32 + 17] done flag; [64] U^-1 published.
constexpr int kSyncTop = 16, kSyncDone = 17, kSyncPub = 64;

__device__ __forceinline__ void tile_of(int t, int& ti, int& tj) {
  // (0,0) (0,1) (0,2) (0,3) (1,1) (1,2) (1,3) (2,2) (2,3) (3,3)
  ti = (t >= 4) + (t >= 7) + (t >= 9);
  const int first = (ti == 0) ? 0 : (ti == 1) ? 4 : (ti == 2) ? 7 : 9;
  tj = ti + (t - first);
}

// Partial Gram P = Ab^T Ab of this workgroup's RW rows; tile t of the 10 is
// written row-major (16x16) to out[t*256 ...].
template <int RW>
__device__ __forceinline__ void gram_partial(const double (*Ab)[LDA_S], double* out, int lane, int wave) {
  const int t0 = (wave == 0) ? 0 : (wave == 1) ? 3 : (wave == 2) ? 6 : 8;  // 3, 3, 2, 2 tiles
  const int nt = (wave < 2) ? 3 : 2;
  for (int q = 0; q < nt; ++q) {
    int ti, tj;
    tile_of(t0 + q, ti, tj);
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < RW; k0 += 4) {
      const double x = Ab[k0 + (lane >> 4)][16 * ti + (lane & 15)];
      const double y = Ab[k0 + (lane >> 4)][16 * tj + (lane & 15)];
      acc = mfma16(x, y, acc);
    }
    // item r = P[16 ti + lane/16 + 4r][16 tj + lane%16]
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(t0 + q) * 256 + ((lane >> 4) + 4 * r) * 16 + (lane & 15)] = acc[r];
  }
}

// Sum of partials src[w * kGramElems + e], w = 0..n-1 (n <= 16), in order.
// Thread tid owns the entry pairs 2 tid + 512 q; eight partials' loads are
// in flight at once.
template <int ELEMS = kGramElems, int STRIDE = ELEMS>
__device__ __forceinline__ void sum_partials(const double* src, int n, double* dst, int tid) {
  constexpr int Q = ELEMS / 512;  // pairs per thread
  constexpr int W = 8;                 // partials per batch
  double2 acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = make_double2(0.0, 0.0);
  for (int w0 = 0; w0 < n; w0 += W) {
    double2 v[W][Q];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const double2* p = reinterpret_cast<const double2*>(src + (size_t)(w0 + w) * STRIDE) + tid;
#pragma unroll
      for (int q = 0; q < Q; ++q) v[w][q] = (w0 + w < n) ? p[256 * q] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (w0 + w < n) {
          acc[q].x += v[w][q].x;
          acc[q].y += v[w][q].y;
        }
  }
  double2* d = reinterpret_cast<double2*>(dst) + tid;
#pragma unroll
  for (int q = 0; q < Q; ++q) d[256 * q] = acc[q];
}

// After this workgroup's partial is in part[g]: the reduced Gram of the
// pass lands in gs for every workgroup.
__device__ __forceinline__ void gram_allreduce(const CqrArgs& a, int pass, int G, int g, int tid, unsigned* slot) {
  unsigned* sw = a.sync + pass * 32;
  double* gs = a.gsum + pass * kGramElems;
  const int grp = g / kGroup, ngroups = (G + kGroup - 1) / kGroup;
  const int gbeg = grp * kGroup, gsize = min(kGroup, G - gbeg);
  if (arrive(&sw[grp], slot) == (unsigned)(gsize - 1)) {
    // last of its group
    if (ngroups == 1) {
      sum_partials(a.part, gsize, gs, tid);
      publish(&sw[kSyncDone]);
    } else {
      sum_partials(a.part + (size_t)gbeg * kGramElems, gsize, a.gpart + (size_t)grp * kGramElems, tid);
      if (arrive(&sw[kSyncTop], slot) == (unsigned)(ngroups - 1)) {
        sum_partials(a.gpart, ngroups, gs, tid);
        publish(&sw[kSyncDone]);
      }
    }
  }
  wait_flag(&sw[kSyncDone], a.abortf);
}

// Load the reduced symmetric Gram into M (full 64x64).  Here and in the
// other global-to-LDS copies every load is issued before the first LDS
// store, so that the loads' round trips overlap.
__device__ __forceinline__ void load_gram(const double* gs, double (*M)[LDA_S], int tid) {
  constexpr int Q = kGramElems / 256;
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = gs[tid + 256 * q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + 256 * q;
    const int t = e >> 8, r = (e >> 4) & 15, c = e & 15;
    int ti, tj;
    tile_of(t, ti, tj);
    if (ti == tj && r > c) continue;  // one writer per entry of a diagonal tile
    M[16 * ti + r][16 * tj + c] = v[q];
    M[16 * tj + c][16 * ti + r] = v[q];
  }
}

// M := the row-major 64x64 block at src (global memory).
__device__ __forceinline__ void load64(const double* src, double (*M)[LDA_S], int tid) {
  constexpr int Q = NB * NB / 256;
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = src[tid + 256 * q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + 256 * q;
    M[e / NB][e % NB] = v[q];
  }
}

// Rows [0, RW) of Ab := Ab * X for an upper triangular X (64x64) on MFMA;
// Ab is overwritten after every wave has read its operands.
template <int RW>
__device__ __forceinline__ void times_upper(double (*Ab)[LDA_S], const double (*X)[LDA_S], int lane, int wave) {
  constexpr int RT = RW / 64;  // row tiles per wave
  f64x4 acc[RT][4];
#pragma unroll
  for (int q = 0; q < RT; ++q) {
    const int t = wave + 4 * q;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      acc[q][c] = f64x4{0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < 16 * (c + 1); k0 += 4) {  // X[k][16c..] = 0 for k >= 16(c+1)
        const double x = Ab[16 * t + (lane & 15)][k0 + (lane >> 4)];
        const double y = X[k0 + (lane >> 4)][16 * c + (lane & 15)];
        acc[q][c] = mfma16(x, y, acc[q][c]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RT; ++q) {
    const int t = wave + 4 * q;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) Ab[16 * t + (lane >> 4) + 4 * r][16 * c + (lane & 15)] = acc[q][c][r];
  }
  __syncthreads();
}

// Wave w's tile row of C = A B for 64x64 upper triangular A, B (row-major in
// LDS; entries of A below its diagonal are treated as zero when mask_a, and
// column k of A is scaled by dcol[k] when dcol is given):
// acc[tj] item r = C[16 w + lane/16 + 4r][16 tj + lane%16].
template <bool TB = false>  // TB: B is given transposed (B[k][j] read as B[j][k])
__device__ __forceinline__ void upper_product_row(const double (*A)[LDA_S], const double (*B)[LDA_S], bool mask_a, const double* dcol,
                                  f64x4 (&acc)[4], int lane, int wave) {
  const int ti = wave;
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    acc[tj] = f64x4{0.0, 0.0, 0.0, 0.0};
    if (tj < ti) continue;
    for (int k0 = 16 * ti; k0 < 16 * (tj + 1); k0 += 4) {
      const int i = 16 * ti + (lane & 15), k = k0 + (lane >> 4);
      double x = A[i][k];
      if (mask_a && k < i) x = 0.0;
      if (dcol) x *= dcol[k];
      const double y = TB ? B[16 * tj + (lane & 15)][k] : B[k][16 * tj + (lane & 15)];
      acc[tj] = mfma16(x, y, acc[tj]);
    }
  }
}

// Largest |M - I| over the workgroup (a NaN entry makes the result NaN).
__device__ __forceinline__ double dev_from_identity(const double (*M)[LDA_S], double* red4, int tid) {
  double d = 0.0;
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e / NB, j = e % NB;
    const double x = fabs(M[i][j] - (i == j ? 1.0 : 0.0));
    d = (x > d || x != x) ? x : d;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(d, off, 64);
    d = (o > d || o != o) ? o : d;
  }
  if ((tid & 63) == 0) red4[tid >> 6] = d;
  __syncthreads();
  double r = red4[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) r = (red4[w] > r || red4[w] != red4[w]) ? red4[w] : r;
  __syncthreads();
  return r;
}

// The column-by-column refactorization of a rejected panel, by one
// workgroup after every other has left the kernel: team_factor
// (Reflector::Generate's rules, RNP/LA/Reflector.hpp:117-158, and the
// Householder application of QR::Factor_unblocked, QR.hpp:78-102) in place,
// the clean V, then T from V's Gram (GenerateBlockTr("F","C"),
// Reflector.hpp:580-626): T(c,c) = tau_c, T(0:c,c) = -tau_c T(0:c,0:c)
// V(:,0:c)^T v_c.  Column-serial, so slow, but rejections are rare.  No
// workgroup reads A after staging its rows, so rewriting A here is safe.
__device__ void refactor_columnwise(const CqrArgs& a, double (*Gs)[LDA_S], double (*Ts)[LDA_S], double* sh, int tid) {
  if (tid == 0) *a.fail = 1u;
  const int mp = a.mp;
  team_factor<double, 256, true>(mp, NB, a.A, (int)a.lda, a.tau, sh);
  __syncthreads();
  for (long e = tid; e < (long)mp * NB; e += 256) {
    const int r = (int)(e % mp), c = (int)(e / mp);
    a.V[(size_t)r + (size_t)c * a.ldv] = (r > c) ? a.A[(size_t)r + (size_t)c * a.lda] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int e = tid; e < NB * NB; e += 256) {
    const int l = e / NB, c = e % NB;
    double gv = 0.0;
    if (l < c)
      for (int r = c; r < mp; ++r) gv = fma(a.V[(size_t)r + (size_t)l * a.ldv], a.V[(size_t)r + (size_t)c * a.ldv], gv);
    Gs[l][c] = gv;
    Ts[l][c] = 0.0;
  }
  __syncthreads();
  for (int c = 0; c < NB; ++c) {
    const double tc = a.tau[c];
    if (tid < c) {
      double sacc = 0.0;
      for (int l = tid; l < c; ++l) sacc = fma(Ts[tid][l], Gs[l][c], sacc);
      Ts[tid][c] = -tc * sacc;
    }
    if (tid == c) Ts[c][c] = tc;
    __syncthreads();
  }
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e % NB, j = e / NB;
    a.T[(size_t)i + (size_t)j * a.ldt] = Ts[i][j];
  }
}

}  // namespace

template <int RW>
__global__ __launch_bounds__(256) void cqr_panel_kernel(CqrArgs a) {
  __shared__ double Ab[RW][LDA_S];
  __shared__ double M1[NB][LDA_S];
  __shared__ double M2[NB][LDA_S];
  __shared__ double rowbuf[NB];
  __shared__ double red4[4];
  __shared__ double gdiag[NB];
  __shared__ double tmpb[small64::kTmpDoubles];
  __shared__ unsigned slot;
  __shared__ int okf;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x, G = gridDim.x;
  const int r0 = g * RW;
  const int mp = a.mp;
  double* rsave = a.scratch;          // R1 (saved and used by workgroup 1)
  double* uinv = a.scratch + 4096;    // U^-1
  double* ulp = a.scratch + 8192;     // U and L packed
  double* dpub = a.scratch + 12288;   // D
  const int tgrp = (G > 2) ? 2 : 1;   // the workgroup that forms T

  // Stage this workgroup's rows (coalesced down each column), zero padding.
  {
    constexpr int PER = RW * NB / 256;
    double v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + 256 * q, r = e % RW, c = e / RW;
      v[q] = (r0 + r < mp) ? a.A[(size_t)(r0 + r) + (size_t)c * a.lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + 256 * q;
      Ab[e % RW][e / RW] = v[q];
    }
  }
  __syncthreads();

  // Pass 1: Q1 = A R1^-1.
  gram_partial<RW>(Ab, a.part + (size_t)g * kGramElems, lane, wave);
  gram_allreduce(a, 0, G, g, tid, &slot);
  load_gram(a.gsum, M1, tid);
  __syncthreads();
  const bool ok1 = small64::chol64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, gdiag, &okf, tid);  // M1 = R1, M2 = R1^-1
  if (!ok1) {
    // Every workgroup factored the same Gram and takes this branch; nothing
    // of the panel has been written.
    if (g == 0) refactor_columnwise(a, M1, M2, rowbuf, tid);
    return;
  }
  if (g == 1)
    for (int e = tid; e < NB * NB; e += 256) rsave[e] = M1[e / NB][e % NB];
  times_upper<RW>(Ab, M2, lane, wave);

  // Pass 2: Q = Q1 R2^-1.
  gram_partial<RW>(Ab, a.part + (size_t)g * kGramElems, lane, wave);
  gram_allreduce(a, 1, G, g, tid, &slot);
  load_gram(a.gsum + kGramElems, M1, tid);
  __syncthreads();
  const double dev = dev_from_identity(M1, red4, tid);
  if (!(dev <= kSecondPass)) {
    if (g == 0) refactor_columnwise(a, M1, M2, rowbuf, tid);
    return;
  }
  if (dev <= kFirstOrder) {
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e / NB, j = e % NB;
      const double E = M1[i][j] - (i == j ? 1.0 : 0.0);
      const double eu = (i < j) ? E : ((i == j) ? 0.5 * E : 0.0);
      M1[i][j] = (i == j ? 1.0 : 0.0) + eu;  // R2
      M2[i][j] = (i == j ? 1.0 : 0.0) - eu;  // R2^-1
    }
    __syncthreads();
  } else {
    small64::chol64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, gdiag, &okf, tid);
  }
  times_upper<RW>(Ab, M2, lane, wave);
  // Ab = Q (this workgroup's rows), M1 = R2.

  if (g == 0) {
    // LU of Q_top - D, published with U^-1.
    for (int e = tid; e < NB * NB; e += 256) M1[e / NB][e % NB] = Ab[e / NB][e % NB];
    __syncthreads();
    small64::lu64<LDA_S>(&M1[0][0], rowbuf, &M2[0][0], tmpb, tid);  // M1 = U and L, M2 = U^-1
    for (int e = tid; e < NB * NB; e += 256) {
      ulp[e] = M1[e / NB][e % NB];
      uinv[e] = M2[e / NB][e % NB];
    }
    if (tid < NB) dpub[tid] = rowbuf[tid];
    publish(&a.sync[kSyncPub]);
    // Rows 64..RW-1 of this workgroup: Y = Q U^-1.
    if constexpr (RW > NB) {
      __syncthreads();
      times_upper<RW>(Ab, M2, lane, wave);
      for (int e = tid; e < (RW - NB) * NB; e += 256) {
        const int r = NB + e % (RW - NB), c = e / (RW - NB);
        const int gr = r0 + r;
        if (gr < mp) {
          const double v = Ab[r][c];
          a.A[(size_t)gr + (size_t)c * a.lda] = v;
          a.V[(size_t)gr + (size_t)c * a.ldv] = v;
        }
      }
    }
    // Top block: L below the diagonal (R above is written by workgroup 1),
    // V's unit lower block, tau.
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const double l = M1[i][j];
      if (i > j) a.A[(size_t)i + (size_t)j * a.lda] = l;
      a.V[(size_t)i + (size_t)j * a.ldv] = (i == j) ? 1.0 : ((i > j) ? l : 0.0);
    }
    if (tid < NB) a.tau[tid] = -M1[tid][tid] * rowbuf[tid];
    return;
  }

  f64x4 racc[4];
  if (g == 1) {
    // R = R2 R1 while workgroup 0 reconstructs.
    for (int e = tid; e < NB * NB; e += 256) M2[e / NB][e % NB] = rsave[e];
    __syncthreads();
    upper_product_row(M1, M2, false, nullptr, racc, lane, wave);
  }
  wait_flag(&a.sync[kSyncPub], a.abortf);
  if (g == 1) {
    // R_householder = D R, the upper triangle of the top block.
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
        if (i <= j) a.A[(size_t)i + (size_t)j * a.lda] = dpub[i] * racc[tj][r];
      }
  }
  // Y = Q U^-1 for this workgroup's rows.
  __syncthreads();
  for (int e = tid; e < NB * NB; e += 256) M2[e / NB][e % NB] = uinv[e];
  __syncthreads();
  times_upper<RW>(Ab, M2, lane, wave);
  for (int e = tid; e < RW * NB; e += 256) {
    const int r = e % RW, c = e / RW;
    const int gr = r0 + r;
    if (gr < mp) {
      const double v = Ab[r][c];
      a.A[(size_t)gr + (size_t)c * a.lda] = v;
      a.V[(size_t)gr + (size_t)c * a.ldv] = v;
    }
  }
  if (g == tgrp) {
    // T = -U D Z^T with Z = L^-1.
    for (int e = tid; e < NB * NB; e += 256) M1[e / NB][e % NB] = ulp[e];
    if (tid < NB) rowbuf[tid] = dpub[tid];
    __syncthreads();
    small64::inv_unit_lower64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, tid);  // M2 = L^-1 (lower blocks)
    // Z^T is upper: stage it transposed in Ab's top rows.
    for (int e = tid; e < NB * NB; e += 256) Ab[e % NB][e / NB] = M2[e / NB][e % NB];
    __syncthreads();
    f64x4 tacc[4];
    upper_product_row(M1, Ab, true, rowbuf, tacc, lane, wave);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
        a.T[(size_t)i + (size_t)j * a.ldt] = (i <= j) ? -tacc[tj][r] : 0.0;
      }
  }
}

// ---------------------------------------------------------------------------
// One outer block of 128 columns in one launch: panel 1 (columns 0..63) by
// CholeskyQR2 + reconstruction as above, its block reflector applied to
// panel 2's columns (W = V1^T A2 by a grid reduction, Z = T1^T W,
// A2 -= V1 Z, all in LDS), panel 2 (rows 64.., columns 64..127) the same
// way over workgroups 1.., and the merged compact-WY form
//   V = [V1 | (0; V2)],  T = [T1, -T1 (V1^T V2) T2; 0, T2]
// with V1^T V2 by a third grid reduction.  This replaces, per outer block,
// two panel launches and the six small products around them.  Workgroup g
// holds rows [64 g, 64 g + 64) of both panels in LDS for the whole launch;
// nothing is written to A before panel 2 has passed its tests, so a
// rejection of either panel leaves A untouched and workgroup 0 refactors
// all 128 columns column by column instead.  Each panel's T is formed by the
// workgroup that factored its LU (still in LDS) while the next grid
// reduction completes: T1 by workgroup 0, T2 and the merge by workgroup 1.
// ---------------------------------------------------------------------------
namespace {

constexpr int kBlkPassWords = 32;                 // per reduction pass
constexpr int kBlkPub1 = 6 * kBlkPassWords;        // panel 1's LU published
constexpr int kBlkPubT1 = kBlkPub1 + 1;            // T1 published
constexpr int kBlkPub2 = kBlkPub1 + 2;             // panel 2's LU published
constexpr int kBlkResident = kBlkPub1 + 3;         // workgroups started (for the gate below)
static_assert(kBlkPub2 < kCqrBlockSyncWords, "block sync words");
constexpr int kCross = 16 * 256;                   // 64x64 product partial (16 tiles)
// scratch (doubles)
constexpr int kBR1a = 0, kBU1i = 4096, kBD1 = 12288, kBTau1 = 12352, kBRF1 = 12416, kBT1 = 16512,
              kBR1b = 20608, kBU2i = 24704, kBD2 = 32896, kBG = 32960, kBTF = 49344;
static_assert(kBTF + 16384 <= kCqrBlockScratch, "block scratch");

// Reduction of Gp partials (ELEMS doubles each, kCross apart from a.part)
// into gs, by reduce-scatter.  Contributors arrive on the pass's first
// counter once their partial is stored.  Each helper workgroup (index hi of
// Hn; hi < 0 for none) then waits for all Gp arrivals, sums its slice of
// the entries over the partials in a fixed order (so the result does not
// depend on timing), stores it and arrives on the second counter; the sum
// is complete once Hn helpers have arrived (blk_wait).  red: 256 doubles of
// LDS scratch.
template <int ELEMS>
__device__ __forceinline__ void blk_contribute(const CqrBlockArgs& a, int pass, bool contrib, int Gp, int hi, int Hn,
                                               double* gs, double* red, int tid, unsigned* slot) {
  unsigned* sw = a.sync + pass * kBlkPassWords;
  if (contrib) arrive(&sw[0], slot);
  if (hi < 0) return;
  wait_count(&sw[0], (unsigned)Gp, a.abortf);
  const int S = (ELEMS + Hn - 1) / Hn;  // slice length
  const int e0 = hi * S, len = min(S, ELEMS - e0);
  if (S <= 128) {
    // P threads per entry, each over every P-th partial; then the P sums
    const int P = 256 / S, j = tid / S, x = tid % S;
    double sum = 0.0;
    if (j < P && x < len) {
      const double* src = a.part + e0 + x;
#pragma unroll 8
      for (int q = j; q < Gp; q += P) sum += src[(size_t)q * kCross];
    }
    red[tid] = sum;
    __syncthreads();
    if (tid < len) {
      double t = red[tid];
      for (int q = 1; q < P; ++q) t += red[q * S + tid];
      gs[e0 + tid] = t;
    }
  } else {
    for (int x = tid; x < len; x += 256) {
      const double* src = a.part + e0 + x;
      double sum = 0.0;
#pragma unroll 8
      for (int q = 0; q < Gp; ++q) sum += src[(size_t)q * kCross];
      gs[e0 + x] = sum;
    }
  }
  arrive(&sw[1], slot);
}
__device__ __forceinline__ void blk_wait(const CqrBlockArgs& a, int pass, int Hn) {
  wait_count(&a.sync[pass * kBlkPassWords + 1], (unsigned)Hn, a.abortf);
}
template <int ELEMS>
__device__ __forceinline__ void blk_allreduce(const CqrBlockArgs& a, int pass, bool contrib, int Gp, int hi, int Hn,
                                              double* gs, double* red, int tid, unsigned* slot) {
  blk_contribute<ELEMS>(a, pass, contrib, Gp, hi, Hn, gs, red, tid, slot);
  blk_wait(a, pass, Hn);
}

// P = X^T Y over 64 rows (16 tiles of 16x16, tile t = 4 ti + tj, row-major).
__device__ __forceinline__ void cross_partial(const double (*X)[LDA_S], const double (*Y)[LDA_S], double* out, int lane,
                                              int wave) {
  const int ti = wave;
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < 64; k0 += 4) {
      const double x = X[k0 + (lane >> 4)][16 * ti + (lane & 15)];
      const double y = Y[k0 + (lane >> 4)][16 * tj + (lane & 15)];
      acc = mfma16(x, y, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(4 * ti + tj) * 256 + ((lane >> 4) + 4 * r) * 16 + (lane & 15)] = acc[r];
  }
}

__device__ __forceinline__ void load_tiles16(const double* gs, double (*M)[LDA_S], int tid) {
  constexpr int Q = kCross / 256;
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = gs[tid + 256 * q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + 256 * q;
    const int t = e >> 8, r = (e >> 4) & 15, c = e & 15;
    M[16 * (t >> 2) + r][16 * (t & 3) + c] = v[q];
  }
}

// acc[tj] = row block `wave` of op(A) B, 64x64 full operands in LDS;
// item r = C[16 wave + lane/16 + 4 r][16 tj + lane%16].
template <bool TA>
__device__ __forceinline__ void prod64(const double (*A)[LDA_S], const double (*B)[LDA_S], f64x4 (&acc)[4], int lane,
                                       int wave) {
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    acc[tj] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < 64; k0 += 4) {
      const int i = 16 * wave + (lane & 15), k = k0 + (lane >> 4);
      const double x = TA ? A[k][i] : A[i][k];
      const double y = B[k][16 * tj + (lane & 15)];
      acc[tj] = mfma16(x, y, acc[tj]);
    }
  }
}

__device__ __forceinline__ void store_acc(const f64x4 (&acc)[4], double (*C)[LDA_S], double scale, int lane, int wave) {
#pragma unroll
  for (int tj = 0; tj < 4; ++tj)
#pragma unroll
    for (int r = 0; r < 4; ++r) C[16 * wave + (lane >> 4) + 4 * r][16 * tj + (lane & 15)] = scale * acc[tj][r];
}

// Rejected block: all 128 columns column by column in place (team_factor),
// the clean V, and T (128 x 128) from V's Gram by the recursion of
// GenerateBlockTr("F","C") (Reflector.hpp:580-626), the Gram and T held in
// scratch.  One workgroup; the others have left.
__device__ __noinline__ void blk_refactor(CqrBlockArgs a, double* sh, int tid) {
  constexpr int K = 2 * NB;
  if (tid == 0) *a.fail = 1u;
  const int mp = a.mp;
  double* Gm = a.scratch + kBG;
  double* Tm = a.scratch + kBTF;
  team_factor<double, 256, true>(mp, K, a.A, (int)a.lda, a.tau, sh);
  __syncthreads();
  for (long e = tid; e < (long)mp * K; e += 256) {
    const int r = (int)(e % mp), c = (int)(e / mp);
    a.V[(size_t)r + (size_t)c * a.ldv] = (r > c) ? a.A[(size_t)r + (size_t)c * a.lda] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int e = tid; e < K * K; e += 256) {
    const int l = e / K, c = e % K;
    double gv = 0.0;
    if (l < c)
      for (int r = c; r < mp; ++r) gv = fma(a.V[(size_t)r + (size_t)l * a.ldv], a.V[(size_t)r + (size_t)c * a.ldv], gv);
    Gm[e] = gv;
    Tm[e] = 0.0;
  }
  __syncthreads();
  for (int c = 0; c < K; ++c) {
    const double tc = a.tau[c];
    if (tid < c) {
      double sacc = 0.0;
      for (int l = tid; l < c; ++l) sacc = fma(Tm[tid * K + l], Gm[l * K + c], sacc);
      Tm[tid * K + c] = -tc * sacc;
    }
    if (tid == c) Tm[c * K + c] = tc;
    __syncthreads();
  }
  for (int e = tid; e < K * K; e += 256) {
    const int i = e % K, j = e / K;
    a.T[(size_t)i + (size_t)j * a.ldt] = Tm[i * K + j];
  }
}

// The reconstruction's R = R2 R1 product (R1 from scratch at r1off) into acc;
// M1 holds R2.
__device__ __forceinline__ void blk_r_product(const double (*M1)[LDA_S], double (*M2)[LDA_S], const double* r1,
                                              f64x4 (&acc)[4], int tid, int lane, int wave) {
  load64(r1, M2, tid);
  __syncthreads();
  upper_product_row(M1, M2, false, nullptr, acc, lane, wave);
}

// T = -U D Z^T, Z = L^-1, from U/L packed in M1 and D in rowbuf (left there
// by lu64); acc holds it (upper).  M2 is overwritten.
__device__ __forceinline__ void blk_t_factor(const double (*M1)[LDA_S], double (*M2)[LDA_S], const double* rowbuf,
                                             double* tmpb, f64x4 (&acc)[4], int tid, int lane, int wave) {
  small64::inv_unit_lower64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, tid);  // M2 = L^-1 (lower blocks)
  upper_product_row<true>(M1, M2, true, rowbuf, acc, lane, wave);
#pragma unroll
  for (int tj = 0; tj < 4; ++tj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
      acc[tj][r] = (i <= j) ? -acc[tj][r] : 0.0;
    }
}

// R2 and R2^-1 from the second pass's Gram in M1 (dev <= kSecondPass):
// first order when |E| <= 2^-40, factored otherwise.
__device__ __forceinline__ void blk_second_r(double (*M1)[LDA_S], double (*M2)[LDA_S], double dev, double* tmpb,
                                             double* gdiag, int* okf, int tid) {
  if (dev <= kFirstOrder) {
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e / NB, j = e % NB;
      const double E = M1[i][j] - (i == j ? 1.0 : 0.0);
      const double eu = (i < j) ? E : ((i == j) ? 0.5 * E : 0.0);
      M1[i][j] = (i == j ? 1.0 : 0.0) + eu;
      M2[i][j] = (i == j ? 1.0 : 0.0) - eu;
    }
    __syncthreads();
  } else {
    small64::chol64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, gdiag, okf, tid);
  }
}

}  // namespace

__global__ __launch_bounds__(256) void cqr_block_kernel(CqrBlockArgs a) {
  RNPX_CQR_GLOBAL(a);
  if (threadIdx.x == 0) __hip_atomic_fetch_add(&a.sync[kBlkResident], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __shared__ double Ab1[64][LDA_S];
  __shared__ double Ab2[64][LDA_S];
  __shared__ double M1[NB][LDA_S];
  __shared__ double M2[NB][LDA_S];
  __shared__ double rowbuf[NB];
  __shared__ double red4[4];
  __shared__ double gdiag[NB];
  __shared__ double tmpb[small64::kTmpDoubles];
  __shared__ unsigned slot;
  __shared__ int okf;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x, G = gridDim.x;
  const int r0 = 64 * g;
  const int mp = a.mp;
  double* sc = a.scratch;
  double* part = a.part + (size_t)g * kCross;

  // Both panels' rows of this workgroup.
  {
    constexpr int PER = 64 * NB / 256;
    double v[2][PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + 256 * q, r = e % 64, c = e / 64;
      const bool in = r0 + r < mp;
      v[0][q] = in ? a.A[(size_t)(r0 + r) + (size_t)c * a.lda] : 0.0;
      v[1][q] = in ? a.A[(size_t)(r0 + r) + (size_t)(c + NB) * a.lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + 256 * q;
      Ab1[e % 64][e / 64] = v[0][q];
      Ab2[e % 64][e / 64] = v[1][q];
    }
  }
  __syncthreads();

  // ---- Panel 1 ----
  gram_partial<64>(Ab1, part, lane, wave);
  blk_allreduce<kGramElems>(a, 0, true, G, g, G, a.gsum, tmpb, tid, &slot);
  load_gram(a.gsum, M1, tid);
  __syncthreads();
  if (!small64::chol64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, gdiag, &okf, tid)) {
    if (g == 0) blk_refactor(a, rowbuf, tid);
    return;
  }
  if (g == 1)
    for (int e = tid; e < NB * NB; e += 256) sc[kBR1a + e] = M1[e / NB][e % NB];
  times_upper<64>(Ab1, M2, lane, wave);
  gram_partial<64>(Ab1, part, lane, wave);
  blk_allreduce<kGramElems>(a, 1, true, G, g, G, a.gsum + kCross, tmpb, tid, &slot);
  load_gram(a.gsum + kCross, M1, tid);
  __syncthreads();
  double dev = dev_from_identity(M1, red4, tid);
  if (!(dev <= kSecondPass)) {
    if (g == 0) blk_refactor(a, rowbuf, tid);
    return;
  }
  blk_second_r(M1, M2, dev, tmpb, gdiag, &okf, tid);
  times_upper<64>(Ab1, M2, lane, wave);  // Ab1 = Q, M1 = R2
  if (g == 0) {
    for (int e = tid; e < NB * NB; e += 256) M1[e / NB][e % NB] = Ab1[e / NB][e % NB];
    __syncthreads();
    small64::lu64<LDA_S>(&M1[0][0], rowbuf, &M2[0][0], tmpb, tid);  // M1 = U and L, M2 = U^-1
    for (int e = tid; e < NB * NB; e += 256) sc[kBU1i + e] = M2[e / NB][e % NB];
    if (tid < NB) {
      sc[kBD1 + tid] = rowbuf[tid];
      sc[kBTau1 + tid] = -M1[tid][tid] * rowbuf[tid];
    }
    publish(&a.sync[kBlkPub1]);
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e / NB, j = e % NB;
      Ab1[i][j] = (i > j) ? M1[i][j] : (i == j ? 1.0 : 0.0);  // V1's top block: L1
    }
    __syncthreads();
  } else {
    f64x4 racc[4];
    if (g == 1) blk_r_product(M1, M2, sc + kBR1a, racc, tid, lane, wave);
    wait_flag(&a.sync[kBlkPub1], a.abortf);
    if (g == 1) {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
          sc[kBRF1 + i * NB + j] = sc[kBD1 + i] * racc[tj][r];
        }
    }
    __syncthreads();
    load64(sc + kBU1i, M2, tid);
    __syncthreads();
    times_upper<64>(Ab1, M2, lane, wave);  // Ab1 = V1 rows
  }

  // ---- Panel 1's reflectors applied to panel 2: A2 -= V1 (T1^T (V1^T A2)) ----
  cross_partial(Ab1, Ab2, part, lane, wave);
  blk_contribute<kCross>(a, 2, true, G, g - 1, G - 1, a.gsum + 2 * kCross, tmpb, tid, &slot);  // g0: T1
  if (g == 0) {
    // T1 while the reduction completes; the LU is still in M1 and rowbuf.
    f64x4 tacc[4];
    blk_t_factor(M1, M2, rowbuf, tmpb, tacc, tid, lane, wave);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
        sc[kBT1 + i * NB + j] = tacc[tj][r];
        a.T[(size_t)i + (size_t)j * a.ldt] = tacc[tj][r];
      }
    publish(&a.sync[kBlkPubT1]);
  }
  blk_wait(a, 2, G - 1);
  wait_flag(&a.sync[kBlkPubT1], a.abortf);
  load_tiles16(a.gsum + 2 * kCross, M2, tid);                                     // W
  load64(sc + kBT1, M1, tid);  // T1
  __syncthreads();
  {
    f64x4 acc[4];
    prod64<true>(M1, M2, acc, lane, wave);  // Z = T1^T W
    __syncthreads();
    store_acc(acc, M1, 1.0, lane, wave);
    __syncthreads();
    prod64<false>(Ab1, M1, acc, lane, wave);  // V1 Z
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) Ab2[16 * wave + (lane >> 4) + 4 * r][16 * tj + (lane & 15)] -= acc[tj][r];
    __syncthreads();
  }

  // ---- Panel 2: rows 64.. (workgroups 1..), LU and T by workgroup 1, R by 2 ----
  const bool p2 = (g >= 1);
  const int gi = g - 1, G2 = G - 1;
  if (p2) gram_partial<64>(Ab2, a.part + (size_t)gi * kCross, lane, wave);
  blk_allreduce<kGramElems>(a, 3, p2, G2, g, G, a.gsum + 3 * kCross, tmpb, tid, &slot);
  load_gram(a.gsum + 3 * kCross, M1, tid);
  __syncthreads();
  if (!small64::chol64<LDA_S>(&M1[0][0], &M2[0][0], tmpb, gdiag, &okf, tid)) {
    if (g == 0) blk_refactor(a, rowbuf, tid);
    return;
  }
  if (g == 2)
    for (int e = tid; e < NB * NB; e += 256) sc[kBR1b + e] = M1[e / NB][e % NB];
  if (p2) times_upper<64>(Ab2, M2, lane, wave);
  else __syncthreads();
  if (p2) gram_partial<64>(Ab2, a.part + (size_t)gi * kCross, lane, wave);
  blk_allreduce<kGramElems>(a, 4, p2, G2, g, G, a.gsum + 4 * kCross, tmpb, tid, &slot);
  load_gram(a.gsum + 4 * kCross, M1, tid);
  __syncthreads();
  dev = dev_from_identity(M1, red4, tid);
  if (!(dev <= kSecondPass)) {
    if (g == 0) blk_refactor(a, rowbuf, tid);
    return;
  }
  blk_second_r(M1, M2, dev, tmpb, gdiag, &okf, tid);

  // Both panels are accepted: panel 1's results go out now.
  if (g == 0) {
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const double v = Ab1[i][j];
      if (i > j) a.A[(size_t)i + (size_t)j * a.lda] = v;
      a.V[(size_t)i + (size_t)j * a.ldv] = v;
      a.A[(size_t)i + (size_t)(j + NB) * a.lda] = Ab2[i][j];  // R12
      a.V[(size_t)i + (size_t)(j + NB) * a.ldv] = 0.0;
      a.T[(size_t)(i + NB) + (size_t)j * a.ldt] = 0.0;
    }
    if (tid < NB) a.tau[tid] = sc[kBTau1 + tid];
  } else {
    for (int e = tid; e < 64 * NB; e += 256) {
      const int r = e % 64, c = e / 64, gr = r0 + r;
      if (gr < mp) {
        const double v = Ab1[r][c];
        a.A[(size_t)gr + (size_t)c * a.lda] = v;
        a.V[(size_t)gr + (size_t)c * a.ldv] = v;
      }
    }
    if (g == 1)
      for (int e = tid; e < NB * NB; e += 256) {
        const int i = e % NB, j = e / NB;
        if (i <= j) a.A[(size_t)i + (size_t)j * a.lda] = sc[kBRF1 + i * NB + j];
      }
  }

  if (p2) times_upper<64>(Ab2, M2, lane, wave);  // Ab2 = Q2 (g >= 1), M1 = R2'
  else __syncthreads();
  if (g == 1) {
    for (int e = tid; e < NB * NB; e += 256) M1[e / NB][e % NB] = Ab2[e / NB][e % NB];
    __syncthreads();
    small64::lu64<LDA_S>(&M1[0][0], rowbuf, &M2[0][0], tmpb, tid);
    for (int e = tid; e < NB * NB; e += 256) sc[kBU2i + e] = M2[e / NB][e % NB];
    if (tid < NB) sc[kBD2 + tid] = rowbuf[tid];
    publish(&a.sync[kBlkPub2]);
    for (int e = tid; e < NB * NB; e += 256) {
      const int i = e % NB, j = e / NB;
      const double l = M1[i][j];
      const double v = (i > j) ? l : (i == j ? 1.0 : 0.0);
      Ab2[i][j] = v;
      if (i > j) a.A[(size_t)(NB + i) + (size_t)(NB + j) * a.lda] = l;
      a.V[(size_t)(NB + i) + (size_t)(NB + j) * a.ldv] = v;
    }
    if (tid < NB) a.tau[NB + tid] = -M1[tid][tid] * rowbuf[tid];
    __syncthreads();
  } else if (g >= 2) {
    f64x4 racc[4];
    if (g == 2) blk_r_product(M1, M2, sc + kBR1b, racc, tid, lane, wave);
    wait_flag(&a.sync[kBlkPub2], a.abortf);
    if (g == 2) {
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
          if (i <= j) a.A[(size_t)(NB + i) + (size_t)(NB + j) * a.lda] = sc[kBD2 + i] * racc[tj][r];
        }
    }
    __syncthreads();
    load64(sc + kBU2i, M2, tid);
    __syncthreads();
    times_upper<64>(Ab2, M2, lane, wave);  // V2 rows
    for (int e = tid; e < 64 * NB; e += 256) {
      const int r = e % 64, c = e / 64, gr = r0 + r;
      if (gr < mp) {
        const double v = Ab2[r][c];
        a.A[(size_t)gr + (size_t)(c + NB) * a.lda] = v;
        a.V[(size_t)gr + (size_t)(c + NB) * a.ldv] = v;
      }
    }
  }

  // ---- Merge: T12 = -T1 (V1^T V2) T2, by workgroup 1 (which holds LU2) ----
  if (!p2) return;
  cross_partial(Ab1, Ab2, a.part + (size_t)gi * kCross, lane, wave);
  blk_contribute<kCross>(a, 5, true, G2, g - 2, G - 2, a.gsum + 5 * kCross, tmpb, tid, &slot);  // g1: T2
  if (g != 1) return;
  {
    f64x4 tacc[4];
    blk_t_factor(M1, M2, rowbuf, tmpb, tacc, tid, lane, wave);  // T2
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
        a.T[(size_t)(NB + i) + (size_t)(NB + j) * a.ldt] = tacc[tj][r];
      }
    store_acc(tacc, Ab2, 1.0, lane, wave);  // Ab2 was last read before the arrival's barrier
  }
  blk_wait(a, 5, G - 2);
  load_tiles16(a.gsum + 5 * kCross, M2, tid);                                     // S
  load64(sc + kBT1, M1, tid);  // T1
  __syncthreads();
  f64x4 acc[4];
  prod64<false>(M1, M2, acc, lane, wave);  // X = T1 S
  __syncthreads();
  store_acc(acc, M1, 1.0, lane, wave);
  __syncthreads();
  prod64<false>(M1, Ab2, acc, lane, wave);  // X T2
#pragma unroll
  for (int tj = 0; tj < 4; ++tj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * wave + (lane >> 4) + 4 * r, j = 16 * tj + (lane & 15);
      a.T[(size_t)i + (size_t)(NB + j) * a.ldt] = -acc[tj][r];
    }
}

int cqr_block_groups(int mp, int max_groups) {
  if (mp < 3 * NB) return 0;
  const int G = (mp + 63) / 64;
  return (G >= 3 && G <= max_groups && G <= 16 * kGroup) ? G : 0;
}

hipError_t launch_cqr_block(const CqrBlockArgs& a, int G, hipStream_t st) {
  if (G < 3 || (long)G * 64 < a.mp || (long)(G - 1) * 64 >= a.mp || a.lda < a.mp || a.ldv < a.mp || a.ldt < 2 * NB ||
      !a.fail || !a.sync || !a.abortf)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(cqr_block_kernel, dim3(G), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Scheduling gate for the look-ahead: one thread waits until all G
// workgroups of the block kernel whose sync words start at block_sync have
// started, so that a GEMM queued behind the gate on another stream cannot
// take the CUs those workgroups need first.  Only the order of dispatch
// depends on it (the GEMM reads nothing the block kernel writes), so the
// wait gives up after 2 ms.  A wait that gives up sets *off, which turns every
// later gate of the handle into a no-op: the runtime can map the two
// streams onto one hardware queue (more streams in the process than
// queues), and then the block kernel sits behind the gate and every wait
// would run to its limit.
__global__ void cqr_gate_kernel(const unsigned* block_sync, unsigned G, unsigned* off) {
  if (threadIdx.x != 0) return;
  if (__hip_atomic_load(off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
  const unsigned* c = block_sync + kBlkResident;
  const uint64_t t0 = wall_clock64();
  while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < G) {
    if (wall_clock64() - t0 > 200000ull) {
      __hip_atomic_store(off, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

hipError_t launch_cqr_gate(const unsigned* block_sync, int G, unsigned* off, hipStream_t st) {
  if (!block_sync || !off || G < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(cqr_gate_kernel, dim3(1), dim3(64), 0, st, block_sync, (unsigned)G, off);
  return hipGetLastError();
}

int cqr_panel_groups(int mp, int max_groups, int* rw) {
  if (mp < 2 * NB) return 0;
  if (max_groups > kCqrMaxGroups) max_groups = kCqrMaxGroups;
  static const int first_rw = [] {
    const char* e = getenv("RNPX_CQR_RW");  // 128: fewer, larger workgroups
    return (e && atoi(e) == 128) ? 128 : 64;
  }();
  for (int r : {first_rw, 64, 128}) {
    const int G = (mp + r - 1) / r;
    if (G >= 2 && G <= max_groups) {
      *rw = r;
      return G;
    }
  }
  return 0;
}

hipError_t launch_cqr_panel(const CqrArgs& a, int G, int rw, hipStream_t st) {
  if (a.mp < 2 * NB || G < 2 || G > kCqrMaxGroups || (long)G * rw < a.mp || (long)(G - 1) * rw >= a.mp ||
      a.lda < a.mp || a.ldv < a.mp || a.ldt < NB)
    return hipErrorInvalidValue;
  if (rw == 64)
    hipLaunchKernelGGL((cqr_panel_kernel<64>), dim3(G), dim3(256), 0, st, a);
  else if (rw == 128)
    hipLaunchKernelGGL((cqr_panel_kernel<128>), dim3(G), dim3(256), 0, st, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace rnpx
