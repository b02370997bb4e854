// rnpx_small64.hpp -- blocked 64x64 factorizations for one workgroup of 256
// threads (4 waves), used by the CholeskyQR2 panel (rnpx_cqr.hip).
//
// The matrices live in LDS, row-major with row stride S (padded).  They are
// processed as 4x4 blocks of 16x16.  Step b of a factorization is a 16-step
// dependency chain on one wave over block row b (one column per lane, with
// v_readlane broadcasts of the pivot row), which yields the diagonal block
// and the blocks right of it at once; the trailing updates and the inverses'
// off-diagonal blocks are 16x16x16 block products on FP64 MFMA spread over
// the four waves, and the diagonal blocks' inverses run on an idle wave
// beside the next chain.
#pragma once
#include "rnpx_kernels.hpp"

namespace rnpx {
namespace small64 {

constexpr int TS = 17;  // row stride of a 16x16 scratch block
// tmp arguments below hold 4 scratch blocks (16 x TS each) and 64 more doubles.
constexpr int kTmpDoubles = 4 * 16 * TS + 64;

// 1/x and 1/sqrt(x) from the hardware approximations refined by two Newton
// steps (a few ulp; x normal).  They keep the division and square root off
// the long dependency chains of the diagonal-block factorizations.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// acc += op(A)(p, q) * op(B)(p2, q2) for 16x16 blocks (block indices in op
// coordinates); op(X) = X^T when the flag is set.  With neg the A operand is
// negated.  acc item r holds C[lane/16 + 4r][lane%16] of the 16x16 result.
template <int SA, int SB>
__device__ __forceinline__ void mma16(f64x4& acc, const double* A, int p, int q, bool tA, const double* B, int p2,
                                      int q2, bool tB, bool neg, int lane) {
  const int i = lane & 15, kk = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * s + kk;
    double a = tA ? A[(16 * q + k) * SA + 16 * p + i] : A[(16 * p + i) * SA + 16 * q + k];
    const double b = tB ? B[(16 * q2 + i) * SB + 16 * p2 + k] : B[(16 * p2 + k) * SB + 16 * q2 + i];
    if (neg) a = -a;
    acc = mfma16(a, b, acc);
  }
}

template <int S>
__device__ __forceinline__ void load_blk(f64x4& acc, const double* C, int br, int bc, int lane) {
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = C[(16 * br + (lane >> 4) + 4 * r) * S + 16 * bc + (lane & 15)];
}
template <int S>
__device__ __forceinline__ void store_blk(const f64x4& acc, double* C, int br, int bc, double scale, int lane) {
#pragma unroll
  for (int r = 0; r < 4; ++r) C[(16 * br + (lane >> 4) + 4 * r) * S + 16 * bc + (lane & 15)] = scale * acc[r];
}

// One wave: X(bx) = inverse of the upper triangle of block (b, b) of M
// (zeros below the diagonal); X is written at block (b, b) of a matrix of
// stride SX.
template <int S, int SX>
__device__ __forceinline__ void inv_upper16(const double* M, int b, double* X, int xb, double* rinv, int lane) {
  // Row l of X per lane; U is constant here, so its rows are read as LDS
  // broadcasts (no v_readlane), with the diagonal reciprocals precomputed.
  const int l = lane & 15;
  if (lane < 16) rinv[l] = rcp_nr(M[(16 * b + l) * S + 16 * b + l]);
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = (j == l) ? 1.0 : 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double xj = acc[j] * rinv[j];
    acc[j] = xj;
#pragma unroll
    for (int jp = j + 1; jp < 16; ++jp) acc[jp] = fma(-M[(16 * b + j) * S + 16 * b + jp], xj, acc[jp]);
  }
  if (lane < 16) {
#pragma unroll
    for (int j = 0; j < 16; ++j) X[(16 * xb + l) * SX + 16 * xb + j] = acc[j];
  }
}

// One wave: Z = inverse of the unit lower triangle below the diagonal of
// block (b, b) of M, written at block (zb, zb) of Z (stride SZ).
template <int S, int SZ>
__device__ __forceinline__ void inv_unit_lower16(const double* M, int b, double* Z, int zb, int lane) {
  // Row l of Z per lane (z_l L = e_l, solved from the last column back);
  // L is constant here, so its rows are read as LDS broadcasts.
  const int l = lane & 15;
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = (j == l) ? 1.0 : 0.0;
#pragma unroll
  for (int j = 15; j >= 0; --j) {
    const double zj = acc[j];
#pragma unroll
    for (int jp = 0; jp < j; ++jp) acc[jp] = fma(-M[(16 * b + j) * S + 16 * b + jp], zj, acc[jp]);
  }
  if (lane < 16) {
#pragma unroll
    for (int j = 0; j < 16; ++j) Z[(16 * zb + l) * SZ + 16 * zb + j] = acc[j];
  }
}

// All waves: off-diagonal blocks of X = U^-1 (block upper triangular) given
// its diagonal blocks and U's blocks above the diagonal in M:
// X_ij = -X_ii sum_{k=i+1..j} U_ik X_kj, by distance j - i = 1, 2, 3.
// tmp: 4 scratch blocks (one per wave).
template <int S>
__device__ __forceinline__ void inv_upper_offdiag(const double* M, double* X, double* tmp, int lane, int wave) {
  double* t = tmp + wave * 16 * TS;
#pragma unroll
  for (int d = 1; d < 4; ++d) {
    if (wave < 4 - d) {
      const int i = wave, j = wave + d;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int k = i + 1; k <= j; ++k) mma16<S, S>(acc, M, i, k, false, X, k, j, false, false, lane);
      store_blk<TS>(acc, t, 0, 0, 1.0, lane);
      f64x4 x = {0.0, 0.0, 0.0, 0.0};
      mma16<S, TS>(x, X, i, i, false, t, 0, 0, false, true, lane);
      store_blk<S>(x, X, i, j, 1.0, lane);
    }
    __syncthreads();
  }
}

// All waves: off-diagonal blocks of Z = L^-1 (block lower triangular) given
// its diagonal blocks and L's blocks below the diagonal in M:
// Z_ij = -Z_ii sum_{k=j..i-1} L_ik Z_kj, by distance i - j = 1, 2, 3.
template <int S>
__device__ __forceinline__ void inv_lower_offdiag(const double* M, double* Z, double* tmp, int lane, int wave) {
  double* t = tmp + wave * 16 * TS;
#pragma unroll
  for (int d = 1; d < 4; ++d) {
    if (wave < 4 - d) {
      const int j = wave, i = wave + d;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int k = j; k < i; ++k) mma16<S, S>(acc, M, i, k, false, Z, k, j, false, false, lane);
      store_blk<TS>(acc, t, 0, 0, 1.0, lane);
      f64x4 z = {0.0, 0.0, 0.0, 0.0};
      mma16<S, TS>(z, Z, i, i, false, t, 0, 0, false, true, lane);
      store_blk<S>(z, Z, i, j, 1.0, lane);
    }
    __syncthreads();
  }
}

// One wave: rows 16b .. 16b+15 of the upper Cholesky factor from the
// partially reduced M, one column per lane across the whole block row, so
// that the diagonal block (zeros below its diagonal) and R_bc for every
// c > b come out of the same 16-step chain.  Lanes left of the block row
// compute on what lies there and store nothing.  Acceptance: every pivot
// d_j > gdiag_j 2^-40 and gdiag_j in [2^-600, 2^600] (gdiag = the original
// diagonal).
template <int S>
__device__ __forceinline__ bool chol_strip(double* M, int b, const double* gdiag, int lane) {
  const int c = lane, c0 = 16 * b;
  double w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = M[(c0 + i) * S + c];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const double d = readlane_t(w[j], c0 + j);
    const double g = gdiag[c0 + j];
    ok = ok && (g >= 0x1p-600) && (g <= 0x1p600) && (d > g * 0x1p-40);
    w[j] *= rsq_nr(d);
#pragma unroll
    for (int i = j + 1; i < 16; ++i) w[i] = fma(-readlane_t(w[j], c0 + i), w[j], w[i]);
  }
  if (c >= c0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) M[(c0 + i) * S + c] = (i <= c - c0) ? w[i] : 0.0;
  }
  return ok;
}

// One wave: the 16-step LU chain of M - D, D_j = -1 when the running pivot
// is positive and +1 otherwise (into dsg), at diagonal block b over its
// block row (ROW: one column per lane; U_bb, L_bb and U_bc for c > b) or its
// block column (one row per lane; L_cb for c > b).  Both chains apply the
// same operations to the diagonal block in the same order (l = w inv, then
// fma(-l, u, w)), so they agree on it bit for bit, pivot signs included;
// the row chain stores it.  w holds what lu_strip_load read; the two loads
// must both precede either chain's stores.
template <int S, bool ROW>
__device__ __forceinline__ void lu_strip_load(const double* M, int b, double (&w)[16], int lane) {
  const int x = lane, x0 = 16 * b;
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = ROW ? M[(x0 + i) * S + x] : M[x * S + x0 + i];
}
template <int S, bool ROW>
__device__ __forceinline__ void lu_strip(double* M, int b, double (&w)[16], double* dsg, int lane) {
  const int x = lane, x0 = 16 * b;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int pj = x0 + j;
    const double p = readlane_t(w[j], pj);
    const double dj = (p > 0.0) ? -1.0 : 1.0;
    const double u = p - dj;  // |u| = |p| + 1 >= 1
    const double inv = rcp_nr(u);
    if (ROW) {
      if (x == pj) w[j] = u;
      if (lane == 0) dsg[pj] = dj;
#pragma unroll
      for (int i = j + 1; i < 16; ++i) {
        const double lij = readlane_t(w[i], pj) * inv;
        w[i] = (x > pj) ? fma(-lij, w[j], w[i]) : ((x == pj) ? lij : w[i]);
      }
    } else {
      const double l = w[j] * inv;
#pragma unroll
      for (int jp = j + 1; jp < 16; ++jp) {
        const double ujp = readlane_t(w[jp], pj);  // row pj is not updated at step j
        w[jp] = (x > pj) ? fma(-l, ujp, w[jp]) : w[jp];
      }
      w[j] = (x > pj) ? l : ((x == pj) ? u : w[j]);
    }
  }
  if (ROW) {
    if (x >= x0) {
#pragma unroll
      for (int i = 0; i < 16; ++i) M[(x0 + i) * S + x] = w[i];
    }
  } else if (x >= x0 + 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) M[x * S + x0 + i] = w[i];
  }
}

// All waves: Cholesky M = R^T R of the symmetric 64x64 M in place (R upper;
// blocks below the diagonal left as they are) and X = R^-1 (upper blocks).
// Wave 0 runs the block-row chain b while wave 1 inverts diagonal block
// b-1; all waves share the trailing updates.  Returns the acceptance test
// (the same in every thread).
template <int S>
__device__ __forceinline__ bool chol64(double* M, double* X, double* tmp, double* gdiag, int* okf, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
  double* rv = tmp + 4 * 16 * TS;  // 4 x 16 reciprocals
  if (tid < 64) gdiag[tid] = M[tid * S + tid];
  if (tid == 0) *okf = 1;
  __syncthreads();
  bool ok = true;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (wave == 0) ok = chol_strip<S>(M, b, gdiag, lane) && ok;
    if (wave == 1 && b > 0) inv_upper16<S, S>(M, b - 1, X, b - 1, rv + 16 * (b - 1), lane);
    __syncthreads();
    if (b < 3) {
      // G_cd -= R_bc^T R_bd for b < c <= d
      const int n = 3 - b, npairs = n * (n + 1) / 2;
      for (int e = wave; e < npairs; e += 4) {
        int c = 0, rem = e;
        while (rem >= n - c) {
          rem -= n - c;
          ++c;
        }
        const int cc = b + 1 + c, dd = cc + rem;
        f64x4 acc;
        load_blk<S>(acc, M, cc, dd, lane);
        mma16<S, S>(acc, M, cc, b, true, M, b, dd, false, true, lane);
        store_blk<S>(acc, M, cc, dd, 1.0, lane);
      }
      __syncthreads();
    }
  }
  if (wave == 1) inv_upper16<S, S>(M, 3, X, 3, rv + 48, lane);
  if (wave == 0 && lane == 0 && !ok) *okf = 0;
  __syncthreads();
  inv_upper_offdiag<S>(M, X, tmp, lane, wave);  // ends with a barrier
  return *okf != 0;
}

// All waves: M (64x64, holding Q_top) := the LU factors of M - D (U on and
// above the diagonal, unit lower L below), D into dsg, X = U^-1 (upper
// blocks).  The block-row and block-column chains of step b run side by
// side on waves 0 and 1 while wave 2 inverts U's diagonal block b-1.
template <int S>
__device__ __forceinline__ void lu64(double* M, double* dsg, double* X, double* tmp, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
  double* rv = tmp + 4 * 16 * TS;  // 4 x 16 reciprocals
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    double w[16];
    // the column chain reads the diagonal block the row chain overwrites
    if (wave == 0) lu_strip_load<S, true>(M, b, w, lane);
    if (wave == 1 && b < 3) lu_strip_load<S, false>(M, b, w, lane);
    if (b < 3) __syncthreads();
    if (wave == 0) lu_strip<S, true>(M, b, w, dsg, lane);
    if (wave == 1 && b < 3) lu_strip<S, false>(M, b, w, dsg, lane);
    if (wave == 2 && b > 0) inv_upper16<S, S>(M, b - 1, X, b - 1, rv + 16 * (b - 1), lane);
    __syncthreads();
    if (b < 3) {
      // W_cd -= L_cb U_bd for c, d > b
      const int n = 3 - b;
      for (int e = wave; e < n * n; e += 4) {
        const int c = b + 1 + e / n, d = b + 1 + e % n;
        f64x4 acc;
        load_blk<S>(acc, M, c, d, lane);
        mma16<S, S>(acc, M, c, b, false, M, b, d, false, true, lane);
        store_blk<S>(acc, M, c, d, 1.0, lane);
      }
      __syncthreads();
    }
  }
  if (wave == 2) inv_upper16<S, S>(M, 3, X, 3, rv + 48, lane);
  __syncthreads();
  inv_upper_offdiag<S>(M, X, tmp, lane, wave);
}

// All waves: Z = L^-1 for the unit lower triangle below the diagonal of M.
template <int S>
__device__ __forceinline__ void inv_unit_lower64(const double* M, double* Z, double* tmp, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
  inv_unit_lower16<S, S>(M, wave, Z, wave, lane);  // each wave one diagonal block
  __syncthreads();
  inv_lower_offdiag<S>(M, Z, tmp, lane, wave);
}

}  // namespace small64
}  // namespace rnpx
