// rnpx_zcqr.hip -- complex<double> panels of 32 columns by
// CholeskyQR2 and Householder reconstruction: the panel step of QR::Factor
// (QR::Factor_unblocked, RNP/LA/QR.hpp:78-102, with Reflector::Generate's
// conventions, Reflector.hpp:117-158) in a handful of stream-ordered kernels
// instead of one grid-wide hand-off per column.
//
//   G1 = A^H A                      (complex MFMA GEMM, split over the rows)
//   R1 = chol(G1), Q1 = A R1^{-1}   (one-workgroup Cholesky; GEMM)
//   G2 = Q1^H Q1, R2 = chol(G2)     (GEMM; one workgroup, with the rest below)
//   Q_top - D = L1 U  (LU without pivoting of the top nb x nb block of
//                      Q = Q1 R2^{-1}, D = diag(+-1) chosen step by step)
//   V = [L1; Q1_bot R2^{-1} U^{-1}],  tau_j = -U_jj D_j,
//   T = -U D L1^{-H},  R = D R2 R1.
// Why this is RNP's factorization: A = Q_c R_c (R_c with a positive real
// diagonal) and RNP's A = Q_h R_h (R_h with the real betas on its diagonal)
// are both QR factorizations, so Q_h = Q_c D with D = diag(+-1); the first nb
// columns of Q_h = I - V T V^H satisfy Q_h - [I; 0] = V (-T V1^H), an LU
// factorization without pivoting, which is the one of (Q_c - [D; 0]) D.
// Reflector::Generate's beta = -sign(Re alpha) |(alpha, x)| makes
// Re tau_j >= 1, which fixes D_j = -sign(Re pivot_j) (+1 for a zero real
// part: RNP's beta > 0 when Re alpha = 0), the choice that also keeps the
// LU's pivots at least 1 in size.
//
// Rejection: a first-pass pivot below 2^-40 of its diagonal, a squared
// column norm outside [2^-600, 2^600], anything not finite, or a second-pass
// Gram further than 2^-27 from I sets the fail flag; then nothing is written
// to A by this path (the write-back kernel is gated on the flag) and the
// column-by-column panel kernel, gated on the same flag, factors the panel.
#include "rnpx_internal.hpp"

#include <algorithm>

namespace rnpx {

namespace zq {

// The small factorizations: one workgroup of LPC lanes per column; thread t
// holds rows (t mod LPC) + LPC r (r < NB / LPC) of column t / LPC of the
// 32 x 32 matrix being worked on in registers, so the lanes of a column sit
// in one wave and share values by shuffles; a matrix other columns read is
// staged in LDS.  A Cholesky or LU step costs one workgroup barrier, a
// triangular inverse or product none.
constexpr int NB = 32;
constexpr int LPC = 8;  // lanes per column (measured: 4 and 16 are slower)
constexpr int kSmall = NB * LPC;
constexpr int R4 = NB / LPC;        // rows per thread
using Mat = zd[NB][NB + 1];

__device__ inline bool zfinite(zd v) { return isfinite(v.x) && isfinite(v.y); }
__device__ inline double zabs2(zd v) { return v.x * v.x + v.y * v.y; }
__device__ inline zd zconj(zd v) { return mkz(v.x, -v.y); }

// v from lane src (0..LPC-1) of this lane's column group.
__device__ inline zd gbcast(zd v, int src) { return mkz(__shfl(v.x, src, LPC), __shfl(v.y, src, LPC)); }

// Column c of the column-major NB x NB matrix at P (leading dimension ld).
__device__ inline void load_col(const zd* __restrict__ P, long ld, int c, int sub, zd (&v)[R4]) {
#pragma unroll
  for (int r = 0; r < R4; ++r) v[r] = P[(size_t)(sub + LPC * r) + (size_t)c * ld];
}
__device__ inline void store_col(zd* __restrict__ P, long ld, int c, int sub, const zd (&v)[R4]) {
#pragma unroll
  for (int r = 0; r < R4; ++r) P[(size_t)(sub + LPC * r) + (size_t)c * ld] = v[r];
}
__device__ inline void to_lds(Mat& M, int c, int sub, const zd (&v)[R4]) {
#pragma unroll
  for (int r = 0; r < R4; ++r) M[sub + LPC * r][c] = v[r];
}

// In registers: column c of the Hermitian G (its upper part, rows <= c) ->
// column c of its upper Cholesky factor R (R^H R = G), zeros below.
// Right-looking; row j of the updated G goes through rowb[j & 1].  Returns
// false (uniformly) on a pivot below rel * its original diagonal or one not
// finite.
__device__ __forceinline__ bool chol_cols(zd (&g)[R4], int c, int sub, double rel, zd (*rowb)[NB], double* dinit) {
#pragma unroll
  for (int r = 0; r < R4; ++r)
    if (sub + LPC * r == c) dinit[c] = g[r].x;
  if (sub == 0) rowb[0][c] = g[0];
  __syncthreads();
#pragma unroll 1
  for (int j = 0; j < NB; ++j) {
    const zd* rb = rowb[j & 1];
    const double d = rb[j].x;
    if (!(d > rel * dinit[j]) || !isfinite(d)) return false;  // every thread reads the same value
    const double rr = sqrt(d), inv = 1.0 / rr;
    const zd rjc = rb[c] * inv;  // R(j, c), c > j
#pragma unroll
    for (int r = 0; r < R4; ++r) {
      const int i = sub + LPC * r;
      if (i == j)
        g[r] = (c > j) ? rjc : (c == j ? mkz(rr, 0.0) : S<zd>::zero());
      else if (i > j && i <= c)
        g[r] = g[r] - zconj(rb[i] * inv) * rjc;
    }
    if (j + 1 < NB) {
#pragma unroll
      for (int r = 0; r < R4; ++r)
        if (sub + LPC * r == j + 1 && c >= j + 1) rowb[(j + 1) & 1][c] = g[r];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R4; ++r)
    if (sub + LPC * r > c) g[r] = S<zd>::zero();
  return true;
}

// Column c of X = U^{-1} for the upper triangular U in LDS (entries below
// its diagonal never read), by the column sweep y = e_c, x_k = y_k / U_kk,
// y_i -= U_ik x_k (i < k), k = NB-1 .. 0.  unit: U_kk = 1.  conjT: U_ik is
// conj(L(k, i)) of the lower triangular L stored there instead.
// The reciprocals of U's diagonal are formed first, in parallel, into rdiag
// (LDS), off the sweep's serial chain.  Whole workgroup.
template <bool UNIT, bool CONJT>
__device__ __forceinline__ void inv_upper_col(const Mat& U, int c, int sub, zd (&x)[R4], zd* rdiag) {
  if constexpr (!UNIT) {
    __syncthreads();  // earlier readers of rdiag are done
    if (threadIdx.x < NB) rdiag[threadIdx.x] = zrecip(U[threadIdx.x][threadIdx.x]);
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R4; ++r) x[r] = (sub + LPC * r == c) ? S<zd>::one() : S<zd>::zero();
  // k = LPC rk + s: rk unrolled (a fixed register), s not
#pragma unroll
  for (int rk = R4 - 1; rk >= 0; --rk) {
#pragma unroll 1
    for (int s8 = LPC - 1; s8 >= 0; --s8) {
      const int k = LPC * rk + s8;
      zd xk = x[rk];
      if constexpr (!UNIT) xk = xk * rdiag[k];
      xk = gbcast(xk, s8);
#pragma unroll
      for (int r = 0; r < R4; ++r) {
        const int i = sub + LPC * r;
        if (i == k)
          x[r] = xk;
        else if (i < k)
          x[r] = x[r] - (CONJT ? zconj(U[k][i]) : U[i][k]) * xk;
      }
    }
  }
}

// Column c of A B for A in LDS (row scale ds[k] on column k of A when
// given) and column c of B in registers.
__device__ __forceinline__ void mul_col(const Mat& A, const zd (&b)[R4], int sub, zd (&out)[R4], const double* ds = nullptr) {
#pragma unroll
  for (int r = 0; r < R4; ++r) out[r] = S<zd>::zero();
#pragma unroll
  for (int rk = 0; rk < R4; ++rk) {
#pragma unroll 2
    for (int s8 = 0; s8 < LPC; ++s8) {
      const int k = LPC * rk + s8;
      const zd bk = gbcast(b[rk], s8);
#pragma unroll
      for (int r = 0; r < R4; ++r) {
        const zd aik = ds ? A[sub + LPC * r][k] * ds[k] : A[sub + LPC * r][k];
        out[r] = out[r] + aik * bk;
      }
    }
  }
}

// Pass 1: R1 = chol(G1) and R1^{-1}, or the fail flag.
__global__ __launch_bounds__(kSmall) void chol_kernel(const zd* __restrict__ G, zd* __restrict__ R1,
                                                      zd* __restrict__ R1i, unsigned* fail) {
  __shared__ Mat Rs;
  __shared__ zd rowb[2][NB];
  __shared__ double dinit[NB];
  const int t = threadIdx.x, c = t / LPC, sub = t % LPC;
  zd g[R4];
  load_col(G, NB, c, sub, g);
  int bad = 0;
#pragma unroll
  for (int r = 0; r < R4; ++r) {
    const int i = sub + LPC * r;
    if (!zfinite(g[r])) bad = 1;
    if (i == c && !(g[r].x >= 0x1p-600 && g[r].x <= 0x1p600)) bad = 1;
  }
  bool ok = __syncthreads_or(bad) == 0;
  if (ok) ok = chol_cols(g, c, sub, 0x1p-40, rowb, dinit);
  if (!ok) {
    if (t == 0) *fail = 1u;
    return;
  }
  store_col(R1, NB, c, sub, g);
  to_lds(Rs, c, sub, g);
  __syncthreads();
  zd x[R4];
  inv_upper_col<false, false>(Rs, c, sub, x, rowb[0]);
  store_col(R1i, NB, c, sub, x);
}

// Pass 2 and the reconstruction.  In: G2 = Q1^H Q1, R1, the top NB rows of
// Q1.  Out: tau, T (ldt), the clean V's top NB rows (ldv), Atop = R (upper)
// and L1 (strictly lower) for the write-back, M = R2^{-1} U^{-1} for the rest of
// V; or the fail flag.
__global__ __launch_bounds__(kSmall) void recon_kernel(const zd* __restrict__ G2, const zd* __restrict__ R1,
                                                       const zd* __restrict__ Q1, long ldq, zd* __restrict__ tau,
                                                       zd* __restrict__ Tm, long ldt, zd* __restrict__ V, long ldv,
                                                       zd* __restrict__ Atop, zd* __restrict__ M, unsigned* fail) {
  __shared__ Mat S1;  // R2
  __shared__ Mat S2;  // R2^{-1}
  __shared__ Mat S3;  // Q1's top rows
  __shared__ Mat S4;  // L1 \ U
  __shared__ zd rowb[2][NB];
  __shared__ double dinit[NB];
  __shared__ double Ds[NB];
  const int t = threadIdx.x, c = t / LPC, sub = t % LPC;
  if (*fail) return;  // pass 1 rejected the panel (uniform)
  zd g[R4];
  load_col(G2, NB, c, sub, g);
  int bad = 0;
#pragma unroll
  for (int r = 0; r < R4; ++r) {
    const int i = sub + LPC * r;
    const zd eij = (i == c) ? mkz(g[r].x - 1.0, g[r].y) : g[r];
    if (!zfinite(g[r]) || zabs2(eij) > 0x1p-54) bad = 1;  // |E_ij| > 2^-27
  }
  // Q1's top rows to LDS (row-major) while the Gram is checked
#pragma unroll
  for (int r = 0; r < R4; ++r) S3[sub + LPC * r][c] = Q1[(size_t)(sub + LPC * r) + (size_t)c * ldq];
  int big = 0;  // |E_ij| > 2^-40: R2 by Cholesky
#pragma unroll
  for (int r = 0; r < R4; ++r) {
    const int i = sub + LPC * r;
    const zd eij = (i == c) ? mkz(g[r].x - 1.0, g[r].y) : g[r];
    if (zabs2(eij) > 0x1p-80) big = 1;
  }
  bool ok = __syncthreads_or(bad) == 0;
  const bool first_order = __syncthreads_or(big) == 0;
  zd x2[R4];
  if (first_order) {
    // R2 = I + E_U, E_U = triu(E) - diag(E)/2, and R2^{-1} = I - E_U, both
    // exact to O(|E|^2) <= 2^-80
#pragma unroll
    for (int r = 0; r < R4; ++r) {
      const int i = sub + LPC * r;
      if (i < c) {
        x2[r] = S<zd>::zero() - g[r];
      } else if (i == c) {
        const double e2 = 0.5 * (g[r].x - 1.0);
        g[r] = mkz(1.0 + e2, 0.0);
        x2[r] = mkz(1.0 - e2, 0.0);
      } else {
        g[r] = S<zd>::zero();
        x2[r] = S<zd>::zero();
      }
    }
    to_lds(S1, c, sub, g);
  } else {
    if (ok) ok = chol_cols(g, c, sub, 0x1p-2, rowb, dinit);
    if (!ok) {
      if (t == 0) *fail = 1u;
      return;
    }
    to_lds(S1, c, sub, g);  // R2
    inv_upper_col<false, false>(S1, c, sub, x2, rowb[0]);  // column c of R2^{-1}
  }
  if (!ok) {
    if (t == 0) *fail = 1u;
    return;
  }
  to_lds(S2, c, sub, x2);
  __syncthreads();
  // q = column c of Q_top = Q1_top R2^{-1}
  zd q[R4];
  mul_col(S3, x2, sub, q);
  // LU without pivoting of Q_top - D: L1 strictly below, U on and above;
  // D_j = -sign(Re pivot) (+1 for a zero real part)
  zd (*Lc)[NB] = rowb;
#pragma unroll
  for (int rj = 0; rj < R4; ++rj)
#pragma unroll 1
  for (int s8 = 0; s8 < LPC; ++s8) {
    const int j = LPC * rj + s8;
    if (c == j) {
      const zd piv = gbcast(q[rj], s8);
      const double dj = piv.x > 0.0 ? -1.0 : 1.0;
      const zd pivp = mkz(piv.x - dj, piv.y);
      const zd rp = zrecip(pivp);
#pragma unroll
      for (int r = 0; r < R4; ++r) {
        const int i = sub + LPC * r;
        if (i == j) {
          q[r] = pivp;
        } else if (i > j) {
          q[r] = q[r] * rp;
          Lc[j & 1][i] = q[r];
        }
      }
      if (sub == 0) Ds[j] = dj;
    }
    __syncthreads();
    if (c > j) {
      const zd u = gbcast(q[rj], s8);  // U(j, c)
#pragma unroll
      for (int r = 0; r < R4; ++r) {
        const int i = sub + LPC * r;
        if (i > j) q[r] = q[r] - Lc[j & 1][i] * u;
      }
    }
  }
  to_lds(S4, c, sub, q);
  // R = D R2 R1 (upper) and L1 (strictly lower) -> Atop; V's top rows = unit L1
  {
    zd r1[R4], rr[R4];
    load_col(R1, NB, c, sub, r1);
    mul_col(S1, r1, sub, rr);
#pragma unroll
    for (int r = 0; r < R4; ++r) {
      const int i = sub + LPC * r;
      Atop[(size_t)i + (size_t)c * NB] = (i <= c) ? rr[r] * Ds[i] : q[r];
      V[(size_t)i + (size_t)c * ldv] = (i > c) ? q[r] : (i == c ? S<zd>::one() : S<zd>::zero());
    }
  }
  __syncthreads();  // S4, Ds complete
  if (t < NB) tau[t] = S<zd>::zero() - S4[t][t] * Ds[t];
  // M = R2^{-1} U^{-1}
  {
    zd ui[R4], mm[R4];
    inv_upper_col<false, false>(S4, c, sub, ui, rowb[0]);
    mul_col(S2, ui, sub, mm);
#pragma unroll
    for (int r = 0; r < R4; ++r)
      if (sub + LPC * r > c) mm[r] = S<zd>::zero();
    store_col(M, NB, c, sub, mm);
  }
  // T = -(U D) L1^{-H}: Y = L1^{-H}, the inverse of the unit upper L1^H
  {
    zd y[R4], tt[R4];
    inv_upper_col<true, true>(S4, c, sub, y, rowb[0]);
    // (U D) with U the upper part of S4: entries below the diagonal masked
    zd u[R4];
#pragma unroll
    for (int r = 0; r < R4; ++r) tt[r] = S<zd>::zero();
#pragma unroll
    for (int rl = 0; rl < R4; ++rl)
#pragma unroll 2
    for (int s8 = 0; s8 < LPC; ++s8) {
      const int l = LPC * rl + s8;
      const zd yl = gbcast(y[rl], s8);
#pragma unroll
      for (int r = 0; r < R4; ++r) {
        const int i = sub + LPC * r;
        u[r] = (l >= i) ? S4[i][l] * Ds[l] : S<zd>::zero();
        tt[r] = tt[r] + u[r] * yl;
      }
    }
#pragma unroll
    for (int r = 0; r < R4; ++r) tt[r] = S<zd>::zero() - tt[r];
    store_col(Tm, ldt, c, sub, tt);
  }
}

// Q1 = A R1^{-1} and the second pass's Gram partials in one pass: a
// workgroup takes 64 rows of the panel (A's rows and R1^{-1} staged in LDS),
// thread t forms columns 8 (t / 64) .. + 7 of row t mod 64 of Q1, then the
// workgroup's partial Q1_rows^H Q1_rows (four entries a thread) goes to
// part + blockIdx 1024.
__global__ __launch_bounds__(256) void q1gram_kernel(int mp, const zd* __restrict__ A, long lda,
                                                     const zd* __restrict__ R1i, zd* __restrict__ Q1,
                                                     zd* __restrict__ part, const unsigned* fail) {
  __shared__ zd Rs[NB][NB + 1];
  __shared__ zd Xs[64][NB + 1];
  __shared__ zd Ys[64][NB + 1];
  if (*fail) return;  // pass 1 rejected the panel (uniform)
  const int t = threadIdx.x, r0 = blockIdx.x * 64;
  for (int e = t; e < NB * NB; e += 256) Rs[e % NB][e / NB] = R1i[e];
  for (int e = t; e < 64 * NB; e += 256) {
    const int rr = e % 64, k = e / 64, r = min(r0 + rr, mp - 1);
    Xs[rr][k] = A[(size_t)r + (size_t)k * lda];
  }
  __syncthreads();
  const int rl = t & 63, cg = t >> 6, r = r0 + rl;
  // the row in registers; R1^{-1}(k, c) is the same address across the wave
  zd x[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) x[k] = Xs[rl][k];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int c = 8 * cg + q;
    zd y = S<zd>::zero();
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (k <= c) y = y + x[k] * Rs[k][c];  // R1^{-1} upper triangular
    if (r < mp) Q1[(size_t)r + (size_t)c * mp] = y;
    Ys[rl][c] = (r < mp) ? y : S<zd>::zero();
  }
  __syncthreads();
  // partial Gram: thread t a 2 x 2 block, rows i0, i0+1 and columns j0, j0+1
  const int i0 = 2 * (t & 15), j0 = 2 * (t >> 4);
  zd g00 = S<zd>::zero(), g10 = S<zd>::zero(), g01 = S<zd>::zero(), g11 = S<zd>::zero();
#pragma unroll 4
  for (int rr = 0; rr < 64; ++rr) {
    const zd a0 = zconj(Ys[rr][i0]), a1 = zconj(Ys[rr][i0 + 1]), b0 = Ys[rr][j0], b1 = Ys[rr][j0 + 1];
    g00 = g00 + a0 * b0;
    g10 = g10 + a1 * b0;
    g01 = g01 + a0 * b1;
    g11 = g11 + a1 * b1;
  }
  zd* pp = part + (size_t)blockIdx.x * NB * NB;
  pp[i0 + j0 * NB] = g00;
  pp[i0 + 1 + j0 * NB] = g10;
  pp[i0 + (j0 + 1) * NB] = g01;
  pp[i0 + 1 + (j0 + 1) * NB] = g11;
}

// G = the sum of ns partial 32 x 32 Grams, in order: 16 workgroups of 64
// threads, one entry a thread.
__global__ __launch_bounds__(64) void gram_sum_kernel(int ns, const zd* __restrict__ part, zd* __restrict__ G,
                                                      const unsigned* fail) {
  if (*fail) return;
  const int e = blockIdx.x * 64 + threadIdx.x;  // < NB * NB
  G[e] = sum_strided(part + e, ns, (size_t)NB * NB);
}

// The rest of V and the write-back in one pass: rows [nb, mp) of V are
// Q1 M (M = R2^{-1} U^{-1}); unless the panel was rejected, A's rows [0, nb)
// become Atop (R and L1) and its rows below become V.  A workgroup takes 64
// rows: their Q1 rows and M staged in LDS, thread t forms columns
// 8 (t / 64) .. + 7 of row t mod 64.
__global__ __launch_bounds__(256) void vcommit_kernel(int mp, const zd* __restrict__ Q1, const zd* __restrict__ Mm,
                                                      const zd* __restrict__ Atop, zd* __restrict__ V, long ldv,
                                                      zd* __restrict__ A, long lda, const unsigned* fail) {
  __shared__ zd Ms[NB][NB + 1];
  __shared__ zd Qs[64][NB + 1];
  const int t = threadIdx.x, r0 = blockIdx.x * 64;
  const bool ok = (*fail == 0u);  // uniform
  for (int e = t; e < NB * NB; e += 256) Ms[e % NB][e / NB] = Mm[e];
  for (int e = t; e < 64 * NB; e += 256) {
    const int rr = e % 64, k = e / 64, r = min(r0 + rr, mp - 1);
    Qs[rr][k] = Q1[(size_t)r + (size_t)k * mp];
  }
  __syncthreads();
  const int rl = t & 63, cg = t >> 6, r = r0 + rl;
  if (r >= mp) return;
  if (r < NB) {
    if (ok) {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        A[(size_t)r + (size_t)(8 * cg + q) * lda] = Atop[(size_t)r + (size_t)(8 * cg + q) * NB];
    }
    return;
  }
  // the row in registers; M(k, c) is the same address across the wave
  zd x[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) x[k] = Qs[rl][k];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int c = 8 * cg + q;
    zd v = S<zd>::zero();
#pragma unroll
    for (int k = 0; k < NB; ++k) v = v + x[k] * Ms[k][c];
    V[(size_t)r + (size_t)c * ldv] = v;
    if (ok) A[(size_t)r + (size_t)c * lda] = v;
  }
}

}  // namespace zq

// The complex panel rows [0, mp) x columns [0, nb) of A (nb = 32,
// mp >= 2 nb): R and V packed into A, tau, the clean V (ldv) and T (ldt), on
// stream st; fail (zero on entry) is set when the panel is rejected, and
// then the column-wise panel (pa, gated on fail) factors it instead.
int zcqr_panel(Handle* h, int mp, int nb, zd* A, long lda, zd* tau, zd* V, long ldv, zd* Tm, long ldt, unsigned* fail,
               const PanelArgs<zd>& pa, int rpt, hipStream_t st) {
  // 32 columns: the reconstruction keeps four 32 x 32 complex matrices in LDS
  if (nb != 32 || mp < 2 * nb) return RNPX_ERR_UNSUPPORTED;
  const size_t nn = (size_t)nb * nb;
  const int ks = std::max(1, std::min(64, mp / 64));  // 64-row slices: short chains per workgroup
  const size_t nrb_max = (size_t)(mp + 63) / 64;
  RNPX_TRY(h->zcqr.ensure(sizeof(zd) * ((size_t)mp * nb + (6 + (size_t)ks + nrb_max) * nn), h));
  zd* Q1 = (zd*)h->zcqr.p;
  zd* G = Q1 + (size_t)mp * nb;
  zd* R1 = G + nn;
  zd* R1i = R1 + nn;
  zd* Atop = R1i + nn;
  zd* M = Atop + nn;
  zd* part = M + nn;              // split-K partials of the first Gram
  zd* part2 = part + (size_t)ks * nn;  // per-64-row partials of the second
  const zd one = S<zd>::one(), zero = S<zd>::zero();
  // pass 1
  RNPX_TRY(gemm<zd>('C', 'N', nb, nb, mp, one, A, lda, A, lda, zero, G, nb, ks, part, st));
  hipLaunchKernelGGL(zq::chol_kernel, dim3(1), dim3(zq::kSmall), 0, st, G, R1, R1i, fail);
  RNPX_TRY(hipGetLastError());
  // Q1 = A R1^{-1} and the pass-2 Gram G2 = Q1^H Q1 (partials per 64 rows)
  const int nrb = (mp + 63) / 64;
  hipLaunchKernelGGL(zq::q1gram_kernel, dim3((unsigned)nrb), dim3(256), 0, st, mp, A, lda, R1i, Q1, part2, fail);
  RNPX_TRY(hipGetLastError());
  hipLaunchKernelGGL(zq::gram_sum_kernel, dim3(zq::NB * zq::NB / 64), dim3(64), 0, st, nrb, part2, G, fail);
  RNPX_TRY(hipGetLastError());
  // pass 2 and the reconstruction
  hipLaunchKernelGGL(zq::recon_kernel, dim3(1), dim3(zq::kSmall), 0, st, G, R1, Q1, (long)mp, tau, Tm, ldt,
                     V, ldv, Atop, M, fail);
  RNPX_TRY(hipGetLastError());
  // the rest of V (Q1_bot R2^{-1} U^{-1}) and, unless rejected, the panel
  // written back
  hipLaunchKernelGGL(zq::vcommit_kernel, dim3((unsigned)((mp + 63) / 64)), dim3(256), 0, st, mp, Q1, M, Atop, V, ldv,
                     A, lda, fail);
  RNPX_TRY(hipGetLastError());
  // a rejected panel: column by column (both launches return at once otherwise)
  PanelArgs<zd> fb = pa;
  fb.run_if = fail;
  RNPX_TRY(launch_panel<zd>(fb, nb, rpt, st));
  RNPX_TRY(launch_tfactor<zd>(fb.gram, nb, std::min(mp, pa.nbp), nb, Tm, (int)ldt, st, fail));
  return RNPX_OK;
}

}  // namespace rnpx
