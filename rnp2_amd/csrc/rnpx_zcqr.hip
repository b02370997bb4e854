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
// to A by this path (the commit kernel is gated on the flag) and the
// column-by-column panel kernel, gated on the same flag, factors the panel.
#include "rnpx_internal.hpp"

#include <algorithm>

namespace rnpx {

namespace zq {

constexpr int kThreads = 256;  // the commit kernel
// The small factorizations: one 256-thread workgroup, the matrices in LDS
// (measured: one wave is slower -- the work, not the barriers, is the cost).
constexpr int kSmall = 256;

__device__ inline bool zfinite(zd v) { return isfinite(v.x) && isfinite(v.y); }
__device__ inline double zabs2(zd v) { return v.x * v.x + v.y * v.y; }
__device__ inline zd zconj(zd v) { return mkz(v.x, -v.y); }
__device__ inline zd zdiv(zd a, zd b) {  // a / b, Smith's scaling
  return a * zrecip(b);
}

// In place: G (Hermitian, NB x NB in LDS) -> its upper Cholesky factor R
// (R^H R = G) in the upper triangle, the lower triangle cleared.  Returns
// false (uniformly) on a pivot below rel * its original diagonal or one not
// finite.  Whole workgroup.
template <int NB>
__device__ bool chol_upper(zd (*G)[NB + 1], double rel, double* dsh) {
  const int t = threadIdx.x;
  if (t < NB) dsh[t] = G[t][t].x;
  __syncthreads();
  for (int j = 0; j < NB; ++j) {
    const double d = G[j][j].x;
    if (!(d > rel * dsh[j]) || !isfinite(d)) return false;  // uniform: every thread reads the same value
    const double r = sqrt(d);
    __syncthreads();
    if (t == 0) G[j][j] = mkz(r, 0.0);
    for (int l = j + 1 + t; l < NB; l += kSmall) G[j][l] = G[j][l] * (1.0 / r);
    __syncthreads();
    // G(i, l) -= conj(R(j, i)) R(j, l) for j < i <= l
    const int rest = NB - j - 1;
    for (int e = t; e < rest * rest; e += kSmall) {
      const int i = j + 1 + e / rest, l = j + 1 + e % rest;  // consecutive threads along a row
      if (i <= l) G[i][l] = G[i][l] - zconj(G[j][i]) * G[j][l];
    }
    __syncthreads();
  }
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e / NB, l = e % NB;
    if (i > l) G[i][l] = S<zd>::zero();
  }
  __syncthreads();
  return true;
}

// X = U^{-1} for upper-triangular U (NB x NB in LDS); thread c back-
// substitutes column c.
template <int NB>
__device__ void inv_upper(const zd (*U)[NB + 1], zd (*X)[NB + 1]) {
  const int c = threadIdx.x;
  if (c < NB) {
    for (int i = c + 1; i < NB; ++i) X[i][c] = S<zd>::zero();
    X[c][c] = zrecip(U[c][c]);
    for (int i = c - 1; i >= 0; --i) {
      zd s = S<zd>::zero();
      for (int k = i + 1; k <= c; ++k) s = s + U[i][k] * X[k][c];
      X[i][c] = S<zd>::zero() - zdiv(s, U[i][i]);
    }
  }
  __syncthreads();
}

// C = A B for NB x NB matrices in LDS (any structure; C distinct from A, B).
template <int NB>
__device__ void mul(const zd (*A)[NB + 1], const zd (*B)[NB + 1], zd (*C)[NB + 1]) {
  for (int e = threadIdx.x; e < NB * NB; e += kSmall) {
    const int i = e / NB, l = e % NB;  // consecutive threads along a row of C and B
    zd s = S<zd>::zero();
    for (int k = 0; k < NB; ++k) s = s + A[i][k] * B[k][l];
    C[i][l] = s;
  }
  __syncthreads();
}

// Pass 1: R1 = chol(G1) and R1^{-1}, or the fail flag.
template <int NB>
__global__ __launch_bounds__(kSmall) void chol_kernel(const zd* __restrict__ G, zd* __restrict__ R1,
                                                        zd* __restrict__ R1i, unsigned* fail) {
  __shared__ zd Gs[NB][NB + 1];
  __shared__ zd Xs[NB][NB + 1];
  __shared__ double dsh[NB];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e % NB, l = e / NB;
    const zd g = G[(size_t)i + (size_t)l * NB];
    Gs[i][l] = g;
    if (!zfinite(g)) bad = 1;
    if (i == l && !(g.x >= 0x1p-600 && g.x <= 0x1p600)) bad = 1;
  }
  __syncthreads();
  bool ok = bad == 0;
  if (ok) ok = chol_upper<NB>(Gs, 0x1p-40, dsh);
  if (!ok) {
    if (t == 0) *fail = 1u;
    return;
  }
  inv_upper<NB>(Gs, Xs);
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e % NB, l = e / NB;
    R1[(size_t)i + (size_t)l * NB] = Gs[i][l];
    R1i[(size_t)i + (size_t)l * NB] = Xs[i][l];
  }
}

// Pass 2 and the reconstruction.  In: G2 = Q1^H Q1, R1, the top NB rows of
// Q1.  Out: tau, T (ldt), the clean V's top NB rows (ldv), Atop = R (upper)
// and L1 (strictly lower) for the commit, M = R2^{-1} U^{-1} for the rest of
// V; or the fail flag.
template <int NB>
__global__ __launch_bounds__(kSmall) void recon_kernel(const zd* __restrict__ G2, const zd* __restrict__ R1,
                                                         const zd* __restrict__ Q1, long ldq, zd* __restrict__ tau,
                                                         zd* __restrict__ Tm, long ldt, zd* __restrict__ V, long ldv,
                                                         zd* __restrict__ Atop, zd* __restrict__ M, unsigned* fail) {
  __shared__ zd S1[NB][NB + 1];  // G2 -> R2
  __shared__ zd S2[NB][NB + 1];  // R2^{-1}
  __shared__ zd S3[NB][NB + 1];  // scratch: R1, then Q_top -> L1\U, ...
  __shared__ zd S4[NB][NB + 1];
  __shared__ double dsh[NB];
  __shared__ double Ds[NB];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (*fail) return;  // pass 1 rejected the panel (uniform)
  if (t == 0) bad = 0;
  __syncthreads();
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e % NB, l = e / NB;
    const zd g = G2[(size_t)i + (size_t)l * NB];
    S1[i][l] = g;
    const zd eij = (i == l) ? mkz(g.x - 1.0, g.y) : g;
    if (!zfinite(g) || zabs2(eij) > 0x1p-54) bad = 1;  // |E_ij| > 2^-27
  }
  __syncthreads();
  bool ok = bad == 0;
  if (ok) ok = chol_upper<NB>(S1, 0x1p-2, dsh);
  if (!ok) {
    if (t == 0) *fail = 1u;
    return;
  }
  inv_upper<NB>(S1, S2);  // S2 = R2^{-1}
  // S4 = Q_top = Q1_top R2^{-1}
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e % NB, l = e / NB;
    S3[i][l] = Q1[(size_t)i + (size_t)l * ldq];
  }
  __syncthreads();
  mul<NB>(S3, S2, S4);
  // LU without pivoting of Q_top - D in S4 (L1 strictly below, U on and above)
  for (int j = 0; j < NB; ++j) {
    const zd piv = S4[j][j];
    const double dj = piv.x > 0.0 ? -1.0 : 1.0;  // uniform
    __syncthreads();
    if (t == 0) {
      Ds[j] = dj;
      S4[j][j] = mkz(piv.x - dj, piv.y);
    }
    __syncthreads();
    const zd rp = zrecip(S4[j][j]);
    for (int i = j + 1 + t; i < NB; i += kSmall) S4[i][j] = S4[i][j] * rp;
    __syncthreads();
    const int rest = NB - j - 1;
    for (int e = t; e < rest * rest; e += kSmall) {
      const int i = j + 1 + e / rest, l = j + 1 + e % rest;
      S4[i][l] = S4[i][l] - S4[i][j] * S4[j][l];
    }
    __syncthreads();
  }
  // R = D R2 R1 -> Atop (upper), L1 (strictly lower) -> Atop
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e % NB, l = e / NB;
    S3[i][l] = R1[(size_t)i + (size_t)l * NB];
  }
  __syncthreads();
  {
    // Atop(i, l) = D_i (R2 R1)(i, l) for i <= l, L1(i, l) for i > l
    for (int e = t; e < NB * NB; e += kSmall) {
      const int i = e / NB, l = e % NB;
      zd v;
      if (i <= l) {
        zd s = S<zd>::zero();
        for (int k = i; k <= l; ++k) s = s + S1[i][k] * S3[k][l];
        v = s * Ds[i];
      } else {
        v = S4[i][l];
      }
      Atop[(size_t)i + (size_t)l * NB] = v;
      // the clean V's top rows: unit lower L1
      V[(size_t)i + (size_t)l * ldv] = (i > l) ? S4[i][l] : (i == l ? S<zd>::one() : S<zd>::zero());
    }
  }
  if (t < NB) tau[t] = S<zd>::zero() - S4[t][t] * Ds[t];
  __syncthreads();
  // S3 = U (upper part of S4), then S1 = U^{-1}
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e / NB, l = e % NB;
    S3[i][l] = (i <= l) ? S4[i][l] : S<zd>::zero();
  }
  __syncthreads();
  inv_upper<NB>(S3, S1);  // S1 = U^{-1} (R2 no longer needed: S2 holds R2^{-1})
  // M = R2^{-1} U^{-1}
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e / NB, l = e % NB;
    zd s = S<zd>::zero();
    for (int k = i; k <= l; ++k) s = s + S2[i][k] * S1[k][l];
    M[(size_t)i + (size_t)l * NB] = (i <= l) ? s : S<zd>::zero();
  }
  __syncthreads();
  // X = L1^{-1} (unit lower) into S2, thread c forward-substitutes column c
  if (t < NB) {
    const int c = t;
    for (int i = 0; i < c; ++i) S2[i][c] = S<zd>::zero();
    S2[c][c] = S<zd>::one();
    for (int i = c + 1; i < NB; ++i) {
      zd s = S<zd>::zero();
      for (int k = c; k < i; ++k) s = s + S4[i][k] * S2[k][c];
      S2[i][c] = S<zd>::zero() - s;
    }
  }
  __syncthreads();
  // T = -(U D) L1^{-H}: T(i, k) = -sum_{l = i..k} U(i, l) D_l conj(X(k, l))
  for (int e = t; e < NB * NB; e += kSmall) {
    const int i = e / NB, k = e % NB;
    zd s = S<zd>::zero();
    if (i <= k)
      for (int l = i; l <= k; ++l) s = s + (S3[i][l] * Ds[l]) * zconj(S2[k][l]);
    Tm[(size_t)i + (size_t)k * ldt] = S<zd>::zero() - s;
  }
}

// A's panel from Atop (top NB rows) and the clean V (rows NB..mp), unless
// the panel was rejected.
__global__ __launch_bounds__(kThreads) void commit_kernel(int mp, int nb, const zd* __restrict__ Atop,
                                                          const zd* __restrict__ V, long ldv, zd* __restrict__ A,
                                                          long lda, const unsigned* fail) {
  if (*fail) return;
  const long total = (long)mp * nb;
  for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < total; e += (long)gridDim.x * kThreads) {
    const int r = (int)(e % mp), c = (int)(e / mp);
    A[(size_t)r + (size_t)c * lda] = r < nb ? Atop[(size_t)r + (size_t)c * nb] : V[(size_t)r + (size_t)c * ldv];
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
  RNPX_TRY(h->zcqr.ensure(sizeof(zd) * ((size_t)mp * nb + (6 + (size_t)ks) * nn), h));
  zd* Q1 = (zd*)h->zcqr.p;
  zd* G = Q1 + (size_t)mp * nb;
  zd* R1 = G + nn;
  zd* R1i = R1 + nn;
  zd* Atop = R1i + nn;
  zd* M = Atop + nn;
  zd* part = M + nn;  // split-K partials of the Gram products
  const zd one = S<zd>::one(), zero = S<zd>::zero();
  // pass 1
  RNPX_TRY(gemm<zd>('C', 'N', nb, nb, mp, one, A, lda, A, lda, zero, G, nb, ks, part, st));
  hipLaunchKernelGGL((zq::chol_kernel<32>), dim3(1), dim3(zq::kSmall), 0, st, G, R1, R1i, fail);
  RNPX_TRY(hipGetLastError());
  RNPX_TRY(gemm<zd>('N', 'N', mp, nb, nb, one, A, lda, R1i, nb, zero, Q1, mp, 1, nullptr, st));
  // pass 2 and the reconstruction
  RNPX_TRY(gemm<zd>('C', 'N', nb, nb, mp, one, Q1, mp, Q1, mp, zero, G, nb, ks, part, st));
  hipLaunchKernelGGL((zq::recon_kernel<32>), dim3(1), dim3(zq::kSmall), 0, st, G, R1, Q1, (long)mp, tau, Tm, ldt,
                     V, ldv, Atop, M, fail);
  RNPX_TRY(hipGetLastError());
  // the rest of V: Q1_bot R2^{-1} U^{-1}
  RNPX_TRY(gemm<zd>('N', 'N', mp - nb, nb, nb, one, Q1 + nb, mp, M, nb, zero, V + nb, ldv, 1, nullptr, st));
  const int cgrid = std::min(1024, (int)(((long)mp * nb + zq::kThreads - 1) / zq::kThreads));
  hipLaunchKernelGGL(zq::commit_kernel, dim3(cgrid), dim3(zq::kThreads), 0, st, mp, nb, Atop, V, ldv, A, lda, fail);
  RNPX_TRY(hipGetLastError());
  // a rejected panel: column by column (both launches return at once otherwise)
  PanelArgs<zd> fb = pa;
  fb.run_if = fail;
  RNPX_TRY(launch_panel<zd>(fb, nb, rpt, st));
  RNPX_TRY(launch_tfactor<zd>(fb.gram, nb, std::min(mp, pa.nbp), nb, Tm, (int)ldt, st, fail));
  return RNPX_OK;
}

}  // namespace rnpx
