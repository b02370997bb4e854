// rnpx_team.hpp -- one workgroup ("team" of NT threads) factoring a matrix
// column by column in place: QR::Factor_unblocked (RNP/LA/QR.hpp:78-102)
// with Reflector::Generate's rules (RNP/LA/Reflector.hpp:117-158).  Used by
// the batched form and by the CholeskyQR2 panel's column-wise refactorization.
#pragma once
#include "rnpx_kernels.hpp"

namespace rnpx {

__device__ inline double team_sum(double v, double* sh, int nt) {
  v = wave_sum(v);
  if (nt == 64) return v;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nt / 64; ++w) s += sh[w];
  return s;
}
__device__ inline double team_max(double v, double* sh, int nt) {
  v = wave_max(v);
  if (nt == 64) return v;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nt / 64; ++w) s = fmax(s, sh[w]);
  return s;
}
template <int NT>
__device__ inline void team_sync() {
  if constexpr (NT == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

// Factor one m x n matrix held at W (leading dimension ld) in place, tau out.
// Team of NT threads (threadIdx.x < NT).  sh: 8 doubles of LDS scratch.
// WAVECOL: the Householder application gives each wave one column at a time,
// its lanes splitting the rows (coalesced, for tall panels); otherwise each
// lane owns whole columns (no cross-lane reduction, for short matrices).
template <typename T, int NT, bool WAVECOL = false>
__device__ void team_factor(int m, int n, T* W, int ld, T* tau, double* sh) {
  const int t = threadIdx.x;
  const int k = min(m, n);
  const double safmin = 0x1p-971;
  for (int c = 0; c < k; ++c) {
    // |x|^2 with x = W(c+1:m, c); scaled exactly like the panel kernel
    double ss = 0.0;
    for (int r = c + 1 + t; r < m; r += NT) ss += S<T>::abs2(W[r + c * ld]);
    ss = team_sum(ss, sh, NT);
    const T alpha = W[c + c * ld];
    const double ar = S<T>::re(alpha), ai = S<T>::im(alpha);
    double xnorm;
    int e = 0;
    if (ss >= 0x1p-600 && ss <= 0x1p600) {
      xnorm = sqrt(ss);
    } else {
      double mx = 0.0;
      for (int r = c + 1 + t; r < m; r += NT) {
        const T x = W[r + c * ld];
        mx = fmax(mx, fmax(fabs(S<T>::re(x)), fabs(S<T>::im(x))));
      }
      mx = team_max(mx, sh, NT);
      if (mx > 0.0 && mx <= 1.7976931348623157e308) e = ilogb(mx);
      double s2 = 0.0;
      for (int r = c + 1 + t; r < m; r += NT) {
        const T x = W[r + c * ld];
        s2 += S<T>::abs2(S<T>::make(ldexp(S<T>::re(x), -e), ldexp(S<T>::im(x), -e)));
      }
      s2 = team_sum(s2, sh, NT);
      ss = s2;
      xnorm = ldexp(sqrt(s2), e);
    }
    team_sync<NT>();
    if (xnorm == 0.0 && ai == 0.0) {
      if (t == 0) tau[c] = S<T>::zero();
      team_sync<NT>();
      continue;
    }
    double beta = hypot3(ar, ai, xnorm);
    if (ar > 0) beta = -beta;
    double sar = ar, sai = ai;
    int escale = 0;
    if (fabs(beta) < safmin) {  // two scalings by 2^971 (Reflector.hpp:139-144)
      escale = 1942;
      sar = ldexp(ar, 1942);
      sai = ldexp(ai, 1942);
      beta = hypot3(sar, sai, ldexp(sqrt(ss), e + 1942));
      if (sar > 0) beta = -beta;
    }
    const T tauc = S<T>::make((beta - sar) / beta, (-sai) / beta);
    const T inv = recip_t(S<T>::make(sar - beta, sai));
    if (escale) beta = (beta * safmin) * safmin;  // safmin * safmin alone underflows to 0
    for (int r = c + 1 + t; r < m; r += NT) {
      T x = W[r + c * ld];
      if (escale) x = S<T>::make(ldexp(S<T>::re(x), escale), ldexp(S<T>::im(x), escale));
      W[r + c * ld] = x * inv;
    }
    team_sync<NT>();
    // apply H^H = I - conj(tau) v v^H to columns c+1..n-1 (column per lane)
    const T ctau = S<T>::conj(tauc);
    if constexpr (WAVECOL) {
      // UNR columns per wave per pass: independent chains in flight.
      constexpr int NW = NT / 64, UNR = 4;
      const int wave = t >> 6, lane = t & 63;
      for (int j0 = c + 1 + wave; j0 < n; j0 += NW * UNR) {
        T p[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) p[u] = S<T>::zero();
        for (int r = c + 1 + lane; r < m; r += 64) {
          const T v = W[r + c * ld];
#pragma unroll
          for (int u = 0; u < UNR; ++u)
            if (j0 + u * NW < n) p[u] = p[u] + S<T>::cmul(v, W[r + (j0 + u * NW) * ld]);
        }
        T cw[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
          const double sr = wave_sum(S<T>::re(p[u]));
          const double si = S<T>::cplx ? wave_sum(S<T>::im(p[u])) : 0.0;
          cw[u] = S<T>::zero();
          if (j0 + u * NW < n) cw[u] = ctau * (W[c + (j0 + u * NW) * ld] + S<T>::make(sr, si));
        }
        if (lane == 0) {
#pragma unroll
          for (int u = 0; u < UNR; ++u)
            if (j0 + u * NW < n) W[c + (j0 + u * NW) * ld] = W[c + (j0 + u * NW) * ld] - cw[u];
        }
        for (int r = c + 1 + lane; r < m; r += 64) {
          const T v = W[r + c * ld];
#pragma unroll
          for (int u = 0; u < UNR; ++u)
            if (j0 + u * NW < n) W[r + (j0 + u * NW) * ld] = W[r + (j0 + u * NW) * ld] - v * cw[u];
        }
      }
    } else {
      // Four independent partial sums (rows r mod 4) and unrolled loops, so
      // that the LDS loads of four rows are in flight at once: with one wave
      // per SIMD nothing else hides their latency.
      const T* v = W + c * ld;
      for (int j = c + 1 + t; j < n; j += NT) {
        T* col = W + j * ld;
        T w0 = col[c], w1 = S<T>::zero(), w2 = S<T>::zero(), w3 = S<T>::zero();
        int r = c + 1;
        for (; r + 4 <= m; r += 4) {
          w0 = w0 + S<T>::cmul(v[r], col[r]);
          w1 = w1 + S<T>::cmul(v[r + 1], col[r + 1]);
          w2 = w2 + S<T>::cmul(v[r + 2], col[r + 2]);
          w3 = w3 + S<T>::cmul(v[r + 3], col[r + 3]);
        }
        for (; r < m; ++r) w0 = w0 + S<T>::cmul(v[r], col[r]);
        const T cw = ctau * ((w0 + w1) + (w2 + w3));
        col[c] = col[c] - cw;
        r = c + 1;
        for (; r + 4 <= m; r += 4) {
          const T c0 = col[r] - v[r] * cw, c1 = col[r + 1] - v[r + 1] * cw;
          const T c2 = col[r + 2] - v[r + 2] * cw, c3 = col[r + 3] - v[r + 3] * cw;
          col[r] = c0;
          col[r + 1] = c1;
          col[r + 2] = c2;
          col[r + 3] = c3;
        }
        for (; r < m; ++r) col[r] = col[r] - v[r] * cw;
      }
    }
    if (t == 0) {
      tau[c] = tauc;
      W[c + c * ld] = S<T>::make(beta, 0.0);
    }
    team_sync<NT>();
  }
}

}  // namespace rnpx
