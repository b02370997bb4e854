// rnpx_batched.hip -- thousands of small, uneven QR factorizations (and
// least-squares solves) in one persistent launch per tier.
//
// Each problem is RNP's QR::Factor on its own m x n matrix (for the sizes of
// this form, m <= 256, that is the column loop of QR::Factor_unblocked,
// RNP/LA/QR.hpp:78-102, with Reflector::Generate's rules, Reflector.hpp:
// 117-158), optionally followed by the least-squares solve Q^H b then R x = y.
//
// Scheduling: the host sorts the problems by cost (largest first) into an
// index list per tier; persistent teams pull the next index from an atomic
// counter until the list is exhausted (every team exits when the counter
// passes the count).  Tier S: one 64-lane wave per problem, the matrix staged
// in LDS with an odd leading dimension (conflict-free column-per-lane
// access).  Tier L: 256 threads per problem working in place in global
// memory (the problem stays L2-resident) for the problems too large for LDS.
// Within a problem, lane t owns columns t, t+NT, ...: the dot products of the
// rank-1 update need no cross-lane reduction; the norm is a team reduction.
// Results do not depend on which team ran a problem, or when.
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
      const int wave = t >> 6, lane = t & 63;
      for (int j = c + 1 + wave; j < n; j += NT / 64) {
        T* col = W + j * ld;
        T p = S<T>::zero();
        for (int r = c + 1 + lane; r < m; r += 64) p = p + S<T>::cmul(W[r + c * ld], col[r]);
        const T w = col[c] + S<T>::make(wave_sum(S<T>::re(p)), wave_sum(S<T>::im(p)));
        const T cw = ctau * w;
        if (lane == 0) col[c] = col[c] - cw;
        for (int r = c + 1 + lane; r < m; r += 64) col[r] = col[r] - W[r + c * ld] * cw;
      }
    } else {
      for (int j = c + 1 + t; j < n; j += NT) {
        T* col = W + j * ld;
        T w = col[c];
        for (int r = c + 1; r < m; ++r) w = w + S<T>::cmul(W[r + c * ld], col[r]);
        const T cw = ctau * w;
        col[c] = col[c] - cw;
        for (int r = c + 1; r < m; ++r) col[r] = col[r] - W[r + c * ld] * cw;
      }
    }
    if (t == 0) {
      tau[c] = tauc;
      W[c + c * ld] = S<T>::make(beta, 0.0);
    }
    team_sync<NT>();
  }
}

// Least squares for m >= n: B (m x nrhs, ldb) <- Q^H B, then R X = B(0:n,:)
// in place (X in the first n rows).  One lane per right-hand side.
template <typename T, int NT>
__device__ void team_solve(int m, int n, const T* W, int ld, const T* tau, int nrhs, T* B, int ldb) {
  const int t = threadIdx.x;
  for (int q = t; q < nrhs; q += NT) {
    T* b = B + (size_t)q * ldb;
    for (int c = 0; c < n; ++c) {
      const T tc = tau[c];
      if (S<T>::is_zero(tc)) continue;
      T w = b[c];
      for (int r = c + 1; r < m; ++r) w = w + S<T>::cmul(W[r + c * ld], b[r]);
      const T cw = S<T>::conj(tc) * w;
      b[c] = b[c] - cw;
      for (int r = c + 1; r < m; ++r) b[r] = b[r] - W[r + c * ld] * cw;
    }
    for (int i = n - 1; i >= 0; --i) {
      T s = b[i];
      for (int j = i + 1; j < n; ++j) s = s - W[i + j * ld] * b[j];
      const T d = W[i + i * ld];
      b[i] = s * recip_t(d);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(64) void batched_small_kernel(int count, const int* __restrict__ order,
                                                           const rnpx_qr_desc* __restrict__ desc, T* A, T* tau, T* B,
                                                           unsigned* counter, int lds_elems) {
  extern __shared__ double smem[];
  __shared__ double sh[8];
  __shared__ int next;
  T* W = reinterpret_cast<T*>(smem);
  while (true) {
    if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1u);
    team_sync<64>();
    const int idx = next;
    team_sync<64>();
    if (idx >= count) break;
    const rnpx_qr_desc d = desc[order[idx]];
    const int m = d.m, n = d.n, ld = d.m | 1;
    T* a = A + d.a_off;
    for (int e = threadIdx.x; e < m * n; e += 64) {
      const int i = e % m, j = e / m;
      W[i + j * ld] = a[(size_t)i + (size_t)j * d.lda];
    }
    team_sync<64>();
    team_factor<T, 64>(m, n, W, ld, tau + d.tau_off, sh);
    for (int e = threadIdx.x; e < m * n; e += 64) {
      const int i = e % m, j = e / m;
      a[(size_t)i + (size_t)j * d.lda] = W[i + j * ld];
    }
    if (d.nrhs > 0 && B && m >= n) team_solve<T, 64>(m, n, W, ld, tau + d.tau_off, d.nrhs, B + d.b_off, d.ldb);
    team_sync<64>();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void batched_large_kernel(int count, const int* __restrict__ order,
                                                            const rnpx_qr_desc* __restrict__ desc, T* A, T* tau,
                                                            T* B, unsigned* counter) {
  __shared__ double sh[8];
  __shared__ int next;
  while (true) {
    if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int idx = next;
    __syncthreads();
    if (idx >= count) break;
    const rnpx_qr_desc d = desc[order[idx]];
    T* a = A + d.a_off;
    team_factor<T, 256>(d.m, d.n, a, d.lda, tau + d.tau_off, sh);
    __threadfence_block();
    __syncthreads();
    if (d.nrhs > 0 && B && d.m >= d.n) team_solve<T, 256>(d.m, d.n, a, d.lda, tau + d.tau_off, d.nrhs, B + d.b_off, d.ldb);
    __syncthreads();
  }
}

// Fallback for a real 64-column panel the CholeskyQR2 kernel rejected: one
// workgroup, launched behind it and gated on its fail flag (it returns at once
// in the common case, so it costs one workgroup slot, not a grid).  When the
// flag is set it factors the mp x ib panel column by column in place with
// team_factor (Reflector::Generate's rules, RNP/LA/Reflector.hpp:117-158, and
// the Householder application of QR::Factor_unblocked, QR.hpp:78-102), writes
// the clean V, and forms T from V's Gram matrix (GenerateBlockTr("F","C"),
// Reflector.hpp:580-626): T(c,c) = tau_c, T(0:c,c) = -tau_c T(0:c,0:c)
// V(:,0:c)^T v_c.  Slow (column-serial over global memory) but exact.
__global__ __launch_bounds__(256) void panel_fallback_kernel(int mp, int ib, double* A, long lda, double* tau,
                                                             double* V, long ldv, double* Tm, long ldt,
                                                             const unsigned* run_if) {
  if (*run_if == 0u) return;
  __shared__ double sh[8];
  __shared__ double Gs[64][65];
  __shared__ double Ts[64][65];
  const int t = threadIdx.x;
  team_factor<double, 256, true>(mp, ib, A, (int)lda, tau, sh);
  __syncthreads();
  // clean V
  for (long e = t; e < (long)mp * ib; e += 256) {
    const int r = (int)(e % mp), c = (int)(e / mp);
    V[(size_t)r + (size_t)c * ldv] = (r > c) ? A[(size_t)r + (size_t)c * lda] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  // Gram of V: G(l,c) = V(:,l)^T V(:,c), l < c
  for (int e = t; e < 64 * 64; e += 256) {
    const int l = e / 64, c = e % 64;
    double g = 0.0;
    if (l < c && c < ib) {
      for (int r = c; r < mp; ++r) g = fma(V[(size_t)r + (size_t)l * ldv], V[(size_t)r + (size_t)c * ldv], g);
    }
    Gs[l][c] = g;
    Ts[l][c] = 0.0;
  }
  __syncthreads();
  for (int c = 0; c < ib; ++c) {
    const double tc = tau[c];
    if (t < c) {
      double s = 0.0;
      for (int l = t; l < c; ++l) s = fma(Ts[t][l], Gs[l][c], s);
      Ts[t][c] = -tc * s;
    }
    if (t == c) Ts[c][c] = tc;
    __syncthreads();
  }
  for (int e = t; e < 64 * 64; e += 256) {
    const int i = e % 64, j = e / 64;
    Tm[(size_t)i + (size_t)j * ldt] = (i < ib && j < ib) ? Ts[i][j] : 0.0;
  }
}

hipError_t launch_panel_fallback(int mp, int ib, double* A, long lda, double* tau, double* V, long ldv, double* Tm,
                                 long ldt, const unsigned* run_if, hipStream_t st) {
  if (mp < 1 || ib < 1 || ib > 64 || lda < mp || ldv < mp || ldt < 64 || !run_if) return hipErrorInvalidValue;
  hipLaunchKernelGGL(panel_fallback_kernel, dim3(1), dim3(256), 0, st, mp, ib, A, lda, tau, V, ldv, Tm, ldt, run_if);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_batched(int tier, int count, const int* order, const rnpx_qr_desc* desc, T* A, T* tau, T* B,
                          unsigned* counter, int teams, int lds_bytes, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (teams < 1) return hipErrorInvalidValue;
  RNPX_LAUNCH_CHECK(launch_zero_u32(counter, 1, st));
  if (tier == 0) {
    if (lds_bytes < 1024 || lds_bytes > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((batched_small_kernel<T>), dim3(teams), dim3(64), lds_bytes, st, count, order, desc, A, tau, B,
                       counter, lds_bytes / (int)sizeof(T));
  } else {
    hipLaunchKernelGGL((batched_large_kernel<T>), dim3(teams), dim3(256), 0, st, count, order, desc, A, tau, B,
                       counter);
  }
  return hipGetLastError();
}

template hipError_t launch_batched<double>(int, int, const int*, const rnpx_qr_desc*, double*, double*, double*,
                                           unsigned*, int, int, hipStream_t);
template hipError_t launch_batched<zd>(int, int, const int*, const rnpx_qr_desc*, zd*, zd*, zd*, unsigned*, int, int,
                                       hipStream_t);

}  // namespace rnpx
