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
#include "rnpx_team.hpp"

#include <cstdlib>

namespace rnpx {

// Least squares for m >= n: B (m x nrhs, ldb) <- Q^H B, then R X = B(0:n,:)
// in place (X in the first n rows), one right-hand side at a time with the
// team's threads splitting the rows: for each reflector a team dot product
// and an axpy (Q^H b in the order QR::MultQ("L","C") applies H_0, H_1, ...),
// then column-oriented back substitution (x_i = b_i / r_ii, then
// b(0:i) -= R(0:i, i) x_i).
template <typename T, int NT>
__device__ void team_solve(int m, int n, const T* W, int ld, const T* tau, int nrhs, T* B, int ldb, double* sh,
                           T* xs) {
  const int t = threadIdx.x;
  for (int q = 0; q < nrhs; ++q) {
    T* b = B + (size_t)q * ldb;
    for (int c = 0; c < n; ++c) {
      const T tc = tau[c];
      if (S<T>::is_zero(tc)) continue;
      T p = S<T>::zero();
      for (int r = c + 1 + t; r < m; r += NT) p = p + S<T>::cmul(W[r + c * ld], b[r]);
      double pr = team_sum(S<T>::re(p), sh, NT);
      double pi = 0.0;
      if constexpr (S<T>::cplx) pi = team_sum(S<T>::im(p), sh, NT);
      const T w = b[c] + S<T>::make(pr, pi);
      const T cw = S<T>::conj(tc) * w;
      team_sync<NT>();
      if (t == 0) b[c] = b[c] - cw;
      for (int r = c + 1 + t; r < m; r += NT) b[r] = b[r] - W[r + c * ld] * cw;
      team_sync<NT>();
    }
    for (int i = n - 1; i >= 0; --i) {
      if (t == 0) {
        const T x = b[i] * recip_t(W[i + i * ld]);
        b[i] = x;
        *xs = x;
      }
      team_sync<NT>();
      const T x = *xs;
      for (int j = t; j < i; j += NT) b[j] = b[j] - W[j + i * ld] * x;
      team_sync<NT>();
    }
  }
}

template <typename T>
__global__ __launch_bounds__(64) void batched_small_kernel(int count, const int* __restrict__ order,
                                                           const rnpx_qr_desc* __restrict__ desc, T* A, T* tau, T* B,
                                                           unsigned* counter, int lds_elems) {
  extern __shared__ double smem[];
  __shared__ double sh[8];
  __shared__ T xs;
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
    if (d.nrhs > 0 && B && m >= n) team_solve<T, 64>(m, n, W, ld, tau + d.tau_off, d.nrhs, B + d.b_off, d.ldb, sh, &xs);
    team_sync<64>();
  }
}

// Tier L: one workgroup per problem, several per CU (real: four of 128
// threads, complex: two of 256): the per-column reductions of a panel leave
// most of a CU idle, so problems share one.  Problems of more than one
// panel's columns take the blocked path below (panels in LDS, block
// reflector updates on MFMA); narrower ones that fit are staged in LDS whole
// (column-major, ld = m: the waves read down columns); the rest (taller than
// the blocked path's panel buffer) are factored in place in global memory.
// The Householder application gives each wave one column at a time with the
// lanes splitting the rows (team_factor's WAVECOL mode).
// Per element type: threads per team, teams per CU and LDS doubles per team
// (real: four 128-thread teams of 39.5 KiB; complex: two 256-thread teams of
// 79.5 KiB), the panel width and the most panel rows the LDS carve-up holds.
template <typename T> struct Large;
template <> struct Large<double> { static constexpr int NT = 128, PER_CU = 4, LDS = 5056, PW = 16, ROWS = 256; };
template <> struct Large<zd> { static constexpr int NT = 256, PER_CU = 2, LDS = 10176, PW = 16, ROWS = 256; };
constexpr int kBlockedMinCols = 17;  // more than one panel: blocked (measured faster than whole-LDS)
template <typename T> using Blk = Large<T>;
static_assert(Large<double>::ROWS * Large<double>::PW + 2 * 16 * 17 <= Large<double>::LDS, "LDS carve-up (real)");
static_assert(2 * (Large<zd>::ROWS * Large<zd>::PW + 2 * 16 * 17) <= Large<zd>::LDS, "LDS carve-up (complex)");
static_assert(Large<double>::PER_CU * (Large<double>::LDS * 8 + 256) <= 160 * 1024, "teams per CU (real)");
static_assert(Large<zd>::PER_CU * (Large<zd>::LDS * 8 + 256) <= 160 * 1024, "teams per CU (complex)");

// acc += a * b on FP64 MFMA for real or complex (4 real products) operands.
struct CAcc {
  f64x4 re, im;
};
__device__ __forceinline__ void cmma(f64x4& acc, double a, double b) { acc = mfma16(a, b, acc); }
__device__ __forceinline__ void cmma(CAcc& acc, zd a, zd b) {
  acc.re = mfma16(a.x, b.x, acc.re);
  acc.re = mfma16(-a.y, b.y, acc.re);
  acc.im = mfma16(a.x, b.y, acc.im);
  acc.im = mfma16(a.y, b.x, acc.im);
}
template <typename T> struct AccOf;
template <> struct AccOf<double> {
  using type = f64x4;
  __device__ static inline type zero() { return f64x4{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline double get(const type& x, int q) { return x[q]; }
  __device__ static inline void set(type& x, int q, double v) { x[q] = v; }
};
template <> struct AccOf<zd> {
  using type = CAcc;
  __device__ static inline type zero() { return CAcc{f64x4{0.0, 0.0, 0.0, 0.0}, f64x4{0.0, 0.0, 0.0, 0.0}}; }
  __device__ static inline zd get(const type& x, int q) { return mkz(x.re[q], x.im[q]); }
  __device__ static inline void set(type& x, int q, zd v) {
    x.re[q] = v.x;
    x.im[q] = v.y;
  }
};

// A problem too large for LDS (m <= Blk<T>::ROWS): blocked by panels of PW
// columns.  Each panel is staged in LDS and factored there (team_factor, the
// QR::Factor_unblocked column loop); T comes from the panel's V as in
// GenerateBlockTr("F","C") (Reflector.hpp:580-626): T(c,c) = tau_c,
// T(0:c,c) = -tau_c T(0:c,0:c) V(:,0:c)^H v_c; the trailing columns then take
// the block reflector at once (ApplyBlock, Reflector.hpp:739-777:
// W = V^H C, Z = T^H W, C -= V Z), each wave 16 columns at a time on FP64
// MFMA with W and Z kept in its accumulators (their layout is the next
// product's operand layout, slice for slice).  The update is issued as
// C^T -= Z^T V^T so that the C tile's loads and stores run down columns.
template <typename T>
__device__ void blocked_factor(int m, int n, T* a, long lda, T* tau, double* lds, double* sh) {
  constexpr int NT = Large<T>::NT, NW = NT / 64, PW = Blk<T>::PW, NB2 = PW / 16, TS = PW + 1;
  using AO = AccOf<T>;
  using Acc = typename AO::type;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  T* Pn = reinterpret_cast<T*>(lds);     // mp x PW, ld mp
  T* Tt = Pn + Blk<T>::ROWS * PW;        // T, row-major, stride TS
  T* Gs = Tt + PW * TS;                  // V's Gram, row-major, stride TS
  const T zero = S<T>::zero();
  const int k = min(m, n);
  for (int p0 = 0; p0 < k; p0 += PW) {
    const int pw = min(PW, k - p0), mp = m - p0;
    T* ap = a + (size_t)p0 + (size_t)p0 * lda;
    for (int e = t; e < mp * pw; e += NT) Pn[e] = ap[(size_t)(e % mp) + (size_t)(e / mp) * lda];
    __syncthreads();
    team_factor<T, NT, true>(mp, pw, Pn, mp, tau + p0, sh);
    __syncthreads();
    for (int e = t; e < mp * pw; e += NT) ap[(size_t)(e % mp) + (size_t)(e / mp) * lda] = Pn[e];
    const int nc = n - p0 - pw;
    __syncthreads();
    if (nc <= 0) continue;
    // Clean V (unit diagonal, zeros above; zero columns beyond pw).
    for (int e = t; e < mp * PW; e += NT) {
      const int r = e % mp, c = e / mp;
      if (c >= pw || r <= c) Pn[e] = (r == c && c < pw) ? S<T>::one() : zero;
    }
    __syncthreads();
    static_assert(PW == 16, "the Gram below is one 16 x 16 MFMA block");
    if (wave == 0) {
      // V's Gram V^H V on MFMA over the panel's rows (the clean V is zero
      // above its unit diagonal, so the full sum is the reference's
      // sum over rows >= c); lane item q holds G(lk + 4q, li)
      const int li = lane & 15, lk = lane >> 4;
      Acc g = AO::zero();
      for (int ks = 0; ks < mp; ks += 4) {
        const int r = ks + lk;
        const T v = (r < mp) ? Pn[r + li * mp] : zero;
        cmma(g, S<T>::conj(v), v);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = lk + 4 * q;
        Gs[l * TS + li] = (l < li && li < pw) ? AO::get(g, q) : zero;
        Tt[l * TS + li] = zero;
      }
    }
    __syncthreads();
    // T column by column: row t of T reads only row t's earlier entries
    // (written by this same thread) and the Gram, so no barrier per column
    if (t < pw) {
      for (int c = t; c < pw; ++c) {
        const T tc = tau[p0 + c];
        if (t == c) {
          Tt[c * TS + c] = tc;
        } else {
          T sacc = zero;
          for (int l = t; l < c; ++l) sacc = sacc + Tt[t * TS + l] * Gs[l * TS + c];
          Tt[t * TS + c] = zero - tc * sacc;
        }
      }
    }
    __syncthreads();
    // Trailing columns [p0 + pw, n), rows [p0, m): 16-column tiles per wave.
    T* cp = a + (size_t)p0;  // row p0 of column 0
    const int jbeg = p0 + pw;
    const int ntiles = (nc + 15) / 16;
    const int li = lane & 15, lk = lane >> 4;
    for (int jt = wave; jt < ntiles; jt += NW) {
      const int j0 = jbeg + 16 * jt;
      Acc w[NB2];
#pragma unroll
      for (int ib = 0; ib < NB2; ++ib) w[ib] = AO::zero();
      {
        // Branch-free loads (indices clamped into the matrix, out-of-range
        // values selected to zero afterwards), eight row slices at a time so
        // that their global loads are in flight together.
        const int j = j0 + li, jc = min(j, n - 1);
        const bool jok = j < n;
        constexpr int U = 8;
        for (int ks0 = 0; ks0 < mp; ks0 += 4 * U) {
          T cv[U], vv[U][NB2];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int rc = min(ks0 + 4 * u + lk, mp - 1);
            cv[u] = cp[(size_t)rc + (size_t)jc * lda];
#pragma unroll
            for (int ib = 0; ib < NB2; ++ib) vv[u][ib] = Pn[rc + (16 * ib + li) * mp];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const bool rok = ks0 + 4 * u + lk < mp;
            const T c = (rok && jok) ? cv[u] : zero;
#pragma unroll
            for (int ib = 0; ib < NB2; ++ib)
              cmma(w[ib], rok ? S<T>::conj(vv[u][ib]) : zero, c);  // W = V^H C: a = V^H[i][k], b = C[k][j]
          }
        }
      }
      Acc z[NB2];
#pragma unroll
      for (int ib = 0; ib < NB2; ++ib) {
        z[ib] = AO::zero();
#pragma unroll
        for (int kb = 0; kb < NB2; ++kb)
#pragma unroll
          for (int sl = 0; sl < 4; ++sl) {
            const T tv = S<T>::conj(Tt[(16 * kb + 4 * sl + lk) * TS + 16 * ib + li]);  // T^H[i][k]
            cmma(z[ib], tv, AO::get(w[kb], sl));
          }
      }
#pragma unroll 2
      for (int r0 = 0; r0 < mp; r0 += 16) {
        const int rr = r0 + li, rc = min(rr, mp - 1);
        const bool rok = rr < mp;
        Acc acc = AO::zero();
        T cl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) cl[q] = cp[(size_t)rc + (size_t)min(j0 + lk + 4 * q, n - 1) * lda];
        T vl[NB2][4];
#pragma unroll
        for (int kb = 0; kb < NB2; ++kb)
#pragma unroll
          for (int sl = 0; sl < 4; ++sl) vl[kb][sl] = Pn[rc + (16 * kb + 4 * sl + lk) * mp];
#pragma unroll
        for (int q = 0; q < 4; ++q) AO::set(acc, q, (rok && j0 + lk + 4 * q < n) ? cl[q] : zero);
#pragma unroll
        for (int kb = 0; kb < NB2; ++kb)
#pragma unroll
          for (int sl = 0; sl < 4; ++sl)
            cmma(acc, zero - AO::get(z[kb], sl), rok ? vl[kb][sl] : zero);  // C^T -= Z^T V^T
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = j0 + lk + 4 * q;
          if (rok && j < n) cp[(size_t)rr + (size_t)j * lda] = AO::get(acc, q);
        }
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(Large<T>::NT) void batched_large_kernel(int count, const int* __restrict__ order,
                                                            const rnpx_qr_desc* __restrict__ desc, T* A, T* tau,
                                                            T* B, unsigned* counter, int bmin) {
  constexpr int NT = Large<T>::NT;
  __shared__ double lds[Large<T>::LDS];
  __shared__ double sh[8];
  __shared__ T xs;
  __shared__ int next;
  constexpr int cap = Large<T>::LDS / (int)(sizeof(T) / sizeof(double));
  T* Ws = reinterpret_cast<T*>(lds);
  const int t = threadIdx.x;
  while (true) {
    if (t == 0) next = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int idx = next;
    __syncthreads();
    if (idx >= count) break;
    const rnpx_qr_desc d = desc[order[idx]];
    const int m = d.m, n = d.n;
    T* a = A + d.a_off;
    const bool solve = d.nrhs > 0 && B && m >= n;
    if (m * n <= cap && (n < bmin || m > Blk<T>::ROWS)) {
      for (int e = t; e < m * n; e += NT) Ws[e] = a[(size_t)(e % m) + (size_t)(e / m) * d.lda];
      __syncthreads();
      team_factor<T, NT, true>(m, n, Ws, m, tau + d.tau_off, sh);
      __syncthreads();
      for (int e = t; e < m * n; e += NT) a[(size_t)(e % m) + (size_t)(e / m) * d.lda] = Ws[e];
      if (solve) team_solve<T, NT>(m, n, Ws, m, tau + d.tau_off, d.nrhs, B + d.b_off, d.ldb, sh, &xs);
    } else {
      if (m <= Blk<T>::ROWS)
        blocked_factor<T>(m, n, a, d.lda, tau + d.tau_off, lds, sh);
      else
        team_factor<T, NT, true>(m, n, a, d.lda, tau + d.tau_off, sh);
      __threadfence_block();
      __syncthreads();
      if (solve) team_solve<T, NT>(m, n, a, d.lda, tau + d.tau_off, d.nrhs, B + d.b_off, d.ldb, sh, &xs);
    }
    __syncthreads();
  }
}

template <typename T>
int batched_large_teams(int count, int num_cus) {
  return std::max(1, std::min(count, Large<T>::PER_CU * num_cus));
}
template int batched_large_teams<double>(int, int);
template int batched_large_teams<zd>(int, int);

template <typename T>
hipError_t launch_batched(int tier, int count, const int* order, const rnpx_qr_desc* desc, T* A, T* tau, T* B,
                          unsigned* counter, int teams, int param, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (teams < 1) return hipErrorInvalidValue;
  RNPX_LAUNCH_CHECK(launch_zero_u32(counter, 1, st));
  if (tier == 0) {
    const int lds_bytes = param;
    if (lds_bytes < 1024 || lds_bytes > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((batched_small_kernel<T>), dim3(teams), dim3(64), lds_bytes, st, count, order, desc, A, tau, B,
                       counter, lds_bytes / (int)sizeof(T));
  } else {
    // Real problems of at least bmin columns take the blocked path even when
    // they would fit in LDS whole (RNPX_BATCH_BLOCKED_MIN, read once,
    // overrides the caller's value).
    static const int env_bmin = [] {
      const char* e = getenv("RNPX_BATCH_BLOCKED_MIN");
      return e ? atoi(e) : 0;
    }();
    const int bmin = env_bmin > 0 ? env_bmin : (param > 0 ? param : kBlockedMinCols);
    hipLaunchKernelGGL((batched_large_kernel<T>), dim3(teams), dim3(Large<T>::NT), 0, st, count, order, desc, A, tau, B,
                       counter, bmin);
  }
  return hipGetLastError();
}

template hipError_t launch_batched<double>(int, int, const int*, const rnpx_qr_desc*, double*, double*, double*,
                                           unsigned*, int, int, hipStream_t);
template hipError_t launch_batched<zd>(int, int, const int*, const rnpx_qr_desc*, zd*, zd*, zd*, unsigned*, int, int,
                                       hipStream_t);

}  // namespace rnpx
