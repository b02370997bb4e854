// rnpx_exact.hip -- QR::Factor (RNP/LA/QR.hpp:129-197) for matrices that
// hold Inf or NaN, and the zero-trimming scans Reflector::LastNonzero*
// (RNP/LA/Reflector.hpp:36-75).
//
// The reference skips rows and columns it finds to be zero before every
// reflector application: Reflector::Apply trims v to its last nonzero entry
// and C to its last nonzero column (Reflector.hpp:196-221), ApplyBlock trims
// V to its last nonzero row and C to its last nonzero column over those rows
// (Reflector.hpp:742-744), and GenerateBlockTr stops each T column at the
// last nonzero of v (Reflector.hpp:581-623).  On finite data the skipped
// products are exact zeros, so the trims change nothing but the sign of a
// zero result, and the fast path (128-column compact-WY blocks) does not
// scan.  Where the matrix holds Inf or NaN the skipped products are 0 * Inf
// = NaN, which of them the reference skips depends on its blocking, and
// geqrf takes this path instead: the reference's own structure (nb-column
// blocks, nb = 32 or what lwork allows, unblocked from the crossover 128 on)
// with every trim it makes, and the zero-skipping of the BLAS contract it
// calls (xgerc and xtrmv skip zero multipliers; xgemv with beta = 0 clears).
//
//  * nonfinite_kernel: any Inf/NaN in an m x n matrix (one pass, a flag).
//  * limit_kernel: LastNonzeroColumnLimit / LastNonzeroRowLimit of a matrix
//    region into device memory: the reference's corner test, else a scan
//    (the same answer as its right-to-left / bottom-up search).
//  * unblocked_kernel: QR::Factor_unblocked (QR.hpp:78-102) by one
//    1024-thread workgroup in place: Reflector::Generate (Reflector.hpp:
//    117-158, BLAS Norm2's Inf/NaN results) and Reflector::Apply with its
//    trims.
//  * block_tr_kernel: Reflector::GenerateBlockTr "F","C" (Reflector.hpp:
//    573-626) by one workgroup.
//  * apply_block_kernel: Reflector::ApplyBlock "L","C","F","C" (Reflector.hpp:
//    739-776) on the trimmed lastv x lastc region, one wave per column of C.
// The host reads lastv / lastc back before each block's apply (this path is
// synchronous; it runs only on non-finite input).
#include "rnpx_internal.hpp"
#include "rnpx_team.hpp"

#include <algorithm>

namespace rnpx {

namespace ex {

constexpr int kNT = 1024;  // threads of the single-workgroup kernels
constexpr int kMaxK = 32;  // widest block (the reference's nb; a short lwork only narrows it)

// T(0) != v, as the reference tests it (NaN counts as nonzero, -0 as zero).
__device__ inline bool nz(double v) { return v != 0.0; }
__device__ inline bool nz(zd v) { return v.x != 0.0 || v.y != 0.0; }
__device__ inline bool finite_t(double v) { return isfinite(v); }
__device__ inline bool finite_t(zd v) { return isfinite(v.x) && isfinite(v.y); }

template <typename T>
__global__ __launch_bounds__(256) void nonfinite_kernel(int m, int n, const T* __restrict__ a, long lda,
                                                        unsigned* flag) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  bool bad = false;
  if (i < m)
    for (int j = blockIdx.y; j < n; j += gridDim.y) bad |= !finite_t(a[(size_t)i + (size_t)j * lda]);
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// which 0: LastNonzeroColumnLimit(m, n, a) = 1 + the last column with a
// nonzero; which 1: LastNonzeroRowLimit(m, n, a) = 1 + the last such row.
// Empty ranges give 0.  *out must be 0 on entry.
template <typename T>
__global__ __launch_bounds__(256) void limit_kernel(int which, int m, int n, const T* __restrict__ a, long lda,
                                                    int* out) {
  if (m <= 0 || n <= 0) return;
  const T* last = a + (size_t)(n - 1) * lda;
  const bool corner = which == 0 ? (nz(last[0]) || nz(last[m - 1])) : (nz(a[m - 1]) || nz(last[m - 1]));
  if (corner) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = which == 0 ? n : m;
    return;
  }
  if (which == 0) {
    for (int j = n - 1 - (int)blockIdx.x; j >= 0; j -= (int)gridDim.x) {
      // a column right of this one is already known to hold a nonzero
      const bool done = threadIdx.x == 0 && *(volatile int*)out > j;
      if (__syncthreads_or(done)) break;
      bool any = false;
      for (int i = threadIdx.x; i < m; i += 256) any |= nz(a[(size_t)i + (size_t)j * lda]);
      if (__syncthreads_or(any)) {
        if (threadIdx.x == 0) atomicMax(out, j + 1);
        break;
      }
    }
  } else {
    int best = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
      bool any = false;
      for (int j = 0; j < n && !any; ++j) any = nz(a[(size_t)i + (size_t)j * lda]);
      if (any) best = i + 1;
    }
    best = (int)wave_max((double)best);
    if ((threadIdx.x & 63) == 0 && best > 0) atomicMax(out, best);
  }
}

__device__ inline void parts(double v, double (&p)[2], int& np) {
  p[0] = v;
  np = 1;
}
__device__ inline void parts(zd v, double (&p)[2], int& np) {
  p[0] = v.x;
  p[1] = v.y;
  np = 2;
}

// BLAS Norm2 of x (len entries, stride 1) as the one-pass scaled sum of
// squares of the reference (real and imaginary parts fed separately) gives
// it: NaN if a part is NaN or two parts are infinite, Inf if one is, else
// max|part| * sqrt(sum (part / max|part|)^2).  Whole team.
template <typename T>
__device__ double norm2_team(int len, const T* x, double* sh) {
  double nan = 0.0, inf = 0.0, mx = 0.0;
  for (int r = threadIdx.x; r < len; r += kNT) {
    double p[2];
    int np;
    parts(x[r], p, np);
    for (int q = 0; q < np; ++q) {
      if (isnan(p[q]))
        nan = 1.0;
      else if (isinf(p[q]))
        inf += 1.0;
      else
        mx = fmax(mx, fabs(p[q]));
    }
  }
  nan = team_max(nan, sh, kNT);
  inf = team_sum(inf, sh, kNT);
  mx = team_max(mx, sh, kNT);
  if (nan > 0.0 || inf > 1.0) return __builtin_nan("");
  if (inf > 0.0) return __builtin_inf();
  if (mx == 0.0) return 0.0;
  double s = 0.0;
  for (int r = threadIdx.x; r < len; r += kNT) {
    double p[2];
    int np;
    parts(x[r], p, np);
    for (int q = 0; q < np; ++q) {
      const double u = p[q] / mx;
      s += u * u;
    }
  }
  s = team_sum(s, sh, kNT);
  return mx * sqrt(s);
}

// LastNonzeroColumnLimit(mm, nn, c) by the whole team: the corner test, else
// columns from the right, 16 at a time, until one holds a nonzero.
template <typename T>
__device__ int lncl_team(int mm, int nn, const T* c, long ldc, double* sh) {
  if (mm <= 0 || nn <= 0) return 0;
  const T* last = c + (size_t)(nn - 1) * ldc;
  if (nz(last[0]) || nz(last[mm - 1])) return nn;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = kNT / 64;
  for (int j0 = nn - 1; j0 >= 0; j0 -= NW) {
    const int j = j0 - wave;
    bool any = false;
    if (j >= 0)
      for (int r = lane; r < mm && !any; r += 64) any = nz(c[(size_t)r + (size_t)j * ldc]);
    const double found = (__ballot(any) != 0ull && j >= 0) ? (double)(j + 1) : 0.0;
    const double best = team_max(found, sh, kNT);
    if (best > 0.0) return (int)best;
  }
  return 0;
}

// QR::Factor_unblocked of the m x n matrix at a (in place), tau out; work
// holds n entries.  One workgroup of kNT threads.
template <typename T>
__global__ __launch_bounds__(kNT) void unblocked_kernel(int m, int n, T* __restrict__ a, long lda, T* __restrict__ tau,
                                                        T* __restrict__ work) {
  __shared__ double sh[kNT / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int NW = kNT / 64;
  const double safmin = 0x1p-971, rsafmn = 0x1p971;
  const int k = min(m, n);
  for (int i = 0; i < k; ++i) {
    T* col = a + (size_t)i + (size_t)i * lda;  // v[0] = a(i,i), x = v[1..]
    const int len = m - i;                     // order of the reflector
    const int lx = len - 1;
    // ---- Reflector::Generate(len, &a(i,i), &a(i+1,i), 1, &tau[i]) ----
    T alpha = col[0];
    double xnorm = norm2_team(lx, col + 1, sh);
    T tv = S<T>::zero();
    if (!(xnorm == 0.0 && S<T>::im(alpha) == 0.0)) {
      double beta = hypot3(S<T>::re(alpha), S<T>::im(alpha), xnorm);
      if (S<T>::re(alpha) > 0) beta = -beta;
      int knt = 0;
      if (fabs(beta) < safmin) {
        do {
          ++knt;
          for (int r = 1 + t; r < len; r += kNT) col[r] = col[r] * rsafmn;
          beta *= rsafmn;
          alpha = alpha * rsafmn;
        } while (fabs(beta) < 1.0);
        __syncthreads();
        xnorm = norm2_team(lx, col + 1, sh);
        beta = hypot3(S<T>::re(alpha), S<T>::im(alpha), xnorm);
        if (S<T>::re(alpha) > 0) beta = -beta;
      }
      tv = S<T>::make((beta - S<T>::re(alpha)) / beta, (-S<T>::im(alpha)) / beta);
      const T inv = recip_t(S<T>::make(S<T>::re(alpha) - beta, S<T>::im(alpha)));
      for (int r = 1 + t; r < len; r += kNT) col[r] = col[r] * inv;
      while (knt-- > 0) beta *= safmin;
      alpha = S<T>::make(beta, 0.0);
    }
    __syncthreads();
    // ---- Reflector::Apply("L", 0, false, len, n-i-1, v, 1, conj(tau), C) with
    //      v(0) = 1 (the reference stores 1 at a(i,i) for the call) ----
    const int nc = n - i - 1;
    T* C = col + lda;
    if (nc > 0 && !S<T>::is_zero(tv)) {
      const T ctau = S<T>::conj(tv);
      // last nonzero of v (v(0) = 1)
      double lv = 1.0;
      for (int r = 1 + t; r < len; r += kNT)
        if (nz(col[r])) lv = fmax(lv, (double)(r + 1));
      const int lenv = (int)team_max(lv, sh, kNT);
      const int lenc = lncl_team(lenv, nc, C, lda, sh);
      if (lenc > 0) {
        // work(j) = sum_{r < lenv} conj(C(r, j)) v(r)  (xgemv "C", beta = 0)
        for (int j = wave; j < lenc; j += NW) {
          const T* cj = C + (size_t)j * lda;
          T s = S<T>::zero();
          for (int r = lane; r < lenv; r += 64) s = s + S<T>::cmul(cj[r], r == 0 ? S<T>::one() : col[r]);
          double sr = wave_sum(S<T>::re(s)), si = 0.0;
          if constexpr (S<T>::cplx) si = wave_sum(S<T>::im(s));
          if (lane == 0) work[j] = S<T>::make(sr, si);
        }
        __syncthreads();
        // C(r, j) += v(r) (-ctau conj(work(j))), columns with work(j) = 0
        // skipped (xgerc)
        const T mtau = S<T>::zero() - ctau;
        for (int j = wave; j < lenc; j += NW) {
          const T wj = work[j];
          if (!nz(wj)) continue;
          const T tmp = mtau * S<T>::conj(wj);
          T* cj = C + (size_t)j * lda;
          for (int r = lane; r < lenv; r += 64) cj[r] = cj[r] + (r == 0 ? S<T>::one() : col[r]) * tmp;
        }
      }
    }
    __syncthreads();
    if (t == 0) {
      tau[i] = tv;
      col[0] = alpha;
    }
    __syncthreads();
  }
}

// Reflector::GenerateBlockTr("F", "C", n, k, v, ldv, tau, t, ldt): v is the
// factored panel (reflector j below the diagonal of column j).  One
// workgroup; k <= kMaxK.
template <typename T>
__global__ __launch_bounds__(kNT) void block_tr_kernel(int n, int k, const T* __restrict__ v, long ldv,
                                                       const T* __restrict__ tau, T* __restrict__ tm, long ldt) {
  __shared__ double sh[kNT / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int NW = kNT / 64;
  if (n <= 0) return;
  int prevlastv = n - 1;
  for (int i = 0; i < k; ++i) {
    if (i > prevlastv) prevlastv = i;
    const T ti = tau[i];
    T* ti_col = tm + (size_t)i * ldt;
    if (!nz(ti)) {
      if (t <= i) ti_col[t] = S<T>::zero();
      __syncthreads();
      continue;
    }
    // lastv: the last row r > i with v(r, i) != 0, else i
    double lv = (double)i;
    for (int r = i + 1 + t; r < n; r += kNT)
      if (nz(v[(size_t)r + (size_t)i * ldv])) lv = fmax(lv, (double)r);
    const int lastv = (int)team_max(lv, sh, kNT);
    const int jlim = min(lastv, prevlastv);
    const T mt = S<T>::zero() - ti;
    // T(j, i) = -tau_i conj(v(i, j)) + (-tau_i) sum_{r=i+1}^{jlim} conj(v(r, j)) v(r, i)
    for (int j = wave; j < i; j += NW) {
      T s = S<T>::zero();
      for (int r = i + 1 + lane; r <= jlim; r += 64)
        s = s + S<T>::cmul(v[(size_t)r + (size_t)j * ldv], v[(size_t)r + (size_t)i * ldv]);
      double sr = wave_sum(S<T>::re(s)), si = 0.0;
      if constexpr (S<T>::cplx) si = wave_sum(S<T>::im(s));
      if (lane == 0) {
        T x = mt * S<T>::conj(v[(size_t)i + (size_t)j * ldv]);
        if (jlim > i) x = x + mt * S<T>::make(sr, si);
        ti_col[j] = x;
      }
    }
    __syncthreads();
    // T(0:i, i) := T(0:i, 0:i) T(0:i, i)  (xtrmv "U","N","N", zero x(j) skipped)
    if (wave == 0 && i > 0) {
      T x = lane < i ? ti_col[lane] : S<T>::zero();
      for (int j = 0; j < i; ++j) {
        const T xj = readlane_t(x, j);
        if (!nz(xj)) continue;
        if (lane < j) x = x + xj * tm[(size_t)lane + (size_t)j * ldt];
        if (lane == j) x = x * tm[(size_t)j + (size_t)j * ldt];
      }
      if (lane < i) ti_col[lane] = x;
    }
    if (t == 0) ti_col[i] = ti;
    if (i > 0) {
      if (lastv > prevlastv) prevlastv = lastv;
    } else {
      prevlastv = lastv;
    }
    __syncthreads();
  }
}

// Reflector::ApplyBlock("L", "C", "F", "C", m, ., k, v, t, c) on the trimmed
// region: rows [0, lastv), columns [0, lastc) of C, lastv >= k.  One wave per
// column of C; lane q < k holds W(j, q) of that column.  k <= kMaxK.
template <typename T>
__global__ __launch_bounds__(256) void apply_block_kernel(int m, int k, int lastv, int lastc, const T* __restrict__ v,
                                                          long ldv, const T* __restrict__ tm, long ldt,
                                                          T* __restrict__ c, long ldc) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= lastc) return;
  T* cj = c + (size_t)j * ldc;
  // W := C1^H
  T w = lane < k ? S<T>::conj(cj[lane]) : S<T>::zero();
  // W := W V1 (xtrmm "R","L","N","U"): w(q) = sum_{l >= q} w(l) V1(l, q), V1(q, q) = 1
  {
    T s = S<T>::zero();
    for (int l = 0; l < k; ++l) {
      const T wl = readlane_t(w, l);
      if (lane < k && l >= lane) s = s + wl * (l == lane ? S<T>::one() : v[(size_t)l + (size_t)lane * ldv]);
    }
    w = s;
  }
  // W += C2^H V2 over rows [k, lastv) (xgemm "C","N")
  if (lastv > k) {
    T acc[kMaxK];
#pragma unroll
    for (int q = 0; q < kMaxK; ++q) acc[q] = S<T>::zero();
    for (int r = k + lane; r < lastv; r += 64) {
      const T cr = cj[r];
#pragma unroll
      for (int q = 0; q < kMaxK; ++q)
        if (q < k) acc[q] = acc[q] + S<T>::cmul(cr, v[(size_t)r + (size_t)q * ldv]);
    }
#pragma unroll
    for (int q = 0; q < kMaxK; ++q) {
      if (q < k) {
        double sr = wave_sum(S<T>::re(acc[q])), si = 0.0;
        if constexpr (S<T>::cplx) si = wave_sum(S<T>::im(acc[q]));
        if (lane == q) w = w + S<T>::make(sr, si);
      }
    }
  }
  // W := W T (xtrmm "R","U","N","N"): w(q) = sum_{l <= q} w(l) T(l, q)
  {
    T s = S<T>::zero();
    for (int l = 0; l < k; ++l) {
      const T wl = readlane_t(w, l);
      if (lane < k && l <= lane) s = s + wl * tm[(size_t)l + (size_t)lane * ldt];
    }
    w = s;
  }
  // C2 -= V2 W^H over rows [k, lastv) (xgemm "N","C", alpha = -1)
  if (m > k) {
    for (int r = k + lane; r < lastv; r += 64) {
      T acc = S<T>::zero();
      for (int q = 0; q < k; ++q) {
        const T bq = S<T>::zero() - S<T>::conj(readlane_t(w, q));
        acc = acc + v[(size_t)r + (size_t)q * ldv] * bq;
      }
      cj[r] = cj[r] + acc;
    }
  }
  // W := W V1^H (xtrmm "R","L","C","U"): w(q) = sum_{l <= q} w(l) conj(V1(q, l))
  {
    T s = S<T>::zero();
    for (int l = 0; l < k; ++l) {
      const T wl = readlane_t(w, l);
      if (lane < k && l <= lane) s = s + wl * (l == lane ? S<T>::one() : S<T>::conj(v[(size_t)lane + (size_t)l * ldv]));
    }
    w = s;
  }
  // C1 -= W^H
  if (lane < k) cj[lane] = cj[lane] - S<T>::conj(w);
}

}  // namespace ex

template <typename T>
int any_nonfinite(Handle* h, int m, int n, const T* A, long lda, bool* out, hipStream_t st) {
  *out = false;
  if (m <= 0 || n <= 0) return RNPX_OK;
  RNPX_TRY(h->exbuf.ensure(64, h));
  unsigned* flag = (unsigned*)h->exbuf.p;
  RNPX_TRY(launch_zero_u32(flag, 1, st));
  const dim3 grid((unsigned)((m + 255) / 256), (unsigned)std::min(n, 2048));
  hipLaunchKernelGGL((ex::nonfinite_kernel<T>), grid, dim3(256), 0, st, m, n, A, lda, flag);
  RNPX_TRY(hipGetLastError());
  unsigned hv = 0;
  RNPX_TRY(hipMemcpyAsync(&hv, flag, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  RNPX_TRY(hipStreamSynchronize(st));
  *out = hv != 0;
  return RNPX_OK;
}

// LastNonzeroColumnLimit (which 0) or LastNonzeroRowLimit (which 1) of the
// m x n device matrix a, read back to the host.
template <typename T>
int last_nonzero_limit(Handle* h, int which, int m, int n, const T* a, long lda, int* out) {
  *out = 0;
  if (m <= 0 || n <= 0) return RNPX_OK;
  RNPX_TRY(h->exbuf.ensure(64, h));
  int* dv = (int*)h->exbuf.p + 1;
  RNPX_TRY(launch_zero_u32((unsigned*)dv, 1, h->stream));
  const int grid = which == 0 ? std::max(1, std::min(n, 2 * h->num_cus)) : std::max(1, std::min((m + 255) / 256, 2 * h->num_cus));
  hipLaunchKernelGGL((ex::limit_kernel<T>), dim3(grid), dim3(256), 0, h->stream, which, m, n, a, lda, dv);
  RNPX_TRY(hipGetLastError());
  RNPX_TRY(hipMemcpyAsync(out, dv, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  RNPX_TRY(hipStreamSynchronize(h->stream));
  return RNPX_OK;
}

// QR::Factor (QR.hpp:129-197) with the reference's blocking: nb = 32 (its
// Tuning default) or lwork / n when lwork is short, nbmin 2, crossover 128.
// lwork 0: the optimal workspace.
template <typename T>
int geqrf_exact(Handle* h, int m, int n, T* A, long lda, T* tau, size_t lwork) {
  const int k = std::min(m, n);
  if (k == 0) return RNPX_OK;
  hipStream_t st = h->stream;
  long nb = 32, nbmin = 2, nx = 0;
  const long ldwork = n;
  if (nb > 1 && nb < k) {
    nx = 128;
    if (nx < k) {
      const long iws = ldwork * nb;
      if (lwork > 0 && (long)lwork < iws) {
        nb = (long)lwork / ldwork;
        nbmin = 2;
      }
    }
  }
  const int ldt = 64;
  RNPX_TRY(h->exbuf.ensure(256 + sizeof(T) * ((size_t)ldt * ldt + (size_t)std::max(n, 1)), h));
  T* Tm = (T*)((char*)h->exbuf.p + 256);
  T* work = Tm + (size_t)ldt * ldt;
  long i = 0;
  if (nb >= nbmin && nb < k && nx < k) {
    for (i = 0; i + nx < k; i += nb) {
      const int ib = (int)std::min((long)k - i, nb);
      T* Ai = A + (size_t)i + (size_t)i * lda;
      hipLaunchKernelGGL((ex::unblocked_kernel<T>), dim3(1), dim3(ex::kNT), 0, st, (int)(m - i), ib, Ai, lda, tau + i,
                         work);
      RNPX_TRY(hipGetLastError());
      if (i + ib < n) {
        hipLaunchKernelGGL((ex::block_tr_kernel<T>), dim3(1), dim3(ex::kNT), 0, st, (int)(m - i), ib, (const T*)Ai, lda,
                           (const T*)(tau + i), Tm, (long)ldt);
        RNPX_TRY(hipGetLastError());
        // ApplyBlock's trims (Reflector.hpp:742-744)
        const int mc = (int)(m - i), nc = (int)(n - i - ib);
        int lastv = 0, lastc = 0;
        RNPX_CHECK(last_nonzero_limit<T>(h, 1, mc, ib, Ai, lda, &lastv));
        if (lastv < ib) lastv = ib;
        T* Ci = Ai + (size_t)ib * lda;
        RNPX_CHECK(last_nonzero_limit<T>(h, 0, lastv, nc, Ci, lda, &lastc));
        if (lastc > 0) {
          hipLaunchKernelGGL((ex::apply_block_kernel<T>), dim3((unsigned)((lastc + 3) / 4)), dim3(256), 0, st, mc, ib,
                             lastv, lastc, (const T*)Ai, lda, (const T*)Tm, (long)ldt, Ci, lda);
          RNPX_TRY(hipGetLastError());
        }
      }
    }
  }
  if (i < k) {
    hipLaunchKernelGGL((ex::unblocked_kernel<T>), dim3(1), dim3(ex::kNT), 0, st, (int)(m - i), (int)(n - i),
                       A + (size_t)i + (size_t)i * lda, lda, tau + i, work);
    RNPX_TRY(hipGetLastError());
  }
  h->exact_calls++;
  return RNPX_OK;
}

template int any_nonfinite<double>(Handle*, int, int, const double*, long, bool*, hipStream_t);
template int any_nonfinite<zd>(Handle*, int, int, const zd*, long, bool*, hipStream_t);
template int last_nonzero_limit<double>(Handle*, int, int, int, const double*, long, int*);
template int last_nonzero_limit<zd>(Handle*, int, int, int, const zd*, long, int*);
template int geqrf_exact<double>(Handle*, int, int, double*, long, double*, size_t);
template int geqrf_exact<zd>(Handle*, int, int, zd*, long, zd*, size_t);

}  // namespace rnpx
