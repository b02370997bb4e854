// rnpx_panel.hip -- the panel kernel: Householder QR of one mp-by-nbp panel
// (the blocked-QR panel step, QR::Factor_unblocked at RNP/LA/QR.hpp:78-102,
// with Reflector::Generate at RNP/LA/Reflector.hpp:117-158 and the rank-1
// Reflector::Apply at Reflector.hpp:225-230 for every column).
//
// Layout: G workgroups of 256 threads; workgroup g owns rows
// [g*RPT*256, (g+1)*RPT*256) of the panel, thread t of it rows g*RPT*256 +
// q*256 + t (q < RPT), and keeps those rows of all NB columns in registers
// for the whole factorization.  Column c costs one grid-wide reduction:
// every thread forms conj(x_r) * a_rj for all j (x = column c below the
// diagonal); slot c of that vector is |x|^2, slots j > c are the dot
// products the rank-1 update needs, slots j < c are the Gram entries
// V(:,j)^H x the T factor needs.  A wave reduce-scatter, a cross-wave LDS
// sum and one bounded-spin grid barrier later every workgroup reads the G
// partial vectors and sums them in a fixed order (bitwise identical in every
// workgroup, and independent of what else runs on the device).
//
// Slow path: when |x|^2 is outside [2^-600, 2^600] (zero, tiny, huge or not
// finite) the fused
// dot products could under/overflow, so the column is redone in three rounds
// -- max |x|, a power-of-two-scaled sum of squares, then dot products with v
// itself -- which follows RNP's scaled Norm2 and its rescale loop
// (Reflector.hpp:139-152; two scalings by 2^971 when |beta| < 2^-971).
//
// Outputs: the panel overwritten by R and V (RNP's packed layout), tau,
// a clean copy of V (unit diagonal, zeros above) for the trailing-update
// GEMMs, and the Gram matrix of V: gram(j,c) = V(:,j)^H V(:,c) for j < c with
// tau_c on the diagonal, from which the T-factor kernel builds T.
#include "rnpx_kernels.hpp"

namespace rnpx {

template <typename T, int N>
__device__ __forceinline__ T select_col(const T (&row)[N], int c) {
  T x = S<T>::zero();
#pragma unroll
  for (int j = 0; j < N; ++j) x = sel(j == c, row[j], x);
  return x;
}

// Wave reduce-scatter of the thread's partial products conj(s_q) * a_qj over
// the rows it owns (rows with use[q] false contribute nothing), then a
// cross-wave LDS sum; the workgroup partial vector lands in slot.
template <typename T, int NB, int RPT, int CH>
__device__ __forceinline__ void reduce_products(const T (&a)[RPT][NB], const T (&s)[RPT], const bool (&use)[RPT],
                                                T (*red)[NB], T* slot, int lane, int wave, int tid) {
  constexpr int NCH = NB / CH;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    T v[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      T acc = S<T>::zero();
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if (use[q]) acc = acc + S<T>::cmul(s[q], a[q][ch * CH + i]);
      v[i] = acc;
    }
    wave_reduce_scatter<T, CH>(v, lane);
    constexpr int PER = (CH >= 64) ? CH / 64 : 1;
    const int base = (CH >= 64) ? lane * PER : lane;
    if (lane < ((CH < 64) ? CH : 64)) {
#pragma unroll
      for (int u = 0; u < PER; ++u) red[wave][ch * CH + base + u] = v[u];
    }
  }
  __syncthreads();
  for (int j = tid; j < NB; j += 256) slot[j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
}

template <typename T> __device__ inline double absmax_t(T v);
template <> __device__ inline double absmax_t<double>(double v) { return fabs(v); }
template <> __device__ inline double absmax_t<zd>(zd v) { return fmax(fabs(v.x), fabs(v.y)); }

template <typename T> __device__ inline T ldexp_t(T v, int e);
template <> __device__ inline double ldexp_t<double>(double v, int e) { return ldexp(v, e); }
template <> __device__ inline zd ldexp_t<zd>(zd v, int e) { return mkz(ldexp(v.x, e), ldexp(v.y, e)); }

// Chunk width of the reduce-scatter (values per lane held at once).
template <typename T> struct RedChunk { static constexpr int value = 64; };
template <> struct RedChunk<zd> { static constexpr int value = 32; };

// ONE: a single workgroup (at most 256 RPT rows); the per-column partial
// vector and the diagonal row stay in LDS instead of crossing global memory.
template <typename T, int NB, int RPT, bool ONE>
__global__ __launch_bounds__(256) void panel_kernel(PanelArgs<T> p) {
  // Gated launch: the whole grid reads the same word and returns together,
  // before any barrier.
  if (p.run_if && *p.run_if == 0u) return;
  constexpr int CH = (RedChunk<T>::value < NB) ? RedChunk<T>::value : NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x, G = ONE ? 1 : (int)gridDim.x;
  const int mp = p.mp, nbp = p.nbp;
  const int kp = (mp < nbp) ? mp : nbp;

  __shared__ T red[4][NB];
  __shared__ double dsh[4];
  __shared__ T lpart[ONE ? 2 * NB : 1];
  __shared__ T lrow[ONE ? 2 * NB : 1];
  T* const part = ONE ? lpart : p.part;
  T* const rowbuf = ONE ? lrow : p.rowbuf;

  T a[RPT][NB];
  int row[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    row[q] = (g * RPT + q) * 256 + tid;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      a[q][j] = S<T>::zero();
      if (row[q] < mp && j < nbp) a[q][j] = p.A[(size_t)row[q] + (size_t)j * p.lda];
    }
  }

  unsigned epoch = 0;
  // After the barrier: all 256 threads fetch the G x NB partials at once
  // (one memory latency instead of G) -- thread t sums groups t/NB,
  // t/NB + SUB, ... of index t%NB -- then every lane j < NB adds the SUB
  // sub-sums in order.  The order is fixed for a given G.
  constexpr int SUB = 256 / NB;
  __shared__ T gred[SUB * NB];
  auto gather_all = [&](T (&Sj)[(NB + 63) / 64]) {
    {
      // MAXQ loads are issued back to back before any is consumed.
      constexpr int MAXQ = 8;
      const int j = tid % NB, sub = tid / NB;
      const T* base = part + (size_t)(epoch & 1u) * (size_t)G * NB + j;
      T s = S<T>::zero();
      for (int g0 = sub; g0 < G; g0 += SUB * MAXQ) {
        T buf[MAXQ];
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
          const int gg = g0 + q * SUB;
          buf[q] = S<T>::zero();
          if (gg < G) buf[q] = base[(size_t)gg * NB];
        }
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) s = s + buf[q];
      }
      gred[sub * NB + j] = s;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < (NB + 63) / 64; ++u) {
      const int j = lane + 64 * u;
      T s = S<T>::zero();
      if (j < NB) {
#pragma unroll
        for (int q = 0; q < SUB; ++q) s = s + gred[q * NB + j];
      }
      Sj[u] = s;
    }
  };
  auto barrier = [&]() {
    ++epoch;
    if (ONE)
      __syncthreads();  // one workgroup: its own partials, no grid hand-off
    else
      grid_barrier(p.counter, p.abortf, epoch * (unsigned)G);
    --epoch;  // gather() reads the slot of the epoch just completed
  };

  for (int c = 0; c < kp; ++c) {
    // ---- round 1: fused |x|^2, dot products and Gram entries ----
    T x[RPT];
    bool below[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      x[q] = select_col<T, NB>(a[q], c);
      below[q] = (row[q] > c) && (row[q] < mp);
    }
    {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (row[q] == c) {
#pragma unroll
          for (int j = 0; j < NB; ++j) rowbuf[(c & 1) * NB + j] = a[q][j];
        }
      }
      reduce_products<T, NB, RPT, CH>(a, x, below, red,
                                      part + (size_t)(epoch & 1u) * (size_t)G * NB + (size_t)g * NB, lane, wave,
                                      tid);
    }
    barrier();
    // Every wave ends up with lane j < NB holding index j (and j+64).
    T Sj[(NB + 63) / 64], rcj[(NB + 63) / 64];
#pragma unroll
    for (int u = 0; u < (NB + 63) / 64; ++u) {
      const int j = lane + 64 * u;
      rcj[u] = S<T>::zero();
      if (j < NB) rcj[u] = rowbuf[(c & 1) * NB + j];
    }
    gather_all(Sj);
    ++epoch;
    const int cu = c >> 6, cl = c & 63;
    T Sc = Sj[0], alpha = rcj[0];
#pragma unroll
    for (int u = 1; u < (NB + 63) / 64; ++u) {
      Sc = sel(cu == u, Sj[u], Sc);
      alpha = sel(cu == u, rcj[u], alpha);
    }
    Sc = readlane_t(Sc, cl);  // c is wave-uniform
    alpha = readlane_t(alpha, cl);
    const double ssq = S<T>::re(Sc);
    const double ar = S<T>::re(alpha), ai = S<T>::im(alpha);

    // |x|^2 == 0 can be underflow of a nonzero x: only the slow path's max
    // round can tell an all-zero column (tau = 0) from a tiny one.
    const bool fast = (ssq >= 0x1p-600 && ssq <= 0x1p600);
    T tau, inv;
    double beta;
    bool identity;
    int escale = 0;  // v = ldexp(x, escale) * inv on the slow path
    if (fast) {
      const double xnorm = sqrt(ssq);
      identity = (xnorm == 0.0 && ai == 0.0);
      beta = hypot3(ar, ai, xnorm);
      if (ar > 0) beta = -beta;
      tau = S<T>::make((beta - ar) / beta, (-ai) / beta);
      inv = recip_t(S<T>::make(ar - beta, ai));
    } else {
      // ---- slow round A: max |x| ----
      double mx = 0.0;
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if (below[q]) mx = fmax(mx, absmax_t(x[q]));
      mx = wave_max(mx);
      if (lane == 0) dsh[wave] = mx;
      __syncthreads();
      {
        double* slot = reinterpret_cast<double*>(part + (size_t)(epoch & 1u) * (size_t)G * NB) + g;
        if (tid == 0) *slot = fmax(fmax(dsh[0], dsh[1]), fmax(dsh[2], dsh[3]));
      }
      barrier();
      double M = 0.0;
      {
        const double* slot = reinterpret_cast<const double*>(part + (size_t)(epoch & 1u) * (size_t)G * NB);
        for (int gg = 0; gg < G; ++gg) M = fmax(M, slot[gg]);
      }
      ++epoch;
      int e = 0;
      if (M > 0.0 && M <= 1.7976931348623157e308) e = ilogb(M);
      // ---- slow round B: sum of (x / 2^e)^2 ----
      double ps = 0.0;
#pragma unroll
      for (int q = 0; q < RPT; ++q)
        if (below[q]) ps += S<T>::abs2(ldexp_t(x[q], -e));
      ps = wave_sum(ps);
      __syncthreads();
      if (lane == 0) dsh[wave] = ps;
      __syncthreads();
      {
        double* slot = reinterpret_cast<double*>(part + (size_t)(epoch & 1u) * (size_t)G * NB) + g;
        if (tid == 0) *slot = (dsh[0] + dsh[1]) + (dsh[2] + dsh[3]);
      }
      barrier();
      double ss = 0.0;
      {
        const double* slot = reinterpret_cast<const double*>(part + (size_t)(epoch & 1u) * (size_t)G * NB);
        for (int gg = 0; gg < G; ++gg) ss += slot[gg];
      }
      ++epoch;
      double xnorm = ldexp(sqrt(ss), e);
      identity = (xnorm == 0.0 && ai == 0.0);
      beta = hypot3(ar, ai, xnorm);
      if (ar > 0) beta = -beta;
      double sar = ar, sai = ai;
      const double safmin = 0x1p-971;
      int knt = 0;
      if (!identity && fabs(beta) < safmin && beta != 0.0) {
        knt = 2;  // two scalings by 2^971 (Reflector.hpp:139-144)
        escale = 1942;
        sar = ldexp(ar, 1942);
        sai = ldexp(ai, 1942);
        xnorm = ldexp(sqrt(ss), e + 1942);
        beta = hypot3(sar, sai, xnorm);
        if (sar > 0) beta = -beta;
      }
      tau = S<T>::make((beta - sar) / beta, (-sai) / beta);
      inv = recip_t(S<T>::make(sar - beta, sai));
      for (int k = 0; k < knt; ++k) beta *= safmin;
      // ---- slow round C: dot products with v itself ----
      if (!identity) {
        T vq[RPT];
#pragma unroll
        for (int q = 0; q < RPT; ++q) vq[q] = ldexp_t(x[q], escale) * inv;
        __syncthreads();
        reduce_products<T, NB, RPT, CH>(a, vq, below, red,
                                        part + (size_t)(epoch & 1u) * (size_t)G * NB + (size_t)g * NB, lane, wave,
                                        tid);
        barrier();
        __syncthreads();
        gather_all(Sj);
        ++epoch;
      }
    }

    // ---- w_j (times conj(tau)) and Gram entries, lane j of every wave ----
    const T ctau = S<T>::conj(tau);
    T ctw[(NB + 63) / 64];
#pragma unroll
    for (int u = 0; u < (NB + 63) / 64; ++u) {
      const int j = lane + 64 * u;
      // fast: S_j = sum conj(x_r) a_rj ; slow: S_j = sum conj(v_r) a_rj
      const T dotv = sel(fast, S<T>::conj(inv) * Sj[u], Sj[u]);
      const T w = rcj[u] + dotv;
      ctw[u] = sel(j > c && j < nbp && !identity, ctau * w, S<T>::zero());
      if (g == 0 && wave == 0 && j < c) {
        // gram(j,c) = conj(V(c,j)) + sum_{r>c} conj(V(r,j)) v_r
        p.gram[(size_t)j + (size_t)c * NB] = sel(identity, S<T>::zero(), S<T>::conj(rcj[u]) + S<T>::conj(dotv));
      }
    }
    if (g == 0 && tid == 0) {
      const T tt = sel(identity, S<T>::zero(), tau);
      p.tau[c] = tt;
      p.gram[(size_t)c + (size_t)c * NB] = tt;
    }

    // ---- apply H^H to the rows this thread owns ----
    // a(r,j) -= coef_r * conj(tau) w_j for j > c (coef = 1 on the diagonal
    // row, v_r below it, 0 above), column c becomes beta / v.  The
    // conj(tau) w_j come from lane j by v_readlane (wave-uniform, SGPRs).
    if (!identity) {
      T coef[RPT], newc[RPT];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const T vq = ldexp_t(x[q], escale) * inv;
        coef[q] = sel(row[q] == c, S<T>::one(), sel(below[q], vq, S<T>::zero()));
        newc[q] = sel(row[q] == c, S<T>::make(beta, 0.0), sel(below[q], vq, x[q]));
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const T cw = readlane_t(ctw[j >> 6], j & 63);
#pragma unroll
        for (int q = 0; q < RPT; ++q) a[q][j] = sel(j == c, newc[q], a[q][j] - coef[q] * cw);
      }
    }
    // No barrier here: red/gred are next written after the next grid
    // barrier, whose __syncthreads orders them after these reads.
  }

  // ---- write back R/V and the clean V ----
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int r = row[q];
    if (r < mp) {
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < nbp) p.A[(size_t)r + (size_t)j * p.lda] = a[q][j];
        if (j < kp && p.V)
          p.V[(size_t)r + (size_t)j * p.ldv] = sel(r > j, a[q][j], sel(r == j, S<T>::one(), S<T>::zero()));
      }
    }
  }
}

template <typename T, int NB, int RPT>
static hipError_t launch_panel_t(const PanelArgs<T>& a, int G, hipStream_t st) {
  if (G == 1)
    hipLaunchKernelGGL((panel_kernel<T, NB, RPT, true>), dim3(1), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((panel_kernel<T, NB, RPT, false>), dim3(G), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_panel(const PanelArgs<T>& a, int nb, int rpt, hipStream_t st) {
  const int rows_per_wg = 256 * rpt;
  const int G = (a.mp + rows_per_wg - 1) / rows_per_wg;
  if (G < 1 || G > kMaxPanelGroups) return hipErrorInvalidValue;
  if (a.nbp > nb || a.nbp < 1) return hipErrorInvalidValue;
  if constexpr (S<T>::cplx) {
    if (nb == 32 && rpt == 1) return launch_panel_t<T, 32, 1>(a, G, st);
    if (nb == 32 && rpt == 2) return launch_panel_t<T, 32, 2>(a, G, st);
    if (nb == 64 && rpt == 1) return launch_panel_t<T, 64, 1>(a, G, st);
  } else {
    if (nb == 32 && rpt == 1) return launch_panel_t<T, 32, 1>(a, G, st);
    if (nb == 32 && rpt == 2) return launch_panel_t<T, 32, 2>(a, G, st);
    if (nb == 32 && rpt == 4) return launch_panel_t<T, 32, 4>(a, G, st);
    if (nb == 64 && rpt == 1) return launch_panel_t<T, 64, 1>(a, G, st);
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_panel<double>(const PanelArgs<double>&, int, int, hipStream_t);
template hipError_t launch_panel<zd>(const PanelArgs<zd>&, int, int, hipStream_t);

}  // namespace rnpx
