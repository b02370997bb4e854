// rnpx_device.hpp -- device-side building blocks for the gfx950 QR kernels:
// the interleaved complex<double> type, scalar helpers shared by the real and
// complex instantiations, the FP64 MFMA wrapper and wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rnpx {

// Interleaved complex<double>, bit-compatible with std::complex<double>.
struct zd {
  double x, y;
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline zd mkz(double re, double im) { zd r; r.x = re; r.y = im; return r; }
__host__ __device__ inline zd operator+(zd a, zd b) { return mkz(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd operator-(zd a, zd b) { return mkz(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd operator-(zd a) { return mkz(-a.x, -a.y); }
__host__ __device__ inline zd operator*(zd a, zd b) { return mkz(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline zd operator*(zd a, double s) { return mkz(a.x * s, a.y * s); }
__host__ __device__ inline zd operator*(double s, zd a) { return mkz(a.x * s, a.y * s); }

// Scalar traits shared by both element types.
template <typename T> struct S;
template <> struct S<double> {
  static constexpr bool cplx = false;
  __host__ __device__ static inline double zero() { return 0.0; }
  __host__ __device__ static inline double one() { return 1.0; }
  __host__ __device__ static inline double conj(double a) { return a; }
  __host__ __device__ static inline double re(double a) { return a; }
  __host__ __device__ static inline double im(double) { return 0.0; }
  __host__ __device__ static inline double make(double r, double) { return r; }
  __host__ __device__ static inline double abs2(double a) { return a * a; }
  __host__ __device__ static inline bool is_zero(double a) { return a == 0.0; }
  // a * conj(b)
  __host__ __device__ static inline double mulc(double a, double b) { return a * b; }
  // conj(a) * b
  __host__ __device__ static inline double cmul(double a, double b) { return a * b; }
};
template <> struct S<zd> {
  static constexpr bool cplx = true;
  __host__ __device__ static inline zd zero() { return mkz(0.0, 0.0); }
  __host__ __device__ static inline zd one() { return mkz(1.0, 0.0); }
  __host__ __device__ static inline zd conj(zd a) { return mkz(a.x, -a.y); }
  __host__ __device__ static inline double re(zd a) { return a.x; }
  __host__ __device__ static inline double im(zd a) { return a.y; }
  __host__ __device__ static inline zd make(double r, double i) { return mkz(r, i); }
  __host__ __device__ static inline double abs2(zd a) { return a.x * a.x + a.y * a.y; }
  __host__ __device__ static inline bool is_zero(zd a) { return a.x == 0.0 && a.y == 0.0; }
  __host__ __device__ static inline zd mulc(zd a, zd b) { return mkz(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
  __host__ __device__ static inline zd cmul(zd a, zd b) { return mkz(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
};

// Complex division (1,0)/(d) with Smith's scaling (the role std::complex's
// operator/ plays at RNP/LA/Reflector.hpp:150).
__device__ inline zd zrecip(zd d) {
  if (fabs(d.y) <= fabs(d.x)) {
    double r = d.y / d.x;
    double t = 1.0 / (d.x + d.y * r);
    return mkz(t, -r * t);
  } else {
    double r = d.x / d.y;
    double t = 1.0 / (d.y + d.x * r);
    return mkz(r * t, -t);
  }
}
__device__ inline double recip_t(double d) { return 1.0 / d; }
__device__ inline zd recip_t(zd d) { return zrecip(d); }

// Traits<real>::hypot3 (RNP/Types.hpp:73-92).
__device__ inline double hypot3(double x, double y, double z) {
  double xa = fabs(x), ya = fabs(y), za = fabs(z);
  double w = fmax(xa, fmax(ya, za));
  if (w == 0.0) return xa + ya + za;
  xa /= w; ya /= w; za /= w;
  return w * sqrt(xa * xa + ya * ya + za * za);
}

// Shuffles of element types across a 64-lane wavefront.
__device__ inline double shfl_xor_t(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ inline zd shfl_xor_t(zd v, int m) { return mkz(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)); }
__device__ inline double shfl_t(double v, int src) { return __shfl(v, src, 64); }
__device__ inline zd shfl_t(zd v, int src) { return mkz(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }

// Component-wise select (a ternary on the struct type would keep arrays of
// it out of registers).
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ zd sel(bool c, zd a, zd b) { return mkz(c ? a.x : b.x, c ? a.y : b.y); }

// ---- VALU cross-lane exchange (no LDS round trip) -------------------------
__device__ __forceinline__ unsigned lo32(double d) { return (unsigned)(unsigned long long)__double_as_longlong(d); }
__device__ __forceinline__ unsigned hi32(double d) {
  return (unsigned)((unsigned long long)__double_as_longlong(d) >> 32);
}
__device__ __forceinline__ double mkd(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One butterfly step of a reduce-scatter: x holds the lane's partials for the
// lower half of the current index block, y for the upper half.  Lanes with
// bit OFF clear keep the lower half, the others the upper half; the result
// is own + partner for the kept half.  OFF 32/16 use gfx950's
// v_permlane32_swap / v_permlane16_swap, OFF 8/4/2/1 DPP within a row
// (row_mirror / row_half_mirror pair the two halves of a 16 / 8 lane group,
// quad_perm the xor-2 / xor-1 partners).  Calling it with x == y adds the
// partner's value to every lane (a full step).
template <int OFF>
__device__ __forceinline__ double bfly(double x, double y, int lane) {
  if constexpr (OFF == 32 || OFF == 16) {
    unsigned xl = lo32(x), xh = hi32(x), yl = lo32(y), yh = hi32(y);
    if constexpr (OFF == 32) {
      auto l = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
      auto h = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
      return mkd(l[0], h[0]) + mkd(l[1], h[1]);
    } else {
      auto l = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
      auto h = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
      return mkd(l[0], h[0]) + mkd(l[1], h[1]);
    }
  } else {
    constexpr int ctrl = (OFF == 8) ? 0x140 : (OFF == 4) ? 0x141 : (OFF == 2) ? 0x4E : 0xB1;
    const bool upper = (lane & OFF) != 0;
    const double send = upper ? x : y;
    const double keep = upper ? y : x;
    const unsigned rl = __builtin_amdgcn_update_dpp(0u, lo32(send), ctrl, 0xF, 0xF, false);
    const unsigned rh = __builtin_amdgcn_update_dpp(0u, hi32(send), ctrl, 0xF, 0xF, false);
    return keep + mkd(rl, rh);
  }
}
template <int OFF>
__device__ __forceinline__ zd bfly(zd x, zd y, int lane) {
  return mkz(bfly<OFF>(x.x, y.x, lane), bfly<OFF>(x.y, y.y, lane));
}

// Broadcast lane `src` (a compile-time or wave-uniform index) to SGPRs.
__device__ __forceinline__ double readlane_t(double v, int src) {
  return mkd(__builtin_amdgcn_readlane(lo32(v), src), __builtin_amdgcn_readlane(hi32(v), src));
}
__device__ __forceinline__ zd readlane_t(zd v, int src) { return mkz(readlane_t(v.x, src), readlane_t(v.y, src)); }

// Reduce-scatter of N per-lane values across the wavefront: afterwards lane
// l holds, in v[0 .. max(N/64,1)), the wave-wide sums of indices
// base(l) .. base(l)+N/64-1 with base(l) = (l * N/64) mod N (for N < 64 the
// upper lanes duplicate the lower ones).  About N butterfly steps per lane,
// all VALU.  Template recursion keeps every index a compile-time constant.
template <typename T, int N, int OFF>
__device__ __forceinline__ void rs_full(T (&v)[N], int lane) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = bfly<OFF>(v[i], v[i], lane);
}
template <typename T, int N, int H, int OFF>
__device__ __forceinline__ void rs_halve(T (&v)[N], int lane) {
  if constexpr (H >= 1 && OFF >= 1) {
#pragma unroll
    for (int i = 0; i < H; ++i) v[i] = bfly<OFF>(v[i], v[i + H], lane);
    rs_halve<T, N, H / 2, OFF / 2>(v, lane);
  }
}
template <typename T, int N>
__device__ __forceinline__ void wave_reduce_scatter(T (&v)[N], int lane) {
  static_assert(N == 8 || N == 16 || N == 32 || N == 64 || N == 128, "unsupported width");
  if constexpr (N >= 64) {
    rs_halve<T, N, N / 2, 32>(v, lane);
  } else if constexpr (N == 32) {
    rs_full<T, N, 32>(v, lane);
    rs_halve<T, N, 16, 16>(v, lane);
  } else if constexpr (N == 16) {
    rs_full<T, N, 32>(v, lane);
    rs_full<T, N, 16>(v, lane);
    rs_halve<T, N, 8, 8>(v, lane);
  } else {
    rs_full<T, N, 32>(v, lane);
    rs_full<T, N, 16>(v, lane);
    rs_full<T, N, 8>(v, lane);
    rs_halve<T, N, 4, 4>(v, lane);
  }
}

// Sum across the 64-lane wavefront (result in every lane).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// One FP64 MFMA: D(16x16) += A(16x4) * B(4x16).
// Operand layout (measured on gfx950): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator item r of lane l holds D[l/16 + 4*r][l%16].
__device__ inline f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace rnpx
