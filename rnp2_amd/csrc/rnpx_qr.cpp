// rnpx_qr.cpp -- host orchestration of the gfx950 QR and the C ABI of
// include/rnpx.h.  Factor / MultQ / GenerateQ follow RNP's blocked
// algorithms (RNP/LA/QR.hpp:129-197, 337-429, 509-570): a panel kernel, the
// T-factor kernel and three MFMA GEMMs per block, all on one HIP stream.
#include "rnpx_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace rnpx {

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------
hipError_t Buffer::ensure(size_t b, Handle* h) {
  if (b <= bytes) return hipSuccess;
  if (h && h->budget && h->held_bytes() - bytes + b > h->budget) return hipErrorOutOfMemory;
  if (p) {
    hipError_t e = hipStreamSynchronize(h ? h->stream : nullptr);
    if (e != hipSuccess) return e;
    hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  size_t want = b + (b >> 3);  // grow with some headroom
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = want;
  return hipSuccess;
}
void Buffer::release() {
  if (p) hipFree(p);
  p = nullptr;
  bytes = 0;
}

size_t Handle::held_bytes() const {
  size_t s = 0;
  for (const Buffer* b : {&vbuf, &vfull, &gram, &tbuf, &wbuf, &zbuf, &part, &ppart, &rowbuf, &sync, &flag, &misc, &stage1, &stage2,
                          &stage3, &vbuf2, &gram2, &tbuf2, &lsbuf, &lsnrm})
    s += b->bytes;
  return s;
}

hipEvent_t Handle::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
int Handle::prof_begin(hipEvent_t* a, hipStream_t s) {
  *a = nullptr;
  if (!prof) return 0;
  *a = take_event();
  if (*a) hipEventRecord(*a, s ? s : stream);
  return 0;
}
void Handle::prof_end(int kid, hipEvent_t a, double flops, double bytes, hipStream_t s) {
  if (!prof || !a) return;
  hipEvent_t b = take_event();
  if (!b) return;
  hipEventRecord(b, s ? s : stream);
  recs.push_back(ProfRec{kid, a, b, flops, bytes});
}
int Handle::ensure_pipeline() {
  if (pstream) return RNPX_OK;
  hipError_t e = hipSuccess;
  if (ext_pstream) {
    pstream = ext_pstream;
  } else {
    e = hipStreamCreateWithFlags(&pstream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_status(e);
  }
  for (hipEvent_t* ev : {&ev_pan[0], &ev_pan[1], &ev_next, &ev_start}) {
    e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e != hipSuccess) return hip_status(e);
  }
  return RNPX_OK;
}

Handle::~Handle() {
  if (stream) hipStreamSynchronize(stream);
  if (batch && batch_free) batch_free(batch);
  for (auto& r : recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : event_pool) hipEventDestroy(e);
  if (pstream) hipStreamSynchronize(pstream);
  for (hipEvent_t ev : {ev_pan[0], ev_pan[1], ev_next, ev_start})
    if (ev) hipEventDestroy(ev);
  if (pstream && pstream != ext_pstream) hipStreamDestroy(pstream);
  for (Buffer* b : {&vbuf, &vfull, &gram, &tbuf, &wbuf, &zbuf, &part, &ppart, &rowbuf, &sync, &flag, &misc, &stage1, &stage2,
                          &stage3, &vbuf2, &gram2, &tbuf2, &lsbuf, &lsnrm})
    b->release();
  if (own_stream && own) hipStreamDestroy(own);
}

int hip_status(hipError_t e) {
  if (e == hipSuccess) return RNPX_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return RNPX_ERR_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return RNPX_ERR_NOGPU;
  if (e == hipErrorInvalidValue) return RNPX_ERR_UNSUPPORTED;
  return RNPX_ERR_HIP;
}

#define RNPX_TRY(x)                             \
  do {                                          \
    hipError_t _e = (x);                        \
    if (_e != hipSuccess) return hip_status(_e); \
  } while (0)

int Handle::check_gpu() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return RNPX_ERR_NOGPU;
  return RNPX_OK;
}

int create_handle(Handle** out, size_t budget) {
  int st = Handle::check_gpu();
  if (st) return st;
  Handle* h = new Handle();
  hipGetDevice(&h->device);
  hipDeviceGetAttribute(&h->num_cus, hipDeviceAttributeMultiprocessorCount, h->device);
  if (h->num_cus < 1) h->num_cus = 1;
  hipError_t e = hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return hip_status(e);
  }
  h->own_stream = true;
  h->stream = h->own;
  h->budget = budget;
  e = h->flag.ensure(64, h);
  if (e == hipSuccess) e = hipMemsetAsync(h->flag.p, 0, 64, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    delete h;
    return hip_status(e);
  }
  *out = h;
  return RNPX_OK;
}

// Thread-local default handle for the drop-in entry points.
Handle* default_handle(int* status) {
  thread_local std::unique_ptr<Handle> h;
  if (!h) {
    Handle* p = nullptr;
    *status = create_handle(&p, 0);
    if (*status) return nullptr;
    h.reset(p);
  }
  *status = RNPX_OK;
  return h.get();
}

// ---------------------------------------------------------------------------
// Tuning
// ---------------------------------------------------------------------------
template <typename T>
static int pick_nb(const Handle* h, int m, int n) {
  int nb = h->tune.nb;
  if (nb == 32 || nb == 64 || (!S<T>::cplx && nb == 128)) {
    // explicit choice
  } else {
    nb = 64;
  }
  // The T-factor kernel holds NB x NB blocks in LDS: 64 is the largest built.
  if (nb > 64) nb = 64;
  (void)m;
  (void)n;
  return nb;
}

static int pick_rpt(const Handle* h, int mp, int nb, bool cplx) {
  int rpt = h->tune.panel_rpt;
  // Built register-resident configurations (no spills): real NB=32 with up
  // to 4 rows per thread, NB=64 with 1; complex NB=32/64 with 1.
  const int maxrpt = cplx ? 1 : (nb <= 32 ? 4 : 1);
  if (rpt < 1 || rpt > maxrpt || (rpt & (rpt - 1))) {
    // default: keep the workgroup count near 16-32
    rpt = 1;
    while (rpt < maxrpt && (mp + 256 * rpt - 1) / (256 * rpt) > 32) rpt *= 2;
  }
  while ((mp + 256 * rpt - 1) / (256 * rpt) > kMaxPanelGroups && rpt < maxrpt) rpt *= 2;
  return rpt;
}

static int pick_ksplit(const Handle* h, int rows, int cols_tiles) {
  if (h->tune.ksplit > 0) return h->tune.ksplit;
  int want = 512 / std::max(1, cols_tiles);
  int maxs = std::max(1, rows / 512);
  return std::max(1, std::min(want, std::min(maxs, 32)));
}

// ---------------------------------------------------------------------------
// Block reflector application (the larfb analogue, Reflector.hpp:725-815)
// with a clean V (unit diagonal, zeros above) and T on the device.
//   side L: C(mc x nc) = op(H) C,   H = I - V T V^H, V is mc x k
//   side R: C(mc x nc) = C op(H),   V is nc x k
// op = H for trans 'N', H^H for trans 'C'.
// ---------------------------------------------------------------------------
template <typename T>
static int apply_block_dev(Handle* h, char side, char trans, int mc, int nc, int k, const T* V, long ldv,
                           const T* Tm, long ldt, T* C, long ldc) {
  if (mc <= 0 || nc <= 0 || k <= 0) return RNPX_OK;
  hipStream_t st = h->stream;
  const T one = S<T>::one(), zero = S<T>::zero(), mone = S<T>::zero() - S<T>::one();
  if (side == 'L') {
    // W = V^H C (k x nc), Z = op(T) W, C -= V Z
    const int tiles = (nc + 127) / 128;
    int ks = pick_ksplit(h, mc, tiles);
    RNPX_TRY(h->wbuf.ensure(sizeof(T) * (size_t)k * nc, h));
    RNPX_TRY(h->zbuf.ensure(sizeof(T) * (size_t)k * nc, h));
    if (ks > 1) RNPX_TRY(h->part.ensure(sizeof(T) * (size_t)ks * k * nc, h));
    T* W = (T*)h->wbuf.p;
    T* Z = (T*)h->zbuf.p;
    const double fl = (S<T>::cplx ? 8.0 : 2.0) * (double)mc * nc * k;
    const double esz = sizeof(T);
    hipEvent_t ev;
    // Mixed precision (real only): the two big products on FP32 MFMA.
    bool f32 = false;
    if constexpr (!S<T>::cplx) f32 = (h->update_prec == 1);
    h->prof_begin(&ev);
    if constexpr (!S<T>::cplx) {
      if (f32) RNPX_TRY(gemm32('C', 'N', k, nc, mc, 1.0, V, ldv, C, ldc, 0.0, W, k, ks, (T*)h->part.p, st));
    }
    if (!f32) RNPX_TRY(gemm<T>('C', 'N', k, nc, mc, one, V, ldv, C, ldc, zero, W, k, ks, (T*)h->part.p, st));
    h->prof_end(K_WGEMM, ev, fl, esz * ((double)mc * nc + (double)mc * k + (double)k * nc));
    RNPX_TRY(gemm<T>(trans == 'N' ? 'N' : 'C', 'N', k, nc, k, one, Tm, ldt, W, k, zero, Z, k, 1, nullptr, st));
    h->prof_begin(&ev);
    if constexpr (!S<T>::cplx) {
      if (f32) RNPX_TRY(gemm32('N', 'N', mc, nc, k, -1.0, V, ldv, Z, k, 1.0, C, ldc, 1, nullptr, st));
    }
    if (!f32) RNPX_TRY(gemm<T>('N', 'N', mc, nc, k, mone, V, ldv, Z, k, one, C, ldc, 1, nullptr, st));
    h->prof_end(K_UPDATE, ev, fl, esz * (2.0 * mc * nc + (double)mc * k + (double)k * nc));
  } else {
    // W = C V (mc x k), Z = W op(T), C -= Z V^H
    const int tiles = (mc + 127) / 128;
    int ks = pick_ksplit(h, nc, tiles);
    RNPX_TRY(h->wbuf.ensure(sizeof(T) * (size_t)k * mc, h));
    RNPX_TRY(h->zbuf.ensure(sizeof(T) * (size_t)k * mc, h));
    if (ks > 1) RNPX_TRY(h->part.ensure(sizeof(T) * (size_t)ks * k * mc, h));
    T* W = (T*)h->wbuf.p;
    T* Z = (T*)h->zbuf.p;
    RNPX_TRY(gemm<T>('N', 'N', mc, k, nc, one, C, ldc, V, ldv, zero, W, mc, ks, (T*)h->part.p, st));
    RNPX_TRY(gemm<T>('N', trans == 'N' ? 'N' : 'C', mc, k, k, one, W, mc, Tm, ldt, zero, Z, mc, 1, nullptr, st));
    RNPX_TRY(gemm<T>('N', 'C', mc, nc, k, mone, Z, mc, V, ldv, one, C, ldc, 1, nullptr, st));
  }
  return RNPX_OK;
}

// T for k <= nb reflectors held in a clean V (n x k): Gram then T kernel.
template <typename T>
static int tfactor_from_v(Handle* h, int n, int k, int nb, const T* V, long ldv, const T* tau, T* Tm, int ldt) {
  const int nchunks = std::max(1, std::min(64, n / 256));
  RNPX_TRY(h->gram.ensure(sizeof(T) * (size_t)nb * nb, h));
  RNPX_TRY(h->misc.ensure(sizeof(T) * (size_t)nchunks * nb * nb, h));
  RNPX_TRY(launch_gram<T>(n, k, V, ldv, tau, (T*)h->gram.p, nb, (T*)h->misc.p, nchunks, h->stream));
  RNPX_TRY(launch_tfactor<T>((const T*)h->gram.p, nb, k, nb, Tm, ldt, h->stream));
  return RNPX_OK;
}

// ---------------------------------------------------------------------------
// Factor (QR.hpp:129-197) on device memory.
// ---------------------------------------------------------------------------
// Whether panel p (rows [i, m), ib columns) takes the CholeskyQR2 +
// reconstruction kernel; G and rows per workgroup in *g / *rw.
template <typename T>
static bool use_cqr(const Handle* h, int mp, int ib, int nb, int* g, int* rw) {
  if constexpr (S<T>::cplx) {
    return false;
  } else {
    if (h->tune.panel == 1 || nb != 64 || ib != 64) return false;
    *g = cqr_panel_groups(mp, std::min(kCqrMaxGroups, h->num_cus), rw);
    return *g > 0;
  }
}

// One panel step.  sync holds, per panel p: [1 + p] the column-wise kernel's
// barrier counter, [1 + npan + p] the CholeskyQR2 kernel's fail flag, and
// kCqrSyncWords words at 1 + 2 npan + kCqrSyncWords p for its reductions.  With the CholeskyQR2 kernel the column-wise kernel is
// launched behind it gated on the fail flag; *gated receives that flag's
// address (the T-factor kernel is gated the same way), or null.
template <typename T>
static int launch_panel_step(Handle* h, int p, int npan, int i, int ib, int m, int nb, T* A, long lda, T* tau, T* V,
                             T* gram, T* Tm, hipStream_t st, const unsigned** gated) {
  const int mp = m - i;
  unsigned* sync = (unsigned*)h->sync.p;
  *gated = nullptr;
  hipEvent_t ev;
  h->prof_begin(&ev, st);
  int G = 0, rw = 0;
  if (use_cqr<T>(h, mp, ib, nb, &G, &rw)) {
    if constexpr (!S<T>::cplx) {
      CqrArgs ca;
      ca.mp = mp;
      ca.A = A + (size_t)i + (size_t)i * lda;
      ca.lda = lda;
      ca.tau = tau + i;
      ca.V = V;
      ca.ldv = mp;
      ca.T = Tm;
      ca.part = (double*)h->cqr_part.p;
      ca.gpart = (double*)h->cqr_gsum.p + 2 * 2560;
      ca.gsum = (double*)h->cqr_gsum.p;
      ca.scratch = (double*)h->cqr_scratch.p;
      ca.sync = sync + 1 + 2 * npan + (size_t)kCqrSyncWords * p;
      ca.abortf = (unsigned*)h->flag.p;
      ca.fail = sync + 1 + npan + p;
      RNPX_TRY(launch_cqr_panel(ca, G, rw, st));
      *gated = ca.fail;
    }
  }
  PanelArgs<T> pa;
  pa.mp = mp;
  pa.nbp = ib;
  pa.A = A + (size_t)i + (size_t)i * lda;
  pa.lda = lda;
  pa.tau = tau + i;
  pa.V = V;
  pa.ldv = mp;
  pa.gram = gram;
  pa.part = (T*)h->ppart.p;
  pa.rowbuf = (T*)h->rowbuf.p;
  pa.counter = sync + 1 + p;
  pa.abortf = (unsigned*)h->flag.p;
  pa.run_if = *gated;
  const int rpt = pick_rpt(h, mp, nb, S<T>::cplx);
  RNPX_TRY(launch_panel<T>(pa, nb, rpt, st));
  // panel traffic: the mp x ib panel read and written once, plus V and tau
  h->prof_end(K_PANEL, ev, (S<T>::cplx ? 4.0 : 1.0) * (2.0 * mp * ib * ib - 2.0 / 3.0 * ib * ib * ib),
              sizeof(T) * (3.0 * mp * ib + ib), st);
  return RNPX_OK;
}

// Factor (QR.hpp:129-197) on device memory.  With look-ahead (the default)
// panel p+1 is factored on a second stream while the trailing update of
// panel p runs: the update of panel p is split into the next panel's columns
// (done first, then the panel stream is released) and the rest.  V, Gram and
// T are double-buffered between consecutive panels.
template <typename T>
int geqrf_dev(Handle* h, int m, int n, T* A, long lda, T* tau) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (lda < std::max(1, m)) return -5;
  const int k = std::min(m, n);
  if (k == 0) return RNPX_OK;
  hipStream_t st = h->stream;
  const int nb = pick_nb<T>(h, m, n);
  const int npan = (k + nb - 1) / nb;
  const int rpt0 = pick_rpt(h, m, nb, S<T>::cplx);
  const int G0 = (m + 256 * rpt0 - 1) / (256 * rpt0);
  // Every panel workgroup must be resident at once (they meet at grid
  // barriers): one 256-thread workgroup per CU at these register counts.
  if (G0 > std::min(kMaxPanelGroups, h->num_cus)) return RNPX_ERR_UNSUPPORTED;
  const bool look = (h->tune.lookahead >= 0) && npan > 1;
  Buffer* vb[2] = {&h->vbuf, &h->vbuf2};
  Buffer* gb[2] = {&h->gram, &h->gram2};
  Buffer* tb[2] = {&h->tbuf, &h->tbuf2};
  for (int q = 0; q < (look ? 2 : 1); ++q) {
    RNPX_TRY(vb[q]->ensure(sizeof(T) * (size_t)m * nb, h));
    RNPX_TRY(gb[q]->ensure(sizeof(T) * (size_t)nb * nb, h));
    RNPX_TRY(tb[q]->ensure(sizeof(T) * (size_t)nb * nb, h));
  }
  RNPX_TRY(h->ppart.ensure(sizeof(T) * 2 * (size_t)kMaxPanelGroups * nb + 64, h));
  RNPX_TRY(h->rowbuf.ensure(sizeof(T) * 2 * (size_t)nb, h));
  const int sync_words = 1 + 2 * npan + kCqrSyncWords * npan;
  RNPX_TRY(h->sync.ensure(sizeof(unsigned) * (size_t)sync_words, h));
  {
    int g0 = 0, rw0 = 0;
    if (use_cqr<T>(h, m, std::min(nb, k), nb, &g0, &rw0)) {
      // later panels may use more (smaller) workgroups than the first
      RNPX_TRY(h->cqr_part.ensure(sizeof(double) * (size_t)std::min(kCqrMaxGroups, h->num_cus) * 2560, h));
      RNPX_TRY(h->cqr_gsum.ensure(sizeof(double) * (2 + 16) * 2560, h));  // gsum, then group sums
      RNPX_TRY(h->cqr_scratch.ensure(sizeof(double) * kCqrScratch, h));
    }
  }
  RNPX_TRY(launch_zero_u32((unsigned*)h->sync.p, sync_words, st));
  const unsigned* gate[2] = {nullptr, nullptr};  // T-factor gate of the panel in each buffer

  auto update = [&](int p, int i, int ib, int c0, int c1) -> int {
    // columns [c0, c1) of the trailing matrix, rows [i, m), reflectors of panel p
    if (c1 <= c0) return RNPX_OK;
    const int q = look ? (p & 1) : 0;
    return apply_block_dev<T>(h, 'L', 'C', m - i, c1 - c0, ib, (const T*)vb[q]->p, m - i, (const T*)tb[q]->p, nb,
                              A + (size_t)i + (size_t)c0 * lda, lda);
  };
  auto tfactor = [&](int p, int ib) -> int {
    const int q = look ? (p & 1) : 0;
    hipEvent_t ev;
    h->prof_begin(&ev);
    RNPX_TRY(launch_tfactor<T>((const T*)gb[q]->p, nb, ib, nb, (T*)tb[q]->p, nb, st, gate[q]));
    h->prof_end(K_TFACTOR, ev, 0.0, 0.0);
    return RNPX_OK;
  };
  auto panel = [&](int p, int i, int ib, hipStream_t ps) -> int {
    const int q = look ? (p & 1) : 0;
    return launch_panel_step<T>(h, p, npan, i, ib, m, nb, A, lda, tau, (T*)vb[q]->p, (T*)gb[q]->p, (T*)tb[q]->p,
                                ps, &gate[q]);
  };

  if (!look) {
    for (int p = 0, i = 0; i < k; i += nb, ++p) {
      const int ib = std::min(nb, k - i);
      int s = panel(p, i, ib, st);
      if (s) return s;
      if (i + ib < n) {
        if ((s = tfactor(p, ib))) return s;
        if ((s = update(p, i, ib, i + ib, n))) return s;
      }
    }
    return RNPX_OK;
  }

  int s = h->ensure_pipeline();
  if (s) return s;
  hipStream_t ps = h->pstream;
  RNPX_TRY(hipEventRecord(h->ev_start, st));
  RNPX_TRY(hipStreamWaitEvent(ps, h->ev_start, 0));
  {
    const int ib0 = std::min(nb, k);
    if ((s = panel(0, 0, ib0, ps))) return s;
    RNPX_TRY(hipEventRecord(h->ev_pan[0], ps));
  }
  for (int p = 0, i = 0; i < k; i += nb, ++p) {
    const int ib = std::min(nb, k - i);
    RNPX_TRY(hipStreamWaitEvent(st, h->ev_pan[p & 1], 0));
    if (i + ib >= n) break;
    if ((s = tfactor(p, ib))) return s;
    const int i1 = i + ib;  // next panel's first column
    if (i1 < k) {
      const int ib1 = std::min(nb, k - i1);
      if ((s = update(p, i, ib, i1, i1 + ib1))) return s;
      RNPX_TRY(hipEventRecord(h->ev_next, st));
      RNPX_TRY(hipStreamWaitEvent(ps, h->ev_next, 0));
      const int q1 = (p + 1) & 1;
      if ((s = panel(p + 1, i1, ib1, ps))) return s;
      RNPX_TRY(hipEventRecord(h->ev_pan[q1], ps));
      if ((s = update(p, i, ib, i1 + ib1, n))) return s;
    } else {
      if ((s = update(p, i, ib, i1, n))) return s;
    }
  }
  return RNPX_OK;
}

// ---------------------------------------------------------------------------
// MultQ (QR.hpp:337-429): T recomputed per block from the stored V.
// ---------------------------------------------------------------------------
template <typename T>
int ormqr_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc) {
  const bool left = (side == 'L');
  if (side != 'L' && side != 'R') return -1;
  if (trans != 'N' && trans != 'C' && trans != 'T') return -2;
  if (S<T>::cplx && trans == 'T') return -2;
  if (trans == 'T') trans = 'C';
  if (m < 0) return -3;
  if (n < 0) return -4;
  const int nq = left ? m : n;
  if (k < 0 || k > nq) return -5;
  if (lda < std::max(1, nq)) return -7;
  if (ldc < std::max(1, m)) return -10;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  const int nb = pick_nb<T>(h, m, n);
  RNPX_TRY(h->vbuf.ensure(sizeof(T) * (size_t)nq * nb, h));
  RNPX_TRY(h->tbuf.ensure(sizeof(T) * (size_t)nb * nb, h));
  T* V = (T*)h->vbuf.p;
  T* Tm = (T*)h->tbuf.p;
  const bool forward = (left && trans == 'C') || (!left && trans == 'N');
  const int nblk = (k + nb - 1) / nb;
  for (int b = 0; b < nblk; ++b) {
    const int i = forward ? b * nb : (nblk - 1 - b) * nb;
    const int ib = std::min(nb, k - i);
    const int len = nq - i;
    RNPX_TRY(launch_extract_v<T>(len, ib, A + (size_t)i + (size_t)i * lda, lda, V, len, h->stream));
    int s = tfactor_from_v<T>(h, len, ib, nb, V, len, tau + i, Tm, nb);
    if (s) return s;
    if (left)
      s = apply_block_dev<T>(h, 'L', trans, len, n, ib, V, len, Tm, nb, C + i, ldc);
    else
      s = apply_block_dev<T>(h, 'R', trans, m, len, ib, V, len, Tm, nb, C + (size_t)i * ldc, ldc);
    if (s) return s;
  }
  return RNPX_OK;
}

// ---------------------------------------------------------------------------
// GenerateQ (QR.hpp:509-570): Q(:,0:n) = H_0 ... H_{k-1} [I; 0].
// The reflectors are copied out first; the identity is then transformed
// block by block, last block first, each block touching columns [i, n).
// ---------------------------------------------------------------------------
template <typename T>
int orgqr_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau) {
  if (m < 0) return -1;
  if (n < 0 || n > m) return -2;
  if (k < 0 || k > n) return -3;
  if (lda < std::max(1, m)) return -5;
  if (n == 0) return RNPX_OK;
  hipStream_t st = h->stream;
  const int nb = pick_nb<T>(h, m, n);
  RNPX_TRY(h->vfull.ensure(sizeof(T) * (size_t)m * std::max(k, 1), h));
  RNPX_TRY(h->tbuf.ensure(sizeof(T) * (size_t)nb * nb, h));
  T* VF = (T*)h->vfull.p;
  if (k > 0) RNPX_TRY(launch_extract_v<T>(m, k, A, lda, VF, m, st));
  RNPX_TRY(launch_set<T>(m, n, S<T>::zero(), S<T>::one(), A, lda, st));
  const int nblk = (k + nb - 1) / nb;
  for (int b = nblk - 1; b >= 0; --b) {
    const int i = b * nb;
    const int ib = std::min(nb, k - i);
    const T* Vb = VF + (size_t)i + (size_t)i * m;
    int s = tfactor_from_v<T>(h, m - i, ib, nb, Vb, m, tau + i, (T*)h->tbuf.p, nb);
    if (s) return s;
    s = apply_block_dev<T>(h, 'L', 'N', m - i, n - i, ib, Vb, m, (const T*)h->tbuf.p, nb,
                           A + (size_t)i + (size_t)i * lda, lda);
    if (s) return s;
  }
  return RNPX_OK;
}

template <typename T>
int larft_dev(Handle* h, int n, int k, const T* V, long ldv, const T* tau, T* Tout, long ldt) {
  if (n < 0) return -2;
  if (k < 1 || k > 64 || k > n) return -3;
  if (ldv < std::max(1, n)) return -5;
  if (ldt < k) return -8;
  const int nb = (k <= 32) ? 32 : 64;
  RNPX_TRY(h->vbuf.ensure(sizeof(T) * (size_t)n * k, h));
  RNPX_TRY(h->tbuf.ensure(sizeof(T) * (size_t)nb * nb, h));
  T* VC = (T*)h->vbuf.p;
  RNPX_TRY(launch_extract_v<T>(n, k, V, ldv, VC, n, h->stream));
  int s = tfactor_from_v<T>(h, n, k, nb, VC, n, tau, (T*)h->tbuf.p, nb);
  if (s) return s;
  RNPX_TRY(launch_copy<T>(k, k, (const T*)h->tbuf.p, nb, Tout, ldt, h->stream));
  return RNPX_OK;
}

template <typename T>
int larfb_dev(Handle* h, char side, char trans, int m, int n, int k, const T* V, long ldv, const T* Tm, long ldt,
              T* C, long ldc) {
  if (side != 'L' && side != 'R') return -2;
  if (trans != 'N' && trans != 'C') return -3;
  if (m < 0) return -4;
  if (n < 0) return -5;
  const int nv = (side == 'L') ? m : n;
  if (k < 0 || k > nv) return -6;
  if (ldv < std::max(1, nv)) return -8;
  if (ldt < std::max(1, k)) return -10;
  if (ldc < std::max(1, m)) return -12;
  if (m == 0 || n == 0 || k == 0) return RNPX_OK;
  RNPX_TRY(h->vfull.ensure(sizeof(T) * (size_t)nv * k, h));
  T* VC = (T*)h->vfull.p;
  RNPX_TRY(launch_extract_v<T>(nv, k, V, ldv, VC, nv, h->stream));
  return apply_block_dev<T>(h, side, trans, m, n, k, VC, nv, Tm, ldt, C, ldc);
}

int prof_collect(Handle* h, int kid, long* launches, double* ms, double* flops, double* bytes) {
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_status(e);
  long n = 0;
  double t = 0, f = 0, b = 0;
  for (auto& r : h->recs) {
    if (r.kid != kid) continue;
    float x = 0;
    if (hipEventElapsedTime(&x, r.a, r.b) == hipSuccess) t += x;
    f += r.flops;
    b += r.bytes;
    ++n;
  }
  *launches = n;
  *ms = t;
  *flops = f;
  *bytes = b;
  return RNPX_OK;
}

void prof_reset(Handle* h) {
  hipStreamSynchronize(h->stream);
  for (auto& r : h->recs) {
    h->event_pool.push_back(r.a);
    h->event_pool.push_back(r.b);
  }
  h->recs.clear();
}

int sync_handle(Handle* h) {
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_status(e);
  unsigned flag = 0;
  e = hipMemcpy(&flag, h->flag.p, sizeof(unsigned), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_status(e);
  if (flag) {
    hipMemset(h->flag.p, 0, sizeof(unsigned));
    return RNPX_ERR_SYNC;
  }
  return RNPX_OK;
}

template int geqrf_dev<double>(Handle*, int, int, double*, long, double*);
template int geqrf_dev<zd>(Handle*, int, int, zd*, long, zd*);
template int ormqr_dev<double>(Handle*, char, char, int, int, int, const double*, long, const double*, double*, long);
template int ormqr_dev<zd>(Handle*, char, char, int, int, int, const zd*, long, const zd*, zd*, long);
template int orgqr_dev<double>(Handle*, int, int, int, double*, long, const double*);
template int orgqr_dev<zd>(Handle*, int, int, int, zd*, long, const zd*);
template int larft_dev<double>(Handle*, int, int, const double*, long, const double*, double*, long);
template int larft_dev<zd>(Handle*, int, int, const zd*, long, const zd*, zd*, long);
template int larfb_dev<double>(Handle*, char, char, int, int, int, const double*, long, const double*, long, double*,
                               long);
template int larfb_dev<zd>(Handle*, char, char, int, int, int, const zd*, long, const zd*, long, zd*, long);

}  // namespace rnpx
