// rnpx_qr.cpp -- host orchestration of the gfx950 QR and the C ABI of
// include/rnpx.h.  Factor / MultQ / GenerateQ follow RNP's blocked
// algorithms (RNP/LA/QR.hpp:129-197, 337-429, 509-570): a panel kernel, the
// T-factor kernel and three MFMA GEMMs per block, all on one HIP stream.
#include "rnpx_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
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
    // The buffer may be in use on any stream of the handle (the look-ahead's
    // panel stream grows its GEMM scratch while earlier panel work is queued):
    // every one of them is drained before the old allocation goes.
    hipError_t e = h ? h->sync_all_streams() : hipDeviceSynchronize();
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
hipError_t HostBuffer::ensure(size_t b, Handle* h) {
  if (b <= bytes) return hipSuccess;
  if (p) {
    hipError_t e = h ? h->sync_all_streams() : hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  const size_t want = std::max<size_t>(b, 4096);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = want;
  return hipSuccess;
}
void HostBuffer::release() {
  if (p) hipHostFree(p);
  p = nullptr;
  bytes = 0;
}
void Buffer::release() {
  if (p) hipFree(p);
  p = nullptr;
  bytes = 0;
}

hipError_t Handle::sync_all_streams() {
  for (hipStream_t s : {stream, pstream, part_ps, part_us}) {
    if (!s) continue;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t Handle::held_bytes() const {
  size_t s = 0;
  for (const Buffer* b : const_cast<Handle*>(this)->buffers()) s += b->bytes;
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
int Handle::prof_begin(int kid, hipEvent_t* a, hipStream_t s) {
  *a = nullptr;
  if (!prof || !((prof_mask >> kid) & 1u)) return 0;
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
void Handle::prof_take(int kid, hipEvent_t* a, hipEvent_t* b) {
  *a = *b = nullptr;
  if (!prof || !((prof_mask >> kid) & 1u)) return;
  *a = take_event();
  *b = take_event();
  if (!*a || !*b) {
    for (hipEvent_t e : {*a, *b})
      if (e) event_pool.push_back(e);
    *a = *b = nullptr;
  }
}
void Handle::prof_push(int kid, hipEvent_t a, hipEvent_t b, double flops, double bytes) {
  if (a && b) recs.push_back(ProfRec{kid, a, b, flops, bytes});
}
// A stream with a CU mask (here every CU) gets a hardware queue of its own,
// where a plain one may share a queue with other streams of the process once
// they outnumber the runtime's queues -- with the update stream among them,
// the look-ahead would run in series.  A plain non-blocking stream where the
// mask is refused.
hipError_t create_own_queue_stream(hipStream_t* s, int num_cus) {
  const int words = (num_cus + 31) / 32;
  std::vector<uint32_t> all(words, 0u);
  for (int c = 0; c < num_cus; ++c) all[c / 32] |= 1u << (c % 32);
  hipError_t e = hipExtStreamCreateWithCUMask(s, (uint32_t)words, all.data());
  if (e == hipSuccess) return e;
  (void)hipGetLastError();
  *s = nullptr;
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

int Handle::ensure_pipeline() {
  if (pstream) return RNPX_OK;
  hipError_t e = hipSuccess;
  if (ext_pstream) {
    pstream = ext_pstream;
  } else {
    e = create_own_queue_stream(&pstream, num_cus);
    if (e != hipSuccess) return hip_status(e);
  }
  for (hipEvent_t* ev : {&ev_pan[0], &ev_pan[1], &ev_next, &ev_start, &ev_rest}) {
    e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e != hipSuccess) return hip_status(e);
  }
  return RNPX_OK;
}

// Two streams on complementary CU masks: the first `cus` CUs for the panel
// stream, the rest for the update stream.  Returns the panel set's size, or 0
// when the runtime refuses masks (the caller then uses plain streams).
int Handle::ensure_partition(int cus) {
  if (part_want == cus) return part_cus;
  for (hipStream_t* s : {&part_ps, &part_us}) {
    if (*s) {
      hipStreamSynchronize(*s);
      hipStreamDestroy(*s);
      *s = nullptr;
    }
  }
  part_want = cus;
  part_cus = 0;
  if (cus <= 0 || cus >= num_cus) return 0;
  const int words = (num_cus + 31) / 32;
  std::vector<uint32_t> mp(words, 0u), mu(words, 0u);
  for (int c = 0; c < num_cus; ++c) (c < cus ? mp : mu)[c / 32] |= 1u << (c % 32);
  hipError_t e1 = hipExtStreamCreateWithCUMask(&part_ps, (uint32_t)words, mp.data());
  if (e1 != hipSuccess) {
    if (getenv("RNPX_REPORT_FALLBACK")) fprintf(stderr, "rnpx: CU mask refused: %s\n", hipGetErrorString(e1));
    part_ps = nullptr;
    (void)hipGetLastError();
    return 0;
  }
  if (hipExtStreamCreateWithCUMask(&part_us, (uint32_t)words, mu.data()) != hipSuccess) {
    part_us = nullptr;
    (void)hipGetLastError();
    hipStreamDestroy(part_ps);
    part_ps = nullptr;
    return 0;
  }
  part_cus = cus;
  if (getenv("RNPX_REPORT_FALLBACK")) fprintf(stderr, "rnpx: CU partition %d + %d of %d\n", cus, num_cus - cus, num_cus);
  return part_cus;
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
  for (hipEvent_t ev : {ev_pan[0], ev_pan[1], ev_next, ev_start, ev_rest})
    if (ev) hipEventDestroy(ev);
  if (pstream && pstream != ext_pstream) hipStreamDestroy(pstream);
  for (hipStream_t ps2 : {part_ps, part_us}) {
    if (ps2) {
      hipStreamSynchronize(ps2);
      hipStreamDestroy(ps2);
    }
  }
  for (Buffer* b : buffers()) b->release();
  hpin.release();
  if (own_stream && own) hipStreamDestroy(own);
}

int hip_status(hipError_t e) {
  if (e == hipSuccess) return RNPX_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return RNPX_ERR_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return RNPX_ERR_NOGPU;
  if (e == hipErrorInvalidValue) return RNPX_ERR_UNSUPPORTED;
  return RNPX_ERR_HIP;
}

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
  // the handle's stream, on a hardware queue of its own (work of other
  // handles in the process does not queue in front of it)
  hipError_t e = create_own_queue_stream(&h->own, h->num_cus);
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
  // Library set-up for the allocation-free batched entry point: this
  // device's address of its static scratch (once per device; a failure is
  // recorded and reported by that entry point only).
  (void)noalloc_setup(h->device);
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
    // Untuned: 64-column panels for real matrices (the CholeskyQR2 block
    // kernel); 32 for complex ones, the only width with a CholeskyQR2 panel
    // (64-column complex panels go column by column, about three times
    // slower at 4096^2).
    nb = S<T>::cplx ? 32 : 64;
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
  // to 4 rows per thread, NB=64 with 1; complex NB=32 with up to 2, NB=64
  // with 1.
  const int maxrpt = nb <= 32 ? (cplx ? 2 : 4) : 1;
  if (rpt < 1 || rpt > maxrpt || (rpt & (rpt - 1))) {
    // default: keep the workgroup count near 16-32
    rpt = 1;
    while (rpt < maxrpt && (mp + 256 * rpt - 1) / (256 * rpt) > 32) rpt *= 2;
  }
  const int cap = std::min(kMaxPanelGroups, h->num_cus);
  while ((mp + 256 * rpt - 1) / (256 * rpt) > cap && rpt < maxrpt) rpt *= 2;
  return rpt;
}

static int pick_ksplit(const Handle* h, int rows, int cols_tiles) {
  // A tuned split applies to the wide products (the trailing update's W);
  // the narrow ones keep the latency-oriented split below whatever the
  // tuning says.
  if (h->tune.ksplit > 0 && cols_tiles > 2) return h->tune.ksplit;
  // About 512 workgroup tiles of 128 x 128 in all (two per CU; measured best
  // for W = V^H C at 4096-8192 rows).  K slices of at least 512, or 128 for
  // the narrow products (one or two column tiles) that sit on the chain of
  // panels, where latency matters more than the partial sums' traffic.
  int want = (cols_tiles <= 2 ? 1024 : 512) / std::max(1, cols_tiles);
  int maxs = std::max(1, rows / (cols_tiles <= 2 ? 128 : 512));
  // narrow products: up to 64 slices (a tall panel's W is otherwise a long
  // chain of K slices on few workgroups)
  return std::max(1, std::min(want, std::min(maxs, cols_tiles <= 2 ? 64 : 32)));
}

// ---------------------------------------------------------------------------
// Block reflector application (the larfb analogue, Reflector.hpp:725-815)
// with a clean V (unit diagonal, zeros above) and T on the device.
//   side L: C(mc x nc) = op(H) C,   H = I - V T V^H, V is mc x k
//   side R: C(mc x nc) = C op(H),   V is nc x k
// op = H for trans 'N', H^H for trans 'C'.
// ---------------------------------------------------------------------------
// Where a block application runs: stream, GEMM scratch, whether it is
// profiled.  The default is the handle's stream and main scratch.
struct GemmWork {
  hipStream_t st;
  Buffer* w;
  Buffer* z;
  Buffer* part;
  bool prof;
};

template <typename T>
static int apply_block_dev(Handle* h, char side, char trans, int mc, int nc, int k, const T* V, long ldv,
                           const T* Tm, long ldt, T* C, long ldc, const GemmWork* gw = nullptr) {
  if (mc <= 0 || nc <= 0 || k <= 0) return RNPX_OK;
  const GemmWork main_work = {h->stream, &h->wbuf, &h->zbuf, &h->part, true};
  if (!gw) gw = &main_work;
  hipStream_t st = gw->st;
  Buffer& wb = *gw->w;
  Buffer& zb = *gw->z;
  Buffer& pb = *gw->part;
  const T one = S<T>::one(), zero = S<T>::zero(), mone = S<T>::zero() - S<T>::one();
  if (side == 'L') {
    // W = V^H C (k x nc), Z = op(T) W, C -= V Z
    const int tiles = (nc + 127) / 128;
    int ks = pick_ksplit(h, mc, tiles);
    RNPX_TRY(wb.ensure(sizeof(T) * (size_t)k * nc, h));
    RNPX_TRY(zb.ensure(sizeof(T) * (size_t)k * nc, h));
    if (ks > 1) RNPX_TRY(pb.ensure(sizeof(T) * (size_t)ks * k * nc, h));
    T* W = (T*)wb.p;
    T* Z = (T*)zb.p;
    const double fl = (S<T>::cplx ? 8.0 : 2.0) * (double)mc * nc * k;
    const double esz = sizeof(T);
    hipEvent_t ev = nullptr;
    // Mixed precision (real only): the two big products on FP32 MFMA.
    bool f32 = false;
    if constexpr (!S<T>::cplx) f32 = (h->update_prec == 1);
    if (gw->prof) h->prof_begin(K_WGEMM, &ev, st);
    if constexpr (!S<T>::cplx) {
      if (f32) RNPX_TRY(gemm32('C', 'N', k, nc, mc, 1.0, V, ldv, C, ldc, 0.0, W, k, ks, (T*)pb.p, st));
    }
    if (!f32) RNPX_TRY(gemm<T>('C', 'N', k, nc, mc, one, V, ldv, C, ldc, zero, W, k, ks, (T*)pb.p, st));
    if (gw->prof) h->prof_end(K_WGEMM, ev, fl, esz * ((double)mc * nc + (double)mc * k + (double)k * nc), st);
    RNPX_TRY(gemm<T>(trans == 'N' ? 'N' : 'C', 'N', k, nc, k, one, Tm, ldt, W, k, zero, Z, k, 1, nullptr, st));
    // The update's timing events ride on its launch (none between kernels).
    hipEvent_t u0 = nullptr, u1 = nullptr;
    if (gw->prof) h->prof_take(K_UPDATE, &u0, &u1);
    if constexpr (!S<T>::cplx) {
      if (f32) {
        if (u0) RNPX_TRY(hipEventRecord(u0, st));
        RNPX_TRY(gemm32('N', 'N', mc, nc, k, -1.0, V, ldv, Z, k, 1.0, C, ldc, 1, nullptr, st));
        if (u0) RNPX_TRY(hipEventRecord(u1, st));
      }
    }
    if (!f32) RNPX_TRY(gemm<T>('N', 'N', mc, nc, k, mone, V, ldv, Z, k, one, C, ldc, 1, nullptr, st, u0, u1));
    h->prof_push(K_UPDATE, u0, u1, fl, esz * (2.0 * mc * nc + (double)mc * k + (double)k * nc));
  } else {
    // W = C V (mc x k), Z = W op(T), C -= Z V^H
    const int tiles = (mc + 127) / 128;
    int ks = pick_ksplit(h, nc, tiles);
    RNPX_TRY(wb.ensure(sizeof(T) * (size_t)k * mc, h));
    RNPX_TRY(zb.ensure(sizeof(T) * (size_t)k * mc, h));
    if (ks > 1) RNPX_TRY(pb.ensure(sizeof(T) * (size_t)ks * k * mc, h));
    T* W = (T*)wb.p;
    T* Z = (T*)zb.p;
    RNPX_TRY(gemm<T>('N', 'N', mc, k, nc, one, C, ldc, V, ldv, zero, W, mc, ks, (T*)pb.p, st));
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
    // Panels of fewer than 128 rows take the column-wise kernel (one
    // workgroup); from 128 rows the CholeskyQR2 kernel (≈ 100 µs for a
    // 128 x 64 panel against ≈ 260 µs column by column).
    if (h->tune.panel == 1 || nb != 64 || ib != 64) return false;
    *g = cqr_panel_groups(mp, std::min(kCqrMaxGroups, h->panel_cap > 0 ? h->panel_cap : h->num_cus), rw);
    return *g > 0;
  }
}

// One inner panel: rows [i, m), columns [i, i + ib) of A factored in place
// (R and V packed, tau), its clean V written at V (leading dimension ldv)
// and its T (ib x ib, leading dimension ldt) at Tm, all on stream st.  sync
// holds, per inner panel p: [1 + p] the column-wise kernel's barrier
// counter, [1 + npan + p] the CholeskyQR2 kernel's fail flag, and
// kCqrSyncWords words at 1 + 2 npan + kCqrSyncWords p for its reductions.
// Where the CholeskyQR2 kernel applies it writes V and T itself (a panel it
// rejects is refactored column by column inside the same launch).
template <typename T>
static int panel_step(Handle* h, int p, int npan, int i, int ib, int m, int nb, T* A, long lda, T* tau, T* V, long ldv,
                      T* Tm, long ldt, hipStream_t st) {
  const int mp = m - i;
  unsigned* sync = (unsigned*)h->sync.p;
  int G = 0, rw = 0;
  if (use_cqr<T>(h, mp, ib, nb, &G, &rw)) {
    if constexpr (!S<T>::cplx) {
      CqrArgs ca;
      ca.mp = mp;
      ca.A = A + (size_t)i + (size_t)i * lda;
      ca.lda = lda;
      ca.tau = tau + i;
      ca.V = V;
      ca.ldv = ldv;
      ca.T = Tm;
      ca.ldt = ldt;
      ca.part = (double*)h->cqr_part.p;
      ca.gpart = (double*)h->cqr_gsum.p + 2 * 2560;
      ca.gsum = (double*)h->cqr_gsum.p;
      ca.scratch = (double*)h->cqr_scratch.p;
      ca.sync = sync + 1 + 2 * npan + (size_t)kCqrSyncWords * p;
      ca.abortf = (unsigned*)h->flag.p;
      ca.fail = sync + 1 + npan + p;
      return hip_status(launch_cqr_panel(ca, G, rw, st));
    }
  }
  PanelArgs<T> pa;
  pa.mp = mp;
  pa.nbp = ib;
  pa.A = A + (size_t)i + (size_t)i * lda;
  pa.lda = lda;
  pa.tau = tau + i;
  pa.V = V;
  pa.ldv = ldv;
  pa.gram = (T*)h->gram.p;
  pa.part = (T*)h->ppart.p;
  pa.rowbuf = (T*)h->rowbuf.p;
  pa.counter = sync + 1 + p;
  pa.abortf = (unsigned*)h->flag.p;
  pa.run_if = nullptr;
  const int rpt = pick_rpt(h, mp, nb, S<T>::cplx);
  if constexpr (S<T>::cplx) {
    // Complex full-width panels of 64 rows or more: CholeskyQR2 and
    // Householder reconstruction (two Gram reductions instead of one grid
    // hand-off, or one workgroup's step, per column: a 256 x 32 panel takes
    // ≈ 130 µs this way against ≈ 186 µs column by column on one
    // workgroup); shorter ones go column by column.
    if (h->tune.panel != 1 && ib == nb && nb == 32 && mp >= 64)
      return zcqr_panel(h, mp, nb, pa.A, lda, pa.tau, V, ldv, Tm, ldt, sync + 1 + npan + p, pa, rpt, st);
  }
  RNPX_TRY(launch_panel<T>(pa, nb, rpt, st));
  RNPX_TRY(launch_tfactor<T>((const T*)h->gram.p, nb, ib, nb, Tm, (int)ldt, st));
  return RNPX_OK;
}

// Factor (QR.hpp:129-197) on device memory.
//
// Outer blocks of nbo = 2 nb columns (the default; nb = 64) are factored as
// two inner panels of nb columns: panel 1, its block reflector applied to
// the next nb columns, panel 2, then the two compact-WY forms are merged,
//   V = [V1 | (0; V2)],  T = [T1, -T1 (V1^H V2) T2; 0, T2],
// so that the trailing matrix sees one rank-nbo update (K = 128 GEMMs run
// half as many passes over it at higher arithmetic intensity than K = 64).
// With nb = 256 tuning, blocks are 256 columns (two 128-column units merged
// the same way, K = 256 updates) while the remaining height exceeds
// RNPX_WIDE_ROWS (4096), and 128 columns below it, where the chain of panels
// is the longer one.
// With look-ahead (the default) the panel phase of block b+1 runs on a
// second stream while the trailing update of block b runs: the update of
// block b is split into block b+1's columns, applied on the panel stream
// right behind panel b so that the chain of panels crosses no stream, and
// the rest, on the update stream.  V and T are double-buffered between
// consecutive outer blocks; V uses leading dimension m.
template <typename T>
int geqrf_dev(Handle* h, int m, int n, T* A, long lda, T* tau, T* B, int nrhs, long ldb) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (lda < std::max(1, m)) return -5;
  if (!B) nrhs = 0;
  if (nrhs > 0 && ldb < std::max(1, m)) return -9;
  const int k = std::min(m, n);
  if (k == 0) return RNPX_OK;
  // Automatic tuning (nb = 0): the autotuned table's entry nearest this
  // shape, for this call only.
  struct TuneRestore {
    Handle* h;
    rnpx_tuning t;
    ~TuneRestore() { h->tune = t; }
  } tune_restore{h, h->tune};
  if (h->tune.nb == 0) {
    int tnb = 0, tks = 0, tcus = 0;
    if (tuning_factor(S<T>::cplx ? 'z' : 'd', m, n, &tnb, &tks, &tcus)) {
      h->tune.nb = tnb;
      if (h->tune.ksplit == 0) h->tune.ksplit = tks;
      if (h->tune.panel_cus == 0) h->tune.panel_cus = tcus;
    }
  }
  hipStream_t st = h->stream;
  int nb = pick_nb<T>(h, m, n);
  // Every panel workgroup must be resident at once (they meet at grid
  // barriers): one 256-thread workgroup per CU at these register counts.
  const int gcap = std::min(kMaxPanelGroups, h->num_cus);
  int rpt0 = pick_rpt(h, m, nb, S<T>::cplx);
  int G0 = (m + 256 * rpt0 - 1) / (256 * rpt0);
  if (G0 > gcap && nb > 32) {
    // Taller than a 64-column panel holds: 32-column panels, whose
    // workgroups keep up to 4 (real) or 2 (complex) rows per thread.
    nb = 32;
    rpt0 = pick_rpt(h, m, nb, S<T>::cplx);
    G0 = (m + 256 * rpt0 - 1) / (256 * rpt0);
  }
  if (G0 > gcap) return RNPX_ERR_UNSUPPORTED;
  const int nbo = (h->tune.nb == 256 ? 4 : ((h->tune.nb == 0 && !S<T>::cplx) || h->tune.nb == 128) ? 2 : 1) * nb;
  // Outer blocks: nbo columns while the remaining height exceeds wide_rows
  // (where the trailing products are the longer chain and K = nbo runs them
  // faster), 2 nb columns below it (where the panels are the longer chain).
  static const int wide_rows = [] {
    const char* e = getenv("RNPX_WIDE_ROWS");
    return e ? atoi(e) : 4096;
  }();
  std::vector<int> bstart;
  {
    const int tailw = std::min(nbo, 2 * nb);
    for (int i = 0; i < k;) {
      bstart.push_back(i);
      i += std::min(k - i, (m - i > wide_rows) ? nbo : tailw);
    }
    bstart.push_back(k);
  }
  auto bw = [&](int b) { return bstart[b + 1] - bstart[b]; };
  const int npan = (k + nb - 1) / nb;     // inner panels
  const int nob = (int)bstart.size() - 1;  // outer blocks
  const bool look = (h->tune.lookahead >= 0) && nob > 1;
  const long ldv = m, ldt = nbo;
  // Inf or NaN in A: the reference's blocking and zero trims decide which
  // 0 * Inf products it forms, so such a matrix is factored their way.  With
  // look-ahead only the first block's columns are checked up front; the
  // rest is checked while the first panel runs, from a copy of those
  // columns that is put back if the factorization has to start over.
  auto factor_nonfinite = [&]() -> int {
    RNPX_CHECK(geqrf_exact<T>(h, m, n, A, lda, tau, h->exact_lwork));
    if (nrhs > 0) RNPX_CHECK(ormqr_dev<T>(h, 'L', 'C', m, nrhs, k, A, lda, tau, B, ldb));
    return RNPX_OK;
  };
  const int w0 = bw(0);
  const bool split_scan = look && n > w0;
  {
    bool bad = false;
    RNPX_CHECK(any_nonfinite<T>(h, m, split_scan ? w0 : n, A, lda, &bad, st));
    if (bad) return factor_nonfinite();
  }
  Buffer* vb[2] = {&h->vbuf, &h->vbuf2};
  Buffer* tb[2] = {&h->tbuf, &h->tbuf2};
  for (int q = 0; q < (look ? 2 : 1); ++q) {
    RNPX_TRY(vb[q]->ensure(sizeof(T) * (size_t)m * nbo, h));
    RNPX_TRY(tb[q]->ensure(sizeof(T) * (size_t)nbo * nbo, h));
  }
  RNPX_TRY(h->gram.ensure(sizeof(T) * (size_t)nb * nb, h));
  RNPX_TRY(h->ppart.ensure(sizeof(T) * 2 * (size_t)kMaxPanelGroups * nb + 64, h));
  RNPX_TRY(h->rowbuf.ensure(sizeof(T) * 2 * (size_t)nb, h));
  const int sync_words = 1 + 2 * npan + kCqrSyncWords * npan;
  RNPX_TRY(h->sync.ensure(sizeof(unsigned) * (size_t)sync_words, h));
  // Outer blocks in one launch (real, 64-column panels, 128-column blocks,
  // tuning panel 0): buffers for the block kernel.
  const bool use_block = !S<T>::cplx && h->tune.panel == 0 && nb == 64 && nbo % (2 * nb) == 0 &&
                         cqr_block_groups(m, std::min(256, h->num_cus)) > 0;
  static_assert(2 * kCqrSyncWords >= kCqrBlockSyncWords, "a block uses its two panels' sync words");
  if (use_block) {
    RNPX_TRY(h->cqr_part.ensure(sizeof(double) * (size_t)cqr_block_groups(m, std::min(256, h->num_cus)) * 4096, h));
    RNPX_TRY(h->cqr_gsum.ensure(sizeof(double) * 6 * 4096, h));
    RNPX_TRY(h->cqr_scratch.ensure(sizeof(double) * kCqrBlockScratch, h));
  }
  {
    int g0 = 0, rw0 = 0;
    if (use_cqr<T>(h, m, std::min(nb, k), nb, &g0, &rw0)) {
      // later panels may use more (smaller) workgroups than the first
      RNPX_TRY(h->cqr_part.ensure(sizeof(double) * (size_t)std::min(kCqrMaxGroups, h->num_cus) * 2560, h));
      RNPX_TRY(h->cqr_gsum.ensure(sizeof(double) * (2 + 16) * 2560, h));  // gsum, then group sums
      RNPX_TRY(h->cqr_scratch.ensure(sizeof(double) * kCqrScratch, h));
    }
  }
  RNPX_TRY(h->misc2.ensure(sizeof(T) * (size_t)std::max(3 * nb * nb, 2 * (nbo / 2) * (nbo / 2)), h));
  RNPX_TRY(launch_zero_u32((unsigned*)h->sync.p, sync_words, st));
  // An outer block of w columns is split into w1 = 2 nb (a block-kernel
  // unit) or nb columns and the rest.  The parts of the merged V and T no
  // panel writes -- V's rows above the second part, T's block below the
  // diagonal -- are zero; they are cleared once, for the full width and the
  // last block's, in both buffers.
  auto split_of = [&](int w) { return w > 2 * nb ? 2 * nb : nb; };
  std::function<int(int, long, T*, T*)> zero_parts = [&](int w, long off, T* Vq, T* Tq) -> int {
    if (w <= nb) return RNPX_OK;
    const int w1 = split_of(w), w2 = w - w1;
    RNPX_TRY(launch_set<T>(w1, w2, S<T>::zero(), S<T>::zero(), Vq + off + (off + w1) * ldv, ldv, st));
    RNPX_TRY(launch_set<T>(w2, w1, S<T>::zero(), S<T>::zero(), Tq + (off + w1) + off * ldt, ldt, st));
    RNPX_CHECK(zero_parts(w1, off, Vq, Tq));
    return zero_parts(w2, off + w1, Vq, Tq);
  };
  if (k > nb) {
    std::vector<int> widths;
    for (int b = 0; b < nob; ++b)
      if (std::find(widths.begin(), widths.end(), bw(b)) == widths.end()) widths.push_back(bw(b));
    for (int q = 0; q < (look ? 2 : 1); ++q)
      for (int w : widths) RNPX_CHECK(zero_parts(w, 0, (T*)vb[q]->p, (T*)tb[q]->p));
  }

  // Columns [i, i + w) (rows [i, m)) factored into V (row 0 = row i,
  // leading dimension ldv) and T (ldt) on stream ps: one block-kernel launch
  // for a real 2 nb-column unit, one panel for nb columns or fewer, else the
  // first part, its reflectors applied to the rest's columns, the rest, and
  // the two compact-WY forms merged,
  //   V = [V1 | (0; V2)],  T = [T1, -T1 (V1^H V2) T2; 0, T2].
  std::function<int(int, int, T*, T*, hipStream_t, const GemmWork*)> phase;
  phase = [&](int i, int w, T* Vp, T* Tp, hipStream_t ps, const GemmWork* gw) -> int {
    if constexpr (!S<T>::cplx) {
      // every workgroup of the block kernel must have a CU of the panel set
      const int bcap = std::min(256, h->panel_cap > 0 ? h->panel_cap : h->num_cus);
      const int Gb = (use_block && w == 2 * nb) ? cqr_block_groups(m - i, bcap) : 0;
      if (Gb > 0) {
        // Both panels, the apply between them and the merge in one launch.
        const int p1 = i / nb;
        CqrBlockArgs ba;
        ba.mp = m - i;
        ba.A = A + (size_t)i + (size_t)i * lda;
        ba.lda = lda;
        ba.tau = tau + i;
        ba.V = Vp;
        ba.ldv = ldv;
        ba.T = Tp;
        ba.ldt = ldt;
        ba.part = (double*)h->cqr_part.p;
        ba.gsum = (double*)h->cqr_gsum.p;
        ba.scratch = (double*)h->cqr_scratch.p;
        ba.sync = (unsigned*)h->sync.p + 1 + 2 * npan + (size_t)kCqrSyncWords * p1;  // panels p1 and p1+1
        ba.abortf = (unsigned*)h->flag.p;
        ba.fail = (unsigned*)h->sync.p + 1 + npan + p1;
        return hip_status(launch_cqr_block(ba, Gb, ps));
      }
    }
    if (w <= nb) return panel_step<T>(h, i / nb, npan, i, w, m, nb, A, lda, tau, Vp, ldv, Tp, ldt, ps);
    const int w1 = split_of(w), w2 = w - w1;
    RNPX_CHECK(phase(i, w1, Vp, Tp, ps, gw));
    // the first part's reflectors applied to the rest's columns
    RNPX_CHECK(apply_block_dev<T>(h, 'L', 'C', m - i, w2, w1, Vp, ldv, Tp, ldt,
                                  A + (size_t)i + (size_t)(i + w1) * lda, lda, gw));
    T* V2 = Vp + (size_t)w1 + (size_t)w1 * ldv;
    T* T2 = Tp + (size_t)w1 + (size_t)w1 * ldt;
    RNPX_CHECK(phase(i + w1, w2, V2, T2, ps, gw));
    // T12 = -T1 (V1^H V2) T2, V1^H V2 over the rows of V2
    const T one = S<T>::one(), zero = S<T>::zero(), mone = S<T>::zero() - S<T>::one();
    T* Sm = (T*)h->misc2.p;
    T* Xm = Sm + (size_t)w1 * w2;
    const int rows2 = m - i - w1;
    const int ks = std::max(1, std::min(32, rows2 / 256));
    RNPX_TRY(gw->part->ensure(sizeof(T) * (size_t)ks * w1 * w2, h));
    RNPX_TRY(gemm<T>('C', 'N', w1, w2, rows2, one, Vp + w1, ldv, V2, ldv, zero, Sm, w1, ks, (T*)gw->part->p, ps));
    RNPX_TRY(gemm<T>('N', 'N', w1, w2, w1, one, Tp, ldt, Sm, w1, zero, Xm, w1, 1, nullptr, ps));
    RNPX_TRY(gemm<T>('N', 'N', w1, w2, w2, mone, Xm, w1, T2, ldt, zero, Tp + (size_t)w1 * ldt, ldt, 1, nullptr, ps));
    return RNPX_OK;
  };
  // The panel phase of outer block b (columns [i, i + wo)) on stream ps.
  auto panel_phase = [&](int b, hipStream_t ps, const GemmWork* gw) -> int {
    const int i = bstart[b], wo = bw(b);
    const int q = look ? (b & 1) : 0;
    hipEvent_t ev;
    h->prof_begin(K_PANEL, &ev, ps);
    RNPX_CHECK(phase(i, wo, (T*)vb[q]->p, (T*)tb[q]->p, ps, gw));
    const double mpd = m - i;
    const double fl = (S<T>::cplx ? 4.0 : 1.0) * (2.0 * mpd * wo * wo - 2.0 / 3.0 * wo * wo * wo);
    h->prof_end(K_PANEL, ev, fl, sizeof(T) * 3.0 * mpd * wo, ps);
    return RNPX_OK;
  };
  // Trailing update by outer block b of columns [c0, c1) (main stream unless
  // gw names another).
  auto update = [&](int b, int c0, int c1, const GemmWork* gw) -> int {
    if (c1 <= c0) return RNPX_OK;
    const int i = bstart[b], wo = bw(b);
    const int q = look ? (b & 1) : 0;
    return apply_block_dev<T>(h, 'L', 'C', m - i, c1 - c0, wo, (const T*)vb[q]->p, ldv, (const T*)tb[q]->p, ldt,
                              A + (size_t)i + (size_t)c0 * lda, lda, gw);
  };

  // Q^H B: block b's reflectors applied to B (rows [i, m)).
  auto update_rhs = [&](int b, const GemmWork* gw) -> int {
    if (nrhs == 0) return RNPX_OK;
    const int i = bstart[b], wo = bw(b);
    const int q = look ? (b & 1) : 0;
    return apply_block_dev<T>(h, 'L', 'C', m - i, nrhs, wo, (const T*)vb[q]->p, ldv, (const T*)tb[q]->p, ldt,
                              B + (size_t)i, ldb, gw);
  };
  static const bool report_fallback = getenv("RNPX_REPORT_FALLBACK") != nullptr;
  auto report = [&]() -> int {
    if (!report_fallback) return RNPX_OK;
    std::vector<unsigned> f(npan);
    RNPX_TRY(hipStreamSynchronize(st));
    RNPX_TRY(hipMemcpy(f.data(), (unsigned*)h->sync.p + 1 + npan, sizeof(unsigned) * npan, hipMemcpyDeviceToHost));
    int nf = 0;
    for (unsigned x : f) nf += (x != 0);
    fprintf(stderr, "rnpx: geqrf %dx%d: %d of %d panels fell back to the column-wise kernel\n", m, n, nf, npan);
    return RNPX_OK;
  };
  const GemmWork main_work = {st, &h->wbuf, &h->zbuf, &h->part, false};
  if (!look) {
    for (int b = 0; b < nob; ++b) {
      int s = panel_phase(b, st, &main_work);
      if (s) return s;
      const int c0 = bstart[b + 1];
      if ((s = update(b, c0, n, nullptr))) return s;
      if ((s = update_rhs(b, nullptr))) return s;
    }
    return report();
  }

  int s = h->ensure_pipeline();
  if (s) return s;
  hipStream_t ps = h->pstream, us = st;
  // CU partition: every panel kernel of this call must fit the panel set
  // (their workgroups meet at grid-wide reductions).
  int pcus = h->tune.panel_cus;
  if (pcus == 0) {
    static const int env_cus = [] {
      const char* e = getenv("RNPX_PANEL_CUS");
      return e ? atoi(e) : 0;
    }();
    pcus = env_cus;
  }
  if (pcus > 0 && G0 <= pcus && !h->ext_pstream && h->ensure_partition(pcus) > 0) {
    ps = h->part_ps;
    us = h->part_us;
    h->panel_cap = pcus;
  }
  struct CapReset {
    Handle* h;
    ~CapReset() { h->panel_cap = 0; }
  } cap_reset{h};
  static const int gate_rows = [] {
    const char* e = getenv("RNPX_GATE_ROWS");
    return e ? atoi(e) : 4096;
  }();
  const GemmWork panel_work = {ps, &h->wbuf2, &h->zbuf2, &h->part2, false};
  const GemmWork upd_work = {us, &h->wbuf, &h->zbuf, &h->part, true};
  T* save = nullptr;
  if (split_scan) {
    RNPX_TRY(h->exsave.ensure(sizeof(T) * (size_t)m * w0, h));
    save = (T*)h->exsave.p;
    RNPX_TRY(hipMemcpy2DAsync(save, sizeof(T) * m, A, sizeof(T) * lda, sizeof(T) * m, w0, hipMemcpyDeviceToDevice, st));
  }
  RNPX_TRY(hipEventRecord(h->ev_start, st));
  RNPX_TRY(hipStreamWaitEvent(ps, h->ev_start, 0));
  if (us != st) RNPX_TRY(hipStreamWaitEvent(us, h->ev_start, 0));
  if ((s = panel_phase(0, ps, &panel_work))) return s;
  RNPX_TRY(hipEventRecord(h->ev_pan[0], ps));
  if (split_scan) {
    // the columns right of the first block (untouched until its update)
    bool bad = false;
    RNPX_CHECK(any_nonfinite<T>(h, m, n - w0, A + (size_t)w0 * lda, lda, &bad, st));
    if (bad) {
      RNPX_TRY(hipStreamSynchronize(ps));
      RNPX_TRY(hipMemcpy2DAsync(A, sizeof(T) * lda, save, sizeof(T) * m, sizeof(T) * m, w0, hipMemcpyDeviceToDevice,
                                st));
      return factor_nonfinite();
    }
  }
  for (int b = 0; b < nob; ++b) {
    RNPX_TRY(hipStreamWaitEvent(us, h->ev_pan[b & 1], 0));
    const int c0 = bstart[b + 1];  // first column after block b
    if (c0 >= n) {
      if ((s = update_rhs(b, &upd_work))) return s;
      break;
    }
    if (b + 1 < nob) {
      const int wn = bw(b + 1);  // block b+1's columns
      // Block b+1's columns are brought up to date on the panel stream, right
      // behind panel b: the rest of block b-1's update (which covered them)
      // is the only work of the other stream it waits for.
      if (b > 0) RNPX_TRY(hipStreamWaitEvent(ps, h->ev_rest, 0));
      if ((s = update(b, c0, c0 + wn, &panel_work))) return s;
      // The rest of block b's update is queued before panel b+1's launches:
      // it depends only on panel b, and queued after them it would start
      // only once the host had issued all of them.
      if constexpr (!S<T>::cplx) {
        // Once the panels are the longer chain, the rest of the update waits
        // until every workgroup of block b+1's kernel has a CU.
        // Not under a panel stream shared with other handles (the server's):
        // there block b+1's kernel can queue behind other requests' work and
        // the gate would hold this stream for nothing.
        const int Gb =
            use_block ? cqr_block_groups(m - c0, std::min(256, h->panel_cap > 0 ? h->panel_cap : h->num_cus)) : 0;
        if (wn >= 2 * nb && Gb > 0 && m - c0 <= gate_rows && c0 + wn < n && !h->ext_pstream)
          RNPX_TRY(launch_cqr_gate((const unsigned*)h->sync.p + 1 + 2 * npan + (size_t)kCqrSyncWords * (c0 / nb), Gb,
                                   (unsigned*)h->flag.p + kGateOffWord, us));
      }
      if ((s = update(b, c0 + wn, n, &upd_work))) return s;
      if ((s = update_rhs(b, &upd_work))) return s;
      // block b's V and T are free again once this stream is past here
      RNPX_TRY(hipEventRecord(h->ev_rest, us));
      if ((s = panel_phase(b + 1, ps, &panel_work))) return s;
      RNPX_TRY(hipEventRecord(h->ev_pan[(b + 1) & 1], ps));
      continue;
    }
    if ((s = update(b, c0, n, &upd_work))) return s;
    if ((s = update_rhs(b, &upd_work))) return s;
    RNPX_TRY(hipEventRecord(h->ev_rest, us));
  }
  if (us != st) {
    // the last panel is ordered before the update stream's last wait
    RNPX_TRY(hipEventRecord(h->ev_next, us));
    RNPX_TRY(hipStreamWaitEvent(st, h->ev_next, 0));
  }
  return report();
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

template int geqrf_dev<double>(Handle*, int, int, double*, long, double*, double*, int, long);
template int geqrf_dev<zd>(Handle*, int, int, zd*, long, zd*, zd*, int, long);
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
