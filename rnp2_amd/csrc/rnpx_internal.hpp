// rnpx_internal.hpp -- host-side types shared by the orchestration files:
// the handle (stream, tuning, device workspace) and the templated device
// entry points behind the C ABI.
#pragma once
#include "rnpx_kernels.hpp"
#include "../../include/rnpx.h"

#include <memory>
#include <vector>

namespace rnpx {

struct Handle;

// A growable device allocation.
struct Buffer {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b, Handle* h);
  void release();
};

// Growable pinned host memory for small read-backs inside loops (a copy into
// pageable memory goes through the runtime's staging buffer).
struct HostBuffer {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b, Handle* h);
  void release();
};

// Live per-kernel timing (HIP events recorded on the handle's stream around
// the launches of one kernel family) with the launches' algorithmic work.
enum KernelId { K_UPDATE = 0, K_WGEMM = 1, K_PANEL = 2, K_TFACTOR = 3, K_COUNT = 4 };
struct ProfRec {
  int kid;
  hipEvent_t a, b;
  double flops, bytes;
};

// Words of the handle's 64-byte flag buffer (zero at creation): [0] the
// sticky abort flag of the bounded waits, [kGateOffWord] set once a
// look-ahead gate has given up (later gates return at once).
constexpr int kGateOffWord = 1;

struct Handle {
  int device = 0;
  int num_cus = 1;
  hipStream_t stream = nullptr;  // stream all work is queued on
  hipStream_t own = nullptr;     // the handle's own stream
  bool own_stream = false;
  size_t budget = 0;             // 0 = unbounded
  rnpx_tuning tune = {0, 0, 0, 0};
  Buffer vbuf, vfull, gram, tbuf, wbuf, zbuf, part, ppart, rowbuf, sync, flag, misc;
  Buffer vbuf2, gram2, tbuf2;    // second set for the look-ahead pipeline
  hipStream_t pstream = nullptr;  // panel stream of the look-ahead pipeline
  hipStream_t ext_pstream = nullptr;  // shared panel stream supplied by the owner (not destroyed)
  hipEvent_t ev_pan[2] = {nullptr, nullptr}, ev_next = nullptr, ev_start = nullptr, ev_rest = nullptr;
  // CU-partitioned look-ahead: the panel stream and the update stream run on
  // disjoint sets of CUs (stream CU masks), so a panel never queues behind
  // the trailing update's workgroups.  part_cus = CUs of the panel set, 0 when
  // the partition is off or could not be created.
  hipStream_t part_ps = nullptr, part_us = nullptr;
  int part_cus = 0, part_want = -1;
  int panel_cap = 0;  // while > 0: the most workgroups a panel kernel may use
  Buffer stage1, stage2, stage3;  // host-argument staging for the drop-in calls
  Buffer lsbuf, lsnrm;            // least-squares refinement scratch
  Buffer cqr_part, cqr_gsum, cqr_scratch;  // CholeskyQR2 panel scratch
  Buffer wbuf2, zbuf2, part2, misc2;       // GEMM scratch of the panel stream
  Buffer trsm_flags;                       // per-block flags of the blocked triangular solve
  Buffer lqbuf;                            // A^H for the LQ routines
  Buffer qlbuf;                            // reversed reflectors and tau (QL, RQ)
  Buffer lubuf;                            // LU panel scratch, Solve's permutation and gather
  Buffer rfbuf;                            // RefineSolution's residual and |op(A)||x| + |b|
  Buffer invbuf;                           // LU::Invert's n x n right-hand side and U's diagonal
  Buffer exbuf;                            // non-finite flag, trim limits, T and work of the exact path
  Buffer exsave;                           // the first block's columns while the rest is checked
  Buffer rinv;                             // mixed LS: R^{-1}, its recursion's products and step vectors
  Buffer zcqr;                             // complex CholeskyQR2 panel: Q1, the Grams, R1, R1^{-1}, M, partials
  Buffer gemvbuf;                          // partial sums of the few-column products (gemv_dev)
  HostBuffer hpin;                         // pinned read-back of the refinement's norms
  unsigned trsm_epoch = 0;                 // value the flags take in the current solve
  // Every buffer above, for accounting and release.
  std::vector<Buffer*> buffers() {
    return {&vbuf, &vfull, &gram, &tbuf, &wbuf, &zbuf, &part, &ppart, &rowbuf, &sync, &flag, &misc,
            &vbuf2, &gram2, &tbuf2, &stage1, &stage2, &stage3, &lsbuf, &lsnrm, &cqr_part, &cqr_gsum,
            &cqr_scratch, &wbuf2, &zbuf2, &part2, &misc2, &trsm_flags, &lqbuf, &qlbuf, &lubuf, &rfbuf, &invbuf, &exbuf, &exsave, &rinv, &zcqr, &gemvbuf};
  }
  hipError_t sync_all_streams();       // the handle's stream, panel stream and partition streams
  void* batch = nullptr;               // batched-call descriptor buffers
  void (*batch_free)(void*) = nullptr;
  int update_prec = 0;                 // 1: trailing-update products on FP32 MFMA
  size_t exact_lwork = 0;              // the workspace the reference's Factor was given (0: optimal)
  long exact_calls = 0;                // factorizations that took the non-finite-input path
  bool prof = false;
  unsigned prof_mask = ~0u;  // kernel ids (bit per KernelId) that are timed
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  // Bracket one launch: call prof_begin before and prof_end after.
  int prof_begin(int kid, hipEvent_t* a, hipStream_t s = nullptr);
  void prof_end(int kid, hipEvent_t a, double flops, double bytes, hipStream_t s = nullptr);
  // Or: two timing events that the launch itself records (a kernel launched
  // with start and stop events, no marker packets around it), then
  // prof_push; both null when kid is not timed.
  void prof_take(int kid, hipEvent_t* a, hipEvent_t* b);
  void prof_push(int kid, hipEvent_t a, hipEvent_t b, double flops, double bytes);
  int ensure_pipeline();
  int ensure_partition(int cus);
  size_t held_bytes() const;
  static int check_gpu();
  ~Handle();
};

int hip_status(hipError_t e);
// A non-blocking stream on a hardware queue of its own where the runtime
// allows (a CU mask holding every CU), else a plain one.
hipError_t create_own_queue_stream(hipStream_t* s, int num_cus);

// Return from the enclosing int-status function on a HIP error (RNPX_TRY) or
// on a nonzero status (RNPX_CHECK).
#define RNPX_TRY(x)                               \
  do {                                            \
    hipError_t _e = (x);                          \
    if (_e != hipSuccess) return hip_status(_e);  \
  } while (0)
#define RNPX_CHECK(x)        \
  do {                       \
    const int _s = (x);      \
    if (_s != 0) return _s;  \
  } while (0)
int create_handle(Handle** out, size_t budget);
Handle* default_handle(int* status);
int sync_handle(Handle* h);
int prof_collect(Handle* h, int kid, long* launches, double* ms, double* flops, double* bytes);
void prof_reset(Handle* h);

// B (m x nrhs, optional): Q^H B is formed alongside, each block reflector
// applied to B right after its trailing update (the least-squares drivers).
template <typename T>
int geqrf_dev(Handle* h, int m, int n, T* A, long lda, T* tau, T* B = nullptr, int nrhs = 0, long ldb = 0);
// Complex panel (nb = 32 columns, mp >= 2 nb) by CholeskyQR2 and
// Householder reconstruction (rnpx_zcqr.hip); a rejected panel sets *fail
// and is factored by the column-wise kernel (pa, rpt) gated on it.
int zcqr_panel(Handle* h, int mp, int nb, zd* A, long lda, zd* tau, zd* V, long ldv, zd* Tm, long ldt, unsigned* fail,
               const PanelArgs<zd>& pa, int rpt, hipStream_t st);
// Inf/NaN anywhere in the m x n matrix: checked on stream st, which is then
// synchronized.
template <typename T> int any_nonfinite(Handle* h, int m, int n, const T* A, long lda, bool* out, hipStream_t st);
// Reflector::LastNonzeroColumnLimit (which 0) / LastNonzeroRowLimit (which 1)
// of a device matrix, read back (synchronous).
template <typename T> int last_nonzero_limit(Handle* h, int which, int m, int n, const T* a, long lda, int* out);
// QR::Factor with the reference's blocking and zero trims (input holding
// Inf or NaN); lwork as given to the reference (0: optimal).
template <typename T> int geqrf_exact(Handle* h, int m, int n, T* A, long lda, T* tau, size_t lwork);
template <typename T> int gels_dev(Handle* h, int m, int n, int nrhs, T* A, long lda, T* tau, T* B, long ldb);
// Y (M x N, 1 <= N <= 4) = alpha op(A) X + beta Y, op(A) M x K, on stream st
// (rnpx_gemv.hip): HBM-bound products with few right-hand sides, slices of the
// reduction added in a fixed order.
template <typename T>
int gemv_dev(Handle* h, char ta, int M, int N, int K, T alpha, const T* A, long lda, const T* X, long ldx, T beta,
             T* Y, long ldy, hipStream_t st);
int gels_mixed_dev(Handle* h, int m, int n, int nrhs, const double* A, long lda, const double* B, long ldb, double* X,
                   long ldx, int max_iter, double tol, int* iters, double* berr);
int ls_berr_dev(Handle* h, int m, int n, int nrhs, const double* A, long lda, const double* B, long ldb,
                const double* X, long ldx, double* berr);
template <typename T> int geqrf_batched(Handle* h, int count, const rnpx_qr_desc* desc, T* A, T* tau, T* B);
// Batched set-up shared by the batched entry points: problems of at most
// kBatchLdsBytes run from LDS in one-wave teams (the small tier), the rest in
// place in 512-thread teams (the large tier); each tier's list is sorted by
// cost, largest first.  Returns -3 on a bad descriptor.
constexpr int kBatchLdsBytes = 40 * 1024;  // per one-wave team: four teams per CU (default)
constexpr int kBatchBlockedMin = 33;       // real problems of at least this many columns: blocked (default)
struct BatchSetup {
  std::vector<int> small, large;
  int lds_bytes = kBatchLdsBytes;       // per one-wave team of the small tier (also the tier threshold)
  int blocked_min = kBatchBlockedMin;   // the large tier's blocked-path threshold
};
// Fills the tiers with bs->lds_bytes as the threshold; call batch_params
// first to take the autotuned values.
template <typename T> int batch_setup(int count, const rnpx_qr_desc* desc, BatchSetup* out);
template <typename T> void batch_params(BatchSetup* out);
// The persisted autotuning table (rnpx_tune.cpp): the entry nearest to an
// m x n factorization of element type dtype ('d' or 'z'), and the batched
// tiers' parameters.  False: no entry (defaults apply).
bool tuning_factor(char dtype, int m, int n, int* nb, int* ksplit, int* panel_cus = nullptr);
bool tuning_batched(char dtype, int* lds, int* blocked_min);
int tuning_load(const char* path);

// Allocation-free batched entry point's set-up: resolves this device's copy of
// the static scratch array once per device (called by create_handle with the
// device current).
int noalloc_setup(int device);
struct BatchedPlan;
template <typename T>
int geqrf_batched_plan(Handle* h, int count, const rnpx_qr_desc* desc, T* A, T* tau, T* B, BatchedPlan** out);
int batched_plan_run(BatchedPlan* p);
int batched_plan_destroy(BatchedPlan* p);
template <typename T>
int ormqr_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc);
template <typename T> int orgqr_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau);
// LQ (rnpx_lq.cpp): LQ::Factor, LQ::MultQ, LQ::GenerateQ on device memory.
template <typename T> int gelqf_dev(Handle* h, int m, int n, T* A, long lda, T* tau);
template <typename T>
int ormlq_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc);
template <typename T> int orglq_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau);
// QL and RQ (rnpx_qlrq.cpp): the QR and LQ paths on the row- and
// column-reversed matrix.
template <typename T> int geqlf_dev(Handle* h, int m, int n, T* A, long lda, T* tau);
template <typename T>
int ormql_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc);
template <typename T> int orgql_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau);
template <typename T> int gerqf_dev(Handle* h, int m, int n, T* A, long lda, T* tau);
template <typename T>
int ormrq_dev(Handle* h, char side, char trans, int m, int n, int k, const T* A, long lda, const T* tau, T* C,
              long ldc);
template <typename T> int orgrq_dev(Handle* h, int m, int n, int k, T* A, long lda, const T* tau);
// Hessenberg::MultQ and GenerateQ (rnpx_hess.cpp): the QR routines on the
// reflector block one row below the diagonal.
template <typename T>
int ormhr_dev(Handle* h, char side, char trans, int m, int n, int ilo, int ihi, const T* A, long lda, const T* tau,
              T* C, long ldc);
template <typename T> int orghr_dev(Handle* h, int n, int ilo, int ihi, T* A, long lda, const T* tau);
// LU::Factor and LU::Solve (rnpx_lu_driver.cpp); pivots 0-based absolute rows on
// the host; getrf_dev synchronizes (its info needs the panels' results).
template <typename T> int getrf_dev(Handle* h, int m, int n, T* A, long lda, int* piv_host, int* info);
template <typename T>
int getrs_dev(Handle* h, char trans, int n, int nrhs, const T* A, long lda, const int* ipiv, T* B, long ldb);
// LU::Invert: inv(A) in place from the factors; *info = 0, or -(i+1) for a
// zero U_ii (A untouched).  Returns a status.
template <typename T> int getri_dev(Handle* h, int n, T* A, long lda, const int* ipiv, int* info);
// LU::RefineSolution (rnpx_lu_driver.cpp); A, AF, B, X on the device, ipiv,
// ferr and berr on the host.
template <typename T>
int gerfs_dev(Handle* h, char trans, int n, int nrhs, const T* A, long lda, const T* AF, long ldaf, const int* ipiv,
              const T* B, long ldb, T* X, long ldx, double* ferr, double* berr);
// op(M) X = B on the handle's stream (rnpx_ls.cpp), M the uplo triangle of R.
template <typename T>
int trsm_tri_dev(Handle* h, char uplo, char op, char diag, int n, int nrhs, const T* R, long ldr, T* B, long ldb);
template <typename T>
int larft_dev(Handle* h, int n, int k, const T* V, long ldv, const T* tau, T* Tout, long ldt);
template <typename T>
int larfb_dev(Handle* h, char side, char trans, int m, int n, int k, const T* V, long ldv, const T* Tm, long ldt,
              T* C, long ldc);

}  // namespace rnpx
