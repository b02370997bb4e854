// rnpx_noalloc.hip -- the batched QR / least-squares entry point that
// allocates no device memory, for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The batched kernels' scratch -- the descriptor table, the cost-sorted
// problem order and the two work-queue counters -- lives in one file-scope
// device array, `scratch`, of kNoAllocScratchBytes (holds up to
// kNoAllocMaxProblems problems).  Each GPU holds its own copy until the
// process ends.  Host code never uses the array's name as an address: at
// set-up (create_handle, once per device, with that device current) its
// device address is resolved with hipGetSymbolAddress and kept per device;
// every copy and kernel argument uses that address.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents: the counters are cleared by a fill kernel at each launch, and
// calls that overlap on one GPU (different streams, or other threads) race.
// Calls queued on one stream are ordered by the stream.
//
// Both tiers run on the handle's stream one after the other (the regular
// batched entry point overlaps them on two streams, whose creation this one
// avoids); the per-problem results are the same bits either way.
#include "rnpx_internal.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

#define RNPX_TRY_HIP(x)                          \
  do {                                           \
    hipError_t _e = (x);                         \
    if (_e != hipSuccess) return hip_status(_e); \
  } while (0)

namespace rnpx {

constexpr int kNoAllocMaxProblems = 16384;
constexpr size_t kNoAllocDescBytes = sizeof(rnpx_qr_desc) * (size_t)kNoAllocMaxProblems;
constexpr size_t kNoAllocOrderBytes = sizeof(int) * (size_t)kNoAllocMaxProblems;
constexpr size_t kNoAllocScratchBytes = kNoAllocDescBytes + kNoAllocOrderBytes + 256;
static_assert(kNoAllocDescBytes % 256 == 0 && kNoAllocOrderBytes % 256 == 0, "scratch sections stay aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};        // the device address of `scratch` on each device
hipError_t g_err[kMaxDevices] = {};   // the lookup's status where it failed
bool g_done[kMaxDevices] = {};
}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return RNPX_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[device]) return g_ptr[device] ? RNPX_OK : hip_status(g_err[device]);
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e == hipSuccess && cur != device) e = hipErrorInvalidDevice;  // must be called with the device current
  void* p = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_done[device] = true;
  g_ptr[device] = (e == hipSuccess) ? p : nullptr;
  g_err[device] = e;
  if (e != hipSuccess) (void)hipGetLastError();
  return e == hipSuccess ? RNPX_OK : hip_status(e);
}

static void* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

template <typename T>
static int geqrf_batched_noalloc(Handle* h, int count, const rnpx_qr_desc* desc, T* A, T* tau, T* B) {
  if (count < 0) return -2;
  if (count == 0) return RNPX_OK;
  if (!desc) return -3;
  if (!A) return -4;
  if (!tau) return -5;
  if (count > kNoAllocMaxProblems) return RNPX_ERR_UNSUPPORTED;
  BatchSetup bs;
  batch_params<T>(&bs);
  int s = batch_setup<T>(count, desc, &bs);
  if (s) return s;
  unsigned char* base = (unsigned char*)scratch_of(h->device);
  if (!base) return RNPX_ERR_HIP;
  rnpx_qr_desc* d_desc = (rnpx_qr_desc*)base;
  int* d_order = (int*)(base + kNoAllocDescBytes);
  unsigned* d_ctr = (unsigned*)(base + kNoAllocDescBytes + kNoAllocOrderBytes);
  std::vector<int> order(bs.large);
  order.insert(order.end(), bs.small.begin(), bs.small.end());
  hipStream_t st = h->stream;
  RNPX_TRY_HIP(hipMemcpyAsync(d_desc, desc, sizeof(rnpx_qr_desc) * (size_t)count, hipMemcpyHostToDevice, st));
  RNPX_TRY_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, st));
  // the host order vector dies with this call
  RNPX_TRY_HIP(hipStreamSynchronize(st));
  const int nl = (int)bs.large.size(), ns = (int)bs.small.size();
  if (nl > 0)
    RNPX_TRY_HIP(launch_batched<T>(1, nl, d_order, d_desc, A, tau, B, d_ctr, batched_large_teams<T>(nl, h->num_cus), bs.blocked_min,
                                   st));
  if (ns > 0) {
    const int per_cu = std::max(1, (160 * 1024) / bs.lds_bytes);
    RNPX_TRY_HIP(launch_batched<T>(0, ns, d_order + nl, d_desc, A, tau, B, d_ctr + 1, std::min(ns, per_cu * h->num_cus),
                                   bs.lds_bytes, st));
  }
  return RNPX_OK;
}

}  // namespace rnpx

extern "C" int rnpx_dgeqrf_batched_noalloc(rnpx_handle h, int count, const rnpx_qr_desc* desc, double* A,
                                           double* tau, double* B) {
  if (!h) return -1;
  return rnpx::geqrf_batched_noalloc<double>(reinterpret_cast<rnpx::Handle*>(h), count, desc, A, tau, B);
}
