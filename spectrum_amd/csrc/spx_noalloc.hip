// spx_fused_peaks_noalloc: spx_fused_peaks for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).  Its
// workspace is a fixed array of device memory, `scratch`, one per GPU, held
// until the process ends; the call itself makes no allocation.
//
// The workspace serves the long-channel search (nfft >= 32768, split path:
// the four-step intermediate, candidate lists and per-block data of a group
// of channels) and the learned-refinement features of the short paths.  Long
// channels run in launch groups of as many channels as the array holds (49 at
// nfft = 65536, 12 at 262144); the results equal spx_fused_peaks' bit for bit.
// Nothing in the array has to start at zero: every kernel that reads part of
// it writes that part first in the same call (the candidate counters are
// cleared by the column kernel).  Calls on one GPU share the array, so calls
// that overlap in time on one GPU (other streams, other contexts, other host
// threads) race and must be ordered by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "spx_host.h"

// 64 MiB of device memory per GPU.
#define SPX_NOALLOC_BYTES (64ull << 20)
static __device__ __align__(256) unsigned char scratch[SPX_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];       // device address of `scratch` on each device
hipError_t g_err[kMaxDevices];  // result of hipGetSymbolAddress on each device
bool g_done[kMaxDevices];
}  // namespace

namespace spx_host {

// Called by spx_create with `device` current: the array's address on this
// device (each GPU holds its own copy at its own address).  Host code uses
// this address, never the array's name, which is a host-side stand-in.
void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) (void)hipGetLastError();
  g_err[device] = (e == hipSuccess && p == nullptr) ? hipErrorInvalidSymbol : e;
  g_ptr[device] = (e == hipSuccess) ? p : nullptr;
  g_done[device] = true;
}

}  // namespace spx_host

extern "C" int spx_fused_peaks_noalloc(spx_ctx* c, const spx_fused_desc* d) {
  using namespace spx_host;
  if (!c) return SPX_EINVAL;
  void* ws = nullptr;
  hipError_t e = hipErrorInvalidDevice;
  if (c->device >= 0 && c->device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[c->device]) {
      ws = g_ptr[c->device];
      e = g_err[c->device];
    }
  }
  if (!ws)
    return fail(c, SPX_EHIP, std::string("spx_fused_peaks_noalloc: the fixed workspace has no device address: ") +
                                 hipGetErrorString(e));
  c->ws_fixed = ws;
  c->ws_fixed_bytes = SPX_NOALLOC_BYTES;
  FusedArgs a;
  int rc = fused_args(c, d, a);
  if (rc == SPX_OK && d->batch != 0) rc = run_fused(c, a);
  c->ws_fixed = nullptr;
  c->ws_fixed_bytes = 0;
  return rc;
}
