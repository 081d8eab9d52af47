// Host-internal interface shared by the C-ABI translation units: the
// context object and its table / scratch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <tuple>

#include "spx_kernels.h"

struct HostStage;  // pinned and device double buffers of the host-streaming entry points

struct spx_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  std::string name;
  uint64_t hbm = 0;
  std::string err;
  std::map<std::tuple<int, uint32_t, uint32_t>, void*> tables;  // (kind, a, b) -> device table
  void* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  // stored spectrum of spx_fused_peaks_top where its search needs one (the
  // transform of those lengths uses `scratch` itself)
  void* spec_buf = nullptr;
  uint64_t spec_bytes = 0;
  const void* refiner = nullptr;  // deployed learned-refiner block (caller-owned device memory)
  void* mlp_partial = nullptr;    // per-workgroup gradient partials
  uint64_t mlp_partial_bytes = 0;
  HostStage* hstage = nullptr;             // created by the first *_host call
  uint64_t host_chunk_bytes = 32ull << 20;  // per-chunk transfer size of the *_host calls
  // set only during spx_fused_peaks_noalloc: the fixed workspace that call
  // uses instead of `scratch`; tables are then looked up, never created
  void* ws_fixed = nullptr;
  uint64_t ws_fixed_bytes = 0;
};

namespace spx_host {

enum TableKind { TW = 0, TW4 = 1, WIN = 2, MIX = 3, LG = 4 };

int fail(spx_ctx* c, int code, const std::string& msg);
int hip_fail(spx_ctx* c, hipError_t e, const char* what);
// SPX_ENOMEM after a failed device or pinned allocation
int oom(spx_ctx* c, const char* what);
// Device table of kind TW (a = length, b = radix order), TW4 (a = length),
// WIN (a = window type, b = length), MIX (a = compute type) or LG (a = length
// 2^19 .. 2^24: the Stockham tables of its LDS passes), generated in FP64 once
// per context.
int get_table(spx_ctx* c, int kind, uint32_t a, uint32_t b, const void** out);
int window_ptr(spx_ctx* c, uint32_t window, uint32_t n_point, const float** w);
// Releases the host-streaming buffers, streams and events.
void host_stage_free(spx_ctx* c);
// Channels per launch group of the split long-channel path (bounds its workspace).
uint32_t fourstep_chunk(uint32_t nfft, uint32_t batch, int learn = 0, uint32_t kcap = 0);
// Validates a fused descriptor into kernel arguments; runs them.
int fused_args(spx_ctx* c, const spx_fused_desc* d, FusedArgs& a);
int run_fused(spx_ctx* c, FusedArgs& a);
// Device address of the fixed workspace array on `device` (spx_noalloc.hip);
// taken once per device at spx_create, with that device current.
void noalloc_init(int device);

}  // namespace spx_host

#define SPX_HIP(ctx, expr)                                        \
  do {                                                            \
    hipError_t e_ = (expr);                                       \
    if (e_ != hipSuccess) return spx_host::hip_fail(ctx, e_, #expr); \
  } while (0)
