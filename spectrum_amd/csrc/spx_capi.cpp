// Host side of the C ABI: contexts, FP64-generated tables, argument checks,
// scratch management and kernel dispatch.  No compute happens on the host:
// every entry point launches HIP kernels on the context stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "spx_kernels.h"

#include "spx_host.h"

namespace spx_host {

int fail(spx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// The runtime's last-error slot is cleared, so that the failure is reported
// by this call only and not again by the next launch check.
int hip_fail(spx_ctx* c, hipError_t e, const char* what) {
  (void)hipGetLastError();
  return fail(c, SPX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int oom(spx_ctx* c, const char* what) {
  (void)hipGetLastError();
  return fail(c, SPX_ENOMEM, what);
}

bool is_pow2(uint32_t n) { return n >= 2 && (n & (n - 1)) == 0; }

// IEEE binary16 and bfloat16 encodings, round to nearest even (finite inputs).
uint16_t f32_to_f16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const int32_t e = (int32_t)((x >> 23) & 0xffu) - 127 + 15;
  uint32_t mant = x & 0x7fffffu;
  if (((x >> 23) & 0xffu) == 0 && mant == 0) return (uint16_t)sign;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {  // subnormal half
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) ++h;
  return (uint16_t)(sign | h);
}
float f16_bits_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float v;
  if (e == 0) v = std::ldexp((float)m, -24);
  else if (e == 31) v = INFINITY;
  else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
  uint32_t x;
  std::memcpy(&x, &v, 4);
  x |= sign;
  std::memcpy(&v, &x, 4);
  return v;
}
uint16_t f32_to_bf16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}
std::vector<float> make_mixed_tables(uint32_t compute);

float bf16_bits_to_f32(uint16_t h) {
  const uint32_t x = (uint32_t)h << 16;
  float v;
  std::memcpy(&v, &x, 4);
  return v;
}

constexpr uint32_t kMaxN = 1u << 24;  // 2^19 .. 2^24: spx_large.hip

// Per-pass twiddle tables of a length-L transform under radix order V (V = 0:
// radix-16 passes, the leftover radix last; V = 1: the leftover first).  Pass
// p >= 1, radix R, sub-length NS = product of the earlier radices, stores
// [k][r] = W_(NS*R)^(r*k), k < NS, r < R; passes concatenated.  FP64, rounded once.
std::vector<float> make_twiddles(uint32_t L, int V) {
  int log = 0;
  while ((1u << log) < L) ++log;
  const int np = (log + 3) / 4;
  const int rx = (log % 4) ? (1 << (log % 4)) : 16;
  auto radix = [&](int p) { return np == 1 ? rx : (V == 0 ? (p == np - 1 ? rx : 16) : (p == 0 ? rx : 16)); };
  std::vector<float> t;
  uint64_t ns = (uint64_t)radix(0);
  for (int p = 1; p < np; ++p) {
    const int R = radix(p);
    for (uint64_t k = 0; k < ns; ++k)
      for (int r = 0; r < R; ++r) {
        const uint64_t m = ns * (uint64_t)R;
        const uint64_t e = (k * (uint64_t)r) % m;
        const double a = 2.0 * M_PI * ((double)e / (double)m);
        t.push_back((float)std::cos(a));
        t.push_back((float)-std::sin(a));
      }
    ns *= (uint64_t)R;
  }
  if (t.empty()) t.assign(2, 0.0f);
  return t;
}

// four-step table: entry [k2 * 256 + n1] = W_N^(n1 * k2), N = 256 * N2.
std::vector<float> make_fourstep(uint32_t N) {
  const uint32_t N1 = 256, N2 = N / N1;
  std::vector<float> t(2 * (size_t)N);
  for (uint32_t k2 = 0; k2 < N2; ++k2)
    for (uint32_t n1 = 0; n1 < N1; ++n1) {
      const uint64_t e = ((uint64_t)n1 * k2) % N;
      const double a = 2.0 * M_PI * ((double)e / (double)N);
      const size_t i = (size_t)k2 * N1 + n1;
      t[2 * i] = (float)std::cos(a);
      t[2 * i + 1] = (float)-std::sin(a);
    }
  return t;
}

// Symmetric windows over n_point samples, denominator n_point - 1
// (include/spectrum/kernels/fft_processor_kernels_rocm.hpp:99-109);
// n_point == 1 gives w = 1 instead of the reference's division by zero.
std::vector<float> make_window(uint32_t type, uint32_t n) {
  std::vector<float> w(n, 1.0f);
  if (n < 2) return w;
  for (uint32_t p = 0; p < n; ++p) {
    const double th = 2.0 * M_PI * (double)p / (double)(n - 1);
    double v = 1.0;
    if (type == SPX_WINDOW_HANN) v = 0.5 * (1.0 - std::cos(th));
    else if (type == SPX_WINDOW_HAMMING) v = 0.54 - 0.46 * std::cos(th);
    else if (type == SPX_WINDOW_BLACKMAN) v = 0.42 - 0.5 * std::cos(th) + 0.08 * std::cos(2.0 * th);
    w[p] = (float)v;
  }
  return w;
}

int get_table(spx_ctx* c, int kind, uint32_t a, uint32_t b, const void** out) {
  auto key = std::make_tuple(kind, a, b);
  auto it = c->tables.find(key);
  if (it != c->tables.end()) {
    *out = it->second;
    return SPX_OK;
  }
  if (c->ws_fixed)
    return fail(c, SPX_EINVAL,
                "spx_fused_peaks_noalloc: the tables of this length and window are not in the context yet; make one "
                "spx_fused_peaks call with the same nfft, n_point and window first (it creates them)");
  std::vector<float> host;
  if (kind == TW) {
    host = make_twiddles(a, (int)b);
    if (a <= 16384) {  // lengths with a compile-time plan: the table must match it
      const int want = spx_twiddle_table_len(a, (int)b);
      if (want < 0 || (size_t)std::max(want, 1) * 2 != host.size())
        return fail(c, SPX_EINVAL, "twiddle table size mismatch");
    }
  }
  else if (kind == TW4) host = make_fourstep(a);
  else if (kind == MIX) host = make_mixed_tables(a);
  else if (kind == LG) {
    host.resize(spx_large_tables(a, nullptr));
    if (host.empty()) return fail(c, SPX_EINVAL, "no large-length plan for this nfft");
    spx_large_tables(a, host.data());
  }
  else host = make_window(a, b);
  void* d = nullptr;
  SPX_HIP(c, hipSetDevice(c->device));
  hipError_t e = hipMalloc(&d, host.size() * sizeof(float));
  if (e != hipSuccess) return oom(c, "table allocation failed");
  SPX_HIP(c, hipMemcpy(d, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice));
  c->tables[key] = d;
  *out = d;
  return SPX_OK;
}

int ensure_scratch(spx_ctx* c, uint64_t bytes) {
  if (c->ws_fixed) {
    if (bytes > c->ws_fixed_bytes)
      return fail(c, SPX_ECAP, "spx_fused_peaks_noalloc: needs " + std::to_string(bytes) +
                                   " bytes of workspace, the fixed workspace holds " +
                                   std::to_string(c->ws_fixed_bytes));
    return SPX_OK;
  }
  if (c->scratch_bytes >= bytes) return SPX_OK;
  SPX_HIP(c, hipSetDevice(c->device));
  if (c->scratch) {
    SPX_HIP(c, hipStreamSynchronize(c->stream));
    SPX_HIP(c, hipFree(c->scratch));
    c->scratch = nullptr;
    c->scratch_bytes = 0;
  }
  hipError_t e = hipMalloc(&c->scratch, bytes);
  if (e != hipSuccess) {
    c->scratch = nullptr;
    return oom(c, "scratch allocation failed");
  }
  c->scratch_bytes = bytes;
  return SPX_OK;
}

int check_fft_shape(spx_ctx* c, const void* in, const void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                    uint32_t window) {
  if (!c) return SPX_EINVAL;
  if (!in || !out) return fail(c, SPX_EINVAL, "null device pointer");
  if (!is_pow2(nfft) || nfft > kMaxN) return fail(c, SPX_EINVAL, "nfft must be a power of two in [2, 2^24]");
  if (n_point == 0 || n_point > nfft) return fail(c, SPX_EINVAL, "n_point must be in [1, nfft]");
  if (window > SPX_WINDOW_BLACKMAN) return fail(c, SPX_EINVAL, "unknown window type");
  (void)batch;
  return SPX_OK;
}

// Twiddle table of a short (one workgroup per channel) fused launch; one
// workgroup per channel.
int prepare(spx_ctx* c, FusedArgs& a, int* grid) {
  const void* t;
  int rc;
  if ((rc = get_table(c, TW, a.nfft, 0, &t))) return rc;
  a.tw_a = (const float2*)t;
  *grid = (int)a.batch;
  return SPX_OK;
}

// Learned refinement on the short paths: the fused kernel writes the
// features of every stored peak (32 bytes each) into the context's scratch,
// then k_mlp_refine updates those records.
int learned_features(spx_ctx* c, FusedArgs& a) {
  a.mlp_feat = nullptr;
  if (a.refine != 2 || a.peak_mode != 1 || a.kcap == 0 || !a.peaks) return SPX_OK;
  int rc = ensure_scratch(c, (uint64_t)a.batch * a.kcap * 32);
  if (rc) return rc;
  a.mlp_feat = c->ws_fixed ? c->ws_fixed : c->scratch;
  return SPX_OK;
}

int learned_refine(spx_ctx* c, const FusedArgs& a) {
  if (!a.mlp_feat) return SPX_OK;
  SPX_HIP(c, spx_launch_mlp_refine(a.mlp, a.mlp_feat, a.peaks, a.counts, a.batch, a.kcap, a.fs / (float)a.nfft,
                                   c->stream));
  return SPX_OK;
}

// Channels per launch group of the split long-channel peak search: bounds the
// workspace (Z, candidate lists, edge rows) to about 2 GiB (6 GiB with the
// learned-refinement buffers).
uint32_t fourstep_chunk(uint32_t nfft, uint32_t batch, int learn, uint32_t kcap) {
  const char* env = std::getenv("SPX_FS_CHUNK");
  uint32_t ch = env ? (uint32_t)std::strtoul(env, nullptr, 10) : 0;
  if (ch == 0) {
    // learned refinement keeps each candidate's neighbourhood (CAND_CAP x 32 B
    // per channel): a larger budget keeps a 1024-channel batch in one group
    const uint64_t per = spx_fs_workspace_bytes_ex(nfft, 1, learn, kcap);
    const uint64_t budget = learn ? (6ull << 30) : (2ull << 30);
    ch = (uint32_t)std::max<uint64_t>(64, budget / std::max<uint64_t>(per, 1));
  }
  return std::min(ch, batch);
}

int run_fused(spx_ctx* c, FusedArgs& a) {
  if (a.batch == 0) return SPX_OK;
  int grid = 0, rc;
  if (spx_large_is(a.nfft)) {
    // 2^19 .. 2^24: three Stockham passes through global memory (two of them
    // L-point transforms in LDS tiles), peak search on the stored spectrum
    int r0, lgs[3], np;
    if (spx_large_plan(a.nfft, &r0, lgs, &np)) return fail(c, SPX_EINVAL, "no large-length plan for this nfft");
    const void* t;
    const float2* twt[3] = {nullptr, nullptr, nullptr};
    if ((rc = get_table(c, LG, a.nfft, 0, &t))) return rc;
    for (int p = 0; p < np; ++p) {
      const void* tp;
      if ((rc = get_table(c, TW, 1u << lgs[p], 0, &tp))) return rc;
      twt[p] = (const float2*)tp;
    }
    const uint32_t lk = (a.refine == 2 && a.mlp != nullptr && a.peak_mode == 1) ? a.kcap : 0u;
    const uint64_t per = spx_large_workspace_bytes(a.nfft, 1, lk);
    const uint64_t budget = c->ws_fixed ? c->ws_fixed_bytes : (2ull << 30);
    uint32_t chunk = (uint32_t)std::min<uint64_t>(a.batch, std::max<uint64_t>(c->ws_fixed ? 0 : 1, budget / per));
    if (chunk == 0) return ensure_scratch(c, per);  // the fixed workspace is too small: reports the shortfall
    if ((rc = ensure_scratch(c, spx_large_workspace_bytes(a.nfft, chunk, lk)))) return rc;
    SPX_HIP(c, hipSetDevice(c->device));
    SPX_HIP(c, spx_launch_large(&a, (const float2*)t, twt, c->ws_fixed ? c->ws_fixed : c->scratch, chunk,
                                c->stream));
    return SPX_OK;
  }
  if (spx_fused_is_large(a.nfft)) {
    // split four-step path (spx_fourstep.hip): three launches for the peak
    // searches, two for the dense outputs
    const void* t;
    if ((rc = get_table(c, TW, 256, 0, &t))) return rc;
    a.tw_a = (const float2*)t;
    if ((rc = get_table(c, TW, a.nfft / 256, 0, &t))) return rc;
    a.tw_b = (const float2*)t;
    if ((rc = get_table(c, TW4, a.nfft, 0, &t))) return rc;
    a.tw_4 = (const float2*)t;
    SPX_HIP(c, hipSetDevice(c->device));
    if (a.peak_mode == 0) {
      const uint64_t per = spx_fs_dense_workspace_bytes(a.nfft, 1);
      const uint32_t chunk = (uint32_t)std::min<uint64_t>(a.batch, std::max<uint64_t>(64, (2ull << 30) / per));
      if ((rc = ensure_scratch(c, spx_fs_dense_workspace_bytes(a.nfft, chunk)))) return rc;
      SPX_HIP(c, spx_launch_fourstep_dense(&a, c->scratch, chunk, c->stream));
      return SPX_OK;
    }
    // learned refinement: the refiner's features and the refiner run per chunk
    const int learn = (a.refine == 2 && a.mlp != nullptr) ? 1 : 0;
    auto wsb = [&](uint32_t n) { return spx_fs_workspace_bytes_ex(a.nfft, n, learn, a.kcap); };
    uint32_t chunk = fourstep_chunk(a.nfft, a.batch, learn, a.kcap);
    if (c->ws_fixed) {
      // as many channels per launch group as the fixed workspace holds
      const uint64_t per = wsb(1) - wsb(0);
      const uint64_t fit = c->ws_fixed_bytes > wsb(0) ? (c->ws_fixed_bytes - wsb(0)) / per : 0;
      if (fit == 0) return ensure_scratch(c, wsb(1));  // reports the shortfall
      chunk = (uint32_t)std::min<uint64_t>(fit, a.batch);
    }
    if ((rc = ensure_scratch(c, wsb(chunk)))) return rc;
    SPX_HIP(c, spx_launch_fourstep(&a, c->ws_fixed ? c->ws_fixed : c->scratch, chunk, c->stream));
    return SPX_OK;
  }
  if ((rc = prepare(c, a, &grid))) return rc;
  if ((rc = learned_features(c, a))) return rc;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_fused(&a, grid, c->stream));
  return learned_refine(c, a);
}

// Cosine coefficients of the symmetric windows, w(n) = c0 - c1 cos(2 pi
// n/(n-1)) + c2 cos(4 pi n/(n-1)) (the formulas of make_window), for kernels
// that evaluate the window instead of reading its table.
void window_coefs(uint32_t window, FusedArgs& a) {
  a.wkind = window;
  a.wc0 = a.wc1 = a.wc2 = 0.0f;
  if (window == SPX_WINDOW_HANN) { a.wc0 = 0.5f; a.wc1 = 0.5f; }
  else if (window == SPX_WINDOW_HAMMING) { a.wc0 = 0.54f; a.wc1 = 0.46f; }
  else if (window == SPX_WINDOW_BLACKMAN) { a.wc0 = 0.42f; a.wc1 = 0.5f; a.wc2 = 0.08f; }
}

int window_ptr(spx_ctx* c, uint32_t window, uint32_t n_point, const float** w) {
  *w = nullptr;
  if (window == SPX_WINDOW_NONE) return SPX_OK;
  const void* t;
  int rc = get_table(c, WIN, window, n_point, &t);
  if (rc) return rc;
  *w = (const float*)t;
  return SPX_OK;
}

int dense_fft(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft, uint32_t window,
              int kind, float scale) {
  int rc = check_fft_shape(c, in, out, batch, n_point, nfft, window);
  if (rc) return rc;
  if (in == out) return fail(c, SPX_EINVAL, "in-place transform is not supported (in == out)");
  FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in;
  a.batch = batch;
  a.n_point = n_point;
  a.nfft = nfft;
  a.dtype = SPX_DTYPE_C64;
  if ((rc = window_ptr(c, window, n_point, &a.window))) return rc;
  if (kind == 0) a.spec_out = (float2*)out;
  else {
    a.mag_out = (float*)out;
    a.mag_kind = (uint32_t)kind;
    a.mag_scale = scale;
  }
  a.peak_mode = 0;
  a.s0 = 0;
  a.s1 = nfft;
  return run_fused(c, a);
}

}  // namespace spx_host

using namespace spx_host;


extern "C" {

const char* spx_version(void) { return "spectrum_mi355x 0.1.0 (gfx950)"; }

int spx_create(int device, spx_ctx** out) {
  if (!out) return SPX_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SPX_EHIP;
  if (device < 0 || device >= n) return SPX_EINVAL;
  spx_ctx* c = new spx_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return SPX_EHIP; }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) { delete c; return SPX_EHIP; }
  c->cu_count = p.multiProcessorCount;
  c->name = p.name;
  c->hbm = p.totalGlobalMem;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SPX_EHIP; }
  c->stream = c->own_stream;
  noalloc_init(device);
  *out = c;
  return SPX_OK;
}

int spx_destroy(spx_ctx* c) {
  if (!c) return SPX_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->tables) (void)hipFree(kv.second);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->spec_buf) (void)hipFree(c->spec_buf);
  if (c->mlp_partial) (void)hipFree(c->mlp_partial);
  host_stage_free(c);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return SPX_OK;
}

int spx_set_stream(spx_ctx* c, void* s) {
  if (!c) return SPX_EINVAL;
  hipStream_t ns = s ? (hipStream_t)s : c->own_stream;
  if (ns == c->stream) return SPX_OK;
  // calls still in flight on the old stream share the context's workspace
  // with the next ones: the new stream waits for them
  SPX_HIP(c, hipSetDevice(c->device));
  hipEvent_t e;
  SPX_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipError_t r = hipEventRecord(e, c->stream);
  if (r == hipSuccess) r = hipStreamWaitEvent(ns, e, 0);
  (void)hipEventDestroy(e);
  SPX_HIP(c, r);
  c->stream = ns;
  return SPX_OK;
}

void* spx_get_stream(spx_ctx* c) { return c ? (void*)c->stream : nullptr; }

int spx_synchronize(spx_ctx* c) {
  if (!c) return SPX_EINVAL;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, hipStreamSynchronize(c->stream));
  return SPX_OK;
}

const char* spx_last_error(const spx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int spx_device_info(spx_ctx* c, char* name, size_t name_len, int* cu_count, uint64_t* hbm_bytes) {
  if (!c) return SPX_EINVAL;
  if (name && name_len) {
    std::strncpy(name, c->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = c->cu_count;
  if (hbm_bytes) *hbm_bytes = c->hbm;
  return SPX_OK;
}

int spx_alloc(spx_ctx* c, uint64_t bytes, void** out) {
  if (!c || !out) return SPX_EINVAL;
  *out = nullptr;
  SPX_HIP(c, hipSetDevice(c->device));
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
    *out = nullptr;
    return oom(c, "hipMalloc failed");
  }
  return SPX_OK;
}

int spx_free(spx_ctx* c, void* p) {
  if (!c) return SPX_EINVAL;
  SPX_HIP(c, hipSetDevice(c->device));
  if (p) SPX_HIP(c, hipFree(p));
  return SPX_OK;
}

int spx_memcpy_h2d(spx_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return SPX_EINVAL;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  SPX_HIP(c, hipStreamSynchronize(c->stream));
  return SPX_OK;
}

int spx_memcpy_d2h(spx_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return SPX_EINVAL;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  SPX_HIP(c, hipStreamSynchronize(c->stream));
  return SPX_OK;
}

int spx_fft_c2c(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                uint32_t window) {
  return dense_fft(c, in, out, batch, n_point, nfft, window, 0, 1.0f);
}

int spx_fft_mag_phase(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                      uint32_t window) {
  return dense_fft(c, in, out, batch, n_point, nfft, window, SPX_MAG_PHASE, 1.0f);
}

int spx_fft_magnitude(spx_ctx* c, const void* in, float* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                      uint32_t window, int squared, float scale) {
  return dense_fft(c, in, out, batch, n_point, nfft, window, squared ? SPX_MAG_SQUARED : SPX_MAG_ABS, scale);
}

int spx_mag_phase(spx_ctx* c, const void* in, void* out, uint64_t total) {
  if (!c) return SPX_EINVAL;
  if (total == 0) return SPX_OK;
  if (!in || !out) return fail(c, SPX_EINVAL, "null device pointer");
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_mag_phase(in, out, total, c->stream));
  return SPX_OK;
}

int spx_magnitude(spx_ctx* c, const void* in, float* out, uint64_t total, float inv_n, int squared) {
  if (!c) return SPX_EINVAL;
  if (total == 0) return SPX_OK;
  if (!in || !out) return fail(c, SPX_EINVAL, "null device pointer");
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_magnitude(in, out, total, inv_n, squared, c->stream));
  return SPX_OK;
}

int spx_peaks_top(spx_ctx* c, const void* spec, void* out, uint32_t batch, uint32_t nfft, uint32_t search_range,
                  float fs, int two_peaks) {
  if (!c) return SPX_EINVAL;
  if (batch == 0) return SPX_OK;
  if (!spec || !out) return fail(c, SPX_EINVAL, "null device pointer");
  if (nfft < 2) return fail(c, SPX_EINVAL, "nfft must be >= 2");
  if (search_range > nfft) search_range = nfft;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_peaks_top(spec, out, batch, nfft, search_range, fs, two_peaks, c->stream));
  return SPX_OK;
}

int spx_fused_peaks_top(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                        uint32_t window, uint32_t search_range, float fs, int two_peaks) {
  int rc = check_fft_shape(c, in, out, batch, n_point, nfft, window);
  if (rc || batch == 0) return rc;
  if (search_range > nfft) search_range = nfft;
  const uint32_t half = search_range / 2;
  FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in;
  a.batch = batch;
  a.n_point = n_point;
  a.nfft = nfft;
  a.dtype = SPX_DTYPE_C64;
  a.fs = fs;
  if ((rc = window_ptr(c, window, n_point, &a.window))) return rc;
  a.peak_mode = two_peaks ? 4u : 3u;
  a.s0 = half;
  a.s1 = nfft - half;
  a.peaks = (spx_max_value*)out;
  if (!spx_fused_is_large(nfft)) {
    // one launch: transform, both argmaxes and the records per workgroup
    int grid = 0;
    if ((rc = prepare(c, a, &grid))) return rc;
    SPX_HIP(c, hipSetDevice(c->device));
    SPX_HIP(c, spx_launch_fused(&a, grid, c->stream));
    return SPX_OK;
  }
  if (!spx_large_is(nfft) && half >= 1) {
    // three launches per chunk of channels, no spectrum stored
    const void* t;
    if ((rc = get_table(c, TW, 256, 0, &t))) return rc;
    a.tw_a = (const float2*)t;
    if ((rc = get_table(c, TW, nfft / 256, 0, &t))) return rc;
    a.tw_b = (const float2*)t;
    if ((rc = get_table(c, TW4, nfft, 0, &t))) return rc;
    a.tw_4 = (const float2*)t;
    const uint64_t per = spx_fs_top_workspace_bytes(nfft, 1);
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(batch, std::max<uint64_t>(64, (2ull << 30) / per));
    if ((rc = ensure_scratch(c, spx_fs_top_workspace_bytes(nfft, chunk)))) return rc;
    SPX_HIP(c, hipSetDevice(c->device));
    SPX_HIP(c, spx_launch_fourstep_top(&a, c->scratch, chunk, c->stream));
    return SPX_OK;
  }
  // 2^19 .. 2^24 (and an empty search range): the spectrum of a chunk of
  // channels into spec_buf, then the search on it
  const uint32_t recs = two_peaks ? 8u : 4u;
  const uint64_t per = (uint64_t)nfft * 8;
  const uint32_t chunk = (uint32_t)std::min<uint64_t>(batch, std::max<uint64_t>(1, (1ull << 30) / per));
  if (c->spec_bytes < per * chunk) {
    SPX_HIP(c, hipSetDevice(c->device));
    if (c->spec_buf) {
      SPX_HIP(c, hipStreamSynchronize(c->stream));
      SPX_HIP(c, hipFree(c->spec_buf));
      c->spec_buf = nullptr;
      c->spec_bytes = 0;
    }
    if (hipMalloc(&c->spec_buf, per * chunk) != hipSuccess) {
      c->spec_buf = nullptr;
      return oom(c, "spectrum buffer allocation failed");
    }
    c->spec_bytes = per * chunk;
  }
  for (uint32_t c0 = 0; c0 < batch; c0 += chunk) {
    const uint32_t nb = std::min(chunk, batch - c0);
    const float2* src = (const float2*)in + (uint64_t)c0 * n_point;
    if ((rc = dense_fft(c, src, c->spec_buf, nb, n_point, nfft, window, 0, 1.0f))) return rc;
    SPX_HIP(c, spx_launch_peaks_top(c->spec_buf, (spx_max_value*)out + (uint64_t)c0 * recs, nb, nfft, search_range, fs,
                                    two_peaks, c->stream));
  }
  return SPX_OK;
}

int spx_all_maxima(spx_ctx* c, const void* spec, void* out, uint32_t* counts, uint32_t batch, uint32_t nfft,
                   float fs, uint32_t s0, uint32_t s1, uint32_t cap) {
  if (!c) return SPX_EINVAL;
  if (batch == 0) return SPX_OK;
  if (!spec || !counts || (!out && cap)) return fail(c, SPX_EINVAL, "null device pointer");
  if (nfft < 3) return fail(c, SPX_EINVAL, "nfft must be >= 3");
  if (s0 < 1 || s1 > nfft - 1 || s0 > s1) return fail(c, SPX_EINVAL, "need 1 <= search_start <= search_end <= nfft-1");
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_all_maxima(spec, out, counts, batch, nfft, fs, s0, s1, cap, c->stream));
  return SPX_OK;
}

}  // extern "C"

namespace spx_host {

// Validates a fused descriptor and fills the kernel arguments (no tables).
int fused_args(spx_ctx* c, const spx_fused_desc* d, FusedArgs& a) {
  std::memset(&a, 0, sizeof(a));
  if (!c) return SPX_EINVAL;
  if (!d) return fail(c, SPX_EINVAL, "null descriptor");
  if (d->batch == 0) return SPX_OK;
  int rc = check_fft_shape(c, d->input, d->peaks_out ? (const void*)d->peaks_out : (const void*)d->counts_out,
                           d->batch, d->n_point, d->nfft, d->window);
  if (rc) return rc;
  if (!d->counts_out || (!d->peaks_out && d->max_peaks)) return fail(c, SPX_EINVAL, "null output pointer");
  if (d->dtype > SPX_DTYPE_C32BF16) return fail(c, SPX_EINVAL, "unknown dtype");
  if (d->peak_mode != SPX_PEAKS_THRESHOLD && d->peak_mode != SPX_PEAKS_ALL_MAXIMA)
    return fail(c, SPX_EINVAL, "unknown peak_mode");
  if (d->refine > SPX_REFINE_LEARNED) return fail(c, SPX_EINVAL, "unknown refine mode");
  if (d->refine == SPX_REFINE_LEARNED) {
    if (!c->refiner) return fail(c, SPX_EINVAL, "learned refinement requested but no refiner is set (spx_set_refiner)");
    if (d->peak_mode != SPX_PEAKS_THRESHOLD || d->nfft < 1024)
      return fail(c, SPX_ENOTSUP, "learned refinement: threshold search and nfft >= 1024 only");
  }
  uint32_t s0 = d->search_start, s1 = d->search_end;
  if (d->peak_mode == SPX_PEAKS_ALL_MAXIMA) {
    // reference range normalisation, all_maxima_pipeline_rocm.cpp:274-276
    if (s0 == 0) s0 = 1;
    if (s1 == 0) s1 = d->nfft / 2;
    if (s1 >= d->nfft) s1 = d->nfft - 1;
  } else if (s0 == 0 && s1 == 0) {
    s1 = d->nfft;
  }
  if (s0 >= s1 || s1 > d->nfft) return fail(c, SPX_EINVAL, "empty or out-of-range search window");
  a.in = d->input;
  a.in_scale = d->in_scale;
  a.batch = d->batch;
  a.n_point = d->n_point;
  a.nfft = d->nfft;
  a.dtype = d->dtype;
  a.fs = d->sample_rate;
  if ((rc = window_ptr(c, d->window, d->n_point, &a.window))) return rc;
  a.peak_mode = d->peak_mode;
  a.nsigma = d->nsigma;
  a.kcap = d->max_peaks;
  a.s0 = s0;
  a.s1 = s1;
  a.refine = d->peak_mode == SPX_PEAKS_ALL_MAXIMA ? 0u : d->refine;
  a.peaks = d->peaks_out;
  a.counts = d->counts_out;
  a.gmax = d->gmax_out;
  a.stats = d->stats_out;
  a.mlp = c->refiner;
  return SPX_OK;
}

// Mixed-precision table block for N = 16384 (layout documented in spx_mixed.hip):
// TB[j1][m2] = 2^-5 W_1024^(m2 j1), TAL[e] = W_16384^e and TAH[e] = W_128^e (e < 128)
// as float2, then the DFT-16 and DFT-32 real-matrix MFMA fragments, hi and lo
// parts, in the compute type (fp16 or bf16).  FP64, each value rounded once.
std::vector<float> make_mixed_tables(uint32_t compute) {
  const size_t bytes = spx_mixed_table_bytes();
  std::vector<float> t((bytes + 3) / 4, 0.0f);
  unsigned char* p = reinterpret_cast<unsigned char*>(t.data());
  auto tw = [](uint64_t e, uint64_t n, float* o) {
    const double a = 2.0 * M_PI * (double)(e % n) / (double)n;
    o[0] = (float)std::cos(a);
    o[1] = (float)-std::sin(a);
  };
  float* TB = reinterpret_cast<float*>(p);
  for (uint32_t j1 = 0; j1 < 32; ++j1)
    for (uint32_t m2 = 0; m2 < 32; ++m2) {
      float* o = TB + 2 * (j1 * 32 + m2);
      tw((uint64_t)m2 * j1, 1024, o);
      o[0] *= 0.03125f;  // the stage-2 scale 2^-5 (exact)
      o[1] *= 0.03125f;
    }
  float* TAL = TB + 2 * 1024;
  float* TAH = TAL + 2 * 128;
  for (uint32_t e = 0; e < 128; ++e) {
    tw(e, 16384, TAL + 2 * e);
    tw(e, 128, TAH + 2 * e);
  }
  // real 2R x 2R DFT matrix [[C, S], [-S, C]]
  auto M = [](int R, int r, int k) -> double {
    const int rr = r % R, kk = k % R;
    const double a = 2.0 * M_PI * (double)((rr * kk) % R) / (double)R;
    if (r < R) return k < R ? std::cos(a) : std::sin(a);
    return k < R ? -std::sin(a) : std::cos(a);
  };
  auto put = [&](uint16_t* hi, uint16_t* lo, double v) {
    if (compute == 0) {
      const uint16_t h = f32_to_f16_bits((float)v);
      *hi = h;
      *lo = f32_to_f16_bits((float)(v - (double)f16_bits_to_f32(h)));
    } else {
      const uint16_t h = f32_to_bf16_bits((float)v);
      *hi = h;
      *lo = f32_to_bf16_bits((float)(v - (double)bf16_bits_to_f32(h)));
    }
  };
  // DFT-16 of column block t (t < 8) with the twiddle factor W_128^(t k1) of
  // its output rows folded in: complex entry W_128^(k1 (8 n1 + t)).  The K = 32
  // operand rows are ordered (n1, re/im): K row q holds component q & 1 of
  // input n1 = q / 2.
  auto M16t = [](int t, int r, int q) -> double {
    const int k1 = r % 16, n1 = q / 2;
    const double a = 2.0 * M_PI * (double)((k1 * (8 * n1 + t)) % 128) / 128.0;
    if (r < 16) return (q & 1) ? std::sin(a) : std::cos(a);
    return (q & 1) ? std::cos(a) : -std::sin(a);
  };
  uint16_t* W16 = reinterpret_cast<uint16_t*>(p + 1024 * 8 + 256 * 8);
  for (int t = 0; t < 8; ++t)
    for (int rb = 0; rb < 2; ++rb)
      for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
          const size_t i = (((size_t)t * 2 + rb) * 64 + l) * 8 + j;
          put(W16 + i, W16 + 8 * 2 * 64 * 8 + i, M16t(t, 16 * rb + l % 16, 8 * (l / 16) + j));
        }
  // DFT-32: K = 64 rows ordered (m, re/im) as well; k-step ks row q holds
  // component q & 1 of input m = 16 ks + q / 2
  uint16_t* W32 = W16 + 2 * 8 * 2 * 64 * 8;
  for (int rb = 0; rb < 4; ++rb)
    for (int ks = 0; ks < 2; ++ks)
      for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
          const size_t i = (((size_t)(rb * 2 + ks)) * 64 + l) * 8 + j;
          const int q = 8 * (l / 16) + j;
          put(W32 + i, W32 + 8 * 64 * 8 + i, M(32, 16 * rb + l % 16, (q & 1) * 32 + 16 * ks + q / 2));
        }
  return t;
}

}  // namespace spx_host

extern "C" {

int spx_fused_peaks(spx_ctx* c, const spx_fused_desc* d) {
  FusedArgs a;
  int rc = fused_args(c, d, a);
  if (rc || d->batch == 0) return rc;
  return run_fused(c, a);
}

// Window coefficients, tables, persistent grid, launch.
static int run_mixed(spx_ctx* c, FusedArgs& a, uint32_t window, uint32_t compute) {
  a.window = nullptr;
  window_coefs(window, a);
  const void* t;
  int rc = get_table(c, MIX, compute, 0, &t);
  if (rc) return rc;
  a.tw_a = (const float2*)t;
  int grid = c->cu_count > 0 ? c->cu_count : 256;
  if ((uint32_t)grid > a.batch) grid = (int)a.batch;
  if ((rc = learned_features(c, a))) return rc;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_mixed(&a, (int)compute, grid, c->stream));
  return learned_refine(c, a);
}

static int mixed_check(spx_ctx* c, uint32_t nfft, uint32_t n_point, uint32_t dtype, uint32_t compute) {
  if (nfft != 16384 || n_point != 16384)
    return fail(c, SPX_ENOTSUP, "mixed-precision path: nfft = n_point = 16384 only");
  if (dtype != SPX_DTYPE_C32F16 && dtype != SPX_DTYPE_C32BF16)
    return fail(c, SPX_EINVAL, "mixed-precision path: input dtype must be fp16 or bf16");
  if (compute > SPX_COMPUTE_BF16) return fail(c, SPX_EINVAL, "unknown compute type");
  return SPX_OK;
}

int spx_fused_peaks_mixed(spx_ctx* c, const spx_fused_desc* d, uint32_t compute) {
  FusedArgs a;
  int rc = fused_args(c, d, a);
  if (rc || d->batch == 0) return rc;
  if ((rc = mixed_check(c, d->nfft, d->n_point, d->dtype, compute))) return rc;
  return run_mixed(c, a, d->window, compute);
}

int spx_fft_mixed(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t nfft, uint32_t window,
                  uint32_t dtype, uint32_t compute) {
  if (!c) return SPX_EINVAL;
  if (batch == 0) return SPX_OK;
  if (!in || !out) return fail(c, SPX_EINVAL, "null pointer");
  int rc = mixed_check(c, nfft, nfft, dtype, compute);
  if (rc) return rc;
  if (window > SPX_WINDOW_BLACKMAN) return fail(c, SPX_EINVAL, "unknown window");
  FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in;
  a.batch = batch;
  a.n_point = nfft;
  a.nfft = nfft;
  a.dtype = dtype;
  a.fs = 1.0f;
  a.spec_out = (float2*)out;
  a.peak_mode = 0;
  return run_mixed(c, a, window, compute);
}


}  // extern "C"

// ---------------------------------------------------------------------------
// learned sub-bin refiner: data set, training, export, inference
// ---------------------------------------------------------------------------
extern "C" {

int spx_mlp_param_count(void) { return spx_mlp_nparam(); }
int spx_mlp_deployed_size(void) { return spx_mlp_deployed_bytes(); }

int spx_set_refiner(spx_ctx* c, const void* deployed) {
  if (!c) return SPX_EINVAL;
  c->refiner = deployed;
  return SPX_OK;
}

// He-uniform initial weights (bound sqrt(6 / fan_in)), zero biases, from a
// SplitMix64 stream; written to params_dev (FP32 master layout).
int spx_mlp_init(spx_ctx* c, float* params_dev, uint64_t seed) {
  if (!c || !params_dev) return SPX_EINVAL;
  const int P = spx_mlp_nparam();
  std::vector<float> p((size_t)P, 0.0f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto uni = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  const int IN = 10, H = 64;
  int o = 0;
  for (int i = 0; i < H * IN; ++i) p[o++] = (float)(uni() * std::sqrt(6.0 / IN));
  o += H;  // b1
  for (int i = 0; i < H * H; ++i) p[o++] = (float)(uni() * std::sqrt(6.0 / H));
  o += H;  // b2
  for (int i = 0; i < H; ++i) p[o++] = (float)(uni() * std::sqrt(6.0 / H));
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, hipMemcpyAsync(params_dev, p.data(), p.size() * 4, hipMemcpyHostToDevice, c->stream));
  SPX_HIP(c, hipStreamSynchronize(c->stream));
  return SPX_OK;
}

static int mlp_groups(spx_ctx* c, uint32_t n) {
  int g = c->cu_count > 0 ? c->cu_count : 256;
  const int steps = (int)((n + 63) / 64);
  return std::max(1, std::min(g, steps));
}

static int mlp_partial(spx_ctx* c, int groups) {
  const uint64_t need = (uint64_t)groups * (uint64_t)(spx_mlp_nparam() + 1) * 4;
  if (c->mlp_partial_bytes >= need) return SPX_OK;
  SPX_HIP(c, hipSetDevice(c->device));
  if (c->mlp_partial) {
    SPX_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->mlp_partial);
    c->mlp_partial = nullptr;
    c->mlp_partial_bytes = 0;
  }
  if (hipMalloc(&c->mlp_partial, need) != hipSuccess) {
    c->mlp_partial = nullptr;
    return oom(c, "mlp partial buffer");
  }
  c->mlp_partial_bytes = need;
  return SPX_OK;
}

// grad_out: NPARAM + 1 floats (gradient of the mean squared error, then the loss)
int spx_mlp_grad(spx_ctx* c, const float* params, const float* X, const float* Y, uint32_t n, float* grad_out) {
  if (!c) return SPX_EINVAL;
  if (!params || !X || !Y || !grad_out || n == 0) return fail(c, SPX_EINVAL, "spx_mlp_grad: bad arguments");
  const int g = mlp_groups(c, n);
  int rc = mlp_partial(c, g);
  if (rc) return rc;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_mlp_grad(params, X, Y, n, (float*)c->mlp_partial, g, grad_out, c->stream));
  return SPX_OK;
}

// Adam on mini-batches: step s (1-based, continuing from step0) uses samples
// [off, off + batch) with off from a SplitMix64 stream (seed, s); grad_buf
// holds NPARAM + 1 floats; loss_out (host, optional) receives each step's loss.
int spx_mlp_train(spx_ctx* c, float* params, float* m, float* v, float* grad_buf, const float* X, const float* Y,
                  uint32_t n, uint32_t batch, uint32_t steps, uint32_t step0, float lr, uint64_t seed,
                  float* loss_out) {
  if (!c) return SPX_EINVAL;
  if (!params || !m || !v || !grad_buf || !X || !Y || n == 0 || batch == 0 || batch > n)
    return fail(c, SPX_EINVAL, "spx_mlp_train: bad arguments");
  const int g = mlp_groups(c, batch);
  int rc = mlp_partial(c, g);
  if (rc) return rc;
  SPX_HIP(c, hipSetDevice(c->device));
  const int P = spx_mlp_nparam();
  for (uint32_t s = 1; s <= steps; ++s) {
    uint64_t z = (seed + (uint64_t)(step0 + s) * 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const uint32_t off = (uint32_t)(z % (uint64_t)(n - batch + 1));
    SPX_HIP(c, spx_launch_mlp_grad(params, X + (uint64_t)off * 10, Y + off, batch, (float*)c->mlp_partial, g,
                                   grad_buf, c->stream));
    SPX_HIP(c, spx_launch_mlp_adam(params, grad_buf, m, v, lr, 0.9f, 0.999f, 1e-8f, (int)(step0 + s), c->stream));
    if (loss_out) {
      SPX_HIP(c, hipMemcpyAsync(loss_out + (s - 1), grad_buf + P, 4, hipMemcpyDeviceToHost, c->stream));
    }
  }
  SPX_HIP(c, hipStreamSynchronize(c->stream));
  return SPX_OK;
}

int spx_mlp_export(spx_ctx* c, const float* params, void* deployed_out) {
  if (!c) return SPX_EINVAL;
  if (!params || !deployed_out) return fail(c, SPX_EINVAL, "spx_mlp_export: null pointer");
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_mlp_export(params, deployed_out, c->stream));
  return SPX_OK;
}

int spx_mlp_predict(spx_ctx* c, const void* deployed, const float* X, uint32_t n, float* out) {
  if (!c) return SPX_EINVAL;
  if (!deployed || !X || !out) return fail(c, SPX_EINVAL, "spx_mlp_predict: null pointer");
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_mlp_predict(deployed, X, n, out, c->stream));
  return SPX_OK;
}

int spx_mlp_dataset(spx_ctx* c, const void* spec, const float* mag16, const spx_max_value* peaks,
                    const uint32_t* counts, uint32_t batch, uint32_t nfft, uint32_t max_peaks, float* feat,
                    float* label, float* parab, uint32_t* valid) {
  if (!c) return SPX_EINVAL;
  if (!spec || !mag16 || !peaks || !counts || !feat || !label || !parab || !valid)
    return fail(c, SPX_EINVAL, "spx_mlp_dataset: null pointer");
  if (nfft < 8 || (nfft & (nfft - 1))) return fail(c, SPX_EINVAL, "spx_mlp_dataset: nfft must be a power of two");
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, spx_launch_mlp_dataset(spec, mag16, peaks, counts, batch, nfft, max_peaks, feat, label, parab, valid,
                                    c->stream));
  return SPX_OK;
}

}  // extern "C"
