// Host-resident inputs and outputs: the batch is streamed through the GPU in
// chunks of channels so that PCIe transfers in both directions overlap each
// other and the kernels.
//
// The reference's host entry points (FFTProcessorROCm::ProcessComplex /
// ProcessMagPhase on a std::vector, fft_processor_rocm.cpp:306-570;
// SpectrumProcessorROCm::ProcessFromCPU / FindAllMaximaFromCPU,
// spectrum_processor_rocm.cpp:263-307, 444-543) upload the whole batch from
// pageable memory, compute, then download the whole result, one step after
// the other; its own profile of that path is ~96 % copies.  Here, per chunk
// i (slot s = i % 3 of three):
//   host thread  copy chunk i of the caller's input into pinned staging
//                (skipped when the caller's buffer is already pinned)
//   stream up    pinned -> device input slot s
//   ctx stream   the device entry point on slot s (spx_fft_c2c, ...,
//                spx_fused_peaks), writing device output slot s
//   stream down  device output slot s -> pinned (or the caller's pinned buffer)
//   host thread  copy chunk i-2's results from pinned staging to the caller
// Events order slot reuse: the upload into slot s waits for the kernels of
// chunk i-3, the kernels wait for the download of chunk i-3.  With two slots
// the host had to wait for chunk i-1's download before it could stage chunk
// i+1, so each chunk took upload + download in turn; the third slot lets the
// host copy out a chunk whose download finished a chunk earlier.  Every call
// returns with its results in host memory (the reference's synchronous
// semantics).
#include <algorithm>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "spx_host.h"
#include "spectrum_mi355x.h"

constexpr int kSlots = 3;

struct HostStage {
  hipStream_t up = nullptr, down = nullptr;
  hipEvent_t ev_up[kSlots] = {}, ev_comp[kSlots] = {}, ev_down[kSlots] = {};
  void* pin_in[kSlots] = {};
  void* pin_out[kSlots] = {};
  void* dev_in[kSlots] = {};
  void* dev_out[kSlots] = {};
  uint64_t in_cap = 0, out_cap = 0;  // bytes per slot
};

namespace spx_host {

void host_stage_free(spx_ctx* c) {
  HostStage* h = c->hstage;
  if (!h) return;
  (void)hipSetDevice(c->device);
  if (h->up) (void)hipStreamSynchronize(h->up);
  if (h->down) (void)hipStreamSynchronize(h->down);
  for (int s = 0; s < kSlots; ++s) {
    if (h->pin_in[s]) (void)hipHostFree(h->pin_in[s]);
    if (h->pin_out[s]) (void)hipHostFree(h->pin_out[s]);
    if (h->dev_in[s]) (void)hipFree(h->dev_in[s]);
    if (h->dev_out[s]) (void)hipFree(h->dev_out[s]);
    for (hipEvent_t e : {h->ev_up[s], h->ev_comp[s], h->ev_down[s]})
      if (e) (void)hipEventDestroy(e);
  }
  if (h->up) (void)hipStreamDestroy(h->up);
  if (h->down) (void)hipStreamDestroy(h->down);
  delete h;
  c->hstage = nullptr;
}

}  // namespace spx_host

namespace {

using namespace spx_host;

// One array of the call: per_chan bytes for each channel, contiguous over
// the batch, in host memory.
struct HostArray {
  const void* in;   // input array (or null)
  void* out;        // output array (or null)
  uint64_t per_chan;
};

// Pinned (page-locked, device-visible) host memory is copied by the DMA
// engines directly; anything else goes through the staging buffers.
bool is_pinned(const void* p) {
  hipPointerAttribute_t at;
  std::memset(&at, 0, sizeof(at));
  const hipError_t e = hipPointerGetAttributes(&at, p);
  (void)hipGetLastError();  // a pageable pointer reports an error: do not leave it behind
  return e == hipSuccess && at.type == hipMemoryTypeHost;
}

// Host memory copy on several threads (one core does not reach the host's
// memory bandwidth, which the PCIe link needs in both directions).
void par_copy(void* dst, const void* src, uint64_t n) {
  constexpr uint64_t kMinPart = 4ull << 20;
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int t = (int)std::min<uint64_t>((uint64_t)std::min(hw, 8), std::max<uint64_t>(1, n / kMinPart));
  if (t <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const uint64_t part = (n / t + 4095) & ~4095ull;
  std::vector<std::thread> th;
  for (int i = 1; i < t; ++i) {
    const uint64_t o = (uint64_t)i * part;
    if (o >= n) break;
    const uint64_t m = std::min(part, n - o);
    th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, m); });
  }
  std::memcpy(dst, src, std::min(part, n));
  for (auto& x : th) x.join();
}

int grow(spx_ctx* c, void** p, uint64_t want, bool pinned) {
  if (pinned) {
    if (*p) SPX_HIP(c, hipHostFree(*p));
    *p = nullptr;
    SPX_HIP(c, hipHostMalloc(p, want, hipHostMallocDefault));
  } else {
    if (*p) SPX_HIP(c, hipFree(*p));
    *p = nullptr;
    SPX_HIP(c, hipMalloc(p, want));
  }
  return SPX_OK;
}

int stage_for(spx_ctx* c, uint64_t in_bytes, uint64_t out_bytes, HostStage** out) {
  SPX_HIP(c, hipSetDevice(c->device));
  if (!c->hstage) {
    HostStage* h = new HostStage();
    c->hstage = h;
    SPX_HIP(c, hipStreamCreateWithFlags(&h->up, hipStreamNonBlocking));
    SPX_HIP(c, hipStreamCreateWithFlags(&h->down, hipStreamNonBlocking));
    for (int s = 0; s < kSlots; ++s) {
      SPX_HIP(c, hipEventCreateWithFlags(&h->ev_up[s], hipEventDisableTiming));
      SPX_HIP(c, hipEventCreateWithFlags(&h->ev_comp[s], hipEventDisableTiming));
      SPX_HIP(c, hipEventCreateWithFlags(&h->ev_down[s], hipEventDisableTiming));
    }
  }
  HostStage* h = c->hstage;
  int rc;
  // the capacity is zero while the slots are being replaced, so that a
  // failed allocation leaves no slot that a later call would take as valid
  if (in_bytes > h->in_cap) {
    h->in_cap = 0;
    for (int s = 0; s < kSlots; ++s) {
      if ((rc = grow(c, &h->pin_in[s], in_bytes, true))) return rc;
      if ((rc = grow(c, &h->dev_in[s], in_bytes, false))) return rc;
    }
    h->in_cap = in_bytes;
  }
  if (out_bytes > h->out_cap) {
    h->out_cap = 0;
    for (int s = 0; s < kSlots; ++s) {
      if ((rc = grow(c, &h->pin_out[s], out_bytes, true))) return rc;
      if ((rc = grow(c, &h->dev_out[s], out_bytes, false))) return rc;
    }
    h->out_cap = out_bytes;
  }
  *out = h;
  return SPX_OK;
}

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Runs `compute` over the batch chunk by chunk.  compute(din, dout, n)
// receives, for each array, the device address of the chunk's first channel
// (din[i] for inputs, dout[j] for outputs, in the order of `arrays`) and
// the chunk's channel count; it enqueues work on the context's stream.
using ChunkFn = std::function<int(const std::vector<const void*>&, const std::vector<void*>&, uint32_t)>;

int stream_batch(spx_ctx* c, uint32_t batch, const std::vector<HostArray>& arrays, const ChunkFn& compute) {
  if (batch == 0) return SPX_OK;
  uint64_t in_per = 0, out_per = 0;
  for (const auto& a : arrays) {
    if (a.in) in_per += a.per_chan;
    if (a.out) out_per += a.per_chan;
  }
  const uint64_t target = std::max<uint64_t>(c->host_chunk_bytes, 1);
  const uint64_t fit = target / std::max<uint64_t>(std::max(in_per, out_per), 1);
  const uint32_t chunk = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(fit, 1), batch);
  // slot layout: each array's chunk at a 256-byte aligned offset
  std::vector<uint64_t> off(arrays.size());
  uint64_t in_bytes = 0, out_bytes = 0;
  for (size_t i = 0; i < arrays.size(); ++i) {
    uint64_t& tot = arrays[i].in ? in_bytes : out_bytes;
    off[i] = tot;
    tot += align256(arrays[i].per_chan * chunk);
  }
  HostStage* h;
  int rc = stage_for(c, std::max<uint64_t>(in_bytes, 256), std::max<uint64_t>(out_bytes, 256), &h);
  if (rc) return rc;
  std::vector<bool> pinned(arrays.size());
  for (size_t i = 0; i < arrays.size(); ++i) pinned[i] = is_pinned(arrays[i].in ? arrays[i].in : arrays[i].out);
  // Uploads straight from one large pinned caller buffer while downloads go
  // straight into another ran one after the other on MI355X (1024 x 64 K
  // complex spectrum: 20-22 ms); staging the input through the two small
  // slots let the directions overlap (13-14 ms).  So with a dense output
  // that is pinned, pinned inputs are staged as well.
  bool dense_pinned_out = false;
  for (size_t i = 0; i < arrays.size(); ++i)
    dense_pinned_out |= arrays[i].out && pinned[i] && 2 * arrays[i].per_chan >= in_per;
  if (dense_pinned_out)
    for (size_t i = 0; i < arrays.size(); ++i)
      if (arrays[i].in) pinned[i] = false;

  const uint32_t nchunks = (batch + chunk - 1) / chunk;
  auto first = [&](uint32_t i) { return (uint64_t)i * chunk; };
  auto count = [&](uint32_t i) { return (uint32_t)std::min<uint64_t>(chunk, batch - first(i)); };
  // results of chunk i from pinned staging to the caller's arrays
  auto drain = [&](uint32_t i) -> int {
    const int s = (int)(i % kSlots);
    bool staged = false;
    for (size_t j = 0; j < arrays.size(); ++j) staged |= arrays[j].out && !pinned[j];
    if (!staged) return SPX_OK;
    SPX_HIP(c, hipEventSynchronize(h->ev_down[s]));
    for (size_t j = 0; j < arrays.size(); ++j) {
      if (!arrays[j].out || pinned[j]) continue;
      par_copy((char*)arrays[j].out + first(i) * arrays[j].per_chan, (char*)h->pin_out[s] + off[j],
               count(i) * arrays[j].per_chan);
    }
    return SPX_OK;
  };

  // every exit, failed or not, leaves no copy or kernel of this call in
  // flight (the staging buffers and the caller's pinned arrays are free)
  auto run = [&]() -> int {
    for (uint32_t i = 0; i < nchunks; ++i) {
      const int s = (int)(i % kSlots);
      const uint32_t n = count(i);
      // 1. input to pinned staging, once the upload of chunk i-3 has read it
      if (i >= kSlots) SPX_HIP(c, hipEventSynchronize(h->ev_up[s]));
      for (size_t j = 0; j < arrays.size(); ++j) {
        if (!arrays[j].in || pinned[j]) continue;
        par_copy((char*)h->pin_in[s] + off[j], (const char*)arrays[j].in + first(i) * arrays[j].per_chan,
                 n * arrays[j].per_chan);
      }
      // 2. upload, once the kernels of chunk i-3 are done with the slot
      if (i >= kSlots) SPX_HIP(c, hipStreamWaitEvent(h->up, h->ev_comp[s], 0));
      std::vector<const void*> din;
      std::vector<void*> dout;
      for (size_t j = 0; j < arrays.size(); ++j) {
        if (arrays[j].in) {
          const void* src = pinned[j] ? (const char*)arrays[j].in + first(i) * arrays[j].per_chan
                                      : (const char*)h->pin_in[s] + off[j];
          void* dst = (char*)h->dev_in[s] + off[j];
          SPX_HIP(c, hipMemcpyAsync(dst, src, n * arrays[j].per_chan, hipMemcpyHostToDevice, h->up));
          din.push_back(dst);
        } else {
          dout.push_back((char*)h->dev_out[s] + off[j]);
        }
      }
      SPX_HIP(c, hipEventRecord(h->ev_up[s], h->up));
      // 3. kernels, once the upload is in and the download of chunk i-3 has
      //    read the output slot
      SPX_HIP(c, hipStreamWaitEvent(c->stream, h->ev_up[s], 0));
      if (i >= kSlots) SPX_HIP(c, hipStreamWaitEvent(c->stream, h->ev_down[s], 0));
      if ((rc = compute(din, dout, n))) return rc;
      SPX_HIP(c, hipEventRecord(h->ev_comp[s], c->stream));
      // 4. download
      SPX_HIP(c, hipStreamWaitEvent(h->down, h->ev_comp[s], 0));
      for (size_t j = 0; j < arrays.size(); ++j) {
        if (!arrays[j].out) continue;
        void* dst = pinned[j] ? (char*)arrays[j].out + first(i) * arrays[j].per_chan : (char*)h->pin_out[s] + off[j];
        SPX_HIP(c, hipMemcpyAsync(dst, (char*)h->dev_out[s] + off[j], n * arrays[j].per_chan, hipMemcpyDeviceToHost,
                                  h->down));
      }
      SPX_HIP(c, hipEventRecord(h->ev_down[s], h->down));
      // 5. the results of the chunk before the previous one to the caller
      //    (its slot is reused by chunk i+1's download)
      if (i >= 2 && (rc = drain(i - 2))) return rc;
    }
    for (uint32_t i = nchunks >= 2 ? nchunks - 2 : 0; i < nchunks; ++i)
      if ((rc = drain(i))) return rc;
    return SPX_OK;
  };
  rc = run();
  const hipError_t e1 = hipStreamSynchronize(h->up);
  const hipError_t e2 = hipStreamSynchronize(h->down);
  const hipError_t e3 = hipStreamSynchronize(c->stream);
  if (rc) return rc;
  SPX_HIP(c, e1);
  SPX_HIP(c, e2);
  SPX_HIP(c, e3);
  return SPX_OK;
}

uint64_t dtype_bytes(uint32_t dtype) { return dtype == SPX_DTYPE_C64 ? 8 : 4; }

int dense_host(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
               uint32_t window, uint64_t out_per_bin,
               const std::function<int(const void*, void*, uint32_t)>& launch) {
  if (!c) return SPX_EINVAL;
  if (batch == 0) return SPX_OK;
  if (!in || !out) return fail(c, SPX_EINVAL, "null host pointer");
  if (nfft == 0 || n_point == 0 || n_point > nfft) return fail(c, SPX_EINVAL, "n_point must be in [1, nfft]");
  (void)window;
  const std::vector<HostArray> arrays = {{in, nullptr, (uint64_t)n_point * 8}, {nullptr, out, nfft * out_per_bin}};
  return stream_batch(c, batch, arrays,
                      [&](const std::vector<const void*>& din, const std::vector<void*>& dout, uint32_t n) {
                        return launch(din[0], dout[0], n);
                      });
}

}  // namespace

extern "C" {

int spx_set_host_chunk(spx_ctx* c, uint64_t bytes) {
  if (!c || bytes == 0) return SPX_EINVAL;
  c->host_chunk_bytes = bytes;
  return SPX_OK;
}

int spx_host_alloc(spx_ctx* c, uint64_t bytes, void** out) {
  if (!c || !out) return SPX_EINVAL;
  *out = nullptr;
  if (bytes == 0) return SPX_OK;
  SPX_HIP(c, hipSetDevice(c->device));
  const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    *out = nullptr;
    return oom(c, "spx_host_alloc: out of pinned host memory");
  }
  SPX_HIP(c, e);
  return SPX_OK;
}

int spx_host_free(spx_ctx* c, void* p) {
  if (!c) return SPX_EINVAL;
  if (p) SPX_HIP(c, hipHostFree(p));
  return SPX_OK;
}

int spx_fft_c2c_host(spx_ctx* c, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                     uint32_t window) {
  return dense_host(c, in, out, batch, n_point, nfft, window, 8, [&](const void* di, void* dout, uint32_t n) {
    return spx_fft_c2c(c, di, dout, n, n_point, nfft, window);
  });
}

int spx_fft_mag_phase_host(spx_ctx* c, const void* in, void* out_float2, uint32_t batch, uint32_t n_point,
                           uint32_t nfft, uint32_t window) {
  return dense_host(c, in, out_float2, batch, n_point, nfft, window, 8, [&](const void* di, void* dout, uint32_t n) {
    return spx_fft_mag_phase(c, di, dout, n, n_point, nfft, window);
  });
}

int spx_fft_magnitude_host(spx_ctx* c, const void* in, float* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                           uint32_t window, int squared, float scale) {
  return dense_host(c, in, out, batch, n_point, nfft, window, 4, [&](const void* di, void* dout, uint32_t n) {
    return spx_fft_magnitude(c, di, (float*)dout, n, n_point, nfft, window, squared, scale);
  });
}

int spx_fused_peaks_host(spx_ctx* c, const spx_fused_desc* d) {
  if (!c) return SPX_EINVAL;
  if (!d) return fail(c, SPX_EINVAL, "null descriptor");
  if (d->batch == 0) return SPX_OK;
  if (!d->input || !d->counts_out || (!d->peaks_out && d->max_peaks))
    return fail(c, SPX_EINVAL, "null host pointer");
  if (d->dtype > SPX_DTYPE_C32BF16) return fail(c, SPX_EINVAL, "unknown dtype");
  std::vector<HostArray> arrays = {{d->input, nullptr, (uint64_t)d->n_point * dtype_bytes(d->dtype)}};
  const int i_scale = d->in_scale ? (int)arrays.size() : -1;
  if (d->in_scale) arrays.push_back({d->in_scale, nullptr, 4});
  const int o_peaks = d->max_peaks ? (int)arrays.size() : -1;
  if (d->max_peaks) arrays.push_back({nullptr, d->peaks_out, (uint64_t)d->max_peaks * sizeof(spx_max_value)});
  arrays.push_back({nullptr, d->counts_out, 8});
  if (d->gmax_out) arrays.push_back({nullptr, d->gmax_out, sizeof(spx_max_value)});
  if (d->stats_out) arrays.push_back({nullptr, d->stats_out, 16});
  return stream_batch(c, d->batch, arrays,
                      [&](const std::vector<const void*>& din, const std::vector<void*>& dout, uint32_t n) {
                        spx_fused_desc x = *d;
                        x.batch = n;
                        x.input = din[0];
                        x.in_scale = i_scale >= 0 ? (const float*)din[1] : nullptr;
                        size_t j = 0;
                        x.peaks_out = o_peaks >= 0 ? (spx_max_value*)dout[j++] : nullptr;
                        x.counts_out = (uint32_t*)dout[j++];
                        x.gmax_out = d->gmax_out ? (spx_max_value*)dout[j++] : nullptr;
                        x.stats_out = d->stats_out ? (float*)dout[j++] : nullptr;
                        return spx_fused_peaks(c, &x);
                      });
}

}  // extern "C"
