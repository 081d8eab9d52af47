// C++ facades with the reference's class names and semantics, implemented on
// the C ABI of libspectrum_mi355x.so.  Host work here is argument checking,
// staging copies and result repacking; every transform and search is a gfx950
// kernel launched through spx_*.  Errors map to the reference's exception
// types: std::invalid_argument for bad sizes/pointers, std::runtime_error for
// HIP failures (src/fft_func/src/fft_processor_rocm.cpp:311-314,359-361).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <spectrum/complex_to_mag_phase_rocm.hpp>
#include <spectrum/factory/spectrum_processor_factory.hpp>
#include <spectrum/fft_processor_rocm.hpp>
#include <spectrum/pipelines/all_maxima_pipeline_rocm.hpp>
#include <spectrum/processors/spectrum_processor_rocm.hpp>
#include <spectrum_mi355x.h>
#include <stdexcept>
#include <string>

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// A stage time as the reference's ROCmProfilingData (MakeROCmDataFromEvents,
// include/spectrum/utils/rocm_profiling_helpers.hpp:60-63): start 0, end and
// complete = the duration; kind 0 upload, 1 compute, 2 download.
drv_gpu_lib::ROCmProfilingData prof_data(double ms, uint32_t kind, const char* op) {
  drv_gpu_lib::ROCmProfilingData d{};
  const uint64_t ns = (uint64_t)(ms * 1e6);
  d.end_ns = ns;
  d.complete_ns = ns;
  d.kind = kind;
  d.op_string = op;
  return d;
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void spx_check(spx_ctx* c, int rc, const char* what) {
  if (rc == SPX_OK) return;
  std::string msg = std::string(what) + ": " + spx_last_error(c);
  if (rc == SPX_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

spx_ctx* make_ctx(drv_gpu_lib::IBackend* backend, const char* who) {
  if (!backend) throw std::invalid_argument(std::string(who) + ": backend cannot be null");
  spx_ctx* c = nullptr;
  int rc = spx_create(backend->GetDeviceIndex(), &c);
  if (rc != SPX_OK) throw std::runtime_error(std::string(who) + ": no usable gfx950 device (spx_create)");
  spx_set_stream(c, backend->GetNativeQueue());
  return c;
}

hipStream_t stream_of(spx_ctx* c) { return (hipStream_t)spx_get_stream(c); }

void reserve(void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes && *buf) return;
  if (*buf) hip_check(hipFree(*buf), "hipFree");
  *buf = nullptr;
  *cap = 0;
  hip_check(hipMalloc(buf, bytes ? bytes : 1), "hipMalloc");
  *cap = bytes;
}

uint32_t next_pow2(uint32_t n) {
  if (n == 0) return 1;
  --n;
  n |= n >> 1; n |= n >> 2; n |= n >> 4; n |= n >> 8; n |= n >> 16;
  return n + 1;
}

bool is_pow2(uint32_t n) { return n && !(n & (n - 1)); }

}  // namespace

// ===========================================================================
// drv_gpu_lib compat
// ===========================================================================
namespace drv_gpu_lib {

ROCmBackend::ROCmBackend() = default;
ROCmBackend::~ROCmBackend() {
  if (stream_) {
    (void)hipSetDevice(device_);
    (void)hipStreamDestroy((hipStream_t)stream_);
  }
}

void ROCmBackend::Initialize(int device_index) {
  int n = 0;
  hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (device_index < 0 || device_index >= n) throw std::invalid_argument("ROCmBackend: bad device index");
  device_ = device_index;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hipDeviceProp_t p;
  hip_check(hipGetDeviceProperties(&p, device_), "hipGetDeviceProperties");
  info_.name = p.name;
  info_.global_mem_bytes = p.totalGlobalMem;
  info_.compute_units = p.multiProcessorCount;
  info_.warp_size = p.warpSize;
  info_.device_index = device_;
  core_.warp_size_ = p.warpSize;
  core_.compute_units_ = p.multiProcessorCount;
  if (stream_) {
    (void)hipStreamDestroy((hipStream_t)stream_);
    stream_ = nullptr;
  }
  hipStream_t s;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  stream_ = s;
}

void* ROCmBackend::Allocate(size_t bytes) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  void* p = nullptr;
  hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
  return p;
}

void ROCmBackend::Free(void* ptr) {
  if (!ptr) return;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipFree(ptr), "hipFree");
}

void ROCmBackend::MemcpyHostToDevice(void* dst, const void* src, size_t bytes) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream_), "hipMemcpyAsync H2D");
  hip_check(hipStreamSynchronize((hipStream_t)stream_), "hipStreamSynchronize");
}

void ROCmBackend::MemcpyDeviceToHost(void* dst, const void* src, size_t bytes) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream_), "hipMemcpyAsync D2H");
  hip_check(hipStreamSynchronize((hipStream_t)stream_), "hipStreamSynchronize");
}

void ROCmBackend::MemcpyDeviceToDevice(void* dst, const void* src, size_t bytes) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream_), "hipMemcpyAsync D2D");
  hip_check(hipStreamSynchronize((hipStream_t)stream_), "hipStreamSynchronize");
}

void ROCmBackend::Synchronize() {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipStreamSynchronize((hipStream_t)stream_), "hipStreamSynchronize");
}

int ROCmCore::GetAvailableDeviceCount() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // namespace drv_gpu_lib

// ===========================================================================
// FFTProcessorROCm
// ===========================================================================
namespace fft_processor {

FFTProcessorROCm::FFTProcessorROCm(drv_gpu_lib::IBackend* backend) : backend_(backend) {
  ctx_ = make_ctx(backend, "FFTProcessorROCm");
}

FFTProcessorROCm::~FFTProcessorROCm() { Release(); }

void FFTProcessorROCm::Release() {
  if (in_) (void)hipFree(in_);
  if (out_) (void)hipFree(out_);
  in_ = out_ = nullptr;
  in_cap_ = out_cap_ = 0;
  if (ctx_) spx_destroy((spx_ctx*)ctx_);
  ctx_ = nullptr;
}

FFTProcessorROCm::FFTProcessorROCm(FFTProcessorROCm&& o) noexcept { *this = std::move(o); }

FFTProcessorROCm& FFTProcessorROCm::operator=(FFTProcessorROCm&& o) noexcept {
  if (this != &o) {
    Release();
    backend_ = o.backend_; ctx_ = o.ctx_; in_ = o.in_; in_cap_ = o.in_cap_; out_ = o.out_; out_cap_ = o.out_cap_;
    nFFT_ = o.nFFT_; n_point_ = o.n_point_; prof_ = o.prof_;
    o.ctx_ = o.in_ = o.out_ = nullptr;
    o.in_cap_ = o.out_cap_ = 0;
  }
  return *this;
}

uint32_t FFTProcessorROCm::NextPowerOf2(uint32_t n) { return next_pow2(n); }

// nFFT = nextPow2(n_point) * repeat_count (fft_processor_rocm.cpp:665-677); the
// gfx950 transforms cover power-of-two lengths 2 .. 2^24.
void FFTProcessorROCm::CalculateNFFT(const FFTProcessorParams& p) {
  if (p.n_point == 0) throw std::invalid_argument("FFTProcessorROCm: n_point must be > 0");
  if (p.repeat_count == 0) throw std::invalid_argument("FFTProcessorROCm: repeat_count must be >= 1");
  const uint64_t n = (uint64_t)next_pow2(p.n_point) * p.repeat_count;
  if (n > (1u << 24) || !is_pow2((uint32_t)n))
    throw std::invalid_argument("FFTProcessorROCm: nFFT must be a power of two <= 2^24 = 16777216 (repeat_count a power of two)");
  n_point_ = p.n_point;
  nFFT_ = (uint32_t)std::max<uint64_t>(n, 2);
}

void FFTProcessorROCm::Reserve(void** buf, size_t* cap, size_t bytes) { reserve(buf, cap, bytes); }

namespace {

std::vector<FFTComplexResult> pack_complex(const std::vector<std::complex<float>>& all, size_t B, uint32_t nfft,
                                           float fs) {
  std::vector<FFTComplexResult> res(B);
  for (size_t b = 0; b < B; ++b) {
    res[b].beam_id = (uint32_t)b;
    res[b].nFFT = nfft;
    res[b].sample_rate = fs;
    res[b].spectrum.assign(all.begin() + b * nfft, all.begin() + (b + 1) * nfft);
  }
  return res;
}

// {|X|, arg X} interleaved -> per-beam magnitude / phase (/ frequency axis)
std::vector<FFTMagPhaseResult> pack_magphase(const std::vector<float>& inter, size_t B, uint32_t nfft,
                                             const FFTProcessorParams& params) {
  const bool freq = params.output_mode == FFTOutputMode::MAGNITUDE_PHASE_FREQ;
  std::vector<FFTMagPhaseResult> res(B);
  for (size_t b = 0; b < B; ++b) {
    auto& r = res[b];
    r.beam_id = (uint32_t)b;
    r.nFFT = nfft;
    r.sample_rate = params.sample_rate;
    r.magnitude.resize(nfft);
    r.phase.resize(nfft);
    const float* p = inter.data() + 2 * b * nfft;
    for (uint32_t k = 0; k < nfft; ++k) {
      r.magnitude[k] = p[2 * k];
      r.phase[k] = p[2 * k + 1];
    }
    if (freq) {
      r.frequency.resize(nfft);
      for (uint32_t k = 0; k < nfft; ++k) r.frequency[k] = (float)k * params.sample_rate / (float)nfft;
    }
  }
  return res;
}

}  // namespace

// Host vector in, host results out: streamed in chunks (spx_fft_c2c_host), so
// the upload, the transforms and the download overlap; their total is
// reported as the FFT time and the upload / download times are 0.
std::vector<FFTComplexResult> FFTProcessorROCm::ProcessComplex(const std::vector<std::complex<float>>& data,
                                                               const FFTProcessorParams& params,
                                                               ROCmProfEvents* prof_events) {
  CalculateNFFT(params);
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (params.beam_count == 0 || data.size() != total)
    throw std::invalid_argument("ProcessComplex: data size (" + std::to_string(data.size()) +
                                ") != beam_count * n_point (" + std::to_string(total) + ")");
  spx_ctx* c = (spx_ctx*)ctx_;
  const size_t B = params.beam_count;
  auto t0 = Clock::now();
  std::vector<std::complex<float>> all(B * nFFT_);
  spx_check(c, spx_fft_c2c_host(c, data.data(), all.data(), (uint32_t)B, params.n_point, nFFT_, SPX_WINDOW_NONE),
            "spx_fft_c2c_host");
  prof_.upload_time_ms = 0.0;
  prof_.fft_time_ms = ms_since(t0);
  prof_.download_time_ms = 0.0;
  auto r = pack_complex(all, B, nFFT_, params.sample_rate);
  prof_.post_processing_time_ms = 0.0;
  prof_.total_time_ms = ms_since(t0);
  if (prof_events) {
    prof_events->push_back({"Upload", prof_data(prof_.upload_time_ms, 0, "Upload")});
    prof_events->push_back({"FFT", prof_data(prof_.fft_time_ms, 1, "FFT")});
    prof_events->push_back({"Download", prof_data(prof_.download_time_ms, 2, "Download")});
  }
  return r;
}

std::vector<FFTComplexResult> FFTProcessorROCm::ProcessComplex(void* gpu_data, const FFTProcessorParams& params,
                                                               size_t gpu_memory_bytes) {
  if (!gpu_data) throw std::invalid_argument("ProcessComplex: gpu_data is null");
  CalculateNFFT(params);
  const size_t B = params.beam_count;
  if (B == 0) throw std::invalid_argument("ProcessComplex: beam_count must be > 0");
  if (gpu_memory_bytes && gpu_memory_bytes < B * params.n_point * 8)
    throw std::invalid_argument("ProcessComplex: gpu_memory_bytes smaller than beam_count * n_point * 8");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&out_, &out_cap_, B * nFFT_ * 8);
  auto t0 = Clock::now();
  spx_check(c, spx_fft_c2c(c, gpu_data, out_, (uint32_t)B, params.n_point, nFFT_, SPX_WINDOW_NONE), "spx_fft_c2c");
  spx_check(c, spx_synchronize(c), "spx_fft_c2c");
  prof_.fft_time_ms = ms_since(t0);
  t0 = Clock::now();
  std::vector<std::complex<float>> all(B * nFFT_);
  hip_check(hipMemcpyAsync(all.data(), out_, B * nFFT_ * 8, hipMemcpyDeviceToHost, stream_of(c)), "download");
  spx_check(c, spx_synchronize(c), "download");
  auto res = pack_complex(all, B, nFFT_, params.sample_rate);
  prof_.download_time_ms = ms_since(t0);
  prof_.post_processing_time_ms = 0.0;
  return res;
}

std::vector<FFTMagPhaseResult> FFTProcessorROCm::ProcessMagPhase(const std::vector<std::complex<float>>& data,
                                                                 const FFTProcessorParams& params,
                                                                 ROCmProfEvents* prof_events) {
  CalculateNFFT(params);
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (params.beam_count == 0 || data.size() != total)
    throw std::invalid_argument("ProcessMagPhase: data size != beam_count * n_point");
  spx_ctx* c = (spx_ctx*)ctx_;
  const size_t B = params.beam_count;
  auto t0 = Clock::now();
  std::vector<float> inter(B * nFFT_ * 2);
  spx_check(c, spx_fft_mag_phase_host(c, data.data(), inter.data(), (uint32_t)B, params.n_point, nFFT_,
                                      SPX_WINDOW_NONE), "spx_fft_mag_phase_host");
  prof_.upload_time_ms = 0.0;
  prof_.fft_time_ms = ms_since(t0);
  prof_.download_time_ms = 0.0;
  auto r = pack_magphase(inter, B, nFFT_, params);
  prof_.total_time_ms = ms_since(t0);
  if (prof_events) {
    prof_events->push_back({"Upload", prof_data(prof_.upload_time_ms, 0, "Upload")});
    prof_events->push_back({"FFT+MagPhase", prof_data(prof_.fft_time_ms, 1, "FFT+MagPhase")});
    prof_events->push_back({"Download", prof_data(prof_.download_time_ms, 2, "Download")});
  }
  return r;
}

std::vector<FFTMagPhaseResult> FFTProcessorROCm::ProcessMagPhase(void* gpu_data, const FFTProcessorParams& params,
                                                                 size_t gpu_memory_bytes) {
  if (!gpu_data) throw std::invalid_argument("ProcessMagPhase: gpu_data is null");
  CalculateNFFT(params);
  const size_t B = params.beam_count;
  if (B == 0) throw std::invalid_argument("ProcessMagPhase: beam_count must be > 0");
  if (gpu_memory_bytes && gpu_memory_bytes < B * params.n_point * 8)
    throw std::invalid_argument("ProcessMagPhase: gpu_memory_bytes too small");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&out_, &out_cap_, B * nFFT_ * 8);
  auto t0 = Clock::now();
  spx_check(c, spx_fft_mag_phase(c, gpu_data, out_, (uint32_t)B, params.n_point, nFFT_, SPX_WINDOW_NONE),
            "spx_fft_mag_phase");
  spx_check(c, spx_synchronize(c), "spx_fft_mag_phase");
  prof_.fft_time_ms = ms_since(t0);
  t0 = Clock::now();
  std::vector<float> inter(B * nFFT_ * 2);
  hip_check(hipMemcpyAsync(inter.data(), out_, inter.size() * 4, hipMemcpyDeviceToHost, stream_of(c)), "download");
  spx_check(c, spx_synchronize(c), "download");
  auto res = pack_magphase(inter, B, nFFT_, params);
  prof_.download_time_ms = ms_since(t0);
  return res;
}

void FFTProcessorROCm::ProcessMagnitudesToGPU(void* gpu_data, void* gpu_out, const FFTProcessorParams& params,
                                              bool squared, WindowType window, ROCmProfEvents* prof_events) {
  if (!gpu_data || !gpu_out) throw std::invalid_argument("ProcessMagnitudesToGPU: null device pointer");
  CalculateNFFT(params);
  if (params.beam_count == 0) throw std::invalid_argument("ProcessMagnitudesToGPU: beam_count must be > 0");
  spx_ctx* c = (spx_ctx*)ctx_;
  auto t0 = Clock::now();
  spx_check(c, spx_fft_magnitude(c, gpu_data, (float*)gpu_out, params.beam_count, params.n_point, nFFT_,
                                 (uint32_t)window, squared ? 1 : 0, 1.0f), "spx_fft_magnitude");
  spx_check(c, spx_synchronize(c), "spx_fft_magnitude");
  prof_.fft_time_ms = ms_since(t0);
  if (prof_events) prof_events->push_back({"FFT+Magnitude", prof_data(prof_.fft_time_ms, 1, "FFT+Magnitude")});
}

// ===========================================================================
// ComplexToMagPhaseROCm
// ===========================================================================
ComplexToMagPhaseROCm::ComplexToMagPhaseROCm(drv_gpu_lib::IBackend* backend) : backend_(backend) {
  ctx_ = make_ctx(backend, "ComplexToMagPhaseROCm");
}
ComplexToMagPhaseROCm::~ComplexToMagPhaseROCm() { Release(); }
void ComplexToMagPhaseROCm::Release() {
  if (in_) (void)hipFree(in_);
  if (out_) (void)hipFree(out_);
  in_ = out_ = nullptr;
  in_cap_ = out_cap_ = 0;
  if (ctx_) spx_destroy((spx_ctx*)ctx_);
  ctx_ = nullptr;
}
ComplexToMagPhaseROCm::ComplexToMagPhaseROCm(ComplexToMagPhaseROCm&& o) noexcept { *this = std::move(o); }
ComplexToMagPhaseROCm& ComplexToMagPhaseROCm::operator=(ComplexToMagPhaseROCm&& o) noexcept {
  if (this != &o) {
    Release();
    backend_ = o.backend_; ctx_ = o.ctx_; in_ = o.in_; in_cap_ = o.in_cap_; out_ = o.out_; out_cap_ = o.out_cap_;
    o.ctx_ = o.in_ = o.out_ = nullptr;
    o.in_cap_ = o.out_cap_ = 0;
  }
  return *this;
}
void ComplexToMagPhaseROCm::Reserve(void** buf, size_t* cap, size_t bytes) { reserve(buf, cap, bytes); }

// complex_to_mag_phase_rocm.cpp:170-176
float ComplexToMagPhaseROCm::ComputeInvN(const MagPhaseParams& p) {
  if (p.norm_coeff < 0.0f) return p.n_point > 0 ? 1.0f / (float)p.n_point : 1.0f;
  if (p.norm_coeff > 0.0f) return p.norm_coeff;
  return 1.0f;
}

static void check_mp(const MagPhaseParams& p, const char* who) {
  if (p.beam_count == 0 || p.n_point == 0) throw std::invalid_argument(std::string(who) + ": empty shape");
}

std::vector<MagPhaseResult> ComplexToMagPhaseROCm::Process(const std::vector<std::complex<float>>& data,
                                                           const MagPhaseParams& params) {
  check_mp(params, "Process");
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (data.size() != total) throw std::invalid_argument("Process: data size != beam_count * n_point");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&in_, &in_cap_, total * 8);
  hip_check(hipMemcpyAsync(in_, data.data(), total * 8, hipMemcpyHostToDevice, stream_of(c)), "upload");
  return Process(in_, params, total * 8);
}

std::vector<MagPhaseResult> ComplexToMagPhaseROCm::Process(void* gpu_data, const MagPhaseParams& params,
                                                           size_t gpu_memory_bytes) {
  check_mp(params, "Process");
  if (!gpu_data) throw std::invalid_argument("Process: gpu_data is null");
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (gpu_memory_bytes && gpu_memory_bytes < total * 8) throw std::invalid_argument("Process: buffer too small");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&out_, &out_cap_, total * 8);
  spx_check(c, spx_mag_phase(c, gpu_data, out_, total), "spx_mag_phase");
  std::vector<float> inter(total * 2);
  hip_check(hipMemcpyAsync(inter.data(), out_, total * 8, hipMemcpyDeviceToHost, stream_of(c)), "download");
  spx_check(c, spx_synchronize(c), "download");
  std::vector<MagPhaseResult> res(params.beam_count);
  for (uint32_t b = 0; b < params.beam_count; ++b) {
    res[b].beam_id = b;
    res[b].n_point = params.n_point;
    res[b].magnitude.resize(params.n_point);
    res[b].phase.resize(params.n_point);
    const float* p = inter.data() + 2 * (size_t)b * params.n_point;
    for (uint32_t i = 0; i < params.n_point; ++i) {
      res[b].magnitude[i] = p[2 * i];
      res[b].phase[i] = p[2 * i + 1];
    }
  }
  return res;
}

void* ComplexToMagPhaseROCm::ProcessToGPU(const std::vector<std::complex<float>>& data, const MagPhaseParams& params) {
  check_mp(params, "ProcessToGPU");
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (data.size() != total) throw std::invalid_argument("ProcessToGPU: data size != beam_count * n_point");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&in_, &in_cap_, total * 8);
  hip_check(hipMemcpyAsync(in_, data.data(), total * 8, hipMemcpyHostToDevice, stream_of(c)), "upload");
  return ProcessToGPU(in_, params, total * 8);
}

void* ComplexToMagPhaseROCm::ProcessToGPU(void* gpu_data, const MagPhaseParams& params, size_t gpu_memory_bytes) {
  check_mp(params, "ProcessToGPU");
  if (!gpu_data) throw std::invalid_argument("ProcessToGPU: gpu_data is null");
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (gpu_memory_bytes && gpu_memory_bytes < total * 8) throw std::invalid_argument("ProcessToGPU: buffer too small");
  spx_ctx* c = (spx_ctx*)ctx_;
  void* out = nullptr;
  hip_check(hipMalloc(&out, total * 8), "hipMalloc");  // caller owns
  int rc = spx_mag_phase(c, gpu_data, out, total);
  if (rc == SPX_OK) rc = spx_synchronize(c);
  if (rc != SPX_OK) {
    (void)hipFree(out);
    spx_check(c, rc, "spx_mag_phase");
  }
  return out;
}

std::vector<MagnitudeResult> ComplexToMagPhaseROCm::ProcessMagnitude(void* gpu_data, const MagPhaseParams& params,
                                                                     size_t gpu_memory_bytes) {
  check_mp(params, "ProcessMagnitude");
  if (!gpu_data) throw std::invalid_argument("ProcessMagnitude: gpu_data is null");
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (gpu_memory_bytes && gpu_memory_bytes < total * 8) throw std::invalid_argument("ProcessMagnitude: buffer too small");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&out_, &out_cap_, total * 4);
  spx_check(c, spx_magnitude(c, gpu_data, (float*)out_, total, ComputeInvN(params), 0), "spx_magnitude");
  std::vector<float> m(total);
  hip_check(hipMemcpyAsync(m.data(), out_, total * 4, hipMemcpyDeviceToHost, stream_of(c)), "download");
  spx_check(c, spx_synchronize(c), "download");
  std::vector<MagnitudeResult> res(params.beam_count);
  for (uint32_t b = 0; b < params.beam_count; ++b) {
    res[b].beam_id = b;
    res[b].n_point = params.n_point;
    res[b].magnitude.assign(m.begin() + (size_t)b * params.n_point, m.begin() + (size_t)(b + 1) * params.n_point);
  }
  return res;
}

void* ComplexToMagPhaseROCm::ProcessMagnitudeToGPU(void* gpu_data, const MagPhaseParams& params,
                                                   size_t gpu_memory_bytes) {
  check_mp(params, "ProcessMagnitudeToGPU");
  if (!gpu_data) throw std::invalid_argument("ProcessMagnitudeToGPU: gpu_data is null");
  const size_t total = (size_t)params.beam_count * params.n_point;
  if (gpu_memory_bytes && gpu_memory_bytes < total * 8) throw std::invalid_argument("ProcessMagnitudeToGPU: buffer too small");
  spx_ctx* c = (spx_ctx*)ctx_;
  void* out = nullptr;
  hip_check(hipMalloc(&out, total * 4), "hipMalloc");  // caller owns
  int rc = spx_magnitude(c, gpu_data, (float*)out, total, ComputeInvN(params), 0);
  if (rc == SPX_OK) rc = spx_synchronize(c);
  if (rc != SPX_OK) {
    (void)hipFree(out);
    spx_check(c, rc, "spx_magnitude");
  }
  return out;
}

void ComplexToMagPhaseROCm::ProcessMagnitudeToBuffer(void* in, void* out, const MagPhaseParams& params) {
  check_mp(params, "ProcessMagnitudeToBuffer");
  if (!in || !out) throw std::invalid_argument("ProcessMagnitudeToBuffer: null device pointer");
  spx_ctx* c = (spx_ctx*)ctx_;
  const size_t total = (size_t)params.beam_count * params.n_point;
  spx_check(c, spx_magnitude(c, in, (float*)out, total, ComputeInvN(params), 0), "spx_magnitude");
  spx_check(c, spx_synchronize(c), "spx_magnitude");
}

}  // namespace fft_processor

// ===========================================================================
// SpectrumProcessorROCm, AllMaximaPipelineROCm, factory
// ===========================================================================
namespace antenna_fft {

namespace {

// AllMaxima range rules of all_maxima_pipeline_rocm.cpp:274-276 and capacity :304-307
void normalise_range(uint32_t nfft, uint32_t& s0, uint32_t& s1) {
  if (s0 == 0) s0 = 1;
  if (s1 == 0) s1 = nfft / 2;
  if (s1 >= nfft) s1 = nfft - 1;
  if (s0 > s1) throw std::invalid_argument("AllMaxima: search_start > search_end");
}

uint32_t capacity(uint32_t s0, uint32_t s1, size_t max_per_beam) {
  return (uint32_t)std::min<size_t>((s1 - s0) / 2, max_per_beam);
}

// Build the result from device records [B][cap] and per-beam found counts.
// Per-beam host results from the records [B][cap] and the uncapped counts.
void fill_cpu_beams(AllMaximaResult& res, const std::vector<MaxValue>& all, const std::vector<uint32_t>& found,
                    uint32_t cap) {
  const uint32_t B = (uint32_t)found.size();
  res.beams.resize(B);
  for (uint32_t b = 0; b < B; ++b) {
    const uint32_t n = std::min(found[b], cap);  // truncation keeps the lowest bins
    res.beams[b].antenna_id = b;
    res.beams[b].num_maxima = n;
    res.beams[b].maxima.assign(all.begin() + (size_t)b * cap, all.begin() + (size_t)b * cap + n);
    res.total_maxima += n;
  }
}

AllMaximaResult finish_all_maxima(spx_ctx* c, const void* rec_dev, const std::vector<uint32_t>& found, uint32_t cap,
                                  OutputDestination dest) {
  const uint32_t B = (uint32_t)found.size();
  AllMaximaResult res;
  res.destination = dest;
  if (dest == OutputDestination::CPU || dest == OutputDestination::ALL) {
    std::vector<MaxValue> all((size_t)B * cap);
    if (cap) {
      hip_check(hipMemcpyAsync(all.data(), rec_dev, all.size() * sizeof(MaxValue), hipMemcpyDeviceToHost,
                               stream_of(c)), "download");
      spx_check(c, spx_synchronize(c), "download");
    }
    fill_cpu_beams(res, all, found, cap);
  } else {
    for (uint32_t f : found) res.total_maxima += f;
  }
  if (dest == OutputDestination::GPU || dest == OutputDestination::ALL) {
    const size_t bytes = (size_t)B * cap * sizeof(MaxValue);
    void* gm = nullptr;
    void* gc = nullptr;
    hip_check(hipMalloc(&gm, bytes ? bytes : 1), "hipMalloc");
    if (hipMalloc(&gc, (size_t)B * 4) != hipSuccess) {
      (void)hipFree(gm);
      throw std::runtime_error("hipMalloc counts failed");
    }
    if (bytes) hip_check(hipMemcpyAsync(gm, rec_dev, bytes, hipMemcpyDeviceToDevice, stream_of(c)), "copy");
    hip_check(hipMemcpyAsync(gc, found.data(), (size_t)B * 4, hipMemcpyHostToDevice, stream_of(c)), "copy");
    spx_check(c, spx_synchronize(c), "copy");
    res.gpu_maxima = gm;
    res.gpu_counts = gc;
    res.gpu_bytes = bytes;
  }
  return res;
}

}  // namespace

SpectrumProcessorROCm::SpectrumProcessorROCm(drv_gpu_lib::IBackend* backend) : backend_(backend) {
  ctx_ = make_ctx(backend, "SpectrumProcessorROCm");
}

SpectrumProcessorROCm::~SpectrumProcessorROCm() {
  if (in_) (void)hipFree(in_);
  if (spec_) (void)hipFree(spec_);
  if (rec_) (void)hipFree(rec_);
  if (ctx_) spx_destroy((spx_ctx*)ctx_);
}

uint32_t SpectrumProcessorROCm::NextPowerOf2(uint32_t n) { return next_pow2(n); }

void SpectrumProcessorROCm::Reserve(void** buf, size_t* cap, size_t bytes) { reserve(buf, cap, bytes); }

// spectrum_processor_rocm.cpp:115-156, 1137-1143; re-plans on every call.
void SpectrumProcessorROCm::Initialize(const SpectrumParams& params) {
  SpectrumParams p = params;
  if (p.antenna_count == 0 || p.n_point == 0) throw std::invalid_argument("Initialize: empty shape");
  if (p.repeat_count == 0) throw std::invalid_argument("Initialize: repeat_count must be >= 1");
  p.base_fft = next_pow2(p.n_point);
  const uint64_t n = (uint64_t)p.base_fft * p.repeat_count;
  if (n > (1u << 24) || !is_pow2((uint32_t)n) || n < 4)
    throw std::invalid_argument("Initialize: nFFT must be a power of two in [4, 2^24 = 16777216]");
  p.nFFT = (uint32_t)n;
  if (p.search_range == 0) p.search_range = p.nFFT / 4;
  params_ = p;
  initialized_ = true;
}

void SpectrumProcessorROCm::RequireInit() const {
  if (!initialized_) throw std::runtime_error("SpectrumProcessorROCm: Initialize() was not called");
}

void SpectrumProcessorROCm::ReallocateBuffersForBatch(size_t n) {
  RequireInit();
  Reserve(&in_, &in_cap_, n * params_.n_point * 8);
  Reserve(&spec_, &spec_cap_, n * params_.nFFT * 8);
}

size_t SpectrumProcessorROCm::CalculateBytesPerAntenna() const {
  RequireInit();
  return (size_t)params_.n_point * 8 + (size_t)params_.nFFT * 8 + 8 * sizeof(MaxValue);
}

std::vector<SpectrumResult> SpectrumProcessorROCm::PeaksFromDevice(const void* signal, size_t beams, size_t start) {
  spx_ctx* c = (spx_ctx*)ctx_;
  const uint32_t N = params_.nFFT;
  const bool two = params_.peak_mode == PeakSearchMode::TWO_PEAKS;
  const size_t R = two ? 8 : 4;
  Reserve(&rec_, &rec_cap_, beams * R * sizeof(MaxValue));
  auto t0 = Clock::now();
  // pad + FFT + both argmaxes + records in the transform's own launch (three
  // launches on the split path, 2^15 .. 2^18), no spectrum in HBM
  spx_check(c, spx_fused_peaks_top(c, signal, rec_, (uint32_t)beams, params_.n_point, N, SPX_WINDOW_NONE,
                                   params_.search_range, params_.sample_rate, two ? 1 : 0), "spx_fused_peaks_top");
  spx_check(c, spx_synchronize(c), "peaks");
  prof_.fft_time_ms = ms_since(t0);
  t0 = Clock::now();
  std::vector<MaxValue> raw(beams * R);
  hip_check(hipMemcpyAsync(raw.data(), rec_, raw.size() * sizeof(MaxValue), hipMemcpyDeviceToHost, stream_of(c)),
            "download");
  spx_check(c, spx_synchronize(c), "download");
  prof_.download_time_ms = ms_since(t0);
  std::vector<SpectrumResult> out;
  out.reserve(beams * (two ? 2 : 1));
  for (size_t i = 0; i < beams; ++i) {
    for (size_t h = 0; h < (two ? 2u : 1u); ++h) {
      SpectrumResult r{};
      r.antenna_id = (uint32_t)(start + i);
      const MaxValue* m = raw.data() + i * R + h * 4;
      r.interpolated = m[0];
      r.left_point = m[1];
      r.center_point = m[2];
      r.right_point = m[3];
      out.push_back(r);
    }
  }
  return out;
}

std::vector<SpectrumResult> SpectrumProcessorROCm::ProcessFromCPU(const std::vector<std::complex<float>>& data,
                                                                  ROCmProfEvents* prof_events) {
  RequireInit();
  const size_t total = (size_t)params_.antenna_count * params_.n_point;
  if (data.size() != total) throw std::invalid_argument("ProcessFromCPU: data size != antenna_count * n_point");
  return ProcessBatch(data, 0, params_.antenna_count, prof_events);
}

std::vector<SpectrumResult> SpectrumProcessorROCm::ProcessBatch(const std::vector<std::complex<float>>& batch_data,
                                                                size_t start_antenna, size_t count,
                                                                ROCmProfEvents* prof_events) {
  RequireInit();
  if (count == 0 || batch_data.size() != count * params_.n_point)
    throw std::invalid_argument("ProcessBatch: data size != batch_antenna_count * n_point");
  spx_ctx* c = (spx_ctx*)ctx_;
  auto t0 = Clock::now();
  Reserve(&in_, &in_cap_, batch_data.size() * 8);
  hip_check(hipMemcpyAsync(in_, batch_data.data(), batch_data.size() * 8, hipMemcpyHostToDevice, stream_of(c)),
            "upload");
  spx_check(c, spx_synchronize(c), "upload");
  prof_.upload_time_ms = ms_since(t0);
  auto r = PeaksFromDevice(in_, count, start_antenna);
  prof_.total_time_ms = ms_since(t0);
  if (prof_events) {
    prof_events->push_back({"Upload", prof_data(prof_.upload_time_ms, 0, "Upload")});
    prof_events->push_back({"FFT+Post", prof_data(prof_.fft_time_ms, 1, "FFT+Post")});
    prof_events->push_back({"Download", prof_data(prof_.download_time_ms, 2, "Download")});
  }
  return r;
}

std::vector<SpectrumResult> SpectrumProcessorROCm::ProcessFromGPU(void* gpu_data, size_t antenna_count,
                                                                  size_t n_point, size_t gpu_memory_bytes) {
  RequireInit();
  if (!gpu_data) throw std::invalid_argument("ProcessFromGPU: gpu_data is null");
  if (antenna_count == 0 || n_point != params_.n_point)
    throw std::invalid_argument("ProcessFromGPU: n_point differs from the initialised n_point");
  if (gpu_memory_bytes && gpu_memory_bytes < antenna_count * n_point * 8)
    throw std::invalid_argument("ProcessFromGPU: gpu_memory_bytes too small");
  return PeaksFromDevice(gpu_data, antenna_count, 0);
}

std::vector<SpectrumResult> SpectrumProcessorROCm::ProcessBatchFromGPU(void* gpu_data, size_t src_offset_bytes,
                                                                       size_t start_antenna, size_t count) {
  RequireInit();
  if (!gpu_data || count == 0) throw std::invalid_argument("ProcessBatchFromGPU: null data or empty batch");
  return PeaksFromDevice((const char*)gpu_data + src_offset_bytes, count, start_antenna);
}

AllMaximaResult SpectrumProcessorROCm::AllMaximaFromSignalDevice(const void* signal, size_t beams,
                                                                 OutputDestination dest, uint32_t s0, uint32_t s1) {
  spx_ctx* c = (spx_ctx*)ctx_;
  const uint32_t N = params_.nFFT;
  normalise_range(N, s0, s1);
  const uint32_t cap = capacity(s0, s1, params_.max_maxima_per_beam);
  Reserve(&rec_, &rec_cap_, beams * std::max<uint32_t>(cap, 1) * sizeof(MaxValue) + beams * 8);
  uint32_t* counts = (uint32_t*)((char*)rec_ + beams * std::max<uint32_t>(cap, 1) * sizeof(MaxValue));
  spx_fused_desc d;
  std::memset(&d, 0, sizeof(d));
  d.input = signal;
  d.batch = (uint32_t)beams;
  d.n_point = params_.n_point;
  d.nfft = N;
  d.window = SPX_WINDOW_NONE;
  d.dtype = SPX_DTYPE_C64;
  d.sample_rate = params_.sample_rate;
  d.peak_mode = SPX_PEAKS_ALL_MAXIMA;
  d.max_peaks = cap;
  d.search_start = s0;
  d.search_end = s1;
  d.refine = SPX_REFINE_NONE;
  d.peaks_out = (spx_max_value*)rec_;
  d.counts_out = counts;
  auto t0 = Clock::now();
  spx_check(c, spx_fused_peaks(c, &d), "spx_fused_peaks");
  std::vector<uint32_t> ct(beams * 2);
  hip_check(hipMemcpyAsync(ct.data(), counts, beams * 8, hipMemcpyDeviceToHost, stream_of(c)), "download");
  spx_check(c, spx_synchronize(c), "spx_fused_peaks");
  prof_.fft_time_ms = ms_since(t0);
  std::vector<uint32_t> found(beams);
  for (size_t b = 0; b < beams; ++b) found[b] = ct[2 * b + 1];
  return finish_all_maxima(c, rec_, found, cap, dest);
}

AllMaximaResult SpectrumProcessorROCm::AllMaximaFromSpectrumDevice(const void* spec, uint32_t beams, uint32_t nfft,
                                                                   float fs, OutputDestination dest, uint32_t s0,
                                                                   uint32_t s1) {
  spx_ctx* c = (spx_ctx*)ctx_;
  if (nfft < 3) throw std::invalid_argument("AllMaxima: nFFT must be >= 3");
  normalise_range(nfft, s0, s1);
  const size_t maxpb = initialized_ ? params_.max_maxima_per_beam : 1000;
  const uint32_t cap = capacity(s0, s1, maxpb);
  Reserve(&rec_, &rec_cap_, (size_t)beams * std::max<uint32_t>(cap, 1) * sizeof(MaxValue) + (size_t)beams * 4);
  uint32_t* counts = (uint32_t*)((char*)rec_ + (size_t)beams * std::max<uint32_t>(cap, 1) * sizeof(MaxValue));
  spx_check(c, spx_all_maxima(c, spec, rec_, counts, beams, nfft, fs, s0, s1, cap), "spx_all_maxima");
  std::vector<uint32_t> found(beams);
  hip_check(hipMemcpyAsync(found.data(), counts, (size_t)beams * 4, hipMemcpyDeviceToHost, stream_of(c)), "download");
  spx_check(c, spx_synchronize(c), "spx_all_maxima");
  return finish_all_maxima(c, rec_, found, cap, dest);
}

AllMaximaResult SpectrumProcessorROCm::FindAllMaximaFromCPU(const std::vector<std::complex<float>>& data,
                                                            OutputDestination dest, uint32_t s0, uint32_t s1,
                                                            ROCmProfEvents* prof_events) {
  RequireInit();
  const size_t total = (size_t)params_.antenna_count * params_.n_point;
  if (data.size() != total) throw std::invalid_argument("FindAllMaximaFromCPU: data size != antenna_count * n_point");
  spx_ctx* c = (spx_ctx*)ctx_;
  auto t0 = Clock::now();
  if (dest == OutputDestination::CPU) {
    // host results only: streamed in chunks (spx_fused_peaks_host), upload and
    // search overlapping
    const uint32_t N = params_.nFFT;
    normalise_range(N, s0, s1);
    const uint32_t cap = capacity(s0, s1, params_.max_maxima_per_beam);
    const size_t beams = params_.antenna_count;
    std::vector<MaxValue> all(beams * cap);
    std::vector<uint32_t> ct(beams * 2);
    spx_fused_desc d;
    std::memset(&d, 0, sizeof(d));
    d.input = data.data();
    d.batch = (uint32_t)beams;
    d.n_point = params_.n_point;
    d.nfft = N;
    d.window = SPX_WINDOW_NONE;
    d.dtype = SPX_DTYPE_C64;
    d.sample_rate = params_.sample_rate;
    d.peak_mode = SPX_PEAKS_ALL_MAXIMA;
    d.max_peaks = cap;
    d.search_start = s0;
    d.search_end = s1;
    d.refine = SPX_REFINE_NONE;
    d.peaks_out = cap ? (spx_max_value*)all.data() : nullptr;
    d.counts_out = ct.data();
    spx_check(c, spx_fused_peaks_host(c, &d), "spx_fused_peaks_host");
    prof_.upload_time_ms = 0.0;
    prof_.fft_time_ms = ms_since(t0);
    std::vector<uint32_t> found(beams);
    for (size_t b = 0; b < beams; ++b) found[b] = ct[2 * b + 1];
    AllMaximaResult r;
    r.destination = dest;
    fill_cpu_beams(r, all, found, cap);
    if (prof_events) {
      prof_events->push_back({"Upload", prof_data(prof_.upload_time_ms, 0, "Upload")});
      prof_events->push_back({"FusedAllMaxima", prof_data(prof_.fft_time_ms, 1, "FusedAllMaxima")});
    }
    return r;
  }
  Reserve(&in_, &in_cap_, total * 8);
  hip_check(hipMemcpyAsync(in_, data.data(), total * 8, hipMemcpyHostToDevice, stream_of(c)), "upload");
  spx_check(c, spx_synchronize(c), "upload");
  prof_.upload_time_ms = ms_since(t0);
  auto r = AllMaximaFromSignalDevice(in_, params_.antenna_count, dest, s0, s1);
  if (prof_events) {
    prof_events->push_back({"Upload", prof_data(prof_.upload_time_ms, 0, "Upload")});
    prof_events->push_back({"FusedAllMaxima", prof_data(prof_.fft_time_ms, 1, "FusedAllMaxima")});
  }
  return r;
}

AllMaximaResult SpectrumProcessorROCm::FindAllMaximaFromGPUPipeline(void* gpu_data, size_t antenna_count,
                                                                    size_t n_point, size_t gpu_memory_bytes,
                                                                    OutputDestination dest, uint32_t s0,
                                                                    uint32_t s1) {
  RequireInit();
  if (!gpu_data) throw std::invalid_argument("FindAllMaximaFromGPUPipeline: gpu_data is null");
  if (antenna_count == 0 || n_point != params_.n_point)
    throw std::invalid_argument("FindAllMaximaFromGPUPipeline: n_point differs from the initialised n_point");
  if (gpu_memory_bytes && gpu_memory_bytes < antenna_count * n_point * 8)
    throw std::invalid_argument("FindAllMaximaFromGPUPipeline: gpu_memory_bytes too small");
  return AllMaximaFromSignalDevice(gpu_data, antenna_count, dest, s0, s1);
}

AllMaximaResult SpectrumProcessorROCm::AllMaximaFromCPU(const std::vector<std::complex<float>>& fft_data,
                                                        uint32_t beam_count, uint32_t nFFT, float fs,
                                                        OutputDestination dest, uint32_t s0, uint32_t s1) {
  if (beam_count == 0 || fft_data.size() != (size_t)beam_count * nFFT)
    throw std::invalid_argument("AllMaximaFromCPU: fft_data size != beam_count * nFFT");
  spx_ctx* c = (spx_ctx*)ctx_;
  Reserve(&spec_, &spec_cap_, fft_data.size() * 8);
  hip_check(hipMemcpyAsync(spec_, fft_data.data(), fft_data.size() * 8, hipMemcpyHostToDevice, stream_of(c)),
            "upload");
  return AllMaximaFromSpectrumDevice(spec_, beam_count, nFFT, fs, dest, s0, s1);
}

AllMaximaResult SpectrumProcessorROCm::FindAllMaxima(void* fft_data, uint32_t beam_count, uint32_t nFFT, float fs,
                                                     OutputDestination dest, uint32_t s0, uint32_t s1) {
  if (!fft_data || beam_count == 0) throw std::invalid_argument("FindAllMaxima: null data or empty batch");
  return AllMaximaFromSpectrumDevice(fft_data, beam_count, nFFT, fs, dest, s0, s1);
}

AllMaximaPipelineROCm::AllMaximaPipelineROCm(drv_gpu_lib::IBackend* backend) : backend_(backend) {
  ctx_ = make_ctx(backend, "AllMaximaPipelineROCm");
}
AllMaximaPipelineROCm::~AllMaximaPipelineROCm() {
  if (ctx_) spx_destroy((spx_ctx*)ctx_);
}

AllMaximaResult AllMaximaPipelineROCm::Execute(void* magnitudes_gpu, void* fft_data_gpu, uint32_t beam_count,
                                               uint32_t nFFT, float fs, OutputDestination dest, uint32_t s0,
                                               uint32_t s1, size_t max_maxima_per_beam) {
  if (!magnitudes_gpu) throw std::invalid_argument("AllMaximaPipelineROCm::Execute: magnitudes_gpu cannot be null");
  if (!fft_data_gpu) throw std::invalid_argument("AllMaximaPipelineROCm::Execute: fft_data_gpu required");
  if (beam_count == 0 || nFFT < 3) throw std::invalid_argument("AllMaximaPipelineROCm::Execute: empty shape");
  spx_ctx* c = (spx_ctx*)ctx_;
  normalise_range(nFFT, s0, s1);
  const uint32_t cap = capacity(s0, s1, max_maxima_per_beam);
  void* rec = nullptr;
  void* cnt = nullptr;
  hip_check(hipMalloc(&rec, (size_t)beam_count * std::max<uint32_t>(cap, 1) * sizeof(MaxValue)), "hipMalloc");
  if (hipMalloc(&cnt, (size_t)beam_count * 4) != hipSuccess) {
    (void)hipFree(rec);
    throw std::runtime_error("hipMalloc failed");
  }
  AllMaximaResult res;
  try {
    spx_check(c, spx_all_maxima(c, fft_data_gpu, rec, (uint32_t*)cnt, beam_count, nFFT, fs, s0, s1, cap),
              "spx_all_maxima");
    std::vector<uint32_t> found(beam_count);
    hip_check(hipMemcpyAsync(found.data(), cnt, (size_t)beam_count * 4, hipMemcpyDeviceToHost, stream_of(c)),
              "download");
    spx_check(c, spx_synchronize(c), "spx_all_maxima");
    res = finish_all_maxima(c, rec, found, cap, dest);
  } catch (...) {
    (void)hipFree(rec);
    (void)hipFree(cnt);
    throw;
  }
  (void)hipFree(rec);
  (void)hipFree(cnt);
  return res;
}

std::unique_ptr<ISpectrumProcessor> SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType t,
                                                                     drv_gpu_lib::IBackend* backend) {
  if (!backend) throw std::invalid_argument("SpectrumProcessorFactory::Create: backend cannot be null");
  switch (t) {
    case drv_gpu_lib::BackendType::ROCm:
    case drv_gpu_lib::BackendType::AUTO:
    case drv_gpu_lib::BackendType::OPENCLandROCm:
      return std::make_unique<SpectrumProcessorROCm>(backend);
    case drv_gpu_lib::BackendType::OPENCL:
    default:
      throw std::runtime_error("SpectrumProcessorFactory::Create: only ROCm backends are supported");
  }
}

}  // namespace antenna_fft
