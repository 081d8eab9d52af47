// C++ drop-in test: the reference's own facade-level test cases
// (tests/test_spectrum_maxima_rocm.hpp:87-250, tests/test_fft_processor_rocm.hpp,
// tests/test_complex_to_mag_phase_rocm.hpp, tests/test_process_magnitude_rocm.hpp
// of the reference), written against the headers under include/ and linked
// to libspectrum_facades.so.  Exit status 0 when every check passes; each
// check prints one line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <spectrum/complex_to_mag_phase_rocm.hpp>
#include <spectrum/factory/spectrum_processor_factory.hpp>
#include <spectrum/fft_processor_rocm.hpp>
#include <spectrum/processors/spectrum_processor_rocm.hpp>
#include <spectrum/utils/cpu_fft.hpp>
#include <stdexcept>
#include <string>
#include <vector>

using namespace antenna_fft;
using cvec = std::vector<std::complex<float>>;

static int g_fail = 0;
static void check(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "[PASS]" : "[FAIL]", what.c_str());
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

static cvec cw(float f, float fs, uint32_t n, uint32_t beams) {
  cvec d((size_t)beams * n);
  for (uint32_t b = 0; b < beams; ++b)
    for (uint32_t t = 0; t < n; ++t) d[(size_t)b * n + t] = {std::sin(2.0f * (float)M_PI * f * t / fs), 0.0f};
  return d;
}
static cvec multi_cw(const std::vector<float>& fr, float fs, uint32_t n, uint32_t beams) {
  cvec d((size_t)beams * n);
  for (uint32_t b = 0; b < beams; ++b)
    for (uint32_t t = 0; t < n; ++t) {
      float v = 0;
      for (float f : fr) v += std::sin(2.0f * (float)M_PI * f * t / fs);
      d[(size_t)b * n + t] = {v, 0.0f};
    }
  return d;
}

int main() {
  drv_gpu_lib::ROCmBackend backend;
  backend.Initialize(0);
  std::printf("device %s\n", backend.GetDeviceName().c_str());

  // ---- spectrum maxima (reference tests 1-5)
  {
    SpectrumParams p;
    p.antenna_count = 4; p.n_point = 1000; p.repeat_count = 1; p.sample_rate = 10000.0f;
    p.peak_mode = PeakSearchMode::ONE_PEAK;
    auto proc = SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::ROCm, &backend);
    proc->Initialize(p);
    auto r = proc->ProcessFromCPU(cw(100.0f, 10000.0f, 1000, 4));
    bool ok = r.size() == 4;
    for (auto& x : r) ok = ok && std::fabs(x.interpolated.refined_frequency - 100.0f) < 5.0f;
    check(ok, "one_peak: 4 antennas, CW 100 Hz within 5 Hz");
  }
  {
    SpectrumParams p;
    p.antenna_count = 2; p.n_point = 1000; p.repeat_count = 1; p.sample_rate = 10000.0f;
    p.peak_mode = PeakSearchMode::TWO_PEAKS;
    auto proc = SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::ROCm, &backend);
    proc->Initialize(p);
    auto r = proc->ProcessFromCPU(multi_cw({100.0f, 300.0f}, 10000.0f, 1000, 2));
    check(r.size() == 4, "two_peaks: two results per antenna");
  }
  {
    SpectrumParams p;
    p.antenna_count = 2; p.n_point = 1024; p.repeat_count = 1; p.sample_rate = 1000.0f;
    p.peak_mode = PeakSearchMode::ONE_PEAK;
    auto proc = SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::ROCm, &backend);
    proc->Initialize(p);
    auto data = multi_cw({50.0f, 120.0f, 200.0f}, 1000.0f, 1024, 2);
    auto r = proc->FindAllMaximaFromCPU(data, OutputDestination::CPU, 1, 0);
    bool ok = r.total_maxima > 0 && r.beams.size() == 2;
    for (auto& b : r.beams) ok = ok && b.num_maxima >= 3;
    check(ok, "find_all_maxima: 3 tones, >= 3 maxima per beam");
    // GPU destination: caller-owned device arrays equal the CPU result
    auto g = proc->FindAllMaximaFromCPU(data, OutputDestination::GPU, 1, 0);
    bool okg = g.gpu_maxima && g.gpu_counts && g.beams.empty();
    if (okg) {
      const uint32_t cap = (uint32_t)(g.gpu_bytes / sizeof(MaxValue) / 2);
      std::vector<uint32_t> cnt(2);
      std::vector<MaxValue> mv((size_t)2 * cap);
      okg = hipMemcpy(cnt.data(), g.gpu_counts, 8, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(mv.data(), g.gpu_maxima, mv.size() * sizeof(MaxValue), hipMemcpyDeviceToHost) == hipSuccess;
      for (int b = 0; b < 2 && okg; ++b) {
        okg = okg && std::min(cnt[b], cap) == r.beams[b].maxima.size();
        for (size_t i = 0; i < r.beams[b].maxima.size() && okg; ++i)
          okg = mv[(size_t)b * cap + i].index == r.beams[b].maxima[i].index;
      }
      (void)hipFree(g.gpu_maxima);
      (void)hipFree(g.gpu_counts);
    }
    check(okg, "find_all_maxima: GPU destination equals CPU destination");
  }
  {
    const uint32_t N = 1024;
    cvec fft(N, {0.1f, 0.0f});
    fft[50] = {100.0f, 0.0f};
    fft[120] = {80.0f, 0.0f};
    fft[200] = {60.0f, 0.0f};
    SpectrumParams p;
    p.antenna_count = 1; p.n_point = N; p.repeat_count = 1; p.sample_rate = 1000.0f;
    auto proc = SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::ROCm, &backend);
    proc->Initialize(p);
    auto r = proc->AllMaximaFromCPU(fft, 1, N, 1000.0f, OutputDestination::CPU, 1, N / 2);
    bool ok = r.beams.size() == 1 && r.beams[0].maxima.size() == 3 && r.beams[0].maxima[0].index == 50 &&
              r.beams[0].maxima[1].index == 120 && r.beams[0].maxima[2].index == 200;
    check(ok, "all_maxima_fft: exactly {50, 120, 200}");
  }
  {
    const uint32_t n = 1000, A = 16;
    cvec d((size_t)A * n);
    for (uint32_t b = 0; b < A; ++b)
      for (uint32_t t = 0; t < n; ++t)
        d[(size_t)b * n + t] = {std::sin(2.0f * (float)M_PI * (100.0f + 10.0f * b) * t / 10000.0f), 0.0f};
    SpectrumParams p;
    p.antenna_count = A; p.n_point = n; p.repeat_count = 1; p.sample_rate = 10000.0f;
    auto proc = SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::ROCm, &backend);
    proc->Initialize(p);
    auto r = proc->ProcessFromCPU(d);
    bool ok = r.size() == A;
    for (auto& x : r) ok = ok && std::fabs(x.interpolated.refined_frequency - (100.0f + 10.0f * x.antenna_id)) < 10.0f;
    check(ok, "batch_16_beams: 100 + 10 b Hz within 10 Hz");
    // device-pointer path gives the same records
    void* dev = nullptr;
    bool okd = hipMalloc(&dev, d.size() * 8) == hipSuccess && hipMemcpy(dev, d.data(), d.size() * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (okd) {
      auto rg = proc->ProcessFromGPU(dev, A, n, d.size() * 8);
      okd = rg.size() == r.size();
      for (size_t i = 0; i < rg.size() && okd; ++i) okd = rg[i].interpolated.index == r[i].interpolated.index;
    }
    if (dev) (void)hipFree(dev);
    check(okd, "ProcessFromGPU equals ProcessFromCPU");
  }
  {
    bool ok = false;
    try {
      SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::ROCm, nullptr);
    } catch (const std::invalid_argument&) {
      ok = true;
    }
    bool ok2 = false;
    try {
      SpectrumProcessorFactory::Create(drv_gpu_lib::BackendType::OPENCL, &backend);
    } catch (const std::runtime_error&) {
      ok2 = true;
    }
    check(ok && ok2, "factory: null backend -> invalid_argument, OpenCL -> runtime_error");
  }

  // ---- FFT processor (reference test_fft_processor_rocm: bin 102 for 100 Hz, fs 1000, N 1024)
  {
    fft_processor::FFTProcessorROCm fft(&backend);
    fft_processor::FFTProcessorParams p;
    p.beam_count = 1; p.n_point = 1024; p.sample_rate = 1000.0f;
    cvec x(1024);
    for (int t = 0; t < 1024; ++t) x[t] = std::polar(1.0f, 2.0f * (float)M_PI * 100.0f * t / 1000.0f);
    auto r = fft.ProcessComplex(x, p);
    int best = 0;
    for (int k = 1; k < 1024; ++k)
      if (std::abs(r[0].spectrum[k]) > std::abs(r[0].spectrum[best])) best = k;
    check(r.size() == 1 && r[0].nFFT == 1024 && best == 102, "ProcessComplex: 100 Hz tone at bin 102");
    auto ref = spectrum_utils::CpuFFT::ForwardC2C(x);
    double e = 0, m = 0;
    for (int k = 0; k < 1024; ++k) {
      e = std::max(e, (double)std::abs(r[0].spectrum[k] - ref[k]));
      m = std::max(m, (double)std::abs(ref[k]));
    }
    check(e / m < 1e-5, "ProcessComplex vs CpuFFT within 1e-5 of the peak");
    auto mp = fft.ProcessMagPhase(x, p);
    bool ok = mp.size() == 1 && mp[0].magnitude.size() == 1024;
    for (int k = 0; k < 1024 && ok; ++k) ok = std::fabs(mp[0].magnitude[k] - std::abs(ref[k])) <= 1e-4f * (float)m;
    check(ok, "ProcessMagPhase magnitudes vs CpuFFT");
    // window + |X|^2 into a caller buffer
    void *din = nullptr, *dout = nullptr;
    bool okw = hipMalloc(&din, 1024 * 8) == hipSuccess && hipMalloc(&dout, 1024 * 4) == hipSuccess &&
               hipMemcpy(din, x.data(), 1024 * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (okw) {
      fft.ProcessMagnitudesToGPU(din, dout, p, true, fft_processor::WindowType::Hann);
      std::vector<float> got(1024);
      okw = hipMemcpy(got.data(), dout, 4096, hipMemcpyDeviceToHost) == hipSuccess;
      cvec xw(1024);
      for (int t = 0; t < 1024; ++t) xw[t] = x[t] * (float)(0.5 - 0.5 * std::cos(2.0 * M_PI * t / 1023.0));
      auto rw = spectrum_utils::CpuFFT::ForwardC2C(xw);
      double mm = 0;
      for (auto& v : rw) mm = std::max(mm, (double)std::norm(v));
      for (int k = 0; k < 1024 && okw; ++k) okw = std::fabs(got[k] - std::norm(rw[k])) <= 1e-4 * mm;
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    check(okw, "ProcessMagnitudesToGPU: Hann, |X|^2");
  }

  // ---- ComplexToMagPhaseROCm: |3+4j| = 5, norm_coeff rules
  {
    fft_processor::ComplexToMagPhaseROCm mp(&backend);
    fft_processor::MagPhaseParams p;
    p.beam_count = 1; p.n_point = 4; p.norm_coeff = 1.0f;
    cvec x = {{3, 4}, {1, 0}, {0, 1}, {-2, 0}};
    auto r = mp.Process(x, p);
    bool ok = r.size() == 1 && std::fabs(r[0].magnitude[0] - 5.0f) < 1e-6f &&
              std::fabs(r[0].phase[1]) < 1e-7f && std::fabs(r[0].phase[2] - (float)M_PI / 2) < 1e-6f;
    check(ok, "ComplexToMagPhase: |3+4j| = 5, phases");
    p.norm_coeff = -1.0f;
    void* d = nullptr;
    bool ok2 = hipMalloc(&d, 32) == hipSuccess && hipMemcpy(d, x.data(), 32, hipMemcpyHostToDevice) == hipSuccess;
    if (ok2) {
      auto m = mp.ProcessMagnitude(d, p, 32);
      ok2 = m.size() == 1 && std::fabs(m[0].magnitude[0] - 1.25f) < 1e-6f;
      void* o = nullptr;
      ok2 = ok2 && hipMalloc(&o, 16) == hipSuccess;
      if (ok2) {
        mp.ProcessMagnitudeToBuffer(d, o, p);
        float h[4];
        ok2 = hipMemcpy(h, o, 16, hipMemcpyDeviceToHost) == hipSuccess && std::fabs(h[3] - 0.5f) < 1e-6f;
        (void)hipFree(o);
      }
    }
    if (d) (void)hipFree(d);
    check(ok2, "ComplexToMagPhase: norm_coeff -1 -> 1/n, ToBuffer");
  }
  // ---- the backend as reference callers use it: device memory through the
  // backend (tests/test_fft_processor_rocm.hpp:211-238,
  // tests/test_complex_to_mag_phase_rocm.hpp:133-223), GetDeviceInfo / GetCore,
  // and the ROCmProfilingData of a profiled call
  {
    const drv_gpu_lib::DeviceInfo info = backend.GetDeviceInfo();
    check(info.name == backend.GetDeviceName() && info.warp_size == 64 && backend.GetCore().GetWarpSize() == 64 &&
              backend.GetGlobalMemorySize() > 0 && info.global_mem_bytes == backend.GetGlobalMemorySize(),
          "backend: GetDeviceInfo().name, GetCore().GetWarpSize() == 64, GetGlobalMemorySize");

    // FFT from a device buffer the backend allocated and filled
    fft_processor::FFTProcessorROCm fft(&backend);
    cvec tone(1024);
    for (int t = 0; t < 1024; ++t) tone[t] = std::polar(1.0f, 2.0f * (float)M_PI * 100.0f * t / 1000.0f);
    const size_t bytes = tone.size() * sizeof(std::complex<float>);
    void* gpu_data = backend.Allocate(bytes);
    backend.MemcpyHostToDevice(gpu_data, tone.data(), bytes);
    fft_processor::FFTProcessorParams p;
    p.beam_count = 1; p.n_point = 1024; p.sample_rate = 1000.0f;
    p.output_mode = fft_processor::FFTOutputMode::COMPLEX;
    auto r = fft.ProcessComplex(gpu_data, p, bytes);
    backend.Free(gpu_data);
    int best = 0;
    for (int k = 1; k < 512; ++k)
      if (std::abs(r[0].spectrum[k]) > std::abs(r[0].spectrum[best])) best = k;
    check(r.size() == 1 && best == (int)(100.0f * r[0].nFFT / 1000.0f),
          "backend.Allocate + MemcpyHostToDevice -> ProcessComplex(void*) -> backend.Free: bin 102");

    // profiling events carry the reference's timestamps
    fft_processor::ROCmProfEvents ev;
    (void)fft.ProcessComplex(tone, p, &ev);
    bool okp = ev.size() == 3;
    for (auto& e : ev)
      okp = okp && e.second.start_ns == 0 && e.second.complete_ns == e.second.end_ns && !e.second.op_string.empty() &&
            e.second.op_string == e.first;
    okp = okp && ev[1].second.kind == 1 && ev[1].second.end_ns > 0;
    check(okp, "ProcessComplex(prof_events): Upload/FFT/Download with end_ns, op_string, kind");

    // ComplexToMagPhaseROCm: GPU input -> CPU output, CPU -> GPU, GPU -> GPU
    fft_processor::ComplexToMagPhaseROCm conv(&backend);
    const uint32_t beams = 4, n = 2048;
    cvec data((size_t)beams * n);
    for (uint32_t b = 0; b < beams; ++b)
      for (uint32_t t = 0; t < n; ++t)
        data[(size_t)b * n + t] = std::polar(1.0f + b, 2.0f * (float)M_PI * 150.0f * t / 1000.0f + 0.3f * b);
    const size_t sz = data.size() * sizeof(std::complex<float>);
    void* gpu_in = backend.Allocate(sz);
    backend.MemcpyHostToDevice(gpu_in, data.data(), sz);
    fft_processor::MagPhaseParams mpp;
    mpp.beam_count = beams; mpp.n_point = n;
    auto res = conv.Process(gpu_in, mpp, sz);
    float worst = 0.f;
    for (uint32_t b = 0; b < beams && res.size() == beams; ++b)
      for (uint32_t t = 0; t < n; ++t) {
        const auto z = data[(size_t)b * n + t];
        worst = std::max(worst, std::fabs(res[b].magnitude[t] - std::abs(z)));
        float pe = std::fabs(res[b].phase[t] - std::arg(z));
        if (pe > (float)M_PI) pe = 2.0f * (float)M_PI - pe;
        worst = std::max(worst, pe);
      }
    check(res.size() == beams && worst < 1e-3f, "ComplexToMagPhase: backend buffer in -> host results (|z|, arg z)");
    void* gpu_out = conv.ProcessToGPU(gpu_in, mpp, sz);
    backend.Free(gpu_in);
    std::vector<float> raw((size_t)beams * n * 2);
    backend.MemcpyDeviceToHost(raw.data(), gpu_out, raw.size() * sizeof(float));
    backend.Free(gpu_out);
    worst = 0.f;
    for (size_t i = 0; i < data.size(); ++i) worst = std::max(worst, std::fabs(std::abs(data[i]) - raw[2 * i]));
    check(worst < 1e-3f, "ComplexToMagPhase: ProcessToGPU(void*) -> backend.MemcpyDeviceToHost -> backend.Free");
    void* gpu_out2 = conv.ProcessToGPU(data, mpp);
    backend.MemcpyDeviceToHost(raw.data(), gpu_out2, raw.size() * sizeof(float));
    backend.Free(gpu_out2);
    worst = 0.f;
    for (size_t i = 0; i < data.size(); ++i) {
      float pe = std::fabs(std::arg(data[i]) - raw[2 * i + 1]);
      if (pe > (float)M_PI) pe = 2.0f * (float)M_PI - pe;
      worst = std::max(worst, pe);
    }
    check(worst < 1e-3f, "ComplexToMagPhase: ProcessToGPU(host vector) phases");
  }
  std::printf("%s (%d failed)\n", g_fail ? "FAILED" : "ALL PASSED", g_fail);
  return g_fail ? 1 : 0;
}
