// Spectrum-maxima processor (drop-in for the reference SpectrumProcessorROCm,
// include/spectrum/processors/spectrum_processor_rocm.hpp:96-273):
//  * ONE_PEAK / TWO_PEAKS: zero-padded FFT on the gfx950 kernels, then the
//    reference argmax + parabolic rules in one kernel (spx_peaks_top);
//  * ALL_MAXIMA from a signal: one fused window->FFT->|X|->local-max->compact
//    launch (spx_fused_peaks, AllMaxima rules); from a spectrum: spx_all_maxima.
// Divergences from the reference, all deliberate: Initialize re-plans whenever a
// parameter changes; max_maxima_per_beam is honoured on every AllMaxima path;
// magnitudes are computed for every beam; argmax ties resolve to the lowest bin.
#pragma once
#include <complex>
#include <core/spx_compat.hpp>
#include <cstdint>
#include <memory>
#include <spectrum/interface/i_spectrum_processor.hpp>
#include <utility>
#include <vector>

namespace antenna_fft {

using ROCmProfEvents = std::vector<std::pair<const char*, drv_gpu_lib::ROCmProfilingData>>;

class SpectrumProcessorROCm : public ISpectrumProcessor {
 public:
  explicit SpectrumProcessorROCm(drv_gpu_lib::IBackend* backend);
  ~SpectrumProcessorROCm() override;
  SpectrumProcessorROCm(const SpectrumProcessorROCm&) = delete;
  SpectrumProcessorROCm& operator=(const SpectrumProcessorROCm&) = delete;

  void Initialize(const SpectrumParams& params) override;
  bool IsInitialized() const override { return initialized_; }
  std::vector<SpectrumResult> ProcessFromCPU(const std::vector<std::complex<float>>& data) override {
    return ProcessFromCPU(data, nullptr);
  }
  std::vector<SpectrumResult> ProcessFromCPU(const std::vector<std::complex<float>>& data,
                                             ROCmProfEvents* prof_events);
  std::vector<SpectrumResult> ProcessFromGPU(void* gpu_data, size_t antenna_count, size_t n_point,
                                             size_t gpu_memory_bytes = 0) override;
  std::vector<SpectrumResult> ProcessBatch(const std::vector<std::complex<float>>& batch_data, size_t start_antenna,
                                           size_t batch_antenna_count) override {
    return ProcessBatch(batch_data, start_antenna, batch_antenna_count, nullptr);
  }
  std::vector<SpectrumResult> ProcessBatch(const std::vector<std::complex<float>>& batch_data, size_t start_antenna,
                                           size_t batch_antenna_count, ROCmProfEvents* prof_events);
  std::vector<SpectrumResult> ProcessBatchFromGPU(void* gpu_data, size_t src_offset_bytes, size_t start_antenna,
                                                  size_t batch_antenna_count) override;
  AllMaximaResult FindAllMaximaFromCPU(const std::vector<std::complex<float>>& data, OutputDestination dest,
                                       uint32_t search_start, uint32_t search_end) override {
    return FindAllMaximaFromCPU(data, dest, search_start, search_end, nullptr);
  }
  AllMaximaResult FindAllMaximaFromCPU(const std::vector<std::complex<float>>& data, OutputDestination dest,
                                       uint32_t search_start, uint32_t search_end, ROCmProfEvents* prof_events);
  AllMaximaResult FindAllMaximaFromGPUPipeline(void* gpu_data, size_t antenna_count, size_t n_point,
                                               size_t gpu_memory_bytes, OutputDestination dest,
                                               uint32_t search_start, uint32_t search_end) override;
  AllMaximaResult AllMaximaFromCPU(const std::vector<std::complex<float>>& fft_data, uint32_t beam_count,
                                   uint32_t nFFT, float sample_rate, OutputDestination dest, uint32_t search_start,
                                   uint32_t search_end) override;
  AllMaximaResult FindAllMaxima(void* fft_data, uint32_t beam_count, uint32_t nFFT, float sample_rate,
                                OutputDestination dest = OutputDestination::CPU, uint32_t search_start = 0,
                                uint32_t search_end = 0) override;
  DriverType GetDriverType() const override { return DriverType::ROCm; }
  ProfilingData GetProfilingData() const override { return prof_; }
  void ReallocateBuffersForBatch(size_t batch_antenna_count) override;
  size_t CalculateBytesPerAntenna() const override;
  void CompilePostKernel() override {}  // kernels are compiled ahead of time
  const SpectrumParams& GetParams() const { return params_; }

 private:
  static uint32_t NextPowerOf2(uint32_t n);
  void RequireInit() const;
  void Reserve(void** buf, size_t* cap, size_t bytes);
  std::vector<SpectrumResult> PeaksFromDevice(const void* signal_dev, size_t beams, size_t start_antenna);
  AllMaximaResult AllMaximaFromSignalDevice(const void* signal_dev, size_t beams, OutputDestination dest,
                                            uint32_t s0, uint32_t s1);
  AllMaximaResult AllMaximaFromSpectrumDevice(const void* spec_dev, uint32_t beams, uint32_t nfft, float fs,
                                              OutputDestination dest, uint32_t s0, uint32_t s1);

  drv_gpu_lib::IBackend* backend_ = nullptr;
  void* ctx_ = nullptr;  // spx_ctx*
  SpectrumParams params_;
  bool initialized_ = false;
  void* in_ = nullptr;
  size_t in_cap_ = 0;
  void* spec_ = nullptr;
  size_t spec_cap_ = 0;
  void* rec_ = nullptr;
  size_t rec_cap_ = 0;
  ProfilingData prof_;
};

}  // namespace antenna_fft
