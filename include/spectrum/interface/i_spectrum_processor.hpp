// Spectrum-maxima processor interface: same pure virtuals and signatures as the
// reference (include/spectrum/interface/i_spectrum_processor.hpp:65-170).
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <spectrum/interface/spectrum_input_data.hpp>
#include <spectrum/interface/spectrum_maxima_types.h>
#include <vector>

namespace antenna_fft {

class ISpectrumProcessor {
 public:
  virtual ~ISpectrumProcessor() = default;
  virtual void Initialize(const SpectrumParams& params) = 0;
  virtual bool IsInitialized() const = 0;
  virtual std::vector<SpectrumResult> ProcessFromCPU(const std::vector<std::complex<float>>& data) = 0;
  virtual std::vector<SpectrumResult> ProcessFromGPU(void* gpu_data, size_t antenna_count, size_t n_point,
                                                     size_t gpu_memory_bytes = 0) = 0;
  virtual std::vector<SpectrumResult> ProcessBatch(const std::vector<std::complex<float>>& batch_data,
                                                   size_t start_antenna, size_t batch_antenna_count) = 0;
  virtual std::vector<SpectrumResult> ProcessBatchFromGPU(void* gpu_data, size_t src_offset_bytes,
                                                          size_t start_antenna, size_t batch_antenna_count) = 0;
  virtual AllMaximaResult FindAllMaximaFromCPU(const std::vector<std::complex<float>>& data, OutputDestination dest,
                                               uint32_t search_start, uint32_t search_end) = 0;
  virtual AllMaximaResult FindAllMaximaFromGPUPipeline(void* gpu_data, size_t antenna_count, size_t n_point,
                                                       size_t gpu_memory_bytes, OutputDestination dest,
                                                       uint32_t search_start, uint32_t search_end) = 0;
  virtual AllMaximaResult AllMaximaFromCPU(const std::vector<std::complex<float>>& fft_data, uint32_t beam_count,
                                           uint32_t nFFT, float sample_rate, OutputDestination dest,
                                           uint32_t search_start, uint32_t search_end) = 0;
  virtual AllMaximaResult FindAllMaxima(void* fft_data, uint32_t beam_count, uint32_t nFFT, float sample_rate,
                                        OutputDestination dest = OutputDestination::CPU, uint32_t search_start = 0,
                                        uint32_t search_end = 0) = 0;
  virtual DriverType GetDriverType() const = 0;
  virtual ProfilingData GetProfilingData() const = 0;
  virtual void ReallocateBuffersForBatch(size_t batch_antenna_count) = 0;
  virtual size_t CalculateBytesPerAntenna() const = 0;
  virtual void CompilePostKernel() = 0;
};

}  // namespace antenna_fft
