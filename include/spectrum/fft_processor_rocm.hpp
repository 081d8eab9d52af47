// Batched zero-padded FFT facade (drop-in for the reference FFTProcessorROCm,
// include/spectrum/fft_processor_rocm.hpp:53-149) on the gfx950 kernels of
// libspectrum_mi355x.so: window + zero padding + Stockham FFT (+ |X|, phase)
// in one launch, no hipFFT.  Every public call is synchronous, as in the reference.
#pragma once
#include <complex>
#include <core/spx_compat.hpp>
#include <cstdint>
#include <spectrum/types/fft_types.hpp>
#include <spectrum/types/window_type.hpp>
#include <utility>
#include <vector>

namespace fft_processor {

using ROCmProfEvents = std::vector<std::pair<const char*, drv_gpu_lib::ROCmProfilingData>>;

class FFTProcessorROCm {
 public:
  explicit FFTProcessorROCm(drv_gpu_lib::IBackend* backend);
  ~FFTProcessorROCm();
  FFTProcessorROCm(const FFTProcessorROCm&) = delete;
  FFTProcessorROCm& operator=(const FFTProcessorROCm&) = delete;
  FFTProcessorROCm(FFTProcessorROCm&& other) noexcept;
  FFTProcessorROCm& operator=(FFTProcessorROCm&& other) noexcept;

  std::vector<FFTComplexResult> ProcessComplex(const std::vector<std::complex<float>>& data,
                                               const FFTProcessorParams& params,
                                               ROCmProfEvents* prof_events = nullptr);
  std::vector<FFTComplexResult> ProcessComplex(void* gpu_data, const FFTProcessorParams& params,
                                               size_t gpu_memory_bytes = 0);
  std::vector<FFTMagPhaseResult> ProcessMagPhase(const std::vector<std::complex<float>>& data,
                                                 const FFTProcessorParams& params,
                                                 ROCmProfEvents* prof_events = nullptr);
  std::vector<FFTMagPhaseResult> ProcessMagPhase(void* gpu_data, const FFTProcessorParams& params,
                                                 size_t gpu_memory_bytes = 0);
  // window + FFT + |X| (or |X|^2) into a caller-owned device buffer [beam][nFFT] float
  void ProcessMagnitudesToGPU(void* gpu_data, void* gpu_out_magnitudes, const FFTProcessorParams& params,
                              bool squared = false, WindowType window = WindowType::None,
                              ROCmProfEvents* prof_events = nullptr);
  FFTProfilingData GetProfilingData() const { return prof_; }
  uint32_t GetNFFT() const { return nFFT_; }

 private:
  static uint32_t NextPowerOf2(uint32_t n);
  void CalculateNFFT(const FFTProcessorParams& params);
  void Reserve(void** buf, size_t* cap, size_t bytes);
  void Release();

  drv_gpu_lib::IBackend* backend_ = nullptr;
  void* ctx_ = nullptr;  // spx_ctx*
  void* in_ = nullptr;
  size_t in_cap_ = 0;
  void* out_ = nullptr;
  size_t out_cap_ = 0;
  uint32_t nFFT_ = 0;
  uint32_t n_point_ = 0;
  FFTProfilingData prof_;
};

}  // namespace fft_processor
