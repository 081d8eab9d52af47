// Elementwise complex -> magnitude/phase facade (drop-in for the reference
// ComplexToMagPhaseROCm, include/spectrum/complex_to_mag_phase_rocm.hpp:83-270).
// *ToGPU methods return a hipMalloc pointer owned by the caller (hipFree).
#pragma once
#include <complex>
#include <core/spx_compat.hpp>
#include <cstdint>
#include <spectrum/types/mag_phase_types.hpp>
#include <vector>

namespace fft_processor {

class ComplexToMagPhaseROCm {
 public:
  explicit ComplexToMagPhaseROCm(drv_gpu_lib::IBackend* backend);
  ~ComplexToMagPhaseROCm();
  ComplexToMagPhaseROCm(const ComplexToMagPhaseROCm&) = delete;
  ComplexToMagPhaseROCm& operator=(const ComplexToMagPhaseROCm&) = delete;
  ComplexToMagPhaseROCm(ComplexToMagPhaseROCm&& other) noexcept;
  ComplexToMagPhaseROCm& operator=(ComplexToMagPhaseROCm&& other) noexcept;

  std::vector<MagPhaseResult> Process(const std::vector<std::complex<float>>& data, const MagPhaseParams& params);
  std::vector<MagPhaseResult> Process(void* gpu_data, const MagPhaseParams& params, size_t gpu_memory_bytes = 0);
  void* ProcessToGPU(const std::vector<std::complex<float>>& data, const MagPhaseParams& params);
  void* ProcessToGPU(void* gpu_data, const MagPhaseParams& params, size_t gpu_memory_bytes = 0);
  std::vector<MagnitudeResult> ProcessMagnitude(void* gpu_data, const MagPhaseParams& params,
                                                size_t gpu_memory_bytes = 0);
  void* ProcessMagnitudeToGPU(void* gpu_data, const MagPhaseParams& params, size_t gpu_memory_bytes = 0);
  void ProcessMagnitudeToBuffer(void* gpu_complex_in, void* gpu_magnitude_out, const MagPhaseParams& params);

 private:
  static float ComputeInvN(const MagPhaseParams& params);
  void Reserve(void** buf, size_t* cap, size_t bytes);
  void Release();

  drv_gpu_lib::IBackend* backend_ = nullptr;
  void* ctx_ = nullptr;  // spx_ctx*
  void* in_ = nullptr;
  size_t in_cap_ = 0;
  void* out_ = nullptr;
  size_t out_cap_ = 0;
};

}  // namespace fft_processor
