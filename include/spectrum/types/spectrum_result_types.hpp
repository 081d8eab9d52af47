#pragma once
#include <core/spx_compat.hpp>
#include <cstddef>
#include <cstdint>
#include <vector>
namespace antenna_fft {
using drv_gpu_lib::OutputDestination;

// 32-byte peak record; field order and size identical to the reference
// (spectrum_result_types.hpp:34-43) and to spx_max_value of the C ABI.
struct MaxValue {
  uint32_t index;
  float real, imag;
  float magnitude;
  float phase;              // degrees
  float freq_offset;
  float refined_frequency;  // (index + freq_offset) * fs / nFFT
  uint32_t pad;
};
static_assert(sizeof(MaxValue) == 32, "MaxValue must stay 32 bytes");

struct SpectrumResult {
  uint32_t antenna_id;
  MaxValue interpolated;
  MaxValue left_point;
  MaxValue center_point;
  MaxValue right_point;
};

struct CPUSpectrumResult {
  SpectrumResult SpectrMax_left, SpectrMax_right;
};

struct AllMaximaBeamResult {
  uint32_t antenna_id;
  uint32_t num_maxima;
  std::vector<MaxValue> maxima;
};

// gpu_maxima / gpu_counts are caller-owned (hipFree) for GPU and ALL destinations.
struct AllMaximaResult {
  std::vector<AllMaximaBeamResult> beams;
  OutputDestination destination = OutputDestination::CPU;
  void* gpu_maxima = nullptr;
  void* gpu_counts = nullptr;
  size_t total_maxima = 0;
  size_t gpu_bytes = 0;
  size_t TotalMaxima() const { return total_maxima; }
};
}  // namespace antenna_fft
