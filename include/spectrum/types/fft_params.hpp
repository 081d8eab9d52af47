#pragma once
#include <cstdint>
#include <spectrum/types/fft_modes.hpp>
namespace fft_processor {
// Same fields, order and defaults as the reference (fft_params.hpp:28-40).
// nFFT = nextPow2(n_point) * repeat_count (a power of two here).
struct FFTProcessorParams {
  uint32_t beam_count = 1;
  uint32_t n_point = 0;
  float sample_rate = 1000.0f;
  FFTOutputMode output_mode = FFTOutputMode::COMPLEX;
  uint32_t repeat_count = 1;
  float memory_limit = 0.80f;
};
}  // namespace fft_processor
