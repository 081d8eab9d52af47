#pragma once
#include <cstddef>
#include <cstdint>
#include <spectrum/types/spectrum_modes.hpp>
namespace antenna_fft {
// Same fields, order and defaults as the reference (spectrum_params.hpp:30-43).
struct SpectrumParams {
  uint32_t antenna_count = 5;
  uint32_t n_point = 1000;
  uint32_t repeat_count = 2;
  float sample_rate = 1000.0f;
  uint32_t search_range = 0;           // 0 -> nFFT / 4
  PeakSearchMode peak_mode = PeakSearchMode::ONE_PEAK;
  float memory_limit = 0.80f;
  size_t max_maxima_per_beam = 1000;
  uint32_t nFFT = 0;                   // computed: nextPow2(n_point) * repeat_count
  uint32_t base_fft = 0;               // computed: nextPow2(n_point)
};
}  // namespace antenna_fft
