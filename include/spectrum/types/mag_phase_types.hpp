#pragma once
#include <cstdint>
#include <vector>
namespace fft_processor {
// ComplexToMagPhaseROCm parameters and results (reference mag_phase_types.hpp:25-52).
// norm_coeff: 0 -> x1, < 0 -> x 1/n_point, > 0 -> x norm_coeff (magnitude-only paths).
struct MagPhaseParams {
  uint32_t beam_count = 1;
  uint32_t n_point = 0;
  float memory_limit = 0.80f;
  float norm_coeff = 1.0f;
};
struct MagPhaseResult {
  uint32_t beam_id = 0;
  uint32_t n_point = 0;
  std::vector<float> magnitude;
  std::vector<float> phase;  // radians
};
struct MagnitudeResult {
  uint32_t beam_id = 0;
  uint32_t n_point = 0;
  std::vector<float> magnitude;
};
}  // namespace fft_processor
