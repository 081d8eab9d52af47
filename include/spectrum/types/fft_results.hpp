#pragma once
#include <complex>
#include <cstdint>
#include <vector>
namespace fft_processor {
// Result records of FFTProcessorROCm (reference fft_results.hpp:27-83).
struct FFTBeamResult {
  uint32_t beam_id = 0;
  uint32_t nFFT = 0;
  float sample_rate = 0.0f;
};
struct FFTComplexResult : FFTBeamResult {
  std::vector<std::complex<float>> spectrum;  // unnormalised, nFFT values
};
struct FFTMagPhaseResult : FFTBeamResult {
  std::vector<float> magnitude;  // |X[k]|
  std::vector<float> phase;      // atan2(im, re), radians
  std::vector<float> frequency;  // k * fs / nFFT (MAGNITUDE_PHASE_FREQ only)
};
struct FFTProfilingData {
  double upload_time_ms = 0.0;
  double fft_time_ms = 0.0;
  double post_processing_time_ms = 0.0;
  double download_time_ms = 0.0;
  double total_time_ms = 0.0;
};
}  // namespace fft_processor
