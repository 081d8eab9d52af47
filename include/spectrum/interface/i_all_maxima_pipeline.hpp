// AllMaxima pipeline interface (reference i_all_maxima_pipeline.hpp:55-81).
#pragma once
#include <cstddef>
#include <cstdint>
#include <spectrum/interface/spectrum_maxima_types.h>
namespace antenna_fft {
class IAllMaximaPipeline {
 public:
  virtual ~IAllMaximaPipeline() = default;
  virtual AllMaximaResult Execute(void* magnitudes_gpu, void* fft_data_gpu, uint32_t beam_count, uint32_t nFFT,
                                  float sample_rate, OutputDestination dest = OutputDestination::CPU,
                                  uint32_t search_start = 1, uint32_t search_end = 0,
                                  size_t max_maxima_per_beam = 1000) = 0;
};
}  // namespace antenna_fft
