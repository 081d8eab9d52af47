// AllMaxima on a spectrum already on the device: strict local maxima of |X| in
// [search_start, search_end), ascending, capped per beam.  Replaces the
// detect -> scan -> compact chain of the reference
// (src/fft_func/src/all_maxima_pipeline_rocm.cpp:258-472) with one gfx950
// kernel launch (spx_all_maxima).
#pragma once
#include <core/spx_compat.hpp>
#include <spectrum/interface/i_all_maxima_pipeline.hpp>
namespace antenna_fft {
class AllMaximaPipelineROCm : public IAllMaximaPipeline {
 public:
  explicit AllMaximaPipelineROCm(drv_gpu_lib::IBackend* backend);
  ~AllMaximaPipelineROCm() override;
  AllMaximaPipelineROCm(const AllMaximaPipelineROCm&) = delete;
  AllMaximaPipelineROCm& operator=(const AllMaximaPipelineROCm&) = delete;
  // magnitudes_gpu is accepted for signature compatibility; magnitudes are
  // recomputed from fft_data_gpu inside the kernel.
  AllMaximaResult Execute(void* magnitudes_gpu, void* fft_data_gpu, uint32_t beam_count, uint32_t nFFT,
                          float sample_rate, OutputDestination dest = OutputDestination::CPU,
                          uint32_t search_start = 1, uint32_t search_end = 0,
                          size_t max_maxima_per_beam = 1000) override;

 private:
  drv_gpu_lib::IBackend* backend_;
  void* ctx_ = nullptr;  // spx_ctx*
};
}  // namespace antenna_fft
