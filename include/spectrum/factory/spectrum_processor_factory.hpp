// Factory Method (reference spectrum_processor_factory.hpp:59): ROCm, AUTO and
// OPENCLandROCm give the gfx950 SpectrumProcessorROCm; OPENCL throws
// std::runtime_error; a null backend throws std::invalid_argument.
#pragma once
#include <core/spx_compat.hpp>
#include <memory>
#include <spectrum/interface/i_spectrum_processor.hpp>
namespace antenna_fft {
class SpectrumProcessorFactory {
 public:
  static std::unique_ptr<ISpectrumProcessor> Create(drv_gpu_lib::BackendType backend_type,
                                                    drv_gpu_lib::IBackend* backend);
};
}  // namespace antenna_fft
