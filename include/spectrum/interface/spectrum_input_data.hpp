#pragma once
#include <core/spx_compat.hpp>
namespace antenna_fft {
using drv_gpu_lib::BackendType;
using drv_gpu_lib::OutputDestination;
using DriverType = drv_gpu_lib::BackendType;
}  // namespace antenna_fft
