#pragma once
#include <spectrum/types/spectrum_modes.hpp>
#include <spectrum/types/spectrum_params.hpp>
#include <spectrum/types/spectrum_profiling.hpp>
#include <spectrum/types/spectrum_result_types.hpp>
