#pragma once
#include <spectrum/types/fft_modes.hpp>
#include <spectrum/types/fft_params.hpp>
#include <spectrum/types/fft_results.hpp>
