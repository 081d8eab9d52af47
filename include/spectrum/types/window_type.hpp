// Window applied before zero padding (symmetric, denominator n_point - 1).
// Same enumerators and values as the reference (include/spectrum/types/window_type.hpp:26-31).
#pragma once
#include <cstdint>
namespace fft_processor {
enum class WindowType : uint32_t { None = 0, Hann = 1, Hamming = 2, Blackman = 3 };
}
