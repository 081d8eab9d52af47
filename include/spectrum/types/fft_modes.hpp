#pragma once
namespace fft_processor {
// Output format of FFTProcessorROCm (reference fft_modes.hpp:24-30).
enum class FFTOutputMode { COMPLEX, MAGNITUDE_PHASE, MAGNITUDE_PHASE_FREQ };
}
