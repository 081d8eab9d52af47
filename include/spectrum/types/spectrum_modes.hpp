#pragma once
namespace antenna_fft {
// Peak search modes (reference spectrum_modes.hpp:21-25).
enum class PeakSearchMode { ONE_PEAK, TWO_PEAKS, ALL_MAXIMA };
}
