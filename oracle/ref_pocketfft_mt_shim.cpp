// Multi-threaded pocketfft c2c<float> over a [batch][n] array (axis 1), built
// from the reference tree's vendored header (third_party/pocketfft) without
// POCKETFFT_NO_MULTITHREADING: the best-effort multi-core CPU baseline.
// Test and baseline infrastructure only; output in oracle/_ref/.
#include <complex>

#include "pocketfft_hdronly.h"

extern "C" int ref_pocketfft_c2c_mt(const float* in, float* out, unsigned long batch, unsigned long n,
                                    unsigned long nthreads) {
  try {
    pocketfft::shape_t shape{batch, n};
    pocketfft::stride_t st{(ptrdiff_t)(n * sizeof(std::complex<float>)), (ptrdiff_t)sizeof(std::complex<float>)};
    pocketfft::shape_t axes{1};
    pocketfft::c2c<float>(shape, st, st, axes, true, reinterpret_cast<const std::complex<float>*>(in),
                          reinterpret_cast<std::complex<float>*>(out), 1.0f, nthreads);
    return 0;
  } catch (...) {
    return 1;
  }
}
