// C entry points over the reference's own CPU FFT (spectrum_utils::CpuFFT,
// include/spectrum/utils/cpu_fft.hpp:68-89 -> pocketfft c2c<float>, single
// thread, plan rebuilt per call).  Compiled from the sources where they lie
// under the reference tree by oracle/Makefile; the output stays in oracle/_ref/.
// Test and baseline infrastructure only.
#include <complex>
#include <cstring>
#include <vector>

#include "cpu_fft.hpp"

extern "C" int ref_cpufft_c2c(const float* in, float* out, unsigned long n) {
  try {
    std::vector<std::complex<float>> v(reinterpret_cast<const std::complex<float>*>(in),
                                       reinterpret_cast<const std::complex<float>*>(in) + n);
    std::vector<std::complex<float>> r = spectrum_utils::CpuFFT::ForwardC2C(v);
    std::memcpy(out, r.data(), n * sizeof(std::complex<float>));
    return 0;
  } catch (...) {
    return 1;
  }
}

extern "C" int ref_cpufft_c2c_batch(const float* in, float* out, unsigned long batch, unsigned long n) {
  for (unsigned long b = 0; b < batch; ++b)
    if (ref_cpufft_c2c(in + 2 * b * n, out + 2 * b * n, n)) return 1;
  return 0;
}
