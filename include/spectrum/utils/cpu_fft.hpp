#pragma once
// spectrum_utils::CpuFFT -- the host-side FFT utility of the reference API
// (include/spectrum/utils/cpu_fft.hpp:57-204 there, exposed to Python as
// cpu_fft_c2c / cpu_ifft_c2c / cpu_fft_r2c / cpu_fft_r2c_full / magnitude,
// python/py_cpu_fft.hpp:28-106).  It is a named CPU utility of that API, not a
// fallback of any GPU entry point: nothing in the GPU library calls it.
//
// Own implementation: any length N; arithmetic in double (power-of-two lengths
// by an iterative radix-2 transform with directly evaluated twiddles, other
// lengths by Bluestein's chirp-z identity over a power-of-two convolution),
// rounded to complex<float> at the end.  Conventions as the reference states
// them: forward unnormalised, inverse scaled by 1/N, R2C returns N/2+1 bins,
// R2C_Full mirrors X[k] = conj(X[N-k]), Magnitude = sqrt(norm) or norm, empty
// input throws std::runtime_error.
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace spectrum_utils {

class CpuFFT {
 public:
  static std::vector<std::complex<float>> ForwardC2C(const std::vector<std::complex<float>>& input) {
    if (input.empty()) throw std::runtime_error("CpuFFT::ForwardC2C: empty input");
    std::vector<cd> a(input.begin(), input.end());
    transform(a, false);
    return to_float(a, 1.0);
  }

  static std::vector<std::complex<float>> InverseC2C(const std::vector<std::complex<float>>& input) {
    if (input.empty()) throw std::runtime_error("CpuFFT::InverseC2C: empty input");
    std::vector<cd> a(input.begin(), input.end());
    transform(a, true);
    return to_float(a, 1.0 / (double)a.size());
  }

  static std::vector<std::complex<float>> ForwardR2C(const std::vector<float>& input) {
    if (input.empty()) throw std::runtime_error("CpuFFT::ForwardR2C: empty input");
    std::vector<cd> a(input.size());
    for (size_t i = 0; i < input.size(); ++i) a[i] = cd(input[i], 0.0);
    transform(a, false);
    a.resize(input.size() / 2 + 1);
    return to_float(a, 1.0);
  }

  static std::vector<std::complex<float>> ForwardR2C_Full(const std::vector<float>& input) {
    auto half = ForwardR2C(input);
    const size_t n = input.size();
    std::vector<std::complex<float>> full(n);
    for (size_t i = 0; i < half.size() && i < n; ++i) full[i] = half[i];
    for (size_t i = half.size(); i < n; ++i) full[i] = std::conj(half[n - i]);
    return full;
  }

  static std::vector<float> Magnitude(const std::vector<std::complex<float>>& spectrum, bool squared = false) {
    std::vector<float> r(spectrum.size());
    for (size_t i = 0; i < spectrum.size(); ++i) {
      const float p = std::norm(spectrum[i]);
      r[i] = squared ? p : std::sqrt(p);
    }
    return r;
  }

 private:
  using cd = std::complex<double>;

  static std::vector<std::complex<float>> to_float(const std::vector<cd>& a, double s) {
    std::vector<std::complex<float>> r(a.size());
    for (size_t i = 0; i < a.size(); ++i) r[i] = std::complex<float>((float)(a[i].real() * s), (float)(a[i].imag() * s));
    return r;
  }

  // exp(sign * 2*pi*i * num/den) with num reduced mod den first (exact for any size)
  static cd unit(uint64_t num, uint64_t den, double sign) {
    num %= den;
    const double ang = sign * 2.0 * M_PI * (double)num / (double)den;
    return cd(std::cos(ang), std::sin(ang));
  }

  static void radix2(std::vector<cd>& a, bool inverse) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {  // bit reversal
      size_t bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(a[i], a[j]);
    }
    const double sign = inverse ? 1.0 : -1.0;
    for (size_t len = 2; len <= n; len <<= 1) {
      const size_t h = len >> 1;
      std::vector<cd> w(h);
      for (size_t k = 0; k < h; ++k) w[k] = unit(k, len, sign);
      for (size_t i = 0; i < n; i += len)
        for (size_t k = 0; k < h; ++k) {
          const cd u = a[i + k], v = a[i + k + h] * w[k];
          a[i + k] = u + v;
          a[i + k + h] = u - v;
        }
    }
  }

  static void transform(std::vector<cd>& a, bool inverse) {
    const size_t n = a.size();
    if ((n & (n - 1)) == 0) {
      radix2(a, inverse);
      return;
    }
    // Bluestein: jk = (j^2 + k^2 - (k-j)^2)/2 gives X[k] = c*[k] sum_j (x[j] c*[j]) c[k-j]
    // with c[j] = exp(+i pi j^2 / n) = unit(j^2, 2n) forward (conjugate sign inverse)
    size_t m = 1;
    while (m < 2 * n - 1) m <<= 1;
    const double sign = inverse ? -1.0 : 1.0;
    std::vector<cd> c(n);
    for (size_t j = 0; j < n; ++j) c[j] = unit((uint64_t)j * j, 2 * (uint64_t)n, sign);
    std::vector<cd> u(m, cd(0, 0)), v(m, cd(0, 0));
    for (size_t j = 0; j < n; ++j) u[j] = a[j] * std::conj(c[j]);
    v[0] = c[0];
    for (size_t j = 1; j < n; ++j) v[j] = v[m - j] = c[j];
    radix2(u, false);
    radix2(v, false);
    for (size_t i = 0; i < m; ++i) u[i] *= v[i];
    radix2(u, true);
    const double s = 1.0 / (double)m;
    for (size_t k = 0; k < n; ++k) a[k] = u[k] * s * std::conj(c[k]);
  }
};

}  // namespace spectrum_utils
