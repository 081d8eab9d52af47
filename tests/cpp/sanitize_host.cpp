// Host-code checks run under AddressSanitizer and UndefinedBehaviorSanitizer
// (built and run by tests/test_sanitize_host.py, CPU only): the CpuFFT utility
// of the drop-in API (include/spectrum/utils/cpu_fft.hpp) and every function
// of the C oracle (oracle/spx_oracle.c), on ordinary and edge-case inputs
// (length 1, non-powers of two, ragged inputs, empty search ranges, capacity 0
// and 1, ranges at the spectrum's ends).  Any out-of-bounds access or
// undefined operation aborts the program; wrong values make it exit non-zero.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "spectrum/utils/cpu_fft.hpp"

extern "C" {
typedef struct orc_max_value {
  uint32_t index;
  float real, imag, magnitude, phase, freq_offset, refined_frequency;
  uint32_t pad;
} orc_max_value;
int orc_fft(const double* in, double* out, int n);
int orc_dft_brute(const double* in, double* out, int n);
int orc_fft_batch(const float* in, double* out, int batch, int n_point, int nfft, const double* window);
void orc_window(int type, int n, double* w);
int orc_fused_peaks(const float* X, int nfft, double fs, int s0, int s1, int mode, double nsigma, int kcap,
                    int refine, orc_max_value* recs, int* stored, int* found, orc_max_value* gmax, double* stats,
                    double* mags_out);
int orc_peaks_top(const float* X, uint32_t nfft, uint32_t search_range, float fs, int two_peaks, orc_max_value* out);
int orc_all_maxima(const float* X, uint32_t nfft, float fs, uint32_t s0, uint32_t s1, uint32_t cap,
                   orc_max_value* out, uint32_t* count);
}

static int failures = 0;
#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
      ++failures;                        \
    }                                    \
  } while (0)

int main() {
  std::mt19937_64 rng(0x5EED);
  std::normal_distribution<double> nd(0.0, 1.0);

  // CpuFFT against the brute-force DFT, any length; inverse(forward(x)) = x
  for (int n : {1, 2, 3, 5, 7, 12, 16, 100, 127, 1000, 1024}) {
    std::vector<std::complex<float>> x(n);
    std::vector<double> xd(2 * n), Xd(2 * n);
    for (int i = 0; i < n; ++i) {
      x[i] = {(float)nd(rng), (float)nd(rng)};
      xd[2 * i] = x[i].real();
      xd[2 * i + 1] = x[i].imag();
    }
    const auto X = spectrum_utils::CpuFFT::ForwardC2C(x);
    orc_dft_brute(xd.data(), Xd.data(), n);
    double e = 0, r = 0;
    for (int k = 0; k < n; ++k) {
      e += std::norm(std::complex<double>(X[k]) - std::complex<double>(Xd[2 * k], Xd[2 * k + 1]));
      r += Xd[2 * k] * Xd[2 * k] + Xd[2 * k + 1] * Xd[2 * k + 1];
    }
    CHECK(std::sqrt(e / r) < 1e-6, "CpuFFT n=%d rel err %g", n, std::sqrt(e / r));
    const auto y = spectrum_utils::CpuFFT::InverseC2C(X);
    double ei = 0;
    for (int i = 0; i < n; ++i) ei = std::max(ei, (double)std::abs(y[i] - x[i]));
    CHECK(ei < 1e-4, "CpuFFT inverse n=%d err %g", n, ei);
    std::vector<float> xr(n);
    for (int i = 0; i < n; ++i) xr[i] = x[i].real();
    const auto h = spectrum_utils::CpuFFT::ForwardR2C(xr);
    const auto f = spectrum_utils::CpuFFT::ForwardR2C_Full(xr);
    CHECK((int)h.size() == n / 2 + 1 && (int)f.size() == n, "R2C sizes n=%d", n);
    const auto m = spectrum_utils::CpuFFT::Magnitude(f, true);
    CHECK((int)m.size() == n, "Magnitude size n=%d", n);
  }

  // orc_fft against the brute-force DFT (powers of two; others rejected)
  for (int n : {1, 2, 4, 64, 4096}) {
    std::vector<double> x(2 * n), a(2 * n), b(2 * n);
    for (auto& v : x) v = nd(rng);
    CHECK(orc_fft(x.data(), a.data(), n) == 0, "orc_fft n=%d", n);
    orc_dft_brute(x.data(), b.data(), n);
    double e = 0, r = 0;
    for (int i = 0; i < 2 * n; ++i) {
      e += (a[i] - b[i]) * (a[i] - b[i]);
      r += b[i] * b[i];
    }
    CHECK(std::sqrt(e / r) < 1e-12, "orc_fft n=%d rel err %g", n, std::sqrt(e / r));
  }
  {
    std::vector<double> x(6), y(6);
    CHECK(orc_fft(x.data(), y.data(), 3) != 0, "orc_fft must reject n=3");
  }

  // windows (n = 1 included) and ragged batches (n_point < nfft)
  for (int type = 0; type <= 3; ++type)
    for (int n : {1, 2, 5, 64}) {
      std::vector<double> w(n);
      orc_window(type, n, w.data());
      for (double v : w) CHECK(std::isfinite(v) && v >= -1e-12 && v <= 1.0 + 1e-12, "window %d n=%d", type, n);
    }
  {
    const int B = 3, np = 37, nfft = 64;
    std::vector<float> in(2 * B * np);
    for (auto& v : in) v = (float)nd(rng);
    std::vector<double> w(np), out(2 * B * nfft);
    orc_window(1, np, w.data());
    CHECK(orc_fft_batch(in.data(), out.data(), B, np, nfft, w.data()) == 0, "orc_fft_batch");
    CHECK(orc_fft_batch(in.data(), out.data(), B, np, nfft, nullptr) == 0, "orc_fft_batch no window");
  }

  // peak rules on a spectrum with tones, at every range edge and capacity
  for (int nfft : {2, 8, 256, 4096}) {
    std::vector<float> X(2 * nfft);
    for (auto& v : X) v = (float)(0.1 * nd(rng));
    for (int t = 0; t < 3 && nfft > 8; ++t) X[2 * ((nfft / 7) * (t + 1))] += 50.0f;
    std::vector<double> mags(nfft);
    for (int mode : {1, 2})
      for (int kcap : {1, 4, 1000}) {
        const int ranges[][2] = {{0, nfft}, {0, 1}, {nfft - 1, nfft}, {nfft / 2, nfft / 2}, {1, nfft - 1}};
        for (const auto& rg : ranges) {
          if (rg[0] > rg[1]) continue;
          std::vector<orc_max_value> recs(kcap);
          orc_max_value g{};
          int stored = -1, found = -1;
          double stats[4];
          CHECK(orc_fused_peaks(X.data(), nfft, 1000.0, rg[0], rg[1], mode, 6.0, kcap, 1, recs.data(), &stored,
                                &found, &g, stats, mags.data()) == 0,
                "orc_fused_peaks");
          CHECK(stored >= 0 && stored <= kcap && stored <= found && found <= rg[1] - rg[0], "counts nfft=%d", nfft);
          for (int i = 1; i < stored; ++i) CHECK(recs[i].index > recs[i - 1].index, "ascending");
        }
      }
    for (uint32_t sr : {0u, 2u, (uint32_t)nfft / 2, (uint32_t)nfft})
      for (int two = 0; two < 2; ++two) {
        orc_max_value out[8];
        CHECK(orc_peaks_top(X.data(), nfft, sr, 1000.0f, two, out) == 0, "orc_peaks_top");
        for (int i = 0; i < (two ? 8 : 4); ++i) CHECK(out[i].index < (uint32_t)nfft, "top index");
      }
    if (nfft >= 3)
      for (uint32_t cap : {0u, 1u, 1000u}) {
        std::vector<orc_max_value> out(cap ? cap : 1);
        uint32_t count = 0;
        CHECK(orc_all_maxima(X.data(), nfft, 1000.0f, 1, nfft - 1, cap, out.data(), &count) == 0, "orc_all_maxima");
        CHECK(count <= (uint32_t)nfft, "all maxima count");
      }
  }

  // known answers: the reference's AllMaxima spikes {50, 120, 200} of 256
  {
    std::vector<float> X(2 * 256, 0.0f);
    for (int k : {50, 120, 200}) X[2 * k] = 10.0f;
    orc_max_value out[8];
    uint32_t count = 0;
    orc_all_maxima(X.data(), 256, 1000.0f, 1, 255, 8, out, &count);
    CHECK(count == 3 && out[0].index == 50 && out[1].index == 120 && out[2].index == 200, "spikes");
  }

  std::printf("%s (%d failures)\n", failures ? "FAILED" : "sanitized host checks passed", failures);
  return failures ? 1 : 0;
}
