// Deterministic synthetic tone + noise generator (host side, FP64).
//
// Counter-based: every (seed, channel, stream) triple gets its own SplitMix64
// sequence (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014), so any subset of channels can be regenerated
// independently and in parallel with identical results.  Uniforms use the top
// 53 bits; Gaussians use Box-Muller.  Signals are
//   x[n] = sum_i A_i exp(j(2 pi f_i n + phi_i)) + w[n],  E|w|^2 = sigma^2,
// evaluated in FP64 (a phasor recurrence re-anchored every 64 samples) and
// rounded once to complex64.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include <spectrum_mi355x.h>

namespace {

struct Rng {
  uint64_t s;
  static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  Rng(uint64_t seed, uint64_t chan, uint64_t stream)
      : s(mix(seed * 0x9E3779B97F4A7C15ull ^ mix(chan + 0x632BE59BD9B4E019ull) ^ mix(stream * 0xD1B54A32D192ED03ull + 1))) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    return mix(s);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  void gauss2(double& a, double& b) {
    const double u1 = uniform(), u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
    a = r * std::cos(2.0 * M_PI * u2);
    b = r * std::sin(2.0 * M_PI * u2);
  }
};

struct Tones {
  std::vector<double> f, a, ph;
  double sigma2 = 0.0;
};

double circ_dist(double f1, double f2) {
  double d = std::fabs(f1 - f2);
  d = d - std::floor(d);
  return d > 0.5 ? 1.0 - d : d;
}

Tones draw(uint32_t kind, Rng& r, uint32_t n) {
  Tones t;
  if (kind == SPX_GEN_THREE_TONES) {
    double amin = 1e300;
    for (int i = 0; i < 3; ++i) {
      t.f.push_back(r.uniform() - 0.5);
      const double a = 0.5 + 1.5 * r.uniform();
      t.a.push_back(a);
      t.ph.push_back(2.0 * M_PI * r.uniform());
      amin = std::fmin(amin, a);
    }
    t.sigma2 = amin * amin / 10.0;  // 10 dB below the weakest tone
  } else if (kind == SPX_GEN_MULTI_TONE) {
    const int nt = 1 + (int)(r.uniform() * 8.0);
    for (int i = 0; i < nt; ++i) {
      t.f.push_back(r.uniform() - 0.5);
      const double snr_db = 20.0 * r.uniform();
      t.a.push_back(std::pow(10.0, snr_db / 20.0));  // unit noise power
      t.ph.push_back(2.0 * M_PI * r.uniform());
    }
    t.sigma2 = 1.0;
  } else if (kind == SPX_GEN_DYNAMIC) {
    const double f1 = r.uniform() - 0.5;
    double f2 = r.uniform() - 0.5;
    for (int tries = 0; tries < 1000 && circ_dist(f1, f2) * n < 64.0; ++tries) f2 = r.uniform() - 0.5;
    t.f = {f1, f2};
    t.a = {1.0, 1e-3};
    t.ph = {2.0 * M_PI * r.uniform(), 2.0 * M_PI * r.uniform()};
    t.sigma2 = 1e-8;
  } else {  // SPX_GEN_FRACTIONAL
    const int nt = 1 + (int)(r.uniform() * 16.0);
    std::vector<uint32_t> ks;
    double amin = 1e300;
    for (int i = 0; i < nt; ++i) {
      uint32_t k = 0;
      bool ok = false;
      for (int tries = 0; tries < 1000 && !ok; ++tries) {
        k = (uint32_t)(r.uniform() * n);
        ok = true;
        for (uint32_t q : ks) {
          uint32_t d = k > q ? k - q : q - k;
          if (d > n - d) d = n - d;
          if (d < 3) ok = false;
        }
      }
      if (!ok) break;
      ks.push_back(k);
      t.f.push_back(((double)k + r.uniform()) / (double)n);
      const double a = 0.5 + 1.5 * r.uniform();
      t.a.push_back(a);
      t.ph.push_back(2.0 * M_PI * r.uniform());
      amin = std::fmin(amin, a);
    }
    t.sigma2 = amin * amin / 100.0;  // 20 dB below the weakest tone
  }
  return t;
}

void gen_channel(uint32_t kind, uint64_t seed, uint64_t chan, uint32_t n, float* out, uint32_t max_tones,
                 double* tf, double* ta) {
  Rng rt(seed, chan, 0), rn(seed, chan, 1);
  Tones t = draw(kind, rt, n);
  std::vector<double> re(n, 0.0), im(n, 0.0);
  for (size_t i = 0; i < t.f.size(); ++i) {
    const double w = 2.0 * M_PI * t.f[i];
    const double sr = std::cos(w), si = std::sin(w);
    double zr = 0, zi = 0;
    for (uint32_t k = 0; k < n; ++k) {
      if ((k & 63) == 0) {
        double fr = t.f[i] * (double)k;
        fr -= std::floor(fr);
        const double ang = 2.0 * M_PI * fr + t.ph[i];
        zr = t.a[i] * std::cos(ang);
        zi = t.a[i] * std::sin(ang);
      }
      re[k] += zr;
      im[k] += zi;
      const double nr = zr * sr - zi * si;
      zi = zr * si + zi * sr;
      zr = nr;
    }
  }
  const double s = std::sqrt(t.sigma2 / 2.0);
  for (uint32_t k = 0; k < n; k += 2) {
    double g0, g1, g2, g3;
    rn.gauss2(g0, g1);
    rn.gauss2(g2, g3);
    re[k] += s * g0;
    im[k] += s * g1;
    if (k + 1 < n) {
      re[k + 1] += s * g2;
      im[k + 1] += s * g3;
    }
  }
  for (uint32_t k = 0; k < n; ++k) {
    out[2 * (size_t)k] = (float)re[k];
    out[2 * (size_t)k + 1] = (float)im[k];
  }
  if (tf || ta) {
    for (uint32_t i = 0; i < max_tones; ++i) {
      const bool has = i < t.f.size();
      if (tf) tf[i] = has ? t.f[i] : 0.0;
      if (ta) ta[i] = has ? t.a[i] : 0.0;
    }
  }
}

}  // namespace

extern "C" int spx_siggen(uint32_t kind, uint64_t seed, uint32_t n_chan, uint32_t n_point, uint32_t chan_offset,
                          void* out_c64, uint32_t max_tones, double* tone_freq, double* tone_amp, int threads) {
  if (kind < SPX_GEN_THREE_TONES || kind > SPX_GEN_FRACTIONAL) return SPX_EINVAL;
  if (!out_c64 || n_point == 0) return SPX_EINVAL;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  if ((uint32_t)threads > n_chan) threads = (int)(n_chan ? n_chan : 1);
  float* out = (float*)out_c64;
  auto work = [&](int t) {
    for (uint32_t c = (uint32_t)t; c < n_chan; c += (uint32_t)threads) {
      gen_channel(kind, seed, (uint64_t)chan_offset + c, n_point, out + 2 * (size_t)c * n_point, max_tones,
                  tone_freq ? tone_freq + (size_t)c * max_tones : nullptr,
                  tone_amp ? tone_amp + (size_t)c * max_tones : nullptr);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  return SPX_OK;
}

extern "C" int spx_siggen_lengths(uint64_t seed, uint32_t n_chan, uint32_t lo_log2, uint32_t hi_log2,
                                  uint32_t* lengths_out) {
  if (!lengths_out || lo_log2 > hi_log2 || hi_log2 > 30) return SPX_EINVAL;
  const uint32_t span = hi_log2 - lo_log2 + 1;
  for (uint32_t c = 0; c < n_chan; ++c) {
    Rng r(seed, c, 2);
    uint32_t k = (uint32_t)(r.uniform() * (double)span);
    if (k >= span) k = span - 1;
    lengths_out[c] = 1u << (lo_log2 + k);
  }
  return SPX_OK;
}

extern "C" int spx_siggen_uneven(uint32_t kind, uint64_t seed, uint32_t n_chan, const uint32_t* lengths,
                                 const uint64_t* offsets, void* out_c64, int threads) {
  if (kind < SPX_GEN_THREE_TONES || kind > SPX_GEN_FRACTIONAL) return SPX_EINVAL;
  if (!out_c64 || !lengths || !offsets) return SPX_EINVAL;
  for (uint32_t c = 0; c < n_chan; ++c)
    if (lengths[c] == 0) return SPX_EINVAL;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  if ((uint32_t)threads > n_chan) threads = (int)(n_chan ? n_chan : 1);
  float* out = (float*)out_c64;
  auto work = [&](int t) {
    for (uint32_t c = (uint32_t)t; c < n_chan; c += (uint32_t)threads)
      gen_channel(kind, seed, c, lengths[c], out + 2 * offsets[c], 0, nullptr, nullptr);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  return SPX_OK;
}
