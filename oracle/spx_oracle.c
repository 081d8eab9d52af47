/*
 * CPU oracle for the FFT -> |X| -> peak-search path.  TEST INFRASTRUCTURE
 * ONLY: tests/, the smoke check and bench.py's cpu_baseline leg may load it,
 * as the checker; the library never does.
 *
 * Plain C restatement, FP64 unless stated:
 *   orc_fft          radix-2 decimation-in-time FFT, forward, unnormalised
 *                    (the transform of CpuFFT::ForwardC2C, include/spectrum/utils/cpu_fft.hpp:68-89,
 *                    computed in double instead of pocketfft's float)
 *   orc_dft_brute    O(N^2) long-double DFT, pins orc_fft on small N
 *   orc_window       symmetric Hann/Hamming/Blackman over n samples, denominator n-1
 *                    (include/spectrum/kernels/fft_processor_kernels_rocm.hpp:99-109)
 *   orc_fused_peaks  the fused threshold search (no reference kernel; rules as in
 *                    spx_fused.hip's header) and the AllMaxima rule set
 *                    (all_maxima_kernel_sources_rocm.hpp:57-264, all_maxima_pipeline_rocm.cpp:274-307)
 *   orc_peaks_top    ONE_PEAK / TWO_PEAKS (fft_kernel_sources_rocm.hpp:122-519), float arithmetic
 *   orc_all_maxima   AllMaxima on a spectrum (all_maxima_kernel_sources_rocm.hpp:57-264), float magnitudes
 * Every function takes the spectrum it judges as an argument, so the same
 * rules can be run on the oracle's own FP64 spectrum or on a GPU spectrum.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_max_value {
  uint32_t index;
  float real, imag, magnitude, phase, freq_offset, refined_frequency;
  uint32_t pad;
} orc_max_value;

/* ---- FFT ----------------------------------------------------------------- */
/* in, out: interleaved re/im doubles, length 2n; n a power of two; in may == out */
int orc_fft(const double* in, double* out, int n) {
  if (n < 1 || (n & (n - 1))) return 1;
  if (out != in) memcpy(out, in, sizeof(double) * 2 * (size_t)n);
  /* bit reversal */
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double tr = out[2 * i], ti = out[2 * i + 1];
      out[2 * i] = out[2 * j];
      out[2 * i + 1] = out[2 * j + 1];
      out[2 * j] = tr;
      out[2 * j + 1] = ti;
    }
  }
  /* twiddles W_n^k for k < n/2, long double once */
  double* w = (double*)malloc(sizeof(double) * (size_t)(n > 1 ? n : 2));
  if (!w) return 2;
  for (int k = 0; k < n / 2; ++k) {
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    w[2 * k] = (double)cosl(a);
    w[2 * k + 1] = (double)-sinl(a);
  }
  for (int len = 2; len <= n; len <<= 1) {
    int step = n / len;
    for (int s = 0; s < n; s += len) {
      for (int k = 0; k < len / 2; ++k) {
        double wr = w[2 * k * step], wi = w[2 * k * step + 1];
        double* a = out + 2 * (s + k);
        double* b = out + 2 * (s + k + len / 2);
        double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
        b[0] = a[0] - br;
        b[1] = a[1] - bi;
        a[0] += br;
        a[1] += bi;
      }
    }
  }
  free(w);
  return 0;
}

/* brute-force DFT in long double (any n) */
int orc_dft_brute(const double* in, double* out, int n) {
  const long double tp = 2.0L * 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    for (int m = 0; m < n; ++m) {
      long long e = ((long long)m * k) % n;
      long double a = tp * (long double)e / (long double)n;
      long double c = cosl(a), s = -sinl(a);
      sr += in[2 * m] * c - in[2 * m + 1] * s;
      si += in[2 * m] * s + in[2 * m + 1] * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
  return 0;
}

/* batched: in [batch][n_point] complex64 (float), zero padded to nfft, windowed
 * (window: double[n_point] or NULL), FFT in double, out double [batch][nfft][2] */
int orc_fft_batch(const float* in, double* out, int batch, int n_point, int nfft, const double* window) {
  for (int b = 0; b < batch; ++b) {
    double* o = out + 2 * (size_t)b * nfft;
    memset(o, 0, sizeof(double) * 2 * (size_t)nfft);
    for (int p = 0; p < n_point; ++p) {
      double wv = window ? window[p] : 1.0;
      o[2 * p] = (double)in[2 * ((size_t)b * n_point + p)] * wv;
      o[2 * p + 1] = (double)in[2 * ((size_t)b * n_point + p) + 1] * wv;
    }
    int rc = orc_fft(o, o, nfft);
    if (rc) return rc;
  }
  return 0;
}

/* ---- windows ---------------------------------------------------------------- */
void orc_window(int type, int n, double* w) {
  for (int p = 0; p < n; ++p) {
    double v = 1.0;
    if (n > 1) {
      double th = 2.0 * M_PI * (double)p / (double)(n - 1);
      if (type == 1) v = 0.5 * (1.0 - cos(th));
      else if (type == 2) v = 0.54 - 0.46 * cos(th);
      else if (type == 3) v = 0.42 - 0.5 * cos(th) + 0.08 * cos(2.0 * th);
    }
    w[p] = v;
  }
}

/* ---- fused peak rules --------------------------------------------------------- */
static double parabolic_d(double yl, double yc, double yr) {
  double den = yl - 2.0 * yc + yr;
  double d = 0.0;
  if (fabs(den) > 1e-10) {
    d = 0.5 * (yl - yr) / den;
    if (d < -0.5) d = -0.5;
    if (d > 0.5) d = 0.5;
  }
  return d;
}

static void fill(orc_max_value* r, uint32_t k, float re, float im, double m, double d, double fs, int nfft) {
  r->index = k;
  r->real = re;
  r->imag = im;
  r->magnitude = (float)m;
  r->phase = (float)(atan2((double)im, (double)re) * 57.29577951);
  r->freq_offset = (float)d;
  r->refined_frequency = (float)(((double)k + d) * fs / (double)nfft);
  r->pad = 0;
}

/*
 * X: float spectrum [nfft][2] of one channel.
 * mode 1: threshold rules; mode 2: AllMaxima rules (no threshold, no refinement,
 *         range normalised by the caller).
 * [s0, s1): search range; stats over that range; neighbours circular.
 * Outputs: recs[kcap], *stored, *found, gmax (mode 1), stats[4] = mean, std, T, mean m^2.
 * mags_out (optional, double[nfft]) returns the magnitudes used.
 */
int orc_fused_peaks(const float* X, int nfft, double fs, int s0, int s1, int mode, double nsigma, int kcap,
                    int refine, orc_max_value* recs, int* stored, int* found, orc_max_value* gmax, double* stats,
                    double* mags_out) {
  double* m = (double*)malloc(sizeof(double) * (size_t)nfft);
  if (!m) return 2;
  for (int k = 0; k < nfft; ++k) {
    double re = X[2 * k], im = X[2 * k + 1];
    m[k] = sqrt(re * re + im * im);
  }
  double s1sum = 0, s2sum = 0, bm = -1.0;
  int bk = -1;
  for (int k = s0; k < s1; ++k) {
    s1sum += m[k];
    s2sum += m[k] * m[k];
    if (m[k] > bm) { bm = m[k]; bk = k; }
  }
  double cnt = (double)(s1 - s0);
  double mu = cnt > 0 ? s1sum / cnt : 0.0;
  double var = cnt > 0 ? s2sum / cnt - mu * mu : 0.0;
  if (var < 0) var = 0;
  double sd = sqrt(var);
  double thr = (mode == 1) ? mu + nsigma * sd : -INFINITY;
  if (stats) {
    stats[0] = mu;
    stats[1] = sd;
    stats[2] = thr;
    stats[3] = cnt > 0 ? s2sum / cnt : 0.0;
  }
  int n = 0;
  for (int k = s0; k < s1; ++k) {
    double yl = m[(k + nfft - 1) % nfft], yr = m[(k + 1) % nfft];
    if (m[k] > yl && m[k] > yr && m[k] > thr) {
      if (n < kcap) {
        double d = (refine && mode == 1) ? parabolic_d(yl, m[k], yr) : 0.0;
        fill(&recs[n], (uint32_t)k, X[2 * k], X[2 * k + 1], m[k], d, fs, nfft);
      }
      ++n;
    }
  }
  *found = n;
  *stored = n < kcap ? n : kcap;
  if (gmax && bk >= 0) {
    double yl = m[(bk + nfft - 1) % nfft], yr = m[(bk + 1) % nfft];
    double d = (refine && mode == 1) ? parabolic_d(yl, m[bk], yr) : 0.0;
    fill(gmax, (uint32_t)bk, X[2 * bk], X[2 * bk + 1], m[bk], d, fs, nfft);
  }
  if (mags_out) memcpy(mags_out, m, sizeof(double) * (size_t)nfft);
  free(m);
  return 0;
}

/* ---- reference ONE_PEAK / TWO_PEAKS (float arithmetic as the reference) ------- */
static float fabsx(float x) { return x < 0 ? -x : x; }
static float cabsf_rn(const float* X, uint32_t i) { return sqrtf(X[2 * i] * X[2 * i] + X[2 * i + 1] * X[2 * i + 1]); }

static void peak4(const float* X, uint32_t nfft, uint32_t c, int mirror, float fs, orc_max_value* out) {
  float bw = fs / (float)nfft;
  float yc = cabsf_rn(X, c);
  float lr = 0, li = 0, rr = 0, ri = 0, yl = 0, yr = 0;
  int hl = c > 0, hr = c < nfft - 1;
  if (hl) { lr = X[2 * (c - 1)]; li = X[2 * (c - 1) + 1]; yl = cabsf_rn(X, c - 1); }
  if (hr) { rr = X[2 * (c + 1)]; ri = X[2 * (c + 1) + 1]; yr = cabsf_rn(X, c + 1); }
  float fo = 0.f, rf = (float)c * bw;
  if (hl && hr) {
    float den = yl - 2.0f * yc + yr;
    if (fabsx(den) > 1e-10f) {
      fo = 0.5f * (yl - yr) / den;
      if (fo < -0.5f) fo = -0.5f;
      if (fo > 0.5f) fo = 0.5f;
      rf = ((float)c + fo) * bw;
    }
  }
  if (mirror) rf = fs - rf;
  float cr = X[2 * c], ci = X[2 * c + 1];
  orc_max_value v;
  v.pad = 0;
  v.index = c; v.real = cr; v.imag = ci; v.magnitude = yc; v.phase = atan2f(ci, cr) * 57.29577951f;
  v.freq_offset = fo; v.refined_frequency = rf; out[0] = v;
  float rfl = hl ? (float)(c - 1) * bw : 0.f;
  if (mirror && hl) rfl = fs - rfl;
  v.index = hl ? c - 1 : 0; v.real = lr; v.imag = li; v.magnitude = yl;
  v.phase = hl ? atan2f(li, lr) * 57.29577951f : 0.f; v.freq_offset = 0.f; v.refined_frequency = rfl; out[1] = v;
  float rfc = (float)c * bw;
  if (mirror) rfc = fs - rfc;
  v.index = c; v.real = cr; v.imag = ci; v.magnitude = yc; v.phase = atan2f(ci, cr) * 57.29577951f;
  v.freq_offset = 0.f; v.refined_frequency = rfc; out[2] = v;
  float rfr = hr ? (float)(c + 1) * bw : 0.f;
  if (mirror && hr) rfr = fs - rfr;
  v.index = hr ? c + 1 : 0; v.real = rr; v.imag = ri; v.magnitude = yr;
  v.phase = hr ? atan2f(ri, rr) * 57.29577951f : 0.f; v.freq_offset = 0.f; v.refined_frequency = rfr; out[3] = v;
}

/* X: [nfft][2] float; out: 4 (one) or 8 (two) records.  Ties: lowest bin. */
int orc_peaks_top(const float* X, uint32_t nfft, uint32_t search_range, float fs, int two_peaks, orc_max_value* out) {
  uint32_t half = search_range / 2, r2 = nfft - half;
  float lm = -1.f, rm = -1.f;
  uint32_t li = 0, ri = r2 < nfft ? r2 : nfft - 1;
  for (uint32_t i = 0; i < half; ++i) {
    float m = cabsf_rn(X, i);
    if (m > lm) { lm = m; li = i; }
  }
  for (uint32_t i = r2; i < nfft; ++i) {
    float m = cabsf_rn(X, i);
    if (m > rm) { rm = m; ri = i; }
  }
  if (!two_peaks) {
    int right = rm > lm;
    peak4(X, nfft, right ? ri : li, right, fs, out);
  } else {
    peak4(X, nfft, li, 0, fs, out);
    peak4(X, nfft, ri, 1, fs, out + 4);
  }
  return 0;
}

/* AllMaxima on a spectrum (range already normalised, 1 <= s0 <= s1 <= nfft-1) */
int orc_all_maxima(const float* X, uint32_t nfft, float fs, uint32_t s0, uint32_t s1, uint32_t cap,
                   orc_max_value* out, uint32_t* count) {
  float bw = fs / (float)nfft;
  uint32_t n = 0;
  for (uint32_t i = s0; i < s1; ++i) {
    float mc = cabsf_rn(X, i), ml = cabsf_rn(X, i - 1), mr = cabsf_rn(X, i + 1);
    if (mc > ml && mc > mr) {
      if (n < cap) {
        orc_max_value v;
        v.index = i; v.real = X[2 * i]; v.imag = X[2 * i + 1]; v.magnitude = mc;
        v.phase = atan2f(X[2 * i + 1], X[2 * i]) * 57.29577951f; v.freq_offset = 0.f;
        v.refined_frequency = (float)i * bw; v.pad = 0;
        out[n] = v;
      }
      ++n;
    }
  }
  *count = n;
  return 0;
}
