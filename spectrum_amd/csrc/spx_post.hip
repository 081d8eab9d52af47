// Kernels that work on a spectrum already in HBM:
//  * elementwise |z| / |z|^2 / {|z|, arg z}      (ComplexToMagPhaseROCm, MagPhaseOp, MagnitudeOp)
//  * ONE_PEAK / TWO_PEAKS argmax with parabolic refinement (SpectrumPostOp)
//  * AllMaxima: strict local maxima in [s0,s1), ascending compaction with a cap
//    (detect + Blelloch scan + compact of the reference, done in one pass per beam)
#include "spx_fused_common.hpp"

namespace spx {

// ---------------------------------------------------------------------------
// elementwise: reference kernels complex_to_mag_phase / complex_to_magnitude(_squared)
// include/spectrum/kernels/fft_processor_kernels_rocm.hpp:123-174
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mag_phase(const float2* __restrict__ in, float2* __restrict__ out,
                                                   uint64_t total) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const float2 z = in[i];
    out[i] = make_float2(sqrtf(z.x * z.x + z.y * z.y), atan2f(z.y, z.x));
  }
}

__global__ void __launch_bounds__(256) k_magnitude(const float2* __restrict__ in, float* __restrict__ out,
                                                   uint64_t total, float inv_n, int squared) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const float2 z = in[i];
    const float p = z.x * z.x + z.y * z.y;
    out[i] = squared ? p * inv_n : sqrtf(p) * inv_n;
  }
}

// ---------------------------------------------------------------------------
// ONE_PEAK / TWO_PEAKS: include/spectrum/kernels/fft_kernel_sources_rocm.hpp:122-519
// Per beam: argmax of |X| over [0, sr/2) and [N - sr/2, N).  Ties resolve to the
// lowest bin (the reference resolves to the lowest thread slot).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_peaks_top(const float2* __restrict__ spec, MaxValue* __restrict__ out,
                                                   uint32_t batch, uint32_t nfft, uint32_t search_range, float fs,
                                                   int two_peaks) {
  const uint32_t beam = blockIdx.x;
  if (beam >= batch) return;
  const float2* X = spec + (uint64_t)beam * nfft;
  const uint32_t half = search_range / 2;
  const uint32_t r2 = nfft - half;
  float lm = -1.f, rm = -1.f;
  uint32_t li = 0xffffffffu, ri = 0xffffffffu;
  for (uint32_t i = threadIdx.x; i < half; i += blockDim.x) {
    const float m = cabs_rn(X[i]);
    if (m > lm) { lm = m; li = i; }
  }
  for (uint32_t i = r2 + threadIdx.x; i < nfft; i += blockDim.x) {
    const float m = cabs_rn(X[i]);
    if (m > rm) { rm = m; ri = i; }
  }
  wave_argmax(lm, li);
  wave_argmax(rm, ri);
  __shared__ float sm[2][4];
  __shared__ uint32_t si[2][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sm[0][w] = lm; si[0][w] = li; sm[1][w] = rm; si[1][w] = ri; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; ++q) {
      if (sm[0][q] > lm || (sm[0][q] == lm && si[0][q] < li)) { lm = sm[0][q]; li = si[0][q]; }
      if (sm[1][q] > rm || (sm[1][q] == rm && si[1][q] < ri)) { rm = sm[1][q]; ri = si[1][q]; }
    }
    // empty ranges keep the reference's initial indices (0 on the left, nFFT - half on the right),
    // clamped into the spectrum
    if (li == 0xffffffffu) li = 0;
    if (ri == 0xffffffffu) ri = r2 < nfft ? r2 : nfft - 1;
    if (!two_peaks) {
      const int right = rm > lm ? 1 : 0;
      write_peak4(out + (uint64_t)beam * 4, GlobalBins{X}, nfft, right ? ri : li, right, fs);
    } else {
      write_peak4(out + (uint64_t)beam * 8, GlobalBins{X}, nfft, li, 0, fs);
      write_peak4(out + (uint64_t)beam * 8 + 4, GlobalBins{X}, nfft, ri, 1, fs);
    }
  }
}

// ---------------------------------------------------------------------------
// AllMaxima on a spectrum: all_maxima_kernel_sources_rocm.hpp:57-264 and
// all_maxima_pipeline_rocm.cpp:274-307.  One workgroup per beam walks [s0, s1)
// in chunks of 1024*8 bins; flags -> block exclusive scan -> records in
// ascending bin order; records beyond `cap` are dropped, count is uncapped.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_all_maxima(const float2* __restrict__ spec, MaxValue* __restrict__ out,
                                                     uint32_t* __restrict__ counts, uint32_t batch, uint32_t nfft,
                                                     float fs, uint32_t s0, uint32_t s1, uint32_t cap) {
  constexpr int T = 1024, E = 8;
  const uint32_t beam = blockIdx.x;
  if (beam >= batch) return;
  const float2* X = spec + (uint64_t)beam * nfft;
  MaxValue* o = out + (uint64_t)beam * cap;
  __shared__ int wsum[16];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const float bw = fs / (float)nfft;
  uint32_t base = 0;  // records found before this chunk
  for (uint32_t c0 = s0; c0 < s1; c0 += T * E) {
    const uint32_t k0 = c0 + (uint32_t)tid * E;
    uint32_t flags = 0;
    int c = 0;
    if (k0 < s1) {
      float prev = cabs_rn(X[k0 - 1]);  // s0 >= 1 so k0 - 1 is valid
      float cur = cabs_rn(X[k0]);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t k = k0 + e;
        if (k < s1) {
          const float nxt = cabs_rn(X[k + 1]);  // s1 <= nfft - 1 so k + 1 is valid
          if (cur > prev && cur > nxt) { flags |= 1u << e; ++c; }
          prev = cur;
          cur = nxt;
        }
      }
    }
    const int inc = wave_incl_scan(c);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i < w) off += wsum[i];
      tot += wsum[i];
    }
    __syncthreads();
    uint32_t pos = base + (uint32_t)(off + inc - c);
    while (flags) {
      const int e = __builtin_ctz(flags);
      flags &= flags - 1;
      if (pos < cap) {
        const uint32_t k = k0 + e;
        const float2 v = X[k];
        MaxValue m;
        m.index = k; m.real = v.x; m.imag = v.y; m.magnitude = cabs_rn(v);
        m.phase = atan2f(v.y, v.x) * 57.29577951f; m.freq_offset = 0.f;
        m.refined_frequency = (float)k * bw; m.pad = 0;
        o[pos] = m;
      }
      ++pos;
    }
    base += (uint32_t)tot;
  }
  if (tid == 0) counts[beam] = base;
}

}  // namespace spx

extern "C" {

static unsigned ew_blocks(uint64_t total) {
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  return blocks == 0 ? 1u : (unsigned)blocks;
}

hipError_t spx_launch_mag_phase(const void* in, void* out, uint64_t total, hipStream_t st) {
  using namespace spx;
  hipLaunchKernelGGL(k_mag_phase, dim3(ew_blocks(total)), dim3(256), 0, st, (const float2*)in, (float2*)out, total);
  return hipGetLastError();
}

hipError_t spx_launch_magnitude(const void* in, float* out, uint64_t total, float inv_n, int squared,
                                hipStream_t st) {
  using namespace spx;
  hipLaunchKernelGGL(k_magnitude, dim3(ew_blocks(total)), dim3(256), 0, st, (const float2*)in, out, total, inv_n,
                     squared);
  return hipGetLastError();
}

hipError_t spx_launch_peaks_top(const void* spec, void* out, uint32_t batch, uint32_t nfft, uint32_t search_range,
                                float fs, int two_peaks, hipStream_t st) {
  using namespace spx;
  hipLaunchKernelGGL(k_peaks_top, dim3(batch), dim3(256), 0, st, (const float2*)spec, (MaxValue*)out, batch, nfft,
                     search_range, fs, two_peaks);
  return hipGetLastError();
}

hipError_t spx_launch_all_maxima(const void* spec, void* out, uint32_t* counts, uint32_t batch, uint32_t nfft,
                                 float fs, uint32_t s0, uint32_t s1, uint32_t cap, hipStream_t st) {
  using namespace spx;
  hipLaunchKernelGGL(k_all_maxima, dim3(batch), dim3(1024), 0, st, (const float2*)spec, (MaxValue*)out, counts,
                     batch, nfft, fs, s0, s1, cap);
  return hipGetLastError();
}

}  // extern "C"
