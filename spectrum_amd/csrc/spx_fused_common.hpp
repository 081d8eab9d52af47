// Device code shared by the fused kernels (one channel per workgroup) and the
// persistent uneven-length batch kernel: input loaders, magnitude and record
// helpers, block reductions, the first-K survivor selection and the
// ONE_PEAK / TWO_PEAKS record writers.
#pragma once
#include <type_traits>
#include "spx_device.hpp"
#include "spx_types.h"
#include "spx_kernels.h"

namespace spx {

using MaxValue = spx_max_value;

// ---------------------------------------------------------------------------
// input loaders (complex64, complex fp16, complex bf16)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Global-memory element at base + a 32-bit byte offset.  The base is cast to
// the global address space before the offset is added, so the load takes the
// scalar-base form (global_load v, v_off, s[base]) with one offset VGPR; an
// element index, or a cast after the add, needs a 64-bit address per load (a
// channel and a window table are far below 4 GiB).
template <class T, bool STREAM = false>
__device__ __forceinline__ T ld_at(const void* base, uint32_t off) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const T gT;
  gT* p = (gT*)((gchar*)base + off);
  if constexpr (STREAM) return __builtin_nontemporal_load(p);
  else return *p;
#else
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
#endif
}

// STREAM: non-temporal complex64 loads
template <int DT, bool STREAM = false>
__device__ __forceinline__ float2 ld_raw(const void* chan_base, uint32_t i) {
  if constexpr (DT == SPX_DTYPE_C64) {
    const stream_float2 v = ld_at<stream_float2, STREAM>(chan_base, i * 8u);
    return make_float2(v.x, v.y);
  } else if constexpr (DT == SPX_DTYPE_C32F16) {
    const uint32_t u = ld_at<uint32_t>(chan_base, i * 4u);
    return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16)));
  } else {
    const uint32_t u = ld_at<uint32_t>(chan_base, i * 4u);
    return make_float2(bf16_to_f((uint16_t)(u & 0xffffu)), bf16_to_f((uint16_t)(u >> 16)));
  }
}

// v[r] = scale * w[n] * x[n] for n = n0 + r*stride, r < R; zero for n >= n_point
// (PAD only).  Offsets are 32-bit from a per-channel scalar base, and indices
// are clamped instead of branched on, so all R loads (and the R window loads)
// are in flight together before the first use.
template <int DT, int R, bool PAD, bool STREAM = false>
__device__ __forceinline__ void load_run_t(const FusedArgs& a, uint64_t ch, uint32_t n0, uint32_t stride, float2* v) {
  constexpr int ES = DT == SPX_DTYPE_C64 ? 8 : 4;
  const char* cb = reinterpret_cast<const char*>(a.in) + ch * (uint64_t)a.n_point * ES;
  const uint32_t last = a.n_point - 1;
  float2 raw[R];
  float wv[R];
  if constexpr (!PAD) {
    // load r of every thread at the uniform base cb + r * stride * ES: one
    // offset VGPR (n0) for all R loads, the r steps in scalar registers
#pragma unroll
    for (int r = 0; r < R; ++r) raw[r] = ld_raw<DT, STREAM>(cb + (uint64_t)r * stride * ES, n0);
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) raw[r] = ld_raw<DT, STREAM>(cb, min(n0 + r * stride, last));
  }
  if (a.window != nullptr) {
    if constexpr (!PAD) {
#pragma unroll
      for (int r = 0; r < R; ++r) wv[r] = ld_at<float>(reinterpret_cast<const char*>(a.window) + 4u * r * stride, 4u * n0);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) wv[r] = ld_at<float>(a.window, 4u * min(n0 + r * stride, last));
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) wv[r] = 1.0f;
  }
  const float sc = (a.in_scale != nullptr) ? gp(a.in_scale)[ch] : 1.0f;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float f = PAD ? ((n0 + r * stride < a.n_point) ? wv[r] * sc : 0.0f) : wv[r] * sc;
    v[r] = c_scale(raw[r], f);
  }
}

// STREAM: non-temporal complex64 loads (the short kernels' dense outputs,
// where they measured 4-5 % faster; the peak search measured no faster)
template <int R, bool STREAM = false>
__device__ __forceinline__ void load_run(const FusedArgs& a, uint64_t ch, uint32_t n0, uint32_t stride, float2* v) {
  if (a.n_point == a.nfft) {
    if (a.dtype == SPX_DTYPE_C64) load_run_t<SPX_DTYPE_C64, R, false, STREAM>(a, ch, n0, stride, v);
    else if (a.dtype == SPX_DTYPE_C32F16) load_run_t<SPX_DTYPE_C32F16, R, false>(a, ch, n0, stride, v);
    else load_run_t<SPX_DTYPE_C32BF16, R, false>(a, ch, n0, stride, v);
  } else {
    if (a.dtype == SPX_DTYPE_C64) load_run_t<SPX_DTYPE_C64, R, true, STREAM>(a, ch, n0, stride, v);
    else if (a.dtype == SPX_DTYPE_C32F16) load_run_t<SPX_DTYPE_C32F16, R, true>(a, ch, n0, stride, v);
    else load_run_t<SPX_DTYPE_C32BF16, R, true>(a, ch, n0, stride, v);
  }
}

// Hide a pointer's value from the optimiser so that loads through it are not
// hoisted out of the persistent channel loop (they are channel-invariant, and
// keeping them all live would spill).
__device__ __forceinline__ int opaque_v(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
template <class P>
__device__ __forceinline__ P* opaque(P* p) {
  uint32_t z = 0;
  asm volatile("" : "+s"(z));
  return p + z;
}

// |X| with the hardware square root (v_sqrt_f32, ~1 ulp)
__device__ __forceinline__ float mag_of(float2 x) { return __builtin_amdgcn_sqrtf(fmaf(x.x, x.x, x.y * x.y)); }

__device__ __forceinline__ float parabolic(float yl, float yc, float yr) {
  float den = yl - 2.0f * yc + yr;
  float d = 0.0f;
  if (fabsf(den) > 1e-10f) {
    d = 0.5f * (yl - yr) / den;
    d = fminf(fmaxf(d, -0.5f), 0.5f);
  }
  return d;
}

// One MaxValue record with sub-bin offset d; bin_hz = fs / nfft.
__device__ __forceinline__ void write_record_d(MaxValue* dst, uint32_t k, float2 x, float yc, float d, float bin_hz) {
  MaxValue mv;
  mv.index = k;
  mv.real = x.x;
  mv.imag = x.y;
  mv.magnitude = yc;
  mv.phase = atan2f(x.y, x.x) * 57.29577951f;
  mv.freq_offset = d;
  mv.refined_frequency = ((float)k + d) * bin_hz;
  mv.pad = 0;
  *gp(dst) = mv;
}

__device__ __forceinline__ void write_record_hz(MaxValue* dst, uint32_t k, float2 x, float yl, float yc, float yr,
                                                uint32_t refine, float bin_hz) {
  write_record_d(dst, k, x, yc, (refine != 0) ? parabolic(yl, yc, yr) : 0.0f, bin_hz);
}

__device__ __forceinline__ void write_record(MaxValue* dst, uint32_t k, float2 x, float yl, float yc, float yr,
                                             const FusedArgs& a) {
  float d = (a.refine != 0) ? parabolic(yl, yc, yr) : 0.0f;
  MaxValue mv;
  mv.index = k;
  mv.real = x.x;
  mv.imag = x.y;
  mv.magnitude = yc;
  mv.phase = atan2f(x.y, x.x) * 57.29577951f;
  mv.freq_offset = d;
  mv.refined_frequency = ((float)k + d) * (a.fs / (float)a.nfft);
  mv.pad = 0;
  *gp(dst) = mv;
}

__device__ __forceinline__ bool in_range(uint32_t k, const FusedArgs& a) { return k >= a.s0 && k < a.s1; }

// dense spectrum / magnitude outputs (phase needs atan2f: kept out of line)
__device__ __noinline__ void store_magphase(float2* dst, float2 x) { *gp(dst) = make_float2(mag_of(x), atan2f(x.y, x.x)); }
// STREAM: non-temporal stores of the complex spectrum (measured faster on the
// split path; for the 4-byte |X| outputs they measured slower and stay plain)
template <bool STREAM = false>
__device__ __forceinline__ void store_dense(const FusedArgs& a, uint64_t i, float2 x) {
  if (a.spec_out) st_f2<STREAM>(a.spec_out + i, x);
  if (a.mag_out) {
    if (a.mag_kind == SPX_MAG_ABS) gp(a.mag_out)[i] = mag_of(x) * a.mag_scale;
    else if (a.mag_kind == SPX_MAG_SQUARED) gp(a.mag_out)[i] = fmaf(x.x, x.x, x.y * x.y) * a.mag_scale;
    else store_magphase(reinterpret_cast<float2*>(a.mag_out) + i, x);
  }
}

// ---------------------------------------------------------------------------
// ONE_PEAK / TWO_PEAKS records (include/spectrum/kernels/fft_kernel_sources_rocm.hpp:122-519)
// ---------------------------------------------------------------------------
// |X| with the correctly rounded square root, as the reference's __fsqrt_rn
__device__ __forceinline__ float cabs_rn(float2 v) { return sqrtf(v.x * v.x + v.y * v.y); }

// Bin accessors for write_peak4: a stored spectrum in global memory, or a
// channel's spectrum in LDS (padded layout LAY).
struct GlobalBins {
  const float2* X;
  __device__ __forceinline__ float2 operator()(uint32_t i) const { return X[i]; }
};
template <class LAY>
struct LdsBins {
  const float2* lds;
  __device__ __forceinline__ float2 operator()(uint32_t i) const { return lds[LAY::idx(0, (int)i)]; }
};

// The four records of one peak at bin c: the parabolic result, then the left
// neighbour, the centre and the right neighbour (non-circular: bins 0 and
// nfft-1 have one neighbour, and the offset stays 0 without both).  mirror:
// right-side peak, frequencies reported as fs - f.
template <class BINS>
__device__ void write_peak4(MaxValue* out, BINS X, uint32_t nfft, uint32_t c, int mirror, float fs) {
  // no fused multiply-adds: fs - (c + d) bw cancels, and every caller (the
  // search on a stored spectrum and the fused kernels) must round alike
#pragma clang fp contract(off)
  const float bw = fs / (float)nfft;
  const float2 cv = X(c);
  const float yc = cabs_rn(cv);
  float2 lv = make_float2(0.f, 0.f), rv = make_float2(0.f, 0.f);
  float yl = 0.f, yr = 0.f;
  const int hl = c > 0 ? 1 : 0;
  const int hr = c < nfft - 1 ? 1 : 0;
  if (hl) { lv = X(c - 1); yl = cabs_rn(lv); }
  if (hr) { rv = X(c + 1); yr = cabs_rn(rv); }
  float fo = 0.f;
  float rf = (float)c * bw;
  if (hl && hr) {
    const float den = yl - 2.0f * yc + yr;
    if (fabsf(den) > 1e-10f) {
      fo = 0.5f * (yl - yr) / den;
      if (fo < -0.5f) fo = -0.5f;
      if (fo > 0.5f) fo = 0.5f;
      rf = ((float)c + fo) * bw;
    }
  }
  if (mirror) rf = fs - rf;
  MaxValue m;
  m.pad = 0;
  m.index = c; m.real = cv.x; m.imag = cv.y; m.magnitude = yc;
  m.phase = atan2f(cv.y, cv.x) * 57.29577951f; m.freq_offset = fo; m.refined_frequency = rf;
  gp(out)[0] = m;
  float rfl = hl ? (float)(c - 1) * bw : 0.f;
  if (mirror && hl) rfl = fs - rfl;
  m.index = hl ? c - 1 : 0; m.real = lv.x; m.imag = lv.y; m.magnitude = yl;
  m.phase = hl ? atan2f(lv.y, lv.x) * 57.29577951f : 0.f; m.freq_offset = 0.f; m.refined_frequency = rfl;
  gp(out)[1] = m;
  float rfc = (float)c * bw;
  if (mirror) rfc = fs - rfc;
  m.index = c; m.real = cv.x; m.imag = cv.y; m.magnitude = yc;
  m.phase = atan2f(cv.y, cv.x) * 57.29577951f; m.freq_offset = 0.f; m.refined_frequency = rfc;
  gp(out)[2] = m;
  float rfr = hr ? (float)(c + 1) * bw : 0.f;
  if (mirror && hr) rfr = fs - rfr;
  m.index = hr ? c + 1 : 0; m.real = rv.x; m.imag = rv.y; m.magnitude = yr;
  m.phase = hr ? atan2f(rv.y, rv.x) * 57.29577951f : 0.f; m.freq_offset = 0.f; m.refined_frequency = rfr;
  gp(out)[3] = m;
}

// Winner of one side, and which record set to write: ONE_PEAK (peak_mode 3)
// writes the larger of the two sides (the left one on ties) at out[0..3];
// TWO_PEAKS (peak_mode 4) writes the left side at out[0..3] and the right
// side at out[4..7].  Empty ranges keep the reference's initial bins (0 on
// the left, nfft - half on the right, clamped into the spectrum).
template <class BINS>
__device__ __forceinline__ void write_top_records(MaxValue* out, BINS X, const FusedArgs& a, float lm, uint32_t li,
                                                  float rm, uint32_t ri) {
  const uint32_t N = a.nfft;
  if (li == 0xffffffffu) li = 0;
  if (ri == 0xffffffffu) ri = a.s1 < N ? a.s1 : N - 1;
  if (a.peak_mode == 3) {
    const int right = rm > lm ? 1 : 0;
    write_peak4(out, X, N, right ? ri : li, right, a.fs);
  } else {
    write_peak4(out, X, N, li, 0, a.fs);
    write_peak4(out + 4, X, N, ri, 1, a.fs);
  }
}

// threshold from the block-reduced sums; -inf when no threshold applies
__device__ __forceinline__ float threshold_of(const FusedArgs& a, double s1, double s2, double cnt, float& mu_out,
                                              float& sd_out) {
  double mu = cnt > 0 ? s1 / cnt : 0.0;
  double var = cnt > 0 ? s2 / cnt - mu * mu : 0.0;
  if (var < 0) var = 0;
  double sd = sqrt(var);
  mu_out = (float)mu;
  sd_out = (float)sd;
  if (a.peak_mode == 1) return (float)(mu + (double)a.nsigma * sd);
  return -__builtin_inff();
}

// Cross-wave combine of the NW <= 16 per-wave partials in LDS, done by every
// wave lane-parallel instead of an NW-step loop in every lane: lane i < NW
// takes wave i's partial and a prefix pass over lanes 0-15 (DPP row shifts
// 1, 2, 4, 8 with the identity shifted in) leaves the combined value of
// lanes 0 .. p in lane p; lane NW - 1 holds the total.
namespace rowred {
template <int CTRL, class V>
__device__ __forceinline__ V shr(V id, V x) {
  if constexpr (sizeof(V) == 8) return dpp::mov<CTRL, 0x1>(id, x);
  else if constexpr (__is_same(V, float)) return dpp::mov<CTRL, 0x1>(id, x);
  else return (V)dpp::mov<CTRL, 0x1>((int)id, (int)x);
}
template <class V>
__device__ __forceinline__ V lane_of(V x, int i) {
  if constexpr (sizeof(V) == 8) {
    const int2 a = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(V, make_int2(__builtin_amdgcn_readlane(a.x, i), __builtin_amdgcn_readlane(a.y, i)));
  } else {
    return __builtin_bit_cast(V, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), i));
  }
}
}  // namespace rowred

// Block-wide sum of two doubles and argmax (lowest index wins ties).
// Result valid in all threads on return.  REUSE: end with a barrier so that
// `red` may be written again (not needed when it is the last use).
template <int T, bool REUSE = true>
__device__ __forceinline__ void block_stats(double& s1, double& s2, float& bm, uint32_t& bk, BlockRed& red) {
  constexpr int NW = (T + 63) / 64;
  static_assert(NW <= 16, "at most 16 waves");
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  wave_argmax(bm, bk);
  if (lane == 0) {
    red.s1[w] = s1;
    red.s2[w] = s2;
    red.m[w] = bm;
    red.k[w] = bk;
  }
  __syncthreads();
  const bool in = lane < NW;
  double a1 = in ? red.s1[lane] : 0.0, a2 = in ? red.s2[lane] : 0.0;
  float m = in ? red.m[lane] : -1.0f;
  uint32_t k = in ? red.k[lane] : 0xffffffffu;
  auto step = [&](auto ctrl) {
    constexpr int C = decltype(ctrl)::value;
    a1 += rowred::shr<C>(0.0, a1);
    a2 += rowred::shr<C>(0.0, a2);
    const float om = rowred::shr<C>(-1.0f, m);
    const uint32_t ok = rowred::shr<C>(0xffffffffu, k);
    const bool up = (om > m) | ((om == m) & (ok < k));
    m = up ? om : m;
    k = up ? ok : k;
  };
  if constexpr (NW > 1) step(std::integral_constant<int, dpp::kShr1>{});
  if constexpr (NW > 2) step(std::integral_constant<int, dpp::kShr2>{});
  if constexpr (NW > 4) step(std::integral_constant<int, dpp::kShr4>{});
  if constexpr (NW > 8) step(std::integral_constant<int, dpp::kShr8>{});
  s1 = rowred::lane_of(a1, NW - 1);
  s2 = rowred::lane_of(a2, NW - 1);
  bm = rowred::lane_of(m, NW - 1);
  bk = rowred::lane_of(k, NW - 1);
  if constexpr (REUSE) __syncthreads();
}

// block_stats without the arg-max: block-wide sums of two doubles, valid in
// all threads on return.
template <int T, bool REUSE = true>
__device__ __forceinline__ void block_sums(double& s1, double& s2, BlockRed& red) {
  constexpr int NW = (T + 63) / 64;
  static_assert(NW <= 16, "at most 16 waves");
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane == 0) {
    red.s1[w] = s1;
    red.s2[w] = s2;
  }
  __syncthreads();
  const bool in = lane < NW;
  double a1 = in ? red.s1[lane] : 0.0, a2 = in ? red.s2[lane] : 0.0;
  auto step = [&](auto ctrl) {
    constexpr int C = decltype(ctrl)::value;
    a1 += rowred::shr<C>(0.0, a1);
    a2 += rowred::shr<C>(0.0, a2);
  };
  if constexpr (NW > 1) step(std::integral_constant<int, dpp::kShr1>{});
  if constexpr (NW > 2) step(std::integral_constant<int, dpp::kShr2>{});
  if constexpr (NW > 4) step(std::integral_constant<int, dpp::kShr4>{});
  if constexpr (NW > 8) step(std::integral_constant<int, dpp::kShr8>{});
  s1 = rowred::lane_of(a1, NW - 1);
  s2 = rowred::lane_of(a2, NW - 1);
  if constexpr (REUSE) __syncthreads();
}

// Block exclusive scan of per-thread counts; returns the offset, total in `total`.
// REUSE: end with a barrier so that `red.c` may be written again (not needed
// when a later barrier comes first).
template <int T, bool REUSE = true>
__device__ __forceinline__ int block_excl_scan(int c, int& total, BlockRed& red) {
  constexpr int NW = (T + 63) / 64;
  static_assert(NW <= 16, "at most 16 waves");
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int inc = wave_incl_scan(c);
  if (lane == 63) red.c[w] = inc;
  __syncthreads();
  // inclusive scan of the wave totals over lanes 0 .. NW - 1
  int x = lane < NW ? red.c[lane] : 0;
  if constexpr (NW > 1) x += rowred::shr<dpp::kShr1>(0, x);
  if constexpr (NW > 2) x += rowred::shr<dpp::kShr2>(0, x);
  if constexpr (NW > 4) x += rowred::shr<dpp::kShr4>(0, x);
  if constexpr (NW > 8) x += rowred::shr<dpp::kShr8>(0, x);
  const int tot = __builtin_amdgcn_readlane(x, NW - 1);
  const int off = w > 0 ? __builtin_amdgcn_readlane(x, w - 1) : 0;
  if constexpr (REUSE) __syncthreads();
  total = tot;
  return off + inc - c;
}

// ===========================================================================
// large path: N = 256 * N2 (N2 = 128 .. 1024), x[n1 + 256*n2]
//   phase 1: column n1, N2-point DFT over n2, times W_N^(n1*k2) -> Z[k2][n1]
//   phase 2: row k2, 256-point DFT over n1               -> X[k2 + N2*k1]
// ===========================================================================
struct Cand {
  uint32_t k;
  float re, im, ml, mc, mr;
};

// Selection of the first K threshold survivors in ascending bin order from an
// unsorted candidate list (local maxima that may clear the threshold), plus the
// counts and stats outputs of channel och.  Survivor bins are marked in a
// bitmap of NW = N/32 words in LDS; a survivor's rank is the number of marked
// bins below it, from the popcount prefix of the bitmap words, so the records
// land in order with no sort and for any K.  Thread t owns words
// [t*WPT, (t+1)*WPT): base[t] holds the survivors below its first word and
// pre[w] the survivors in its earlier words (at most 32*WPT, 16 bits).
// cand(i) returns candidate i (mc = NaN: no survivor); hook(i, rank, q) runs
// after the record of candidate i is written at `rank`.  Every thread of the
// T-thread workgroup calls it; bits, pre, base are LDS.
struct NoRecordHook {
  __device__ __forceinline__ void operator()(uint32_t, uint32_t, const Cand&) const {}
};
template <int T, int NW, class SRC, class HOOK = NoRecordHook>
__device__ __forceinline__ void select_first_k(const FusedArgs& a, SRC cand, uint32_t ncand, float thr, float mu,
                                               float sd, double P, uint64_t och, uint32_t* bits, uint16_t* pre,
                                               int* base, BlockRed& red, HOOK hook = HOOK()) {
  constexpr int WPT = NW / T;
  static_assert(NW % T == 0 && 32 * WPT < 65536, "bitmap words per thread");
  const int tid = threadIdx.x;
  for (int i = opaque_v(tid); i < NW; i += T) bits[i] = 0u;
  __syncthreads();
  for (uint32_t i = opaque_v(tid); i < ncand; i += T) {
    const Cand q = cand(i);
    if (q.mc > thr) atomicOr(&bits[q.k >> 5], 1u << (q.k & 31));
  }
  __syncthreads();
  int run = 0;
  for (int w = 0; w < WPT; ++w) {
    pre[tid * WPT + w] = (uint16_t)run;
    run += __popc(bits[tid * WPT + w]);
  }
  int total;
  base[tid] = block_excl_scan<T>(run, total, red);
  __syncthreads();
  const uint32_t kcap = a.kcap;
  MaxValue* out = a.peaks + och * kcap;
  if (kcap > 0 && total > 0) {
    for (uint32_t i = opaque_v(tid); i < ncand; i += T) {
      const Cand q = cand(i);
      if (q.mc > thr) {
        const uint32_t w = q.k >> 5;
        const uint32_t rank = (uint32_t)base[w / WPT] + pre[w] + __popc(bits[w] & ((1u << (q.k & 31)) - 1u));
        if (rank < kcap) {
          write_record(out + rank, q.k, make_float2(q.re, q.im), q.ml, q.mc, q.mr, a);
          hook(i, rank, q);
        }
      }
    }
  }
  if (tid == 0) {
    if (a.counts) {
      gp(a.counts)[2 * och + 0] = min((uint32_t)total, kcap);
      gp(a.counts)[2 * och + 1] = (uint32_t)total;
    }
    if (a.stats) {
      gp(a.stats)[4 * och + 0] = mu;
      gp(a.stats)[4 * och + 1] = sd;
      gp(a.stats)[4 * och + 2] = thr;
      gp(a.stats)[4 * och + 3] = (float)P;
    }
  }
}

}  // namespace spx
