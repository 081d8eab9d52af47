// Device code shared by the fused kernels (one channel per workgroup) and the
// persistent uneven-length batch kernel: input loaders, magnitude and record
// helpers, block reductions and the whole large-N (four-step) channel body.
#pragma once
#include "spx_device.hpp"
#include "spx_types.h"
#include "spx_kernels.h"

namespace spx {

using MaxValue = spx_max_value;

// ---------------------------------------------------------------------------
// input loaders (complex64, complex fp16, complex bf16)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// STREAM: non-temporal complex64 loads
template <int DT, bool STREAM = false>
__device__ __forceinline__ float2 ld_raw(const void* chan_base, uint32_t i) {
  if constexpr (DT == SPX_DTYPE_C64) {
    return ld_f2<STREAM>(reinterpret_cast<const float2*>(chan_base) + i);
  } else if constexpr (DT == SPX_DTYPE_C32F16) {
    const __half2 h = gp(reinterpret_cast<const __half2*>(chan_base))[i];
    return __half22float2(h);
  } else {
    const uint32_t u = gp(reinterpret_cast<const uint32_t*>(chan_base))[i];
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
  const void* cb = reinterpret_cast<const char*>(a.in) + ch * (uint64_t)a.n_point * ES;
  const uint32_t last = a.n_point - 1;
  float2 raw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) raw[r] = ld_raw<DT, STREAM>(cb, PAD ? min(n0 + r * stride, last) : n0 + r * stride);
  float wv[R];
  if (a.window != nullptr) {
#pragma unroll
    for (int r = 0; r < R; ++r) wv[r] = gp(a.window)[PAD ? min(n0 + r * stride, last) : n0 + r * stride];
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

// Block-wide sum of two doubles and argmax (lowest index wins ties).
// Result valid in all threads on return.  REUSE: end with a barrier so that
// `red` may be written again (not needed when it is the last use).
template <int T, bool REUSE = true>
__device__ __forceinline__ void block_stats(double& s1, double& s2, float& bm, uint32_t& bk, BlockRed& red) {
  constexpr int NW = (T + 63) / 64;
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
  s1 = 0;
  s2 = 0;
  bm = -1.0f;
  bk = 0xffffffffu;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    s1 += red.s1[i];
    s2 += red.s2[i];
    if (red.m[i] > bm || (red.m[i] == bm && red.k[i] < bk)) {
      bm = red.m[i];
      bk = red.k[i];
    }
  }
  if constexpr (REUSE) __syncthreads();
}

// Block exclusive scan of per-thread counts; returns the offset, total in `total`.
// REUSE: end with a barrier so that `red.c` may be written again (not needed
// when a later barrier comes first).
template <int T, bool REUSE = true>
__device__ __forceinline__ int block_excl_scan(int c, int& total, BlockRed& red) {
  constexpr int NW = (T + 63) / 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int inc = wave_incl_scan(c);
  if (lane == 63) red.c[w] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (i < w) off += red.c[i];
    tot += red.c[i];
  }
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

__device__ __forceinline__ void push_cand(Cand* cand, uint32_t* ncand, uint32_t cap, uint32_t k, float2 x, float yl,
                                          float yc, float yr) {
  const uint32_t p = atomicAdd(ncand, 1u);
  if (p < cap) {
    Cand cd;
    cd.k = k; cd.re = x.x; cd.im = x.y; cd.ml = yl; cd.mc = yc; cd.mr = yr;
    gp(cand)[p] = cd;
  }
}

// Selection of the first K threshold survivors in ascending bin order from a
// candidate list (local maxima that may clear the threshold), plus the counts
// and stats outputs of channel och.  One bucket per thread (T buckets of
// 2^BSH bins) is histogrammed and scanned to find the bucket holding the K-th
// survivor; survivors up to that bucket are gathered into gat (GMAX entries of
// LDS) and bitonic-sorted by bin.  hist: T ints of LDS.  cand(i) returns
// candidate i (a candidate with mc = NaN is no survivor).  Every thread of the
// T-thread workgroup calls it.
// candidate i of a list in global memory, as stored
struct CandGlobal {
  const Cand* c;
  __device__ __forceinline__ Cand operator()(uint32_t i) const { return gp(c)[i]; }
};

template <int T, int BSH, int GMAX, class SRC>
__device__ __forceinline__ void select_first_k(const FusedArgs& a, SRC cand, uint32_t ncand, float thr, float mu,
                                               float sd, double P, uint64_t och, int* hist, uint2* gat, int& s_bstar,
                                               uint32_t& s_ngather, BlockRed& red) {
  const int tid0 = threadIdx.x;
  for (int i = opaque_v(tid0); i < T; i += T) hist[i] = 0;
  __syncthreads();
  for (uint32_t i0 = opaque_v(tid0); i0 < ncand; i0 += 4 * T) {
    Cand cd[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cd[q] = cand(min(i0 + q * T, ncand - 1));
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (i0 + q * T < ncand && cd[q].mc > thr) atomicAdd(&hist[cd[q].k >> BSH], 1);
  }
  __syncthreads();
  // exclusive scan over buckets (one bucket per thread)
  int hv = hist[tid0];
  int total;
  int excl = block_excl_scan<T>(hv, total, red);
  const int kcap = (int)a.kcap;
  if (tid0 == 0) { s_bstar = T - 1; s_ngather = 0; }
  __syncthreads();
  if (total > kcap && excl < kcap && excl + hv >= kcap) s_bstar = tid0;
  __syncthreads();
  const int bstar = s_bstar;
  // gather (key = bin, value = candidate index) of survivors in buckets <= bstar
  for (uint32_t i0 = opaque_v(tid0); i0 < ncand; i0 += 4 * T) {
    Cand cd[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cd[q] = cand(min(i0 + q * T, ncand - 1));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t i = i0 + q * T;
      if (i < ncand && cd[q].mc > thr && (int)(cd[q].k >> BSH) <= bstar) {
        const uint32_t p = atomicAdd(&s_ngather, 1u);
        if (p < (uint32_t)GMAX) gat[p] = make_uint2(cd[q].k, i);
      }
    }
  }
  __syncthreads();
  const int ng = min((int)s_ngather, GMAX);
  int np2 = 1;
  while (np2 < ng) np2 <<= 1;
  for (int i = ng + opaque_v(tid0); i < np2; i += T) gat[i] = make_uint2(0xffffffffu, 0u);
  __syncthreads();
  // bitonic sort by key
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = opaque_v(tid0); i < np2 / 2; i += T) {
        const int lo = 2 * i - (i & (stride - 1));
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        uint2 x = gat[lo], y = gat[hi];
        if ((x.x > y.x) == up) { gat[lo] = y; gat[hi] = x; }
      }
      __syncthreads();
    }
  }
  const int nout = min(total, kcap);
  MaxValue* out = a.peaks + (uint64_t)och * a.kcap;
  for (int i = opaque_v(tid0); i < nout && i < ng; i += T) {
    const Cand cd = cand(gat[i].y);
    write_record(out + i, cd.k, make_float2(cd.re, cd.im), cd.ml, cd.mc, cd.mr, a);
  }
  if (tid0 == 0) {
    if (a.counts) {
      gp(a.counts)[2 * och + 0] = (uint32_t)nout;
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

template <int N2_, int T_ = 1024>
struct LargeCfg {
  static constexpr int N1 = 256;                // phase-2 row length (radix 16 x 16)
  static constexpr int N2 = N2_;                // phase-1 column length
  static constexpr int N = N1 * N2;
  static constexpr int T = T_;                  // workgroup size (1024, or 512 for two workgroups per CU)
  static constexpr int PPT = 16;
  static constexpr int TILE = T * PPT;          // points per tile
  static constexpr int CB = TILE / N2;          // phase-1 columns per tile
  static constexpr int RB = TILE / N1;          // phase-2 rows per tile
  static constexpr int NT1 = N1 / CB;           // phase-1 tiles
  static constexpr int NT2 = N2 / RB;           // phase-2 tiles
  using LAY1 = LayInter<N2, CB>;
  using LAY2 = LayContig<N1, RB>;
  static constexpr int LDS = LAY2::SIZE > LAY1::SIZE ? LAY2::SIZE : LAY1::SIZE;
  // scratch slot layout (bytes)
  static constexpr uint64_t Z_BYTES = (uint64_t)N * 8;
  static constexpr uint64_t CAND_CAP = N / 2 + 2;
  static constexpr uint64_t CAND_BYTES = (CAND_CAP * sizeof(Cand) + 255) / 256 * 256;
  static constexpr uint64_t EDGE_M_BYTES = (uint64_t)NT2 * 4 * N1 * 4;
  static constexpr uint64_t EDGE_X_BYTES = (uint64_t)NT2 * 2 * N1 * 8;
  static constexpr uint64_t SLOT_BYTES = Z_BYTES + CAND_BYTES + EDGE_M_BYTES + EDGE_X_BYTES;
  static constexpr int NBK = T;                 // selection buckets, one per thread
  static constexpr int BSH = ilog2(N) - ilog2(T);  // bucket = k >> BSH
  static_assert(NT1 >= 1 && NT2 >= 1 && CB >= 8 && RB >= 4, "tile shape");
  static_assert(Plan<N1>::RL == 16, "rows end on a radix-16 pass");
};

// TWL: the per-pass twiddle tables of the 256-point rows and of the column
// length are copied per channel into LDS (padded rows) instead of being read
// from global memory.
constexpr int kTwRows = Plan<256>::TWSIZE_LDS;
template <int N2>
constexpr int tw_cols_lds() {
  return Plan<N2, 0>::TWSIZE_LDS > Plan<N2, 1>::TWSIZE_LDS ? Plan<N2, 0>::TWSIZE_LDS : Plan<N2, 1>::TWSIZE_LDS;
}

template <int N2, int T = 1024, bool TWL = true>
struct LargeShared {
  using C = LargeCfg<N2, T>;
  float2 tile[C::LDS];
  float2 twa[TWL ? kTwRows : 1];
  float2 twb[TWL ? tw_cols_lds<N2>() : 1];
  BlockRed red;
  int hist[C::NBK];
  uint32_t ncand;
  uint32_t ngather;
  int bstar;
};

// One channel of length N = 256*N2 through the three phases, run by a whole
// workgroup of TW threads.  ich indexes the input (a.in + ich*n_point), och the
// outputs; slot is this workgroup's scratch slot.  V picks the radix order of
// the N2-point column transform (Plan<N2, V>; tw_b must hold its tables).
// MODE 0: dense outputs only (spectrum / |X| / |X|^2 / mag+phase); MODE 1: peaks only.
template <int N2, int MODE, int V, int TW = 1024, bool TWL = true>
__device__ __forceinline__ void large_channel(const FusedArgs& a, uint64_t ich, uint64_t och, unsigned char* slot,
                                              LargeShared<N2, TW, TWL>& sh) {
  constexpr bool kLargeTwLds = TWL;
  using C = LargeCfg<N2, TW>;
  constexpr int N = C::N, N1 = C::N1, T = C::T, PPT = C::PPT, CB = C::CB, RB = C::RB;
  constexpr int NT1 = C::NT1, NT2 = C::NT2;
  constexpr int R01 = Plan<N2, V>::R0, RL1 = Plan<N2, V>::RL;
  using LAY1 = typename C::LAY1;
  using LAY2 = typename C::LAY2;

  float2* lds = sh.tile;
  float* ldsf = reinterpret_cast<float*>(sh.tile);
  const int tid0 = threadIdx.x;

  float2* Z = reinterpret_cast<float2*>(slot);
  Cand* cand = reinterpret_cast<Cand*>(slot + C::Z_BYTES);
  float* edgeM = reinterpret_cast<float*>(slot + C::Z_BYTES + C::CAND_BYTES);
  float2* edgeX = reinterpret_cast<float2*>(slot + C::Z_BYTES + C::CAND_BYTES + C::EDGE_M_BYTES);

  constexpr bool peaks = (MODE == 1);
  const bool full_band = (a.s0 == 0 && a.s1 == (uint32_t)N);

  const float2* twb_src = a.tw_b;  // column tables: LDS copy or global
  const float2* twa_src = a.tw_a;  // row tables
  if constexpr (kLargeTwLds) {
    static_assert(Plan<N2, V>::TWSIZE_LDS <= tw_cols_lds<N2>() && Plan<256>::TWSIZE_LDS <= kTwRows, "LDS tables");
    __syncthreads();  // the previous channel may still read them
    copy_tw_to_lds<N2, V, T>(a.tw_b, sh.twb, opaque_v(tid0));
    copy_tw_to_lds<256, 0, T>(a.tw_a, sh.twa, opaque_v(tid0));
    __syncthreads();
    twb_src = sh.twb;
    twa_src = sh.twa;
  }
  {
    int tid = opaque_v(tid0);
    FusedArgs ai = a;
    // ------------------------------------------------------------------
    // phase 1
    // ------------------------------------------------------------------
    float energy = 0.f;
#pragma unroll 1
    for (int t = 0; t < NT1; ++t) {
      tid = opaque_v(tid0);
      ai.tw_b = opaque(a.tw_b);
      ai.tw_4 = opaque(a.tw_4);
      ai.window = opaque(a.window);
      float2 v[PPT];
      if constexpr (R01 == 16) {
        const int c = tid % CB, jj = tid / CB;  // jj in [0, N2/16)
        const int n1 = t * CB + c;
        load_run<16>(ai, ich, (uint32_t)(n1 + N1 * jj), (uint32_t)(N1 * (N2 / 16)), v);
      } else {
#pragma unroll
        for (int b = 0; b < PPT / R01; ++b) {
          const int j = tid + b * T;
          const int c = j % CB, jj = j / CB;
          const int n1 = t * CB + c;
          load_run<R01>(ai, ich, (uint32_t)(n1 + N1 * jj), (uint32_t)(N1 * (N2 / R01)), &v[b * R01]);
        }
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) energy = fmaf(v[r].x, v[r].x, fmaf(v[r].y, v[r].y, energy));
      if constexpr (kLargeTwLds) fft_tile<LAY1, N2, CB, T, PPT, V, true>(lds, v, twb_src, tid);
      else fft_tile<LAY1, N2, CB, T, PPT, V>(lds, v, ai.tw_b, tid);
      // v[b*RL1 + r] = Y[column s][k2 = jj' + r*N2/RL1] for j = tid + b*T
#pragma unroll
      for (int b = 0; b < PPT / RL1; ++b) {
        int s, jb;
        LAY1::template map<RL1>(tid + b * T, s, jb);
        const uint32_t col = (uint32_t)(t * CB + s);
#pragma unroll
        for (int r = 0; r < RL1; ++r) {
          const uint32_t e = (uint32_t)(jb + r * (N2 / RL1)) * N1 + col;
          gp(Z)[e] = c_mul(v[b * RL1 + r], gp(ai.tw_4)[e]);
        }
      }
    }
    __syncthreads();

    // block reduce the input energy -> P = mean |X|^2 (Parseval)
    double P;
    {
      double e = energy, dummy = 0.0;
      float bmx = -1.f;
      uint32_t bkx = 0;
      block_stats<T>(e, dummy, bmx, bkx, sh.red);
      P = e;  // sum_n |x_w[n]|^2 = (1/N) sum_k |X[k]|^2
    }
    float lb = 0.f;
    if (a.peak_mode == 1 && full_band && a.nsigma >= 0.f) lb = (float)(fmin(1.0, (double)a.nsigma) * sqrt(P) * 0.999);
    if (tid == 0) sh.ncand = 0;

    // ------------------------------------------------------------------
    // phase 2
    // ------------------------------------------------------------------
    float fs1 = 0.f, fs2 = 0.f;
    float bm = -1.f;
    uint32_t bk = 0xffffffffu;
    float2 bx = make_float2(0.f, 0.f);
    float bl = 0.f, br = 0.f;

#pragma unroll 1
    for (int t = 0; t < NT2; ++t) {
      tid = opaque_v(tid0);
      ai.tw_a = opaque(a.tw_a);
      const int s_row = tid / (N1 / 16), jj_row = tid % (N1 / 16);  // radix-16 mapping of both passes
      float2 v[PPT];
      {
        const float2* zr = Z + (uint32_t)((t * RB + s_row) * N1 + jj_row);
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = gp(zr)[r * (N1 / 16)];
      }
      if constexpr (kLargeTwLds) fft_tile<LAY2, N1, RB, T, PPT, 0, true>(lds, v, twa_src, tid);
      else fft_tile<LAY2, N1, RB, T, PPT>(lds, v, ai.tw_a, tid);
      // v[r] = X[k2 = t*RB + s_row][k1 = jj_row + 16 r]  ->  bin k2 + N2*k1
      const int k2 = t * RB + s_row;
      float m[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        m[r] = mag_of(v[r]);
        if (full_band || in_range((uint32_t)k2 + (uint32_t)N2 * (uint32_t)(jj_row + 16 * r), a)) {
          fs1 += m[r];
          fs2 = fmaf(m[r], m[r], fs2);
        }
      }
      __syncthreads();  // LDS free again
      if constexpr (!peaks) {
        // transpose through LDS so that natural-order runs of RB bins are written
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int e = (jj_row + 16 * r) * RB + s_row;
          lds[e + (e >> 4)] = v[r];
        }
        __syncthreads();
        const uint64_t cbase = (uint64_t)och * N;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int e = tid + i * T;
          const int k1 = e / RB, s = e % RB;
          store_dense(a, cbase + (uint64_t)(t * RB + s) + (uint64_t)N2 * k1, lds[e + (e >> 4)]);
        }
      } else {
        // magnitude tile [s][k1] and the edge rows 0, 1, RB-2, RB-1
#pragma unroll
        for (int r = 0; r < 16; ++r) ldsf[s_row * N1 + jj_row + 16 * r] = m[r];
        {
          int slotm = -1, slotx = -1;
          if (s_row == 0) { slotm = 0; slotx = 0; }
          else if (s_row == 1) slotm = 1;
          else if (s_row == RB - 2) slotm = 2;
          else if (s_row == RB - 1) { slotm = 3; slotx = 1; }
          if (slotm >= 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              const int k1 = jj_row + 16 * r;
              gp(edgeM)[(t * 4 + slotm) * N1 + k1] = m[r];
              if (slotx >= 0) gp(edgeX)[(t * 2 + slotx) * N1 + k1] = v[r];
            }
          }
        }
        __syncthreads();
        if (s_row >= 1 && s_row <= RB - 2) {
          // best-bin tracking with selects; local maxima that can clear the
          // threshold are rare and handled one by one afterwards
          uint32_t mask = 0;
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int k1 = jj_row + 16 * r;
            const uint32_t k = (uint32_t)k2 + (uint32_t)N2 * (uint32_t)k1;
            const float yl = ldsf[(s_row - 1) * N1 + k1];
            const float yr = ldsf[(s_row + 1) * N1 + k1];
            const bool ok = full_band || in_range(k, a);
            const bool upd = ok && (m[r] > bm || (m[r] == bm && k < bk));
            bm = upd ? m[r] : bm;
            bk = upd ? k : bk;
            bx.x = upd ? v[r].x : bx.x;
            bx.y = upd ? v[r].y : bx.y;
            bl = upd ? yl : bl;
            br = upd ? yr : br;
            if (ok && m[r] > yl && m[r] > yr && m[r] >= lb) mask |= 1u << r;
          }
          while (mask) {
            const int r = __builtin_ctz(mask);
            mask &= mask - 1;
            const int k1 = jj_row + 16 * r;
            float2 x = v[0];
            float mc = m[0];
#pragma unroll
            for (int q = 1; q < 16; ++q) {
              x = (q == r) ? v[q] : x;
              mc = (q == r) ? m[q] : mc;
            }
            push_cand(cand, &sh.ncand, (uint32_t)C::CAND_CAP, (uint32_t)k2 + (uint32_t)N2 * (uint32_t)k1, x,
                      ldsf[(s_row - 1) * N1 + k1], mc, ldsf[(s_row + 1) * N1 + k1]);
          }
        }
        __syncthreads();
      }
    }

    if constexpr (peaks) {
      // phase 2b: tile-boundary rows from the edge buffers
      constexpr int NBND = 2 * NT2 * N1;
      for (int i = opaque_v(tid0); i < NBND; i += T) {
        const int k1 = i % N1;
        const int q = i / N1;            // 0 .. 2*NT2-1
        const int t = q >> 1;
        const bool last = (q & 1) != 0;  // row RB-1 of tile t, else row 0
        float yl, yc, yr;
        float2 x;
        uint32_t k2;
        if (!last) {
          k2 = (uint32_t)(t * RB);
          yc = gp(edgeM)[(t * 4 + 0) * N1 + k1];
          x = gp(edgeX)[(t * 2 + 0) * N1 + k1];
          yr = gp(edgeM)[(t * 4 + 1) * N1 + k1];
          if (t > 0) yl = gp(edgeM)[((t - 1) * 4 + 3) * N1 + k1];
          else yl = gp(edgeM)[((NT2 - 1) * 4 + 3) * N1 + (k1 + N1 - 1) % N1];  // k-1 wraps to row N2-1, column k1-1
        } else {
          k2 = (uint32_t)(t * RB + RB - 1);
          yc = gp(edgeM)[(t * 4 + 3) * N1 + k1];
          x = gp(edgeX)[(t * 2 + 1) * N1 + k1];
          yl = gp(edgeM)[(t * 4 + 2) * N1 + k1];
          if (t < NT2 - 1) yr = gp(edgeM)[((t + 1) * 4 + 0) * N1 + k1];
          else yr = gp(edgeM)[(0 * 4 + 0) * N1 + (k1 + 1) % N1];  // k+1 wraps to row 0, column k1+1
        }
        const uint32_t k = k2 + (uint32_t)N2 * (uint32_t)k1;
        if (full_band || in_range(k, a)) {
          if (yc > bm || (yc == bm && k < bk)) { bm = yc; bk = k; bx = x; bl = yl; br = yr; }
          if (yc > yl && yc > yr && yc >= lb) push_cand(cand, &sh.ncand, (uint32_t)C::CAND_CAP, k, x, yl, yc, yr);
        }
      }
      __syncthreads();

      // ------------------------------------------------------------------
      // phase 3: statistics, threshold, selection of the first K survivors
      // ------------------------------------------------------------------
      double s1 = fs1, s2 = fs2;
      float gm = bm;
      uint32_t gk = bk;
      block_stats<T>(s1, s2, gm, gk, sh.red);
      const double cnt = (double)(a.s1 - a.s0);
      float mu, sd;
      const float thr = threshold_of(a, s1, s2, cnt, mu, sd);
      if (a.gmax && gk != 0xffffffffu && bk == gk && bm == gm) write_record(a.gmax + och, bk, bx, bl, bm, br, a);

      const uint32_t ncand = min(sh.ncand, (uint32_t)C::CAND_CAP);
      select_first_k<T, C::BSH, (C::LDS > 4096 ? 4096 : C::LDS)>(a, CandGlobal{cand}, ncand, thr, mu, sd, P, och,
                                                                sh.hist,
                                                                reinterpret_cast<uint2*>(lds), sh.bstar,
                                                                sh.ngather, sh.red);
    }
    __syncthreads();
    }
}

}  // namespace spx
