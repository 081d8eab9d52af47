// Fused window -> FFT -> |X| -> statistics -> local maxima -> threshold ->
// ascending compaction -> sub-bin refinement, for gfx950.
//
// One workgroup per channel (no cross-workgroup synchronisation):
//
//  * N <= 16384: the whole channel lives in LDS.  The first Stockham
//    pass reads the input straight from HBM in butterfly order (window applied
//    in registers), later passes go through LDS, the last pass stays in VGPRs.
//    |X|, mean/std, local maxima, the block scan and the records are produced
//    from the LDS copy of the spectrum; nothing but records leaves the chip.
//
// Longer channels (N = 256 * N2 >= 32768) run on the split four-step path
// (spx_fourstep.hip).
//
// Peak rules (peak_mode 1, the fused north-star path; no reference kernel):
//   m[k] = |X[k]| (FP32), range [s0,s1) (default the whole band, circular
//   neighbours), mean/std over the range (population std), T = mean + nsigma*std,
//   peak iff m[k] > m[k-1] && m[k] > m[k+1] && m[k] > T, first K in ascending k,
//   parabolic delta = 0.5(yl-yr)/(yl-2yc+yr) (|den| > 1e-10, clamped to +-0.5)
//   as in include/spectrum/kernels/fft_kernel_sources_rocm.hpp:228-239.
// peak_mode 2 is the AllMaxima rule set (all_maxima_kernel_sources_rocm.hpp:57-87,
// 205-264): no threshold, record freq_offset 0 and f = k*fs/N, count uncapped.
// peak_modes 3 / 4 are the reference ONE_PEAK / TWO_PEAKS searches
// (fft_kernel_sources_rocm.hpp:122-519) taken on the spectrum the kernel has
// just computed, without storing it.
#include "spx_fused_common.hpp"
#include "spx_mlp.hpp"

namespace spx {

// ===========================================================================
// small path
// ===========================================================================
template <int L>
struct SmallCfg {
  static constexpr int T = (L / 16 > 64) ? L / 16 : 64;
  static constexpr int PPT = (L >= 16) ? 16 : L;
  static constexpr int E = (L >= T) ? L / T : 1;   // bins per owner thread in the peak stage
  static constexpr int OWN = L / E;                 // owner threads
  static constexpr int WPE = 4;
  using LAY = LayContig<L, 1>;
};

// LEARN: the records get the parabolic offset and the features of every
// stored peak go to a.mlp_feat; the learned refinement (k_mlp_refine) runs
// after the kernel and replaces the offset where its prediction is valid.
template <int L, int MODE, bool LEARN = false>
__global__ void __launch_bounds__(SmallCfg<L>::T) __attribute__((amdgpu_waves_per_eu(SmallCfg<L>::WPE))) k_fused_small(FusedArgs a) {
  using C = SmallCfg<L>;
  using PL = Plan<L, 0>;
  using LAY = typename C::LAY;
  constexpr int T = C::T, PPT = C::PPT;
  constexpr int R0 = PL::R0, RL = PL::RL;
  __shared__ float2 lds[LAY::SIZE];
  __shared__ BlockRed red;
  const int tid = threadIdx.x;

  // one channel per workgroup: a grid-stride loop here would let the compiler
  // hoist every (channel-invariant) twiddle and window load out of the loop
  // and keep them all live in VGPRs.
  {
    const uint32_t ch = blockIdx.x;
    if (ch >= a.batch) return;
    float2 v[PPT];
    // ---- load in first-pass butterfly order: x[jj + r*L/R0]
#pragma unroll
    for (int b = 0; b < PPT / R0; ++b) {
      const int j = tid + b * T;
      if (j < L / R0) {
        load_run<R0, MODE == 0>(a, ch, (uint32_t)j, (uint32_t)(L / R0), &v[b * R0]);
      }
    }
    fft_tile_pf<LAY, L, 1, T, PPT>(lds, v, a.tw_a, tid);
    // v[b*RL + r] = X[j + r*L/RL], j = tid + b*T < L/RL

    const uint64_t cbase = (uint64_t)ch * L;
    if constexpr (MODE == 0) {
#pragma unroll
      for (int b = 0; b < PPT / RL; ++b) {
        const int j = tid + b * T;
        if (j < L / RL) {
#pragma unroll
          for (int r = 0; r < RL; ++r) store_dense(a, cbase + j + r * (L / RL), v[b * RL + r]);
        }
      }
    }

    if constexpr (MODE == 1) {
      __syncthreads();
#pragma unroll
      for (int b = 0; b < PPT / RL; ++b) {
        const int j = tid + b * T;
        if (j < L / RL) {
#pragma unroll
          for (int r = 0; r < RL; ++r) lds[LAY::idx(0, j + r * (L / RL))] = v[b * RL + r];
        }
      }
      __syncthreads();

      constexpr int E = C::E;
      float m[E];
      double s1 = 0.0, s2 = 0.0;
      float fs1 = 0.f, fs2 = 0.f;
      float bm = -1.0f;
      uint32_t bk = 0xffffffffu;
      const bool owner = tid < C::OWN;
      const uint32_t k0 = (uint32_t)tid * E;
      // whole band (the default): no per-bin range test
      const bool full = a.s0 == 0 && a.s1 >= (uint32_t)L;
      float ml = 0.f, mr = 0.f;
      if constexpr (E == 16) {
        // every thread owns a run of 16 bins, which shares one pad slot and
        // so is contiguous in LDS; the run and both circular neighbours are
        // read before any of them is used
        const float2* run = lds + LAY::idx(0, k0);
        float2 xv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) xv[e] = run[e];
        const float2 xl = lds[LAY::idx(0, (k0 + L - 1) % L)];
        const float2 xr = lds[LAY::idx(0, (k0 + E) % L)];
#pragma unroll
        for (int e = 0; e < E; ++e) m[e] = mag_of(xv[e]);
        if (full && a.gmax == nullptr) {
          // no global-maximum record asked for: sums only
#pragma unroll
          for (int e = 0; e < E; ++e) {
            fs1 += m[e];
            fs2 = fmaf(m[e], m[e], fs2);
          }
        } else if (full) {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            fs1 += m[e];
            fs2 = fmaf(m[e], m[e], fs2);
            const bool up = m[e] > bm;
            bm = up ? m[e] : bm;
            bk = up ? k0 + e : bk;
          }
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const bool in = in_range(k0 + e, a);
            const float mv = in ? m[e] : 0.0f;
            fs1 += mv;
            fs2 = fmaf(mv, mv, fs2);
            const bool up = in && m[e] > bm;
            bm = up ? m[e] : bm;
            bk = up ? k0 + e : bk;
          }
        }
        ml = mag_of(xl);
        mr = mag_of(xr);
      } else if (owner) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const uint32_t k = k0 + e;
          m[e] = mag_of(lds[LAY::idx(0, k)]);
          if (in_range(k, a)) {
            fs1 += m[e];
            fs2 = fmaf(m[e], m[e], fs2);
            if (m[e] > bm) { bm = m[e]; bk = k; }
          }
        }
        ml = mag_of(lds[LAY::idx(0, (k0 + L - 1) % L)]);
        mr = mag_of(lds[LAY::idx(0, (k0 + E) % L)]);
      }
      s1 = fs1;
      s2 = fs2;
      // one channel per workgroup: `red` is written once per field, so
      // neither reduction needs its closing barrier
      if (a.gmax != nullptr) block_stats<T, false>(s1, s2, bm, bk, red);
      else block_sums<T, false>(s1, s2, red);
      const double cnt = (double)(a.s1 - a.s0);
      float mu, sd;
      const float thr = threshold_of(a, s1, s2, cnt, mu, sd);

      // local maxima of the owned run (circular neighbours)
      uint32_t flags = 0;
      if (owner) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const float yl = (e == 0) ? ml : m[e - 1];
          const float yr = (e == E - 1) ? mr : m[e + 1];
          const bool pk = m[e] > yl && m[e] > yr && m[e] > thr;
          flags |= pk ? (1u << e) : 0u;
        }
        if (!full) {
#pragma unroll
          for (int e = 0; e < E; ++e) flags &= in_range(k0 + e, a) ? ~0u : ~(1u << e);
        }
      }
      const int c = __builtin_popcount(flags);
      int total;
      int off = block_excl_scan<T, false>(c, total, red);
      if (owner && flags && off < (int)a.kcap) {
        MaxValue* out = a.peaks + (uint64_t)ch * a.kcap;
        int pos = off;
        while (flags && pos < (int)a.kcap) {
          const int e = __builtin_ctz(flags);
          flags &= flags - 1;
          const uint32_t k = tid * E + e;
          const float yc = mag_of(lds[LAY::idx(0, k)]);
          const float yl = mag_of(lds[LAY::idx(0, (k + L - 1) % L)]);
          const float yr = mag_of(lds[LAY::idx(0, (k + 1) % L)]);
          write_record(out + pos, k, lds[LAY::idx(0, k)], yl, yc, yr, a);
          if constexpr (LEARN)
            mlp_store_features([&](uint32_t q) { return lds[LAY::idx(0, q)]; }, k, L, mlp_feat_row(a, ch, pos));
          ++pos;
        }
      }
      if (tid == 0) {
        if (a.counts) {
          gp(a.counts)[2 * ch + 0] = (uint32_t)min(total, (int)a.kcap);
          gp(a.counts)[2 * ch + 1] = (uint32_t)total;
        }
        if (a.stats) {
          double pe = cnt > 0 ? s2 / cnt : 0.0;
          gp(a.stats)[4 * ch + 0] = mu;
          gp(a.stats)[4 * ch + 1] = sd;
          gp(a.stats)[4 * ch + 2] = thr;
          gp(a.stats)[4 * ch + 3] = (float)pe;
        }
        if (a.gmax && bk != 0xffffffffu) {
          const float yl = mag_of(lds[LAY::idx(0, (bk + L - 1) % L)]);
          const float yr = mag_of(lds[LAY::idx(0, (bk + 1) % L)]);
          write_record(a.gmax + ch, bk, lds[LAY::idx(0, bk)], yl, bm, yr, a);
        }
      }
    }

    if constexpr (MODE == 3) {
      // reference ONE_PEAK / TWO_PEAKS on the spectrum in LDS: argmax of |X|
      // over [0, half) and [nfft - half, nfft) (a.s0 = half, a.s1 = nfft -
      // half), lowest bin on ties, then the records of write_top_records
      __syncthreads();
#pragma unroll
      for (int b = 0; b < PPT / RL; ++b) {
        const int j = tid + b * T;
        if (j < L / RL) {
#pragma unroll
          for (int r = 0; r < RL; ++r) lds[LAY::idx(0, j + r * (L / RL))] = v[b * RL + r];
        }
      }
      __syncthreads();
      constexpr int E = C::E;
      float lm = -1.f, rm = -1.f;
      uint32_t li = 0xffffffffu, ri = 0xffffffffu;
      if (tid < C::OWN) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const uint32_t k = tid * E + e;
          const float m = cabs_rn(lds[LAY::idx(0, k)]);
          if (k < a.s0 && m > lm) { lm = m; li = k; }
          if (k >= a.s1 && m > rm) { rm = m; ri = k; }
        }
      }
      double d0 = 0.0, d1 = 0.0;
      block_stats<T, true>(d0, d1, lm, li, red);
      block_stats<T, false>(d0, d1, rm, ri, red);
      if (tid == 0)
        write_top_records(a.peaks + (uint64_t)ch * (a.peak_mode == 3 ? 4 : 8), LdsBins<LAY>{lds}, a, lm, li, rm, ri);
    }
  }
}

// ===========================================================================
// dispatch
// ===========================================================================
template <int L>
static hipError_t launch_small(const FusedArgs& a, int grid, hipStream_t st) {
  if constexpr (L >= 1024) {
    if (a.peak_mode == 1 && a.refine == 2) {
      hipLaunchKernelGGL((k_fused_small<L, 1, true>), dim3(grid), dim3(SmallCfg<L>::T), 0, st, a);
      return hipGetLastError();
    }
  }
  if (a.peak_mode >= 3)
    hipLaunchKernelGGL((k_fused_small<L, 3>), dim3(grid), dim3(SmallCfg<L>::T), 0, st, a);
  else if (a.peak_mode != 0)
    hipLaunchKernelGGL((k_fused_small<L, 1>), dim3(grid), dim3(SmallCfg<L>::T), 0, st, a);
  else
    hipLaunchKernelGGL((k_fused_small<L, 0>), dim3(grid), dim3(SmallCfg<L>::T), 0, st, a);
  return hipGetLastError();
}

}  // namespace spx

extern "C" {

int spx_fused_is_large(uint32_t nfft) { return nfft >= (1u << 15) ? 1 : 0; }

// concatenated per-pass twiddle table length (float2 entries) of plan (L, V)
int spx_twiddle_table_len(uint32_t L, int V) {
  using namespace spx;
  switch (L) {
#define SPX_TWL(x) case x: return V ? Plan<x, 1>::TWSIZE : Plan<x, 0>::TWSIZE;
    SPX_TWL(2) SPX_TWL(4) SPX_TWL(8) SPX_TWL(16) SPX_TWL(32) SPX_TWL(64) SPX_TWL(128) SPX_TWL(256) SPX_TWL(512)
    SPX_TWL(1024) SPX_TWL(2048) SPX_TWL(4096) SPX_TWL(8192) SPX_TWL(16384)
#undef SPX_TWL
    default: return -1;
  }
}

hipError_t spx_launch_fused(const FusedArgs* a, int grid, hipStream_t st) {
  using namespace spx;
  switch (a->nfft) {
    case 1u << 1: return launch_small<2>(*a, grid, st);
    case 1u << 2: return launch_small<4>(*a, grid, st);
    case 1u << 3: return launch_small<8>(*a, grid, st);
    case 1u << 4: return launch_small<16>(*a, grid, st);
    case 1u << 5: return launch_small<32>(*a, grid, st);
    case 1u << 6: return launch_small<64>(*a, grid, st);
    case 1u << 7: return launch_small<128>(*a, grid, st);
    case 1u << 8: return launch_small<256>(*a, grid, st);
    case 1u << 9: return launch_small<512>(*a, grid, st);
    case 1u << 10: return launch_small<1024>(*a, grid, st);
    case 1u << 11: return launch_small<2048>(*a, grid, st);
    case 1u << 12: return launch_small<4096>(*a, grid, st);
    case 1u << 13: return launch_small<8192>(*a, grid, st);
    case 1u << 14: return launch_small<16384>(*a, grid, st);
    default: return hipErrorInvalidValue;
  }
}

}  // extern "C"
