// Transforms and peak searches of lengths N = 2^19 .. 2^24 (beyond the
// split four-step path, whose row and column lengths stop at 256 x 1024).
//
// The transform is a Stockham FFT over global memory: ceil(log2 N / 4)
// passes, radix 16 except a leftover radix (2, 4 or 8) in the last pass, one
// butterfly per thread, the window, zero padding and (16-bit storage) the
// per-channel load scale applied as the first pass reads the input, and the
// dense output kinds (spectrum, |X|, |X|^2 scaled,
// {|X|, arg}) written by the last pass.  Between passes the data go through a
// ping-pong workspace, so a transform moves 16 B per sample per pass: these
// lengths are served for completeness (the reference plans any power of two
// with hipFFT), not at the rate of the fused lengths.
//
// The peak searches (spx_fused_peaks at these lengths) run on the stored
// spectrum with the same rules and records as the fused kernels:
//   k_ls_stats   many blocks per channel: sums of |X| and |X|^2 over the
//                search range and the largest |X| (lowest bin on ties)
//   k_ls_select  one block per channel: statistics, threshold, global-maximum
//                record, then every bin in ascending order: strict local
//                maxima (circular neighbours) above the threshold are
//                counted and the first K written in order (block scan of
//                per-thread counts, a running total across tiles)
#include "spx_fused_common.hpp"

namespace spx {

struct GsArgs {
  FusedArgs a;             // input, window, n_point, dense outputs (first / last pass)
  const float2* src;       // [nch][N] (middle and last passes)
  float2* dst;             // [nch][N] (first and middle passes)
  const float2* tw;        // this pass's table [k][r] = W_(NS*R)^(r*k), or null when NS = 1
  uint32_t N, NS;
  uint64_t ch0;            // first caller channel of the chunk
};

// FIRST: read the caller's input (window, zero padding); LAST: write the dense outputs.
template <int R, bool FIRST, bool LAST>
__global__ void __launch_bounds__(256) k_gs_pass(GsArgs g) {
  const uint32_t N = g.N, NB = N / R;
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= NB) return;
  const uint64_t c = blockIdx.y;
  float2 v[R];
  if constexpr (FIRST) {
    const FusedArgs& a = g.a;
    const uint64_t ch = g.ch0 + c;
    const uint32_t es = a.dtype == SPX_DTYPE_C64 ? 8 : 4;
    const void* in = reinterpret_cast<const char*>(a.in) + ch * (uint64_t)a.n_point * es;
    const float sc = a.in_scale != nullptr ? gp(a.in_scale)[ch] : 1.0f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t n = j + (uint32_t)r * NB;
      const uint32_t ni = min(n, a.n_point - 1);
      const float w = a.window != nullptr ? gp(a.window)[ni] : 1.0f;
      const float2 x = a.dtype == SPX_DTYPE_C64 ? ld_raw<SPX_DTYPE_C64>(in, ni)
                       : a.dtype == SPX_DTYPE_C32F16 ? ld_raw<SPX_DTYPE_C32F16>(in, ni)
                                                     : ld_raw<SPX_DTYPE_C32BF16>(in, ni);
      v[r] = c_scale(x, n < a.n_point ? w * sc : 0.0f);
    }
  } else {
    const float2* src = g.src + c * (uint64_t)N;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = gp(src)[j + (uint32_t)r * NB];
  }
  const uint32_t NS = g.NS;
  const uint32_t k = j % NS;
  if (NS > 1) {
    float2 w[R];
    load_tw_row<R, false>(g.tw + (uint64_t)k * R, w);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], w[r]);
  }
  Dft<R>::run(v);
  const uint32_t d0 = (j / NS) * NS * R + k;
  if constexpr (LAST) {
    const uint64_t base = (g.ch0 + c) * (uint64_t)N;
#pragma unroll
    for (int r = 0; r < R; ++r) store_dense(g.a, base + d0 + (uint32_t)r * NS, v[r]);
  } else {
    float2* dst = g.dst + c * (uint64_t)N;
#pragma unroll
    for (int r = 0; r < R; ++r) gp(dst)[d0 + (uint32_t)r * NS] = v[r];
  }
}

template <int R>
static void launch_gs(const GsArgs& g, bool first, bool last, uint32_t nch, hipStream_t st) {
  const dim3 grid((g.N / R + 255) / 256, nch);
  if (first) hipLaunchKernelGGL((k_gs_pass<R, true, false>), grid, dim3(256), 0, st, g);
  else if (last) hipLaunchKernelGGL((k_gs_pass<R, false, true>), grid, dim3(256), 0, st, g);
  else hipLaunchKernelGGL((k_gs_pass<R, false, false>), grid, dim3(256), 0, st, g);
}

// ---------------------------------------------------------------------------
// peak search on a stored spectrum
// ---------------------------------------------------------------------------
struct LsPart {
  double s1, s2;  // sum |X|, sum |X|^2 over the block's in-range bins
  float m;        // largest |X|, -1 if none
  uint32_t k;     // its bin (lowest on ties)
};

constexpr int LS_T = 256;
constexpr int LS_BPT = 64;                    // bins per thread in k_ls_stats
constexpr int LS_BLK = LS_T * LS_BPT;         // bins per k_ls_stats block

__global__ void __launch_bounds__(LS_T) k_ls_stats(FusedArgs a, const float2* spec, LsPart* part, uint32_t nblk) {
  __shared__ BlockRed red;
  const uint64_t c = blockIdx.y;
  const uint32_t N = a.nfft;
  const float2* X = spec + c * (uint64_t)N;
  const uint32_t k0 = blockIdx.x * LS_BLK;
  float fs1 = 0.f, fs2 = 0.f, bm = -1.f;
  uint32_t bk = 0xffffffffu;
  for (int i = 0; i < LS_BPT; ++i) {
    const uint32_t k = k0 + (uint32_t)i * LS_T + threadIdx.x;  // coalesced; ascending per thread
    if (k < N && in_range(k, a)) {
      const float m = mag_of(gp(X)[k]);
      fs1 += m;
      fs2 = fmaf(m, m, fs2);
      if (m > bm) {
        bm = m;
        bk = k;
      }
    }
  }
  double s1 = fs1, s2 = fs2;
  block_stats<LS_T, false>(s1, s2, bm, bk, red);
  if (threadIdx.x == 0) {
    LsPart p;
    p.s1 = s1;
    p.s2 = s2;
    p.m = bm;
    p.k = bk;
    gp(part)[c * nblk + blockIdx.x] = p;
  }
}

// |X| of bin k of a channel's stored spectrum (k taken modulo N)
__device__ __forceinline__ float ls_mag(const float2* X, uint32_t k, uint32_t N) { return mag_of(gp(X)[k & (N - 1)]); }

__global__ void __launch_bounds__(LS_T) k_ls_select(FusedArgs a, const float2* spec, const LsPart* part,
                                                    uint32_t nblk) {
  __shared__ BlockRed red;
  const uint64_t c = blockIdx.x;  // channel of the chunk (the host offsets the output pointers)
  const uint32_t N = a.nfft;
  const float2* X = spec + c * (uint64_t)N;
  const int tid = threadIdx.x;
  // statistics and the global maximum from the block partials
  double s1 = 0.0, s2 = 0.0;
  float bm = -1.f;
  uint32_t bk = 0xffffffffu;
  for (uint32_t i = tid; i < nblk; i += LS_T) {
    const LsPart p = gp(part)[c * nblk + i];
    s1 += p.s1;
    s2 += p.s2;
    if (p.k != 0xffffffffu && (p.m > bm || (p.m == bm && p.k < bk))) {
      bm = p.m;
      bk = p.k;
    }
  }
  block_stats<LS_T, true>(s1, s2, bm, bk, red);
  const double cnt = (double)(a.s1 - a.s0);
  float mu, sd;
  const float thr = threshold_of(a, s1, s2, cnt, mu, sd);
  // global-maximum record
  if (a.gmax && tid == 0 && bk != 0xffffffffu)
    write_record(a.gmax + c, bk, gp(X)[bk], ls_mag(X, bk + N - 1, N), bm, ls_mag(X, bk + 1, N), a);
  // ascending sweep: 16 consecutive bins per thread per tile
  constexpr int E = 16;
  uint32_t total = 0;
  const uint32_t kcap = a.kcap;
  spx_max_value* out = a.peaks + c * (uint64_t)kcap;
  for (uint32_t base = 0; base < N; base += LS_T * E) {
    const uint32_t k0 = base + (uint32_t)tid * E;
    float m[E + 2];
#pragma unroll
    for (int e = 0; e < E + 2; ++e) m[e] = ls_mag(X, k0 + N - 1 + (uint32_t)e, N);
    uint32_t mask = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t k = k0 + e;
      const bool ok = k < N && in_range(k, a) && m[e + 1] > m[e] && m[e + 1] > m[e + 2] && m[e + 1] > thr;
      mask |= (ok ? 1u : 0u) << e;
    }
    int tot;
    const int off = block_excl_scan<LS_T>(__popc(mask), tot, red);
    uint32_t rank = total + (uint32_t)off;
    while (mask) {
      const int e = __builtin_ctz(mask);
      mask &= mask - 1;
      if (rank < kcap) {
        const uint32_t k = k0 + e;
        write_record(out + rank, k, gp(X)[k], m[e], m[e + 1], m[e + 2], a);
      }
      ++rank;
    }
    total += (uint32_t)tot;
  }
  if (tid == 0) {
    if (a.counts) {
      gp(a.counts)[2 * c + 0] = min(total, kcap);
      gp(a.counts)[2 * c + 1] = total;
    }
    if (a.stats) {
      gp(a.stats)[4 * c + 0] = mu;
      gp(a.stats)[4 * c + 1] = sd;
      gp(a.stats)[4 * c + 2] = thr;
      gp(a.stats)[4 * c + 3] = (float)(cnt > 0 ? s2 / cnt : 0.0);
    }
  }
}

}  // namespace spx

extern "C" {

int spx_large_is(uint32_t nfft) { return nfft >= (1u << 19) && nfft <= (1u << 24) && (nfft & (nfft - 1)) == 0; }

// Workspace of the large-length path for `chunk` channels: two ping-pong
// spectra, plus the peak-search partials.
uint64_t spx_large_workspace_bytes(uint32_t nfft, uint32_t chunk) {
  using namespace spx;
  const uint64_t nblk = (nfft + LS_BLK - 1) / LS_BLK;
  return 2ull * chunk * nfft * 8 + (uint64_t)chunk * nblk * sizeof(LsPart) + 512;
}

// Transform (and, for a->peak_mode != 0, the peak search) of a->batch
// channels of length a->nfft in chunks of `chunk` channels.  tw: the
// concatenated per-pass tables of length nfft (radix 16 first, the leftover
// radix last), as make_twiddles builds them.
hipError_t spx_launch_large(const FusedArgs* a0, const float2* tw, void* ws, uint32_t chunk, hipStream_t st) {
  using namespace spx;
  const uint32_t N = a0->nfft;
  if (!spx_large_is(N) || chunk == 0) return hipErrorInvalidValue;
  int lg = 0;
  while ((1u << lg) < N) ++lg;
  const int np = (lg + 3) / 4;
  const int rx = (lg % 4) ? (1 << (lg % 4)) : 16;
  const uint64_t nblk = (N + LS_BLK - 1) / LS_BLK;
  float2* buf[2] = {reinterpret_cast<float2*>(ws), reinterpret_cast<float2*>(ws) + (uint64_t)chunk * N};
  LsPart* part = reinterpret_cast<LsPart*>(reinterpret_cast<unsigned char*>(ws) + 2ull * chunk * N * 8);
  const bool peaks = a0->peak_mode != 0;
  for (uint32_t c0 = 0; c0 < a0->batch; c0 += chunk) {
    const uint32_t nch = min(chunk, a0->batch - c0);
    GsArgs g;
    g.a = *a0;
    g.ch0 = c0;
    g.N = N;
    // pass p reads buf[(p+1) % 2] and writes buf[p % 2]; for the peak search
    // the last pass writes the chunk's complex spectrum into the buffer it
    // does not read
    float2* spec = buf[(np - 1) % 2];
    if (peaks) {
      g.a.spec_out = spec - (uint64_t)c0 * N;
      g.a.mag_out = nullptr;
    }
    uint64_t ns = 1, off = 0;
    for (int p = 0; p < np; ++p) {
      const int R = (p == np - 1) ? rx : 16;
      g.NS = (uint32_t)ns;
      g.tw = p == 0 ? nullptr : tw + off;
      g.src = buf[(p + 1) % 2];
      g.dst = buf[p % 2];
      const bool first = p == 0, last = p == np - 1;
      switch (R) {
        case 2: launch_gs<2>(g, first, last, nch, st); break;
        case 4: launch_gs<4>(g, first, last, nch, st); break;
        case 8: launch_gs<8>(g, first, last, nch, st); break;
        default: launch_gs<16>(g, first, last, nch, st); break;
      }
      if (p > 0) off += ns * (uint64_t)R;
      ns *= (uint64_t)R;
    }
    if (peaks) {
      FusedArgs b = *a0;
      b.batch = nch;
      b.peaks = a0->peaks ? a0->peaks + (uint64_t)c0 * a0->kcap : nullptr;
      b.counts = a0->counts + 2ull * c0;
      b.gmax = a0->gmax ? a0->gmax + c0 : nullptr;
      b.stats = a0->stats ? a0->stats + 4ull * c0 : nullptr;
      hipLaunchKernelGGL(k_ls_stats, dim3((uint32_t)nblk, nch), dim3(LS_T), 0, st, b, spec, part, (uint32_t)nblk);
      hipLaunchKernelGGL(k_ls_select, dim3(nch), dim3(LS_T), 0, st, b, spec, part, (uint32_t)nblk);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // extern "C"
