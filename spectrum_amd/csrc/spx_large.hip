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
// spectrum with the same rules and records as the fused kernels, in four
// launches that each spread a channel over many workgroups except the two
// per-channel steps that touch a few values:
//   k_ls_stats   blocks of 16384 bins: sums of |X| and |X|^2 over the search
//                range and the largest |X| (lowest bin on ties)
//   k_ls_thresh  one block per channel: statistics, threshold, the
//                global-maximum record and the stats output
//   k_ls_cand    blocks of 16384 bins: strict local maxima (circular
//                neighbours) above the threshold, their bins in ascending
//                order per block (block scans of per-thread counts)
//   k_ls_gather  one block per channel: offsets of the blocks' lists, then
//                the first K bins overall as records, in ascending order
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

constexpr int LS_CAP = LS_BLK / 2;  // candidate slots per block: strict maxima are never adjacent

struct LsChan {
  float thr;
};

__global__ void __launch_bounds__(LS_T) k_ls_thresh(FusedArgs a, const float2* spec, const LsPart* part,
                                                    uint32_t nblk, LsChan* chs) {
  __shared__ BlockRed red;
  const uint64_t c = blockIdx.x;  // channel of the chunk (the host offsets the output pointers)
  const uint32_t N = a.nfft;
  const float2* X = spec + c * (uint64_t)N;
  const int tid = threadIdx.x;
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
  block_stats<LS_T, false>(s1, s2, bm, bk, red);
  const double cnt = (double)(a.s1 - a.s0);
  float mu, sd;
  const float thr = threshold_of(a, s1, s2, cnt, mu, sd);
  if (tid == 0) {
    gp(chs)[c].thr = thr;
    if (a.gmax && bk != 0xffffffffu)
      write_record(a.gmax + c, bk, gp(X)[bk], ls_mag(X, bk + N - 1, N), bm, ls_mag(X, bk + 1, N), a);
    if (a.stats) {
      gp(a.stats)[4 * c + 0] = mu;
      gp(a.stats)[4 * c + 1] = sd;
      gp(a.stats)[4 * c + 2] = thr;
      gp(a.stats)[4 * c + 3] = (float)(cnt > 0 ? s2 / cnt : 0.0);
    }
  }
}

// cnt[c * nblk + b]: the number of bins listed by block b; idx[(c * nblk +
// b) * LS_CAP + i]: its i-th bin, ascending
__global__ void __launch_bounds__(LS_T) k_ls_cand(FusedArgs a, const float2* spec, const LsChan* chs,
                                                  uint32_t nblk, uint32_t* cnt, uint32_t* idx) {
  __shared__ BlockRed red;
  const uint64_t c = blockIdx.y;
  const uint32_t N = a.nfft;
  const float2* X = spec + c * (uint64_t)N;
  const int tid = threadIdx.x;
  const float thr = gp(chs)[c].thr;
  const uint64_t slot = c * nblk + blockIdx.x;
  uint32_t* out = idx + slot * LS_CAP;
  constexpr int E = 16;
  uint32_t total = 0;
  for (uint32_t tile = 0; tile < (uint32_t)LS_BLK; tile += LS_T * E) {
    const uint32_t k0 = blockIdx.x * LS_BLK + tile + (uint32_t)tid * E;  // N is a multiple of LS_BLK
    float m[E + 2];
#pragma unroll
    for (int e = 0; e < E + 2; ++e) m[e] = ls_mag(X, k0 + N - 1 + (uint32_t)e, N);
    uint32_t mask = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const bool ok = in_range(k0 + e, a) && m[e + 1] > m[e] && m[e + 1] > m[e + 2] && m[e + 1] > thr;
      mask |= (ok ? 1u : 0u) << e;
    }
    int tot;
    uint32_t rank = total + (uint32_t)block_excl_scan<LS_T>(__popc(mask), tot, red);
    while (mask) {
      const int e = __builtin_ctz(mask);
      mask &= mask - 1;
      gp(out)[rank++] = k0 + e;
    }
    total += (uint32_t)tot;
  }
  if (tid == 0) gp(cnt)[slot] = total;
}

constexpr int LS_MAXBLK = (1 << 24) / LS_BLK;

__global__ void __launch_bounds__(LS_T) k_ls_gather(FusedArgs a, const float2* spec, const uint32_t* cnt,
                                                    const uint32_t* idx, uint32_t nblk) {
  __shared__ BlockRed red;
  __shared__ uint32_t off[LS_MAXBLK];
  const uint64_t c = blockIdx.x;
  const uint32_t N = a.nfft;
  const float2* X = spec + c * (uint64_t)N;
  const int tid = threadIdx.x;
  // exclusive prefix of the block counts: thread t owns blocks [t*PER, (t+1)*PER)
  constexpr int PER = LS_MAXBLK / LS_T;
  uint32_t own[PER], run = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t b = (uint32_t)tid * PER + i;
    own[i] = b < nblk ? gp(cnt)[c * nblk + b] : 0u;
    run += own[i];
  }
  int tot;
  uint32_t base = (uint32_t)block_excl_scan<LS_T>((int)run, tot, red);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t b = (uint32_t)tid * PER + i;
    if (b < nblk) off[b] = base;
    base += own[i];
  }
  __syncthreads();
  const uint32_t total = (uint32_t)tot, kcap = a.kcap;
  const uint32_t nout = min(total, kcap);
  spx_max_value* out = a.peaks + c * (uint64_t)kcap;
  for (uint32_t s = (uint32_t)tid; s < nout; s += LS_T) {
    // the listing block: the last b with off[b] <= s (blocks without bins
    // share their successor's offset and are passed over)
    uint32_t b = 0;
    for (uint32_t step = LS_MAXBLK / 2; step > 0; step >>= 1)
      if (b + step < nblk && off[b + step] <= s) b += step;
    const uint32_t k = gp(idx)[(c * nblk + b) * LS_CAP + (s - off[b])];
    write_record(out + s, k, gp(X)[k], ls_mag(X, k + N - 1, N), ls_mag(X, k, N), ls_mag(X, k + 1, N), a);
  }
  if (tid == 0 && a.counts) {
    gp(a.counts)[2 * c + 0] = nout;
    gp(a.counts)[2 * c + 1] = total;
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
  return 2ull * chunk * nfft * 8 + (uint64_t)chunk * nblk * sizeof(LsPart) + (uint64_t)chunk * sizeof(LsChan) +
         (uint64_t)chunk * nblk * 4 * (1 + LS_CAP) + 1024;
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
  unsigned char* tail = reinterpret_cast<unsigned char*>(ws) + 2ull * chunk * N * 8;
  LsPart* part = reinterpret_cast<LsPart*>(tail);
  tail += (uint64_t)chunk * nblk * sizeof(LsPart);
  LsChan* chs = reinterpret_cast<LsChan*>(tail);
  tail += ((uint64_t)chunk * sizeof(LsChan) + 255) & ~255ull;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(tail);
  tail += ((uint64_t)chunk * nblk * 4 + 255) & ~255ull;
  uint32_t* idx = reinterpret_cast<uint32_t*>(tail);
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
      hipLaunchKernelGGL(k_ls_thresh, dim3(nch), dim3(LS_T), 0, st, b, spec, part, (uint32_t)nblk, chs);
      hipLaunchKernelGGL(k_ls_cand, dim3((uint32_t)nblk, nch), dim3(LS_T), 0, st, b, spec, chs, (uint32_t)nblk, cnt,
                         idx);
      hipLaunchKernelGGL(k_ls_gather, dim3(nch), dim3(LS_T), 0, st, b, spec, cnt, idx, (uint32_t)nblk);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // extern "C"
