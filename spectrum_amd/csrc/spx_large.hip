// Transforms and peak searches of lengths N = 2^19 .. 2^24 (beyond the
// split four-step path, whose row and column lengths stop at 256 x 1024).
//
// The transform is a Stockham FFT over global memory in three passes: a
// register pass of radix R0 = 2^r0 (one butterfly per thread; it applies the
// window, zero padding and, for 16-bit storage, the per-channel load scale
// as it reads the input), then two passes of radix L1 and L2 (256, 512 or
// 1024 points, N = R0 L1 L2) whose butterflies are whole L-point transforms
// in an LDS tile: a workgroup takes 16 or 32 consecutive butterflies j (the
// columns of the tile, so every load and store moves runs of 16 or 32
// consecutive complex values), multiplies element e of butterfly j by the Stockham
// twiddle W_(NS L)^(e (j mod NS)), transforms the 16 columns with the
// LDS-tile FFT of the split path's column kernel and stores element q at
// (j div NS) NS L + (j mod NS) + q NS.  The last pass writes the dense output
// kinds (spectrum, |X|, |X|^2 scaled, {|X|, arg}).  Between passes the data
// go through a ping-pong workspace: 16 B per sample per pass, three passes.
// A thread holds the elements e = jj + r L/16 (r < 16) of one column, so the
// twiddle splits as W_(NS L)^(jj k) W_(16 NS)^(r k) (k = j mod NS): the first
// factor from a [jj][k] table, the second from the powers 1, 2, 3, 4, 8 and
// 12 of W_(16 NS)^k (an [i][k] table) and their products; both tables are
// laid out so that the columns of a workgroup read consecutive entries.
//
// The peak searches (spx_fused_peaks at these lengths) run on the stored
// spectrum with the same rules and records as the fused kernels: the
// transform's last pass also writes per-workgroup statistics, then three
// launches, the candidate search spread over many workgroups per channel and
// the two per-channel steps touching a few values:
//   (last pass)  per workgroup: sums of |X| and |X|^2 over the search range
//                and the largest |X| (lowest bin on ties)
//   k_ls_thresh  one block per channel: statistics, threshold, the
//                global-maximum record and the stats output
//   k_ls_cand    blocks of 16384 bins: strict local maxima (circular
//                neighbours) above the threshold, their bins in ascending
//                order per block (block scans of per-thread counts)
//   k_ls_gather  one block per channel: offsets of the blocks' lists, then
//                the first K bins overall as records, in ascending order
#include "spx_fused_common.hpp"
#include "spx_mlp.hpp"

#include <cmath>

namespace spx {

struct GsArgs {
  FusedArgs a;             // input, window, n_point
  float2* dst;             // [nch][N]
  uint32_t N;
  uint64_t ch0;            // first caller channel of the chunk
};

// First pass, radix R (sub-transform length 1, no twiddles): butterfly j
// reads the caller's input x[j + r N/R] (window, zero padding, load scale)
// and its outputs go to y[j R + q].  The 256 butterflies of a workgroup
// cover one contiguous run of 256 R outputs, stored through LDS so that
// consecutive lanes write consecutive values.
template <int R, int DT>
__global__ void __launch_bounds__(256) k_gs_first(GsArgs g) {
  constexpr int T = 256, E = T * R;
  constexpr uint32_t ES = DT == SPX_DTYPE_C64 ? 8 : 4;
  __shared__ float2 xs[E + E / 16];  // element i at i + i/16
  const uint32_t N = g.N, NB = N / R;
  const int tid = threadIdx.x;
  const uint32_t j = blockIdx.x * T + (uint32_t)tid;  // NB is a multiple of T
  const uint64_t c = blockIdx.y;
  const FusedArgs& a = g.a;
  const uint64_t ch = g.ch0 + c;
  const void* in = reinterpret_cast<const char*>(a.in) + ch * (uint64_t)a.n_point * ES;
  const uint32_t last = a.n_point - 1;
  // every load is issued before any value is used (no branch between them:
  // a per-sample branch made each load wait for the previous one)
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = ld_raw<DT, true>(in, min(j + (uint32_t)r * NB, last));
  const float sc = a.in_scale != nullptr ? gp(a.in_scale)[ch] : 1.0f;
  // the window from its FP64-generated table, read after the signal loads
  // are in flight (evaluating it from its cosine coefficients with the
  // hardware cosine measured 1-4 % faster from 2^22 up, but 2-6e-7 off the
  // FP64 window instead of the table's rounding of 3e-8)
  float w[R];
  if (a.window != nullptr) {  // uniform
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = gp(a.window)[min(j + (uint32_t)r * NB, last)];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = 1.0f;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = c_scale(v[r], j + (uint32_t)r * NB < a.n_point ? w[r] * sc : 0.0f);
  Dft<R>::run(v);
  // y[j R + q] = v[q]: local element tid R + q
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = tid * R + q;
    xs[i + (i >> 4)] = v[q];
  }
  __syncthreads();
  float2* dst = g.dst + c * (uint64_t)N + (uint64_t)blockIdx.x * E;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int i = tid + m * T;
    gp(dst)[i] = xs[i + (i >> 4)];
  }
}

template <int R>
static void launch_gs_first(const GsArgs& g, uint32_t nch, hipStream_t st) {
  const dim3 grid(g.N / R / 256, nch);
  if (g.a.dtype == SPX_DTYPE_C64) hipLaunchKernelGGL((k_gs_first<R, SPX_DTYPE_C64>), grid, dim3(256), 0, st, g);
  else if (g.a.dtype == SPX_DTYPE_C32F16)
    hipLaunchKernelGGL((k_gs_first<R, SPX_DTYPE_C32F16>), grid, dim3(256), 0, st, g);
  else hipLaunchKernelGGL((k_gs_first<R, SPX_DTYPE_C32BF16>), grid, dim3(256), 0, st, g);
}

// ---------------------------------------------------------------------------
// Stockham pass of radix L through an LDS tile
// ---------------------------------------------------------------------------
// Partial statistics of a block of bins for the threshold search.
struct LsPart {
  double s1, s2;  // sum |X|, sum |X|^2 over the block's in-range bins
  float m;        // largest |X|, -1 if none
  uint32_t k;     // its bin (lowest on ties)
};

struct GlArgs {
  FusedArgs a;             // the dense outputs (LAST)
  const float2* src;       // [nch][N]
  float2* dst;             // [nch][N] (not LAST)
  const float2* tw_tile;   // plan (L, 0) tables of the L-point transform
  const float2* ta;        // [jj][k] = W_(NS L)^(jj k), jj < L/16, k < NS
  const float2* tb;        // [i][k] = W_(16 NS)^(m_i k), m = 1, 2, 3, 4, 8, 12
  uint32_t N, NS;          // NS: product of the earlier radices, a power of two >= 2
  uint64_t ch0;            // LAST: channel index of the chunk's first channel in the dense outputs
  LsPart* part;            // LAST, threshold search: [nch][gridDim.x] statistics of the workgroups' bins
};

// CB columns (consecutive butterflies) per workgroup of T = CB L/16
// threads, 16 elements per thread.
template <int L, int CB, bool LAST>
__global__ void __launch_bounds__(CB * L / 16) k_gl_pass(GlArgs g) {
  constexpr int T = CB * L / 16;
  using LAY = LayInter<L, CB>;
  using PL = Plan<L, 0>;
  constexpr int RL = PL::RL;
  static_assert(PL::R0 == 16, "the tile's first pass takes the 16 elements a thread loads");
  __shared__ __align__(16) float2 tile[LAY::SIZE];
  __shared__ __align__(16) float2 twl[PL::TWSIZE_LDS];
  const int tid = threadIdx.x;
  copy_tw_to_lds<L, 0, T>(g.tw_tile, twl, tid);
  const uint32_t N = g.N, NS = g.NS, M = N / L;
  const uint64_t c = blockIdx.y;
  // column s = butterfly j; the thread's elements e = jj + r L/16 (the
  // tile's first-pass inputs for thread tid)
  const int s = tid % CB, jj = tid / CB;
  const uint32_t j = blockIdx.x * (uint32_t)CB + (uint32_t)s;
  const uint32_t k = j & (NS - 1);
  const float2* src = g.src + c * (uint64_t)N;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = ld_f2<true>(src + j + (uint32_t)(jj + r * (L / 16)) * M);
  {
    // W_(NS L)^(e k) = W_(NS L)^(jj k) U^r, U = W_(16 NS)^k; U^r from
    // U, U^2, U^3, U^4, U^8, U^12 (at most one product)
    const float2 wj = gp(g.ta)[(uint32_t)jj * NS + k];
    float2 u[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) u[i] = gp(g.tb)[(uint32_t)i * NS + k];
    v[0] = c_mul(v[0], wj);
#pragma unroll
    for (int r = 1; r < 16; ++r) {
      const int lo = r & 3, hi = r >> 2;
      const float2 ur = hi == 0 ? u[lo - 1] : (lo == 0 ? u[2 + hi] : c_mul(u[lo - 1], u[2 + hi]));
      v[r] = c_mul(v[r], c_mul(wj, ur));
    }
  }
  fft_tile<LAY, L, CB, T, 16, 0, true>(tile, v, twl, tid);
  // v[b RL + r] = X[s][jj + b L/16 + r L/RL]  (butterfly tid + b T of the
  // last tile pass: column s, index jj + b T/CB)
  const uint32_t base = (j & ~(NS - 1)) * (uint32_t)L + k;
#pragma unroll
  for (int b = 0; b < 16 / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const uint32_t q = (uint32_t)(jj + b * (L / 16) + r * (L / RL));
      if constexpr (LAST) store_dense<true>(g.a, (g.ch0 + c) * (uint64_t)N + base + q * NS, v[b * RL + r]);
      else gp(g.dst)[c * (uint64_t)N + base + q * NS] = v[b * RL + r];
    }
  if constexpr (LAST) {
    if (g.part != nullptr) {  // uniform
      // sums of |X|, |X|^2 and the largest |X| (lowest bin on ties) over the
      // workgroup's in-range bins (in the last pass NS = N/L and base = j:
      // the bins j + q N/L of its CB consecutive j)
      float fs1 = 0.f, fs2 = 0.f, bm = -1.f;
      uint32_t bk = 0xffffffffu;
#pragma unroll
      for (int b = 0; b < 16 / RL; ++b)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const uint32_t bin = base + (uint32_t)(jj + b * (L / 16) + r * (L / RL)) * NS;
          if (in_range(bin, g.a)) {
            const float m = mag_of(v[b * RL + r]);
            fs1 += m;
            fs2 = fmaf(m, m, fs2);
            if (m > bm || (m == bm && bin < bk)) {
              bm = m;
              bk = bin;
            }
          }
        }
      __shared__ BlockRed red;
      double s1 = fs1, s2 = fs2;
      block_stats<T, false>(s1, s2, bm, bk, red);
      if (tid == 0) {
        LsPart p;
        p.s1 = s1;
        p.s2 = s2;
        p.m = bm;
        p.k = bk;
        gp(g.part)[c * gridDim.x + blockIdx.x] = p;
      }
    }
  }
}

template <int L, int CB>
static void launch_gl(const GlArgs& g, bool last, uint32_t nch, hipStream_t st) {
  const dim3 grid(g.N / L / CB, nch);
  if (last) hipLaunchKernelGGL((k_gl_pass<L, CB, true>), grid, dim3(CB * L / 16), 0, st, g);
  else hipLaunchKernelGGL((k_gl_pass<L, CB, false>), grid, dim3(CB * L / 16), 0, st, g);
}

// columns per workgroup of the 2^lg-point LDS pass, measured (512 MB of
// input, complex spectrum): 256 points 32 columns (256-byte runs, two
// workgroups of 66 KiB per CU; 2^20: 0.595 -> 0.559 ms against 16 columns,
// 0.609 ms with 64), 512 points 16 (32 columns: one workgroup per CU,
// slower), 1024 points 16 (one workgroup per CU; 8 columns, two per CU with
// 64-byte runs, were slower: the two passes at 2^24 took 356 and 419 us
// against 270 and 302 us)
static inline uint32_t gl_cols(int lg) { return lg == 8 ? 32u : 16u; }
static void launch_gl_any(int lg, const GlArgs& g, bool last, uint32_t nch, hipStream_t st) {
  if (lg == 8) launch_gl<256, 32>(g, last, nch, st);
  else if (lg == 9) launch_gl<512, 16>(g, last, nch, st);
  else launch_gl<1024, 16>(g, last, nch, st);
}

// ---------------------------------------------------------------------------
// peak search on a stored spectrum
// ---------------------------------------------------------------------------
constexpr int LS_T = 256;
constexpr int LS_BLK = 16384;                 // bins per k_ls_cand block

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
  constexpr int E = 16, TB = LS_T * E;
  // |X| of the tile's bins t0 - 1 .. t0 + TB (circular), read with
  // consecutive lanes on consecutive bins; value i at i + i/16
  __shared__ float ms[TB + 2 + (TB + 2) / 16];
  uint32_t total = 0;
  for (uint32_t tile = 0; tile < (uint32_t)LS_BLK; tile += TB) {
    const uint32_t t0 = blockIdx.x * LS_BLK + tile;  // N is a multiple of LS_BLK
    {
      // all E loads in flight before the first use (non-temporal: the
      // spectrum is read once here), then the two values past the tile
      float2 xv[E];
#pragma unroll
      for (int e = 0; e < E; ++e) xv[e] = ld_f2<true>(X + ((t0 + N - 1 + (uint32_t)(tid + e * LS_T)) & (N - 1)));
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = tid + e * LS_T;
        ms[i + (i >> 4)] = mag_of(xv[e]);
      }
      if (tid < 2) {
        const int i = TB + tid;
        ms[i + (i >> 4)] = ls_mag(X, t0 + N - 1 + (uint32_t)i, N);
      }
    }
    __syncthreads();
    const uint32_t k0 = t0 + (uint32_t)tid * E;
    float m[E + 2];
#pragma unroll
    for (int e = 0; e < E + 2; ++e) {
      const int i = tid * E + e;
      m[e] = ms[i + (i >> 4)];
    }
    uint32_t mask = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const bool ok = in_range(k0 + e, a) && m[e + 1] > m[e] && m[e + 1] > m[e + 2] && m[e + 1] > thr;
      mask |= (ok ? 1u : 0u) << e;
    }
    int tot;
    // the scan's barriers also end every read of ms before the next tile
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
    // learned refinement: the record's complex neighbourhood from the stored
    // spectrum, for k_mlp_refine after this launch
    if (a.mlp_feat != nullptr)
      mlp_store_features([&](uint32_t i) { return gp(X)[i]; }, k, N, mlp_feat_row(a, c, (int)s));
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
// spectra, plus the peak-search partials (statistics of the last pass's
// workgroups, at most N/4096 per channel; candidate lists of the blocks of
// LS_BLK bins) and, for learned refinement, 32 bytes of features per record
// slot (learn_kcap slots per channel, 0 without).
uint64_t spx_large_workspace_bytes(uint32_t nfft, uint32_t chunk, uint32_t learn_kcap) {
  using namespace spx;
  const uint64_t nblk = (nfft + LS_BLK - 1) / LS_BLK;
  return 2ull * chunk * nfft * 8 + (uint64_t)chunk * (nfft / 4096) * sizeof(LsPart) + (uint64_t)chunk * sizeof(LsChan) +
         (uint64_t)chunk * nblk * 4 * (1 + LS_CAP) + (uint64_t)chunk * learn_kcap * 32 + 1024;
}

// Radix plan of N = 2^lg: the register pass 2^r0, then np LDS passes of
// 2^lgs[i] points (r0 + sum lgs = lg).  Three passes at every length: four
// passes at 2^23 and 2^24 (16 x 256 x 128 x 16, 16 x 256 x 256 x 16, the
// last a 16-point pass over 256 columns) instead of the one-workgroup-per-CU
// 1024-point passes measured slower (512 MiB: complex spectrum 0.846 and
// 0.864 ms against 0.744 and 0.803 ms).
int spx_large_plan(uint32_t nfft, int* r0, int* lgs, int* np) {
  if (!spx_large_is(nfft)) return -1;
  int lg = 0;
  while ((1u << lg) < nfft) ++lg;
  static const int tab[6][5] = {{3, 2, 8, 8, 0}, {4, 2, 8, 8, 0}, {4, 2, 9, 8, 0},
                                {4, 2, 9, 9, 0}, {4, 2, 10, 9, 0}, {4, 2, 10, 10, 0}};
  const int* t = tab[lg - 19];
  *r0 = t[0];
  *np = t[1];
  for (int i = 0; i < 3; ++i) lgs[i] = t[2 + i];
  return 0;
}

// Stockham twiddles of the LDS passes, in float pairs (cos, -sin), generated
// in FP64 and rounded once: per pass (NS = product of the earlier radices, L
// its radix) the table A[jj][k] = W_(NS L)^(jj k) (jj < max(L/16, 1),
// k < NS), then B[i][k] = W_(16 NS)^(m_i k) with m = 1, 2, 3, 4, 8, 12.
// Returns the number of floats; fills `out` when it is not null.
uint64_t spx_large_tables(uint32_t nfft, float* out) {
  int r0, lgs[3], np;
  if (spx_large_plan(nfft, &r0, lgs, &np)) return 0;
  uint64_t n = 0;
  auto put = [&](uint64_t e, uint64_t m) {
    if (out) {
      const double a = 2.0 * M_PI * ((double)(e % m) / (double)m);
      out[2 * n] = (float)std::cos(a);
      out[2 * n + 1] = (float)-std::sin(a);
    }
    ++n;
  };
  uint64_t ns = 1ull << r0;
  for (int p = 0; p < np; ++p) {
    const uint64_t L = 1ull << lgs[p];
    for (uint64_t jj = 0; jj < (L >= 16 ? L / 16 : 1); ++jj)
      for (uint64_t k = 0; k < ns; ++k) put(jj * k, ns * L);
    static const int pw[6] = {1, 2, 3, 4, 8, 12};
    for (int i = 0; i < 6; ++i)
      for (uint64_t k = 0; k < ns; ++k) put((uint64_t)pw[i] * k, 16 * ns);
    ns *= L;
  }
  return 2 * n;
}

// Transform (and, for a->peak_mode != 0, the peak search) of a->batch
// channels of length a->nfft in chunks of `chunk` channels.  tw: the
// Stockham tables of spx_large_tables; twt[i]: the plan (L, 0) table of LDS
// pass i's L-point transform, as make_twiddles builds it.
hipError_t spx_launch_large(const FusedArgs* a0, const float2* tw, const float2* const* twt, void* ws, uint32_t chunk,
                            hipStream_t st) {
  using namespace spx;
  const uint32_t N = a0->nfft;
  int r0, lgs[3], np;
  if (spx_large_plan(N, &r0, lgs, &np) || chunk == 0) return hipErrorInvalidValue;
  const uint32_t R0 = 1u << r0;
  const uint64_t nblk = (N + LS_BLK - 1) / LS_BLK;
  const int lgl = lgs[np - 1];
  const uint32_t nparts = N / ((1u << lgl) * gl_cols(lgl));  // last-pass workgroups per channel
  float2* buf[2] = {reinterpret_cast<float2*>(ws), reinterpret_cast<float2*>(ws) + (uint64_t)chunk * N};
  unsigned char* tail = reinterpret_cast<unsigned char*>(ws) + 2ull * chunk * N * 8;
  LsPart* part = reinterpret_cast<LsPart*>(tail);
  tail += ((uint64_t)chunk * (N / 4096) * sizeof(LsPart) + 255) & ~255ull;
  LsChan* chs = reinterpret_cast<LsChan*>(tail);
  tail += ((uint64_t)chunk * sizeof(LsChan) + 255) & ~255ull;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(tail);
  tail += ((uint64_t)chunk * nblk * 4 + 255) & ~255ull;
  uint32_t* idx = reinterpret_cast<uint32_t*>(tail);
  tail += ((uint64_t)chunk * nblk * 4 * LS_CAP + 255) & ~255ull;
  const bool peaks = a0->peak_mode != 0;
  // learned refinement (threshold search): features of the chunk's records
  const bool learn = peaks && a0->peak_mode == 1 && a0->refine == 2 && a0->mlp != nullptr && a0->kcap > 0 &&
                     a0->peaks != nullptr;
  void* feat = learn ? reinterpret_cast<void*>(tail) : nullptr;
  // table offsets (float2 units) of each LDS pass: A | B
  uint64_t ta[3], tb[3], off = 0, ns = R0;
  for (int p = 0; p < np; ++p) {
    const uint64_t L = 1ull << lgs[p];
    ta[p] = off;
    off += ns * (L >= 16 ? L / 16 : 1);
    tb[p] = off;
    off += 6 * ns;
    ns *= L;
  }
  for (uint32_t c0 = 0; c0 < a0->batch; c0 += chunk) {
    const uint32_t nch = min(chunk, a0->batch - c0);
    // pass 0 -> buf[0]; LDS pass p (1-based) reads buf[(p-1) % 2] and writes
    // buf[p % 2], the last one the dense outputs or, for the peak search,
    // the chunk's complex spectrum into the buffer it does not read
    GsArgs g;
    g.a = *a0;
    g.ch0 = c0;
    g.N = N;
    g.dst = buf[0];
    switch (R0) {
      case 8: launch_gs_first<8>(g, nch, st); break;
      default: launch_gs_first<16>(g, nch, st); break;
    }
    float2* spec = buf[np % 2];
    uint32_t nsp = R0;
    for (int p = 0; p < np; ++p) {
      const bool last = p == np - 1;
      GlArgs h;
      h.a = *a0;
      h.ch0 = c0;
      h.N = N;
      h.src = buf[p % 2];
      h.dst = last ? nullptr : buf[(p + 1) % 2];
      h.NS = nsp;
      h.tw_tile = twt[p];
      h.ta = tw + ta[p];
      h.tb = tw + tb[p];
      h.part = (last && peaks) ? part : nullptr;
      if (last && peaks) {  // the chunk's spectrum, indexed from its first channel
        h.a.spec_out = spec;
        h.a.mag_out = nullptr;
        h.ch0 = 0;
      }
      launch_gl_any(lgs[p], h, last, nch, st);
      nsp <<= lgs[p];
    }
    if (peaks) {
      FusedArgs b = *a0;
      b.batch = nch;
      b.peaks = a0->peaks ? a0->peaks + (uint64_t)c0 * a0->kcap : nullptr;
      b.counts = a0->counts + 2ull * c0;
      b.gmax = a0->gmax ? a0->gmax + c0 : nullptr;
      b.stats = a0->stats ? a0->stats + 4ull * c0 : nullptr;
      b.mlp_feat = feat;
      hipLaunchKernelGGL(k_ls_thresh, dim3(nch), dim3(LS_T), 0, st, b, spec, part, nparts, chs);
      hipLaunchKernelGGL(k_ls_cand, dim3((uint32_t)nblk, nch), dim3(LS_T), 0, st, b, spec, chs, (uint32_t)nblk, cnt,
                         idx);
      hipLaunchKernelGGL(k_ls_gather, dim3(nch), dim3(LS_T), 0, st, b, spec, cnt, idx, (uint32_t)nblk);
      if (learn) {
        const hipError_t e = spx_launch_mlp_refine(a0->mlp, feat, b.peaks, b.counts, nch, a0->kcap,
                                                   a0->fs / (float)N, st);
        if (e != hipSuccess) return e;
      }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // extern "C"
