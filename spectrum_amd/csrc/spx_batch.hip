// Persistent uneven-length batch: every channel of a batch whose length is
// 2^10 .. 2^14 runs through the fused window -> FFT -> |X| -> threshold peak
// search in persistent launches (the host plan captures them in a hipGraph
// together with the split-path launches of the longer lengths).
//
// The host builds a work list per tile shape, sorted by decreasing cost (tiles
// of the longest length first: a longest-processing-time-first schedule).  A
// tile shape is TILE complex points in LDS held by TILE/16 threads: 16384
// (1024 threads, one workgroup per CU), 8192 (512 threads, two per CU) or
// 4096 (256 threads, four per CU); the grid of a shape fills every CU.  Each
// workgroup takes the next item from its launch's cursor (one relaxed atomic
// per item, no other cross-workgroup traffic); an item is a tile of
// G = TILE/L whole channels: the same Stockham passes as the single-channel
// kernel but over G sequences, then per-channel (segmented) statistics, peak
// flags and scan, so every channel's records equal the single-channel
// kernel's.  The tile shape and radix order (Plan<L, V>) of each length are
// chosen by the host autotuner; the order is carried in the item word.
#include "spx_fused_common.hpp"
#include "spx_mlp.hpp"

namespace spx {

template <int DT, int R>
__device__ __forceinline__ void load_pts_t(const void* cb, const float* win, uint32_t n0, uint32_t stride, float2* v) {
  float2 raw[R];
  float wv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) raw[r] = ld_raw<DT>(cb, n0 + r * stride);
  if (win != nullptr) {
#pragma unroll
    for (int r = 0; r < R; ++r) wv[r] = gp(win)[n0 + r * stride];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) wv[r] = 1.0f;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = c_scale(raw[r], wv[r]);
}

template <int R>
__device__ __forceinline__ void load_pts(const void* cb, uint32_t dtype, const float* win, uint32_t n0, uint32_t stride,
                                         float2* v) {
  if (dtype == SPX_DTYPE_C64) load_pts_t<SPX_DTYPE_C64, R>(cb, win, n0, stride, v);
  else if (dtype == SPX_DTYPE_C32F16) load_pts_t<SPX_DTYPE_C32F16, R>(cb, win, n0, stride, v);
  else load_pts_t<SPX_DTYPE_C32BF16, R>(cb, win, n0, stride, v);
}

__device__ __forceinline__ const void* chan_base(const BatchArgs& b, uint32_t ch) {
  const uint64_t es = b.dtype == SPX_DTYPE_C64 ? 8 : 4;
  return reinterpret_cast<const char*>(b.in) + gp(b.offsets)[ch] * es;
}

// G = TILE/L channels of length L, TILE/16 threads, 16 points per thread.
template <int L, int V, int TILE>
__device__ __forceinline__ void batch_small_tile(const BatchArgs& b, const uint4 it, float2* lds, BlockRed& red) {
  using PL = Plan<L, V>;
  constexpr int T = TILE / 16, PPT = 16, G = TILE / L;
  constexpr int TPC = L / 16;      // threads per channel in the peak stage (16 bins each)
  constexpr int WPC = TPC / 64;    // waves per channel
  constexpr int R0 = PL::R0, RL = PL::RL;
  using LAY = LayContig<L, G>;
  static_assert(WPC >= 1 && G * L == T * PPT, "tile shape");
  const int tid = opaque_v((int)threadIdx.x);
  const uint32_t bucket = it.x & 0xffu;
  const float2* tw = opaque(b.bk[bucket].tw);
  const float* win = opaque(b.bk[bucket].window);
  const uint32_t first = it.y, count = it.z;

  float2 v[PPT];
#pragma unroll
  for (int q = 0; q < PPT / R0; ++q) {
    const int j = tid + q * T;
    const int s = j / (L / R0), jj = j % (L / R0);
    if ((uint32_t)s < count) {
      const uint32_t ch = gp(b.perm)[first + s];
      load_pts<R0>(chan_base(b, ch), b.dtype, win, (uint32_t)jj, (uint32_t)(L / R0), &v[q * R0]);
    } else {
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = make_float2(0.f, 0.f);
    }
  }
  fft_tile_pf<LAY, L, G, T, PPT, V>(lds, v, tw, tid);
  __syncthreads();  // the last pass read LDS
#pragma unroll
  for (int q = 0; q < PPT / RL; ++q) {
    int s, jj;
    LAY::template map<RL>(tid + q * T, s, jj);
#pragma unroll
    for (int r = 0; r < RL; ++r) lds[LAY::idx(s, jj + r * (L / RL))] = v[q * RL + r];
  }
  __syncthreads();

  // peak stage: thread owns bins k0 .. k0+15 of channel slot s
  const int s = tid / TPC;
  const uint32_t k0 = (uint32_t)(tid % TPC) * 16u;
  const bool active = (uint32_t)s < count;
  const int w = tid >> 6, lane = tid & 63;
  float m[16];
  float fs1 = 0.f, fs2 = 0.f, bm = -1.0f;
  uint32_t bk = 0xffffffffu;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const uint32_t k = k0 + e;
    m[e] = mag_of(lds[LAY::idx(s, k)]);
    fs1 += m[e];
    fs2 = fmaf(m[e], m[e], fs2);
    if (m[e] > bm) { bm = m[e]; bk = k; }
  }
  double s1 = wave_sum((double)fs1), s2 = wave_sum((double)fs2);
  wave_argmax(bm, bk);
  if (lane == 0) {
    red.s1[w] = s1;
    red.s2[w] = s2;
    red.m[w] = bm;
    red.k[w] = bk;
  }
  __syncthreads();
  s1 = 0.0;
  s2 = 0.0;
  bm = -1.0f;
  bk = 0xffffffffu;
#pragma unroll
  for (int i = 0; i < WPC; ++i) {
    const int wi = s * WPC + i;
    s1 += red.s1[wi];
    s2 += red.s2[wi];
    if (red.m[wi] > bm || (red.m[wi] == bm && red.k[wi] < bk)) {
      bm = red.m[wi];
      bk = red.k[wi];
    }
  }
  __syncthreads();
  const double mu = s1 / (double)L;
  double var = s2 / (double)L - mu * mu;
  if (var < 0) var = 0;
  const double sd = sqrt(var);
  const float thr = (float)(mu + (double)b.nsigma * sd);

  const float ml = mag_of(lds[LAY::idx(s, (k0 + L - 1) % L)]);
  const float mr = mag_of(lds[LAY::idx(s, (k0 + 16) % L)]);
  uint32_t flags = 0;
  int c = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const float yl = (e == 0) ? ml : m[e - 1];
    const float yr = (e == 15) ? mr : m[e + 1];
    if (m[e] > yl && m[e] > yr && m[e] > thr) {
      flags |= 1u << e;
      ++c;
    }
  }
  // segmented exclusive scan over the channel's waves
  const int inc = wave_incl_scan(c);
  if (lane == 63) red.c[w] = inc;
  __syncthreads();
  int off = 0, total = 0;
#pragma unroll
  for (int i = 0; i < WPC; ++i) {
    const int wi = s * WPC + i;
    if (wi < w) off += red.c[wi];
    total += red.c[wi];
  }
  off += inc - c;
  const float bin_hz = b.fs / (float)L;
  if (active) {
    const uint32_t ch = gp(b.perm)[first + s];
    if (flags && off < (int)b.kcap) {
      MaxValue* out = b.peaks + (uint64_t)ch * b.kcap;
      int pos = off;
      while (flags && pos < (int)b.kcap) {
        const int e = __builtin_ctz(flags);
        flags &= flags - 1;
        const uint32_t k = k0 + e;
        const float yc = mag_of(lds[LAY::idx(s, k)]);
        const float yl = mag_of(lds[LAY::idx(s, (k + L - 1) % L)]);
        const float yr = mag_of(lds[LAY::idx(s, (k + 1) % L)]);
        write_record_hz(out + pos, k, lds[LAY::idx(s, k)], yl, yc, yr, b.refine, bin_hz);
        if (b.mlp_feat != nullptr)
          mlp_store_features([&](uint32_t q) { return lds[LAY::idx(s, (int)q)]; }, k, (uint32_t)L,
                             reinterpret_cast<_Float16*>(b.mlp_feat) +
                                 ((uint64_t)(first + s - b.feat_base) * b.kcap + (uint32_t)pos) * 16);
        ++pos;
      }
    }
    if ((tid % TPC) == 0) {
      if (b.counts) {
        gp(b.counts)[2 * (uint64_t)ch + 0] = (uint32_t)min(total, (int)b.kcap);
        gp(b.counts)[2 * (uint64_t)ch + 1] = (uint32_t)total;
      }
      if (b.stats) {
        gp(b.stats)[4 * (uint64_t)ch + 0] = (float)mu;
        gp(b.stats)[4 * (uint64_t)ch + 1] = (float)sd;
        gp(b.stats)[4 * (uint64_t)ch + 2] = thr;
        gp(b.stats)[4 * (uint64_t)ch + 3] = (float)(s2 / (double)L);
      }
      if (b.gmax && bk != 0xffffffffu) {
        const float yl = mag_of(lds[LAY::idx(s, (bk + L - 1) % L)]);
        const float yr = mag_of(lds[LAY::idx(s, (bk + 1) % L)]);
        write_record_hz(b.gmax + ch, bk, lds[LAY::idx(s, bk)], yl, bm, yr, b.refine, bin_hz);
      }
    }
  }
  __syncthreads();  // LDS and red are reused by the next item
}

__global__ void k_batch_reset(uint32_t* counter) {
  if (threadIdx.x == 0) gp(counter)[0] = 0u;
}

// LDS of a tile: every length's LayContig<L, TILE/L> has the same size
template <int TILE>
struct BatchShared {
  float2 tile[LayContig<TILE, 1>::SIZE];
  BlockRed red;
};
static_assert(LayContig<16384, 1>::SIZE == LayContig<1024, 16>::SIZE, "one tile size for every length");

// one tile item of length L (item word: bucket | variant << 8)
template <int TILE, int L>
__device__ __forceinline__ void batch_item_len(const BatchArgs& b, const uint4 it, BatchShared<TILE>& sh) {
  if constexpr (L <= TILE) {
    if ((it.x >> 8) & 1u) batch_small_tile<L, 1, TILE>(b, it, sh.tile, sh.red);
    else batch_small_tile<L, 0, TILE>(b, it, sh.tile, sh.red);
  }
}

// Every TILE/16-thread shape keeps four waves per SIMD (128 VGPRs): one
// 1024-thread, two 512-thread or four 256-thread workgroups per CU.
template <int TILE>
__global__ void __launch_bounds__(TILE / 16) __attribute__((amdgpu_waves_per_eu(4))) k_batch(BatchArgs b) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ uint32_t item_s;
  BatchShared<TILE>& sh = *reinterpret_cast<BatchShared<TILE>*>(smem_raw);
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(b.counter, 1u);
    __syncthreads();
    const uint32_t i = __builtin_amdgcn_readfirstlane(item_s);
    __syncthreads();
    if (i >= b.n_items) break;  // uniform: every thread read the same i
    uint4 it = gp(b.items)[i];
    it.x = __builtin_amdgcn_readfirstlane(it.x);
    it.y = __builtin_amdgcn_readfirstlane(it.y);
    it.z = __builtin_amdgcn_readfirstlane(it.z);
    // only orders the plan has tables for are ever put in an item (the host
    // validates every item word: variant 1 only where the length has one)
    switch (it.x & 0xffu) {
      case 0: batch_item_len<TILE, 1024>(b, it, sh); break;
      case 1: batch_item_len<TILE, 2048>(b, it, sh); break;
      case 2: batch_item_len<TILE, 4096>(b, it, sh); break;
      case 3: batch_item_len<TILE, 8192>(b, it, sh); break;
      case 4: batch_item_len<TILE, 16384>(b, it, sh); break;
      default: break;
    }
  }
}

}  // namespace spx

extern "C" {

// Dynamic LDS of the batch kernel of a tile shape (one tile).
static size_t batch_smem(int tile) {
  using namespace spx;
  return tile == 16384 ? sizeof(BatchShared<16384>) : tile == 8192 ? sizeof(BatchShared<8192>) : sizeof(BatchShared<4096>);
}

// Allows the batch kernels their dynamic LDS; call once per process before
// the first launch and outside any stream capture.
hipError_t spx_batch_kernel_init(void) {
  using namespace spx;
  hipError_t e = hipFuncSetAttribute((const void*)k_batch<16384>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)batch_smem(16384));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_batch<8192>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)batch_smem(8192));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_batch<4096>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)batch_smem(4096));
  return e;
}

int spx_batch_tile_ok(int tile) { return tile == 16384 || tile == 8192 || tile == 4096; }

// b->counter must point at a word of its own per launch (reset here).
hipError_t spx_launch_batch(const BatchArgs* b, int tile, int grid, hipStream_t st) {
  using namespace spx;
  if (!spx_batch_tile_ok(tile) || grid <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, b->counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (tile == 16384) hipLaunchKernelGGL(k_batch<16384>, dim3(grid), dim3(1024), batch_smem(tile), st, *b);
  else if (tile == 8192) hipLaunchKernelGGL(k_batch<8192>, dim3(grid), dim3(512), batch_smem(tile), st, *b);
  else hipLaunchKernelGGL(k_batch<4096>, dim3(grid), dim3(256), batch_smem(tile), st, *b);
  return hipGetLastError();
}

}  // extern "C"
