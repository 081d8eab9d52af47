// Device code of the long-channel (N = 256 * N2 >= 32768) kernels
// (spx_fourstep.hip).  Four-step N = 256 * N2, x[n1 + 256 n2] -> X[k2 + N2 k1]:
// N2-point column transforms times W_N^(n1 k2) give Z[k2][n1], then 256-point
// row transforms give the bins k2 + N2 k1.
#pragma once
#include "spx_fused_common.hpp"

namespace spx {

struct FsRowPart {
  double s1, s2;      // sum |X|, sum |X|^2 over the block's rows (in band)
  float m;            // largest |X| of the block, -1 if none
  uint32_t k;         // its bin
  float xr, xi, ml, mr;  // X and the neighbours' |X| (NaN: outside the block)
};

struct FsWork {
  float2* Z;          // [nch][N]
  Cand* cand;         // [nch][N/2 + 2]: the high list from the front, the low list from the back
  uint32_t* ncand;    // [nch][2] lengths of the high and the low list
  double* pe;         // [nch][NT1] input energy per column block
  FsRowPart* part;    // [nch][NT2]
  float* edge;        // [nch][NT2][2][256]  |X| of block rows 0 and RB-1
  // learned refinement only: X at bins k-2, k-1, k+1, k+2 of every candidate
  // (NaN: in another row block) and, when records past the 64th are kept,
  // X of block rows 0, 1, RB-2, RB-1 (otherwise a neighbour in another block
  // is recomputed from Z in the finish)
  float2* candx;      // [nch][CAND_CAP][4]
  float2* edgex;      // [nch][NT2][4][256] or null
  uint32_t ch0, nch;  // chunk: channels ch0 .. ch0+nch-1 of the launch
  // optional indirection (uneven batches): channel i of the launch is channel
  // chan[i] of the caller, whose input starts inoff[chan[i]] elements into a.in
  const uint32_t* chan;
  const uint64_t* inoff;
};

template <int N2, int TC>
struct FsCfg {
  static constexpr int N1 = 256, N = N1 * N2, PPT = 16;
  static constexpr int CB = TC * PPT / N2;     // columns per k_fs_cols workgroup
  static constexpr int NT1 = N1 / CB;
  static constexpr int TR = 256;               // k_fs_rows workgroup (its launch bound)
  static constexpr int RB = TR * PPT / N1;     // rows per k_fs_rows workgroup (16)
  static constexpr int NT2 = N2 / RB;
  static constexpr int TF = 64;                // k_fs_finish workgroup: one wave (its launch bound)
  static constexpr uint64_t CAND_CAP = N / 2 + 2;
  // where the column blocks are at least 16 wide (N2 = 128 .. 512) the
  // column kernel reads the signal and the row kernels read Z with
  // non-temporal loads, and the dense path stores the spectrum
  // non-temporally: measured faster there, slower at N2 = 1024 (8-column
  // blocks).  Z itself is stored plainly: the row kernels find part of it
  // still in the Infinity Cache.
  static constexpr bool STREAM_Z = (CB >= 16);
  // channels per column workgroup, and the waves per SIMD the column kernel
  // is compiled for.  N2 = 512: one channel, so that nothing stays live
  // across channels and the kernel fits 128 registers; two of its 512-thread
  // workgroups then share a CU (LDS 76 KiB each).  Measured at 512 x 131072:
  // peak search 0.349 -> 0.334 ms, complex spectrum 0.424 -> 0.408 ms.  At
  // N2 = 1024 the same form was slower (0.489 -> 0.510 ms), at N2 = 128 too
  // (2048 x 32768: 0.290 -> 0.305 ms), and capping the
  // four-channel form at 128 registers spills (0.353 -> 0.396 ms at 2^17).
  static constexpr int MC = (N2 == 512) ? 1 : 4;
  static constexpr int COLS_WPE = (N2 == 512) ? 4 : 1;
  using LAY1 = LayInter<N2, CB>;
  using LAY2 = LayContig<N1, RB>;
  static_assert(CB >= 8 && NT1 >= 1 && RB >= 4 && NT2 >= 1, "tile shape");
  static_assert(Plan<N2, 0>::R0 == 16, "column plan starts with radix 16");
};

// Candidate pre-filter of a channel, identical in the row and finish
// kernels.  By Parseval mean |X|^2 = P is known before any row transform, and
// mean + k std >= min(1, k) sqrt(P), so maxima below lb (0.999 of that) can
// never clear the threshold.  Candidates at or above hi = 3 lb go to the high
// list: when the threshold turns out >= hi, the low list is not read at all.
struct FsBounds {
  float lb, hi;
};
__device__ __forceinline__ FsBounds fs_bounds(const FusedArgs& a, double P, bool full_band) {
  FsBounds b{0.f, 0.f};
  if (a.peak_mode == 1 && full_band && a.nsigma >= 0.f) {
    b.lb = (float)(fmin(1.0, (double)a.nsigma) * sqrt(P) * 0.999);
    b.hi = 3.0f * b.lb;
  }
  return b;
}

template <int T>
__device__ __forceinline__ double block_sum_d(double x, double* red) {
  constexpr int NW = T / 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  x = wave_sum(x);
  if (lane == 0) red[w] = x;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

template <int N2, int TC>
__device__ __forceinline__ double fs_energy(const FsWork& w, uint32_t ci) {
  using C = FsCfg<N2, TC>;
  double P = 0.0;
  for (int i = 0; i < C::NT1; ++i) P += gp(w.pe)[(uint64_t)ci * C::NT1 + i];
  return P;
}

// 256-point transforms of 16 interleaved columns (thread t: column s = t%16,
// jj = t/16; in: v[r] = x[s][jj + 16 r]; out: v[r] = X[s][jj + 16 r]), with
// the LDS exchange done in halves through a float tile of 5104 entries.
// Element e = 16 a + b of column s is at 320 a + 16 b + s + 16 (a & 3) (rows
// of 320 floats, so the skew never reaches the next row): the pass-0 writes
// (a = jj, b = r) and the pass-1 reads (a = r, b = jj) of each half-wave both
// fall on 32 distinct banks (4-byte accesses go to 32 banks, a half-wave per
// cycle).
template <bool TWL>
__device__ __forceinline__ void fft256_cols16_halfx(float* lf, float2 (&v)[16], const float2* tw, int tid) {
  using PL = Plan<256, 0>;
  static_assert(PL::NP == 2 && PL::radix(0) == 16 && PL::radix(1) == 16, "plan");
  const int s = tid & 15, jj = tid >> 4;
  Dft<16>::run(v);
  float* wb = lf + 320 * jj + s + 16 * (jj & 3);
  const float* rb = lf + 16 * jj + s;
#pragma unroll
  for (int r = 0; r < 16; ++r) wb[16 * r] = v[r].x;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r].x = rb[320 * r + 16 * (r & 3)];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) wb[16 * r] = v[r].y;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r].y = rb[320 * r + 16 * (r & 3)];
  float2 w[16];
  if constexpr (TWL) load_tw_row<16, true>(tw + PL::twoff_lds(1) + jj * 18, w);
  else load_tw_row<16, false>(tw + PL::twoff(1) + (uint32_t)(jj * 16), w);
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = c_mul(v[r], w[r]);
  Dft<16>::run(v);
}

// Orders a wave's LDS writes before its later LDS reads (a wave's LDS
// operations complete in order; this only stops the compiler from moving
// them), for exchanges that stay inside one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 256-point transforms of the RB rows of a LayContig<256, RB> tile, 16
// points per thread, with the one LDS exchange between the two radix-16
// passes done in halves (real parts, then imaginary parts): the tile needs 4
// bytes per point.  In: v[r] = x[s][jj + 16 r] for s = tid/16, jj = tid%16.
// Out: v[r] = X[s][jj + 16 r].  tw: the twiddle tables of Plan<256, 0>, in
// global memory or (TWL) the padded LDS copy of copy_tw_to_lds.
// A row is 16 consecutive lanes, so every exchange stays inside one wave
// (rows 4w .. 4w+3 of the tile belong to wave w) and needs no workgroup
// barrier.
template <int RB, int T, bool TWL>
__device__ __forceinline__ void fft256_rows_halfx(float* lf, float2 (&v)[16], const float2* tw, int tid) {
  using LAY = LayContig<256, RB>;
  using PL = Plan<256, 0>;
  static_assert(PL::NP == 2 && PL::radix(0) == 16 && PL::radix(1) == 16 && T * 16 == 256 * RB, "plan");
  Dft<16>::run(v);  // pass 0: sub-transform length 1, no twiddles
  int s, jj;
  LAY::template map<16>(tid, s, jj);
  float* wb = lf + LAY::idx(s, 16 * jj);  // pass-0 outputs y[s][16 jj + r]
  const float* rbase = lf + LAY::idx(s, jj);  // pass-1 inputs y[s][jj + 16 r], stride 17
#pragma unroll
  for (int r = 0; r < 16; ++r) wb[r] = v[r].x;
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r].x = rbase[17 * r];
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 16; ++r) wb[r] = v[r].y;
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r].y = rbase[17 * r];
  // pass 1: sub-transform length 16, twiddles W_256^(r (jj % 16))
  float2 w[16];
  if constexpr (TWL) load_tw_row<16, true>(tw + PL::twoff_lds(1) + jj * 18, w);
  else load_tw_row<16, false>(tw + PL::twoff(1) + (uint32_t)(jj * 16), w);
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = c_mul(v[r], w[r]);
  Dft<16>::run(v);
}

// |X| of the outer neighbour of bin k (side -1: k-1 from the previous block's
// last row; side +1: k+1 from the next block's first row; both wrap)
// Edge layout [N2/RB blocks][2 rows: 0 and RB-1][256 columns].
template <int N2, int RB>
__device__ __forceinline__ float fs_outer_rb(const float* edge, uint32_t k, int side) {
  constexpr int N1 = 256, NT2 = N2 / RB;
  const uint32_t k2 = k & (N2 - 1), k1 = k / N2;
  if (side < 0) {
    if (k2 > 0) return gp(edge)[((k2 / RB - 1) * 2 + 1) * N1 + k1];
    return gp(edge)[((NT2 - 1) * 2 + 1) * N1 + (k1 + N1 - 1) % N1];  // k-1 wraps to row N2-1, column k1-1
  }
  if (k2 < (uint32_t)N2 - 1) return gp(edge)[((k2 / RB + 1) * 2 + 0) * N1 + k1];
  return gp(edge)[0 * N1 + (k1 + 1) % N1];  // k+1 wraps to row 0, column k1+1
}
template <int N2, int TC>
__device__ __forceinline__ float fs_outer(const float* edge, uint32_t k, int side) {
  return fs_outer_rb<N2, FsCfg<N2, TC>::RB>(edge, k, side);
}

// Column-kernel workgroup size per column length: 16 columns per workgroup
// (128-byte row segments) up to N2 = 512; N2 = 1024 gets 8 (16 would need
// 1024 threads, fewer registers than its transform needs).
template <int N2>
struct FsShape { static constexpr int TC = N2 >= 512 ? 512 : 256; };

}  // namespace spx
