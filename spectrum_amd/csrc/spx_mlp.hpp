// Learned sub-bin refiner: a 10 -> 64 -> 64 -> 1 ReLU network that maps the
// 5-bin complex neighbourhood of a spectral peak to its frequency offset in
// bins.  Shared by the training kernels (spx_mlp.hip) and the fused peak
// kernel, so training-time evaluation and fused inference run the same code.
//
// Input features (10 reals): z_j = X[k+j] / X[k], j = -2..2 (circular), as
// (re, im) pairs in order j = -2, -1, 0, 1, 2; dividing by the complex peak
// value normalises magnitude and phase (z_0 = 1 + 0i).
//
// Parameters (FP32 master copy, 4929 values): W1[64][10] b1[64] W2[64][64]
// b2[64] w3[64] b3.  Deployed copy (inference): FP16 MFMA B-operand
// fragments of W1^T (K padded 10 -> 32) and W2^T plus FP16-valued biases and
// w3 held as float; FP32 accumulation (v_mfma_f32_16x16x32_f16).
#pragma once
#include "spx_device.hpp"
#include "spx_kernels.h"

namespace spx {
namespace mlp {
constexpr int IN = 10, H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;  // 4929
// deployed block (bytes)
constexpr int D_W1 = 0;                      // [cb 4][lane 64][8 x fp16]
constexpr int D_W2 = D_W1 + 4 * 64 * 16;     // [ks 2][cb 4][lane 64][8 x fp16]
constexpr int D_B1 = D_W2 + 8 * 64 * 16;     // float[64]
constexpr int D_B2 = D_B1 + 256;
constexpr int D_W3 = D_B2 + 256;
constexpr int D_B3 = D_W3 + 256;             // float, padded to 16 B
constexpr int D_BYTES = D_B3 + 16;
}  // namespace mlp

typedef _Float16 mlp_h8 __attribute__((ext_vector_type(8)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// features of the peak at k from any accessor X(k) (indices already circular)
template <class XF>
__device__ __forceinline__ void mlp_features(XF X, uint32_t k, uint32_t n, float* f) {
  const float2 c = X(k);
  const float den = fmaf(c.x, c.x, c.y * c.y);
  const float inv = den > 0.f ? 1.0f / den : 0.f;
#pragma unroll
  for (int j = -2; j <= 2; ++j) {
    const float2 x = X((k + n + (uint32_t)j) % n);
    // x * conj(c) / |c|^2
    f[2 * (j + 2) + 0] = fmaf(x.x, c.x, x.y * c.y) * inv;
    f[2 * (j + 2) + 1] = fmaf(x.y, c.x, -x.x * c.y) * inv;
  }
}

// The deployed weights one lane of the evaluating wave needs, in registers:
// loaded by mlp_load well before mlp_rows16 so that their latency hides
// behind other work (the global loads stay in flight across barriers).
struct MlpRegs {
  mlp_h8 w1[4];      // W1^T fragments, column blocks 0..3
  mlp_h8 w2[2][4];   // W2^T fragments, K-steps 0..1, column blocks 0..3
  float b1[4], b2[4], w3[4];
  float b3;
};

__device__ __forceinline__ MlpRegs mlp_load(const unsigned char* wts) {
  using namespace mlp;
  const int l = threadIdx.x & 63, c = l & 15;
  const mlp_h8* W1F = reinterpret_cast<const mlp_h8*>(wts + D_W1);
  const mlp_h8* W2F = reinterpret_cast<const mlp_h8*>(wts + D_W2);
  MlpRegs r;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    r.w1[cb] = gp(W1F)[cb * 64 + l];
    r.w2[0][cb] = gp(W2F)[(0 * 4 + cb) * 64 + l];
    r.w2[1][cb] = gp(W2F)[(1 * 4 + cb) * 64 + l];
    r.b1[cb] = gp(reinterpret_cast<const float*>(wts + D_B1))[16 * cb + c];
    r.b2[cb] = gp(reinterpret_cast<const float*>(wts + D_B2))[16 * cb + c];
    r.w3[cb] = gp(reinterpret_cast<const float*>(wts + D_W3))[16 * cb + c];
  }
  r.b3 = gp(reinterpret_cast<const float*>(wts + D_B3))[0];
  return r;
}

// Feature rows of the fused kernels' learned refinement: stored peak `pos`
// of channel ch at a.mlp_feat + (ch * kcap + pos) * 16 halves.
__device__ __forceinline__ _Float16* mlp_feat_row(const FusedArgs& a, uint64_t ch, int pos) {
  return reinterpret_cast<_Float16*>(a.mlp_feat) + (ch * a.kcap + (uint32_t)pos) * 16;
}

// Features of the peak at k as one 32-byte row of 16 halves (the 10
// features, then zeros): the row layout mlp_rows16 reads.
template <class XF>
__device__ __forceinline__ void mlp_store_features(XF X, uint32_t k, uint32_t n, _Float16* row) {
  float f[mlp::IN];
  mlp_features(X, k, n, f);
  mlp_h8 lo, hi;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    lo[j] = (_Float16)f[j];
    hi[j] = (_Float16)(j + 8 < mlp::IN ? f[j + 8] : 0.f);
  }
  mlp_h8* r = reinterpret_cast<mlp_h8*>(row);
  gp(r)[0] = lo;
  gp(r)[1] = hi;
}

// One wave evaluates the network for 16 rows.  feat: 16 rows x 16 halves in
// LDS (features in columns 0..9, zeros in 10..15; K columns 16..31 of the
// MFMA operand are zero), hbuf: 16 x 64 halves of LDS scratch; out[row]
// receives the offsets (rows >= the caller's count are computed but not
// meaningful).  All 64 lanes of the wave must call it, with the registers
// mlp_load gave each of them.
__device__ __forceinline__ void mlp_rows16(const _Float16* feat, _Float16* hbuf, float* out, const MlpRegs& w) {
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  // layer 1: A = feat (row l%16, k = 8g + j), B = W1^T fragments
  mlp_h8 a1 = {0, 0, 0, 0, 0, 0, 0, 0};
  if (g < 2) a1 = *reinterpret_cast<const mlp_h8*>(feat + c * 16 + 8 * g);
  mlp_f4 h[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const mlp_f4 z = {0.f, 0.f, 0.f, 0.f};
    h[cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, w.w1[cb], z, 0, 0, 0);
  }
  // lane holds h[cb][i] = H1[row 4g+i][unit 16cb+c]: bias, ReLU, stage as fp16 rows
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) hbuf[(4 * g + i) * 64 + 16 * cb + c] = (_Float16)fmaxf(h[cb][i] + w.b1[cb], 0.f);
  }
  __builtin_amdgcn_s_waitcnt(0);  // LDS writes of this wave complete before the reads below
  __builtin_amdgcn_wave_barrier();
  // layer 2: A = H1 rows (two K-steps of 32), B = W2^T fragments
  const mlp_h8 a20 = *reinterpret_cast<const mlp_h8*>(hbuf + c * 64 + 8 * g);
  const mlp_h8 a21 = *reinterpret_cast<const mlp_h8*>(hbuf + c * 64 + 32 + 8 * g);
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const mlp_f4 z = {0.f, 0.f, 0.f, 0.f};
    mlp_f4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a20, w.w2[0][cb], z, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a21, w.w2[1][cb], d, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) part[i] = fmaf(fmaxf(d[i] + w.b2[cb], 0.f), w.w3[cb], part[i]);
  }
  // layer 3: sum the 16 unit-lanes of each row group
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) part[i] += __shfl_xor(part[i], o, 64);
  }
  if (c == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[4 * g + i] = part[i] + w.b3;
  }
}

// Learned refinement of one channel's n stored records by one wave, 16 per
// MFMA pass, from the feature rows the peak kernel wrote
// (mlp_store_features): a prediction inside [-0.5, 0.5] replaces the
// record's parabolic offset and refined frequency; anything else (outside,
// not finite) keeps them.  fbuf: 16 x 16 halves, hbuf: 16 x 64 halves, res:
// 16 floats of this wave's LDS.
__device__ __forceinline__ void mlp_refine_rows(const MlpRegs& w, const _Float16* feat, spx_max_value* rec, uint32_t n,
                                                float bin_hz, _Float16* fbuf, _Float16* hbuf, float* res) {
  const int l = threadIdx.x & 63;
  const mlp_h8* src = reinterpret_cast<const mlp_h8*>(feat);
  for (uint32_t r0 = 0; r0 < n; r0 += 16) {
    // rows r0 .. r0+15 as 32 half-rows of 16 bytes, zeros past the last peak
    if (l < 32) {
      const uint32_t row = r0 + (uint32_t)(l >> 1);
      mlp_h8 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (row < n) v = gp(src)[2 * row + (l & 1)];
      reinterpret_cast<mlp_h8*>(fbuf)[l] = v;
    }
    __builtin_amdgcn_s_waitcnt(0);  // LDS writes of this wave complete before mlp_rows16 reads them
    __builtin_amdgcn_wave_barrier();
    mlp_rows16(fbuf, hbuf, res, w);
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (l < 16 && r0 + l < n) {
      const float dl = res[l];
      if (fabsf(dl) <= 0.5f) {
        spx_max_value* r = rec + r0 + l;
        const uint32_t k = gp(&r->index)[0];
        gp(&r->freq_offset)[0] = dl;
        gp(&r->refined_frequency)[0] = ((float)k + dl) * bin_hz;
      }
    }
    __builtin_amdgcn_wave_barrier();  // res and fbuf are rewritten by the next pass
  }
}

// Same, loading the weights itself (callers with nothing to overlap).
__device__ __forceinline__ void mlp_rows16(const _Float16* feat, _Float16* hbuf, float* out,
                                           const unsigned char* wts) {
  const MlpRegs w = mlp_load(wts);
  mlp_rows16(feat, hbuf, out, w);
}

}  // namespace spx
