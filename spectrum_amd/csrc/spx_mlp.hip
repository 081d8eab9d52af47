// Training-side kernels of the learned sub-bin refiner (network and feature
// definitions in spx_mlp.hpp):
//   k_mlp_dataset  features + zero-padded-FFT labels for detected peaks
//   k_mlp_grad     forward + backward over a mini-batch, per-workgroup partial
//                  gradients held in registers (no atomics; deterministic)
//   k_mlp_reduce   sums the partials, k_mlp_adam updates the FP32 master
//   k_mlp_export   FP32 master -> deployed FP16 fragment block
//   k_mlp_predict  inference with the deployed block (the fused kernel's code)
#include "spx_fused_common.hpp"
#include "spx_mlp.hpp"

namespace spx {

using namespace mlp;

// ---------------------------------------------------------------------------
// dataset: one thread per (channel, record slot)
//   spec  [B][N] complex64 windowed spectrum (same window as the search)
//   mag16 [B][16N] |X| of the 16x zero-padded transform of the same window
//   label = fine-grid position of the padded maximum within +-11 fine bins of
//           16k, refined by a parabola through its neighbours, in coarse bins
//           relative to k; invalid when the maximum sits on the search edge
// ---------------------------------------------------------------------------
__global__ void k_mlp_dataset(const float2* spec, const float* mag16, const spx_max_value* peaks,
                              const uint32_t* counts, uint32_t B, uint32_t N, uint32_t K, float* feat, float* label,
                              float* parab, uint32_t* valid) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)B * K) return;
  const uint32_t b = (uint32_t)(i / K), slot = (uint32_t)(i % K);
  uint32_t ok = 0;
  float f[IN];
#pragma unroll
  for (int j = 0; j < IN; ++j) f[j] = 0.f;
  float lab = 0.f, par = 0.f;
  if (slot < gp(counts)[2 * b]) {
    const spx_max_value pk = gp(peaks)[(uint64_t)b * K + slot];
    const uint32_t k = pk.index;
    const float2* X = spec + (uint64_t)b * N;
    mlp_features([&](uint32_t q) { return gp(X)[q]; }, k, N, f);
    par = pk.freq_offset;
    const uint32_t NF = 16 * N;
    const float* M = mag16 + (uint64_t)b * NF;
    int best = 0;
    float bm = -1.f;
    for (int o = -11; o <= 11; ++o) {
      const float m = gp(M)[(uint64_t)((int64_t)(16ull * k) + NF + o) % NF];
      if (m > bm) { bm = m; best = o; }
    }
    if (best > -11 && best < 11) {
      const float yl = gp(M)[(uint64_t)((int64_t)(16ull * k) + NF + best - 1) % NF];
      const float yr = gp(M)[(uint64_t)((int64_t)(16ull * k) + NF + best + 1) % NF];
      const float den = yl - 2.0f * bm + yr;
      const float d = fabsf(den) > 0.f ? 0.5f * (yl - yr) / den : 0.f;
      lab = ((float)best + d) * (1.0f / 16.0f);
      ok = fabsf(lab) <= 0.75f ? 1u : 0u;
    }
  }
#pragma unroll
  for (int j = 0; j < IN; ++j) gp(feat)[i * IN + j] = f[j];
  gp(label)[i] = lab;
  gp(parab)[i] = par;
  gp(valid)[i] = ok;
}

// ---------------------------------------------------------------------------
// forward + backward, 64 samples per step, 256 threads, persistent over the
// batch.  Gradient ownership: thread t accumulates dW2[t/4][16(t%4)..+16];
// threads t < 64 accumulate dW1[t][*], db1[t], db2[t], dw3[t]; thread 0 db3.
// loss = mean over the n samples of (out - y)^2.
// ---------------------------------------------------------------------------
constexpr int GS = 64;     // samples per step
constexpr int HP = H + 1;  // padded row length

__global__ void __launch_bounds__(256) k_mlp_grad(const float* params, const float* X, const float* Y, uint32_t n,
                                                  float* partial /* [grid][NPARAM + 1] */) {
  __shared__ float w[NPARAM];
  __shared__ float xs[GS][IN + 1];
  __shared__ float h1[GS][HP], h2[GS][HP], g2[GS][HP], g1[GS][HP];
  __shared__ float dout[GS], lossv[GS];
  const int t = threadIdx.x;
  for (int i = t; i < NPARAM; i += 256) w[i] = gp(params)[i];
  float aW2[16], aW1[IN], aB1 = 0.f, aB2 = 0.f, aW3 = 0.f, aB3 = 0.f, aL = 0.f;
#pragma unroll
  for (int q = 0; q < 16; ++q) aW2[q] = 0.f;
#pragma unroll
  for (int q = 0; q < IN; ++q) aW1[q] = 0.f;
  const float inv_n = 1.0f / (float)max(n, 1u);
  const uint32_t nsteps = (n + GS - 1) / GS;
  __syncthreads();
  for (uint32_t st = blockIdx.x; st < nsteps; st += gridDim.x) {
    const uint32_t s0 = st * GS;
    for (int q = t; q < GS * IN; q += 256) {
      const int s = q / IN, j = q % IN;
      xs[s][j] = (s0 + s < n) ? gp(X)[(uint64_t)(s0 + s) * IN + j] : 0.f;
    }
    __syncthreads();
    const int s = t >> 2, hb = (t & 3) * 16;
    // H1 = relu(W1 x + b1)
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int hh = hb + q;
      float acc = w[OFF_B1 + hh];
#pragma unroll
      for (int j = 0; j < IN; ++j) acc = fmaf(w[OFF_W1 + hh * IN + j], xs[s][j], acc);
      h1[s][hh] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    // H2 = relu(W2 h1 + b2)
#pragma unroll 2
    for (int q = 0; q < 16; ++q) {
      const int hh = hb + q;
      float acc = w[OFF_B2 + hh];
#pragma unroll 8
      for (int j = 0; j < H; ++j) acc = fmaf(w[OFF_W2 + hh * H + j], h1[s][j], acc);
      h2[s][hh] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    if (t < GS) {
      float o = w[OFF_B3];
      for (int j = 0; j < H; ++j) o = fmaf(w[OFF_W3 + j], h2[t][j], o);
      const bool live = s0 + t < n;
      const float e = live ? o - gp(Y)[s0 + t] : 0.f;
      dout[t] = 2.0f * e * inv_n;
      lossv[t] = e * e;
    }
    __syncthreads();
    // dH2 = dout * w3 * [h2 > 0]
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int hh = hb + q;
      g2[s][hh] = h2[s][hh] > 0.f ? dout[s] * w[OFF_W3 + hh] : 0.f;
    }
    __syncthreads();
    // dW2[r][k] += sum_s dH2[s][r] h1[s][k]
    {
      const int r = t >> 2, kb = (t & 3) * 16;
      for (int ss = 0; ss < GS; ++ss) {
        const float gg = g2[ss][r];
#pragma unroll
        for (int q = 0; q < 16; ++q) aW2[q] = fmaf(gg, h1[ss][kb + q], aW2[q]);
      }
    }
    // dH1 = (W2^T dH2) * [h1 > 0]
#pragma unroll 2
    for (int q = 0; q < 16; ++q) {
      const int kk = hb + q;
      float acc = 0.f;
#pragma unroll 8
      for (int r = 0; r < H; ++r) acc = fmaf(w[OFF_W2 + r * H + kk], g2[s][r], acc);
      g1[s][kk] = h1[s][kk] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    if (t < H) {
      for (int ss = 0; ss < GS; ++ss) {
        const float gg1 = g1[ss][t];
#pragma unroll
        for (int j = 0; j < IN; ++j) aW1[j] = fmaf(gg1, xs[ss][j], aW1[j]);
        aB1 += gg1;
        aB2 += g2[ss][t];
        aW3 = fmaf(dout[ss], h2[ss][t], aW3);
      }
      if (t == 0) {
        for (int ss = 0; ss < GS; ++ss) {
          aB3 += dout[ss];
          aL += lossv[ss];
        }
      }
    }
    __syncthreads();
  }
  float* out = partial + (uint64_t)blockIdx.x * (NPARAM + 1);
  {
    const int r = t >> 2, kb = (t & 3) * 16;
#pragma unroll
    for (int q = 0; q < 16; ++q) gp(out)[OFF_W2 + r * H + kb + q] = aW2[q];
  }
  if (t < H) {
#pragma unroll
    for (int j = 0; j < IN; ++j) gp(out)[OFF_W1 + t * IN + j] = aW1[j];
    gp(out)[OFF_B1 + t] = aB1;
    gp(out)[OFF_B2 + t] = aB2;
    gp(out)[OFF_W3 + t] = aW3;
  }
  if (t == 0) {
    gp(out)[OFF_B3] = aB3;
    gp(out)[NPARAM] = aL * inv_n;
  }
}

// grad[p] = sum over workgroups; grad[NPARAM] = loss
__global__ void k_mlp_reduce(const float* partial, uint32_t groups, float* grad) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > (uint32_t)NPARAM) return;
  float s = 0.f;
  for (uint32_t gI = 0; gI < groups; ++gI) s += gp(partial)[(uint64_t)gI * (NPARAM + 1) + p];
  gp(grad)[p] = s;
}

__global__ void k_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)NPARAM) return;
  const float g = gp(grad)[p];
  const float mm = fmaf(b1, gp(m)[p], (1.f - b1) * g);
  const float vv = fmaf(b2, gp(v)[p], (1.f - b2) * g * g);
  gp(m)[p] = mm;
  gp(v)[p] = vv;
  gp(params)[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// FP32 master -> deployed block (FP16 fragments, FP16-valued biases / w3)
__global__ void k_mlp_export(const float* params, unsigned char* dep) {
  const int t = threadIdx.x;  // 256 threads
  _Float16* w1f = reinterpret_cast<_Float16*>(dep + D_W1);
  _Float16* w2f = reinterpret_cast<_Float16*>(dep + D_W2);
  // W1 fragments: [cb][lane][j] = W1[16cb + lane%16][8(lane/16) + j] (0 beyond the 10 inputs)
  for (int q = t; q < 4 * 64 * 8; q += 256) {
    const int j = q & 7, lane = (q >> 3) & 63, cb = q >> 9;
    const int unit = 16 * cb + (lane & 15), in = 8 * (lane >> 4) + j;
    gp(w1f)[q] = (_Float16)(in < IN ? gp(params)[OFF_W1 + unit * IN + in] : 0.f);
  }
  // W2 fragments: [ks][cb][lane][j] = W2[16cb + lane%16][32ks + 8(lane/16) + j]
  for (int q = t; q < 8 * 64 * 8; q += 256) {
    const int j = q & 7, lane = (q >> 3) & 63, f = q >> 9, cb = f & 3, ks = f >> 2;
    const int unit = 16 * cb + (lane & 15), in = 32 * ks + 8 * (lane >> 4) + j;
    gp(w2f)[q] = (_Float16)gp(params)[OFF_W2 + unit * H + in];
  }
  float* b1 = reinterpret_cast<float*>(dep + D_B1);
  float* b2 = reinterpret_cast<float*>(dep + D_B2);
  float* w3 = reinterpret_cast<float*>(dep + D_W3);
  float* b3 = reinterpret_cast<float*>(dep + D_B3);
  if (t < H) {
    gp(b1)[t] = (float)(_Float16)gp(params)[OFF_B1 + t];
    gp(b2)[t] = (float)(_Float16)gp(params)[OFF_B2 + t];
    gp(w3)[t] = (float)(_Float16)gp(params)[OFF_W3 + t];
  }
  if (t == 0) gp(b3)[0] = (float)(_Float16)gp(params)[OFF_B3];
}

// inference with the deployed block: each wave takes 16 rows
__global__ void __launch_bounds__(256) k_mlp_predict(const unsigned char* dep, const float* X, uint32_t n, float* out) {
  __shared__ __align__(16) _Float16 feat[4][16 * 16];
  __shared__ __align__(16) _Float16 hbuf[4][16 * 64];
  __shared__ float res[4][16];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const uint64_t r0 = ((uint64_t)blockIdx.x * 4 + wv) * 16;
  for (int q = l; q < 16 * 16; q += 64) {
    const int row = q >> 4, col = q & 15;
    const uint64_t rr = r0 + row;
    feat[wv][q] = (_Float16)((col < IN && rr < n) ? gp(X)[rr * IN + col] : 0.f);
  }
  __builtin_amdgcn_wave_barrier();
  mlp_rows16(feat[wv], hbuf[wv], res[wv], dep);
  __builtin_amdgcn_wave_barrier();
  if (l < 16 && r0 + l < n) gp(out)[r0 + l] = res[wv][l];
}

// Learned refinement of the stored peak records of the short fused kernels:
// one wave per channel, 16 peaks per MFMA pass, features from the rows the
// fused kernel wrote (mlp_store_features).  A prediction inside [-0.5, 0.5]
// replaces the record's parabolic offset and refined frequency; anything
// else (outside, not finite) keeps them.
// Feature rows [batch][kcap][16]; row ch refines the records of channel
// och = chan ? chan[ch] : ch (peaks[och][kcap], counts[och]), whose bin width
// is fs / lens[och] when lens is given, else bin_hz.
__global__ void __launch_bounds__(256) k_mlp_refine(const unsigned char* dep, const _Float16* feat,
                                                    MaxValue* peaks, const uint32_t* counts, uint32_t batch,
                                                    uint32_t kcap, float bin_hz, const uint32_t* chan,
                                                    const uint32_t* lens, float fs) {
  __shared__ __align__(16) _Float16 fbuf[4][16 * 16];
  __shared__ __align__(16) _Float16 hbuf[4][16 * 64];
  __shared__ float res[4][16];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const uint32_t ch = blockIdx.x * 4 + wv;
  if (ch >= batch) return;  // whole waves; the kernel has no workgroup barrier
  const uint64_t och = chan != nullptr ? (uint64_t)gp(chan)[ch] : (uint64_t)ch;
  const uint32_t n = min(gp(counts)[2 * och], kcap);
  if (n == 0) return;
  const float bw = lens != nullptr ? fs / (float)gp(lens)[och] : bin_hz;
  const MlpRegs w = mlp_load(dep);
  mlp_refine_rows(w, feat + (uint64_t)ch * kcap * 16, peaks + och * kcap, n, bw, fbuf[wv], hbuf[wv], res[wv]);
}

}  // namespace spx

extern "C" {

hipError_t spx_launch_mlp_refine(const void* dep, const void* feat, void* peaks, const uint32_t* counts,
                                 uint32_t batch, uint32_t kcap, float bin_hz, hipStream_t st, const uint32_t* chan,
                                 const uint32_t* lens, float fs) {
  if (batch == 0 || kcap == 0) return hipSuccess;
  hipLaunchKernelGGL(spx::k_mlp_refine, dim3((batch + 3) / 4), dim3(256), 0, st, (const unsigned char*)dep,
                     (const _Float16*)feat, (spx::MaxValue*)peaks, counts, batch, kcap, bin_hz, chan, lens, fs);
  return hipGetLastError();
}

int spx_mlp_nparam(void) { return spx::mlp::NPARAM; }
int spx_mlp_deployed_bytes(void) { return spx::mlp::D_BYTES; }

hipError_t spx_launch_mlp_dataset(const void* spec, const float* mag16, const void* peaks, const uint32_t* counts,
                                  uint32_t B, uint32_t N, uint32_t K, float* feat, float* label, float* parab,
                                  uint32_t* valid, hipStream_t st) {
  const uint64_t tot = (uint64_t)B * K;
  if (tot == 0) return hipSuccess;
  hipLaunchKernelGGL(spx::k_mlp_dataset, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     (const float2*)spec, mag16, (const spx_max_value*)peaks, counts, B, N, K, feat, label, parab,
                     valid);
  return hipGetLastError();
}

// partial: [groups][NPARAM + 1]; grad: [NPARAM + 1] (last entry = loss)
hipError_t spx_launch_mlp_grad(const float* params, const float* X, const float* Y, uint32_t n, float* partial,
                               int groups, float* grad, hipStream_t st) {
  hipLaunchKernelGGL(spx::k_mlp_grad, dim3(groups), dim3(256), 0, st, params, X, Y, n, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(spx::k_mlp_reduce, dim3((spx::mlp::NPARAM + 1 + 255) / 256), dim3(256), 0, st,
                     (const float*)partial, (uint32_t)groups, grad);
  return hipGetLastError();
}

hipError_t spx_launch_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                               float eps, int step, hipStream_t st) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(spx::k_mlp_adam, dim3((spx::mlp::NPARAM + 255) / 256), dim3(256), 0, st, params, grad, m, v,
                     lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t spx_launch_mlp_export(const float* params, void* dep, hipStream_t st) {
  hipLaunchKernelGGL(spx::k_mlp_export, dim3(1), dim3(256), 0, st, params, (unsigned char*)dep);
  return hipGetLastError();
}

hipError_t spx_launch_mlp_predict(const void* dep, const float* X, uint32_t n, float* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(spx::k_mlp_predict, dim3((n + 63) / 64), dim3(256), 0, st, (const unsigned char*)dep, X, n, out);
  return hipGetLastError();
}

}  // extern "C"
