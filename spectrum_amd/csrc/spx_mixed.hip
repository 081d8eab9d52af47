// Mixed-precision FFT -> |X| -> peak search for N = 16384 = 16 x 32 x 32 on
// the gfx950 matrix cores: every butterfly stage is a real DFT-as-GEMM on
// MFMA (v_mfma_f32_16x16x32_{f16,bf16}) with FP32 accumulation; twiddles,
// scaling, |X| and the peak rules run in FP32 on the vector units.
//
// Index map (Cooley-Tukey, natural-order input):
//   n = 1024*n1 + n2,  n2 = 32*m1 + m2;   k = k1 + 16*(j1 + 32*j2)
//   A [k1][n2]     = sum_n1 x[1024 n1 + n2] W16^(n1 k1)          (stage 1, DFT-16)
//   A'[k1][n2]     = A W_16384^(n2 k1)
//   B [k1][j1][m2] = sum_m1 A'[k1][32 m1 + m2] W32^(m1 j1)       (stage 2, DFT-32)
//   B'             = B W_1024^(m2 j1)
//   X[k]           = sum_m2 B'[k1][j1][m2] W32^(m2 j2)            (stage 3, DFT-32)
// A complex DFT-R is the real 2R x 2R matrix [[C, S], [-S, C]] (C = cos,
// S = sin of 2 pi r q / R) acting on [Re; Im]; the data of 16 sequences are
// the B operand (K = 32 real inputs per MFMA), the DFT matrix the A operand,
// with its K columns permuted to the order (input, re/im): a lane's eight K
// rows are then four complex inputs exactly as they sit in memory (half2
// words), with no splitting into re and im vectors.
// MFMA 16x16x32 lane layout (checked on the device): lane l holds
// A[l%16][8(l/16) + j], B[8(l/16) + j][l%16], D[4(l/16) + i][l%16].
//
// Precision: the input (fp16 or bf16 storage) is multiplied by the window and
// a per-channel power-of-two pre-scale s (max component * s < 2^11), and
// every stage's operands are fp16: stage-1 outputs stay below 16*sqrt(2)*2^11,
// stage-2 outputs are scaled by 2^-5 (exact) before rounding, so no operand
// exceeds the fp16 range.  The fp16 variant uses the fp16-rounded DFT
// matrices directly (data rounding, 2^-11 per stage, dominates); the bf16
// variant splits both operands into hi + lo (hi*hi + hi*lo + lo*hi, three
// MFMAs), which reproduces the fp16 data and a 16-bit matrix.  Twiddles
// W_16384^e come from two 128-entry LDS tables (e = 128 e_hi + e_lo), the
// stage-2 twiddles from a 1024-entry LDS table; every table is generated in
// FP64 on the host and rounded once.  The window is evaluated in FP32 from
// its cosine coefficients: a three-term cosine recurrence over each lane's
// samples, started from the hardware cosine (in revolutions), with the
// pre-scale folded into the coefficients.
#include "spx_fused_common.hpp"
#include "spx_mlp.hpp"

namespace spx {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace mx {
constexpr int N = 16384;
constexpr int T = 512;             // 8 waves
// stage-2 operand row stride in half2 words (144 B: 16-byte aligned rows for
// the operand reads; the stage-1 stores, one word per lane, then meet 4-way
// bank conflicts, which measured cheaper than a conflict-free layout read
// one word at a time)
constexpr int PSTR = 36;
constexpr int PING_BYTES = 512 * PSTR * 4;     // [k1][m2] rows of 32 (+4) half2 {re, im}
// [j1][k1] rows of 32 (+4) half2, row r shifted by 16 (r >> 6) words
constexpr int PONG_BYTES = (512 * PSTR + 112) * 4;
// spectrum plane, bin k at k + k/32 + 16 (k >> 11): conflict-free per-thread
// runs for the sweep and conflict-free stage-3 stores
constexpr int XPAD = N + N / 32 + 112;
constexpr int TB_LDS = PING_BYTES + PONG_BYTES;   // 2^-5 W_1024^(m2 j1), [j1][m2] float2
constexpr int TAL_LDS = TB_LDS + 1024 * 8;        // W_16384^e, e < 128
constexpr int TAH_LDS = TAL_LDS + 128 * 8;        // W_128^e, e < 128
constexpr int LDS_BYTES = TAH_LDS + 128 * 8;
static_assert(XPAD * 4 <= PING_BYTES && XPAD * 4 <= PONG_BYTES, "FP32 spectrum planes fit the operand buffers");
// global table block (bytes): TB [32][32] float2 | TAL [128] float2 | TAH [128] float2 | W16 frags | W32 frags
constexpr int TB_OFF = 0;
constexpr int TAL_OFF = TB_OFF + 1024 * 8;
constexpr int TAH_OFF = TAL_OFF + 128 * 8;
constexpr int W16_OFF = TAH_OFF + 128 * 8;
constexpr int W16_BYTES = 2 * 8 * 2 * 64 * 16;  // [hi|lo][block t 8][rb 2][lane 64][8 x 16 bit]
constexpr int W32_OFF = W16_OFF + W16_BYTES;
constexpr int W32_BYTES = 2 * 8 * 64 * 16;      // [hi|lo][rb 4 x ks 2][lane 64][8 x 16 bit]
constexpr int TABLE_BYTES = W32_OFF + W32_BYTES;
}  // namespace mx

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// D += W * x for one 16x16x32 block; W = whi (+ wlo).
template <int CT>
struct Mma;
template <>
struct Mma<0> {  // fp16 operands; the fp16 DFT matrix (2^-12 rounding) is used unsplit
  using W = f16x8;
  static constexpr bool SPLIT = false;
  __device__ __forceinline__ static f32x4 run(const W& whi, const W&, const f16x8& x, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(whi, x, acc, 0, 0, 0);
  }
};
template <>
struct Mma<1> {  // bf16 operands, both split into hi + lo
  using W = bf16x8;
  static constexpr bool SPLIT = true;
  __device__ __forceinline__ static f32x4 run(const W& whi, const W& wlo, const f16x8& x, f32x4 acc) {
    bf16x8 xh, xl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float f = (float)x[j];
      xh[j] = (__bf16)f;
      xl[j] = (__bf16)(f - (float)xh[j]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(whi, xh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(whi, xl, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(wlo, xh, acc, 0, 0, 0);
  }
};

template <int DT>
__device__ __forceinline__ float2 unpack_in(uint32_t u) {
  if constexpr (DT == SPX_DTYPE_C32F16) {
    return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16)));
  } else {
    return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
  }
}

__device__ __forceinline__ uint32_t pack_h2(float re, float im) {
  const f16x2 h = {(_Float16)re, (_Float16)im};
  return __builtin_bit_cast(uint32_t, h);
}

// DFT-32 operands of lane group g for one 16-column block: the K = 64 rows
// are ordered (m, re/im), so k-step 0 holds inputs 4g .. 4g+3 and k-step 1
// inputs 16 + 4g .. 16 + 4g + 3 of the row, each as its half2 {re, im} word
// (row: 32 words, 16-byte aligned)
__device__ __forceinline__ void load_ops(const uint32_t* row, int g, f16x8& b0, f16x8& b1) {
  b0 = __builtin_bit_cast(f16x8, *reinterpret_cast<const uint4*>(row + 4 * g));
  b1 = __builtin_bit_cast(f16x8, *reinterpret_cast<const uint4*>(row + 16 + 4 * g));
}

// one channel's stage-1 input for this lane: words x[1024 n1 + n2] for
// n1 = 4g + m (m < 4), n2 = 16(w + 8t) + c, as 32-bit byte offsets from the
// channel's scalar base (the (t, m) step in the base, the lane part in one
// offset VGPR); every word is loaded by exactly one lane
__device__ __forceinline__ void load_raw(const FusedArgs& a, uint32_t ch, int w, int g, int c, uint32_t (&raw)[8][4]) {
  const char* base = reinterpret_cast<const char*>(a.in) + (uint64_t)ch * (mx::N * 4u);
  const uint32_t lo = 4u * ((uint32_t)g * 4096u + 16u * (uint32_t)w + (uint32_t)c);
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m) raw[t][m] = ld_at<uint32_t>(base + (4096u * m + 512u * t), lo);
}

// Threshold peak search over the FP32 spectrum planes (bin k at k + k/32 +
// 16 (k >> 11));
// thread t owns natural bins 32t .. 32t+31.  LEARN: the records get the
// parabolic offset and the features of every stored peak go to a.mlp_feat,
// for the learned refinement that runs after the kernel.
template <bool LEARN>
__device__ __forceinline__ void mixed_peaks(const FusedArgs& a, uint32_t ch, const float* xre, const float* xim,
                                            float lb, BlockRed& red) {
  constexpr int N = mx::N, T = mx::T, E = N / T;
  const int tid = threadIdx.x;
  auto adr = [](uint32_t k) -> int { return (int)(k + (k >> 5) + 16 * (k >> 11)); };
  auto xat = [&](uint32_t k) -> float2 { const int q = adr(k); return make_float2(xre[q], xim[q]); };
  const uint32_t k0 = (uint32_t)tid * E;
  const bool full = (a.s0 == 0 && a.s1 == (uint32_t)N);  // uniform
  // one sweep: sums and argmax over the range, and the bins that are local
  // maxima above the Parseval bound lb (the threshold is not known yet); the
  // threshold test then re-reads only those bins, in noise a handful per
  // channel instead of a third of all bins
  float fs1 = 0.f, fs2 = 0.f, bm = -1.0f;
  uint32_t bk = 0xffffffffu;
  uint32_t lmax = 0;
  {
    // the thread's 32 bins sit at 33 t + e in both planes (one pad slot per
    // 32 bins): all reads are issued before the first use, then the sums,
    // argmax and local-maximum bits are selects, with no per-bin branch on
    // the (uniform) whole-band case
    float m[E];
#pragma unroll
    for (int e = 0; e < E; ++e) m[e] = mag_of(make_float2(xre[adr(k0) + e], xim[adr(k0) + e]));
    const float ml = mag_of(xat((k0 + N - 1) % N));
    const float mr = mag_of(xat((k0 + E) % N));
    if (full && a.gmax == nullptr) {
      // no global-maximum record asked for: sums and maxima only
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float yl = e == 0 ? ml : m[e - 1], yr = e == E - 1 ? mr : m[e + 1];
        fs1 += m[e];
        fs2 = fmaf(m[e], m[e], fs2);
        lmax |= (m[e] > yl && m[e] > yr && m[e] > lb) ? 1u << e : 0u;
      }
    } else if (full) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float yl = e == 0 ? ml : m[e - 1], yr = e == E - 1 ? mr : m[e + 1];
        fs1 += m[e];
        fs2 = fmaf(m[e], m[e], fs2);
        const bool up = m[e] > bm;
        bm = up ? m[e] : bm;
        bk = up ? k0 + e : bk;
        lmax |= (m[e] > yl && m[e] > yr && m[e] > lb) ? 1u << e : 0u;
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float yl = e == 0 ? ml : m[e - 1], yr = e == E - 1 ? mr : m[e + 1];
        const bool in = in_range(k0 + e, a);
        const float mv = in ? m[e] : 0.0f;
        fs1 += mv;
        fs2 = fmaf(mv, mv, fs2);
        const bool up = in && m[e] > bm;
        bm = up ? m[e] : bm;
        bk = up ? k0 + e : bk;
        lmax |= (in && m[e] > yl && m[e] > yr && m[e] > lb) ? 1u << e : 0u;
      }
    }
  }
  double s1 = fs1, s2 = fs2;
  // the scan's first barrier follows every read of the sums, and the caller's
  // barrier after the peaks every read of the scan totals
  if (a.gmax != nullptr) block_stats<T, false>(s1, s2, bm, bk, red);
  else block_sums<T, false>(s1, s2, red);
  const double cnt = (double)(a.s1 - a.s0);
  float mu, sd;
  const float thr = threshold_of(a, s1, s2, cnt, mu, sd);
  uint32_t flags = 0;
  int cpk = 0;
  for (uint32_t q = lmax; q; q &= q - 1) {
    const int e = __builtin_ctz(q);
    if (mag_of(xat(k0 + e)) > thr) {
      flags |= 1u << e;
      ++cpk;
    }
  }
  int total;
  const int off = block_excl_scan<T, false>(cpk, total, red);
  if (flags && off < (int)a.kcap) {
    MaxValue* out = a.peaks + (uint64_t)ch * a.kcap;
    int pos = off;
    while (flags && pos < (int)a.kcap) {
      const int e = __builtin_ctz(flags);
      flags &= flags - 1;
      const uint32_t k = k0 + e;
      const float yc = mag_of(xat(k));
      const float yl = mag_of(xat((k + N - 1) % N));
      const float yr = mag_of(xat((k + 1) % N));
      write_record(out + pos, k, xat(k), yl, yc, yr, a);
      if constexpr (LEARN) mlp_store_features(xat, k, (uint32_t)N, mlp_feat_row(a, ch, pos));
      ++pos;
    }
  }
  if (tid == 0) {
    if (a.counts) {
      gp(a.counts)[2 * ch + 0] = (uint32_t)min(total, (int)a.kcap);
      gp(a.counts)[2 * ch + 1] = (uint32_t)total;
    }
    if (a.stats) {
      gp(a.stats)[4 * ch + 0] = mu;
      gp(a.stats)[4 * ch + 1] = sd;
      gp(a.stats)[4 * ch + 2] = thr;
      gp(a.stats)[4 * ch + 3] = (float)(cnt > 0 ? s2 / cnt : 0.0);
    }
    if (a.gmax && bk != 0xffffffffu) {
      const float yl = mag_of(xat((bk + N - 1) % N));
      const float yr = mag_of(xat((bk + 1) % N));
      write_record(a.gmax + ch, bk, xat(bk), yl, bm, yr, a);
    }
  }
}

// MODE 0: dense FP32 spectrum to a.spec_out; MODE 1: threshold peak search.
// Persistent: workgroup b runs channels b, b + grid, ...; the next channel's
// input is loaded while the current one goes through stages 2, 3 and the peaks.
template <int DT, int CT, int MODE, bool LEARN = false>
__global__ void __launch_bounds__(512) k_mixed16k(FusedArgs a) {
  using namespace mx;
  using WT = typename Mma<CT>::W;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ BlockRed red;
  uint32_t* ping = reinterpret_cast<uint32_t*>(smem);
  uint32_t* pong = reinterpret_cast<uint32_t*>(smem + PING_BYTES);
  float* xre = reinterpret_cast<float*>(smem);                // spectrum planes, natural order padded
  float* xim = reinterpret_cast<float*>(smem + PING_BYTES);
  float2* tb = reinterpret_cast<float2*>(smem + TB_LDS);
  float2* tal = reinterpret_cast<float2*>(smem + TAL_LDS);
  float2* tah = reinterpret_cast<float2*>(smem + TAH_LDS);

  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
  const unsigned char* tabs = reinterpret_cast<const unsigned char*>(a.tw_a);
  {
    const float2* gt = reinterpret_cast<const float2*>(tabs);   // TB | TAL | TAH are contiguous
    for (int i = tid; i < 1024 + 256; i += T) tb[i] = gp(gt)[i];
  }
  // per-wave largest |re|, |im| of the words in raw: |re| and |im| as 16-bit
  // patterns (sign cleared) order like their values for fp16 and bf16
  // alike, so one packed integer max per word.  Posted for the next channel
  // ahead of the barrier that ends the current one, so the pre-scale needs no
  // barrier of its own.
  __shared__ float pmax[T / 64];
  auto post_max = [&](const uint32_t (&rw)[8][4], int wv, int ln) {
    u16x2 mxp = {0, 0};
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int m = 0; m < 4; ++m) mxp = __builtin_elementwise_max(mxp, __builtin_bit_cast(u16x2, rw[t][m] & 0x7fff7fffu));
    const uint32_t mxb = (uint32_t)max(mxp[0], mxp[1]);
    float mv = DT == SPX_DTYPE_C32F16 ? (float)__builtin_bit_cast(_Float16, (uint16_t)mxb) : __uint_as_float(mxb << 16);
    mv = wave_maxf(mv);
    if (ln == 0) pmax[wv] = mv;
  };
  uint32_t ch = blockIdx.x;
  if (ch >= a.batch) return;
  uint32_t raw[8][4];
  load_raw(a, ch, w, g, c, raw);
  post_max(raw, w, l);
  __syncthreads();

  for (; ch < a.batch; ch += gridDim.x) {
    // recompute every index from a laundered thread id so that the
    // channel-invariant window values and twiddles are not hoisted out of
    // the loop and kept live (they would not fit in registers)
    const int tid = opaque_v((int)threadIdx.x);
    const int w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    const WT* W16 = opaque(reinterpret_cast<const WT*>(tabs + W16_OFF));
    const WT* W32 = opaque(reinterpret_cast<const WT*>(tabs + W32_OFF));
    // keep the prefetched words packed until they are used
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int m = 0; m < 4; ++m) raw[t][m] = (uint32_t)opaque_v((int)raw[t][m]);
    // ---- per-channel power-of-two pre-scale from the largest component
    // (the waves' maxima were posted before the barrier that ended the
    // previous channel)
    const float isc = (a.in_scale != nullptr) ? gp(a.in_scale)[ch] : 1.0f;
    float M = 0.f;
#pragma unroll
    for (int i = 0; i < T / 64; ++i) M = fmaxf(M, pmax[i]);
    M *= fabsf(isc);
    int e2 = 0;
    if (M > 0.f && M < __builtin_inff()) (void)frexpf(M, &e2);
    e2 = max(min(e2, 100), -100);
    const float s_in = ldexpf(isc, 11 - e2);   // max component * s_in < 2^11
    const float s_out = ldexpf(1.0f, e2 - 6);  // undoes s_in / isc and the stage-2 2^-5

    // ---- stage 1: DFT-16 over n1, twiddle W_16384^(n2 k1), store half2 rows [k1][m2][m1]
    {
      // The twiddle W_16384^(n2 k1), n2 = 128 t + 16 w + c, is W_128^(t k1)
      // (one factor per output row k1 of block t: folded into block t's
      // DFT-16 matrix, W16^(n1 k1) W_128^(t k1) = W_128^(k1 (8 n1 + t)))
      // times the lane's column factor pc[i] = W_16384^((16 w + c)(4 g + i)),
      // the same for every block.
      float2 pc[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = (16 * w + c) * (4 * g + i);  // < 2048
        pc[i] = c_mul(tal[e & 127], tah[e >> 7]);
      }
      // window times pre-scale, w(n) s = c0 s - c1 s cos(th) + c2 s cos(2 th),
      // th = 2 pi n / (N - 1), as the polynomial P(cos th) = (c0 - c2) s +
      // cos th (-c1 s + 2 c2 s cos th); the lane's 4 samples of block t are
      // n0 + 1024 m, so cos th follows the recurrence
      // cos((m+1) d) = 2 cos(d) cos(m d) - cos((m-1) d) (FP32, 3 steps)
      const float winv = 1.0f / (float)(N - 1);
      const float dr = 1024.0f * winv;  // revolutions per step
      const float two_cd = 2.0f * __builtin_amdgcn_cosf(dr);
      const bool win = a.wkind != 0;
      const float p0 = win ? (a.wc0 - a.wc2) * s_in : s_in;
      const float p1 = win ? -a.wc1 * s_in : 0.0f;
      const float p2 = win ? 2.0f * a.wc2 * s_in : 0.0f;
      float en = 0.f;  // energy of the fp16 operands (MODE 1)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int n2 = 16 * (w + 8 * t) + c;
        const float r0 = (float)((4 * g) * 1024 + n2) * winv;
        float cp = __builtin_amdgcn_cosf(r0 - dr), cc = __builtin_amdgcn_cosf(r0);
        // K = 32 operand rows ordered (n1, re/im): the lane's rows 8g .. 8g+7
        // are re, im of its four samples, i.e. its four words times w(n) s
        // (FP32 product, one rounding to fp16)
        uint32_t bw[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const float ws = fmaf(fmaf(p2, cc, p1), cc, p0);
          const float2 x = unpack_in<DT>(raw[t][m]);
          const f16x2 h = {(_Float16)__builtin_fmaf(x.x, ws, 0.0f), (_Float16)__builtin_fmaf(x.y, ws, 0.0f)};
          bw[m] = __builtin_bit_cast(uint32_t, h);
          if constexpr (MODE == 1) en = __builtin_amdgcn_fdot2(h, h, en, false);
          const float cn = fmaf(two_cd, cc, -cp);
          cp = cc;
          cc = cn;
        }
        const f16x8 b = __builtin_bit_cast(f16x8, make_uint4(bw[0], bw[1], bw[2], bw[3]));
        const WT w0h = gp(W16)[(2 * t + 0) * 64 + l], w1h = gp(W16)[(2 * t + 1) * 64 + l];
        WT w0l = w0h, w1l = w1h;
        if constexpr (Mma<CT>::SPLIT) {
          w0l = gp(W16)[1024 + (2 * t + 0) * 64 + l];
          w1l = gp(W16)[1024 + (2 * t + 1) * 64 + l];
        }
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        const f32x4 dre = Mma<CT>::run(w0h, w0l, b, z);
        const f32x4 dim = Mma<CT>::run(w1h, w1l, b, z);
        const int m1 = n2 >> 5, m2 = n2 & 31;
        // row k1 = 4g + i of [k1][m2][m1]: base + i * 32 * PSTR (an immediate)
        uint32_t* pb = ping + (128 * g + m2) * PSTR + m1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float2 y = c_mul(make_float2(dre[i], dim[i]), pc[i]);
          pb[i * 32 * PSTR] = pack_h2(y.x, y.y);
        }
      }
      if constexpr (MODE == 1) {
        en = dpp::lane63(dpp::reduce63((double)en, [](double x, double y) { return x + y; }));
        if (l == 0) red.s2[w] = en;
      }
    }
    __syncthreads();
    // lower bound of the threshold (full band, threshold mode): by Parseval
    // mean |X|^2 = E 2^(2(e2 - 11)) (E = energy of the fp16 operands, the
    // transform's X = DFT(operands) 2^(e2 - 11)), and
    // mean + k std >= min(1, k) sqrt(mean |X|^2); 0.99 of it leaves room for
    // the transform's rounding (~1e-3 in energy).  -1: no bound.
    float lb = -1.0f;
    if constexpr (MODE == 1) {
      if (a.peak_mode == 1 && a.s0 == 0 && a.s1 == (uint32_t)N && a.nsigma >= 0.0f) {
        double E = 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i) E += red.s2[i];
        const double bnd = 0.99 * fmin(1.0, (double)a.nsigma) * sqrt(E) * ldexp(1.0, e2 - 11);
        lb = (bnd >= 0.0 && bnd < 3.0e38) ? (float)bnd : -1.0f;
      }
    }
    // the next channel's input streams in while the rest of this channel runs
    // (from here for the dense output; after stage 3 for the peak search,
    // whose registers it would otherwise crowd)
    if constexpr (MODE == 0) {
      const uint32_t nx = ch + gridDim.x;
      if (nx < a.batch) load_raw(a, nx, w, g, c, raw);
    }

    // DFT-32 matrix fragments: [rb 0..3][ks 0..1]
    WT wh[8], wl[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      wh[f] = gp(W32)[f * 64 + l];
      wl[f] = wh[f];
      if constexpr (Mma<CT>::SPLIT) wl[f] = gp(W32)[512 + f * 64 + l];
    }

    // ---- stage 2: DFT-32 over m1 for the 512 columns (k1, m2), twiddle, store rows [j1][k1][m2]
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int cb = w + 8 * t;
      const int k1 = cb >> 1, m2 = 16 * (cb & 1) + c;
      f16x8 b0, b1;
      load_ops(ping + (k1 * 32 + m2) * PSTR, g, b0, b1);
      float2 twb[8];  // stage-2 twiddles of the block's outputs, read before its stores
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) twb[4 * h + i] = tb[(16 * h + 4 * g + i) * 32 + m2];
      f32x4 d[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        d[rb] = Mma<CT>::run(wh[rb * 2 + 0], wl[rb * 2 + 0], b0, z);
        d[rb] = Mma<CT>::run(wh[rb * 2 + 1], wl[rb * 2 + 1], b1, d[rb]);
      }
      // row j1 = 16h + 4g + i of [j1][k1][m2]: base + (16h + i) * 16 * PSTR
      // row j1 * 16 + k1 shifted by 16 (j1 >> 2) = 64 h + 16 g words: the lane
      // groups g land 16 banks apart (4-way conflicts without the shift)
      uint32_t* qb = pong + (64 * g + k1) * PSTR + m2 + 16 * g;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float2 y = c_mul(make_float2(d[h][i], d[2 + h][i]), twb[4 * h + i]);  // tb carries the 2^-5
          qb[(16 * h + i) * 16 * PSTR + 64 * h] = pack_h2(y.x, y.y);
        }
    }
    __syncthreads();

    // ---- stage 3: DFT-32 over m2 for the 512 columns (j1, k1).  The re plane of
    // the spectrum goes to the (free) stage-2 buffer at once; the im plane
    // overlays the stage-3 operands and is held until every wave has read them.
    f32x4 dim3[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j1 = w + 8 * t;
      f16x8 b0, b1;
      load_ops(pong + (j1 * 16 + c) * PSTR + 16 * (j1 >> 2), g, b0, b1);
      f32x4 d[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        d[rb] = Mma<CT>::run(wh[rb * 2 + 0], wl[rb * 2 + 0], b0, z);
        d[rb] = Mma<CT>::run(wh[rb * 2 + 1], wl[rb * 2 + 1], b1, d[rb]);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int j2 = 16 * h + 4 * g + i;
          if constexpr (MODE == 0) {
            gp(a.spec_out + (uint64_t)ch * N)[c + 16 * j1 + 512 * j2] =
                make_float2(d[h][i] * s_out, d[2 + h][i] * s_out);
          } else {
            // k = c + 16 j1 + 512 j2 sits at k + k/32 + 16 (k >> 11)
            //   = c + 16 j1 + j1/2 + 528 j2 + 64 h + 16 g  (j2 = 16 h + 4 g + i)
            xre[c + 16 * j1 + (j1 >> 1) + (528 * 4 + 16) * g + 528 * (16 * h + i) + 64 * h] = d[h][i] * s_out;
          }
        }
      dim3[t][0] = d[2];
      dim3[t][1] = d[3];
    }
    if constexpr (MODE == 0) post_max(raw, w, l);  // the next channel's words (loaded after stage 1)
    __syncthreads();  // every stage-3 operand has been read
    if constexpr (MODE == 1) {
      {
        const uint32_t nx = ch + gridDim.x;
        if (nx < a.batch) load_raw(a, nx, w, g, c, raw);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int j1 = w + 8 * t;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            xim[c + 16 * j1 + (j1 >> 1) + (528 * 4 + 16) * g + 528 * (16 * h + i) + 64 * h] = dim3[t][h][i] * s_out;
      }
      __syncthreads();
      mixed_peaks<LEARN>(a, ch, xre, xim, lb, red);
      post_max(raw, w, l);  // the next channel's words (loaded after stage 3)
      __syncthreads();  // the planes are overwritten by the next channel
    }
  }
}
template <int DT, int CT, int MODE, bool LEARN>
static hipError_t launch_mixed_t(const FusedArgs& a, int grid, hipStream_t st) {
  static bool attr_done = false;
  if (!attr_done) {
    hipError_t e = hipFuncSetAttribute((const void*)k_mixed16k<DT, CT, MODE, LEARN>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, mx::LDS_BYTES);
    if (e != hipSuccess) return e;
    attr_done = true;
  }
  hipLaunchKernelGGL((k_mixed16k<DT, CT, MODE, LEARN>), dim3(grid), dim3(mx::T), mx::LDS_BYTES, st, a);
  return hipGetLastError();
}

template <int DT, int CT>
static hipError_t launch_mixed_m(const FusedArgs& a, int grid, hipStream_t st) {
  if (a.peak_mode == 0) return launch_mixed_t<DT, CT, 0, false>(a, grid, st);
  if (a.refine == 2) return launch_mixed_t<DT, CT, 1, true>(a, grid, st);
  return launch_mixed_t<DT, CT, 1, false>(a, grid, st);
}

}  // namespace spx

extern "C" {

uint64_t spx_mixed_table_bytes(void) { return (uint64_t)spx::mx::TABLE_BYTES; }

// compute: 0 = fp16 MFMA, 1 = bf16 MFMA (split operands).  a->tw_a holds the
// mixed-precision table block; a->nfft must be 16384 and a->dtype fp16/bf16;
// grid = persistent workgroups (one per CU).
hipError_t spx_launch_mixed(const FusedArgs* a, int compute, int grid, hipStream_t st) {
  using namespace spx;
  if (a->nfft != (uint32_t)mx::N || a->n_point != (uint32_t)mx::N || grid <= 0) return hipErrorInvalidValue;
  if (a->dtype == SPX_DTYPE_C32F16)
    return compute ? launch_mixed_m<SPX_DTYPE_C32F16, 1>(*a, grid, st) : launch_mixed_m<SPX_DTYPE_C32F16, 0>(*a, grid, st);
  if (a->dtype == SPX_DTYPE_C32BF16)
    return compute ? launch_mixed_m<SPX_DTYPE_C32BF16, 1>(*a, grid, st)
                   : launch_mixed_m<SPX_DTYPE_C32BF16, 0>(*a, grid, st);
  return hipErrorInvalidValue;
}

}  // extern "C"
