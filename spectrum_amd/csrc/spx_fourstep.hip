// Threshold peak search of long channels (N = 256*N2 >= 32768) as three
// launches over a chunk of channels, many small workgroups per CU:
//   k_fs_cols   column transforms (N2 points) times W_N^(n1*k2) -> Z[k2][n1],
//               per-workgroup input energy (Parseval)
//   k_fs_rows   256-point row transforms, |X|, per-block sums, best bin and
//               the local maxima that can clear the threshold; a maximum on
//               a block's first or last row is kept with its outer neighbour
//               pending (NaN), and those two |X| rows go to an edge buffer
//   k_fs_finish one workgroup per channel: resolves pending neighbours,
//               statistics, threshold, first K survivors in ascending order
// The kernel boundaries order every exchange between workgroups (no flags,
// no waits), and several resident workgroups per CU overlap one workgroup's
// loads with another's arithmetic.
#include "spx_fs_common.hpp"
#include "spx_mlp.hpp"

#include <cstring>

namespace spx {



// Column transforms of MC channels of the same column block
// (CB columns of N2 rows) per workgroup: the window values and the four-step
// twiddles W_N^(n1 k2) of the block are read once into registers and serve
// every channel (they are the same for all channels, and re-reading them per
// channel put as many bytes through the vector-memory pipeline as the input
// itself).  The arithmetic of a channel does not depend on its neighbours.
// N2 = 256 uses the half-width exchange fft256_cols16_halfx, other lengths
// the generic column tile; both read their pass twiddles from LDS (twl).
template <int N2, int DT, int TC>
__device__ __forceinline__ void fs_cols_multi(const FusedArgs& a, const FsWork& w, float2* tile, const float2* twl,
                                              double* red) {
  using C = FsCfg<N2, TC>;
  constexpr int N1 = C::N1, CB = C::CB, NT1 = C::NT1, PPT = C::PPT, MC = C::MC;
  constexpr int RL1 = Plan<N2, 0>::RL;
  constexpr bool HALFX = (N2 == 256 && CB == 16);
  using LAY1 = typename C::LAY1;
  constexpr uint32_t ES = DT == SPX_DTYPE_C64 ? 8 : 4;
  const int tid = threadIdx.x;
  const uint32_t cg = blockIdx.x / NT1, t = blockIdx.x % NT1;
  const int c = tid % CB, jj = tid / CB;  // jj in [0, N2/16)
  const uint32_t n0 = t * CB + c + N1 * jj, stride = N1 * (N2 / 16);
  const uint32_t last = a.n_point - 1;
  float wv[PPT];
#pragma unroll
  for (int r = 0; r < PPT; ++r) wv[r] = a.window != nullptr ? gp(a.window)[min(n0 + r * stride, last)] : 1.0f;
  // output positions of v[b*RL1 + r]: Z element (jb + r*N2/RL1)*256 + column
  uint32_t zoff[PPT / RL1];
  float2 t4[PPT];
  if constexpr (RL1 <= 4) {
    // W_N^(n1 (jb + r N2/RL1)) = W_N^(n1 jb) W_N^(n1 r N2/RL1): one table
    // entry per b and RL1 - 1 factors per thread (its column n1 is fixed),
    // 9 or 7 loads instead of 16 (256 x 2^18 peak search 0.521 -> 0.507 ms,
    // complex spectrum 0.593 -> 0.574 ms; 2^17 within noise)
    float2 ur[RL1];
    {
      int s, jb;
      LAY1::template map<RL1>(tid, s, jb);
#pragma unroll
      for (int r = 1; r < RL1; ++r) ur[r] = gp(a.tw_4)[t * CB + s + (uint32_t)(r * (N2 / RL1)) * N1];
    }
#pragma unroll
    for (int b = 0; b < PPT / RL1; ++b) {
      int s, jb;
      LAY1::template map<RL1>(tid + b * TC, s, jb);
      zoff[b] = (uint32_t)jb * N1 + t * CB + s;
      const float2 w0 = gp(a.tw_4)[zoff[b]];
      t4[b * RL1] = w0;
#pragma unroll
      for (int r = 1; r < RL1; ++r) t4[b * RL1 + r] = c_mul(w0, ur[r]);
    }
  } else {
#pragma unroll
    for (int b = 0; b < PPT / RL1; ++b) {
      int s, jb;
      LAY1::template map<RL1>(tid + b * TC, s, jb);
      zoff[b] = (uint32_t)jb * N1 + t * CB + s;
#pragma unroll
      for (int r = 0; r < RL1; ++r) t4[b * RL1 + r] = gp(a.tw_4)[zoff[b] + (uint32_t)(r * (N2 / RL1)) * N1];
    }
  }
#pragma unroll 1
  for (int m = 0; m < MC; ++m) {
    const uint32_t ci = cg * MC + m;
    if (ci >= w.nch) break;  // uniform over the workgroup
    const uint64_t ich = (uint64_t)w.ch0 + ci;
    // the caller's channel id: through the indirection when there is one
    const uint64_t uch = (w.chan != nullptr) ? (uint64_t)gp(w.chan)[ich] : ich;
    const char* cb = (w.chan != nullptr) ? reinterpret_cast<const char*>(a.in) + gp(w.inoff)[uch] * ES
                                         : reinterpret_cast<const char*>(a.in) + ich * (uint64_t)a.n_point * ES;
    const float sc = (a.in_scale != nullptr) ? gp(a.in_scale)[uch] : 1.0f;
    float2 v[PPT];
#pragma unroll
    for (int r = 0; r < PPT; ++r) {
      const uint32_t n = n0 + r * stride;
      const float2 raw = ld_raw<DT, C::STREAM_Z>(cb, min(n, last));
      v[r] = c_scale(raw, n < a.n_point ? wv[r] * sc : 0.0f);
    }
    float e = 0.f;
#pragma unroll
    for (int r = 0; r < PPT; ++r) e = fmaf(v[r].x, v[r].x, fmaf(v[r].y, v[r].y, e));
    if (m > 0) __syncthreads();  // the previous channel's exchange reads are done
    // laundered thread id and table pointers: the transform's twiddle rows
    // are re-read per channel instead of being kept live across the loop
    if constexpr (HALFX) fft256_cols16_halfx<true>(reinterpret_cast<float*>(tile), v, opaque(twl), opaque_v(tid));
    else fft_tile<LAY1, N2, CB, TC, PPT, 0, true>(tile, v, opaque(twl), opaque_v(tid));
    float2* Z = w.Z + (uint64_t)ci * C::N;
#pragma unroll
    for (int b = 0; b < PPT / RL1; ++b)
#pragma unroll
      for (int r = 0; r < RL1; ++r)
        gp(Z)[zoff[b] + (uint32_t)(r * (N2 / RL1)) * N1] = c_mul(v[b * RL1 + r], t4[b * RL1 + r]);
    const double E = block_sum_d<TC>((double)e, red);
    if (tid == 0) {
      gp(w.pe)[(uint64_t)ci * NT1 + t] = E;
      if (t == 0) {
        gp(w.ncand)[2 * ci + 0] = 0u;
        gp(w.ncand)[2 * ci + 1] = 0u;
      }
    }
  }
}

// ---------------------------------------------------------------------------
template <int N2, int TC>
__global__ void __launch_bounds__(TC) __attribute__((amdgpu_waves_per_eu(FsCfg<N2, TC>::COLS_WPE)))
k_fs_cols(FusedArgs a, FsWork w) {
  using C = FsCfg<N2, TC>;
  using LAY1 = typename C::LAY1;
  constexpr bool HALFX = (N2 == 256 && C::CB == 16);
  __shared__ __align__(16) float2 tile[HALFX ? 5104 / 2 : LAY1::SIZE];
  // the column plan's twiddle rows (read only after the first exchange's
  // barriers; from L2 per channel they would load the vector-memory
  // address unit as much as the input itself)
  __shared__ __align__(16) float2 twl[Plan<N2, 0>::TWSIZE_LDS];
  __shared__ double red[TC / 64];
  copy_tw_to_lds<N2, 0, TC>(a.tw_b, twl, threadIdx.x);
  if (a.dtype == SPX_DTYPE_C64) fs_cols_multi<N2, SPX_DTYPE_C64, TC>(a, w, tile, twl, red);
  else if (a.dtype == SPX_DTYPE_C32F16) fs_cols_multi<N2, SPX_DTYPE_C32F16, TC>(a, w, tile, twl, red);
  else fs_cols_multi<N2, SPX_DTYPE_C32BF16, TC>(a, w, tile, twl, red);
}


// ---------------------------------------------------------------------------
// LEARN: also the complex neighbourhood of every candidate (w.candx) and the
// complex rows 0, 1, RB-2, RB-1 of the block (w.edgex), for the refiner's
// features in the finish.
template <int N2, int TC, bool LEARN = false>
__global__ void __launch_bounds__(256) k_fs_rows(FusedArgs a, FsWork w) {
  using C = FsCfg<N2, TC>;
  constexpr int N1 = C::N1, RB = C::RB, NT2 = C::NT2, PPT = C::PPT, T = C::TR;
  using LAY2 = typename C::LAY2;
  // LEARN: the transform's float tile aliases the complex tile, which then
  // holds X of the block (|X| is recomputed from it, bit for bit the same)
  __shared__ __align__(16) float ldsf[LEARN ? 1 : LAY2::SIZE];
  __shared__ __align__(16) float2 cxt[LEARN ? LAY2::SIZE : 1];
  float* const tile = LEARN ? reinterpret_cast<float*>(cxt) : ldsf;
  auto mt = [&](int row, int k1) -> float {
    if constexpr (LEARN) return mag_of(cxt[LAY2::idx(row, k1)]);
    else return ldsf[LAY2::idx(row, k1)];
  };
  __shared__ __align__(16) float2 twl[Plan<256, 0>::TWSIZE_LDS];
  __shared__ BlockRed red;
  const int tid = threadIdx.x;
  const uint32_t ci = blockIdx.x / NT2, t = blockIdx.x % NT2;
  const int s_row = tid / (N1 / 16), jj_row = tid % (N1 / 16);
  float2 v[PPT];
  {
    const float2* zr = w.Z + (uint64_t)ci * C::N + (uint32_t)((t * RB + s_row) * N1 + jj_row);
#pragma unroll
    for (int r = 0; r < PPT; ++r) v[r] = ld_f2<C::STREAM_Z>(zr + r * (N1 / 16));
  }
  const bool full_band = (a.s0 == 0 && a.s1 == (uint32_t)C::N);
  const FsBounds bnd = fs_bounds(a, fs_energy<N2, TC>(w, ci), full_band);
  const float lb = bnd.lb;
  copy_tw_to_lds<256, 0, T>(a.tw_a, twl, tid);
  __syncthreads();
  fft256_rows_halfx<RB, T, true>(tile, v, twl, tid);
  // v[r] = X[k2 = t*RB + s_row][k1 = jj_row + 16 r]  ->  bin k2 + N2*k1 (ascending in r)
  const uint32_t k2 = t * RB + s_row;
  if constexpr (LEARN) {
    __syncthreads();  // every wave is done with the float tile
#pragma unroll
    for (int r = 0; r < PPT; ++r) cxt[LAY2::idx(s_row, jj_row + 16 * r)] = v[r];
    const int er = s_row < 2 ? s_row : (s_row >= RB - 2 ? s_row - (RB - 4) : -1);
    if (er >= 0 && w.edgex != nullptr) {
      float2* ex = w.edgex + (((uint64_t)ci * NT2 + t) * 4 + er) * N1;
#pragma unroll
      for (int r = 0; r < PPT; ++r) gp(ex)[jj_row + 16 * r] = v[r];
    }
  }
  float m[PPT];
  float fs1 = 0.f, fs2 = 0.f;
  float bm = -1.f;  // largest in-band |X| of this thread (first r on ties: lowest bin)
  int rb = -1;
  if (full_band) {
#pragma unroll
    for (int r = 0; r < PPT; ++r) {
      m[r] = mag_of(v[r]);
      fs1 += m[r];
      fs2 = fmaf(m[r], m[r], fs2);
      const bool u = m[r] > bm;
      bm = u ? m[r] : bm;
      rb = u ? r : rb;
    }
  } else {
#pragma unroll
    for (int r = 0; r < PPT; ++r) {
      m[r] = mag_of(v[r]);
      if (in_range(k2 + (uint32_t)N2 * (uint32_t)(jj_row + 16 * r), a)) {
        fs1 += m[r];
        fs2 = fmaf(m[r], m[r], fs2);
        const bool u = m[r] > bm;
        bm = u ? m[r] : bm;
        rb = u ? r : rb;
      }
    }
  }
  if constexpr (!LEARN) {
    wave_lds_sync();  // the wave's rows of the tile are free again
#pragma unroll
    for (int r = 0; r < PPT; ++r) ldsf[LAY2::idx(s_row, jj_row + 16 * r)] = m[r];
  }
  if (s_row == 0 || s_row == RB - 1) {
    float* er = w.edge + (((uint64_t)ci * NT2 + t) * 2 + (s_row == 0 ? 0 : 1)) * N1;
#pragma unroll
    for (int r = 0; r < PPT; ++r) gp(er)[jj_row + 16 * r] = m[r];
  }
  __syncthreads();
  // neighbours k -/+ 1 are rows s_row -/+ 1 of the same column; outside the
  // block they are left pending (NaN) for the finish
  const bool top = (s_row == 0), bot = (s_row == RB - 1);
  const int rl = top ? s_row : s_row - 1, rr = bot ? s_row : s_row + 1;
  const float qnan = __builtin_nanf("");
  if (bm >= lb) {  // some bin of this thread may clear the threshold
    Cand* cand = w.cand + (uint64_t)ci * C::CAND_CAP;
    uint32_t mhigh = 0, mlow = 0;
    // only the bins at or above lb are tested (usually a few of the 16); a
    // bin's own |X| and its neighbours come back from the tile
    uint32_t pend = 0;
#pragma unroll
    for (int r = 0; r < PPT; ++r) pend |= (m[r] >= lb ? 1u : 0u) << r;
    while (pend) {
      const int r = __builtin_ctz(pend);
      pend &= pend - 1;
      const int k1 = jj_row + 16 * r;
      const float mc = mt(s_row, k1);
      const float yl = mt(rl, k1);
      const float yr = mt(rr, k1);
      const bool ok = full_band || in_range(k2 + (uint32_t)N2 * (uint32_t)k1, a);
      if (ok && (top || mc > yl) && (bot || mc > yr)) {
        if (mc >= bnd.hi) mhigh |= 1u << r;
        else mlow |= 1u << r;
      }
    }
    // the two lists are filled in separate loops, each with one counter
    // address for the whole wave, so that the increments of a wave are
    // combined into one atomic
    auto push = [&](uint32_t mask, uint32_t* ctr, bool high) {
      while (mask) {
        const int r = __builtin_ctz(mask);
        mask &= mask - 1;
        const int k1 = jj_row + 16 * r;
        float2 x = v[0];
        float mc = m[0];
#pragma unroll
        for (int q = 1; q < PPT; ++q) {
          x = (q == r) ? v[q] : x;
          mc = (q == r) ? m[q] : mc;
        }
        const uint32_t p = atomicAdd(ctr, 1u);
        if (p < (uint32_t)C::CAND_CAP) {
          Cand cd;
          cd.k = k2 + (uint32_t)N2 * (uint32_t)k1;
          cd.re = x.x;
          cd.im = x.y;
          cd.ml = top ? qnan : mt(rl, k1);
          cd.mc = mc;
          cd.mr = bot ? qnan : mt(rr, k1);
          const uint32_t slot = high ? p : (uint32_t)C::CAND_CAP - 1 - p;
          gp(cand)[slot] = cd;
          if constexpr (LEARN) {
            float2* cx4 = w.candx + ((uint64_t)ci * C::CAND_CAP + slot) * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int row = s_row + (j < 2 ? j - 2 : j - 1);
              gp(cx4)[j] = (row >= 0 && row < RB) ? cxt[LAY2::idx(row, k1)] : make_float2(qnan, qnan);
            }
          }
        }
      }
    };
    push(mhigh, gp(w.ncand) + 2 * ci + 0, true);
    push(mlow, gp(w.ncand) + 2 * ci + 1, false);
  }
  const uint32_t bk = rb >= 0 ? k2 + (uint32_t)N2 * (uint32_t)(jj_row + 16 * rb) : 0xffffffffu;
  double s1 = fs1, s2 = fs2;
  float gm = bm;
  uint32_t gk = bk;
  block_stats<T, false>(s1, s2, gm, gk, red);
  FsRowPart* pp = w.part + (uint64_t)ci * NT2 + t;
  if (tid == 0) {
    gp(&pp->s1)[0] = s1;
    gp(&pp->s2)[0] = s2;
    gp(&pp->m)[0] = gm;
    gp(&pp->k)[0] = gk;
  }
  if (gk != 0xffffffffu && bk == gk && bm == gm) {
    float2 x = v[0];
#pragma unroll
    for (int q = 1; q < PPT; ++q) x = (q == rb) ? v[q] : x;
    const int k1 = jj_row + 16 * rb;
    gp(&pp->xr)[0] = x.x;
    gp(&pp->xi)[0] = x.y;
    gp(&pp->ml)[0] = top ? qnan : mt(rl, k1);
    gp(&pp->mr)[0] = bot ? qnan : mt(rr, k1);
  }
}

// ---------------------------------------------------------------------------

typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));

// A refiner feature pair still to be filled: the neighbour of a stored
// record (feature row `row`, pair `slot`) that lies in another row block.
struct FsNbrReq {
  uint32_t row, k, slot;
  float cr, ci;
};
constexpr int FS_NBR_CAP = 128;  // two per record, records 0 .. 63

template <int N2, int TC, bool LEARN>
struct FsFinishShared {
  static constexpr int NW = FsCfg<N2, TC>::N / 32;  // survivor bitmap words
  uint32_t bits[NW];
  uint16_t pre[NW];
  int base[FsCfg<N2, TC>::TF];
  BlockRed red;
  FsNbrReq req[LEARN ? FS_NBR_CAP : 1];  // learned refinement only
  uint32_t nreq;
};

// X at bins kb[0..G) of the channel whose four-step intermediate is z
// (Z[k2][n1], X[k2 + N2 k1] = sum_n1 Z[k2][n1] W_256^(n1 k1)), the 256 terms
// of each shared by the wave's 64 lanes (all loads issued first); the
// results in every lane.  The twiddles come from the hardware sine and
// cosine (in revolutions): a refiner feature is rounded to fp16, so their
// ~1e-6 error does not matter there.
template <int N2, int G>
__device__ __forceinline__ void fs_bins_from_z(const float2* z, const uint32_t (&kb)[G], float2 (&out)[G]) {
  const int l = threadIdx.x & 63;
  float2 v[G][4];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const float2* row = z + (uint64_t)(kb[g] & (N2 - 1)) * 256;
#pragma unroll
    for (int q = 0; q < 4; ++q) v[g][q] = gp(row)[(uint32_t)l + 64u * q];
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const uint32_t k1 = kb[g] / N2;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t n1 = (uint32_t)l + 64u * q;
      const float t = (float)((n1 * k1) & 255u) * (1.0f / 256.0f);  // revolutions
      const float c = __builtin_amdgcn_cosf(t), sn = __builtin_amdgcn_sinf(t);
      // v * (c - i sn)
      re = fmaf(v[g][q].x, c, fmaf(v[g][q].y, sn, re));
      im = fmaf(v[g][q].y, c, fmaf(-v[g][q].x, sn, im));
    }
    auto add = [](float a, float b) { return a + b; };
    re = dpp::reduce63(re, add);
    im = dpp::reduce63(im, add);
    out[g] = make_float2(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, re), 63)),
                         __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, im), 63)));
  }
}

// The finish of one channel, one wave (TF = 64 threads) per channel: sums and
// best bin of the row blocks (one block per lane), threshold, then the first
// K survivors of the candidate list in ascending bin order.  A maximum on a
// block's first or last row has its outer neighbour pending (NaN): it is
// looked up here, and the candidate dropped (mc = NaN) when it is no maximum.
// Up to 64 candidates (one per lane) are selected with ballots, without
// barriers or LDS; longer lists go through the bitmap selection (select_first_k).
// X at bin (k + d) mod N (|d| <= 2) from the complex edge rows 0, 1, RB-2,
// RB-1 of its row block (a neighbour that lies outside the candidate's block)
template <int N2, int RB>
__device__ __forceinline__ float2 fs_outer_x(const float2* edgex, uint32_t k, int d) {
  constexpr uint32_t N = 256u * N2;
  const uint32_t kk = (k + N + (uint32_t)d) % N;
  const uint32_t k2 = kk & (N2 - 1), k1 = kk / N2, row = k2 % RB;
  const uint32_t e = row < 2 ? row : row - (RB - 4);
  return gp(edgex)[((k2 / RB) * 4 + e) * 256 + k1];
}

// LEARN: every stored record also gets its refiner feature row (a.mlp_feat,
// row ci * kcap + rank of the chunk), from the candidate's neighbourhood.
template <int N2, int TC, bool LEARN = false>
__global__ void __launch_bounds__(64) k_fs_finish(FusedArgs a, FsWork w) {
  using C = FsCfg<N2, TC>;
  constexpr int NT1 = C::NT1, NT2 = C::NT2, T = C::TF;
  static_assert(T == 64 && NT1 <= 64 && NT2 <= 64 && C::CAND_CAP >= 64, "one value per lane");
  __shared__ FsFinishShared<N2, TC, LEARN> sh;
  const int lane = threadIdx.x;
  const uint32_t ci = blockIdx.x;
  if constexpr (LEARN) {
    if (lane == 0) sh.nreq = 0u;
  }
  const uint64_t och = w.chan != nullptr ? (uint64_t)gp(w.chan)[w.ch0 + ci] : (uint64_t)w.ch0 + ci;
  const Cand* cand = w.cand + (uint64_t)ci * C::CAND_CAP;
  const float* edge = w.edge + (uint64_t)ci * NT2 * 2 * C::N1;
  // every load of the channel at once: its high-list candidate (speculative),
  // list lengths, energy and row-block part
  Cand c = gp(cand)[lane];
  const uint32_t nh = min(gp(w.ncand)[2 * ci + 0], (uint32_t)C::CAND_CAP);
  const uint32_t nl = min(gp(w.ncand)[2 * ci + 1], (uint32_t)C::CAND_CAP - nh);
  const double P = fs_energy<N2, TC>(w, ci);
  double s1 = 0.0, s2 = 0.0;
  float bm = -1.f, bl = 0.f, br = 0.f;
  uint32_t bk = 0xffffffffu;
  float2 bx = make_float2(0.f, 0.f);
  if (lane < NT2) {
    const FsRowPart* pp = w.part + (uint64_t)ci * NT2 + lane;
    s1 = gp(&pp->s1)[0];
    s2 = gp(&pp->s2)[0];
    bm = gp(&pp->m)[0];
    bk = gp(&pp->k)[0];
    bx = make_float2(gp(&pp->xr)[0], gp(&pp->xi)[0]);
    bl = gp(&pp->ml)[0];
    br = gp(&pp->mr)[0];
    if (bk == 0xffffffffu) bm = -1.f;
  }
  auto resolve = [edge](Cand& q) {
    const bool pl = __builtin_isnan(q.ml), pr = __builtin_isnan(q.mr);
    if (pl || pr) {
      const float nb = fs_outer<N2, TC>(edge, q.k, pl ? -1 : 1);
      if (!(q.mc > nb)) q.mc = __builtin_nanf("");
      if (pl) q.ml = nb;
      else q.mr = nb;
    }
  };
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  float gm = bm;
  uint32_t gk = bk;
  wave_argmax(gm, gk);
  const double cnt = (double)(a.s1 - a.s0);
  float mu, sd;
  const float thr = threshold_of(a, s1, s2, cnt, mu, sd);
  // the candidate list to select from: the high list, and the low list
  // (entries CAND_CAP-1, CAND_CAP-2, ...) unless the threshold is above it
  const bool full_band = (a.s0 == 0 && a.s1 == (uint32_t)C::N);
  const FsBounds bnd = fs_bounds(a, P, full_band);
  const uint32_t nc = (thr >= bnd.hi && bnd.hi > 0.f) ? nh : nh + nl;
  auto at = [cand, nh](uint32_t i) { return gp(cand)[i < nh ? i : (uint32_t)C::CAND_CAP - 1 - (i - nh)]; };
  if ((uint32_t)lane >= nh && (uint32_t)lane < nc) c = at(lane);
  if ((uint32_t)lane < nc) resolve(c);
  if (a.gmax && gk != 0xffffffffu && bk == gk && bm == gm) {
    if (__builtin_isnan(bl)) bl = fs_outer<N2, TC>(edge, bk, -1);
    if (__builtin_isnan(br)) br = fs_outer<N2, TC>(edge, bk, 1);
    write_record(a.gmax + och, bk, bx, bl, bm, br, a);
  }
  // refiner features of candidate i (list position), stored at record `rank`
  auto features = [&](uint32_t i, uint32_t rank, const Cand& q) {
    if constexpr (LEARN) {
      const uint32_t slot = i < nh ? i : (uint32_t)C::CAND_CAP - 1 - (i - nh);
      const float2* cx4 = w.candx + ((uint64_t)ci * C::CAND_CAP + slot) * 4;
      const float2* ex = w.edgex + (uint64_t)ci * NT2 * 4 * C::N1;
      float2 x[5];
      x[2] = make_float2(q.re, q.im);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int d = j < 2 ? j - 2 : j - 1;
        float2 v = gp(cx4)[j];
        if (__builtin_isnan(v.x)) {
          if (w.edgex != nullptr) {
            v = fs_outer_x<N2, C::RB>(ex, q.k, d);
          } else {
            // in another row block: its feature pair is filled after the
            // records, from Z (the pair is written as zero meanwhile)
            const uint32_t s = atomicAdd(&sh.nreq, 1u);
            if (s < (uint32_t)FS_NBR_CAP)
              sh.req[s] = FsNbrReq{rank, (q.k + (uint32_t)C::N + (uint32_t)d) % (uint32_t)C::N, (uint32_t)(d + 2),
                                   q.re, q.im};
            v = make_float2(0.f, 0.f);
          }
        }
        x[d + 2] = v;
      }
      const uint32_t k0 = q.k;
      mlp_store_features([&](uint32_t qq) { return x[(qq + (uint32_t)C::N + 2u - k0) % (uint32_t)C::N]; }, q.k,
                         (uint32_t)C::N, mlp_feat_row(a, ci, (int)rank));
    }
  };
  // the feature pairs of neighbours in other row blocks, one bin at a time
  // by the whole wave (at most two per record; records 0 .. 63 without edge
  // rows, the host's choice whenever max_peaks <= 64)
  auto fill_requests = [&]() {
    if constexpr (LEARN) {
      if (w.edgex != nullptr) return;
      __builtin_amdgcn_s_waitcnt(0);  // the wave's LDS appends are done
      __builtin_amdgcn_wave_barrier();
      const uint32_t n = min(sh.nreq, (uint32_t)FS_NBR_CAP);
      const float2* z = w.Z + (uint64_t)ci * C::N;
      constexpr int G = 4;  // bins per round, their loads in flight together
      for (uint32_t i0 = 0; i0 < n; i0 += G) {
        uint32_t kb[G];
#pragma unroll
        for (int g = 0; g < G; ++g) kb[g] = sh.req[min(i0 + g, n - 1)].k;
        float2 xv[G];
        fs_bins_from_z<N2, G>(z, kb, xv);
        if (lane < G && i0 + (uint32_t)lane < n) {
          const FsNbrReq r = sh.req[i0 + lane];
          float2 x = xv[0];
#pragma unroll
          for (int g = 1; g < G; ++g) x = lane == g ? xv[g] : x;
          // x conj(c) / |c|^2, as mlp_features computes it
          const float den = fmaf(r.cr, r.cr, r.ci * r.ci);
          const float inv = den > 0.f ? 1.0f / den : 0.f;
          const f16x2_t h = {(_Float16)(fmaf(x.x, r.cr, x.y * r.ci) * inv),
                             (_Float16)(fmaf(x.y, r.cr, -x.x * r.ci) * inv)};
          *gp(reinterpret_cast<f16x2_t*>(mlp_feat_row(a, ci, (int)r.row) + 2 * r.slot)) = h;
        }
      }
    }
  };
  if (nc <= 64u) {
    const bool surv = (uint32_t)lane < nc && c.mc > thr;
    uint64_t smask = __ballot(surv);
    const int total = __popcll(smask);
    // rank = survivors with a lower bin (bins are distinct)
    int rank = 0;
    while (smask) {
      const int j = __builtin_ctzll(smask);
      smask &= smask - 1;
      const uint32_t kj = __builtin_amdgcn_readlane(c.k, j);
      rank += (kj < c.k) ? 1 : 0;
    }
    const int kcap = (int)a.kcap;
    if (surv && rank < kcap) {
      write_record(a.peaks + och * a.kcap + rank, c.k, make_float2(c.re, c.im), c.ml, c.mc, c.mr, a);
      features((uint32_t)lane, (uint32_t)rank, c);
    }
    fill_requests();
    if (lane == 0) {
      if (a.counts) {
        gp(a.counts)[2 * och + 0] = (uint32_t)min(total, kcap);
        gp(a.counts)[2 * och + 1] = (uint32_t)total;
      }
      if (a.stats) {
        gp(a.stats)[4 * och + 0] = mu;
        gp(a.stats)[4 * och + 1] = sd;
        gp(a.stats)[4 * och + 2] = thr;
        gp(a.stats)[4 * och + 3] = (float)P;
      }
    }
  } else {
    select_first_k<T, FsFinishShared<N2, TC, LEARN>::NW>(a, [at, resolve](uint32_t i) {
      Cand q = at(i);
      resolve(q);
      return q;
    }, nc, thr, mu, sd, P, och, sh.bits, sh.pre, sh.base, sh.red, features);
    fill_requests();
  }
}


// Dense outputs of long channels (spectrum, |X|, |X|^2 or {|X|, arg X}) after
// k_fs_cols: the row transforms of k_fs_rows, then a transpose through the
// float tile (real parts, then imaginary parts) so that 16 lanes write 16
// consecutive bins k2 .. k2+15 of one k1 (bin k2 + N2 k1).  Element (s, k1)
// of the block sits at 17 k1 + s.
template <int N2, int TC>
__global__ void __launch_bounds__(256) k_fs_rows_dense(FusedArgs a, FsWork w) {
  using C = FsCfg<N2, TC>;
  constexpr int N1 = C::N1, RB = C::RB, NT2 = C::NT2, PPT = C::PPT, T = C::TR;
  using LAY2 = typename C::LAY2;
  static_assert(RB == 16 && 17 * (N1 - 1) + RB <= LAY2::SIZE, "transpose fits the tile");
  __shared__ __align__(16) float ldsf[LAY2::SIZE];
  __shared__ __align__(16) float2 twl[Plan<256, 0>::TWSIZE_LDS];
  const int tid = threadIdx.x;
  const uint32_t ci = blockIdx.x / NT2, t = blockIdx.x % NT2;
  const uint64_t och = (uint64_t)w.ch0 + ci;
  const int s_row = tid / (N1 / 16), jj_row = tid % (N1 / 16);
  float2 v[PPT];
  {
    const float2* zr = w.Z + (uint64_t)ci * C::N + (uint32_t)((t * RB + s_row) * N1 + jj_row);
#pragma unroll
    for (int r = 0; r < PPT; ++r) v[r] = ld_f2<C::STREAM_Z>(zr + r * (N1 / 16));
  }
  copy_tw_to_lds<256, 0, T>(a.tw_a, twl, tid);
  __syncthreads();
  fft256_rows_halfx<RB, T, true>(ldsf, v, twl, tid);
  // v[r] = X[k2 = t*RB + s_row][k1 = jj_row + 16 r]; after the transpose thread
  // tid holds X[k2 = t*RB + tid%16][k1 = tid/16 + 16 q], q < 16
  const int s_out = tid % 16, k1_out = tid / 16;
  float2 x[PPT];
  __syncthreads();  // every wave is past its transform's exchange
#pragma unroll
  for (int r = 0; r < PPT; ++r) ldsf[17 * (jj_row + 16 * r) + s_row] = v[r].x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PPT; ++q) x[q].x = ldsf[17 * (k1_out + 16 * q) + s_out];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PPT; ++r) ldsf[17 * (jj_row + 16 * r) + s_row] = v[r].y;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PPT; ++q) x[q].y = ldsf[17 * (k1_out + 16 * q) + s_out];
  const uint64_t cbase = och * (uint64_t)C::N + (uint64_t)(t * RB + s_out);
#pragma unroll
  for (int q = 0; q < PPT; ++q) store_dense<C::STREAM_Z>(a, cbase + (uint64_t)N2 * (uint32_t)(k1_out + 16 * q), x[q]);
}

template <int N2, int TC>
static hipError_t launch_fs_dense(const FusedArgs& a, const FsWork& w0, uint32_t chunk, hipStream_t st) {
  using C = FsCfg<N2, TC>;
  for (uint32_t c0 = 0; c0 < a.batch; c0 += chunk) {
    FsWork w = w0;
    w.ch0 = c0;
    w.nch = min(chunk, a.batch - c0);
    const uint32_t groups = (w.nch + C::MC - 1) / C::MC;
    hipLaunchKernelGGL((k_fs_cols<N2, TC>), dim3(groups * C::NT1), dim3(TC), 0, st, a, w);
    hipLaunchKernelGGL((k_fs_rows_dense<N2, TC>), dim3(w.nch * C::NT2), dim3(C::TR), 0, st, a, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Reference ONE_PEAK / TWO_PEAKS (peak_mode 3 / 4) on long channels, taken on
// the row transforms without storing the spectrum: k_fs_cols, then
// k_fs_rows_top (per row block and side the largest |X|, lowest bin on ties,
// with X there and at both neighbours; the block's first and last rows as
// complex edge rows), then k_fs_top_finish (one wave per channel: the
// winners over the blocks, neighbours outside their block looked up in the
// edge rows, the records of write_top_records).
// ---------------------------------------------------------------------------
struct FsTopSide {
  float m;          // largest |X| of the side in the block, -1 if none
  uint32_t k;       // its bin, 0xffffffff if none
  float2 x, xl, xr; // X at k, k-1, k+1 (xl.x / xr.x NaN: the neighbour is in another block)
};
struct FsTopPart {
  FsTopSide s[2];   // left range [0, s0), right range [s1, nfft)
};

template <int N2, int TC>
__global__ void __launch_bounds__(256) k_fs_rows_top(FusedArgs a, FsWork w) {
  using C = FsCfg<N2, TC>;
  constexpr int N1 = C::N1, RB = C::RB, NT2 = C::NT2, PPT = C::PPT, T = C::TR;
  using LAY2 = typename C::LAY2;
  // the transform's float tile first, then the block's complex spectrum
  __shared__ __align__(16) float2 cx[LAY2::SIZE];
  __shared__ __align__(16) float2 twl[Plan<256, 0>::TWSIZE_LDS];
  __shared__ BlockRed red;
  const int tid = threadIdx.x;
  const uint32_t ci = blockIdx.x / NT2, t = blockIdx.x % NT2;
  const int s_row = tid / (N1 / 16), jj_row = tid % (N1 / 16);
  float2 v[PPT];
  {
    const float2* zr = w.Z + (uint64_t)ci * C::N + (uint32_t)((t * RB + s_row) * N1 + jj_row);
#pragma unroll
    for (int r = 0; r < PPT; ++r) v[r] = ld_f2<C::STREAM_Z>(zr + r * (N1 / 16));
  }
  copy_tw_to_lds<256, 0, T>(a.tw_a, twl, tid);
  __syncthreads();
  fft256_rows_halfx<RB, T, true>(reinterpret_cast<float*>(cx), v, twl, tid);
  __syncthreads();  // every wave is done with the float tile
  // v[r] = X[k2 = t*RB + s_row][k1 = jj_row + 16 r]  ->  bin k2 + N2*k1
  const uint32_t k2 = t * RB + s_row;
  float2* edge = reinterpret_cast<float2*>(w.edge) + ((uint64_t)ci * NT2 + t) * 2 * N1;
#pragma unroll
  for (int r = 0; r < PPT; ++r) cx[LAY2::idx(s_row, jj_row + 16 * r)] = v[r];
  if (s_row == 0 || s_row == RB - 1) {
#pragma unroll
    for (int r = 0; r < PPT; ++r) gp(edge)[(s_row == 0 ? 0 : N1) + jj_row + 16 * r] = v[r];
  }
  // per side the largest |X| of the thread (ascending k: the first wins
  // ties), with X there; v is only ever indexed with unrolled constants
  float lm = -1.f, rm = -1.f;
  uint32_t li = 0xffffffffu, ri = 0xffffffffu;
  float2 lx = make_float2(0.f, 0.f), rx = lx;
#pragma unroll
  for (int r = 0; r < PPT; ++r) {
    const uint32_t k = k2 + (uint32_t)N2 * (uint32_t)(jj_row + 16 * r);
    const float m = cabs_rn(v[r]);
    const bool ul = k < a.s0 && m > lm, ur = k >= a.s1 && m > rm;
    lm = ul ? m : lm;
    li = ul ? k : li;
    lx = ul ? v[r] : lx;
    rm = ur ? m : rm;
    ri = ur ? k : ri;
    rx = ur ? v[r] : rx;
  }
  float glm = lm, grm = rm;
  uint32_t gli = li, gri = ri;
  double d0 = 0.0, d1 = 0.0;
  block_stats<T, true>(d0, d1, glm, gli, red);  // its closing barrier also orders the cx writes
  block_stats<T, false>(d0, d1, grm, gri, red);
  FsTopPart* pp = reinterpret_cast<FsTopPart*>(w.part) + (uint64_t)ci * NT2 + t;
  const float qnan = __builtin_nanf("");
  auto put = [&](int side, float m, uint32_t k, float2 x) {
    FsTopSide sd;
    sd.m = m;
    sd.k = k;
    sd.x = x;
    const int k1 = (int)(k / (uint32_t)N2);  // the bin's column: its row in the block is s_row
    sd.xl = s_row > 0 ? cx[LAY2::idx(s_row - 1, k1)] : make_float2(qnan, qnan);
    sd.xr = s_row < RB - 1 ? cx[LAY2::idx(s_row + 1, k1)] : make_float2(qnan, qnan);
    gp(&pp->s[side])[0] = sd;
  };
  auto none = [&](int side) {
    FsTopSide sd;
    sd.m = -1.f;
    sd.k = 0xffffffffu;
    sd.x = sd.xl = sd.xr = make_float2(0.f, 0.f);
    gp(&pp->s[side])[0] = sd;
  };
  if (gli == 0xffffffffu) {
    if (tid == 0) none(0);
  } else if (li == gli) {
    put(0, lm, li, lx);
  }
  if (gri == 0xffffffffu) {
    if (tid == 0) none(1);
  } else if (ri == gri) {
    put(1, rm, ri, rx);
  }
}

// X at a neighbour of bin k that lies in another row block (side -1: k - 1,
// the previous block's last row; side +1: k + 1, the next block's first row;
// a channel's bins 0 and nfft-1 have no outer neighbour, callers never ask)
template <int N2, int RB>
__device__ __forceinline__ float2 fs_outer_cx(const float2* edge, uint32_t k, int side) {
  constexpr int N1 = 256, NT2 = N2 / RB;
  const uint32_t k2 = k & (N2 - 1), k1 = k / N2;
  if (side < 0) {
    if (k2 > 0) return gp(edge)[((k2 / RB - 1) * 2 + 1) * N1 + k1];
    return gp(edge)[((NT2 - 1) * 2 + 1) * N1 + (k1 + N1 - 1) % N1];  // row N2-1, column k1-1
  }
  if (k2 < (uint32_t)N2 - 1) return gp(edge)[((k2 / RB + 1) * 2 + 0) * N1 + k1];
  return gp(edge)[0 * N1 + (k1 + 1) % N1];  // row 0, column k1+1
}

// The three bins around a peak, for write_peak4.
struct TripletBins {
  uint32_t c;
  float2 xl, xc, xr;
  __device__ __forceinline__ float2 operator()(uint32_t i) const { return i == c ? xc : (i < c ? xl : xr); }
};

template <int N2, int TC>
__global__ void __launch_bounds__(64) k_fs_top_finish(FusedArgs a, FsWork w) {
  using C = FsCfg<N2, TC>;
  constexpr int NT2 = C::NT2, RB = C::RB;
  static_assert(NT2 <= 64, "one row block per lane");
  const int lane = threadIdx.x;
  const uint32_t ci = blockIdx.x;
  const uint64_t och = (uint64_t)w.ch0 + ci;
  const FsTopPart* parts = reinterpret_cast<const FsTopPart*>(w.part) + (uint64_t)ci * NT2;
  const float2* edge = reinterpret_cast<const float2*>(w.edge) + (uint64_t)ci * NT2 * 2 * 256;
  const uint32_t N = a.nfft;
  FsTopSide s0, s1;
  s0.m = s1.m = -1.f;
  s0.k = s1.k = 0xffffffffu;
  s0.x = s0.xl = s0.xr = s1.x = s1.xl = s1.xr = make_float2(0.f, 0.f);
  if (lane < NT2) {
    s0 = gp(&parts[lane].s[0])[0];
    s1 = gp(&parts[lane].s[1])[0];
  }
  // the side's winner over the row blocks (lowest bin on ties), its values
  // from the owning lane, a neighbour in another block from the edge rows
  auto side_bins = [&](const FsTopSide& sd, int side, float& gm) {
    gm = sd.m;
    uint32_t gk = sd.k;
    wave_argmax(gm, gk);
    const uint64_t own = __ballot(lane < NT2 && sd.k == gk && gk != 0xffffffffu);
    const int src = own ? __builtin_ctzll(own) : 0;
    auto rd = [&](float f) { return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, f), src)); };
    TripletBins t3;
    t3.c = gk;
    t3.xc = make_float2(rd(sd.x.x), rd(sd.x.y));
    t3.xl = make_float2(rd(sd.xl.x), rd(sd.xl.y));
    t3.xr = make_float2(rd(sd.xr.x), rd(sd.xr.y));
    if (gk != 0xffffffffu) {
      if (__builtin_isnan(t3.xl.x) && t3.c > 0) t3.xl = fs_outer_cx<N2, RB>(edge, t3.c, -1);
      if (__builtin_isnan(t3.xr.x) && t3.c < N - 1) t3.xr = fs_outer_cx<N2, RB>(edge, t3.c, 1);
    } else {
      // no bin compared greater than -1 (|X| NaN throughout the side): the
      // side's first bin, as the sequential rule leaves it, with NaN values
      const float qnan = __builtin_nanf("");
      t3.c = side == 0 ? 0u : min(a.s1, N - 1);
      t3.xc = t3.xl = t3.xr = make_float2(qnan, qnan);
    }
    return t3;
  };
  float gm0, gm1;
  const TripletBins t0 = side_bins(s0, 0, gm0);
  const TripletBins t1 = side_bins(s1, 1, gm1);
  if (lane == 0) {
    MaxValue* out = a.peaks + och * (a.peak_mode == 3 ? 4 : 8);
    if (a.peak_mode == 3) {
      if (gm1 > gm0) write_peak4(out, t1, N, t1.c, 1, a.fs);
      else write_peak4(out, t0, N, t0.c, 0, a.fs);
    } else {
      write_peak4(out, t0, N, t0.c, 0, a.fs);
      write_peak4(out + 4, t1, N, t1.c, 1, a.fs);
    }
  }
}

template <int N2>
static uint64_t fs_top_channel_bytes() {
  using C = FsCfg<N2, FsShape<N2>::TC>;
  auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
  return r((uint64_t)C::N * 8) + 256 + r((uint64_t)C::NT1 * 8) + r((uint64_t)C::NT2 * sizeof(FsTopPart)) +
         r((uint64_t)C::NT2 * 2 * 256 * 8);
}

template <int N2>
static hipError_t launch_fs_top_n(const FusedArgs& a, void* ws, uint32_t chunk, hipStream_t st) {
  using C = FsCfg<N2, FsShape<N2>::TC>;
  constexpr int TC = FsShape<N2>::TC;
  auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
  unsigned char* p = reinterpret_cast<unsigned char*>(ws);
  FsWork w;
  std::memset(&w, 0, sizeof(w));
  w.Z = reinterpret_cast<float2*>(p);            p += r((uint64_t)chunk * C::N * 8);
  w.ncand = reinterpret_cast<uint32_t*>(p);      p += r((uint64_t)chunk * 8);
  w.pe = reinterpret_cast<double*>(p);           p += r((uint64_t)chunk * C::NT1 * 8);
  w.part = reinterpret_cast<FsRowPart*>(p);      p += r((uint64_t)chunk * C::NT2 * sizeof(FsTopPart));
  w.edge = reinterpret_cast<float*>(p);
  for (uint32_t c0 = 0; c0 < a.batch; c0 += chunk) {
    w.ch0 = c0;
    w.nch = min(chunk, a.batch - c0);
    const uint32_t groups = (w.nch + C::MC - 1) / C::MC;
    hipLaunchKernelGGL((k_fs_cols<N2, TC>), dim3(groups * C::NT1), dim3(TC), 0, st, a, w);
    hipLaunchKernelGGL((k_fs_rows_top<N2, TC>), dim3(w.nch * C::NT2), dim3(C::TR), 0, st, a, w);
    hipLaunchKernelGGL((k_fs_top_finish<N2, TC>), dim3(w.nch), dim3(64), 0, st, a, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <int N2>
static uint64_t fs_dense_channel_bytes() {
  using C = FsCfg<N2, FsShape<N2>::TC>;
  auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
  return r((uint64_t)C::N * 8) + 256 + r((uint64_t)C::NT1 * 8);
}

template <int N2>
static hipError_t launch_fs_dense_n(const FusedArgs& a, void* ws, uint32_t chunk, hipStream_t st) {
  using C = FsCfg<N2, FsShape<N2>::TC>;
  auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
  unsigned char* p = reinterpret_cast<unsigned char*>(ws);
  FsWork w;
  std::memset(&w, 0, sizeof(w));
  w.Z = reinterpret_cast<float2*>(p);            p += r((uint64_t)chunk * C::N * 8);
  w.ncand = reinterpret_cast<uint32_t*>(p);      p += r((uint64_t)chunk * 8);
  w.pe = reinterpret_cast<double*>(p);
  return launch_fs_dense<N2, FsShape<N2>::TC>(a, w, chunk, st);
}

template <int N2, int TC>
static hipError_t launch_fs(const FusedArgs& a, const FsWork& w0, uint32_t chunk, hipStream_t st) {
  using C = FsCfg<N2, TC>;
  for (uint32_t c0 = 0; c0 < a.batch; c0 += chunk) {
    FsWork w = w0;
    w.ch0 = c0;
    w.nch = min(chunk, a.batch - c0);
    const uint32_t groups = (w.nch + C::MC - 1) / C::MC;
    hipLaunchKernelGGL((k_fs_cols<N2, TC>), dim3(groups * C::NT1), dim3(TC), 0, st, a, w);
    if (w.candx != nullptr) {
      // learned refinement: feature rows of the chunk's records, then the
      // refiner over them (before the chunk's workspace is reused)
      hipLaunchKernelGGL((k_fs_rows<N2, TC, true>), dim3(w.nch * C::NT2), dim3(C::TR), 0, st, a, w);
      hipLaunchKernelGGL((k_fs_finish<N2, TC, true>), dim3(w.nch), dim3(C::TF), 0, st, a, w);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      // feature rows are chunk-local; with channel indirection the records
      // are the caller's channels chan[c0 + i]
      if (w.chan != nullptr)
        e = spx_launch_mlp_refine(a.mlp, a.mlp_feat, a.peaks, a.counts, w.nch, a.kcap, a.fs / (float)a.nfft, st,
                                  w.chan + c0);
      else
        e = spx_launch_mlp_refine(a.mlp, a.mlp_feat, a.peaks + (uint64_t)c0 * a.kcap,
                                  a.counts + 2 * (uint64_t)c0, w.nch, a.kcap, a.fs / (float)a.nfft, st);
      if (e != hipSuccess) return e;
    } else {
      hipLaunchKernelGGL((k_fs_rows<N2, TC>), dim3(w.nch * C::NT2), dim3(C::TR), 0, st, a, w);
      hipLaunchKernelGGL((k_fs_finish<N2, TC>), dim3(w.nch), dim3(C::TF), 0, st, a, w);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}


// learned refinement on the split path: threshold search with a refiner set
__host__ __device__ inline bool fs_learn(const FusedArgs& a) {
  return a.refine == 2 && a.peak_mode == 1 && a.mlp != nullptr && a.kcap > 0;
}

template <int N2>
static uint64_t fs_channel_bytes(bool learn, uint32_t kcap) {
  using C = FsCfg<N2, FsShape<N2>::TC>;
  auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
  uint64_t b = r((uint64_t)C::N * 8) + r(C::CAND_CAP * sizeof(Cand)) + 256 + r((uint64_t)C::NT1 * 8) +
               r((uint64_t)C::NT2 * sizeof(FsRowPart)) + r((uint64_t)C::NT2 * 2 * 256 * 4);
  if (learn)
    b += r(C::CAND_CAP * 4 * 8) + (kcap > 64 ? r((uint64_t)C::NT2 * 4 * 256 * 8) : 0) + (uint64_t)kcap * 32;
  return b;
}

template <int N2>
static hipError_t launch_fs_n(const FusedArgs& a, void* ws, uint32_t chunk, const uint32_t* chan,
                              const uint64_t* inoff, hipStream_t st) {
  using C = FsCfg<N2, FsShape<N2>::TC>;
  auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
  unsigned char* p = reinterpret_cast<unsigned char*>(ws);
  FsWork w;
  w.Z = reinterpret_cast<float2*>(p);            p += r((uint64_t)chunk * C::N * 8);
  w.cand = reinterpret_cast<Cand*>(p);           p += r((uint64_t)chunk * C::CAND_CAP * sizeof(Cand));
  w.ncand = reinterpret_cast<uint32_t*>(p);      p += r((uint64_t)chunk * 8);
  w.pe = reinterpret_cast<double*>(p);           p += r((uint64_t)chunk * C::NT1 * 8);
  w.part = reinterpret_cast<FsRowPart*>(p);      p += r((uint64_t)chunk * C::NT2 * sizeof(FsRowPart));
  w.edge = reinterpret_cast<float*>(p);          p += r((uint64_t)chunk * C::NT2 * 2 * 256 * 4);
  w.candx = nullptr;
  w.edgex = nullptr;
  FusedArgs al = a;
  if (fs_learn(a)) {
    w.candx = reinterpret_cast<float2*>(p);      p += r((uint64_t)chunk * C::CAND_CAP * 4 * 8);
    if (a.kcap > 64) {
      // more than 64 records: the outer neighbours come from the edge rows
      w.edgex = reinterpret_cast<float2*>(p);    p += r((uint64_t)chunk * C::NT2 * 4 * 256 * 8);
    }
    al.mlp_feat = p;  // chunk feature rows: [chunk][kcap][16 halves]
  }
  w.ch0 = 0;
  w.nch = 0;
  w.chan = chan;
  w.inoff = inoff;
  return launch_fs<N2, FsShape<N2>::TC>(al, w, chunk, st);
}

}  // namespace spx

extern "C" {

// Workspace of the split long-channel peak search for `chunk` channels of
// length nfft (0 when nfft is not 2^15 .. 2^18).
uint64_t spx_fs_workspace_bytes(uint32_t nfft, uint32_t chunk) { return spx_fs_workspace_bytes_ex(nfft, chunk, 0, 0); }

// ... with the learned-refinement buffers when learn != 0 (kcap records per channel)
uint64_t spx_fs_workspace_bytes_ex(uint32_t nfft, uint32_t chunk, int learn, uint32_t kcap) {
  using namespace spx;
  uint64_t per = 0;
  switch (nfft) {
    case 1u << 15: per = fs_channel_bytes<128>(learn != 0, kcap); break;
    case 1u << 16: per = fs_channel_bytes<256>(learn != 0, kcap); break;
    case 1u << 17: per = fs_channel_bytes<512>(learn != 0, kcap); break;
    case 1u << 18: per = fs_channel_bytes<1024>(learn != 0, kcap); break;
    default: return 0;
  }
  return per * (uint64_t)chunk + 9 * 256;
}

// Runs the split path over a.batch channels in chunks of `chunk` channels,
// reusing the workspace `ws` (spx_fs_workspace_bytes(nfft, chunk) bytes).
// chan / inoff (device arrays, or both null): channel i of the launch is the
// caller's channel chan[i], its input at a.in + inoff[chan[i]] elements, its
// outputs at index chan[i].
hipError_t spx_launch_fourstep_ex(const FusedArgs* a, void* ws, uint32_t chunk, const uint32_t* chan,
                                  const uint64_t* inoff, hipStream_t st) {
  using namespace spx;
  if (chunk == 0 || (chan == nullptr) != (inoff == nullptr)) return hipErrorInvalidValue;
  switch (a->nfft) {
    case 1u << 15: return launch_fs_n<128>(*a, ws, chunk, chan, inoff, st);
    case 1u << 16: return launch_fs_n<256>(*a, ws, chunk, chan, inoff, st);
    case 1u << 17: return launch_fs_n<512>(*a, ws, chunk, chan, inoff, st);
    case 1u << 18: return launch_fs_n<1024>(*a, ws, chunk, chan, inoff, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t spx_launch_fourstep(const FusedArgs* a, void* ws, uint32_t chunk, hipStream_t st) {
  return spx_launch_fourstep_ex(a, ws, chunk, nullptr, nullptr, st);
}

// Dense outputs (a.spec_out / a.mag_out) of long channels on the split path.
uint64_t spx_fs_dense_workspace_bytes(uint32_t nfft, uint32_t chunk) {
  using namespace spx;
  uint64_t per = 0;
  switch (nfft) {
    case 1u << 15: per = fs_dense_channel_bytes<128>(); break;
    case 1u << 16: per = fs_dense_channel_bytes<256>(); break;
    case 1u << 17: per = fs_dense_channel_bytes<512>(); break;
    case 1u << 18: per = fs_dense_channel_bytes<1024>(); break;
    default: return 0;
  }
  return per * (uint64_t)chunk + 3 * 256;
}

// ONE_PEAK / TWO_PEAKS (a->peak_mode 3 / 4) of long channels on the split path.
uint64_t spx_fs_top_workspace_bytes(uint32_t nfft, uint32_t chunk) {
  using namespace spx;
  uint64_t per = 0;
  switch (nfft) {
    case 1u << 15: per = fs_top_channel_bytes<128>(); break;
    case 1u << 16: per = fs_top_channel_bytes<256>(); break;
    case 1u << 17: per = fs_top_channel_bytes<512>(); break;
    case 1u << 18: per = fs_top_channel_bytes<1024>(); break;
    default: return 0;
  }
  return per * (uint64_t)chunk + 5 * 256;
}

hipError_t spx_launch_fourstep_top(const FusedArgs* a, void* ws, uint32_t chunk, hipStream_t st) {
  using namespace spx;
  if (chunk == 0 || (a->peak_mode != 3 && a->peak_mode != 4)) return hipErrorInvalidValue;
  switch (a->nfft) {
    case 1u << 15: return launch_fs_top_n<128>(*a, ws, chunk, st);
    case 1u << 16: return launch_fs_top_n<256>(*a, ws, chunk, st);
    case 1u << 17: return launch_fs_top_n<512>(*a, ws, chunk, st);
    case 1u << 18: return launch_fs_top_n<1024>(*a, ws, chunk, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t spx_launch_fourstep_dense(const FusedArgs* a, void* ws, uint32_t chunk, hipStream_t st) {
  using namespace spx;
  if (chunk == 0) return hipErrorInvalidValue;
  switch (a->nfft) {
    case 1u << 15: return launch_fs_dense_n<128>(*a, ws, chunk, st);
    case 1u << 16: return launch_fs_dense_n<256>(*a, ws, chunk, st);
    case 1u << 17: return launch_fs_dense_n<512>(*a, ws, chunk, st);
    case 1u << 18: return launch_fs_dense_n<1024>(*a, ws, chunk, st);
    default: return hipErrorInvalidValue;
  }
}

}  // extern "C"
