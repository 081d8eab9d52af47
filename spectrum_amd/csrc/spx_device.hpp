// Device building blocks for the gfx950 FFT -> |X| -> peak-search kernels.
//
// Everything here is written for CDNA4 wave64: compile-time radix-2/4/8/16
// DFT butterflies held in VGPRs, LDS-staged Stockham passes (in place, one
// butterfly set per thread per pass), and wave/block reductions and scans
// built on 64-lane shuffles.
//
// Transform convention (matches the reference's hipFFT/pocketfft forward C2C,
// src/fft_func/src/fft_processor_rocm.cpp:353, include/spectrum/utils/cpu_fft.hpp:82-86):
//   X[k] = sum_n x[n] * exp(-2*pi*i*n*k/N), unnormalised.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spx {

// ---------------------------------------------------------------------------
// complex helpers (float2 = {re, im})
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 w) {
  return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
}
__device__ __forceinline__ float2 c_mul_mj(float2 a) { return make_float2(a.y, -a.x); }  // * (-j)
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// Pointers that travel inside a by-value argument struct are generic to the
// compiler; dereferencing them as such emits flat_* instructions (both
// waitcounts, 64-bit addresses, serialised waits).  gp() asserts the global
// address space so loads and stores become global_* with scalar bases.
// (The host compilation pass parses device code too but has no address
// spaces; there the pointer is returned unchanged.)
template <class T>
__device__ __forceinline__ auto gp(T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (__attribute__((address_space(1))) T*)p;
#else
  return p;
#endif
}

// Streaming (non-temporal) complex64 access for data a launch touches once
// and nothing reads again soon (the signal, the four-step intermediate Z as
// the row kernels read it, the dense spectrum they write).  On MI355X plain
// reads of contiguous rows ran 14 % faster with non-temporal loads.
typedef float stream_float2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  const stream_float2 v =
      __builtin_nontemporal_load((__attribute__((address_space(1))) const stream_float2*)p);
  return make_float2(v.x, v.y);
#else
  return *p;
#endif
}
__device__ __forceinline__ void st_stream(float2* p, float2 x) {
#if defined(__HIP_DEVICE_COMPILE__)
  stream_float2 v;
  v.x = x.x;
  v.y = x.y;
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) stream_float2*)p);
#else
  *p = x;
#endif
}

template <bool STREAM>
__device__ __forceinline__ float2 ld_f2(const float2* p) {
  if constexpr (STREAM) return ld_stream(p);
  else return *gp(p);
}
template <bool STREAM>
__device__ __forceinline__ void st_f2(float2* p, float2 x) {
  if constexpr (STREAM) st_stream(p, x);
  else *gp(p) = x;
}


constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
constexpr int ipow16(int p) { return p <= 0 ? 1 : 16 * ipow16(p - 1); }

// cos(pi/8), sin(pi/8), sqrt(1/2): the only irrational constants of a radix-16 butterfly.
#define SPX_C8 0.92387953251128674f
#define SPX_S8 0.38268343236508978f
#define SPX_R2 0.70710678118654752f

// multiply by W16^E = exp(-2*pi*i*E/16), E known at compile time.
template <int E>
__device__ __forceinline__ float2 tw16(float2 a) {
  constexpr int e = E & 15;
  if constexpr (e == 0) return a;
  else if constexpr (e == 4) return make_float2(a.y, -a.x);
  else if constexpr (e == 8) return make_float2(-a.x, -a.y);
  else if constexpr (e == 12) return make_float2(-a.y, a.x);
  else if constexpr (e == 2) return make_float2(SPX_R2 * (a.x + a.y), SPX_R2 * (a.y - a.x));
  else if constexpr (e == 6) return make_float2(SPX_R2 * (a.y - a.x), -SPX_R2 * (a.x + a.y));
  else if constexpr (e == 10) return make_float2(-SPX_R2 * (a.x + a.y), SPX_R2 * (a.x - a.y));
  else if constexpr (e == 14) return make_float2(SPX_R2 * (a.x - a.y), SPX_R2 * (a.x + a.y));
  else {
    // W16^e = (cos t, -sin t), t = 2*pi*e/16
    constexpr float c = (e == 1 || e == 15) ? SPX_C8 : (e == 3 || e == 13) ? SPX_S8
                       : (e == 5 || e == 11) ? -SPX_S8 : -SPX_C8;
    constexpr float s = (e == 1 || e == 7) ? SPX_S8 : (e == 3 || e == 5) ? SPX_C8
                       : (e == 9 || e == 15) ? -SPX_S8 : -SPX_C8;
    // (x + iy)(c - is) = (xc + ys) + i(yc - xs)
    return make_float2(fmaf(a.x, c, a.y * s), fmaf(a.y, c, -a.x * s));
  }
}

// tw16<E> for an exponent that is a constant only after loop unrolling
__device__ __forceinline__ float2 tw16_rt(int e, float2 a) {
  switch (e & 15) {
    case 0: return a;
    case 1: return tw16<1>(a);
    case 2: return tw16<2>(a);
    case 3: return tw16<3>(a);
    case 4: return tw16<4>(a);
    case 5: return tw16<5>(a);
    case 6: return tw16<6>(a);
    case 7: return tw16<7>(a);
    case 8: return tw16<8>(a);
    case 9: return tw16<9>(a);
    case 10: return tw16<10>(a);
    case 11: return tw16<11>(a);
    case 12: return tw16<12>(a);
    case 13: return tw16<13>(a);
    case 14: return tw16<14>(a);
    default: return tw16<15>(a);
  }
}

// ---------------------------------------------------------------------------
// DFT butterflies, natural-order in and out, in registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = c_add(t, b);
  b = c_sub(t, b);
}

__device__ __forceinline__ void dft4(float2& a, float2& b, float2& c, float2& d) {
  float2 t0 = c_add(a, c), t1 = c_sub(a, c), t2 = c_add(b, d), t3 = c_mul_mj(c_sub(b, d));
  a = c_add(t0, t2);
  c = c_sub(t0, t2);
  b = c_add(t1, t3);
  d = c_sub(t1, t3);
}

template <int R> struct Dft;
template <> struct Dft<1> { __device__ __forceinline__ static void run(float2*) {} };
template <> struct Dft<2> { __device__ __forceinline__ static void run(float2* v) { dft2(v[0], v[1]); } };
template <> struct Dft<4> {
  __device__ __forceinline__ static void run(float2* v) { dft4(v[0], v[1], v[2], v[3]); }
};
// 8 = 2 x 4: n = 2*n2 + n1, k = k1 + 4*k2
template <> struct Dft<8> {
  __device__ __forceinline__ static void run(float2* v) {
    dft4(v[0], v[2], v[4], v[6]);
    dft4(v[1], v[3], v[5], v[7]);
    // A[n1][k1] lives at v[n1 + 2*k1]; twiddle W8^(n1*k1) = W16^(2*n1*k1)
    v[3] = tw16<2>(v[3]);
    v[5] = tw16<4>(v[5]);
    v[7] = tw16<6>(v[7]);
    float2 o[8];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 a = v[2 * k1], b = v[2 * k1 + 1];
      o[k1] = c_add(a, b);
      o[k1 + 4] = c_sub(a, b);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = o[i];
  }
};
// 16 = 4 x 4: n = 4*n2 + n1, k = k1 + 4*k2
template <> struct Dft<16> {
  __device__ __forceinline__ static void run(float2* v) {
    dft4(v[0], v[4], v[8], v[12]);
    dft4(v[1], v[5], v[9], v[13]);
    dft4(v[2], v[6], v[10], v[14]);
    dft4(v[3], v[7], v[11], v[15]);
    // A[n1][k1] at v[n1 + 4*k1]; twiddle W16^(n1*k1)
    v[5] = tw16<1>(v[5]);
    v[9] = tw16<2>(v[9]);
    v[13] = tw16<3>(v[13]);
    v[6] = tw16<2>(v[6]);
    v[10] = tw16<4>(v[10]);
    v[14] = tw16<6>(v[14]);
    v[7] = tw16<3>(v[7]);
    v[11] = tw16<6>(v[11]);
    v[15] = tw16<9>(v[15]);
    dft4(v[0], v[1], v[2], v[3]);
    dft4(v[4], v[5], v[6], v[7]);
    dft4(v[8], v[9], v[10], v[11]);
    dft4(v[12], v[13], v[14], v[15]);
    // X[k1 + 4*k2] sits at v[4*k1 + k2]: transpose (register renaming only)
    float2 o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = o[i];
  }
};

// ---------------------------------------------------------------------------
// Radix plan for a power-of-two length L = 2^a: ceil(a/4) passes, all radix 16
// except one leftover radix RX (2, 4 or 8) when 4 does not divide a.
//   V = 0: radix-16 passes first, the leftover last (every LDS access of every
//          pass is then base + compile-time stride, see the layouts);
//   V = 1: the leftover first, then radix 16 (an alternative the batch
//          autotuner may pick per length).
// Twiddles: pass p >= 1 (sub-length NS = product of the earlier radices, radix
// R) reads a table laid out [k][r] = W_(NS*R)^(r*k), k < NS, r < R, so one
// thread's R twiddles are contiguous (R/2 16-byte loads from one address).
// The tables of all passes are concatenated; twoff(p) is the offset of pass p
// in float2 units.
// ---------------------------------------------------------------------------
template <int L, int V = 0>
struct Plan {
  static constexpr int LOG = ilog2(L);
  static constexpr int NP = (LOG + 3) / 4;
  static constexpr int RX = (LOG % 4) ? (1 << (LOG % 4)) : 16;
  static constexpr int radix(int p) { return NP == 1 ? RX : (V == 0 ? (p == NP - 1 ? RX : 16) : (p == 0 ? RX : 16)); }
  // loops, not recursion: a recursive constexpr function evaluated outside a
  // constant expression is emitted as a real device call (s_swappc plus the
  // caller's register saves around it)
  static constexpr int ns(int p) {
    int n = 1;
    for (int i = 0; i < p; ++i) n *= radix(i);
    return n;
  }
  static constexpr int twoff(int p) {
    int o = 0;
    for (int i = 1; i < p; ++i) o += ns(i) * radix(i);
    return o;
  }
  static constexpr int TWSIZE = NP <= 1 ? 1 : twoff(NP - 1) + ns(NP - 1) * radix(NP - 1);
  // LDS copies pad every [k][r] row by two entries (a 16-byte multiple that
  // spreads the rows of neighbouring k over the banks)
  static constexpr int twoff_lds(int p) {
    int o = 0;
    for (int i = 1; i < p; ++i) o += ns(i) * (radix(i) + 2);
    return o;
  }
  static constexpr int TWSIZE_LDS = NP <= 1 ? 1 : twoff_lds(NP - 1) + ns(NP - 1) * (radix(NP - 1) + 2);
  static constexpr int R0 = (NP == 1) ? RX : (V == 0 ? 16 : RX);   // first-pass radix
  static constexpr int RL = (NP == 1) ? RX : (V == 0 ? RX : 16);   // last-pass radix
};

// ---------------------------------------------------------------------------
// LDS layouts for a tile of NSEQ sequences of length L.
//  Contig: element i of sequence s at pad(s*L + i), pad(a) = a + a/16
//          (one spare slot per 16 complex breaks the 128-byte stride of the
//          radix-16 writes across the LDS banks).  Because reads step by
//          L/R = 16^k and writes by NS = 1 or 16^k, pad() stays linear:
//          element r of a butterfly is at base + r * (compile-time stride).
//  Inter:  element i of sequence s at i*NSEQ + s (sequences interleaved; the
//          lanes of a wave walk s, so every access is unit-stride).
// Thread mapping of butterfly j (radix R): Contig walks jj fastest, Inter walks s.
// ---------------------------------------------------------------------------
template <int L, int NSEQ>
struct LayContig {
  static constexpr bool CONTIG = true;
  static constexpr int SIZE = L * NSEQ + (L * NSEQ) / 16;
  __device__ __forceinline__ static int idx(int s, int i) {
    int a = s * L + i;
    return a + (a >> 4);
  }
  static constexpr int rstride(int R) { return (L / R) % 16 == 0 ? (L / R) + (L / R) / 16 : -1; }
  static constexpr int wstride(int NS) { return NS == 1 ? 1 : (NS % 16 == 0 ? NS + NS / 16 : -1); }
  template <int R>
  __device__ __forceinline__ static void map(int j, int& s, int& jj) {
    s = j / (L / R);
    jj = j % (L / R);
  }
};

template <int L, int NSEQ>
struct LayInter {
  static constexpr bool CONTIG = false;
  static constexpr int SIZE = L * NSEQ;
  __device__ __forceinline__ static int idx(int s, int i) { return i * NSEQ + s; }
  static constexpr int rstride(int R) { return (L / R) * NSEQ; }
  static constexpr int wstride(int NS) { return NS * NSEQ; }
  template <int R>
  __device__ __forceinline__ static void map(int j, int& s, int& jj) {
    s = j % NSEQ;
    jj = j / NSEQ;
  }
};

// R contiguous twiddles (a row of a [k][r] table) with 16-byte loads; TWLDS:
// the table lives in LDS (plain loads) instead of global memory.
template <int R, bool TWLDS = false>
__device__ __forceinline__ void load_tw_row(const float2* row, float2* w) {
  if constexpr (TWLDS) {
    if constexpr (R == 2) {
      w[0] = make_float2(1.f, 0.f);
      w[1] = row[1];
    } else {
      const float4* q = reinterpret_cast<const float4*>(row);
#pragma unroll
      for (int i = 0; i < R / 2; ++i) {
        const float4 t = q[i];
        w[2 * i] = make_float2(t.x, t.y);
        w[2 * i + 1] = make_float2(t.z, t.w);
      }
    }
    return;
  }
  if constexpr (R == 2) {
    w[0] = make_float2(1.f, 0.f);
    w[1] = gp(row)[1];
  } else {
    const float4* q = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      const float4 t = gp(q)[i];
      w[2 * i] = make_float2(t.x, t.y);
      w[2 * i + 1] = make_float2(t.z, t.w);
    }
  }
}

// One Stockham pass (radix R, current sub-transform length NS) over a tile of
// NSEQ sequences of length L held by T threads with PPT points each.
//  in : v[b*R + r] = x[s][jj + r*L/R]   (loaded from LDS unless FIRST)
//  out: LDS y[s][(jj/NS)*NS*R + jj%NS + r*NS]           (unless LAST)
//       or, for LAST, v[b*R + r] = X[s][jj + r*L/R] in registers.
// Twiddles W_(NS*R)^(r*(jj%NS)) from the pass table twp[k*R + r].
template <class LAY, int L, int NSEQ, int R, int NS, int T, int PPT, bool FIRST, bool LAST, bool TWLDS = false>
__device__ __forceinline__ void fft_pass(float2* __restrict__ lds, float2 (&v)[PPT],
                                         const float2* __restrict__ twp, int tid) {
  constexpr int BPT = PPT / R;
  constexpr int NB = NSEQ * L / R;            // butterflies in the tile
  constexpr bool FULL = (NB == T * BPT);      // every thread busy
  constexpr int RS = LAY::rstride(R);
  constexpr int WS = LAY::wstride(NS);
  if constexpr (!FIRST) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = tid + b * T;
      if (FULL || j < NB) {
        int s, jj;
        LAY::template map<R>(j, s, jj);
        if constexpr (RS > 0) {
          const float2* base = lds + LAY::idx(s, jj);
#pragma unroll
          for (int r = 0; r < R; ++r) v[b * R + r] = base[r * RS];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) v[b * R + r] = lds[LAY::idx(s, jj + r * (L / R))];
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = tid + b * T;
    if (FULL || j < NB) {
      if constexpr (NS > 1) {
        int s, jj;
        LAY::template map<R>(j, s, jj);
        const int k = jj % NS;
        float2 w[R];
        load_tw_row<R, TWLDS>(twp + (uint32_t)(k * (TWLDS ? R + 2 : R)), w);
#pragma unroll
        for (int r = 1; r < R; ++r) v[b * R + r] = c_mul(v[b * R + r], w[r]);
      }
      Dft<R>::run(&v[b * R]);
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = tid + b * T;
      if (FULL || j < NB) {
        int s, jj;
        LAY::template map<R>(j, s, jj);
        const int d0 = (jj / NS) * NS * R + (jj % NS);
        if constexpr (WS > 0) {
          float2* base = lds + LAY::idx(s, d0);
#pragma unroll
          for (int r = 0; r < R; ++r) base[r * WS] = v[b * R + r];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) lds[LAY::idx(s, d0 + r * NS)] = v[b * R + r];
        }
      }
    }
    __syncthreads();
  }
}

template <class LAY, int L, int NSEQ, int T, int PPT, int P, int V, bool TWLDS>
struct FftPasses {
  __device__ __forceinline__ static void run(float2* lds, float2 (&v)[PPT], const float2* tw, int tid) {
    using PL = Plan<L, V>;
    constexpr int OFF = TWLDS ? PL::twoff_lds(P) : PL::twoff(P);
    fft_pass<LAY, L, NSEQ, PL::radix(P), PL::ns(P), T, PPT, P == 0, P == PL::NP - 1, TWLDS>(lds, v, tw + OFF, tid);
    if constexpr (P + 1 < PL::NP) FftPasses<LAY, L, NSEQ, T, PPT, P + 1, V, TWLDS>::run(lds, v, tw, tid);
  }
};

// The same passes with each pass's twiddles loaded one step early: right
// after the previous pass has written its outputs to LDS (its VGPRs are free
// then) and before the barrier, so the table reads' latency runs under the
// barrier wait and the LDS reads instead of after them.  Fewer bytes are read
// than the R-entry rows hold (the vector-memory path, not the table, limits):
//  TW_ROWS  w[b*R + r] = W_(NS*R)^(r*(jj%NS)), the table rows as they are;
//  TW_LAST  last pass (NS = L/R) of a contiguous tile whose butterfly
//           index j = tid + b*T gives k = j mod L/R = tid mod L/R + q_b with
//           q_b = (b*T) mod L/R a multiple of L/16 (T and L/R powers of two,
//           L <= 16 T): W_L^(r k) = W_L^(r (tid mod L/R)) W16^(r q_b 16/L), so
//           one row is read (w[r]) and the W16 factors are compile-time
//           rotations;
//  TW_FACT  radix 16, one butterfly per thread: W^(r k) = W^((r&3) k)
//           W^((r&12) k), so entries 1, 2, 3, 4, 8, 12 of the row are read
//           (w[0..5] = W^(k), W^(2k), W^(3k), W^(4k), W^(8k), W^(12k)) and the
//           other nine are products.
enum { TW_ROWS = 0, TW_LAST = 1, TW_FACT = 2 };

template <class LAY, int L, int NSEQ, int R, int NS, int T, int PPT, bool LAST>
constexpr int tw_mode() {
  if constexpr (NS <= 1) return TW_ROWS;
  else if constexpr (LAST && LAY::CONTIG && PPT / R > 1 && NS * R == L && (16 * T) % L == 0 &&
                     ((L / R) % T == 0 || T % (L / R) == 0))
    return TW_LAST;
  else if constexpr (R == 16 && PPT == 16) return TW_FACT;
  else return TW_ROWS;
}

template <class LAY, int L, int NSEQ, int R, int NS, int T, int PPT, bool TWLDS, int TWM>
__device__ __forceinline__ void pass_twiddles(const float2* __restrict__ twp, int tid, float2 (&w)[PPT]) {
  constexpr int BPT = PPT / R;
  constexpr int NB = NSEQ * L / R;
  constexpr bool FULL = (NB == T * BPT);
  constexpr int RW = TWLDS ? R + 2 : R;  // row pitch
  if constexpr (TWM == TW_LAST) {
    load_tw_row<R, TWLDS>(twp + (uint32_t)((tid % (L / R)) * RW), w);
  } else if constexpr (TWM == TW_FACT) {
    int s, jj;
    LAY::template map<R>(tid, s, jj);
    const float2* row = twp + (uint32_t)((jj % NS) * RW);
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = TWLDS ? row[1 + i] : *gp(row + 1 + i);
    w[4] = TWLDS ? row[8] : *gp(row + 8);
    w[5] = TWLDS ? row[12] : *gp(row + 12);
  } else {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = tid + b * T;
      if (FULL || j < NB) {
        int s, jj;
        LAY::template map<R>(j, s, jj);
        load_tw_row<R, TWLDS>(twp + (uint32_t)((jj % NS) * RW), &w[b * R]);
      }
    }
  }
}

// twiddle of element r of butterfly b from what pass_twiddles read
template <int L, int R, int T, int TWM>
__device__ __forceinline__ float2 pass_tw(const float2* w, int b, int r) {
  if constexpr (TWM == TW_LAST) {
    return tw16_rt(r * (((b * T) % (L / R)) * 16 / L), w[r]);
  } else if constexpr (TWM == TW_FACT) {
    const int lo = r & 3, hi = r >> 2;  // w[lo - 1] = W^(lo k), w[2 + hi] = W^(4 hi k)
    if (hi == 0) return w[lo - 1];
    if (lo == 0) return w[2 + hi];
    return c_mul(w[lo - 1], w[2 + hi]);
  } else {
    return w[b * R + r];
  }
}

// fft_pass with the twiddles given and without the closing barrier (the
// caller issues the next pass's twiddle loads first).
template <class LAY, int L, int NSEQ, int R, int NS, int T, int PPT, bool FIRST, bool LAST, int TWM>
__device__ __forceinline__ void fft_pass_w(float2* __restrict__ lds, float2 (&v)[PPT], const float2 (&w)[PPT], int tid) {
  constexpr int BPT = PPT / R;
  constexpr int NB = NSEQ * L / R;
  constexpr bool FULL = (NB == T * BPT);
  constexpr int RS = LAY::rstride(R);
  constexpr int WS = LAY::wstride(NS);
  if constexpr (!FIRST) {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = tid + b * T;
      if (FULL || j < NB) {
        int s, jj;
        LAY::template map<R>(j, s, jj);
        if constexpr (RS > 0) {
          const float2* base = lds + LAY::idx(s, jj);
#pragma unroll
          for (int r = 0; r < R; ++r) v[b * R + r] = base[r * RS];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) v[b * R + r] = lds[LAY::idx(s, jj + r * (L / R))];
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = tid + b * T;
    if (FULL || j < NB) {
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[b * R + r] = c_mul(v[b * R + r], pass_tw<L, R, T, TWM>(w, b, r));
      }
      Dft<R>::run(&v[b * R]);
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = tid + b * T;
      if (FULL || j < NB) {
        int s, jj;
        LAY::template map<R>(j, s, jj);
        const int d0 = (jj / NS) * NS * R + (jj % NS);
        if constexpr (WS > 0) {
          float2* base = lds + LAY::idx(s, d0);
#pragma unroll
          for (int r = 0; r < R; ++r) base[r * WS] = v[b * R + r];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) lds[LAY::idx(s, d0 + r * NS)] = v[b * R + r];
        }
      }
    }
  }
}

template <class LAY, int L, int NSEQ, int T, int PPT, int P, int V, bool TWLDS>
struct FftPassesPF {
  __device__ __forceinline__ static void run(float2* lds, float2 (&v)[PPT], const float2 (&w)[PPT], const float2* tw,
                                             int tid) {
    using PL = Plan<L, V>;
    constexpr int R = PL::radix(P), NS = PL::ns(P);
    constexpr bool LAST = P == PL::NP - 1;
    fft_pass_w<LAY, L, NSEQ, R, NS, T, PPT, P == 0, LAST, tw_mode<LAY, L, NSEQ, R, NS, T, PPT, LAST>()>(lds, v, w, tid);
    if constexpr (P + 1 < PL::NP) {
      constexpr int RN = PL::radix(P + 1), NSN = PL::ns(P + 1);
      constexpr bool LASTN = P + 1 == PL::NP - 1;
      constexpr int OFFN = TWLDS ? PL::twoff_lds(P + 1) : PL::twoff(P + 1);
      float2 wn[PPT];
      pass_twiddles<LAY, L, NSEQ, RN, NSN, T, PPT, TWLDS, tw_mode<LAY, L, NSEQ, RN, NSN, T, PPT, LASTN>()>(
          tw + OFFN, tid, wn);
      __syncthreads();
      FftPassesPF<LAY, L, NSEQ, T, PPT, P + 1, V, TWLDS>::run(lds, v, wn, tw, tid);
    }
  }
};

// Full length-L transform of the tile.  Before the call v holds the first
// pass's inputs (x[s][jj + r*L/R0], R0 = Plan<L,V>::R0, for the mapping of
// j = tid + b*T); after it v holds X[s][jj + r*L/RL] for the radix-RL mapping
// (RL = Plan<L,V>::RL).  tw: the concatenated per-pass twiddle tables of plan
// (L, V) (Plan<L,V>::TWSIZE entries).
template <class LAY, int L, int NSEQ, int T, int PPT, int V = 0, bool TWLDS = false>
__device__ __forceinline__ void fft_tile(float2* lds, float2 (&v)[PPT], const float2* tw, int tid) {
  FftPasses<LAY, L, NSEQ, T, PPT, 0, V, TWLDS>::run(lds, v, tw, tid);
}

// fft_tile with every pass's twiddles loaded before the barrier that starts
// the pass (same arguments and result).
template <class LAY, int L, int NSEQ, int T, int PPT, int V = 0, bool TWLDS = false>
__device__ __forceinline__ void fft_tile_pf(float2* lds, float2 (&v)[PPT], const float2* tw, int tid) {
  float2 w0[PPT];  // pass 0 has sub-transform length 1: no twiddles
  FftPassesPF<LAY, L, NSEQ, T, PPT, 0, V, TWLDS>::run(lds, v, w0, tw, tid);
}

// ---------------------------------------------------------------------------
// wave64 / block reductions and scans
// ---------------------------------------------------------------------------
// Cross-lane steps on DPP (data moved by the vector ALU, a few cycles per
// step, instead of ds_bpermute through the LDS crossbar).  The scan is the
// row-shift / row-broadcast sequence: within each row of 16 lanes, shifts by
// 1, 2, 4, 8 give inclusive prefixes (a lane without a source lane keeps the
// identity `id`); row_bcast15 then adds row 0's total to row 1 and row 2's
// to row 3, and row_bcast31 adds rows 0-1's total to rows 2 and 3, so lane i
// ends with the inclusive prefix over lanes 0..i.
namespace dpp {
constexpr int kShr1 = 0x111, kShr2 = 0x112, kShr4 = 0x114, kShr8 = 0x118, kBcast15 = 0x142, kBcast31 = 0x143;
template <int CTRL, int ROWS>
__device__ __forceinline__ int mov(int id, int x) {
  return __builtin_amdgcn_update_dpp(id, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double mov(double id, double x) {
  const int2 a = __builtin_bit_cast(int2, x), o = __builtin_bit_cast(int2, id);
  return __builtin_bit_cast(double, make_int2(mov<CTRL, ROWS>(o.x, a.x), mov<CTRL, ROWS>(o.y, a.y)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float mov(float id, float x) {
  return __builtin_bit_cast(float, mov<CTRL, ROWS>(__builtin_bit_cast(int, id), __builtin_bit_cast(int, x)));
}
// inclusive scan of x with the associative, commutative op f and identity id
template <class V, class F>
__device__ __forceinline__ V scan(V x, V id, F f) {
  x = f(x, mov<kShr1, 0xf>(id, x));
  x = f(x, mov<kShr2, 0xf>(id, x));
  x = f(x, mov<kShr4, 0xf>(id, x));
  x = f(x, mov<kShr8, 0xf>(id, x));
  x = f(x, mov<kBcast15, 0xa>(id, x));
  x = f(x, mov<kBcast31, 0xc>(id, x));
  return x;
}
// The same six moves for a reduction read only in lane 63: the lanes a move
// leaves unwritten (row_shr across a row start, rows outside the row mask)
// keep whatever the destination held, so no identity has to be moved in
// first.  Lane 63's chain reads only written lanes: position p of a row reads
// p - k only for p >= k, row 3 takes lane 47 and lane 31, row 1 lane 15.
template <int CTRL, int ROWS>
__device__ __forceinline__ int movu(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float movu(float x) {
  return __builtin_bit_cast(float, movu<CTRL, ROWS>(__builtin_bit_cast(int, x)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double movu(double x) {
  const int2 a = __builtin_bit_cast(int2, x);
  return __builtin_bit_cast(double, make_int2(movu<CTRL, ROWS>(a.x), movu<CTRL, ROWS>(a.y)));
}
// reduction of x with f; the result is valid in lane 63 only
template <class V, class F>
__device__ __forceinline__ V reduce63(V x, F f) {
  x = f(x, movu<kShr1, 0xf>(x));
  x = f(x, movu<kShr2, 0xf>(x));
  x = f(x, movu<kShr4, 0xf>(x));
  x = f(x, movu<kShr8, 0xf>(x));
  x = f(x, movu<kBcast15, 0xa>(x));
  x = f(x, movu<kBcast31, 0xc>(x));
  return x;
}
__device__ __forceinline__ double lane63(double x) {
  const int2 a = __builtin_bit_cast(int2, x);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(a.x, 63), __builtin_amdgcn_readlane(a.y, 63)));
}
}  // namespace dpp

// sum over the wave, the same value in every lane
__device__ __forceinline__ double wave_sum(double x) {
  return dpp::lane63(dpp::reduce63(x, [](double a, double b) { return a + b; }));
}
// maximum over the wave (fmaxf: NaN lanes ignored), the same value in every lane
__device__ __forceinline__ float wave_maxf(float x) {
  x = dpp::reduce63(x, [](float a, float b) { return fmaxf(a, b); });
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}
// argmax with lowest index on ties; (m, k) pairs, the result in every lane
__device__ __forceinline__ void wave_argmax(float& m, uint32_t& k) {
  auto step = [&](float om, uint32_t ok) {
    if (om > m || (om == m && ok < k)) { m = om; k = ok; }
  };
  // only lane 63 is read back (see reduce63)
  step(dpp::movu<dpp::kShr1, 0xf>(m), (uint32_t)dpp::movu<dpp::kShr1, 0xf>((int)k));
  step(dpp::movu<dpp::kShr2, 0xf>(m), (uint32_t)dpp::movu<dpp::kShr2, 0xf>((int)k));
  step(dpp::movu<dpp::kShr4, 0xf>(m), (uint32_t)dpp::movu<dpp::kShr4, 0xf>((int)k));
  step(dpp::movu<dpp::kShr8, 0xf>(m), (uint32_t)dpp::movu<dpp::kShr8, 0xf>((int)k));
  step(dpp::movu<dpp::kBcast15, 0xa>(m), (uint32_t)dpp::movu<dpp::kBcast15, 0xa>((int)k));
  step(dpp::movu<dpp::kBcast31, 0xc>(m), (uint32_t)dpp::movu<dpp::kBcast31, 0xc>((int)k));
  m = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, m), 63));
  k = (uint32_t)__builtin_amdgcn_readlane((int)k, 63);
}
// inclusive scan of ints across the wave
__device__ __forceinline__ int wave_incl_scan(int x) {
  return dpp::scan(x, 0, [](int a, int b) { return a + b; });
}

struct BlockRed {
  double s1[16], s2[16];
  float m[16];
  uint32_t k[16];
  int c[16];
  int total;
};

}  // namespace spx

namespace spx {
// Copies the per-pass tables of plan (L, V) from global memory into the padded
// LDS layout (rows of R + 2 entries); all T threads of the block take part.
template <int L, int V, int T>
__device__ __forceinline__ void copy_tw_to_lds(const float2* glob, float2* lds, int tid) {
  using PL = Plan<L, V>;
  for (int i = tid; i < PL::TWSIZE; i += T) {
    int p = 1, base = 0;
    while (p + 1 < PL::NP && i >= PL::twoff(p + 1)) ++p;
    base = i - PL::twoff(p);
    const int R = PL::radix(p);
    lds[PL::twoff_lds(p) + (base / R) * (R + 2) + base % R] = gp(glob)[i];
  }
}
}  // namespace spx
