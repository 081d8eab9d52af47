// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include "spx_types.h"

#define SPX_MAG_ABS 1
#define SPX_MAG_SQUARED 2
#define SPX_MAG_PHASE 3

// Arguments of the fused window -> FFT -> |X| -> peaks kernels.
struct FusedArgs {
  const void* in;          // [batch][n_point] complex (dtype)
  const float* in_scale;   // optional per-channel scale applied on load (FP16/BF16 pre-scaling)
  uint32_t batch, n_point, nfft, dtype;
  const float* window;     // optional [n_point] real window
  float fs;
  // optional dense outputs
  float2* spec_out;        // [batch][nfft] complex64
  float* mag_out;          // [batch][nfft] float (ABS/SQUARED) or float2 (PHASE: {|X|, arg X})
  uint32_t mag_kind;
  float mag_scale;
  // peak search: 0 none, 1 threshold (mean + nsigma*std), 2 local maxima only,
  // 3 / 4 the reference ONE_PEAK / TWO_PEAKS (s0 = search_range/2, s1 = nfft - s0)
  uint32_t peak_mode;
  float nsigma;
  uint32_t kcap;
  uint32_t s0, s1;
  uint32_t refine;         // 0 none, 1 parabolic, 2 learned (MLP) with parabolic fallback
  spx_max_value* peaks;    // [batch][kcap]
  uint32_t* counts;        // [batch][2] {stored, found}
  spx_max_value* gmax;     // [batch]
  float* stats;            // [batch][4] {mean, std, threshold, mean |X|^2}
  // tables
  const float2* tw_a;      // per-pass twiddle tables: small path length N, large path length 256
  const float2* tw_b;      // large path: per-pass tables of the column length N2 = N/256
  const float2* tw_4;      // large path: W_N^(n1*k2) at [k2*256 + n1]
  const void* mlp;         // learned refiner weights (refine == 2)
  void* mlp_feat;          // refine == 2 on the short paths: [batch][kcap][16] fp16 features of the stored peaks
  // window as w[n] = c0 - c1 cos(2 pi n/(n_point-1)) + c2 cos(4 pi n/(n_point-1))
  // (kernels that evaluate it instead of reading `window`; wkind 0 = none)
  uint32_t wkind;
  float wc0, wc1, wc2;
};

// Uneven-length batch.  Bucket b holds length L = 2^(10+b): b 0..4 run in
// persistent launches (one per tile shape TILE = 16384, 8192 or 4096 points)
// over work lists of tiles of TILE/L whole channels in LDS; b 5..7 run on the
// split four-step path, one launch sequence per length (or per chunk of its
// channels) with channel indirection (spx_launch_fourstep_ex).
#define SPX_BATCH_BUCKETS 8
#define SPX_BATCH_LOG_MIN 10
#define SPX_BATCH_FIRST_LARGE 5

struct BatchBucket {
  const float2* tw;        // tables of plan (L, V)
  const float* window;     // [L] or null
};

struct BatchArgs {
  const void* in;          // packed channels (dtype)
  const uint64_t* offsets; // [n_chan] element offset of each channel in `in`
  const uint32_t* perm;    // channel ids in work-list order
  const uint4* items;      // {bucket | variant << 8, first index into perm, channel count, 0}
  uint32_t* counter;       // work-list cursor of this launch (zeroed by spx_launch_batch before the batch kernel)
  BatchBucket bk[SPX_BATCH_FIRST_LARGE];
  uint32_t n_items, dtype, kcap, refine;
  float fs, nsigma;
  spx_max_value* peaks;    // [n_chan][kcap]
  uint32_t* counts;        // [n_chan][2]
  spx_max_value* gmax;     // [n_chan] or null
  float* stats;            // [n_chan][4] or null
  // learned refinement (refine 2): feature rows of the stored peaks of the
  // channel at perm index i at mlp_feat + ((i - feat_base) * kcap + pos) * 16 halves
  void* mlp_feat;
  uint32_t feat_base;
};

extern "C" {
hipError_t spx_batch_kernel_init(void);
uint64_t spx_mixed_table_bytes(void);
int spx_mlp_nparam(void);
int spx_mlp_deployed_bytes(void);
hipError_t spx_launch_mlp_dataset(const void* spec, const float* mag16, const void* peaks, const uint32_t* counts,
                                  uint32_t B, uint32_t N, uint32_t K, float* feat, float* label, float* parab,
                                  uint32_t* valid, hipStream_t st);
hipError_t spx_launch_mlp_grad(const float* params, const float* X, const float* Y, uint32_t n, float* partial,
                               int groups, float* grad, hipStream_t st);
hipError_t spx_launch_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                               float eps, int step, hipStream_t st);
hipError_t spx_launch_mlp_export(const float* params, void* dep, hipStream_t st);
hipError_t spx_launch_mlp_predict(const void* dep, const float* X, uint32_t n, float* out, hipStream_t st);
// feature row i refines channel chan ? chan[i] : i; its bin width fs / lens[channel] when lens is given
hipError_t spx_launch_mlp_refine(const void* dep, const void* feat, void* peaks, const uint32_t* counts,
                                 uint32_t batch, uint32_t kcap, float bin_hz, hipStream_t st,
                                 const uint32_t* chan = nullptr, const uint32_t* lens = nullptr, float fs = 0.0f);
hipError_t spx_launch_mixed(const FusedArgs* a, int compute, int grid, hipStream_t st);
int spx_batch_tile_ok(int tile);
hipError_t spx_launch_batch(const BatchArgs* b, int tile, int grid, hipStream_t st);
hipError_t spx_launch_fused(const FusedArgs* a, int grid, hipStream_t st);
int spx_fused_is_large(uint32_t nfft);
uint64_t spx_fs_workspace_bytes(uint32_t nfft, uint32_t chunk);
uint64_t spx_fs_workspace_bytes_ex(uint32_t nfft, uint32_t chunk, int learn, uint32_t kcap);
hipError_t spx_launch_fourstep(const FusedArgs* a, void* ws, uint32_t chunk, hipStream_t st);
hipError_t spx_launch_fourstep_ex(const FusedArgs* a, void* ws, uint32_t chunk, const uint32_t* chan,
                                  const uint64_t* inoff, hipStream_t st);
uint64_t spx_fs_dense_workspace_bytes(uint32_t nfft, uint32_t chunk);
hipError_t spx_launch_fourstep_dense(const FusedArgs* a, void* ws, uint32_t chunk, hipStream_t st);
int spx_large_is(uint32_t nfft);
uint64_t spx_large_workspace_bytes(uint32_t nfft, uint32_t chunk, uint32_t learn_kcap);
int spx_large_plan(uint32_t nfft, int* r0, int* lgs, int* np);
uint64_t spx_large_tables(uint32_t nfft, float* out);
hipError_t spx_launch_large(const FusedArgs* a, const float2* tw, const float2* const* twt, void* ws, uint32_t chunk,
                            hipStream_t st);
int spx_twiddle_table_len(uint32_t L, int V);
hipError_t spx_launch_mag_phase(const void* in, void* out, uint64_t total, hipStream_t st);
hipError_t spx_launch_magnitude(const void* in, float* out, uint64_t total, float inv_n, int squared,
                                hipStream_t st);
// ONE_PEAK / TWO_PEAKS (peak_mode 3 / 4, s0 >= 1) on the split path, nfft 2^15 .. 2^18:
// k_fs_cols, k_fs_rows_top, k_fs_top_finish per chunk of channels, no spectrum stored
uint64_t spx_fs_top_workspace_bytes(uint32_t nfft, uint32_t chunk);
hipError_t spx_launch_fourstep_top(const FusedArgs* a, void* ws, uint32_t chunk, hipStream_t st);
hipError_t spx_launch_peaks_top(const void* spec, void* out, uint32_t batch, uint32_t nfft, uint32_t search_range,
                                float fs, int two_peaks, hipStream_t st);
hipError_t spx_launch_all_maxima(const void* spec, void* out, uint32_t* counts, uint32_t batch, uint32_t nfft,
                                 float fs, uint32_t s0, uint32_t s1, uint32_t cap, hipStream_t st);
}
