/*
 * spectrum_mi355x.h -- C ABI of the MI355X (gfx950) batched FFT -> |X| ->
 * spectrum-maxima library (libspectrum_mi355x.so).
 *
 * Plain pointers and sizes only; every device pointer is a HIP device
 * allocation owned by the caller (the library never hands back memory it
 * owns).  All compute entry points are asynchronous on the context's stream
 * (spx_set_stream); call spx_synchronize() before reading results.
 * Status: 0 = SPX_OK, otherwise one of SPX_E*; spx_last_error() gives text.
 *
 * Transform convention: X[k] = sum_n x[n] exp(-2 pi i n k / N), unnormalised,
 * as hipfftExecC2C(HIPFFT_FORWARD) in the reference
 * (src/fft_func/src/fft_processor_rocm.cpp:353) and pocketfft c2c forward with
 * fct = 1 in its CPU path (include/spectrum/utils/cpu_fft.hpp:68-89).
 * Layouts: input [batch][n_point] complex64 (beam-major), spectra
 * [batch][nfft], records spx_max_value[batch][cap] (32 bytes each).
 * nfft must be a power of two, 2 .. 2^24, nfft >= n_point.  Lengths up to
 * 2^18 run on the fused kernels; 2^19 .. 2^24 run as three Stockham
 * passes through global memory with the peak search on the stored spectrum
 * (same rules and records).
 */
#ifndef SPECTRUM_MI355X_H
#define SPECTRUM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPX_OK 0
#define SPX_EINVAL 1   /* bad size, null pointer, unsupported length */
#define SPX_ENOMEM 2   /* device allocation failed */
#define SPX_EHIP 3     /* HIP runtime error */
#define SPX_ECAP 4     /* a capacity limit was exceeded */
#define SPX_ENOTSUP 5  /* feature not available for these arguments */

/* window applied to the n_point input samples before zero padding:
 * symmetric, denominator n_point - 1 (fft_processor_kernels_rocm.hpp:99-109);
 * values generated once in FP64 and rounded to FP32 */
#define SPX_WINDOW_NONE 0
#define SPX_WINDOW_HANN 1
#define SPX_WINDOW_HAMMING 2
#define SPX_WINDOW_BLACKMAN 3

#define SPX_DTYPE_C64 0      /* complex64: interleaved float re, im */
#define SPX_DTYPE_C32F16 1   /* complex fp16: interleaved half re, im */
#define SPX_DTYPE_C32BF16 2  /* complex bf16: interleaved bfloat16 re, im */

/* peak record, identical in size, order and meaning to the reference's
 * MaxValue (include/spectrum/types/spectrum_result_types.hpp:34-43):
 * phase in degrees, refined_frequency = (index + freq_offset) * fs / nfft */
typedef struct spx_max_value {
  uint32_t index;
  float real, imag;
  float magnitude;
  float phase;
  float freq_offset;
  float refined_frequency;
  uint32_t pad;
} spx_max_value;

typedef struct spx_ctx spx_ctx;

/* ---- context ----------------------------------------------------------
 * Replaces the IBackend/GpuContext pair the reference facades are built on
 * (include/spectrum/fft_processor_rocm.hpp:59, spectrum_processor_rocm.hpp:98).
 * A context owns mutable state (workspace, tables, host staging, the last
 * error text): one host thread at a time may call into it.  Threads that
 * work at once take a context each; several contexts may share a device. */
int spx_create(int device, spx_ctx** out);
int spx_destroy(spx_ctx* ctx);
/* borrow a caller stream (hipStream_t); NULL selects the context's own stream.
 * Work already queued on the previous stream is ordered before anything
 * queued on the new one (the new stream waits on an event of the old), so
 * the context's workspace is never shared by calls in flight. */
int spx_set_stream(spx_ctx* ctx, void* hip_stream);
void* spx_get_stream(spx_ctx* ctx);
int spx_synchronize(spx_ctx* ctx);
const char* spx_last_error(const spx_ctx* ctx);
int spx_device_info(spx_ctx* ctx, char* name, size_t name_len, int* cu_count, uint64_t* hbm_bytes);
/* caller-owned device memory helpers (hipMalloc/hipFree/hipMemcpy on the context device) */
int spx_alloc(spx_ctx* ctx, uint64_t bytes, void** out);
int spx_free(spx_ctx* ctx, void* ptr);
int spx_memcpy_h2d(spx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int spx_memcpy_d2h(spx_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- FFT front half ----------------------------------------------------
 * Zero-pad n_point -> nfft, optional window, forward C2C, unnormalised.
 * Replaces PadDataOp + hipfftExecC2C inside FFTProcessorROCm::ProcessComplex
 * (fft_processor_rocm.cpp:306-429, pad_data_op.hpp:82-133). in != out. */
int spx_fft_c2c(spx_ctx* ctx, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                uint32_t window);
/* ... and interleaved {|X|, atan2(im,re) radians} float2 output
 * (FFTProcessorROCm::ProcessMagPhase, fft_processor_rocm.cpp:435-570, MagPhaseOp) */
int spx_fft_mag_phase(spx_ctx* ctx, const void* in, void* out_float2, uint32_t batch, uint32_t n_point,
                      uint32_t nfft, uint32_t window);
/* ... and |X|*scale or |X|^2*scale float output
 * (FFTProcessorROCm::ProcessMagnitudesToGPU, fft_processor_rocm.cpp:579-659) */
int spx_fft_magnitude(spx_ctx* ctx, const void* in, float* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                      uint32_t window, int squared, float scale);

/* ---- elementwise (no FFT) ------------------------------------------------
 * {|z|, arg z} interleaved (complex_to_mag_phase, fft_processor_kernels_rocm.hpp:123-137;
 * ComplexToMagPhaseROCm::Process*, complex_to_mag_phase_rocm.cpp:182-300) */
int spx_mag_phase(spx_ctx* ctx, const void* in, void* out_float2, uint64_t total);
/* |z|*inv_n or |z|^2*inv_n (complex_to_magnitude(_squared), fft_processor_kernels_rocm.hpp:143-174;
 * ComplexToMagPhaseROCm::ProcessMagnitude*, complex_to_mag_phase_rocm.cpp:322-410) */
int spx_magnitude(spx_ctx* ctx, const void* in, float* out, uint64_t total, float inv_n, int squared);

/* ---- reference peak searches on a spectrum ---------------------------------
 * ONE_PEAK (4 records per beam) / TWO_PEAKS (8): argmax over [0, sr/2) and
 * [nfft - sr/2, nfft) with parabolic refinement and right-side mirroring
 * (post_kernel_one_peak / _two_peaks, fft_kernel_sources_rocm.hpp:122-519;
 * SpectrumPostOp::Execute, spectrum_post_op.hpp:77-106). */
int spx_peaks_top(spx_ctx* ctx, const void* spectrum, void* maxval_out, uint32_t batch, uint32_t nfft,
                  uint32_t search_range, float sample_rate, int two_peaks);
/* ... and the same search fused into the transform: zero-pad n_point ->
 * nfft, optional window, FFT, both argmaxes and the records, with no spectrum
 * stored for nfft <= 2^18 (one launch per batch up to 2^14, three per chunk
 * of channels from 2^15 to 2^18; from 2^19 the spectrum of a chunk goes
 * through a context buffer).  Same records as spx_fft_c2c + spx_peaks_top.
 * Replaces SpectrumProcessorROCm::ProcessFromCPU / ProcessFromGPU's
 * pad -> hipfftExecC2C -> SpectrumPostOp chain (spectrum_processor_rocm.cpp:
 * 179-420: ExecutePadKernel, ExecuteFFT, ExecutePostKernel; post_kernel_one_peak /
 * _two_peaks, fft_kernel_sources_rocm.hpp:122-519). */
int spx_fused_peaks_top(spx_ctx* ctx, const void* signal, void* maxval_out, uint32_t batch, uint32_t n_point,
                        uint32_t nfft, uint32_t window, uint32_t search_range, float sample_rate, int two_peaks);
/* AllMaxima: strict local maxima of |X| in [s0, s1) (s0 >= 1, s1 <= nfft - 1),
 * ascending, at most cap records per beam, counts[b] = number found (uncapped)
 * (AllMaximaPipelineROCm::Execute, all_maxima_pipeline_rocm.cpp:258-472;
 * detect/scan/compact kernels, all_maxima_kernel_sources_rocm.hpp:57-264). */
int spx_all_maxima(spx_ctx* ctx, const void* spectrum, void* maxval_out, uint32_t* counts, uint32_t batch,
                   uint32_t nfft, float sample_rate, uint32_t s0, uint32_t s1, uint32_t cap);

/* ---- fused window -> FFT -> |X| -> peaks (one launch) ---------------------
 * peak_mode SPX_PEAKS_THRESHOLD: the north-star search.  mean and population
 * std of |X| over [s0, s1) (default the whole band), threshold
 * T = mean + nsigma*std, peaks = strict local maxima (circular neighbours)
 * with |X| > T, the first max_peaks in ascending bin order, parabolic (or
 * learned) sub-bin refinement, plus the global-maximum record.
 * peak_mode SPX_PEAKS_ALL_MAXIMA: the AllMaxima rule set on a raw signal
 * (SpectrumProcessorROCm::FindAllMaximaFromCPU / FromGPUPipeline,
 * spectrum_processor_rocm.cpp:444-608): no threshold, no refinement,
 * counts[2b+1] uncapped. */
#define SPX_PEAKS_THRESHOLD 1
#define SPX_PEAKS_ALL_MAXIMA 2
#define SPX_REFINE_NONE 0
#define SPX_REFINE_PARABOLIC 1
#define SPX_REFINE_LEARNED 2

typedef struct spx_fused_desc {
  const void* input;        /* [batch][n_point], dtype below */
  const float* in_scale;    /* optional per-channel scale applied on load (FP16/BF16 pre-scaling) */
  uint32_t batch, n_point, nfft;
  uint32_t window;          /* SPX_WINDOW_* */
  uint32_t dtype;           /* SPX_DTYPE_* */
  float sample_rate;
  uint32_t peak_mode;       /* SPX_PEAKS_* */
  float nsigma;             /* threshold multiplier (SPX_PEAKS_THRESHOLD) */
  uint32_t max_peaks;       /* records kept per channel */
  uint32_t search_start, search_end; /* [s0, s1); 0, 0 = whole band */
  uint32_t refine;          /* SPX_REFINE_* */
  spx_max_value* peaks_out; /* [batch][max_peaks] */
  uint32_t* counts_out;     /* [batch][2] = {stored, found} */
  spx_max_value* gmax_out;  /* optional [batch] global maximum record */
  float* stats_out;         /* optional [batch][4] = {mean, std, threshold, mean |X|^2} */
} spx_fused_desc;

int spx_fused_peaks(spx_ctx* ctx, const spx_fused_desc* desc);

/* spx_fused_peaks without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * workspace is a fixed 64 MiB device array per GPU, held until the process
 * ends; long channels (nfft >= 32768) run in launch groups that fit it, and a
 * learned-refinement call on the short path needs batch * max_peaks * 32
 * bytes of it (more: SPX_ECAP).  The context's tables for the call's nfft, n_point and window
 * must exist already (any earlier spx_fused_peaks call with them creates
 * them; otherwise SPX_EINVAL, nothing allocated).  Results equal
 * spx_fused_peaks' bit for bit.  All calls on one GPU share the array:
 * calls that overlap in time on one GPU race. */
int spx_fused_peaks_noalloc(spx_ctx* ctx, const spx_fused_desc* desc);

/* ---- host-resident data ------------------------------------------------------
 * The same operations on HOST buffers (every pointer, including all of the
 * descriptor's, is a host address; pageable or pinned).  The batch streams
 * through the GPU in chunks of channels: the upload of one chunk, the kernels
 * of the previous one and the download of the one before overlap on three
 * streams, with pinned staging (skipped for buffers that are already pinned).
 * Results equal the device entry points' bit for bit; the call returns once
 * they are in host memory.  Replace the host-vector paths of
 * FFTProcessorROCm::ProcessComplex / ProcessMagPhase
 * (fft_processor_rocm.cpp:306-570) and SpectrumProcessorROCm::ProcessFromCPU /
 * FindAllMaximaFromCPU (spectrum_processor_rocm.cpp:263-307, 444-543), which
 * upload, compute and download one after the other. */
int spx_fft_c2c_host(spx_ctx* ctx, const void* in, void* out, uint32_t batch, uint32_t n_point, uint32_t nfft,
                     uint32_t window);
int spx_fft_mag_phase_host(spx_ctx* ctx, const void* in, void* out_float2, uint32_t batch, uint32_t n_point,
                           uint32_t nfft, uint32_t window);
int spx_fft_magnitude_host(spx_ctx* ctx, const void* in, float* out, uint32_t batch, uint32_t n_point,
                           uint32_t nfft, uint32_t window, int squared, float scale);
int spx_fused_peaks_host(spx_ctx* ctx, const spx_fused_desc* desc);
/* Bytes per transfer chunk of the *_host calls (default 32 MiB; the larger of
 * a chunk's input and output).  The context keeps three chunks of pinned and
 * three of device memory per direction. */
int spx_set_host_chunk(spx_ctx* ctx, uint64_t bytes);
/* Pinned (page-locked) host memory for the *_host calls' buffers: transfers
 * from and to it need no staging copy. */
int spx_host_alloc(spx_ctx* ctx, uint64_t bytes, void** out);
int spx_host_free(spx_ctx* ctx, void* ptr);

/* ---- learned sub-bin refiner ---------------------------------------------
 * No counterpart in the reference (its refinement is parabolic only,
 * include/spectrum/kernels/fft_kernel_sources_rocm.hpp:228-239).  A
 * 10 -> 64 -> 64 -> 1 ReLU network maps the 5-bin complex neighbourhood of a
 * peak (X[k+j]/X[k], j = -2..2) to its offset in bins.  Training runs on the
 * GPU (own forward/backward and Adam kernels, FP32 master weights); the
 * deployed block holds FP16 weights for MFMA inference with FP32
 * accumulation.  With spx_set_refiner, spx_fused_peaks accepts
 * refine = SPX_REFINE_LEARNED (threshold search, nfft >= 1024): the fused
 * kernel stores the parabolic records and the features of each stored peak
 * (on the split path, nfft 32768 .. 262144, the row kernel keeps each
 * candidate's complex neighbourhood and the finish kernel writes the
 * features; from 2^19 the gather kernel reads them from the stored chunk
 * spectrum), and a second MFMA kernel replaces the offset wherever the
 * network's prediction lies in [-0.5, 0.5]. */
int spx_mlp_param_count(void);    /* 4929 FP32 master parameters */
int spx_mlp_deployed_size(void);  /* bytes of the deployed block */
int spx_mlp_init(spx_ctx* ctx, float* params, uint64_t seed);
/* grad: param_count + 1 floats (d mean-squared-error / d params, then the loss) */
int spx_mlp_grad(spx_ctx* ctx, const float* params, const float* X, const float* Y, uint32_t n, float* grad);
int spx_mlp_train(spx_ctx* ctx, float* params, float* adam_m, float* adam_v, float* grad_buf, const float* X,
                  const float* Y, uint32_t n, uint32_t batch, uint32_t steps, uint32_t step0, float lr,
                  uint64_t seed, float* loss_out_host);
int spx_mlp_export(spx_ctx* ctx, const float* params, void* deployed_out);
int spx_mlp_predict(spx_ctx* ctx, const void* deployed, const float* X, uint32_t n, float* out);
int spx_set_refiner(spx_ctx* ctx, const void* deployed);
/* Features, zero-padded-FFT labels and parabolic offsets of stored peaks:
 * spec [batch][nfft] complex64 (windowed), mag16 [batch][16 nfft] |X| of the
 * 16x zero-padded transform of the same windowed signal, peaks/counts from
 * spx_fused_peaks; outputs per (channel, slot): feat [..][10], label, parab,
 * valid (1 when the slot holds a peak whose padded maximum lies inside
 * +-11 fine bins of 16k). */
int spx_mlp_dataset(spx_ctx* ctx, const void* spec, const float* mag16, const spx_max_value* peaks,
                    const uint32_t* counts, uint32_t batch, uint32_t nfft, uint32_t max_peaks, float* feat,
                    float* label, float* parab, uint32_t* valid);

/* ---- mixed precision: FP16/BF16 storage, DFT-as-GEMM on the matrix cores --
 * No counterpart in the reference (its FFT is hipFFT FP32 C2C,
 * src/fft_func/src/fft_processor_rocm.cpp:353).  N = 16384 = 16 x 32 x 32:
 * every butterfly stage is a real DFT-as-GEMM on MFMA with FP32 accumulation;
 * the input is fp16 or bf16 complex ([batch][16384], re/im interleaved),
 * multiplied by the window and a per-channel power-of-two pre-scale, stage
 * operands are fp16 (compute F16) or split bf16 hi+lo (compute BF16).
 * spx_fft_mixed writes the FP32 spectrum (complex64); spx_fused_peaks_mixed
 * runs the spx_fused_peaks rules on it (desc->dtype fp16/bf16, n_point =
 * nfft = 16384; refine NONE, PARABOLIC or LEARNED (the refiner set with
 * spx_set_refiner, one extra launch)). */
#define SPX_COMPUTE_F16 0
#define SPX_COMPUTE_BF16 1
int spx_fft_mixed(spx_ctx* ctx, const void* in, void* out_c64, uint32_t batch, uint32_t nfft, uint32_t window,
                  uint32_t dtype, uint32_t compute);
int spx_fused_peaks_mixed(spx_ctx* ctx, const spx_fused_desc* desc, uint32_t compute);

/* ---- uneven-length batch: one hipGraph per plan ----
 * No counterpart in the reference, which builds one hipFFT plan per (nFFT,
 * beam_count) and runs lengths one after another
 * (src/fft_func/src/spectrum_processor_rocm.cpp:1129-1150).  Channels of
 * lengths 1024 .. 131072 (powers of two, any mix, any order) are searched with
 * the SPX_PEAKS_THRESHOLD rules of spx_fused_peaks (whole band, nfft = length):
 * lengths up to 16384 in persistent launches over cost-sorted work lists of
 * LDS tiles (one launch per tile shape in use: 16384, 8192 or 4096 points per
 * workgroup, i.e. 1024, 512 or 256 threads and 16384/tile workgroups per CU,
 * each tile holding tile/length whole channels), longer lengths by the split
 * four-step launches of spx_fused_peaks reading each channel at its own
 * offset, in one or more launch groups; the plan captures all of it in one
 * hipGraph.  The radix order and tile shape of every length up to 16384 and
 * the launch groups of every longer length come from an autotuner cache
 * (JSON, written by spx_batch_autotune); without one every length uses the
 * default order, the 16384-point tile and one launch group.  A plan binds its
 * buffers: execute re-runs the captured graph on whatever the input buffer
 * holds at that time. */
typedef struct spx_batch_plan spx_batch_plan;

#define SPX_BATCH_NO_GRAPH 1u   /* launch the kernels directly on every execute */

typedef struct spx_batch_desc {
  const void* input;          /* device: packed channels (dtype) */
  const uint64_t* offsets;    /* host [n_chan]: element offset of channel c in input */
  const uint32_t* lengths;    /* host [n_chan]: length of channel c (power of two, 1024..131072) */
  uint32_t n_chan;
  uint32_t window;            /* SPX_WINDOW_* (symmetric over the channel's length) */
  uint32_t dtype;             /* SPX_DTYPE_* */
  float sample_rate;
  float nsigma;
  uint32_t max_peaks;         /* records kept per channel (>= 1) */
  uint32_t refine;            /* SPX_REFINE_NONE, _PARABOLIC or _LEARNED (the refiner set with
                                 spx_set_refiner when the plan is created; it must stay allocated
                                 while the plan exists) */
  spx_max_value* peaks_out;   /* device [n_chan][max_peaks] */
  uint32_t* counts_out;       /* device [n_chan][2] = {stored, found} */
  spx_max_value* gmax_out;    /* optional device [n_chan] */
  float* stats_out;           /* optional device [n_chan][4] */
  const char* tune_cache;     /* optional path of the autotuner's JSON cache */
  uint32_t flags;             /* SPX_BATCH_* */
} spx_batch_desc;

/* A plan keeps pointers to its context (stream, twiddle and window tables):
 * destroy every plan before the context it was created on. */
int spx_batch_plan_create(spx_ctx* ctx, const spx_batch_desc* desc, spx_batch_plan** out);
int spx_batch_execute(spx_batch_plan* plan);
int spx_batch_plan_destroy(spx_batch_plan* plan);
/* JSON description of the plan (per tile launch: shape, grid, items; per
 * length: radix order, tile shape or launch groups, items, scratch). */
int spx_batch_plan_info(spx_batch_plan* plan, char* buf, size_t len);
/* Times, for every length present in desc (on desc's own buffers, each length
 * alone), every candidate: up to 16384 each radix order (where the length has
 * two) in each tile shape that holds the length, above 16384 one, two or four
 * launch groups; keeps the fastest and writes the choices with every measured
 * time to cache_path as JSON. */
int spx_batch_autotune(spx_ctx* ctx, const spx_batch_desc* desc, const char* cache_path);

/* ---- synthetic signals (host side, deterministic) --------------------------
 * Sum of complex tones plus complex white Gaussian noise, generated in FP64
 * with counter-based SplitMix64 streams keyed by (seed, channel) and rounded
 * to complex64.  kind selects the workload's tone rules (SPX_GEN_*). Tones
 * drawn per channel are returned in tone_freq/tone_amp ([n_chan][max_tones],
 * cycles/sample; unused slots 0) when non-NULL. */
#define SPX_GEN_THREE_TONES 1   /* 3 tones, f~U[-.5,.5), A~U[.5,2], noise 10 dB below the weakest */
#define SPX_GEN_MULTI_TONE 2    /* 1-8 tones, per-tone SNR ~U[0,20] dB vs unit noise power */
#define SPX_GEN_DYNAMIC 3       /* tone A=1, tone -60 dB, noise -80 dB, >= 64 bins apart */
#define SPX_GEN_FRACTIONAL 4    /* 1-16 tones at (k+u)/N, A~U[.5,2], >= 3 bins apart, noise 20 dB below weakest */
int spx_siggen(uint32_t kind, uint64_t seed, uint32_t n_chan, uint32_t n_point, uint32_t chan_offset,
               void* out_c64, uint32_t max_tones, double* tone_freq, double* tone_amp, int threads);
/* Per-channel lengths 2^(lo_log2 + U{0 .. hi_log2-lo_log2}) from the (seed, channel) stream. */
int spx_siggen_lengths(uint64_t seed, uint32_t n_chan, uint32_t lo_log2, uint32_t hi_log2, uint32_t* lengths_out);
/* spx_siggen for channels of individual lengths: channel c (generator channel
 * index c) is written at out_c64 + offsets[c] complex elements. */
int spx_siggen_uneven(uint32_t kind, uint64_t seed, uint32_t n_chan, const uint32_t* lengths,
                      const uint64_t* offsets, void* out_c64, int threads);

/* ---- version ---- */
const char* spx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SPECTRUM_MI355X_H */
