/*
 * Extensions of the toolkit's C ABI (no counterpart in the reference).
 *
 * Plain C, plain pointers and sizes.  "d_" pointers are device (HBM)
 * pointers; these entry points enqueue on the library stream and return
 * without waiting unless stated, so a caller can time or overlap them.  The
 * library's own stream is a blocking stream (ordered with the legacy default
 * stream); work on other non-blocking streams must be ordered by the caller,
 * or the caller's stream passed with cut_set_stream.  They
 * exist for callers whose data is already resident in HBM (benchmarks,
 * pipelines that produce RF on the GPU), for per-stage testing, and to select
 * the build's precision/interpolation variants.  All return false and print a
 * message on std::cerr on failure, like the reference API.
 */
#ifndef CUT_EXT_H
#define CUT_EXT_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif
#include "cuda_beamformer_parameters.h"

#define CUT_EXT_EXPORT __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---------------------------------------------------------- */
/* Use the caller's hipStream_t (NULL restores the library's own stream). */
CUT_EXT_EXPORT void cut_set_stream(void* hip_stream);
CUT_EXT_EXPORT bool cut_synchronize(void);
CUT_EXT_EXPORT const char* cut_version(void);

/* ---- variants ----------------------------------------------------------- */
/* decode: 0 auto (int8 MFMA for int16 input, FP32 MFMA GEMM otherwise),
 *         1 FP32 MFMA GEMM, 2 int8 MFMA (int16 input only). */
CUT_EXT_EXPORT bool cut_set_decode_mode(int mode);
/* DAS: cubic 1 = Catmull-Rom as the reference, 0 = linear;
 * cf_hercules 1 = apply the coherence factor in HERCULES too (the reference
 * computes it but does not apply it); iq_format: IQ handed from the Hilbert
 * transform to the DAS by beamform calls, 0 = complex f32, 1 = FP16 pairs
 * holding value * 2^-k (k fixed by the match filter so int16 data cannot
 * overflow; int16 input only), 2 = BF16 pairs.  Staged formats are HERCULES
 * with Hann apodization; accumulation stays FP32. */
CUT_EXT_EXPORT bool cut_set_das_options(int cubic, int cf_hercules, int iq_format);

/* ---- whole path on device pointers (async) ------------------------------ */
/* Same semantics as cuda_toolkit::beamform (cuda_toolkit/src/api/cuda_toolkit.cpp:109-199)
 * with the raw frame and the volume already in HBM. */
CUT_EXT_EXPORT bool cut_beamform_device(const void* d_raw, size_t raw_bytes, void* d_volume, size_t volume_bytes,
                                        const CudaBeamformerParameters* bp);

/* ---- batches of frames of any sizes ------------------------------------- */
typedef struct CutFrame {
    const void* d_raw;       /* raw frame in device memory, as cuda_toolkit::beamform's input */
    size_t raw_bytes;
    void* d_volume;          /* complex64 [Nz][Ny][Nx] */
    size_t volume_bytes;
    const CudaBeamformerParameters* bp;
} CutFrame;
/* Beamform n frames (each with its own transmit count, sample count, grid...)
 * and wait for them.  mode 0: every frame's decode and Hilbert, then one
 * persistent HERCULES DAS launch over the work of the whole batch, captured
 * into a HIP graph that is replayed while pointers, sizes and options stay the
 * same (frames must be HERCULES and share focus type and apodization); mode
 * 1: the single-frame path frame after frame; mode 2: as mode 0 with the
 * frames' work kept in frame order (lower per-frame latency, the launch grid
 * still tuned).  Each frame keeps its own RF
 * state and buffers across calls.  frame_ms (may be NULL): per frame, the time
 * from the start of the batch on the GPU to that frame's completion. */
CUT_EXT_EXPORT bool cut_beamform_batch(const CutFrame* frames, int n, int mode, double* frame_ms);
/* Timed autotuning.  The batch DAS launch shape (grid size, work order) is
 * tuned per batch shape by default: its candidates compute the same bits.
 * The choices that change the rounding (~1e-7 relative) -- the single-frame
 * HERCULES transmit split (1, 2, 4, 8 or 16 workgroups per tile, partial sums
 * added in split order), for FORCES, UFORCES and RCA_TPW the waves per 64-voxel
 * workgroup (1 to 16, each taking a share of the voxel's outer loop), and the
 * FFT tile -- come from fixed heuristics by default, so every process
 * computes the same bits, as the reference does.  on = 1 (or the environment
 * variable CUT_AUTOTUNE=1) also times those on first use per shape (that call
 * synchronises) and keeps the fastest; on = 0 turns off the batch tuning too. */
CUT_EXT_EXPORT bool cut_set_autotune(int on);
/* FFT tile of the Hilbert transforms (rows of up to 4096 samples): values
 * per thread in the LDS passes, 8 or 16; 0 (default) = tuned per row length,
 * row count, input kind, filter and output format on first use (with
 * autotuning on), else 16. */
CUT_EXT_EXPORT bool cut_set_fft_tile(int values_per_thread);
/* The tuned launch shapes (DAS split, waves per voxel group, FFT tile) differ
 * in the rounding of the results (~1e-7 relative), so two processes that tune
 * independently (timed tuning on) can differ in the last bits.  cut_tuning_save writes this
 * process's choices to a text file, cut_tuning_load reads such a file (its
 * choices replace tuning for those shapes); with the environment variable
 * CUT_TUNING_FILE set, the library reads that file on first use and rewrites
 * it after every new choice, so processes sharing it compute the same bits. */
CUT_EXT_EXPORT bool cut_tuning_save(const char* path);
CUT_EXT_EXPORT bool cut_tuning_load(const char* path);
/* Launch shape the last graph-mode batch used: resident workgroups per CU
 * (0 = one workgroup per tile) and work order (0 heaviest first, 1 frame by frame). */
CUT_EXT_EXPORT bool cut_batch_launch_shape(int* grid_per_cu, int* order);

/* ---- stages on device pointers (async), one RF state of their own ------- */
/* RfProcessor::init (cuda_toolkit/src/rf_processing/rf_processor.cpp:21-71):
 * raw_dims = {samples*transmits + pad, raw channels}; dec_dims = {S, C, T}. */
CUT_EXT_EXPORT bool cut_rf_init(const uint* raw_dims, const uint* dec_dims, int encoding);
CUT_EXT_EXPORT bool cut_rf_set_channel_mapping(const int16_t* map, uint count);
CUT_EXT_EXPORT bool cut_rf_set_match_filter(const float* taps, uint length);
/* convert (+ decode when decode != 0) into d_out[(t*C+c)*S+s]*out_stride (floats);
 * out_stride 2 writes the real parts of a complex buffer, as the reference does. */
CUT_EXT_EXPORT bool cut_rf_convert_decode(const void* d_raw, int type, float* d_out, int out_stride, int decode);
/* Bandpass input (replaces RfProcessor::convert_demod,
 * cuda_toolkit/src/rf_processing/rf_processor.cpp:225-246, and
 * DataConverter::convert_and_demod, data_conversion/data_converter.cu:64-131):
 * row map[c] holds (I, Q) pairs, sample s of transmit t at 2*(t*S + s);
 * d_out[t][c][s] (complex) = (I - iQ) exp(-2 pi i s demod_freq / sample_freq),
 * filtered, when match-filter taps are set, by the centred FIR
 * sum_i v[s + i - L/2] h[L-1-i] over the taps inside the row, divided by
 * their count.  Needs rf_raw_dim[0] >= 2*T*S. */
CUT_EXT_EXPORT bool cut_rf_convert_demod(const void* d_raw, int type, void* d_out, float demod_freq,
                                         float sample_freq);
/* Hilbert (+ match filter) of rows read at d_in[(row*S+s)*in_stride] -> complex d_out. */
CUT_EXT_EXPORT bool cut_rf_hilbert(const float* d_in, int in_stride, void* d_out);
/* DAS of complex IQ [T][C][S] into a complex volume [Nz][Ny][Nx]
 * (cuda_toolkit/src/beamformer/beamformer.cu:139-197). */
CUT_EXT_EXPORT bool cut_das_device(const void* d_iq, void* d_volume, const CudaBeamformerParameters* bp);
/* cut_das_device without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture or a tight serving loop.  The workspace (transmit-split
 * partial sums, RCA_TPW angle table) is a fixed 128 MiB device array per GPU,
 * held until the process ends; the HERCULES transmit split is limited to what
 * it holds (8 splits at 128^3 without the coherence factor, 4 with it), and
 * larger volumes run unsplit.  No launch shape is tuned by this call.  Calls
 * that overlap on one GPU (enqueued on two streams at once) race on the
 * array. */
CUT_EXT_EXPORT bool cut_das_device_noalloc(const void* d_iq, void* d_volume, const CudaBeamformerParameters* bp);
/* The same two stages with IQ in a given storage format (see
 * cut_set_das_options): the Hilbert output is written as value * scale
 * (*scale_out receives the scale used: 2^-k for FP16, 1 otherwise) and the
 * DAS reads IQ stored that way. */
CUT_EXT_EXPORT bool cut_rf_hilbert_format(const float* d_in, int in_stride, void* d_out, int format,
                                          float* scale_out);
CUT_EXT_EXPORT bool cut_das_device_format(const void* d_iq, int format, float scale, void* d_volume,
                                          const CudaBeamformerParameters* bp);

/* Number of (voxel, transmit, channel) taps whose apodization passes the
 * sequence's threshold: the work one DAS launch does (host computation). */
CUT_EXT_EXPORT bool cut_das_active_taps(const CudaBeamformerParameters* bp, unsigned long long* taps);

/* ---- 3-D block matching of envelope volumes ----------------------------- */
/* ZNCC of 16^3 template blocks (at `stride`, raster order x fastest) of d_tpl
 * against d_src ([nz][ny][nx] f32) over shifts |d| <= 8 that keep the source
 * block inside the volume.  d_centre == NULL: full search; otherwise only
 * |d - centre[b]| <= radius (radius <= 3, refinement).  d_block_ids (may be
 * NULL = all blocks in order) selects n_blocks blocks; outputs d_disp int
 * [n_blocks][3] (dx, dy, dz) and d_peak f32 [n_blocks].  The 2-D reference
 * counterpart is ImageProcessor::ncc_block_match
 * (cuda_toolkit/src/image_processing/image_processor.cpp:10-56). */
CUT_EXT_EXPORT bool cut_ncc3d_search(const float* d_tpl, const float* d_src, int nx, int ny, int nz, int stride,
                                     const int* d_block_ids, int n_blocks, const int* d_centre, int radius,
                                     int* d_disp, float* d_peak);
/* |IQ| of count complex samples into floats. */
CUT_EXT_EXPORT bool cut_envelope(const void* d_iq, size_t count, float* d_out);

/* ---- learned displacement proposal (no reference counterpart) ----------
 * A 3-D CNN over a pair of 16^3 blocks (template block, source block at the
 * same position, each normalised to zero mean and unit variance) predicts the
 * block's displacement in voxels.  Layers: conv 2->16 s2, 16->32 s2, 32->32
 * s1, 32->32 s2 (3^3 kernels, zero border 1, ReLU), FC 256->3.  Parameters are
 * FP32, packed w1 b1 w2 b2 w3 b3 w4 b4 wf bf, conv weights
 * [co][ci][kz][ky][kx], FC weights [3][256] over input c*8 + position.  All
 * pointers are device pointers; calls are asynchronous on the library stream
 * unless stated. */
CUT_EXT_EXPORT int cut_pnet_param_count(void);
/* Normalised block pairs for training: desc[i] = {pair, block, ox, oy, oz,
 * flips} selects template block `block` (blocks in (z, y, x) raster order at
 * `stride`) of volume pair `pair` (vols[2 pair], vols[2 pair + 1] are its
 * template and source volumes, [nz][ny][nx] f32) and the source block at the
 * same position plus (ox, oy, oz); bits 0/1/2 of flips mirror x/y/z.
 * out: f32 [n][2][16][16][16]. */
CUT_EXT_EXPORT bool cut_pnet_gather(const uint64_t* d_vols, int nx, int ny, int nz, int stride, const int* d_desc,
                                    int n, float* d_out);
/* Products of the training path (forward and gradient): 0 = BF16 MFMA with
 * FP32 accumulation (default), 1 = FP32 operands (verification).  The fused
 * inference always runs BF16 MFMA. */
CUT_EXT_EXPORT bool cut_pnet_set_precision(int mode);
/* Network input of a block: 0 = the 16^3 template block and the 16^3 source
 * block at the same position; 1 = the 32^3 neighbourhoods [b - 8, b + 24) of
 * both (the +-8 search region), each averaged 2x2x2 to 16^3. */
CUT_EXT_EXPORT bool cut_pnet_set_input_mode(int mode);
/* Training-path forward: x [B][2][16^3] -> y [B][3]. */
CUT_EXT_EXPORT bool cut_pnet_forward(const float* d_params, const float* d_x, int B, float* d_y);
/* loss = (1/B) sum_b 0.5 |y_b - t_b|^2 (d_loss: one f32, may be NULL) and its
 * gradient with respect to every parameter (d_grad: f32 [param_count]). */
CUT_EXT_EXPORT bool cut_pnet_loss_grad(const float* d_params, const float* d_x, const float* d_target, int B,
                                       float* d_grad, float* d_loss);
/* One Adam step (bias-corrected, step counts from 1; wd: decoupled weight decay). */
CUT_EXT_EXPORT bool cut_pnet_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, int n, float lr,
                                  float beta1, float beta2, float eps, int step, float wd);
/* Fused inference for blocks d_ids[0..n) (NULL: blocks 0..n-1) of a volume
 * pair: d_pred f32 [n][3] (dx, dy, dz); d_centre (may be NULL) int [n][3] =
 * clamp(rint(pred), -8, 8). */
CUT_EXT_EXPORT bool cut_pnet_infer(const float* d_params, const float* d_tpl, const float* d_src, int nx, int ny,
                                   int nz, int stride, const int* d_ids, int n, float* d_pred, int* d_centre);
/* Learned block matching of every block of the grid: CNN proposal, ZNCC
 * search within +-radius (<= 3) of the rounded proposal, and the full +-8
 * search for every block whose refined peak is below `threshold` or (when
 * boundary_check) whose refined maximum sits on a window face that is not a
 * limit of the full search.  d_disp int [nb][3], d_peak f32 [nb] as
 * cut_ncc3d_search; d_path (may be NULL) int [nb]: 1 where the full search
 * ran; d_pred (may be NULL) f32 [nb][3]: the proposals.  h_fallbacks (may be
 * NULL): host int receiving the number of full searches (the call then
 * synchronises). */
CUT_EXT_EXPORT bool cut_motion3d_learned(const float* d_params, const float* d_tpl, const float* d_src, int nx,
                                         int ny, int nz, int stride, int radius, float threshold, int boundary_check,
                                         int* d_disp, float* d_peak, int* d_path, float* d_pred, int* h_fallbacks);

/* ---- GL seam ----------------------------------------------------------- */
/* Register plain device buffers in place of GL buffers so the C API of
 * cuda_toolkit_ogl.h can run where no GL context exists (tests, headless
 * servers).  rf_bytes[i] is the size of rf_ptrs[i]. */
CUT_EXT_EXPORT bool cut_gl_register_device_buffers(void* const* rf_ptrs, const size_t* rf_bytes, uint count,
                                                   void* raw_ptr, size_t raw_bytes);

/* ---- synthetic acquisitions (test and benchmark inputs) ----------------- */
typedef struct CutSynthSpec {
    int geometry;          /* 0: linear array, FORCES delay model; 1: row-column array, HERCULES model */
    int T, C, S;           /* transmits (encoding order), channels, samples */
    int n_scatterers;
    float pitch;           /* [m] element pitch (both axes) */
    float c, fc, fs;       /* [m/s], [Hz], [Hz] */
    float box_min[3], box_max[3];  /* scatterer box [m] */
    float amp_min, amp_max;
    float noise_frac;      /* noise std as a fraction of the peak single-element response */
    float peak;            /* int16 full scale of the encoded data */
    int encoding;          /* NONE, HADAMARD or WALSH */
    int cycles;            /* pulse length in periods */
    unsigned long long seed;
    float shift[3];        /* rigid displacement applied to every scatterer [m] */
    float decorrelation;   /* fraction of scatterer amplitude redrawn (0 = identical speckle) */
    unsigned long long amp_seed; /* seed of the redrawn amplitudes */
} CutSynthSpec;
/* raw: int16 [C][T*S] (row r = raw channel r).  map (C entries, may be NULL)
 * receives decoded channel -> raw row, a permutation drawn from the seed. */
CUT_EXT_EXPORT bool cut_synth_rf(const CutSynthSpec* spec, int16_t* raw, int16_t* map);

#ifdef __cplusplus
}
#endif
#endif /* CUT_EXT_H */
