// Shared host-side helpers: error convention, device buffers, kernel-side
// parameter blocks.  Error convention follows the reference
// (cuda_toolkit/src/defs.h:117-178): failing calls print one line on std::cerr
// and the entry point returns false.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <iostream>
#include <vector>

#include "../../include/cuda_beamformer_parameters.h"

#define CUT_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t cut_err_ = (expr);                                                    \
        if (cut_err_ != hipSuccess) {                                                    \
            std::cerr << "HIP error " << hipGetErrorName(cut_err_) << " '"              \
                      << hipGetErrorString(cut_err_) << "' at " << __FILE__ << ':'       \
                      << __LINE__ << std::endl;                                          \
            return false;                                                                \
        }                                                                                \
    } while (0)

#define CUT_FAIL(msg)                                 \
    do {                                              \
        std::cerr << "cuda_toolkit: " << msg << std::endl; \
        return false;                                 \
    } while (0)

namespace cut {

// Device allocation reused while the requested size is unchanged (the
// reference's size-keyed buffer cache, cuda_toolkit/src/api/cuda_toolkit.cpp:43-82).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    bool ensure(size_t n) {
        if (n == bytes && ptr) return true;
        release();
        if (n == 0) return true;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            std::cerr << "cuda_toolkit: hipMalloc(" << n << ") failed: " << hipGetErrorString(e) << std::endl;
            ptr = nullptr;
            return false;
        }
        bytes = n;
        return true;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// Library stream: every entry point enqueues on it.  Callers may substitute
// their own stream (cut_set_stream) so that their events bracket our kernels.
hipStream_t stream();
void set_stream(hipStream_t s);

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).
// ---------------------------------------------------------------------------

// Convert (cuda_toolkit/src/rf_processing/data_conversion/data_conversion_kernels.cuh:25-41):
// out[(t*C + c)*S + s] * out_stride = raw[map[c]*R0 + t*S + s].
bool launch_convert(const void* raw, int type, const int16_t* d_map, size_t R0, int T, int C, int S,
                    float* out, int out_stride, hipStream_t st);

// Bandpass input (data_conversion_kernels.cuh:14-102): (I, Q) pairs at
// raw[map[c]*R0 + 2*(t*S + s)], v = (I - iQ) exp(-2 pi i s fd/fs), then with
// L > 0 taps the centred FIR sum_i v[s + i - L/2] h[L-1-i] over the taps
// inside the row, divided by their count.  out: complex [T][C][S].
bool launch_convert_demod(const void* raw, int type, const int16_t* d_map, size_t R0, int T, int C, int S,
                          double fd_fs, const float* d_h, int L, float2* out, hipStream_t st);

// Decode (hadamard_decoder.cpp:49-78): out[j][c][s] = scale * sum_k H[j][k] raw[map[c]][k*S+s].
// mode 0: FP32 MFMA GEMM (any input type); mode 1: int8 MFMA on hi/lo byte
// splits of int16 input with int32 accumulation.  d_h8: int8 H padded to
// Kp x Kp (Kp = T rounded up to 32); d_hf: float H, same padding; d_rowsum:
// int32 row sums of H (length Kp).
bool launch_decode(const void* raw, int type, const int16_t* d_map, size_t R0, int T, int C, int S,
                   const int8_t* d_h8, const float* d_hf, const int* d_rowsum, float scale,
                   float* out, int out_stride, int mode, hipStream_t st);

// Hilbert transform with optional match filter (hilbert_handler.cu:132-155).
// Input row r of length S is read from a float array at in[(r*S + s)*in_stride]
// (in_kind 0), or, convert fused, from raw int16/f32 with channel map
// (in_kind 1 = i16, 2 = f32; row r = t*C + c).
struct HilbertInput {
    const void* ptr;
    int kind;        // 0 float strided, 1 raw i16, 2 raw f32
    int stride;      // kind 0
    const int16_t* map;
    size_t R0;
    int C;
};
bool launch_twiddles(float2* d_tw, int S, hipStream_t st);
bool launch_filter_spectrum(const float* d_h, int L, int S, float2* d_F, hipStream_t st);
// Storage formats of IQ (analytic-signal) samples between the Hilbert
// transform and the DAS: complex f32, half2 holding value * scale (scale a
// power of two chosen so int16-derived data cannot overflow), bf16x2.
enum IqFormat { IQ_F32 = 0, IQ_F16 = 1, IQ_BF16 = 2 };
inline size_t iq_sample_bytes(int fmt) { return fmt == IQ_F32 ? 8 : 4; }
// vpt: the FFT tile, values per thread in the LDS transforms (8 or 16; see
// RfPipeline::hilbert for how it is chosen).  Both compute the same transform
// with a different pass structure (rounding differs at the 1e-7 level).
bool launch_hilbert(const HilbertInput& in, int rows, int S, const float2* d_tw, const float2* d_F, void* out,
                    int out_fmt, float out_scale, int vpt, hipStream_t st);
bool hilbert_supported(int S);  // 1 <= S <= 4096, powers of two up to 65536, others up to 32768
bool hilbert_pow2(int S);       // powers of two 16..4096 take the radix-16 kernel
// Other lengths: the operator as a cyclic convolution, evaluated with
// M = 2^conv_log2(S) >= 2S - 1 point FFTs of the zero-padded rows; H = the
// M-point spectrum of the operator's kernel h (with weights, filter and 1/M),
// built in FP64 by launch_conv_tables (d_tmp: conv_tables_tmp_bytes(S)); d_h,
// L: the match filter taps (L = 0: none).  pair: the table of the LDS row-pair
// kernel, M float4 (H_r, H_i) = the spectra of Re h and Im h; else M float2
// H (the long path).
int conv_log2(int S);
bool hilbert_conv(int S);     // the length takes the convolution form
int hilbert_fft_len(int S);   // transform length: S, or 2^conv_log2(S) for the convolution form
size_t conv_tables_tmp_bytes(int S);
bool launch_conv_tables(const float* d_h, int L, int S, void* d_tmp, void* d_H, bool pair, hipStream_t st);
bool launch_hilbert_conv(const HilbertInput& in, int rows, int S, const float2* d_tw, const float4* d_H, void* out,
                         int ofmt, float oscale, int vpt, hipStream_t st);
// Rows of more than 4096 samples: N-point transforms (N = S for powers of
// two, else the convolution length 2^conv_log2(S)) as N / 4096 x 4096 steps
// through HBM, N <= HILBERT_LONG_MAX_N.  Tables: twN = N-point twiddles, tw2 =
// 4096-point twiddles, H = the convolution spectrum (other lengths).  The
// steps run in place in the output when it is complex f32 and S is a power
// of two; otherwise in a workspace of ws_rows rows of N complex values
// (hilbert_long_chunk_rows), rows taken in chunks of ws_rows.
constexpr int HILBERT_LONG_MAX_N = 65536;
bool hilbert_long(int S);       // S > 4096
int hilbert_long_n(int S);
size_t hilbert_long_chunk_rows(int S, int rows);
bool hilbert_long_in_place(int S, int ofmt);
struct LongHilbertTables {
    const float2* twN;
    const float2* tw2;
    const float2* H;
};
bool launch_hilbert_long(const HilbertInput& in, int rows, int S, const LongHilbertTables& t, const float2* d_F,
                         void* out, int ofmt, float oscale, float2* ws, size_t ws_rows, hipStream_t st);

// DAS parameters in kernel form (beamformer.cu:17-98).
struct DasParams {
    int S, C, T;
    int Nx, Ny, Nz;
    float min_x, min_y, min_z;
    float res_x, res_y, res_z;
    float xdc_min_x, xdc_min_y, xdc_max_x, xdc_max_y;
    float pitch_x, pitch_y;
    float spm;
    int delay;
    float focal_z;
    int focus_type;  // 0 none, 1 XZ_PLANE, 2 YZ_PLANE, 3 XZ_FOCUS, 4 YZ_FOCUS
    float fn_tx, fn_rx, cw;
    int apo_type, to_power;
    int readi_groups, encoding;
    unsigned long long readi_bits;
    int sequence;
    // per-transmit geometry of UFORCES and RCA_TPW (beamformer_kernels.cu):
    // UFORCES: tx_tab[t].x = sparse_elements[t] - sparse_elements[0];
    // RCA_TPW: tx_tab[t].x = steering angle foci[3t+2] [rad]
    double fc_fs;  // centre / sampling frequency (RCA_TPW phase rotation)
    float2 tx_tab[MAX_ACQ_COUNT];
};
bool das_params_from_bp(const CudaBeamformerParameters& bp, DasParams& p);

struct DasOptions {
    int cubic = 1;         // 1 = Catmull-Rom (reference), 0 = linear
    int cf_hercules = 0;   // apply the coherence factor in HERCULES as well
};
// HERCULES splits the transmit loop over workgroups and needs a workspace for
// the partial sums (das_workspace_bytes; 0 when no split is used).
int das_splits(long long tiles, int gt_total);
// Autotuning of the transmit split (on by default; the workspace then covers
// the largest candidate).
void set_das_autotune(bool on);
size_t das_workspace_bytes(const DasParams& p);
// iq_fmt: an IqFormat; the stored samples are value * iq_scale.  Staged
// (F16/BF16) IQ is supported by the HERCULES kernel with Hann apodization.
// The HERCULES transmit split is limited to what the workspace holds.
// allow_tune false: no launch-shape timing (and no HIP events created) on
// this call; a shape tuned earlier for the same key is still used.
bool launch_das(const void* iq, int iq_fmt, float iq_scale, const DasParams& p, const DasOptions& o,
                float2* vol, void* workspace, size_t workspace_bytes, hipStream_t st, bool allow_tune = true);

// Batched HERCULES DAS: frames of any sizes in one launch.  Items are
// (frame, tile) pairs, tile = wz * nxt * nyt + wy * nxt + wx with wz counted
// from the deepest tile row.  d_next: the item counter; d_done[f]: tiles of
// frame f finished; d_finish[f]: wall-clock time at which frame f completed.
// All frames share the template variant (focus type, apodization, IQ format).
struct BatchFrameDev {
    DasParams p;
    const void* iq;
    float2* vol;
    float iq_inv;
    int tiles, nxt, nyt, nzt;
};
void das_tile_grid(const DasParams& p, int& nxt, int& nyt, int& nzt);
// HERCULES launches the generic-apodization kernels: the sine apodization, or
// a Hann volume with a voxel at z <= 0 or a negative F-number (exact weight).
bool herc_generic_apo(const DasParams& p);
bool launch_batch_start(int* d_counters, int n, long long* d_start, hipStream_t st);
bool launch_das_batch(const BatchFrameDev* d_frames, const int2* d_items, int n_items, int grid, int* d_next,
                      int* d_done, long long* d_finish, const DasOptions& o, bool focus, bool sinapo, int fmt,
                      hipStream_t st);

}  // namespace cut
