// Host-side pipeline objects shared by the API layers.
#pragma once

#include <mutex>
#include <vector>

#include "common.hpp"

namespace cut {

// Row-major Hadamard matrix of order n in the reference's construction and
// row order (cuda_toolkit/src/rf_processing/hadamard/hadamard_decoder.cpp:110-222).
// Returns false (with a message) for unsupported orders.
bool hadamard_matrix(int n, int ordering, std::vector<int>& H);

// Library-wide options (extensions; defaults reproduce the reference).
struct Options {
    int decode_mode = 0;    // 0 auto (int8 MFMA for int16 input), 1 FP32 GEMM, 2 int8 MFMA
    int das_cubic = 1;      // 1 Catmull-Rom (reference), 0 linear
    int das_cf_hercules = 0;
    int iq_format = 0;      // IqFormat of the IQ handed from the Hilbert transform to the DAS
    int fft_vpt = 0;        // FFT tile of the Hilbert kernels: 0 tuned, 8 or 16 values per thread
};
Options& options();

// RF front end: convert/remap, decode, Hilbert.  Mirrors the reference's
// RfProcessor (cuda_toolkit/src/rf_processing/rf_processor.h:19-74): state is
// cached by (raw dims, decoded dims, encoding) and rebuilt when they change.
class RfPipeline {
public:
    bool init(uint raw_len, uint raw_channels, uint S, uint C, uint T, int encoding);
    bool deinit();
    bool initialized() const { return init_; }
    bool set_channel_mapping(const int16_t* map, size_t count);
    bool set_match_filter(const float* h, size_t length);
    // decode (encoding != NONE) or plain convert into out[(t*C+c)*S+s]*stride
    bool convert_decode(const void* d_raw, int type, float* d_out, int out_stride, bool decode, hipStream_t st);
    bool hilbert(const HilbertInput& in, float2* d_out, hipStream_t st) { return hilbert(in, d_out, IQ_F32, 1.0f, st); }
    // out in IqFormat fmt, holding value * scale
    bool hilbert(const HilbertInput& in, void* d_out, int fmt, float scale, hipStream_t st);
    // Bandpass input: (I, Q) pairs per sample, demodulated by exp(-2 pi i s
    // fd/fs) and filtered by the current taps as a centred FIR
    // (RfProcessor::convert_demod, rf_processing/rf_processor.cpp:225-246)
    bool convert_demod(const void* d_raw, int type, float2* d_out, double fd_fs, hipStream_t st);
    // Hilbert that converts raw input on the fly (decode == NONE fast path)
    bool convert_hilbert(const void* d_raw, int type, void* d_out, int fmt, float scale, hipStream_t st);
    // Power-of-two scale for FP16-staged IQ of int16 input: the analytic
    // signal of int16 data filtered by the current taps stays below 32768.
    float fp16_scale() const;
    bool check_map(uint raw_channels) const;
    bool build_conv_table();

    uint S() const { return S_; }
    uint C() const { return C_; }
    uint T() const { return T_; }
    uint raw_len() const { return R0_; }
    uint raw_channels() const { return RC_; }
    int encoding() const { return enc_; }

private:
    bool init_ = false;
    uint R0_ = 0, RC_ = 0, S_ = 0, C_ = 0, T_ = 0;
    int enc_ = HADAMARD;
    DeviceBuffer map_;
    std::vector<int16_t> map_host_;
    DeviceBuffer h8_, hf_, rowsum_;
    float scale_ = 1.0f;
    DeviceBuffer tw_;
    DeviceBuffer conv_H_, conv_tmp_;  // convolution kernel spectrum (lengths that are not powers of two)
    DeviceBuffer tw2_, long_ws_;    // rows > 4096: 4096-point twiddles, workspace of the HBM-staged steps
    DeviceBuffer filter_taps_, filter_spec_;
    std::vector<float> filter_host_;
    bool filter_on_ = false;
};

// Whole path raw -> volume on device pointers (the body of beamform()).
struct BeamformBuffers {
    DeviceBuffer raw, dec, iq, vol, das_ws;
};
// Every check beamform() makes before touching the GPU (host only).
bool validate_beamform(const CudaBeamformerParameters& bp, size_t raw_bytes, size_t vol_bytes);
bool beamform_device(RfPipeline& rf, BeamformBuffers& bufs, const void* d_raw, size_t raw_bytes, float2* d_vol,
                     size_t vol_bytes, const CudaBeamformerParameters& bp, hipStream_t st);
// The same in two halves: host-side set-up of one frame (validation, RF state,
// DAS constants, buffers; may synchronise), then the decode + Hilbert launches
// (launches only, so they can be captured into a graph).
struct FramePlan {
    DasParams p;
    DasOptions opts;
    int fmt = IQ_F32;
    float scale = 1.0f;
};
bool prepare_frame(RfPipeline& rf, BeamformBuffers& bufs, const CudaBeamformerParameters& bp, size_t raw_bytes,
                   size_t vol_bytes, FramePlan& plan);
bool frame_to_iq(RfPipeline& rf, BeamformBuffers& bufs, const void* d_raw, const CudaBeamformerParameters& bp,
                 const FramePlan& plan, hipStream_t st);

unsigned long long das_active_taps(const DasParams& p);

std::mutex& api_mutex();

}  // namespace cut
