// C++ span API (cuda_toolkit.hpp), MATLAB shims (cuda_toolkit_matlab.h) and
// the extension C ABI (cut_ext.h).
//
// Replaces cuda_toolkit/src/api/cuda_toolkit.cpp:109-199 (beamform) and
// cuda_toolkit/src/api/cuda_toolkit_matlab.cpp:9-37.  State lives in
// process-lifetime singletons (never destroyed, so no HIP call runs during
// static destruction) and every entry point holds one mutex: the reference
// is not thread safe, this build serialises callers instead.
#include "../../include/cuda_toolkit.hpp"
#include "../../include/cuda_toolkit_matlab.h"
#include "../../include/cut_ext.h"
#include "pipeline.hpp"

#include <exception>

namespace {

struct BeamformState {
    cut::RfPipeline rf;
    cut::BeamformBuffers bufs;
};

BeamformState& bf_state()
{
    static BeamformState* s = new BeamformState();
    return *s;
}

struct StageState {
    cut::RfPipeline rf;
    cut::DeviceBuffer das_ws;
};

StageState& stage_state()
{
    static StageState* s = new StageState();
    return *s;
}

template <typename F>
bool guarded(const char* what, F&& f)
{
    try {
        return f();
    } catch (const std::exception& e) {
        std::cerr << "cuda_toolkit: " << what << ": " << e.what() << std::endl;
    } catch (...) {
        std::cerr << "cuda_toolkit: " << what << ": unknown exception" << std::endl;
    }
    return false;
}

}  // namespace

bool cuda_toolkit::beamform(std::span<const uint8_t> input_data, std::span<uint8_t> output_data,
                            const CudaBeamformerParameters& bp)
{
    return guarded("beamform", [&] {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        auto& s = bf_state();
        hipStream_t st = cut::stream();
        const size_t nvox = (size_t)bp.output_points[0] * bp.output_points[1] * bp.output_points[2];
        const size_t vol_bytes = nvox * sizeof(float2);
        if (nvox == 0) CUT_FAIL("Empty output grid.");
        if (output_data.size() < vol_bytes)
            CUT_FAIL("Output span holds " << output_data.size() << " bytes, the volume needs " << vol_bytes << ".");
        if (input_data.empty()) CUT_FAIL("Empty input span.");
        if (!cut::validate_beamform(bp, input_data.size(), output_data.size())) {
            std::cerr << "Beamforming failed." << std::endl;
            return false;
        }
        if (!s.bufs.raw.ensure(input_data.size()) || !s.bufs.vol.ensure(vol_bytes)) {
            std::cerr << "Failed to update beamformer buffers." << std::endl;
            return false;
        }
        CUT_HIP_CHECK(hipMemcpyAsync(s.bufs.raw.ptr, input_data.data(), input_data.size(), hipMemcpyHostToDevice, st));
        if (!cut::beamform_device(s.rf, s.bufs, s.bufs.raw.ptr, input_data.size(), s.bufs.vol.as<float2>(), vol_bytes,
                                  bp, st)) {
            (void)hipStreamSynchronize(st);
            std::cerr << "Beamforming failed." << std::endl;
            return false;
        }
        CUT_HIP_CHECK(hipMemcpyAsync(output_data.data(), s.bufs.vol.ptr, vol_bytes, hipMemcpyDeviceToHost, st));
        CUT_HIP_CHECK(hipStreamSynchronize(st));
        return true;
    });
}

// ---------------------------------------------------------------------------
// MATLAB shims (cuda_toolkit_matlab.cpp:9-37): spans sized from the BP.
// ---------------------------------------------------------------------------
extern "C" void beamform_i16_matlab(const short* data, CudaBeamformerParameters bp, float* output)
{
    const size_t in_bytes = (size_t)bp.rf_raw_dim[0] * bp.rf_raw_dim[1] * sizeof(short);
    const size_t out_bytes = (size_t)bp.output_points[0] * bp.output_points[1] * bp.output_points[2] * 2 * sizeof(float);
    bp.data_type = TYPE_I16;
    if (!cuda_toolkit::beamform({reinterpret_cast<const uint8_t*>(data), in_bytes},
                                {reinterpret_cast<uint8_t*>(output), out_bytes}, bp))
        std::cerr << "Error: Beamforming failed." << std::endl;
}

extern "C" void beamform_f32_matlab(const float* data, CudaBeamformerParameters bp, float* output)
{
    const size_t in_bytes = (size_t)bp.rf_raw_dim[0] * bp.rf_raw_dim[1] * sizeof(float);
    const size_t out_bytes = (size_t)bp.output_points[0] * bp.output_points[1] * bp.output_points[2] * 2 * sizeof(float);
    if (!cuda_toolkit::beamform({reinterpret_cast<const uint8_t*>(data), in_bytes},
                                {reinterpret_cast<uint8_t*>(output), out_bytes}, bp))
        std::cerr << "Error: Beamforming failed." << std::endl;
}

// ---------------------------------------------------------------------------
// Extension C ABI
// ---------------------------------------------------------------------------
extern "C" {

void cut_set_stream(void* hip_stream)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    cut::set_stream((hipStream_t)hip_stream);
}

bool cut_synchronize(void)
{
    CUT_HIP_CHECK(hipStreamSynchronize(cut::stream()));
    return true;
}

const char* cut_version(void) { return "cuda_ultrasound_toolkit_amd 0.1 (gfx950)"; }

bool cut_set_decode_mode(int mode)
{
    if (mode < 0 || mode > 2) CUT_FAIL("decode mode must be 0, 1 or 2");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    cut::options().decode_mode = mode;
    return true;
}

bool cut_set_das_options(int cubic, int cf_hercules, int iq_format)
{
    if (iq_format < cut::IQ_F32 || iq_format > cut::IQ_BF16) CUT_FAIL("IQ format " << iq_format << " not available");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    cut::options().das_cubic = cubic ? 1 : 0;
    cut::options().das_cf_hercules = cf_hercules ? 1 : 0;
    cut::options().iq_format = iq_format;
    return true;
}

bool cut_set_fft_tile(int values_per_thread)
{
    if (values_per_thread != 0 && values_per_thread != 8 && values_per_thread != 16)
        CUT_FAIL("FFT tile must be 0 (tuned), 8 or 16 values per thread");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    cut::options().fft_vpt = values_per_thread;
    return true;
}

bool cut_beamform_device(const void* d_raw, size_t raw_bytes, void* d_volume, size_t volume_bytes,
                         const CudaBeamformerParameters* bp)
{
    return guarded("cut_beamform_device", [&] {
        if (!d_raw || !d_volume || !bp) CUT_FAIL("null argument");
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        auto& s = bf_state();
        return cut::beamform_device(s.rf, s.bufs, d_raw, raw_bytes, (float2*)d_volume, volume_bytes, *bp, cut::stream());
    });
}

bool cut_rf_init(const uint* raw_dims, const uint* dec_dims, int encoding)
{
    if (!raw_dims || !dec_dims) CUT_FAIL("null argument");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    return stage_state().rf.init(raw_dims[0], raw_dims[1], dec_dims[0], dec_dims[1], dec_dims[2], encoding);
}

bool cut_rf_set_channel_mapping(const int16_t* map, uint count)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    return stage_state().rf.set_channel_mapping(map, count);
}

bool cut_rf_set_match_filter(const float* taps, uint length)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    return stage_state().rf.set_match_filter(taps, length);
}

bool cut_rf_convert_decode(const void* d_raw, int type, float* d_out, int out_stride, int decode)
{
    if (!d_raw || !d_out) CUT_FAIL("null argument");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    return stage_state().rf.convert_decode(d_raw, type, d_out, out_stride, decode != 0, cut::stream());
}

bool cut_rf_convert_demod(const void* d_raw, int type, void* d_out, float demod_freq, float sample_freq)
{
    if (!d_raw || !d_out) CUT_FAIL("null argument");
    if (!(sample_freq > 0.0f)) CUT_FAIL("sample frequency must be positive");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    return stage_state().rf.convert_demod(d_raw, type, (float2*)d_out, (double)demod_freq / (double)sample_freq,
                                          cut::stream());
}

bool cut_rf_hilbert(const float* d_in, int in_stride, void* d_out)
{
    if (!d_in || !d_out) CUT_FAIL("null argument");
    if (in_stride != 1 && in_stride != 2) CUT_FAIL("input stride must be 1 or 2");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    auto& rf = stage_state().rf;
    cut::HilbertInput in{d_in, 0, in_stride, nullptr, 0, (int)rf.C()};
    return rf.hilbert(in, (float2*)d_out, cut::stream());
}

bool cut_rf_hilbert_format(const float* d_in, int in_stride, void* d_out, int format, float* scale_out)
{
    if (!d_in || !d_out) CUT_FAIL("null argument");
    if (in_stride != 1 && in_stride != 2) CUT_FAIL("input stride must be 1 or 2");
    if (format < cut::IQ_F32 || format > cut::IQ_BF16) CUT_FAIL("IQ format " << format << " not available");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    auto& rf = stage_state().rf;
    const float scale = format == cut::IQ_F16 ? rf.fp16_scale() : 1.0f;
    if (scale_out) *scale_out = scale;
    cut::HilbertInput in{d_in, 0, in_stride, nullptr, 0, (int)rf.C()};
    return rf.hilbert(in, d_out, format, scale, cut::stream());
}

bool cut_das_device_format(const void* d_iq, int format, float scale, void* d_volume,
                           const CudaBeamformerParameters* bp)
{
    return guarded("cut_das_device_format", [&] {
        if (!d_iq || !d_volume || !bp) CUT_FAIL("null argument");
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        cut::DasParams p;
        if (!cut::das_params_from_bp(*bp, p)) return false;
        cut::DasOptions o;
        o.cubic = cut::options().das_cubic;
        o.cf_hercules = cut::options().das_cf_hercules;
        auto& ws = stage_state().das_ws;
        const size_t need = cut::das_workspace_bytes(p);
        if (need && !ws.ensure(std::max(need, ws.bytes))) return false;
        return cut::launch_das(d_iq, format, scale, p, o, (float2*)d_volume, ws.ptr, ws.bytes, cut::stream());
    });
}

bool cut_das_device(const void* d_iq, void* d_volume, const CudaBeamformerParameters* bp)
{
    return guarded("cut_das_device", [&] {
        if (!d_iq || !d_volume || !bp) CUT_FAIL("null argument");
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        cut::DasParams p;
        if (!cut::das_params_from_bp(*bp, p)) return false;
        cut::DasOptions o;
        o.cubic = cut::options().das_cubic;
        o.cf_hercules = cut::options().das_cf_hercules;
        auto& ws = stage_state().das_ws;
        const size_t need = cut::das_workspace_bytes(p);
        if (need && !ws.ensure(std::max(need, ws.bytes))) return false;
        return cut::launch_das(d_iq, 0, 1.0f, p, o, (float2*)d_volume, ws.ptr, ws.bytes, cut::stream());
    });
}

bool cut_das_active_taps(const CudaBeamformerParameters* bp, unsigned long long* taps)
{
    if (!bp || !taps) CUT_FAIL("null argument");
    cut::DasParams p;
    if (!cut::das_params_from_bp(*bp, p)) return false;
    *taps = cut::das_active_taps(p);
    return true;
}

}  // extern "C"
