// C API over OpenGL buffers (cuda_toolkit_ogl.h).
//
// Replaces cuda_toolkit/src/api/cuda_toolkit_ogl.cpp:1-273.  GL buffers are
// registered with hipGraphicsGLRegisterBuffer and mapped around each call, as
// the reference does with CUDA-GL interop.  A second backing, plain device
// buffers registered through cut_gl_register_device_buffers, lets the same
// entry points run where no GL context exists.
#include <array>

#include "pipeline.hpp"
// after hip_runtime.h (pipeline.hpp), which declares the types it uses
#include <hip/hip_gl_interop.h>

#include "../../include/cuda_toolkit_ogl.h"
#include "../../include/cut_ext.h"

namespace {

constexpr uint kMaxBuffers = 16;

struct Slot {
    uint ssbo = 0;
    hipGraphicsResource_t res = nullptr;
    void* dev = nullptr;  // device-buffer backing
    size_t bytes = 0;
};

struct GraphicsSession {
    cut::RfPipeline rf;
    Slot raw;
    std::array<Slot, kMaxBuffers> rf_bufs;
    uint rf_count = 0;
    bool buffers_init = false;
};

GraphicsSession& session()
{
    static GraphicsSession* s = new GraphicsSession();
    return *s;
}

bool unregister_all()
{
    auto& s = session();
    bool ok = true;
    auto drop = [&](Slot& slot) {
        if (slot.res) {
            hipError_t e = hipGraphicsUnregisterResource(slot.res);
            if (e != hipSuccess) {
                std::cerr << "Failed to unregister OpenGL buffer: " << hipGetErrorString(e) << std::endl;
                ok = false;
            }
        }
        slot = Slot{};
    };
    drop(s.raw);
    for (auto& b : s.rf_bufs) drop(b);
    s.rf_count = 0;
    s.buffers_init = false;
    return ok;
}

bool map_slot(Slot& slot, void** ptr, size_t* bytes)
{
    if (slot.dev) {
        *ptr = slot.dev;
        *bytes = slot.bytes;
        return true;
    }
    if (!slot.res) CUT_FAIL("OpenGL resource is null.");
    CUT_HIP_CHECK(hipGraphicsMapResources(1, &slot.res, cut::stream()));
    CUT_HIP_CHECK(hipGraphicsResourceGetMappedPointer(ptr, bytes, slot.res));
    if (!*ptr) CUT_FAIL("Failed to map OpenGL buffer.");
    return true;
}

void unmap_slot(Slot& slot)
{
    if (slot.res) (void)hipGraphicsUnmapResources(1, &slot.res, cut::stream());
}

}  // namespace

extern "C" {

bool init_cuda_configuration(const uint* input_dims, const uint* decoded_dims)
{
    if (!input_dims || !decoded_dims) CUT_FAIL("null argument");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    if (!session().rf.init(input_dims[0], input_dims[1], decoded_dims[0], decoded_dims[1], decoded_dims[2], HADAMARD)) {
        std::cerr << "Failed to initialize CUDA session." << std::endl;
        return false;
    }
    return true;
}

void deinit_cuda_configuration(void)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    (void)hipStreamSynchronize(cut::stream());
    session().rf.deinit();
    unregister_all();
}

bool register_cuda_buffers(const uint* rf_data_ssbos, uint rf_buffer_count, uint raw_data_ssbo)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    unregister_all();
    auto& s = session();
    if (rf_buffer_count > kMaxBuffers) CUT_FAIL("Too many RF data buffers. Maximum is " << kMaxBuffers << ".");
    if (rf_buffer_count > 0 && !rf_data_ssbos) CUT_FAIL("null buffer list");
    s.raw.ssbo = raw_data_ssbo;
    CUT_HIP_CHECK(hipGraphicsGLRegisterBuffer(&s.raw.res, raw_data_ssbo, hipGraphicsRegisterFlagsNone));
    for (uint i = 0; i < rf_buffer_count; ++i) {
        const uint ssbo = rf_data_ssbos[i];
        if (ssbo == 0) {
            std::cerr << "Invalid SSBO at index " << i << "." << std::endl;
            unregister_all();
            return false;
        }
        hipError_t e = hipGraphicsGLRegisterBuffer(&s.rf_bufs[i].res, ssbo, hipGraphicsRegisterFlagsNone);
        if (e != hipSuccess) {
            std::cerr << "HIP error " << hipGetErrorString(e) << " registering SSBO " << ssbo << std::endl;
            unregister_all();
            return false;
        }
        s.rf_bufs[i].ssbo = ssbo;
    }
    s.rf_count = rf_buffer_count;
    s.buffers_init = true;
    return true;
}

bool cut_gl_register_device_buffers(void* const* rf_ptrs, const size_t* rf_bytes, uint count, void* raw_ptr,
                                    size_t raw_bytes)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    unregister_all();
    auto& s = session();
    if (count > kMaxBuffers) CUT_FAIL("Too many RF data buffers. Maximum is " << kMaxBuffers << ".");
    if (!raw_ptr || (count && (!rf_ptrs || !rf_bytes))) CUT_FAIL("null argument");
    s.raw.dev = raw_ptr;
    s.raw.bytes = raw_bytes;
    for (uint i = 0; i < count; ++i) {
        if (!rf_ptrs[i]) {
            unregister_all();
            CUT_FAIL("Invalid buffer at index " << i << ".");
        }
        s.rf_bufs[i].dev = rf_ptrs[i];
        s.rf_bufs[i].bytes = rf_bytes[i];
    }
    s.rf_count = count;
    s.buffers_init = true;
    return true;
}

bool cuda_set_channel_mapping(const i16 channel_mapping[MAX_CHANNEL_COUNT])
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    if (!session().rf.set_channel_mapping(channel_mapping, MAX_CHANNEL_COUNT)) {
        std::cerr << "Failed to set channel mapping." << std::endl;
        return false;
    }
    return true;
}

bool cuda_set_match_filter(const float* match_filter, uint length)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    if (!session().rf.set_match_filter(match_filter, length)) {
        std::cerr << "Failed to set match filter." << std::endl;
        return false;
    }
    return true;
}

bool cuda_decode(size_t input_offset, uint output_buffer_idx)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    auto& s = session();
    if (!s.buffers_init) CUT_FAIL("OGL buffers not registered.");
    if (output_buffer_idx >= s.rf_count) CUT_FAIL("Output buffer index " << output_buffer_idx << " not registered.");
    auto& rf = s.rf;
    if (!rf.initialized()) CUT_FAIL("Session not initialized.");
    if (input_offset % sizeof(int16_t)) CUT_FAIL("Input offset must be a multiple of 2 bytes.");
    Slot& in = s.raw;
    Slot& out = s.rf_bufs[output_buffer_idx];
    void *pin = nullptr, *pout = nullptr;
    size_t bin = 0, bout = 0;
    if (!map_slot(in, &pin, &bin)) CUT_FAIL("Failed to map input OpenGL buffer.");
    if (!map_slot(out, &pout, &bout)) {
        unmap_slot(in);
        CUT_FAIL("Failed to map output OpenGL buffer.");
    }
    const size_t need_in = input_offset + (size_t)rf.raw_len() * rf.raw_channels() * sizeof(int16_t);
    const size_t need_out = (size_t)rf.S() * rf.C() * rf.T() * sizeof(float2);
    bool ok = true;
    if (bin < need_in) { std::cerr << "cuda_toolkit: raw buffer too small for the configured dims" << std::endl; ok = false; }
    if (bout < need_out) { std::cerr << "cuda_toolkit: RF buffer too small for the configured dims" << std::endl; ok = false; }
    if (ok) {
        const int16_t* src = (const int16_t*)((const char*)pin + input_offset);
        ok = rf.convert_decode(src, TYPE_I16, (float*)pout, 2, rf.encoding() != NONE, cut::stream());
        if (!ok) std::cerr << "Failed to decode data." << std::endl;
    }
    if (ok) ok = hipStreamSynchronize(cut::stream()) == hipSuccess;
    unmap_slot(in);
    unmap_slot(out);
    return ok;
}

bool cuda_hilbert(uint input_buffer_idx, uint output_buffer_idx)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    auto& s = session();
    if (!s.buffers_init) CUT_FAIL("OGL buffers not registered.");
    if (input_buffer_idx >= s.rf_count || output_buffer_idx >= s.rf_count) CUT_FAIL("Buffer index not registered.");
    auto& rf = s.rf;
    if (!rf.initialized()) CUT_FAIL("Session not initialized.");
    if (input_buffer_idx == output_buffer_idx) CUT_FAIL("Hilbert input and output must be different buffers.");
    Slot& in = s.rf_bufs[input_buffer_idx];
    Slot& out = s.rf_bufs[output_buffer_idx];
    void *pin = nullptr, *pout = nullptr;
    size_t bin = 0, bout = 0;
    if (!map_slot(in, &pin, &bin)) CUT_FAIL("Failed to map input OpenGL buffer.");
    if (!map_slot(out, &pout, &bout)) {
        unmap_slot(in);
        CUT_FAIL("Failed to map output OpenGL buffer.");
    }
    const size_t need = (size_t)rf.S() * rf.C() * rf.T() * sizeof(float2);
    bool ok = bin >= need && bout >= need;
    if (!ok) std::cerr << "cuda_toolkit: RF buffer too small for the configured dims" << std::endl;
    if (ok) {
        cut::HilbertInput hin{pin, 0, 2, nullptr, 0, (int)rf.C()};
        ok = rf.hilbert(hin, (float2*)pout, cut::stream());
    }
    if (ok) ok = hipStreamSynchronize(cut::stream()) == hipSuccess;
    unmap_slot(in);
    unmap_slot(out);
    return ok;
}

}  // extern "C"
