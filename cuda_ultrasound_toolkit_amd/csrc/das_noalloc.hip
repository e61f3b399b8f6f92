// DAS entry point that allocates nothing: cut_das_device_noalloc.
//
// The same launch as cut_das_device (launch_das, same kernels), with the
// workspace taken from a fixed device array instead of a cached hipMalloc:
// the HERCULES transmit-split partial sums (float2, or float4 with the
// coherence factor, per voxel and split) and the RCA_TPW angle table.  The
// split count is limited to what the array holds (the largest power of two
// not above the tuned or default choice); a volume too large for two splits
// runs with the transmit loop unsplit, which needs no workspace.  No launch
// shape is tuned here (tuning creates HIP events); a shape tuned earlier by
// another entry point for the same key is used.
//
// Every call on a device shares that device's array; nothing in it needs to
// start at zero (each split's workgroups overwrite every voxel of their
// tiles, the angle table is rewritten on every call).  Calls that overlap on
// one device (work enqueued on two streams at once) race on the array.
#include "../../include/cut_ext.h"
#include "pipeline.hpp"

#include <exception>
#include <mutex>

constexpr size_t DAS_SCRATCH_BYTES = size_t(128) << 20;  // 128 MiB per device

static __device__ __align__(256) unsigned char scratch[DAS_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).  The array's name is
// a host-side stand-in and is never passed to the device.
void* scratch_ptr()
{
    static std::mutex mu;
    static void* ptr[MAX_DEVICES] = {};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        std::cerr << "cuda_toolkit: hipGetDevice failed: " << hipGetErrorString(e) << std::endl;
        return nullptr;
    }
    if (dev < 0 || dev >= MAX_DEVICES) {
        std::cerr << "cuda_toolkit: device " << dev << " beyond the " << MAX_DEVICES << " supported" << std::endl;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::cerr << "cuda_toolkit: hipGetSymbolAddress(scratch) failed: " << hipGetErrorString(e) << std::endl;
            return nullptr;
        }
        ptr[dev] = p;
    }
    return ptr[dev];
}

}  // namespace

extern "C" bool cut_das_device_noalloc(const void* d_iq, void* d_volume, const CudaBeamformerParameters* bp)
{
    try {
        if (!d_iq || !d_volume || !bp) CUT_FAIL("null argument");
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        cut::DasParams p;
        if (!cut::das_params_from_bp(*bp, p)) return false;
        // the only workspace a call cannot do without: the RCA_TPW angle table
        const size_t need = p.sequence == RCA_TPW ? (size_t)p.T * sizeof(double2) : 0;
        if (need > DAS_SCRATCH_BYTES)
            CUT_FAIL("cut_das_device_noalloc: " << need << " bytes of workspace needed, the fixed array holds "
                     << DAS_SCRATCH_BYTES);
        void* ws = scratch_ptr();
        if (!ws) return false;
        cut::DasOptions o;
        o.cubic = cut::options().das_cubic;
        o.cf_hercules = cut::options().das_cf_hercules;
        return cut::launch_das(d_iq, cut::IQ_F32, 1.0f, p, o, (float2*)d_volume, ws, DAS_SCRATCH_BYTES, cut::stream(),
                               false);
    } catch (const std::exception& e) {
        std::cerr << "cuda_toolkit: cut_das_device_noalloc: " << e.what() << std::endl;
    }
    return false;
}
