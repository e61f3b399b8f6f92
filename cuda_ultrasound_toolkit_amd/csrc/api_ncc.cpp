// cuda_toolkit::motion_detection / corr_images and their MATLAB shims.
//
// Replaces cuda_toolkit/src/api/cuda_toolkit.cpp:201-330 and
// cuda_toolkit_matlab.cpp:39-75.  Host spans in, host spans out, synchronous;
// parameter errors return false with a message before any GPU work.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/cuda_toolkit.hpp"
#include "../../include/cuda_toolkit_matlab.h"
#include "../../include/cut_ext.h"
#include "ncc.hpp"
#include "pipeline.hpp"

namespace {

struct NccState {
    cut::DeviceBuffer images, maps, tpl, src, corr, i1, i2;
};

NccState& ncc_state()
{
    static NccState* s = new NccState();
    return *s;
}

}  // namespace

namespace cut {

// Host-side checks of motion_detection (image_processor.cpp:10-56 leaves these unchecked).
bool validate_motion(const NccMotionParameters& p, size_t image_bytes, size_t map_bytes, NccDev& d)
{
    d.W = (int)p.image_dims[0];
    d.H = (int)p.image_dims[1];
    d.frames = (int)p.frame_count;
    d.gx = (int)p.motion_grid_dims[0];
    d.gy = (int)p.motion_grid_dims[1];
    d.patch = (int)p.patch_size;
    d.spacing = (int)p.motion_grid_spacing;
    d.mx = (int)p.search_margins[0];
    d.my = (int)p.search_margins[1];
    d.abs_thr = p.abs_cor_threshold;
    d.rel_thr = p.rel_cor_threshold;
    d.min_var = p.min_patch_variance;
    d.subpixel = p.use_subpixel ? 1 : 0;
    d.ref = (int)p.reference_frame;
    d.neighbour = p.neighbour_compare ? 1 : 0;
    if (d.W <= 0 || d.H <= 0 || d.frames <= 0) CUT_FAIL("motion_detection: empty images");
    if (d.gx <= 0 || d.gy <= 0 || d.patch <= 0) CUT_FAIL("motion_detection: empty motion grid or patch");
    if ((uint)d.ref >= (uint)d.frames) CUT_FAIL("motion_detection: reference frame out of range");
    if (image_bytes < (size_t)d.W * d.H * d.frames * sizeof(float))
        CUT_FAIL("motion_detection: image span holds " << image_bytes << " bytes, "
                 << (size_t)d.W * d.H * d.frames * 4 << " needed");
    if (map_bytes < (size_t)d.gx * d.gy * d.frames * sizeof(float4))
        CUT_FAIL("motion_detection: motion map span holds " << map_bytes << " bytes, "
                 << (size_t)d.gx * d.gy * d.frames * 16 << " needed");
    if (d.mx < 0 || d.my < 0) CUT_FAIL("motion_detection: negative search margins");
    const long long vw = std::min(2LL * d.mx + 1, (long long)d.W), vh = std::min(2LL * d.my + 1, (long long)d.H);
    if (vw * vh > 0x7fffffffLL) CUT_FAIL("motion_detection: a search of " << vw << " x " << vh << " offsets is too large");
    return true;
}

}  // namespace cut

bool cuda_toolkit::motion_detection(std::span<const uint8_t> images, std::span<uint8_t> motion_maps,
                                    const NccMotionParameters& params)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        cut::NccDev d;
        if (!cut::validate_motion(params, images.size(), motion_maps.size(), d)) return false;
        auto& s = ncc_state();
        const size_t ib = (size_t)d.W * d.H * d.frames * sizeof(float);
        const size_t mb = (size_t)d.gx * d.gy * d.frames * sizeof(float4);
        if (!s.images.ensure(ib) || !s.maps.ensure(mb)) return false;
        hipStream_t st = cut::stream();
        CUT_HIP_CHECK(hipMemcpyAsync(s.images.ptr, images.data(), ib, hipMemcpyHostToDevice, st));
        if (!cut::launch_block_match2d(s.images.as<float>(), d, s.maps.as<float4>(), st)) return false;
        CUT_HIP_CHECK(hipMemcpyAsync(motion_maps.data(), s.maps.ptr, mb, hipMemcpyDeviceToHost, st));
        CUT_HIP_CHECK(hipStreamSynchronize(st));
        return true;
    } catch (...) {
        std::cerr << "cuda_toolkit: motion_detection: exception" << std::endl;
        return false;
    }
}

bool cuda_toolkit::corr_images(std::span<const float> template_image, std::span<const float> source_image,
                               std::span<float> corr_map, const uint template_dims[2], const uint source_dims[2])
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!template_dims || !source_dims) CUT_FAIL("corr_images: null dims");
        const uint tw = template_dims[0], th = template_dims[1], sw = source_dims[0], sh = source_dims[1];
        if (tw == 0 || th == 0 || sw == 0 || sh == 0) CUT_FAIL("Invalid image dimensions for corr_images.");
        if (tw > sw || th > sh) CUT_FAIL("Template dimensions must be <= source dimensions for corr_images.");
        const size_t vw = sw - tw + 1, vh = sh - th + 1;
        if (template_image.size() < (size_t)tw * th || source_image.size() < (size_t)sw * sh)
            CUT_FAIL("Input image spans are smaller than the provided dimensions for corr_images.");
        if (corr_map.size() < vw * vh) CUT_FAIL("Output corr_map span is too small for corr_images.");
        // template minus its mean, and its energy, in FP64 on the host
        const size_t n = (size_t)tw * th;
        double mean = 0.0;
        for (size_t i = 0; i < n; ++i) mean += template_image[i];
        mean /= (double)n;
        std::vector<double> tc(n);
        double energy = 0.0;
        for (size_t i = 0; i < n; ++i) {
            tc[i] = template_image[i] - mean;
            energy += tc[i] * tc[i];
        }
        auto& s = ncc_state();
        if (!s.tpl.ensure(n * sizeof(double)) || !s.src.ensure((size_t)sw * sh * sizeof(float)) ||
            !s.corr.ensure(vw * vh * sizeof(float)))
            return false;
        hipStream_t st = cut::stream();
        CUT_HIP_CHECK(hipMemcpyAsync(s.tpl.ptr, tc.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        CUT_HIP_CHECK(hipMemcpyAsync(s.src.ptr, source_image.data(), (size_t)sw * sh * sizeof(float),
                                     hipMemcpyHostToDevice, st));
        if (!cut::launch_corr_images(s.tpl.as<double>(), energy, (int)tw, (int)th, s.src.as<float>(), (int)sw,
                                     (int)sh, s.corr.as<float>(), st))
            return false;
        CUT_HIP_CHECK(hipMemcpyAsync(corr_map.data(), s.corr.ptr, vw * vh * sizeof(float), hipMemcpyDeviceToHost, st));
        CUT_HIP_CHECK(hipStreamSynchronize(st));
        return true;
    } catch (...) {
        std::cerr << "cuda_toolkit: corr_images: exception" << std::endl;
        return false;
    }
}

extern "C" void motion_detect_f32_matlab(const float* images, NCCMotionParameters params, float* motion_maps)
{
    const size_t in = (size_t)params.image_dims[0] * params.image_dims[1] * params.frame_count * sizeof(float);
    const size_t out =
        (size_t)params.motion_grid_dims[0] * params.motion_grid_dims[1] * 4 * sizeof(float) * params.frame_count;
    if (!cuda_toolkit::motion_detection({reinterpret_cast<const uint8_t*>(images), in},
                                        {reinterpret_cast<uint8_t*>(motion_maps), out}, params))
        std::cerr << "Error: Motion detection failed." << std::endl;
}

extern "C" void corr_images_f32_matlab(const float* template_image, const float* source_image,
                                       CorrImagesParameters params, float* corr_map)
{
    const size_t tn = (size_t)params.template_dims[0] * params.template_dims[1];
    const size_t sn = (size_t)params.source_dims[0] * params.source_dims[1];
    const size_t cw = params.source_dims[0] >= params.template_dims[0] ? params.source_dims[0] - params.template_dims[0] + 1 : 0;
    const size_t ch = params.source_dims[1] >= params.template_dims[1] ? params.source_dims[1] - params.template_dims[1] + 1 : 0;
    if (!cuda_toolkit::corr_images({template_image, tn}, {source_image, sn}, {corr_map, cw * ch},
                                   params.template_dims, params.source_dims))
        std::cerr << "Error: Correlation failed." << std::endl;
}

// ---------------------------------------------------------------------------
// 3-D block matching (extension entry points, device pointers, async)
// ---------------------------------------------------------------------------
extern "C" bool cut_ncc3d_search(const float* d_tpl, const float* d_src, int nx, int ny, int nz, int stride,
                                 const int* d_block_ids, int n_blocks, const int* d_centre, int radius, int* d_disp,
                                 float* d_peak)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_tpl || !d_src || !d_disp || !d_peak) CUT_FAIL("ncc3d: null argument");
        if (nx < 16 || ny < 16 || nz < 16 || stride <= 0) CUT_FAIL("ncc3d: volume smaller than a 16^3 block");
        const long long total = (long long)((nx - 16) / stride + 1) * ((ny - 16) / stride + 1) * ((nz - 16) / stride + 1);
        if (!d_block_ids && n_blocks != total) CUT_FAIL("ncc3d: " << n_blocks << " blocks given, the grid has " << total);
        if (n_blocks < 0 || n_blocks > total) CUT_FAIL("ncc3d: bad block count");
        if (d_centre && (radius < 0 || radius > 3)) CUT_FAIL("ncc3d: window radius must be 0..3");
        auto& s = ncc_state();
        const size_t wb = cut::ncc3d_boxsum_doubles(nx, ny, nz) * sizeof(double);
        const size_t tb = cut::ncc3d_boxsum_tmp_doubles(nx, ny, nz) * sizeof(double);
        if (!s.i1.ensure(std::max(wb, s.i1.bytes)) || !s.i2.ensure(std::max(tb, s.i2.bytes))) return false;
        hipStream_t st = cut::stream();
        if (!cut::ncc3d_boxsums(d_src, nx, ny, nz, s.i1.as<double>(), s.i2.as<double>(), st)) return false;
        return cut::launch_ncc3d(d_tpl, d_src, s.i1.as<double>(), nx, ny, nz, stride, d_block_ids, n_blocks, nullptr,
                                 false, d_centre, radius, d_disp, d_peak, st);
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_ncc3d_search: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_envelope(const void* d_iq, size_t count, float* d_out)
{
    if (!d_iq || !d_out) CUT_FAIL("envelope: null argument");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    return cut::launch_envelope((const float2*)d_iq, count, d_out, cut::stream());
}
