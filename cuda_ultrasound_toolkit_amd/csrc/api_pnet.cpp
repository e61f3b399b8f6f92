// Extension entry points of the displacement-proposal network and of learned
// 3-D block matching (no reference counterpart).  Device pointers in and out,
// asynchronous on the library stream; parameter errors return false with a
// message before any GPU work.
#include <algorithm>

#include "../../include/cut_ext.h"
#include "ncc.hpp"
#include "pipeline.hpp"
#include "pnet.hpp"

namespace {

struct PnetState {
    cut::PnetTrainer trainer;
    int input_mode = 1;
    cut::DeviceBuffer frags, pred, centre, ids, count, boxsum, boxtmp, pool;
};

PnetState& pnet_state()
{
    static PnetState* s = new PnetState();
    return *s;
}

bool check_volume(int nx, int ny, int nz, int stride)
{
    if (nx < 16 || ny < 16 || nz < 16 || stride <= 0) CUT_FAIL("pnet: volume smaller than a 16^3 block or bad stride");
    return true;
}

int grid_blocks(int nx, int ny, int nz, int stride)
{
    return ((nx - 16) / stride + 1) * ((ny - 16) / stride + 1) * ((nz - 16) / stride + 1);
}

}  // namespace

extern "C" int cut_pnet_param_count(void) { return cut::PN_PARAMS; }

extern "C" bool cut_pnet_set_precision(int mode)
{
    if (mode != 0 && mode != 1) CUT_FAIL("pnet: precision mode must be 0 (BF16 MFMA) or 1 (FP32)");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    pnet_state().trainer.set_fp32(mode == 1);
    return true;
}

extern "C" bool cut_pnet_set_input_mode(int mode)
{
    if (mode != 0 && mode != 1) CUT_FAIL("pnet: input mode must be 0 (block) or 1 (pooled search region)");
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    pnet_state().input_mode = mode;
    return true;
}

extern "C" bool cut_pnet_gather(const uint64_t* d_vols, int nx, int ny, int nz, int stride, const int* d_desc, int n,
                                float* d_out)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_vols || !d_desc || !d_out || n < 0) CUT_FAIL("pnet gather: null argument");
        if (!check_volume(nx, ny, nz, stride)) return false;
        return cut::pnet_gather(d_vols, nx, ny, nz, stride, d_desc, n, pnet_state().input_mode, d_out, cut::stream());
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_pnet_gather: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_pnet_forward(const float* d_params, const float* d_x, int B, float* d_y)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_params || !d_x || !d_y || B <= 0) CUT_FAIL("pnet forward: null argument or empty batch");
        return pnet_state().trainer.forward(d_params, d_x, B, d_y, cut::stream());
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_pnet_forward: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_pnet_loss_grad(const float* d_params, const float* d_x, const float* d_target, int B,
                                   float* d_grad, float* d_loss)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_params || !d_x || !d_target || !d_grad || B <= 0)
            CUT_FAIL("pnet loss_grad: null argument or empty batch");
        return pnet_state().trainer.loss_grad(d_params, d_x, d_target, B, d_grad, d_loss, cut::stream());
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_pnet_loss_grad: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_pnet_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, int n, float lr,
                              float beta1, float beta2, float eps, int step, float wd)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_params || !d_grad || !d_m || !d_v || n <= 0) CUT_FAIL("pnet adam: null argument");
        return cut::pnet_adam(d_params, d_grad, d_m, d_v, n, lr, beta1, beta2, eps, step, wd, cut::stream());
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_pnet_adam: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_pnet_infer(const float* d_params, const float* d_tpl, const float* d_src, int nx, int ny, int nz,
                               int stride, const int* d_ids, int n, float* d_pred, int* d_centre)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_params || !d_tpl || !d_src || !d_pred || n < 0) CUT_FAIL("pnet infer: null argument");
        if (!check_volume(nx, ny, nz, stride)) return false;
        if (!d_ids && n > grid_blocks(nx, ny, nz, stride)) CUT_FAIL("pnet infer: more blocks than the grid holds");
        auto& s = pnet_state();
        const size_t pb = 2 * cut::pnet_pool_floats(nx, ny, nz) * sizeof(float);
        if (!s.frags.ensure(cut::PN_FRAG_TOTAL * sizeof(uint16_t)) || !s.pool.ensure(std::max(pb, s.pool.bytes)))
            return false;
        hipStream_t st = cut::stream();
        if (!cut::pnet_pack(d_params, s.frags.as<uint16_t>(), st)) return false;
        return cut::pnet_infer(d_params, s.frags.as<uint16_t>(), d_tpl, d_src, nx, ny, nz, stride, d_ids, n,
                               s.input_mode, s.pool.as<float>(), d_pred, d_centre, st);
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_pnet_infer: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_motion3d_learned(const float* d_params, const float* d_tpl, const float* d_src, int nx, int ny,
                                     int nz, int stride, int radius, float threshold, int boundary_check, int* d_disp,
                                     float* d_peak, int* d_path, float* d_pred, int* h_fallbacks)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!d_params || !d_tpl || !d_src || !d_disp || !d_peak) CUT_FAIL("motion3d learned: null argument");
        if (!check_volume(nx, ny, nz, stride)) return false;
        if (radius < 0 || radius > 3) CUT_FAIL("motion3d learned: window radius must be 0..3");
        const int nb = grid_blocks(nx, ny, nz, stride);
        auto& s = pnet_state();
        const size_t wb = cut::ncc3d_boxsum_doubles(nx, ny, nz) * sizeof(double);
        const size_t tb = cut::ncc3d_boxsum_tmp_doubles(nx, ny, nz) * sizeof(double);
        if (!s.frags.ensure(cut::PN_FRAG_TOTAL * sizeof(uint16_t)) || !s.pred.ensure((size_t)nb * 3 * sizeof(float)) ||
            !s.centre.ensure((size_t)nb * 3 * sizeof(int)) || !s.ids.ensure((size_t)nb * sizeof(int)) ||
            !s.count.ensure(sizeof(int)) || !s.boxsum.ensure(std::max(wb, s.boxsum.bytes)) ||
            !s.boxtmp.ensure(std::max(tb, s.boxtmp.bytes)) ||
            !s.pool.ensure(std::max(2 * cut::pnet_pool_floats(nx, ny, nz) * sizeof(float), s.pool.bytes)))
            return false;
        hipStream_t st = cut::stream();
        float* pred = d_pred ? d_pred : s.pred.as<float>();
        int* centre = s.centre.as<int>();
        const double* W = s.boxsum.as<double>();
        if (!cut::pnet_pack(d_params, s.frags.as<uint16_t>(), st) ||
            !cut::ncc3d_boxsums(d_src, nx, ny, nz, s.boxsum.as<double>(), s.boxtmp.as<double>(), st) ||
            !cut::pnet_infer(d_params, s.frags.as<uint16_t>(), d_tpl, d_src, nx, ny, nz, stride, nullptr, nb,
                             s.input_mode, s.pool.as<float>(), pred, centre, st) ||
            !cut::launch_ncc3d(d_tpl, d_src, W, nx, ny, nz, stride, nullptr, nb, nullptr, false, centre, radius, d_disp,
                               d_peak, st) ||
            !cut::launch_ncc3d_fallback(d_disp, d_peak, centre, nx, ny, nz, stride, nb, radius, threshold,
                                        boundary_check != 0, s.ids.as<int>(), s.count.as<int>(), d_path, st) ||
            !cut::launch_ncc3d(d_tpl, d_src, W, nx, ny, nz, stride, s.ids.as<int>(), nb, s.count.as<int>(), true,
                               nullptr, 0, d_disp, d_peak, st))
            return false;
        if (h_fallbacks) {
            CUT_HIP_CHECK(hipMemcpyAsync(h_fallbacks, s.count.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
            CUT_HIP_CHECK(hipStreamSynchronize(st));
        }
        return true;
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_motion3d_learned: exception" << std::endl;
        return false;
    }
}
