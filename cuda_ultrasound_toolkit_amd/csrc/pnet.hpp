// Displacement-proposal network: layer table, parameter packing and launchers.
#pragma once

#include "common.hpp"

namespace cut {

// Conv layers: input channels, output channels; FP32 parameter offsets
// (w1 b1 w2 b2 w3 b3 w4 b4 wf bf, conv weights [co][ci][kz][ky][kx]).
__host__ __device__ constexpr int pn_ci(int l) { return l == 0 ? 2 : (l == 1 ? 16 : 32); }
__host__ __device__ constexpr int pn_co(int l) { return l == 0 ? 16 : 32; }
__host__ __device__ constexpr int pn_w_off(int l)
{
    return l == 0 ? 0 : pn_w_off(l - 1) + pn_co(l - 1) * pn_ci(l - 1) * 27 + pn_co(l - 1);
}
__host__ __device__ constexpr int pn_b_off(int l) { return pn_w_off(l) + pn_co(l) * pn_ci(l) * 27; }
constexpr int PN_WF_OFF = pn_w_off(4);
constexpr int PN_BF_OFF = PN_WF_OFF + 3 * 256;
constexpr int PN_PARAMS = PN_BF_OFF + 3;
static_assert(PN_WF_OFF == 70096 && PN_PARAMS == 70867, "parameter packing");

// MFMA weight fragments of the fused inference kernel (BF16): per layer
// k-steps of 32 x output-channel tiles of 16 x 64 lanes x 8.
__host__ __device__ constexpr int pn_steps(int l) { return (pn_ci(l) * 27 + 31) / 32; }
__host__ __device__ constexpr int pn_co_tiles(int l) { return pn_co(l) / 16; }
__host__ __device__ constexpr int pn_frag_size(int l) { return pn_steps(l) * pn_co_tiles(l) * 64 * 8; }
__host__ __device__ constexpr int pn_frag_off(int l) { return l == 0 ? 0 : pn_frag_off(l - 1) + pn_frag_size(l - 1); }
constexpr int PN_FRAG_TOTAL = pn_frag_off(4);

// rows of the weight-gradient product handled per split-K workgroup
constexpr int PN_SPLIT_ROWS = 2048;

// Batched training path (forward, loss and gradient); owns its workspace.
class PnetTrainer {
  public:
    bool forward(const float* d_params, const float* d_x, int B, float* d_y, hipStream_t st);
    bool loss_grad(const float* d_params, const float* d_x, const float* d_t, int B, float* d_grad, float* d_loss,
                   hipStream_t st);
    // false: BF16 MFMA products (default); true: FP32 operands on the vector
    // ALUs, for verifying the layer structure against an FP64 reference
    void set_fp32(bool on) { fp32_ = on; }

  private:
    bool reserve(int B);
    DeviceBuffer ws_;
    int B_ = 0;
    bool fp32_ = false;
    void* col_[4] = {};
    float* act_[4] = {};
    float* g_[4] = {};
    float* dcol_ = nullptr;
    float* part_ = nullptr;
    float* y_ = nullptr;
    float* dy_ = nullptr;
    uint16_t* pbf_ = nullptr;
};

bool pnet_adam(float* d_p, const float* d_g, float* d_m, float* d_v, int n, float lr, float b1, float b2, float eps,
               int step, float wd, hipStream_t st);
bool pnet_gather(const uint64_t* d_vols, int nx, int ny, int nz, int stride, const int* d_desc, int n, int mode,
                 float* d_out, hipStream_t st);
bool pnet_pack(const float* d_params, uint16_t* d_frags, hipStream_t st);
// Pooled volumes of input mode 1: 2 * pnet_pool_floats floats of scratch (d_pool may be
// null: the blocks are then pooled on the fly).
size_t pnet_pool_floats(int nx, int ny, int nz);
bool pnet_infer(const float* d_params, const uint16_t* d_frags, const float* d_tpl, const float* d_src, int nx, int ny,
                int nz, int stride, const int* d_ids, int n, int mode, float* d_pool, float* d_pred, int* d_centre,
                hipStream_t st);

}  // namespace cut
