// Block-matching launchers and their parameter blocks.
#pragma once

#include "common.hpp"

namespace cut {

// Largest ZNCC map (valid offsets per grid point) one workgroup holds in LDS.
constexpr int NCC2D_MAP_MAX = 8192;

// NccMotionParameters in kernel form (cuda_beamformer_parameters.h:91-107).
struct NccDev {
    int W, H, frames;
    int gx, gy;
    int patch, spacing, mx, my;
    float abs_thr, rel_thr, min_var;
    int subpixel;
    int ref, neighbour;
};

bool validate_motion(const NccMotionParameters& p, size_t image_bytes, size_t map_bytes, NccDev& d);
bool launch_block_match2d(const float* d_images, const NccDev& p, float4* d_maps, hipStream_t st);
bool launch_corr_images(const double* d_tpl_c, double tpl_energy, int tw, int th, const float* d_src, int sw, int sh,
                        float* d_out, hipStream_t st);

}  // namespace cut
