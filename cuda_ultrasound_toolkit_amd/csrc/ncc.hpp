// Block-matching launchers and their parameter blocks.
#pragma once

#include "common.hpp"

namespace cut {

// Longest ZNCC map prefix (valid offsets per grid point) a workgroup holds in
// LDS; entries beyond it that the peak search reads are evaluated on demand.
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

// 3-D search (16^3 blocks, |d| <= 8).  FP64 box sums of the source first
// (ncc3d_boxsums), then the full search (centre == nullptr) or the windowed
// search |d - centre| <= radius over `nblocks` blocks (ids == nullptr: blocks
// 0..nblocks-1).  count (full search only, may be null): a device block count
// that bounds the grid.  by_id: results (and window centres) are indexed by
// block id rather than by list position.
size_t ncc3d_boxsum_doubles(int nx, int ny, int nz);
size_t ncc3d_boxsum_tmp_doubles(int nx, int ny, int nz);
bool ncc3d_boxsums(const float* d_src, int nx, int ny, int nz, double* d_out, double* d_tmp, hipStream_t st);
bool launch_ncc3d(const float* d_tpl, const float* d_src, const double* d_w, int nx, int ny, int nz, int stride,
                  const int* d_ids, int nblocks, const int* d_count, bool by_id, const int* d_centre, int radius,
                  int* d_disp, float* d_peak, hipStream_t st);
// Select the blocks that need the full search after a windowed search (see
// ncc3d_fallback_kernel); *d_count receives their number, d_ids their ids.
bool launch_ncc3d_fallback(const int* d_disp, const float* d_peak, const int* d_centre, int nx, int ny, int nz,
                           int stride, int nb, int radius, float thr, bool boundary, int* d_ids, int* d_count,
                           int* d_path, hipStream_t st);
bool launch_envelope(const float2* d_iq, size_t n, float* d_out, hipStream_t st);

}  // namespace cut
