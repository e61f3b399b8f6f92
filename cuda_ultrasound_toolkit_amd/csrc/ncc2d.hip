// 2-D NCC block matching and full-image ZNCC.
//
// motion_detection: cuda_toolkit/src/image_processing/image_processor.cpp:10-56,132-229
//   with block_match_kernels.cu:7-240 (NPP valid-region ZNCC, CUB top-32 sort,
//   3x3 paraboloid peak test).  The reference launches three kernels per grid
//   point from a host loop; here one workgroup owns one (frame, grid point):
//   it computes the ZNCC map entries the search reads into LDS (FP64 window
//   sums and energy normalisation), selects the 32 largest of the first 2048 map entries in the
//   sort's order (value descending, index ascending on ties, missing entries
//   -1), and runs the peak test on lanes 0..31 of the first wave with the same
//   32-lane shuffle-down max tree, in FP32 like the reference.
// corr_images: image_processor.cpp:58-130, one thread per map entry.
#include "common.hpp"
#include "ncc.hpp"

namespace cut {

constexpr int BM_THREADS = 256;
constexpr int BM_CAND = 32;
constexpr int BM_SORT = 2048;

// 3x3 least-squares quadratic fit (block_match_kernels.cuh:130-154): rows give
// a (x^2), b (y^2), c (xy), d (x), e (y), f (1).  The entries are the
// pseudo-inverse of the 9 x 6 design matrix on offsets {-1, 0, 1}^2.
__constant__ float kP3[6][9] = {
    {1.f / 6, -1.f / 3, 1.f / 6, 1.f / 6, -1.f / 3, 1.f / 6, 1.f / 6, -1.f / 3, 1.f / 6},
    {1.f / 6, 1.f / 6, 1.f / 6, -1.f / 3, -1.f / 3, -1.f / 3, 1.f / 6, 1.f / 6, 1.f / 6},
    {0.25f, 0.f, -0.25f, 0.f, 0.f, 0.f, -0.25f, 0.f, 0.25f},
    {-1.f / 6, 0.f, 1.f / 6, -1.f / 6, 0.f, 1.f / 6, -1.f / 6, 0.f, 1.f / 6},
    {-1.f / 6, -1.f / 6, -1.f / 6, 0.f, 0.f, 0.f, 1.f / 6, 1.f / 6, 1.f / 6},
    {-1.f / 9, 2.f / 9, -1.f / 9, 2.f / 9, 5.f / 9, 2.f / 9, -1.f / 9, 2.f / 9, -1.f / 9},
};

struct Roi {
    int tl, tt, tw, th, sl, st, sw, sh;
};

__device__ __forceinline__ Roi block_roi(const NccDev& p, int gi, int gj)
{
    const int half = (int)(p.patch / 2u);
    int tl = gj * p.spacing - half, tt = gi * p.spacing - half;
    int tr = tl + p.patch - 1, tb = tt + p.patch - 1;
    int sl = tl - p.mx, sr = tr + p.mx, st = tt - p.my, sb = tb + p.my;
    tl = max(tl, 0); tr = min(tr, p.W - 1); sl = max(sl, 0); sr = min(sr, p.W - 1);
    tt = max(tt, 0); tb = min(tb, p.H - 1); st = max(st, 0); sb = min(sb, p.H - 1);
    return {tl, tt, tr - tl + 1, tb - tt + 1, sl, st, sr - sl + 1, sb - st + 1};
}

__device__ __forceinline__ double block_sum_d(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < BM_THREADS / 64; ++w) s += red[w];
    return s;
}

__global__ __launch_bounds__(BM_THREADS) void block_match2d_kernel(const float* __restrict__ images, const NccDev p,
                                                                  float4* __restrict__ maps)
{
    __shared__ float map[NCC2D_MAP_MAX];
    __shared__ unsigned char taken[BM_SORT];
    __shared__ double red[BM_THREADS / 64];
    __shared__ float best_v[BM_THREADS / 64];
    __shared__ int best_i[BM_THREADS / 64];
    __shared__ float cand_v[BM_CAND];
    __shared__ int cand_i[BM_CAND];

    const int tid = threadIdx.x;
    const int gpt = blockIdx.x;
    const int f = blockIdx.y;
    const int gi = gpt / p.gx, gj = gpt - gi * p.gx;
    float4* out = maps + ((size_t)f * p.gy + gi) * p.gx + gj;
    int ref = p.ref;
    if (f == ref) {
        if (tid == 0) *out = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    if (p.neighbour) ref = ref > f ? f + 1 : f - 1;
    const Roi r = block_roi(p, gi, gj);
    if (r.tw <= 0 || r.th <= 0 || r.sw < r.tw || r.sh < r.th) {  // NPP rejects the ROI: entry stays 0
        if (tid == 0) *out = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const int vw = r.sw - r.tw + 1, vh = r.sh - r.th + 1, total = vw * vh;
    const float* tpl = images + (size_t)ref * p.W * p.H + (size_t)r.tt * p.W + r.tl;
    const float* src = images + (size_t)f * p.W * p.H + (size_t)r.st * p.W + r.sl;
    const int n = r.tw * r.th;

    // template mean and energy (FP64)
    double ts = 0.0;
    for (int k = tid; k < n; k += BM_THREADS) ts += tpl[(size_t)(k / r.tw) * p.W + k % r.tw];
    const double tmean = block_sum_d(ts, red) / n;
    double te = 0.0;
    for (int k = tid; k < n; k += BM_THREADS) {
        const double d = tpl[(size_t)(k / r.tw) * p.W + k % r.tw] - tmean;
        te += d * d;
    }
    te = block_sum_d(te, red);

    // ZNCC map entry s (offset (s % vw, s / vw)), the same arithmetic wherever
    // it is evaluated
    auto entry = [&](int s) -> float {
        const int u = s % vw, v = s / vw;
        double ss = 0.0, s2 = 0.0, ct = 0.0;
        for (int y = 0; y < r.th; ++y) {
            const float* srow = src + (size_t)(v + y) * p.W + u;
            const float* trow = tpl + (size_t)y * p.W;
            for (int x = 0; x < r.tw; ++x) {
                const double sv = srow[x];
                ss += sv;
                s2 = fma(sv, sv, s2);
                ct = fma(sv, (double)trow[x] - tmean, ct);
            }
        }
        const double se = s2 - ss * ss / n;
        const double den = sqrt(te * (se > 0.0 ? se : 0.0));
        return den > 0.0 ? (float)(ct / den) : 0.0f;
    };
    // The peak search reads the first 2048 entries, the 3 x 3 neighbourhoods
    // of its candidates (at most one map row further) and the no-shift entry:
    // the map prefix [0, P) is held in LDS, any other entry that is read is
    // evaluated where it is needed, so maps of any size are searched.
    const int P = min(total, min(NCC2D_MAP_MAX, BM_SORT + vw + 1));
    for (int s = tid; s < P; s += BM_THREADS) map[s] = entry(s);
    for (int i = tid; i < BM_SORT; i += BM_THREADS) taken[i] = 0;
    __syncthreads();
    auto map_at = [&](int s) { return s < P ? map[s] : entry(s); };

    // top 32 of the first 2048 entries (missing ones -1), value desc, index asc
    for (int c = 0; c < BM_CAND; ++c) {
        float bv = -3.0e38f;
        int bi = 0x7fffffff;
        for (int i = tid * (BM_SORT / BM_THREADS); i < (tid + 1) * (BM_SORT / BM_THREADS); ++i) {
            if (taken[i]) continue;
            const float v = i < total ? map[i] : -1.0f;
            if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_down(bv, o, 64);
            const int oi = __shfl_down(bi, o, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if ((tid & 63) == 0) { best_v[tid >> 6] = bv; best_i[tid >> 6] = bi; }
        __syncthreads();
        if (tid == 0) {
            float v0 = best_v[0];
            int i0 = best_i[0];
            for (int w = 1; w < BM_THREADS / 64; ++w)
                if (best_v[w] > v0 || (best_v[w] == v0 && best_i[w] < i0)) { v0 = best_v[w]; i0 = best_i[w]; }
            cand_v[c] = v0;
            cand_i[c] = i0;
            taken[i0] = 1;
        }
        __syncthreads();
    }

    // peak test (block_match_kernels.cu:67-185) on lanes 0..31
    if (tid >= BM_CAND) return;
    const int nsx = r.tl - r.sl, nsy = r.tt - r.st;
    const int px = cand_i[tid] % vw, py = cand_i[tid] / vw;
    float vals[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int yy = py + k / 3 - 1, xx = px + k % 3 - 1;
        vals[k] = (yy < 0 || yy >= vh || xx < 0 || xx >= vw) ? 0.0f : map_at(yy * vw + xx);
    }
    float co[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 9; ++j) s += kP3[i][j] * vals[j];
        co[i] = s;
    }
    const float root = sqrtf(co[0] * co[0] + co[1] * co[1] + co[2] * co[2] - 2 * co[0] * co[1]);
    const float s0 = co[0] + co[1] + root, s1 = co[0] + co[1] - root;
    const float min_sharp = fminf(fabsf(s0), fabsf(s1));
    const float den = co[2] * co[2] - 4 * co[1] * co[0];
    float sx = 0.0f, sy = 0.0f;
    if (den != 0.0f) {
        sx = (2 * co[1] * co[3] - co[2] * co[4]) / den;
        sy = (2 * co[0] * co[4] - co[2] * co[3]) / den;
    }
    const float use = p.subpixel ? 1.0f : 0.0f;
    sx = fminf(fmaxf(sx, -1.0f), 1.0f) * use;
    sy = fminf(fmaxf(sy, -1.0f), 1.0f) * use;
    float ox = sx + (float)px - (float)nsx, oy = sy + (float)py - (float)nsy;
    const float no_shift_peak = map_at(nsy * vw + nsx);
    const float thr = no_shift_peak * (1.0f + p.rel_thr);
    float peak = cand_v[tid];
    if (min_sharp < p.min_var || s0 >= 0.0f || peak > 1.0f || peak < thr) peak = -2.0f;
    for (int o = 16; o > 0; o >>= 1) {
        const float v2 = __shfl_down(peak, o, 32);
        const float x2 = __shfl_down(ox, o, 32);
        const float y2 = __shfl_down(oy, o, 32);
        if (v2 > peak) { peak = v2; ox = x2; oy = y2; }
    }
    if (tid == 0) {
        if (peak < p.abs_thr || peak <= no_shift_peak) {
            ox = oy = 0.0f;
            if (peak < thr) peak = -1.0f;
        }
        *out = make_float4(ox, oy, peak, no_shift_peak);
    }
}

bool launch_block_match2d(const float* d_images, const NccDev& p, float4* d_maps, hipStream_t st)
{
    dim3 grid(p.gx * p.gy, p.frames);
    block_match2d_kernel<<<grid, BM_THREADS, 0, st>>>(d_images, p, d_maps);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

// Full-image ZNCC: one thread per valid offset; tpl_c holds template - mean (FP64).
__global__ __launch_bounds__(256) void corr_images_kernel(const double* __restrict__ tpl_c, double tpl_energy, int tw,
                                                          int th, const float* __restrict__ src, int sw, int vw,
                                                          int vh, float* __restrict__ out)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (long long)vw * vh) return;
    const int u = (int)(s % vw), v = (int)(s / vw);
    const int n = tw * th;
    double ss = 0.0, s2 = 0.0, ct = 0.0;
    for (int y = 0; y < th; ++y) {
        const float* srow = src + (size_t)(v + y) * sw + u;
        const double* trow = tpl_c + (size_t)y * tw;
        for (int x = 0; x < tw; ++x) {
            const double sv = srow[x];
            ss += sv;
            s2 = fma(sv, sv, s2);
            ct = fma(sv, trow[x], ct);
        }
    }
    const double se = s2 - ss * ss / n;
    const double den = sqrt(tpl_energy * (se > 0.0 ? se : 0.0));
    out[s] = den > 0.0 ? (float)(ct / den) : 0.0f;
}

bool launch_corr_images(const double* d_tpl_c, double tpl_energy, int tw, int th, const float* d_src, int sw, int sh,
                        float* d_out, hipStream_t st)
{
    const int vw = sw - tw + 1, vh = sh - th + 1;
    const long long n = (long long)vw * vh;
    corr_images_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(d_tpl_c, tpl_energy, tw, th, d_src, sw, vw, vh, d_out);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

}  // namespace cut
