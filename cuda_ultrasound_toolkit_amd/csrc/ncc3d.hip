// 3-D ZNCC block matching of envelope volumes: full search and windowed
// (refinement) search.
//
// The reference matches 2-D images (image_processor.cpp:132-229); this is the
// same zero-mean normalised cross-correlation on 16^3 blocks of a volume.  A
// template block at (bx, by, bz) is compared with the source block at
// (bx+dx, by+dy, bz+dz) for every |d| <= margin that keeps the source block
// inside the volume; the best shift (first maximum in (dz, dy, dx) raster
// order) and its ZNCC are returned.
//
// Window sums of S and S^2 over every 16^3 box of the source come from FP64
// box-sum volumes, built once per source volume by three separable passes;
// the energy normalisation is done in FP64.
//
// Full search: one workgroup per block, one thread per (dz, dy) shift pair
// holding the 17 dx correlations.  The 32^3 source region is streamed through
// LDS one z-plane at a time (double-buffered); at source plane Z the thread
// correlates template plane Z - dz with it, each 16-sample template row
// against the 32-sample source row it meets (16 x 17 FMAs per row pair).
//
// Windowed search (|d - centre| <= radius): the (16 + 2r)^3 source region
// sits in LDS; threads own a (dz, dy) pair and a subset of template planes
// and hold the 2r+1 dx correlations; partial sums are added in a fixed order.
#include "common.hpp"
#include "ncc.hpp"

#include <cmath>

namespace cut {

constexpr int NB = 16;           // block edge
constexpr int NM = 8;            // search margin
constexpr int NS = 2 * NM + 1;   // shifts per axis
constexpr int NROW = NB + 2 * NM;
constexpr int NF_THREADS = 320;  // 17 dz x 17 dy = 289 used
constexpr int RS = 36;           // LDS row stride (floats) of a staged source plane
constexpr int TPS = 260;         // LDS plane stride (floats) of the template

// ---------------------------------------------------------------------------
// box sums: W[z][y][x] = sum over [z, z+16) x [y, y+16) x [x, x+16)
// ---------------------------------------------------------------------------
__global__ void boxsum_x_kernel(const float* __restrict__ src, int nx, int ny, int nz, double* __restrict__ a1,
                                double* __restrict__ a2)
{
    const int wx = nx - NB + 1;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nz * ny * wx) return;
    const int x = (int)(i % wx);
    const size_t row = i / wx;
    const float* s = src + row * nx + x;
    double p = 0.0, q = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const double v = s[k];
        p += v;
        q += v * v;
    }
    a1[i] = p;
    a2[i] = q;
}

// Sums of 16 consecutive entries along an axis of stride `step` (elements):
// output (outer, k, inner_i) = sum_{j<16} input (outer, k + j, inner_i), for
// k < n_out_axis.  A thread forms BS_RUN consecutive outputs along the axis
// from the BS_RUN + 15 inputs they cover, held in registers (19 loads per 4
// outputs instead of 64; longer runs leave too few threads to hide the load
// latency), each sum taken in the order j = 0..15 from 0.
constexpr int BS_RUN = 4;
__global__ __launch_bounds__(256) void boxsum_axis_kernel(const double* __restrict__ a1, const double* __restrict__ a2,
                                                          size_t n_runs, int inner, int n_in_axis, int n_out_axis,
                                                          size_t step, double* __restrict__ b1,
                                                          double* __restrict__ b2)
{
    // run index = (outer, run, inner_i); outputs k = run * BS_RUN .. + BS_RUN - 1
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_runs) return;
    const int runs_axis = (n_out_axis + BS_RUN - 1) / BS_RUN;
    const size_t in_i = i % inner;
    const size_t rest = i / inner;
    const int k0 = (int)(rest % runs_axis) * BS_RUN;
    const size_t outer = rest / runs_axis;
    const size_t base = (outer * n_in_axis + k0) * inner + in_i;
    const size_t obase = (outer * n_out_axis + k0) * inner + in_i;
    const int nk = min(BS_RUN, n_out_axis - k0);  // outputs of this run
    const int nin = nk + NB - 1;                  // inputs they cover (k0 + nin <= n_in_axis)
    double v1[BS_RUN + NB - 1], v2[BS_RUN + NB - 1];
#pragma unroll
    for (int j = 0; j < BS_RUN + NB - 1; ++j) {
        v1[j] = j < nin ? a1[base + j * step] : 0.0;
        v2[j] = j < nin ? a2[base + j * step] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < BS_RUN; ++k) {
        if (k >= nk) break;
        double p = 0.0, q = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            p += v1[k + j];
            q += v2[k + j];
        }
        b1[obase + k * step] = p;
        b2[obase + k * step] = q;
    }
}

// d_out: 2 * nz * ny * (nx - 15) doubles, receives [W1 | W2] (each (nz-15)(ny-15)(nx-15));
// d_tmp: 2 * nz * (ny - 15) * (nx - 15) doubles.
bool ncc3d_boxsums(const float* d_src, int nx, int ny, int nz, double* d_out, double* d_tmp, hipStream_t st)
{
    const int wx = nx - NB + 1, wy = ny - NB + 1, wz = nz - NB + 1;
    const size_t na = (size_t)nz * ny * wx, nbq = (size_t)nz * wy * wx, nw = (size_t)wz * wy * wx;
    boxsum_x_kernel<<<(unsigned)((na + 255) / 256), 256, 0, st>>>(d_src, nx, ny, nz, d_out, d_out + na);
    // y: (z, y, x) -> (z, y', x); z: (z, y', x) -> (z', y', x)
    const size_t ry = (size_t)nz * ((wy + BS_RUN - 1) / BS_RUN) * wx;
    const size_t rz = (size_t)((wz + BS_RUN - 1) / BS_RUN) * wy * wx;
    boxsum_axis_kernel<<<(unsigned)((ry + 255) / 256), 256, 0, st>>>(d_out, d_out + na, ry, wx, ny, wy, (size_t)wx,
                                                                      d_tmp, d_tmp + nbq);
    boxsum_axis_kernel<<<(unsigned)((rz + 255) / 256), 256, 0, st>>>(d_tmp, d_tmp + nbq, rz, wy * wx, nz, wz,
                                                                      (size_t)wy * wx, d_out, d_out + nw);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

size_t ncc3d_boxsum_doubles(int nx, int ny, int nz) { return 2 * (size_t)nz * ny * (nx - NB + 1); }
size_t ncc3d_boxsum_tmp_doubles(int nx, int ny, int nz) { return 2 * (size_t)nz * (ny - NB + 1) * (nx - NB + 1); }

struct N3Geom {
    int nx, ny, nz, stride, gx, gy;
    int wx, wy;  // box-sum volume row length and rows per plane
};

__device__ __forceinline__ void block_origin(const N3Geom& g, int b, int& bx, int& by, int& bz)
{
    bx = (b % g.gx) * g.stride;
    by = ((b / g.gx) % g.gy) * g.stride;
    bz = (b / (g.gx * g.gy)) * g.stride;
}

// template minus its mean into LDS (plane stride TPS); returns the FP64 energy (all threads)
__device__ double load_template(const float* __restrict__ tpl, const N3Geom& g, int bx, int by, int bz, float* tl,
                                double* red)
{
    const int tid = threadIdx.x, nt = blockDim.x, nw = (nt + 63) / 64;
    double s = 0.0;
    for (int k = tid; k < NB * NB * NB; k += nt) {
        const int x = k % NB, y = (k / NB) % NB, z = k / (NB * NB);
        const float v = tpl[((size_t)(bz + z) * g.ny + by + y) * g.nx + bx + x];
        tl[z * TPS + y * NB + x] = v;
        s += v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < nw; ++w) tot += red[w];
    const double mean = tot / (NB * NB * NB);
    __syncthreads();
    double e = 0.0;
    for (int k = tid; k < NB * NB * NB; k += nt) {
        const int i = (k >> 8) * TPS + (k & 255);
        const double d = (double)tl[i] - mean;
        tl[i] = (float)d;
        e += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = e;
    __syncthreads();
    double te = 0.0;
    for (int w = 0; w < nw; ++w) te += red[w];
    __syncthreads();
    return te;
}

__device__ __forceinline__ bool better(float v, int i, float bv, int bi) { return v > bv || (v == bv && i < bi); }

// workgroup argmax (value, shift index) -> thread 0 holds the result
__device__ void block_argmax(float& bv, int& bi, float* rv, int* ri)
{
    const int tid = threadIdx.x;
    for (int o = 32; o > 0; o >>= 1) {
        const float v = __shfl_down(bv, o, 64);
        const int i = __shfl_down(bi, o, 64);
        if (better(v, i, bv, bi)) { bv = v; bi = i; }
    }
    if ((tid & 63) == 0) { rv[tid >> 6] = bv; ri[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < (int)((blockDim.x + 63) / 64); ++w)
            if (better(rv[w], ri[w], bv, bi)) { bv = rv[w]; bi = ri[w]; }
    }
}

__device__ __forceinline__ float zncc_from(float cross, double te, const double* W1, const double* W2,
                                           const N3Geom& g, int sx, int sy, int sz)
{
    const size_t i = ((size_t)sz * g.wy + sy) * g.wx + sx;
    const double n = (double)NB * NB * NB;
    const double s1 = W1[i], s2 = W2[i];
    const double se = s2 - s1 * s1 / n;
    const double den = sqrt(te * (se > 0.0 ? se : 0.0));
    return den > 0.0 ? (float)((double)cross / den) : 0.0f;
}

__device__ __forceinline__ void write_result(int o, int bi, float bv, int* disp, float* peak)
{
    if (bi == 0x7fffffff) {  // no valid shift
        disp[3 * o] = 0; disp[3 * o + 1] = 0; disp[3 * o + 2] = 0;
        peak[o] = -INFINITY;
    } else {
        disp[3 * o] = bi % NS - NM;
        disp[3 * o + 1] = (bi / NS) % NS - NM;
        disp[3 * o + 2] = bi / (NS * NS) - NM;
        peak[o] = bv;
    }
}

// ids: block list (nullptr = blocks 0..); count: device block count that
// bounds the grid (nullptr = grid size); by_id: write results at the block id
// instead of the list position.
__global__ __launch_bounds__(NF_THREADS) void ncc3d_full_kernel(const float* __restrict__ tpl,
                                                                const float* __restrict__ src,
                                                                const double* __restrict__ W1,
                                                                const double* __restrict__ W2, const N3Geom g,
                                                                const int* __restrict__ ids,
                                                                const int* __restrict__ count, int by_id,
                                                                int* __restrict__ disp, float* __restrict__ peak)
{
    __shared__ __attribute__((aligned(16))) float tl[NB * TPS];
    __shared__ __attribute__((aligned(16))) float sp[2][NROW * RS];
    __shared__ double red[NF_THREADS / 64];
    __shared__ float rv[NF_THREADS / 64];
    __shared__ int ri[NF_THREADS / 64];
    if (count && (int)blockIdx.x >= *count) return;
    const int b = ids ? ids[blockIdx.x] : (int)blockIdx.x;
    int bx, by, bz;
    block_origin(g, b, bx, by, bz);
    const double te = load_template(tpl, g, bx, by, bz, tl, red);

    const int tid = threadIdx.x;
    const int dzi = tid / NS, dyi = tid % NS;
    const int dy = dyi - NM, dz = dzi - NM;
    const int dxlo = max(-NM, -bx), dxhi = min(NM, g.nx - NB - bx);
    const int dylo = max(-NM, -by), dyhi = min(NM, g.ny - NB - by);
    const int dzlo = max(-NM, -bz), dzhi = min(NM, g.nz - NB - bz);
    const bool active = tid < NS * NS && dy >= dylo && dy <= dyhi && dz >= dzlo && dz <= dzhi;

    // staging: 1024 samples of a 32 x 32 source plane, clamped to the volume
    float pre[4];
    auto fetch = [&](int zr) {
        const int Z = min(max(bz - NM + zr, 0), g.nz - 1);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = tid + i * NF_THREADS;
            if (k < NROW * NROW) {
                const int Y = min(max(by - NM + (k >> 5), 0), g.ny - 1);
                const int X = min(max(bx - NM + (k & 31), 0), g.nx - 1);
                pre[i] = src[((size_t)Z * g.ny + Y) * g.nx + X];
            }
        }
    };
    auto put = [&](float* dst) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = tid + i * NF_THREADS;
            if (k < NROW * NROW) dst[(k >> 5) * RS + (k & 31)] = pre[i];
        }
    };

    float acc[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) acc[i] = 0.0f;
    fetch(0);
    put(sp[0]);
    __syncthreads();
    for (int zr = 0; zr < NROW; ++zr) {
        if (zr + 1 < NROW) fetch(zr + 1);
        const int z = zr - dzi;  // template plane meeting source plane zr at this dz
        if (active && z >= 0 && z < NB) {
            const float* pl = sp[zr & 1];
            for (int y = 0; y < NB; ++y) {
                const float4* s4 = reinterpret_cast<const float4*>(pl + (y + dyi) * RS);
                const float4* t4 = reinterpret_cast<const float4*>(tl + z * TPS + y * NB);
                float r[NROW], t[NB];
#pragma unroll
                for (int k = 0; k < NROW / 4; ++k) {
                    const float4 v = s4[k];
                    r[4 * k] = v.x; r[4 * k + 1] = v.y; r[4 * k + 2] = v.z; r[4 * k + 3] = v.w;
                }
#pragma unroll
                for (int k = 0; k < NB / 4; ++k) {
                    const float4 v = t4[k];
                    t[4 * k] = v.x; t[4 * k + 1] = v.y; t[4 * k + 2] = v.z; t[4 * k + 3] = v.w;
                }
#pragma unroll
                for (int x = 0; x < NB; ++x)
#pragma unroll
                    for (int d = 0; d < NS; ++d) acc[d] = fmaf(t[x], r[x + d], acc[d]);
            }
        }
        if (zr + 1 < NROW) put(sp[(zr + 1) & 1]);
        __syncthreads();
    }
    float bv = -INFINITY;
    int bi = 0x7fffffff;
    if (active) {
        const int sy = by + dy, sz = bz + dz;
#pragma unroll
        for (int d = 0; d < NS; ++d) {
            const int dx = d - NM;
            if (dx < dxlo || dx > dxhi) continue;
            const float v = zncc_from(acc[d], te, W1, W2, g, bx + dx, sy, sz);
            const int idx = (dzi * NS + dyi) * NS + d;
            if (better(v, idx, bv, bi)) { bv = v; bi = idx; }
        }
    }
    block_argmax(bv, bi, rv, ri);
    if (tid == 0) write_result(by_id ? b : (int)blockIdx.x, bi, bv, disp, peak);
}

// Windowed search |d - centre| <= R (R <= 3).  One thread per (dz, dy) pair and
// template plane (whole waves: 448 threads at R = 2), so every thread runs one
// plane and no group of planes waits for a longer one.
constexpr int NW_MAXR = 3;
constexpr int nw_threads(int R) { return ((2 * R + 1) * (2 * R + 1) * NB + 63) / 64 * 64; }
template <int R>
__global__ __launch_bounds__(nw_threads(R)) void ncc3d_window_kernel(const float* __restrict__ tpl,
                                                                  const float* __restrict__ src,
                                                                  const double* __restrict__ W1,
                                                                  const double* __restrict__ W2, const N3Geom g,
                                                                  const int* __restrict__ ids,
                                                                  const int* __restrict__ centre, int by_id,
                                                                  int* __restrict__ disp, float* __restrict__ peak)
{
    constexpr int NWT = nw_threads(R);
    constexpr int w = 2 * R + 1, e = NB + 2 * R, pairs = w * w;
    constexpr int Q = (NWT / pairs) < NB ? (NWT / pairs) : NB;
    static_assert(Q * pairs * w <= e * e * e, "partial sums reuse the region buffer");
    __shared__ __attribute__((aligned(16))) float tl[NB * TPS];
    __shared__ __attribute__((aligned(16))) float reg[e * e * e];
    __shared__ double red[NWT / 64];
    __shared__ float rv[NWT / 64];
    __shared__ int ri[NWT / 64];
    const int b = ids ? ids[blockIdx.x] : (int)blockIdx.x;
    const int o = by_id ? b : (int)blockIdx.x;
    int bx, by, bz;
    block_origin(g, b, bx, by, bz);
    const int tid = threadIdx.x;
    const int cx = centre[3 * o], cy = centre[3 * o + 1], cz = centre[3 * o + 2];
    // region sample 0 sits at source (bx + cx - R, by + cy - R, bz + cz - R); clamped reads
    // only feed shifts that are excluded below
    const int ox = bx + cx - R, oy = by + cy - R, oz = bz + cz - R;
    for (int k = tid; k < e * e * e; k += NWT) {
        const int X = min(max(ox + k % e, 0), g.nx - 1);
        const int Y = min(max(oy + (k / e) % e, 0), g.ny - 1);
        const int Z = min(max(oz + k / (e * e), 0), g.nz - 1);
        reg[k] = src[((size_t)Z * g.ny + Y) * g.nx + X];
    }
    const double te = load_template(tpl, g, bx, by, bz, tl, red);  // synchronises

    // thread -> (dz, dy) pair and a group of template planes
    const int pr = tid % pairs, q = tid / pairs;
    const int dyr = pr % w, dzr = pr / w;
    float acc[w];
#pragma unroll
    for (int i = 0; i < w; ++i) acc[i] = 0.0f;
    if (q < Q) {
        for (int z = q; z < NB; z += Q)
            for (int y = 0; y < NB; ++y) {
                const float* srow = reg + ((z + dzr) * e + (y + dyr)) * e;
                const float* t = tl + z * TPS + y * NB;
                float r[NB + 2 * R], tv16[NB];
                if constexpr (e % 4 == 0) {  // 16-byte aligned rows
#pragma unroll
                    for (int k = 0; k < e / 4; ++k) {
                        const float4 q4 = reinterpret_cast<const float4*>(srow)[k];
                        r[4 * k] = q4.x; r[4 * k + 1] = q4.y; r[4 * k + 2] = q4.z; r[4 * k + 3] = q4.w;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NB + 2 * R; ++k) r[k] = srow[k];
                }
#pragma unroll
                for (int k = 0; k < NB / 4; ++k) {
                    const float4 q4 = reinterpret_cast<const float4*>(t)[k];
                    tv16[4 * k] = q4.x; tv16[4 * k + 1] = q4.y; tv16[4 * k + 2] = q4.z; tv16[4 * k + 3] = q4.w;
                }
#pragma unroll
                for (int x = 0; x < NB; ++x) {
                    const float tv = tv16[x];
#pragma unroll
                    for (int d = 0; d < w; ++d) acc[d] = fmaf(tv, r[x + d], acc[d]);
                }
            }
    }
    __syncthreads();  // the region is no longer read: reuse it for the partial sums
    float* part = reg;
    if (q < Q) {
#pragma unroll
        for (int d = 0; d < w; ++d) part[(q * pairs + pr) * w + d] = acc[d];
    }
    __syncthreads();
    float bv = -INFINITY;
    int bi = 0x7fffffff;
    for (int s = tid; s < pairs * w; s += NWT) {
        const int dxr = s % w, p2 = s / w;  // p2 = dzr * w + dyr
        const int dx = cx - R + dxr, dy = cy - R + p2 % w, dz = cz - R + p2 / w;
        if (dx < -NM || dx > NM || dy < -NM || dy > NM || dz < -NM || dz > NM) continue;
        const int sx = bx + dx, sy = by + dy, sz = bz + dz;
        if (sx < 0 || sy < 0 || sz < 0 || sx + NB > g.nx || sy + NB > g.ny || sz + NB > g.nz) continue;
        float c = 0.0f;
        for (int qq = 0; qq < Q; ++qq) c += part[(qq * pairs + p2) * w + dxr];
        const float v = zncc_from(c, te, W1, W2, g, sx, sy, sz);
        const int idx = ((dz + NM) * NS + dy + NM) * NS + dx + NM;
        if (better(v, idx, bv, bi)) { bv = v; bi = idx; }
    }
    block_argmax(bv, bi, rv, ri);
    if (tid == 0) write_result(o, bi, bv, disp, peak);
}

bool launch_ncc3d(const float* d_tpl, const float* d_src, const double* d_w, int nx, int ny, int nz, int stride,
                  const int* d_ids, int nblocks, const int* d_count, bool by_id, const int* d_centre, int radius,
                  int* d_disp, float* d_peak, hipStream_t st)
{
    if (nblocks <= 0) return true;
    const int wx = nx - NB + 1, wy = ny - NB + 1, wz = nz - NB + 1;
    const double* W1 = d_w;
    const double* W2 = d_w + (size_t)wx * wy * wz;
    N3Geom g{nx, ny, nz, stride, (nx - NB) / stride + 1, (ny - NB) / stride + 1, wx, wy};
    if (d_centre) {
        if (radius < 0 || radius > NW_MAXR) CUT_FAIL("ncc3d: window radius must be 0.." << NW_MAXR);
        if (d_count) CUT_FAIL("ncc3d: device block counts apply to the full search only");
        const int bid = by_id ? 1 : 0;
        switch (radius) {
        case 0: ncc3d_window_kernel<0><<<nblocks, nw_threads(0), 0, st>>>(d_tpl, d_src, W1, W2, g, d_ids, d_centre, bid, d_disp, d_peak); break;
        case 1: ncc3d_window_kernel<1><<<nblocks, nw_threads(1), 0, st>>>(d_tpl, d_src, W1, W2, g, d_ids, d_centre, bid, d_disp, d_peak); break;
        case 2: ncc3d_window_kernel<2><<<nblocks, nw_threads(2), 0, st>>>(d_tpl, d_src, W1, W2, g, d_ids, d_centre, bid, d_disp, d_peak); break;
        default: ncc3d_window_kernel<3><<<nblocks, nw_threads(3), 0, st>>>(d_tpl, d_src, W1, W2, g, d_ids, d_centre, bid, d_disp, d_peak); break;
        }
    } else {
        ncc3d_full_kernel<<<nblocks, NF_THREADS, 0, st>>>(d_tpl, d_src, W1, W2, g, d_ids, d_count, by_id ? 1 : 0,
                                                          d_disp, d_peak);
    }
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

// Blocks that need the full search after a windowed search around `centre`:
// peak below the threshold (or no valid shift), or, with `boundary`, a maximum
// on a window face that is not a limit of the full search.  Appends them to
// ids (order unspecified; each block's result does not depend on it).
__global__ void ncc3d_fallback_kernel(const int* __restrict__ disp, const float* __restrict__ peak,
                                      const int* __restrict__ centre, int nb, const N3Geom g, int radius, float thr,
                                      int boundary, int* __restrict__ ids, int* __restrict__ count,
                                      int* __restrict__ path)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    bool need = !(peak[b] >= thr);
    if (!need && boundary) {
        int o[3];
        block_origin(g, b, o[0], o[1], o[2]);
        const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int d = disp[3 * b + a], c = centre[3 * b + a];
            const int lo = max(-NM, -o[a]), hi = min(NM, n[a] - NB - o[a]);
            if ((d == c - radius && d > lo) || (d == c + radius && d < hi)) need = true;
        }
    }
    if (path) path[b] = need ? 1 : 0;
    if (need) ids[atomicAdd(count, 1)] = b;
}

__global__ void zero_int_kernel(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

bool launch_ncc3d_fallback(const int* d_disp, const float* d_peak, const int* d_centre, int nx, int ny, int nz,
                           int stride, int nb, int radius, float thr, bool boundary, int* d_ids, int* d_count,
                           int* d_path, hipStream_t st)
{
    N3Geom g{nx, ny, nz, stride, (nx - NB) / stride + 1, (ny - NB) / stride + 1, nx - NB + 1, ny - NB + 1};
    zero_int_kernel<<<1, 64, 0, st>>>(d_count, 1);
    ncc3d_fallback_kernel<<<(nb + 255) / 256, 256, 0, st>>>(d_disp, d_peak, d_centre, nb, g, radius, thr,
                                                            boundary ? 1 : 0, d_ids, d_count, d_path);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

// |IQ| -> float volume
__global__ void envelope_kernel(const float2* __restrict__ iq, size_t n, float* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 v = iq[i];
        out[i] = sqrtf(v.x * v.x + v.y * v.y);
    }
}

bool launch_envelope(const float2* d_iq, size_t n, float* d_out, hipStream_t st)
{
    envelope_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(d_iq, n, d_out);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

}  // namespace cut
