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
// Full search: one workgroup per block.  The template minus its mean sits in
// LDS.  Thread (dy, dz pair) walks the template rows (y, z) and correlates
// each 16-sample template row with the 32-sample source row it meets, giving
// all 17 dx shifts at once (16 x 17 FMAs per 32 loaded samples); consecutive z
// shifts share source rows, so each row is loaded once for two shifts.  Window
// sums of S and S^2 come from FP64 integral volumes (built once per source
// volume) and the energy normalisation is done in FP64.
#include "common.hpp"
#include "ncc.hpp"

#include <cmath>

namespace cut {

constexpr int NB = 16;           // block edge
constexpr int NM = 8;            // search margin
constexpr int NS = 2 * NM + 1;   // shifts per axis
constexpr int NROW = NB + 2 * NM;
constexpr int N3_THREADS = 160;  // 17 dy x 9 dz-pairs = 153 used

// Integral volume with a zero border: I[z+1][y+1][x+1] = sum over [0..z][0..y][0..x].
__global__ void integral_x_kernel(const float* __restrict__ src, int nx, int ny, int nz, double* __restrict__ i1,
                                  double* __restrict__ i2)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;  // (z, y) row
    const int P = nx + 1, Q = ny + 1;
    if (r >= (nz + 1) * (ny + 1)) return;
    const int z = r / Q, y = r % Q;
    double* o1 = i1 + (size_t)r * P;
    double* o2 = i2 + (size_t)r * P;
    o1[0] = 0.0;
    o2[0] = 0.0;
    if (z == 0 || y == 0) {
        for (int x = 1; x < P; ++x) { o1[x] = 0.0; o2[x] = 0.0; }
        return;
    }
    const float* s = src + ((size_t)(z - 1) * ny + (y - 1)) * nx;
    double a = 0.0, b = 0.0;
    for (int x = 0; x < nx; ++x) {
        const double v = s[x];
        a += v;
        b += v * v;
        o1[x + 1] = a;
        o2[x + 1] = b;
    }
}

__global__ void integral_scan_kernel(double* __restrict__ i1, double* __restrict__ i2, int P, int Q, int R, int axis)
{
    // prefix along y (axis 1) or z (axis 2) of a (R)(Q)(P) array
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (axis == 1) {
        if (t >= R * P) return;
        const int z = t / P, x = t % P;
        double a = 0.0, b = 0.0;
        for (int y = 0; y < Q; ++y) {
            const size_t i = ((size_t)z * Q + y) * P + x;
            a += i1[i]; b += i2[i];
            i1[i] = a; i2[i] = b;
        }
    } else {
        if (t >= Q * P) return;
        double a = 0.0, b = 0.0;
        for (int z = 0; z < R; ++z) {
            const size_t i = (size_t)z * Q * P + t;
            a += i1[i]; b += i2[i];
            i1[i] = a; i2[i] = b;
        }
    }
}

__device__ __forceinline__ double box_sum(const double* __restrict__ I, int P, int Q, int x, int y, int z, int B)
{
    auto at = [&](int zz, int yy, int xx) { return I[((size_t)zz * Q + yy) * P + xx]; };
    return at(z + B, y + B, x + B) - at(z, y + B, x + B) - at(z + B, y, x + B) - at(z + B, y + B, x) + at(z, y, x + B) +
           at(z, y + B, x) + at(z + B, y, x) - at(z, y, x);
}

struct N3Geom {
    int nx, ny, nz, stride, gx, gy;
};

__device__ __forceinline__ void block_origin(const N3Geom& g, int b, int& bx, int& by, int& bz)
{
    bx = (b % g.gx) * g.stride;
    by = ((b / g.gx) % g.gy) * g.stride;
    bz = (b / (g.gx * g.gy)) * g.stride;
}

// template minus its mean into LDS; returns the FP64 energy (all threads)
__device__ double load_template(const float* __restrict__ tpl, const N3Geom& g, int bx, int by, int bz, float* tl,
                                double* red)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    double s = 0.0;
    for (int k = tid; k < NB * NB * NB; k += nt) {
        const int x = k % NB, y = (k / NB) % NB, z = k / (NB * NB);
        const float v = tpl[((size_t)(bz + z) * g.ny + by + y) * g.nx + bx + x];
        tl[k] = v;
        s += v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < (nt + 63) / 64; ++w) tot += red[w];
    const double mean = tot / (NB * NB * NB);
    __syncthreads();
    double e = 0.0;
    for (int k = tid; k < NB * NB * NB; k += nt) {
        const double d = (double)tl[k] - mean;
        tl[k] = (float)d;
        e += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = e;
    __syncthreads();
    double te = 0.0;
    for (int w = 0; w < (nt + 63) / 64; ++w) te += red[w];
    __syncthreads();
    return te;
}

__device__ __forceinline__ bool better(float v, int i, float bv, int bi) { return v > bv || (v == bv && i < bi); }

// workgroup argmax (value, shift index) -> lane 0 of wave 0 holds the result
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

__device__ __forceinline__ float zncc_from(float cross, double te, const double* I1, const double* I2, const N3Geom& g,
                                           int sx, int sy, int sz)
{
    const int P = g.nx + 1, Q = g.ny + 1;
    const double n = (double)NB * NB * NB;
    const double s1 = box_sum(I1, P, Q, sx, sy, sz, NB);
    const double s2 = box_sum(I2, P, Q, sx, sy, sz, NB);
    const double se = s2 - s1 * s1 / n;
    const double den = sqrt(te * (se > 0.0 ? se : 0.0));
    return den > 0.0 ? (float)((double)cross / den) : 0.0f;
}

__global__ __launch_bounds__(N3_THREADS) void ncc3d_full_kernel(const float* __restrict__ tpl,
                                                                const float* __restrict__ src,
                                                                const double* __restrict__ I1,
                                                                const double* __restrict__ I2, const N3Geom g,
                                                                const int* __restrict__ ids, int* __restrict__ disp,
                                                                float* __restrict__ peak)
{
    __shared__ float tl[NB * NB * NB];
    __shared__ double red[4];
    __shared__ float rv[4];
    __shared__ int ri[4];
    const int b = ids ? ids[blockIdx.x] : (int)blockIdx.x;
    int bx, by, bz;
    block_origin(g, b, bx, by, bz);
    const double te = load_template(tpl, g, bx, by, bz, tl, red);

    const int tid = threadIdx.x;
    const int dyi = tid % NS, dzp = tid / NS;
    const int dy = dyi - NM, dz0 = 2 * dzp - NM;
    // valid shift ranges (source block inside the volume)
    const int dxlo = max(-NM, -bx), dxhi = min(NM, g.nx - NB - bx);
    const int dylo = max(-NM, -by), dyhi = min(NM, g.ny - NB - by);
    const int dzlo = max(-NM, -bz), dzhi = min(NM, g.nz - NB - bz);
    const bool active = dzp < (NS + 1) / 2 && dy >= dylo && dy <= dyhi && !(dz0 > dzhi) && !(dz0 + 1 < dzlo);

    float acc0[NS], acc1[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) { acc0[i] = 0.0f; acc1[i] = 0.0f; }
    if (active) {
        const int Y0 = by + dy;
        for (int y = 0; y < NB; ++y) {
            const int Y = Y0 + y;
            for (int z = 0; z <= NB; ++z) {
                // source row at Z = bz + z + dz0 (serves dz0 at z and dz0+1 at z-1)
                const int Z = min(max(bz + z + dz0, 0), g.nz - 1);
                const float* srow = src + ((size_t)Z * g.ny + Y) * g.nx;
                float r[NROW];
#pragma unroll
                for (int k = 0; k < NROW; ++k) r[k] = srow[min(max(bx - NM + k, 0), g.nx - 1)];
                if (z < NB) {
                    const float* t = tl + (z * NB + y) * NB;
#pragma unroll
                    for (int x = 0; x < NB; ++x) {
                        const float tv = t[x];
#pragma unroll
                        for (int d = 0; d < NS; ++d) acc0[d] = fmaf(tv, r[x + d], acc0[d]);
                    }
                }
                if (z > 0) {
                    const float* t = tl + ((z - 1) * NB + y) * NB;
#pragma unroll
                    for (int x = 0; x < NB; ++x) {
                        const float tv = t[x];
#pragma unroll
                        for (int d = 0; d < NS; ++d) acc1[d] = fmaf(tv, r[x + d], acc1[d]);
                    }
                }
            }
        }
    }
    // normalise and pick this thread's best
    float bv = -3.0e38f;
    int bi = 0x7fffffff;
    if (active) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int dz = dz0 + h;
            if (dz < dzlo || dz > dzhi) continue;
#pragma unroll
            for (int d = 0; d < NS; ++d) {
                const int dx = d - NM;
                if (dx < dxlo || dx > dxhi) continue;
                const float v = zncc_from(h ? acc1[d] : acc0[d], te, I1, I2, g, bx + dx, by + dy, bz + dz);
                const int idx = ((dz + NM) * NS + dyi) * NS + d;
                if (better(v, idx, bv, bi)) { bv = v; bi = idx; }
            }
        }
    }
    block_argmax(bv, bi, rv, ri);
    if (tid == 0) {
        const int d = bi % NS, dyy = (bi / NS) % NS, dzz = bi / (NS * NS);
        disp[3 * blockIdx.x] = d - NM;
        disp[3 * blockIdx.x + 1] = dyy - NM;
        disp[3 * blockIdx.x + 2] = dzz - NM;
        peak[blockIdx.x] = bv;
    }
}

// Windowed search |d - centre| <= radius (radius <= 3): one thread per shift.
__global__ __launch_bounds__(N3_THREADS) void ncc3d_window_kernel(const float* __restrict__ tpl,
                                                                  const float* __restrict__ src,
                                                                  const double* __restrict__ I1,
                                                                  const double* __restrict__ I2, const N3Geom g,
                                                                  const int* __restrict__ ids,
                                                                  const int* __restrict__ centre, int radius,
                                                                  int* __restrict__ disp, float* __restrict__ peak)
{
    __shared__ float tl[NB * NB * NB];
    __shared__ double red[4];
    __shared__ float rv[4];
    __shared__ int ri[4];
    const int b = ids ? ids[blockIdx.x] : (int)blockIdx.x;
    int bx, by, bz;
    block_origin(g, b, bx, by, bz);
    const double te = load_template(tpl, g, bx, by, bz, tl, red);
    const int w = 2 * radius + 1;
    const int tid = threadIdx.x;
    const int cx = centre[3 * blockIdx.x], cy = centre[3 * blockIdx.x + 1], cz = centre[3 * blockIdx.x + 2];
    float bv = -3.0e38f;
    int bi = 0x7fffffff;
    for (int s = tid; s < w * w * w; s += blockDim.x) {
        const int dx = cx + s % w - radius, dy = cy + (s / w) % w - radius, dz = cz + s / (w * w) - radius;
        if (dx < -NM || dx > NM || dy < -NM || dy > NM || dz < -NM || dz > NM) continue;
        const int sx = bx + dx, sy = by + dy, sz = bz + dz;
        if (sx < 0 || sy < 0 || sz < 0 || sx + NB > g.nx || sy + NB > g.ny || sz + NB > g.nz) continue;
        float c = 0.0f;
        for (int z = 0; z < NB; ++z)
            for (int y = 0; y < NB; ++y) {
                const float* srow = src + ((size_t)(sz + z) * g.ny + sy + y) * g.nx + sx;
                const float* t = tl + (z * NB + y) * NB;
#pragma unroll
                for (int x = 0; x < NB; ++x) c = fmaf(t[x], srow[x], c);
            }
        const float v = zncc_from(c, te, I1, I2, g, sx, sy, sz);
        const int idx = ((dz + NM) * NS + dy + NM) * NS + dx + NM;
        if (better(v, idx, bv, bi)) { bv = v; bi = idx; }
    }
    block_argmax(bv, bi, rv, ri);
    if (tid == 0) {
        if (bi == 0x7fffffff) {  // no valid shift in the window
            disp[3 * blockIdx.x] = cx; disp[3 * blockIdx.x + 1] = cy; disp[3 * blockIdx.x + 2] = cz;
            peak[blockIdx.x] = -3.0e38f;
        } else {
            disp[3 * blockIdx.x] = bi % NS - NM;
            disp[3 * blockIdx.x + 1] = (bi / NS) % NS - NM;
            disp[3 * blockIdx.x + 2] = bi / (NS * NS) - NM;
            peak[blockIdx.x] = bv;
        }
    }
}

bool ncc3d_integrals(const float* d_src, int nx, int ny, int nz, double* d_i1, double* d_i2, hipStream_t st)
{
    const int P = nx + 1, Q = ny + 1, R = nz + 1;
    integral_x_kernel<<<(R * Q + 127) / 128, 128, 0, st>>>(d_src, nx, ny, nz, d_i1, d_i2);
    integral_scan_kernel<<<(R * P + 127) / 128, 128, 0, st>>>(d_i1, d_i2, P, Q, R, 1);
    integral_scan_kernel<<<(Q * P + 127) / 128, 128, 0, st>>>(d_i1, d_i2, P, Q, R, 2);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool launch_ncc3d(const float* d_tpl, const float* d_src, const double* d_i1, const double* d_i2, int nx, int ny,
                  int nz, int stride, const int* d_ids, int nblocks, const int* d_centre, int radius, int* d_disp,
                  float* d_peak, hipStream_t st)
{
    if (nblocks <= 0) return true;
    N3Geom g{nx, ny, nz, stride, (nx - NB) / stride + 1, (ny - NB) / stride + 1};
    if (d_centre) {
        if (radius < 0 || radius > 3) CUT_FAIL("ncc3d: window radius must be 0..3");
        ncc3d_window_kernel<<<nblocks, N3_THREADS, 0, st>>>(d_tpl, d_src, d_i1, d_i2, g, d_ids, d_centre, radius,
                                                            d_disp, d_peak);
    } else {
        ncc3d_full_kernel<<<nblocks, N3_THREADS, 0, st>>>(d_tpl, d_src, d_i1, d_i2, g, d_ids, d_disp, d_peak);
    }
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
