// Displacement-proposal network: a small 3-D CNN mapping a pair of 16^3
// envelope blocks (template block, source block at the same position) to a
// 3-D displacement in voxels.  BF16 MFMA products with FP32 accumulation.
//
//   L1 conv 2->16,  stride 2, 16^3 -> 8^3, ReLU
//   L2 conv 16->32, stride 2,  8^3 -> 4^3, ReLU
//   L3 conv 32->32, stride 1,  4^3 -> 4^3, ReLU
//   L4 conv 32->32, stride 2,  4^3 -> 2^3, ReLU
//   FC 256 -> 3 (input index c * 8 + position), FP32
// 3^3 kernels with a zero border of one voxel.  FP32 parameters packed as
// w1 b1 w2 b2 w3 b3 w4 b4 wf bf, conv weights [co][ci][kz][ky][kx].
//
// Training path (batched): im2col to BF16, one MFMA GEMM kernel for the
// forward product, the weight gradient (split-K, partials reduced in a fixed
// order) and the column gradient; gather-form col2im with the ReLU mask;
// FP32 FC layer, loss and Adam.  Every reduction has a fixed order, so a
// training run is reproducible.
//
// Inference path: one fused kernel, one workgroup per block: extract and
// normalise the block pair into LDS (BF16, channel-last), then each conv layer
// as an implicit GEMM on 16x16x32 BF16 MFMAs (rows = output positions,
// columns = output channels, k = (tap, input channel)), activations kept in
// LDS, weights read as pre-packed MFMA fragments, FC in FP32.
#include "common.hpp"
#include "pnet.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace cut {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t to_bf(float f) { return f2bf(f); }
__device__ __forceinline__ uint16_t to_bf(uint16_t h) { return h; }

// ---------------------------------------------------------------------------
// block extraction: zero-mean, unit-variance template and source blocks
// ---------------------------------------------------------------------------
// Sum over a 256-thread workgroup (all threads receive the total).
__device__ __forceinline__ float wg_sum256(float v, float* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Thread t of 256 loads row (z, y) = (t / 16, t % 16) of the network input of
// a block whose first voxel is (x0, y0, z0), with the axes in `flips`
// mirrored, and normalises it over the 16^3 input; coordinates are clamped to
// the volume.  mode 0: the 16^3 block itself; mode 1: the 32^3 neighbourhood
// [x0 - 8, x0 + 24) (the +-8 search region of the block), averaged 2 x 2 x 2.
__device__ __forceinline__ void load_norm_row(const float* __restrict__ vol, int nx, int ny, int nz, int x0, int y0,
                                              int z0, int flips, int mode, float v[16], float* red)
{
    const int t = threadIdx.x;
    int z = t >> 4, y = t & 15;
    if (flips & 4) z = 15 - z;
    if (flips & 2) y = 15 - y;
    float s = 0.0f;
    if (mode == 0) {
        const int Z = min(max(z0 + z, 0), nz - 1), Y = min(max(y0 + y, 0), ny - 1);
        const float* row = vol + ((size_t)Z * ny + Y) * nx;
#pragma unroll
        for (int x = 0; x < 16; ++x) {
            const int xx = (flips & 1) ? 15 - x : x;
            v[x] = row[min(max(x0 + xx, 0), nx - 1)];
            s += v[x];
        }
    } else {
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = 0.0f;
        for (int q = 0; q < 4; ++q) {
            const int Z = min(max(z0 - 8 + 2 * z + (q >> 1), 0), nz - 1);
            const int Y = min(max(y0 - 8 + 2 * y + (q & 1), 0), ny - 1);
            const float* row = vol + ((size_t)Z * ny + Y) * nx;
#pragma unroll
            for (int x = 0; x < 16; ++x) {
                const int xx = (flips & 1) ? 15 - x : x;
                const int X = x0 - 8 + 2 * xx;
                v[x] += row[min(max(X, 0), nx - 1)] + row[min(max(X + 1, 0), nx - 1)];
            }
        }
#pragma unroll
        for (int x = 0; x < 16; ++x) {
            v[x] *= 0.125f;
            s += v[x];
        }
    }
    const float mean = wg_sum256(s, red) * (1.0f / 4096.0f);
    float q = 0.0f;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        v[x] -= mean;
        q += v[x] * v[x];
    }
    const float var = wg_sum256(q, red) * (1.0f / 4096.0f);
    const float inv = var > 0.0f ? 1.0f / sqrtf(var) : 0.0f;
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] *= inv;
}

__device__ __forceinline__ void block_xyz(int b, int gx, int gy, int stride, int& x, int& y, int& z)
{
    x = (b % gx) * stride;
    y = ((b / gx) % gy) * stride;
    z = (b / (gx * gy)) * stride;
}

// 2 x 2 x 2 averages of a volume with clamped coordinates, over the padded
// grid P[k] = mean of vol[clamp(2k - 8 + {0, 1})] per axis, k < pn = n / 2 + 8:
// the input of input mode 1 for a block at x0 (x0 even) is P[x0/2 .. x0/2 + 16).
__global__ void pn_pool_kernel(const float* __restrict__ vol, int nx, int ny, int nz, int pnx, int pny, int pnz,
                               float* __restrict__ P)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)pnx * pny * pnz) return;
    const int kx = (int)(i % pnx), ky = (int)((i / pnx) % pny), kz = (int)(i / ((size_t)pnx * pny));
    float s = 0.0f;
    for (int q = 0; q < 4; ++q) {
        const int Z = min(max(2 * kz - 8 + (q >> 1), 0), nz - 1);
        const int Y = min(max(2 * ky - 8 + (q & 1), 0), ny - 1);
        const float* row = vol + ((size_t)Z * ny + Y) * nx;
        s += row[min(max(2 * kx - 8, 0), nx - 1)] + row[min(max(2 * kx - 7, 0), nx - 1)];
    }
    P[i] = s * 0.125f;
}

// Row (z, y) of the 16^3 window of a pooled volume starting at (px, py, pz),
// normalised over the window.  vec4: rows are 16-byte aligned.
__device__ __forceinline__ void load_norm_pooled(const float* __restrict__ P, int pnx, int pny, int px, int py, int pz,
                                                 bool vec4, float v[16], float* red)
{
    const int t = threadIdx.x;
    const float* row = P + ((size_t)(pz + (t >> 4)) * pny + py + (t & 15)) * pnx + px;
    if (vec4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 q = reinterpret_cast<const float4*>(row)[k];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = row[x];
    }
    float s = 0.0f;
#pragma unroll
    for (int x = 0; x < 16; ++x) s += v[x];
    const float mean = wg_sum256(s, red) * (1.0f / 4096.0f);
    float q = 0.0f;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        v[x] -= mean;
        q += v[x] * v[x];
    }
    const float var = wg_sum256(q, red) * (1.0f / 4096.0f);
    const float inv = var > 0.0f ? 1.0f / sqrtf(var) : 0.0f;
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] *= inv;
}

// desc[i] = {pair, block, ox, oy, oz, flips}: template block `block` of pair
// `pair`, source block at the same position plus (ox, oy, oz).
__global__ __launch_bounds__(256) void pn_gather_kernel(const uint64_t* __restrict__ vols, int nx, int ny, int nz,
                                                        int stride, const int* __restrict__ desc, int mode,
                                                        float* __restrict__ out)
{
    __shared__ float red[4];
    const int* d = desc + 6 * blockIdx.x;
    const float* tpl = reinterpret_cast<const float*>(vols[2 * d[0]]);
    const float* src = reinterpret_cast<const float*>(vols[2 * d[0] + 1]);
    const int gx = (nx - 16) / stride + 1, gy = (ny - 16) / stride + 1;
    int bx, by, bz;
    block_xyz(d[1], gx, gy, stride, bx, by, bz);
    float v[16];
    float* o = out + (size_t)blockIdx.x * 8192 + threadIdx.x * 16;
    load_norm_row(tpl, nx, ny, nz, bx, by, bz, d[5], mode, v, red);
#pragma unroll
    for (int x = 0; x < 16; ++x) o[x] = v[x];
    load_norm_row(src, nx, ny, nz, bx + d[2], by + d[3], bz + d[4], d[5], mode, v, red);
#pragma unroll
    for (int x = 0; x < 16; ++x) o[4096 + x] = v[x];
}

// ---------------------------------------------------------------------------
// training path
// ---------------------------------------------------------------------------
// C[m][n] = sum_k A(m, k) B(k, n) over k in this workgroup's K chunk, BF16
// operands (converted on the way into LDS), FP32 accumulation.  Tile 64 x 64
// x 32; wave w owns the 32 x 32 sub-tile (w & 1, w >> 1).  EPI 1 stores
// max(C + bias[n], 0).  Split-K partials go to C + blockIdx.z * zstride.
template <typename TA, typename TB, int EPI>
__global__ __launch_bounds__(256) void pn_gemm_kernel(int M, int N, int K, const TA* __restrict__ A, long sam,
                                                      long sak, const TB* __restrict__ B, long sbk, long sbn,
                                                      float* __restrict__ C, long scm, long scn, int kchunk,
                                                      long zstride, const float* __restrict__ bias)
{
    constexpr int LD = 40;  // LDS row stride (bf16): 80 bytes keeps 16-byte reads aligned and spread over banks
    __shared__ __attribute__((aligned(16))) uint16_t As[64 * LD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * LD];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
    const long m0 = (long)blockIdx.x * 64, n0 = (long)blockIdx.y * 64;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    const bool a_mfast = sam == 1, b_nfast = sbn == 1;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += 32) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + i * 256;
            const int am = a_mfast ? (e & 63) : (e >> 5), ak = a_mfast ? (e >> 6) : (e & 31);
            const long gm = m0 + am;
            const int gk = k0 + ak;
            As[am * LD + ak] = (gm < M && gk < ke) ? to_bf(A[gm * sam + (long)gk * sak]) : (uint16_t)0;
            const int bn = b_nfast ? (e & 63) : (e >> 5), bk = b_nfast ? (e >> 6) : (e & 31);
            const long gn = n0 + bn;
            const int gk2 = k0 + bk;
            Bs[bn * LD + bk] = (gn < N && gk2 < ke) ? to_bf(B[(long)gk2 * sbk + gn * sbn]) : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(As + (wm * 32 + (lane & 31)) * LD + kk + 8 * (lane >> 5));
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bs + (wn * 32 + (lane & 31)) * LD + kk + 8 * (lane >> 5));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    C += (long)blockIdx.z * zstride;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long n = n0 + wn * 32 + (lane & 31);
        if (m < M && n < N) {
            float v = acc[r];
            if (EPI == 1) v = fmaxf(v + bias[n], 0.0f);
            C[m * scm + n * scn] = v;
        }
    }
}

// The same product with FP32 operands on the vector ALUs (verification
// mode): thread (tm, tn) of 16 x 16 owns rows tm + 16 i and columns tn + 16 j.
template <int EPI>
__global__ __launch_bounds__(256) void pn_gemm_f32_kernel(int M, int N, int K, const float* __restrict__ A, long sam,
                                                          long sak, const float* __restrict__ B, long sbk, long sbn,
                                                          float* __restrict__ C, long scm, long scn, int kchunk,
                                                          long zstride, const float* __restrict__ bias)
{
    __shared__ float As[32][65];
    __shared__ float Bs[32][65];
    const int tid = threadIdx.x, tm = tid & 15, tn = tid >> 4;
    const long m0 = (long)blockIdx.x * 64, n0 = (long)blockIdx.y * 64;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    const bool a_mfast = sam == 1, b_nfast = sbn == 1;
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        for (int i = 0; i < 8; ++i) {
            const int e = tid + i * 256;
            const int am = a_mfast ? (e & 63) : (e >> 5), ak = a_mfast ? (e >> 6) : (e & 31);
            const long gm = m0 + am;
            const int gk = k0 + ak;
            As[ak][am] = (gm < M && gk < ke) ? A[gm * sam + (long)gk * sak] : 0.0f;
            const int bn = b_nfast ? (e & 63) : (e >> 5), bk = b_nfast ? (e >> 6) : (e & 31);
            const long gn = n0 + bn;
            const int gk2 = k0 + bk;
            Bs[bk][bn] = (gn < N && gk2 < ke) ? B[(long)gk2 * sbk + gn * sbn] : 0.0f;
        }
        __syncthreads();
        for (int k = 0; k < 32; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(As[k][tm + 16 * i], Bs[k][tn + 16 * j], acc[i][j]);
        __syncthreads();
    }
    C += (long)blockIdx.z * zstride;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long m = m0 + tm + 16 * i, n = n0 + tn + 16 * j;
            if (m < M && n < N) {
                float v = acc[i][j];
                if (EPI == 1) v = fmaxf(v + bias[n], 0.0f);
                C[m * scm + n * scn] = v;
            }
        }
}

// col[row][ci * 27 + tap] (BF16, or FP32 in verification mode) for row = b * Do^3 + output position; input
// element (b, c, z, y, x) at X[b * sb + c * sc + ((z * D + y) * D + x) * sp].
__device__ __forceinline__ void store_col(uint16_t* p, float v) { *p = f2bf(v); }
__device__ __forceinline__ void store_col(float* p, float v) { *p = v; }

template <typename TC>
__global__ void pn_im2col_kernel(const float* __restrict__ X, long sb, long sc, long sp, int D, int s, int Do, int ci,
                                 long total, TC* __restrict__ col)
{
    const int K = ci * 27, P = Do * Do * Do;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i % K);
        const long row = i / K;
        const int c = k / 27, tap = k % 27;
        const long b = row / P;
        const int pos = (int)(row % P);
        const int iz = (pos / (Do * Do)) * s + tap / 9 - 1;
        const int iy = ((pos / Do) % Do) * s + (tap / 3) % 3 - 1;
        const int ix = (pos % Do) * s + tap % 3 - 1;
        float v = 0.0f;
        if (iz >= 0 && iy >= 0 && ix >= 0 && iz < D && iy < D && ix < D)
            v = X[b * sb + c * sc + ((long)(iz * D + iy) * D + ix) * sp];
        store_col(col + i, v);
    }
}

// G[b][pos][c] = [act > 0] * sum of the column gradients that read input (b, pos, c).
__global__ void pn_col2im_kernel(const float* __restrict__ dcol, int D, int s, int Do, int ci, long total,
                                 const float* __restrict__ act, float* __restrict__ G)
{
    const int K = ci * 27, P = Do * Do * Do, DI = D * D * D;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % ci);
        const long rest = i / ci;
        const int pos = (int)(rest % DI);
        const long b = rest / DI;
        const int iz = pos / (D * D), iy = (pos / D) % D, ix = pos % D;
        float sum = 0.0f;
        for (int tap = 0; tap < 27; ++tap) {
            const int tz = iz + 1 - tap / 9, ty = iy + 1 - (tap / 3) % 3, tx = ix + 1 - tap % 3;
            if (tz < 0 || ty < 0 || tx < 0 || tz % s || ty % s || tx % s) continue;
            const int oz = tz / s, oy = ty / s, ox = tx / s;
            if (oz >= Do || oy >= Do || ox >= Do) continue;
            sum += dcol[(b * P + (oz * Do + oy) * Do + ox) * K + c * 27 + tap];
        }
        G[i] = act[i] > 0.0f ? sum : 0.0f;
    }
}

// out[j] (+)= sum over z of part[z * stride + j], z ascending
__global__ void pn_splitk_reduce_kernel(const float* __restrict__ part, int splits, long stride, long n,
                                        float* __restrict__ out)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[z * stride + j];
    out[j] = s;
}

// db[c] = sum over rows of G[row][c]; one workgroup per channel
__global__ __launch_bounds__(256) void pn_bias_grad_kernel(const float* __restrict__ G, long rows, int co,
                                                           float* __restrict__ db)
{
    __shared__ float red[4];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (long r = threadIdx.x; r < rows; r += 256) s += G[r * co + c];
    s = wg_sum256(s, red);
    if (threadIdx.x == 0) db[c] = s;
}

// y[b][o] = bf[o] + sum_{c, p} wf[o][c * 8 + p] act4[(b * 8 + p) * 32 + c]
__global__ void pn_fc_forward_kernel(const float* __restrict__ act4, const float* __restrict__ wf,
                                     const float* __restrict__ bfc, int B, float* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * 3) return;
    const int b = i / 3, o = i % 3;
    float s = bfc[o];
    for (int j = 0; j < 256; ++j) s = fmaf(wf[o * 256 + j], act4[(b * 8 + (j & 7)) * 32 + (j >> 3)], s);
    y[i] = s;
}

// dy = (y - t) / B; loss = sum 0.5 (y - t)^2 / B  (one workgroup)
__global__ __launch_bounds__(256) void pn_loss_kernel(const float* __restrict__ y, const float* __restrict__ t, int B,
                                                      float* __restrict__ dy, float* __restrict__ loss)
{
    __shared__ float red[4];
    float s = 0.0f;
    const float inv = 1.0f / (float)B;
    for (int i = threadIdx.x; i < 3 * B; i += 256) {
        const float e = y[i] - t[i];
        dy[i] = e * inv;
        s += 0.5f * e * e * inv;
    }
    s = wg_sum256(s, red);
    if (threadIdx.x == 0 && loss) *loss = s;
}

// dwf, dbf (thread per weight, batch loop in order) and G4 = [act4 > 0] * dact4
__global__ void pn_fc_backward_w_kernel(const float* __restrict__ act4, const float* __restrict__ dy, int B,
                                        float* __restrict__ dwf, float* __restrict__ dbf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * 257) return;
    const int o = i / 257, j = i % 257;
    float s = 0.0f;
    if (j == 256) {
        for (int b = 0; b < B; ++b) s += dy[b * 3 + o];
        dbf[o] = s;
    } else {
        for (int b = 0; b < B; ++b) s = fmaf(dy[b * 3 + o], act4[(b * 8 + (j & 7)) * 32 + (j >> 3)], s);
        dwf[o * 256 + j] = s;
    }
}

__global__ void pn_fc_backward_x_kernel(const float* __restrict__ act4, const float* __restrict__ dy,
                                        const float* __restrict__ wf, int B, float* __restrict__ G4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * 256) return;
    const int b = i >> 8, p = (i >> 5) & 7, c = i & 31;
    const int j = c * 8 + p;
    const float g = dy[b * 3] * wf[j] + dy[b * 3 + 1] * wf[256 + j] + dy[b * 3 + 2] * wf[512 + j];
    G4[i] = act4[i] > 0.0f ? g : 0.0f;
}

__global__ void pn_to_bf16_kernel(const float* __restrict__ p, int n, uint16_t* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(p[i]);
}

__global__ void pn_adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                               float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                               float c2, float wd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi * c1) / (sqrtf(vi * c2) + eps) + wd * p[i]);
}

// ---------------------------------------------------------------------------
// fused inference
// ---------------------------------------------------------------------------
// MFMA fragments of the conv weights (BF16), one 16-byte group per lane:
// frag[step][co_tile][lane] holds W[co][ci][tap] for co = 16 co_tile + lane % 16
// and k = 32 step + 8 (lane / 16) + j, j < 8, with (tap, ci) = kmap(layer, k).
__device__ __forceinline__ void pn_kmap(int layer, int k, int& tap, int& ci)
{
    if (layer == 0) { tap = k >> 1; ci = k & 1; }                              // k = 2 tap + ci
    else if (layer == 1) { tap = k >> 4; ci = k & 15; }                        // k = 16 tap + ci
    else { tap = k >> 5; ci = k & 31; }                                        // k = 32 tap + ci
}

__global__ void pn_pack_frags_kernel(const float* __restrict__ params, uint16_t* __restrict__ frags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= PN_FRAG_TOTAL) return;
    int layer = 3;
    while (layer > 0 && i < pn_frag_off(layer)) --layer;
    const int r = i - pn_frag_off(layer);
    const int j = r & 7, lane = (r >> 3) & 63, ct = (r >> 9) % pn_co_tiles(layer), step = (r >> 9) / pn_co_tiles(layer);
    const int co = ct * 16 + (lane & 15), k = step * 32 + 8 * (lane >> 4) + j;
    int tap, ci;
    pn_kmap(layer, k, tap, ci);
    float w = 0.0f;
    if (tap < 27 && ci < pn_ci(layer)) w = params[pn_w_off(layer) + (co * pn_ci(layer) + ci) * 27 + tap];
    frags[i] = f2bf(w);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 as_bf(u32x4 v)
{
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ float relu_b(float v, float b) { return fmaxf(v + b, 0.0f); }

// One conv layer with 16-row position tiles.  act_in: BF16 channel-last
// [Din^3][CI]; act_out: BF16 [Do^3][32].  Wave w takes position tile w and
// both output-channel tiles.
template <int LAYER, int CI, int DIN, int S>
__device__ __forceinline__ void pn_conv_tile(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                             const uint16_t* __restrict__ frags, const float* __restrict__ bias,
                                             int tile, int lane)
{
    constexpr int DO = DIN / S;
    constexpr int STEPS = (CI * 27 + 31) / 32;
    const int m = lane & 15, g = lane >> 4;
    const int pos = tile * 16 + m;
    const int oz = pos / (DO * DO), oy = (pos / DO) % DO, ox = pos % DO;
    f32x4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
    const u32x4* fr = reinterpret_cast<const u32x4*>(frags);
#pragma unroll 3
    for (int st = 0; st < STEPS; ++st) {
        int tap, ci;
        pn_kmap(LAYER, st * 32 + 8 * g, tap, ci);
        const int iz = oz * S + tap / 9 - 1, iy = oy * S + (tap / 3) % 3 - 1, ix = ox * S + tap % 3 - 1;
        u32x4 a = {0u, 0u, 0u, 0u};
        if (tap < 27 && iz >= 0 && iy >= 0 && ix >= 0 && iz < DIN && iy < DIN && ix < DIN)
            a = *reinterpret_cast<const u32x4*>(in + ((iz * DIN + iy) * DIN + ix) * CI + ci);
        const u32x4 b0 = fr[(st * 2 + 0) * 64 + lane];
        const u32x4 b1 = fr[(st * 2 + 1) * 64 + lane];
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(b0), acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(b1), acc1, 0, 0, 0);
    }
    // D layout: lane holds rows 4 (lane / 16) + r, column lane % 16
    const int co = lane & 15;
    const float bb0 = bias[co], bb1 = bias[16 + co];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int p = tile * 16 + 4 * g + r;
        if (p < DO * DO * DO) {
            out[p * 32 + co] = f2bf(relu_b(acc0[r], bb0));
            out[p * 32 + 16 + co] = f2bf(relu_b(acc1[r], bb1));
        }
    }
}

// pred[i][3] for blocks ids[i] (ids == nullptr: i); centre (may be null):
// clamp(rint(pred), -8, 8).
__global__ __launch_bounds__(256) void pn_infer_kernel(const float* __restrict__ tpl, const float* __restrict__ src,
                                                       int nx, int ny, int nz, int stride,
                                                       const int* __restrict__ ids,
                                                       const float* __restrict__ params,
                                                       const uint16_t* __restrict__ frags, int mode,
                                                       const float* __restrict__ ptpl, const float* __restrict__ psrc,
                                                       int pnx, int pny, int pvec,
                                                       float* __restrict__ pred, int* __restrict__ centre)
{
    __shared__ __attribute__((aligned(16))) uint32_t xin[4096];  // (template, source) BF16 pairs, [z][y][x]
    __shared__ __attribute__((aligned(16))) uint16_t a1[512 * 16];
    __shared__ __attribute__((aligned(16))) uint16_t a2[64 * 32];
    __shared__ __attribute__((aligned(16))) uint16_t a3[64 * 32];
    __shared__ float red[4];
    __shared__ float fc[3][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int b = ids ? ids[blockIdx.x] : (int)blockIdx.x;
    const int gx = (nx - 16) / stride + 1, gy = (ny - 16) / stride + 1;
    int bx, by, bz;
    block_xyz(b, gx, gy, stride, bx, by, bz);

    // input pair, normalised, as BF16 (template in the low half)
    {
        float v[16], w[16];
        if (ptpl) {
            load_norm_pooled(ptpl, pnx, pny, bx >> 1, by >> 1, bz >> 1, pvec != 0, v, red);
            load_norm_pooled(psrc, pnx, pny, bx >> 1, by >> 1, bz >> 1, pvec != 0, w, red);
        } else {
            load_norm_row(tpl, nx, ny, nz, bx, by, bz, 0, mode, v, red);
            load_norm_row(src, nx, ny, nz, bx, by, bz, 0, mode, w, red);
        }
#pragma unroll
        for (int x = 0; x < 16; ++x) xin[tid * 16 + x] = (uint32_t)f2bf(v[x]) | ((uint32_t)f2bf(w[x]) << 16);
    }
    __syncthreads();

    // L1: 32 position tiles of 16, K = 54 (2 k-steps), 16 output channels
    {
        const u32x4* fr = reinterpret_cast<const u32x4*>(frags + pn_frag_off(0));
        const u32x4 w0 = fr[lane], w1 = fr[64 + lane];
        const int m = lane & 15, g = lane >> 4;
        const float bias = params[pn_b_off(0) + m];
        for (int tile = wv; tile < 32; tile += 4) {
            const int pos = tile * 16 + m;
            const int oz = pos >> 6, oy = (pos >> 3) & 7, ox = pos & 7;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int st = 0; st < 2; ++st) {
                u32x4 a;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int tap = 16 * st + 4 * g + q;
                    const int iz = 2 * oz + tap / 9 - 1, iy = 2 * oy + (tap / 3) % 3 - 1, ix = 2 * ox + tap % 3 - 1;
                    uint32_t e = 0u;
                    if (tap < 27 && iz >= 0 && iy >= 0 && ix >= 0 && iz < 16 && iy < 16 && ix < 16)
                        e = xin[(iz * 16 + iy) * 16 + ix];
                    a[q] = e;
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(st ? w1 : w0), acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) a1[(tile * 16 + 4 * g + r) * 16 + m] = f2bf(relu_b(acc[r], bias));
        }
    }
    __syncthreads();
    // L2: 4 position tiles (one per wave), K = 432 (14 k-steps)
    pn_conv_tile<1, 16, 8, 2>(a1, a2, frags + pn_frag_off(1), params + pn_b_off(1), wv, lane);
    __syncthreads();
    // L3: 4 position tiles, K = 864 (27 k-steps)
    pn_conv_tile<2, 32, 4, 1>(a2, a3, frags + pn_frag_off(2), params + pn_b_off(2), wv, lane);
    __syncthreads();
    // L4: 8 positions (one half-empty tile); the 27 k-steps are split over the waves
    float* part = reinterpret_cast<float*>(a1);  // [4 waves][8 positions][32 channels]
    {
        const u32x4* fr = reinterpret_cast<const u32x4*>(frags + pn_frag_off(3));
        const int m = lane & 15, g = lane >> 4;
        const int oz = m >> 2, oy = (m >> 1) & 1, ox = m & 1;
        f32x4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int st = wv; st < 27; st += 4) {
            const int tap = st, ci = 8 * g;
            const int iz = 2 * oz + tap / 9 - 1, iy = 2 * oy + (tap / 3) % 3 - 1, ix = 2 * ox + tap % 3 - 1;
            u32x4 a = {0u, 0u, 0u, 0u};
            if (m < 8 && iz >= 0 && iy >= 0 && ix >= 0 && iz < 4 && iy < 4 && ix < 4)
                a = *reinterpret_cast<const u32x4*>(a3 + ((iz * 4 + iy) * 4 + ix) * 32 + ci);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(fr[(st * 2) * 64 + lane]), acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(fr[(st * 2 + 1) * 64 + lane]), acc1, 0, 0, 0);
        }
        const int co = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int p = 4 * g + r;
            if (p < 8) {
                part[(wv * 8 + p) * 32 + co] = acc0[r];
                part[(wv * 8 + p) * 32 + 16 + co] = acc1[r];
            }
        }
    }
    __syncthreads();
    // L4 bias + ReLU, then FC 256 -> 3 in FP32: thread t holds input j = t = c * 8 + p
    {
        const int c = tid >> 3, p = tid & 7;
        const int e = p * 32 + c;
        const float s = ((part[e] + part[256 + e]) + (part[512 + e] + part[768 + e]));
        const float a4 = relu_b(s, params[pn_b_off(3) + c]);
        const float* wf = params + PN_WF_OFF;
        float y0 = a4 * wf[tid], y1 = a4 * wf[256 + tid], y2 = a4 * wf[512 + tid];
        for (int o = 32; o > 0; o >>= 1) {
            y0 += __shfl_xor(y0, o, 64);
            y1 += __shfl_xor(y1, o, 64);
            y2 += __shfl_xor(y2, o, 64);
        }
        if (lane == 0) { fc[0][wv] = y0; fc[1][wv] = y1; fc[2][wv] = y2; }
        __syncthreads();
        if (tid < 3) {
            const float y = params[PN_BF_OFF + tid] + ((fc[tid][0] + fc[tid][1]) + (fc[tid][2] + fc[tid][3]));
            pred[3 * blockIdx.x + tid] = y;
            if (centre) centre[3 * blockIdx.x + tid] = (int)fminf(fmaxf(rintf(y), -8.0f), 8.0f);
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {
constexpr int PN_DIN[4] = {16, 8, 4, 4};
constexpr int PN_S[4] = {2, 2, 1, 2};
constexpr int pn_do(int l) { return PN_DIN[l] / PN_S[l]; }
constexpr int pn_p(int l) { return pn_do(l) * pn_do(l) * pn_do(l); }

unsigned grid_for(long n, int bs = 256) { return (unsigned)std::min<long>((n + bs - 1) / bs, 1 << 20); }

template <typename TA, typename TB, int EPI>
bool gemm(int M, int N, int K, const TA* A, long sam, long sak, const TB* B, long sbk, long sbn, float* C, long scm,
          long scn, int splits, long zstride, const float* bias, hipStream_t st)
{
    const int kchunk = ((K + splits - 1) / splits + 31) / 32 * 32;
    splits = (K + kchunk - 1) / kchunk;
    dim3 grid((M + 63) / 64, (N + 63) / 64, splits);
    if constexpr (std::is_same_v<TA, float> && std::is_same_v<TB, float>)
        pn_gemm_f32_kernel<EPI><<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, kchunk, zstride,
                                                      bias);
    else
        pn_gemm_kernel<TA, TB, EPI><<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, kchunk,
                                                          zstride, bias);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}
}  // namespace

bool PnetTrainer::reserve(int B)
{
    if (B <= B_) return true;
    size_t col = 0, act = 0, dcol = 0, part = 0;
    for (int l = 0; l < 4; ++l) {
        const size_t rows = (size_t)B * pn_p(l), K = (size_t)pn_ci(l) * 27;
        col += rows * K;
        act += rows * pn_co(l);
        if (l > 0) dcol = std::max(dcol, rows * K);
        const int splits = (int)((rows + PN_SPLIT_ROWS - 1) / PN_SPLIT_ROWS);
        part = std::max(part, (size_t)splits * pn_co(l) * K);
    }
    const size_t bytes = col * 4 + 2 * act * 4 + dcol * 4 + part * 4 + (size_t)B * 6 * 4 + PN_PARAMS * 2 + 32 * 256;
    if (!ws_.ensure(bytes)) return false;
    char* p = static_cast<char*>(ws_.ptr);
    auto take = [&](size_t n) {
        char* r = p;
        p += (n + 255) / 256 * 256;
        return r;
    };
    for (int l = 0; l < 4; ++l) {
        const size_t rows = (size_t)B * pn_p(l), K = (size_t)pn_ci(l) * 27;
        col_[l] = take(rows * K * 4);  // FP32 in verification mode, BF16 otherwise
    }
    for (int l = 0; l < 4; ++l) act_[l] = reinterpret_cast<float*>(take((size_t)B * pn_p(l) * pn_co(l) * 4));
    for (int l = 0; l < 4; ++l) g_[l] = reinterpret_cast<float*>(take((size_t)B * pn_p(l) * pn_co(l) * 4));
    dcol_ = reinterpret_cast<float*>(take(dcol * 4));
    part_ = reinterpret_cast<float*>(take(part * 4));
    y_ = reinterpret_cast<float*>(take((size_t)B * 3 * 4));
    dy_ = reinterpret_cast<float*>(take((size_t)B * 3 * 4));
    pbf_ = reinterpret_cast<uint16_t*>(take(PN_PARAMS * 2));
    if ((size_t)(p - static_cast<char*>(ws_.ptr)) > ws_.bytes) CUT_FAIL("pnet: workspace layout overflow");
    B_ = B;
    return true;
}

bool PnetTrainer::forward(const float* d_params, const float* d_x, int B, float* d_y, hipStream_t st)
{
    if (B <= 0) CUT_FAIL("pnet: empty batch");
    if (!reserve(B)) return false;
    if (!fp32_) pn_to_bf16_kernel<<<grid_for(PN_PARAMS), 256, 0, st>>>(d_params, PN_PARAMS, pbf_);
    for (int l = 0; l < 4; ++l) {
        const int ci = pn_ci(l), co = pn_co(l), D = PN_DIN[l], K = ci * 27;
        const long rows = (long)B * pn_p(l);
        long sb, sc, sp;
        const float* X;
        if (l == 0) { X = d_x; sb = 8192; sc = 4096; sp = 1; }
        else { X = act_[l - 1]; sb = (long)D * D * D * ci; sc = 1; sp = ci; }
        const long total = rows * K;
        bool ok;
        if (fp32_) {
            float* col = static_cast<float*>(col_[l]);
            pn_im2col_kernel<float><<<grid_for(total), 256, 0, st>>>(X, sb, sc, sp, D, PN_S[l], pn_do(l), ci, total, col);
            ok = gemm<float, float, 1>((int)rows, co, K, col, K, 1, d_params + pn_w_off(l), 1, K, act_[l], co, 1, 1, 0,
                                       d_params + pn_b_off(l), st);
        } else {
            uint16_t* col = static_cast<uint16_t*>(col_[l]);
            pn_im2col_kernel<uint16_t><<<grid_for(total), 256, 0, st>>>(X, sb, sc, sp, D, PN_S[l], pn_do(l), ci, total,
                                                                        col);
            ok = gemm<uint16_t, uint16_t, 1>((int)rows, co, K, col, K, 1, pbf_ + pn_w_off(l), 1, K, act_[l], co, 1, 1,
                                             0, d_params + pn_b_off(l), st);
        }
        if (!ok) return false;
    }
    pn_fc_forward_kernel<<<grid_for(B * 3), 256, 0, st>>>(act_[3], d_params + PN_WF_OFF, d_params + PN_BF_OFF, B,
                                                          d_y ? d_y : y_);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool PnetTrainer::loss_grad(const float* d_params, const float* d_x, const float* d_t, int B, float* d_grad,
                            float* d_loss, hipStream_t st)
{
    if (!forward(d_params, d_x, B, nullptr, st)) return false;
    pn_loss_kernel<<<1, 256, 0, st>>>(y_, d_t, B, dy_, d_loss);
    pn_fc_backward_w_kernel<<<grid_for(3 * 257), 256, 0, st>>>(act_[3], dy_, B, d_grad + PN_WF_OFF,
                                                               d_grad + PN_BF_OFF);
    pn_fc_backward_x_kernel<<<grid_for((long)B * 256), 256, 0, st>>>(act_[3], dy_, d_params + PN_WF_OFF, B, g_[3]);
    CUT_HIP_CHECK(hipGetLastError());
    for (int l = 3; l >= 0; --l) {
        const int ci = pn_ci(l), co = pn_co(l), K = ci * 27;
        const long rows = (long)B * pn_p(l);
        // dW[co][k] = sum_rows G[row][co] col[row][k]
        const int splits = (int)((rows + PN_SPLIT_ROWS - 1) / PN_SPLIT_ROWS);
        const bool ok = fp32_ ? gemm<float, float, 0>(co, K, (int)rows, g_[l], 1, co, static_cast<float*>(col_[l]), K,
                                                      1, part_, K, 1, splits, (long)co * K, nullptr, st)
                              : gemm<float, uint16_t, 0>(co, K, (int)rows, g_[l], 1, co,
                                                         static_cast<uint16_t*>(col_[l]), K, 1, part_, K, 1, splits,
                                                         (long)co * K, nullptr, st);
        if (!ok) return false;
        const int kchunk = (((int)rows + splits - 1) / splits + 31) / 32 * 32;
        const int used = (int)((rows + kchunk - 1) / kchunk);
        pn_splitk_reduce_kernel<<<grid_for((long)co * K), 256, 0, st>>>(part_, used, (long)co * K, (long)co * K,
                                                                         d_grad + pn_w_off(l));
        pn_bias_grad_kernel<<<co, 256, 0, st>>>(g_[l], rows, co, d_grad + pn_b_off(l));
        CUT_HIP_CHECK(hipGetLastError());
        if (l == 0) break;
        // dcol[row][k] = sum_co G[row][co] W[co][k]; then col2im into the previous layer
        const bool ok2 = fp32_ ? gemm<float, float, 0>((int)rows, K, co, g_[l], co, 1, d_params + pn_w_off(l), K, 1,
                                                       dcol_, K, 1, 1, 0, nullptr, st)
                               : gemm<float, uint16_t, 0>((int)rows, K, co, g_[l], co, 1, pbf_ + pn_w_off(l), K, 1,
                                                          dcol_, K, 1, 1, 0, nullptr, st);
        if (!ok2) return false;
        const int D = PN_DIN[l];
        const long total = (long)B * D * D * D * ci;
        pn_col2im_kernel<<<grid_for(total), 256, 0, st>>>(dcol_, D, PN_S[l], pn_do(l), ci, total, act_[l - 1],
                                                          g_[l - 1]);
        CUT_HIP_CHECK(hipGetLastError());
    }
    return true;
}

bool pnet_adam(float* d_p, const float* d_g, float* d_m, float* d_v, int n, float lr, float b1, float b2, float eps,
               int step, float wd, hipStream_t st)
{
    if (step < 1) CUT_FAIL("pnet: Adam steps count from 1");
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)step));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)step));
    pn_adam_kernel<<<grid_for(n), 256, 0, st>>>(d_p, d_g, d_m, d_v, n, lr, b1, b2, eps, c1, c2, wd);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool pnet_gather(const uint64_t* d_vols, int nx, int ny, int nz, int stride, const int* d_desc, int n, int mode,
                 float* d_out, hipStream_t st)
{
    if (n <= 0) return true;
    pn_gather_kernel<<<n, 256, 0, st>>>(d_vols, nx, ny, nz, stride, d_desc, mode, d_out);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool pnet_pack(const float* d_params, uint16_t* d_frags, hipStream_t st)
{
    pn_pack_frags_kernel<<<grid_for(PN_FRAG_TOTAL), 256, 0, st>>>(d_params, d_frags);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

size_t pnet_pool_floats(int nx, int ny, int nz) { return (size_t)(nx / 2 + 8) * (ny / 2 + 8) * (nz / 2 + 8); }

bool pnet_infer(const float* d_params, const uint16_t* d_frags, const float* d_tpl, const float* d_src, int nx, int ny,
                int nz, int stride, const int* d_ids, int n, int mode, float* d_pool, float* d_pred, int* d_centre,
                hipStream_t st)
{
    if (n <= 0) return true;
    const float* pt = nullptr;
    const float* ps = nullptr;
    const int pnx = nx / 2 + 8, pny = ny / 2 + 8, pnz = nz / 2 + 8;
    if (mode == 1 && d_pool && stride % 2 == 0) {
        // pool both volumes once; every block then reads a 16^3 window
        const size_t np = (size_t)pnx * pny * pnz;
        float* P = d_pool;
        pn_pool_kernel<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(d_tpl, nx, ny, nz, pnx, pny, pnz, P);
        pn_pool_kernel<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(d_src, nx, ny, nz, pnx, pny, pnz, P + np);
        pt = P;
        ps = P + np;
    }
    const int pvec = (pnx % 4 == 0 && stride % 8 == 0) ? 1 : 0;
    pn_infer_kernel<<<n, 256, 0, st>>>(d_tpl, d_src, nx, ny, nz, stride, d_ids, d_params, d_frags, mode, pt, ps, pnx,
                                       pny, pvec, d_pred, d_centre);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

}  // namespace cut
