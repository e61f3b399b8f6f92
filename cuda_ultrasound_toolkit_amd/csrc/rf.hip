// RF conversion and Hadamard decode kernels for gfx950.
//
// convert: cuda_toolkit/src/rf_processing/data_conversion/data_conversion_kernels.cuh:25-41
// decode:  cuda_toolkit/src/rf_processing/hadamard/hadamard_decoder.cpp:49-78 (cuBLAS SGEMM)
//
// The decode is a GEMM per channel c: Out_c[j][s] = scale * sum_k H[j][k] X_c[k][s]
// with X_c[k][s] = raw[map[c]][k*S + s].  A workgroup owns one channel and 64
// consecutive samples, all output rows.  The int16 tile is read once from HBM
// (16-byte loads), split into bytes and transposed into LDS so that each lane
// holds 16 consecutive k for one sample column, the layout MFMA wants for its
// B operand.  Two MFMA paths:
//   mode 1 (int8): x = 256*hi + (lo_u - 128) + 128 with hi, lo_u-128 in int8;
//                  sum_k H x = 256*(H.hi) + H.(lo-128) + 128*rowsum(H), all in
//                  int32, exact for every int16 input and T <= 1024.
//   mode 0 (fp32): v_mfma_f32_32x32x2_f32 on the converted floats; exact for
//                  int16 input and power-of-two T (every partial sum is an
//                  integer below 2^24), so both modes agree bit for bit.
#include "common.hpp"

namespace cut {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// convert (no decode)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void convert_kernel(const T* __restrict__ raw, const int16_t* __restrict__ map,
                                                      size_t R0, int C, int S, float* __restrict__ out,
                                                      int out_stride)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    const int t = blockIdx.z;
    if (s >= S) return;
    const T* src = raw + (size_t)(uint16_t)map[c] * R0 + (size_t)t * S;
    out[(((size_t)t * C + c) * S + s) * out_stride] = static_cast<float>(src[s]);
}

bool launch_convert(const void* raw, int type, const int16_t* d_map, size_t R0, int T, int C, int S,
                    float* out, int out_stride, hipStream_t st)
{
    dim3 grid((S + 255) / 256, C, T);
    if (type == TYPE_I16)
        convert_kernel<int16_t><<<grid, 256, 0, st>>>((const int16_t*)raw, d_map, R0, C, S, out, out_stride);
    else if (type == TYPE_F32)
        convert_kernel<float><<<grid, 256, 0, st>>>((const float*)raw, d_map, R0, C, S, out, out_stride);
    else
        CUT_FAIL("data converter: unsupported input data type " << type);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

// ---------------------------------------------------------------------------
// bandpass input: convert + demodulation + centred FIR in one pass
// (data_conversion_kernels.cuh:14-21, :46-64, :66-102; data_converter.cu:64-131)
// ---------------------------------------------------------------------------
// A workgroup owns BP_SEG outputs of one (t, c) row.  It demodulates the
// input samples those outputs need (the segment plus the filter's reach on
// both sides) into LDS, then each output is the FIR over the taps that fall
// inside the row, divided by their count (taps are finite).  The phase s fd/fs is reduced to
// revolutions in FP64 and evaluated with v_sin/v_cos.  LDS holds the samples
// from org = s0 - L/2 on, zeros where the window leaves the row (and past
// its end, as far as the products below read), so every output sums all L
// taps (the zeros add nothing).
//
// The FIR runs on the matrix cores as a banded (Toeplitz) product.  Outputs
// go in blocks of 16 consecutive samples; output 16 b + m of the segment is
// sum_k T[m][k] X[k][b] with X[k][b] = v[16 b + k] (the window from the
// block's first sample on) and T[m][k] = hr[k - m], zero unless
// 0 <= k - m < L (hr = the taps in application order, hr[i] = h[L-1-i]).
// A wave computes 16 blocks (256 outputs) with v_mfma_f32_16x16x4_f32 over
// K = L + 15 rounded up to 16 (80 for 64 taps: 80 % of the products are
// taps), the I and Q parts as two products sharing the tap operand, the
// operands of four K steps read from LDS ahead of their eight products.  The
// FP32 matrix rate is twice the non-packed FMA rate; every product is the
// same single-rounded FP32 product as an FMA's, summed in another order.
// Sample j sits at LDS entry j + j/16, so the 16 blocks a B operand spans
// (16 samples apart) fall on different banks.  Workgroups go transmit-fastest
// (t = blockIdx.x % T): consecutive workgroups read consecutive stretches of
// a channel's raw row.
constexpr int BP_SEG = 1024;
// LDS (sized for the filter at launch): the window's entries, samples
// 0 .. BP_SEG + L + 13 (the last read is BP_SEG + L + 13 at most), then the
// taps at 16 .. 16 + L - 1 with zeros around (reads 1 .. L + 45)
__host__ __device__ inline int bp_window(int L) { return BP_SEG + (L > 0 ? L + 14 : 0); }
__host__ __device__ inline int bp_entries(int n) { return n + n / 16 + 1; }
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
__global__ __launch_bounds__(256) void convert_demod_kernel(const T* __restrict__ raw, const int16_t* __restrict__ map,
                                                            size_t R0, int C, int S, double fd_fs,
                                                            const float* __restrict__ h, int L,
                                                            float2* __restrict__ out)
{
    static_assert(BP_SEG == 4 * 256, "four waves of 16 blocks of 16 outputs");
    extern __shared__ float2 bp_lds[];
    float2* const vt = bp_lds;
    float* const hp = reinterpret_cast<float*>(bp_lds + bp_entries(bp_window(L)));
    const int nT = gridDim.x / C;
    const int t = blockIdx.x % nT, c = blockIdx.x / nT;
    const int row = t * C + c;               // output row
    const int s0 = blockIdx.y * BP_SEG;
    const int half = L > 0 ? L / 2 : 0;
    const int org = s0 - half;
    const int nwin = BP_SEG + (L > 0 ? L - 1 : 0);  // window samples org .. org + nwin - 1
    const int nfill = bp_window(L);                  // samples written (zeros past nwin)
    const int Kp = (L + 30) & ~15;                    // K = L + 15 rounded up to 16
    const int tid = threadIdx.x;
    // sample org + j of the row (j >= 0)
    auto v = [&](int j) -> float2& { return vt[j + (j >> 4)]; };
    const T* r = raw + (size_t)(uint16_t)map[c] * R0 + 2 * (size_t)t * S;
    // every thread issues all its loads first
    constexpr int FILL = (BP_SEG + MAX_FILTER_LENGTH + 14 + 255) / 256;
    struct alignas(2 * sizeof(T)) IQPair { T i, q; };
    // (I, Q) as one load when the row start allows it (uniform per workgroup)
    const bool paired = (reinterpret_cast<uintptr_t>(r) & (2 * sizeof(T) - 1)) == 0;
    // the taps (entries 16 .. 16 + L - 1 of hp) load with the samples
    constexpr int HFILL = (MAX_FILTER_LENGTH + 46 + 255) / 256;
    float hv[HFILL];
#pragma unroll
    for (int u = 0; u < HFILL; ++u) {
        const int i = tid + 256 * u;
        hv[u] = (L > 0 && i >= 16 && i < 16 + L) ? h[L - 1 - (i - 16)] : 0.0f;
    }
    float Iv[FILL], Qv[FILL];
#pragma unroll
    for (int u = 0; u < FILL; ++u) {
        const int js = org + tid + 256 * u;  // sample index in the row
        Iv[u] = Qv[u] = 0.0f;
        if (tid + 256 * u < nwin && js >= 0 && js < S) {
            if (paired) {
                const IQPair w = reinterpret_cast<const IQPair*>(r)[js];
                Iv[u] = static_cast<float>(w.i);
                Qv[u] = static_cast<float>(w.q);
            } else {
                Iv[u] = static_cast<float>(r[2 * (size_t)js]);
                Qv[u] = static_cast<float>(r[2 * (size_t)js + 1]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < FILL; ++u) {
        const int j = tid + 256 * u, js = org + j;
        if (j < nfill) {
            float2 y = make_float2(0.0f, 0.0f);
            if (j < nwin && js >= 0 && js < S) {
                const double ph = (double)js * fd_fs;
                const float f = (float)(ph - floor(ph));
                const float cs = __builtin_amdgcn_cosf(f), sn = __builtin_amdgcn_sinf(f);
                const float I = Iv[u], Q = Qv[u];
                y = make_float2(I * cs - Q * sn, -I * sn - Q * cs);  // (I - iQ) exp(-2 pi i f)
            }
            v(j) = y;
        }
    }
    if (L > 0)
#pragma unroll
        for (int u = 0; u < HFILL; ++u)
            if (tid + 256 * u < L + 46) hp[tid + 256 * u] = hv[u];
    __syncthreads();
    float2* const orow = out + (size_t)row * S;
    if (L <= 0) {
        const int sb = s0 + 4 * tid;
        for (int q = 0; q < 4 && sb + q < S; ++q) orow[sb + q] = v(4 * tid + q);
        return;
    }
    const int lane = tid & 63;
    if (s0 + 256 * (tid >> 6) >= S) return;         // the wave's outputs are past the row (uniform)
    const int b = (tid >> 6) * 16 + (lane & 15);    // the lane's block: B and D column
    const int kr = lane >> 4;                       // the lane's K row of A and B
    const float* hl = hp + 16 + kr - (lane & 15);   // A[m][kr] = T[m][k0 + kr], m = lane & 15
    int j = 16 * b + kr;                            // B[kr][b] = X[k0 + kr][b]
    v4f yi = {0.0f, 0.0f, 0.0f, 0.0f}, yq = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < Kp; k0 += 16, j += 16) {
        float a[4];
        float2 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = hl[k0 + 4 * u];
            x[u] = v(j + 4 * u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            yi = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], x[u].x, yi, 0, 0, 0);
            yq = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], x[u].y, yq, 0, 0, 0);
        }
    }
    // D[m][b], m = 4 kr + q: the lane holds outputs 16 b + 4 kr + q, q = 0..3
    const int sb = s0 + 16 * b + 4 * kr;
    if (sb >= S) return;
    const int nout = min(4, S - sb);
    // divided by the number of taps inside the row: i with 0 <= s - half + i < S
    float2 y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = sb + q;
        const int ia = max(0, half - s), ib = min(L - 1, S - 1 - s + half);
        const float inv = 1.0f / (float)max(1, ib - ia + 1);
        y[q] = make_float2(yi[q] * inv, yq[q] * inv);
    }
    if (nout == 4 && (reinterpret_cast<uintptr_t>(orow + sb) & 15) == 0) {
        float4* o4 = reinterpret_cast<float4*>(orow + sb);
        o4[0] = make_float4(y[0].x, y[0].y, y[1].x, y[1].y);
        o4[1] = make_float4(y[2].x, y[2].y, y[3].x, y[3].y);
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < nout) orow[sb + q] = y[q];
    }
}

bool launch_convert_demod(const void* raw, int type, const int16_t* d_map, size_t R0, int T, int C, int S,
                          double fd_fs, const float* d_h, int L, float2* out, hipStream_t st)
{
    if (L < 0 || L > MAX_FILTER_LENGTH || L > S) CUT_FAIL("bandpass filter length " << L << " not in [0, min(S, " << MAX_FILTER_LENGTH << ")]");
    if (L > 0 && !d_h) CUT_FAIL("bandpass filter taps missing");
    dim3 grid(T * C, (S + BP_SEG - 1) / BP_SEG);
    const size_t lds = (size_t)bp_entries(bp_window(L)) * sizeof(float2) + (L > 0 ? (size_t)(L + 46) * sizeof(float) : 0);
    if (type == TYPE_I16)
        convert_demod_kernel<int16_t><<<grid, 256, lds, st>>>((const int16_t*)raw, d_map, R0, C, S, fd_fs, d_h, L, out);
    else if (type == TYPE_F32)
        convert_demod_kernel<float><<<grid, 256, lds, st>>>((const float*)raw, d_map, R0, C, S, fd_fs, d_h, L, out);
    else
        CUT_FAIL("Data converter: Unsupported input data type.");
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

// ---------------------------------------------------------------------------
// decode
// ---------------------------------------------------------------------------
constexpr int DEC_NT = 64;      // samples per workgroup
constexpr int DEC_KC = 128;     // k rows staged per chunk
constexpr int DEC_ROWB = 144;   // LDS bytes per sample column of a byte tile (16-byte padded)
constexpr int DEC_ROWF = 65;    // LDS floats per k row of the fp32 tile (bank padded)
constexpr int DEC_STREAM_WG_PER_CU = 3;  // resident workgroups of the streaming int8 decode (<= 168 VGPRs)

// Load 8 consecutive samples (k row, sample s) as int32 values.
template <typename T, bool VEC>
__device__ __forceinline__ void load8(const T* __restrict__ row, int s, int S, int v[8])
{
    if constexpr (VEC && sizeof(T) == 2) {
        if (s + 8 <= S) {
            v4i q = *(const v4i*)(row + s);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[2 * e] = (int)(int16_t)(q[e] & 0xffff);
                v[2 * e + 1] = (int)(int16_t)((unsigned)q[e] >> 16);
            }
            return;
        }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (s + e < S) ? (int)row[s + e] : 0;
}

template <bool VEC, int OUT_STRIDE>
__global__ __launch_bounds__(256) void decode_i8_kernel(const int16_t* __restrict__ raw,
                                                        const int16_t* __restrict__ map, size_t R0, int T,
                                                        int C, int S, int Kp, const int8_t* __restrict__ h8,
                                                        const int* __restrict__ rowsum, float scale,
                                                        float* __restrict__ out)
{
    __shared__ __attribute__((aligned(16))) int8_t bhi[DEC_NT * DEC_ROWB];
    __shared__ __attribute__((aligned(16))) int8_t blo[DEC_NT * DEC_ROWB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s0 = blockIdx.x * DEC_NT;
    const int c = blockIdx.y;
    const int mbase = blockIdx.z * 128 + wave * 32;
    const bool mvalid = mbase < Kp;
    const int16_t* src = raw + (size_t)(uint16_t)map[c] * R0;

    v16i acc_hi[2], acc_lo[2];
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int r = 0; r < 16; ++r) { acc_hi[n][r] = 0; acc_lo[n][r] = 0; }

    const int kq = tid >> 3, sg = tid & 7;
    for (int k0 = 0; k0 < Kp; k0 += DEC_KC) {
        const int kc = min(DEC_KC, Kp - k0);
        __syncthreads();
        if (kq * 4 < kc) {
            int v[4][8];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + kq * 4 + j;
                if (k < T) load8<int16_t, VEC>(src + (size_t)k * S, s0 + sg * 8, S, v[j]);
                else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[j][e] = 0;
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                unsigned hi = 0, lo = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const unsigned x = (unsigned)v[j][e];
                    hi |= ((x >> 8) & 0xffu) << (8 * j);
                    lo |= ((x & 0xffu) ^ 0x80u) << (8 * j);
                }
                const int off = (sg * 8 + e) * DEC_ROWB + kq * 4;
                *(unsigned*)&bhi[off] = hi;
                *(unsigned*)&blo[off] = lo;
            }
        }
        __syncthreads();
        if (mvalid) {
            const int8_t* hrow = h8 + (size_t)(mbase + (lane & 31)) * Kp + k0 + (lane >> 5) * 16;
            for (int kk = 0; kk < kc; kk += 32) {
                const v4i a = *(const v4i*)(hrow + kk);
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    const int off = (nt * 32 + (lane & 31)) * DEC_ROWB + kk + (lane >> 5) * 16;
                    const v4i bh = *(const v4i*)&bhi[off];
                    const v4i bl = *(const v4i*)&blo[off];
                    acc_hi[nt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bh, acc_hi[nt], 0, 0, 0);
                    acc_lo[nt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bl, acc_lo[nt], 0, 0, 0);
                }
            }
        }
    }
    if (!mvalid) return;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int s = s0 + nt * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = mbase + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (m < T && s < S) {
                const int total = 256 * acc_hi[nt][r] + acc_lo[nt][r] + 128 * rowsum[m];
                out[(((size_t)m * C + c) * S + s) * OUT_STRIDE] = (float)total * scale;
            }
        }
    }
}

// Streaming form of decode_i8_kernel for T <= 128 (one k chunk) and 16-byte
// aligned rows: a fixed grid of workgroups walks the (channel, NT-sample)
// tiles in order (tile = blockIdx.x + i * gridDim.x), keeps H in registers,
// and issues the next tile's int16 loads before the current tile's MFMA and
// stores, so HBM reads stay in flight while a workgroup computes and writes.
// Staging: thread (kq, sg) loads samples sg*8 .. sg*8+7 of rows kq*RPT ..
// kq*RPT+RPT-1 (16-byte loads; NT/8 threads cover a row segment of NT
// samples).  Results are identical to decode_i8_kernel (same products, same
// integer sums).
// Rows that do not start on 16 bytes (S not a multiple of 8: row k of a
// channel starts at sample k*S): the two aligned 16-byte chunks covering a
// thread's 8 samples are loaded through a buffer resource bounded to the
// channel's row (R0 samples from a 16-byte aligned start, R0 a multiple of 8,
// so a chunk is either inside the row or wholly past it and reads zeros), and
// the 8 samples are shifted out of the 16 where they are consumed, so the
// loads stay in flight during the previous tile's work.
constexpr int DEC_RSRC_WORD3 = 0x00020000;  // raw buffer, 32-bit format: byte-offset range check

// 8 int16 starting at sample sh (0..7) of the 16 in (lo, hi), as 4 packed pairs
__device__ __forceinline__ v4i shift_samples(v4i lo, v4i hi, unsigned sh)
{
    const unsigned q = sh >> 1, rb = (sh & 1u) * 2u;  // whole dwords, then 0 or 2 bytes
    // dword q + i of the 8 (lo then hi), i = 0..4, by a two-level select on q
    auto pick = [&](unsigned a0, unsigned a1, unsigned a2, unsigned a3) {
        const unsigned x = (q & 1u) ? a1 : a0;
        const unsigned y = (q & 1u) ? a3 : a2;
        return (q & 2u) ? y : x;
    };
    const unsigned d0 = lo[0], d1 = lo[1], d2 = lo[2], d3 = lo[3], d4 = hi[0], d5 = hi[1], d6 = hi[2],
                   d7 = hi[3];
    const unsigned e0 = pick(d0, d1, d2, d3), e1 = pick(d1, d2, d3, d4), e2 = pick(d2, d3, d4, d5),
                   e3 = pick(d3, d4, d5, d6), e4 = pick(d4, d5, d6, d7);
    return v4i{(int)__builtin_amdgcn_alignbyte(e1, e0, rb), (int)__builtin_amdgcn_alignbyte(e2, e1, rb),
               (int)__builtin_amdgcn_alignbyte(e3, e2, rb), (int)__builtin_amdgcn_alignbyte(e4, e3, rb)};
}

template <int OUT_STRIDE, int NT, bool ALIGNED>
__global__ __launch_bounds__(256, 3) void decode_i8_stream_kernel(const int16_t* __restrict__ raw,
                                                               const int16_t* __restrict__ map, size_t R0, int T,
                                                               int C, int S, int Kp, int nsx,
                                                               const int8_t* __restrict__ h8,
                                                               const int* __restrict__ rowsum, float scale,
                                                               float* __restrict__ out)
{
    constexpr int SG = NT / 8, KQ = 256 / SG, RPT = DEC_KC / KQ;
    static_assert(RPT % 4 == 0 && RPT * KQ == DEC_KC, "staging covers the k chunk in groups of 4 rows");
    // output tile [transmit][sample] (OTS floats per transmit row, padded) when
    // the output is contiguous (OUT_STRIDE 1): written from the MFMA layout,
    // then stored as float4 per lane, whole row segments per instruction
    constexpr int OTS = NT + 4;
    __shared__ __attribute__((aligned(16))) int8_t bhi[NT * DEC_ROWB];
    __shared__ __attribute__((aligned(16))) int8_t blo[NT * DEC_ROWB];
    __shared__ __attribute__((aligned(16))) float otile[OUT_STRIDE == 1 ? DEC_KC * OTS : 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mbase = wave * 32;
    const bool mvalid = mbase < Kp;
    const int ntiles = nsx * C;
    const int kq = tid / SG, sg = tid % SG;

    // A operand (H rows mbase.., 32 k per fragment) for the whole K
    v4i afr[DEC_KC / 32];
#pragma unroll
    for (int kk = 0; kk < DEC_KC / 32; ++kk) {
        afr[kk] = v4i{0, 0, 0, 0};
        if (mvalid && kk * 32 < Kp)
            afr[kk] = *(const v4i*)(h8 + (size_t)(mbase + (lane & 31)) * Kp + kk * 32 + (lane >> 5) * 16);
    }

    // rows kq*RPT + j, samples sg*8 .. sg*8+7 of a tile, as packed int16 pairs
    // (unaligned rows: the covering chunks cur, curh, shifted when consumed)
    v4i cur[RPT];
    v4i curh[ALIGNED ? 1 : RPT];
    auto load_tile = [&](int tile) {
        const int c = tile / nsx, sx = tile - c * nsx;
        const int16_t* src = raw + (size_t)(uint16_t)map[c] * R0;
        const int s = sx * NT + sg * 8;
        if constexpr (ALIGNED) {
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int k = kq * RPT + j;
                cur[j] = (k < T && s < S) ? *(const v4i*)(src + (size_t)k * S + s) : v4i{0, 0, 0, 0};
            }
        } else {
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<int16_t*>(src), 0, (int)(R0 * 2), DEC_RSRC_WORD3);
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int k = kq * RPT + j;
                cur[j] = v4i{0, 0, 0, 0};
                curh[j] = v4i{0, 0, 0, 0};
                if (k < T && s < S) {
                    const unsigned boff = (((unsigned)k * (unsigned)S + (unsigned)s) & ~7u) * 2u;
                    cur[j] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, boff, 0, 0));
                    curh[j] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, boff + 16u, 0, 0));
                }
            }
        }
    };
    int tile = blockIdx.x;
    if (tile < ntiles) load_tile(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        const int c = tile / nsx, s0 = (tile - c * nsx) * NT;
        __syncthreads();  // the previous tile's MFMA reads of LDS are done
        if constexpr (!ALIGNED) {
            const int s = s0 + sg * 8;
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int k = kq * RPT + j;
                cur[j] = shift_samples(cur[j], curh[j], ((unsigned)k * (unsigned)S + (unsigned)s) & 7u);
            }
        }
        if (kq * RPT < Kp) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
#pragma unroll
                for (int j0 = 0; j0 < RPT; j0 += 4) {
                    unsigned hi = 0, lo = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const unsigned x = ((unsigned)cur[j0 + j][e >> 1] >> (16 * (e & 1))) & 0xffffu;
                        hi |= ((x >> 8) & 0xffu) << (8 * j);
                        lo |= ((x & 0xffu) ^ 0x80u) << (8 * j);
                    }
                    const int off = (sg * 8 + e) * DEC_ROWB + kq * RPT + j0;
                    *(unsigned*)&bhi[off] = hi;
                    *(unsigned*)&blo[off] = lo;
                }
            }
        }
        const int next = tile + gridDim.x;
        if (next < ntiles) load_tile(next);
        __syncthreads();
        // one accumulator per column block: H.hi, times 256, then + H.lo
        // (|H.hi| <= 128 T, so 256 H.hi + H.lo stays far inside int32)
#pragma unroll
        for (int nt = 0; nt < NT / 32; ++nt) {
            if (!mvalid) break;
            v16i acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0;
            const int col = (nt * 32 + (lane & 31)) * DEC_ROWB + (lane >> 5) * 16;
#pragma unroll
            for (int kk = 0; kk < DEC_KC / 32; ++kk) {
                if (kk * 32 >= Kp) break;
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(afr[kk], *(const v4i*)&bhi[col + kk * 32], acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] *= 256;
#pragma unroll
            for (int kk = 0; kk < DEC_KC / 32; ++kk) {
                if (kk * 32 >= Kp) break;
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(afr[kk], *(const v4i*)&blo[col + kk * 32], acc, 0, 0, 0);
            }
            const int sl = nt * 32 + (lane & 31), s = s0 + sl;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = mbase + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const float v = (float)(acc[r] + 128 * rowsum[m < T ? m : 0]) * scale;
                if constexpr (OUT_STRIDE == 1) {
                    otile[m * OTS + sl] = v;
                } else if (m < T && s < S) {
                    out[(((size_t)m * C + c) * S + s) * OUT_STRIDE] = v;
                }
            }
        }
        if constexpr (OUT_STRIDE == 1) {
            __syncthreads();
            if constexpr (ALIGNED) {
                // S is a multiple of 8 here, so a float4 starting below S lies inside it
                constexpr int Q4 = NT / 4;
                for (int q = tid; q < T * Q4; q += 256) {
                    const int m = q / Q4, s = s0 + (q - m * Q4) * 4;
                    if (s < S)
                        *reinterpret_cast<float4*>(out + ((size_t)m * C + c) * S + s) =
                            *reinterpret_cast<const float4*>(&otile[m * OTS + (s - s0)]);
                }
            } else if (S % 4 == 0 && (uintptr_t)out % 16 == 0) {
                // rows of a multiple of 4 samples: float4 segments as above
                constexpr int Q4 = NT / 4;
                for (int q = tid; q < T * Q4; q += 256) {
                    const int m = q / Q4, s = s0 + (q - m * Q4) * 4;
                    if (s < S)
                        *reinterpret_cast<float4*>(out + ((size_t)m * C + c) * S + s) =
                            *reinterpret_cast<const float4*>(&otile[m * OTS + (s - s0)]);
                }
            } else if (S % 2 == 0 && (uintptr_t)out % 8 == 0) {
                constexpr int Q2 = NT / 2;
                for (int q = tid; q < T * Q2; q += 256) {
                    const int m = q / Q2, s = s0 + (q - m * Q2) * 2;
                    if (s < S)
                        *reinterpret_cast<float2*>(out + ((size_t)m * C + c) * S + s) =
                            *reinterpret_cast<const float2*>(&otile[m * OTS + (s - s0)]);
                }
            } else {
                // one float per lane, consecutive lanes on consecutive samples
                for (int q = tid; q < T * NT; q += 256) {
                    const int m = q / NT, sl = q - m * NT, s = s0 + sl;
                    if (s < S) out[((size_t)m * C + c) * S + s] = otile[m * OTS + sl];
                }
            }
        }
    }
}

template <typename TIN, bool VEC, int OUT_STRIDE>
__global__ __launch_bounds__(256) void decode_f32_kernel(const TIN* __restrict__ raw,
                                                         const int16_t* __restrict__ map, size_t R0, int T,
                                                         int C, int S, int Kp, const float* __restrict__ hf,
                                                         float scale, float* __restrict__ out)
{
    __shared__ float bt[DEC_KC * DEC_ROWF];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s0 = blockIdx.x * DEC_NT;
    const int c = blockIdx.y;
    const int mbase = blockIdx.z * 128 + wave * 32;
    const bool mvalid = mbase < Kp;
    const TIN* src = raw + (size_t)(uint16_t)map[c] * R0;

    v16f acc[2];
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[n][r] = 0.0f;

    const int kq = tid >> 3, sg = tid & 7;
    for (int k0 = 0; k0 < Kp; k0 += DEC_KC) {
        const int kc = min(DEC_KC, Kp - k0);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int kl = kq * 4 + j;
            if (kl >= kc) break;
            const int k = k0 + kl;
            float* dst = bt + kl * DEC_ROWF + sg * 8;
            if (k < T) {
                if constexpr (sizeof(TIN) == 2) {
                    int v[8];
                    load8<int16_t, VEC>((const int16_t*)src + (size_t)k * S, s0 + sg * 8, S, v);
#pragma unroll
                    for (int e = 0; e < 8; ++e) dst[e] = (float)v[e];
                } else {
                    const float* row = (const float*)src + (size_t)k * S;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int s = s0 + sg * 8 + e;
                        dst[e] = s < S ? row[s] : 0.0f;
                    }
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) dst[e] = 0.0f;
            }
        }
        __syncthreads();
        if (mvalid) {
            const float* hrow = hf + (size_t)(mbase + (lane & 31)) * Kp + k0 + (lane >> 5);
            for (int kk = 0; kk < kc; kk += 2) {
                const float a = hrow[kk];
                const float* brow = bt + (kk + (lane >> 5)) * DEC_ROWF + (lane & 31);
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, brow[nt * 32], acc[nt], 0, 0, 0);
            }
        }
    }
    if (!mvalid) return;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int s = s0 + nt * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = mbase + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (m < T && s < S) out[(((size_t)m * C + c) * S + s) * OUT_STRIDE] = acc[nt][r] * scale;
        }
    }
}

bool launch_decode(const void* raw, int type, const int16_t* d_map, size_t R0, int T, int C, int S,
                   const int8_t* d_h8, const float* d_hf, const int* d_rowsum, float scale, float* out,
                   int out_stride, int mode, hipStream_t st)
{
    if (out_stride != 1 && out_stride != 2) CUT_FAIL("decode: output stride must be 1 or 2");
    const int Kp = (T + 31) / 32 * 32;
    dim3 grid((S + DEC_NT - 1) / DEC_NT, C, (Kp + 127) / 128);
    const bool vec = ((uintptr_t)raw % 16 == 0) && (R0 % 8 == 0) && (S % 8 == 0);
    // streaming form: channel rows that start on 16 bytes (any S; 16-byte
    // loads and float4 stores when S is a multiple of 8 and the output
    // aligned, else the shifted chunks and one float per lane)
    const bool rows_al = ((uintptr_t)raw % 16 == 0) && (R0 % 8 == 0) && R0 <= (size_t)0x3fffffff;
    if (type == TYPE_I16 && mode == 1 && rows_al && Kp <= DEC_KC) {
        const bool al = (S % 8 == 0) && ((uintptr_t)out % 16 == 0);
        // streaming form: a fixed grid of workgroups per CU walks the tiles
        static int cus = 0;
        if (cus == 0) {
            int dev = 0;
            CUT_HIP_CHECK(hipGetDevice(&dev));
            CUT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        }
#ifndef CUT_DEC_SNT
#define CUT_DEC_SNT 64
#endif
        constexpr int SNT = CUT_DEC_SNT;
        const int nsx = (S + SNT - 1) / SNT;
        const int ntiles = nsx * C;
        const int wgs = std::min(ntiles, cus * DEC_STREAM_WG_PER_CU);
        const int16_t* r = (const int16_t*)raw;
#define CUT_DEC_S(O, A) decode_i8_stream_kernel<O, SNT, A><<<wgs, 256, 0, st>>>(r, d_map, R0, T, C, S, Kp, nsx, d_h8, d_rowsum, scale, out)
        if (al) { if (out_stride == 1) CUT_DEC_S(1, true); else CUT_DEC_S(2, true); }
        else { if (out_stride == 1) CUT_DEC_S(1, false); else CUT_DEC_S(2, false); }
#undef CUT_DEC_S
    } else if (type == TYPE_I16 && mode == 1) {
        const int16_t* r = (const int16_t*)raw;
#define CUT_DEC_I8(V, O) decode_i8_kernel<V, O><<<grid, 256, 0, st>>>(r, d_map, R0, T, C, S, Kp, d_h8, d_rowsum, scale, out)
        if (vec) { if (out_stride == 1) CUT_DEC_I8(true, 1); else CUT_DEC_I8(true, 2); }
        else { if (out_stride == 1) CUT_DEC_I8(false, 1); else CUT_DEC_I8(false, 2); }
#undef CUT_DEC_I8
    } else if (type == TYPE_I16) {
        const int16_t* r = (const int16_t*)raw;
#define CUT_DEC_F(V, O) decode_f32_kernel<int16_t, V, O><<<grid, 256, 0, st>>>(r, d_map, R0, T, C, S, Kp, d_hf, scale, out)
        if (vec) { if (out_stride == 1) CUT_DEC_F(true, 1); else CUT_DEC_F(true, 2); }
        else { if (out_stride == 1) CUT_DEC_F(false, 1); else CUT_DEC_F(false, 2); }
#undef CUT_DEC_F
    } else if (type == TYPE_F32) {
        const float* r = (const float*)raw;
        if (out_stride == 1)
            decode_f32_kernel<float, false, 1><<<grid, 256, 0, st>>>(r, d_map, R0, T, C, S, Kp, d_hf, scale, out);
        else
            decode_f32_kernel<float, false, 2><<<grid, 256, 0, st>>>(r, d_map, R0, T, C, S, Kp, d_hf, scale, out);
    } else {
        CUT_FAIL("data converter: unsupported input data type " << type);
    }
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

}  // namespace cut
