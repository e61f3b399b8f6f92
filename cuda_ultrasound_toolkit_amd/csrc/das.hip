// Delay-and-sum beamforming kernels for gfx950.
//
// HERCULES (3-D, row-column array): templated_beamformer.cuh:102-216.
// FORCES   (2-D synthetic aperture): templated_beamformer.cuh:219-324.
// Helpers: beamformer_utils.cuh:99-213, templated_beamformer.cuh:16-99.
//
// HERCULES design.  A workgroup owns a 16 x 16 x 8 voxel tile; each thread owns
// one (x, y) column and its 8 voxels along z, so per (t, c) the lateral
// distance rho^2 = dx^2 + dy^2 is computed once and shared by 8 voxels (the
// Hann weight is a polynomial in rho^2 (fn / 2z)^2).  For every transmit the
// workgroup derives, from the tile's bounding box and the F-number aperture,
// the channel range that can have apo > 0.1 anywhere in the tile; channels
// outside are skipped without touching memory.  Channels inside are processed
// 16 (linear) or 8 (cubic) at a time: for each, the span of sample indices any
// voxel of the tile can read is bounded from the box (min/max receive and
// transmit path), and that window (<= 256 samples) is staged into a shared
// LDS pool as interpolation coefficients per sample, so a tap is one (two)
// 16-byte LDS reads and a few FMAs instead of gathers from L2.  A window that
// is too long or does not fit the pool falls back to direct global reads for
// that channel.
#include "common.hpp"
#include "tuning.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

namespace cut {

constexpr int HT_X = 16, HT_Y = 16, HT_Z = 8;  // voxel tile
constexpr int HCB = 16;                         // most channels staged per pass
constexpr int HSB = 64;                         // channels per block of window tables (one per lane)
constexpr int HW = 256;                         // longest window staged in LDS (samples)
// LDS window pool of a workgroup: 32 KiB, shared by the channels of a pass at
// offsets given by a prefix sum of their window lengths.  Each sample i of a
// window is stored as the coefficients of the interpolant on [i, i+1]:
// linear, one float4 (p[i], p[i+1] - p[i]); cubic, two float4 (in two planes)
// holding the Catmull-Rom polynomial c0 + c1 t + c2 t^2 + c3 t^3 (16 channels per pass for
// linear, 8 for cubic).
constexpr int HPOOL_BYTES = 32768;
// apo > 0.1  <=>  fn*rho/(2z) < acos(sqrt(0.1))/pi for the Hann window.
constexpr float HANN_A_THR = 0.39758361765043326f;
constexpr float HANN_S_THR = HANN_A_THR * HANN_A_THR;
// cos^2(pi a) as a polynomial in s = a^2 on [0, HANN_S_THR] (where the
// threshold keeps the weight): near-minimax degree 4 with the constant term
// fixed at 1, max error 2.6e-6 in FP32 Horner form; replaces v_cos_f32, which
// issues at a quarter of the VALU rate.
constexpr float HANN_P1 = -9.86909866f, HANN_P2 = 32.4382935f, HANN_P3 = -42.1193924f, HANN_P4 = 25.4091187f;
__device__ __forceinline__ float hann_poly(float s)
{
    return fmaf(fmaf(fmaf(fmaf(HANN_P4, s, HANN_P3), s, HANN_P2), s, HANN_P1), s, 1.0f);
}

__device__ __forceinline__ float cos_2pi(float x) { return __builtin_amdgcn_cosf(x); }
// v_sqrt_f32 (1 ulp): the IEEE-exact sqrtf expansion costs ~15 VALU ops per call.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// cos^2(pi * min(a, 0.5)) = 0.5 + 0.5 cos(2 pi min(a, 0.5))  (beamformer_utils.cuh:142-154)
__device__ __forceinline__ float hann_from_a(float a)
{
    a = fminf(a, 0.5f);
    return fmaf(0.5f, cos_2pi(a), 0.5f);
}

// templated_beamformer.cuh:86-99
__device__ __forceinline__ float sin_apo_to(float pos, float peak, int power)
{
    float a = 3.14159265358979f * pos / peak;
    a = fminf(fmaxf(a, -3.14159265358979f), 3.14159265358979f);
    a = sinf(a);
    float out = 1.0f;
    for (int i = 0; i < power; ++i) out *= a;
    return out;
}

__device__ __forceinline__ float2 lerp2(float2 p0, float2 p1, float t)
{
    return make_float2(fmaf(t, p1.x - p0.x, p0.x), fmaf(t, p1.y - p0.y, p0.y));
}

// Catmull-Rom with tangents (p[i+1]-p[i-1])/2 (beamformer_utils.cuh:99-125).
__device__ __forceinline__ float2 cubic2(float2 pm1, float2 p0, float2 p1, float2 p2, float t)
{
    const float t2 = t * t;
    const float h00 = fmaf(fmaf(2.0f, t, -3.0f), t2, 1.0f);
    const float h10 = fmaf(t - 2.0f, t2, t);
    const float h01 = fmaf(-2.0f, t, 3.0f) * t2;
    const float h11 = (t - 1.0f) * t2;
    const float m0x = 0.5f * (p1.x - pm1.x), m0y = 0.5f * (p1.y - pm1.y);
    const float m1x = 0.5f * (p2.x - p0.x), m1y = 0.5f * (p2.y - p0.y);
    return make_float2(fmaf(h00, p0.x, fmaf(h01, p1.x, fmaf(h10, m0x, h11 * m1x))),
                       fmaf(h00, p0.y, fmaf(h01, p1.y, fmaf(h10, m0y, h11 * m1y))));
}

// IQ sample i of a buffer in format IQK (an IqFormat), as complex f32
// (staged formats hold value * scale; inv = 1 / scale, a power of two).
template <int IQK>
__device__ __forceinline__ float2 iq_at(const void* __restrict__ base, size_t i, float inv)
{
    if constexpr (IQK == IQ_F32) {
        return static_cast<const float2*>(base)[i];
    } else if constexpr (IQK == IQ_F16) {
        const float2 f = __half22float2(static_cast<const __half2*>(base)[i]);
        return make_float2(f.x * inv, f.y * inv);
    } else {
        const uint32_t u = static_cast<const uint32_t*>(base)[i];
        return make_float2(__uint_as_float(u << 16) * inv, __uint_as_float(u & 0xffff0000u) * inv);
    }
}

// Cubic interpolation at i + t from an FP32 IQ buffer of n samples: p[i-1] ..
// p[i+2], p[i+2] clamped to p[n-1] (1 <= i <= n-2), fetched as two 16-byte
// loads of sample pairs (8-byte aligned; the hardware takes unaligned 16-byte
// loads).  On the 128^3 RCA_TPW volume: 10.5 -> 10.1 ms against four 8-byte
// loads.  (One 16-byte load for the linear pair measured no better.)
__device__ __forceinline__ float2 cubic_pairs(const float2* __restrict__ iq, int i, int n, float t)
{
    float4 ab, cd;
    __builtin_memcpy(&ab, iq + (i - 1), sizeof(float4));
    __builtin_memcpy(&cd, iq + min(i + 1, n - 2), sizeof(float4));
    if (i + 2 >= n) cd = make_float4(cd.z, cd.w, cd.z, cd.w);
    return cubic2(make_float2(ab.x, ab.y), make_float2(ab.z, ab.w), make_float2(cd.x, cd.y), make_float2(cd.z, cd.w), t);
}

// Read p[i] of a row, index clamped into [0, S-1]: the reference's cubic form
// touches p[S] at x = S-2 with weight exactly 0, any finite value works.
template <int IQK>
__device__ __forceinline__ float2 row_at(const void* __restrict__ base, size_t row, int i, int S, float inv)
{
    i = min(max(i, 0), S - 1);
    return iq_at<IQK>(base, row + i, inv);
}

template <bool CUBIC, int IQK = IQ_F32>
__device__ __forceinline__ float2 interp_global(const void* __restrict__ base, size_t row, float x, int S,
                                                float inv = 1.0f)
{
    const float fi = floorf(x);
    const float t = x - fi;
    const int i = (int)fi;
    if constexpr (CUBIC)
        return cubic2(row_at<IQK>(base, row, i - 1, S, inv), row_at<IQK>(base, row, i, S, inv),
                      row_at<IQK>(base, row, i + 1, S, inv), row_at<IQK>(base, row, i + 2, S, inv), t);
    else
        return lerp2(row_at<IQK>(base, row, i, S, inv), row_at<IQK>(base, row, i + 1, S, inv), t);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// One weight for every tap of a channel (the sine apodization).
struct ConstW {
    float w;
};

// Weight of one tap: apodization when it passes the reference's threshold
// (apo > 0.1), else 0.  (Lanes of voxels outside the grid compute in-range
// values that are never stored.)
// rho2 = squared lateral distance, kz2j = (fn / (2 z))^2 of the voxel's depth.
template <bool SINAPO>
__device__ __forceinline__ float tap_weight(float rho2, float kz2j, float apo_c)
{
    if constexpr (SINAPO) {
        return apo_c > 0.1f ? apo_c : 0.0f;
    } else {
        // cos^2(pi a) > 0.1 <=> a < acos(sqrt(0.1))/pi on [0, 0.5]; for those a
        // the reference's min(a, 0.5) clamp is inactive
        const float s = rho2 * kz2j;
        return s < HANN_S_THR ? hann_poly(s) : 0.0f;
    }
}

// The reference's F-number Hann weight for any sign of the depth and the
// F-number: cos^2(pi min(a, 1/2)) with a = fn rho / (2 z)
// (beamformer_utils.cuh:143-154), and 0 unless it passes the threshold
// (apo > 0.1).  For a voxel above the array (z < 0) a is negative and
// unbounded below, and the weight comes back above 0.1 periodically in |a|.
// cos^2 has period 1 in a, so a is reduced to r = a - rint(a) in [-1/2, 1/2]
// and the weight is the polynomial form in r^2 with its threshold test, the
// same decision and value as the z > 0 path (a cosine's error would widen
// the band of taps whose inclusion differs from the reference's).  fminf maps
// a NaN ratio to 1/2 (weight 0), as the reference's does.  kzj = fn / (2 z).
__device__ __forceinline__ float hann_weight_any(float rho2, float kzj)
{
    const float a = fminf(kzj * fsqrt(rho2), 0.5f);
    const float r = a - rintf(a);
    const float s = r * r;
    return s < HANN_S_THR ? hann_poly(s) : 0.0f;
}

// Fractional sample index (txd + |rx| + fz) * spm + delay, clamped to [1, S-2].
__device__ __forceinline__ float tap_index(float r2, float spm, float basej, float hi_clamp)
{
    return __builtin_amdgcn_fmed3f(fmaf(fsqrt(r2), spm, basej), 1.0f, hi_clamp);
}

struct MinMax {
    float lo, hi;
};

__device__ __forceinline__ MinMax dist_range(float e, float a, float b)
{
    // |e - v| for v in [a, b]
    const float dmin = (e < a) ? a - e : (e > b ? e - b : 0.0f);
    const float dmax = fmaxf(fabsf(e - a), fabsf(e - b));
    return {dmin, dmax};
}

// Work decomposition: tiles (nxt x nyt x nzt) times `splits` contiguous ranges
// of the (group, transmit) loop, one workgroup each, deepest tiles first (the
// active-tap count grows with depth, so heavy workgroups start first and the
// light ones fill the tail).  With splits > 1 every workgroup writes its partial
// sums and hercules_reduce adds them in split order: (acc.re, acc.im) as a
// float2 per voxel, or (acc.re, acc.im, incoherent, 0) as a float4 when the
// coherence factor needs the incoherent sum.
struct HercGrid {
    int nxt, nyt, nzt, splits, gt_per_split;
};

// One 16 x 16 x 8 voxel tile at (bx, by, bz) over the (group, transmit)
// range [gt_lo, gt_hi).  Writes the final voxels (coherence factor applied
// when CF), or, when part is not null, the partial sums to part[voxel]
// (float2 acc, or float4 (acc, incoherent, 0) when CF).  All threads of the workgroup must call it (it synchronises).
template <bool CUBIC, bool FOCUS, bool CF, bool SINAPO, int IQK, bool CLAMP>
__device__ __forceinline__ void hercules_tile(const void* __restrict__ iq, float iq_inv, const DasParams& p, int bx,
                                              int by, int bz, int gt_lo, int gt_hi, float2* __restrict__ vol,
                                              void* __restrict__ part, float4* win, int* wlo_s, int* wmode_s,
                                              int* woff_s, int* wlen_s)
{
    // window entry stride (bytes), pool capacity (entries), channels per pass
    // and staging threads per channel.  Cubic entries are split into two
    // planes of float4 (c0, c1) and (c2, c3), the second HPOOL entries after
    // the first, so that both reads of a tap have the 16-byte lane stride of
    // the linear form (a 32-byte stride measured ~1.4 LDS bank-conflict
    // cycles per read).
    constexpr int WIN_ESZ = (int)sizeof(float4);
    constexpr int HPOOL = HPOOL_BYTES / (CUBIC ? 2 * WIN_ESZ : WIN_ESZ);
    constexpr int NCB = CUBIC ? HCB / 2 : HCB;
    constexpr int TPC = 256 / NCB;
    const int tid = threadIdx.x;
    // each wave owns an 8 x 8 patch of columns: a compact footprint keeps the
    // wave's aperture test coherent, so more channels are skipped whole
    const int wv = tid >> 6, ln = tid & 63;
    const int lx = (wv & 1) * 8 + (ln & 7), ly = (wv >> 1) * 8 + (ln >> 3);
    const int ix = bx + lx, iy = by + ly;
    const bool col_valid = ix < p.Nx && iy < p.Ny;

    // voxel coordinates rounded once from FP64 (as the depths below)
    const double vy_d = (double)p.min_y + (double)min(iy, p.Ny - 1) * (double)p.res_y;
    const float vx = (float)((double)p.min_x + (double)min(ix, p.Nx - 1) * (double)p.res_x);
    const float vy = (float)vy_d;
    // the wave's patch corners: lane 0 (first column, first row), lane 7 (last
    // column) and lane 56 (last row), the same FP32 values the lanes hold
    const float wvx0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vx), 0));
    const float wvx1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vx), 7));
    const float wvy0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vy), 0));
    const float wvy1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vy), 56));

    // tile bounding box
    const float xa = p.min_x + (float)bx * p.res_x, xb = p.min_x + (float)min(bx + HT_X - 1, p.Nx - 1) * p.res_x;
    const float ya = p.min_y + (float)by * p.res_y, yb = p.min_y + (float)min(by + HT_Y - 1, p.Ny - 1) * p.res_y;
    const float za = p.min_z + (float)bz * p.res_z, zb = p.min_z + (float)min(bz + HT_Z - 1, p.Nz - 1) * p.res_z;
    const float X0 = fminf(xa, xb), X1 = fmaxf(xa, xb);
    const float Y0 = fminf(ya, yb), Y1 = fmaxf(ya, yb);
    const float Z0 = fminf(za, zb), Z1 = fmaxf(za, zb);
    const MinMax zr = dist_range(0.0f, Z0, Z1);  // |z| range

    const float fz = p.focal_z;
    // the tile's 8 depths are the same for every lane: z-only values go to
    // scalar registers (readfirstlane), which frees 16 VGPRs
    auto uni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
    auto uni_d = [](double v) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll)), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    };
    // per-voxel constants.  The depth is formed in FP64 (min_z + iz res_z from
    // the float grid constants): a float depth is off by up to half an ulp of
    // ~0.06 m, the same error for every tap of a plane (2e-5 samples on both
    // paths), which does not average out over the taps as rounding does.
    float z2[HT_Z], kz2[HT_Z], base[HT_Z];
    float kzs[HT_Z];  // fn / (2 z), signed: the weight of volumes with a voxel at z <= 0
    double base_d[HT_Z];  // the same in FP64, anchored per channel block below
    bool zvalid[HT_Z];
#pragma unroll
    for (int j = 0; j < HT_Z; ++j) {
        const int iz = bz + j;
        zvalid[j] = iz < p.Nz;
        const double zd = (double)p.min_z + (double)min(iz, p.Nz - 1) * (double)p.res_z;
        const float z = (float)zd;
        z2[j] = uni((float)(zd * zd));
        const float kz = p.fn_rx / z * 0.5f;
        kz2[j] = uni(kz * kz);
        kzs[j] = kz;
        double txd;
        if constexpr (FOCUS) {
            const double ty = vy_d, tzz = zd - (double)fz;
            txd = copysign(sqrt(ty * ty + tzz * tzz), tzz) + (double)fz;
        } else {
            txd = zd;  // (0, y sin 0, z cos 0): steering angle zero
        }
        base_d[j] = (txd + (double)fz) * (double)p.spm + (double)p.delay;
        base[j] = (float)base_d[j];
        // txd = z: depth-only, held in scalar registers (24 VGPRs fewer; the
        // cubic kernel with its pipelined taps 10.27 -> 10.20 ms on the 128^3
        // frame, where before the pipelining it measured slower so)
        if constexpr (!FOCUS) {
            base[j] = uni(base[j]);
            base_d[j] = uni_d(base_d[j]);
        }
    }
    // Smallest and largest (fn/2z)^2 over the tile's depths (the deepest and
    // the shallowest when the tile lies below the array, z > 0), and from them
    // two radii for rho^2, uniform over the tile: r2_any, below which some
    // depth of a column can be inside the aperture (0.2 % margin), and
    // r2_full, below which every depth is, in the same FP32 product
    // rho2 * kz2[j] < HANN_S_THR that the taps test (rho2 < r2_full gives
    // rho2 * kz2_hi <= r2_full * kz2_hi < HANN_S_THR, products being monotonic).
    float kz2_lo = kz2[0], kz2_hi = kz2[0];
#pragma unroll
    for (int j = 1; j < HT_Z; ++j) { kz2_lo = fminf(kz2_lo, kz2[j]); kz2_hi = fmaxf(kz2_hi, kz2[j]); }
    constexpr float HANN_S_THR_W = (HANN_A_THR * 1.001f) * (HANN_A_THR * 1.001f);
    const float r2_any = uni(HANN_S_THR_W / kz2_lo);
    float r2_full = HANN_S_THR / kz2_hi;
    while (r2_full > 0.0f && !(r2_full * kz2_hi < HANN_S_THR)) r2_full = __int_as_float(__float_as_int(r2_full) - 1);
    r2_full = uni(r2_full);
    // transmit path bounds over the tile
    float txd_lo, txd_hi;
    if constexpr (FOCUS) {
        const MinMax yr = dist_range(0.0f, Y0, Y1);
        float cand[4];
        int n = 0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const float yy = a ? yr.hi : yr.lo;
                const float zz = (b ? Z1 : Z0) - fz;
                cand[n++] = copysignf(sqrtf(yy * yy + zz * zz), zz) + fz;
            }
        txd_lo = fminf(fminf(cand[0], cand[1]), fminf(cand[2], cand[3]));
        txd_hi = fmaxf(fmaxf(cand[0], cand[1]), fmaxf(cand[2], cand[3]));
    } else {
        txd_lo = Z0;
        txd_hi = Z1;
    }

    // Aperture radius (Hann): rho < 2*A*z/fn.  Unbounded when fn <= 0 or z <= 0.
    // The generic instantiation (SINAPO) serves the sine apodization and, for
    // Hann, volumes the polynomial form does not cover (a voxel at z <= 0 or a
    // negative F-number: the launch routes them here): the exact weight per tap.
    const bool hann_exact = SINAPO && p.apo_type != RX_TO_SIN;
    const bool aperture_bounded = !SINAPO && p.fn_rx > 0.0f && Z0 > 0.0f && p.pitch_x > 0.0f;
    const bool unbounded = !(p.fn_rx > 0.0f);  // no F-number: every lane is inside the aperture
    const float rmax_ap = aperture_bounded ? 2.0f * (HANN_A_THR * 1.001f) * Z1 / p.fn_rx + 1e-6f : 0.0f;

    float2 acc[HT_Z];
    float inc[HT_Z];
#pragma unroll
    for (int j = 0; j < HT_Z; ++j) { acc[j] = make_float2(0.0f, 0.0f); inc[j] = 0.0f; }

    const int G = p.readi_groups > 1 ? p.readi_groups : 1;
    const float lo_clamp = 1.0f, hi_clamp = (float)p.S - 2.0f;
    const size_t rowlen = (size_t)p.S;

    for (int gt = gt_lo; gt < gt_hi; ++gt) {
        {
            const int g = gt / p.T, t = gt - g * p.T;
            float cur = (float)t;
            bool neg = false;
            if (G > 1) {
                cur = p.encoding == WALSH ? (float)t * (float)G + (float)g : (float)g * (float)p.T + (float)t;
                const int bit = p.encoding == WALSH ? ((t & 1) ? G - 1 - g : g) : g;
                neg = (p.readi_bits >> bit) & 1ull;
            }
            const float ey = fmaf(cur, p.pitch_y, p.xdc_min_y);
            const MinMax dyr = dist_range(ey, Y0, Y1);
            // a negative READI sign is applied by accumulating this transmit into -acc
            auto flip = [&]() {
#pragma unroll
                for (int j = 0; j < HT_Z; ++j) { acc[j].x = -acc[j].x; acc[j].y = -acc[j].y; }
            };
            int c_lo = 0, c_hi = p.C - 1;
            if (aperture_bounded) {
                if (dyr.lo >= rmax_ap) continue;
                const float hw = sqrtf(rmax_ap * rmax_ap - dyr.lo * dyr.lo);
                c_lo = max(0, (int)floorf((X0 - hw - p.xdc_min_x) / p.pitch_x) - 1);
                c_hi = min(p.C - 1, (int)ceilf((X1 + hw - p.xdc_min_x) / p.pitch_x) + 1);
                if (c_lo > c_hi) continue;
            }
            const float dy = ey - vy;
            const float dy2 = dy * dy;
            // largest dy^2 of the wave's lanes (one of its two edge rows, the same
            // product), and a lower bound of the smallest
            const float dya = ey - wvy0, dyb = ey - wvy1;
            const float wdy2_hi = fmaxf(dya * dya, dyb * dyb);
            const bool ey_in = ey >= fminf(wvy0, wvy1) && ey <= fmaxf(wvy0, wvy1);
            const float wdy2_lo = ey_in ? 0.0f : fminf(dya * dya, dyb * dyb);
            const size_t toff = (size_t)t * p.C * rowlen;
            if (neg) flip();
            float apo_y = 1.0f;
            if constexpr (SINAPO) if (!hann_exact) apo_y = sin_apo_to(ey, p.xdc_max_y, p.to_power);

            // Channels go in blocks of up to 64.  Wave 0 bounds the windows of a
            // block's channels, one per lane (the sample span any voxel of the
            // tile can read), and places them in the LDS pool pass by pass (an
            // exclusive prefix sum over each NCB-lane group; a pass is NCB
            // channels): one computation for a block instead of one per wave
            // and pass.  Each wave meanwhile sorts the block's channels for its
            // own patch (below).
            for (int cb64 = c_lo; cb64 <= c_hi; cb64 += HSB) {
              const int nC64 = min(HSB, c_hi - cb64 + 1);
              const int lane = tid & 63;
              __syncthreads();  // the previous block is done with the tables and the pool
              if (tid < 64) {
                  const int kb = lane & (NCB - 1);
                  int wlo = 0, len = 0;
                  if (lane < nC64) {
                      const int c = cb64 + lane;
                      const float ex = fmaf((float)c, p.pitch_x, p.xdc_min_x);
                      const MinMax dxr = dist_range(ex, X0, X1);
                      // v_sqrt_f32 (1 ulp, ~1e-4 samples here) against the windows'
                      // 2-sample margins: the IEEE sqrtf expansion is not needed
                      const float rmin = fsqrt(dxr.lo * dxr.lo + dyr.lo * dyr.lo + zr.lo * zr.lo);
                      const float rmax = fsqrt(dxr.hi * dxr.hi + dyr.hi * dyr.hi + zr.hi * zr.hi);
                      float ilo = fmaf(txd_lo + rmin + fz, p.spm, (float)p.delay);
                      float ihi = fmaf(txd_hi + rmax + fz, p.spm, (float)p.delay);
                      if (p.spm < 0.0f) { const float tt = ilo; ilo = ihi; ihi = tt; }
                      ilo = fminf(fmaxf(ilo, lo_clamp), hi_clamp);
                      ihi = fminf(fmaxf(ihi, lo_clamp), hi_clamp);
                      wlo = max(0, (int)floorf(ilo) - 2);
                      // the linear entry of sample i holds p[i + 1] - p[i], so a window
                      // ends at S-2 (indices are clamped to [1, S-2]; entry S-1 is never
                      // read); the cubic entries clamp their neighbours themselves
                      const int whi = min(CUBIC ? p.S - 1 : p.S - 2, (int)floorf(ihi) + 3);
                      len = whi - wlo + 1;
                  }
                  const bool fits = lane < nC64 && len > 0 && len <= HW;
                  const int need = fits ? len : 0;
                  int scan = need;
#pragma unroll
                  for (int off = 1; off < NCB; off <<= 1) {
                      const int v = __shfl_up(scan, off, NCB);
                      if (kb >= off) scan += v;
                  }
                  wlo_s[lane] = wlo;
                  wlen_s[lane] = len;
                  woff_s[lane] = scan - need;
                  wmode_s[lane] = (fits && scan <= HPOOL) ? 0 : 1;
              }
              // lane L holds the element x of channel cb64 + L
              const float my_ex = fmaf((float)(cb64 + lane), p.pitch_x, p.xdc_min_x);
              // Lane L sorts channel cb64 + L for the whole wave from the patch's
              // extreme lanes: the largest rho2 of any lane is that of a corner lane
              // (the largest |dx| and dy^2, same FP32 expression, monotonic), and
              // the patch box bounds the smallest from below.
              unsigned long long m_any64 = 0, m_full64 = 0;
              float my_apo = 0.0f;  // sine apodization: lane L's channel weight
              if constexpr (!SINAPO) {
                  const float dxa = my_ex - wvx0, dxb = my_ex - wvx1;
                  const float dx_hi = fmaxf(fabsf(dxa), fabsf(dxb));
                  const bool ex_in = my_ex >= fminf(wvx0, wvx1) && my_ex <= fmaxf(wvx0, wvx1);
                  const float dx_lo = ex_in ? 0.0f : fminf(fabsf(dxa), fabsf(dxb));
                  const bool any_in = lane < nC64 && (unbounded || fmaf(dx_lo, dx_lo, wdy2_lo) < r2_any);
                  const bool all_in = fmaf(dx_hi, dx_hi, wdy2_hi) < r2_full;
                  m_any64 = __builtin_amdgcn_ballot_w64(any_in);
                  m_full64 = __builtin_amdgcn_ballot_w64(all_in) & m_any64;
              } else if (!hann_exact) {
                  // sine apodization of channel cb64 + lane (as chan() forms it)
                  my_apo = sin_apo_to(my_ex, p.xdc_max_x, p.to_power) * apo_y;
                  m_full64 = __builtin_amdgcn_ballot_w64(lane < nC64 && my_apo > 0.1f);
                  m_any64 = m_full64;
              }
              __syncthreads();  // tables ready
              for (int sub = 0; sub < nC64; sub += NCB) {
                const int cb = cb64 + sub;
                const int nC = min(NCB, nC64 - sub);
                if (sub > 0) __syncthreads();  // the previous pass is done with the pool
                {
                    // threads [wi*TPC, (wi+1)*TPC) stage channel cb + wi
                    const int wi = tid / TPC, part = tid % TPC;
                    if (wi < nC && wmode_s[sub + wi] == 0) {
                        const int c_wlo = wlo_s[sub + wi], c_len = wlen_s[sub + wi], c_off = woff_s[sub + wi];
                        {
                            const size_t row = toff + (size_t)(cb + wi) * rowlen;
                            if constexpr (CUBIC) {
                                // Catmull-Rom on [i, i+1] (tangents (p[i+1] - p[i-1]) / 2):
                                // c0 = p0, c1 = (p1 - pm1)/2, c2 = pm1 - 2.5 p0 + 2 p1 - p2/2,
                                // c3 = (p2 - pm1)/2 + 1.5 (p0 - p1); neighbours outside the row
                                // are clamped into it: the entries that need them are read
                                // only at t = 0 (sample S-2 at the clamp) or never
                                float4* w4 = win + c_off;
                                for (int q = part; q < c_len; q += TPC) {
                                    const int i = c_wlo + q;
                                    const float2 pm = iq_at<IQK>(iq, row + max(i - 1, 0), iq_inv);
                                    const float2 p0 = iq_at<IQK>(iq, row + i, iq_inv);
                                    const float2 p1 = iq_at<IQK>(iq, row + min(i + 1, p.S - 1), iq_inv);
                                    const float2 p2 = iq_at<IQK>(iq, row + min(i + 2, p.S - 1), iq_inv);
                                    const float c1x = 0.5f * (p1.x - pm.x), c1y = 0.5f * (p1.y - pm.y);
                                    const float c2x = fmaf(-0.5f, p2.x, fmaf(2.0f, p1.x, fmaf(-2.5f, p0.x, pm.x)));
                                    const float c2y = fmaf(-0.5f, p2.y, fmaf(2.0f, p1.y, fmaf(-2.5f, p0.y, pm.y)));
                                    const float c3x = fmaf(1.5f, p0.x - p1.x, 0.5f * (p2.x - pm.x));
                                    const float c3y = fmaf(1.5f, p0.y - p1.y, 0.5f * (p2.y - pm.y));
                                    w4[q] = make_float4(p0.x, p0.y, c1x, c1y);
                                    w4[HPOOL + q] = make_float4(c2x, c2y, c3x, c3y);
                                }
                            } else {
                                // (p[i], p[i+1] - p[i]), i <= S-2: both samples from one address.
                                // Four entries per thread are requested before any is written, so a
                                // window's loads overlap instead of each waiting on the last (when
                                // the rows come from HBM rather than L2 the wait would otherwise
                                // exceed a pass's arithmetic).
                                float4* w4 = win + c_off;
                                constexpr int U = 4;
                                for (int q0 = part; q0 < c_len; q0 += U * TPC) {
                                    float2 a[U], b[U];
#pragma unroll
                                    for (int u = 0; u < U; ++u) {
                                        const int q = min(q0 + u * TPC, c_len - 1);
                                        const size_t i = row + (size_t)(c_wlo + q);
                                        a[u] = iq_at<IQK>(iq, i, iq_inv);
                                        b[u] = iq_at<IQK>(iq, i + 1, iq_inv);
                                    }
#pragma unroll
                                    for (int u = 0; u < U; ++u) {
                                        const int q = q0 + u * TPC;
                                        if (q < c_len) w4[q] = make_float4(a[u].x, a[u].y, b[u].x - a[u].x, b[u].y - a[u].y);
                                    }
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                // per channel: lateral distance, apodization inputs, and whether any
                // lane of the wave can be inside the aperture (uniform skip)
                auto chan = [&](float ex, float& rho2, float& apo_c) -> bool {
                    const float dx = ex - vx;
                    rho2 = fmaf(dx, dx, dy2);
                    apo_c = 0.0f;
                    if constexpr (SINAPO) {
                        if (hann_exact) return true;  // per tap, below
                        apo_c = sin_apo_to(ex, p.xdc_max_x, p.to_power) * apo_y;
                        return apo_c > 0.1f;  // uniform: depends on the element only
                    } else {
                        // the depth with the smallest (fn/2z)^2 has the widest aperture
                        return unbounded || __builtin_amdgcn_ballot_w64(rho2 * kz2_lo < HANN_S_THR_W) != 0;
                    }
                };
                // Sample indices of this block of channels are taken relative to
                // the smallest window start of the block (an integer, so fractions
                // are unchanged and indices stay small and non-negative); the
                // per-channel window offset then moves into the uniform LDS base
                // (a byte offset that may be negative: every index read is at
                // least the window's start).  Lane cc of every wave holds channel
                // cc's element x and LDS base, bit cc of lds_mask whether its
                // window is in LDS; the channel loops read them with readlane, so
                // their bookkeeping stays in scalar registers.
                const bool my_lds = lane < nC && wmode_s[sub + lane] == 0;
                const int my_wlo = my_lds ? wlo_s[sub + lane] : 0x7fffffff;
                const unsigned long long lds_mask = __builtin_amdgcn_ballot_w64(my_lds);
                int org = my_wlo;
#pragma unroll
                for (int off = 8; off > 0; off >>= 1) org = min(org, __shfl_xor(org, off, 16));
                org = __builtin_amdgcn_readfirstlane(org);
                const int my_base = my_lds ? (woff_s[sub + lane] - (my_wlo - org)) * (int)WIN_ESZ : 0;
                const float forg = (float)org;
                const float lo_r = lo_clamp - forg, hi_r = hi_clamp - forg;
                float base_r[HT_Z];
#pragma unroll
                for (int j = 0; j < HT_Z; ++j) base_r[j] = (float)(base_d[j] - (double)org);
                const char* wbytes = reinterpret_cast<const char*>(win);
                // one tap of channel cc (LDS window base wb, lateral distance^2
                // rho2): sample index, window entry, interpolation, accumulation
                auto tap_at = [&](int j, float w, int wb, float rho2) {
                    // the window covers every voxel of the tile with a 2-sample margin,
                    // so the index needs no clamp beyond the reference's [1, S-2], and
                    // none at all in tiles whose every index lies inside it
                    const float xr = fmaf(fsqrt(rho2 + z2[j]), p.spm, base_r[j]);
                    const float x = CLAMP ? __builtin_amdgcn_fmed3f(xr, lo_r, hi_r) : xr;
                    const float tt = __builtin_amdgcn_fractf(x);
                    const int addr = wb + (int)x * (int)WIN_ESZ;  // window entry of block-relative index
                    float2 v;
                    if constexpr (CUBIC) {
                        const float4 e0 = *reinterpret_cast<const float4*>(wbytes + addr);
                        const float4 e1 = *reinterpret_cast<const float4*>(wbytes + addr + HPOOL * WIN_ESZ);
                        v = make_float2(fmaf(fmaf(fmaf(e1.z, tt, e1.x), tt, e0.z), tt, e0.x),
                                        fmaf(fmaf(fmaf(e1.w, tt, e1.y), tt, e0.w), tt, e0.y));
                    } else {
                        const float4 e = *reinterpret_cast<const float4*>(wbytes + addr);
                        v = make_float2(fmaf(tt, e.z, e.x), fmaf(tt, e.w, e.y));
                    }
                    acc[j].x = fmaf(w, v.x, acc[j].x);
                    acc[j].y = fmaf(w, v.y, acc[j].y);
                    if constexpr (CF) inc[j] = fmaf(v.x, v.x, fmaf(v.y, v.y, inc[j]));  // w != 0 here
                };
                auto chan_ex = [&](int cc) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_ex), sub + cc)); };
                // The 8 taps of a channel are software-pipelined: the window
                // entry (entries, cubic) of depth j+1 is requested before depth j
                // is accumulated, so each wave keeps LDS reads in flight across
                // its arithmetic (each tap waiting on its own read left the
                // kernel latency-bound at four waves per SIMD).  mode: std::false_type
                // (Hann weight, every tap inside the aperture), std::true_type (Hann
                // weight and threshold), ConstW (one weight for the channel: sine
                // apodization).
                auto taps8 = [&](int wb, float rho2, auto mode) {
                    struct Ent {
                        float4 e0, e1;
                        float tt;
                    };
                    auto fetch = [&](int j) {
                        Ent r;
                        const float xr = fmaf(fsqrt(rho2 + z2[j]), p.spm, base_r[j]);
                        const float x = CLAMP ? __builtin_amdgcn_fmed3f(xr, lo_r, hi_r) : xr;
                        r.tt = __builtin_amdgcn_fractf(x);
                        const int addr = wb + (int)x * (int)WIN_ESZ;
                        r.e0 = *reinterpret_cast<const float4*>(wbytes + addr);
                        if constexpr (CUBIC) r.e1 = *reinterpret_cast<const float4*>(wbytes + addr + HPOOL * WIN_ESZ);
                        return r;
                    };
                    Ent cur = fetch(0);
#pragma unroll
                    for (int j = 0; j < HT_Z; ++j) {
                        Ent nxt;
                        if (j + 1 < HT_Z) nxt = fetch(j + 1);
                        float w;
                        if constexpr (std::is_same_v<decltype(mode), ConstW>) w = mode.w;
                        else if constexpr (decltype(mode)::value) w = tap_weight<false>(rho2, kz2[j], 0.0f);
                        else w = hann_poly(rho2 * kz2[j]);
                        const float tt = cur.tt;
                        float2 v;
                        if constexpr (CUBIC)
                            v = make_float2(fmaf(fmaf(fmaf(cur.e1.z, tt, cur.e1.x), tt, cur.e0.z), tt, cur.e0.x),
                                            fmaf(fmaf(fmaf(cur.e1.w, tt, cur.e1.y), tt, cur.e0.w), tt, cur.e0.y));
                        else
                            v = make_float2(fmaf(tt, cur.e0.z, cur.e0.x), fmaf(tt, cur.e0.w, cur.e0.y));
                        acc[j].x = fmaf(w, v.x, acc[j].x);
                        acc[j].y = fmaf(w, v.y, acc[j].y);
                        // the incoherent sum takes the taps inside the aperture only
                        // (a weight that passes the threshold is at least 0.1)
                        if constexpr (CF) {
                            const float in = fmaf(v.x, v.x, fmaf(v.y, v.y, inc[j]));
                            if constexpr (std::is_same_v<decltype(mode), std::true_type>) inc[j] = w != 0.0f ? in : inc[j];
                            else inc[j] = in;
                        }
                        if (j + 1 < HT_Z) cur = nxt;
                    }
                };
                if constexpr (!SINAPO) {
                    // Hann apodization (with or without the coherence factor).  The
                    // pass's channels are sorted first into two scalar masks: those where
                    // every lane of the wave is inside the aperture at every depth of
                    // the tile (rho2 < r2_full; most wave-channels of a volume), whose
                    // taps need no threshold test, and those where only some are
                    // (rho2 < r2_any for some lane).  Two plain loops then keep the
                    // accumulators in place (a branch between the two tap forms inside
                    // one loop measured slower: the compiler moves the accumulators
                    // between register sets at its join).
                    const unsigned long long m_any = (m_any64 >> sub) & lds_mask;
                    unsigned long long m_full = (m_full64 >> sub) & m_any;
                    unsigned long long m_part = m_any & ~m_full;
                    while (m_full) {
                        const int cc = __builtin_ctzll(m_full);
                        m_full &= m_full - 1;
                        const int wb = __builtin_amdgcn_readlane(my_base, cc);
                        const float dx = chan_ex(cc) - vx;
                        taps8(wb, fmaf(dx, dx, dy2), std::false_type{});
                    }
                    while (m_part) {
                        const int cc = __builtin_ctzll(m_part);
                        m_part &= m_part - 1;
                        const int wb = __builtin_amdgcn_readlane(my_base, cc);
                        const float dx = chan_ex(cc) - vx;
                        taps8(wb, fmaf(dx, dx, dy2), std::true_type{});
                    }
                } else if (!hann_exact) {
                    // sine apodization: one weight per channel (and transmit), the same
                    // for every lane and depth; channels whose weight passes 0.1
                    unsigned long long m_full = (m_full64 >> sub) & lds_mask;
                    while (m_full) {
                        const int cc = __builtin_ctzll(m_full);
                        m_full &= m_full - 1;
                        const int wb = __builtin_amdgcn_readlane(my_base, cc);
                        const float dx = chan_ex(cc) - vx;
                        const float wc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_apo), sub + cc));
                        taps8(wb, fmaf(dx, dx, dy2), ConstW{wc});
                    }
                } else {
                    // the exact Hann weight of volumes reaching z <= 0: channels whose
                    // window sits in LDS; a single path keeps the accumulators in place
                    // across iterations
                    for (int cc = 0; cc < nC; ++cc) {
                        if (!((lds_mask >> cc) & 1ull)) continue;
                        const int wb = __builtin_amdgcn_readlane(my_base, cc);
                        const float dx = chan_ex(cc) - vx;
                        const float rho2 = fmaf(dx, dx, dy2);
                        // Taps outside the aperture get weight 0, or, with the coherence
                        // factor, are masked off (and skipped when no lane of the wave is
                        // inside).
#pragma unroll
                        for (int j = 0; j < HT_Z; ++j) {
                            const float w = hann_weight_any(rho2, kzs[j]);
                            if (CF && w == 0.0f) continue;
                            tap_at(j, w, wb, rho2);
                        }
                    }
                }
                // channels whose window did not fit: direct reads
                for (int cc = 0; cc < nC; ++cc) {
                    if ((lds_mask >> cc) & 1ull) continue;
                    float rho2, apo_c;
                    if (!chan(chan_ex(cc), rho2, apo_c))
                        continue;
                    const size_t grow = toff + (size_t)(cb + cc) * rowlen;
#pragma unroll
                    for (int j = 0; j < HT_Z; ++j) {
                        const float w = SINAPO && hann_exact ? hann_weight_any(rho2, kzs[j])
                                                             : tap_weight<SINAPO>(rho2, kz2[j], apo_c);
                        if (w == 0.0f) continue;
                        const float x = tap_index(rho2 + z2[j], p.spm, base[j], hi_clamp);
                        const float2 v = interp_global<CUBIC, IQK>(iq, grow, x, p.S, iq_inv);
                        acc[j].x = fmaf(w, v.x, acc[j].x);
                        acc[j].y = fmaf(w, v.y, acc[j].y);
                        if constexpr (CF) inc[j] = fmaf(v.x, v.x, fmaf(v.y, v.y, inc[j]));
                    }
                }
              }
            }
            if (neg) flip();
        }
    }
    if (!col_valid) return;
#pragma unroll
    for (int j = 0; j < HT_Z; ++j) {
        const int iz = bz + j;
        if (iz >= p.Nz) break;
        const size_t vi = ((size_t)iz * p.Ny + iy) * p.Nx + ix;
        float2 a = acc[j];
        if (part) {
            if constexpr (CF) static_cast<float4*>(part)[vi] = make_float4(a.x, a.y, inc[j], 0.0f);
            else static_cast<float2*>(part)[vi] = a;
            continue;
        }
        if constexpr (CF) {
            float cf = powf((a.x * a.x + a.y * a.y) / inc[j], p.cw);
            if (isnan(cf)) cf = 0.0f;
            a.x *= cf;
            a.y *= cf;
        }
        vol[vi] = a;
    }
}


// Whether every sample index any (transmit, channel, voxel) of the tile can
// form lies inside [2, S-3]: then the reference's clamp to [1, S-2] never
// acts (1-sample margin for rounding) and the tap loop drops it.  Bounds:
// transmit path over the tile
// box as in hercules_tile, receive path between the tile's depth range and
// the distance to the farthest element of the aperture.
template <bool FOCUS>
__device__ __forceinline__ bool tile_never_clamps(const DasParams& p, int bx, int by, int bz)
{
    if (!(p.spm > 0.0f)) return false;
    const float xa = p.min_x + (float)bx * p.res_x, xb = p.min_x + (float)min(bx + HT_X - 1, p.Nx - 1) * p.res_x;
    const float ya = p.min_y + (float)by * p.res_y, yb = p.min_y + (float)min(by + HT_Y - 1, p.Ny - 1) * p.res_y;
    const float za = p.min_z + (float)bz * p.res_z, zb = p.min_z + (float)min(bz + HT_Z - 1, p.Nz - 1) * p.res_z;
    const float X0 = fminf(xa, xb), X1 = fmaxf(xa, xb), Y0 = fminf(ya, yb), Y1 = fmaxf(ya, yb);
    const float Z0 = fminf(za, zb), Z1 = fmaxf(za, zb);
    const MinMax zr = dist_range(0.0f, Z0, Z1);
    const float fz = p.focal_z;
    float txd_lo = Z0, txd_hi = Z1;
    if constexpr (FOCUS) {
        const MinMax yr = dist_range(0.0f, Y0, Y1);
        txd_lo = 1e30f;
        txd_hi = -1e30f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const float yy = a ? yr.hi : yr.lo, zz = (b ? Z1 : Z0) - fz;
                const float v = copysignf(sqrtf(yy * yy + zz * zz), zz) + fz;
                txd_lo = fminf(txd_lo, v);
                txd_hi = fmaxf(txd_hi, v);
            }
    }
    const int G = p.readi_groups > 1 ? p.readi_groups : 1;
    const float ex0 = p.xdc_min_x, ex1 = fmaf((float)(p.C - 1), p.pitch_x, p.xdc_min_x);
    const float ey0 = p.xdc_min_y, ey1 = fmaf((float)(G * p.T - 1), p.pitch_y, p.xdc_min_y);
    const float dx = fmaxf(fmaxf(fabsf(ex0 - X0), fabsf(ex0 - X1)), fmaxf(fabsf(ex1 - X0), fabsf(ex1 - X1)));
    const float dy = fmaxf(fmaxf(fabsf(ey0 - Y0), fabsf(ey0 - Y1)), fmaxf(fabsf(ey1 - Y0), fabsf(ey1 - Y1)));
    const float rmax = sqrtf(dx * dx + dy * dy + zr.hi * zr.hi);
    const float ilo = fmaf(txd_lo + zr.lo + fz, p.spm, (float)p.delay);
    const float ihi = fmaf(txd_hi + rmax + fz, p.spm, (float)p.delay);
    return ilo >= 2.0f && ihi <= (float)p.S - 3.0f;
}

// Four workgroups per CU (four waves per SIMD): a budget of 128 VGPRs, which
// the linear loops meet without spilling.  The cubic loops (two-plane
// coefficient reads, Horner steps) want ~150 and spill 10-28 registers at
// 128; with three workgroups per CU they do not spill but ran slower on the
// 128^3 frame (13.9 against 12.6 ms), so the cubic kernels keep four too.

template <bool CUBIC, bool FOCUS, bool CF, bool SINAPO, int IQK>
__global__ __launch_bounds__(256, 4) void hercules_kernel(const void* __restrict__ iq, float iq_inv, const DasParams p,
                                                       const HercGrid hg, float2* __restrict__ vol,
                                                       void* __restrict__ part)
{
    __shared__ float4 win[HPOOL_BYTES / sizeof(float4)];
    __shared__ int wlo_s[HSB];
    __shared__ int wmode_s[HSB];
    __shared__ int woff_s[HSB];
    __shared__ int wlen_s[HSB];
    const int per_z = hg.nxt * hg.nyt * hg.splits;
    const int wz = blockIdx.x / per_z;
    const int wr = blockIdx.x - wz * per_z;
    const int split = wr % hg.splits;
    const int wxy = wr / hg.splits;
    const int bx = (wxy % hg.nxt) * HT_X, by = (wxy / hg.nxt) * HT_Y, bz = (hg.nzt - 1 - wz) * HT_Z;
    const int G = p.readi_groups > 1 ? p.readi_groups : 1;
    const int gt_lo = split * hg.gt_per_split;
    const int gt_hi = min(G * p.T, gt_lo + hg.gt_per_split);
    constexpr size_t PART_ESZ = CF ? sizeof(float4) : sizeof(float2);
    void* pp = hg.splits > 1 ? static_cast<char*>(part) + (size_t)split * p.Nx * p.Ny * p.Nz * PART_ESZ : nullptr;
    if (tile_never_clamps<FOCUS>(p, bx, by, bz))
        hercules_tile<CUBIC, FOCUS, CF, SINAPO, IQK, false>(iq, iq_inv, p, bx, by, bz, gt_lo, gt_hi, vol, pp, win,
                                                            wlo_s, wmode_s, woff_s, wlen_s);
    else
        hercules_tile<CUBIC, FOCUS, CF, SINAPO, IQK, true>(iq, iq_inv, p, bx, by, bz, gt_lo, gt_hi, vol, pp, win,
                                                           wlo_s, wmode_s, woff_s, wlen_s);
}

// A batch of frames of any sizes in one launch.  Work items (frame, tile) are
// listed in `items` (ordered by the host); workgroups take them in order from
// a device counter (persistent form: gridDim.x < n_items) or one each
// (gridDim.x == n_items).  Each item runs the whole transmit loop of its tile
// (no split).  When a frame's last tile is done, its completion time
// (wall clock) is recorded.
template <bool CUBIC, bool FOCUS, bool CF, bool SINAPO, int IQK>
__global__ __launch_bounds__(256, 4) void hercules_batch_kernel(const BatchFrameDev* __restrict__ frames,
                                                             const int2* __restrict__ items, int n_items,
                                                             int* __restrict__ next, int* __restrict__ done,
                                                             long long* __restrict__ finish)
{
    __shared__ float4 win[HPOOL_BYTES / sizeof(float4)];
    __shared__ int wlo_s[HSB];
    __shared__ int wmode_s[HSB];
    __shared__ int woff_s[HSB];
    __shared__ int wlen_s[HSB];
    __shared__ int item_s;
    const bool queue = gridDim.x < (unsigned)n_items;
    int it = blockIdx.x;
    while (true) {
        if (queue) {
            __syncthreads();
            if (threadIdx.x == 0) item_s = atomicAdd(next, 1);
            __syncthreads();
            it = item_s;
        }
        if (it >= n_items) break;
        const int2 w = items[__builtin_amdgcn_readfirstlane(it)];
        const int f = __builtin_amdgcn_readfirstlane(w.x), tile = __builtin_amdgcn_readfirstlane(w.y);
        const BatchFrameDev& fr = frames[f];
        const DasParams p = fr.p;
        const int per_z = fr.nxt * fr.nyt;
        const int wz = tile / per_z, wxy = tile - wz * per_z;
        const int bx = (wxy % fr.nxt) * HT_X, by = (wxy / fr.nxt) * HT_Y, bz = (fr.nzt - 1 - wz) * HT_Z;
        const int G = p.readi_groups > 1 ? p.readi_groups : 1;
        if (tile_never_clamps<FOCUS>(p, bx, by, bz))
            hercules_tile<CUBIC, FOCUS, CF, SINAPO, IQK, false>(fr.iq, fr.iq_inv, p, bx, by, bz, 0, G * p.T, fr.vol,
                                                                nullptr, win, wlo_s, wmode_s, woff_s, wlen_s);
        else
            hercules_tile<CUBIC, FOCUS, CF, SINAPO, IQK, true>(fr.iq, fr.iq_inv, p, bx, by, bz, 0, G * p.T, fr.vol,
                                                               nullptr, win, wlo_s, wmode_s, woff_s, wlen_s);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(done + f, 1) == fr.tiles - 1) finish[f] = wall_clock64();
        }
        if (!queue) break;
    }
}

// Zero the batch counters and record the start time.
__global__ void batch_start_kernel(int* counters, int n, long long* start)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
    if (threadIdx.x == 0) *start = wall_clock64();
}

// Sum of the per-split partials in split order (deterministic), then the
// coherence factor when it applies (CF: float4 partials with the incoherent
// sum; otherwise float2).
template <bool CF>
__global__ __launch_bounds__(256) void hercules_reduce(const void* __restrict__ part, int splits, size_t nvox,
                                                       float cw, float2* __restrict__ vol)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float ax = 0.0f, ay = 0.0f, inc = 0.0f;
    for (int s = 0; s < splits; ++s) {
        if constexpr (CF) {
            const float4 q = static_cast<const float4*>(part)[(size_t)s * nvox + v];
            ax += q.x;
            ay += q.y;
            inc += q.z;
        } else {
            const float2 q = static_cast<const float2*>(part)[(size_t)s * nvox + v];
            ax += q.x;
            ay += q.y;
        }
    }
    if constexpr (CF) {
        float cf = powf((ax * ax + ay * ay) / inc, cw);
        if (isnan(cf)) cf = 0.0f;
        ax *= cf;
        ay *= cf;
    }
    vol[v] = make_float2(ax, ay);
}

// The per-voxel kernels below (FORCES, UFORCES, RCA_TPW) work on groups of
// 64 voxels (one per lane).  A group's sum is split over nw waves of one
// workgroup (an outer loop of the sum: wave w takes iterations w, w + nw,
// ...), so that small 2-D grids still fill the chip; the partial sums meet
// in LDS and are added in wave order (deterministic).  A workgroup has
// max(4, nw) waves, i.e. 4 / nw groups when nw < 4.
constexpr int VOX_MAX_WAVES = 16;

inline unsigned vox_block(int nw) { return 64u * (unsigned)(nw < 4 ? 4 : nw); }
inline unsigned vox_grid(long long nvox, int nw)
{
    const long long groups = (nvox + 63) / 64, per_block = nw < 4 ? 4 / nw : 1;
    return (unsigned)((groups + per_block - 1) / per_block);
}

// Calls f(std::integral_constant<int, nw>) for nw in {1, 2, 4, 8, 16}.
template <class F>
void with_waves(int nw, F&& f)
{
    switch (nw) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: f(std::integral_constant<int, 16>{}); break;
    }
}

struct VoxelLane {
    long long v;  // voxel (clamped to the grid; `valid` says whether it is stored)
    bool valid;
    int w, nw;    // wave within the voxel group, waves per voxel group
};

__device__ __forceinline__ VoxelLane voxel_lane(long long nvox, int nw)
{
    VoxelLane r;
    const int wb = threadIdx.x >> 6;  // wave of the workgroup
    const int per_block = (int)(blockDim.x >> 6) / nw;
    r.v = ((long long)blockIdx.x * per_block + wb / nw) * 64 + (threadIdx.x & 63);
    r.valid = r.v < nvox;
    if (!r.valid) r.v = nvox - 1;
    r.w = wb % nw;
    r.nw = nw;
    return r;
}

// Voxel groups as 4 x 16 (x, y) patches of one depth plane (RCA_TPW): the 64
// lanes of a group then read samples a few cache lines apart (the receive
// delay of a column does not depend on y, the transmit delay moves slowly
// with y), where 64 voxels along x spread their reads over ~8 lines.  On the
// 128^3 volume 4 x 16 measured best (cubic / linear 8.3 / 5.5 ms; 8 x 8:
// 8.6 / 5.9, 2 x 32: 8.15 / 5.7, 16 x 4: 8.8 / 6.4, 64 x 1: 10.1 / 6.4).
// Lanes past the grid's x or y edge compute a clamped voxel and store nothing.
inline long long patch_groups(int Nx, int Ny, int Nz) { return (long long)((Nx + 3) / 4) * ((Ny + 15) / 16) * Nz; }

__device__ __forceinline__ VoxelLane voxel_lane_patch(int Nx, int Ny, int Nz, int nw, int& ix, int& iy, int& iz)
{
    VoxelLane r;
    const int wb = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per_block = (int)(blockDim.x >> 6) / nw;
    const long long g = (long long)blockIdx.x * per_block + wb / nw;
    const int npx = (Nx + 3) / 4, npy = (Ny + 15) / 16;
    const long long pz = g / ((long long)npx * npy);
    const int rem = (int)(g - pz * npx * npy);
    const int py = rem / npx, px = rem - py * npx;
    ix = px * 4 + (lane & 3);
    iy = py * 16 + (lane >> 2);
    r.valid = ix < Nx && iy < Ny && pz < Nz;
    ix = min(ix, Nx - 1);
    iy = min(iy, Ny - 1);
    iz = (int)min(pz, (long long)Nz - 1);
    r.v = ((long long)iz * Ny + iy) * Nx + ix;
    r.w = wb % nw;
    r.nw = nw;
    return r;
}

__device__ __forceinline__ float2 apply_cf(float2 acc, float inc, float cw)
{
    float cf = powf((acc.x * acc.x + acc.y * acc.y) / inc, cw);
    if (isnan(cf)) cf = 0.0f;
    return make_float2(acc.x * cf, acc.y * cf);
}

// Adds the waves' partials of each voxel (wave 0, 1, ... in order), applies
// the coherence factor and stores.  Every thread of the workgroup calls it.
__device__ __forceinline__ void store_voxel(const VoxelLane& vl, float2 acc, float inc, float cw,
                                            float2* __restrict__ vol)
{
    __shared__ float4 part[VOX_MAX_WAVES * 64];
    if (vl.nw > 1) {
        const int l = threadIdx.x & 63, wb = threadIdx.x >> 6;
        part[wb * 64 + l] = make_float4(acc.x, acc.y, inc, 0.0f);
        __syncthreads();
        if (vl.w != 0) return;
        acc = make_float2(0.0f, 0.0f);
        inc = 0.0f;
        for (int k = 0; k < vl.nw; ++k) {
            const float4 q = part[(wb + k) * 64 + l];
            acc.x += q.x;
            acc.y += q.y;
            inc += q.z;
        }
    }
    if (vl.valid) vol[vl.v] = apply_cf(acc, inc, cw);
}

// FORCES (templated_beamformer.cuh:219-324): the volume is y-invariant in the
// reference's arithmetic (y only selects the output row).  The waves split
// the receive-channel loop.
template <bool CUBIC, int NW>
__global__ __launch_bounds__(1024) void forces_kernel(const void* __restrict__ iq, const DasParams p,
                                                      float2* __restrict__ vol)
{
    const VoxelLane vl = voxel_lane((long long)p.Nx * p.Ny * p.Nz, NW);
    const int ix = (int)(vl.v % p.Nx);
    const int iz = (int)(vl.v / ((long long)p.Nx * p.Ny));
    const float vx = p.min_x + (float)ix * p.res_x;
    const float vz = p.min_z + (float)iz * p.res_z;
    const int G = p.readi_groups > 1 ? p.readi_groups : 1;
    float2 acc = make_float2(0.0f, 0.0f);
    float inc = 0.0f;
    const float vz2 = vz * vz;
    for (int c = vl.w; c < p.C; c += NW) {
        const float ec = fmaf((float)c, p.pitch_x, p.xdc_min_x);
        const float rx = ec - vx;
        const float rxd = fsqrt(fmaf(rx, rx, vz2));
        float rx_apo;
        if (p.apo_type == RX_HANN || p.apo_type == TX_TO_SIN) rx_apo = hann_from_a(p.fn_rx * (fabsf(rx) / vz) * 0.5f);
        else rx_apo = sin_apo_to(ec, p.xdc_max_x, p.to_power);
        if (!(rx_apo > 0.0f)) continue;
        for (int g = 0; g < G; ++g) {
            const bool neg = G > 1 && ((p.readi_bits >> g) & 1ull);
            for (int t = 0; t < p.T; ++t) {
                const float et = fmaf((float)(g * p.T + t), p.pitch_x, p.xdc_min_x);
                const float tx = vx - et;
                float tx_apo;
                if (p.apo_type == BOTH_TO_SIN || p.apo_type == TX_TO_SIN) tx_apo = sin_apo_to(et, p.xdc_max_x, p.to_power);
                else tx_apo = hann_from_a(p.fn_tx * (fabsf(tx) / vz) * 0.5f);
                float apo = rx_apo * tx_apo;
                float x = fmaf(fsqrt(fmaf(tx, tx, vz2)) + rxd, p.spm, (float)p.delay);
                x = fminf(fmaxf(x, 1.0f), (float)p.S - 2.0f);
                const float2 val = interp_global<CUBIC>(iq, ((size_t)t * p.C + c) * p.S, x, p.S);
                if (neg) apo = -apo;
                acc.x = fmaf(apo, val.x, acc.x);
                acc.y = fmaf(apo, val.y, acc.y);
                inc = fmaf(val.x, val.x, fmaf(val.y, val.y, inc));
            }
        }
    }
    store_voxel(vl, acc, inc, p.cw, vol);
}

// ---------------------------------------------------------------------------
// UFORCES and RCA_TPW (beamformer_kernels.cu:94-172, :489-546): one thread
// per voxel, receive element e at x_min + p/2 + e p on the x axis (x_min is
// already the first element's centre, so the reference's extra half pitch is
// kept), coherence factor over the apodized taps.
// ---------------------------------------------------------------------------

// Weight of the first used UFORCES transmit (beamformer_kernels.cu:148-151).
constexpr float UF_FIRST_SCALE = 0.088388347648318f;


// UFORCES: data row block t + 1 (block 0 is a discarded flash) was fired from
// element sparse_elements[t].  The transmit vector is calc_tx_distance(vox,
// focal, type) (beamformer_utils.cuh:8-32) with focal = (x_min + p/2, vox.y,
// 0), shifted by tx_tab[t].x pitches in x; path |rx| + |tx| sign(vox.z),
// index clamped to [0, S-2], Catmull-Rom over the flat buffer, Hann receive
// weight without a threshold (taps of weight cos^2(pi/2) are skipped).
//
// Precision as in the RCA_TPW kernel below: per voxel and transmit the
// channel-independent part (z + |tx| sign) spm + delay is FP64 and split
// into an integer and a fraction; per channel only (|rx| - z) spm =
// rx_x^2 / (|rx| + z) spm is FP32.  The waves split the transmit loop.
template <bool CUBIC, int NW>
__global__ __launch_bounds__(1024) void uforces_kernel(const float2* __restrict__ iq, const DasParams p,
                                                       float2* __restrict__ vol)
{
    const VoxelLane vl = voxel_lane((long long)p.Nx * p.Ny * p.Nz, NW);
    const long long v = vl.v;
    const int ix = (int)(v % p.Nx), iy = (int)((v / p.Nx) % p.Ny), iz = (int)(v / ((long long)p.Nx * p.Ny));
    const double vxd = (double)p.min_x + ix * (double)p.res_x;
    const double vyd = (double)p.min_y + iy * (double)p.res_y;
    const double vzd = (double)p.min_z + iz * (double)p.res_z;
    const double half = 0.5 * (double)p.pitch_x;
    const double fx = (double)p.xdc_min_x + half;
    double tx0 = 0.0, ty = 0.0, tz = 0.0;
    if (p.focus_type == 3) {         // XZ_FOCUS at depth 0
        tx0 = fx - vxd;
        tz = -vzd;
    } else if (p.focus_type == 4) {  // YZ_FOCUS: focal.y = vox.y
        tz = -vzd;
    } else if (p.focus_type == 1) {  // XZ_PLANE: focal.x taken as the angle
        tx0 = vxd * sin(fx);
        tz = vzd * cos(fx);
    } else if (p.focus_type == 2) {  // YZ_PLANE: focal.y taken as the angle
        ty = vyd * sin(vyd);
        tz = vzd * cos(vyd);
    }
    const double sgn = copysign(1.0, vzd), tyz2 = fma(ty, ty, tz * tz);
    const float vz = (float)vzd, vz2 = vz * vz;
    const float rx0 = (float)((double)p.xdc_min_x - vxd + half);
    const int n = p.T * p.C * p.S;
    float2 acc = make_float2(0.0f, 0.0f);
    float inc = 0.0f;
    for (int t = vl.w; t + 1 < p.T; t += NW) {
        const double txx = fma((double)p.tx_tab[t].x, (double)p.pitch_x, tx0);
        const double X0 = (vzd + sqrt(fma(txx, txx, tyz2)) * sgn) * (double)p.spm + (double)p.delay;
        const double I0 = fmin(fmax(floor(X0), -1048576.0), 1048576.0);
        const float f0 = (float)(X0 - I0), lo = (float)(-I0), hi = (float)((double)p.S - 2.0 - I0);
        const int i0 = (int)I0;
        const float sc = t == 0 ? UF_FIRST_SCALE : 1.0f;
        const int row0 = (t + 1) * p.C * p.S;
        for (int e = 0; e < p.C; ++e) {
            const float rxx = fmaf((float)e, p.pitch_x, rx0);
            const float a = p.fn_rx * (fabsf(rxx) / vz) * 0.5f;
            const bool in = a < 0.5f;
            if (!__any(in)) continue;
            const float w = in ? hann_from_a(a) * sc : 0.0f;
            const float rr = rxx * rxx;
            const float den = fsqrt(rr + vz2) + vz;
            const float excess = den > 0.0f ? rr * __builtin_amdgcn_rcpf(den) * p.spm : 0.0f;
            const float x = __builtin_amdgcn_fmed3f(f0 + excess, lo, hi);
            const float fl = floorf(x);
            const float tt = x - fl;
            // sample in [0, S-2] of row (t + 1, e); the integer clamp only acts
            // on non-finite input; j = -1 reads the previous row's last sample
            const int b = row0 + e * p.S + clampi(i0 + (int)fl, 0, p.S - 2);
            float2 val;
            if constexpr (CUBIC)
                val = cubic_pairs(iq, b, n, tt);
            else
                val = lerp2(iq[b], iq[b + 1], tt);
            const float qx = val.x * w, qy = val.y * w;
            acc.x += qx;
            acc.y += qy;
            inc = fmaf(qx, qx, fmaf(qy, qy, inc));
        }
    }
    store_voxel(vl, acc, inc, p.cw, vol);
}

// Steering table of RCA_TPW in FP64: (sin, cos) of each transmit's angle.
__global__ void tpw_table_kernel(const DasParams p, double2* __restrict__ sc)
{
    for (int t = threadIdx.x; t < p.T; t += blockDim.x) {
        double s, c;
        sincos((double)p.tx_tab[t].x, &s, &c);
        sc[t] = make_double2(s, c);
    }
}

constexpr int TPW_QB = 8;  // angles per register block

// RCA_TPW: tilted plane waves with steering (sin, cos) = sc[a] along y
// (YZ_PLANE) or x (XZ_PLANE); channels of Hann weight below 0.1 skipped; index
// clamped to [1, S-2]; Catmull-Rom read; the sample is rotated by
// exp(+2 pi i x fc/fs) (rotate_iq, :482-487) before the weight.
//
// Precision: the index is about (z + lat sin + z cos) spm, up to a few
// thousand, where an FP32 sum would round to ~1e-4 samples.  Per voxel and
// angle the part that does not depend on the channel, X0 = (z + lat sin +
// z cos) spm + delay, is formed in FP64 and split into an integer I0 and a
// fraction; per channel only the receive excess over the depth,
// (|rx| - z) spm = rx_x^2 / (|rx| + z) spm (at most a few hundred samples),
// is FP32.  The tap's index is I0 + (fraction + excess), the clamp bounds are
// moved by I0 (integers, exact), and the rotation's phase is frac(I0 fc/fs)
// (FP64) + (fraction + excess) fc/fs.  The waves split the channel loop.
template <bool CUBIC, int NW>
__global__ __launch_bounds__(1024) void tpw_kernel(const float2* __restrict__ iq, const DasParams p,
                                                   const double2* __restrict__ sc, float2* __restrict__ vol)
{
    int ix, iy, iz;
    const VoxelLane vl = voxel_lane_patch(p.Nx, p.Ny, p.Nz, NW, ix, iy, iz);
    const double vxd = (double)p.min_x + ix * (double)p.res_x;
    const double vyd = (double)p.min_y + iy * (double)p.res_y;
    const double vzd = (double)p.min_z + iz * (double)p.res_z;
    const double latd = p.focus_type == 2 ? vyd : vxd;
    const double spmd = p.spm;
    const float vz = (float)vzd, vz2 = vz * vz;
    const float rx0 = (float)((double)p.xdc_min_x + 0.5 * (double)p.pitch_x - vxd);
    const float fcfs = (float)p.fc_fs;
    const int n = p.T * p.C * p.S;
    float2 acc = make_float2(0.0f, 0.0f);
    float inc = 0.0f;
    for (int q0 = 0; q0 < p.T; q0 += TPW_QB) {
        const int nq = min(TPW_QB, p.T - q0);
        float f0[TPW_QB], ph0[TPW_QB], lo[TPW_QB], hi[TPW_QB];
        int i0[TPW_QB], row[TPW_QB];
#pragma unroll
        for (int j = 0; j < TPW_QB; ++j) {
            const int q = min(q0 + j, p.T - 1);
            const double2 s = sc[q];
            const double X0 = fma(latd, s.x, fma(vzd, s.y, vzd)) * spmd + (double)p.delay;
            // bounded so that every integer below stays in range for any input
            const double I0 = fmin(fmax(floor(X0), -1048576.0), 1048576.0);
            const double pI = I0 * p.fc_fs;
            f0[j] = (float)(X0 - I0);
            ph0[j] = (float)(pI - floor(pI));
            lo[j] = (float)(1.0 - I0);
            hi[j] = (float)((double)p.S - 2.0 - I0);
            i0[j] = (int)I0;
            row[j] = q * p.C * p.S;
        }
        for (int c = vl.w; c < p.C; c += NW) {
            const float rxx = fmaf((float)c, p.pitch_x, rx0);
            const float a = p.fn_rx * (fabsf(rxx) / vz) * 0.5f;
            // cos^2(pi min(a, 1/2)) >= 0.1  <=>  a < acos(sqrt(0.1)) / pi
            const bool in = a < HANN_A_THR;
            if (!__any(in)) continue;
            const float w = in ? hann_from_a(a) : 0.0f;
            const float rr = rxx * rxx;
            const float den = fsqrt(rr + vz2) + vz;
            const float excess = den > 0.0f ? rr * __builtin_amdgcn_rcpf(den) * p.spm : 0.0f;
            const int crow = c * p.S;
#pragma unroll
            for (int j = 0; j < TPW_QB; ++j) {
                if (j >= nq) break;
                const float x = __builtin_amdgcn_fmed3f(f0[j] + excess, lo[j], hi[j]);
                const float fl = floorf(x);
                const float t = x - fl;
                // sample in [1, S-2] of row (q0 + j, c); the integer clamp only
                // acts on non-finite input
                const int i = row[j] + crow + clampi(i0[j] + (int)fl, 1, p.S - 2);
                float2 val;
                if constexpr (CUBIC) val = cubic_pairs(iq, i, n, t);
                else val = lerp2(iq[i], iq[i + 1], t);
                const float ph = __builtin_amdgcn_fractf(fmaf(x, fcfs, ph0[j]));
                const float cs = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
                const float qx = (val.x * cs - val.y * sn) * w, qy = (val.x * sn + val.y * cs) * w;
                acc.x += qx;
                acc.y += qy;
                inc = fmaf(qx, qx, fmaf(qy, qy, inc));
            }
        }
    }
    store_voxel(vl, acc, inc, p.cw, vol);
}

// Launch tuning (tuning.hpp): the transmit split of the single-frame HERCULES
// launch (the heuristic das_splits, or, with timed tuning on, the fastest of
// 1, 2, 4, 8 and 16 splits) and the waves per voxel group of the per-voxel
// kernels (vox_waves, or the fastest of 1 to 16), measured the first time a
// shape is seen (outside stream capture).
namespace {
constexpr int TUNE_MAX_SPLITS = 16;

int max_splits(int gt_total) { return std::max(1, std::min(TUNE_MAX_SPLITS, gt_total / 8)); }

std::vector<int> pow2_upto(int n)
{
    std::vector<int> c;
    for (int v = 1; v <= n; v *= 2) c.push_back(v);
    return c;
}

// Waves per 64-voxel group: enough for ~8192 waves in flight (32 per CU),
// at most VOX_MAX_WAVES.
int vox_waves(long long nvox)
{
    const long long groups = (nvox + 63) / 64;
    int nw = 1;
    while (nw < VOX_MAX_WAVES && groups * nw < 8192) nw *= 2;
    return nw;
}
}  // namespace

bool herc_generic_apo(const DasParams& p)
{
    if (p.apo_type == RX_TO_SIN) return true;
    // the depths as the kernel forms them (FP64 grid, rounded to float)
    const float z0 = (float)((double)p.min_z), z1 = (float)((double)p.min_z + (double)(p.Nz - 1) * (double)p.res_z);
    return !(p.fn_rx >= 0.0f) || !(fminf(z0, z1) > 0.0f);
}

size_t das_workspace_bytes(const DasParams& p)
{
    if (p.sequence == RCA_TPW) return (size_t)MAX_ACQ_COUNT * sizeof(double2);
    if (p.sequence != HERCULES) return 0;
    const long long tiles = (long long)((p.Nx + HT_X - 1) / HT_X) * ((p.Ny + HT_Y - 1) / HT_Y) * ((p.Nz + HT_Z - 1) / HT_Z);
    const int G = p.readi_groups > 1 ? p.readi_groups : 1;
    int splits = das_splits(tiles, G * p.T);
    if (autotune_enabled()) splits = std::max(splits, max_splits(G * p.T));
    return splits > 1 ? (size_t)splits * p.Nx * p.Ny * p.Nz * sizeof(float4) : 0;
}

int das_splits(long long tiles, int gt_total)
{
    // aim for >= 16384 workgroups (4 resident per CU x 256 CUs x 16 rounds:
    // the deepest tiles carry ~z^2 more taps, so many small items balance the
    // chip), at least 8 transmits per split.  On the 128^3 frame (1024 tiles,
    // 128 transmits) this is 16, the timed tuner's choice too (16 / 8 / 4 / 2 /
    // 1 splits: 7.39 / 8.2 / 9.3 / 10.2 / 10.5 ms).
    long long s = (16384 + tiles - 1) / tiles;
    s = std::min<long long>(s, std::max(1, gt_total / 8));
    return (int)std::max<long long>(1, std::min<long long>(s, 16));
}

namespace {
template <bool CU, bool FO, bool CFV, bool SA>
bool launch_herc_fmt(int fmt, dim3 grid, hipStream_t st, const void* iq, float inv, const DasParams& p,
                     const HercGrid& hg, float2* vol, void* part)
{
    if (fmt == IQ_F32) {
        hercules_kernel<CU, FO, CFV, SA, IQ_F32><<<grid, 256, 0, st>>>(iq, inv, p, hg, vol, part);
        return true;
    }
    if constexpr (SA) {
        CUT_FAIL("DAS: staged IQ formats support Hann apodization of volumes below the array (z > 0, F-number >= 0) only");
    } else {
        if (fmt == IQ_F16) hercules_kernel<CU, FO, CFV, SA, IQ_F16><<<grid, 256, 0, st>>>(iq, inv, p, hg, vol, part);
        else hercules_kernel<CU, FO, CFV, SA, IQ_BF16><<<grid, 256, 0, st>>>(iq, inv, p, hg, vol, part);
        return true;
    }
}

template <bool CU, bool FO, bool CFV>
bool launch_herc_sa(bool sa, int fmt, dim3 grid, hipStream_t st, const void* iq, float inv, const DasParams& p,
                    const HercGrid& hg, float2* vol, void* part)
{
    return sa ? launch_herc_fmt<CU, FO, CFV, true>(fmt, grid, st, iq, inv, p, hg, vol, part)
              : launch_herc_fmt<CU, FO, CFV, false>(fmt, grid, st, iq, inv, p, hg, vol, part);
}

template <bool CU, bool FO>
bool launch_herc_cf(bool cf, bool sa, int fmt, dim3 grid, hipStream_t st, const void* iq, float inv,
                    const DasParams& p, const HercGrid& hg, float2* vol, void* part)
{
    return cf ? launch_herc_sa<CU, FO, true>(sa, fmt, grid, st, iq, inv, p, hg, vol, part)
              : launch_herc_sa<CU, FO, false>(sa, fmt, grid, st, iq, inv, p, hg, vol, part);
}

template <bool CU>
bool launch_herc_fo(bool fo, bool cf, bool sa, int fmt, dim3 grid, hipStream_t st, const void* iq, float inv,
                    const DasParams& p, const HercGrid& hg, float2* vol, void* part)
{
    return fo ? launch_herc_cf<CU, true>(cf, sa, fmt, grid, st, iq, inv, p, hg, vol, part)
              : launch_herc_cf<CU, false>(cf, sa, fmt, grid, st, iq, inv, p, hg, vol, part);
}
}  // namespace

namespace {
template <bool CU, bool FO, bool CFV, bool SA>
bool launch_batch_fmt(int fmt, int grid, hipStream_t st, const BatchFrameDev* fr, const int2* items, int n, int* next,
                      int* done, long long* finish)
{
    if (fmt == IQ_F32) {
        hercules_batch_kernel<CU, FO, CFV, SA, IQ_F32><<<grid, 256, 0, st>>>(fr, items, n, next, done, finish);
        return true;
    }
    if constexpr (SA) {
        CUT_FAIL("DAS: staged IQ formats support Hann apodization of volumes below the array (z > 0, F-number >= 0) only");
    } else {
        if (fmt == IQ_F16)
            hercules_batch_kernel<CU, FO, CFV, SA, IQ_F16><<<grid, 256, 0, st>>>(fr, items, n, next, done, finish);
        else
            hercules_batch_kernel<CU, FO, CFV, SA, IQ_BF16><<<grid, 256, 0, st>>>(fr, items, n, next, done, finish);
        return true;
    }
}
template <bool CU, bool FO, bool CFV>
bool launch_batch_sa(bool sa, int fmt, int grid, hipStream_t st, const BatchFrameDev* fr, const int2* items, int n,
                     int* next, int* done, long long* finish)
{
    return sa ? launch_batch_fmt<CU, FO, CFV, true>(fmt, grid, st, fr, items, n, next, done, finish)
              : launch_batch_fmt<CU, FO, CFV, false>(fmt, grid, st, fr, items, n, next, done, finish);
}
template <bool CU, bool FO>
bool launch_batch_cf(bool cf, bool sa, int fmt, int grid, hipStream_t st, const BatchFrameDev* fr, const int2* items,
                     int n, int* next, int* done, long long* finish)
{
    return cf ? launch_batch_sa<CU, FO, true>(sa, fmt, grid, st, fr, items, n, next, done, finish)
              : launch_batch_sa<CU, FO, false>(sa, fmt, grid, st, fr, items, n, next, done, finish);
}
template <bool CU>
bool launch_batch_fo(bool fo, bool cf, bool sa, int fmt, int grid, hipStream_t st, const BatchFrameDev* fr,
                     const int2* items, int n, int* next, int* done, long long* finish)
{
    return fo ? launch_batch_cf<CU, true>(cf, sa, fmt, grid, st, fr, items, n, next, done, finish)
              : launch_batch_cf<CU, false>(cf, sa, fmt, grid, st, fr, items, n, next, done, finish);
}
}  // namespace

void das_tile_grid(const DasParams& p, int& nxt, int& nyt, int& nzt)
{
    nxt = (p.Nx + HT_X - 1) / HT_X;
    nyt = (p.Ny + HT_Y - 1) / HT_Y;
    nzt = (p.Nz + HT_Z - 1) / HT_Z;
}

bool launch_batch_start(int* d_counters, int n, long long* d_start, hipStream_t st)
{
    batch_start_kernel<<<1, 256, 0, st>>>(d_counters, n, d_start);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool launch_das_batch(const BatchFrameDev* d_frames, const int2* d_items, int n_items, int grid, int* d_next,
                      int* d_done, long long* d_finish, const DasOptions& o, bool focus, bool sinapo, int fmt,
                      hipStream_t st)
{
    if (n_items <= 0) return true;
    if (fmt != IQ_F32 && fmt != IQ_F16 && fmt != IQ_BF16) CUT_FAIL("DAS: unknown IQ format " << fmt);
    grid = std::max(1, std::min(grid, n_items));
    const bool cf = o.cf_hercules != 0;
    const bool ok = o.cubic ? launch_batch_fo<true>(focus, cf, sinapo, fmt, grid, st, d_frames, d_items, n_items,
                                                    d_next, d_done, d_finish)
                            : launch_batch_fo<false>(focus, cf, sinapo, fmt, grid, st, d_frames, d_items, n_items,
                                                     d_next, d_done, d_finish);
    if (!ok) return false;
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool launch_das(const void* iq, int iq_fmt, float iq_scale, const DasParams& p, const DasOptions& o, float2* vol,
                void* workspace, size_t workspace_bytes, hipStream_t st, bool allow_tune)
{
    if (iq_fmt != IQ_F32 && iq_fmt != IQ_F16 && iq_fmt != IQ_BF16) CUT_FAIL("DAS: unknown IQ format " << iq_fmt);
    if (!(iq_scale > 0.0f)) CUT_FAIL("DAS: IQ scale must be positive");
    const float inv = 1.0f / iq_scale;
    if (p.Nx <= 0 || p.Ny <= 0 || p.Nz <= 0) CUT_FAIL("DAS: empty output grid");
    if (p.sequence == HERCULES) {
        if (p.focus_type != 2 && p.focus_type != 4) CUT_FAIL("Invalid focus for HERCULES.");
        HercGrid hg;
        hg.nxt = (p.Nx + HT_X - 1) / HT_X;
        hg.nyt = (p.Ny + HT_Y - 1) / HT_Y;
        hg.nzt = (p.Nz + HT_Z - 1) / HT_Z;
        const long long tiles = (long long)hg.nxt * hg.nyt * hg.nzt;
        const int G = p.readi_groups > 1 ? p.readi_groups : 1;
        const size_t nvox = (size_t)p.Nx * p.Ny * p.Nz;
        const bool focus = p.focus_type == 4;
        const bool sinapo = herc_generic_apo(p);
        const bool cf = o.cf_hercules != 0;
        const int dflt = das_splits(tiles, G * p.T);
        auto fits = [&](int sp) {
            return sp <= 1 || (workspace && workspace_bytes >= (size_t)sp * nvox * (cf ? sizeof(float4) : sizeof(float2)));
        };
        auto run = [&](int splits) -> bool {
            // a choice tuned with a larger workspace falls back to the largest
            // power-of-two split that fits
            while (splits > 1 && !fits(splits)) splits >>= 1;
            hg.splits = splits;
            hg.gt_per_split = (G * p.T + splits - 1) / splits;
            void* part = splits > 1 ? workspace : nullptr;
            if (tiles * splits > 0x7fffffffLL) CUT_FAIL("DAS: grid too large");
            dim3 grid((unsigned)(tiles * splits));
            const bool ok = o.cubic ? launch_herc_fo<true>(focus, cf, sinapo, iq_fmt, grid, st, iq, inv, p, hg, vol, part)
                                    : launch_herc_fo<false>(focus, cf, sinapo, iq_fmt, grid, st, iq, inv, p, hg, vol, part);
            if (!ok) return false;
            CUT_HIP_CHECK(hipGetLastError());
            if (splits > 1)
            {
                const unsigned rg = (unsigned)((nvox + 255) / 256);
                if (cf) hercules_reduce<true><<<rg, 256, 0, st>>>(part, splits, nvox, p.cw, vol);
                else hercules_reduce<false><<<rg, 256, 0, st>>>(part, splits, nvox, p.cw, vol);
            }
            return true;
        };
        const TuneKey key{p.Nx, p.Ny, p.Nz, p.T, p.C, p.S, G, iq_fmt, o.cubic, cf, focus, sinapo};
        const int cand = fits(max_splits(G * p.T)) ? max_splits(G * p.T) : 1;
        if (!tuned_run(TUNE_DAS_SPLITS, key, dflt, pow2_upto(cand), st, allow_tune, run)) return false;
    } else if (p.sequence == FORCES) {
        if (p.readi_groups > 1 && p.encoding != HADAMARD) CUT_FAIL("FORCES only supports Hadamard encoding at this time.");
        if (iq_fmt != IQ_F32) CUT_FAIL("DAS: staged IQ formats are supported for HERCULES only");
        const long long nvox = (long long)p.Nx * p.Ny * p.Nz;
        auto run = [&](int nw) -> bool {
            const unsigned grid = vox_grid(nvox, nw), block = vox_block(nw);
            with_waves(nw, [&](auto w) {
                if (o.cubic) forces_kernel<true, w()><<<grid, block, 0, st>>>(iq, p, vol);
                else forces_kernel<false, w()><<<grid, block, 0, st>>>(iq, p, vol);
            });
            CUT_HIP_CHECK(hipGetLastError());
            return true;
        };
        const TuneKey key{FORCES, p.Nx, p.Ny, p.Nz, p.T, p.C, p.S, p.readi_groups, o.cubic, p.apo_type, 0, 0};
        if (!tuned_run(TUNE_VOX_WAVES, key, vox_waves(nvox), pow2_upto(VOX_MAX_WAVES), st, allow_tune, run))
            return false;
    } else if (p.sequence == UFORCES || p.sequence == RCA_TPW) {
        if (iq_fmt != IQ_F32) CUT_FAIL("DAS: staged IQ formats are supported for HERCULES only");
        if (p.T > MAX_ACQ_COUNT) CUT_FAIL("DAS: at most " << MAX_ACQ_COUNT << " transmits");
        if ((long long)p.T * p.C * p.S > 0x7fffffffLL) CUT_FAIL("DAS: IQ buffer too large");
        if (p.sequence == RCA_TPW && p.focus_type != 1 && p.focus_type != 2)
            CUT_FAIL("TPW beamforming requires a plane focus.");
        if (p.sequence == UFORCES && p.focus_type == 0) CUT_FAIL("Beamformer: UFORCES needs a transmit orientation.");
        const long long nvox = (long long)p.Nx * p.Ny * p.Nz;
        const float2* in = static_cast<const float2*>(iq);
        double2* sc = nullptr;
        if (p.sequence == RCA_TPW) {
            if (!workspace || workspace_bytes < (size_t)p.T * sizeof(double2))
                CUT_FAIL("DAS: workspace of " << (size_t)p.T * sizeof(double2) << " bytes needed");
            sc = static_cast<double2*>(workspace);
            tpw_table_kernel<<<1, 256, 0, st>>>(p, sc);
            CUT_HIP_CHECK(hipGetLastError());
        }
        auto run = [&](int nw) -> bool {
            const unsigned grid = p.sequence == UFORCES ? vox_grid(nvox, nw)
                                                        : vox_grid(64 * patch_groups(p.Nx, p.Ny, p.Nz), nw);
            const unsigned block = vox_block(nw);
            with_waves(nw, [&](auto w) {
                if (p.sequence == UFORCES) {
                    if (o.cubic) uforces_kernel<true, w()><<<grid, block, 0, st>>>(in, p, vol);
                    else uforces_kernel<false, w()><<<grid, block, 0, st>>>(in, p, vol);
                } else {
                    if (o.cubic) tpw_kernel<true, w()><<<grid, block, 0, st>>>(in, p, sc, vol);
                    else tpw_kernel<false, w()><<<grid, block, 0, st>>>(in, p, sc, vol);
                }
            });
            CUT_HIP_CHECK(hipGetLastError());
            return true;
        };
        const TuneKey key{p.sequence, p.Nx, p.Ny, p.Nz, p.T, p.C, p.S, p.focus_type, o.cubic, 0, 0, 0};
        if (!tuned_run(TUNE_VOX_WAVES, key, vox_waves(nvox), pow2_upto(VOX_MAX_WAVES), st, allow_tune, run))
            return false;
    } else {
        CUT_FAIL("Beamformer: Unsupported sequence ID " << p.sequence);
    }
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

}  // namespace cut
