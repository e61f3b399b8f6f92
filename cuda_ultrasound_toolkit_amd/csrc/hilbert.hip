// Hilbert transform with optional match filter.
//
// Semantics: cuda_toolkit/src/rf_processing/rf_ffts/hilbert_handler.cu:132-155
// and hilbert_kernels.cuh:17-40.  Per row x of length S:
//   X = DFT_S(x);  Y[k] = (X[k] * w_k) * F[k] for k <= S/2, 0 above;
//   w_0 = 0.75/S, w_k = 1.5/S (k >= 1, the Nyquist bin included);
//   y = unnormalised inverse DFT_S(Y).
//
// Implementation: Stockham auto-sort FFTs held in LDS.  Each thread owns 16
// (or 8: the FFT tile, tuned per length) complex values per pass (radix-16
// passes, one radix-2/4/8 pass when log2 S is not a multiple of 4), so
// S = 4096 takes three passes per transform.
// Power-of-two rows are taken in pairs: one forward transform of
// x_a + i x_b yields both spectra, then one inverse transform per row
// (three transforms per pair instead of four).  The first forward pass reads
// the input rows straight from HBM (coalesced; optionally converting raw
// int16/f32 through the channel map, so convert+Hilbert is one kernel when no
// decode is needed), the spectral weight and filter are applied while the
// two spectra are separated, and the last inverse pass writes each complex
// row straight to HBM.  4096-point rows (16 values per thread) run as a
// fixed grid of workgroups walking the row pairs, each requesting the next
// pair's rows while it transforms the current one (hilbert_pair_stream_kernel).
// Other lengths are a cyclic convolution with the
// operator's kernel, evaluated by power-of-two FFTs of the zero-padded row
// (hilbert_conv_kernel); rows longer than 4096 take HBM-staged four-step
// transforms.  LDS rows are padded by one complex per 16 so the radix-16
// transposes are free of bank conflicts.  Twiddles come from a table
// generated in FP64.
#include "common.hpp"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

namespace cut {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <bool INV> __device__ __forceinline__ float2 cmul_tw(float2 a, float2 w)
{
    if constexpr (INV) w.y = -w.y;
    return cmul(a, w);
}
// Twiddles of one radix-R butterfly, w_r = W^(r e) for r = 1..R-1 (W = e^(-2 pi i / S),
// table tw[j] = W^j): the powers W^e, W^2e, W^4e, W^8e come from the table
// (requested before the pass's data loads and barrier, so their latency is
// hidden), the others are products of at most three of them (error
// ~3e-7 relative against the table's 0.5 ulp).  Fifteen table reads per
// butterfly after the barrier left each pass waiting on L2.
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
template <int R>
struct TwPow {
    float2 p[ilog2c(R) > 0 ? ilog2c(R) : 1];
};
template <int R>
__device__ __forceinline__ TwPow<R> tw_pow_load(const float2* __restrict__ tw, int e, int mask)
{
    TwPow<R> t;
#pragma unroll
    for (int q = 0; q < ilog2c(R); ++q) t.p[q] = tw[(e << q) & mask];
    return t;
}
template <int R>
__device__ __forceinline__ void tw_pow_expand(const TwPow<R>& t, float2* w)
{
    w[1] = t.p[0];
    if constexpr (R > 2) {
        w[2] = t.p[1];
        w[3] = cmul(w[1], w[2]);
    }
    if constexpr (R > 4) {
        w[4] = t.p[2];
#pragma unroll
        for (int r = 5; r < 8; ++r) w[r] = cmul(w[r - 4], w[4]);
    }
    if constexpr (R > 8) {
        w[8] = t.p[3];
#pragma unroll
        for (int r = 9; r < 16; ++r) w[r] = cmul(w[r - 8], w[8]);
    }
}

// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a)
{
    if constexpr (INV) return make_float2(-a.y, a.x);
    else return make_float2(a.y, -a.x);
}

template <bool INV> __device__ __forceinline__ void dft2(float2& a, float2& b)
{
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3)
{
    float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
    float2 s13 = cadd(a1, a3), d13 = mul_mi<INV>(csub(a1, a3));
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, d13);
    a3 = csub(d02, d13);
}

// Exact constants of the 8th and 16th roots of unity.
#define CUT_R2 0.70710678118654752440f
#define CUT_C1 0.92387953251128675613f
#define CUT_S1 0.38268343236508977173f

template <bool INV> __device__ __forceinline__ float2 w16(float2 a, int e)
{
    // a * exp(-+ 2 pi i e / 16) for the exponents used by the 4x4 split
    float2 w;
    switch (e) {
        case 1: w = make_float2(CUT_C1, -CUT_S1); break;
        case 2: w = make_float2(CUT_R2, -CUT_R2); break;
        case 3: w = make_float2(CUT_S1, -CUT_C1); break;
        case 4: w = make_float2(0.0f, -1.0f); break;
        case 6: w = make_float2(-CUT_R2, -CUT_R2); break;
        case 9: w = make_float2(-CUT_C1, CUT_S1); break;
        default: return a;
    }
    return cmul_tw<INV>(a, w);
}

template <int R, bool INV> __device__ __forceinline__ void dft(float2* u)
{
    if constexpr (R == 2) {
        dft2<INV>(u[0], u[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(u[0], u[1], u[2], u[3]);
    } else if constexpr (R == 8) {
        // n = g + 2h, k = q + 4m
        float2 v0[4] = {u[0], u[2], u[4], u[6]};
        float2 v1[4] = {u[1], u[3], u[5], u[7]};
        dft4<INV>(v0[0], v0[1], v0[2], v0[3]);
        dft4<INV>(v1[0], v1[1], v1[2], v1[3]);
        // twiddle W8^(q) on the g = 1 branch
        v1[1] = cmul_tw<INV>(v1[1], make_float2(CUT_R2, -CUT_R2));
        v1[2] = mul_mi<INV>(v1[2]);
        v1[3] = cmul_tw<INV>(v1[3], make_float2(-CUT_R2, -CUT_R2));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u[q] = cadd(v0[q], v1[q]);
            u[q + 4] = csub(v0[q], v1[q]);
        }
    } else if constexpr (R == 16) {
        // n = g + 4h, k = q + 4m: inner DFT4 over h, twiddle W16^(g q), outer DFT4 over g
        float2 v[4][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            v[g][0] = u[g]; v[g][1] = u[g + 4]; v[g][2] = u[g + 8]; v[g][3] = u[g + 12];
            dft4<INV>(v[g][0], v[g][1], v[g][2], v[g][3]);
        }
#pragma unroll
        for (int g = 1; g < 4; ++g)
#pragma unroll
            for (int q = 1; q < 4; ++q) v[g][q] = w16<INV>(v[g][q], g * q);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float2 a0 = v[0][q], a1 = v[1][q], a2 = v[2][q], a3 = v[3][q];
            dft4<INV>(a0, a1, a2, a3);
            u[q] = a0; u[q + 4] = a1; u[q + 8] = a2; u[q + 12] = a3;
        }
    }
}

__device__ __forceinline__ int lds_idx(int i) { return i + (i >> 4); }

struct HilbertDev {
    const void* ptr;
    int stride;
    const int16_t* map;
    size_t R0;
    int C;
};

template <int KIND>
__device__ __forceinline__ float load_in(const HilbertDev& in, int row, int S, int idx)
{
    if constexpr (KIND == 0) {
        return ((const float*)in.ptr)[((size_t)row * S + idx) * in.stride];
    } else {
        const int t = row / in.C, c = row - t * in.C;
        const size_t off = (size_t)(uint16_t)in.map[c] * in.R0 + (size_t)t * S + idx;
        if constexpr (KIND == 1) return (float)((const int16_t*)in.ptr)[off];
        else return ((const float*)in.ptr)[off];
    }
}

// Start of a row's input (uniform when a workgroup holds one row, so that
// the per-element loads take a scalar base and a 32-bit offset) and the
// element at idx from it.
template <int KIND>
__device__ __forceinline__ const void* row_base(const HilbertDev& in, int row, int S)
{
    if constexpr (KIND == 0) {
        return (const float*)in.ptr + (size_t)row * S * in.stride;
    } else {
        const int t = row / in.C, c = row - t * in.C;
        const size_t off = (size_t)(uint16_t)in.map[c] * in.R0 + (size_t)t * S;
        if constexpr (KIND == 1) return (const int16_t*)in.ptr + off;
        else return (const float*)in.ptr + off;
    }
}

template <int KIND>
__device__ __forceinline__ float load_at(const void* base, unsigned stride, unsigned idx)
{
    if constexpr (KIND == 0) return ((const float*)base)[idx * stride];
    else if constexpr (KIND == 1) return (float)((const int16_t*)base)[idx];
    else return ((const float*)base)[idx];
}

// Value at index idx of a pass's input; a loader that also takes the slot
// m * R + r of the value in the thread's tile is a row held in registers.
template <typename LoadF>
__device__ __forceinline__ float2 pass_load(LoadF& load, int idx, int slot)
{
    if constexpr (std::is_invocable_v<LoadF&, int, int>) return load(idx, slot);
    else return load(idx);
}

// Twiddle powers of the thread's butterflies in a pass: from registers when
// the caller holds them (PRE: tpre, one per butterfly), else from the table.
template <int S, int R, int VPT, bool PRE>
__device__ __forceinline__ void pass_twiddles(int tid, int p, const float2* __restrict__ tw,
                                              const TwPow<R>* tpre, TwPow<R>* tp)
{
    constexpr int P = S / VPT;
    constexpr int PER = VPT / R;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int i = tid + m * P;
        if constexpr (PRE) tp[m] = tpre[m];
        else tp[m] = tw_pow_load<R>(tw, (i & (p - 1)) * (S / (p * R)), S - 1);
    }
}

// One Stockham pass over an S-point row held by P = S/VPT threads.  The
// thread holds VPT values (the FFT tile: 16 or 8): VPT/R butterflies of radix
// R.  Load phase -> (barrier) -> twiddle + DFT + store phase.
template <int S, int R, bool INV, int VPT, bool PRE = false, typename LoadF, typename StoreF>
__device__ __forceinline__ void stockham_pass(int tid, int p, const float2* __restrict__ tw, LoadF load,
                                              StoreF store, bool sync_between, const TwPow<R>* tpre = nullptr)
{
    constexpr int P = S / VPT;
    constexpr int NB = S / R;
    constexpr int PER = VPT / R;
    TwPow<R> tp[PER];
    if (p > 1) pass_twiddles<S, R, VPT, PRE>(tid, p, tw, tpre, tp);
    float2 u[VPT];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int i = tid + m * P;
#pragma unroll
        for (int r = 0; r < R; ++r) u[m * R + r] = pass_load(load, i + r * NB, m * R + r);
    }
    if (sync_between) __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int i = tid + m * P;
        const int k = i & (p - 1);
        if (p > 1) {
            float2 w[R];
            tw_pow_expand<R>(tp[m], w);
#pragma unroll
            for (int r = 1; r < R; ++r) u[m * R + r] = cmul_tw<INV>(u[m * R + r], w[r]);
        }
        dft<R, INV>(u + m * R);
        const int base = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) store(base + r * p, u[m * R + r]);
    }
}

// round-to-nearest-even FP32 -> BF16 bits
__device__ __forceinline__ uint32_t bf16_bits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// All passes of one N-point transform held in LDS: radix-VPT passes, then one
// radix 2..VPT/2 pass when log2 N is not a multiple of log2 VPT.  The first
// pass reads through first_load (global memory, or LDS when first_lds), the
// last writes through last_store; a barrier follows every pass whose output
// later passes or transforms read.
// With PRE, tpre holds the thread's twiddle powers of the radix-VPT passes
// after the first (one TwPow per pass: a radix-VPT pass is one butterfly per
// thread).
template <int LOGN, bool INV, int VPT = 16, bool PRE = false, typename LoadF, typename StoreF>
__device__ __forceinline__ void fft_passes(int tid, const float2* __restrict__ tw, float2* lds, LoadF first_load,
                                           bool first_lds, StoreF last_store, bool last_global,
                                           const TwPow<VPT>* tpre = nullptr)
{
    static_assert(VPT == 8 || VPT == 16, "FFT tile: 8 or 16 values per thread");
    constexpr int LV = VPT == 16 ? 4 : 3;
    constexpr int N = 1 << LOGN;
    constexpr int NPV = LOGN / LV;
    constexpr int RT = 1 << (LOGN % LV);
    auto load_lds = [&](int i) { return lds[lds_idx(i)]; };
    auto store_lds = [&](int i, float2 v) { lds[lds_idx(i)] = v; };
    int p = 1;
#pragma unroll
    for (int q = 0; q < NPV; ++q) {
        const bool last = (q == NPV - 1) && RT == 1;
        if (q == 0) {
            if (last) stockham_pass<N, VPT, INV, VPT>(tid, p, tw, first_load, last_store, first_lds);
            else stockham_pass<N, VPT, INV, VPT>(tid, p, tw, first_load, store_lds, first_lds);
        } else {
            const TwPow<VPT>* tq = PRE ? tpre + (q - 1) : nullptr;
            if (last) stockham_pass<N, VPT, INV, VPT, PRE>(tid, p, tw, load_lds, last_store, true, tq);
            else stockham_pass<N, VPT, INV, VPT, PRE>(tid, p, tw, load_lds, store_lds, true, tq);
        }
        if (!(last && last_global)) __syncthreads();
        p *= VPT;
    }
    if constexpr (RT > 1) {
        if (NPV == 0) stockham_pass<N, RT, INV, VPT>(tid, p, tw, first_load, last_store, first_lds);
        else stockham_pass<N, RT, INV, VPT>(tid, p, tw, load_lds, last_store, true);
        if (!last_global) __syncthreads();
    }
}

// Two transforms of the same length and direction in lockstep: one set of
// passes, barriers and twiddle loads for both, two independent butterfly
// chains per thread.
template <int S, int R, bool INV, int VPT, bool PRE = false, typename LA, typename LB, typename SA, typename SB>
__device__ __forceinline__ void stockham_pass2(int tid, int p, const float2* __restrict__ tw, LA la, LB lb, SA sa,
                                               SB sb, bool sync_between, const TwPow<R>* tpre = nullptr)
{
    constexpr int P = S / VPT;
    constexpr int NB = S / R;
    constexpr int PER = VPT / R;
    TwPow<R> tp[PER];
    if (p > 1) pass_twiddles<S, R, VPT, PRE>(tid, p, tw, tpre, tp);
    float2 ua[VPT], ub[VPT];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int i = tid + m * P;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ua[m * R + r] = la(i + r * NB);
            ub[m * R + r] = lb(i + r * NB);
        }
    }
    if (sync_between) __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int i = tid + m * P;
        const int k = i & (p - 1);
        if (p > 1) {
            float2 w[R];
            tw_pow_expand<R>(tp[m], w);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                ua[m * R + r] = cmul_tw<INV>(ua[m * R + r], w[r]);
                ub[m * R + r] = cmul_tw<INV>(ub[m * R + r], w[r]);
            }
        }
        dft<R, INV>(ua + m * R);
        dft<R, INV>(ub + m * R);
        const int base = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sa(base + r * p, ua[m * R + r]);
            sb(base + r * p, ub[m * R + r]);
        }
    }
}

// fft_passes for two transforms in lockstep (rows in lds_a and lds_b).
template <int LOGN, bool INV, int VPT, bool PRE = false, typename LA, typename LB, typename SA, typename SB>
__device__ __forceinline__ void fft_passes2(int tid, const float2* __restrict__ tw, float2* lds_a, float2* lds_b,
                                            LA first_a, LB first_b, bool first_lds, SA last_a, SB last_b,
                                            bool last_global, const TwPow<VPT>* tpre = nullptr)
{
    static_assert(VPT == 8 || VPT == 16, "FFT tile: 8 or 16 values per thread");
    constexpr int LV = VPT == 16 ? 4 : 3;
    constexpr int N = 1 << LOGN;
    constexpr int NPV = LOGN / LV;
    constexpr int RT = 1 << (LOGN % LV);
    auto la = [&](int i) { return lds_a[lds_idx(i)]; };
    auto lb = [&](int i) { return lds_b[lds_idx(i)]; };
    auto sa = [&](int i, float2 v) { lds_a[lds_idx(i)] = v; };
    auto sb = [&](int i, float2 v) { lds_b[lds_idx(i)] = v; };
    int p = 1;
#pragma unroll
    for (int q = 0; q < NPV; ++q) {
        const bool last = (q == NPV - 1) && RT == 1;
        if (q == 0) {
            if (last) stockham_pass2<N, VPT, INV, VPT>(tid, p, tw, first_a, first_b, last_a, last_b, first_lds);
            else stockham_pass2<N, VPT, INV, VPT>(tid, p, tw, first_a, first_b, sa, sb, first_lds);
        } else {
            const TwPow<VPT>* tq = PRE ? tpre + (q - 1) : nullptr;
            if (last) stockham_pass2<N, VPT, INV, VPT, PRE>(tid, p, tw, la, lb, last_a, last_b, true, tq);
            else stockham_pass2<N, VPT, INV, VPT, PRE>(tid, p, tw, la, lb, sa, sb, true, tq);
        }
        if (!(last && last_global)) __syncthreads();
        p *= VPT;
    }
    if constexpr (RT > 1) {
        if (NPV == 0) stockham_pass2<N, RT, INV, VPT>(tid, p, tw, first_a, first_b, last_a, last_b, first_lds);
        else stockham_pass2<N, RT, INV, VPT>(tid, p, tw, la, lb, last_a, last_b, true);
        if (!last_global) __syncthreads();
    }
}

// Threads of a workgroup for rows of 2^LOG points at VPT values per thread:
// at least 256 (several rows of short transforms share a workgroup).
constexpr int fft_block(int LOG, int VPT) { return (1 << LOG) / VPT > 256 ? (1 << LOG) / VPT : 256; }

// Two rows a, b per transform (power-of-two S): one forward FFT of the
// complex row z = x_a + i x_b gives both real rows' spectra,
//   X_a[k] = (Z[k] + conj Z[S-k]) / 2,   X_b[k] = -i (Z[k] - conj Z[S-k]) / 2,
// so a pair costs one forward and two inverse transforms instead of two and
// two.  Only k = 0..S/2 are formed (the weight is 0 above); the inverse's
// first pass reads zeros for k > S/2.  Y_a goes back into the row buffer,
// Y_b into a second half-length buffer: each index k <= S/2 is read (as Z[k]
// or Z[S-k]) only by the thread that overwrites it, so the split needs no
// barrier of its own.  A lone last row (odd row count) pairs with zeros.
// Rounding of the shared transform scales with the larger row of the pair
// (FP32, ~1e-7 of its magnitude).
template <int LOGS, int KIND, bool FILTER, int VPT>
__global__ __launch_bounds__(fft_block(LOGS, VPT)) void hilbert_pair_kernel(HilbertDev in, int rows,
                                                                            const float2* __restrict__ tw,
                                                                            const float2* __restrict__ F,
                                                                            void* __restrict__ out, int ofmt,
                                                                            float oscale)
{
    constexpr int S = 1 << LOGS;
    constexpr int H = S / 2;
    constexpr int P = S / VPT;
    constexpr int RPW = P >= 256 ? 1 : 256 / P;     // row pairs per workgroup
    constexpr int PADDED = S + S / 16;
    // 4096-point rows run the two inverse transforms in lockstep (a full second
    // row buffer; 0.356 -> 0.336 ms for 16384 rows), shorter rows one after
    // the other (half-length second buffer; lockstep measured no gain there)
    constexpr bool DUAL = LOGS >= 12;
    constexpr int PADDED2 = DUAL ? PADDED : H + H / 16 + 1;  // lds_idx(H) + 1
    __shared__ float2 buf[RPW * PADDED];
    __shared__ float2 buf2[RPW * PADDED2];

    const int tid = threadIdx.x % P;
    const int rsub = threadIdx.x / P;
    const int pair = blockIdx.x * RPW + rsub;
    const int ra = 2 * pair, rb = 2 * pair + 1;
    const bool act_a = ra < rows, act_b = rb < rows;
    const int rra = act_a ? ra : 0, rrb = act_b ? rb : 0;
    float2* lds = buf + rsub * PADDED;
    float2* lds2 = buf2 + rsub * PADDED2;

    const void* base_a = row_base<KIND>(in, rra, S);
    const void* base_b = row_base<KIND>(in, rrb, S);
    const unsigned istride = (unsigned)in.stride;
    auto load_pair = [&](int idx) {
        return make_float2(act_a ? load_at<KIND>(base_a, istride, (unsigned)idx) : 0.0f,
                           act_b ? load_at<KIND>(base_b, istride, (unsigned)idx) : 0.0f);
    };
    auto store_lds = [&](int idx, float2 v) { lds[lds_idx(idx)] = v; };
    fft_passes<LOGS, false, VPT>(tid, tw, lds, load_pair, false, store_lds, false);

    // split into the two one-sided spectra, weight (w_k / 2) and filter
    const float half_w = 0.75f / (float)S;
    for (int k = tid; k <= H; k += P) {
        const float2 zk = lds[lds_idx(k)];
        const float2 zn = lds[lds_idx((S - k) & (S - 1))];
        const float w = k == 0 ? 0.5f * half_w : half_w;
        float2 ya = make_float2((zk.x + zn.x) * w, (zk.y - zn.y) * w);
        float2 yb = make_float2((zk.y + zn.y) * w, (zn.x - zk.x) * w);
        if constexpr (FILTER) {
            const float2 f = F[k];
            ya = cmul(ya, f);
            yb = cmul(yb, f);
        }
        lds[lds_idx(k)] = ya;
        lds2[lds_idx(k)] = yb;
    }
    __syncthreads();

    // output rows addressed from (uniform) row starts with 32-bit offsets
    const size_t esz = ofmt == IQ_F32 ? sizeof(float2) : sizeof(uint32_t);
    char* const out_a = static_cast<char*>(out) + (size_t)rra * S * esz;
    char* const out_b = static_cast<char*>(out) + (size_t)rrb * S * esz;
    auto store_row = [&](bool act, char* orow, int idx, float2 v) {
        if (!act) return;
        if (ofmt == IQ_F32) {
            reinterpret_cast<float2*>(orow)[(unsigned)idx] = v;
        } else if (ofmt == IQ_F16) {
            reinterpret_cast<__half2*>(orow)[(unsigned)idx] = __floats2half2_rn(v.x * oscale, v.y * oscale);
        } else {
            reinterpret_cast<uint32_t*>(orow)[(unsigned)idx] = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
        }
    };
    auto load_ya = [&](int idx) { return idx > H ? make_float2(0.0f, 0.0f) : lds[lds_idx(idx)]; };
    auto load_yb = [&](int idx) { return idx > H ? make_float2(0.0f, 0.0f) : lds2[lds_idx(idx)]; };
    auto store_a = [&](int idx, float2 v) { store_row(act_a, out_a, idx, v); };
    auto store_b = [&](int idx, float2 v) { store_row(act_b, out_b, idx, v); };
    if constexpr (DUAL) {
        // the forward's twiddles reloaded through an opaque copy of the table
        // pointer rather than kept live
        const float2* tw_i = tw;
        __asm__ volatile("" : "+s"(tw_i));
        fft_passes2<LOGS, true, VPT>(tid, tw_i, lds, lds2, load_ya, load_yb, true, store_a, store_b, true);
    } else {
        // Each transform reads the twiddles through its own opaque copy of the
        // table pointer, so the compiler reloads them instead of keeping one
        // transform's twiddle registers live through the next (196 VGPRs, two
        // workgroups per CU, without; 162 and three with).
        const float2* tw_a = tw;
        __asm__ volatile("" : "+s"(tw_a));
        fft_passes<LOGS, true, VPT>(tid, tw_a, lds, load_ya, true, store_a, true);
        __syncthreads();  // the last pass of a read the row buffer that b's first pass writes
        const float2* tw_b = tw;
        __asm__ volatile("" : "+s"(tw_b));
        fft_passes<LOGS, true, VPT>(tid, tw_b, lds, load_yb, false, store_b, true);
    }
}

// hilbert_pair_kernel for 4096-point rows and 16 values per thread as a fixed
// grid of workgroups, each walking row pairs blockIdx.x, + gridDim.x, ...
// Each iteration requests the next pair's input rows once the first forward
// pass has consumed the current pair's (into the same registers); the next
// iteration's first pass consumes them.  Vector-memory loads retire in order,
// so a load issued after the prefetch would make the wave wait for the
// prefetch as well: the twiddle powers of the two twiddled passes (p = 16,
// 256; the inverse uses their conjugates) depend on the thread alone and are
// loaded once and held in registers, and the filter values of the split are
// requested (L2 hits) just before the prefetch.  Arithmetic, LDS
// layout and rounding are those of hilbert_pair_kernel<12, KIND, FILTER, 16>.
template <int KIND>
using HilRaw = std::conditional_t<KIND == 1, int, float>;

template <int KIND>
__device__ __forceinline__ HilRaw<KIND> load_raw(const void* base, unsigned stride, unsigned idx)
{
    if constexpr (KIND == 0) return ((const float*)base)[idx * stride];
    else if constexpr (KIND == 1) return (int)((const int16_t*)base)[idx];
    else return ((const float*)base)[idx];
}

constexpr int HIL_STREAM_LOGS = 12;
// Word 3 of a raw (untyped, stride 0) buffer resource on gfx9-family GPUs:
// 32-bit data format, so that the range check compares byte offsets with
// the resource's size and drops stores at or beyond it.
constexpr int HIL_RSRC_WORD3 = 0x00020000;

template <int KIND, bool FILTER, int OFMT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void hilbert_pair_stream_kernel(HilbertDev in, int rows, const float2* __restrict__ tw,
                                const float2* __restrict__ F, void* __restrict__ out, float oscale)
{
    constexpr int LOGS = HIL_STREAM_LOGS;
    constexpr int S = 1 << LOGS;
    constexpr int H = S / 2;
    constexpr int VPT = 16;
    constexpr int P = S / VPT;
    constexpr int NB = S / VPT;
    constexpr int PADDED = S + S / 16;
    static_assert(P == 256, "one row pair per 256-thread workgroup");
    __shared__ float2 buf[PADDED];
    __shared__ float2 buf2[PADDED];
    float2* const lds = buf;
    float2* const lds2 = buf2;

    const int tid0 = threadIdx.x;
    const int npairs = (rows + 1) / 2;
    TwPow<VPT> tpre[LOGS / 4 - 1];
#pragma unroll
    for (int q = 1, p = VPT; q < LOGS / 4; ++q, p *= VPT)
        tpre[q - 1] = tw_pow_load<VPT>(tw, (tid0 & (p - 1)) * (S / (p * VPT)), S - 1);
    // (the scheduler keeps the order of these loads and the rows' below)
    __builtin_amdgcn_sched_barrier(0);

    const unsigned istride = (unsigned)in.stride;
    using Raw = HilRaw<KIND>;
    // Loads of a pair's two rows, unconditional so that the wait for them
    // has one count on every path: past the last pair the last pair is
    // fetched again (and not used); a lone last row's partner is read from
    // the row itself and zeroed where it is consumed.
    auto fetch = [&](int pair, Raw* xa, Raw* xb, int tid) {
        const int pc = pair < npairs ? pair : npairs - 1;
        const int ra = 2 * pc, rb = ra + 1 < rows ? ra + 1 : ra;
        const void* base_a = row_base<KIND>(in, ra, S);
        const void* base_b = row_base<KIND>(in, rb, S);
#pragma unroll
        for (int r = 0; r < VPT; ++r) xa[r] = load_raw<KIND>(base_a, istride, (unsigned)(tid + r * NB));
#pragma unroll
        for (int r = 0; r < VPT; ++r) xb[r] = load_raw<KIND>(base_b, istride, (unsigned)(tid + r * NB));
    };
    // The walk is rotated: an iteration runs its pair's transforms from the
    // second forward pass on, then the next pair's first pass, which consumes
    // the input rows (ca, cb) requested one iteration earlier, and requests
    // the following pair's rows into the same registers, after the filter
    // values of the next split (older loads: waiting for them leaves the
    // rows in flight).  The wait for the rows then follows this pair's
    // output stores on every path, so they stay in flight too.
    Raw ca[VPT], cb[VPT];
    float2 fk[H / P];
    float2 fH = make_float2(1.0f, 0.0f);
    auto store_lds = [&](int idx, float2 v) { lds[lds_idx(idx)] = v; };
    auto load_lds = [&](int idx) { return lds[lds_idx(idx)]; };
    auto first_and_prefetch = [&](int pair, int tid) {
        const bool act_b = 2 * pair + 1 < rows;
        auto first = [&](int, int slot) { return make_float2((float)ca[slot], act_b ? (float)cb[slot] : 0.0f); };
        stockham_pass<S, VPT, false, VPT>(tid, 1, tw, first, store_lds, false);
        if constexpr (FILTER) {
#pragma unroll
            for (int j = 0; j < H / P; ++j) fk[j] = F[tid + j * P];
            fH = F[H];
        }
        __builtin_amdgcn_sched_barrier(0);
        fetch(pair + gridDim.x, ca, cb, tid);
        __syncthreads();
    };
    fetch(blockIdx.x, ca, cb, tid0);
    first_and_prefetch(blockIdx.x, tid0);

    const float half_w = 0.75f / (float)S;
    constexpr unsigned ESZ = OFMT == IQ_F32 ? sizeof(float2) : sizeof(uint32_t);
    // the grid holds at most one workgroup per pair: every workgroup has a
    // first pair (a loop without an entry test, past which the loads above
    // would be sunk)
    int pair = blockIdx.x;
    do {
        // the pass addresses derived from the thread index are formed anew in
        // each iteration rather than hoisted out of the walk and held (their
        // registers spilled)
        int tid = tid0;
        __asm__ volatile("" : "+v"(tid));
        tid &= P - 1;

#pragma unroll
        for (int q = 1, p = VPT; q < LOGS / 4; ++q, p *= VPT) {
            stockham_pass<S, VPT, false, VPT, true>(tid, p, tw, load_lds, store_lds, true, tpre + (q - 1));
            __syncthreads();
        }

        auto split = [&](int k, float2 f) {
            const float2 zk = lds[lds_idx(k)];
            const float2 zn = lds[lds_idx((S - k) & (S - 1))];
            const float w = k == 0 ? 0.5f * half_w : half_w;
            float2 ya = make_float2((zk.x + zn.x) * w, (zk.y - zn.y) * w);
            float2 yb = make_float2((zk.y + zn.y) * w, (zn.x - zk.x) * w);
            if constexpr (FILTER) {
                ya = cmul(ya, f);
                yb = cmul(yb, f);
            }
            lds[lds_idx(k)] = ya;
            lds2[lds_idx(k)] = yb;
        };
#pragma unroll
        for (int j = 0; j < H / P; ++j) split(tid + j * P, fk[j]);
        if (tid == 0) split(H, fH);
        __syncthreads();

        // Rows a and b are adjacent in the output: one buffer resource spans
        // them, sized to row a alone when b is the lone last row's zero
        // partner, so that b's stores are dropped by the range check rather
        // than skipped by a branch.
        const int ra = 2 * pair;
        const bool act_b = ra + 1 < rows;
        char* const out_a = static_cast<char*>(out) + (size_t)ra * S * ESZ;
        const __amdgpu_buffer_rsrc_t orsrc =
            __builtin_amdgcn_make_buffer_rsrc(out_a, 0, (act_b ? 2 : 1) * S * ESZ, HIL_RSRC_WORD3);
        auto store_out = [&](unsigned byte_off, float2 v) {
            if constexpr (OFMT == IQ_F32) {
                __builtin_amdgcn_raw_buffer_store_b64(
                    __builtin_bit_cast(HIP_vector_base<unsigned, 2>::Native_vec_, v), orsrc, byte_off, 0, 0);
            } else if constexpr (OFMT == IQ_F16) {
                const __half2 h = __floats2half2_rn(v.x * oscale, v.y * oscale);
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, h), orsrc, byte_off, 0, 0);
            } else {
                __builtin_amdgcn_raw_buffer_store_b32(bf16_bits(v.x) | (bf16_bits(v.y) << 16), orsrc, byte_off, 0,
                                                      0);
            }
        };
        auto load_ya = [&](int idx) { return idx > H ? make_float2(0.0f, 0.0f) : lds[lds_idx(idx)]; };
        auto load_yb = [&](int idx) { return idx > H ? make_float2(0.0f, 0.0f) : lds2[lds_idx(idx)]; };
        auto store_a = [&](int idx, float2 v) { store_out((unsigned)idx * ESZ, v); };
        auto store_b = [&](int idx, float2 v) { store_out((unsigned)(S + idx) * ESZ, v); };
        fft_passes2<LOGS, true, VPT, true>(tid, tw, lds, lds2, load_ya, load_yb, true, store_a, store_b, true, tpre);
        __syncthreads();  // the next pair's first pass writes the row buffer this pair's last pass read
        pair += gridDim.x;
        if (pair < npairs) first_and_prefetch(pair, tid);
    } while (pair < npairs);
}

// Workgroups of hilbert_pair_stream_kernel resident on the device at once
// (per device, from the occupancy of the instantiation launched).
template <int KIND, bool FILTER, int OFMT>
static bool hil_stream_grid(int pairs, int& grid)
{
    static int wgs[64] = {};
    int dev = 0;
    CUT_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) CUT_FAIL("Hilbert: device ordinal " << dev << " out of range");
    if (wgs[dev] == 0) {
        int cus = 0, per_cu = 0;
        CUT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        CUT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&hilbert_pair_stream_kernel<KIND, FILTER, OFMT>), 256, 0));
        wgs[dev] = std::max(1, cus * std::max(1, per_cu));
    }
    grid = std::min(pairs, wgs[dev]);
    return true;
}

template <int KIND, bool FILTER, int OFMT>
static bool launch_hilbert_stream_f(const HilbertDev& d, int rows, const float2* tw, const float2* F, void* out,
                                    float oscale, hipStream_t st)
{
    int grid = 0;
    if (!hil_stream_grid<KIND, FILTER, OFMT>((rows + 1) / 2, grid)) return false;
    hilbert_pair_stream_kernel<KIND, FILTER, OFMT><<<grid, 256, 0, st>>>(d, rows, tw, F, out, oscale);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

template <int KIND, bool FILTER>
static bool launch_hilbert_stream(const HilbertDev& d, int rows, const float2* tw, const float2* F, void* out,
                                  int ofmt, float oscale, hipStream_t st)
{
    if (ofmt == IQ_F32) return launch_hilbert_stream_f<KIND, FILTER, IQ_F32>(d, rows, tw, F, out, oscale, st);
    if (ofmt == IQ_F16) return launch_hilbert_stream_f<KIND, FILTER, IQ_F16>(d, rows, tw, F, out, oscale, st);
    return launch_hilbert_stream_f<KIND, FILTER, IQ_BF16>(d, rows, tw, F, out, oscale, st);
}

// Hilbert transform of rows of any length S <= 4096 that is not a power of
// two.  The reference's operator, y = IDFT_S(G . DFT_S(x)) with G[k] = w_k F[k]
// for k <= S/2 and 0 above (unnormalised inverse), is a cyclic convolution of
// length S with g = IDFT_S(G): y[n] = sum_m x[m] g[(n - m) mod S].  With the row
// zero-padded to M = 2^conv_log2(S) >= 2S - 1 and h the kernel laid out over
// the offsets -(S-1)..S-1 (h[j mod M] = g[j mod S]), the cyclic convolution of
// length M equals it for n < S (no offset wraps around):
//   y[n] = IFFT_M(FFT_M(x) . H)[n],  H = FFT_M(h) / M.
// Rows are taken in pairs a, b.  With h = h_r + i h_i (both real) and
// H_r = FFT_M(h_r) / M, H_i = FFT_M(h_i) / M, one forward transform of
// x = x_a + i x_b and two inverse ones give both rows:
//   P = IFFT_M(X . H_r) = conv(x_a, h_r) + i conv(x_b, h_r),
//   Q = IFFT_M(X . H_i) = conv(x_a, h_i) + i conv(x_b, h_i),
//   y_a = Re P + i Re Q,  y_b = Im P + i Im Q
// (three transforms per pair instead of four).  HP[k] = (H_r[k], H_i[k]) is
// built once per length and filter in FP64 (conv_tables).  The first pass of
// P's inverse reads X[k] from LDS and keeps X[k] H_i[k] in registers: the
// first pass of Q's inverse loads the same indices in the same order, as the
// last passes of the two inverses store the same indices, so P's outputs stay
// in registers until Q's are formed and each output element is written once.
// A lone last row (odd row count) pairs with zeros; rounding of the shared
// transforms scales with the larger row of the pair (FP32, ~1e-7).
template <int LOGM, int KIND, int VPT>
__global__ __launch_bounds__(fft_block(LOGM, VPT)) void hilbert_conv_kernel(HilbertDev in, int rows, int S,
                                                                           const float2* __restrict__ tw,
                                                                           const float4* __restrict__ HP,
                                                                           void* __restrict__ out, int ofmt,
                                                                           float oscale)
{
    constexpr int M = 1 << LOGM;
    constexpr int P = M / VPT;
    constexpr int RPW = P >= 256 ? 1 : 256 / P;  // row pairs per workgroup
    constexpr int PADDED = M + M / 16;
    __shared__ float2 buf[RPW * PADDED];

    const int tid = threadIdx.x % P;
    const int rsub = threadIdx.x / P;
    const int pair = blockIdx.x * RPW + rsub;
    const int ra = 2 * pair, rb = 2 * pair + 1;
    const bool act_a = ra < rows, act_b = rb < rows;
    const int rra = act_a ? ra : 0, rrb = act_b ? rb : 0;
    float2* lds = buf + rsub * PADDED;
    const void* base_a = row_base<KIND>(in, rra, S);
    const void* base_b = row_base<KIND>(in, rrb, S);
    const unsigned istride = (unsigned)in.stride;

    auto load_x = [&](int n) {
        const bool in_row = n < S;
        return make_float2(act_a && in_row ? load_at<KIND>(base_a, istride, (unsigned)n) : 0.0f,
                           act_b && in_row ? load_at<KIND>(base_b, istride, (unsigned)n) : 0.0f);
    };
    auto store_lds = [&](int i, float2 v) { lds[lds_idx(i)] = v; };
    fft_passes<LOGM, false, VPT>(tid, tw, lds, load_x, false, store_lds, false);

    // P: X H_r from LDS, X H_i kept for Q (each pass's loads are fully unrolled,
    // so the counters below index registers)
    float2 xq[VPT];
    int nq = 0;
    auto load_p = [&](int k) {
        const float2 x = lds[lds_idx(k)];
        const float4 h = HP[k];
        xq[nq++] = cmul(x, make_float2(h.z, h.w));
        return cmul(x, make_float2(h.x, h.y));
    };
    float2 pv[VPT];
    int np = 0;
    auto store_p = [&](int n, float2 v) { pv[np++] = v; };
    const float2* tw_p = tw;
    __asm__ volatile("" : "+s"(tw_p));  // twiddles reloaded per transform (see hilbert_pair_kernel)
    fft_passes<LOGM, true, VPT>(tid, tw_p, lds, load_p, true, store_p, true);

    // Q from the kept products; its last pass writes both rows
    int lq = 0;
    auto load_q = [&](int) { return xq[lq++]; };
    const size_t esz = ofmt == IQ_F32 ? sizeof(float2) : sizeof(uint32_t);
    char* const out_a = static_cast<char*>(out) + (size_t)rra * S * esz;
    char* const out_b = static_cast<char*>(out) + (size_t)rrb * S * esz;
    auto put = [&](char* orow, int n, float2 v) {
        if (ofmt == IQ_F32) {
            reinterpret_cast<float2*>(orow)[(unsigned)n] = v;
        } else if (ofmt == IQ_F16) {
            reinterpret_cast<__half2*>(orow)[(unsigned)n] = __floats2half2_rn(v.x * oscale, v.y * oscale);
        } else {
            reinterpret_cast<uint32_t*>(orow)[(unsigned)n] = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
        }
    };
    int sq = 0;
    auto store_q = [&](int n, float2 q) {
        const float2 p = pv[sq++];
        if (n >= S) return;
        if (act_a) put(out_a, n, make_float2(p.x, q.x));
        if (act_b) put(out_b, n, make_float2(p.y, q.y));
    };
    const float2* tw_q = tw;
    __asm__ volatile("" : "+s"(tw_q));
    // first_lds: the barrier inside Q's first pass orders its LDS stores after
    // the loads of P's last pass
    fft_passes<LOGM, true, VPT>(tid, tw_q, lds, load_q, true, store_q, true);
}

// Kernel spectrum of the convolution form, in FP64 (exact integer angle
// reduction):  G[q] = w_q F[q] for q <= S/2 (F = DFT_S of the taps, 1 without
// a filter; w_0 = 0.75/S, w_q = 1.5/S),  g[n] = sum_q G[q] exp(+2 pi i q n / S),
//   H[k] = (1/M) sum_{j=-(S-1)}^{S-1} g[j mod S] exp(-2 pi i j k / M).
__global__ void conv_G_kernel(const float* __restrict__ h, int L, int S, double2* __restrict__ G)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > S / 2) return;
    double re = 1.0, im = 0.0;
    if (L > 0) {
        re = 0.0;
        for (int n = 0; n < L; ++n) {
            const long long e = ((long long)q * n) % S;
            double sn, cs;
            sincospi(2.0 * (double)e / (double)S, &sn, &cs);
            re += (double)h[n] * cs;
            im -= (double)h[n] * sn;
        }
    }
    const double w = (q == 0 ? 0.75 : 1.5) / (double)S;
    G[q] = make_double2(re * w, im * w);
}

__global__ void conv_g_kernel(const double2* __restrict__ G, int S, double2* __restrict__ g)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= S) return;
    double re = 0.0, im = 0.0;
    for (int q = 0; q <= S / 2; ++q) {
        const long long e = ((long long)q * n) % S;
        double sn, cs;
        sincospi(2.0 * (double)e / (double)S, &sn, &cs);
        re += G[q].x * cs - G[q].y * sn;
        im += G[q].x * sn + G[q].y * cs;
    }
    g[n] = make_double2(re, im);
}

__global__ void conv_H_kernel(const double2* __restrict__ g, int S, int M, float2* __restrict__ H)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    double re = 0.0, im = 0.0;
    for (int j = -(S - 1); j < S; ++j) {
        const double2 v = g[j < 0 ? j + S : j];
        const unsigned long long e = ((unsigned long long)(j < 0 ? j + M : j) * (unsigned long long)k) & (M - 1);
        double sn, cs;
        sincospi(2.0 * (double)e / (double)M, &sn, &cs);
        re += v.x * cs + v.y * sn;
        im += v.y * cs - v.x * sn;
    }
    H[k] = make_float2((float)(re / M), (float)(im / M));
}

// The same spectrum split by the real and imaginary parts of the kernel, for
// the row-pair kernel: HP[k] = (H_r[k], H_i[k]), H_r = FFT_M(Re h) / M,
// H_i = FFT_M(Im h) / M (so H = H_r + i H_i).
__global__ void conv_HP_kernel(const double2* __restrict__ g, int S, int M, float4* __restrict__ HP)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    double rr = 0.0, ri = 0.0, ir = 0.0, ii = 0.0;
    for (int j = -(S - 1); j < S; ++j) {
        const double2 v = g[j < 0 ? j + S : j];
        const unsigned long long e = ((unsigned long long)(j < 0 ? j + M : j) * (unsigned long long)k) & (M - 1);
        double sn, cs;
        sincospi(2.0 * (double)e / (double)M, &sn, &cs);
        rr += v.x * cs;
        ri -= v.x * sn;
        ir += v.y * cs;
        ii -= v.y * sn;
    }
    HP[k] = make_float4((float)(rr / M), (float)(ri / M), (float)(ir / M), (float)(ii / M));
}

int conv_log2(int S)
{
    int lg = 4;
    while ((1 << lg) < 2 * S - 1) ++lg;
    return lg;
}

size_t conv_tables_tmp_bytes(int S) { return (size_t)(S / 2 + 1 + S) * sizeof(double2); }

bool launch_conv_tables(const float* d_h, int L, int S, void* d_tmp, void* d_H, bool pair, hipStream_t st)
{
    const int M = 1 << conv_log2(S);
    double2* G = static_cast<double2*>(d_tmp);
    double2* g = G + (S / 2 + 1);
    conv_G_kernel<<<(S / 2 + 1 + 127) / 128, 128, 0, st>>>(d_h, L, S, G);
    conv_g_kernel<<<(S + 127) / 128, 128, 0, st>>>(G, S, g);
    if (pair) conv_HP_kernel<<<(M + 127) / 128, 128, 0, st>>>(g, S, M, static_cast<float4*>(d_H));
    else conv_H_kernel<<<(M + 127) / 128, 128, 0, st>>>(g, S, M, static_cast<float2*>(d_H));
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

// Twiddle table tw[m] = exp(-2 pi i m / S), generated in FP64.
__global__ void twiddle_kernel(float2* tw, int S)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= S) return;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)S, &sn, &cs);
    tw[m] = make_float2((float)cs, (float)-sn);
}

// Match-filter spectrum F[k] = sum_{n<L} h[n] exp(-2 pi i k n / S) in FP64
// (hilbert_handler.cu:55-78 takes the forward DFT of the zero-padded taps).
__global__ void filter_spectrum_kernel(const float* h, int L, int S, float2* F)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    double re = 0.0, im = 0.0;
    for (int n = 0; n < L; ++n) {
        const long long e = ((long long)k * n) % S;
        double sn, cs;
        sincospi(2.0 * (double)e / (double)S, &sn, &cs);
        re += (double)h[n] * cs;
        im -= (double)h[n] * sn;
    }
    F[k] = make_float2((float)re, (float)im);
}

bool launch_twiddles(float2* d_tw, int S, hipStream_t st)
{
    twiddle_kernel<<<(S + 255) / 256, 256, 0, st>>>(d_tw, S);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool launch_filter_spectrum(const float* d_h, int L, int S, float2* d_F, hipStream_t st)
{
    filter_spectrum_kernel<<<(S + 255) / 256, 256, 0, st>>>(d_h, L, S, d_F);
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

bool hilbert_pow2(int S) { return S >= 16 && (S & (S - 1)) == 0 && S <= 4096; }
bool hilbert_long(int S) { return S > 4096; }
int hilbert_long_n(int S) { return ((S & (S - 1)) == 0) ? S : 1 << conv_log2(S); }
// Lengths that take the convolution form: every supported length except the
// powers of two 16..4096 (pair kernel) and the powers of two beyond 4096
// (long path with the spectral weight); so 1, 2, 4 and 8 take it too.
bool hilbert_conv(int S)
{
    return hilbert_supported(S) && !hilbert_pow2(S) && !(hilbert_long(S) && (S & (S - 1)) == 0);
}
int hilbert_fft_len(int S) { return hilbert_conv(S) ? 1 << conv_log2(S) : S; }
bool hilbert_supported(int S) { return S >= 1 && (S <= 4096 || hilbert_long_n(S) <= HILBERT_LONG_MAX_N); }

template <int LOGM, int VPT>
static bool launch_conv_m(const HilbertInput& in, int rows, int S, const float2* tw, const float4* H, void* out,
                          int ofmt, float oscale, hipStream_t st)
{
    constexpr int M = 1 << LOGM;
    constexpr int P = M / VPT;
    constexpr int RPW = P >= 256 ? 1 : 256 / P;  // row pairs per workgroup
    HilbertDev d{in.ptr, in.stride, in.map, in.R0, in.C};
    const int pairs = (rows + 1) / 2;
    dim3 grid((pairs + RPW - 1) / RPW);
    dim3 block(P * RPW);
#define CUT_CONV(K) hilbert_conv_kernel<LOGM, K, VPT><<<grid, block, 0, st>>>(d, rows, S, tw, H, out, ofmt, oscale)
    if (in.kind == 0) CUT_CONV(0); else if (in.kind == 1) CUT_CONV(1); else CUT_CONV(2);
#undef CUT_CONV
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

template <int VPT>
static bool launch_conv_vpt(const HilbertInput& in, int rows, int S, const float2* tw, const float4* H, void* out,
                            int ofmt, float oscale, hipStream_t st)
{
    switch (conv_log2(S)) {
        case 4: return launch_conv_m<4, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 5: return launch_conv_m<5, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 6: return launch_conv_m<6, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 7: return launch_conv_m<7, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 8: return launch_conv_m<8, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 9: return launch_conv_m<9, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 10: return launch_conv_m<10, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 11: return launch_conv_m<11, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 12: return launch_conv_m<12, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        case 13: return launch_conv_m<13, VPT>(in, rows, S, tw, H, out, ofmt, oscale, st);
        default: CUT_FAIL("Hilbert: sample count " << S << " not supported by the LDS transforms (1..4096)");
    }
}

bool launch_hilbert_conv(const HilbertInput& in, int rows, int S, const float2* d_tw, const float4* d_H, void* out,
                         int ofmt, float oscale, int vpt, hipStream_t st)
{
    if (rows <= 0) return true;
    if (!hilbert_conv(S) || hilbert_long(S)) CUT_FAIL("Hilbert: " << S << " samples do not take the convolution kernel");
    if (!d_tw || !d_H) CUT_FAIL("Hilbert: convolution tables of " << S << " samples missing");
    if (ofmt != IQ_F32 && ofmt != IQ_F16 && ofmt != IQ_BF16) CUT_FAIL("Hilbert: unknown output format " << ofmt);
    if (vpt == 8) return launch_conv_vpt<8>(in, rows, S, d_tw, d_H, out, ofmt, oscale, st);
    if (vpt == 16) return launch_conv_vpt<16>(in, rows, S, d_tw, d_H, out, ofmt, oscale, st);
    CUT_FAIL("Hilbert: FFT tile of " << vpt << " values per thread (8 or 16)");
}

template <int LOGS, int VPT>
static bool launch_hilbert_s(const HilbertInput& in, int rows, const float2* tw, const float2* F, void* out, int ofmt,
                             float oscale, hipStream_t st)
{
    constexpr int S = 1 << LOGS;
    constexpr int P = S / VPT;
    constexpr int RPW = P >= 256 ? 1 : 256 / P;     // row pairs per workgroup
    HilbertDev d{in.ptr, in.stride, in.map, in.R0, in.C};
    if constexpr (LOGS == HIL_STREAM_LOGS && VPT == 16) {
        if (F) {
            if (in.kind == 0) return launch_hilbert_stream<0, true>(d, rows, tw, F, out, ofmt, oscale, st);
            if (in.kind == 1) return launch_hilbert_stream<1, true>(d, rows, tw, F, out, ofmt, oscale, st);
            return launch_hilbert_stream<2, true>(d, rows, tw, F, out, ofmt, oscale, st);
        }
        if (in.kind == 0) return launch_hilbert_stream<0, false>(d, rows, tw, F, out, ofmt, oscale, st);
        if (in.kind == 1) return launch_hilbert_stream<1, false>(d, rows, tw, F, out, ofmt, oscale, st);
        return launch_hilbert_stream<2, false>(d, rows, tw, F, out, ofmt, oscale, st);
    }
    const int pairs = (rows + 1) / 2;
    dim3 grid((pairs + RPW - 1) / RPW);
    dim3 block(P * RPW);
#define CUT_HIL(K, FL) \
    hilbert_pair_kernel<LOGS, K, FL, VPT><<<grid, block, 0, st>>>(d, rows, tw, F, out, ofmt, oscale)
    if (F) {
        if (in.kind == 0) CUT_HIL(0, true); else if (in.kind == 1) CUT_HIL(1, true); else CUT_HIL(2, true);
    } else {
        if (in.kind == 0) CUT_HIL(0, false); else if (in.kind == 1) CUT_HIL(1, false); else CUT_HIL(2, false);
    }
#undef CUT_HIL
    CUT_HIP_CHECK(hipGetLastError());
    return true;
}

template <int VPT>
static bool launch_hilbert_vpt(const HilbertInput& in, int rows, int S, const float2* d_tw, const float2* d_F,
                               void* out, int ofmt, float oscale, hipStream_t st)
{
    switch (S) {
        case 16: return launch_hilbert_s<4, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 32: return launch_hilbert_s<5, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 64: return launch_hilbert_s<6, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 128: return launch_hilbert_s<7, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 256: return launch_hilbert_s<8, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 512: return launch_hilbert_s<9, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 1024: return launch_hilbert_s<10, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 2048: return launch_hilbert_s<11, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        case 4096: return launch_hilbert_s<12, VPT>(in, rows, d_tw, d_F, out, ofmt, oscale, st);
        default: CUT_FAIL("Hilbert: sample count " << S << " not supported (powers of two 16..4096)");
    }
}

bool launch_hilbert(const HilbertInput& in, int rows, int S, const float2* d_tw, const float2* d_F, void* out,
                    int ofmt, float oscale, int vpt, hipStream_t st)
{
    if (rows <= 0) return true;
    if (!hilbert_pow2(S)) CUT_FAIL("Hilbert: " << S << " samples do not take the row-pair kernel");
    if (!d_tw) CUT_FAIL("Hilbert: twiddle table of " << S << " samples missing");
    if (ofmt != IQ_F32 && ofmt != IQ_F16 && ofmt != IQ_BF16) CUT_FAIL("Hilbert: unknown output format " << ofmt);
    if (vpt == 8) return launch_hilbert_vpt<8>(in, rows, S, d_tw, d_F, out, ofmt, oscale, st);
    if (vpt == 16) return launch_hilbert_vpt<16>(in, rows, S, d_tw, d_F, out, ofmt, oscale, st);
    CUT_FAIL("Hilbert: FFT tile of " << vpt << " values per thread (8 or 16)");
}


// ---------------------------------------------------------------------------
// Rows longer than 4096 samples.  The N-point transforms (N = S for powers of
// two, else the convolution length M = 2^conv_log2(S) of hilbert_conv_kernel)
// are split as N = N1 x 4096, N1 = N / 4096 <= 16, in Bailey's four-step form
// with the data in HBM between the steps: with n = n2 + 4096 n1 and
// k = k1 + N1 k2,
//   X[k1 + N1 k2] = sum_n2 W_4096^(n2 k2) [W_N^(n2 k1) sum_n1 x[n2 + 4096 n1] W_N1^(n1 k1)].
// long_fwd_pass_kernel does the inner length-N1 DFTs and the twiddles
// (thread = one n2 of one row, loads coalesced across n2) into B[k1][n2];
// long_row_kernel transforms each 4096-point segment B[k1][.] in LDS,
// multiplies the spectrum (index k = k1 + N1 k2) by the spectral weight and
// filter (powers of two) or by the convolution kernel's spectrum H (other
// lengths, rows zero-padded to N), and transforms back; long_inv_pass_kernel
// applies the conjugate twiddles and the inverse length-N1 DFTs, which land
// on natural sample order, and writes the row's first S samples.  Every step
// reads and writes the same addresses per thread or workgroup, so the steps
// run in place in one row-major buffer.  Three launches per chunk of rows.
// ---------------------------------------------------------------------------
constexpr int LN2_LOG = 12;
constexpr int LN2 = 1 << LN2_LOG;

enum { LROW_HILBERT = 0, LROW_MUL = 1 };

__device__ __forceinline__ void store_fmt(void* out, size_t i, float2 v, int ofmt, float oscale)
{
    if (ofmt == IQ_F32) static_cast<float2*>(out)[i] = v;
    else if (ofmt == IQ_F16) static_cast<__half2*>(out)[i] = __floats2half2_rn(v.x * oscale, v.y * oscale);
    else static_cast<uint32_t*>(out)[i] = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
}

// grid (LN2 / 256, rows of the chunk); rows row0 + blockIdx.y; samples n >= S
// of the zero-padded row read as 0.
template <int N1, int KIND>
__global__ __launch_bounds__(256) void long_fwd_pass_kernel(HilbertDev in, int row0, int S,
                                                            const float2* __restrict__ twN, float2* buf)
{
    constexpr int N = N1 * LN2;
    const int n2 = blockIdx.x * 256 + threadIdx.x;
    float2* brow = buf + (size_t)blockIdx.y * N;
    float2 u[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int n = n2 + LN2 * n1;
        u[n1] = make_float2(n < S ? load_in<KIND>(in, row0 + (int)blockIdx.y, S, n) : 0.0f, 0.0f);
    }
    dft<N1, false>(u);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        float2 v = u[k1];
        if (k1 > 0) v = cmul(v, twN[n2 * k1]);  // n2 k1 < N
        brow[k1 * LN2 + n2] = v;
    }
}

// grid (N1, rows of the chunk), 256 threads: segment k1 = blockIdx.x of a row.
template <int MODE, bool FILTER>
__global__ __launch_bounds__(256) void long_row_kernel(float2* buf, int N1, int S, const float2* __restrict__ tw,
                                                       const float2* __restrict__ G, const float2* __restrict__ F)
{
    __shared__ float2 lds[LN2 + LN2 / 16];
    const int N = N1 * LN2;
    const int k1 = blockIdx.x;
    float2* seg = buf + (size_t)blockIdx.y * N + (size_t)k1 * LN2;
    const int tid = threadIdx.x;
    auto load_g = [&](int i) { return seg[i]; };
    const float w1 = 1.5f / (float)S, w0 = 0.75f / (float)S;
    const int half = S / 2;
    auto store_mul = [&](int k2, float2 v) {
        const int k = k1 + N1 * k2;
        if constexpr (MODE == LROW_HILBERT) {
            if (k > half) {
                v = make_float2(0.0f, 0.0f);
            } else {
                const float w = k == 0 ? w0 : w1;
                v = make_float2(v.x * w, v.y * w);
                if constexpr (FILTER) v = cmul(v, F[k]);
            }
        } else {
            v = cmul(v, G[k]);
        }
        lds[lds_idx(k2)] = v;
    };
    fft_passes<LN2_LOG, false>(tid, tw, lds, load_g, false, store_mul, false);
    auto load_lds = [&](int i) { return lds[lds_idx(i)]; };
    auto store_g = [&](int i, float2 v) { seg[i] = v; };
    const float2* tw2 = tw;
    __asm__ volatile("" : "+s"(tw2));
    fft_passes<LN2_LOG, true>(tid, tw2, lds, load_lds, true, store_g, true);
}

// grid (LN2 / 256, rows of the chunk): samples n < S of each row to out in
// format ofmt.
template <int N1>
__global__ __launch_bounds__(256) void long_inv_pass_kernel(int row0, int S, const float2* __restrict__ twN,
                                                            const float2* buf, void* out, int ofmt, float oscale)
{
    constexpr int N = N1 * LN2;
    const int n2 = blockIdx.x * 256 + threadIdx.x;
    const float2* brow = buf + (size_t)blockIdx.y * N;
    float2 u[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        const float2 v = brow[k1 * LN2 + n2];
        u[k1] = k1 > 0 ? cmul_tw<true>(v, twN[n2 * k1]) : v;
    }
    dft<N1, true>(u);
    const size_t orow = (size_t)(row0 + (int)blockIdx.y) * S;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int n = n2 + LN2 * n1;
        if (n < S) store_fmt(out, orow + n, u[n1], ofmt, oscale);
    }
}

size_t hilbert_long_chunk_rows(int S, int rows)
{
    const size_t per_row = (size_t)hilbert_long_n(S) * sizeof(float2);
    const size_t cap = (size_t)512 << 20;  // workspace bound
    return std::max<size_t>(1, std::min<size_t>((size_t)rows, cap / per_row));
}

bool hilbert_long_in_place(int S, int ofmt) { return ofmt == IQ_F32 && (S & (S - 1)) == 0; }

template <int N1>
static bool launch_long_n1(const HilbertInput& in, int rows, int S, const LongHilbertTables& t, const float2* F,
                           void* out, int ofmt, float oscale, float2* ws, size_t ws_rows, hipStream_t st)
{
    HilbertDev d{in.ptr, in.stride, in.map, in.R0, in.C};
    const bool pow2 = (S & (S - 1)) == 0;
    const bool inplace = hilbert_long_in_place(S, ofmt);
    const int chunk = inplace ? rows : (int)ws_rows;
    if (!inplace && (!ws || chunk <= 0)) CUT_FAIL("Hilbert: no workspace for rows of " << S << " samples");
    if (!pow2 && !t.H) CUT_FAIL("Hilbert: no convolution table for rows of " << S << " samples");
    for (int r0 = 0; r0 < rows; r0 += chunk) {
        const int nr = std::min(chunk, rows - r0);
        float2* buf = inplace ? static_cast<float2*>(out) + (size_t)r0 * S : ws;
        const dim3 gp(LN2 / 256, nr), gr(N1, nr);
        if (in.kind == 0) long_fwd_pass_kernel<N1, 0><<<gp, 256, 0, st>>>(d, r0, S, t.twN, buf);
        else if (in.kind == 1) long_fwd_pass_kernel<N1, 1><<<gp, 256, 0, st>>>(d, r0, S, t.twN, buf);
        else long_fwd_pass_kernel<N1, 2><<<gp, 256, 0, st>>>(d, r0, S, t.twN, buf);
        if (!pow2) long_row_kernel<LROW_MUL, false><<<gr, 256, 0, st>>>(buf, N1, S, t.tw2, t.H, nullptr);
        else if (F) long_row_kernel<LROW_HILBERT, true><<<gr, 256, 0, st>>>(buf, N1, S, t.tw2, nullptr, F);
        else long_row_kernel<LROW_HILBERT, false><<<gr, 256, 0, st>>>(buf, N1, S, t.tw2, nullptr, nullptr);
        long_inv_pass_kernel<N1><<<gp, 256, 0, st>>>(r0, S, t.twN, buf, out, ofmt, oscale);
        CUT_HIP_CHECK(hipGetLastError());
    }
    return true;
}

bool launch_hilbert_long(const HilbertInput& in, int rows, int S, const LongHilbertTables& t, const float2* d_F,
                         void* out, int ofmt, float oscale, float2* ws, size_t ws_rows, hipStream_t st)
{
    if (rows <= 0) return true;
    if (ofmt != IQ_F32 && ofmt != IQ_F16 && ofmt != IQ_BF16) CUT_FAIL("Hilbert: unknown output format " << ofmt);
    if (!hilbert_long(S) || !hilbert_supported(S)) CUT_FAIL("Hilbert: sample count " << S << " not supported");
    if (!t.twN || !t.tw2) CUT_FAIL("Hilbert: twiddle tables of " << S << " samples missing");
    switch (hilbert_long_n(S) / LN2) {
        case 2: return launch_long_n1<2>(in, rows, S, t, d_F, out, ofmt, oscale, ws, ws_rows, st);
        case 4: return launch_long_n1<4>(in, rows, S, t, d_F, out, ofmt, oscale, ws, ws_rows, st);
        case 8: return launch_long_n1<8>(in, rows, S, t, d_F, out, ofmt, oscale, ws, ws_rows, st);
        case 16: return launch_long_n1<16>(in, rows, S, t, d_F, out, ofmt, oscale, ws, ws_rows, st);
        default: CUT_FAIL("Hilbert: sample count " << S << " not supported");
    }
}

}  // namespace cut
