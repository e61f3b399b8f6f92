/*
 * FP64 CPU oracle.  TEST INFRASTRUCTURE ONLY: see oracle.h.  The library never
 * links or calls this file; tests, smoke() and bench.py's cpu_baseline leg do.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Hadamard matrices.                                                        */
/* ------------------------------------------------------------------------ */

/* Quadratic character of a (mod q), q prime: +1 residue, -1 non-residue, 0. */
static int qchar(int a, int q)
{
    a %= q;
    if (a < 0) a += q;
    if (a == 0) return 0;
    for (int x = 1; x < q; ++x)
        if ((x * x) % q == a) return 1;
    return -1;
}

/* Base blocks of order 12 and 20 (Paley construction from the quadratic
 * character mod 11 and mod 19).  The arrangement (circulant in j-i for 12,
 * back-circulant in i+j for 20, unit first row and column) reproduces the
 * tables in hadamard_decoder.cpp:8-45 entry for entry; the golden fixture
 * tests/golden/hadamard_base_tables.json pins that. */
static double base_entry(int base, int i, int j)
{
    if (i == 0 || j == 0) return 1.0;
    if (base == 12) {
        if (i == j) return -1.0;
        return (double)(-qchar(j - i, 11));
    }
    int m = (i + j - 2) % 19;
    if (m == 0) return -1.0;
    return (double)(-qchar(m, 19));
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* hadamard_decoder.cpp:110-189 (generate_hadamard) and :191-222 (_sort_walsh).
 * Sylvester doubling of a base block placed top-left: entry (i,j) is the base
 * entry times (-1)^popcount((i/base) & (j/base)).  WALSH orders rows by the
 * number of sign changes along the row (stable sort; the reference's
 * std::sort only differs on ties, which occur for non-power-of-two orders). */
int or_hadamard(int n, int ordering, double* H)
{
    if (n <= 0 || n > 1024) return -1;
    int base;
    if (is_pow2(n)) base = 1;
    else if (n % 12 == 0 && is_pow2(n / 12)) base = 12;
    else if (n % 20 == 0 && is_pow2(n / 20)) base = 20;
    else return -1;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double b = base == 1 ? 1.0 : base_entry(base, i % base, j % base);
            int bits = __builtin_popcount((unsigned)((i / base) & (j / base)));
            H[(size_t)i * n + j] = (bits & 1) ? -b : b;
        }
    if (ordering == WALSH) {
        int* changes = (int*)calloc(n, sizeof(int));
        int* order = (int*)malloc(n * sizeof(int));
        double* tmp = (double*)malloc((size_t)n * n * sizeof(double));
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j + 1 < n; ++j)
                changes[i] += (H[(size_t)i * n + j] * H[(size_t)i * n + j + 1] < 0);
            order[i] = i;
        }
        for (int i = 1; i < n; ++i) { /* insertion sort: stable */
            int v = order[i], j = i - 1;
            while (j >= 0 && changes[order[j]] > changes[v]) { order[j + 1] = order[j]; --j; }
            order[j + 1] = v;
        }
        for (int i = 0; i < n; ++i)
            memcpy(tmp + (size_t)i * n, H + (size_t)order[i] * n, n * sizeof(double));
        memcpy(H, tmp, (size_t)n * n * sizeof(double));
        free(changes); free(order); free(tmp);
    }
    return 0;
}

/* hadamard_decoder.cpp:49-78.  The column-major SGEMM there computes
 * out[j][i] = (1/T) * sum_k H[j][k] * conv[k][i] with i = c*S + s. */
void or_decode(const double* conv, const double* H, int T, size_t cols, double* out)
{
    const double inv = 1.0 / (double)T;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)cols; ++i) {
        for (int j = 0; j < T; ++j) {
            double acc = 0.0;
            for (int k = 0; k < T; ++k) acc += H[(size_t)j * T + k] * conv[(size_t)k * cols + i];
            out[(size_t)j * cols + i] = acc * inv;
        }
    }
}

/* data_conversion_kernels.cuh:25-41: out[t][c][s] = raw[map[c]][t*S + s]. */
void or_convert(const void* raw, int type, const int16_t* map, size_t R0, int T, int C, int S, double* out)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < T; ++t)
        for (int c = 0; c < C; ++c) {
            size_t src = (size_t)(uint16_t)map[c] * R0 + (size_t)t * S;
            double* dst = out + ((size_t)t * C + c) * S;
            if (type == TYPE_I16) {
                const int16_t* r = (const int16_t*)raw + src;
                for (int s = 0; s < S; ++s) dst[s] = r[s];
            } else {
                const float* r = (const float*)raw + src;
                for (int s = 0; s < S; ++s) dst[s] = r[s];
            }
        }
}

/* Bandpass input (data_conversion_kernels.cuh:14-21, :46-64, :66-102;
 * data_converter.cu:64-131): every two samples of a raw row are an (I, Q)
 * pair; output sample s of transmit t, channel c is
 *   v = (I - iQ) exp(-2 pi i s fd / fs),  I, Q = raw[map[c]][2 (t S + s) + 0, 1],
 * then, with L > 0 taps, the centred FIR
 *   y[s] = sum_{i < L, 0 <= s + i - L/2 < S} v[s + i - L/2] h[L - 1 - i]  /  (number of such i),
 * else y = v.  out: interleaved complex [T][C][S]. */
void or_convert_demod(const void* raw, int type, const int16_t* map, size_t R0, int T, int C, int S, double fd,
                      double fs, const double* h, int L, double* out)
{
#pragma omp parallel for schedule(static)
    for (int tc = 0; tc < T * C; ++tc) {
        const int t = tc / C, c = tc % C;
        const size_t src = (size_t)(uint16_t)map[c] * R0 + 2 * (size_t)t * S;
        double* v = (double*)malloc(2 * (size_t)S * sizeof(double));
        for (int k = 0; k < S; ++k) {
            double I, Q;
            if (type == TYPE_I16) {
                I = ((const int16_t*)raw)[src + 2 * (size_t)k];
                Q = ((const int16_t*)raw)[src + 2 * (size_t)k + 1];
            } else {
                I = ((const float*)raw)[src + 2 * (size_t)k];
                Q = ((const float*)raw)[src + 2 * (size_t)k + 1];
            }
            const double ph = 2.0 * M_PI * fmod((double)k * fd / fs, 1.0);
            const double cr = cos(ph), ci = -sin(ph);
            v[2 * k] = I * cr + Q * ci;        /* (I - iQ)(cr + i ci) */
            v[2 * k + 1] = I * ci - Q * cr;
        }
        double* y = out + 2 * ((size_t)t * C + c) * S;
        if (L <= 0) {
            memcpy(y, v, 2 * (size_t)S * sizeof(double));
        } else {
            const int half = L / 2;
            for (int k = 0; k < S; ++k) {
                double ar = 0.0, ai = 0.0;
                int n = 0;
                for (int i = 0; i < L; ++i) {
                    const int j = k + i - half;
                    if (j < 0 || j >= S) continue;
                    ar += v[2 * j] * h[L - 1 - i];
                    ai += v[2 * j + 1] * h[L - 1 - i];
                    ++n;
                }
                y[2 * k] = ar / n;
                y[2 * k + 1] = ai / n;
            }
        }
        free(v);
    }
}

/* ------------------------------------------------------------------------ */
/* DFT / FFT / Hilbert.                                                      */
/* ------------------------------------------------------------------------ */

/* exp(sign * 2*pi*i * m / n) with the angle reduced exactly in integers. */
static void cis(long long m, long long n, int sign, double* c, double* s)
{
    m %= n;
    if (m < 0) m += n;
    double a = 2.0 * M_PI * (double)m / (double)n;
    *c = cos(a);
    *s = sign * sin(a);
}

/* Textbook O(n^2) DFT: X[k] = sum_n x[n] exp(sign*2*pi*i*k*n/N). */
void or_dft(const double* re, const double* im, int n, int sign, double* ore, double* oim)
{
    for (int k = 0; k < n; ++k) {
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < n; ++j) {
            double c, s;
            cis((long long)k * j, n, sign, &c, &s);
            double xr = re[j], xi = im ? im[j] : 0.0;
            ar += xr * c - xi * s;
            ai += xr * s + xi * c;
        }
        ore[k] = ar;
        oim[k] = ai;
    }
}

static void fft_pow2(double* re, double* im, int n, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double wr, wi;
            cis(k, len, sign, &wr, &wi);
            for (int b = 0; b < n; b += len) {
                int p = b + k, q = p + half;
                double xr = re[q] * wr - im[q] * wi;
                double xi = re[q] * wi + im[q] * wr;
                re[q] = re[p] - xr; im[q] = im[p] - xi;
                re[p] += xr; im[p] += xi;
            }
        }
    }
}

/* In-place unnormalised DFT of any length: radix-2 for powers of two,
 * Bluestein's chirp-z (via a power-of-two convolution) otherwise. */
void or_fft(double* re, double* im, int n, int sign)
{
    if (n <= 1) return;
    if (is_pow2(n)) { fft_pow2(re, im, n, sign); return; }
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    double* ar = (double*)calloc(m, sizeof(double));
    double* ai = (double*)calloc(m, sizeof(double));
    double* br = (double*)calloc(m, sizeof(double));
    double* bi = (double*)calloc(m, sizeof(double));
    double* wr = (double*)malloc(n * sizeof(double));
    double* wi = (double*)malloc(n * sizeof(double));
    for (int k = 0; k < n; ++k) { /* w_k = exp(sign*i*pi*k^2/n) */
        cis((long long)k * k, 2LL * n, sign, &wr[k], &wi[k]);
    }
    for (int k = 0; k < n; ++k) {
        ar[k] = re[k] * wr[k] - im[k] * wi[k];
        ai[k] = re[k] * wi[k] + im[k] * wr[k];
        br[k] = wr[k]; bi[k] = -wi[k];
        if (k) { br[m - k] = wr[k]; bi[m - k] = -wi[k]; }
    }
    fft_pow2(ar, ai, m, -1);
    fft_pow2(br, bi, m, -1);
    for (int k = 0; k < m; ++k) {
        double r = ar[k] * br[k] - ai[k] * bi[k];
        double i = ar[k] * bi[k] + ai[k] * br[k];
        ar[k] = r; ai[k] = i;
    }
    fft_pow2(ar, ai, m, 1);
    for (int k = 0; k < n; ++k) {
        double r = ar[k] / m, i = ai[k] / m;
        re[k] = r * wr[k] - i * wi[k];
        im[k] = r * wi[k] + i * wr[k];
    }
    free(ar); free(ai); free(br); free(bi); free(wr); free(wi);
}

/* hilbert_handler.cu:55-78: F = forward DFT_S of the taps zero-padded to S. */
void or_filter_spectrum(const double* h, int L, int S, double* spec)
{
    double* re = (double*)calloc(S, sizeof(double));
    double* im = (double*)calloc(S, sizeof(double));
    for (int i = 0; i < L && i < S; ++i) re[i] = h[i];
    or_fft(re, im, S, -1);
    for (int k = 0; k < S; ++k) { spec[2 * k] = re[k]; spec[2 * k + 1] = im[k]; }
    free(re); free(im);
}

/* hilbert_handler.cu:132-155 + hilbert_kernels.cuh:17-40.  Per row:
 * X = DFT_S(x) (R2C, bins 0..S/2), Y[k] = w_k X[k] F[k] for k <= S/2 with
 * w_0 = 0.75/S and w_k = 1.5/S otherwise (the Nyquist bin is not halved: the
 * kernel halves bin S/2+1, which is already zero), Y[k] = 0 above S/2, and
 * y = unnormalised inverse DFT_S(Y).  out is interleaved complex [rows][S]. */
void or_hilbert(const double* x, size_t rows, int S, const double* h, int L, double* out)
{
    double* F = NULL;
    if (L > 0) {
        F = (double*)malloc(2 * (size_t)S * sizeof(double));
        or_filter_spectrum(h, L, S, F);
    }
    const int half = S / 2;
#pragma omp parallel
    {
        double* re = (double*)malloc(S * sizeof(double));
        double* im = (double*)malloc(S * sizeof(double));
#pragma omp for schedule(static)
        for (long long r = 0; r < (long long)rows; ++r) {
            const double* xr = x + (size_t)r * S;
            for (int i = 0; i < S; ++i) { re[i] = xr[i]; im[i] = 0.0; }
            or_fft(re, im, S, -1);
            for (int k = 0; k < S; ++k) {
                if (k > half) { re[k] = 0.0; im[k] = 0.0; continue; }
                double w = (k == 0 ? 0.75 : 1.5) / (double)S;
                double a = re[k] * w, b = im[k] * w;
                if (F) {
                    double fr = F[2 * k], fi = F[2 * k + 1];
                    re[k] = a * fr - b * fi;
                    im[k] = a * fi + b * fr;
                } else {
                    re[k] = a; im[k] = b;
                }
            }
            or_fft(re, im, S, 1);
            double* o = out + 2 * (size_t)r * S;
            for (int i = 0; i < S; ++i) { o[2 * i] = re[i]; o[2 * i + 1] = im[i]; }
        }
        free(re); free(im);
    }
    free(F);
}

/* ------------------------------------------------------------------------ */
/* Delay and sum.                                                            */
/* ------------------------------------------------------------------------ */

/* beamformer.cu:17-98 (_params_to_constants), evaluated in float exactly as
 * the reference host code does. */
int or_das_constants(const CudaBeamformerParameters* bp, or_das_consts* k)
{
    memset(k, 0, sizeof(*k));
    k->S = (int)bp->dec_data_dim[0];
    k->C = (int)bp->dec_data_dim[1];
    k->T = (int)bp->dec_data_dim[2];
    k->spm = bp->sampling_frequency / bp->speed_of_sound;
    float px = bp->xdc_element_pitch[0], py = bp->xdc_element_pitch[1];
    k->pitch_x = px;
    k->pitch_y = py;
    k->xdc_min_x = -bp->xdc_transform[12] + px * 0.5f;
    k->xdc_min_y = -bp->xdc_transform[13] + py * 0.5f;
    k->xdc_max_x = bp->xdc_transform[12] - px * 0.5f;
    k->xdc_max_y = bp->xdc_transform[13] - py * 0.5f;
    k->delay = (int)(bp->time_offset * bp->sampling_frequency);
    k->sequence = bp->das_shader_id;
    k->Nx = (int)bp->output_points[0];
    k->Ny = (int)bp->output_points[1];
    k->Nz = (int)bp->output_points[2];
    k->min_x = bp->output_min_coordinate[0];
    k->min_y = bp->output_min_coordinate[1];
    k->min_z = bp->output_min_coordinate[2];
    k->res_x = (bp->output_max_coordinate[0] - bp->output_min_coordinate[0]) / (float)(bp->output_points[0] - 1);
    k->res_y = k->res_x;
    k->res_z = (bp->output_max_coordinate[2] - bp->output_min_coordinate[2]) / (float)(bp->output_points[2] - 1);
    k->fn_rx = bp->fn_rx;
    k->fn_tx = bp->fn_tx;
    k->readi_groups = (uint8_t)bp->readi_group_count;
    if (k->readi_groups == 0) k->readi_groups = 1;
    k->readi_id = (uint8_t)bp->readi_group_id;
    k->encoding = bp->decode;
    float fz = bp->foci[2];
    if (bp->tx_orientation == ORIENT_NONE) {
        k->focus_type = 0;
    } else if (isinf(fz) || bp->das_shader_id == RCA_TPW) {
        fz = 0.0f;
        k->focus_type = bp->tx_orientation == bp->rx_orientation ? 1 : 2;
    } else {
        k->focus_type = bp->tx_orientation == bp->rx_orientation ? 3 : 4;
    }
    k->focal_z = fz;
    float cw = bp->coherency_weighting;
    k->cw = cw < 0.0f ? 0.0f : (cw > 1.0f ? 1.0f : cw);
    k->apo_type = bp->apo_type;
    k->to_power = bp->to_power;
    k->fs = bp->sampling_frequency;
    k->fc = bp->center_frequency;
    for (int t = 0; t < MAX_ACQ_COUNT; ++t) {
        k->tx_angle[t] = bp->foci[3 * t + 2];
        k->tx_element[t] = bp->sparse_elements[t];
    }
    return 0;
}

/* READI sign row (beamformer.cu:238-252 / :355-372): bit i set when entry i of
 * row readi_group_id of the group-sized decode matrix is negative. */
int or_readi_row_bits(const or_das_consts* k, uint64_t* bits)
{
    *bits = 0;
    if (k->readi_groups <= 1) return 0;
    int g = k->readi_groups;
    double* H = (double*)malloc((size_t)g * g * sizeof(double));
    if (or_hadamard(g, k->encoding, H) != 0) { free(H); return -1; }
    for (int i = 0; i < g && i < 64; ++i)
        if (H[(size_t)k->readi_id * g + i] < 0) *bits |= (1ULL << i);
    free(H);
    return 0;
}

/* beamformer_utils.cuh:142-154 */
static double hann_apo(double lateral, double depth, double fnum)
{
    double a = fnum * (lateral / depth) / 2.0;
    if (a > 0.5) a = 0.5;
    a = cos(M_PI * a);
    return a * a;
}

/* templated_beamformer.cuh:86-99 */
static double sin_to_apo(double pos, double peak, int power)
{
    double a = M_PI * pos / peak;
    if (a < -M_PI) a = -M_PI;
    if (a > M_PI) a = M_PI;
    a = sin(a);
    double out = 1.0;
    for (int i = 0; i < power; ++i) out *= a;
    return out;
}

/* beamformer_utils.cuh:99-140: Catmull-Rom (tangents (p[i+1]-p[i-1])/2) or
 * linear read of an interleaved complex row.  The cubic form reads one sample
 * past the row at x = S-2 with zero weight; that sample is taken as 0. */
static void interp_row(const double* row, int S, double x, int linear, double* vr, double* vi)
{
    int i = (int)floor(x);
    double t = x - i;
#define P(j, c) (((j) >= 0 && (j) < S) ? row[2 * (size_t)(j) + (c)] : 0.0)
    if (linear) {
        *vr = P(i, 0) + t * (P(i + 1, 0) - P(i, 0));
        *vi = P(i, 1) + t * (P(i + 1, 1) - P(i, 1));
        return;
    }
    double t2 = t * t;
    double h00 = (2.0 * t - 3.0) * t2 + 1.0;
    double h10 = (t - 2.0) * t2 + t;
    double h01 = (-2.0 * t + 3.0) * t2;
    double h11 = (t - 1.0) * t2;
    for (int c = 0; c < 2; ++c) {
        double m0 = 0.5 * (P(i + 1, c) - P(i - 1, c));
        double m1 = 0.5 * (P(i + 2, c) - P(i, c));
        double v = h00 * P(i, c) + h01 * P(i + 1, c) + h10 * m0 + h11 * m1;
        if (c == 0) *vr = v; else *vi = v;
    }
#undef P
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* beamformer_utils.cuh:142-154 with fminf's NaN rule (a NaN ratio, from a
 * voxel at depth 0 under the element, clamps to 1/2). */
static double hann_fmin(double lateral, double depth, double fnum)
{
    double a = fmin(fnum * (lateral / depth) / 2.0, 0.5);
    a = cos(M_PI * a);
    return a * a;
}

/* beamformer_utils.cuh:56-97 (cubic_spline) and :127-140 (lerp_read) as the
 * UFORCES kernel calls them: sample j of the row at `base` is element
 * base + j of the flat [T][C][S] buffer, so j = -1 reads the previous row's
 * last sample.  Elements past the buffer's end (reached only with zero
 * weight, at x = S-2) read 0. */
static void interp_flat(const double* iq, size_t n, size_t base, double x, int linear, double* vr, double* vi)
{
    long long i = (long long)floor(x);
    double t = x - (double)i;
#define Q(j, c) ((long long)base + (j) >= 0 && (size_t)((long long)base + (j)) < n ? iq[2 * (size_t)((long long)base + (j)) + (c)] : 0.0)
    if (linear) {
        *vr = Q(i, 0) + t * (Q(i + 1, 0) - Q(i, 0));
        *vi = Q(i, 1) + t * (Q(i + 1, 1) - Q(i, 1));
        return;
    }
    double t2 = t * t;
    double h00 = (2.0 * t - 3.0) * t2 + 1.0;
    double h10 = (t - 2.0) * t2 + t;
    double h01 = (-2.0 * t + 3.0) * t2;
    double h11 = (t - 1.0) * t2;
    for (int c = 0; c < 2; ++c) {
        double m0 = 0.5 * (Q(i + 1, c) - Q(i - 1, c));
        double m1 = 0.5 * (Q(i + 2, c) - Q(i, c));
        double v = h00 * Q(i, c) + h01 * Q(i + 1, c) + h10 * m0 + h11 * m1;
        if (c == 0) *vr = v; else *vi = v;
    }
#undef Q
}

/* Taps of a UFORCES voxel with nonzero Hann weight (ratio below 1/2): the
 * ones the GPU kernel evaluates.  The others carry cos^2(pi/2), 0 to within
 * rounding in the reference. */
static int fnum_ratio_inside(double lateral, double depth, double fnum)
{
    return fnum * (lateral / depth) / 2.0 < 0.5;
}

/* UFORCES voxel (beamformer_kernels.cu:94-172).  Transmit t + 1 (row block 0
 * is a discarded flash) is fired from element sparse_elements[t]; the
 * transmit vector starts as calc_tx_distance(vox, focal, type)
 * (beamformer_utils.cuh:8-32) with focal = (x_min + p/2, vox.y, 0) and moves
 * by (sparse_elements[t] - sparse_elements[0]) pitches in x; receive element e
 * sits at x_min + p/2 + e p.  Path = |rx| + |tx| sign(vox.z), index clamped to
 * [0, S-2], Catmull-Rom over the flat buffer, 1/sqrt(128) on the first used
 * transmit, Hann receive apodization without a threshold, and the incoherent
 * sum over the apodized values. */
static void uforces_voxel(const double* iq, const or_das_consts* k, double vx, double vy, double vz, int interp,
                          double* ar, double* ai, double* inc, long long* act)
{
    const int S = k->S, C = k->C, T = k->T;
    const size_t n = (size_t)T * C * S;
    const double px = k->pitch_x;
    const double fx = (double)k->xdc_min_x + px / 2.0;
    double tx0 = 0.0, ty = 0.0, tz = 0.0;
    switch (k->focus_type) {
        case 3: tx0 = fx - vx; tz = -vz; break;                       /* XZ_FOCUS, focus depth 0 */
        case 4: tz = -vz; break;                                      /* YZ_FOCUS: focal.y = vox.y */
        case 1: tx0 = vx * sin(fx); tz = vz * cos(fx); break;         /* XZ_PLANE: focal.x as the angle */
        case 2: ty = vy * sin(vy); tz = vz * cos(vy); break;          /* YZ_PLANE: focal.y as the angle */
        default: break;
    }
    const double sgn = copysign(1.0, vz);
    const double rx0 = (double)k->xdc_min_x - vx + px / 2.0;
    for (int t = 0; t + 1 < T; ++t) {
        const double txx = tx0 + (double)(k->tx_element[t] - k->tx_element[0]) * px;
        const double txd = sqrt(txx * txx + ty * ty + tz * tz) * sgn;
        const double scale = t == 0 ? 1.0 / sqrt(128.0) : 1.0;
        for (int e = 0; e < C; ++e) {
            const double rxx = rx0 + (double)e * px;
            const double apo = hann_fmin(fabs(rxx), vz, k->fn_rx);
            const double idx = clampd((sqrt(rxx * rxx + vz * vz) + txd) * (double)k->spm + k->delay, 0.0, S - 2.0);
            double pr, pi;
            interp_flat(iq, n, ((size_t)(t + 1) * C + e) * S, idx, interp, &pr, &pi);
            pr *= scale * apo;
            pi *= scale * apo;
            *ar += pr; *ai += pi;
            *inc += pr * pr + pi * pi;
            if (fnum_ratio_inside(fabs(rxx), vz, k->fn_rx)) ++*act;
        }
    }
}

/* RCA_TPW voxel (beamformer_kernels.cu:489-546): tilted plane waves with
 * steering angle foci[3a+2] along y (YZ_PLANE) or x (XZ_PLANE); receive
 * element c at x_min + p/2 + c p on the x axis.  Channels with Hann
 * apodization below 0.1 are skipped; index clamped to [1, S-2]; Catmull-Rom
 * read; the sample is rotated by exp(+2 pi i index fc / fs) (rotate_iq,
 * :482-487) before apodization; incoherent sum over the apodized values. */
static void tpw_voxel(const double* iq, const or_das_consts* k, double vx, double vy, double vz, int interp,
                      double* ar, double* ai, double* inc, long long* act)
{
    const int S = k->S, C = k->C, T = k->T;
    const double px = k->pitch_x;
    const double lat = k->focus_type == 2 ? vy : vx;
    const double rx0 = (double)k->xdc_min_x - vx + px / 2.0;
    const double f = (double)k->fc / (double)k->fs;
    for (int c = 0; c < C; ++c) {
        const double rxx = rx0 + (double)c * px;
        const double rxd = sqrt(rxx * rxx + vz * vz);
        const double apo = hann_fmin(fabs(rxx), vz, k->fn_rx);
        if (apo < 0.1) continue;
        for (int a = 0; a < T; ++a) {
            const double th = k->tx_angle[a];
            const double txd = lat * sin(th) + vz * cos(th);
            const double idx = clampd((rxd + txd) * (double)k->spm + k->delay, 1.0, S - 2.0);
            double pr, pi;
            interp_row(iq + 2 * ((size_t)a * C + c) * S, S, idx, interp, &pr, &pi);
            const double ph = 2.0 * M_PI * idx * f;
            const double cr = cos(ph), ci = sin(ph);
            const double qr = (pr * cr - pi * ci) * apo, qi = (pr * ci + pi * cr) * apo;
            *ar += qr; *ai += qi;
            *inc += qr * qr + qi * qi;
            ++*act;
        }
    }
}

/* HERCULES: templated_beamformer.cuh:102-216; FORCES: :219-324. */
int or_das(const double* iq, const or_das_consts* k, int interp, int cf_mode, double* out, long long* active)
{
    return or_das_planes(iq, k, interp, cf_mode, NULL, k->Nz, out, active);
}

int or_das_planes(const double* iq, const or_das_consts* k, int interp, int cf_mode, const int* planes, int n_planes,
                  double* out, long long* active)
{
    const int S = k->S, C = k->C, T = k->T;
    const int G = k->readi_groups;
    uint64_t row_bits = 0;
    if (G > 1 && or_readi_row_bits(k, &row_bits) != 0) return -1;
    const int hercules = k->sequence == HERCULES;
    const int legacy = k->sequence == UFORCES || k->sequence == RCA_TPW;
    if (!hercules && !legacy && k->sequence != FORCES) return -1;
    if (hercules && k->focus_type != 2 && k->focus_type != 4) return -1;
    if (k->sequence == FORCES && G > 1 && k->encoding != HADAMARD) return -1;
    if (legacy && (T > MAX_ACQ_COUNT || k->focus_type == 0)) return -1;
    if (k->sequence == RCA_TPW && k->focus_type != 1 && k->focus_type != 2) return -1;
    if (k->sequence == UFORCES && k->tx_element[0] == -1) return -1;
    if (planes) {
        for (int q = 0; q < n_planes; ++q)
            if (planes[q] < 0 || planes[q] >= k->Nz) return -1;
    } else if (n_planes != k->Nz) {
        return -1;
    }
    const long long nvox = (long long)k->Nx * k->Ny * n_planes;
    long long act_total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : act_total)
    for (long long v = 0; v < nvox; ++v) {
        int ix = (int)(v % k->Nx);
        int iy = (int)((v / k->Nx) % k->Ny);
        int pq = (int)(v / ((long long)k->Nx * k->Ny));
        int iz = planes ? planes[pq] : pq;
        double vx = (double)k->min_x + (double)ix * (double)k->res_x;
        double vy = (double)k->min_y + (double)iy * (double)k->res_y;
        double vz = (double)k->min_z + (double)iz * (double)k->res_z;
        double ar = 0.0, ai = 0.0, inc = 0.0;
        long long act = 0;
        if (k->sequence == UFORCES) {
            uforces_voxel(iq, k, vx, vy, vz, interp, &ar, &ai, &inc, &act);
        } else if (k->sequence == RCA_TPW) {
            tpw_voxel(iq, k, vx, vy, vz, interp, &ar, &ai, &inc, &act);
        } else if (hercules) {
            double fz = k->focal_z;
            double tx_y = 0.0, tx_z = vz; /* YZ_PLANE with zero steering */
            if (k->focus_type == 4) { tx_y = vy; tx_z = vz - fz; }
            double txd = copysign(sqrt(tx_y * tx_y + tx_z * tx_z), tx_z) + fz;
            for (int g = 0; g < (G > 1 ? G : 1); ++g)
                for (int t = 0; t < T; ++t) {
                    double cur = t;
                    if (G > 1) cur = k->encoding == WALSH ? (double)t * G + g : (double)g * T + t;
                    double ey = (double)k->xdc_min_y + cur * k->pitch_y;
                    double ry = ey - vy;
                    int neg = 0;
                    if (G > 1) {
                        int bit = k->encoding == WALSH ? ((t & 1) ? G - 1 - g : g) : g;
                        neg = (int)((row_bits >> bit) & 1ULL);
                    }
                    for (int c = 0; c < C; ++c) {
                        double ex = (double)k->xdc_min_x + (double)c * k->pitch_x;
                        double rx = ex - vx;
                        double apo;
                        if (k->apo_type == RX_TO_SIN)
                            apo = sin_to_apo(ex, k->xdc_max_x, k->to_power) * sin_to_apo(ey, k->xdc_max_y, k->to_power);
                        else
                            apo = hann_apo(sqrt(rx * rx + ry * ry), vz, k->fn_rx);
                        if (!(apo > 0.1)) continue;
                        ++act;
                        double idx = (txd + sqrt(rx * rx + ry * ry + vz * vz) + fz) * (double)k->spm + k->delay;
                        idx = clampd(idx, 1.0, S - 2.0);
                        double pr, pi;
                        interp_row(iq + 2 * ((size_t)t * C + c) * S, S, idx, interp, &pr, &pi);
                        double a = neg ? -apo : apo;
                        ar += a * pr; ai += a * pi;
                        inc += pr * pr + pi * pi;
                    }
                }
        } else {
            for (int c = 0; c < C; ++c) {
                double ec = (double)k->xdc_min_x + (double)c * k->pitch_x;
                double rx = ec - vx;
                double rxd = sqrt(rx * rx + vz * vz);
                double rx_apo;
                if (k->apo_type == RX_HANN || k->apo_type == TX_TO_SIN)
                    rx_apo = hann_apo(fabs(rx), vz, k->fn_rx);
                else
                    rx_apo = sin_to_apo(ec, k->xdc_max_x, k->to_power);
                if (!(rx_apo > 0.0)) continue;
                for (int g = 0; g < G; ++g) {
                    int neg = G > 1 ? (int)((row_bits >> g) & 1ULL) : 0;
                    for (int t = 0; t < T; ++t) {
                        double et = (double)k->xdc_min_x + (double)(g * T + t) * k->pitch_x;
                        double tx = vx - et;
                        double tx_apo;
                        if (k->apo_type == BOTH_TO_SIN || k->apo_type == TX_TO_SIN)
                            tx_apo = sin_to_apo(et, k->xdc_max_x, k->to_power);
                        else
                            tx_apo = hann_apo(fabs(tx), vz, k->fn_tx);
                        double apo = rx_apo * tx_apo;
                        ++act;
                        double idx = (sqrt(tx * tx + vz * vz) + rxd) * (double)k->spm + k->delay;
                        idx = clampd(idx, 1.0, S - 2.0);
                        double pr, pi;
                        interp_row(iq + 2 * ((size_t)t * C + c) * S, S, idx, interp, &pr, &pi);
                        double a = neg ? -apo : apo;
                        ar += a * pr; ai += a * pi;
                        inc += pr * pr + pi * pi;
                    }
                }
            }
        }
        if (!hercules || cf_mode == 1) {
            double cf = pow((ar * ar + ai * ai) / inc, (double)k->cw);
            if (isnan(cf)) cf = 0.0;
            ar *= cf; ai *= cf;
        }
        out[2 * v] = ar;
        out[2 * v + 1] = ai;
        act_total += act;
    }
    if (active) *active = act_total;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 2-D block matching.                                                       */
/* ------------------------------------------------------------------------ */

/* Zero-mean normalised cross-correlation over the valid region
 * (nppiCrossCorrValid_NormLevel, called at block_match_kernels.cu:214-217):
 * out[v][u] = sum (T - mean T)(S_uv - mean S_uv) / sqrt(sum (T-mT)^2 sum (S_uv-mS_uv)^2),
 * with the window S_uv = S[v .. v+th-1][u .. u+tw-1].  A zero denominator gives 0. */
void or_zncc_valid(const double* src, int sw, int sh, int sstride,
                   const double* tpl, int tw, int th, int tstride, double* out)
{
    const int vw = sw - tw + 1, vh = sh - th + 1;
    const double n = (double)tw * th;
    double tm = 0.0;
    for (int y = 0; y < th; ++y)
        for (int x = 0; x < tw; ++x) tm += tpl[(size_t)y * tstride + x];
    tm /= n;
    double te = 0.0;
    for (int y = 0; y < th; ++y)
        for (int x = 0; x < tw; ++x) {
            double d = tpl[(size_t)y * tstride + x] - tm;
            te += d * d;
        }
    for (int v = 0; v < vh; ++v)
        for (int u = 0; u < vw; ++u) {
            double sm = 0.0;
            for (int y = 0; y < th; ++y)
                for (int x = 0; x < tw; ++x) sm += src[(size_t)(v + y) * sstride + u + x];
            sm /= n;
            double cross = 0.0, se = 0.0;
            for (int y = 0; y < th; ++y)
                for (int x = 0; x < tw; ++x) {
                    double s = src[(size_t)(v + y) * sstride + u + x] - sm;
                    double t = tpl[(size_t)y * tstride + x] - tm;
                    cross += s * t;
                    se += s * s;
                }
            double den = sqrt(te * se);
            out[(size_t)v * vw + u] = den > 0.0 ? cross / den : 0.0;
        }
}

/* 3x3 least-squares quadratic fit matrix (block_match_kernels.cuh:130-154):
 * rows give a (x^2), b (y^2), c (xy), d (x), e (y), f (1) of
 * a x^2 + b y^2 + c xy + d x + e y + f fitted to the 3x3 neighbourhood
 * (row-major, x fastest, offsets -1..1). */
static const double P3[6][9] = {
    {1.0 / 6, -1.0 / 3, 1.0 / 6, 1.0 / 6, -1.0 / 3, 1.0 / 6, 1.0 / 6, -1.0 / 3, 1.0 / 6},
    {1.0 / 6, 1.0 / 6, 1.0 / 6, -1.0 / 3, -1.0 / 3, -1.0 / 3, 1.0 / 6, 1.0 / 6, 1.0 / 6},
    {0.25, 0.0, -0.25, 0.0, 0.0, 0.0, -0.25, 0.0, 0.25},
    {-1.0 / 6, 0.0, 1.0 / 6, -1.0 / 6, 0.0, 1.0 / 6, -1.0 / 6, 0.0, 1.0 / 6},
    {-1.0 / 6, -1.0 / 6, -1.0 / 6, 0.0, 0.0, 0.0, 1.0 / 6, 1.0 / 6, 1.0 / 6},
    {-1.0 / 9, 2.0 / 9, -1.0 / 9, 2.0 / 9, 5.0 / 9, 2.0 / 9, -1.0 / 9, 2.0 / 9, -1.0 / 9},
};

/* One grid point: block_match_pipeline (block_match_kernels.cu:190-240) =
 * ZNCC map -> find_peaks_kernel (:7-64: the first 2048 map entries, missing
 * ones as -1, sorted descending with ties in index order, top 32) ->
 * test_peaks (:67-185) including its 32-lane shuffle-down max tree. */
static void block_match_one(const double* tpl_img, const double* src_img, int W, int H,
                             const NCCMotionParameters* p, int gi, int gj, float* out4)
{
    const int P = (int)p->patch_size, sp = (int)p->motion_grid_spacing;
    const int mx = (int)p->search_margins[0], my = (int)p->search_margins[1];
    const int half = (int)(p->patch_size / 2u);
    int tl = gj * sp - half, tt = gi * sp - half;
    int tr = tl + P - 1, tb = tt + P - 1;
    int sl = tl - mx, sr = tr + mx, st = tt - my, sbm = tb + my;
    if (tl < 0) tl = 0;
    if (tr > W - 1) tr = W - 1;
    if (sl < 0) sl = 0;
    if (sr > W - 1) sr = W - 1;
    if (tt < 0) tt = 0;
    if (tb > H - 1) tb = H - 1;
    if (st < 0) st = 0;
    if (sbm > H - 1) sbm = H - 1;
    out4[0] = out4[1] = out4[2] = out4[3] = 0.0f;
    const int tw = tr - tl + 1, th = tb - tt + 1, sw = sr - sl + 1, sh = sbm - st + 1;
    if (tw <= 0 || th <= 0 || sw < tw || sh < th) return; /* NPP rejects the ROI: entry stays 0 */
    const int vw = sw - tw + 1, vh = sh - th + 1;
    double* map = (double*)malloc((size_t)vw * vh * sizeof(double));
    or_zncc_valid(src_img + (size_t)st * W + sl, sw, sh, W, tpl_img + (size_t)tt * W + tl, tw, th, W, map);
    const int nsx = tl - sl, nsy = tt - st;
    const int total = vw * vh;
    /* top 32 of 2048 virtual entries, stable descending */
    int idx[32];
    double val[32];
    char* taken = (char*)calloc(2048, 1);
    for (int r = 0; r < 32; ++r) {
        int best = -1;
        double bv = 0.0;
        for (int i = 0; i < 2048; ++i) {
            if (taken[i]) continue;
            double v = i < total ? map[i] : -1.0;
            if (best < 0 || v > bv) { best = i; bv = v; }
        }
        taken[best] = 1;
        idx[r] = best;
        val[r] = bv;
    }
    free(taken);
    const double no_shift_peak = map[(size_t)nsy * vw + nsx];
    double pk[32], ox[32], oy[32];
    for (int c = 0; c < 32; ++c) {
        const int px = idx[c] % vw, py = idx[c] / vw;
        double v9[9];
        int k = 0;
        for (int y = -1; y <= 1; ++y)
            for (int x = -1; x <= 1; ++x, ++k) {
                int yy = py + y, xx = px + x;
                v9[k] = (yy < 0 || yy >= vh || xx < 0 || xx >= vw) ? 0.0 : map[(size_t)yy * vw + xx];
            }
        double co[6];
        for (int i = 0; i < 6; ++i) {
            co[i] = 0.0;
            for (int j = 0; j < 9; ++j) co[i] += P3[i][j] * v9[j];
        }
        const double root = sqrt(co[0] * co[0] + co[1] * co[1] + co[2] * co[2] - 2 * co[0] * co[1]);
        const double s0 = co[0] + co[1] + root, s1 = co[0] + co[1] - root;
        const double min_sharp = fmin(fabs(s0), fabs(s1));
        const double den = co[2] * co[2] - 4 * co[1] * co[0];
        double sx = 0.0, sy = 0.0;
        if (den != 0.0) {
            sx = (2 * co[1] * co[3] - co[2] * co[4]) / den;
            sy = (2 * co[0] * co[4] - co[2] * co[3]) / den;
        }
        sx = fmin(fmax(sx, -1.0), 1.0) * (p->use_subpixel ? 1.0 : 0.0);
        sy = fmin(fmax(sy, -1.0), 1.0) * (p->use_subpixel ? 1.0 : 0.0);
        ox[c] = sx + px - nsx;
        oy[c] = sy + py - nsy;
        const double thr = no_shift_peak * (1.0 + p->rel_cor_threshold);
        double peak = val[c];
        if (min_sharp < p->min_patch_variance || s0 >= 0.0 || peak > 1.0 || peak < thr) peak = -2.0;
        pk[c] = peak;
    }
    /* __shfl_down_sync tree over 32 lanes: lanes past the end keep their own value */
    for (int off = 16; off > 0; off >>= 1) {
        double npk[32], nox[32], noy[32];
        for (int l = 0; l < 32; ++l) {
            const int s = l + off < 32 ? l + off : l;
            if (pk[s] > pk[l]) { npk[l] = pk[s]; nox[l] = ox[s]; noy[l] = oy[s]; }
            else { npk[l] = pk[l]; nox[l] = ox[l]; noy[l] = oy[l]; }
        }
        memcpy(pk, npk, sizeof pk); memcpy(ox, nox, sizeof ox); memcpy(oy, noy, sizeof oy);
    }
    double peak = pk[0], dx = ox[0], dy = oy[0];
    const double thr = no_shift_peak * (1.0 + p->rel_cor_threshold);
    if (peak < p->abs_cor_threshold || peak <= no_shift_peak) {
        dx = dy = 0.0;
        if (peak < thr) peak = -1.0;
    }
    out4[0] = (float)dx;
    out4[1] = (float)dy;
    out4[2] = (float)peak;
    out4[3] = (float)no_shift_peak;
    free(map);
}

/* ImageProcessor::ncc_block_match (image_processor.cpp:10-56) over f32 frames
 * [frame][H][W]; maps: float4 [frame][grid_y][grid_x], the reference frame's zero. */
int or_motion_detection(const float* images, const NCCMotionParameters* p, float* maps)
{
    const int W = (int)p->image_dims[0], H = (int)p->image_dims[1], F = (int)p->frame_count;
    const int gx = (int)p->motion_grid_dims[0], gy = (int)p->motion_grid_dims[1];
    memset(maps, 0, sizeof(float) * 4 * (size_t)gx * gy * F);
    double* imgs = (double*)malloc(sizeof(double) * (size_t)W * H * F);
    for (size_t i = 0; i < (size_t)W * H * F; ++i) imgs[i] = images[i];
    for (int f = 0; f < F; ++f) {
        int ref = (int)p->reference_frame;
        if (f == ref) continue;
        if (p->neighbour_compare) ref = ref > f ? f + 1 : f - 1;
        const double* tpl = imgs + (size_t)ref * W * H;
        const double* src = imgs + (size_t)f * W * H;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int gi = 0; gi < gy; ++gi)
            for (int gj = 0; gj < gx; ++gj)
                block_match_one(tpl, src, W, H, p, gi, gj, maps + 4 * (((size_t)f * gy + gi) * gx + gj));
    }
    free(imgs);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3-D ZNCC block matching.                                                  */
/* ------------------------------------------------------------------------ */
double or_zncc3d_at(const float* T, const float* S, int nx, int ny, int nz, int bx, int by, int bz, int B,
                    int dx, int dy, int dz)
{
    const int sx = bx + dx, sy = by + dy, sz = bz + dz;
    if (sx < 0 || sy < 0 || sz < 0 || sx + B > nx || sy + B > ny || sz + B > nz) return -INFINITY;
    const double n = (double)B * B * B;
    double tm = 0.0, sm = 0.0;
    for (int z = 0; z < B; ++z)
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                tm += T[((size_t)(bz + z) * ny + by + y) * nx + bx + x];
                sm += S[((size_t)(sz + z) * ny + sy + y) * nx + sx + x];
            }
    tm /= n;
    sm /= n;
    double c = 0.0, te = 0.0, se = 0.0;
    for (int z = 0; z < B; ++z)
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                const double t = T[((size_t)(bz + z) * ny + by + y) * nx + bx + x] - tm;
                const double s = S[((size_t)(sz + z) * ny + sy + y) * nx + sx + x] - sm;
                c += t * s;
                te += t * t;
                se += s * s;
            }
    const double d = sqrt(te * se);
    return d > 0.0 ? c / d : 0.0;
}

static void ncc3d_search(const float* T, const float* S, int nx, int ny, int nz, int B, int stride, int margin,
                         const int* centre, int radius, int* disp, float* peak)
{
    const int gx = (nx - B) / stride + 1, gy = (ny - B) / stride + 1, gz = (nz - B) / stride + 1;
    const long long nb = (long long)gx * gy * gz;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long b = 0; b < nb; ++b) {
        const int ix = (int)(b % gx), iy = (int)((b / gx) % gy), iz = (int)(b / ((long long)gx * gy));
        const int bx = ix * stride, by = iy * stride, bz = iz * stride;
        int c0[3] = {0, 0, 0};
        int r = margin;
        if (centre) { c0[0] = centre[3 * b]; c0[1] = centre[3 * b + 1]; c0[2] = centre[3 * b + 2]; r = radius; }
        double best = -INFINITY;
        int bd[3] = {0, 0, 0};
        for (int dz = c0[2] - r; dz <= c0[2] + r; ++dz)
            for (int dy = c0[1] - r; dy <= c0[1] + r; ++dy)
                for (int dx = c0[0] - r; dx <= c0[0] + r; ++dx) {
                    if (dx < -margin || dx > margin || dy < -margin || dy > margin || dz < -margin || dz > margin)
                        continue;
                    const double v = or_zncc3d_at(T, S, nx, ny, nz, bx, by, bz, B, dx, dy, dz);
                    if (v > best) { best = v; bd[0] = dx; bd[1] = dy; bd[2] = dz; }
                }
        disp[3 * b] = bd[0];
        disp[3 * b + 1] = bd[1];
        disp[3 * b + 2] = bd[2];
        peak[b] = (float)best;
    }
}

void or_ncc3d_full(const float* T, const float* S, int nx, int ny, int nz, int B, int stride, int margin, int* disp,
                   float* peak)
{
    ncc3d_search(T, S, nx, ny, nz, B, stride, margin, NULL, 0, disp, peak);
}

void or_ncc3d_window(const float* T, const float* S, int nx, int ny, int nz, int B, int stride, int margin,
                     const int* centre, int radius, int* disp, float* peak)
{
    ncc3d_search(T, S, nx, ny, nz, B, stride, margin, centre, radius, disp, peak);
}

/* ------------------------------------------------------------------------ */
/* Displacement-proposal CNN, FP64 forward and backward.                     */
/* ------------------------------------------------------------------------ */
typedef struct { int ci, co, din, stride; } conv_spec;
static const conv_spec PNET[4] = {{2, 16, 16, 2}, {16, 32, 8, 2}, {32, 32, 4, 1}, {32, 32, 4, 2}};

int or_pnet_param_count(void)
{
    int n = 0;
    for (int l = 0; l < 4; ++l) n += PNET[l].co * PNET[l].ci * 27 + PNET[l].co;
    return n + 3 * 256 + 3;
}

static void conv_fwd(const conv_spec* c, const double* w, const double* b, const double* in, double* out)
{
    const int D = c->din, Do = D / c->stride, s = c->stride;
    for (int co = 0; co < c->co; ++co)
        for (int oz = 0; oz < Do; ++oz)
            for (int oy = 0; oy < Do; ++oy)
                for (int ox = 0; ox < Do; ++ox) {
                    double a = b[co];
                    for (int ci = 0; ci < c->ci; ++ci)
                        for (int k = 0; k < 27; ++k) {
                            const int iz = oz * s + k / 9 - 1, iy = oy * s + (k / 3) % 3 - 1, ix = ox * s + k % 3 - 1;
                            if (iz < 0 || iy < 0 || ix < 0 || iz >= D || iy >= D || ix >= D) continue;
                            a += w[(co * c->ci + ci) * 27 + k] * in[((size_t)(ci * D + iz) * D + iy) * D + ix];
                        }
                    out[((size_t)(co * Do + oz) * Do + oy) * Do + ox] = a > 0.0 ? a : 0.0;
                }
}

/* dout is the gradient w.r.t. the ReLU output; act is that output */
static void conv_bwd(const conv_spec* c, const double* w, const double* in, const double* act, const double* dout,
                     double* dw, double* db, double* din)
{
    const int D = c->din, Do = D / c->stride, s = c->stride;
    if (din) memset(din, 0, sizeof(double) * c->ci * D * D * D);
    for (int co = 0; co < c->co; ++co)
        for (int o = 0; o < Do * Do * Do; ++o) {
            const size_t oi = (size_t)co * Do * Do * Do + o;
            if (!(act[oi] > 0.0)) continue;
            const double g = dout[oi];
            const int oz = o / (Do * Do), oy = (o / Do) % Do, ox = o % Do;
            db[co] += g;
            for (int ci = 0; ci < c->ci; ++ci)
                for (int k = 0; k < 27; ++k) {
                    const int iz = oz * s + k / 9 - 1, iy = oy * s + (k / 3) % 3 - 1, ix = ox * s + k % 3 - 1;
                    if (iz < 0 || iy < 0 || ix < 0 || iz >= D || iy >= D || ix >= D) continue;
                    const size_t ii = ((size_t)(ci * D + iz) * D + iy) * D + ix;
                    dw[(co * c->ci + ci) * 27 + k] += g * in[ii];
                    if (din) din[ii] += g * w[(co * c->ci + ci) * 27 + k];
                }
        }
}

static size_t act_size(int l) { const int Do = PNET[l].din / PNET[l].stride; return (size_t)PNET[l].co * Do * Do * Do; }

static void pnet_run(const double* P, const double* x, double** acts, double* y)
{
    const double* p = P;
    const double* in = x;
    for (int l = 0; l < 4; ++l) {
        const double* w = p;
        p += PNET[l].co * PNET[l].ci * 27;
        const double* b = p;
        p += PNET[l].co;
        conv_fwd(&PNET[l], w, b, in, acts[l]);
        in = acts[l];
    }
    for (int o = 0; o < 3; ++o) {
        double a = p[3 * 256 + o];
        for (int i = 0; i < 256; ++i) a += p[o * 256 + i] * in[i];
        y[o] = a;
    }
}

void or_pnet_forward(const double* P, const double* in, int B, double* out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < B; ++b) {
        double* acts[4];
        for (int l = 0; l < 4; ++l) acts[l] = (double*)malloc(sizeof(double) * act_size(l));
        pnet_run(P, in + (size_t)b * 2 * 4096, acts, out + 3 * b);
        for (int l = 0; l < 4; ++l) free(acts[l]);
    }
}

double or_pnet_loss_grad(const double* P, const double* in, const double* target, int B, double* grad)
{
    const int np = or_pnet_param_count();
    memset(grad, 0, sizeof(double) * np);
    double loss = 0.0;
#pragma omp parallel
    {
        double* g = (double*)calloc(np, sizeof(double));
        double* acts[4];
        double* dacts[4];
        for (int l = 0; l < 4; ++l) {
            acts[l] = (double*)malloc(sizeof(double) * act_size(l));
            dacts[l] = (double*)malloc(sizeof(double) * act_size(l));
        }
        double lsum = 0.0;
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < B; ++b) {
            const double* x = in + (size_t)b * 2 * 4096;
            double y[3];
            pnet_run(P, x, acts, y);
            /* offsets of each parameter block */
            size_t off[10], o = 0;
            for (int l = 0; l < 4; ++l) {
                off[2 * l] = o; o += PNET[l].co * PNET[l].ci * 27;
                off[2 * l + 1] = o; o += PNET[l].co;
            }
            off[8] = o; off[9] = o + 768;
            double dy[3];
            for (int i = 0; i < 3; ++i) {
                const double e = y[i] - target[3 * b + i];
                lsum += 0.5 * e * e / B;
                dy[i] = e / B;
            }
            for (int i = 0; i < 3; ++i) {
                g[off[9] + i] += dy[i];
                for (int k = 0; k < 256; ++k) g[off[8] + i * 256 + k] += dy[i] * acts[3][k];
            }
            for (int k = 0; k < 256; ++k) {
                double a = 0.0;
                for (int i = 0; i < 3; ++i) a += dy[i] * P[off[8] + i * 256 + k];
                dacts[3][k] = a;
            }
            for (int l = 3; l >= 0; --l) {
                const double* lin = l ? acts[l - 1] : x;
                conv_bwd(&PNET[l], P + off[2 * l], lin, acts[l], dacts[l], g + off[2 * l], g + off[2 * l + 1],
                         l ? dacts[l - 1] : NULL);
            }
        }
#pragma omp critical
        {
            for (int i = 0; i < np; ++i) grad[i] += g[i];
            loss += lsum;
        }
        free(g);
        for (int l = 0; l < 4; ++l) { free(acts[l]); free(dacts[l]); }
    }
    return loss;
}
