// Synthetic transmit-encoded RF acquisitions (host, multi-threaded).
//
// Produces the inputs the benchmarks and tests run on: point scatterers, each
// echo a Gaussian-windowed tone burst at the delay of the beamformer's own
// geometry model, per single-element transmit; white noise; transmit encoding
// with the decode matrix (raw_k = sum_t H[t][k] r_t, so decoding with H gives
// r back); scaling to an int16 full scale.  Delay models:
//   linear array (FORCES):   (|p - e_t| + |p - e_c|) / c, elements on x
//   row-column (HERCULES):   (z_p + |p - (x_c, y_t, 0)|) / c
// Deterministic for a given spec: every random stream is derived from the seed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/cut_ext.h"
#include "pipeline.hpp"

namespace {

inline uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }
inline double normal(uint64_t& s)
{
    double u1 = uniform(s), u2 = uniform(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

template <typename F>
void parallel_for(long long n, F&& f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nt; ++w)
        th.emplace_back([&, w] {
            for (long long i = w; i < n; i += nt) f(i);
        });
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" bool cut_synth_rf(const CutSynthSpec* sp, int16_t* raw, int16_t* map)
{
    if (!sp || !raw) CUT_FAIL("synth: null argument");
    const int T = sp->T, C = sp->C, S = sp->S;
    if (T <= 0 || C <= 0 || S <= 0 || sp->n_scatterers < 0) CUT_FAIL("synth: bad dimensions");
    std::vector<int> H;
    if (sp->encoding != NONE && !cut::hadamard_matrix(T, sp->encoding, H)) return false;

    // scatterers
    uint64_t rng = sp->seed * 0x2545F4914F6CDD1Dull + 1;
    const int N = sp->n_scatterers;
    std::vector<double> px(N), py(N), pz(N), amp(N);
    for (int i = 0; i < N; ++i) {
        px[i] = sp->box_min[0] + (sp->box_max[0] - sp->box_min[0]) * uniform(rng) + sp->shift[0];
        py[i] = sp->box_min[1] + (sp->box_max[1] - sp->box_min[1]) * uniform(rng) + sp->shift[1];
        pz[i] = sp->box_min[2] + (sp->box_max[2] - sp->box_min[2]) * uniform(rng) + sp->shift[2];
        amp[i] = sp->amp_min + (sp->amp_max - sp->amp_min) * uniform(rng);
    }
    if (sp->decorrelation > 0.0f) {
        uint64_t r2 = sp->amp_seed * 0x9E3779B97F4A7C15ull + 7;
        for (int i = 0; i < N; ++i) {
            const double fresh = sp->amp_min + (sp->amp_max - sp->amp_min) * uniform(r2);
            amp[i] = (1.0 - sp->decorrelation) * amp[i] + sp->decorrelation * fresh;
        }
    }
    // channel map: decoded channel c lives in raw row map[c]
    std::vector<int16_t> perm(C);
    for (int c = 0; c < C; ++c) perm[c] = (int16_t)c;
    for (int c = C - 1; c > 0; --c) {
        const int j = (int)(splitmix(rng) % (uint64_t)(c + 1));
        std::swap(perm[c], perm[j]);
    }
    if (map) std::memcpy(map, perm.data(), C * sizeof(int16_t));

    const double half = 0.5 * sp->pitch;
    const double sigma = sp->cycles / (2.3548200450309493 * sp->fc);  // envelope FWHM = cycles periods
    const double support = 4.0 * sigma;
    std::vector<float> resp((size_t)T * C * S, 0.0f);
    const bool rca = sp->geometry == 1;
    const double ex0 = -0.5 * C * sp->pitch + half;
    const double ey0 = -0.5 * T * sp->pitch + half;

    parallel_for((long long)T * C, [&](long long row) {
        const int t = (int)(row / C), c = (int)(row % C);
        float* r = resp.data() + (size_t)row * S;
        const double xc = ex0 + c * sp->pitch;
        for (int i = 0; i < N; ++i) {
            double tau;
            if (rca) {
                const double yt = ey0 + t * sp->pitch;
                const double dx = px[i] - xc, dy = py[i] - yt;
                tau = (pz[i] + std::sqrt(dx * dx + dy * dy + pz[i] * pz[i])) / sp->c;
            } else {
                const double xt = ex0 + t * sp->pitch;
                const double d1 = std::hypot(px[i] - xt, pz[i]);
                const double d2 = std::hypot(px[i] - xc, pz[i]);
                tau = (d1 + d2) / sp->c;
            }
            const int s0 = std::max(0, (int)std::ceil((tau - support) * sp->fs));
            const int s1 = std::min(S - 1, (int)std::floor((tau + support) * sp->fs));
            for (int s = s0; s <= s1; ++s) {
                const double dt = s / (double)sp->fs - tau;
                r[s] += (float)(amp[i] * std::exp(-0.5 * dt * dt / (sigma * sigma)) * std::cos(2.0 * M_PI * sp->fc * dt));
            }
        }
    });
    float peak = 0.0f;
    for (float v : resp) peak = std::max(peak, std::fabs(v));
    if (sp->noise_frac > 0.0f) {
        const double sd = sp->noise_frac * (peak > 0 ? peak : 1.0f);
        // the noise stream also depends on amp_seed, so the two frames of a
        // motion pair (same seed, different amp_seed) carry independent noise
        const uint64_t nseed = sp->seed ^ (sp->amp_seed * 0xA24BAED4963EE407ull);
        parallel_for((long long)T * C, [&](long long row) {
            uint64_t s = nseed ^ (0xD1B54A32D192ED03ull * (uint64_t)(row + 1));
            float* r = resp.data() + (size_t)row * S;
            for (int i = 0; i < S; ++i) r[i] += (float)(sd * normal(s));
        });
    }
    // encode: raw_k[c][s] = sum_t H[t][k] r_t[c][s]
    std::vector<float> enc;
    const float* src = resp.data();
    if (sp->encoding != NONE) {
        enc.assign((size_t)T * C * S, 0.0f);
        const size_t CS = (size_t)C * S;
        std::vector<float> Ht((size_t)T * T);  // Ht[k][t] = H[t][k]
        for (int t = 0; t < T; ++t)
            for (int k = 0; k < T; ++k) Ht[(size_t)k * T + t] = (float)H[(size_t)t * T + k];
        // Sylvester order (H[t][k] = (-1)^popcount(t & k)): fast Walsh-Hadamard
        // transform in double, T log T per sample instead of T^2
        bool sylvester = (T & (T - 1)) == 0;
        for (int t = 0; sylvester && t < T; ++t)
            for (int k = 0; k < T; ++k)
                if (H[(size_t)t * T + k] != ((__builtin_popcount(t & k) & 1) ? -1 : 1)) {
                    sylvester = false;
                    break;
                }
        if (sylvester) {
            parallel_for((long long)C, [&](long long c) {
                std::vector<double> col(T);
                for (int s = 0; s < S; ++s) {
                    for (int t = 0; t < T; ++t) col[t] = resp[(size_t)t * CS + (size_t)c * S + s];
                    for (int h = 1; h < T; h <<= 1)
                        for (int i = 0; i < T; i += 2 * h)
                            for (int j = i; j < i + h; ++j) {
                                const double a = col[j], b = col[j + h];
                                col[j] = a + b;
                                col[j + h] = a - b;
                            }
                    for (int k = 0; k < T; ++k) enc[(size_t)k * CS + (size_t)c * S + s] = (float)col[k];
                }
            });
        } else parallel_for((long long)C, [&](long long c) {
            std::vector<float> col(T);
            for (int s = 0; s < S; ++s) {
                for (int t = 0; t < T; ++t) col[t] = resp[(size_t)t * CS + (size_t)c * S + s];
                for (int k = 0; k < T; ++k) {
                    const float* h = Ht.data() + (size_t)k * T;
                    double a = 0.0;
                    for (int t = 0; t < T; ++t) a += (double)(h[t] * col[t]);
                    enc[(size_t)k * CS + (size_t)c * S + s] = (float)a;
                }
            }
        });
        src = enc.data();
    }
    float emax = 0.0f;
    for (size_t i = 0; i < (size_t)T * C * S; ++i) emax = std::max(emax, std::fabs(src[i]));
    const double scale = emax > 0 ? sp->peak / emax : 1.0;
    const size_t R0 = (size_t)T * S;
    parallel_for((long long)T * C, [&](long long row) {
        const int t = (int)(row / C), c = (int)(row % C);
        const float* r = src + (size_t)row * S;
        int16_t* o = raw + (size_t)perm[c] * R0 + (size_t)t * S;
        for (int s = 0; s < S; ++s) {
            double v = std::nearbyint(r[s] * scale);
            v = std::min(32767.0, std::max(-32768.0, v));
            o[s] = (int16_t)v;
        }
    });
    return true;
}
