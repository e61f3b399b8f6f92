// Host orchestration of the RF -> volume path.
//
// Replaces cuda_toolkit/src/rf_processing/rf_processor.cpp:21-192 (RfProcessor),
// hadamard_decoder.cpp:80-222 (matrix generation), hilbert_handler.cu:30-78
// (plans, filter spectrum) and beamformer.cu:17-197 (constants, dispatch).
// Differences from the reference, all on error paths: dimensions, buffer
// sizes, channel-map entries and filter length are validated before any
// kernel runs, and nothing synchronises the device inside the pipeline.
#include "pipeline.hpp"
#include "tuning.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

namespace cut {

static hipStream_t g_user_stream = nullptr;
static bool g_user_stream_set = false;

hipStream_t stream()
{
    if (g_user_stream_set) return g_user_stream;
    // A blocking stream: it is ordered with the legacy default stream, as the
    // reference's calls are, so device buffers a caller fills or reads there
    // (torch's default stream, for one) need no extra synchronisation around
    // the device-pointer entry points.  A non-blocking stream raced such
    // callers: a fill of the output issued before the call could land after
    // the library's kernels.
    static hipStream_t s = [] {
        hipStream_t h = nullptr;
        if (hipStreamCreateWithFlags(&h, hipStreamDefault) != hipSuccess) h = nullptr;
        return h;
    }();
    return s;
}

void set_stream(hipStream_t s)
{
    g_user_stream = s;
    g_user_stream_set = s != nullptr;
}

Options& options()
{
    static Options o;
    return o;
}

std::mutex& api_mutex()
{
    static std::mutex m;
    return m;
}

// ---------------------------------------------------------------------------
// Hadamard matrices
// ---------------------------------------------------------------------------
static int quadratic_character(int a, int q)
{
    a %= q;
    if (a < 0) a += q;
    if (a == 0) return 0;
    for (int x = 1; x < q; ++x)
        if ((x * x) % q == a) return 1;
    return -1;
}

// Order-12 and order-20 blocks of the Paley construction (quadratic character
// mod 11 / mod 19) in the arrangement of the reference's tables
// (hadamard_decoder.cpp:8-45): unit first row and column; order 12 circulant
// in (j - i), order 20 back-circulant in (i + j).
static int paley_entry(int base, int i, int j)
{
    if (i == 0 || j == 0) return 1;
    if (base == 12) return i == j ? -1 : -quadratic_character(j - i, 11);
    const int m = (i + j - 2) % 19;
    return m == 0 ? -1 : -quadratic_character(m, 19);
}

static bool pow2(uint n) { return n && !(n & (n - 1)); }

bool hadamard_matrix(int n, int ordering, std::vector<int>& H)
{
    if (n <= 0 || n > 1024) {
        std::cerr << "Maximum hadamard size (1024) exceeded. " << n << " requested." << std::endl;
        return false;
    }
    int base;
    if (pow2((uint)n)) base = 1;
    else if (n % 12 == 0 && pow2((uint)(n / 12))) base = 12;
    else if (n % 20 == 0 && pow2((uint)(n / 20))) base = 20;
    else {
        std::cerr << "Hadamard: Size '" << n << "' not supported (2^n, 12*2^n or 20*2^n)" << std::endl;
        return false;
    }
    H.assign((size_t)n * n, 0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const int b = base == 1 ? 1 : paley_entry(base, i % base, j % base);
            const int par = __builtin_popcount((unsigned)((i / base) & (j / base))) & 1;
            H[(size_t)i * n + j] = par ? -b : b;
        }
    if (ordering == WALSH) {
        std::vector<int> changes(n, 0), order(n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j + 1 < n; ++j) changes[i] += H[(size_t)i * n + j] != H[(size_t)i * n + j + 1];
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return changes[a] < changes[b]; });
        std::vector<int> sorted((size_t)n * n);
        for (int i = 0; i < n; ++i)
            std::copy_n(H.begin() + (size_t)order[i] * n, n, sorted.begin() + (size_t)i * n);
        H.swap(sorted);
    }
    return true;
}

// ---------------------------------------------------------------------------
// RfPipeline
// ---------------------------------------------------------------------------
bool RfPipeline::init(uint raw_len, uint raw_channels, uint S, uint C, uint T, int encoding)
{
    if (init_ && R0_ == raw_len && RC_ == raw_channels && S_ == S && C_ == C && T_ == T && enc_ == encoding)
        return true;
    const long long channel_count = (long long)C * T;
    if (S == 0 || channel_count <= 0 || S > 0x7fffffffu || channel_count > 0x7fffffffLL) {
        std::cerr << "Invalid dimensions for RF raw data: " << S << " samples, " << channel_count << " channels."
                  << std::endl;
        return false;
    }
    init_ = false;
    R0_ = raw_len; RC_ = raw_channels; S_ = S; C_ = C; T_ = T; enc_ = encoding;
    if (encoding != NONE) {
        std::vector<int> H;
        if (!hadamard_matrix((int)T, encoding, H)) {
            std::cerr << "Failed to generate Hadamard matrix." << std::endl;
            return false;
        }
        const int Kp = ((int)T + 31) / 32 * 32;
        std::vector<int8_t> h8((size_t)Kp * Kp, 0);
        std::vector<float> hf((size_t)Kp * Kp, 0.0f);
        std::vector<int> rs(Kp, 0);
        for (uint j = 0; j < T; ++j)
            for (uint k = 0; k < T; ++k) {
                const int v = H[(size_t)j * T + k];
                h8[(size_t)j * Kp + k] = (int8_t)v;
                hf[(size_t)j * Kp + k] = (float)v;
                rs[j] += v;
            }
        if (!h8_.ensure(h8.size()) || !hf_.ensure(hf.size() * 4) || !rowsum_.ensure(rs.size() * 4)) return false;
        CUT_HIP_CHECK(hipMemcpy(h8_.ptr, h8.data(), h8.size(), hipMemcpyHostToDevice));
        CUT_HIP_CHECK(hipMemcpy(hf_.ptr, hf.data(), hf.size() * 4, hipMemcpyHostToDevice));
        CUT_HIP_CHECK(hipMemcpy(rowsum_.ptr, rs.data(), rs.size() * 4, hipMemcpyHostToDevice));
        scale_ = 1.0f / (float)T;
    }
    // FFT plan: twiddles of the transform length (S for powers of two, the
    // convolution length M >= 2S - 1 otherwise, whose kernel spectrum is
    // built below and with every filter change); rows longer than 4096 also
    // take the 4096-point twiddles of their row steps (unsupported lengths
    // fail at hilbert())
    conv_H_.release(); tw2_.release(); long_ws_.release();
    if (hilbert_supported((int)S)) {
        if (hilbert_long((int)S) &&
            (!tw2_.ensure((size_t)4096 * sizeof(float2)) || !launch_twiddles(tw2_.as<float2>(), 4096, stream())))
            return false;
        const size_t N = (size_t)hilbert_fft_len((int)S);
        if (!tw_.ensure(N * sizeof(float2)) || !launch_twiddles(tw_.as<float2>(), (int)N, stream())) return false;
        if (hilbert_conv((int)S) && !conv_H_.ensure(N * (hilbert_long((int)S) ? sizeof(float2) : sizeof(float4))))
            return false;
    } else {
        tw_.release();
    }
    CUT_HIP_CHECK(hipStreamSynchronize(stream()));
    init_ = true;
    if (filter_on_) {
        // the spectrum of a loaded filter depends on the sample count: rebuild it
        std::vector<float> h;
        h.swap(filter_host_);
        filter_on_ = false;
        if (h.size() <= S) return set_match_filter(h.data(), h.size());
    }
    return build_conv_table();
}

// Kernel spectrum of the convolution form (lengths that are not powers of
// two) for the current filter.
bool RfPipeline::build_conv_table()
{
    if (!conv_H_.ptr) return true;
    if (!conv_tmp_.ensure(std::max(conv_tmp_.bytes, conv_tables_tmp_bytes((int)S_)))) return false;
    const int L = filter_on_ ? (int)filter_host_.size() : 0;
    if (!launch_conv_tables(filter_taps_.as<float>(), L, (int)S_, conv_tmp_.ptr, conv_H_.ptr, !hilbert_long((int)S_),
                            stream()))
        return false;
    CUT_HIP_CHECK(hipStreamSynchronize(stream()));
    return true;
}

bool RfPipeline::deinit()
{
    if (!init_) {
        std::cerr << "Session not initialized." << std::endl;
        return false;
    }
    init_ = false;
    h8_.release(); hf_.release(); rowsum_.release(); tw_.release();
    conv_H_.release(); conv_tmp_.release(); tw2_.release(); long_ws_.release();
    filter_spec_.release(); filter_taps_.release();
    filter_host_.clear();
    filter_on_ = false;
    R0_ = RC_ = S_ = C_ = T_ = 0;
    return true;
}

bool RfPipeline::convert_demod(const void* d_raw, int type, float2* d_out, double fd_fs, hipStream_t st)
{
    if (!init_) CUT_FAIL("Session not initialized.");
    if (!check_map(RC_)) return false;
    if ((size_t)R0_ < 2 * (size_t)T_ * S_)
        CUT_FAIL("rf_raw_dim[0] = " << R0_ << " is shorter than 2*transmits*samples for (I, Q) pairs.");
    const int L = filter_on_ ? (int)filter_host_.size() : 0;
    return launch_convert_demod(d_raw, type, map_.as<int16_t>(), R0_, (int)T_, (int)C_, (int)S_, fd_fs,
                                filter_taps_.as<float>(), L, d_out, st);
}

bool RfPipeline::set_channel_mapping(const int16_t* map, size_t count)
{
    if (!map || count == 0) CUT_FAIL("Failed to load channel mapping: empty map");
    if (map_host_.size() == count && std::equal(map_host_.begin(), map_host_.end(), map) && map_.ptr) return true;
    map_host_.assign(map, map + count);
    if (!map_.ensure(count * sizeof(int16_t))) return false;
    CUT_HIP_CHECK(hipMemcpy(map_.ptr, map, count * sizeof(int16_t), hipMemcpyHostToDevice));
    return true;
}

bool RfPipeline::check_map(uint raw_channels) const
{
    if (!map_.ptr) CUT_FAIL("Channel mapping not set.");
    if (map_host_.size() < C_) CUT_FAIL("Channel mapping has " << map_host_.size() << " entries, " << C_ << " needed.");
    for (uint c = 0; c < C_; ++c) {
        const int m = map_host_[c];
        if (m < 0 || (uint)m >= raw_channels)
            CUT_FAIL("Channel mapping entry " << c << " = " << m << " outside the " << raw_channels << " raw channels.");
    }
    return true;
}

bool RfPipeline::set_match_filter(const float* h, size_t length)
{
    if (!init_) CUT_FAIL("Session not initialized.");
    if (length == 0 || !h) {
        const bool was_on = filter_on_;
        filter_on_ = false;
        filter_host_.clear();
        return was_on ? build_conv_table() : true;
    }
    if (length > S_) CUT_FAIL("Failed to load match filter: length " << length << " exceeds " << S_ << " samples.");
    if (filter_on_ && filter_host_.size() == length && std::equal(filter_host_.begin(), filter_host_.end(), h) &&
        filter_spec_.bytes == (size_t)S_ * sizeof(float2))
        return true;
    filter_host_.assign(h, h + length);
    if (!filter_taps_.ensure(length * sizeof(float)) || !filter_spec_.ensure((size_t)S_ * sizeof(float2))) return false;
    CUT_HIP_CHECK(hipMemcpyAsync(filter_taps_.ptr, h, length * sizeof(float), hipMemcpyHostToDevice, stream()));
    if (!launch_filter_spectrum(filter_taps_.as<float>(), (int)length, (int)S_, filter_spec_.as<float2>(), stream()))
        return false;
    CUT_HIP_CHECK(hipStreamSynchronize(stream()));
    filter_on_ = true;
    return build_conv_table();
}

bool RfPipeline::convert_decode(const void* d_raw, int type, float* d_out, int out_stride, bool decode, hipStream_t st)
{
    if (!init_) CUT_FAIL("Session not initialized.");
    if (!check_map(RC_)) return false;
    if (type != TYPE_I16 && type != TYPE_F32) CUT_FAIL("Data converter: Unsupported input data type.");
    if ((size_t)R0_ < (size_t)T_ * S_) CUT_FAIL("Raw row length " << R0_ << " is shorter than transmits*samples.");
    if (decode && enc_ != NONE) {
        int mode = options().decode_mode;
        if (mode == 0) mode = (type == TYPE_I16) ? 2 : 1;
        if (type != TYPE_I16) mode = 1;
        return launch_decode(d_raw, type, map_.as<int16_t>(), R0_, (int)T_, (int)C_, (int)S_, h8_.as<int8_t>(),
                             hf_.as<float>(), rowsum_.as<int>(), scale_, d_out, out_stride, mode == 2 ? 1 : 0, st);
    }
    return launch_convert(d_raw, type, map_.as<int16_t>(), R0_, (int)T_, (int)C_, (int)S_, d_out, out_stride, st);
}

bool RfPipeline::hilbert(const HilbertInput& in, void* d_out, int fmt, float scale, hipStream_t st)
{
    if (!init_) CUT_FAIL("Session not initialized.");
    if (!hilbert_supported((int)S_)) CUT_FAIL("Hilbert: sample count " << S_ << " is not supported by the FFT.");
    const float2* F = filter_on_ ? filter_spec_.as<float2>() : nullptr;
    if (hilbert_long((int)S_)) {
        const int rows = (int)(C_ * T_);
        size_t ws_rows = 0;
        if (!hilbert_long_in_place((int)S_, fmt)) {
            ws_rows = hilbert_long_chunk_rows((int)S_, rows);
            if (!long_ws_.ensure(std::max(long_ws_.bytes, ws_rows * hilbert_long_n((int)S_) * sizeof(float2))))
                return false;
        }
        const LongHilbertTables t{tw_.as<float2>(), tw2_.as<float2>(), conv_H_.as<float2>()};
        return launch_hilbert_long(in, rows, (int)S_, t, F, d_out, fmt, scale, long_ws_.as<float2>(), ws_rows, st);
    }
    // FFT tile (values per thread of the LDS transforms): forced by the
    // options, else tuned per (length, rows, input, filter, format)
    const int rows = (int)(C_ * T_);
    const bool pow2 = !hilbert_conv((int)S_);  // powers of two 16..4096
    auto run = [&](int vpt) {
        return pow2 ? launch_hilbert(in, rows, (int)S_, tw_.as<float2>(), F, d_out, fmt, scale, vpt, st)
                    : launch_hilbert_conv(in, rows, (int)S_, tw_.as<float2>(), conv_H_.as<float4>(), d_out, fmt, scale,
                                          vpt, st);
    };
    if (options().fft_vpt) return run(options().fft_vpt);
    const TuneKey key{(long long)S_, rows, in.kind, in.stride, F != nullptr, fmt, 0, 0, 0, 0, 0, 0};
    return tuned_run(TUNE_FFT_TILE, key, 16, {16, 8}, st, true, run);
}

bool RfPipeline::convert_hilbert(const void* d_raw, int type, void* d_out, int fmt, float scale, hipStream_t st)
{
    if (!init_) CUT_FAIL("Session not initialized.");
    if (!check_map(RC_)) return false;
    if (type != TYPE_I16 && type != TYPE_F32) CUT_FAIL("Data converter: Unsupported input data type.");
    HilbertInput in{d_raw, type == TYPE_I16 ? 1 : 2, 1, map_.as<int16_t>(), (size_t)R0_, (int)C_};
    return hilbert(in, d_out, fmt, scale, st);
}

float RfPipeline::fp16_scale() const
{
    // |decoded| <= 32768; filtering multiplies the bound by sum |h|; the
    // analytic signal of a length-S row (w_k = 0.75/S, 1.5/S) stays below
    // 0.75 (1 + (2/pi) ln S + 1) times the filtered bound, under 8 for every
    // supported S (6.8 at 65536).  scale = 2^-k with 32768 * 8 * max(1, sum |h|) * scale <= 32768.
    double l1 = 0.0;
    if (filter_on_)
        for (float v : filter_host_) l1 += std::fabs((double)v);
    const double gain = 8.0 * std::max(1.0, l1);
    return (float)std::ldexp(1.0, -(int)std::ceil(std::log2(gain)));
}

// ---------------------------------------------------------------------------
// DAS constants (beamformer.cu:17-98), evaluated in float like the reference.
// ---------------------------------------------------------------------------
bool das_params_from_bp(const CudaBeamformerParameters& bp, DasParams& p)
{
    std::memset(&p, 0, sizeof(p));
    p.S = (int)bp.dec_data_dim[0];
    p.C = (int)bp.dec_data_dim[1];
    p.T = (int)bp.dec_data_dim[2];
    p.spm = bp.sampling_frequency / bp.speed_of_sound;
    const float px = bp.xdc_element_pitch[0], py = bp.xdc_element_pitch[1];
    p.pitch_x = px;
    p.pitch_y = py;
    p.xdc_min_x = -bp.xdc_transform[12] + px * 0.5f;
    p.xdc_min_y = -bp.xdc_transform[13] + py * 0.5f;
    p.xdc_max_x = bp.xdc_transform[12] - px * 0.5f;
    p.xdc_max_y = bp.xdc_transform[13] - py * 0.5f;
    p.delay = (int)(bp.time_offset * bp.sampling_frequency);
    p.sequence = bp.das_shader_id;
    p.Nx = (int)bp.output_points[0];
    p.Ny = (int)bp.output_points[1];
    p.Nz = (int)bp.output_points[2];
    p.min_x = bp.output_min_coordinate[0];
    p.min_y = bp.output_min_coordinate[1];
    p.min_z = bp.output_min_coordinate[2];
    p.res_x = (bp.output_max_coordinate[0] - bp.output_min_coordinate[0]) / (float)(bp.output_points[0] - 1);
    p.res_y = p.res_x;
    p.res_z = (bp.output_max_coordinate[2] - bp.output_min_coordinate[2]) / (float)(bp.output_points[2] - 1);
    p.fn_rx = bp.fn_rx;
    p.fn_tx = bp.fn_tx;
    p.readi_groups = (uint8_t)bp.readi_group_count;
    if (p.readi_groups == 0) p.readi_groups = 1;
    const int readi_id = (uint8_t)bp.readi_group_id;
    p.encoding = bp.decode;
    float fz = bp.foci[2];
    if (bp.tx_orientation == ORIENT_NONE) {
        p.focus_type = 0;
    } else if (std::isinf(fz) || bp.das_shader_id == RCA_TPW) {
        fz = 0.0f;
        p.focus_type = bp.tx_orientation == bp.rx_orientation ? 1 : 2;
    } else {
        p.focus_type = bp.tx_orientation == bp.rx_orientation ? 3 : 4;
    }
    p.focal_z = fz;
    p.cw = std::min(std::max(bp.coherency_weighting, 0.0f), 1.0f);
    p.apo_type = bp.apo_type;
    p.to_power = bp.to_power;
    p.fc_fs = (double)bp.center_frequency / (double)bp.sampling_frequency;
    if (p.sequence == UFORCES || p.sequence == RCA_TPW) {
        if (p.T > MAX_ACQ_COUNT) CUT_FAIL("Beamformer: at most " << MAX_ACQ_COUNT << " transmits.");
        for (int t = 0; t < p.T; ++t) {
            if (p.sequence == UFORCES) {
                p.tx_tab[t] = make_float2((float)(bp.sparse_elements[t] - bp.sparse_elements[0]), 0.0f);
            } else {
                p.tx_tab[t] = make_float2(bp.foci[3 * t + 2], 0.0f);
            }
        }
    }
    p.readi_bits = 0;
    if (p.readi_groups > 1) {
        if (p.encoding == NONE) CUT_FAIL("Beamformer: READI groups need an encoding matrix.");
        std::vector<int> H;
        if (!hadamard_matrix(p.readi_groups, p.encoding, H)) CUT_FAIL("Beamformer: Failed to generate Hadamard matrix.");
        if (readi_id >= p.readi_groups) CUT_FAIL("Beamformer: READI group id out of range.");
        for (int i = 0; i < p.readi_groups && i < 64; ++i)
            if (H[(size_t)readi_id * p.readi_groups + i] < 0) p.readi_bits |= 1ull << i;
    }
    return true;
}


// Taps (voxel, transmit, channel) whose apodization passes the sequence's
// threshold (apo > 0.1 for HERCULES, rx_apo > 0 for FORCES): the work the DAS
// kernels do per volume.  Host only; used to state achieved FLOP rates.
unsigned long long das_active_taps(const DasParams& p)
{
    const double A = 0.39758361765043326;  // acos(sqrt(0.1)) / pi
    const int G = p.readi_groups > 1 ? p.readi_groups : 1;
    auto hann = [](double lat, double z, double fn) {
        double a = std::min(fn * (lat / z) / 2.0, 0.5);
        double c = std::cos(M_PI * a);
        return c * c;
    };
    if (p.sequence == FORCES) {
        unsigned long long n = 0;
        for (int iz = 0; iz < p.Nz; ++iz)
            for (int ix = 0; ix < p.Nx; ++ix) {
                const double vx = (double)p.min_x + ix * (double)p.res_x, vz = (double)p.min_z + iz * (double)p.res_z;
                for (int c = 0; c < p.C; ++c) {
                    const double ec = (double)p.xdc_min_x + c * (double)p.pitch_x;
                    double rx_apo = (p.apo_type == RX_HANN || p.apo_type == TX_TO_SIN) ? hann(std::fabs(ec - vx), vz, p.fn_rx) : 1.0;
                    if (rx_apo > 0.0) n += (unsigned long long)G * p.T;
                }
            }
        return n * p.Ny;
    }
    if (p.sequence == UFORCES || p.sequence == RCA_TPW) {
        // receive element c at x_min + p/2 + c p; UFORCES evaluates every tap
        // of nonzero Hann weight over T - 1 transmits, RCA_TPW the channels of
        // weight >= 0.1 over all T angles
        const bool uf = p.sequence == UFORCES;
        unsigned long long n = 0;
        for (int iz = 0; iz < p.Nz; ++iz)
            for (int ix = 0; ix < p.Nx; ++ix) {
                const double vx = (double)p.min_x + ix * (double)p.res_x, vz = (double)p.min_z + iz * (double)p.res_z;
                for (int c = 0; c < p.C; ++c) {
                    const double rx = (double)p.xdc_min_x - vx + 0.5 * p.pitch_x + c * (double)p.pitch_x;
                    const double a = p.fn_rx * (std::fabs(rx) / vz) / 2.0;
                    if (uf ? a < 0.5 : hann(std::fabs(rx), vz, p.fn_rx) >= 0.1) ++n;
                }
            }
        return n * p.Ny * (unsigned long long)(uf ? std::max(0, p.T - 1) : p.T);
    }
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<unsigned long long> part(nt, 0);
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nt; ++w)
        th.emplace_back([&, w] {
            unsigned long long n = 0;
            for (int gt = (int)w; gt < G * p.T; gt += (int)nt) {
                const int g = gt / p.T, t = gt % p.T;
                double cur = t;
                if (G > 1) cur = p.encoding == WALSH ? (double)t * G + g : (double)g * p.T + t;
                const double ey = (double)p.xdc_min_y + cur * (double)p.pitch_y;
                for (int c = 0; c < p.C; ++c) {
                    const double ex = (double)p.xdc_min_x + c * (double)p.pitch_x;
                    if (p.apo_type == RX_TO_SIN) continue;  // counted below
                    for (int iy = 0; iy < p.Ny; ++iy)
                        for (int ix = 0; ix < p.Nx; ++ix) {
                            const double dx = ex - ((double)p.min_x + ix * (double)p.res_x);
                            const double dy = ey - ((double)p.min_y + iy * (double)p.res_y);
                            const double rho = std::sqrt(dx * dx + dy * dy);
                            for (int iz = 0; iz < p.Nz; ++iz) {
                                const double z = (double)p.min_z + iz * (double)p.res_z;
                                const double a = p.fn_rx * rho / (2.0 * z);
                                if (z > 0 && p.fn_rx >= 0.0f) {
                                    if (a < A) n += 1;  // cos^2 decreasing on [0, 1/2]
                                } else {
                                    // beamformer_utils.cuh:143-154 for any sign: cos^2(pi min(a, 1/2)),
                                    // NaN (0/0) mapped to 1/2 as by fminf
                                    const double c = std::cos(M_PI * std::fmin(a, 0.5));
                                    if (c * c > 0.1) n += 1;
                                }
                            }
                        }
                }
            }
            part[w] = n;
        });
    for (auto& t : th) t.join();
    unsigned long long n = 0;
    for (auto v : part) n += v;
    if (p.apo_type == RX_TO_SIN) {
        auto sinto = [&](double e, double peak) {
            double a = std::clamp(M_PI * e / peak, -M_PI, M_PI);
            double s = std::sin(a), o = 1.0;
            for (int i = 0; i < p.to_power; ++i) o *= s;
            return o;
        };
        unsigned long long k = 0;
        for (int gt = 0; gt < G * p.T; ++gt) {
            const int g = gt / p.T, t = gt % p.T;
            double cur = G > 1 ? (p.encoding == WALSH ? (double)t * G + g : (double)g * p.T + t) : t;
            const double ey = (double)p.xdc_min_y + cur * (double)p.pitch_y;
            for (int c = 0; c < p.C; ++c)
                if (sinto((double)p.xdc_min_x + c * (double)p.pitch_x, p.xdc_max_x) * sinto(ey, p.xdc_max_y) > 0.1) ++k;
        }
        n = k * (unsigned long long)p.Nx * p.Ny * p.Nz;
    }
    return n;
}

// ---------------------------------------------------------------------------
// beamform body (api/cuda_toolkit.cpp:109-199 on device pointers)
// ---------------------------------------------------------------------------
bool validate_beamform(const CudaBeamformerParameters& bp, size_t raw_bytes, size_t vol_bytes)
{
    const uint R0 = bp.rf_raw_dim[0], RC = bp.rf_raw_dim[1];
    const uint S = bp.dec_data_dim[0], C = bp.dec_data_dim[1], T = bp.dec_data_dim[2];
    const bool bandpass = bp.sample_type == SAMPLE_BANDPASS;
    if (bp.sample_type != SAMPLE_NORMAL && !bandpass) CUT_FAIL("Unsupported sample type: " << bp.sample_type);
    if (bp.data_type != TYPE_I16 && bp.data_type != TYPE_F32) CUT_FAIL("Data converter: Unsupported input data type.");
    if (S == 0 || C == 0 || T == 0) CUT_FAIL("Invalid dimensions for RF raw data: " << S << " samples, " << C * T << " channels.");
    if (C > MAX_CHANNEL_COUNT) CUT_FAIL("Channel count " << C << " exceeds " << MAX_CHANNEL_COUNT << ".");
    if (bandpass) {
        // (I, Q) pairs, demodulated and FIR-filtered; no Hilbert transform, no decode
        if (bp.decode != NONE) CUT_FAIL("Bandpass input takes no transmit decoding (decode must be NONE).");
        if ((size_t)R0 < 2 * (size_t)T * S)
            CUT_FAIL("rf_raw_dim[0] = " << R0 << " is shorter than 2*transmits*samples for (I, Q) pairs.");
        if (!(bp.sampling_frequency > 0.0f)) CUT_FAIL("Bandpass input needs a positive sampling frequency.");
        if (options().iq_format != IQ_F32) CUT_FAIL("Staged IQ formats are not available for bandpass input.");
    } else {
        if (!hilbert_supported((int)S)) CUT_FAIL("Sample count " << S << " is not supported by the FFT.");
        if ((size_t)R0 < (size_t)T * S) CUT_FAIL("rf_raw_dim[0] = " << R0 << " is shorter than transmits*samples.");
    }
    const size_t elem = bp.data_type == TYPE_I16 ? 2 : 4;
    if (raw_bytes < (size_t)R0 * RC * elem)
        CUT_FAIL("Input holds " << raw_bytes << " bytes, rf_raw_dim needs " << (size_t)R0 * RC * elem << ".");
    if (bp.decode != NONE && bp.decode != HADAMARD && bp.decode != WALSH) CUT_FAIL("Unknown encoding " << bp.decode);
    if (bp.decode != NONE) {
        std::vector<int> H;
        if (!hadamard_matrix((int)T, bp.decode, H)) CUT_FAIL("Failed to generate Hadamard matrix.");
    }
    for (uint c = 0; c < C; ++c)
        if (bp.channel_mapping[c] < 0 || (uint)bp.channel_mapping[c] >= RC)
            CUT_FAIL("Channel mapping entry " << c << " = " << bp.channel_mapping[c] << " outside the " << RC << " raw channels.");
    if (bp.filter_length > MAX_FILTER_LENGTH || bp.filter_length > S)
        CUT_FAIL("Filter length " << bp.filter_length << " exceeds " << std::min<uint>(S, MAX_FILTER_LENGTH) << ".");
    const size_t nvox = (size_t)bp.output_points[0] * bp.output_points[1] * bp.output_points[2];
    if (nvox == 0) CUT_FAIL("Empty output grid.");
    if (vol_bytes < nvox * sizeof(float2))
        CUT_FAIL("Output holds " << vol_bytes << " bytes, the volume needs " << nvox * sizeof(float2) << ".");
    DasParams p;
    if (!das_params_from_bp(bp, p)) return false;
    if (p.sequence == HERCULES) {
        if (p.focus_type != 2 && p.focus_type != 4) CUT_FAIL("Invalid focus for HERCULES.");
    } else if (p.sequence == FORCES) {
        if (p.readi_groups > 1 && p.encoding != HADAMARD) CUT_FAIL("FORCES only supports Hadamard encoding at this time.");
    } else if (p.sequence == UFORCES) {
        if (bp.sparse_elements[0] == -1) CUT_FAIL("Beamformer: UFORCES requires sparse elements to be defined.");
        if (p.focus_type == 0) CUT_FAIL("Beamformer: UFORCES needs a transmit orientation.");
    } else if (p.sequence == RCA_TPW) {
        if (p.focus_type != 1 && p.focus_type != 2) CUT_FAIL("TPW beamforming requires a plane focus.");
    } else {
        CUT_FAIL("Beamformer: Unsupported sequence ID " << p.sequence);
    }
    return true;
}

bool prepare_frame(RfPipeline& rf, BeamformBuffers& bufs, const CudaBeamformerParameters& bp, size_t raw_bytes,
                   size_t vol_bytes, FramePlan& plan)
{
    if (!validate_beamform(bp, raw_bytes, vol_bytes)) return false;
    const uint R0 = bp.rf_raw_dim[0], RC = bp.rf_raw_dim[1];
    const uint S = bp.dec_data_dim[0], C = bp.dec_data_dim[1], T = bp.dec_data_dim[2];

    if (!rf.init(R0, RC, S, C, T, bp.decode)) CUT_FAIL("Failed to initialize RF processor.");
    if (!rf.set_channel_mapping(bp.channel_mapping, C)) CUT_FAIL("Failed to set channel mapping.");
    if (!rf.set_match_filter(bp.rf_filter, bp.filter_length)) CUT_FAIL("Failed to set match filter.");
    if (!das_params_from_bp(bp, plan.p)) return false;

    plan.fmt = options().iq_format;
    if (plan.fmt != IQ_F32 && plan.p.sequence != HERCULES) CUT_FAIL("Staged IQ formats are supported for HERCULES only.");
    if (plan.fmt == IQ_F16 && bp.data_type != TYPE_I16) CUT_FAIL("FP16-staged IQ needs int16 input data.");
    plan.scale = plan.fmt == IQ_F16 ? rf.fp16_scale() : 1.0f;
    plan.opts.cubic = options().das_cubic;
    plan.opts.cf_hercules = options().das_cf_hercules;
    const size_t count = (size_t)S * C * T;
    if (!bufs.iq.ensure(count * iq_sample_bytes(plan.fmt))) return false;
    if (bp.decode != NONE && !bufs.dec.ensure(count * sizeof(float))) return false;
    return true;
}

bool frame_to_iq(RfPipeline& rf, BeamformBuffers& bufs, const void* d_raw, const CudaBeamformerParameters& bp,
                 const FramePlan& plan, hipStream_t st)
{
    const uint C = bp.dec_data_dim[1];
    void* iq = bufs.iq.ptr;
    if (bp.sample_type == SAMPLE_BANDPASS) {
        const double fd_fs = (double)bp.center_frequency / (double)bp.sampling_frequency;
        if (!rf.convert_demod(d_raw, bp.data_type, static_cast<float2*>(iq), fd_fs, st))
            CUT_FAIL("Failed to convert and demodulate data.");
    } else if (bp.decode != NONE) {
        if (!rf.convert_decode(d_raw, bp.data_type, bufs.dec.as<float>(), 1, true, st)) CUT_FAIL("Failed to decode RF data.");
        HilbertInput in{bufs.dec.ptr, 0, 1, nullptr, 0, (int)C};
        if (!rf.hilbert(in, iq, plan.fmt, plan.scale, st)) CUT_FAIL("Failed to apply Hilbert transform.");
    } else {
        if (!rf.convert_hilbert(d_raw, bp.data_type, iq, plan.fmt, plan.scale, st))
            CUT_FAIL("Failed to apply Hilbert transform.");
    }
    return true;
}

bool beamform_device(RfPipeline& rf, BeamformBuffers& bufs, const void* d_raw, size_t raw_bytes, float2* d_vol,
                     size_t vol_bytes, const CudaBeamformerParameters& bp, hipStream_t st)
{
    FramePlan plan;
    if (!prepare_frame(rf, bufs, bp, raw_bytes, vol_bytes, plan)) return false;
    if (!frame_to_iq(rf, bufs, d_raw, bp, plan, st)) return false;
    const size_t ws = das_workspace_bytes(plan.p);
    if (ws && !bufs.das_ws.ensure(std::max(ws, bufs.das_ws.bytes))) return false;
    if (!launch_das(bufs.iq.ptr, plan.fmt, plan.scale, plan.p, plan.opts, d_vol, bufs.das_ws.ptr, bufs.das_ws.bytes, st))
        CUT_FAIL("Beamforming failed.");
    return true;
}

}  // namespace cut
