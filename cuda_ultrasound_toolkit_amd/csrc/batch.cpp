// Batches of frames of any sizes in one call (extension; the reference
// beamforms one frame per cuda_toolkit::beamform call).
//
// Graph mode: every frame has its own RF state and buffers (set up on the
// host first); the decode and Hilbert launches of all frames and one
// persistent HERCULES DAS launch over the (frame, tile) work list of the whole
// batch are captured into a HIP graph, which is replayed while the batch's
// pointers, sizes and options stay the same.  The DAS launch shape (grid size
// and work order) is chosen by an autotuner the first time a batch shape is
// seen: every candidate computes the same tiles with the same code, so the
// choice changes only the time.  Sequential mode runs the plain single-frame
// path frame after frame on the same per-frame state, for comparison.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "../../include/cut_ext.h"
#include "pipeline.hpp"

namespace {

struct Slot {
    cut::RfPipeline rf;
    cut::BeamformBuffers bufs;
    cut::FramePlan plan;
    int nxt = 0, nyt = 0, nzt = 0;
};

// grid: 0 = one workgroup per work item, k > 0 = k resident workgroups per CU
// pulling items from the device counter; order: 0 heaviest tiles first across
// the batch, 1 frame after frame (deepest tiles first within a frame).
struct LaunchShape {
    int grid;
    int order;
};
constexpr LaunchShape CANDIDATES[] = {{0, 0}, {0, 1}, {2, 0}, {4, 0}, {8, 0}, {4, 1}};

struct BatchState {
    std::vector<std::unique_ptr<Slot>> slots;
    cut::DeviceBuffer frames, items, counters, times;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graph_sig;
    std::map<uint64_t, LaunchShape> tuned;
    int autotune = 1;
    LaunchShape last{4, 0};
    int cus = 0;
    double ticks_per_ms = 0.0;
};

BatchState& batch_state()
{
    static BatchState* s = new BatchState();
    return *s;
}

uint64_t mix(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

uint64_t hash_bytes(const void* p, size_t n)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

void drop_graph(BatchState& s)
{
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    s.exec = nullptr;
    s.graph = nullptr;
    s.graph_sig.clear();
}

// Work list in the given order.
std::vector<int2> make_items(const BatchState& s, int n, int order)
{
    std::vector<int2> items;
    std::vector<double> cost;
    for (int f = 0; f < n; ++f) {
        const Slot& sl = *s.slots[f];
        const cut::DasParams& p = sl.plan.p;
        const int per_z = sl.nxt * sl.nyt;
        for (int wz = 0; wz < sl.nzt; ++wz) {
            // tile rows counted from the deepest; the active aperture grows with depth
            const int bz = (sl.nzt - 1 - wz) * 8;
            const double z = p.min_z + (bz + 4) * p.res_z;
            const double c = (double)p.T * p.C * (1.0 + std::max(0.0, z) / (std::fabs(p.res_z) * 32.0 + 1e-30));
            for (int t = 0; t < per_z; ++t) {
                items.push_back(make_int2(f, wz * per_z + t));
                cost.push_back(c);
            }
        }
    }
    if (order == 0) {
        std::vector<size_t> idx(items.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
        std::vector<int2> out(items.size());
        for (size_t i = 0; i < idx.size(); ++i) out[i] = items[idx[i]];
        return out;
    }
    return items;
}

bool upload_items(BatchState& s, const std::vector<int2>& items)
{
    if (!s.items.ensure(std::max(items.size() * sizeof(int2), s.items.bytes))) return false;
    CUT_HIP_CHECK(hipMemcpy(s.items.ptr, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
    return true;
}

int grid_of(const BatchState& s, LaunchShape ls, int n_items)
{
    return ls.grid == 0 ? n_items : std::min(n_items, ls.grid * s.cus);
}

// decode + Hilbert of every frame, then the batch DAS (launches only)
bool enqueue_batch(BatchState& s, const CutFrame* frames, int n, int n_items, LaunchShape ls, bool focus,
                   bool sinapo, hipStream_t st)
{
    int* cnt = s.counters.as<int>();
    long long* tm = s.times.as<long long>();
    if (!cut::launch_batch_start(cnt, n + 1, tm, st)) return false;
    for (int f = 0; f < n; ++f) {
        Slot& sl = *s.slots[f];
        if (!cut::frame_to_iq(sl.rf, sl.bufs, frames[f].d_raw, *frames[f].bp, sl.plan, st)) return false;
    }
    const cut::FramePlan& p0 = s.slots[0]->plan;
    return cut::launch_das_batch(s.frames.as<cut::BatchFrameDev>(), s.items.as<int2>(), n_items,
                                 grid_of(s, ls, n_items), cnt, cnt + 1, tm + 1, p0.opts, focus, sinapo, p0.fmt, st);
}

}  // namespace

extern "C" bool cut_set_autotune(int on)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    batch_state().autotune = on ? 1 : 0;
    cut::set_das_autotune(on != 0);
    return true;
}

extern "C" bool cut_beamform_batch(const CutFrame* frames, int n, int mode, double* frame_ms)
{
    try {
        std::lock_guard<std::mutex> lock(cut::api_mutex());
        if (!frames || n <= 0) CUT_FAIL("batch: no frames");
        if (mode < 0 || mode > 2)
            CUT_FAIL("batch: mode must be 0 (graph), 1 (sequential) or 2 (graph, frames in order)");
        auto& s = batch_state();
        hipStream_t st = cut::stream();
        if (!s.cus) {
            int dev = 0, khz = 0;
            CUT_HIP_CHECK(hipGetDevice(&dev));
            CUT_HIP_CHECK(hipDeviceGetAttribute(&s.cus, hipDeviceAttributeMultiprocessorCount, dev));
            CUT_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            s.ticks_per_ms = khz;
        }
        while ((int)s.slots.size() < n) s.slots.emplace_back(new Slot());
        // host-side set-up of every frame (cached per slot by the RF state)
        for (int f = 0; f < n; ++f) {
            const CutFrame& fr = frames[f];
            if (!fr.d_raw || !fr.d_volume || !fr.bp) CUT_FAIL("batch: frame " << f << " has a null pointer");
            Slot& sl = *s.slots[f];
            if (!cut::prepare_frame(sl.rf, sl.bufs, *fr.bp, fr.raw_bytes, fr.volume_bytes, sl.plan)) {
                std::cerr << "batch: frame " << f << " rejected" << std::endl;
                return false;
            }
            cut::das_tile_grid(sl.plan.p, sl.nxt, sl.nyt, sl.nzt);
        }

        if (mode == 1) {
            std::vector<hipEvent_t> ev(n + 1);
            for (auto& e : ev) CUT_HIP_CHECK(hipEventCreate(&e));
            bool ok = hipEventRecord(ev[0], st) == hipSuccess;
            for (int f = 0; ok && f < n; ++f) {
                Slot& sl = *s.slots[f];
                ok = cut::beamform_device(sl.rf, sl.bufs, frames[f].d_raw, frames[f].raw_bytes,
                                          (float2*)frames[f].d_volume, frames[f].volume_bytes, *frames[f].bp, st) &&
                     hipEventRecord(ev[f + 1], st) == hipSuccess;
            }
            if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
            for (int f = 0; ok && frame_ms && f < n; ++f) {
                float ms = 0.0f;
                ok = hipEventElapsedTime(&ms, ev[0], ev[f + 1]) == hipSuccess;
                frame_ms[f] = ms;
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
            if (!ok) CUT_FAIL("batch: sequential run failed");
            return true;
        }

        // graph mode: HERCULES frames sharing one kernel variant
        const cut::FramePlan& p0 = s.slots[0]->plan;
        const bool focus = p0.p.focus_type == 4, sinapo = cut::herc_generic_apo(p0.p);
        for (int f = 0; f < n; ++f) {
            const cut::FramePlan& pl = s.slots[f]->plan;
            if (pl.p.sequence != HERCULES) CUT_FAIL("batch: graph mode runs HERCULES frames (frame " << f << ")");
            if ((pl.p.focus_type == 4) != focus || cut::herc_generic_apo(pl.p) != sinapo)
                CUT_FAIL("batch: the frames of a batch must share focus type and apodization kernel (sine or exact "
                         "Hann for volumes reaching z <= 0, or the Hann polynomial) (frame " << f << ")");
        }
        std::vector<cut::BatchFrameDev> fd(n);
        int n_items = 0;
        uint64_t shape_sig = mix(0, (uint64_t)n);
        std::vector<uint64_t> sig;
        for (int f = 0; f < n; ++f) {
            const Slot& sl = *s.slots[f];
            cut::BatchFrameDev& d = fd[f];
            d.p = sl.plan.p;
            d.iq = sl.bufs.iq.ptr;
            d.vol = (float2*)frames[f].d_volume;
            d.iq_inv = 1.0f / sl.plan.scale;
            d.nxt = sl.nxt;
            d.nyt = sl.nyt;
            d.nzt = sl.nzt;
            d.tiles = sl.nxt * sl.nyt * sl.nzt;
            n_items += d.tiles;
            shape_sig = mix(shape_sig, hash_bytes(&d.p, sizeof(d.p)));
            sig.push_back((uint64_t)(uintptr_t)frames[f].d_raw);
            sig.push_back((uint64_t)(uintptr_t)frames[f].d_volume);
            sig.push_back((uint64_t)(uintptr_t)sl.bufs.iq.ptr);
            sig.push_back((uint64_t)(uintptr_t)sl.bufs.dec.ptr);
            sig.push_back(hash_bytes(frames[f].bp, sizeof(CudaBeamformerParameters)));
        }
        shape_sig = mix(shape_sig, (uint64_t)p0.fmt * 131 + p0.opts.cubic * 7 + p0.opts.cf_hercules);
        // mode 2 keeps the frames in order (each frame's tiles before the next
        // frame's): early frames complete early; mode 0 takes the fastest order
        shape_sig = mix(shape_sig, (uint64_t)mode);
        sig.push_back(shape_sig);

        if (!s.frames.ensure(std::max(fd.size() * sizeof(cut::BatchFrameDev), s.frames.bytes)) ||
            !s.counters.ensure(std::max((size_t)(n + 1) * sizeof(int), s.counters.bytes)) ||
            !s.times.ensure(std::max((size_t)(n + 1) * sizeof(long long), s.times.bytes)))
            return false;

        // launch shape: tuned once per batch shape
        auto tuned = s.tuned.find(shape_sig);
        LaunchShape ls = tuned != s.tuned.end() ? tuned->second : LaunchShape{4, mode == 2 ? 1 : 0};
        if (s.autotune && tuned == s.tuned.end()) {
            drop_graph(s);
            CUT_HIP_CHECK(hipMemcpy(s.frames.ptr, fd.data(), fd.size() * sizeof(cut::BatchFrameDev),
                                    hipMemcpyHostToDevice));
            float best = 1e30f;
            hipEvent_t a, b;
            CUT_HIP_CHECK(hipEventCreate(&a));
            CUT_HIP_CHECK(hipEventCreate(&b));
            for (const LaunchShape& c : CANDIDATES) {
                if (mode == 2 && c.order != 1) continue;
                if (!upload_items(s, make_items(s, n, c.order))) return false;
                int* cnt = s.counters.as<int>();
                long long* tm = s.times.as<long long>();
                float ms = 0.0f;
                bool ok = cut::launch_batch_start(cnt, n + 1, tm, st);
                for (int f = 0; ok && f < n; ++f)
                    ok = cut::frame_to_iq(s.slots[f]->rf, s.slots[f]->bufs, frames[f].d_raw, *frames[f].bp,
                                          s.slots[f]->plan, st);
                ok = ok && hipEventRecord(a, st) == hipSuccess &&
                     cut::launch_das_batch(s.frames.as<cut::BatchFrameDev>(), s.items.as<int2>(), n_items,
                                           grid_of(s, c, n_items), cnt, cnt + 1, tm + 1, p0.opts, focus, sinapo,
                                           p0.fmt, st) &&
                     hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
                     hipEventElapsedTime(&ms, a, b) == hipSuccess;
                if (!ok) {
                    (void)hipEventDestroy(a);
                    (void)hipEventDestroy(b);
                    CUT_FAIL("batch: autotuning run failed");
                }
                if (ms < best) {
                    best = ms;
                    ls = c;
                }
            }
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            s.tuned[shape_sig] = ls;
        }
        s.last = ls;

        if (!s.exec || sig != s.graph_sig || s.graph_sig.empty()) {
            drop_graph(s);
            CUT_HIP_CHECK(hipMemcpy(s.frames.ptr, fd.data(), fd.size() * sizeof(cut::BatchFrameDev),
                                    hipMemcpyHostToDevice));
            if (!upload_items(s, make_items(s, n, ls.order))) return false;
            CUT_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const bool ok = enqueue_batch(s, frames, n, n_items, ls, focus, sinapo, st);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (!ok || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                CUT_FAIL("batch: graph capture failed");
            }
            s.graph = g;
            CUT_HIP_CHECK(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
            s.graph_sig = sig;
        }
        CUT_HIP_CHECK(hipGraphLaunch(s.exec, st));
        CUT_HIP_CHECK(hipStreamSynchronize(st));
        if (frame_ms) {
            std::vector<long long> t(n + 1);
            CUT_HIP_CHECK(hipMemcpy(t.data(), s.times.ptr, (n + 1) * sizeof(long long), hipMemcpyDeviceToHost));
            for (int f = 0; f < n; ++f) frame_ms[f] = (double)(t[f + 1] - t[0]) / s.ticks_per_ms;
        }
        return true;
    } catch (...) {
        std::cerr << "cuda_toolkit: cut_beamform_batch: exception" << std::endl;
        return false;
    }
}

extern "C" bool cut_batch_launch_shape(int* grid_per_cu, int* order)
{
    std::lock_guard<std::mutex> lock(cut::api_mutex());
    const auto& s = batch_state();
    if (grid_per_cu) *grid_per_cu = s.last.grid;
    if (order) *order = s.last.order;
    return true;
}
