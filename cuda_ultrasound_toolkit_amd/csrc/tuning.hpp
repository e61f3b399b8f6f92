// Launch-shape autotuning shared by the DAS and Hilbert launchers.
//
// A choice (HERCULES transmit split, waves per voxel group of the per-voxel
// DAS kernels, FFT tile of the Hilbert kernels) is looked up per shape key;
// the first time a key is seen, with autotuning on and the stream not
// capturing a graph, every candidate is timed and the fastest kept.  The
// candidates compute the same result with a different work split, so their
// outputs differ only in the rounding (~1e-7).  Choices are per process; a
// process that loads another's choices (cut_tuning_load, or the file named by
// CUT_TUNING_FILE, read on first use and rewritten after every new choice)
// launches the same shapes and so computes the same bits.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "common.hpp"

namespace cut {

using TuneKey = std::array<long long, 12>;
enum TuneTable { TUNE_DAS_SPLITS = 0, TUNE_VOX_WAVES = 1, TUNE_FFT_TILE = 2 };

bool autotune_enabled();
bool tune_lookup(int table, const TuneKey& key, int& choice);
void tune_store(int table, const TuneKey& key, int choice);
std::string tuning_dump();
bool tuning_load(const std::string& text);

// Runs run(choice) with the choice tuned for `key`, else `dflt`; when the key
// is new, autotuning is on, allow_tune is set and the stream is not
// capturing, runs dflt once untimed (first-launch costs), then every
// candidate twice, timed, keeps the fastest and runs it once more unless it
// was the last one timed, so that the output is the chosen shape's, the same
// bits as every later call.
template <class Run>
bool tuned_run(int table, const TuneKey& key, int dflt, const std::vector<int>& cands, hipStream_t st, bool allow_tune,
               Run&& run)
{
    int choice = dflt;
    const bool known = tune_lookup(table, key, choice);
    const bool tune = !known && allow_tune && autotune_enabled() && cands.size() > 1;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (!tune || hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return run(choice);
    if (!run(dflt)) return false;
    hipEvent_t a, b;
    CUT_HIP_CHECK(hipEventCreate(&a));
    CUT_HIP_CHECK(hipEventCreate(&b));
    float best = 1e30f;
    int best_c = dflt, last_c = dflt;
    for (int rep = 0; rep < 2; ++rep)
        for (int cand : cands) {
            float ms = 0.0f;
            const bool ok = hipEventRecord(a, st) == hipSuccess && run(cand) && hipEventRecord(b, st) == hipSuccess &&
                            hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess;
            if (!ok) {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
                CUT_FAIL("launch tuning failed");
            }
            if (ms < best) {
                best = ms;
                best_c = cand;
            }
            last_c = cand;
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    tune_store(table, key, best_c);
    return best_c == last_c || run(best_c);
}

}  // namespace cut
