// Tuned launch-shape table: storage, the autotune switch and persistence
// (see tuning.hpp).
#include "tuning.hpp"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <utility>

#include "../../include/cut_ext.h"

namespace cut {
namespace {

// Timed tuning of the choices that change the rounding (DAS split, waves per
// voxel group, FFT tile) is off unless CUT_AUTOTUNE is set to a non-zero
// value or cut_set_autotune(1) is called: by default every process takes the
// same heuristic choices (or those of a loaded tuning file) and so computes
// the same bits.
bool autotune_from_env()
{
    const char* v = std::getenv("CUT_AUTOTUNE");
    return v && *v && std::strcmp(v, "0") != 0;
}

struct TuneState {
    std::mutex mu;
    std::map<std::pair<int, TuneKey>, int> table;
    bool autotune = autotune_from_env();
    bool file_read = false;
};

TuneState& tune_state()
{
    static TuneState* s = new TuneState();
    return *s;
}

const char* tuning_file() { return std::getenv("CUT_TUNING_FILE"); }

// Text form: one line per choice, "<table> <12 key values> <choice>".
std::string dump_locked(const TuneState& s)
{
    std::ostringstream o;
    for (const auto& [k, v] : s.table) {
        o << k.first;
        for (long long x : k.second) o << ' ' << x;
        o << ' ' << v << '\n';
    }
    return o.str();
}

bool load_locked(TuneState& s, const std::string& text)
{
    std::istringstream in(text);
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream ls(line);
        int table = -1, choice = 0;
        TuneKey key{};
        ls >> table;
        for (auto& x : key) ls >> x;
        ls >> choice;
        if (!ls || table < TUNE_DAS_SPLITS || table > TUNE_FFT_TILE || choice <= 0)
            CUT_FAIL("tuning table: malformed line " << n + 1 << ": '" << line << "'");
        s.table[{table, key}] = choice;
        ++n;
    }
    return true;
}

void read_file_locked(TuneState& s)
{
    if (s.file_read) return;
    s.file_read = true;
    const char* path = tuning_file();
    if (!path) return;
    std::ifstream f(path);
    if (!f) return;  // not written yet: this process tunes and writes it
    std::stringstream b;
    b << f.rdbuf();
    (void)load_locked(s, b.str());
}

void write_file_locked(const TuneState& s)
{
    const char* path = tuning_file();
    if (!path) return;
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) return;
        f << dump_locked(s);
    }
    (void)std::rename(tmp.c_str(), path);
}

}  // namespace

bool autotune_enabled()
{
    auto& s = tune_state();
    std::lock_guard<std::mutex> lock(s.mu);
    return s.autotune;
}

void set_das_autotune(bool on)
{
    auto& s = tune_state();
    std::lock_guard<std::mutex> lock(s.mu);
    s.autotune = on;
}

bool tune_lookup(int table, const TuneKey& key, int& choice)
{
    auto& s = tune_state();
    std::lock_guard<std::mutex> lock(s.mu);
    read_file_locked(s);
    auto it = s.table.find({table, key});
    if (it == s.table.end()) return false;
    choice = it->second;
    return true;
}

void tune_store(int table, const TuneKey& key, int choice)
{
    auto& s = tune_state();
    std::lock_guard<std::mutex> lock(s.mu);
    read_file_locked(s);
    s.table[{table, key}] = choice;
    write_file_locked(s);
}

std::string tuning_dump()
{
    auto& s = tune_state();
    std::lock_guard<std::mutex> lock(s.mu);
    read_file_locked(s);
    return dump_locked(s);
}

bool tuning_load(const std::string& text)
{
    auto& s = tune_state();
    std::lock_guard<std::mutex> lock(s.mu);
    read_file_locked(s);
    return load_locked(s, text);
}

}  // namespace cut

extern "C" bool cut_tuning_save(const char* path)
{
    if (!path) CUT_FAIL("cut_tuning_save: null path");
    std::ofstream f(path, std::ios::trunc);
    if (!f) CUT_FAIL("cut_tuning_save: cannot write " << path);
    f << cut::tuning_dump();
    return (bool)f;
}

extern "C" bool cut_tuning_load(const char* path)
{
    if (!path) CUT_FAIL("cut_tuning_load: null path");
    std::ifstream f(path);
    if (!f) CUT_FAIL("cut_tuning_load: cannot read " << path);
    std::stringstream b;
    b << f.rdbuf();
    return cut::tuning_load(b.str());
}
