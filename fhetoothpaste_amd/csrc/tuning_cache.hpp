// Tuning cache file of tp_autotune_cached (tuning_cache.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tp {

struct TuneEntry {
    // key
    std::string arch;
    int cus = 0;
    uint32_t N = 0, n = 0, b = 0, l = 0, method = 0;
    uint64_t count = 0;
    int fft_opt_in = 0;
    // settings tp_autotune chose and the times it measured (ms, -1 where not built)
    int waves = 2;
    uint64_t split_max = 0;
    int engine = 0, fft_waves = 1, grid = 0;
    int split_halves = 1;  // digit-parallel form: 1 = two waves per transform, 0 = one
    std::vector<float> ms;
    bool same_key(const TuneEntry& o) const;
};

// A missing file is an empty cache (returns 0); a malformed one returns -1 with err set.
int read_tuning_cache(const char* path, std::vector<TuneEntry>& out, std::string& err);
// Rewrites the whole file (through path.tmp and a rename).
int write_tuning_cache(const char* path, const std::vector<TuneEntry>& entries, std::string& err);

}  // namespace tp
