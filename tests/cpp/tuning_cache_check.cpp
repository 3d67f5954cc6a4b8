// Round trip of the tuning cache file (fhetoothpaste_amd/csrc/tuning_cache.cpp): entries written
// and read back field for field, an entry without the later "split_halves" key read with its
// default, a missing file as an empty cache, malformed files rejected.  Prints "ok" and exits 0.
#include <cstdio>
#include <string>
#include <vector>

#include "tuning_cache.hpp"

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("check failed line %d: %s\n", __LINE__, #c);    \
            ++fails;                                                    \
        }                                                               \
    } while (0)

static void write_text(const std::string& path, const char* text) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    std::fputs(text, f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    std::string err;
    std::vector<tp::TuneEntry> in, out;
    // a missing file is an empty cache
    CHECK(tp::read_tuning_cache((dir + "/absent.json").c_str(), out, err) == 0 && out.empty());
    for (int k = 0; k < 3; ++k) {
        tp::TuneEntry e;
        e.arch = k == 2 ? "gfx950:sramecc+:xnack-" : "gfx950";
        e.cus = 256 - k;
        e.N = 1024u << (k & 1);
        e.n = 630 + k;
        e.b = 12;
        e.l = 3;
        e.method = (uint32_t)k;
        e.count = 4096ull + (uint64_t)k * 1000000000ull;
        e.fft_opt_in = k & 1;
        e.waves = 1 + (k & 1);
        e.split_max = 512 + (uint64_t)k;
        e.engine = k & 1;
        e.fft_waves = 2 - (k & 1);
        e.grid = 2048 * k;
        e.split_halves = k == 1 ? 0 : 1;
        e.ms = {81.25f, -1.0f, 128.5f + (float)k, -1.0f, -1.0f, 80.75f};
        in.push_back(e);
    }
    const std::string path = dir + "/tuning.json";
    CHECK(tp::write_tuning_cache(path.c_str(), in, err) == 0);
    CHECK(tp::read_tuning_cache(path.c_str(), out, err) == 0);
    CHECK(out.size() == in.size());
    for (size_t k = 0; k < in.size() && k < out.size(); ++k) {
        const tp::TuneEntry &a = in[k], &b = out[k];
        CHECK(a.same_key(b) && b.same_key(a));
        CHECK(a.waves == b.waves && a.split_max == b.split_max && a.engine == b.engine);
        CHECK(a.fft_waves == b.fft_waves && a.grid == b.grid && a.split_halves == b.split_halves);
        CHECK(a.ms == b.ms);
    }
    // the key: a different count, CU count or architecture is another entry
    tp::TuneEntry other = in[0];
    other.count += 1;
    CHECK(!other.same_key(in[0]));
    other = in[0];
    other.arch = "gfx942";
    CHECK(!other.same_key(in[0]));
    // an entry written before "split_halves" existed reads with the default (1)
    write_text(dir + "/old.json",
               "{\"entries\": [{\"arch\": \"gfx950\", \"cus\": 256, \"N\": 1024, \"n\": 630, \"b\": 12, \"l\": 3, "
               "\"method\": 0, \"count\": 64, \"fft_opt_in\": 0, \"waves\": 2, \"split_max\": 512, \"engine\": 0, "
               "\"fft_waves\": 1, \"grid\": 0, \"ms\": [1.5, 2.5]}]}\n");
    CHECK(tp::read_tuning_cache((dir + "/old.json").c_str(), out, err) == 0);
    CHECK(out.size() == 1 && out[0].split_halves == 1 && out[0].count == 64 && out[0].ms.size() == 2);
    // malformed files are rejected with a message
    const char* bad[] = {"{\"entries\": [{\"arch\": \"gfx950\", \"N\": }]}", "{\"entries\": [", "not json",
                         "{\"entries\": [{}]}"};
    for (const char* text : bad) {
        write_text(dir + "/bad.json", text);
        err.clear();
        CHECK(tp::read_tuning_cache((dir + "/bad.json").c_str(), out, err) != 0 && !err.empty());
    }
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
