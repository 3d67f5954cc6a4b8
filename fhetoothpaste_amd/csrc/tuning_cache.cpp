// Tuning cache of tp_autotune_cached: a JSON file holding one entry per (device architecture, CU
// count, parameter set, launch size, engine opt-in) with the settings tp_autotune chose and the
// times it measured.  Written by this file only; read back with a small parser for exactly that
// shape (an object whose "entries" array holds flat objects of numbers, strings and one array of
// numbers).
#include "tuning_cache.hpp"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tp {
namespace {

struct Cursor {
    const std::string& s;
    size_t i = 0;
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    bool str(std::string& out) {
        ws();
        if (i >= s.size() || s[i] != '"') return false;
        const size_t j = s.find('"', i + 1);
        if (j == std::string::npos) return false;
        out = s.substr(i + 1, j - i - 1);
        i = j + 1;
        return true;
    }
    bool num(double& out) {
        ws();
        const char* p = s.c_str() + i;
        char* end = nullptr;
        out = std::strtod(p, &end);
        if (end == p) return false;
        i += (size_t)(end - p);
        return true;
    }
};

bool parse_entry(Cursor& c, TuneEntry& e) {
    if (!c.eat('{')) return false;
    if (c.eat('}')) return false;
    do {
        std::string key;
        if (!c.str(key) || !c.eat(':')) return false;
        if (key == "arch") {
            if (!c.str(e.arch)) return false;
        } else if (key == "ms") {
            if (!c.eat('[')) return false;
            e.ms.clear();
            if (!c.eat(']')) {
                do {
                    double v;
                    if (!c.num(v)) return false;
                    e.ms.push_back((float)v);
                } while (c.eat(','));
                if (!c.eat(']')) return false;
            }
        } else {
            double v;
            if (!c.num(v)) return false;
            const long long x = (long long)v;
            if (key == "cus") e.cus = (int)x;
            else if (key == "N") e.N = (uint32_t)x;
            else if (key == "n") e.n = (uint32_t)x;
            else if (key == "b") e.b = (uint32_t)x;
            else if (key == "l") e.l = (uint32_t)x;
            else if (key == "method") e.method = (uint32_t)x;
            else if (key == "count") e.count = (uint64_t)x;
            else if (key == "fft_opt_in") e.fft_opt_in = (int)x;
            else if (key == "waves") e.waves = (int)x;
            else if (key == "split_max") e.split_max = (uint64_t)x;
            else if (key == "engine") e.engine = (int)x;
            else if (key == "fft_waves") e.fft_waves = (int)x;
            else if (key == "grid") e.grid = (int)x;
            else if (key == "split_halves") e.split_halves = (int)x;
            // other keys are ignored
        }
    } while (c.eat(','));
    return c.eat('}');
}

}  // namespace

bool TuneEntry::same_key(const TuneEntry& o) const {
    return arch == o.arch && cus == o.cus && N == o.N && n == o.n && b == o.b && l == o.l && method == o.method &&
           count == o.count && fft_opt_in == o.fft_opt_in;
}

int read_tuning_cache(const char* path, std::vector<TuneEntry>& out, std::string& err) {
    out.clear();
    std::ifstream f(path);
    if (!f) return 0;  // no file yet: an empty cache
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    Cursor c{s};
    c.ws();
    if (c.i == s.size()) return 0;
    const size_t at = s.find("\"entries\"");
    if (at == std::string::npos) {
        err = "tuning cache has no \"entries\" array";
        return -1;
    }
    c.i = at + 9;
    if (!c.eat(':') || !c.eat('[')) {
        err = "tuning cache: malformed \"entries\"";
        return -1;
    }
    if (c.eat(']')) return 0;
    do {
        TuneEntry e;
        if (!parse_entry(c, e)) {
            err = "tuning cache: malformed entry";
            return -1;
        }
        out.push_back(e);
    } while (c.eat(','));
    if (!c.eat(']')) {
        err = "tuning cache: unterminated \"entries\"";
        return -1;
    }
    return 0;
}

int write_tuning_cache(const char* path, const std::vector<TuneEntry>& entries, std::string& err) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) {
        err = std::string("cannot write ") + tmp;
        return -1;
    }
    std::fprintf(f, "{\"format\": \"toothpaste_amd tuning cache\", \"version\": 1, \"entries\": [\n");
    for (size_t k = 0; k < entries.size(); ++k) {
        const TuneEntry& e = entries[k];
        std::fprintf(f,
                     "{\"arch\": \"%s\", \"cus\": %d, \"N\": %u, \"n\": %u, \"b\": %u, \"l\": %u, \"method\": %u, "
                     "\"count\": %llu, \"fft_opt_in\": %d, \"waves\": %d, \"split_max\": %llu, \"engine\": %d, "
                     "\"fft_waves\": %d, \"grid\": %d, \"split_halves\": %d, \"ms\": [",
                     e.arch.c_str(), e.cus, e.N, e.n, e.b, e.l, e.method, (unsigned long long)e.count, e.fft_opt_in,
                     e.waves, (unsigned long long)e.split_max, e.engine, e.fft_waves, e.grid, e.split_halves);
        for (size_t v = 0; v < e.ms.size(); ++v) std::fprintf(f, "%s%.4f", v ? ", " : "", (double)e.ms[v]);
        std::fprintf(f, "]}%s\n", k + 1 < entries.size() ? "," : "");
    }
    std::fprintf(f, "]}\n");
    if (std::fclose(f) != 0 || std::rename(tmp.c_str(), path) != 0) {
        err = std::string("cannot replace ") + path;
        return -1;
    }
    return 0;
}

}  // namespace tp
