// Host check of the integer modulus-switch emulation (fhetoothpaste_amd/csrc/modswitch.hpp)
// against x87 long double, which is the reference's arithmetic (modulus_switching.cpp:36-93) on
// x86-64.  Random inputs plus near-tie inputs v ~ (2k+1) sM / (2 tM) and near-integer inputs
// v ~ k sM / tM, for several target moduli and all three rounding kinds.  Prints the number of
// values checked; exits 1 on the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "../../fhetoothpaste_amd/csrc/modswitch.hpp"

static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static uint64_t ld_switch(uint64_t v, uint64_t sM, uint64_t tM, int kind) {
    const long double x = ((long double)tM * (long double)v) / (long double)sM;
    uint64_t r = kind == 0 ? (uint64_t)roundl(x) : kind == 1 ? (uint64_t)floorl(x) : (uint64_t)ceill(x);
    if (r >= tM) r -= tM;
    return r;
}

int main() {
    const uint64_t moduli[] = {0xFFFFFFFF00000001ULL, 0xFFFFFFFFFFFFFFC5ULL, 0x8000000000000001ULL};
    const int ts[] = {1, 11, 12, 13, 20, 32, 62};
    uint64_t st = 7, checked = 0;
    for (uint64_t sM : moduli)
        for (int t : ts) {
            const uint64_t mu = tp::ms_reciprocal(sM);
            const uint64_t tM = 1ULL << t;
            for (int i = 0; i < 40000; ++i) {
                uint64_t v;
                const int mode = i % 4;
                const unsigned __int128 k = splitmix(st) % tM;
                if (mode == 0) v = splitmix(st) % sM;
                else if (mode == 1) {  // near a half-integer quotient
                    const unsigned __int128 num = (unsigned __int128)(2 * k + 1) * sM;
                    v = (uint64_t)(num / (2 * (unsigned __int128)tM)) + (splitmix(st) % 5) - 2;
                } else if (mode == 2) {  // near an integer quotient
                    v = (uint64_t)(((unsigned __int128)k * sM) / tM) + (splitmix(st) % 5) - 2;
                } else {
                    v = splitmix(st) >> (splitmix(st) % 64);  // small magnitudes
                }
                if (v >= sM) v %= sM;
                for (int kind = 0; kind < 3; ++kind) {
                    const uint64_t a = tp::ms_switch(v, sM, mu, t, kind), b = ld_switch(v, sM, tM, kind);
                    ++checked;
                    if (a != b) {
                        std::printf("mismatch sM=%llx t=%d kind=%d v=%llu: emulated %llu, long double %llu\n",
                                    (unsigned long long)sM, t, kind, (unsigned long long)v, (unsigned long long)a,
                                    (unsigned long long)b);
                        return 1;
                    }
                }
            }
        }
    std::printf("checked %llu\n", (unsigned long long)checked);
    return 0;
}
