// Host check of the instruction-level Goldilocks primitives (fhetoothpaste_amd/csrc/gl_valu.hpp):
// the host build runs the same op sequences as the device build, lane by lane.  Every primitive is
// compared with 128-bit integer arithmetic on edge values (chosen to reach each carry/borrow case
// its proof names) and on random values.  Prints "ok <checks>" and exits 0, or the first mismatch.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include "../../fhetoothpaste_amd/csrc/gl_valu.hpp"

typedef unsigned __int128 u128;
static const uint64_t Q = glv::kQ;
static uint64_t g_state = 0x9E3779B97F4A7C15ULL;
static uint64_t rnd() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static uint64_t modq(u128 x) { return (uint64_t)(x % Q); }
static uint64_t mulq(uint64_t a, uint64_t b) { return modq((u128)a * b); }
static uint64_t powq(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= Q;
    while (e) {
        if (e & 1) r = mulq(r, b);
        b = mulq(b, b);
        e >>= 1;
    }
    return r;
}
static long g_checks = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        ++g_checks;                                        \
        if (!(cond)) {                                     \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                  \
            fprintf(stderr, "\n");                         \
            exit(1);                                       \
        }                                                  \
    } while (0)

static std::vector<uint64_t> edge_values() {
    std::vector<uint64_t> v = {0, 1, 2, 3, 0xFFFFFFFFULL, 0x100000000ULL, 0x100000001ULL, 0x1FFFFFFFFULL,
                               0x7FFFFFFFFFFFFFFFULL, 0x8000000000000000ULL, Q - 2, Q - 1, Q, Q + 1, Q + 2,
                               0xFFFFFFFF00000000ULL, 0xFFFFFFFFFFFFFFFEULL, 0xFFFFFFFFFFFFFFFFULL,
                               0xFFFFFFFE00000000ULL, 0xFFFFFFFE00000001ULL, 0x00000000FFFFFFFEULL,
                               0xFFFFFFFF7FFFFFFFULL, 0x80000000FFFFFFFFULL};
    return v;
}

int main(int argc, char** argv) {
    const long n_rand = argc > 1 ? atol(argv[1]) : 2000000;
    const uint64_t qm1 = glv::qm1_const();
    const std::vector<uint64_t> E = edge_values();
    auto rnd_any = [&](int i) -> uint64_t {
        // a mix of uniform words, words near Q / 2^64 and words with zero or full halves
        switch (i & 7) {
            case 0: return rnd();
            case 1: return Q + (rnd() % (0xFFFFFFFFULL));
            case 2: return Q - 1 - (rnd() & 0xFFFF);
            case 3: return rnd() & 0xFFFFFFFFULL;
            case 4: return rnd() | 0xFFFFFFFF00000000ULL;
            case 5: return E[rnd() % E.size()];
            case 6: return rnd() >> (rnd() % 64);
            default: return ~(rnd() >> (rnd() % 64));
        }
    };

    // mul128: exact product
    auto check_mul128 = [&](uint64_t a, uint64_t b) {
        uint64_t lo, hi;
        glv::mul128(a, b, lo, hi);
        const u128 p = (u128)a * b;
        CHECK(lo == (uint64_t)p && hi == (uint64_t)(p >> 64), "mul128 %llx %llx", (unsigned long long)a,
              (unsigned long long)b);
    };
    // reduce_canon: any 128-bit value -> canonical residue
    auto check_reduce = [&](uint64_t lo, uint64_t hi) {
        const uint64_t r = glv::reduce_canon(lo, hi, qm1);
        const u128 x = ((u128)hi << 64) | lo;
        CHECK(r < Q && r == modq(x), "reduce_canon lo=%llx hi=%llx -> %llx", (unsigned long long)lo,
              (unsigned long long)hi, (unsigned long long)r);
    };
    auto check_pairs = [&](uint64_t u, uint64_t t) {
        check_mul128(u, t);
        check_reduce(u, t);
        const uint64_t m = glv::mul_canon(u, t, qm1);
        CHECK(m < Q && m == mulq(u, t), "mul_canon");
        const uint64_t tc = t % Q;
        const uint64_t a = glv::add_lc(u, tc);
        CHECK(a % Q == modq((u128)u + tc), "add_lc %llx %llx", (unsigned long long)u, (unsigned long long)tc);
        const uint64_t s = glv::sub_lc(u, tc);
        CHECK(s % Q == modq((u128)u + Q - tc), "sub_lc %llx %llx", (unsigned long long)u, (unsigned long long)tc);
        const uint64_t uc = u % Q;
        const uint64_t ac = glv::add_cc(uc, tc, qm1);
        CHECK(ac < Q && ac == modq((u128)uc + tc), "add_cc");
        const uint64_t sc = glv::sub_cc(uc, tc);
        CHECK(sc < Q && sc == modq((u128)uc + Q - tc), "sub_cc %llx %llx", (unsigned long long)uc,
              (unsigned long long)tc);
    };

    // the four (carry, borrow) cases of reduce_canon / fix_canon, and the r >= Q case
    check_reduce(~0ULL, 1ULL | (0xFFFFFFFFULL << 32));     // c = 1, b = 1
    check_reduce(0, 5ULL << 32);                           // c = 0, b = 1
    check_reduce(~0ULL, 0x00000000FFFFFFFFULL);            // c = 1, b = 0
    check_reduce(Q + 5, 0);                                // c = 0, b = 0, r >= Q
    check_reduce(~0ULL, ~0ULL);
    for (uint64_t a : E)
        for (uint64_t b : E) {
            check_pairs(a, b);
            for (int k = 0; k < 8; ++k) check_reduce(a ^ (rnd() & 0xFFFF), b ^ (rnd() & 0xFFFF));
        }
    for (long i = 0; i < n_rand; ++i) check_pairs(rnd_any((int)i), rnd_any((int)(i >> 3)));

    // products by 2^e for every e in [0, 192)
    for (int e = 0; e < 192; ++e) {
        const uint64_t p2 = powq(2, e);
        auto one = [&](uint64_t x) {
            bool neg;
            const uint64_t r = glv::mul_pow2(x, e, neg, qm1);
            const uint64_t want = mulq(x % Q, p2);
            const uint64_t got = neg ? (Q - r % Q) % Q : r % Q;
            CHECK(got == want, "mul_pow2 e=%d x=%llx", e, (unsigned long long)x);
            if (e % 96 != 0) CHECK(r < Q, "mul_pow2 not canonical e=%d x=%llx", e, (unsigned long long)x);
        };
        for (uint64_t x : E) one(x);
        for (long i = 0; i < n_rand / 64; ++i) one(rnd_any((int)i));
    }

    // multiply-accumulate over T terms
    for (int T = 1; T <= 16; ++T) {
        for (long it = 0; it < n_rand / 64 + 64; ++it) {
            glv::Acc acc{glv::acc_init_word((uint32_t)T), (uint32_t)T};
            uint64_t want = 0;
            const bool extreme = it < 64;
            for (int j = 0; j < T; ++j) {
                const uint64_t a = extreme ? (it & 1 ? ~0ULL : E[(it + j) % E.size()]) : rnd_any((int)(it + j));
                const uint64_t b = extreme ? (it & 2 ? ~0ULL : E[(it * 7 + j) % E.size()]) : rnd_any((int)(it * 3 + j));
                glv::acc_mac(acc, a, b);
                want = modq((u128)want + mulq(a, b));
                CHECK(acc.C <= (uint32_t)(3 * T), "acc count out of range T=%d C=%u", T, acc.C);
            }
            const uint64_t got = glv::acc_value(acc, qm1);
            CHECK(got < Q && got == want, "acc T=%d it=%ld", T, it);
        }
    }
    // two-word lazy multiply-accumulate chains
    for (long it = 0; it < n_rand / 8 + 256; ++it) {
        const int T = 1 + (int)(it % 12);
        const bool extreme = it < 256;
        uint64_t acc = extreme ? E[it % E.size()] : rnd_any((int)it);
        uint64_t want = acc % Q;
        for (int j = 0; j < T; ++j) {
            const uint64_t a = extreme ? (it & 1 ? ~0ULL : E[(it + j) % E.size()]) : rnd_any((int)(it + 5 * j));
            const uint64_t b = extreme ? (it & 2 ? ~0ULL : E[(it * 5 + j) % E.size()]) : rnd_any((int)(it * 7 + j));
            acc = glv::mac_lazy(acc, a, b);
            want = modq((u128)want + mulq(a, b));
            CHECK(acc % Q == want, "mac_lazy it=%ld j=%d", it, j);
        }
    }
    // group forms with mixed members (M = 4, different exponents per member)
    for (long it = 0; it < n_rand / 16 + 64; ++it) {
        uint64_t x[4], y[4], r[4], lo[4], hi[4];
        for (int m = 0; m < 4; ++m) { x[m] = rnd_any((int)(it * 4 + m)); y[m] = rnd_any((int)(it * 3 + m)); }
        glv::mul_canon_g<4>(x, y, r, qm1);
        for (int m = 0; m < 4; ++m) CHECK(r[m] == mulq(x[m], y[m]), "mul_canon_g");
        glv::mul128_g<4>(x, y, lo, hi);
        for (int m = 0; m < 4; ++m) CHECK(lo[m] == (uint64_t)((u128)x[m] * y[m]) && hi[m] == (uint64_t)(((u128)x[m] * y[m]) >> 64), "mul128_g");
        const int elo[4] = {3, 17, 30, 1}, emid[4] = {33, 48, 63, 42}, sneg[4] = {1, 12, 24, 32};
        glv::mul_2e_lo_g<4>(x, elo, r, qm1);
        for (int m = 0; m < 4; ++m) CHECK(r[m] == mulq(x[m] % Q, powq(2, elo[m])), "mul_2e_lo_g");
        glv::mul_2e_mid_g<4>(x, emid, r, qm1);
        for (int m = 0; m < 4; ++m) CHECK(r[m] == mulq(x[m] % Q, powq(2, emid[m])), "mul_2e_mid_g");
        glv::mul_2e_neg_g<4>(x, sneg, r);
        for (int m = 0; m < 4; ++m) CHECK(r[m] == mulq(x[m] % Q, powq(2, 192 - sneg[m])), "mul_2e_neg_g");
        uint64_t acc[4];
        for (int m = 0; m < 4; ++m) acc[m] = rnd_any((int)(it + m));
        uint64_t want[4];
        for (int m = 0; m < 4; ++m) want[m] = modq((u128)(acc[m] % Q) + mulq(x[m], y[m]));
        glv::mac_lazy_g<4>(acc, x, y);
        for (int m = 0; m < 4; ++m) CHECK(acc[m] % Q == want[m], "mac_lazy_g");
    }
    printf("ok %ld\n", g_checks);
    return 0;
}
