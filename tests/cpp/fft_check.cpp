// Host check of the exact floating-point external product (fhetoothpaste_amd/csrc/fft_key.cpp):
//  * HostNtt::inverse undoes HostNtt::forward;
//  * the transform-domain key of device_key(), multiplied by the folded transform of digit
//    polynomials in double precision (the device's factorisation and position order) and brought
//    back by the inverse transform, gives every limb convolution sum_r d_r * K_r^(k) within the
//    rounding bound of its exact integer value (computed here by schoolbook negacyclic
//    convolution over int64), and the limbs recombine to the exact product mod Q (computed with
//    the NTT);
//  * the adversarial case (every digit at its maximum, every key limb at -2^15 or 2^15-ish) stays
//    within the bound as well;
//  * the device's recombination (fft_limbs.hpp, host build of its op sequence) of rounded limb
//    patterns equals acc + sum_k 2^(16 k) c_k mod Q over 128-bit integers, at the extreme limb
//    values the bound allows and at random ones.
// Prints "ok <max error> <bound>" and exits 0, or the first failure.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../fhetoothpaste_amd/csrc/fft_key.hpp"
#include "../../fhetoothpaste_amd/csrc/fft_limbs.hpp"
#include "../../fhetoothpaste_amd/csrc/gl64.hpp"
#include "../../fhetoothpaste_amd/csrc/host_ntt.hpp"

static const uint64_t Q = gl::kQ;
static uint64_t g_state = 0x243F6A8885A308D3ULL;
static uint64_t rnd() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
#define FAIL(...)                              \
    do {                                       \
        fprintf(stderr, "FAIL: " __VA_ARGS__); \
        fprintf(stderr, "\n");                 \
        exit(1);                               \
    } while (0)

constexpr int N = 1024, H = 512, LIMBS = 4;

struct Cd {
    double re, im;
};

// double-precision folded transform in the device order (in place), and its inverse (x 512)
static void fwd_d(std::vector<Cd>& z, const std::vector<double>& zt) {
    for (int s = 0; s < 9; ++s) {
        const int len = H >> s, h = len / 2;
        for (int b = 0; b < (1 << s); ++b) {
            const double wr = zt[2 * ((2 << s) + b)], wi = zt[2 * ((2 << s) + b) + 1];
            for (int j = b * len; j < b * len + h; ++j) {
                const Cd v = z[j + h], u = z[j];
                const double tr = std::fma(v.re, wr, -(v.im * wi)), ti = std::fma(v.re, wi, v.im * wr);
                z[j] = {u.re + tr, u.im + ti};
                z[j + h] = {u.re - tr, u.im - ti};
            }
        }
    }
}
static void inv_d(std::vector<Cd>& z, const std::vector<double>& zt) {
    for (int s = 8; s >= 0; --s) {
        const int len = H >> s, h = len / 2;
        for (int b = 0; b < (1 << s); ++b) {
            const double wr = zt[2 * ((2 << s) + b)], wi = zt[2 * ((2 << s) + b) + 1];
            for (int j = b * len; j < b * len + h; ++j) {
                const Cd u = z[j], v = z[j + h];
                const Cd d = {u.re - v.re, u.im - v.im};
                z[j] = {u.re + v.re, u.im + v.im};
                z[j + h] = {std::fma(d.re, wr, d.im * wi), std::fma(d.im, wr, -(d.re * wi))};
            }
        }
    }
}

static int64_t centred(uint64_t x) { return x > Q / 2 ? (int64_t)(x - Q) : (int64_t)x; }

// balanced limbs, as device_key() splits them
static void limbs_of(uint64_t x, int32_t (&l)[LIMBS]) {
    int64_t v = centred(x);
    for (int k = 0; k < LIMBS - 1; ++k) {
        int64_t d = v & 0xFFFF;
        if (d >= 0x8000) d -= 0x10000;
        l[k] = (int32_t)d;
        v = (v - d) >> 16;
    }
    l[LIMBS - 1] = (int32_t)v;
}

static uint64_t mod_q(__int128 v) {
    __int128 r = v % (__int128)Q;
    if (r < 0) r += Q;
    return (uint64_t)r;
}

int main(int argc, char** argv) {
    const int trials = argc > 1 ? atoi(argv[1]) : 3;
    const tp::HostNtt T(N);
    // NTT round trip
    for (int it = 0; it < 4; ++it) {
        std::vector<uint64_t> a(N), b;
        for (auto& x : a) x = rnd() % Q;
        b = a;
        T.forward(b.data());
        T.inverse(b.data());
        if (a != b) FAIL("HostNtt inverse does not undo forward");
    }
    const std::vector<double> zt = tp::fftk::zeta_table();
    const int b = 12, l = 3, D = 6, first_shift = 60;
    double worst = 0.0, bound_any = tp::fftk::rounding_bound(b, l, tp::fftk::kKeyInfWorst), bound_key = 0;
    for (int trial = 0; trial < trials + 1; ++trial) {
        const bool adversarial = trial == trials;
        // one step of a key: rows r = h' l + j, halves h, coefficient-domain words
        std::vector<uint64_t> coef(2 * l * 2 * N), brk(2 * l * 2 * N);
        for (size_t m = 0; m < coef.size(); ++m) {
            if (adversarial) {
                // limbs (-2^15, 2^15 - 1, -2^15, 2^15) alternate in sign pattern per coefficient
                coef[m] = (m & 1) ? mod_q((__int128)0x7FFF7FFF7FFF7FFFLL) : mod_q(-(__int128)0x7FFF7FFF7FFF8000LL);
            } else {
                coef[m] = rnd() % Q;
            }
        }
        for (size_t p = 0; p < coef.size() / N; ++p) {
            std::copy(coef.begin() + p * N, coef.begin() + (p + 1) * N, brk.begin() + p * N);
            T.forward(brk.data() + p * N);
        }
        double kinf = 0;
        const std::vector<double> dk = tp::fftk::device_key(brk.data(), 1, l, &kinf);
        bound_key = std::max(bound_key, tp::fftk::rounding_bound(b, l, kinf));
        // digits of random canonical words (the top digit is partial), or all at their maximum
        std::vector<std::vector<int32_t>> d(2 * l, std::vector<int32_t>(N));
        for (int r = 0; r < 2 * l; ++r) {
            const int j = r % l;
            const int shift = first_shift - j * b;
            const int bits = std::min(b, 64 - shift);
            for (int m = 0; m < N; ++m) d[r][m] = adversarial ? (1 << bits) - 1 : (int32_t)((rnd() % Q) >> shift & ((1u << bits) - 1));
        }
        (void)D;
        for (int h = 0; h < 2; ++h) {
            // exact: sum_r NTT(d_r) . K_{r,h} -> INTT
            std::vector<uint64_t> acc(N, 0), tmp(N);
            for (int r = 0; r < 2 * l; ++r) {
                for (int m = 0; m < N; ++m) tmp[m] = (uint64_t)d[r][m];
                T.forward(tmp.data());
                const uint64_t* K = brk.data() + ((size_t)r * 2 + h) * N;
                for (int m = 0; m < N; ++m) acc[m] = gl::add(acc[m], gl::mul(tmp[m], K[m]));
            }
            T.inverse(acc.data());
            // floating point, limb by limb
            std::vector<uint64_t> got(N, 0);
            std::vector<std::array<double, LIMBS>> limb_val(N);
            for (int k = 0; k < LIMBS; ++k) {
                std::vector<Cd> F(H, Cd{0, 0});
                for (int r = 0; r < 2 * l; ++r) {
                    std::vector<Cd> z(H);
                    for (int m = 0; m < H; ++m) z[m] = {(double)d[r][m], (double)d[r][m + H]};
                    fwd_d(z, zt);
                    const double* K = dk.data() + 2 * ((((size_t)r * 2 + h) * LIMBS + k) * H);
                    for (int p = 0; p < H; ++p) {
                        const int L = p >> 3, t = p & 7;
                        const double kr = K[2 * (64 * t + L)], ki = K[2 * (64 * t + L) + 1];
                        F[p].re = std::fma(z[p].re, kr, F[p].re);
                        F[p].re = std::fma(-z[p].im, ki, F[p].re);
                        F[p].im = std::fma(z[p].re, ki, F[p].im);
                        F[p].im = std::fma(z[p].im, kr, F[p].im);
                    }
                }
                inv_d(F, zt);
                // exact limb convolution over the integers (schoolbook, negacyclic)
                std::vector<int64_t> ex(N, 0);
                for (int r = 0; r < 2 * l; ++r) {
                    const uint64_t* cf = coef.data() + ((size_t)r * 2 + h) * N;
                    std::vector<int32_t> lk(N);
                    for (int m = 0; m < N; ++m) {
                        int32_t lm[LIMBS];
                        limbs_of(cf[m], lm);
                        lk[m] = lm[k];
                    }
                    for (int a = 0; a < N; ++a) {
                        if (!d[r][a]) continue;
                        for (int c = 0; c < N; ++c) {
                            const int e = a + c;
                            const int64_t pr = (int64_t)d[r][a] * lk[c];
                            if (e < N) ex[e] += pr;
                            else ex[e - N] -= pr;
                        }
                    }
                }
                for (int m = 0; m < N; ++m) {
                    const double v = (m < H ? F[m].re : F[m - H].im) / 512.0;
                    const double err = std::fabs(v - (double)ex[m]);
                    if (err > worst) worst = err;
                    const int64_t rv = (int64_t)std::nearbyint(v);
                    if (rv != ex[m]) FAIL("limb %d coefficient %d rounds to %lld, exact %lld", k, m, (long long)rv,
                                          (long long)ex[m]);
                    limb_val[m][k] = v;
                }
            }
            const uint64_t qm1 = glv::qm1_const();
            for (int m = 0; m < N; ++m) {
                uint64_t bits[LIMBS];
                for (int k = 0; k < LIMBS; ++k) {
                    const double y = std::fma(limb_val[m][k], 1.0, 6755399441055744.0);  // v + 1.5 2^52
                    memcpy(&bits[k], &y, 8);
                }
                got[m] = tp::fftl::combine(0, bits, qm1);
                if (got[m] != acc[m]) FAIL("recombined coefficient %d (half %d, trial %d) differs", m, h, trial);
            }
        }
    }
    // recombination against 128-bit integers: limbs at +-2^40 (beyond any certified gadget's
    // 2LD N 2^12 2^15 <= 2^40) and random ones, accumulator random canonical
    {
        const uint64_t qm1 = glv::qm1_const();
        for (long it = 0; it < 400000; ++it) {
            int64_t c[LIMBS];
            for (int k = 0; k < LIMBS; ++k) {
                c[k] = (int64_t)(rnd() % (1ULL << 41)) - (1LL << 40);
                if (it < 64) c[k] = ((it >> k) & 1) ? (1LL << 40) : -(1LL << 40);
            }
            const uint64_t a = it < 64 ? (it & 1 ? Q - 1 : 0) : rnd() % Q;
            uint64_t bits[LIMBS];
            for (int k = 0; k < LIMBS; ++k) bits[k] = tp::fftl::kMagicBits + (uint64_t)c[k];
            __int128 want = a;
            for (int k = 0; k < LIMBS; ++k) want += (__int128)c[k] << (16 * k);
            if (tp::fftl::combine(a, bits, qm1) != mod_q(want)) FAIL("recombination differs at it %ld", it);
        }
    }
    if (worst >= bound_any) FAIL("observed error %.3g exceeds the bound %.3g", worst, bound_any);
    if (bound_any >= 0.5) FAIL("bound %.3g does not certify rounding", bound_any);
    printf("ok %.3g %.3g %.3g\n", worst, bound_any, bound_key);
    return 0;
}
