// Host-side key generation in the reference byte layouts, from one of two word streams:
//
//   seed == TP_SEED_OS_ENTROPY: ChaCha20 (RFC 8439, section 2.3) keyed with 256 bits and a 96-bit
//     nonce read from the operating system (getrandom), re-keyed before its 32-bit block counter
//     wraps.  This is the stream for real keys (the drop-in classes use it unless SetSeed was
//     called): ChaCha20's output does not reveal its key, so a published key word says nothing
//     about the noise drawn after it.
//   any other seed: a reproducible stream for tests and parity checks, SplitMix64 with state
//     seed ^ tag*0xD1B54A32D192ED03 ^ 0xA0761D6478BD642F.  Its outputs can be inverted to its
//     state; keys made from it are for tests only.
//
// The reference draws keys and encryption randomness from OpenFHE samplers that cannot be
// seeded (lib/src/base_crypto.cpp:34-38,139-142).
//   uniform mod Q: rejection of 64-bit draws >= Q;
//   Gaussian(sigma), sigma <= 21: inversion of a 63-bit cumulative table of exp(-k^2 / 2 sigma^2)
//     over |k| <= ceil(12 sigma), magnitude from the top 63 bits of a draw, sign from bit 0;
//   Gaussian(sigma), sigma > 21: the exact rejection sampler of Canonne, Kamath and Steinke
//     ("The Discrete Gaussian for Differential Privacy", 2020, Algorithms 1-3): a discrete
//     Laplace proposal of scale t = floor(sigma) + 1 accepted with probability
//     exp(-(|y| - sigma^2/t)^2 / (2 sigma^2)); Bernoulli(exp(-g)) draws compare a 53-bit uniform
//     with exp(-g).  No tail is cut off.
// Layouts: RGSW rows r < l encrypt -s m g_r, rows l + r encrypt m g_r, produced in the order of
// RLWEEncryptor::MakeRGSW (base_crypto.cpp:60-97); RLWE = [a | a s + NTT(e) + m] (:27-58);
// LWE = [a | <a, s> + e + m] (:137-155).
#include <sys/random.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/toothpaste_amd.h"
#include "gl64.hpp"
#include "host_ntt.hpp"

namespace {

// ChaCha20 block function (RFC 8439, section 2.3): 20 rounds over the state
// [constants | key | block counter | nonce], output = rounds(state) + state.
struct ChaCha20 {
    uint32_t key[8];
    uint32_t nonce[3];
    uint32_t counter = 0;
    uint32_t block[16];
    int pos = 16;  // next unread word of block

    static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
    static void qr(uint32_t* x, int a, int b, int c, int d) {
        x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
        x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
        x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
        x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
    }
    static void block_fn(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3], uint32_t out[16]) {
        uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) in[4 + i] = key[i];
        in[12] = counter;
        for (int i = 0; i < 3; ++i) in[13 + i] = nonce[i];
        uint32_t x[16];
        std::memcpy(x, in, sizeof x);
        for (int r = 0; r < 10; ++r) {
            qr(x, 0, 4, 8, 12); qr(x, 1, 5, 9, 13); qr(x, 2, 6, 10, 14); qr(x, 3, 7, 11, 15);
            qr(x, 0, 5, 10, 15); qr(x, 1, 6, 11, 12); qr(x, 2, 7, 8, 13); qr(x, 3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
    }
    bool rekey() {
        uint8_t buf[44];
        size_t got = 0;
        while (got < sizeof buf) {
            const ssize_t r = getrandom(buf + got, sizeof buf - got, 0);
            if (r < 0) {
                if (errno == EINTR) continue;
                return false;
            }
            got += (size_t)r;
        }
        std::memcpy(key, buf, 32);
        std::memcpy(nonce, buf + 32, 12);
        std::memset(buf, 0, sizeof buf);
        counter = 0;
        pos = 16;
        return true;
    }
    uint64_t next() {
        if (pos + 2 > 16) {
            if (counter == 0xFFFFFFFFu) rekey();
            block_fn(key, counter++, nonce, block);
            pos = 0;
        }
        const uint64_t lo = block[pos], hi = block[pos + 1];
        pos += 2;
        return lo | (hi << 32);
    }
};

struct Rng {
    bool os = false;
    bool ok = true;
    uint64_t s = 0;
    ChaCha20 cc;
    Rng(uint64_t seed, uint64_t tag) {
        if (seed == TP_SEED_OS_ENTROPY) {
            os = true;
            ok = cc.rekey();
        } else {
            s = seed ^ (tag * 0xD1B54A32D192ED03ULL) ^ 0xA0761D6478BD642FULL;
        }
    }
    uint64_t next() {
        if (os) return cc.next();
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    uint64_t uniform_q() {
        uint64_t x;
        do { x = next(); } while (x >= gl::kQ);
        return x;
    }
    // uniform in [0, n), n >= 1 (rejection from the top bits)
    uint64_t below(uint64_t n) {
        if (n <= 1) return 0;
        const uint64_t lim = ~0ULL - (~0ULL % n);
        uint64_t x;
        do { x = next(); } while (x >= lim);
        return x % n;
    }
    double unit() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1), 53 bits
};

// Canonne-Kamath-Steinke building blocks (sigma > 21).
bool bernoulli_exp(Rng& g, double gamma) { return g.unit() < std::exp(-gamma); }

int64_t discrete_laplace(Rng& g, uint64_t t) {
    for (;;) {
        const uint64_t u = g.below(t);
        if (!bernoulli_exp(g, (double)u / (double)t)) continue;
        uint64_t v = 0;
        while (bernoulli_exp(g, 1.0)) ++v;
        const bool b = g.next() >> 63;
        if (b && u == 0 && v == 0) continue;
        const int64_t z = (int64_t)(u + t * v);
        return b ? -z : z;
    }
}

struct Gauss {
    static constexpr double kTableMax = 21.0;  // 12 sigma fits the 256-entry table
    double sigma = 0;
    uint32_t K = 0;
    uint64_t T[256];
    explicit Gauss(double sg) : sigma(sg) {
        if (sigma <= 0) { T[0] = 1ULL << 63; return; }
        if (sigma > kTableMax) return;
        const uint32_t k_max = (uint32_t)std::ceil(sigma * 12.0);
        double w[256], tot = 0;
        for (uint32_t k = 0; k <= k_max; ++k) {
            w[k] = std::exp(-(double)(k * k) / (2.0 * sigma * sigma));
            tot += (k ? 2.0 * w[k] : w[k]);
        }
        double c = 0;
        for (uint32_t k = 0; k <= k_max; ++k) {
            c += (k ? 2.0 * w[k] : w[k]);
            const double v = std::ldexp(c / tot, 63);
            T[k] = (k == k_max || v >= 9223372036854775808.0) ? (1ULL << 63) : (uint64_t)v;
        }
        K = k_max;
    }
    int64_t sample_int(Rng& g) const {
        if (sigma > kTableMax) {
            const uint64_t t = (uint64_t)std::floor(sigma) + 1;
            const double s2 = sigma * sigma;
            for (;;) {
                const int64_t y = discrete_laplace(g, t);
                const double d = std::fabs((double)y) - s2 / (double)t;
                if (bernoulli_exp(g, d * d / (2.0 * s2))) return y;
            }
        }
        const uint64_t x = g.next();
        const uint64_t u = x >> 1;
        uint32_t k = 0;
        while (k < K && u >= T[k]) ++k;
        return (x & 1) ? -(int64_t)k : (int64_t)k;
    }
    uint64_t sample(Rng& g) const {
        const int64_t v = sample_int(g);
        return v >= 0 ? (uint64_t)v : gl::neg((uint64_t)(-v));
    }
};

uint32_t max_digits(uint32_t b) { return 64 / b + (64 % b ? 1 : 0); }

void gadget(uint64_t* g, uint32_t b, uint32_t l, int kind) {
    const uint32_t D = max_digits(b);
    for (uint32_t j = 0; j < l; ++j) {
        if (kind == 0) g[j] = 1ULL << (b * (D - 1 - j));
        else g[j] = (j == D - 1) ? 1 : (1ULL << (64 - b * (j + 1)));
    }
}

void make_rlwe(const tp::HostNtt& T, uint64_t* out, const uint64_t* msg_ntt, const uint64_t* s_ntt, Rng& g,
               const Gauss& G, uint64_t* e) {
    const uint32_t N = T.N;
    for (uint32_t i = 0; i < N; ++i) out[i] = g.uniform_q();
    for (uint32_t i = 0; i < N; ++i) e[i] = G.sample(g);
    T.forward(e);
    for (uint32_t i = 0; i < N; ++i) out[N + i] = gl::add(gl::add(gl::mul(out[i], s_ntt[i]), e[i]), msg_ntt[i]);
}

void make_rgsw_const(const tp::HostNtt& T, uint64_t* out, uint64_t m, const uint64_t* gad, uint32_t l,
                     const uint64_t* s_ntt, Rng& g, const Gauss& G, std::vector<uint64_t>& ws) {
    const uint32_t N = T.N;
    uint64_t* msg = ws.data();
    uint64_t* msgs = msg + N;
    uint64_t* e = msgs + N;
    for (uint32_t t = 0; t < l; ++t) {
        const uint32_t r = l - 1 - t;
        const uint64_t mg = gl::mul(m, gad[r]);
        for (uint32_t i = 0; i < N; ++i) {
            msg[i] = mg;
            msgs[i] = gl::neg(gl::mul(mg, s_ntt[i]));
        }
        make_rlwe(T, out + (size_t)r * 2 * N, msgs, s_ntt, g, G, e);
        make_rlwe(T, out + (size_t)r * 2 * N + (size_t)2 * N * l, msg, s_ntt, g, G, e);
    }
}

}  // namespace

extern "C" int tp_keygen_secret(uint64_t* out, uint32_t len, int dist, uint64_t seed, uint64_t tag) {
    if (!out || (dist != 0 && dist != 1)) return TP_E_INVALID;
    Rng g(seed, tag);
    if (!g.ok) return TP_E_ENTROPY;
    for (uint32_t i = 0; i < len; ++i) {
        if (dist == 0) out[i] = g.next() >> 63;
        else {
            const uint64_t r = g.next() >> 62;
            out[i] = r < 2 ? 0 : (r == 2 ? 1 : gl::kQ - 1);
        }
    }
    return TP_OK;
}

extern "C" int tp_keygen_brk(uint64_t* brk, const uint64_t* lwe_key, const uint64_t* rlwe_key, uint32_t N,
                             uint32_t n, uint32_t b, uint32_t l, int method, int gadget_kind, double sigma,
                             uint64_t seed) {
    if (!brk || !lwe_key || !rlwe_key || b == 0 || b > 32 || l == 0 || l > max_digits(b)) return TP_E_INVALID;
    if (N < 16 || (N & (N - 1))) return TP_E_INVALID;
    if (method == TP_METHOD_BMMP && (n & 1)) return TP_E_INVALID;
    tp::HostNtt T(N);
    Rng g(seed, 3);
    if (!g.ok) return TP_E_ENTROPY;
    Gauss G(sigma);
    std::vector<uint64_t> s_ntt(rlwe_key, rlwe_key + N), ws(3 * N);
    T.forward(s_ntt.data());
    uint64_t gad[64];
    gadget(gad, b, l, gadget_kind);
    const size_t rg = (size_t)4 * N * l;
    for (uint32_t i = 0; i < n; ++i) {
        if (method == TP_METHOD_CGGI_BINARY) {
            make_rgsw_const(T, brk + i * rg, lwe_key[i], gad, l, s_ntt.data(), g, G, ws);
        } else if (method == TP_METHOD_CGGI_TERNARY) {
            const uint64_t s = lwe_key[i];
            // cggi_blind_rotation.cpp:294-305: 0 -> (0,0), 1 -> (1,0), anything else -> (0,1)
            make_rgsw_const(T, brk + 2 * i * rg, s == 1, gad, l, s_ntt.data(), g, G, ws);
            make_rgsw_const(T, brk + (2 * i + 1) * rg, s != 0 && s != 1, gad, l, s_ntt.data(), g, G, ws);
        } else if (method == TP_METHOD_BMMP) {
            if (i & 1) continue;
            uint64_t* chunk = brk + (size_t)3 * (i >> 1) * rg;
            make_rgsw_const(T, chunk, lwe_key[i], gad, l, s_ntt.data(), g, G, ws);
            make_rgsw_const(T, chunk + rg, lwe_key[i + 1], gad, l, s_ntt.data(), g, G, ws);
            make_rgsw_const(T, chunk + 2 * rg, gl::mul(lwe_key[i], lwe_key[i + 1]), gad, l, s_ntt.data(), g, G,
                            ws);
        } else {
            return TP_E_INVALID;
        }
    }
    return TP_OK;
}

extern "C" int tp_keygen_lwe_ksk(uint64_t* ksk, const uint64_t* src, const uint64_t* dst, uint32_t n_in,
                                 uint32_t n_out, uint32_t b, uint32_t d, double sigma, uint64_t seed) {
    if (!ksk || !src || !dst || b == 0 || b > 32 || d == 0) return TP_E_INVALID;
    Rng g(seed, 4);
    if (!g.ok) return TP_E_ENTROPY;
    Gauss G(sigma);
    for (uint32_t i = 0; i < n_in; ++i) {
        uint64_t gij = 1ULL << (64 - b);
        for (uint32_t j = 0; j < d; ++j) {
            uint64_t* out = ksk + ((size_t)i * d + j) * (n_out + 1);
            for (uint32_t k = 0; k < n_out; ++k) out[k] = g.uniform_q();
            uint64_t bb = G.sample(g);
            for (uint32_t k = 0; k < n_out; ++k) bb = gl::add(bb, gl::mul(out[k], dst[k]));
            out[n_out] = gl::add(bb, gl::mul(gij, src[i]));
            gij >>= b;
        }
    }
    return TP_OK;
}

// Automorphism keys (automorphism_evaluation.cpp:43-63 with rlwe_conversion.cpp:187-218): key j-1
// is RLWEtoRLWEConverter's key from s(X^t) to s(X), t = 2^j + 1; row i = RLWE(g_i s(X^t)),
// g_i = 2^(64 - b(i+1)) (kept >= 1).  One stream (tag 5) over j = 1..log2 N in order.
extern "C" int tp_keygen_auto_keys(uint64_t* keys, const uint64_t* s, uint32_t N, uint32_t b, uint32_t d,
                                   double sigma, uint64_t seed) {
    if (!keys || !s || b == 0 || b > 32 || d == 0 || N < 16 || (N & (N - 1))) return TP_E_INVALID;
    tp::HostNtt T(N);
    Rng g(seed, 5);
    if (!g.ok) return TP_E_ENTROPY;
    Gauss G(sigma);
    std::vector<uint64_t> t_ntt(s, s + N), s_auto(N), msg(N), e(N);
    T.forward(t_ntt.data());
    for (uint32_t j = 1; j <= T.logN; ++j) {
        const uint64_t t = (1ULL << j) + 1;
        for (uint64_t i = 0; i < N; ++i) {
            const uint64_t p = (i * t) & (2 * (uint64_t)N - 1);
            s_auto[p & (N - 1)] = p >= N ? gl::neg(s[i]) : s[i];
        }
        uint64_t gij = 1ULL << (64 - b);
        for (uint32_t i = 0; i < d; ++i) {
            for (uint32_t k = 0; k < N; ++k) msg[k] = gl::mul(s_auto[k], gij);
            T.forward(msg.data());
            make_rlwe(T, keys + ((size_t)(j - 1) * d + i) * 2 * N, msg.data(), t_ntt.data(), g, G, e.data());
            gij >>= b;
            if (gij == 0) gij = 1;
        }
    }
    return TP_OK;
}

extern "C" int tp_keygen_rlwe_ksk(uint64_t* key, const uint64_t* src, const uint64_t* dst, uint32_t N, uint32_t b,
                                  uint32_t d, double sigma, uint64_t seed, uint64_t tag) {
    if (!key || !src || !dst || b == 0 || b > 32 || d == 0 || N < 16 || (N & (N - 1))) return TP_E_INVALID;
    tp::HostNtt T(N);
    Rng g(seed, tag);
    if (!g.ok) return TP_E_ENTROPY;
    Gauss G(sigma);
    std::vector<uint64_t> t_ntt(dst, dst + N), msg(N), e(N);
    T.forward(t_ntt.data());
    uint64_t gij = 1ULL << (64 - b);
    for (uint32_t i = 0; i < d; ++i) {
        for (uint32_t k = 0; k < N; ++k) msg[k] = gl::mul(src[k], gij);
        T.forward(msg.data());
        make_rlwe(T, key + (size_t)i * 2 * N, msg.data(), t_ntt.data(), g, G, e.data());
        gij >>= b;
        if (gij == 0) gij = 1;
    }
    return TP_OK;
}

extern "C" int tp_sample_gaussian(int64_t* out, size_t count, double sigma, uint64_t seed, uint64_t tag) {
    if (!out || !(sigma >= 0) || std::isinf(sigma)) return TP_E_INVALID;
    Rng g(seed, tag);
    if (!g.ok) return TP_E_ENTROPY;
    Gauss G(sigma);
    for (size_t i = 0; i < count; ++i) out[i] = G.sample_int(g);
    return TP_OK;
}

extern "C" int tp_chacha20_block(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3], uint32_t out[16]) {
    if (!key || !nonce || !out) return TP_E_INVALID;
    ChaCha20::block_fn(key, counter, nonce, out);
    return TP_OK;
}
