// CPU port of the batched binary-CGGI bootstrap, vectorised with AVX2 (four Goldilocks words per
// 256-bit register), multi-threaded over ciphertexts.
//
// TEST INFRASTRUCTURE ONLY: this is the CPU baseline that bench.py times beside the GPU path, and
// tests/test_cpu_port.py checks it word for word against the scalar oracle (toothpaste_oracle.c).
// The product never loads it.
//
// It follows the oracle's (and the reference's) order of operations, vectorised:
//   EvalFunc (lib/src/operators/negacyclic_function_evaluator.cpp:92-101)
//   -> CGGIBlindRotator::BlindRotate, binary key (lib/src/operators/cggi_blind_rotation.cpp:313-386):
//      acc in the NTT domain, per step the monomial product NTT(X^-a) - NTT(1), two inverse
//      transforms, the unsigned digits of both halves, 2l forward transforms and the
//      multiply-accumulate with the key (lib/src/operators/mux_operator.cpp:30-84,136-152)
//   -> SampleExtract (lib/include/static/sample_extraction.h:20-36)
//   -> LWEtoLWEConverter::Convert, signed digits (lib/src/operators/lwe_conversion.cpp:205-239).
// The NTT is the reference backend's convention (lib/src/backend/solinas.cpp:228-313): radix-2
// Cooley-Tukey forward with psi_rev twiddles, Gentleman-Sande inverse with psi_inv_rev and N^-1.
// Every value is kept canonical, so the words equal the scalar oracle's.
#include <immintrin.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kQ = 0xFFFFFFFF00000001ULL;
constexpr uint64_t kEps = 0xFFFFFFFFULL;  // 2^64 mod Q

// ---------------------------------------------------------------- scalar field (tables, tails)
inline uint64_t s_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    if (s < a) s += kEps;  // a + b - 2^64 <= Q - 2 - eps: no second wrap
    return s >= kQ ? s - kQ : s;
}
inline uint64_t s_sub(uint64_t a, uint64_t b) { return a >= b ? a - b : a - b + kQ; }
inline uint64_t s_neg(uint64_t a) { return a ? kQ - a : 0; }
inline uint64_t s_mul(uint64_t a, uint64_t b) {
    const unsigned __int128 p = (unsigned __int128)a * b;
    return (uint64_t)(p % kQ);
}
uint64_t s_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = s_mul(r, a);
        a = s_mul(a, a);
        e >>= 1;
    }
    return r;
}
uint32_t brv(uint32_t x, unsigned bits) {
    uint32_t r = 0;
    for (unsigned i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// ---------------------------------------------------------------- AVX2 field
typedef __m256i V;
inline V vset(uint64_t x) { return _mm256_set1_epi64x((long long)x); }
inline V vld(const uint64_t* p) { return _mm256_loadu_si256((const V*)p); }
inline void vst(uint64_t* p, V x) { _mm256_storeu_si256((V*)p, x); }
// unsigned a < b, as a lane mask (AVX2 compares are signed: flip the sign bits)
inline V ltu(V a, V b) {
    const V s = vset(0x8000000000000000ULL);
    return _mm256_cmpgt_epi64(_mm256_xor_si256(b, s), _mm256_xor_si256(a, s));
}
// [0, 2^64) -> [0, Q): x >= Q becomes x - Q = x + eps (mod 2^64)
inline V canon(V x) {
    const V s = vset(0x8000000000000000ULL);
    const V ge = _mm256_cmpgt_epi64(_mm256_xor_si256(x, s), vset((kQ - 1) ^ 0x8000000000000000ULL));
    return _mm256_add_epi64(x, _mm256_and_si256(ge, vset(kEps)));
}
// canonical + canonical -> canonical
inline V vadd(V a, V b) {
    V s = _mm256_add_epi64(a, b);
    s = _mm256_add_epi64(s, _mm256_and_si256(ltu(s, a), vset(kEps)));
    return canon(s);
}
// canonical - canonical -> canonical: a borrow gives a - b + 2^64, and a - b + Q is that minus eps
inline V vsub(V a, V b) {
    const V d = _mm256_sub_epi64(a, b);
    return _mm256_sub_epi64(d, _mm256_and_si256(ltu(a, b), vset(kEps)));
}
// canonical * canonical -> canonical: the 128-bit product from four 32 x 32 -> 64 products, then
// lo + h0 2^64 + h1 2^96 = lo + h0 eps - h1 (2^64 = eps, 2^96 = -1 mod Q)
inline V vmul(V a, V b) {
    const V ah = _mm256_srli_epi64(a, 32), bh = _mm256_srli_epi64(b, 32);
    const V p00 = _mm256_mul_epu32(a, b), p01 = _mm256_mul_epu32(a, bh);
    const V p10 = _mm256_mul_epu32(ah, b), p11 = _mm256_mul_epu32(ah, bh);
    const V mid = _mm256_add_epi64(p01, p10);
    const V cm = _mm256_and_si256(ltu(mid, p01), vset(1ULL << 32));  // carry of mid: 2^96 = 2^32 in hi
    const V lo = _mm256_add_epi64(p00, _mm256_slli_epi64(mid, 32));
    const V cl = ltu(lo, p00);                                           // carry into hi (mask = -1)
    V hi = _mm256_add_epi64(p11, _mm256_srli_epi64(mid, 32));
    hi = _mm256_sub_epi64(_mm256_add_epi64(hi, cm), cl);
    const V h0 = _mm256_and_si256(hi, vset(kEps)), h1 = _mm256_srli_epi64(hi, 32);
    V t = _mm256_sub_epi64(lo, h1);
    t = _mm256_sub_epi64(t, _mm256_and_si256(ltu(lo, h1), vset(kEps)));
    const V u = _mm256_sub_epi64(_mm256_slli_epi64(h0, 32), h0);  // h0 eps < 2^64
    V r = _mm256_add_epi64(t, u);
    r = _mm256_add_epi64(r, _mm256_and_si256(ltu(r, u), vset(kEps)));
    return canon(r);
}

// ---------------------------------------------------------------- ring
struct Ring {
    uint32_t N = 0, logN = 0;
    uint64_t n_inv = 0;
    std::vector<uint64_t> psi_rev, psi_inv_rev, mono;  // mono: rows NTT(X^i), i < N
    explicit Ring(uint32_t n);
    void fwd(uint64_t* a) const;
    void inv(uint64_t* a) const;
};

// Cooley-Tukey level with half distance t and m blocks, four butterflies per instruction group.
// t >= 4: the two halves of a block are contiguous vectors; t = 2 and t = 1 regroup two or four
// blocks into (u, v) vectors with lane permutes.
void ct_level(uint64_t* a, uint32_t m, uint32_t t, const uint64_t* w) {
    if (t >= 4) {
        for (uint32_t i = 0; i < m; ++i) {
            const V wv = vset(w[i]);
            uint64_t* p = a + 2 * (size_t)i * t;
            for (uint32_t j = 0; j < t; j += 4) {
                const V u = vld(p + j), v = vmul(vld(p + j + t), wv);
                vst(p + j, vadd(u, v));
                vst(p + j + t, vsub(u, v));
            }
        }
    } else if (t == 2) {
        for (uint32_t i = 0; i < m; i += 2) {
            uint64_t* p = a + 4 * (size_t)i;
            const V A = vld(p), B = vld(p + 4);  // [u0 u1 v0 v1], [u2 u3 v2 v3]
            const V U = _mm256_permute2x128_si256(A, B, 0x20), Vv = _mm256_permute2x128_si256(A, B, 0x31);
            const V W = _mm256_setr_epi64x((long long)w[i], (long long)w[i], (long long)w[i + 1], (long long)w[i + 1]);
            const V x = vmul(Vv, W), s = vadd(U, x), d = vsub(U, x);
            vst(p, _mm256_permute2x128_si256(s, d, 0x20));
            vst(p + 4, _mm256_permute2x128_si256(s, d, 0x31));
        }
    } else {
        for (uint32_t i = 0; i < m; i += 4) {
            uint64_t* p = a + 2 * (size_t)i;
            const V A = vld(p), B = vld(p + 4);  // [u0 v0 u1 v1], [u2 v2 u3 v3]
            const V U = _mm256_unpacklo_epi64(A, B), Vv = _mm256_unpackhi_epi64(A, B);  // [u0 u2 u1 u3]
            const V W = _mm256_setr_epi64x((long long)w[i], (long long)w[i + 2], (long long)w[i + 1], (long long)w[i + 3]);
            const V x = vmul(Vv, W), s = vadd(U, x), d = vsub(U, x);
            vst(p, _mm256_unpacklo_epi64(s, d));
            vst(p + 4, _mm256_unpackhi_epi64(s, d));
        }
    }
}

// Gentleman-Sande level: (u, v) -> (u + v, (u - v) w).
void gs_level(uint64_t* a, uint32_t m, uint32_t t, const uint64_t* w) {
    if (t >= 4) {
        for (uint32_t i = 0; i < m; ++i) {
            const V wv = vset(w[i]);
            uint64_t* p = a + 2 * (size_t)i * t;
            for (uint32_t j = 0; j < t; j += 4) {
                const V u = vld(p + j), v = vld(p + j + t);
                vst(p + j, vadd(u, v));
                vst(p + j + t, vmul(vsub(u, v), wv));
            }
        }
    } else if (t == 2) {
        for (uint32_t i = 0; i < m; i += 2) {
            uint64_t* p = a + 4 * (size_t)i;
            const V A = vld(p), B = vld(p + 4);
            const V U = _mm256_permute2x128_si256(A, B, 0x20), Vv = _mm256_permute2x128_si256(A, B, 0x31);
            const V W = _mm256_setr_epi64x((long long)w[i], (long long)w[i], (long long)w[i + 1], (long long)w[i + 1]);
            const V s = vadd(U, Vv), d = vmul(vsub(U, Vv), W);
            vst(p, _mm256_permute2x128_si256(s, d, 0x20));
            vst(p + 4, _mm256_permute2x128_si256(s, d, 0x31));
        }
    } else {
        for (uint32_t i = 0; i < m; i += 4) {
            uint64_t* p = a + 2 * (size_t)i;
            const V A = vld(p), B = vld(p + 4);
            const V U = _mm256_unpacklo_epi64(A, B), Vv = _mm256_unpackhi_epi64(A, B);
            const V W = _mm256_setr_epi64x((long long)w[i], (long long)w[i + 2], (long long)w[i + 1], (long long)w[i + 3]);
            const V s = vadd(U, Vv), d = vmul(vsub(U, Vv), W);
            vst(p, _mm256_unpacklo_epi64(s, d));
            vst(p + 4, _mm256_unpackhi_epi64(s, d));
        }
    }
}

Ring::Ring(uint32_t n) : N(n) {
    while ((1u << logN) < N) ++logN;
    const uint64_t psi = s_pow(7, (kQ - 1) >> (logN + 1));  // 7 generates Z_Q^* (solinas.cpp:175-176)
    const uint64_t psi_inv = s_pow(psi, kQ - 2);
    n_inv = s_pow(N, kQ - 2);
    psi_rev.resize(N);
    psi_inv_rev.resize(N);
    for (uint32_t i = 0; i < N; ++i) {
        psi_rev[i] = s_pow(psi, brv(i, logN));
        psi_inv_rev[i] = s_pow(psi_inv, brv(i, logN));
    }
    mono.assign((size_t)N * N, 0);
    for (uint32_t i = 0; i < N; ++i) {
        mono[(size_t)i * N + i] = 1;
        fwd(mono.data() + (size_t)i * N);
    }
}

void Ring::fwd(uint64_t* a) const {
    for (uint32_t m = 1, t = N / 2; m < N; m <<= 1, t >>= 1) ct_level(a, m, t, psi_rev.data() + m);
}

void Ring::inv(uint64_t* a) const {
    for (uint32_t m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1) gs_level(a, m, t, psi_inv_rev.data() + m);
    const V ni = vset(n_inv);
    for (uint32_t j = 0; j < N; j += 4) vst(a + j, vmul(vld(a + j), ni));
}

// o = a * b, o = o + a * b, o = a + b over n words (n a multiple of 4)
inline void v_mul(uint64_t* o, const uint64_t* a, const uint64_t* b, uint32_t n) {
    for (uint32_t i = 0; i < n; i += 4) vst(o + i, vmul(vld(a + i), vld(b + i)));
}
inline void v_add(uint64_t* o, const uint64_t* a, const uint64_t* b, uint32_t n) {
    for (uint32_t i = 0; i < n; i += 4) vst(o + i, vadd(vld(a + i), vld(b + i)));
}

uint32_t max_digits(uint32_t b) { return 64 / b + (64 % b ? 1 : 0); }

// Working memory of one thread.
struct Work {
    std::vector<uint64_t> acc, s, tmp, mono, scratch, ext, ks;
    Work(uint32_t N, uint32_t n_out)
        : acc(2 * N), s(2 * N), tmp(2 * N), mono(N), scratch(N), ext(N + 1), ks(n_out + 1 + 4) {}
};

// RLWEPrimeProduct (mux_operator.cpp:30-84): unsigned digits of poly, forward transforms, and the
// products with the key rows accumulated into acc (overwritten when !add on the first digit).
void prime_product(const Ring& R, uint64_t* acc, const uint64_t* rp, const uint64_t* poly, bool add, uint32_t b,
                   uint32_t l, uint64_t* scratch) {
    const uint32_t N = R.N;
    const uint32_t first = b * (max_digits(b) - 1);
    const V mask = vset((1ULL << b) - 1);
    for (uint32_t j = 0; j < l; ++j) {
        const __m128i sh = _mm_cvtsi32_si128((int)(first - j * b));
        for (uint32_t i = 0; i < N; i += 4) vst(scratch + i, _mm256_and_si256(_mm256_srl_epi64(vld(poly + i), sh), mask));
        R.fwd(scratch);
        const uint64_t* ra = rp + (size_t)j * 2 * N;
        if (!add && j == 0) {
            v_mul(acc, ra, scratch, N);
            v_mul(acc + N, ra + N, scratch, N);
        } else {
            for (uint32_t i = 0; i < N; i += 4) {
                const V d = vld(scratch + i);
                vst(acc + i, vadd(vld(acc + i), vmul(vld(ra + i), d)));
                vst(acc + N + i, vadd(vld(acc + N + i), vmul(vld(ra + N + i), d)));
            }
        }
    }
}

// Blind rotation, binary key, into coefficients (cggi_blind_rotation.cpp:313-386 as the oracle's
// or_blind_rotate restates it).
void blind_rotate(const Ring& R, Work& W, const uint64_t* lwe, const uint64_t* tv, const uint64_t* brk, uint32_t n,
                  uint32_t b, uint32_t l) {
    const uint32_t N = R.N;
    uint64_t* acc = W.acc.data();
    memcpy(acc, tv, sizeof(uint64_t) * 2 * N);
    R.fwd(acc);
    R.fwd(acc + N);
    // times X^b (NTT domain): a signed monomial row
    uint64_t bb = lwe[n];
    const bool flip = bb >= N;
    if (flip) bb -= N;
    for (uint32_t i = 0; i < 2 * N; ++i) W.tmp[i] = flip ? s_neg(acc[i]) : acc[i];
    v_mul(acc, W.tmp.data(), R.mono.data() + bb * N, N);
    v_mul(acc + N, W.tmp.data() + N, R.mono.data() + bb * N, N);
    const size_t rg = (size_t)4 * N * l;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t ai = lwe[i];
        if (ai == 0) continue;
        // NTT(X^(2N - a_i)) - NTT(1)
        const uint64_t e = 2 * N - ai;
        const uint64_t* row = R.mono.data() + (e >= N ? e - N : e) * N;
        const V one = vset(1);
        if (e >= N) {
            for (uint32_t k = 0; k < N; k += 4) vst(W.mono.data() + k, vsub(vsub(vset(0), vld(row + k)), one));
        } else {
            for (uint32_t k = 0; k < N; k += 4) vst(W.mono.data() + k, vsub(vld(row + k), one));
        }
        // BinaryCMux (mux_operator.cpp:136-152)
        v_mul(W.s.data(), acc, W.mono.data(), N);
        v_mul(W.s.data() + N, acc + N, W.mono.data(), N);
        R.inv(W.s.data());
        R.inv(W.s.data() + N);
        const uint64_t* rgsw = brk + i * rg;
        prime_product(R, W.tmp.data(), rgsw, W.s.data(), false, b, l, W.scratch.data());
        prime_product(R, W.tmp.data(), rgsw + (size_t)2 * N * l, W.s.data() + N, true, b, l, W.scratch.data());
        v_add(acc, acc, W.tmp.data(), 2 * N);
    }
    R.inv(acc);
    R.inv(acc + N);
}

// LWE key switch with signed digits (lwe_conversion.cpp:205-239), overwrite semantics; columns
// in vectors of four (the last partial vector through a padded copy).
void keyswitch(uint64_t* out, const uint64_t* in, const uint64_t* ksk, uint32_t n_in, uint32_t n_out, uint32_t b,
               uint32_t d, uint64_t* accbuf) {
    const uint32_t D = max_digits(b);
    uint64_t C = 0;  // corrector (gadget_decomp.h:16-28)
    int64_t shift0 = 64 - (int64_t)b;
    while (shift0 > 0) {
        C += 1ULL << (shift0 + b - 1);
        shift0 -= b;
    }
    const uint64_t least = 1ULL << (shift0 + b - 1);
    C = (C + least) % kQ;
    const uint64_t mask = (1ULL << b) - 1;
    const uint32_t W = n_out + 1, Wv = W & ~3u;
    for (uint32_t k = 0; k < W + 4; ++k) accbuf[k] = 0;
    uint64_t pad[4];
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint64_t ai = s_add(in[i], C);
        const uint64_t* chunk = ksk + (size_t)i * d * W;
        uint32_t shift = 64 - b;
        for (uint32_t j = 0; j < d; ++j, chunk += W) {
            const uint64_t delta = (j != D - 1) ? (1ULL << (b - 1)) : least;
            const V dij = vset(s_sub((ai >> shift) & mask, delta));
            uint32_t k = 0;
            for (; k < Wv; k += 4) vst(accbuf + k, vadd(vld(accbuf + k), vmul(vld(chunk + k), dij)));
            if (k < W) {
                for (uint32_t q = 0; q < 4; ++q) pad[q] = k + q < W ? chunk[k + q] : 0;
                vst(accbuf + k, vadd(vld(accbuf + k), vmul(vld(pad), dij)));
            }
            shift -= b;
        }
    }
    for (uint32_t k = 0; k < n_out; ++k) out[k] = s_neg(accbuf[k]);
    out[n_out] = s_sub(in[n_in], accbuf[n_out]);
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

// 1 when this build's vector path can run here (AVX2).
int cp_available(void) { return __builtin_cpu_supports("avx2") ? 1 : 0; }

// In-place forward (inverse != 0: inverse) transforms of count polynomials of N words.
void cp_ntt(uint32_t N, uint64_t* data, size_t count, int inverse) {
    const Ring R(N);
    for (size_t c = 0; c < count; ++c) {
        if (inverse) R.inv(data + c * N);
        else R.fwd(data + c * N);
    }
}

// Seconds per transform of one polynomial (forward, then inverse), over reps transforms.
void cp_ntt_time(uint32_t N, int reps, double* fwd_s, double* inv_s) {
    const Ring R(N);
    std::vector<uint64_t> a(N);
    for (uint32_t i = 0; i < N; ++i) a[i] = s_mul(i + 1, 0x9E3779B97F4A7C15ULL % kQ);
    double t0 = now();
    for (int r = 0; r < reps; ++r) R.fwd(a.data());
    double t1 = now();
    for (int r = 0; r < reps; ++r) R.inv(a.data());
    double t2 = now();
    *fwd_s = (t1 - t0) / reps;
    *inv_s = (t2 - t1) / reps;
}

// Batched bootstrap: blind rotation (binary key, brk in the reference's NTT-domain layout),
// sample extraction, and the LWE key switch when ksk is given (out: count x (n_out + 1), else
// count x (N + 1)).  Ciphertexts are split into contiguous ranges over `threads` threads;
// per_ct (optional, count doubles) receives each bootstrap's own duration.  Returns the elapsed
// seconds of the whole batch, or -1 on bad arguments.
double cp_bootstrap_batch(uint32_t N, uint64_t* out, const uint64_t* lwes, size_t count, const uint64_t* tv,
                          const uint64_t* brk, uint32_t n, uint32_t b, uint32_t l, const uint64_t* ksk, uint32_t n_out,
                          uint32_t ks_b, uint32_t ks_d, int threads, double* per_ct) {
    if (N < 16 || (N & (N - 1)) || b == 0 || b > 32 || l == 0 || b * (max_digits(b) - 1) < (l - 1) * b) return -1.0;
    if (threads < 1) threads = 1;
    const Ring R(N);
    const double t0 = now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const size_t c0 = count * t / threads, c1 = count * (t + 1) / threads;
        pool.emplace_back([&, c0, c1]() {
            Work W(N, n_out);
            for (size_t c = c0; c < c1; ++c) {
                const double s0 = per_ct ? now() : 0.0;
                blind_rotate(R, W, lwes + c * (n + 1), tv, brk, n, b, l);
                // SampleExtract at index 0
                const uint64_t* a = W.acc.data();
                W.ext[0] = a[0];
                for (uint32_t i = 1; i < N; ++i) W.ext[i] = s_neg(a[N - i]);
                W.ext[N] = a[N];
                if (ksk) keyswitch(out + c * (n_out + 1), W.ext.data(), ksk, N, n_out, ks_b, ks_d, W.ks.data());
                else memcpy(out + c * (N + 1), W.ext.data(), sizeof(uint64_t) * (N + 1));
                if (per_ct) per_ct[c] = now() - s0;
            }
        });
    }
    for (auto& th : pool) th.join();
    return now() - t0;
}

}  // extern "C"
