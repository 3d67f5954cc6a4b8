// Goldilocks field arithmetic, Q = 2^64 - 2^32 + 1, for gfx950 device code and host code.
//
// CDNA4 has no 64x64-bit multiply: a 64x64->128 product is four v_mad_u64_u32
// (32x32+64 -> 64) operations.  The 128-bit product is folded with the Solinas
// identities 2^64 = 2^32 - 1 and 2^96 = -1 (mod Q), the same identities the
// reference's scalar path uses (lib/include/backend/solinas_internal.h:61-73).
// Every function returns a canonical value (< Q) when its inputs are canonical.
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define GL_FN __host__ __device__ __forceinline__
#else
#define GL_FN inline
#endif

namespace gl {

constexpr uint64_t kQ = 0xFFFFFFFF00000001ULL;
constexpr uint64_t kEps = 0xFFFFFFFFULL;  // 2^64 mod Q

GL_FN uint64_t canon(uint64_t x) { return x >= kQ ? x - kQ : x; }

GL_FN uint64_t add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    // On carry, s + 2^64 - Q = s + eps; either way subtracting Q mod 2^64 is right.
    return (s < a || s >= kQ) ? s - kQ : s;
}

GL_FN uint64_t sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    return a < b ? d + kQ : d;
}

GL_FN uint64_t neg(uint64_t a) { return a ? kQ - a : 0; }

// Full 64x64 -> 128 product from 32-bit halves (4 multiply-adds, no overflow in any step).
GL_FN void mul_wide(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
    const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    const uint64_t p00 = a0 * b0;
    const uint64_t t = a0 * b1 + (p00 >> 32);
    const uint64_t u = a1 * b0 + (uint32_t)t;
    lo = (u << 32) | (uint32_t)p00;
    hi = a1 * b1 + (t >> 32) + (u >> 32);
}

// Any 128-bit value -> canonical.  x = lo + 2^64 h0 + 2^96 h1 = lo + eps h0 - h1.
GL_FN uint64_t reduce128(uint64_t hi, uint64_t lo) {
    const uint64_t h1 = hi >> 32, h0 = (uint32_t)hi;
    uint64_t t0 = lo - h1;
    if (lo < h1) t0 -= kEps;  // borrow: -2^64 = -eps
    const uint64_t t1 = (h0 << 32) - h0;
    uint64_t r = t0 + t1;
    if (r < t1) r += kEps;  // carry: +2^64 = +eps (cannot carry again)
    return canon(r);
}

GL_FN uint64_t mul(uint64_t a, uint64_t b) {
    uint64_t hi, lo;
    mul_wide(a, b, hi, lo);
    return reduce128(hi, lo);
}

// x * 2^-s for 0 < s <= 32:  x = q 2^s + r  ->  q + r 2^-s,  2^-s = -2^(96-s) = -eps 2^(32-s).
GL_FN uint64_t div_pow2(uint64_t x, int s) {
    const uint64_t q = x >> s;
    const uint64_t r = x & ((1ULL << s) - 1);
    const uint64_t v = r << (32 - s);  // < 2^32
    const uint64_t ve = (v << 32) - v;  // v * eps < Q
    return sub(q, ve);
}

// x * 2^e for 0 <= e < 192 (2^96 = -1, 2^192 = 1).  With a constant e every branch folds.
GL_FN uint64_t mul_pow2(uint64_t x, int e) {
    bool negate = false;
    if (e >= 96) { e -= 96; negate = true; }
    uint64_t r;
    if (e == 0) r = x;
    else if (e < 64) r = reduce128(x >> (64 - e), x << e);
    else { r = div_pow2(x, 96 - e); negate = !negate; }  // 2^e = -2^(e-96) = -2^-(96-e)
    return negate ? neg(r) : r;
}

// ------------------------------------------------------------------ lazy forms
// "Lazy" values are any 64-bit integers standing for their residue mod Q; they let
// butterflies skip canonicalisation.  Digit extraction and outputs use canonical values.

// Any 128-bit (hi, lo) -> lazy r == x mod Q (no final canonicalisation).
GL_FN uint64_t reduce128_lazy(uint64_t hi, uint64_t lo) {
    const uint64_t h0 = (uint32_t)hi, h1 = hi >> 32;
    uint64_t s, d;
    const bool c = __builtin_add_overflow(h0 * kEps, lo, &s);  // + 2^64 == + eps
    const bool b = __builtin_sub_overflow(s, h1, &d);          // - 2^64 == - eps
    return d + (uint64_t)(((int64_t)c - (int64_t)b) * (int64_t)kEps);
}

GL_FN uint64_t mul_lazy(uint64_t a, uint64_t b) {
    uint64_t hi, lo;
    mul_wide(a, b, hi, lo);
    return reduce128_lazy(hi, lo);
}

// Lazy a * b + c for any 64-bit a, b, c: the 128-bit sum is below 2^128 ((2^64 - 1)^2 + 2^64 - 1),
// so it is reduced once.  The 128-bit form lets the compiler chain the products and the addend
// (fewer instructions than a product reduction followed by a lazy add).
GL_FN uint64_t mac_lazy(uint64_t a, uint64_t b, uint64_t c) {
    const unsigned __int128 p = (unsigned __int128)a * b + c;
    return reduce128_lazy((uint64_t)(p >> 64), (uint64_t)p);
}

// a lazy, b < Q: one carry correction suffices (a + b < 2^64 + Q).
GL_FN uint64_t add_lc(uint64_t a, uint64_t b) {
    uint64_t s;
    const bool c = __builtin_add_overflow(a, b, &s);
    return s + (c ? kEps : 0);
}

// a lazy, b <= Q: one borrow correction suffices.
GL_FN uint64_t sub_lc(uint64_t a, uint64_t b) {
    uint64_t d;
    const bool c = __builtin_sub_overflow(a, b, &d);
    return d - (c ? kEps : 0);
}

// both lazy
GL_FN uint64_t add_ll(uint64_t a, uint64_t b) {
    uint64_t s, s2;
    const bool c = __builtin_add_overflow(a, b, &s);
    const bool c2 = __builtin_add_overflow(s, c ? kEps : 0, &s2);
    return s2 + (c2 ? kEps : 0);
}

GL_FN uint64_t sub_ll(uint64_t a, uint64_t b) {
    uint64_t d, d2;
    const bool c = __builtin_sub_overflow(a, b, &d);
    const bool c2 = __builtin_sub_overflow(d, c ? kEps : 0, &d2);
    return d2 - (c2 ? kEps : 0);
}

// x * 2^e for 0 < e < 32, x lazy: (x << e) + (x >> (64 - e)) eps.
GL_FN uint64_t mul_2e_small(uint64_t x, int e) {
    uint64_t s;
    const bool c = __builtin_add_overflow((x >> (64 - e)) * kEps, x << e, &s);
    return s + (c ? kEps : 0);
}

// x * 2^32 = x1 eps + (x0 << 32), x lazy.
GL_FN uint64_t mul_2_32(uint64_t x) {
    uint64_t s;
    const bool c = __builtin_add_overflow((x >> 32) * kEps, x << 32, &s);
    return s + (c ? kEps : 0);
}

// x * 2^-s for 0 < s <= 32, x lazy (q + r 2^-s with r 2^-s = -r eps 2^(32-s)).
GL_FN uint64_t div_2e_lazy(uint64_t x, int s) {
    const uint64_t q = x >> s;
    const uint64_t v = (x & ((1ULL << s) - 1)) << (32 - s);
    return sub_lc(q, (v << 32) - v);
}

// Lazy x * 2^e for 0 <= e < 192 when e is a compile-time constant after unrolling; the sign
// of 2^e (2^96 = -1) is returned separately so butterflies can swap outputs instead.
GL_FN uint64_t mul_2e_unsigned(uint64_t x, int e, bool& negate) {
    negate = false;
    if (e >= 96) { e -= 96; negate = true; }
    if (e == 0) return x;
    if (e < 32) return mul_2e_small(x, e);
    if (e == 32) return mul_2_32(x);
    if (e < 64) return mul_2e_small(mul_2_32(x), e - 32);
    negate = !negate;                 // 2^e = -2^-(96 - e)
    return div_2e_lazy(x, 96 - e);
}

GL_FN uint64_t pow(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b = canon(b);
    while (e) {
        if (e & 1) r = mul(r, b);
        b = mul(b, b);
        e >>= 1;
    }
    return r;
}

GL_FN uint64_t inv(uint64_t x) { return pow(x, kQ - 2); }

// Exponents e with psi_rev[k] = 2^e for k in [1, 32) (identical for every N >= 64, since these
// are 64th roots of unity and 8 = 2^3 has order 64): forward, and the inverse-root table.
constexpr int kFwdPow2[32] = {0,   48,  120, 168, 156, 12,  84,  132, 78,  126, 6,
                              54,  42,  90,  162, 18,  39,  87,  159, 15,  3,   51,
                              123, 171, 117, 165, 45,  93,  81,  129, 9,   57};
constexpr int kInvPow2[32] = {0,   144, 72,  24,  36,  180, 108, 60,  114, 66,  186,
                              138, 150, 102, 30,  174, 153, 105, 33,  177, 189, 141,
                              69,  21,  75,  27,  147, 99,  111, 63,  183, 135};

}  // namespace gl
