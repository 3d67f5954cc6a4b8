// Modulus switching v -> round(tM * v / sM) exactly as the reference computes it in x87 extended
// precision (lib/src/static/modulus_switching.cpp:36-93): `long double` has a 64-bit significand,
// rounding to nearest-even.  With sM >= 2^63 (exactly representable) and tM = 2^t, tM * v is
// exact, and the only rounding is the quotient's: fl(v / sM) = m * 2^-(64 + s), m the
// round-to-nearest-even 64-bit significand.  It is computed here with integers, so the result
// matches the reference's bit for bit, including the near-tie inputs where round-to-nearest of the
// exact rational would differ.  Then ROUND = floor(x + 1/2) (roundl: ties away from zero),
// FLOOR = floor(x), CEIL = ceil(x), and a result >= tM has tM subtracted once.
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define MS_FN __host__ __device__ __forceinline__
#else
#define MS_FN inline
#endif

namespace tp {

enum ModSwitchKind { kMsRound = 0, kMsFloor = 1, kMsCeil = 2 };

MS_FN int ms_clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll(x);
#else
    return __builtin_clzll(x);
#endif
}

// mu = floor((2^128 - 1) / sM) - 2^64, the part of the 65-bit reciprocal of sM below 2^64
// (sM >= 2^63), computed once per modulus on the host.
inline uint64_t ms_reciprocal(uint64_t sM) {
    return (uint64_t)((~(unsigned __int128)0) / sM - ((unsigned __int128)1 << 64));
}

// Round-to-nearest-even 64-bit significand of v / sM for 0 < v < sM, sM >= 2^63:
// v / sM = m * 2^-(64 + s), 2^63 <= m < 2^64.
MS_FN void ms_quotient(uint64_t v, uint64_t sM, uint64_t mu, uint64_t& m, int& s) {
    s = ms_clz64(v);
    uint64_t w = v << s;  // top bit set
    if (w >= sM) {        // then v 2^(s-1) lies in [sM/2, sM)
        --s;
        w = v << s;
    }
    // q = floor(w 2^64 / sM), w < sM, by the reciprocal: w (2^64 + mu) / 2^64 underestimates it by
    // less than 2, so at most two corrections follow; r = w 2^64 - q sM ends in [0, sM)
    uint64_t q = w + (uint64_t)(((unsigned __int128)w * mu) >> 64);
    unsigned __int128 rr = ((unsigned __int128)w << 64) - (unsigned __int128)q * sM;
    if (rr >= sM) {
        rr -= sM;
        ++q;
    }
    if (rr >= sM) {
        rr -= sM;
        ++q;
    }
    const uint64_t r = (uint64_t)rr;
    // nearest-even on the remainder: compare 2r with sM (2r may need 65 bits)
    const bool above = (r >> 63) || (r << 1) > sM;
    const bool tie = !(r >> 63) && (r << 1) == sM;
    if (above || (tie && (q & 1))) {
        ++q;
        if (q == 0) {  // carried out: v / sM = 2^-s exactly
            q = 1ULL << 63;
            --s;
        }
    }
    m = q;
}

// v < sM, sM >= 2^63, mu = ms_reciprocal(sM), tM = 2^t with 1 <= t <= 62.
MS_FN uint64_t ms_switch(uint64_t v, uint64_t sM, uint64_t mu, int t, int kind) {
    if (v == 0) return 0;
    uint64_t m;
    int s;
    ms_quotient(v, sM, mu, m, s);
    // x = tM fl(v/sM) = m 2^-(64 + s - t); sh = 64 + s - t >= 2 fractional bits
    const int sh = 64 + s - t;
    uint64_t ip, frac_half, frac_rest;  // integer part, first fractional bit, any later bit
    if (sh >= 65) {  // x < 2^64 2^-65 = 1/2
        ip = 0;
        frac_half = 0;
        frac_rest = 1;  // m != 0
    } else if (sh == 64) {
        ip = 0;
        frac_half = m >> 63;
        frac_rest = (m << 1) != 0;
    } else {
        ip = m >> sh;
        frac_half = (m >> (sh - 1)) & 1;
        frac_rest = (m & ((1ULL << (sh - 1)) - 1)) != 0;
    }
    uint64_t r = ip;
    if (kind == kMsRound) r += frac_half;
    else if (kind == kMsCeil) r += (frac_half | frac_rest) ? 1 : 0;
    const uint64_t tM = 1ULL << t;
    if (r >= tM) r -= tM;
    return r;
}

}  // namespace tp
