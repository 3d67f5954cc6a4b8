// Key-switching key as int8 matrix-core operands (host side of kernels_ks_mfma.hip).
//
// The LWE key switch out[c][col] = sum_ij u_ij(c) K[i][j][col] (unsigned digits u < 2^b, then the
// per-key correction and the negation of lwe_keyswitch_kernel) is a product of a (count x n_in 16)
// digit matrix and a (n_in 16 x W) key matrix: word i of a ciphertext fills 16 digit slots
// (digits j < d, the rest zero).  Each key word is split into eight bytes, K = sum_t 2^(8t) K_t, and
// each byte is stored minus 128 so that it fits a signed 8-bit operand; the kernel adds
// 128 sum_ij u_ij back per ciphertext.  The matrix is stored in the order the kernel reads it:
// column group g (32 columns), k-step s (two words, 32 slots), byte t, then one 16-byte operand per
// lane of a 64-lane wave (lane l: column 32 g + l % 32, word 2 s + l / 32, its 16 slots).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <vector>

namespace tp {

// Digit held by 16-byte slot p of a word's operand, or -1 for a zero slot.  b = 4: the kernel
// unpacks the nibbles of (x + C) >> (64 - 4 d) with two masks (slots 0..7: even nibbles, 8..15: odd
// ones), nibble q holding digit d - 1 - q; other bases: slot p holds digit p.
inline int ks_slot_digit(int b, int d, int p) {
    if (b == 4) {
        const int q = p < 8 ? 2 * p : 2 * (p - 8) + 1;
        return q < d ? d - 1 - q : -1;
    }
    return p < d ? p : -1;
}

// Whether the matrix-core key switch covers (n_in, b, d): digits fit the 16 slots and a signed
// byte, the k loop splits into whole chunks of 8 words, and every 32-bit partial sum is exact
// (n_in d (2^b - 1) 255 < 2^31).
inline bool ks_mfma_fits(uint32_t n_in, uint32_t b, uint32_t d) {
    if (b < 1 || b > 7 || d < 1 || d > 16 || b * d > 64 || n_in % 8 != 0) return false;
    return (uint64_t)n_in * d * ((1u << b) - 1) * 255 < (1ULL << 31);
}

inline size_t ks_limb_bytes(uint32_t n_in, uint32_t n_out) {
    const size_t groups = (n_out + 1 + 31) / 32;
    return groups * (n_in / 2) * 8 * 1024;
}

// ksk: [n_in][d][n_out + 1] (reference layout).  out: ks_limb_bytes(n_in, n_out) bytes.
inline void build_ks_limbs(const uint64_t* ksk, uint32_t n_in, uint32_t n_out, uint32_t b, uint32_t d,
                           int8_t* out) {
    const size_t W = n_out + 1, groups = (W + 31) / 32, steps = n_in / 2;
    int slot_j[16];
    for (int p = 0; p < 16; ++p) slot_j[p] = ks_slot_digit((int)b, (int)d, p);
    size_t o = 0;
    for (size_t g = 0; g < groups; ++g)
        for (size_t s = 0; s < steps; ++s)
            for (int t = 0; t < 8; ++t)
                for (int lane = 0; lane < 64; ++lane) {
                    const size_t col = 32 * g + (lane & 31), i = 2 * s + (lane >> 5);
                    for (int p = 0; p < 16; ++p) {
                        const int j = slot_j[p];
                        const uint64_t v = (j >= 0 && col < W) ? ksk[(i * d + (size_t)j) * W + col] : 0;
                        out[o++] = (int8_t)((int)((v >> (8 * t)) & 255) - 128);
                    }
                }
}

}  // namespace tp
