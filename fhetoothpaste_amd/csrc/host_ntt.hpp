// Host-side negacyclic NTT tables and a scalar forward transform (key generation only;
// the hot path runs on the GPU).  Same convention as the reference
// (lib/src/backend/solinas.cpp:168-261): psi = 7^((Q-1)/2N), psi_rev[i] = psi^brv(i),
// natural order in, bit-reversed order out.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <vector>

#include "gl64.hpp"

namespace tp {

struct HostNtt {
    uint32_t N = 0, logN = 0;
    std::vector<uint64_t> psi_rev, psi_inv_rev;

    explicit HostNtt(uint32_t n) : N(n) {
        while ((1u << logN) < N) ++logN;
        const uint64_t psi = gl::pow(7, (gl::kQ - 1) >> (logN + 1));
        const uint64_t psi_inv = gl::inv(psi);
        std::vector<uint64_t> pw(N), pwi(N);
        pw[0] = pwi[0] = 1;
        for (uint32_t i = 1; i < N; ++i) {
            pw[i] = gl::mul(pw[i - 1], psi);
            pwi[i] = gl::mul(pwi[i - 1], psi_inv);
        }
        psi_rev.resize(N);
        psi_inv_rev.resize(N);
        for (uint32_t i = 0; i < N; ++i) {
            uint32_t r = 0, x = i;
            for (uint32_t b = 0; b < logN; ++b) { r = (r << 1) | (x & 1); x >>= 1; }
            psi_rev[i] = pw[r];
            psi_inv_rev[i] = pwi[r];
        }
    }

    // Tables of the twisted forward transform at N = 1024 (2N words, placed after psi_rev):
    // [0, N): rho_j^i, rho_j = psi^(2 brv4(j) + 1), for coefficient i of block j (64 consecutive
    //   NTT indices) in key-row pair order: word 2 (64 c + L) + q holds lane L's k = 2c + q, the
    //   pass-2 element i = 4 k + (L & 3) of block j = L >> 2;
    // [N, 2N): at N + 256 + q, w^(2 brv4(q & 15)), and at N + 512 + q, w^(brv5(q & 31)),
    //   w = psi^32: the cyclic 64-point transform's last two levels at the places where the
    //   plain transform keeps its levels 8 and 9; the rest is zero.
    // The inverse transform's tables (inverse = true) hold the inverses at the same places, the
    // untwist rho_j^-i times N^-1 (the inverse transform's only scaling).
    std::vector<uint64_t> twisted64(bool inverse = false) const {
        std::vector<uint64_t> t(2 * (size_t)N, 0);
        if (N != 1024) return t;
        auto brv = [](uint32_t x, int bits) {
            uint32_t r = 0;
            for (int b = 0; b < bits; ++b) { r = (r << 1) | (x & 1); x >>= 1; }
            return r;
        };
        const uint64_t psi = inverse ? gl::inv(psi_rev[N / 2]) : psi_rev[N / 2];
        const uint64_t w = gl::pow(psi, 32);
        const uint64_t scale = inverse ? gl::inv(N) : 1;
        for (uint32_t c = 0; c < 8; ++c)
            for (uint32_t L = 0; L < 64; ++L)
                for (uint32_t q = 0; q < 2; ++q) {
                    const uint32_t k = 2 * c + q, j = L >> 2, i = 4 * k + (L & 3);
                    const uint64_t rho = gl::pow(psi, 2 * brv(j, 4) + 1);
                    t[2 * (64 * c + L) + q] = gl::mul(gl::pow(rho, i), scale);
                }
        for (uint32_t q = 0; q < 256; ++q) t[N + 256 + q] = gl::pow(w, 2 * brv(q & 15, 4));
        for (uint32_t q = 0; q < 512; ++q) t[N + 512 + q] = gl::pow(w, brv(q & 31, 5));
        return t;
    }

    // Tables of the two-wave N = 1024 transforms (2N words): [0, N) the twist rho_j^i at
    // (64 q + l) 8 + k for position 512 q + ((l & 7) | k << 3 | (l >> 3) << 6) (block j, index i
    // inside it); [N, 2N) at (64 q + l) 8 + s the cyclic levels' twiddles of lane l, with m = l & 7:
    // s = 0: w^(4 brv3(m)) (t3), s = 1, 2: w^(2 brv4(2m + s - 1)) (t4), s = 3..6: w^(brv5(4m + s - 3))
    // (t5), s = 7 unused.  inverse = true: the inverses at the same places, the untwist times N^-1.
    std::vector<uint64_t> half_tables(bool inverse = false) const {
        std::vector<uint64_t> t(2 * (size_t)N, 0);
        if (N != 1024) return t;
        auto brv = [](uint32_t x, int bits) {
            uint32_t r = 0;
            for (int b = 0; b < bits; ++b) { r = (r << 1) | (x & 1); x >>= 1; }
            return r;
        };
        const uint64_t psi = inverse ? gl::inv(psi_rev[N / 2]) : psi_rev[N / 2];
        const uint64_t w = gl::pow(psi, 32);
        const uint64_t scale = inverse ? gl::inv(N) : 1;
        for (uint32_t q = 0; q < 2; ++q)
            for (uint32_t l = 0; l < 64; ++l) {
                const size_t base = (size_t)(64 * q + l) * 8;
                for (uint32_t k = 0; k < 8; ++k) {
                    const uint32_t j = 8 * q + (l >> 3), i = (l & 7) | (k << 3);
                    t[base + k] = gl::mul(gl::pow(gl::pow(psi, 2 * brv(j, 4) + 1), i), scale);
                }
                const uint32_t m = l & 7;
                t[N + base + 0] = gl::pow(w, 4 * brv(m, 3));
                for (uint32_t s2 = 1; s2 < 3; ++s2) t[N + base + s2] = gl::pow(w, 2 * brv(2 * m + s2 - 1, 4));
                for (uint32_t s2 = 3; s2 < 7; ++s2) t[N + base + s2] = gl::pow(w, brv(4 * m + s2 - 3, 5));
            }
        return t;
    }

    void forward(uint64_t* a) const {
        for (uint32_t m = 1, t = N / 2; m < N; m <<= 1, t >>= 1)
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t w = psi_rev[m + i];
                uint64_t* p = a + 2 * i * t;
                for (uint32_t j = 0; j < t; ++j) {
                    const uint64_t u = p[j], v = gl::mul(p[j + t], w);
                    p[j] = gl::add(u, v);
                    p[j + t] = gl::sub(u, v);
                }
            }
    }

    // Inverse of forward(): bit-reversed order in, natural order out (Gentleman-Sande with
    // psi_inv_rev, then the factor N^-1).
    void inverse(uint64_t* a) const {
        for (uint32_t m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1)
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t w = psi_inv_rev[m + i];
                uint64_t* p = a + 2 * i * t;
                for (uint32_t j = 0; j < t; ++j) {
                    const uint64_t u = p[j], v = p[j + t];
                    p[j] = gl::add(u, v);
                    p[j + t] = gl::mul(gl::sub(u, v), w);
                }
            }
        const uint64_t ninv = gl::inv(N);
        for (uint32_t i = 0; i < N; ++i) a[i] = gl::mul(a[i], ninv);
    }
};

}  // namespace tp
