// CGGI / BMMP blind-rotation kernel for gfx950 (shared by the translation units that
// instantiate it for two and for one wave per SIMD).
#pragma once
#include <hip/hip_runtime.h>
#include "gl64.hpp"
#include "ntt_wave.hpp"
#include "kernels.hpp"

namespace tp {
// Position of NTT index 16 L + k inside a device key row: pairs (2c, 2c+1) of lane L are
// adjacent and lanes are contiguous, so a wave reads one row with 16-byte coalesced loads.
template <int LOGN>
__device__ __forceinline__ int key_pos(int L, int k) {
    return ((k >> 1) * (2 * Shape<LOGN>::P)) + 2 * L + (k & 1);
}

template <int LOGN>
__device__ __forceinline__ void load_key_row(uint64_t (&kv)[kE], const uint64_t* __restrict__ row, int L) {
#pragma unroll
    for (int c = 0; c < kE / 2; ++c) {
        const U64x2 v = ld_pair(row, c * Shape<LOGN>::P + L);
        kv[2 * c] = v.x;
        kv[2 * c + 1] = v.y;
    }
}

// ------------------------------------------------------------------ blind rotation

// Two ciphertexts' waves share a SIMD, and the issue arbiter favours the older wave of equal
// priority: left alone, the older one runs ahead at nearly one wave's issue rate and the younger
// then finishes alone.  Cycling each wave's priority with its step index (3, 2, 1, 0) hands the
// arbiter to the wave that has fallen behind on most steps, so the pair stays level and finishes
// together (n=630, N=1024, 4096 ciphertexts: 88.2 -> 81.3 ms; results unchanged, priority only
// orders issue).
__device__ __forceinline__ void step_priority(int i) {
    switch (i & 3) {
        case 0: __builtin_amdgcn_s_setprio(3); break;
        case 1: __builtin_amdgcn_s_setprio(2); break;
        case 2: __builtin_amdgcn_s_setprio(1); break;
        default: __builtin_amdgcn_s_setprio(0); break;
    }
}

// acc_h[k] = coefficient idx1(L,k) of X^e * src(X) (negacyclic), e in [0, 2N).
template <int LOGN>
__device__ __forceinline__ void rotate_from(uint64_t (&acc)[kE], const uint64_t* src, uint32_t e, int L,
                                            bool lds_src) {
    constexpr int N = Shape<LOGN>::N;
    const bool flip = e >= (uint32_t)N;
    const int r = (int)(e & (N - 1));
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        int j = idx1<LOGN>(L, k) - r;
        bool ng = flip;
        if (j < 0) { j += N; ng = !ng; }
        const uint64_t v = lds_src ? src[phys(j)] : src[j];
        acc[k] = ng ? gl::neg(v) : v;
    }
}

// Blind rotation, one ciphertext per work-group of P = N/16 lanes (one wave at N = 1024).
// The accumulator is kept in the coefficient domain, in registers (lane L holds
// coefficients L + P k of both halves); every step is exact in Z_Q, so the result is
// bit-identical to the reference's NTT-domain formulation:
//
// CGGI binary (cggi_blind_rotation.cpp:313-386, mux_operator.cpp:30-152):
//   acc <- acc + INTT( sum_r NTT(udigit_r((X^-a_i - 1) acc)) * BRK_i[r] )   for a_i != 0
//   (the reference multiplies NTT(acc) by NTT(X^-a_i) - NTT(1) before its INTT; here the
//   monomial product is a signed rotation of the coefficient vector).
// CGGI ternary (cggi_blind_rotation.cpp:388-472, mux_operator.cpp:156-202) and BMMP
// (bmmp_blind_rotation.cpp:354-419, mux_operator.cpp:204-294):
//   acc <- acc + INTT( sum_r NTT(sdigit_r(acc)) * sum_c w_c * RGSW_c[r] )
//   with w = (X^-a - 1, X^a - 1) (ternary) or (X^a0 - 1, X^a1 - 1, X^(a0+a1) - X^a0 - X^a1 + 1)
//   (BMMP); sum_c w_c (RGSW_c [x] acc) equals the reference's per-control products, weighted
//   and summed, in Z_Q.
// WPE = waves per SIMD the register budget is sized for: 2 caps the kernel at 256 VGPRs (two
// ciphertexts' waves interleave on a SIMD), 1 allows 512 (no spills for the ternary/BMMP MAC).
template <int LOGN, int LD, int METHOD, int WPE>
__global__ void __launch_bounds__(Shape<LOGN>::P) __attribute__((amdgpu_waves_per_eu(WPE)))
br_kernel(BrArgs A) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    constexpr int NC = METHOD == 0 ? 1 : (METHOD == 1 ? 2 : 3);  // RGSWs per step
    // binary: transpose buffer and parked digit source side by side, so that together they hold
    // the 2N-word negacyclic extension the rotation reads (below)
    __shared__ uint64_t lds[METHOD == 0 ? 2 * S::LDS_WORDS : S::LDS_WORDS];
    uint64_t* const buf = lds;
    uint64_t* const src = METHOD == 0 ? lds + S::LDS_WORDS : lds;
    const int L = threadIdx.x;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    const size_t rgsw_words = (size_t)4 * N * LD;

    // first-level constants of the signed digits: offsets 2^(b-1) and `least` (the last digit)
    uint64_t kp_n = 0, km_n = 0, kp_l = 0, km_l = 0;
    if constexpr (METHOD != 0) {
        signed_first_consts(1ULL << (A.base_log2 - 1), kp_n, km_n);
        signed_first_consts(A.least, kp_l, km_l);
    }

    // N = 2048 at two waves per SIMD: a SIMD holds waves of two work-groups, each of whose two waves
    // meet at LDS barriers in every transform.  Half of the work-groups issue at a higher priority, so
    // the two work-groups of a SIMD drift apart in phase instead of reaching their barriers together
    // (5,180 ternary ciphertexts, n = 800: 783 -> 748 ms; priority orders issue, results unchanged).
    if constexpr (S::WAVES == 2 && WPE == 2)
        if ((blockIdx.x >> 9) & 1) __builtin_amdgcn_s_setprio(1);
    for (int ct = blockIdx.x; ct < A.count; ct += gridDim.x) {
        const int slot = A.active ? (int)A.active[ct] : ct;
        const uint64_t* lwe = A.lwe + (size_t)slot * (A.n + 1);
        const uint64_t* tv = A.tv + (A.tv_index ? (size_t)A.tv_index[slot] * 2 * N : (size_t)slot * A.tv_stride);
        const int total = METHOD == 2 ? A.n / 2 : A.n;
        const int i0 = A.slot_step ? (int)A.slot_step[slot] : 0;
        const int i1 = (A.chunk_steps > 0 && i0 + A.chunk_steps < total) ? i0 + A.chunk_steps : total;
        uint64_t acc[2][kE];
        if (i0 == 0) {
            const uint32_t bexp = (uint32_t)lwe[A.n];
            rotate_from<LOGN>(acc[0], tv, bexp, L, false);
            rotate_from<LOGN>(acc[1], tv + N, bexp, L, false);
        } else {
            const uint64_t* st = A.state + (size_t)slot * 2 * N;
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                acc[0][k] = st[idx1<LOGN>(L, k)];
                acc[1][k] = st[N + idx1<LOGN>(L, k)];
            }
        }

        for (int i = i0; i < i1; ++i) {
            // monomial exponents of this step (NTT(X^e) rows of the table, e in [0, 2N))
            uint32_t e0 = 0, e1 = 0, e2 = 0;
            if constexpr (METHOD == 0 || METHOD == 1) {
                const uint32_t ai = (uint32_t)lwe[i];
                if (ai == 0) continue;
                e0 = 2 * N - ai;
                e1 = ai;
            } else {
                e0 = (2 * N - (uint32_t)lwe[2 * i]) & (2 * N - 1);
                e1 = (2 * N - (uint32_t)lwe[2 * i + 1]) & (2 * N - 1);
                e2 = (e0 + e1) & (2 * N - 1);
            }
            const uint64_t* key = A.brk + (size_t)i * NC * rgsw_words;
            // (N = 1024: one wave per polynomial; at N = 2048 a polynomial's two waves meet at LDS
            // barriers and the ternary uneven batch ran 6% slower with it)
            if constexpr (S::WAVES == 1) step_priority(i);
            // binary: the 2 LD products of each output coefficient, summed in lazy words; ternary and
            // BMMP: one such sum per control (RGSW_c), weighted once the 2 LD rows are summed
            uint64_t out[2][kE];
            uint64_t outc[METHOD == 0 ? 1 : NC][2][kE];
#pragma unroll
            for (int k = 0; k < kE; ++k) out[0][k] = out[1][k] = 0;
            if constexpr (METHOD != 0) {
#pragma unroll
                for (int r = 0; r < NC; ++r)
#pragma unroll
                    for (int k = 0; k < kE; ++k) outc[r][0][k] = outc[r][1][k] = 0;
            }

#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                if constexpr (METHOD == 0) {
                    // the polynomial the digits are taken from, X^e0 acc_h - acc_h, is formed once
                    // per half and parked in LDS, where every digit re-reads it (kept out of
                    // registers).  With e' = e0 mod N and E = (-acc_h | acc_h) (e0 < N) or
                    // (acc_h | -acc_h) (e0 >= N, X^e0 = -X^e'), coefficient i of X^e0 acc_h is
                    // E[i + N - e']: lane L reads one base address plus constant offsets.
                    constexpr uint32_t kStride = S::P + S::P / 16;  // phys(idx1(L, k + 1)) - phys(idx1(L, k))
                    constexpr uint32_t kHi = N + N / 16;             // phys(j + N) - phys(j)
                    uint64_t na[kE], zero[kE];
#pragma unroll
                    for (int k = 0; k < kE; ++k) zero[k] = 0;
                    glv::sub_lc_g<kE>(zero, acc[0], na);  // -acc_h, canonical
                    int Lh = L;  // re-read per half: the addresses are formed here
                    asm volatile("" : "+v"(Lh));
                    const uint32_t pl = (uint32_t)Lh + ((uint32_t)Lh >> 4);  // phys(idx1(L, 0))
                    // which half of E receives acc_h and which -acc_h: a wave-uniform choice of
                    // base address, not a per-lane select of values
                    const uint32_t off_a = e0 >= (uint32_t)N ? 0u : kHi, off_n = kHi - off_a;
                    poly_sync<S::WAVES>();
#pragma unroll
                    for (int k = 0; k < kE; ++k) {
                        lds[pl + off_a + k * kStride] = acc[0][k];
                        lds[pl + off_n + k * kStride] = na[k];
                    }
                    poly_sync<S::WAVES>();
                    const uint32_t j0 = (uint32_t)Lh + (uint32_t)N - (e0 & (uint32_t)(N - 1));
                    const uint32_t p0 = j0 + (j0 >> 4);
                    uint64_t u[kE], t[kE];
#pragma unroll
                    for (int k = 0; k < kE; ++k) u[k] = lds[p0 + k * kStride];
                    glv::sub_lc_g<kE>(u, acc[0], t);  // both canonical: canonical
                    poly_sync<S::WAVES>();
#pragma unroll
                    for (int k = 0; k < kE; ++k) src[pl + k * kStride] = t[k];
                    poly_sync<S::WAVES>();
                }
                uint64_t xc[METHOD == 0 ? 1 : kE];  // x + C mod Q of the half (signed digits)
                if constexpr (METHOD != 0) {
                    uint64_t cv[kE];
#pragma unroll
                    for (int k = 0; k < kE; ++k) cv[k] = A.corrector;
                    glv::add_cc_g<kE>(acc[0], cv, xc, glv::qm1_const());
                }
#pragma unroll 1
                for (int j = 0; j < LD; ++j) {
                    uint64_t d[kE];
                    if constexpr (METHOD == 0) {
                        const int shift = A.first_shift - j * A.base_log2;
                        // the lane index re-read per digit: its LDS address is formed here, not
                        // held in a register across the digit loop.  Coefficient L + P k sits at
                        // word (L + L / 16) + k (P + P / 16): one base plus constant offsets.
                        uint32_t Lj = (uint32_t)L;
                        asm volatile("" : "+v"(Lj));
                        const uint32_t base = Lj + (Lj >> 4);
                        constexpr uint32_t kStride = S::P + S::P / 16;
                        const uint32_t* src32 = (const uint32_t*)src;
                        if (shift >= 32 && A.base_log2 < 32) {
                            // the digit lies in the high word: one 32-bit read and a bit-field extract
                            // (bits past the word's top are zero in both forms)
#pragma unroll
                            for (int k = 0; k < kE; ++k)
                                d[k] = __builtin_amdgcn_ubfe(src32[2 * (base + k * kStride) + 1], (uint32_t)(shift - 32),
                                                             (uint32_t)A.base_log2);
                        } else if (shift + A.base_log2 <= 32 && A.base_log2 < 32) {
#pragma unroll
                            for (int k = 0; k < kE; ++k)
                                d[k] = __builtin_amdgcn_ubfe(src32[2 * (base + k * kStride)], (uint32_t)shift,
                                                             (uint32_t)A.base_log2);
                        } else {
#pragma unroll
                            for (int k = 0; k < kE; ++k) d[k] = (src[base + k * kStride] >> shift) & mask;
                        }
                        // the first transform level on the digits themselves: cheap while they are small
                        if (A.base_log2 <= 16) {
#pragma unroll
                            for (int k = 0; k < kE / 2; ++k) ct_bfly_small_first(d[k], d[k + kE / 2]);
                        } else {
#pragma unroll
                            for (int k = 0; k < kE / 2; ++k) ct_bfly_pow2(d[k], d[k + kE / 2], gl::kFwdPow2[1]);
                        }
                    } else {
                        // signed digits of x + C mod Q (gadget_decomp.h:169-198), x + C formed once per half
                        const bool last = (j == A.max_digits - 1);
                        const int shift = last ? 0 : 64 - (j + 1) * A.base_log2;
                        const uint64_t dm = last ? (2 * A.least - 1) : mask;
                        const uint64_t delta = last ? A.least : (1ULL << (A.base_log2 - 1));
                        if (shift >= 32 && A.base_log2 < 32) {
#pragma unroll
                            for (int k = 0; k < kE; ++k)
                                d[k] = __builtin_amdgcn_ubfe(glv::hi32(xc[k]), (uint32_t)(shift - 32), (uint32_t)A.base_log2);
                        } else {
#pragma unroll
                            for (int k = 0; k < kE; ++k) d[k] = (xc[k] >> shift) & dm;
                        }
                        // the first transform level folds the offset in while the digits are small
                        if (dm < (1ULL << 16)) {
                            const uint64_t kp = last ? kp_l : kp_n, km = last ? km_l : km_n;
#pragma unroll
                            for (int k = 0; k < kE / 2; ++k) ct_bfly_signed_first(d[k], d[k + kE / 2], kp, km);
                        } else {
#pragma unroll
                            for (int k = 0; k < kE; ++k) d[k] = gl::sub(d[k], delta);
#pragma unroll
                            for (int k = 0; k < kE / 2; ++k) ct_bfly_pow2(d[k], d[k + kE / 2], gl::kFwdPow2[1]);
                        }
                    }
                    const size_t row = (size_t)(h * LD + j) * 2 * N;
                    if constexpr (METHOD == 0) {
                        // key pairs of both rows, two at a time: the first two are requested during
                        // the transform's last level, each next two while the current two are used
                        const uint64_t* ka = opaque(key + row);
                        const uint64_t* kb = opaque(key + row + N);
                        U64x2 va[2], vb[2];
                        auto load_group = [&](U64x2 (&ga)[2], U64x2 (&gb)[2], int c0) __attribute__((always_inline)) {
#pragma unroll
                            for (int q = 0; q < 2; ++q) {
                                ga[q] = ld_pair(ka, (c0 + q) * S::P + L);
                                gb[q] = ld_pair(kb, (c0 + q) * S::P + L);
                            }
                            __builtin_amdgcn_sched_barrier(0);
                        };
                        ntt_fwd<LOGN, 1>(d, buf, A.tw, L, [&]() __attribute__((always_inline)) {
                            load_group(va, vb, 0);
                        });
                        static_for<4>([&](auto G) __attribute__((always_inline)) {
                            constexpr int c0 = 2 * G;
                            U64x2 na[2], nb[2];
                            if constexpr (G < 3) load_group(na, nb, c0 + 2);
                            // coefficients 2 c0 .. 2 c0 + 3 against both key rows, as two groups of 4
                            const uint64_t dv[4] = {d[2 * c0], d[2 * c0 + 1], d[2 * c0 + 2], d[2 * c0 + 3]};
                            const uint64_t kav[4] = {va[0].x, va[0].y, va[1].x, va[1].y};
                            const uint64_t kbv[4] = {vb[0].x, vb[0].y, vb[1].x, vb[1].y};
                            uint64_t g0[4], g1[4];
#pragma unroll
                            for (int m = 0; m < 4; ++m) {
                                g0[m] = out[0][2 * c0 + m];
                                g1[m] = out[1][2 * c0 + m];
                            }
                            {
                                // both key rows as one group of eight members (each carry is read
                                // eight instructions after it is written)
                                uint64_t g8[8], d8[8], k8[8];
#pragma unroll
                                for (int m = 0; m < 4; ++m) {
                                    g8[m] = g0[m];
                                    g8[m + 4] = g1[m];
                                    d8[m] = d8[m + 4] = dv[m];
                                    k8[m] = kav[m];
                                    k8[m + 4] = kbv[m];
                                }
                                glv::mac_lazy_g<8>(g8, d8, k8);
#pragma unroll
                                for (int m = 0; m < 4; ++m) {
                                    g0[m] = g8[m];
                                    g1[m] = g8[m + 4];
                                }
                            }
#pragma unroll
                            for (int m = 0; m < 4; ++m) {
                                out[0][2 * c0 + m] = g0[m];
                                out[1][2 * c0 + m] = g1[m];
                            }
                            if constexpr (G < 3) {
#pragma unroll
                                for (int q = 0; q < 2; ++q) {
                                    va[q] = na[q];
                                    vb[q] = nb[q];
                                }
                            }
                        });
                    } else {
                        // the reference's MultiMux order (mux_operator.cpp:156-202,264-291): every control
                        // c sums d_row * RGSW_c[row] over the 2 LD rows, and its weight multiplies the
                        // sum once per step (below); key pairs of pair c + 1 are requested while pair c
                        // is used (the first ones during the transform's last level)
                        const uint64_t* krow = opaque(key + row);
                        struct PairLoads {
                            U64x2 k[NC][2];  // key pairs of RGSW_c, output halves 0 and 1
                        };
                        auto load_pair = [&](PairLoads& pl, int c) __attribute__((always_inline)) {
#pragma unroll
                            for (int r = 0; r < NC; ++r)
#pragma unroll
                                for (int p = 0; p < 2; ++p)
                                    pl.k[r][p] = ld_pair(krow + r * rgsw_words + p * N, c * S::P + L);
                            __builtin_amdgcn_sched_barrier(0);
                        };
                        PairLoads cur;
                        ntt_fwd<LOGN, 1>(d, buf, A.tw, L, [&]() __attribute__((always_inline)) {
                            load_pair(cur, 0);
                        });
                        static_for<kE / 2>([&](auto C) __attribute__((always_inline)) {
                            constexpr int c = C;
                            PairLoads nxt;
                            if constexpr (c + 1 < kE / 2) load_pair(nxt, c + 1);
                            // members (control, half, coefficient 2c or 2c + 1), in groups of at most 8
                            static_for<(NC + 1) / 2>([&](auto G) __attribute__((always_inline)) {
                                constexpr int r0 = 2 * G;
                                constexpr int nr = (r0 + 2 <= NC) ? 2 : 1;
                                constexpr int M = 4 * nr;
                                uint64_t g[M], dd[M], kk[M];
                                static_for<M>([&](auto Mi) __attribute__((always_inline)) {
                                    constexpr int m = Mi;
                                    constexpr int r = r0 + m / 4, p = (m / 2) % 2, q = m % 2;
                                    g[m] = outc[r][p][2 * c + q];
                                    dd[m] = d[2 * c + q];
                                    kk[m] = q ? cur.k[r][p].y : cur.k[r][p].x;
                                });
                                glv::mac_lazy_g<M>(g, dd, kk);
                                static_for<M>([&](auto Mi) __attribute__((always_inline)) {
                                    constexpr int m = Mi;
                                    constexpr int r = r0 + m / 4, p = (m / 2) % 2, q = m % 2;
                                    outc[r][p][2 * c + q] = g[m];
                                });
                            });
                            if constexpr (c + 1 < kE / 2) cur = nxt;
                        });
                    }
                }
                // the half being decomposed is always acc[0]: exchange the halves in registers
#pragma unroll
                for (int k = 0; k < kE; ++k) {
                    const uint64_t t = acc[0][k];
                    acc[0][k] = acc[1][k];
                    acc[1][k] = t;
                }
            }
            if constexpr (METHOD != 0) {
                // out = sum_c w_c sum_rows: w = (X^-a - 1, X^a - 1) for ternary
                // (cggi_blind_rotation.cpp:388-472), (X^a0 - 1, X^a1 - 1, X^(a0+a1) - X^a0 - X^a1 + 1)
                // for BMMP (bmmp_blind_rotation.cpp:280-323); NTT(X^e) rows come from the monomial
                // table, X^e = -X^(e-N) for e >= N; the rows of pair c + 1 are requested while pair c
                // is used
                constexpr int NQ = NC == 3 ? 3 : 2;
                const uint64_t* mono = opaque_after(A.mono, outc[0][0][0]);
                const uint32_t es[3] = {e0, e1, e2};
                auto load_m = [&](U64x2 (&m)[NQ], int c) __attribute__((always_inline)) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) m[q] = ld_pair(mono + (size_t)(es[q] & (N - 1)) * N, c * S::P + L);
                    __builtin_amdgcn_sched_barrier(0);
                };
                auto mono_val = [&](const U64x2& v, uint32_t e) __attribute__((always_inline)) {
                    return e >= (uint32_t)N ? U64x2{gl::neg(v.x), gl::neg(v.y)} : v;
                };
                U64x2 mc[NQ];
                load_m(mc, 0);
                static_for<kE / 2>([&](auto C) __attribute__((always_inline)) {
                    constexpr int c = C;
                    U64x2 mn[NQ];
                    if constexpr (c + 1 < kE / 2) load_m(mn, c + 1);
                    const U64x2 m0 = mono_val(mc[0], e0), m1 = mono_val(mc[1], e1);
                    U64x2 w[NC];
                    w[0] = U64x2{gl::sub(m0.x, 1), gl::sub(m0.y, 1)};  // X^e0 - 1
                    w[1] = U64x2{gl::sub(m1.x, 1), gl::sub(m1.y, 1)};  // X^e1 - 1
                    if constexpr (NC == 3) {
                        const U64x2 m2 = mono_val(mc[NQ - 1], e2);
                        w[NC - 1] = U64x2{gl::add(gl::sub(gl::sub(m2.x, m0.x), m1.x), 1),
                                          gl::add(gl::sub(gl::sub(m2.y, m0.y), m1.y), 1)};
                    }
                    // members (half, coefficient 2c or 2c + 1), one multiply-accumulate per control
                    uint64_t o[4] = {0, 0, 0, 0};
                    static_for<NC>([&](auto R) __attribute__((always_inline)) {
                        constexpr int r = R;
                        uint64_t ww[4], vv[4];
                        static_for<4>([&](auto Mi) __attribute__((always_inline)) {
                            constexpr int m = Mi;
                            constexpr int p = m / 2, q = m % 2;
                            ww[m] = q ? w[r].y : w[r].x;
                            vv[m] = outc[r][p][2 * c + q];
                        });
                        glv::mac_lazy_g<4>(o, ww, vv);
                    });
                    static_for<4>([&](auto Mi) __attribute__((always_inline)) {
                        constexpr int m = Mi;
                        out[m / 2][2 * c + m % 2] = o[m];
                    });
                    if constexpr (c + 1 < kE / 2) {
#pragma unroll
                        for (int q = 0; q < NQ; ++q) mc[q] = mn[q];
                    }
                });
            }
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                if constexpr (METHOD == 0) {
                    // inverse transform (canonical out) -> acc + result, canonical
                    const uint64_t qm1 = glv::qm1_const();
                    uint64_t t[kE];
                    ntt_inv<LOGN>(out[0], buf, A.itw, L);
                    glv::add_cc_g<kE>(out[0], acc[0], t, qm1);
#pragma unroll
                    for (int k = 0; k < kE; ++k) {
                        acc[0][k] = acc[1][k];
                        acc[1][k] = t[k];
                        out[0][k] = out[1][k];
                    }
                } else {
                    ntt_inv<LOGN>(out[0], buf, A.itw, L);
#pragma unroll
                    for (int k = 0; k < kE; ++k) {
                        const uint64_t t = gl::canon(gl::add_lc(out[0][k], acc[0][k]));
                        acc[0][k] = acc[1][k];
                        acc[1][k] = t;
                        out[0][k] = out[1][k];
                    }
                }
            }
        }

        if (A.slot_step && L == 0) A.slot_step[slot] = (uint32_t)i1;
        if (i1 < total) {
            // chunk boundary: park the accumulator for the next launch
            uint64_t* st = A.state + (size_t)slot * 2 * N;
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                st[idx1<LOGN>(L, k)] = acc[0][k];
                st[N + idx1<LOGN>(L, k)] = acc[1][k];
            }
            poly_sync<S::WAVES>();
            continue;
        }

        // outputs: sample extraction at index 0 (sample_extraction.h:20-36) and optional RLWE
        if (A.out_lwe) {
            uint64_t* o = A.out_lwe + (size_t)slot * (N + 1);
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                const int j = idx1<LOGN>(L, k);
                o[j == 0 ? 0 : N - j] = j == 0 ? acc[0][k] : gl::neg(acc[0][k]);
                if (j == 0) o[N] = acc[1][k];
            }
        }
        if (A.out_rlwe) {
            // one half at a time through a reference, so acc is never indexed at run time
            auto emit = [&](uint64_t (&x)[kE], uint64_t* o) __attribute__((always_inline)) {
                if (A.out_rlwe_ntt) {
                    ntt_fwd<LOGN>(x, buf, A.tw, L);
#pragma unroll
                    for (int k = 0; k < kE; ++k) o[idx3(L, k)] = gl::canon(x[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < kE; ++k) o[idx1<LOGN>(L, k)] = x[k];
                }
            };
            emit(acc[0], A.out_rlwe + (size_t)slot * 2 * N);
            emit(acc[1], A.out_rlwe + (size_t)slot * 2 * N + N);
        }
        poly_sync<S::WAVES>();
    }
}


template <int LOGN, int LD, int METHOD, int WPE>
inline void launch_br(const BrArgs& A, int grid, hipStream_t s) {
    hipLaunchKernelGGL((br_kernel<LOGN, LD, METHOD, WPE>), dim3(grid), dim3(Shape<LOGN>::P), 0, s, A);
}

}  // namespace tp
