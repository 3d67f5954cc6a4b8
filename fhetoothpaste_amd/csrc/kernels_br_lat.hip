// Digit-parallel blind rotation at N = 1024: 2 l waves per ciphertext (the one-bootstrap-per-call
// path of BlindRotator::BlindRotate, lib/include/interfaces/blindrotation_operator.h:29-40, and
// launches of at most a few ciphertexts per CU).  br_kernel runs a ciphertext's 2 l digit products
// one after the other in one wave; here one work-group of 2 l waves runs them side by side: wave
// (h, j) forms digit j of accumulator half h, its transform and its products with key row h l + j
// (for ternary and BMMP with the combined key), and writes its two partial products to LDS; all
// waves then add the 2 l partial products, a slice of the words each, and waves 0 and 1
// inverse-transform the sums of output half 0 and 1 and update that half of the accumulator, which
// lives in LDS.  The arithmetic is br_kernel's, exact in Z_Q, so the result is bit-identical.
// Work-groups stride over the ciphertexts of the launch with br_kernel's slot, test-vector and
// chunking rules (BrArgs).
#include <hip/hip_runtime.h>
#include "br_kernel.hpp"

namespace tp {

namespace {

constexpr int kLatLogN = 10;

template <int LD, int METHOD>
__global__ void __launch_bounds__(2 * LD * 64) br_latency_kernel(BrArgs A) {
    constexpr int LOGN = kLatLogN;
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    static_assert(S::P == 64, "one wave per polynomial");
    constexpr int NC = METHOD == 2 ? 3 : (METHOD == 1 ? 2 : 1);
    constexpr int G = 2 * LD;
    // accL: ternary/BMMP: the accumulator halves (coefficients, phys(idx1) positions); binary: the
    // halves' digit sources X^e0 acc_h - acc_h (the accumulator itself stays in the registers of
    // waves 0 and 1).  red: the partial products of every wave (NTT domain, idx3 positions; the
    // first words of a wave's region are its transpose buffer in the product phase, and in the
    // binary step prologue red[2 + p] holds wave p's negacyclic extension of acc_p).
    __shared__ uint64_t accL[2][S::LDS_WORDS];
    __shared__ uint64_t red[G][2 * N];
    static_assert(S::LDS_WORDS <= 2 * N, "transpose buffer inside the product region");
    static_assert(G >= 4, "binary prologue uses red[2] and red[3]");

    const int w = threadIdx.x >> 6;
    const int h = w / LD, j = w % LD;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    const size_t rgsw_words = (size_t)4 * N * LD;
    const int total = METHOD == 2 ? A.n / 2 : A.n;

    for (int ct = blockIdx.x; ct < A.count; ct += gridDim.x) {
        // the lane index is re-read per ciphertext so that its LDS and key addresses are formed
        // inside the loop, not held in registers across it
        int L = threadIdx.x & 63;
        asm volatile("" : "+v"(L));
        const int slot = A.active ? (int)A.active[ct] : ct;
        const uint64_t* lwe = A.lwe + (size_t)slot * (A.n + 1);
        const uint64_t* tv =
            A.tv + (A.tv_index ? (size_t)A.tv_index[slot] * 2 * N : (size_t)slot * A.tv_stride);
        const int i0 = A.slot_step ? (int)A.slot_step[slot] : 0;
        const int i1 = (A.chunk_steps > 0 && i0 + A.chunk_steps < total) ? i0 + A.chunk_steps : total;
        uint64_t accR[kE];  // binary: wave p < 2 holds accumulator half p (coefficients idx1(L, k))
        if (w < 2) {
            if (i0 == 0) {
                rotate_from<LOGN>(accR, tv + w * N, (uint32_t)lwe[A.n], L, false);
            } else {
                const uint64_t* st = A.state + (size_t)slot * 2 * N + (size_t)w * N;
#pragma unroll
                for (int k = 0; k < kE; ++k) accR[k] = st[idx1<LOGN>(L, k)];
            }
            if constexpr (METHOD != 0) {
#pragma unroll
                for (int k = 0; k < kE; ++k) accL[w][phys(idx1<LOGN>(L, k))] = accR[k];
            }
        }
        __syncthreads();

        for (int i = i0; i < i1; ++i) {
            uint32_t e0 = 0, e1 = 0, e2 = 0;
            if constexpr (METHOD == 0 || METHOD == 1) {
                const uint32_t ai = (uint32_t)lwe[i];
                if (ai == 0) continue;  // uniform over the work-group
                e0 = 2 * N - ai;
                e1 = ai;
            } else {
                e0 = (2 * N - (uint32_t)lwe[2 * i]) & (2 * N - 1);
                e1 = (2 * N - (uint32_t)lwe[2 * i + 1]) & (2 * N - 1);
                e2 = (e0 + e1) & (2 * N - 1);
            }
            const uint64_t* key = A.brk + (size_t)i * NC * rgsw_words;
            uint64_t* tb = red[w];

            if constexpr (METHOD == 0) {
                // step prologue: wave p parks X^e0 acc_p - acc_p in accL[p].  With e' = e0 mod N and
                // E = (-acc_p | acc_p) (e0 < N) or (acc_p | -acc_p) (e0 >= N) in red[2 + p] (unpadded:
                // consecutive lanes touch consecutive words), coefficient i of X^e0 acc_p is
                // E[i + N - e'].
                if (w < 2) {
                    uint64_t na[kE], zero[kE], u[kE], t[kE];
#pragma unroll
                    for (int k = 0; k < kE; ++k) zero[k] = 0;
                    glv::sub_lc_g<kE>(zero, accR, na);  // -acc_p, canonical
                    uint64_t* E = red[2 + w];
                    const uint32_t off_a = e0 >= (uint32_t)N ? 0u : (uint32_t)N, off_n = (uint32_t)N - off_a;
#pragma unroll
                    for (int k = 0; k < kE; ++k) {
                        E[off_a + (uint32_t)idx1<LOGN>(L, k)] = accR[k];
                        E[off_n + (uint32_t)idx1<LOGN>(L, k)] = na[k];
                    }
                    poly_sync<S::WAVES>();
                    const uint32_t j0 = (uint32_t)L + (uint32_t)N - (e0 & (uint32_t)(N - 1));
#pragma unroll
                    for (int k = 0; k < kE; ++k) u[k] = E[j0 + (uint32_t)(k * S::P)];
                    glv::sub_lc_g<kE>(u, accR, t);  // both canonical: canonical
#pragma unroll
                    for (int k = 0; k < kE; ++k) accL[w][phys(idx1<LOGN>(L, k))] = t[k];
                }
                __syncthreads();
            }

            // product phase: digit j of half h
            uint64_t d[kE];
            if constexpr (METHOD == 0) {
                const int shift = A.first_shift - j * A.base_log2;
                const uint32_t* src32 = (const uint32_t*)accL[h];
                const uint32_t base = (uint32_t)L + ((uint32_t)L >> 4);
                constexpr uint32_t kStride = S::P + S::P / 16;
                if (shift >= 32 && A.base_log2 < 32) {
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
                    for (int k = 0; k < kE; ++k) d[k] = (accL[h][base + k * kStride] >> shift) & mask;
                }
                if (A.base_log2 <= 16) {
#pragma unroll
                    for (int k = 0; k < kE / 2; ++k) ct_bfly_small_first(d[k], d[k + kE / 2]);
                } else {
#pragma unroll
                    for (int k = 0; k < kE / 2; ++k) ct_bfly_pow2(d[k], d[k + kE / 2], gl::kFwdPow2[1]);
                }
            } else {
                const bool last = (j == A.max_digits - 1);
                const int shift = last ? 0 : 64 - (j + 1) * A.base_log2;
                const uint64_t dm = last ? (2 * A.least - 1) : mask;
                const uint64_t delta = last ? A.least : (1ULL << (A.base_log2 - 1));
#pragma unroll
                for (int k = 0; k < kE; ++k)
                    d[k] = (gl::add(accL[h][phys(idx1<LOGN>(L, k))], A.corrector) >> shift) & dm;
                if (dm < (1ULL << 16)) {
                    uint64_t kp, km;
                    signed_first_consts(delta, kp, km);
#pragma unroll
                    for (int k = 0; k < kE / 2; ++k) ct_bfly_signed_first(d[k], d[k + kE / 2], kp, km);
                } else {
#pragma unroll
                    for (int k = 0; k < kE; ++k) d[k] = gl::sub(d[k], delta);
#pragma unroll
                    for (int k = 0; k < kE / 2; ++k) ct_bfly_pow2(d[k], d[k + kE / 2], gl::kFwdPow2[1]);
                }
            }
            uint64_t out[2][kE];
            const size_t row = (size_t)(h * LD + j) * 2 * N;
            if constexpr (METHOD == 0) {
                // key pairs of both rows, two at a time: the first two requested during the
                // transform's last level, each next two while the current two are used
                const uint64_t* ka = key + row;
                const uint64_t* kb = key + row + N;
                U64x2 va[2], vb[2];
                auto load_group = [&](U64x2 (&ga)[2], U64x2 (&gb)[2], int c0) __attribute__((always_inline)) {
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        ga[q] = ld_pair(ka, (c0 + q) * S::P + L);
                        gb[q] = ld_pair(kb, (c0 + q) * S::P + L);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                };
                ntt_fwd<LOGN, 1>(d, tb, A.tw, L, [&]() __attribute__((always_inline)) { load_group(va, vb, 0); });
                static_for<4>([&](auto G) __attribute__((always_inline)) {
                    constexpr int c0 = 2 * G;
                    U64x2 na[2], nb[2];
                    if constexpr (G < 3) load_group(na, nb, c0 + 2);
                    // coefficients 2 c0 .. 2 c0 + 3 against both key rows, one group of eight
                    uint64_t g8[8], d8[8], k8[8];
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        g8[m] = g8[m + 4] = 0;
                        d8[m] = d8[m + 4] = d[2 * c0 + m];
                    }
                    k8[0] = va[0].x; k8[1] = va[0].y; k8[2] = va[1].x; k8[3] = va[1].y;
                    k8[4] = vb[0].x; k8[5] = vb[0].y; k8[6] = vb[1].x; k8[7] = vb[1].y;
                    glv::mac_lazy_g<8>(g8, d8, k8);
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        out[0][2 * c0 + m] = g8[m];
                        out[1][2 * c0 + m] = g8[m + 4];
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
                // combined key, pair by pair; the monomial rows and key pairs of pair c + 1 are
                // requested while pair c is used (the first during the transform's last level)
                const uint64_t* krow = key + row;
                const uint64_t* mono = A.mono;
                struct PairLoads {
                    U64x2 m[3];
                    U64x2 k[NC][2];
                };
                auto load_pair = [&](PairLoads& pl, int c) __attribute__((always_inline)) {
                    const uint32_t es[3] = {e0, e1, e2};
#pragma unroll
                    for (int q = 0; q < (NC == 3 ? 3 : 2); ++q)
                        pl.m[q] = ld_pair(mono + (size_t)(es[q] & (N - 1)) * N, c * S::P + L);
#pragma unroll
                    for (int r = 0; r < NC; ++r)
#pragma unroll
                        for (int p = 0; p < 2; ++p) pl.k[r][p] = ld_pair(krow + r * rgsw_words + p * N, c * S::P + L);
                    __builtin_amdgcn_sched_barrier(0);
                };
                auto mono_val = [&](const U64x2& v, uint32_t e) __attribute__((always_inline)) {
                    return e >= (uint32_t)N ? U64x2{gl::neg(v.x), gl::neg(v.y)} : v;
                };
                PairLoads cur;
                ntt_fwd<LOGN, 1>(d, tb, A.tw, L, [&]() __attribute__((always_inline)) { load_pair(cur, 0); });
                static_for<kE / 2>([&](auto C) __attribute__((always_inline)) {
                    constexpr int c = C;
                    PairLoads nxt;
                    if constexpr (c + 1 < kE / 2) load_pair(nxt, c + 1);
                    const U64x2 m0 = mono_val(cur.m[0], e0), m1 = mono_val(cur.m[1], e1);
                    const U64x2 w0{gl::sub(m0.x, 1), gl::sub(m0.y, 1)};  // X^e0 - 1
                    const U64x2 w1{gl::sub(m1.x, 1), gl::sub(m1.y, 1)};  // X^e1 - 1
                    U64x2 w2{0, 0};
                    if constexpr (NC == 3) {
                        // X^(e0+e1) - X^e0 - X^e1 + 1 (bmmp_blind_rotation.cpp:280-323)
                        const U64x2 m2 = mono_val(cur.m[2], e2);
                        w2 = U64x2{gl::add(gl::sub(gl::sub(m2.x, m0.x), m1.x), 1),
                                   gl::add(gl::sub(gl::sub(m2.y, m0.y), m1.y), 1)};
                    }
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const U64x2 a0 = cur.k[0][p], a1 = cur.k[1][p];
                        uint64_t cx = gl::mac_lazy(a1.x, w1.x, gl::mul_lazy(a0.x, w0.x));
                        uint64_t cy = gl::mac_lazy(a1.y, w1.y, gl::mul_lazy(a0.y, w0.y));
                        if constexpr (NC == 3) {
                            const U64x2 a2 = cur.k[NC - 1][p];
                            cx = gl::mac_lazy(a2.x, w2.x, cx);
                            cy = gl::mac_lazy(a2.y, w2.y, cy);
                        }
                        out[p][2 * c] = gl::mul_lazy(d[2 * c], cx);
                        out[p][2 * c + 1] = gl::mul_lazy(d[2 * c + 1], cy);
                    }
                    if constexpr (c + 1 < kE / 2) cur = nxt;
                });
            }
            // the wave's transpose buffer is done with (its own barrier inside the transform)
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                tb[idx3(L, k)] = out[0][k];
                tb[N + idx3(L, k)] = out[1][k];
            }
            __syncthreads();

            // reduction phase: every wave sums the G partial products of a slice of the 2 N words
            // into the first wave's region (each word is read and written by one thread only)
            for (int x = threadIdx.x; x < 2 * N; x += G * 64) {
                uint64_t s = red[0][x];
#pragma unroll
                for (int g = 1; g < G; ++g) s = gl::add_ll(s, red[g][x]);
                red[0][x] = s;
            }
            __syncthreads();

            // wave p inverse-transforms the sum of output half p and updates that half
            if (w < 2) {
                uint64_t s[kE];
#pragma unroll
                for (int k = 0; k < kE; ++k) s[k] = red[0][w * N + idx3(L, k)];
                if constexpr (METHOD == 0) {
                    uint64_t t[kE];
                    ntt_inv<LOGN>(s, red[2 + w], A.itw, L);  // canonical out
                    glv::add_cc_g<kE>(s, accR, t, glv::qm1_const());
#pragma unroll
                    for (int k = 0; k < kE; ++k) accR[k] = t[k];
                } else {
                    uint64_t prev[kE];
#pragma unroll
                    for (int k = 0; k < kE; ++k) prev[k] = accL[w][phys(idx1<LOGN>(L, k))];
                    ntt_inv<LOGN>(s, accL[w], A.itw, L);
                    poly_sync<S::WAVES>();
#pragma unroll
                    for (int k = 0; k < kE; ++k)
                        accL[w][phys(idx1<LOGN>(L, k))] = gl::canon(gl::add_lc(s[k], prev[k]));
                }
            }
            __syncthreads();
        }

        if (A.slot_step && threadIdx.x == 0) A.slot_step[slot] = (uint32_t)i1;
        if (w < 2) {
            uint64_t x[kE];
#pragma unroll
            for (int k = 0; k < kE; ++k) x[k] = METHOD == 0 ? accR[k] : accL[w][phys(idx1<LOGN>(L, k))];
            if (i1 < total) {
                // chunk boundary: park the accumulator half for the next launch
                uint64_t* st = A.state + (size_t)slot * 2 * N + (size_t)w * N;
#pragma unroll
                for (int k = 0; k < kE; ++k) st[idx1<LOGN>(L, k)] = x[k];
            } else {
                // outputs (one half per wave): sample extraction at index 0 and optional RLWE
                if (A.out_lwe) {
                    uint64_t* o = A.out_lwe + (size_t)slot * (N + 1);
#pragma unroll
                    for (int k = 0; k < kE; ++k) {
                        const int jj = idx1<LOGN>(L, k);
                        if (w == 0) o[jj == 0 ? 0 : N - jj] = jj == 0 ? x[k] : gl::neg(x[k]);
                        else if (jj == 0) o[N] = x[k];
                    }
                }
                if (A.out_rlwe) {
                    uint64_t* o = A.out_rlwe + (size_t)slot * 2 * N + (size_t)w * N;
                    if (A.out_rlwe_ntt) {
                        poly_sync<S::WAVES>();
                        ntt_fwd<LOGN>(x, accL[w], A.tw, L);
#pragma unroll
                        for (int k = 0; k < kE; ++k) o[idx3(L, k)] = gl::canon(x[k]);
                    } else {
#pragma unroll
                        for (int k = 0; k < kE; ++k) o[idx1<LOGN>(L, k)] = x[k];
                    }
                }
            }
        }
        // the next ciphertext's prologue rewrites the accumulator halves
        __syncthreads();
    }
}

}  // namespace

// N = 1024, digits 2..4, A.count ciphertexts over `grid` work-groups of 2 * digits waves.
hipError_t launch_blind_rotate_latency(int method, int digits, const BrArgs& A, int grid, hipStream_t s) {
    if (grid <= 0) return hipErrorInvalidValue;
#define TP_LAT_CASE(LD, M) \
    if (digits == LD && method == M) { \
        hipLaunchKernelGGL((br_latency_kernel<LD, M>), dim3(grid), dim3(2 * LD * 64), 0, s, A); \
        return hipGetLastError(); \
    }
    TP_LAT_CASE(2, 0) TP_LAT_CASE(2, 1) TP_LAT_CASE(2, 2)
    TP_LAT_CASE(3, 0) TP_LAT_CASE(3, 1) TP_LAT_CASE(3, 2)
    TP_LAT_CASE(4, 0) TP_LAT_CASE(4, 1) TP_LAT_CASE(4, 2)
#undef TP_LAT_CASE
    return hipErrorInvalidValue;
}

}  // namespace tp
