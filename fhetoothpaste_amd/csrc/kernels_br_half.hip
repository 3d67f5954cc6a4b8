// Binary CGGI blind rotation at N = 1024 for one bootstrap per call (BlindRotator::BlindRotate,
// lib/include/interfaces/blindrotation_operator.h:29-40, cggi_blind_rotation.cpp:313-386) and for
// launches of a few ciphertexts per CU: one work-group of 4 l waves per ciphertext, every
// transform split over two waves (ntt_half.hpp), so that the 2 l digit transforms of a step run as
// three waves per SIMD and the two inverse transforms as one wave on each of the four SIMDs.
//   prologue  waves 0..3 (accumulator half o = w >> 1, index half w & 1; the accumulator lives in
//             LDS) park X^e0 acc_o - acc_o in LDS (through the 2N-word negacyclic extension);
//   digits    wave w: digit j of half h (2 l digit polynomials, w >> 1 = h l + j), index half w & 1:
//             the first level from both halves' digits, the rest of the transform inside the wave,
//             products with key rows (h l + j) of both outputs, partial products to LDS;
//   reduce    all threads sum the 2 l partial products of each output word;
//   inverse   waves 0..3: output half o, index half w & 1; the last level exchanges halves in LDS;
//             acc_o += INTT(sum).
// Exact in Z_Q like br_kernel, so the outputs are bit-identical.  Slot, test-vector and chunking
// rules as br_kernel (BrArgs).
#include <hip/hip_runtime.h>
#include "br_kernel.hpp"
#include "ntt_half.hpp"

namespace tp {

namespace {

constexpr int kN = 1024;
constexpr int kRegion = 1024;  // words of a wave's LDS region
static_assert(half::kBufWords <= kRegion, "a wave's transpose buffer fits its region");

// coefficient (index) of word k of lane l in index half qh: 512 qh + H1 position
__device__ __forceinline__ int coef(int qh, int l, int k) { return (qh << 9) | half::pos<half::HL::k1>(l, k); }

template <int LD>
__global__ void __launch_bounds__(4 * LD * 64) br_half_kernel(BrArgs A) {
    constexpr int N = kN;
    constexpr int G = 4 * LD;
    static_assert(G >= 8, "extension (4 regions) and exchange (2 regions) regions");
    // red[w]: wave w's transpose buffer, then its partial products [output o][512]; in the
    // prologue red[0..3] hold the two 2N-word extensions, in the inverse red[4..5] the halves
    // exchanged for the last level.  par: the parked digit sources X^e0 acc_h - acc_h (phys
    // positions), then the reduced sums [o][N] (plain positions).
    __shared__ uint64_t red[G][kRegion];
    __shared__ uint64_t par[2][Shape<10>::LDS_WORDS];
    __shared__ uint64_t acc[2][kN];  // the accumulator (coefficients, canonical)
    static_assert(2 * Shape<10>::LDS_WORDS >= 2 * kN, "sums fit the parked region");

    const int w = threadIdx.x >> 6;
    const int dw = w >> 1, h = dw / LD, j = dw % LD, q = w & 1;  // digit phase
    const int o = w >> 1;                                        // waves 0..3: accumulator half
    const size_t rgsw_words = (size_t)4 * N * LD;
    const int total = A.n;
    uint64_t* const sums = &par[0][0];
    uint64_t* const xchg = &red[4][0];

    for (int ct = blockIdx.x; ct < A.count; ct += gridDim.x) {
        int l = threadIdx.x & 63;
        asm volatile("" : "+v"(l));
        const int slot = A.active ? (int)A.active[ct] : ct;
        const uint64_t* lwe = A.lwe + (size_t)slot * (A.n + 1);
        const uint64_t* tv =
            A.tv + (A.tv_index ? (size_t)A.tv_index[slot] * 2 * N : (size_t)slot * A.tv_stride);
        const int i0 = A.slot_step ? (int)A.slot_step[slot] : 0;
        const int i1 = (A.chunk_steps > 0 && i0 + A.chunk_steps < total) ? i0 + A.chunk_steps : total;
        if (w < 4) {
            // wave w sets coefficients coef(q, l, k) of accumulator half o
            uint64_t accR[half::kH];
            if (i0 == 0) {
                // X^b tv_o (negacyclic)
                const uint32_t e = (uint32_t)lwe[A.n];
                const bool flip = e >= (uint32_t)N;
                const int r = (int)(e & (N - 1));
                const uint64_t* src = tv + (size_t)o * N;
#pragma unroll
                for (int k = 0; k < half::kH; ++k) {
                    int jj = coef(q, l, k) - r;
                    bool ng = flip;
                    if (jj < 0) {
                        jj += N;
                        ng = !ng;
                    }
                    const uint64_t v = src[jj];
                    accR[k] = ng ? gl::neg(v) : v;
                }
            } else {
                const uint64_t* st = A.state + (size_t)slot * 2 * N + (size_t)o * N;
#pragma unroll
                for (int k = 0; k < half::kH; ++k) accR[k] = st[coef(q, l, k)];
            }
#pragma unroll
            for (int k = 0; k < half::kH; ++k) acc[o][coef(q, l, k)] = accR[k];
        }
        __syncthreads();

        for (int i = i0; i < i1; ++i) {
            const uint32_t ai = (uint32_t)lwe[i];
            if (ai == 0) continue;  // uniform over the work-group
            const uint32_t e0 = 2 * N - ai;
            const uint64_t* key = A.brk + (size_t)i * rgsw_words;
            // addresses are formed from the lane index inside the step, not held across steps
            asm volatile("" : "+v"(l));

            // prologue: with e' = e0 mod N, E_o = (-acc_o | acc_o) (e0 < N) or (acc_o | -acc_o)
            // (e0 >= N) holds coefficient c of X^e0 acc_o at E_o[c + N - e']
            uint64_t* E = &red[2 * o][0];
            if (w < 4) {
                uint64_t accR[half::kH], na[half::kH], zero[half::kH];
#pragma unroll
                for (int k = 0; k < half::kH; ++k) {
                    accR[k] = acc[o][coef(q, l, k)];
                    zero[k] = 0;
                }
                glv::sub_lc_g<half::kH>(zero, accR, na);  // -acc, canonical
                const uint32_t off_a = e0 >= (uint32_t)N ? 0u : (uint32_t)N, off_n = (uint32_t)N - off_a;
#pragma unroll
                for (int k = 0; k < half::kH; ++k) {
                    E[off_a + (uint32_t)coef(q, l, k)] = accR[k];
                    E[off_n + (uint32_t)coef(q, l, k)] = na[k];
                }
            }
            __syncthreads();
            if (w < 4) {
                uint64_t accR[half::kH], u[half::kH], t[half::kH];
                const uint32_t j0 = (uint32_t)N - (e0 & (uint32_t)(N - 1));
#pragma unroll
                for (int k = 0; k < half::kH; ++k) {
                    accR[k] = acc[o][coef(q, l, k)];
                    u[k] = E[j0 + (uint32_t)coef(q, l, k)];
                }
                glv::sub_lc_g<half::kH>(u, accR, t);  // both canonical: canonical
#pragma unroll
                for (int k = 0; k < half::kH; ++k) par[o][phys(coef(q, l, k))] = t[k];
            }
            __syncthreads();

            // digits: digit j of half h, index half q; the first level needs positions p and p + 512
            __builtin_amdgcn_s_setprio(3);  // lowered as the wave progresses (half::forward)
            {
                const int shift = A.first_shift - j * A.base_log2;
                const uint32_t* src32 = (const uint32_t*)par[h];
                const uint32_t bl = (uint32_t)A.base_log2;
                uint64_t dlo[half::kH], dhi[half::kH];
                auto digit = [&](int c) __attribute__((always_inline)) -> uint64_t {
                    const uint32_t pc = (uint32_t)phys(c);
                    if (shift >= 32 && bl < 32) return __builtin_amdgcn_ubfe(src32[2 * pc + 1], (uint32_t)(shift - 32), bl);
                    if (shift + (int)bl <= 32 && bl < 32) return __builtin_amdgcn_ubfe(src32[2 * pc], (uint32_t)shift, bl);
                    return (par[h][pc] >> shift) & ((1ULL << bl) - 1);
                };
#pragma unroll
                for (int k = 0; k < half::kH; ++k) {
                    dlo[k] = digit(coef(0, l, k));
                    dhi[k] = digit(coef(1, l, k));
                }
                uint64_t x[half::kH];
                if (bl <= 16) {
                    // u + v 2^48 = (u, v << 16), u - v 2^48 = (u + 1, ~(v << 16)) (ct_bfly_small_first)
#pragma unroll
                    for (int k = 0; k < half::kH; ++k) {
                        const uint32_t a = (uint32_t)dlo[k], b16 = (uint32_t)dhi[k] << 16;
                        x[k] = q == 0 ? glv::pk(a, b16) : glv::pk(a + 1, ~b16);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < half::kH; ++k) {
                        uint64_t u = dlo[k], v = dhi[k];
                        ct_bfly_pow2(u, v, gl::kFwdPow2[1]);
                        x[k] = q == 0 ? u : v;
                    }
                }
                // key words of output positions 512 q + 8 l + k: device row pair (k16 >> 1) 64 + L16
                // with L16 = 32 q + (l >> 1), k16 = 8 (l & 1) + k
                const uint64_t* ka = key + (size_t)(h * LD + j) * 2 * N;
                const uint64_t* kb = ka + N;
                const uint32_t L16 = (uint32_t)(32 * q + (l >> 1)), c0 = (uint32_t)(4 * (l & 1));
                // key pair c (words 2c, 2c + 1) of both rows, one pair ahead of its use (the
                // first during the transform's last level)
                auto load_keys = [&](U64x2& ga, U64x2& gb, int c) __attribute__((always_inline)) {
                    ga = ld_pair(ka, (c0 + c) * 64 + L16);
                    gb = ld_pair(kb, (c0 + c) * 64 + L16);
                    __builtin_amdgcn_sched_barrier(0);
                };
                U64x2 va, vb;
                uint64_t* buf = &red[w][0];
                auto first_keys = [&]() __attribute__((always_inline)) { load_keys(va, vb, 0); };
                if (q == 0) half::forward<0>(x, buf, A.tw + 3 * N, l, first_keys);
                else half::forward<1>(x, buf, A.tw + 3 * N, l, first_keys);
                half::wave_sync();  // the transpose buffer's last reads are done
                __builtin_amdgcn_s_setprio(0);
                static_for<4>([&](auto C) __attribute__((always_inline)) {
                    constexpr int c = C;
                    U64x2 na, nb;
                    if constexpr (c < 3) load_keys(na, nb, c + 1);
                    // words 2c, 2c + 1 against both rows
                    uint64_t g4[4] = {0, 0, 0, 0};
                    const uint64_t d4[4] = {x[2 * c], x[2 * c + 1], x[2 * c], x[2 * c + 1]};
                    const uint64_t k4[4] = {va.x, va.y, vb.x, vb.y};
                    glv::mac_lazy_g<4>(g4, d4, k4);
                    buf[half::pos<half::HL::k3>(l, 2 * c)] = g4[0];
                    buf[half::pos<half::HL::k3>(l, 2 * c + 1)] = g4[1];
                    buf[half::kHalf + half::pos<half::HL::k3>(l, 2 * c)] = g4[2];
                    buf[half::kHalf + half::pos<half::HL::k3>(l, 2 * c + 1)] = g4[3];
                    if constexpr (c < 3) {
                        va = na;
                        vb = nb;
                    }
                });
            }
            __builtin_amdgcn_s_setprio(3);
            __syncthreads();

            // reduce: sums[o N + g] = sum over the 2 l digit polynomials (waves 2 d + (g >> 9))
            for (int x = threadIdx.x; x < 2 * N; x += G * 64) {
                const int oo = x >> 10, g = x & (N - 1);
                const int off = oo * half::kHalf + (g & (half::kHalf - 1));
                uint64_t s = red[g >> 9][off];
#pragma unroll
                for (int d = 1; d < 2 * LD; ++d) s = gl::add_ll(s, red[2 * d + (g >> 9)][off]);
                sums[x] = s;
            }
            __syncthreads();

            // inverse, all but the last level: waves 0..3
            if (w < 4) {
                uint64_t x[half::kH];
#pragma unroll
                for (int k = 0; k < half::kH; ++k) x[k] = sums[o * N + (q << 9) + half::pos<half::HL::k3>(l, k)];
                if (q == 0) half::inverse_part1<0>(x, &red[w][0], A.itw + 3 * N, l);
                else half::inverse_part1<1>(x, &red[w][0], A.itw + 3 * N, l);
#pragma unroll
                for (int k = 0; k < half::kH; ++k) xchg[(2 * o + q) * half::kHalf + half::pos<half::HL::k1>(l, k)] = x[k];
            }
            __syncthreads();
            // the last level and the accumulator update
            if (w < 4) {
                uint64_t lo[half::kH], hi[half::kH], y[half::kH], t[half::kH], accR[half::kH];
#pragma unroll
                for (int k = 0; k < half::kH; ++k) {
                    lo[k] = xchg[(2 * o) * half::kHalf + half::pos<half::HL::k1>(l, k)];
                    hi[k] = xchg[(2 * o + 1) * half::kHalf + half::pos<half::HL::k1>(l, k)];
                    accR[k] = acc[o][coef(q, l, k)];
                }
                if (q == 0) half::inverse_part2<0>(y, lo, hi);
                else half::inverse_part2<1>(y, lo, hi);
                glv::add_cc_g<half::kH>(y, accR, t, glv::qm1_const());
#pragma unroll
                for (int k = 0; k < half::kH; ++k) acc[o][coef(q, l, k)] = t[k];
            }
            // the next step's prologue writes red[0..3] and par, read above only before the
            // barriers that precede this point; red[4..5] (read just now) are next written after
            // the next prologue's barriers; each accumulator word is read and written by one lane
        }

        if (A.slot_step && threadIdx.x == 0) A.slot_step[slot] = (uint32_t)i1;
        if (w < 4) {
            uint64_t accR[half::kH];
#pragma unroll
            for (int k = 0; k < half::kH; ++k) accR[k] = acc[o][coef(q, l, k)];
            if (i1 < total) {
                // chunk boundary: park the accumulator for the next launch
                uint64_t* st = A.state + (size_t)slot * 2 * N + (size_t)o * N;
#pragma unroll
                for (int k = 0; k < half::kH; ++k) st[coef(q, l, k)] = accR[k];
            } else {
                if (A.out_lwe) {
                    // sample extraction at index 0
                    uint64_t* out = A.out_lwe + (size_t)slot * (N + 1);
#pragma unroll
                    for (int k = 0; k < half::kH; ++k) {
                        const int c = coef(q, l, k);
                        if (o == 0) out[c == 0 ? 0 : N - c] = c == 0 ? accR[k] : gl::neg(accR[k]);
                        else if (c == 0) out[N] = accR[k];
                    }
                }
                if (A.out_rlwe && !A.out_rlwe_ntt) {
                    uint64_t* out = A.out_rlwe + (size_t)slot * 2 * N + (size_t)o * N;
#pragma unroll
                    for (int k = 0; k < half::kH; ++k) out[coef(q, l, k)] = accR[k];
                }
            }
        }
        if (i1 >= total && A.out_rlwe && A.out_rlwe_ntt) {
            // NTT-domain output (reference order): the halves through LDS to waves 0 and 1, which
            // transform them with the one-wave transform
            __syncthreads();
            if (w < 2) {
                uint64_t x16[kE];
#pragma unroll
                for (int k = 0; k < kE; ++k) x16[k] = acc[w][idx1<10>(l, k)];
                ntt_fwd<10>(x16, &red[2 * w][0], A.tw, l);  // a 1088-word buffer: two regions
                uint64_t* out = A.out_rlwe + (size_t)slot * 2 * N + (size_t)w * N;
#pragma unroll
                for (int k = 0; k < kE; ++k) out[idx3(l, k)] = gl::canon(x16[k]);
            }
        }
        // the next ciphertext's set-up and prologue rewrite what this one read
        __syncthreads();
    }
}

}  // namespace

// Binary blind rotation at N = 1024 with 2..3 digits: A.count ciphertexts over `grid`
// work-groups of 4 * digits waves.
hipError_t launch_blind_rotate_half(int digits, const BrArgs& A, int grid, hipStream_t s) {
    if (grid <= 0) return hipErrorInvalidValue;
    if (digits == 2) {
        hipLaunchKernelGGL(br_half_kernel<2>, dim3(grid), dim3(4 * 2 * 64), 0, s, A);
        return hipGetLastError();
    }
    if (digits == 3) {
        hipLaunchKernelGGL(br_half_kernel<3>, dim3(grid), dim3(4 * 3 * 64), 0, s, A);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

}  // namespace tp
