// HIP kernels for gfx950: batched NTT, CGGI blind rotation + sample extraction,
// batched LWE->LWE key switching.
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

struct alignas(16) U64x2 {
    uint64_t x, y;
};

template <int LOGN>
__device__ __forceinline__ void load_key_row(uint64_t (&kv)[kE], const uint64_t* __restrict__ row, int L) {
    const U64x2* r = reinterpret_cast<const U64x2*>(row);
#pragma unroll
    for (int c = 0; c < kE / 2; ++c) {
        const U64x2 v = r[c * Shape<LOGN>::P + L];
        kv[2 * c] = v.x;
        kv[2 * c + 1] = v.y;
    }
}

// NTT(X^e) at the lane's NTT positions (2c, 2c+1), e in [0, 2N): the table holds NTT(X^e)
// for e < N in key-row order; X^e = -X^(e-N) for e >= N.
template <int LOGN>
__device__ __forceinline__ U64x2 mono_pair(const uint64_t* __restrict__ mono, uint32_t e, int c, int L) {
    constexpr int N = Shape<LOGN>::N;
    const U64x2 v = reinterpret_cast<const U64x2*>(mono + (size_t)(e & (N - 1)) * N)[c * Shape<LOGN>::P + L];
    if (e >= (uint32_t)N) return U64x2{gl::neg(v.x), gl::neg(v.y)};
    return v;
}

// NTT(X^e) - 1
template <int LOGN>
__device__ __forceinline__ U64x2 weight_pair(const uint64_t* __restrict__ mono, uint32_t e, int c, int L) {
    const U64x2 v = mono_pair<LOGN>(mono, e, c, L);
    return U64x2{gl::sub(v.x, 1), gl::sub(v.y, 1)};
}

// BMMP third weight X^(a0+a1) - X^a0 - X^a1 + 1 (bmmp_blind_rotation.cpp:280-323).
template <int LOGN>
__device__ __forceinline__ U64x2 bmmp_w2_pair(const uint64_t* __restrict__ mono, uint32_t e0, uint32_t e1,
                                              uint32_t e2, int c, int L) {
    const U64x2 m2 = mono_pair<LOGN>(mono, e2, c, L), m0 = mono_pair<LOGN>(mono, e0, c, L),
                m1 = mono_pair<LOGN>(mono, e1, c, L);
    return U64x2{gl::add(gl::sub(gl::sub(m2.x, m0.x), m1.x), 1), gl::add(gl::sub(gl::sub(m2.y, m0.y), m1.y), 1)};
}

// ------------------------------------------------------------------ batched NTT

template <int LOGN, bool INVERSE>
__global__ void __launch_bounds__(256) ntt_batch_kernel(uint64_t* __restrict__ out, const uint64_t* __restrict__ in,
                                                        size_t count, const uint64_t* __restrict__ tw) {
    using S = Shape<LOGN>;
    constexpr int POLYS = 256 / S::P;
    __shared__ uint64_t lds[POLYS][S::LDS_WORDS];
    const int g = threadIdx.x / S::P, L = threadIdx.x % S::P;
    for (size_t base = (size_t)blockIdx.x * POLYS; base < count; base += (size_t)gridDim.x * POLYS) {
        const size_t p = base + g;
        const bool live = p < count;
        const uint64_t* src = in + (live ? p : 0) * S::N;
        uint64_t x[kE];
        if (!INVERSE) {
#pragma unroll
            for (int k = 0; k < kE; ++k) x[k] = src[idx1<LOGN>(L, k)];
            ntt_forward<LOGN>(x, lds[g], tw, L);
            if (live)
#pragma unroll
                for (int k = 0; k < kE; ++k) out[p * S::N + idx3(L, k)] = gl::canon(x[k]);
        } else {
#pragma unroll
            for (int k = 0; k < kE; ++k) x[k] = src[idx3(L, k)];
            ntt_inverse<LOGN>(x, lds[g], tw, L);
            if (live)
#pragma unroll
                for (int k = 0; k < kE; ++k) out[p * S::N + idx1<LOGN>(L, k)] = gl::canon(x[k]);
        }
    }
}

// ------------------------------------------------------------------ blind rotation

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
template <int LOGN, int LD, int METHOD>
__global__ void __launch_bounds__(Shape<LOGN>::P) __attribute__((amdgpu_waves_per_eu(2)))
br_kernel(BrArgs A) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    constexpr int NC = METHOD == 0 ? 1 : (METHOD == 1 ? 2 : 3);  // RGSWs per step
    __shared__ uint64_t buf[S::LDS_WORDS];
    __shared__ uint64_t src[METHOD == 0 ? S::LDS_WORDS : 1];
    const int L = threadIdx.x;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    const size_t rgsw_words = (size_t)4 * N * LD;

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
            uint64_t out[2][kE];
#pragma unroll
            for (int k = 0; k < kE; ++k) out[0][k] = out[1][k] = 0;

#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                if constexpr (METHOD == 0) {
                    // the half being decomposed is parked in LDS so that every digit can re-read
                    // its rotation X^e0 acc_h - acc_h (kept out of registers)
                    poly_sync<S::WAVES>();
#pragma unroll
                    for (int k = 0; k < kE; ++k) src[phys(idx1<LOGN>(L, k))] = acc[0][k];
                    poly_sync<S::WAVES>();
                }
#pragma unroll 1
                for (int j = 0; j < LD; ++j) {
                    uint64_t d[kE];
                    if constexpr (METHOD == 0) {
                        const int shift = A.first_shift - j * A.base_log2;
                        rotate_from<LOGN>(d, src, e0, L, true);
#pragma unroll
                        for (int k = 0; k < kE; ++k) d[k] = (gl::sub(d[k], acc[0][k]) >> shift) & mask;
                    } else {
                        // signed digits of x + C mod Q (gadget_decomp.h:169-198)
                        const bool last = (j == A.max_digits - 1);
                        const int shift = last ? 0 : 64 - (j + 1) * A.base_log2;
                        const uint64_t dm = last ? (2 * A.least - 1) : mask;
                        const uint64_t delta = last ? A.least : (1ULL << (A.base_log2 - 1));
#pragma unroll
                        for (int k = 0; k < kE; ++k)
                            d[k] = gl::sub((gl::add(acc[0][k], A.corrector) >> shift) & dm, delta);
                    }
                    ntt_forward<LOGN>(d, buf, A.tw, L);
                    const size_t row = (size_t)(h * LD + j) * 2 * N;
                    if constexpr (METHOD == 0) {
                        // key pairs are read next to their use, not hoisted over the transform
                        const U64x2* ka = reinterpret_cast<const U64x2*>(opaque(key + row));
                        const U64x2* kb = reinterpret_cast<const U64x2*>(opaque(key + row + N));
#pragma unroll
                        for (int c = 0; c < kE / 2; ++c) {
                            const U64x2 va = ka[c * S::P + L], vb = kb[c * S::P + L];
                            out[0][2 * c] = gl::add_ll(out[0][2 * c], gl::mul_lazy(d[2 * c], va.x));
                            out[0][2 * c + 1] = gl::add_ll(out[0][2 * c + 1], gl::mul_lazy(d[2 * c + 1], va.y));
                            out[1][2 * c] = gl::add_ll(out[1][2 * c], gl::mul_lazy(d[2 * c], vb.x));
                            out[1][2 * c + 1] = gl::add_ll(out[1][2 * c + 1], gl::mul_lazy(d[2 * c + 1], vb.y));
                        }
                    } else {
                        // combined key sum_c w_c RGSW_c[row], formed pair by pair next to its use
                        const uint64_t* krow = opaque(key + row);
                        const uint64_t* mono = opaque_after(A.mono, d[0]);  // weights re-read per digit
#pragma unroll
                        for (int c = 0; c < kE / 2; ++c) {
                            const U64x2 w0 = weight_pair<LOGN>(mono, e0, c, L);
                            const U64x2 w1 = weight_pair<LOGN>(mono, e1, c, L);
                            U64x2 w2{0, 0};
                            if constexpr (NC == 3) w2 = bmmp_w2_pair<LOGN>(mono, e0, e1, e2, c, L);
#pragma unroll
                            for (int p = 0; p < 2; ++p) {
                                const U64x2* k0 = reinterpret_cast<const U64x2*>(krow + p * N);
                                const U64x2* k1 = reinterpret_cast<const U64x2*>(krow + rgsw_words + p * N);
                                const U64x2 a0 = k0[c * S::P + L], a1 = k1[c * S::P + L];
                                uint64_t cx = gl::add_ll(gl::mul_lazy(a0.x, w0.x), gl::mul_lazy(a1.x, w1.x));
                                uint64_t cy = gl::add_ll(gl::mul_lazy(a0.y, w0.y), gl::mul_lazy(a1.y, w1.y));
                                if constexpr (NC == 3) {
                                    const U64x2* k2 = reinterpret_cast<const U64x2*>(krow + 2 * rgsw_words + p * N);
                                    const U64x2 a2 = k2[c * S::P + L];
                                    cx = gl::add_ll(cx, gl::mul_lazy(a2.x, w2.x));
                                    cy = gl::add_ll(cy, gl::mul_lazy(a2.y, w2.y));
                                }
                                if (p == 0) {
                                    out[0][2 * c] = gl::add_ll(out[0][2 * c], gl::mul_lazy(d[2 * c], cx));
                                    out[0][2 * c + 1] = gl::add_ll(out[0][2 * c + 1], gl::mul_lazy(d[2 * c + 1], cy));
                                } else {
                                    out[1][2 * c] = gl::add_ll(out[1][2 * c], gl::mul_lazy(d[2 * c], cx));
                                    out[1][2 * c + 1] = gl::add_ll(out[1][2 * c + 1], gl::mul_lazy(d[2 * c + 1], cy));
                                }
                            }
                        }
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
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                ntt_inverse<LOGN>(out[0], buf, A.itw, L);
#pragma unroll
                for (int k = 0; k < kE; ++k) {
                    const uint64_t t = gl::canon(gl::add_lc(out[0][k], acc[0][k]));
                    acc[0][k] = acc[1][k];
                    acc[1][k] = t;
                    out[0][k] = out[1][k];
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
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint64_t* o = A.out_rlwe + (size_t)slot * 2 * N + h * N;
                if (A.out_rlwe_ntt) {
                    ntt_forward<LOGN>(acc[h], buf, A.tw, L);
#pragma unroll
                    for (int k = 0; k < kE; ++k) o[idx3(L, k)] = gl::canon(acc[h][k]);
                } else {
#pragma unroll
                    for (int k = 0; k < kE; ++k) o[idx1<LOGN>(L, k)] = acc[h][k];
                }
            }
        }
        poly_sync<S::WAVES>();
    }
}

// ------------------------------------------------------------------ LWE key switch

// out[c][k] = -(sum_ij d_ij(c) KSK[i][j][k]) (k < n_out), out[c][n_out] = in[c][n_in] - (...)
// (lwe_conversion.cpp:205-239).  With unsigned digits u_ij = d_ij + delta_j the sum is
// sum u_ij K - corr[k], corr[k] = sum_ij delta_j K[i][j][k] precomputed per key.  u_ij < 2^b, so
// u K_lo and u K_hi are accumulated exactly in 64 bits and reduced once at the end.
// GEMM layout: one ciphertext per lane (its digits are per-lane work), a tile of kKsCols key
// columns per block; the key row segment K[i][j][col0 .. col0 + kKsCols) is wave-uniform and is
// read through the scalar cache, so each MAC is two v_mad_u64_u32 with a scalar operand.
constexpr int kKsCols = 16;
constexpr int kKsBlock = 256;

__global__ void __launch_bounds__(kKsBlock) lwe_keyswitch_kernel(KsArgs A) {
    const int col0 = blockIdx.x * kKsCols;
    const int c = blockIdx.y * kKsBlock + threadIdx.x;
    const int W = A.n_out + 1;
    const bool live = c < A.count;
    const int ct = live ? (A.rows ? (int)A.rows[c] : c) : 0;
    const uint64_t* in = A.in + (size_t)ct * A.in_stride;
    uint64_t lo[kKsCols], hi[kKsCols];
#pragma unroll
    for (int t = 0; t < kKsCols; ++t) lo[t] = hi[t] = 0;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    // the device key carries kKsCols words of zero padding, so a partial last tile reads in
    // bounds; its extra columns are discarded
    const uint64_t* krow = A.ksk + col0;
    for (int i = 0; i < A.n_in; ++i) {
        const uint64_t x = gl::add(in[i], A.corrector);
        const uint64_t* kr = krow + (size_t)i * A.digits * W;
        for (int j = 0; j < A.digits; ++j, kr += W) {
            const uint32_t u = (uint32_t)((x >> (64 - (j + 1) * A.base_log2)) & mask);
#pragma unroll
            for (int t = 0; t < kKsCols; ++t) {
                const uint64_t kv = kr[t];
                lo[t] += (uint64_t)u * (uint32_t)kv;
                hi[t] += (uint64_t)u * (uint32_t)(kv >> 32);
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int t = 0; t < kKsCols; ++t) {
        const int col = col0 + t;
        if (col >= W) break;
        // total = hi * 2^32 + lo, both < 2^62
        const uint64_t t_lo = (hi[t] << 32) + lo[t];
        const uint64_t carry = t_lo < lo[t];
        const uint64_t t_hi = (hi[t] >> 32) + carry;
        const uint64_t acc = gl::sub(gl::reduce128(t_hi, t_lo), A.corr[col]);
        uint64_t* o = A.out + (size_t)c * W;  // compact: output row c reads input row rows[c]
        if (col < A.n_out) o[col] = gl::neg(acc);
        else o[col] = gl::sub(in[A.n_in], acc);
    }
}

// ------------------------------------------------------------------ launchers

template <int LOGN>
static void launch_ntt(uint64_t* out, const uint64_t* in, size_t count, const uint64_t* tw, bool inverse,
                       hipStream_t s) {
    constexpr int POLYS = 256 / Shape<LOGN>::P;
    size_t blocks = (count + POLYS - 1) / POLYS;
    if (blocks > 65535) blocks = 65535;
    if (inverse)
        hipLaunchKernelGGL((ntt_batch_kernel<LOGN, true>), dim3((unsigned)blocks), dim3(256), 0, s, out, in, count, tw);
    else
        hipLaunchKernelGGL((ntt_batch_kernel<LOGN, false>), dim3((unsigned)blocks), dim3(256), 0, s, out, in, count, tw);
}

hipError_t launch_ntt_batch(int logN, uint64_t* out, const uint64_t* in, size_t count, const uint64_t* tw,
                            bool inverse, hipStream_t s) {
    switch (logN) {
        case 10: launch_ntt<10>(out, in, count, tw, inverse, s); break;
        case 11: launch_ntt<11>(out, in, count, tw, inverse, s); break;
        case 12: launch_ntt<12>(out, in, count, tw, inverse, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LOGN, int LD, int METHOD>
static void launch_br(const BrArgs& A, int grid, hipStream_t s) {
    hipLaunchKernelGGL((br_kernel<LOGN, LD, METHOD>), dim3(grid), dim3(Shape<LOGN>::P), 0, s, A);
}

hipError_t launch_blind_rotate(int logN, int method, int digits, const BrArgs& A, int grid, hipStream_t s) {
#define TP_BR_CASE(LG, LD, M) \
    if (logN == LG && digits == LD && method == M) { launch_br<LG, LD, M>(A, grid, s); return hipGetLastError(); }
#define TP_BR_METHODS(LG, LD) TP_BR_CASE(LG, LD, 0) TP_BR_CASE(LG, LD, 1) TP_BR_CASE(LG, LD, 2)
    TP_BR_METHODS(10, 2) TP_BR_METHODS(10, 3) TP_BR_METHODS(10, 4)
    TP_BR_METHODS(11, 2) TP_BR_METHODS(11, 3) TP_BR_METHODS(11, 4)
    TP_BR_METHODS(12, 3) TP_BR_METHODS(12, 4)
#undef TP_BR_METHODS
#undef TP_BR_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_lwe_keyswitch(const KsArgs& A, hipStream_t s) {
    dim3 grid((A.n_out + 1 + kKsCols - 1) / kKsCols, (A.count + kKsBlock - 1) / kKsBlock);
    hipLaunchKernelGGL(lwe_keyswitch_kernel, grid, dim3(kKsBlock), 0, s, A);
    return hipGetLastError();
}

}  // namespace tp
