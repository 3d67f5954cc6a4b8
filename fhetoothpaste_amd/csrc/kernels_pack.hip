// HIP kernels for gfx950: RLWE key switching, automorphisms, trace and n-LWE -> RLWE packing.
// One RLWE per work-group of P = N/16 lanes (lane L holds coefficients L + P k, k < 16).
#include <hip/hip_runtime.h>
#include "gl64.hpp"
#include "ntt_wave.hpp"
#include "kernels.hpp"

namespace tp {

namespace {

// x[k] <- coefficient idx1(L,k) of tau_t(p), p given by the same lanes' registers (coefficients,
// canonical): out[(i t) mod N] = +-p[i] (utils/math_utils.cpp:70-85), through LDS.
template <int LOGN>
__device__ __forceinline__ void automorphism_regs(uint64_t (&x)[kE], uint64_t* lds, uint32_t t, int L) {
    using S = Shape<LOGN>;
    constexpr uint32_t N = S::N;
    poly_sync<S::WAVES>();
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        const uint32_t i = idx1<LOGN>(L, k);
        const uint32_t e = (i * t) & (2 * N - 1);
        lds[phys(e & (N - 1))] = e >= N ? gl::neg(x[k]) : x[k];
    }
    poly_sync<S::WAVES>();
#pragma unroll
    for (int k = 0; k < kE; ++k) x[k] = lds[phys(idx1<LOGN>(L, k))];
}

// x[k] <- coefficient idx1(L,k) of X^s p (negacyclic), 0 <= s < 2N, through LDS.
template <int LOGN>
__device__ __forceinline__ void rotate_regs(uint64_t (&x)[kE], uint64_t* lds, uint32_t s, int L) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    poly_sync<S::WAVES>();
#pragma unroll
    for (int k = 0; k < kE; ++k) lds[phys(idx1<LOGN>(L, k))] = x[k];
    poly_sync<S::WAVES>();
    const bool flip = s >= (uint32_t)N;
    const int r = (int)(s & (N - 1));
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        int j = idx1<LOGN>(L, k) - r;
        bool ng = flip;
        if (j < 0) { j += N; ng = !ng; }
        const uint64_t v = lds[phys(j)];
        x[k] = ng ? gl::neg(v) : v;
    }
}

// RLWEtoRLWEConverter::Convert core (rlwe_conversion.cpp:224-249): out_p = sum_j
// NTT(sdigit_j(a)) * KSK_j.p, p = a, b (lazy, NTT domain, pass-3 layout).  a: coefficients,
// canonical, pass-1 layout.  ksk: d rows of [2N] in key-row order.
template <int LOGN>
__device__ __forceinline__ void rlwe_ks_core(uint64_t (&out_a)[kE], uint64_t (&out_b)[kE], const uint64_t (&a)[kE],
                                             uint64_t* lds, const KeySwitchArgs& A, const uint64_t* ksk, int L) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
#pragma unroll
    for (int k = 0; k < kE; ++k) out_a[k] = out_b[k] = 0;
    // x + C mod Q (canonical a: the corrector sum is formed once, not per digit) and the first-level
    // constants of the signed digits (offsets 2^(b-1) and `least`)
    uint64_t xc[kE], cv[kE];
#pragma unroll
    for (int k = 0; k < kE; ++k) cv[k] = A.corrector;
    {
        uint64_t ac[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) ac[k] = gl::canon(a[k]);
        glv::add_cc_g<kE>(ac, cv, xc, glv::qm1_const());
    }
    uint64_t kp_n, km_n, kp_l, km_l;
    signed_first_consts(1ULL << (A.base_log2 - 1), kp_n, km_n);
    signed_first_consts(A.least, kp_l, km_l);
#pragma unroll 1
    for (int j = 0; j < A.digits; ++j) {
        const bool last = (j == A.max_digits - 1);
        const int shift = last ? 0 : 64 - (j + 1) * A.base_log2;
        const uint64_t dm = last ? (2 * A.least - 1) : mask;
        const uint64_t delta = last ? A.least : (1ULL << (A.base_log2 - 1));
        uint64_t d[kE];
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
        // key pairs of both rows, two at a time: the first two are requested during the transform's
        // last level, each next two while the current two are used
        const uint64_t* ka = opaque(ksk + (size_t)j * 2 * N);
        const uint64_t* kb = opaque(ksk + (size_t)j * 2 * N + N);
        U64x2 va[2], vb[2];
        auto load_group = [&](U64x2 (&ga)[2], U64x2 (&gb)[2], int c0) __attribute__((always_inline)) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                ga[q] = ld_pair(ka, (c0 + q) * S::P + L);
                gb[q] = ld_pair(kb, (c0 + q) * S::P + L);
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        ntt_fwd<LOGN, 1>(d, lds, A.tw, L, [&]() __attribute__((always_inline)) { load_group(va, vb, 0); });
        static_for<4>([&](auto G) __attribute__((always_inline)) {
            constexpr int c0 = 2 * G;
            U64x2 na[2], nb[2];
            if constexpr (G < 3) load_group(na, nb, c0 + 2);
            // coefficients 2 c0 .. 2 c0 + 3 against both key rows, one group of eight
            uint64_t g8[8], d8[8], k8[8];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                g8[m] = out_a[2 * c0 + m];
                g8[m + 4] = out_b[2 * c0 + m];
                d8[m] = d8[m + 4] = d[2 * c0 + m];
            }
            k8[0] = va[0].x; k8[1] = va[0].y; k8[2] = va[1].x; k8[3] = va[1].y;
            k8[4] = vb[0].x; k8[5] = vb[0].y; k8[6] = vb[1].x; k8[7] = vb[1].y;
            glv::mac_lazy_g<8>(g8, d8, k8);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                out_a[2 * c0 + m] = g8[m];
                out_b[2 * c0 + m] = g8[m + 4];
            }
            if constexpr (G < 3) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    va[q] = na[q];
                    vb[q] = nb[q];
                }
            }
        });
    }
}

// AutomorphismEvaluator::Eval on registers (automorphism_evaluation.cpp:150-159):
// [a|b] coefficients -> [a'|b'] NTT domain (canonical, pass-3 layout) with
// a' = -sum NTT(sdigit(tau(a))) K.a,  b' = NTT(tau(b)) - sum NTT(sdigit(tau(a))) K.b.
template <int LOGN>
__device__ __forceinline__ void automorphism_eval(uint64_t (&a)[kE], uint64_t (&b)[kE], uint64_t* lds, uint32_t t,
                                                  const KeySwitchArgs& A, const uint64_t* ksk, int L) {
    automorphism_regs<LOGN>(a, lds, t, L);
    automorphism_regs<LOGN>(b, lds, t, L);
    ntt_fwd<LOGN>(b, lds, A.tw, L);
    uint64_t oa[kE], ob[kE];
    rlwe_ks_core<LOGN>(oa, ob, a, lds, A, ksk, L);
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        a[k] = gl::neg(gl::canon(oa[k]));
        b[k] = gl::sub(gl::canon(b[k]), gl::canon(ob[k]));
    }
}

}  // namespace

// Batched automorphism: in [count][2N] coefficients, out [count][2N] NTT domain.  b is read and
// transformed after the key switch of a, so its registers are not live across the digit loop (two
// waves per SIMD fit); a sync before the writes keeps an aliased in == out row intact.
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) __attribute__((amdgpu_waves_per_eu(2)))
automorphism_kernel(KeySwitchArgs A, const uint64_t* in, uint64_t* out, int count, uint32_t t, const uint64_t* ksk) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    for (int ct = blockIdx.x; ct < count; ct += gridDim.x) {
        // the lane index re-read per RLWE: its addresses are formed inside the loop, not held in
        // registers across it
        int L = threadIdx.x;
        asm volatile("" : "+v"(L));
        const uint64_t* src = in + (size_t)ct * 2 * N;
        uint64_t a[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) a[k] = src[idx1<LOGN>(L, k)];
        automorphism_regs<LOGN>(a, lds, t, L);
        uint64_t oa[kE], ob[kE];
        rlwe_ks_core<LOGN>(oa, ob, a, lds, A, ksk, L);
        uint64_t b[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) b[k] = src[N + idx1<LOGN>(L, k)];
        automorphism_regs<LOGN>(b, lds, t, L);
        ntt_fwd<LOGN>(b, lds, A.tw, L);
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            oa[k] = gl::neg(gl::canon(oa[k]));
            ob[k] = gl::sub(gl::canon(b[k]), gl::canon(ob[k]));
        }
        // coalesced writes (NTT index L + P k); the transposes' syncs also order every read of in
        // before the writes
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(oa, lds, L);
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(ob, lds, L);
        uint64_t* dst = out + (size_t)ct * 2 * N;
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            dst[idx1<LOGN>(L, k)] = oa[k];
            dst[N + idx1<LOGN>(L, k)] = ob[k];
        }
        poly_sync<S::WAVES>();
    }
}

// RLWE key switch with optional automorphism, one ciphertext per work-group:
//   mode 0 (RLWEtoRLWEConverter::Convert, rlwe_conversion.cpp:224-249): in = [a coef | b NTT],
//          r = (-sum NTT(sdigit_j(a)) K_j.a, b - sum NTT(sdigit_j(a)) K_j.b);
//   mode 1 (AutomorphismEvaluator::Eval, automorphism_evaluation.cpp:150-159): in = [a|b] coef,
//          a' = tau_t(a), b' = NTT(tau_t(b)), then as mode 0;
//   mode 2 (AutomorphismEvaluator::EvalSwitchOnly, :171-176): a' = tau_t(a), b' = 0, then as mode 0.
// out (NTT domain) = r, or out + r when accumulating (the reference subtracts into its output).
// b is read after the key switch of a (registers, as in automorphism_kernel).
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) __attribute__((amdgpu_waves_per_eu(2)))
rlwe_keyswitch_kernel(KeySwitchArgs A, const uint64_t* in, uint64_t* out, int count, uint32_t t, int mode,
                      int accumulate, const uint64_t* ksk) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    for (int ct = blockIdx.x; ct < count; ct += gridDim.x) {
        // the lane index re-read per RLWE: its addresses are formed inside the loop, not held in
        // registers across it
        int L = threadIdx.x;
        asm volatile("" : "+v"(L));
        const uint64_t* src = in + (size_t)ct * 2 * N;
        uint64_t a[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) a[k] = src[idx1<LOGN>(L, k)];
        if (mode != 0) automorphism_regs<LOGN>(a, lds, t, L);
        uint64_t oa[kE], ob[kE];
        rlwe_ks_core<LOGN>(oa, ob, a, lds, A, ksk, L);
        uint64_t b[kE];
        if (mode == 0) {
#pragma unroll
            for (int k = 0; k < kE; ++k) b[k] = src[N + idx3(L, k)];
        } else if (mode == 1) {
#pragma unroll
            for (int k = 0; k < kE; ++k) b[k] = src[N + idx1<LOGN>(L, k)];
            automorphism_regs<LOGN>(b, lds, t, L);
            ntt_fwd<LOGN>(b, lds, A.tw, L);
        } else {
#pragma unroll
            for (int k = 0; k < kE; ++k) b[k] = 0;
        }
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            oa[k] = gl::neg(gl::canon(oa[k]));
            ob[k] = gl::sub(gl::canon(b[k]), gl::canon(ob[k]));
        }
        // coalesced writes (NTT index L + P k); the transposes' syncs also order every read of in
        // before the writes
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(oa, lds, L);
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(ob, lds, L);
        uint64_t* dst = out + (size_t)ct * 2 * N;
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            const int q = idx1<LOGN>(L, k);
            uint64_t ra = oa[k], rb = ob[k];
            if (accumulate) {
                ra = gl::add(dst[q], ra);
                rb = gl::add(dst[N + q], rb);
            }
            dst[q] = ra;
            dst[N + q] = rb;
        }
        poly_sync<S::WAVES>();
    }
}

// Automorphism key switch of an NTT-domain RLWE [a|b] (reference order), as the SAP tree applies it
// to its nodes (sap_decomposer.cpp:240-248, :287-291: BackwardNTT then EvalAuto).  a is brought to
// coefficients in registers; b stays in the evaluation domain, where tau_t is a permutation:
// NTT(b(X^t))[p] = NTT(b)[q] with 2 brv(q) + 1 = t (2 brv(p) + 1) mod 2N, so it is gathered, not
// transformed.  y = (-KS(a).a, b' - KS(a).b), canonical, pass-3 layout (NTT index idx3(L, k)),
// equal to rlwe_keyswitch_kernel mode 1 on the inverse transform of src.  Ends with a sync after
// its last read of src.
template <int LOGN>
__device__ __forceinline__ void auto_ntt_core(uint64_t (&ya)[kE], uint64_t (&yb)[kE], const uint64_t* src,
                                              uint64_t* lds, uint32_t t, const KeySwitchArgs& A,
                                              const uint64_t* ksk, int L) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    uint64_t a[kE];
#pragma unroll
    for (int k = 0; k < kE; ++k) a[k] = src[idx3(L, k)];
    ntt_inv<LOGN>(a, lds, A.itw, L);
#pragma unroll
    for (int k = 0; k < kE; ++k) a[k] = gl::canon(a[k]);
    automorphism_regs<LOGN>(a, lds, t, L);
    uint64_t oa[kE], ob[kE];
    rlwe_ks_core<LOGN>(oa, ob, a, lds, A, ksk, L);
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        const uint32_t p = (uint32_t)idx3(L, k);
        const uint32_t w = 2 * (__builtin_bitreverse32(p) >> (32 - LOGN)) + 1;
        const uint32_t e = (t * w) & (2 * N - 1);
        const uint64_t b = src[N + (__builtin_bitreverse32((e - 1) >> 1) >> (32 - LOGN))];
        ya[k] = gl::neg(gl::canon(oa[k]));
        yb[k] = gl::sub(gl::canon(b), gl::canon(ob[k]));
    }
    poly_sync<S::WAVES>();
}

// out (+)= tau_t(in) key-switched (auto_ntt_core), one RLWE per work-group; in may equal out.
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) __attribute__((amdgpu_waves_per_eu(2)))
auto_ntt_kernel(KeySwitchArgs A, const uint64_t* in, uint64_t* out, int count, uint32_t t, int accumulate,
                const uint64_t* ksk) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    for (int ct = blockIdx.x; ct < count; ct += gridDim.x) {
        // the lane index re-read per RLWE: its addresses are formed inside the loop, not held in
        // registers across it
        int L = threadIdx.x;
        asm volatile("" : "+v"(L));
        uint64_t ya[kE], yb[kE];
        auto_ntt_core<LOGN>(ya, yb, in + (size_t)ct * 2 * N, lds, t, A, ksk, L);
        // coalesced writes (NTT index L + P k)
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(ya, lds, L);
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(yb, lds, L);
        uint64_t* dst = out + (size_t)ct * 2 * N;
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            const int q = idx1<LOGN>(L, k);
            uint64_t ra = ya[k], rb = yb[k];
            if (accumulate) {
                ra = gl::add(dst[q], ra);
                rb = gl::add(dst[N + q], rb);
            }
            dst[q] = ra;
            dst[N + q] = rb;
        }
        poly_sync<S::WAVES>();
    }
}

// One level of the SAP HomTrunc tree (sap_decomposer.cpp:266-311) for all its nodes and inputs:
// RLWE r = e * count + i of X is node e of input i (NTT domain); y = tau_t(x) key-switched
// (t = 2^level + 1).  mode 0 (inner node, :288-302): children x + y -> node 2e and
// (x - y) * polys0 -> node 2e + 1 of out ([2 nodes][count] RLWEs).  mode 1 (level 1, both children
// leaves, :266-277): (x + y) * polys0[2e] + (x - y) * polys1[e] -> out[r]; the caller sums over e.
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) __attribute__((amdgpu_waves_per_eu(2)))
sap_tree_kernel(KeySwitchArgs A, const uint64_t* X, uint64_t* out, const uint64_t* polys0, const uint64_t* polys1,
                int nodes, int count, uint32_t t, int mode, const uint64_t* ksk) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    const int total = nodes * count;
    for (int r = blockIdx.x; r < total; r += gridDim.x) {
        // the lane index re-read per RLWE: its addresses are formed inside the loop, not held in
        // registers across it
        int L = threadIdx.x;
        asm volatile("" : "+v"(L));
        const int e = r / count, i = r - e * count;
        const uint64_t* x = X + (size_t)r * 2 * N;
        uint64_t ya[kE], yb[kE];
        auto_ntt_core<LOGN>(ya, yb, x, lds, t, A, ksk, L);
        // to the lane-contiguous layout (NTT index L + P k), so that the node arithmetic's reads of
        // x and the polynomials and its writes are coalesced
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(ya, lds, L);
        transpose<LOGN, Layout::kPass3, Layout::kPass1>(yb, lds, L);
        // node arithmetic on canonical values with the VALU-layer primitives, eight coefficients of
        // one half (a or b) per group
        const uint64_t qm1 = glv::qm1_const();
        if (mode == 0) {
            uint64_t* ev = out + ((size_t)(2 * e) * count + i) * 2 * N;
            uint64_t* od = out + ((size_t)(2 * e + 1) * count + i) * 2 * N;
            static_for<4>([&](auto G) __attribute__((always_inline)) {
                constexpr int h = G >> 1, k0 = 8 * (G & 1);
                uint64_t xv[8], yv[8], sv[8], dv[8], mv[8], pv[8];
                static_for<8>([&](auto M_) __attribute__((always_inline)) {
                    const int q = idx1<LOGN>(L, k0 + M_);
                    xv[M_] = x[h * N + q];
                    yv[M_] = h ? yb[k0 + M_] : ya[k0 + M_];
                    mv[M_] = polys0[q];
                });
                glv::add_cc_g<8>(xv, yv, sv, qm1);
                glv::sub_lc_g<8>(xv, yv, dv);
                glv::mul_canon_g<8>(dv, mv, pv, qm1);
                static_for<8>([&](auto M_) __attribute__((always_inline)) {
                    const int q = idx1<LOGN>(L, k0 + M_);
                    ev[h * N + q] = sv[M_];
                    od[h * N + q] = pv[M_];
                });
            });
        } else {
            uint64_t* o = out + (size_t)r * 2 * N;
            const uint64_t* me = polys0 + (size_t)(2 * e) * N;
            const uint64_t* mo = polys1 + (size_t)e * N;
            static_for<4>([&](auto G) __attribute__((always_inline)) {
                constexpr int h = G >> 1, k0 = 8 * (G & 1);
                uint64_t xv[8], yv[8], sv[8], dv[8], m0[8], m1[8], p0[8], p1[8], ov[8];
                static_for<8>([&](auto M_) __attribute__((always_inline)) {
                    const int q = idx1<LOGN>(L, k0 + M_);
                    xv[M_] = x[h * N + q];
                    yv[M_] = h ? yb[k0 + M_] : ya[k0 + M_];
                    m0[M_] = me[q];
                    m1[M_] = mo[q];
                });
                glv::add_cc_g<8>(xv, yv, sv, qm1);
                glv::sub_lc_g<8>(xv, yv, dv);
                glv::mul_canon_g<8>(sv, m0, p0, qm1);
                glv::mul_canon_g<8>(dv, m1, p1, qm1);
                glv::add_cc_g<8>(p0, p1, ov, qm1);
                static_for<8>([&](auto M_) __attribute__((always_inline)) { o[h * N + idx1<LOGN>(L, k0 + M_)] = ov[M_]; });
            });
        }
        poly_sync<S::WAVES>();
    }
}

// Batched trace (trace_evaluation.cpp:135-162): o = in / N; for t = N+1, N/2+1, ..., 3:
// o += INTT(Auto_t(o)); out = NTT(o).  in coefficients, out NTT domain.  keys: log2 N key blocks,
// block j-1 for t = 2^j + 1.
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) trace_kernel(KeySwitchArgs A, const uint64_t* in, uint64_t* out,
                                                               int count, const uint64_t* keys) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    const size_t key_words = (size_t)A.digits * 2 * N;
    for (int ct = blockIdx.x; ct < count; ct += gridDim.x) {
        // the lane index re-read per RLWE: its addresses are formed inside the loop, not held in
        // registers across it
        int L = threadIdx.x;
        asm volatile("" : "+v"(L));
        uint64_t o[2][kE];
        const uint64_t* src = in + (size_t)ct * 2 * N;
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            o[0][k] = gl::div_pow2(src[idx1<LOGN>(L, k)], LOGN);
            o[1][k] = gl::div_pow2(src[N + idx1<LOGN>(L, k)], LOGN);
        }
#pragma unroll 1
        for (int j = LOGN; j >= 1; --j) {
            uint64_t a[kE], b[kE];
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                a[k] = o[0][k];
                b[k] = o[1][k];
            }
            automorphism_eval<LOGN>(a, b, lds, (1u << j) + 1, A, keys + (size_t)(j - 1) * key_words, L);
            ntt_inv<LOGN>(a, lds, A.itw, L);
            ntt_inv<LOGN>(b, lds, A.itw, L);
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                o[0][k] = gl::canon(gl::add_lc(a[k], o[0][k]));
                o[1][k] = gl::canon(gl::add_lc(b[k], o[1][k]));
            }
        }
        uint64_t* dst = out + (size_t)ct * 2 * N;
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            uint64_t y[kE];
#pragma unroll
            for (int k = 0; k < kE; ++k) y[k] = h ? o[1][k] : o[0][k];
            ntt_fwd<LOGN>(y, lds, A.tw, L);
            transpose<LOGN, Layout::kPass3, Layout::kPass1>(y, lds, L);  // coalesced writes
#pragma unroll
            for (int k = 0; k < kE; ++k) dst[h * N + idx1<LOGN>(L, k)] = gl::canon(y[k]);
        }
        poly_sync<S::WAVES>();
    }
}

// Packing leaves (lwe_to_rlwe_packing.cpp:238-251): node k = L0Setup(lwes[perm((len-1) ^ brv(k))]):
// a_0 = lwe_0, a_i = -lwe_(N-i), b_0 = lwe_N, all scaled by N^-1, other b coefficients 0.
// Leaves whose input index is >= n_inputs are zero (PackConsecutively's "fake" nodes,
// lwe_to_rlwe_packing.cpp:260-396: merging zeros yields zeros exactly, so the full tree over
// them equals the reference's pruned stack walk).
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) pack_leaf_kernel(const uint64_t* lwes, const uint32_t* perm,
                                                                   uint64_t* nodes, int len, int log_len,
                                                                   int n_inputs) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    const int L = threadIdx.x;
    const uint64_t ninv = gl::inv(N);
    for (int leaf = blockIdx.x; leaf < len; leaf += gridDim.x) {
        uint32_t r = 0, x = (uint32_t)leaf;
        for (int b = 0; b < log_len; ++b) { r = (r << 1) | (x & 1); x >>= 1; }
        uint32_t src = (uint32_t)(len - 1) ^ r;
        uint64_t* dst = nodes + (size_t)leaf * 2 * N;
        if (src >= (uint32_t)n_inputs) {
            // leaf past the inputs (consecutive packing): a zero RLWE, which every merge keeps exact
#pragma unroll
            for (int k = 0; k < kE; ++k) dst[idx1<LOGN>(L, k)] = dst[N + idx1<LOGN>(L, k)] = 0;
            continue;
        }
        if (perm) src = perm[src];
        const uint64_t* lwe = lwes + (size_t)src * (N + 1);
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            const int i = idx1<LOGN>(L, k);
            const uint64_t a = i == 0 ? lwe[0] : gl::neg(lwe[N - i]);
            dst[i] = gl::mul(a, ninv);
            dst[N + i] = i == 0 ? gl::mul(lwe[N], ninv) : 0;
        }
    }
}

// One level of the packing tree: node j of level lvl = (R + X^s L) + tau_(2^lvl+1)(R - X^s L),
// s = N >> lvl, with L, R the level lvl-1 nodes 2j, 2j+1 (stored at leaf slots 2j 2^(lvl-1),
// (2j+1) 2^(lvl-1)); the result goes to the left slot (lwe_to_rlwe_packing.cpp:170-224).
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) pack_merge_kernel(KeySwitchArgs A, uint64_t* nodes, int merges,
                                                                    int lvl, const uint64_t* ksk) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    const int L = threadIdx.x;
    const uint32_t shift = (uint32_t)N >> lvl;
    for (int j = blockIdx.x; j < merges; j += gridDim.x) {
        uint64_t* left = nodes + ((size_t)(2 * j) << (lvl - 1)) * 2 * N;
        const uint64_t* right = nodes + ((size_t)(2 * j + 1) << (lvl - 1)) * 2 * N;
        uint64_t p[2][kE], m[2][kE];
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            uint64_t x[kE];
#pragma unroll
            for (int k = 0; k < kE; ++k) x[k] = left[h * N + idx1<LOGN>(L, k)];
            rotate_regs<LOGN>(x, lds, shift, L);
#pragma unroll
            for (int k = 0; k < kE; ++k) {
                const uint64_t r = right[h * N + idx1<LOGN>(L, k)];
                const uint64_t s = gl::add(r, x[k]), d = gl::sub(r, x[k]);
                if (h) { p[1][k] = s; m[1][k] = d; } else { p[0][k] = s; m[0][k] = d; }
            }
        }
        automorphism_eval<LOGN>(m[0], m[1], lds, (1u << lvl) + 1, A, ksk, L);
        ntt_inv<LOGN>(m[0], lds, A.itw, L);
        ntt_inv<LOGN>(m[1], lds, A.itw, L);
        poly_sync<S::WAVES>();
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            left[idx1<LOGN>(L, k)] = gl::canon(gl::add_lc(m[0][k], p[0][k]));
            left[N + idx1<LOGN>(L, k)] = gl::canon(gl::add_lc(m[1][k], p[1][k]));
        }
    }
}

// Trace steps after the tree when len < N: o += INTT(Auto_(2^lvl+1)(o)), lvl = logN .. log len + 1.
template <int LOGN>
__global__ void __launch_bounds__(Shape<LOGN>::P) pack_tail_kernel(KeySwitchArgs A, uint64_t* node, int lvl_stop,
                                                                   const uint64_t* keys) {
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    __shared__ uint64_t lds[S::LDS_WORDS];
    const int L = threadIdx.x;
    const size_t key_words = (size_t)A.digits * 2 * N;
    uint64_t o[2][kE];
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        o[0][k] = node[idx1<LOGN>(L, k)];
        o[1][k] = node[N + idx1<LOGN>(L, k)];
    }
#pragma unroll 1
    for (int lvl = LOGN; lvl > lvl_stop; --lvl) {
        uint64_t a[kE], b[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            a[k] = o[0][k];
            b[k] = o[1][k];
        }
        automorphism_eval<LOGN>(a, b, lds, (1u << lvl) + 1, A, keys + (size_t)(lvl - 1) * key_words, L);
        ntt_inv<LOGN>(a, lds, A.itw, L);
        ntt_inv<LOGN>(b, lds, A.itw, L);
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            o[0][k] = gl::canon(gl::add_lc(a[k], o[0][k]));
            o[1][k] = gl::canon(gl::add_lc(b[k], o[1][k]));
        }
    }
    poly_sync<S::WAVES>();
#pragma unroll
    for (int k = 0; k < kE; ++k) {
        node[idx1<LOGN>(L, k)] = o[0][k];
        node[N + idx1<LOGN>(L, k)] = o[1][k];
    }
}

// ------------------------------------------------------------------ launchers

#define TP_LOGN_SWITCH(LOGN_VAR, CALL)                    \
    switch (LOGN_VAR) {                                   \
        case 10: { constexpr int LG = 10; CALL; break; }  \
        case 11: { constexpr int LG = 11; CALL; break; }  \
        case 12: { constexpr int LG = 12; CALL; break; }  \
        default: return hipErrorInvalidValue;             \
    }

hipError_t launch_automorphism(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                               uint32_t t, const uint64_t* ksk, hipStream_t s) {
    TP_LOGN_SWITCH(logN, hipLaunchKernelGGL(automorphism_kernel<LG>, dim3(count), dim3(Shape<LG>::P), 0, s, A, in,
                                            out, count, t, ksk));
    return hipGetLastError();
}

hipError_t launch_rlwe_keyswitch(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                                uint32_t t, int mode, int accumulate, const uint64_t* ksk, hipStream_t s) {
    TP_LOGN_SWITCH(logN, hipLaunchKernelGGL(rlwe_keyswitch_kernel<LG>, dim3(count), dim3(Shape<LG>::P), 0, s, A, in,
                                            out, count, t, mode, accumulate, ksk));
    return hipGetLastError();
}

hipError_t launch_auto_ntt(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                           uint32_t t, int accumulate, const uint64_t* ksk, hipStream_t s) {
    TP_LOGN_SWITCH(logN, hipLaunchKernelGGL(auto_ntt_kernel<LG>, dim3(count), dim3(Shape<LG>::P), 0, s, A, in, out,
                                            count, t, accumulate, ksk));
    return hipGetLastError();
}

hipError_t launch_sap_tree(int logN, const KeySwitchArgs& A, const uint64_t* X, uint64_t* out, const uint64_t* polys0,
                           const uint64_t* polys1, int nodes, int count, uint32_t t, int mode, const uint64_t* ksk,
                           hipStream_t s) {
    const long long total = (long long)nodes * count;
    if (total <= 0) return hipSuccess;
    if (total > 0x7fffffff) return hipErrorInvalidValue;
    TP_LOGN_SWITCH(logN, hipLaunchKernelGGL(sap_tree_kernel<LG>, dim3((unsigned)total), dim3(Shape<LG>::P), 0, s, A, X,
                                            out, polys0, polys1, nodes, count, t, mode, ksk));
    return hipGetLastError();
}

hipError_t launch_trace(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                        const uint64_t* keys, hipStream_t s) {
    TP_LOGN_SWITCH(logN, hipLaunchKernelGGL(trace_kernel<LG>, dim3(count), dim3(Shape<LG>::P), 0, s, A, in, out,
                                            count, keys));
    return hipGetLastError();
}

hipError_t launch_pack(int logN, const KeySwitchArgs& A, const uint64_t* lwes, const uint32_t* perm, int len,
                       int n_inputs, uint64_t* nodes, const uint64_t* keys, hipStream_t s) {
    int log_len = 0;
    while ((1 << log_len) < len) ++log_len;
    const size_t key_words = (size_t)A.digits * 2 * ((size_t)1 << logN);
    TP_LOGN_SWITCH(logN, {
        hipLaunchKernelGGL(pack_leaf_kernel<LG>, dim3(len), dim3(Shape<LG>::P), 0, s, lwes, perm, nodes, len, log_len,
                           n_inputs);
        for (int lvl = 1; lvl <= log_len; ++lvl)
            hipLaunchKernelGGL(pack_merge_kernel<LG>, dim3(len >> lvl), dim3(Shape<LG>::P), 0, s, A, nodes,
                               len >> lvl, lvl, keys + (size_t)(lvl - 1) * key_words);
        if (log_len < logN)
            hipLaunchKernelGGL(pack_tail_kernel<LG>, dim3(1), dim3(Shape<LG>::P), 0, s, A, nodes, log_len, keys);
    });
    return hipGetLastError();
}

}  // namespace tp
