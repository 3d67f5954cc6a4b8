// N = 1024 negacyclic transforms split over two waves: wave q holds the 512 positions whose index
// bit 9 equals q, 8 words per lane.  Same results as ntt_fwd / ntt_inv (the twisted forms of
// ntt_wave.hpp: four power-of-two levels, the twist rho_j^i, a cyclic 64-point transform per block
// j of 64 positions), with every level inside one wave except the inverse's last (bit 9), which
// exchanges halves through LDS.
//
// Layouts of a wave's half (position p = i & 511, lane l, word k):
//   H1  p = l | k << 6                       levels 0..3 act on bits 9..6 (bit 9 = the wave)
//   H2  p = (l & 7) | k << 3 | (l >> 3) << 6   cyclic levels t0..t2 on bits 5..3 (block j = bits 9..6)
//   H3  p = k | l << 3                       cyclic levels t3..t5 on bits 2..0 (8 consecutive)
// The power-of-two levels have compile-time twiddles (the wave's half is a template argument);
// t3..t5, whose sub-blocks vary from lane to lane, read per-lane twiddles.
//
// Tables (N words each, built by HostNtt::half_tables): [0, N) twist rho_j^i at (64 q + l) 8 + k
// for the H2 position of lane l, word k; [N, 2N) the t3/t4/t5 twiddles of lane l at
// (64 q + l) 8 + s: s = 0 for t3, 1 + (k >> 2) for t4, 3 + (k >> 1) for t5.  The inverse tables
// hold the inverses at the same places.
#pragma once
#include "ntt_wave.hpp"

namespace tp {
namespace half {

constexpr int kH = 8;         // words per lane
constexpr int kHalf = 512;    // positions per wave
constexpr int kBufWords = kHalf + kHalf / 16;  // one padded transpose buffer

enum class HL { k1, k2, k3 };

template <HL Y>
__device__ __forceinline__ int pos(int l, int k) {
    if constexpr (Y == HL::k1) return l | (k << 6);
    else if constexpr (Y == HL::k2) return (l & 7) | (k << 3) | ((l >> 3) << 6);
    else return k | (l << 3);
}

__device__ __forceinline__ void wave_sync() { poly_sync<1>(); }

// H_A -> H_B through a wave-private LDS buffer (kBufWords words).
template <HL A, HL B>
__device__ __forceinline__ void transpose(uint64_t (&x)[kH], uint64_t* buf, int l) {
    wave_sync();
#pragma unroll
    for (int k = 0; k < kH; ++k) buf[phys(pos<A>(l, k))] = x[k];
    wave_sync();
#pragma unroll
    for (int k = 0; k < kH; ++k) x[k] = buf[phys(pos<B>(l, k))];
}

// ---------------------------------------------------------------- four-butterfly levels
//
// A level pairs x[bu(H, b)] with x[bu(H, b) + H], b < 4 (H = 4, 2, 1).

constexpr int bu(int H, int b) { return (b / H) * 2 * H + b % H; }

template <class E>
constexpr KindList kinds4(glv::P2 kind) {
    KindList kl{0, {}, {}};
    for (int b = 0; b < 4; ++b) {
        const glv::Pow2Kind pk = glv::pow2_kind(E::e(b));
        if (pk.kind == kind) {
            kl.b[kl.n] = b;
            kl.arg[kl.n] = pk.arg;
            ++kl.n;
        }
    }
    return kl;
}

template <class E, glv::P2 K>
__device__ __forceinline__ void pow2_group4(const uint64_t (&v)[4], uint64_t (&t)[4], uint64_t qm1) {
    constexpr KindList kl = kinds4<E>(K);
    if constexpr (kl.n > 0) {
        constexpr int n = kl.n;
        uint64_t xv[n], r[n];
        int arg[n];
        static_for<n>([&](auto M_) __attribute__((always_inline)) {
            xv[M_] = v[kl.b[M_]];
            arg[M_] = kl.arg[M_];
        });
        if constexpr (K == glv::P2::kLo) glv::mul_2e_lo_g<n>(xv, arg, r, qm1);
        else if constexpr (K == glv::P2::kMid) glv::mul_2e_mid_g<n>(xv, arg, r, qm1);
        else if constexpr (K == glv::P2::kNeg) glv::mul_2e_neg_g<n>(xv, arg, r);
        else {
            static_for<n>([&](auto M_) __attribute__((always_inline)) { r[M_] = xv[M_]; });
            glv::canon_g<n>(r, qm1);
        }
        static_for<n>([&](auto M_) __attribute__((always_inline)) { t[kl.b[M_]] = r[M_]; });
    }
}

// t[b] = v[b] 2^E::e(b) up to the sign E::neg(b), canonical
template <class E>
__device__ __forceinline__ void pow2_products4(const uint64_t (&v)[4], uint64_t (&t)[4], uint64_t qm1) {
    pow2_group4<E, glv::P2::kLo>(v, t, qm1);
    pow2_group4<E, glv::P2::kMid>(v, t, qm1);
    pow2_group4<E, glv::P2::kNeg>(v, t, qm1);
    pow2_group4<E, glv::P2::kOne>(v, t, qm1);
}

// Cooley-Tukey outputs from canonical products t (NEG::neg(b): the product is -t[b]).
template <int H, class NEG>
__device__ __forceinline__ void ct_out4(uint64_t (&x)[kH], const uint64_t (&t)[4]) {
    uint64_t u[4], s_[4], d_[4];
    static_for<4>([&](auto B) __attribute__((always_inline)) { u[B] = x[bu(H, B)]; });
    glv::add_lc_g<4>(u, t, s_);
    glv::sub_lc_g<4>(u, t, d_);
    static_for<4>([&](auto B) __attribute__((always_inline)) {
        constexpr bool ng = NEG::neg(B);
        x[bu(H, B)] = ng ? d_[B] : s_[B];
        x[bu(H, B) + H] = ng ? s_[B] : d_[B];
    });
}

template <class E>
__device__ __forceinline__ void ct_pow2_4(uint64_t (&x)[kH], uint64_t qm1) {
    uint64_t v[4], t[4];
    static_for<4>([&](auto B) __attribute__((always_inline)) { v[B] = x[bu(E::H, B) + E::H]; });
    pow2_products4<E>(v, t, qm1);
    ct_out4<E::H, E>(x, t);
}

// w[b]: the twiddle of butterfly b
template <int H>
__device__ __forceinline__ void ct_tw4(uint64_t (&x)[kH], const uint64_t (&w)[4], uint64_t qm1) {
    uint64_t v[4], t[4];
    static_for<4>([&](auto B) __attribute__((always_inline)) { v[B] = x[bu(H, B) + H]; });
    glv::mul_canon_g<4>(v, w, t, qm1);
    ct_out4<H, NoNeg>(x, t);
}

// Gentleman-Sande on canonical values: x[u] = u + v (canonical), d[b] = u - v (operands swapped
// where NEG::neg(b)), canonical.
template <int H, class NEG>
__device__ __forceinline__ void gs_sd4(uint64_t (&x)[kH], uint64_t (&d)[4], uint64_t qm1) {
    uint64_t u[4], v[4], a_[4], b_[4], s_[4];
    static_for<4>([&](auto B) __attribute__((always_inline)) {
        u[B] = x[bu(H, B)];
        v[B] = x[bu(H, B) + H];
        a_[B] = NEG::neg(B) ? v[B] : u[B];
        b_[B] = NEG::neg(B) ? u[B] : v[B];
    });
    glv::add_cc_g<4>(u, v, s_, qm1);
    glv::sub_lc_g<4>(a_, b_, d);
    static_for<4>([&](auto B) __attribute__((always_inline)) { x[bu(H, B)] = s_[B]; });
}

template <class E>
__device__ __forceinline__ void gs_pow2_4(uint64_t (&x)[kH], uint64_t qm1) {
    uint64_t d[4], t[4];
    gs_sd4<E::H, E>(x, d, qm1);
    pow2_products4<E>(d, t, qm1);
    static_for<4>([&](auto B) __attribute__((always_inline)) { x[bu(E::H, B) + E::H] = t[B]; });
}

template <int H>
__device__ __forceinline__ void gs_tw4(uint64_t (&x)[kH], const uint64_t (&w)[4], uint64_t qm1) {
    uint64_t d[4], t[4];
    gs_sd4<H, NoNeg>(x, d, qm1);
    glv::mul_canon_g<4>(d, w, t, qm1);
    static_for<4>([&](auto B) __attribute__((always_inline)) { x[bu(H, B) + H] = t[B]; });
}

// ---------------------------------------------------------------- exponents
//
// Level t of the power-of-two levels (bit 9 - t; H1: t = 1, 2, 3 -> H = 4, 2, 1) in wave half Q:
// block (Q << (t - 1)) | (k >> (4 - t)) of the level, twiddle psi_rev[2^t + block].
template <int T, int Q>
struct FwdH {
    static constexpr int H = 8 >> T;
    static constexpr int e(int b) { return gl::kFwdPow2[(1 << T) + ((Q << (T - 1)) | (bu(H, b) >> (4 - T)))]; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};
template <int T, int Q>
struct InvH {
    static constexpr int H = 8 >> T;
    static constexpr int e(int b) { return gl::kInvPow2[(1 << T) + ((Q << (T - 1)) | (bu(H, b) >> (4 - T)))]; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};
// Cyclic levels t1, t2 in H2 (H = 2, 1): sub-block k >> (3 - t).
template <int T>
struct DftH {
    static constexpr int H = 4 >> T;
    static constexpr int e(int b) { return kDft64Pow2[T][bu(H, b) >> (3 - T)]; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};
template <int T>
struct InvDftH {
    static constexpr int H = 4 >> T;
    static constexpr int e(int b) { return (192 - kDft64Pow2[T][bu(H, b) >> (3 - T)]) % 192; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};

// per-lane table words: 8 consecutive words at (64 q + l) 8
__device__ __forceinline__ void ld_lane8(uint64_t (&v)[kH], const uint64_t* t, int ql) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const U64x2 p = ld_pair(t, (uint32_t)(ql * 4 + c));
        v[2 * c] = p.x;
        v[2 * c + 1] = p.y;
    }
}

// twist / untwist: x[k] <- x[k] f[k], canonical
__device__ __forceinline__ void scale8(uint64_t (&x)[kH], const uint64_t (&f)[kH], uint64_t qm1) {
    uint64_t o[kH];
    glv::mul_canon_g<kH>(x, f, o, qm1);
#pragma unroll
    for (int k = 0; k < kH; ++k) x[k] = o[k];
}

// ---------------------------------------------------------------- forward
//
// x: H1 values after level 0 (lazy); out: H3 values, lazy, position 512 Q + pos<k3>(l, k).
// tab = the forward tables; buf = the wave's transpose buffer; at_last runs before level t5.
// The wave's priority drops from 3 (set by the caller) to 2 and then 1 as it passes the two
// transposes, so among the waves of a SIMD the one that is furthest behind issues first and
// they finish the phase together.
template <int Q, class AtLast>
__device__ __forceinline__ void forward(uint64_t (&x)[kH], uint64_t* buf, const uint64_t* tab, int l,
                                        AtLast&& at_last) {
    const uint64_t qm1 = glv::qm1_const();
    tab = opaque(tab);
    const int ql = Q * 64 + l;
    ct_pow2_4<FwdH<1, Q>>(x, qm1);
    ct_pow2_4<FwdH<2, Q>>(x, qm1);
    ct_pow2_4<FwdH<3, Q>>(x, qm1);
    uint64_t f[kH];
    ld_lane8(f, opaque_after(tab, x[0]), ql);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(2);  // see forward's comment
    transpose<HL::k1, HL::k2>(x, buf, l);
    scale8(x, f, qm1);
    {
        // t0: twiddle 1 on canonical values
        uint64_t t[4];
        static_for<4>([&](auto B) __attribute__((always_inline)) { t[B] = x[B + 4]; });
        ct_out4<4, NoNeg>(x, t);
    }
    ct_pow2_4<DftH<1>>(x, qm1);
    ct_pow2_4<DftH<2>>(x, qm1);
    uint64_t w[kH];
    ld_lane8(w, opaque_after(tab + 1024, x[0]), ql);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(1);
    transpose<HL::k2, HL::k3>(x, buf, l);
    {
        const uint64_t w3[4] = {w[0], w[0], w[0], w[0]};
        ct_tw4<4>(x, w3, qm1);  // t3: lower words 0..3
        const uint64_t w4[4] = {w[1], w[1], w[2], w[2]};
        ct_tw4<2>(x, w4, qm1);  // t4: lower words 0, 1, 4, 5
        at_last();
        const uint64_t w5[4] = {w[3], w[4], w[5], w[6]};
        ct_tw4<1>(x, w5, qm1);  // t5: lower words 0, 2, 4, 6
    }
}

// ---------------------------------------------------------------- inverse
//
// Part 1 on wave half Q: x = H3 values (position 512 Q + pos<k3>(l, k), any 64-bit), out: the
// values before the last level (bit 9), canonical, H1.  Part 2 (after the caller has exchanged
// halves): lo = position p, hi = position p + 512 (both canonical) -> wave Q's output coefficient
// at 512 Q + p, canonical, with N^-1 applied (the untwist table carries it).
template <int Q>
__device__ __forceinline__ void inverse_part1(uint64_t (&x)[kH], uint64_t* buf, const uint64_t* tab, int l) {
    const uint64_t qm1 = glv::qm1_const();
    glv::canon_g<kH>(x, qm1);
    tab = opaque(tab);
    const int ql = Q * 64 + l;
    uint64_t w[kH];
    ld_lane8(w, opaque_after(tab + 1024, x[0]), ql);
    __builtin_amdgcn_sched_barrier(0);
    {
        const uint64_t w5[4] = {w[3], w[4], w[5], w[6]};
        gs_tw4<1>(x, w5, qm1);
        const uint64_t w4[4] = {w[1], w[1], w[2], w[2]};
        gs_tw4<2>(x, w4, qm1);
        const uint64_t w3[4] = {w[0], w[0], w[0], w[0]};
        gs_tw4<4>(x, w3, qm1);
    }
    uint64_t f[kH];
    ld_lane8(f, opaque_after(tab, x[0]), ql);
    __builtin_amdgcn_sched_barrier(0);
    transpose<HL::k3, HL::k2>(x, buf, l);
    gs_pow2_4<InvDftH<2>>(x, qm1);
    gs_pow2_4<InvDftH<1>>(x, qm1);
    {
        // t0: twiddle 1, the differences go to the untwist as they are
        uint64_t d[4];
        gs_sd4<4, NoNeg>(x, d, qm1);
        static_for<4>([&](auto B) __attribute__((always_inline)) { x[B + 4] = d[B]; });
    }
    scale8(x, f, qm1);
    transpose<HL::k2, HL::k1>(x, buf, l);
    gs_pow2_4<InvH<3, Q>>(x, qm1);
    gs_pow2_4<InvH<2, Q>>(x, qm1);
    gs_pow2_4<InvH<1, Q>>(x, qm1);
}

// The last level: half 0 takes lo + hi, half 1 (lo - hi) 2^e1, e1 the level's exponent (N^-1 is
// already in, from the untwist).
struct InvLast4 {
    static constexpr int e(int) { return gl::kInvPow2[1]; }
    static constexpr bool neg(int) { return glv::pow2_kind(e(0)).neg; }
};
template <int Q>
__device__ __forceinline__ void inverse_part2(uint64_t (&x)[kH], const uint64_t (&lo)[kH], const uint64_t (&hi)[kH]) {
    const uint64_t qm1 = glv::qm1_const();
    static_for<2>([&](auto G) __attribute__((always_inline)) {
        uint64_t a[4], b[4], r[4];
        static_for<4>([&](auto M_) __attribute__((always_inline)) {
            a[M_] = lo[4 * G + M_];
            b[M_] = hi[4 * G + M_];
        });
        if constexpr (Q == 0) {
            glv::add_cc_g<4>(a, b, r, qm1);
        } else {
            uint64_t d[4];
            if constexpr (InvLast4::neg(0)) glv::sub_lc_g<4>(b, a, d);
            else glv::sub_lc_g<4>(a, b, d);
            pow2_group4<InvLast4, glv::pow2_kind(InvLast4::e(0)).kind>(d, r, qm1);
        }
        static_for<4>([&](auto M_) __attribute__((always_inline)) { x[4 * G + M_] = r[M_]; });
    });
}

}  // namespace half
}  // namespace tp
