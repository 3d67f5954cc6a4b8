// Negacyclic Goldilocks NTT of one polynomial held in registers by P = N/16 lanes
// (one wave for N = 1024, two waves for N = 2048), 16 coefficients per lane.
//
// Result convention (identical to the reference, lib/src/backend/solinas.cpp:228-313):
//   forward:  natural order in, out[k] = a(psi^(2 brv(k) + 1)) (bit-reversed order) out,
//             psi = 7^((Q-1)/2N);
//   inverse:  bit-reversed in, natural out, scaled by N^-1.
// The radix-2 Cooley-Tukey levels are grouped into three register passes with two
// LDS transposes in between:
//   pass 1  levels m = 1..8     lane L holds idx = L + P k          (twiddles are powers of 2)
//   pass 2  levels m = 16..128  lane L holds idx = (L>>lb)<<(lb+4) | k<<lb | L&(2^lb-1)
//   pass 3  the last lb levels  lane L holds idx = 16 L + k
// with lb = log2(N) - 8.  The inverse runs the same passes backwards (Gentleman-Sande).
// A level with m blocks uses twiddle psi_rev[m + block] (psi_inv_rev for the inverse).
#pragma once
#include <hip/hip_runtime.h>
#include <utility>
#include "gl64.hpp"
#include "gl_valu.hpp"

namespace tp {

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>): an unrolled loop whose index
// is a compile-time constant in every copy, whatever the unroller decides.  Register arrays
// indexed by it stay in registers (a loop the unroller gives up on would index them at run time
// and move them to scratch memory).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int kE = 16;  // coefficients per lane
// Butterflies per instruction group: the eight of a level's lane share each op, so a carry is read
// eight instructions after it is written (the compiler adds a wait state between an inline-asm
// carry write and a read fewer than five instructions later).
constexpr int kGW = 8;

template <int LOGN>
struct Shape {
    static constexpr int N = 1 << LOGN;
    static constexpr int P = N / kE;        // lanes per polynomial
    static constexpr int LB = LOGN - 8;     // levels in pass 3
    static constexpr int WAVES = P / 64;    // waves per polynomial
    static constexpr int LDS_WORDS = N + N / 16;
};

// LDS address of logical index idx: one pad word every 16 breaks the power-of-two strides.
__device__ __forceinline__ int phys(int idx) { return idx + (idx >> 4); }

template <int LOGN>
__device__ __forceinline__ int idx1(int L, int k) { return L + (k << (LOGN - 4)); }
template <int LOGN>
__device__ __forceinline__ int idx2(int L, int k) {
    constexpr int lb = LOGN - 8;
    return ((L >> lb) << (lb + 4)) | (k << lb) | (L & ((1 << lb) - 1));
}
__device__ __forceinline__ int idx3(int L, int k) { return (L << 4) | k; }

// Order LDS traffic between the lanes that share one polynomial.
template <int WAVES>
__device__ __forceinline__ void poly_sync() {
    if constexpr (WAVES == 1) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

enum class Layout { kPass1, kPass2, kPass3 };

template <int LOGN, Layout LY>
__device__ __forceinline__ int layout_idx(int L, int k) {
    if constexpr (LY == Layout::kPass1) return idx1<LOGN>(L, k);
    else if constexpr (LY == Layout::kPass2) return idx2<LOGN>(L, k);
    else return idx3(L, k);
}

// Move the polynomial from layout A to layout B through LDS.  A sync precedes the
// writes so that an earlier reader of the same buffer is done with it.
template <int LOGN, Layout A, Layout B>
__device__ __forceinline__ void transpose(uint64_t (&x)[kE], uint64_t* lds, int L) {
    using S = Shape<LOGN>;
    poly_sync<S::WAVES>();
#pragma unroll
    for (int k = 0; k < kE; ++k) lds[phys(layout_idx<LOGN, A>(L, k))] = x[k];
    poly_sync<S::WAVES>();
#pragma unroll
    for (int k = 0; k < kE; ++k) x[k] = lds[phys(layout_idx<LOGN, B>(L, k))];
}

// Butterflies, written with the instruction-level primitives of gl_valu.hpp.
// Forward (Cooley-Tukey): u and v lazy, the product t = v w canonical, u + t and u - t lazy.
// Inverse (Gentleman-Sande): every value canonical (inputs made canonical on entry).
// qm1 = Q - 1, the comparison constant of the canonical forms, lives in an SGPR pair.
__device__ __forceinline__ void ct_bfly_t(uint64_t& u, uint64_t& v, uint64_t t, bool negate) {
    const uint64_t a = u;
    const uint64_t s = glv::add_lc(a, t), d = glv::sub_lc(a, t);
    u = negate ? d : s;  // u + (-t) = u - t: swap where the results go
    v = negate ? s : d;
}

__device__ __forceinline__ void ct_bfly_pow2(uint64_t& u, uint64_t& v, int e) {
    bool ng;
    const uint64_t t = glv::mul_pow2(v, e, ng, glv::qm1_const());
    ct_bfly_t(u, v, t, ng);
}

// First forward level on inputs below 2^16 (unsigned gadget digits): its twiddle is 2^48
// (gl::kFwdPow2[1]), so u + v 2^48 < 2^64 is the word (u, v << 16) and u - v 2^48 is represented
// exactly by u + Q - v 2^48, the word (u + 1, ~(v << 16)) (Q = (1, 2^32 - 1) in 32-bit halves).
__device__ __forceinline__ void ct_bfly_small_first(uint64_t& u, uint64_t& v) {
    const uint32_t a = (uint32_t)u, b = (uint32_t)v << 16;
    u = glv::pk(a, b);
    v = glv::pk(a + 1, ~b);
}

// First forward level on signed gadget digits held as u - delta with 0 <= u < 2^16: the pair
// (u_a - delta) +- (u_b - delta) 2^48 is formed from the small u's (as above) plus two constants of
// the digit, kp = -delta (1 + 2^48) and km = delta (2^48 - 1) mod Q (both canonical).
__device__ __forceinline__ void ct_bfly_signed_first(uint64_t& u, uint64_t& v, uint64_t kp, uint64_t km) {
    uint64_t a = u, b = v;
    ct_bfly_small_first(a, b);
    u = glv::add_lc(a, kp);
    v = glv::add_lc(b, km);
}

// The two constants of ct_bfly_signed_first for a digit offset delta < Q.
__device__ __forceinline__ void signed_first_consts(uint64_t delta, uint64_t& kp, uint64_t& km) {
    const uint64_t d48 = gl::mul_pow2(delta, 48);
    kp = gl::neg(gl::add(delta, d48));
    km = gl::sub(d48, delta);
}

__device__ __forceinline__ void ct_bfly_mul(uint64_t& u, uint64_t& v, uint64_t w) {
    ct_bfly_t(u, v, glv::mul_canon(v, w, glv::qm1_const()), false);
}

// canonical u, v: (u + v, (u - v) w)
__device__ __forceinline__ void gs_bfly_mul(uint64_t& u, uint64_t& v, uint64_t w) {
    const uint64_t qm1 = glv::qm1_const();
    const uint64_t a = u, b = v;
    u = glv::add_cc(a, b, qm1);
    v = glv::mul_canon(glv::sub_cc(a, b), w, qm1);
}

// Sign of 2^e as mul_pow2 reports it (2^96 = -1; 2^e' = -2^-(96 - e') for 64 < e' < 96).
constexpr bool pow2_negative(int e) { return (e >= 96) != (e % 96 > 64); }

// canonical u, v: (u + v, (u - v) 2^e); a negative 2^e swaps the operands of the difference.
__device__ __forceinline__ void gs_bfly_pow2(uint64_t& u, uint64_t& v, int e) {
    const uint64_t qm1 = glv::qm1_const();
    const uint64_t a = u, b = v;
    u = glv::add_cc(a, b, qm1);
    const bool neg = pow2_negative(e);
    const uint64_t diff = neg ? glv::sub_cc(b, a) : glv::sub_cc(a, b);
    bool ng;
    v = glv::mul_pow2(diff, e, ng, qm1);
}

// Lazy -> canonical (3 VALU): x >= Q becomes x + eps (mod 2^64) = x - Q.
__device__ __forceinline__ uint64_t canon_v(uint64_t x) {
    glv::lmask c;
    return glv::mad_eps(glv::cnd01(glv::ge_q(x, glv::qm1_const())), x, c);
}

// Twiddle tables are read inside the step loops rather than held in registers across them.
__device__ __forceinline__ const uint64_t* opaque(const uint64_t* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// Same, but ordered after the value v is computed: a level's twiddles are fetched when the
// previous level is done, so only one level's twiddles are live at a time.
__device__ __forceinline__ const uint64_t* opaque_after(const uint64_t* p, uint64_t v) {
    asm volatile("" : "+s"(p) : "v"(v));
    return p;
}

// Global-memory views of table pointers.  The pointers above come out of an asm statement, which
// hides that they point to global memory; without the cast the loads would be flat loads, whose
// completion is counted with the LDS operations' (every LDS wait would also wait for them).
typedef const __attribute__((address_space(1))) char g_char;
typedef const __attribute__((address_space(1))) uint64_t g_u64;

// Element idx of a table at a wave-uniform base: a 32-bit byte offset lets the load use the
// scalar base register plus one offset register, with the constant part in the immediate.
__device__ __forceinline__ uint64_t ld_u64(const uint64_t* base, uint32_t idx) {
    return *(g_u64*)((g_char*)base + (size_t)(idx * 8u));
}

// Two adjacent words (16-byte aligned) of a global table: pair i = words 2i, 2i + 1.
struct alignas(16) U64x2 {
    uint64_t x, y;
};
typedef const __attribute__((address_space(1))) U64x2 g_u64x2;
__device__ __forceinline__ U64x2 ld_pair(const uint64_t* base, uint32_t i) {
    g_u64x2* p = (g_u64x2*)((g_char*)base + (size_t)(i * 16u));
    U64x2 r;
    r.x = p->x;
    r.y = p->y;
    return r;
}

// The twiddles of one transform level: w[b] = table[first + b], b < nb (nb <= 8), issued together
// ahead of the level's butterflies (one memory latency per level instead of one per load).
template <int NB>
__device__ __forceinline__ void ld_level(uint64_t (&w)[8], const uint64_t* t, uint32_t first) {
#pragma unroll
    for (int b = 0; b < NB; ++b) w[b] = ld_u64(t, first + b);
    __builtin_amdgcn_sched_barrier(0);
}

// ---------------------------------------------------------------- level helpers
//
// A transform level holds 8 butterflies per lane; butterfly b pairs x[bf_u(H, b)] with
// x[bf_u(H, b) + H] (H = half distance).  The butterflies are issued as groups of members of the
// gl_valu.hpp primitives (op by op across the group), never one after the other.

constexpr int bf_u(int H, int b) { return (b / H) * 2 * H + b % H; }

// Butterflies of one level whose products by 2^e are of one kind (gl_valu.hpp P2), with the
// shift argument of each.
struct KindList {
    int n;
    int b[8];
    int arg[8];
};
template <class E>
constexpr KindList kind_list(glv::P2 kind) {
    KindList kl{0, {}, {}};
    for (int b = 0; b < 8; ++b) {
        const glv::Pow2Kind pk = glv::pow2_kind(E::e(b));
        if (pk.kind == kind) {
            kl.b[kl.n] = b;
            kl.arg[kl.n] = pk.arg;
            ++kl.n;
        }
    }
    return kl;
}

// t[b] = v[b] 2^(E::e(b)) up to the sign glv::pow2_kind(E::e(b)).neg, canonical, for the 8
// butterflies of a level, one group per kind.
template <class E, glv::P2 K>
__device__ __forceinline__ void pow2_kind_group(const uint64_t (&v)[8], uint64_t (&t)[8], uint64_t qm1) {
    constexpr KindList kl = kind_list<E>(K);
    if constexpr (kl.n > 0) {
        constexpr int n = kl.n;
        uint64_t xv[n], r[n];
        int arg[n];
        static_for<n>([&](auto M_) __attribute__((always_inline)) {
            constexpr int m = M_;
            xv[m] = v[kl.b[m]];
            arg[m] = kl.arg[m];
        });
        if constexpr (K == glv::P2::kLo) glv::mul_2e_lo_g<n>(xv, arg, r, qm1);
        else if constexpr (K == glv::P2::kMid) glv::mul_2e_mid_g<n>(xv, arg, r, qm1);
        else glv::mul_2e_neg_g<n>(xv, arg, r);
        static_for<n>([&](auto M_) __attribute__((always_inline)) {
            constexpr int m = M_;
            t[kl.b[m]] = r[m];
        });
    }
}
template <class E>
__device__ __forceinline__ void pow2_products(const uint64_t (&v)[8], uint64_t (&t)[8], uint64_t qm1) {
    pow2_kind_group<E, glv::P2::kLo>(v, t, qm1);
    pow2_kind_group<E, glv::P2::kMid>(v, t, qm1);
    pow2_kind_group<E, glv::P2::kNeg>(v, t, qm1);
    // twiddle 1: the product is v itself, made canonical for the butterfly's sum and difference
    constexpr KindList one = kind_list<E>(glv::P2::kOne);
    if constexpr (one.n > 0) {
        uint64_t xv[one.n];
        static_for<one.n>([&](auto M_) __attribute__((always_inline)) { xv[M_] = v[one.b[M_]]; });
        glv::canon_g<one.n>(xv, qm1);
        static_for<one.n>([&](auto M_) __attribute__((always_inline)) { t[one.b[M_]] = xv[M_]; });
    }
}

// Cooley-Tukey outputs of the 8 butterflies from their canonical products t (neg(b): the
// product is -t[b], so the sum and difference trade places); two groups of 4.
template <int H, class NEG>
__device__ __forceinline__ void ct_outputs(uint64_t (&x)[kE], const uint64_t (&t)[8]) {
    static_for<8 / kGW>([&](auto G) __attribute__((always_inline)) {
        uint64_t u[kGW], tt[kGW], s_[kGW], d_[kGW];
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) {
            constexpr int b = kGW * G + M_;
            u[M_] = x[bf_u(H, b)];
            tt[M_] = t[b];
        });
        glv::add_lc_g<kGW>(u, tt, s_);
        glv::sub_lc_g<kGW>(u, tt, d_);
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) {
            constexpr int b = kGW * G + M_;
            constexpr bool ng = NEG::neg(b);
            x[bf_u(H, b)] = ng ? d_[M_] : s_[M_];
            x[bf_u(H, b) + H] = ng ? s_[M_] : d_[M_];
        });
    });
}

// ---------------------------------------------------------------- forward

template <int LV>
struct FwdP1 {
    static constexpr int H = 8 >> LV;
    static constexpr int e(int b) { return gl::kFwdPow2[(1 << LV) + (bf_u(H, b) >> (4 - LV))]; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};

template <int LOGN, int FIRST_LV = 0>
__device__ __forceinline__ void fwd_pass1(uint64_t (&x)[kE]) {
    const uint64_t qm1 = glv::qm1_const();
    static_for<4 - FIRST_LV>([&](auto I) __attribute__((always_inline)) {
        constexpr int lv = FIRST_LV + I;
        using E = FwdP1<lv>;
        uint64_t v[8], t[8];
        static_for<8>([&](auto B) __attribute__((always_inline)) { v[B] = x[bf_u(E::H, B) + E::H]; });
        pow2_products<E>(v, t, qm1);
        ct_outputs<E::H, E>(x, t);
    });
}

// A level with table twiddles: butterfly b uses w[bf_u(H, b) >> SS]; products in two groups of 4.
struct NoNeg {
    static constexpr bool neg(int) { return false; }
};
template <int H, int SS>
__device__ __forceinline__ void ct_level_tw(uint64_t (&x)[kE], const uint64_t (&w)[8]) {
    const uint64_t qm1 = glv::qm1_const();
    uint64_t t[8];
    static_for<8 / kGW>([&](auto G) __attribute__((always_inline)) {
        uint64_t v[kGW], ww[kGW], r[kGW];
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) {
            constexpr int b = kGW * G + M_;
            v[M_] = x[bf_u(H, b) + H];
            ww[M_] = w[bf_u(H, b) >> SS];
        });
        glv::mul_canon_g<kGW>(v, ww, r, qm1);
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) { t[kGW * G + M_] = r[M_]; });
    });
    ct_outputs<H, NoNeg>(x, t);
}

// Levels 4 .. LOGN - 1 take their twiddles from the table: level lv of pass 2 (lv < 8, layout
// idx2) has 2^(lv - 4) distinct blocks per lane, level lv of pass 3 (layout idx3) 16 >> (LOGN - lv);
// butterfly k of the lane uses w[tw_slot(k)] of the level's list.
template <int LOGN, int LV>
struct Level {
    static constexpr int NB = LV < 8 ? (1 << (LV - 4)) : (16 >> (LOGN - LV));
    static constexpr int HALF = LV < 8 ? (8 >> (LV - 4)) : (1 << (LOGN - 1 - LV));
    static constexpr int SLOT_SHIFT = LV < 8 ? 8 - LV : LOGN - LV;
    __device__ static uint32_t first(int L) {
        if constexpr (LV < 8) return (1u << LV) + (((uint32_t)L >> (LOGN - 8)) << (LV - 4));
        else return (1u << LV) + (((uint32_t)L << 4) >> (LOGN - LV));
    }
};

// Natural-order coefficients (pass-1 layout, lazy or canonical) -> lazy NTT values in
// pass-3 layout.  FIRST_LV = 1 starts after the first level (a caller that has done it).  Each
// table level's twiddles are requested while the level before it computes (level 4's before the
// first transpose), so their memory latency overlaps arithmetic.
struct NoPrefetch {
    __device__ void operator()() const {}
};

// `at_last` runs at the start of the last level (a caller's loads for what follows the transform
// are then in flight during that level).
template <int LOGN, int FIRST_LV = 0, class AtLast = NoPrefetch>
__device__ __forceinline__ void ntt_forward(uint64_t (&x)[kE], uint64_t* lds, const uint64_t* tw, int L,
                                            AtLast&& at_last = AtLast{}) {
    tw = opaque(tw);
    fwd_pass1<LOGN, FIRST_LV>(x);
    uint64_t w[8];
    ld_level<Level<LOGN, 4>::NB>(w, opaque_after(tw, x[0]), Level<LOGN, 4>::first(L));
    transpose<LOGN, Layout::kPass1, Layout::kPass2>(x, lds, L);
    static_for<LOGN - 4>([&](auto I) __attribute__((always_inline)) {
        constexpr int lv = 4 + I;
        using Lv = Level<LOGN, lv>;
        if constexpr (lv == 8) transpose<LOGN, Layout::kPass2, Layout::kPass3>(x, lds, L);
        uint64_t wn[8];
        if constexpr (lv + 1 < LOGN)
            ld_level<Level<LOGN, lv + 1>::NB>(wn, opaque_after(tw, x[0]), Level<LOGN, lv + 1>::first(L));
        else
            at_last();
        ct_level_tw<Lv::HALF, Lv::SLOT_SHIFT>(x, w);
        if constexpr (lv + 1 < LOGN) {
#pragma unroll
            for (int b = 0; b < 8; ++b) w[b] = wn[b];
        }
    });
}

// ---------------------------------------------------------------- forward, N = 1024, twisted
//
// After the four power-of-two levels, block j (64 consecutive NTT indices) holds a polynomial B_j
// modulo X^64 - rho_j^64, rho_j = psi^(2 brv4(j) + 1), and its outputs are B_j(rho_j w^brv6(t)),
// w = psi^32 = 7^((Q-1)/64), a 64th root of unity: a cyclic 64-point transform of the coefficients
// times rho_j^i.  So the six levels with table twiddles become one product per coefficient (the
// twist, table) and a cyclic transform whose twiddles w^(brv_t(m) (32 >> t)) are powers of two
// (w = 2^39): its first four levels run with shift-based products on compile-time exponents, its
// last two (where a lane's sub-blocks differ from lane to lane) with table twiddles.  The outputs
// are the same words in the same places as ntt_forward's.

// Exponents of 2 of the cyclic 64-point transform's twiddles, level t < 4, sub-block m.
constexpr int kDft64Pow2[4][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0}, {0, 48, 0, 0, 0, 0, 0, 0}, {0, 48, 120, 168, 0, 0, 0, 0},
    {0, 48, 120, 168, 156, 12, 84, 132}};

// Level t of the cyclic transform in the pass-2 layout (lane L holds i = 4 k + (L & 3) of block
// L >> 2): distance 8 >> t in k, sub-block k >> (4 - t).
template <int T>
struct DftP2 {
    static constexpr int H = 8 >> T;
    static constexpr int e(int b) { return kDft64Pow2[T][bf_u(H, b) >> (4 - T)]; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};

// t64: [0, N) twist factors rho_j^i in key-row order (lane L, pair c: words 2 (64 c + L) + {0, 1}
// for k = 2c, 2c + 1), [N, 2N) the twiddles of the cyclic transform's last two levels at the
// indices ntt_forward reads its levels 8 and 9 from.
template <int FIRST_LV = 0, class AtLast = NoPrefetch>
__device__ __forceinline__ void ntt_forward_t64(uint64_t (&x)[kE], uint64_t* lds, const uint64_t* t64, int L,
                                                AtLast&& at_last = AtLast{}) {
    constexpr int LOGN = 10;
    constexpr int N = 1 << LOGN;
    const uint64_t qm1 = glv::qm1_const();
    t64 = opaque(t64);
    fwd_pass1<LOGN, FIRST_LV>(x);
    // twist factors requested before the transpose
    const uint64_t* tf = opaque_after(t64, x[0]);
    U64x2 rf[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) rf[c] = ld_pair(tf, c * 64 + L);
    __builtin_amdgcn_sched_barrier(0);
    transpose<LOGN, Layout::kPass1, Layout::kPass2>(x, lds, L);
    {
        uint64_t r[kE], rr[kE];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            rr[2 * c] = rf[c].x;
            rr[2 * c + 1] = rf[c].y;
        }
        static_for<2>([&](auto G) __attribute__((always_inline)) {
            uint64_t a8[8], b8[8], o8[8];
            static_for<8>([&](auto M_) __attribute__((always_inline)) {
                a8[M_] = x[8 * G + M_];
                b8[M_] = rr[8 * G + M_];
            });
            glv::mul_canon_g<8>(a8, b8, o8, qm1);
            static_for<8>([&](auto M_) __attribute__((always_inline)) { r[8 * G + M_] = o8[M_]; });
        });
#pragma unroll
        for (int k = 0; k < kE; ++k) x[k] = r[k];
    }
    // level t = 0: twiddle 1 on canonical values
    {
        using E = DftP2<0>;
        uint64_t t[8];
        static_for<8>([&](auto B) __attribute__((always_inline)) { t[B] = x[bf_u(E::H, B) + E::H]; });
        ct_outputs<E::H, E>(x, t);
    }
    uint64_t w[8];
    ld_level<Level<LOGN, 8>::NB>(w, opaque_after(t64 + N, x[0]), Level<LOGN, 8>::first(L));
    static_for<3>([&](auto I) __attribute__((always_inline)) {
        using E = DftP2<1 + I>;
        uint64_t v[8], t[8];
        static_for<8>([&](auto B) __attribute__((always_inline)) { v[B] = x[bf_u(E::H, B) + E::H]; });
        pow2_products<E>(v, t, qm1);
        ct_outputs<E::H, E>(x, t);
    });
    transpose<LOGN, Layout::kPass2, Layout::kPass3>(x, lds, L);
    {
        using Lv = Level<LOGN, 8>;
        uint64_t wn[8];
        ld_level<Level<LOGN, 9>::NB>(wn, opaque_after(t64 + N, x[0]), Level<LOGN, 9>::first(L));
        ct_level_tw<Lv::HALF, Lv::SLOT_SHIFT>(x, w);
#pragma unroll
        for (int b = 0; b < 8; ++b) w[b] = wn[b];
    }
    {
        using Lv = Level<LOGN, 9>;
        at_last();
        ct_level_tw<Lv::HALF, Lv::SLOT_SHIFT>(x, w);
    }
}

// The forward transform the kernels run: the twisted form at N = 1024 (tw then holds its tables
// after psi_rev, N words in), ntt_forward otherwise.
template <int LOGN, int FIRST_LV = 0, class AtLast = NoPrefetch>
__device__ __forceinline__ void ntt_fwd(uint64_t (&x)[kE], uint64_t* lds, const uint64_t* tw, int L,
                                        AtLast&& at_last = AtLast{}) {
    if constexpr (LOGN == 10) ntt_forward_t64<FIRST_LV>(x, lds, tw + (1 << LOGN), L, at_last);
    else ntt_forward<LOGN, FIRST_LV>(x, lds, tw, L, at_last);
}

// ---------------------------------------------------------------- inverse

// Gentleman-Sande level on canonical values: x[u] = x[u] + x[v], and the difference (operands
// swapped where the product by the twiddle is negative) goes to d[b] for the caller's products.
template <int H, class NEG>
__device__ __forceinline__ void gs_sums_diffs(uint64_t (&x)[kE], uint64_t (&d)[8], uint64_t qm1) {
    static_for<8 / kGW>([&](auto G) __attribute__((always_inline)) {
        uint64_t u[kGW], v[kGW], s_[kGW], a_[kGW], b_[kGW], r[kGW];
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) {
            constexpr int b = kGW * G + M_;
            u[M_] = x[bf_u(H, b)];
            v[M_] = x[bf_u(H, b) + H];
            a_[M_] = NEG::neg(b) ? v[M_] : u[M_];
            b_[M_] = NEG::neg(b) ? u[M_] : v[M_];
        });
        glv::add_cc_g<kGW>(u, v, s_, qm1);
        glv::sub_lc_g<kGW>(a_, b_, r);
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) {
            constexpr int b = kGW * G + M_;
            x[bf_u(H, b)] = s_[M_];
            d[b] = r[M_];
        });
    });
}

template <int LV>
struct InvP1 {
    static constexpr int H = 8 >> LV;
    static constexpr int e(int b) { return gl::kInvPow2[(1 << LV) + (bf_u(H, b) >> (4 - LV))]; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};

// The last inverse level with N^-1 = 2^-S merged in (S = LOGN, or 0 where a table carries N^-1):
// x[k] = (a + b) 2^-S and x[k + 8] = (a - b) 2^(e1 - S), e1 the level's exponent.
template <int S>
struct InvLast {
    static constexpr int e(int) { return (gl::kInvPow2[1] + 192 - S) % 192; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};

// Levels m = 8..1 with power-of-two inverse twiddles on canonical values; with SCALE,
// N^-1 = 2^(192 - LOGN) is merged into the last level (both of its factors are powers of two).
template <int LOGN, bool SCALE = true>
__device__ __forceinline__ void inv_pass1(uint64_t (&x)[kE]) {
    const uint64_t qm1 = glv::qm1_const();
    static_for<3>([&](auto I) __attribute__((always_inline)) {
        constexpr int lv = 3 - I;
        using E = InvP1<lv>;
        uint64_t d[8], t[8];
        gs_sums_diffs<E::H, E>(x, d, qm1);
        pow2_products<E>(d, t, qm1);
        static_for<8>([&](auto B) __attribute__((always_inline)) { x[bf_u(E::H, B) + E::H] = t[B]; });
    });
    using E = InvLast<SCALE ? LOGN : 0>;
    uint64_t d[8], t[8];
    gs_sums_diffs<8, E>(x, d, qm1);
    if constexpr (SCALE) {
        static_for<2>([&](auto G) __attribute__((always_inline)) {
            uint64_t s_[4], r[4];
            int sh[4];
            static_for<4>([&](auto M_) __attribute__((always_inline)) {
                s_[M_] = x[4 * G + M_];
                sh[M_] = LOGN;
            });
            glv::mul_2e_neg_g<4>(s_, sh, r);
            static_for<4>([&](auto M_) __attribute__((always_inline)) { x[4 * G + M_] = r[M_]; });
        });
    }
    pow2_products<E>(d, t, qm1);
    static_for<8>([&](auto B) __attribute__((always_inline)) { x[B + 8] = t[B]; });
}

// A Gentleman-Sande level with table twiddles on canonical values: butterfly b uses
// w[bf_u(H, b) >> SS].
template <int H, int SS>
__device__ __forceinline__ void gs_level_tw(uint64_t (&x)[kE], const uint64_t (&w)[8]) {
    const uint64_t qm1 = glv::qm1_const();
    uint64_t d[8];
    gs_sums_diffs<H, NoNeg>(x, d, qm1);
    static_for<8 / kGW>([&](auto G) __attribute__((always_inline)) {
        uint64_t dd[kGW], ww[kGW], r[kGW];
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) {
            constexpr int b = kGW * G + M_;
            dd[M_] = d[b];
            ww[M_] = w[bf_u(H, b) >> SS];
        });
        glv::mul_canon_g<kGW>(dd, ww, r, qm1);
        static_for<kGW>([&](auto M_) __attribute__((always_inline)) { x[bf_u(H, kGW * G + M_) + H] = r[M_]; });
    });
}

// NTT values in pass-3 layout (lazy, or canonical when CANON_IN) -> canonical natural-order
// coefficients in pass-1 layout (table levels LOGN - 1 .. 4 with the next level's twiddles
// requested during the current one).
template <int LOGN, bool CANON_IN = false>
__device__ __forceinline__ void ntt_inverse(uint64_t (&x)[kE], uint64_t* lds, const uint64_t* itw, int L) {
    if constexpr (!CANON_IN) glv::canon_g<kE>(x, glv::qm1_const());
    itw = opaque(itw);
    uint64_t w[8];
    ld_level<Level<LOGN, LOGN - 1>::NB>(w, opaque_after(itw, x[0]), Level<LOGN, LOGN - 1>::first(L));
    static_for<LOGN - 4>([&](auto I) __attribute__((always_inline)) {
        constexpr int lv = LOGN - 1 - I;
        using Lv = Level<LOGN, lv>;
        if constexpr (lv == 7) transpose<LOGN, Layout::kPass3, Layout::kPass2>(x, lds, L);
        uint64_t wn[8];
        if constexpr (lv > 4)
            ld_level<Level<LOGN, lv - 1>::NB>(wn, opaque_after(itw, x[0]), Level<LOGN, lv - 1>::first(L));
        gs_level_tw<Lv::HALF, Lv::SLOT_SHIFT>(x, w);
        if constexpr (lv > 4) {
#pragma unroll
            for (int b = 0; b < 8; ++b) w[b] = wn[b];
        }
    });
    transpose<LOGN, Layout::kPass2, Layout::kPass1>(x, lds, L);
    inv_pass1<LOGN>(x);
}

// ---------------------------------------------------------------- inverse, N = 1024, twisted
//
// The inverse of ntt_forward_t64: the cyclic transform's last two levels with table twiddles, its
// first four Gentleman-Sande levels with inverse power-of-two twiddles w^-(brv_t(m) (32 >> t)),
// then the untwist rho_j^-i N^-1 (table: N^-1 rides on the untwist, so the last power-of-two level
// scales nothing).  Six unscaled Gentleman-Sande levels are 2^6 times the inverse of the six
// forward levels, as are the cyclic transform's, so no factor moves.

template <int T>
struct InvDftP2 {
    static constexpr int H = 8 >> T;
    static constexpr int e(int b) { return (192 - kDft64Pow2[T][bf_u(H, b) >> (4 - T)]) % 192; }
    static constexpr bool neg(int b) { return glv::pow2_kind(e(b)).neg; }
};

// it64: [0, N) rho_j^-i N^-1 in the forward's pair order, [N, 2N) the inverse twiddles of the cyclic
// transform's last two levels at ntt_inverse's level-9 and level-8 indices.
template <bool CANON_IN = false>
__device__ __forceinline__ void ntt_inverse_t64(uint64_t (&x)[kE], uint64_t* lds, const uint64_t* it64, int L) {
    constexpr int LOGN = 10;
    constexpr int N = 1 << LOGN;
    const uint64_t qm1 = glv::qm1_const();
    if constexpr (!CANON_IN) glv::canon_g<kE>(x, qm1);
    it64 = opaque(it64);
    using L9 = Level<LOGN, 9>;
    using L8 = Level<LOGN, 8>;
    uint64_t w[8];
    ld_level<L9::NB>(w, opaque_after(it64 + N, x[0]), L9::first(L));
    {
        uint64_t wn[8];
        ld_level<L8::NB>(wn, opaque_after(it64 + N, x[0]), L8::first(L));
        gs_level_tw<L9::HALF, L9::SLOT_SHIFT>(x, w);
#pragma unroll
        for (int b = 0; b < 8; ++b) w[b] = wn[b];
    }
    gs_level_tw<L8::HALF, L8::SLOT_SHIFT>(x, w);
    transpose<LOGN, Layout::kPass3, Layout::kPass2>(x, lds, L);
    // untwist factors, in flight during the four levels below
    const uint64_t* tf = opaque_after(it64, x[0]);
    U64x2 rf[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) rf[c] = ld_pair(tf, c * 64 + L);
    __builtin_amdgcn_sched_barrier(0);
    static_for<3>([&](auto I) __attribute__((always_inline)) {
        using E = InvDftP2<3 - I>;
        uint64_t d[8], t[8];
        gs_sums_diffs<E::H, E>(x, d, qm1);
        pow2_products<E>(d, t, qm1);
        static_for<8>([&](auto B) __attribute__((always_inline)) { x[bf_u(E::H, B) + E::H] = t[B]; });
    });
    {
        // level t = 0: twiddle 1, the differences go to the untwist as they are
        uint64_t d[8];
        gs_sums_diffs<8, NoNeg>(x, d, qm1);
        static_for<8>([&](auto B) __attribute__((always_inline)) { x[B + 8] = d[B]; });
    }
    {
        uint64_t rr[kE];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            rr[2 * c] = rf[c].x;
            rr[2 * c + 1] = rf[c].y;
        }
        static_for<2>([&](auto G) __attribute__((always_inline)) {
            uint64_t a8[8], b8[8], o8[8];
            static_for<8>([&](auto M_) __attribute__((always_inline)) {
                a8[M_] = x[8 * G + M_];
                b8[M_] = rr[8 * G + M_];
            });
            glv::mul_canon_g<8>(a8, b8, o8, qm1);
            static_for<8>([&](auto M_) __attribute__((always_inline)) { x[8 * G + M_] = o8[M_]; });
        });
    }
    transpose<LOGN, Layout::kPass2, Layout::kPass1>(x, lds, L);
    inv_pass1<LOGN, false>(x);
}

// The inverse transform the kernels run: the twisted form at N = 1024 (itw then holds its tables
// after psi_inv_rev, N words in), ntt_inverse otherwise.
template <int LOGN, bool CANON_IN = false>
__device__ __forceinline__ void ntt_inv(uint64_t (&x)[kE], uint64_t* lds, const uint64_t* itw, int L) {
    if constexpr (LOGN == 10) ntt_inverse_t64<CANON_IN>(x, lds, itw + (1 << LOGN), L);
    else ntt_inverse<LOGN, CANON_IN>(x, lds, itw, L);
}

}  // namespace tp
