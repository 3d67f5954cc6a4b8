// Folded complex FFT of one ring element, N = 1024 real coefficients held as n = 512 complex
// doubles by one wave (8 per lane), used by the exact floating-point external product of the
// binary blind rotation (kernels_br_fft.hip).
//
// Folding: z_m = a_m + i a_{m + 512}.  The transform evaluates z(Y) at the 512 roots of
// Y^512 = i, which are the roots psi^(1 + 4j) of X^1024 + 1 (psi = e^(i pi / 1024)): there
// a(psi^(1+4j)) = z(psi^(1+4j)), and the other 512 roots are their conjugates, so for real a the
// 512 values determine the negacyclic product exactly as the 1024-point transform would.
//
// Structure: the in-place Cooley-Tukey factorisation of Y^512 - i, i.e. the first half of the
// tree of the 1024-point negacyclic transform: level s (s = 0..8) has 2^s contiguous blocks of
// 512 >> s positions; block b pairs positions j and j + len / 2 with twiddle
// zeta[2^(s+1) + b], zeta[k] = psi^brv10(k).  Since brv10(2k + 1) = brv10(2k) + 512 and
// psi^512 = i, zeta[2k + 1] = i zeta[2k]: the odd-indexed twiddles are never loaded (a product by
// i is a swap with a sign flip, folded into the butterfly's additions).
//
// Register passes (three levels each) with two LDS transposes:
//   pass 1  levels 0..2  lane L holds positions L + 64 t          (twiddles: 4 wave-uniform values)
//   pass 2  levels 3..5  lane L holds 64 (L >> 3) + (L & 7) + 8 t (4 twiddles per lane)
//   pass 3  levels 6..8  lane L holds 8 L + t                     (4 twiddles per lane)
// (the twiddles of a lane depend on L only: loaded once per kernel, see Tw).
// The inverse (Gentleman-Sande with conjugate twiddles) runs the passes backwards and returns
// 512 z (the 1/512 is applied by the caller when it rounds).
//
// Rounding: each level is a butterfly u +- w v with |w| = 1, the structure of Higham's
// Theorem 24.2 (Accuracy and Stability of Numerical Algorithms, 2nd ed.): the computed
// transform has normwise relative error at most 9 eta / (1 - 9 eta), eta = mu + gamma_4
// (sqrt 2 + mu), mu the twiddle error (correctly rounded table: mu <= u / sqrt 2).
#pragma once
#include <hip/hip_runtime.h>
#include "ntt_wave.hpp"

namespace tp {
namespace cf {

struct alignas(16) C {
    double re, im;
};

constexpr int kN = 512;         // complex points
constexpr int kPer = 8;         // per lane
constexpr int kSlots = 576;     // LDS slots of one transpose buffer (one pad slot per 8)

__device__ __forceinline__ int phys(int p) { return p + (p >> 3); }

__device__ __forceinline__ int pos1(int L, int t) { return L + 64 * t; }
__device__ __forceinline__ int pos2(int L, int t) { return 64 * (L >> 3) + (L & 7) + 8 * t; }
__device__ __forceinline__ int pos3(int L, int t) { return 8 * L + t; }

typedef const __attribute__((address_space(1))) C g_c;
__device__ __forceinline__ C ldc(const C* base, uint32_t i) {
    g_c* p = (g_c*)((const __attribute__((address_space(1))) char*)base + (size_t)(i * 16u));
    C r;
    r.re = p->re;
    r.im = p->im;
    return r;
}

// v w
__device__ __forceinline__ C cmul(C v, C w) {
    C r;
    r.re = fma(v.re, w.re, -(v.im * w.im));
    r.im = fma(v.re, w.im, v.im * w.re);
    return r;
}
// v conj(w)
__device__ __forceinline__ C cmulc(C v, C w) {
    C r;
    r.re = fma(v.re, w.re, v.im * w.im);
    r.im = fma(v.im, w.re, -(v.re * w.im));
    return r;
}

// Cooley-Tukey: (u, v) -> (u + W v, u - W v), W = w (I = false) or i w (I = true).
template <bool I>
__device__ __forceinline__ void ct(C& u, C& v, C w) {
    const C t = cmul(v, w);
    const C a = u;
    if constexpr (!I) {
        u = C{a.re + t.re, a.im + t.im};
        v = C{a.re - t.re, a.im - t.im};
    } else {
        u = C{a.re - t.im, a.im + t.re};
        v = C{a.re + t.im, a.im - t.re};
    }
}

// Gentleman-Sande: (u, v) -> (u + v, (u - v) conj(W)), conj(i w) = -i conj(w).
template <bool I>
__device__ __forceinline__ void gs(C& u, C& v, C w) {
    const C a = u, b = v;
    u = C{a.re + b.re, a.im + b.im};
    const C t = cmulc(C{a.re - b.re, a.im - b.im}, w);
    v = I ? C{t.im, -t.re} : t;
}

// Half distance of level lv within a pass (in t units) and the first element of butterfly b.
constexpr int half_of(int lv) { return 4 >> (lv % 3); }
constexpr int bf_u8(int H, int b) { return (b / H) * 2 * H + b % H; }

// Twiddle index (into zeta) of the butterfly of level lv whose first element is t, for lane L:
// 2^(lv+1) + (block of the pass) * 2^p + (t >> (3 - p)), p = lv % 3 (pass 1 has block 0, pass 2
// block L >> 3, pass 3 block L).
__device__ __forceinline__ uint32_t tw_index(int lv, int t, int L) {
    const int p = lv % 3;
    const uint32_t base = lv < 3 ? 0u : (lv < 6 ? (uint32_t)(L >> 3) : (uint32_t)L);
    return (2u << lv) + (base << p) + (uint32_t)(t >> (3 - p));
}

// The twiddles of one lane for all nine levels, loaded once and kept in registers: per pass,
// v[0] for its first level (index may be odd: loaded as it is), v[1] for its second (an even
// index; the butterflies with t >> 2 odd use i times it) and v[2], v[3] for its third (even
// indices of t = 0 and t = 4; the butterflies with t >> 1 odd use i times them).  Pass 1's are the
// same for every lane.
struct Tw {
    C v[3][4];
};

__device__ __forceinline__ Tw load_tw(const C* zt, int L) {
    Tw w;
#pragma unroll
    for (int pass = 0; pass < 3; ++pass) {
        const int lv = 3 * pass;
        w.v[pass][0] = ldc(zt, tw_index(lv, 0, L));
        w.v[pass][1] = ldc(zt, tw_index(lv + 1, 0, L));
        w.v[pass][2] = ldc(zt, tw_index(lv + 2, 0, L));
        w.v[pass][3] = ldc(zt, tw_index(lv + 2, 4, L));
    }
    return w;
}

template <int LV>
__device__ __forceinline__ void ld_tw(C (&w)[2], const Tw& tw, int) {
    constexpr int pass = LV / 3, p = LV % 3;
    w[0] = tw.v[pass][p];
    if constexpr (p == 2) w[1] = tw.v[pass][3];
}

// The same twiddles with passes 2 and 3 in an LDS table of 8 x 64 values (value v of pass q + 1 of
// lane L at (4 q + v) 64 + L, one table for every wave of a work-group) and pass 1 in registers:
// eight registers instead of 32 per lane, at one LDS read per level.
struct TwLds {
    C p1[4];
    const C* tab;
};

__device__ __forceinline__ void fill_tw_lds(C* tab, const C* zt, int L) {
#pragma unroll
    for (int pass = 1; pass < 3; ++pass) {
        const int lv = 3 * pass;
        tab[(4 * (pass - 1) + 0) * 64 + L] = ldc(zt, tw_index(lv, 0, L));
        tab[(4 * (pass - 1) + 1) * 64 + L] = ldc(zt, tw_index(lv + 1, 0, L));
        tab[(4 * (pass - 1) + 2) * 64 + L] = ldc(zt, tw_index(lv + 2, 0, L));
        tab[(4 * (pass - 1) + 3) * 64 + L] = ldc(zt, tw_index(lv + 2, 4, L));
    }
}

__device__ __forceinline__ TwLds make_tw_lds(const C* tab, const C* zt) {
    TwLds w;
    w.p1[0] = ldc(zt, tw_index(0, 0, 0));
    w.p1[1] = ldc(zt, tw_index(1, 0, 0));
    w.p1[2] = ldc(zt, tw_index(2, 0, 0));
    w.p1[3] = ldc(zt, tw_index(2, 4, 0));
    w.tab = tab;
    return w;
}

template <int LV>
__device__ __forceinline__ void ld_tw(C (&w)[2], const TwLds& tw, int L) {
    constexpr int pass = LV / 3, p = LV % 3;
    if constexpr (pass == 0) {
        w[0] = tw.p1[p];
        if constexpr (p == 2) w[1] = tw.p1[3];
    } else {
        constexpr int q = 4 * (pass - 1) + p;
        w[0] = tw.tab[q * 64 + L];
        if constexpr (p == 2) w[1] = tw.tab[(q + 1) * 64 + L];
    }
}

template <int LV, bool INV>
__device__ __forceinline__ void level(C (&x)[kPer], const C (&w)[2], int L) {
    constexpr int H = half_of(LV);
    constexpr int p = LV % 3;
    static_for<4>([&](auto B) __attribute__((always_inline)) {
        constexpr int u = bf_u8(H, B);
        constexpr bool odd = p == 1 ? ((u >> 2) & 1) != 0 : (p == 2 ? ((u >> 1) & 1) != 0 : false);
        constexpr int q = p == 2 ? (u >> 2) : 0;
        if constexpr (!INV) ct<odd>(x[u], x[u + H], w[q]);
        else gs<odd>(x[u], x[u + H], w[q]);
    });
}

template <int A, int B>
__device__ __forceinline__ void transpose(C (&x)[kPer], C* lds, int L) {
    poly_sync<1>();
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
        const int p = A == 1 ? pos1(L, t) : (A == 2 ? pos2(L, t) : pos3(L, t));
        lds[phys(p)] = x[t];
    }
    poly_sync<1>();
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
        const int p = B == 1 ? pos1(L, t) : (B == 2 ? pos2(L, t) : pos3(L, t));
        x[t] = lds[phys(p)];
    }
}

// Pass-1 layout (z_{L + 64 t}) -> transform values in pass-3 layout (position 8 L + t).  Each
// level's twiddles are fetched while the level before it computes.
template <class TW, class AtLast = NoPrefetch>
__device__ __forceinline__ void forward(C (&x)[kPer], C* lds, const TW& tw, int L, AtLast&& at_last = AtLast{}) {
    C w[2];
    ld_tw<0>(w, tw, L);
    static_for<9>([&](auto I) __attribute__((always_inline)) {
        constexpr int lv = I;
        if constexpr (lv == 3) transpose<1, 2>(x, lds, L);
        if constexpr (lv == 6) transpose<2, 3>(x, lds, L);
        if constexpr (lv == 8) at_last();
        C wn[2];
        if constexpr (lv + 1 < 9) ld_tw<lv + 1>(wn, tw, L);
        level<lv, false>(x, w, L);
        if constexpr (lv + 1 < 9) {
            w[0] = wn[0];
            w[1] = wn[1];
        }
    });
}

// Transform values in pass-3 layout -> 512 z in pass-1 layout.
template <class TW>
__device__ __forceinline__ void inverse(C (&x)[kPer], C* lds, const TW& tw, int L) {
    C w[2];
    ld_tw<8>(w, tw, L);
    static_for<9>([&](auto I) __attribute__((always_inline)) {
        constexpr int lv = 8 - I;
        C wn[2];
        if constexpr (lv > 0) ld_tw<lv - 1>(wn, tw, L);
        level<lv, true>(x, w, L);
        if constexpr (lv == 6) transpose<3, 2>(x, lds, L);
        if constexpr (lv == 3) transpose<2, 1>(x, lds, L);
        if constexpr (lv > 0) {
            w[0] = wn[0];
            w[1] = wn[1];
        }
    });
}

}  // namespace cf
}  // namespace tp
