// Goldilocks arithmetic (Q = 2^64 - 2^32 + 1) written instruction by instruction for the gfx950
// VALU, for the NTT butterflies and multiply-accumulates of the blind rotation.
//
// Every primitive below is a fixed sequence of the ops in the "instruction layer": in device code
// each op is one inline-asm VALU instruction, whose carry-out lands in an SGPR pair (a lane mask
// the compiler combines with scalar ops); in host code the same op is a plain C function on one
// lane's values.  The host build of this header therefore executes exactly the device sequence,
// lane by lane, and is what the CPU tests check against 128-bit integer arithmetic (edge values
// and random values, including every carry/borrow combination each proof below names).
//
// Representations:
//   lazy      any 64-bit integer x, standing for x mod Q;
//   canonical x < Q.
// Building blocks and their VALU counts (the point of this layer: LLVM's own selection of the
// same arithmetic spends 64-bit compares and selects on every carry):
//   mul128          8   a*b as a 128-bit (lo, hi), any 64-bit a, b
//   reduce_canon    8   any 128-bit (lo, hi) -> canonical
//   add_lc          4   lazy u + canonical t -> lazy
//   sub_lc          5   lazy u - canonical t -> lazy
//   add_cc          5   canonical + canonical -> canonical
//   sub_cc          5   canonical - canonical -> canonical
//   pow2 products   5-10  any x * 2^e -> canonical (sign returned separately)
// The identities are the Solinas ones the reference's scalar path uses
// (lib/include/backend/solinas_internal.h:61-73): 2^64 = 2^32 - 1 (= eps) and 2^96 = -1 mod Q.
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define GV_FN __device__ __forceinline__
#else
#define GV_FN inline
#endif

namespace glv {

constexpr uint64_t kQ = 0xFFFFFFFF00000001ULL;
constexpr uint64_t kQm1 = 0xFFFFFFFF00000000ULL;

GV_FN uint32_t lo32(uint64_t x) { return (uint32_t)x; }
GV_FN uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
GV_FN uint64_t pk(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

// ------------------------------------------------------------------ instruction layer
#if defined(__HIP_DEVICE_COMPILE__)
// A lane mask (carry, borrow, compare result) held in an SGPR pair.
typedef uint64_t lmask;
GV_FN lmask m_or(lmask a, lmask b) { return a | b; }
GV_FN lmask m_and(lmask a, lmask b) { return a & b; }
GV_FN lmask m_andn(lmask a, lmask b) { return a & ~b; }
GV_FN lmask m_xnor(lmask a, lmask b) { return ~(a ^ b); }

// v_mad_u64_u32: d = a * b + c (mod 2^64), co = bit 64 of the 65-bit sum
GV_FN uint64_t mad(uint32_t a, uint32_t b, uint64_t c, lmask& co) {
    uint64_t d;
    asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(co) : "v"(a), "v"(b), "v"(c));
    return d;
}
// Dead carry-outs rotate over four sink registers (sink(m) for member m): consecutive writes of
// one register by inline asm make the compiler insert a wait state between them.
#define GV_SINK_SWITCH(m, BODY) \
    switch ((m) & 3) { \
        case 0: BODY("s[94:95]", "s94", "s95"); break; \
        case 1: BODY("s[96:97]", "s96", "s97"); break; \
        case 2: BODY("s[98:99]", "s98", "s99"); break; \
        default: BODY("vcc", "vcc_lo", "vcc_hi"); break; \
    }
GV_FN uint64_t mad0(uint32_t a, uint32_t b, int m = 3) {
    uint64_t d;
#define B_(R, C1, C2) asm volatile("v_mad_u64_u32 %0, " R ", %1, %2, 0" : "=v"(d) : "v"(a), "v"(b) : C1, C2)
    GV_SINK_SWITCH(m, B_)
#undef B_
    return d;
}
GV_FN uint64_t mad_eps_x(uint32_t a, uint64_t c, int m = 3) {  // a * eps + c, carry known to be zero
    uint64_t d;
#define B_(R, C1, C2) asm volatile("v_mad_u64_u32 %0, " R ", %1, -1, %2" : "=&v"(d) : "v"(a), "v"(c) : C1, C2)
    GV_SINK_SWITCH(m, B_)
#undef B_
    return d;
}
// a * eps + c (eps = 2^32 - 1, the inline constant -1 as an unsigned 32-bit operand)
GV_FN uint64_t mad_eps(uint32_t a, uint64_t c, lmask& co) {
    uint64_t d;
    asm volatile("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(d), "=s"(co) : "v"(a), "v"(c));
    return d;
}
GV_FN uint64_t mad_eps0(uint32_t a, int m = 3) {
    uint64_t d;
#define B_(R, C1, C2) asm volatile("v_mad_u64_u32 %0, " R ", %1, -1, 0" : "=v"(d) : "v"(a) : C1, C2)
    GV_SINK_SWITCH(m, B_)
#undef B_
    return d;
}
GV_FN uint32_t add_co(uint32_t a, uint32_t b, lmask& co) {
    uint32_t d;
    asm volatile("v_add_co_u32 %0, %1, %2, %3" : "=v"(d), "=s"(co) : "v"(a), "v"(b));
    return d;
}
GV_FN uint32_t addc(uint32_t a, uint32_t b, lmask ci, lmask& co) {
    uint32_t d;
    asm volatile("v_addc_co_u32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(co) : "v"(a), "v"(b), "s"(ci));
    return d;
}
GV_FN uint32_t addc0(uint32_t a, lmask ci, int m = 3) {  // a + ci
    uint32_t d;
#define B_(R, C1, C2) asm volatile("v_addc_co_u32 %0, " R ", %1, 0, %2" : "=v"(d) : "v"(a), "s"(ci) : C1, C2)
    GV_SINK_SWITCH(m, B_)
#undef B_
    return d;
}
GV_FN uint32_t addc0c(uint32_t a, lmask ci, lmask& co) {  // a + ci, with its carry-out
    uint32_t d;
    asm volatile("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(d), "=s"(co) : "v"(a), "s"(ci));
    return d;
}
GV_FN uint32_t subb0x(uint32_t a, lmask bi, int m = 3) {  // a - bi, borrow-out dead
    uint32_t d;
#define B_(R, C1, C2) asm volatile("v_subb_co_u32 %0, " R ", %1, 0, %2" : "=v"(d) : "v"(a), "s"(bi) : C1, C2)
    GV_SINK_SWITCH(m, B_)
#undef B_
    return d;
}
GV_FN uint32_t sub_co(uint32_t a, uint32_t b, lmask& bo) {
    uint32_t d;
    asm volatile("v_sub_co_u32 %0, %1, %2, %3" : "=v"(d), "=s"(bo) : "v"(a), "v"(b));
    return d;
}
GV_FN uint32_t subb(uint32_t a, uint32_t b, lmask bi, lmask& bo) {
    uint32_t d;
    asm volatile("v_subb_co_u32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(bo) : "v"(a), "v"(b), "s"(bi));
    return d;
}
GV_FN uint32_t subb0(uint32_t a, lmask bi, lmask& bo) {  // a - bi
    uint32_t d;
    asm volatile("v_subb_co_u32 %0, %1, %2, 0, %3" : "=v"(d), "=s"(bo) : "v"(a), "s"(bi));
    return d;
}
GV_FN uint32_t cnd(uint32_t f, uint32_t t, lmask m) {  // m ? t : f
    uint32_t d;
    asm volatile("v_cndmask_b32 %0, %1, %2, %3" : "=v"(d) : "v"(f), "v"(t), "s"(m));
    return d;
}
GV_FN uint32_t cnd01(lmask m) {  // m ? 1 : 0
    uint32_t d;
    asm volatile("v_cndmask_b32 %0, 0, 1, %1" : "=v"(d) : "s"(m));
    return d;
}
GV_FN uint32_t cnd0f(lmask m) {  // m ? 0xFFFFFFFF : 0
    uint32_t d;
    asm volatile("v_cndmask_b32 %0, 0, -1, %1" : "=v"(d) : "s"(m));
    return d;
}
GV_FN uint32_t cndx1(uint32_t f, lmask m) {  // m ? 0xFFFFFFFF : f
    uint32_t d;
    asm volatile("v_cndmask_b32 %0, %1, -1, %2" : "=v"(d) : "v"(f), "s"(m));
    return d;
}
// x > Q - 1, i.e. x >= Q (qm1 = Q - 1 in an SGPR pair)
GV_FN lmask ge_q(uint64_t x, uint64_t qm1) {
    lmask m;
    asm volatile("v_cmp_lt_u64 %0, %1, %2" : "=s"(m) : "s"(qm1), "v"(x));
    return m;
}
GV_FN uint64_t add64(uint64_t a, uint64_t b) {  // (a + b) mod 2^64, no carry out
    uint64_t d;
    asm volatile("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(d) : "v"(a), "v"(b));
    return d;
}
GV_FN uint64_t shl64(uint64_t a, int s) {
    uint64_t d;
    asm volatile("v_lshlrev_b64 %0, %1, %2" : "=v"(d) : "i"(s), "v"(a));
    return d;
}
GV_FN uint64_t shr64(uint64_t a, int s) {
    uint64_t d;
    asm volatile("v_lshrrev_b64 %0, %1, %2" : "=v"(d) : "i"(s), "v"(a));
    return d;
}
GV_FN uint32_t shl32(uint32_t a, int s) {
    uint32_t d;
    asm volatile("v_lshlrev_b32 %0, %1, %2" : "=v"(d) : "i"(s), "v"(a));
    return d;
}
GV_FN uint32_t shr32(uint32_t a, int s) {
    uint32_t d;
    asm volatile("v_lshrrev_b32 %0, %1, %2" : "=v"(d) : "i"(s), "v"(a));
    return d;
}
GV_FN uint64_t qm1_const() { return kQm1; }
#else
// Host: the same ops on one lane (in a HIP build the host pass sees these in place of the asm).
typedef bool lmask;
GV_FN lmask m_or(lmask a, lmask b) { return a || b; }
GV_FN lmask m_and(lmask a, lmask b) { return a && b; }
GV_FN lmask m_andn(lmask a, lmask b) { return a && !b; }
GV_FN lmask m_xnor(lmask a, lmask b) { return a == b; }
GV_FN uint64_t mad(uint32_t a, uint32_t b, uint64_t c, lmask& co) {
    const unsigned __int128 t = (unsigned __int128)((uint64_t)a * b) + c;
    co = (bool)(t >> 64);
    return (uint64_t)t;
}
GV_FN uint64_t mad0(uint32_t a, uint32_t b, int = 0) { return (uint64_t)a * b; }
GV_FN uint64_t mad_eps(uint32_t a, uint64_t c, lmask& co) { return mad(a, 0xFFFFFFFFu, c, co); }
GV_FN uint64_t mad_eps0(uint32_t a, int = 0) { return (uint64_t)a * 0xFFFFFFFFu; }
GV_FN uint64_t mad_eps_x(uint32_t a, uint64_t c, int = 0) { return (uint64_t)a * 0xFFFFFFFFu + c; }
GV_FN uint32_t add_co(uint32_t a, uint32_t b, lmask& co) {
    const uint64_t t = (uint64_t)a + b;
    co = (bool)(t >> 32);
    return (uint32_t)t;
}
GV_FN uint32_t addc(uint32_t a, uint32_t b, lmask ci, lmask& co) {
    const uint64_t t = (uint64_t)a + b + (ci ? 1 : 0);
    co = (bool)(t >> 32);
    return (uint32_t)t;
}
GV_FN uint32_t addc0(uint32_t a, lmask ci, int = 0) { return a + (ci ? 1u : 0u); }
GV_FN uint32_t addc0c(uint32_t a, lmask ci, lmask& co) { return addc(a, 0, ci, co); }
GV_FN uint32_t sub_co(uint32_t a, uint32_t b, lmask& bo) {
    bo = a < b;
    return a - b;
}
GV_FN uint32_t subb(uint32_t a, uint32_t b, lmask bi, lmask& bo) {
    const uint64_t s = (uint64_t)b + (bi ? 1 : 0);
    bo = (uint64_t)a < s;
    return (uint32_t)((uint64_t)a - s);
}
GV_FN uint32_t subb0(uint32_t a, lmask bi, lmask& bo) { return subb(a, 0, bi, bo); }
GV_FN uint32_t subb0x(uint32_t a, lmask bi, int = 0) { return a - (bi ? 1u : 0u); }
GV_FN uint32_t cnd(uint32_t f, uint32_t t, lmask m) { return m ? t : f; }
GV_FN uint32_t cnd01(lmask m) { return m ? 1u : 0u; }
GV_FN uint32_t cnd0f(lmask m) { return m ? 0xFFFFFFFFu : 0u; }
GV_FN uint32_t cndx1(uint32_t f, lmask m) { return m ? 0xFFFFFFFFu : f; }
GV_FN lmask ge_q(uint64_t x, uint64_t qm1) { return qm1 < x; }
GV_FN uint64_t add64(uint64_t a, uint64_t b) { return a + b; }
GV_FN uint64_t shl64(uint64_t a, int s) { return a << s; }
GV_FN uint64_t shr64(uint64_t a, int s) { return a >> s; }
GV_FN uint32_t shl32(uint32_t a, int s) { return a << s; }
GV_FN uint32_t shr32(uint32_t a, int s) { return a >> s; }
GV_FN uint64_t qm1_const() { return kQm1; }
#endif

// ------------------------------------------------------------------ field primitives
//
// Every primitive is written for a group of M independent operands ("members"): each op of its
// sequence is issued for all M members before the next op, so consecutive VALU instructions are
// independent and a member's dependent ops are M instructions apart (gfx950 needs a wait state
// between dependent back-to-back instructions; the device ops are volatile asm, so this order is
// the schedule).  M = 1 gives the scalar forms.

#if defined(__HIP__)
#define GV_FOR(m) _Pragma("unroll") for (int m = 0; m < M; ++m)
#else
#define GV_FOR(m) for (int m = 0; m < M; ++m)
#endif

// Full 128-bit product of any 64-bit a, b (8 VALU).  X = a1 b0 + a0 b1 keeps its carry cX (bit
// 96 of the product); lo = a0 b0 + X0 2^32, hi = a1 b1 + X1 + carry + cX 2^32 (< 2^64 since the
// product is below 2^128).
template <int M>
GV_FN void mul128_g(const uint64_t (&a)[M], const uint64_t (&b)[M], uint64_t (&lo)[M], uint64_t (&hi)[M]) {
    uint64_t X[M], L[M], H[M];
    lmask cX[M], k[M];
    GV_FOR(m) X[m] = mad0(hi32(a[m]), lo32(b[m]), m);
    GV_FOR(m) L[m] = mad0(lo32(a[m]), lo32(b[m]), m);
    GV_FOR(m) X[m] = mad(lo32(a[m]), hi32(b[m]), X[m], cX[m]);
    GV_FOR(m) H[m] = mad0(hi32(a[m]), hi32(b[m]), m);
    uint32_t l1[M], h0[M], h1[M];
    GV_FOR(m) l1[m] = add_co(hi32(L[m]), lo32(X[m]), k[m]);
    GV_FOR(m) h0[m] = addc(lo32(H[m]), hi32(X[m]), k[m], k[m]);
    GV_FOR(m) h1[m] = addc0(hi32(H[m]), k[m], m);
    GV_FOR(m) h1[m] = addc0(h1[m], cX[m], m);
    GV_FOR(m) {
        lo[m] = pk(lo32(L[m]), l1[m]);
        hi[m] = pk(h0[m], h1[m]);
    }
}

// The final step of every reduction below: r stands for r + (c - b) eps with the side conditions
// c -> r <= 2^64 - 2^33 and b -> r >= 2^64 - 2^32 + 1; returns the canonical value (5 VALU plus
// scalar mask logic).  With p = c & ~b | (c == b) & (r >= Q) and n = b & ~c the added word
// is eps when p (r + eps <= Q - 1 after a carry; r + eps wraps to r - Q when r >= Q), Q = -eps
// mod 2^64 when n (r - eps lies in [2^64 - 2^33 + 2, Q - 2]), else 0.
template <int M>
GV_FN void fix_canon_g(uint64_t (&r)[M], const lmask (&c)[M], const lmask (&b)[M], uint64_t qm1) {
    lmask ge[M], p[M], n[M];
    uint32_t t[M], dlo[M], dhi[M];
    GV_FOR(m) ge[m] = ge_q(r[m], qm1);
    GV_FOR(m) {
        p[m] = m_or(m_andn(c[m], b[m]), m_and(m_xnor(c[m], b[m]), ge[m]));
        n[m] = m_andn(b[m], c[m]);
    }
    GV_FOR(m) t[m] = cnd01(n[m]);
    GV_FOR(m) dhi[m] = cnd0f(n[m]);
    GV_FOR(m) dlo[m] = cndx1(t[m], p[m]);
    GV_FOR(m) r[m] = add64(r[m], pk(dlo[m], dhi[m]));
}

// Any 128-bit (lo, hi) -> canonical (8 VALU): lo + h0 2^64 + h1 2^96 = lo + h0 eps - h1.
//   R + c 2^64 = lo + h0 eps         (c -> R <= 2^64 - 1 + (2^32 - 1)^2 - 2^64 = 2^64 - 2^33)
//   r - b 2^64 = R - h1              (b -> R < h1 < 2^32, so r >= 2^64 - 2^32 + 1)
template <int M>
GV_FN void reduce_canon_g(const uint64_t (&lo)[M], const uint64_t (&hi)[M], uint64_t (&r)[M], uint64_t qm1) {
    lmask c[M], b[M];
    uint64_t R[M];
    uint32_t r0[M], r1[M];
    GV_FOR(m) R[m] = mad_eps(lo32(hi[m]), lo[m], c[m]);
    GV_FOR(m) r0[m] = sub_co(lo32(R[m]), hi32(hi[m]), b[m]);
    GV_FOR(m) r1[m] = subb0(hi32(R[m]), b[m], b[m]);
    GV_FOR(m) r[m] = pk(r0[m], r1[m]);
    fix_canon_g<M>(r, c, b, qm1);
}

template <int M>
GV_FN void mul_canon_g(const uint64_t (&a)[M], const uint64_t (&b)[M], uint64_t (&r)[M], uint64_t qm1) {
    uint64_t lo[M], hi[M];
    mul128_g<M>(a, b, lo, hi);
    reduce_canon_g<M>(lo, hi, r, qm1);
}

// lazy u + canonical t -> lazy (4 VALU).  A carry means s = u + t - 2^64 <= t - 1 <= Q - 2, so
// s + eps <= 2^64 - 1: one correction, added with a multiply-add by eps.
template <int M>
GV_FN void add_lc_g(const uint64_t (&u)[M], const uint64_t (&t)[M], uint64_t (&r)[M]) {
    lmask k[M], c[M];
    uint32_t s0[M], s1[M], e[M];
    GV_FOR(m) s0[m] = add_co(lo32(u[m]), lo32(t[m]), k[m]);
    GV_FOR(m) s1[m] = addc(hi32(u[m]), hi32(t[m]), k[m], c[m]);
    GV_FOR(m) e[m] = cnd01(c[m]);
    GV_FOR(m) r[m] = mad_eps_x(e[m], pk(s0[m], s1[m]), m);
}

// lazy u - canonical t -> lazy (5 VALU).  A borrow means d = u - t + 2^64 >= 2^64 - Q + 1 = eps,
// so d - eps >= 0: one correction.  With u canonical too the result is canonical: a borrow gives
// d - eps = u - t + Q in [1, Q - 1].
template <int M>
GV_FN void sub_lc_g(const uint64_t (&u)[M], const uint64_t (&t)[M], uint64_t (&r)[M]) {
    lmask k[M], b[M], x[M];
    uint32_t d0[M], d1[M], e[M];
    GV_FOR(m) d0[m] = sub_co(lo32(u[m]), lo32(t[m]), k[m]);
    GV_FOR(m) d1[m] = subb(hi32(u[m]), hi32(t[m]), k[m], b[m]);
    GV_FOR(m) e[m] = cnd0f(b[m]);
    GV_FOR(m) d0[m] = sub_co(d0[m], e[m], x[m]);
    GV_FOR(m) d1[m] = subb0x(d1[m], x[m], m);
    GV_FOR(m) r[m] = pk(d0[m], d1[m]);
}

// canonical + canonical -> canonical (5 VALU).  u + t <= 2Q - 2: a carry (s = u + t - 2^64 <=
// Q - eps - 1) or s >= Q both take s + eps (mod 2^64), which is u + t - Q; both cannot hold.
template <int M>
GV_FN void add_cc_g(const uint64_t (&u)[M], const uint64_t (&t)[M], uint64_t (&r)[M], uint64_t qm1) {
    lmask k[M], c[M], ge[M];
    uint32_t s0[M], s1[M], e[M];
    GV_FOR(m) s0[m] = add_co(lo32(u[m]), lo32(t[m]), k[m]);
    GV_FOR(m) s1[m] = addc(hi32(u[m]), hi32(t[m]), k[m], c[m]);
    GV_FOR(m) ge[m] = ge_q(pk(s0[m], s1[m]), qm1);
    GV_FOR(m) e[m] = cnd01(m_or(c[m], ge[m]));
    GV_FOR(m) r[m] = mad_eps_x(e[m], pk(s0[m], s1[m]), m);
}

// Lazy -> canonical (3 VALU): x >= Q becomes x + eps (mod 2^64) = x - Q.
template <int M>
GV_FN void canon_g(uint64_t (&x)[M], uint64_t qm1) {
    lmask ge[M];
    uint32_t e[M];
    GV_FOR(m) ge[m] = ge_q(x[m], qm1);
    GV_FOR(m) e[m] = cnd01(ge[m]);
    GV_FOR(m) x[m] = mad_eps_x(e[m], x[m], m);
}

// ------------------------------------------------------------------ products by 2^e
//
// x * 2^e for compile-time exponents (one per member, all of one kind), any 64-bit x ->
// canonical.  2^96 = -1; the kinds below cover e in (0, 96) except 64, which callers avoid
// (the exponents of the transforms are multiples of 3).

// 0 < e < 32 (6 VALU).  x 2^e = lo + h 2^64 with h = x >> (64 - e) < 2^31, so R + c 2^64 =
// lo + h eps; a carry leaves R < 2^63 and R + eps < Q, otherwise R >= Q is brought below Q by
// the same + eps (mod 2^64).
template <int M>
GV_FN void mul_2e_lo_g(const uint64_t (&x)[M], const int (&e)[M], uint64_t (&r)[M], uint64_t qm1) {
    uint64_t lo[M], R[M];
    uint32_t h[M], f[M];
    lmask c[M], ge[M];
    GV_FOR(m) lo[m] = shl64(x[m], e[m]);
    GV_FOR(m) h[m] = shr32(hi32(x[m]), 32 - e[m]);
    GV_FOR(m) R[m] = mad_eps(h[m], lo[m], c[m]);
    GV_FOR(m) ge[m] = ge_q(R[m], qm1);
    GV_FOR(m) f[m] = cnd01(m_or(c[m], ge[m]));
    GV_FOR(m) r[m] = mad_eps_x(f[m], R[m], m);
}

// 32 < e < 64 (10 VALU): the 128-bit x << e, reduced.
template <int M>
GV_FN void mul_2e_mid_g(const uint64_t (&x)[M], const int (&e)[M], uint64_t (&r)[M], uint64_t qm1) {
    uint64_t lo[M], hi[M];
    GV_FOR(m) lo[m] = shl64(x[m], e[m]);
    GV_FOR(m) hi[m] = shr64(x[m], 64 - e[m]);
    reduce_canon_g<M>(lo, hi, r, qm1);
}

// x * 2^-s for 0 < s < 32 (8 VALU).  x = q 2^s + w, w 2^-s = -w 2^(96 - s) and w 2^(96 - s) =
// v 2^64 = v eps with v = w << (32 - s) < 2^32, so x 2^-s = q - v eps.  A borrow (q < v eps)
// gives d = q - v eps + 2^64 >= 2^33 - 1 >= eps, and d - eps = q - v eps + Q < Q; no borrow
// leaves q - v eps <= q < 2^(64 - s).
template <int M>
GV_FN void mul_2e_neg_g(const uint64_t (&x)[M], const int (&s)[M], uint64_t (&r)[M]) {
    uint64_t q[M], ve[M];
    uint32_t v[M], d0[M], d1[M], e[M];
    lmask k[M], b[M], z[M];
    GV_FOR(m) q[m] = shr64(x[m], s[m]);
    GV_FOR(m) v[m] = shl32(lo32(x[m]), 32 - s[m]);
    GV_FOR(m) ve[m] = mad_eps0(v[m], m);
    GV_FOR(m) d0[m] = sub_co(lo32(q[m]), lo32(ve[m]), k[m]);
    GV_FOR(m) d1[m] = subb(hi32(q[m]), hi32(ve[m]), k[m], b[m]);
    GV_FOR(m) e[m] = cnd0f(b[m]);
    GV_FOR(m) d0[m] = sub_co(d0[m], e[m], z[m]);
    GV_FOR(m) d1[m] = subb0x(d1[m], z[m], m);
    GV_FOR(m) r[m] = pk(d0[m], d1[m]);
}

// Kind of a product by 2^e (0 <= e < 192): its exponent in (0, 96) after taking out the sign
// of 2^96 = -1, the sign of the returned value (2^e' = -2^-(96 - e') for 64 < e' < 96) and the
// sequence that computes it.
enum class P2 { kLo, kMid, kNeg, kOne };
struct Pow2Kind {
    P2 kind;
    int arg;   // shift of the sequence (e' for kLo/kMid, 96 - e' for kNeg)
    bool neg;  // the product is minus the sequence's result
};
constexpr Pow2Kind pow2_kind(int e) {
    bool neg = e >= 96;
    const int ep = e % 96;
    if (ep == 0) return {P2::kOne, 0, neg};
    if (ep < 32) return {P2::kLo, ep, neg};
    if (ep < 64) return {P2::kMid, ep, neg};
    return {P2::kNeg, 96 - ep, !neg};  // ep == 64 is not used by the transforms
}

// Scalar forms (M = 1).
GV_FN void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    uint64_t A[1] = {a}, B[1] = {b}, LO[1], HI[1];
    mul128_g<1>(A, B, LO, HI);
    lo = LO[0];
    hi = HI[0];
}
GV_FN uint64_t reduce_canon(uint64_t lo, uint64_t hi, uint64_t qm1) {
    uint64_t L[1] = {lo}, H[1] = {hi}, R[1];
    reduce_canon_g<1>(L, H, R, qm1);
    return R[0];
}
GV_FN uint64_t mul_canon(uint64_t a, uint64_t b, uint64_t qm1) {
    uint64_t A[1] = {a}, B[1] = {b}, R[1];
    mul_canon_g<1>(A, B, R, qm1);
    return R[0];
}
GV_FN uint64_t add_lc(uint64_t u, uint64_t t) {
    uint64_t U[1] = {u}, T[1] = {t}, R[1];
    add_lc_g<1>(U, T, R);
    return R[0];
}
GV_FN uint64_t sub_lc(uint64_t u, uint64_t t) {
    uint64_t U[1] = {u}, T[1] = {t}, R[1];
    sub_lc_g<1>(U, T, R);
    return R[0];
}
GV_FN uint64_t add_cc(uint64_t u, uint64_t t, uint64_t qm1) {
    uint64_t U[1] = {u}, T[1] = {t}, R[1];
    add_cc_g<1>(U, T, R, qm1);
    return R[0];
}
GV_FN uint64_t sub_cc(uint64_t u, uint64_t t) { return sub_lc(u, t); }
GV_FN uint64_t canon(uint64_t x, uint64_t qm1) {
    uint64_t X[1] = {x};
    canon_g<1>(X, qm1);
    return X[0];
}

// x * 2^e for a compile-time 0 <= e < 192, e % 96 != 64 -> canonical (x itself when e % 96 == 0),
// with the sign returned separately so that butterflies exchange outputs instead of negating.
GV_FN uint64_t mul_pow2(uint64_t x, int e, bool& neg, uint64_t qm1) {
    const Pow2Kind k = pow2_kind(e);
    neg = k.neg;
    uint64_t X[1] = {x}, R[1];
    const int A[1] = {k.arg};
    if (k.kind == P2::kOne) return x;
    if (k.kind == P2::kLo) mul_2e_lo_g<1>(X, A, R, qm1);
    else if (k.kind == P2::kMid) mul_2e_mid_g<1>(X, A, R, qm1);
    else mul_2e_neg_g<1>(X, A, R);
    return R[0];
}

// ------------------------------------------------------------------ multiply-accumulate

// Accumulator of a sum of products: value = A + C 2^64 (mod Q) with a 32-bit count C.  Each term
// adds its 128-bit product as L = a0 b0 + A (carry cL, weight 2^64), the cross terms and a1 b1 as
// in mul128_g, and folds the product's top words with one multiply-add by eps (carry c1) and one
// subtraction of h1 (borrow b): C moves by cL + c1 - b in [-1, 2] per term (14 VALU per term).
// Start from A = acc_init_word(T), C = T for T terms (A + T 2^64 = 0 mod Q), so C stays in
// [0, 3T] and never wraps.
struct Acc {
    uint64_t A;
    uint32_t C;
};

constexpr uint64_t acc_init_word(uint32_t terms) {
    // Q - terms * eps (terms < 2^31), which is -terms 2^64 mod Q
    return kQ - (uint64_t)terms * 0xFFFFFFFFULL;
}

template <int M>
GV_FN void acc_mac_g(Acc (&acc)[M], const uint64_t (&a)[M], const uint64_t (&b)[M]) {
    uint64_t X[M], L[M], H[M], R[M];
    lmask cX[M], cL[M], k[M], c1[M], bb[M];
    uint32_t l1[M], h0[M], h1[M], r0[M], r1[M], C[M];
    GV_FOR(m) X[m] = mad0(hi32(a[m]), lo32(b[m]), m);
    GV_FOR(m) L[m] = mad(lo32(a[m]), lo32(b[m]), acc[m].A, cL[m]);
    GV_FOR(m) X[m] = mad(lo32(a[m]), hi32(b[m]), X[m], cX[m]);
    GV_FOR(m) H[m] = mad0(hi32(a[m]), hi32(b[m]), m);
    GV_FOR(m) l1[m] = add_co(hi32(L[m]), lo32(X[m]), k[m]);
    GV_FOR(m) h0[m] = addc(lo32(H[m]), hi32(X[m]), k[m], k[m]);
    GV_FOR(m) h1[m] = addc0(hi32(H[m]), k[m], m);
    GV_FOR(m) h1[m] = addc0(h1[m], cX[m], m);
    GV_FOR(m) R[m] = mad_eps(h0[m], pk(lo32(L[m]), l1[m]), c1[m]);
    GV_FOR(m) r0[m] = sub_co(lo32(R[m]), h1[m], bb[m]);
    GV_FOR(m) r1[m] = subb0(hi32(R[m]), bb[m], bb[m]);
    GV_FOR(m) C[m] = addc0(acc[m].C, cL[m], m);
    GV_FOR(m) C[m] = addc0(C[m], c1[m], m);
    GV_FOR(m) C[m] = subb0x(C[m], bb[m], m);
    GV_FOR(m) {
        acc[m].A = pk(r0[m], r1[m]);
        acc[m].C = C[m];
    }
}

// Value of each accumulator, canonical (4 VALU): A + C eps with C eps < 2^38.
template <int M>
GV_FN void acc_value_g(const Acc (&acc)[M], uint64_t (&v)[M], uint64_t qm1) {
    uint64_t R[M];
    lmask c[M], ge[M];
    uint32_t f[M];
    GV_FOR(m) R[m] = mad_eps(acc[m].C, acc[m].A, c[m]);
    GV_FOR(m) ge[m] = ge_q(R[m], qm1);
    GV_FOR(m) f[m] = cnd01(m_or(c[m], ge[m]));
    GV_FOR(m) v[m] = mad_eps_x(f[m], R[m], m);
}

// Two-word accumulator (16 VALU per term): acc (lazy) + a b -> lazy.  The addend rides in the
// low product, L = a0 b0 + acc (carry cL, weight 2^64 = 2^32 of the cross term X), so cL is added
// to X1 (carry c2, weight 2^96 like cX; cX and c2 never both occur: cX means X <= 2^64 - 2^34 + 2,
// so X1 + 1 does not carry).  The 128-bit total (< 2^128 since a b + acc < 2^128) is reduced
// lazily: r + (c - b) eps with the corrections of fix_canon but no comparison with Q.
template <int M>
GV_FN void mac_lazy_g(uint64_t (&acc)[M], const uint64_t (&a)[M], const uint64_t (&b)[M]) {
    uint64_t X[M], L[M], H[M], R[M];
    lmask cX[M], cL[M], c2[M], k[M], c1[M], bb[M];
    uint32_t x1[M], l1[M], h0[M], h1[M], r0[M], r1[M], t[M], dlo[M], dhi[M];
    GV_FOR(m) X[m] = mad0(hi32(a[m]), lo32(b[m]), m);
    GV_FOR(m) L[m] = mad(lo32(a[m]), lo32(b[m]), acc[m], cL[m]);
    GV_FOR(m) X[m] = mad(lo32(a[m]), hi32(b[m]), X[m], cX[m]);
    GV_FOR(m) H[m] = mad0(hi32(a[m]), hi32(b[m]), m);
    GV_FOR(m) x1[m] = addc0c(hi32(X[m]), cL[m], c2[m]);
    GV_FOR(m) l1[m] = add_co(hi32(L[m]), lo32(X[m]), k[m]);
    GV_FOR(m) h0[m] = addc(lo32(H[m]), x1[m], k[m], k[m]);
    GV_FOR(m) h1[m] = addc0(hi32(H[m]), k[m], m);
    GV_FOR(m) h1[m] = addc0(h1[m], m_or(cX[m], c2[m]), m);
    GV_FOR(m) R[m] = mad_eps(h0[m], pk(lo32(L[m]), l1[m]), c1[m]);
    GV_FOR(m) r0[m] = sub_co(lo32(R[m]), h1[m], bb[m]);
    GV_FOR(m) r1[m] = subb0(hi32(R[m]), bb[m], bb[m]);
    GV_FOR(m) t[m] = cnd01(m_andn(bb[m], c1[m]));
    GV_FOR(m) dhi[m] = cnd0f(m_andn(bb[m], c1[m]));
    GV_FOR(m) dlo[m] = cndx1(t[m], m_andn(c1[m], bb[m]));
    GV_FOR(m) acc[m] = add64(pk(r0[m], r1[m]), pk(dlo[m], dhi[m]));
}

GV_FN void acc_mac(Acc& acc, uint64_t a, uint64_t b) {
    Acc X[1] = {acc};
    const uint64_t A[1] = {a}, B[1] = {b};
    acc_mac_g<1>(X, A, B);
    acc = X[0];
}
GV_FN uint64_t mac_lazy(uint64_t acc, uint64_t a, uint64_t b) {
    uint64_t X[1] = {acc};
    const uint64_t A[1] = {a}, B[1] = {b};
    mac_lazy_g<1>(X, A, B);
    return X[0];
}
GV_FN uint64_t acc_value(const Acc& acc, uint64_t qm1) {
    const Acc X[1] = {acc};
    uint64_t V[1];
    acc_value_g<1>(X, V, qm1);
    return V[0];
}

#undef GV_FOR

}  // namespace glv
