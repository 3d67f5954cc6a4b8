/*
 * CPU oracle for the CGGI/BMMP bootstrapping path over the Goldilocks prime.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against: only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it.  Nothing in the product links it.
 *
 * It restates the reference algorithm (LSchild/FHEToothpaste) in plain C with
 * exact arithmetic modulo Q = 2^64 - 2^32 + 1, step by step in the reference's
 * own operation order (NTT-domain monomial products, digit order, Rep2 rows),
 * so that a GPU implementation that reorganises the computation can be checked
 * for bit equality.  Each function cites the reference file:line it follows.
 *
 * Pinning (what makes this restatement trustworthy, see tests/test_oracle_*):
 *   - field ops against unsigned __int128 '%' (textbook) on edge values;
 *   - NTT against the O(N^2) closed form out[k] = sum_j a_j psi^((2 brv(k)+1) j);
 *   - NTT against schoolbook negacyclic convolution (convolution theorem);
 *   - decomposers against the recomposition identity sum_j g_j d_j;
 *   - blind rotation with sigma = 0 keys against the exact phase relation
 *     phase(out) = X^(b - <a,s>) * phase(acc) (reference test
 *     tests/cggi_blind_rotation_test.cpp:118-123);
 *   - key switching with sigma = 0 against the exact phase relation.
 * Functions at the HEXL call sites with the 64-bit modulus use exact mod-Q
 * semantics; parity with HEXL there is unpinned (HEXL is absent and unpinned).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>
#include <time.h>

typedef unsigned __int128 u128;
#define OR_Q 0xFFFFFFFF00000001ULL
#define OR_EPS 0xFFFFFFFFULL

#define OR_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Field arithmetic: lib/include/backend/solinas_internal.h:39-111          */
/* ------------------------------------------------------------------------ */

OR_EXPORT uint64_t or_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    int carry = s < a;
    if (carry || s >= OR_Q) s -= OR_Q; /* on carry: s + 2^64 - Q = s + eps */
    return s;
}

OR_EXPORT uint64_t or_sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    if (a < b) d += OR_Q;
    return d;
}

/* x = lo + 2^64 (h0 + 2^32 h1) = lo + eps*h0 - h1  (solinas_internal.h:61-73) */
OR_EXPORT uint64_t or_reduce128(uint64_t hi, uint64_t lo) {
    uint64_t h1 = hi >> 32, h0 = hi & OR_EPS;
    uint64_t t0 = lo - h1;
    if (lo < h1) t0 -= OR_EPS;
    uint64_t t1 = (h0 << 32) - h0;
    uint64_t r = t0 + t1;
    if (r < t0) r += OR_EPS;
    if (r >= OR_Q) r -= OR_Q;
    return r;
}

OR_EXPORT uint64_t or_mul(uint64_t a, uint64_t b) {
    u128 p = (u128)a * b;
    return or_reduce128((uint64_t)(p >> 64), (uint64_t)p);
}

/* Textbook reference for the tests: 128-bit remainder. */
OR_EXPORT uint64_t or_mul_textbook(uint64_t a, uint64_t b) {
    return (uint64_t)(((u128)a * b) % OR_Q);
}

OR_EXPORT uint64_t or_pow(uint64_t base, uint64_t e) {
    uint64_t r = 1;
    base %= OR_Q;
    while (e) {
        if (e & 1) r = or_mul(r, base);
        base = or_mul(base, base);
        e >>= 1;
    }
    return r;
}

/* Fermat inverse; inv(0) = 0 as in solinas_internal.h:98-111. */
OR_EXPORT uint64_t or_inv(uint64_t x) { return or_pow(x, OR_Q - 2); }

static uint64_t neg(uint64_t x) { return x ? OR_Q - x : 0; }

/* ------------------------------------------------------------------------ */
/* NTT: lib/src/backend/solinas.cpp:168-313                                 */
/* ------------------------------------------------------------------------ */

static uint32_t brv(uint32_t x, unsigned bits) {
    uint32_t r = 0;
    for (unsigned i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

typedef struct {
    uint32_t N, logN;
    uint64_t psi, n_inv;
    uint64_t *psi_rev, *psi_inv_rev; /* psi^brv(i), psi^-brv(i)  (solinas.cpp:180-192) */
    uint64_t *mono;                  /* N*(N+2): rows NTT(X^i), then 2 zero rows (cggi_blind_rotation.cpp:474-485) */
} or_ring;

static void ntt_fwd(const or_ring* R, uint64_t* a);
static void ntt_inv(const or_ring* R, uint64_t* a);

OR_EXPORT or_ring* or_ring_create(uint32_t N) {
    if (N < 2 || (N & (N - 1))) return NULL;
    or_ring* R = (or_ring*)calloc(1, sizeof(or_ring));
    R->N = N;
    while ((1u << R->logN) < N) R->logN++;
    /* psi = 7^((Q-1)/2N): 7 generates Z_Q^*  (solinas.cpp:175-176) */
    R->psi = or_pow(7, (OR_Q - 1) >> (R->logN + 1));
    uint64_t psi_inv = or_inv(R->psi);
    R->n_inv = or_inv(N);
    R->psi_rev = (uint64_t*)malloc(sizeof(uint64_t) * N);
    R->psi_inv_rev = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (uint32_t i = 0; i < N; ++i) {
        uint32_t r = brv(i, R->logN);
        R->psi_rev[i] = or_pow(R->psi, r);
        R->psi_inv_rev[i] = or_pow(psi_inv, r);
    }
    R->mono = (uint64_t*)calloc((size_t)N * (N + 2), sizeof(uint64_t));
    for (uint32_t i = 0; i < N; ++i) {
        R->mono[(size_t)i * N + i] = 1;
        ntt_fwd(R, R->mono + (size_t)i * N);
    }
    return R;
}

OR_EXPORT void or_ring_destroy(or_ring* R) {
    if (!R) return;
    free(R->psi_rev); free(R->psi_inv_rev); free(R->mono); free(R);
}

OR_EXPORT uint64_t or_ring_psi(const or_ring* R) { return R->psi; }

/* Cooley-Tukey, natural in, bit-reversed out; twiddle of block i at level m
 * is psi_rev[m + i] (solinas.cpp:228-261, same result as its fused levels). */
static void ntt_fwd(const or_ring* R, uint64_t* a) {
    const uint32_t N = R->N;
    for (uint32_t m = 1, t = N / 2; m < N; m <<= 1, t >>= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            const uint64_t w = R->psi_rev[m + i];
            uint64_t* p = a + 2 * i * t;
            for (uint32_t j = 0; j < t; ++j) {
                uint64_t u = p[j], v = or_mul(p[j + t], w);
                p[j] = or_add(u, v);
                p[j + t] = or_sub(u, v);
            }
        }
    }
}

/* Gentleman-Sande, bit-reversed in, natural out, times N^-1 (solinas.cpp:263-313). */
static void ntt_inv(const or_ring* R, uint64_t* a) {
    const uint32_t N = R->N;
    for (uint32_t m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            const uint64_t w = R->psi_inv_rev[m + i];
            uint64_t* p = a + 2 * i * t;
            for (uint32_t j = 0; j < t; ++j) {
                uint64_t u = p[j], v = p[j + t];
                p[j] = or_add(u, v);
                p[j + t] = or_mul(or_sub(u, v), w);
            }
        }
    }
    for (uint32_t j = 0; j < N; ++j) a[j] = or_mul(a[j], R->n_inv);
}

OR_EXPORT void or_ntt_forward(const or_ring* R, uint64_t* out, const uint64_t* in, uint32_t count) {
    for (uint32_t c = 0; c < count; ++c) {
        if (out + (size_t)c * R->N != in + (size_t)c * R->N)
            memcpy(out + (size_t)c * R->N, in + (size_t)c * R->N, sizeof(uint64_t) * R->N);
        ntt_fwd(R, out + (size_t)c * R->N);
    }
}

OR_EXPORT void or_ntt_inverse(const or_ring* R, uint64_t* out, const uint64_t* in, uint32_t count) {
    for (uint32_t c = 0; c < count; ++c) {
        if (out + (size_t)c * R->N != in + (size_t)c * R->N)
            memcpy(out + (size_t)c * R->N, in + (size_t)c * R->N, sizeof(uint64_t) * R->N);
        ntt_inv(R, out + (size_t)c * R->N);
    }
}

/* Closed form, O(N^2): out[k] = sum_j a_j psi^((2 brv(k) + 1) j).  Used only to pin ntt_fwd. */
OR_EXPORT void or_ntt_closed_form(const or_ring* R, uint64_t* out, const uint64_t* in) {
    const uint32_t N = R->N;
    for (uint32_t k = 0; k < N; ++k) {
        uint64_t x = or_pow(R->psi, 2 * (uint64_t)brv(k, R->logN) + 1);
        uint64_t acc = 0, xp = 1;
        for (uint32_t j = 0; j < N; ++j) {
            acc = or_add(acc, or_mul(in[j], xp));
            xp = or_mul(xp, x);
        }
        out[k] = acc;
    }
}

/* Schoolbook negacyclic product c = a*b mod (X^N + 1). */
OR_EXPORT void or_negacyclic_mul(uint32_t N, uint64_t* c, const uint64_t* a, const uint64_t* b) {
    memset(c, 0, sizeof(uint64_t) * N);
    for (uint32_t i = 0; i < N; ++i)
        for (uint32_t j = 0; j < N; ++j) {
            uint64_t p = or_mul(a[i], b[j]);
            uint32_t k = i + j;
            if (k < N) c[k] = or_add(c[k], p);
            else c[k - N] = or_sub(c[k - N], p);
        }
}

static void vmul(uint64_t* o, const uint64_t* a, const uint64_t* b, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) o[i] = or_mul(a[i], b[i]);
}
static void vadd(uint64_t* o, const uint64_t* a, const uint64_t* b, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) o[i] = or_add(a[i], b[i]);
}
static void vsub(uint64_t* o, const uint64_t* a, const uint64_t* b, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) o[i] = or_sub(a[i], b[i]);
}

/* ------------------------------------------------------------------------ */
/* Gadget decomposition                                                     */
/* ------------------------------------------------------------------------ */

/* ceil(64 / b): IntLog2(Q) = 64 (utils/math_utils.cpp:11-25). */
static uint32_t max_digits(uint32_t b) { return 64 / b + (64 % b ? 1 : 0); }

/* Corrector C and least-digit correction (static/gadget_decomp.h:16-28). */
OR_EXPORT uint64_t or_signed_corrector(uint32_t b, uint64_t* least_out) {
    uint64_t c = 0;
    int64_t shift = 64 - (int64_t)b;
    while (shift > 0) { c += 1ULL << (shift + b - 1); shift -= b; }
    uint64_t least = 1ULL << (shift + b - 1);
    c += least;
    if (least_out) *least_out = least;
    return c % OR_Q;
}

/* Unsigned digits of mux_operator.cpp:23-26,46-48,62-64:
 * first_shift = b (ceil(64/b) - 1); digit j = (x >> (first_shift - j b)) & (2^b - 1).
 * out: l rows of n. */
OR_EXPORT void or_decompose_unsigned(uint64_t* out, const uint64_t* x, uint32_t n, uint32_t b, uint32_t l) {
    const uint32_t first = b * (max_digits(b) - 1);
    const uint64_t mask = (1ULL << b) - 1;
    for (uint32_t j = 0; j < l; ++j)
        for (uint32_t i = 0; i < n; ++i) out[(size_t)j * n + i] = (x[i] >> (first - j * b)) & mask;
}

/* Signed digits of gadget_decomp.h:156-205 (coefficient domain, before the NTT):
 * x' = x + C mod Q; digit j = ((x' >> (64 - b(j+1))) & mask) - 2^(b-1) mod Q; when j is the
 * (ceil(64/b)-1)-th digit it uses shift 0, the least-digit correction and a matching mask. */
OR_EXPORT void or_decompose_signed(uint64_t* out, const uint64_t* x, uint32_t n, uint32_t b, uint32_t l) {
    uint64_t least;
    const uint64_t C = or_signed_corrector(b, &least);
    const uint32_t D = max_digits(b);
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t xp = or_add(x[i], C);
        int64_t shift = 64 - (int64_t)b;
        for (uint32_t j = 0; j < l; ++j) {
            uint64_t delta = 1ULL << (b - 1), mask = (1ULL << b) - 1;
            if (j == D - 1) { delta = least; shift = 0; mask = (delta << 1) - 1; }
            out[(size_t)j * n + i] = or_sub((xp >> shift) & mask, delta);
            shift -= b;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* MuxOperator: lib/src/operators/mux_operator.cpp                          */
/* ------------------------------------------------------------------------ */

/* RLWEPrimeProduct (mux_operator.cpp:30-84), unsigned digits. */
static void rlwe_prime_product(const or_ring* R, uint64_t* acc, const uint64_t* rp, const uint64_t* poly,
                               int add, uint32_t b, uint32_t l, uint64_t* scratch) {
    const uint32_t N = R->N;
    const uint32_t first = b * (max_digits(b) - 1);
    const uint64_t mask = (1ULL << b) - 1;
    for (uint32_t j = 0; j < l; ++j) {
        const uint32_t shift = first - j * b;
        for (uint32_t i = 0; i < N; ++i) scratch[i] = (poly[i] >> shift) & mask;
        ntt_fwd(R, scratch);
        const uint64_t* ra = rp + (size_t)j * 2 * N;
        if (!add && j == 0) {
            vmul(acc, ra, scratch, N);
            vmul(acc + N, ra + N, scratch, N);
        } else {
            for (uint32_t i = 0; i < N; ++i) {
                acc[i] = or_add(acc[i], or_mul(ra[i], scratch[i]));
                acc[N + i] = or_add(acc[N + i], or_mul(ra[N + i], scratch[i]));
            }
        }
    }
}

/* ExternalProduct (mux_operator.cpp:114-121): rows [0,l) x digits(a), rows [l,2l) x digits(b). */
OR_EXPORT void or_external_product(const or_ring* R, uint64_t* result, const uint64_t* rgsw, const uint64_t* rhs_coef,
                                   int add, uint32_t b, uint32_t l) {
    uint64_t* scratch = (uint64_t*)malloc(sizeof(uint64_t) * R->N);
    rlwe_prime_product(R, result, rgsw, rhs_coef, add, b, l, scratch);
    rlwe_prime_product(R, result, rgsw + (size_t)2 * R->N * l, rhs_coef + R->N, 1, b, l, scratch);
    free(scratch);
}

/* BinaryCMux (mux_operator.cpp:136-152): monomial product in the NTT domain,
 * INTT x2, external product, add.  acc is NTT-domain [a|b]. */
static void binary_cmux(const or_ring* R, uint64_t* acc, const uint64_t* rgsw, const uint64_t* mono,
                        uint32_t b, uint32_t l, uint64_t* ws /* 5N */) {
    const uint32_t N = R->N;
    uint64_t* s = ws;           /* 2N */
    uint64_t* tmp = ws + 2 * N; /* 2N */
    uint64_t* scratch = ws + 4 * N;
    vmul(s, acc, mono, N);
    vmul(s + N, acc + N, mono, N);
    ntt_inv(R, s);
    ntt_inv(R, s + N);
    rlwe_prime_product(R, tmp, rgsw, s, 0, b, l, scratch);
    rlwe_prime_product(R, tmp, rgsw + (size_t)2 * N * l, s + N, 1, b, l, scratch);
    vadd(acc, acc, tmp, 2 * N);
}

OR_EXPORT void or_binary_cmux(const or_ring* R, uint64_t* acc, const uint64_t* rgsw, const uint64_t* mono,
                              uint32_t b, uint32_t l) {
    uint64_t* ws = (uint64_t*)malloc(sizeof(uint64_t) * 5 * R->N);
    binary_cmux(R, acc, rgsw, mono, b, l, ws);
    free(ws);
}

/* SignedGadgetDecomposeRepNTT<1> (gadget_decomp.h:156-205): row j = [NTT(d_j) | NTT(d_j)]. */
static void signed_rep2_ntt(const or_ring* R, uint64_t* out, const uint64_t* x, uint32_t b, uint32_t l) {
    const uint32_t N = R->N;
    uint64_t* d = (uint64_t*)malloc(sizeof(uint64_t) * N * l);
    or_decompose_signed(d, x, N, b, l);
    for (uint32_t j = 0; j < l; ++j) {
        uint64_t* row = out + (size_t)j * 2 * N;
        memcpy(row, d + (size_t)j * N, sizeof(uint64_t) * N);
        ntt_fwd(R, row);
        memcpy(row + N, row, sizeof(uint64_t) * N);
    }
    free(d);
}

/* sum_k w_k (RGSW_k [x] acc) with one shared signed decomposition: the common core of
 * TernaryCMux (mux_operator.cpp:156-202, k = 2) and MultiMux (:204-294, k = 3).
 * Row r of RGSW_k multiplies the Rep2 digit row r; rows are summed, the sum is multiplied
 * by the weight w_k (NTT domain) and added to acc. */
static void multi_mux(const or_ring* R, uint64_t* acc, uint32_t k, const uint64_t* rgsws, const uint64_t* weights,
                      uint32_t b, uint32_t l, uint64_t* ws) {
    const uint32_t N = R->N;
    const size_t rp = (size_t)2 * N * l, rg = 2 * rp;
    uint64_t* in = ws;                 /* 2N */
    uint64_t* dig = ws + 2 * N;        /* 4Nl */
    uint64_t* prod = dig + rg;         /* 4Nl */
    uint64_t* sum = prod + rg;         /* 2N */
    ntt_inv(R, memcpy(in, acc, sizeof(uint64_t) * N));
    ntt_inv(R, memcpy(in + N, acc + N, sizeof(uint64_t) * N));
    signed_rep2_ntt(R, dig, in, b, l);
    signed_rep2_ntt(R, dig + rp, in + N, b, l);
    for (uint32_t c = 0; c < k; ++c) {
        vmul(prod, rgsws + c * rg, dig, (uint32_t)rg);
        memcpy(sum, prod, sizeof(uint64_t) * 2 * N);
        for (uint32_t r = 1; r < 2 * l; ++r) vadd(sum, sum, prod + (size_t)r * 2 * N, 2 * N);
        vmul(sum, sum, weights + (size_t)c * N, N);
        vmul(sum + N, sum + N, weights + (size_t)c * N, N);
        vadd(acc, acc, sum, 2 * N);
    }
}

static size_t multi_mux_ws(uint32_t N, uint32_t l) { return (size_t)4 * N + (size_t)8 * N * l; }

OR_EXPORT void or_multi_mux(const or_ring* R, uint64_t* acc, uint32_t k, const uint64_t* rgsws,
                            const uint64_t* weights, uint32_t b, uint32_t l) {
    uint64_t* ws = (uint64_t*)malloc(sizeof(uint64_t) * multi_mux_ws(R->N, l));
    multi_mux(R, acc, k, rgsws, weights, b, l, ws);
    free(ws);
}

/* ------------------------------------------------------------------------ */
/* Blind rotation                                                           */
/* ------------------------------------------------------------------------ */

enum { OR_BINARY = 0, OR_TERNARY = 1, OR_BMMP = 2 };

/* acc <- acc * X^b in the NTT domain (cggi_blind_rotation.cpp:336-355). */
static void mul_by_xb(const or_ring* R, uint64_t* acc, uint64_t bb, uint64_t* tmp) {
    const uint32_t N = R->N;
    if (bb >= N) {
        bb -= N;
        for (uint32_t i = 0; i < 2 * N; ++i) tmp[i] = neg(acc[i]);
        vmul(acc, tmp, R->mono + bb * N, N);
        vmul(acc + N, tmp + N, R->mono + bb * N, N);
    } else {
        vmul(acc, acc, R->mono + bb * N, N);
        vmul(acc + N, acc + N, R->mono + bb * N, N);
    }
}

/* NTT(+-X^e) for e in [0, 2N). */
static void signed_mono(const or_ring* R, uint64_t* out, uint64_t e) {
    const uint32_t N = R->N;
    if (e >= N) for (uint32_t i = 0; i < N; ++i) out[i] = neg(R->mono[(e - N) * N + i]);
    else memcpy(out, R->mono + e * N, sizeof(uint64_t) * N);
}

/* BMMP step polynomials (bmmp_blind_rotation.cpp:280-323):
 * [X^a0 - 1 | X^a1 - 1 | X^(a0+a1) - X^a0 - X^a1 + 1], NTT domain. */
static void bmmp_step_polys(const or_ring* R, uint64_t* p, uint64_t a0, uint64_t a1) {
    const uint32_t N = R->N;
    const uint64_t sum = (a0 + a1) & (2 * N - 1);
    signed_mono(R, p, a0);
    signed_mono(R, p + N, a1);
    signed_mono(R, p + 2 * N, sum);
    vsub(p + 2 * N, p + 2 * N, p, N);
    vsub(p + 2 * N, p + 2 * N, p + N, N);
    vsub(p, p, R->mono, N);
    vsub(p + N, p + N, R->mono, N);
    vadd(p + 2 * N, p + 2 * N, R->mono, N);
}

/* BlindRotate (cggi_blind_rotation.cpp:236-262, 313-472; bmmp_blind_rotation.cpp:325-419).
 * lwe: n+1 values in Z_2N (b at index n).  acc_coef: 2N coefficient-domain [a|b].
 * result: 2N, NTT domain unless as_coef.  result may alias acc_coef. */
OR_EXPORT int or_blind_rotate(const or_ring* R, uint64_t* result, const uint64_t* lwe, const uint64_t* acc_coef,
                              const uint64_t* brk, uint32_t n, uint32_t b, uint32_t l, int method, int as_coef) {
    const uint32_t N = R->N;
    if (method == OR_BMMP && (n & 1)) return -1;
    uint64_t* ws = (uint64_t*)malloc(sizeof(uint64_t) * (multi_mux_ws(N, l) + 8 * (size_t)N));
    uint64_t* mono = ws + multi_mux_ws(N, l); /* 3N */
    uint64_t* tmp = mono + 3 * N;             /* 2N */
    if (result != acc_coef) memcpy(result, acc_coef, sizeof(uint64_t) * 2 * N);
    ntt_fwd(R, result);
    ntt_fwd(R, result + N);
    mul_by_xb(R, result, lwe[n], tmp);
    if (method == OR_BINARY) {
        const size_t rg = (size_t)4 * N * l;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t ai = lwe[i];
            if (ai == 0) continue;
            signed_mono(R, mono, 2 * N - ai);            /* X^(2N - a_i) */
            vsub(mono, mono, R->mono, N);                 /* - NTT(1) */
            binary_cmux(R, result, brk + i * rg, mono, b, l, ws);
        }
    } else if (method == OR_TERNARY) {
        const size_t rg = (size_t)8 * N * l;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t ai = lwe[i];
            if (ai == 0) continue;
            signed_mono(R, mono, 2 * N - ai);            /* X^(-a_i) */
            signed_mono(R, mono + N, ai);                 /* X^(a_i) */
            vsub(mono, mono, R->mono, N);
            vsub(mono + N, mono + N, R->mono, N);
            multi_mux(R, result, 2, brk + i * rg, mono, b, l, ws);
        }
    } else if (method == OR_BMMP) {
        const size_t rg = (size_t)12 * N * l;
        const uint64_t m2 = 2 * N - 1;
        for (uint32_t i = 0; i < n; i += 2) {
            const uint64_t a0 = (2 * N - lwe[i]) & m2, a1 = (2 * N - lwe[i + 1]) & m2;
            bmmp_step_polys(R, mono, a0, a1);
            multi_mux(R, result, 3, brk + (i >> 1) * rg, mono, b, l, ws);
        }
    } else {
        free(ws);
        return -2;
    }
    if (as_coef) {
        ntt_inv(R, result);
        ntt_inv(R, result + N);
    }
    free(ws);
    return 0;
}

/* SampleExtract at index 0 (static/sample_extraction.h:20-36); zero coefficients
 * stay zero (the reference skips the write into a zeroed output). */
OR_EXPORT void or_sample_extract(uint64_t* out, const uint64_t* rlwe_coef, uint32_t N) {
    out[0] = rlwe_coef[0];
    for (uint32_t i = 1; i < N; ++i) out[i] = neg(rlwe_coef[N - i]);
    out[N] = rlwe_coef[N];
}

/* LWEtoLWEConverter::Convert (lwe_conversion.cpp:205-239) with overwrite semantics:
 * out[k] = -sum d_ij KSK[i][j][k], out[n_out] = in[n_in] - sum d_ij KSK[i][j][n_out]. */
OR_EXPORT int or_lwe_keyswitch(uint64_t* out, const uint64_t* in, const uint64_t* ksk, uint32_t n_in,
                               uint32_t n_out, uint32_t b, uint32_t d) {
    if (d >= max_digits(b) && 64 % b) return -1; /* reference shift would underflow */
    uint64_t least;
    const uint64_t C = or_signed_corrector(b, &least);
    const uint32_t D = max_digits(b);
    const uint64_t mask = (1ULL << b) - 1;
    uint64_t* acc = (uint64_t*)calloc(n_out + 1, sizeof(uint64_t));
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint64_t ai = or_add(in[i], C);
        const uint64_t* chunk = ksk + (size_t)i * d * (n_out + 1);
        uint32_t shift = 64 - b;
        for (uint32_t j = 0; j < d; ++j, chunk += n_out + 1) {
            const uint64_t delta = (j != D - 1) ? (1ULL << (b - 1)) : least;
            const uint64_t dij = or_sub((ai >> shift) & mask, delta);
            for (uint32_t k = 0; k <= n_out; ++k) acc[k] = or_add(acc[k], or_mul(chunk[k], dij));
            shift -= b;
        }
    }
    for (uint32_t k = 0; k < n_out; ++k) out[k] = neg(acc[k]);
    out[n_out] = or_sub(in[n_in], acc[n_out]);
    free(acc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Seeded key generation (replaces the OpenFHE samplers, which have no seed  */
/* API: base_crypto.cpp:34-38,139-142).  Byte layouts follow                */
/* base_crypto.cpp:27-97 and cggi_blind_rotation.cpp:264-311.               */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t s; } or_rng;

static uint64_t rng_next(or_rng* g) {
    uint64_t z = (g->s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static or_rng rng_make(uint64_t seed, uint64_t tag) {
    or_rng g;
    g.s = seed ^ (tag * 0xD1B54A32D192ED03ULL) ^ 0xA0761D6478BD642FULL;
    return g;
}
static uint64_t rng_uniform_q(or_rng* g) {
    uint64_t x;
    do { x = rng_next(g); } while (x >= OR_Q);
    return x;
}

/* Discrete Gaussian.  sigma <= 21: inversion of a 63-bit cumulative table over |x| <= 12 sigma.
 * sigma > 21: the rejection sampler of Canonne, Kamath and Steinke ("The Discrete Gaussian for
 * Differential Privacy", 2020, Algorithms 1-3) -- a discrete Laplace proposal of scale
 * t = floor(sigma) + 1 accepted with probability exp(-(|y| - sigma^2/t)^2 / (2 sigma^2)),
 * Bernoulli(exp(-g)) as a 53-bit uniform below exp(-g).  Same draws, in the same order, as the
 * library's keygen (both consume the seeded stream above). */
typedef struct { double sigma; uint32_t K; uint64_t T[256]; } or_gauss;
#define OR_GAUSS_TABLE_MAX 21.0
static void gauss_init(or_gauss* G, double sigma) {
    G->sigma = sigma;
    G->K = 0;
    if (sigma <= 0) { G->T[0] = 1ULL << 63; return; }
    if (sigma > OR_GAUSS_TABLE_MAX) return;
    uint32_t K = (uint32_t)ceil(sigma * 12.0);
    double w[256], tot = 0;
    for (uint32_t k = 0; k <= K; ++k) {
        w[k] = exp(-(double)(k * k) / (2.0 * sigma * sigma));
        tot += (k ? 2.0 * w[k] : w[k]);
    }
    double c = 0;
    for (uint32_t k = 0; k <= K; ++k) {
        c += (k ? 2.0 * w[k] : w[k]);
        double v = ldexp(c / tot, 63);
        G->T[k] = (k == K || v >= 9223372036854775808.0) ? (1ULL << 63) : (uint64_t)v;
    }
    G->K = K;
}
static double rng_unit(or_rng* g) { return (double)(rng_next(g) >> 11) * 0x1.0p-53; }
static uint64_t rng_below(or_rng* g, uint64_t n) {
    if (n <= 1) return 0;
    const uint64_t lim = ~0ULL - (~0ULL % n);
    uint64_t x;
    do { x = rng_next(g); } while (x >= lim);
    return x % n;
}
static int bern_exp(or_rng* g, double gamma) { return rng_unit(g) < exp(-gamma); }
static int64_t dlaplace(or_rng* g, uint64_t t) {
    for (;;) {
        const uint64_t u = rng_below(g, t);
        if (!bern_exp(g, (double)u / (double)t)) continue;
        uint64_t v = 0;
        while (bern_exp(g, 1.0)) ++v;
        const int b = (int)(rng_next(g) >> 63);
        if (b && u == 0 && v == 0) continue;
        const int64_t z = (int64_t)(u + t * v);
        return b ? -z : z;
    }
}
static int64_t gauss_int(const or_gauss* G, or_rng* g) {
    if (G->sigma > OR_GAUSS_TABLE_MAX) {
        const uint64_t t = (uint64_t)floor(G->sigma) + 1;
        const double s2 = G->sigma * G->sigma;
        for (;;) {
            const int64_t y = dlaplace(g, t);
            const double d = fabs((double)y) - s2 / (double)t;
            if (bern_exp(g, d * d / (2.0 * s2))) return y;
        }
    }
    const uint64_t x = rng_next(g);
    const uint64_t u = x >> 1;
    uint32_t k = 0;
    while (k < G->K && u >= G->T[k]) ++k;
    return (x & 1) ? -(int64_t)k : (int64_t)k;
}
static uint64_t gauss_sample(const or_gauss* G, or_rng* g) {
    const int64_t v = gauss_int(G, g);
    return v >= 0 ? (uint64_t)v : neg((uint64_t)(-v));
}
OR_EXPORT void or_sample_gaussian(int64_t* out, size_t count, double sigma, uint64_t seed, uint64_t tag) {
    or_rng g = rng_make(seed, tag);
    or_gauss G;
    gauss_init(&G, sigma);
    for (size_t i = 0; i < count; ++i) out[i] = gauss_int(&G, &g);
}

/* Secret keys.  dist 0: binary {0,1}; 1: ternary P(0)=1/2, P(1)=P(-1)=1/4 with -1 stored as Q-1. */
OR_EXPORT void or_keygen_secret(uint64_t* out, uint32_t len, int dist, uint64_t seed, uint64_t tag) {
    or_rng g = rng_make(seed, tag);
    for (uint32_t i = 0; i < len; ++i) {
        if (dist == 0) out[i] = rng_next(&g) >> 63;
        else {
            const uint64_t r = rng_next(&g) >> 62;
            out[i] = r < 2 ? 0 : (r == 2 ? 1 : OR_Q - 1);
        }
    }
}

/* MakeRLWE (base_crypto.cpp:27-58): a uniform (used as NTT domain), e Gaussian,
 * b = a * s + NTT(e) + msg_ntt. */
static void make_rlwe(const or_ring* R, uint64_t* out, const uint64_t* msg_ntt, const uint64_t* s_ntt,
                      or_rng* g, const or_gauss* G, uint64_t* e) {
    const uint32_t N = R->N;
    for (uint32_t i = 0; i < N; ++i) out[i] = rng_uniform_q(g);
    for (uint32_t i = 0; i < N; ++i) e[i] = gauss_sample(G, g);
    ntt_fwd(R, e);
    for (uint32_t i = 0; i < N; ++i)
        out[N + i] = or_add(or_add(or_mul(out[i], s_ntt[i]), e[i]), msg_ntt[i]);
}

/* MakeRGSW (base_crypto.cpp:60-97) for a constant message m, with gadget weights g[0..l):
 * row r < l encrypts -s m g_r, row l + r encrypts m g_r.  Rows are produced in the reference's
 * order (r = l-1 down to 0, the -s m row first), which fixes the random-draw order. */
static void make_rgsw_const(const or_ring* R, uint64_t* out, uint64_t m, const uint64_t* gad, uint32_t l,
                            const uint64_t* s_ntt, or_rng* g, const or_gauss* G, uint64_t* ws /* 3N */) {
    const uint32_t N = R->N;
    uint64_t* msg = ws;      /* m g (NTT of a constant is the constant in every slot) */
    uint64_t* msgs = ws + N; /* -s m g */
    uint64_t* e = ws + 2 * N;
    for (uint32_t t = 0; t < l; ++t) {
        const uint32_t r = l - 1 - t;
        const uint64_t mg = or_mul(m, gad[r]);
        for (uint32_t i = 0; i < N; ++i) { msg[i] = mg; msgs[i] = neg(or_mul(mg, s_ntt[i])); }
        make_rlwe(R, out + (size_t)r * 2 * N, msgs, s_ntt, g, G, e);
        make_rlwe(R, out + (size_t)r * 2 * N + (size_t)2 * N * l, msg, s_ntt, g, G, e);
    }
}

/* Gadget weights.  kind 0 (reference MakeRGSW, base_crypto.cpp:68,85-95): L^(D-1-j).
 * kind 1 (signed decomposer, gadget_decomp.h:176-193): 2^(64 - b(j+1)), the j = D-1 digit weight 1.
 * The two agree whenever b divides 64. */
OR_EXPORT void or_gadget(uint64_t* g, uint32_t b, uint32_t l, int kind) {
    const uint32_t D = max_digits(b);
    for (uint32_t j = 0; j < l; ++j) {
        if (kind == 0) g[j] = 1ULL << (b * (D - 1 - j));
        else g[j] = (j == D - 1) ? 1 : (1ULL << (64 - b * (j + 1)));
    }
}

/* Blind-rotation key.  method 0 binary: n x RGSW(s_i); 1 ternary: n x [RGSW(s_i==1) | RGSW(s_i==-1)];
 * 2 BMMP: n/2 x [RGSW(s_2i) | RGSW(s_2i+1) | RGSW(s_2i s_2i+1)].  gadget_kind as or_gadget. */
OR_EXPORT int or_keygen_brk(const or_ring* R, uint64_t* brk, const uint64_t* lwe_key, const uint64_t* rlwe_key,
                            uint32_t n, uint32_t b, uint32_t l, int method, int gadget_kind, double sigma,
                            uint64_t seed) {
    const uint32_t N = R->N;
    or_rng g = rng_make(seed, 3);
    or_gauss G;
    gauss_init(&G, sigma);
    uint64_t* s_ntt = (uint64_t*)malloc(sizeof(uint64_t) * 4 * N);
    uint64_t* ws = s_ntt + N;
    uint64_t gad[64];
    or_gadget(gad, b, l, gadget_kind);
    memcpy(s_ntt, rlwe_key, sizeof(uint64_t) * N);
    ntt_fwd(R, s_ntt);
    const size_t rg = (size_t)4 * N * l;
    for (uint32_t i = 0; i < n; ++i) {
        if (method == OR_BINARY) {
            make_rgsw_const(R, brk + i * rg, lwe_key[i], gad, l, s_ntt, &g, &G, ws);
        } else if (method == OR_TERNARY) {
            const uint64_t s = lwe_key[i];
            const uint64_t p = (s == 1), m = (s != 0 && s != 1); /* cggi_blind_rotation.cpp:294-305 */
            make_rgsw_const(R, brk + 2 * i * rg, p, gad, l, s_ntt, &g, &G, ws);
            make_rgsw_const(R, brk + (2 * i + 1) * rg, m, gad, l, s_ntt, &g, &G, ws);
        } else if (method == OR_BMMP) {
            if (i & 1) continue;
            uint64_t* chunk = brk + (size_t)3 * (i >> 1) * rg;
            make_rgsw_const(R, chunk, lwe_key[i], gad, l, s_ntt, &g, &G, ws);
            make_rgsw_const(R, chunk + rg, lwe_key[i + 1], gad, l, s_ntt, &g, &G, ws);
            make_rgsw_const(R, chunk + 2 * rg, or_mul(lwe_key[i], lwe_key[i + 1]), gad, l, s_ntt, &g, &G, ws);
        }
    }
    free(s_ntt);
    return 0;
}

/* MakeLWE (base_crypto.cpp:137-155): a uniform, e Gaussian, b = e + <a, s> + m. */
static void make_lwe(uint64_t* out, uint64_t m, const uint64_t* s, uint32_t n, or_rng* g, const or_gauss* G) {
    for (uint32_t i = 0; i < n; ++i) out[i] = rng_uniform_q(g);
    uint64_t bb = gauss_sample(G, g);
    for (uint32_t i = 0; i < n; ++i) bb = or_add(bb, or_mul(out[i], s[i]));
    out[n] = or_add(bb, m);
}

/* LWE->LWE key-switching key (lwe_conversion.cpp:174-203): [n_in][d][n_out+1],
 * entry (i, j) = LWE_t(2^(64 - b(j+1)) * s_src[i]). */
OR_EXPORT void or_keygen_lwe_ksk(uint64_t* ksk, const uint64_t* src, const uint64_t* dst, uint32_t n_in,
                                 uint32_t n_out, uint32_t b, uint32_t d, double sigma, uint64_t seed) {
    or_rng g = rng_make(seed, 4);
    or_gauss G;
    gauss_init(&G, sigma);
    for (uint32_t i = 0; i < n_in; ++i) {
        uint64_t gij = 1ULL << (64 - b);
        for (uint32_t j = 0; j < d; ++j) {
            make_lwe(ksk + ((size_t)i * d + j) * (n_out + 1), or_mul(gij, src[i]), dst, n_out, &g, &G);
            gij >>= b;
        }
    }
}

/* RLWE encryption of a coefficient-domain message under s (coefficients); output NTT domain. */
OR_EXPORT void or_encrypt_rlwe(const or_ring* R, uint64_t* out, const uint64_t* msg_coef, const uint64_t* s_coef,
                               double sigma, uint64_t seed, uint64_t tag) {
    const uint32_t N = R->N;
    or_rng g = rng_make(seed, tag);
    or_gauss G;
    gauss_init(&G, sigma);
    uint64_t* ws = (uint64_t*)malloc(sizeof(uint64_t) * 3 * N);
    memcpy(ws, s_coef, sizeof(uint64_t) * N);
    ntt_fwd(R, ws);
    memcpy(ws + N, msg_coef, sizeof(uint64_t) * N);
    ntt_fwd(R, ws + N);
    make_rlwe(R, out, ws + N, ws, &g, &G, ws + 2 * N);
    free(ws);
}

/* Uniform values in [0, bound), bound a power of two (LWE inputs in Z_2N). */
OR_EXPORT void or_random_mod_pow2(uint64_t* out, size_t len, uint64_t bound, uint64_t seed, uint64_t tag) {
    or_rng g = rng_make(seed, tag);
    for (size_t i = 0; i < len; ++i) out[i] = rng_next(&g) & (bound - 1);
}

/* Gaussian samples (as values mod Q), for tests of the sampler. */
OR_EXPORT void or_gauss_samples(uint64_t* out, size_t len, double sigma, uint64_t seed, uint64_t tag) {
    or_rng g = rng_make(seed, tag);
    or_gauss G;
    gauss_init(&G, sigma);
    for (size_t i = 0; i < len; ++i) out[i] = gauss_sample(&G, &g);
}

/* PhaseRLWE (base_crypto.cpp:99-108): coefficients of b - a s, for an NTT-domain RLWE. */
OR_EXPORT void or_phase_rlwe(const or_ring* R, uint64_t* out, const uint64_t* rlwe_ntt, const uint64_t* s_coef) {
    const uint32_t N = R->N;
    uint64_t* s = (uint64_t*)malloc(sizeof(uint64_t) * N);
    memcpy(s, s_coef, sizeof(uint64_t) * N);
    ntt_fwd(R, s);
    for (uint32_t i = 0; i < N; ++i) out[i] = or_sub(rlwe_ntt[N + i], or_mul(rlwe_ntt[i], s[i]));
    ntt_inv(R, out);
    free(s);
}

/* PhaseLWE (base_crypto.cpp:157-169): b - <a, s> mod Q. */
OR_EXPORT uint64_t or_phase_lwe(const uint64_t* lwe, const uint64_t* s, uint32_t n) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) acc = or_add(acc, or_mul(lwe[i], s[i]));
    return or_sub(lwe[n], acc);
}

/* ------------------------------------------------------------------------ */
/* RLWE key switching, automorphism, trace, packing                         */
/* ------------------------------------------------------------------------ */

/* RLWEtoRLWEConverter key (rlwe_conversion.cpp:187-218): d rows [2N], row i = RLWE_t(g_i s_src),
 * g_i = 2^(64 - b(i+1)) (kept >= 1), message in coefficients, NTT-domain output. */
static void keygen_rlwe_ksk_rng(const or_ring* R, uint64_t* ksk, const uint64_t* src, const uint64_t* t_ntt,
                                uint32_t b, uint32_t d, or_rng* g, const or_gauss* G) {
    const uint32_t N = R->N;
    uint64_t* ws = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    uint64_t gij = 1ULL << (64 - b);
    for (uint32_t i = 0; i < d; ++i) {
        for (uint32_t k = 0; k < N; ++k) ws[k] = or_mul(src[k], gij);
        ntt_fwd(R, ws);                       /* NTT(e + m) = NTT(e) + NTT(m) */
        make_rlwe(R, ksk + (size_t)i * 2 * N, ws, t_ntt, g, G, ws + N);
        gij >>= b;
        if (gij == 0) gij = 1;
    }
    free(ws);
}

OR_EXPORT void or_keygen_rlwe_ksk(const or_ring* R, uint64_t* ksk, const uint64_t* src, const uint64_t* dst,
                                  uint32_t b, uint32_t d, double sigma, uint64_t seed, uint64_t tag) {
    or_rng g = rng_make(seed, tag);
    or_gauss G;
    gauss_init(&G, sigma);
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * R->N);
    memcpy(t, dst, sizeof(uint64_t) * R->N);
    ntt_fwd(R, t);
    keygen_rlwe_ksk_rng(R, ksk, src, t, b, d, &g, &G);
    free(t);
}

/* EvalNegacyclicAutomorphism (utils/math_utils.cpp:70-85): out[(i t) mod N] = +-in[i],
 * negated when (i t mod 2N) >= N. */
OR_EXPORT void or_automorphism_poly(uint64_t* out, const uint64_t* in, uint64_t t, uint32_t N) {
    for (uint64_t i = 0; i < N; ++i) {
        const uint64_t p = (i * t) & (2 * (uint64_t)N - 1);
        out[p & (N - 1)] = p >= N ? neg(in[i]) : in[i];
    }
}

/* Automorphism keys for t = 2^j + 1, j = 1..log2 N (trace_evaluation.cpp:50-62): source s(X^t),
 * target s(X), one seeded stream, keys in order of j.  keys: log2(N) x d x 2N. */
OR_EXPORT void or_keygen_auto_keys(const or_ring* R, uint64_t* keys, const uint64_t* s, uint32_t b, uint32_t d,
                                   double sigma, uint64_t seed) {
    const uint32_t N = R->N;
    or_rng g = rng_make(seed, 5);
    or_gauss G;
    gauss_init(&G, sigma);
    uint64_t* t_ntt = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    uint64_t* s_auto = t_ntt + N;
    memcpy(t_ntt, s, sizeof(uint64_t) * N);
    ntt_fwd(R, t_ntt);
    for (uint32_t j = 1; j <= R->logN; ++j) {
        or_automorphism_poly(s_auto, s, (1ULL << j) + 1, N);
        keygen_rlwe_ksk_rng(R, keys + (size_t)(j - 1) * d * 2 * N, s_auto, t_ntt, b, d, &g, &G);
    }
    free(t_ntt);
}

/* RLWEtoRLWEConverter::Convert (rlwe_conversion.cpp:224-249): in = [a (coefficients) | b (NTT)];
 * out -= sum_i NTT(sdigit_i(a)) * KSK_i (both halves); out_b += b. */
OR_EXPORT void or_rlwe_keyswitch(const or_ring* R, uint64_t* out, const uint64_t* in, const uint64_t* ksk,
                                 uint32_t b, uint32_t d) {
    const uint32_t N = R->N;
    uint64_t* dig = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N * d + sizeof(uint64_t) * 2 * N);
    uint64_t* acc = dig + (size_t)2 * N * d;
    signed_rep2_ntt(R, dig, in, b, d);
    vmul(dig, dig, ksk, 2 * N * d);
    memcpy(acc, dig, sizeof(uint64_t) * 2 * N);
    for (uint32_t i = 1; i < d; ++i) vadd(acc, acc, dig + (size_t)i * 2 * N, 2 * N);
    vsub(out, out, acc, 2 * N);
    vadd(out + N, out + N, in + N, N);
    free(dig);
}

/* AutomorphismEvaluator::Eval (automorphism_evaluation.cpp:150-159, 183-207): input [a|b] in
 * coefficients; a' = tau_t(a), b' = NTT(tau_t(b)); out (zeroed here) = Convert([a'|b']). */
OR_EXPORT void or_automorphism(const or_ring* R, uint64_t* out, const uint64_t* in, uint64_t t, const uint64_t* ksk,
                               uint32_t b, uint32_t d) {
    const uint32_t N = R->N;
    uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    or_automorphism_poly(buf, in, t, N);
    or_automorphism_poly(buf + N, in + N, t, N);
    ntt_fwd(R, buf + N);
    memset(out, 0, sizeof(uint64_t) * 2 * N);
    or_rlwe_keyswitch(R, out, buf, ksk, b, d);
    free(buf);
}

/* TraceEvaluator::Eval (trace_evaluation.cpp:135-162): in coefficients, out NTT:
 * o = in / N; for t = N+1, N/2+1, ..., 3: o += INTT(Auto_t(o)); out = NTT(o). */
OR_EXPORT void or_trace(const or_ring* R, uint64_t* out, const uint64_t* in, const uint64_t* keys, uint32_t b,
                        uint32_t d) {
    const uint32_t N = R->N;
    uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    for (uint32_t i = 0; i < 2 * N; ++i) out[i] = or_mul(in[i], R->n_inv);
    for (int j = (int)R->logN; j >= 1; --j) {
        or_automorphism(R, buf, out, (1ULL << j) + 1, keys + (size_t)(j - 1) * d * 2 * N, b, d);
        ntt_inv(R, buf);
        ntt_inv(R, buf + N);
        vadd(out, out, buf, 2 * N);
    }
    ntt_fwd(R, out);
    ntt_fwd(R, out + N);
}

/* LWEtoRLWEPacker::L0Setup (lwe_to_rlwe_packing.cpp:238-251): LWE (dimension N) -> RLWE in
 * coefficients with b in the constant coefficient, scaled by N^-1. */
static void pack_l0(const or_ring* R, uint64_t* out, const uint64_t* lwe) {
    const uint32_t N = R->N;
    memset(out, 0, sizeof(uint64_t) * 2 * N);
    out[0] = lwe[0];
    for (uint32_t i = 1; i < N; ++i) out[i] = neg(lwe[N - i]);
    out[N] = lwe[N];
    for (uint32_t i = 0; i <= N; ++i) out[i] = or_mul(out[i], R->n_inv);
}

/* X^s p for 0 <= s < N, both halves (coefficients). */
static void rot_poly2(uint64_t* out, const uint64_t* in, uint32_t s, uint32_t N) {
    for (int h = 0; h < 2; ++h)
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t k = i + s;
            if (k < N) out[h * N + k] = in[h * N + i];
            else out[h * N + k - N] = neg(in[h * N + i]);
        }
}

/* LWEtoRLWEPacker::Pack (lwe_to_rlwe_packing.cpp:129-236): len LWEs (dimension N, len a power of
 * two >= 2) -> one RLWE in coefficients.  Leaf k is input[(len-1) ^ brv(k)]; two nodes of level
 * l (left = earlier, right = later) merge into
 *   (R + X^(N/2^(l+1)) L) + tau_(2^(l+1)+1)(R - X^(N/2^(l+1)) L);
 * then the trace steps t = N+1, ..., 2^(log len + 1) + 1 follow when len < N. */
OR_EXPORT int or_pack(const or_ring* R, uint64_t* out, const uint64_t* lwes, uint32_t len, const uint64_t* keys,
                      uint32_t b, uint32_t d) {
    const uint32_t N = R->N;
    uint32_t log_l = 0;
    while ((1u << log_l) < len) ++log_l;
    if (len < 2 || (1u << log_l) != len || len > N) return -1;
    uint64_t* nodes = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N * len);
    uint64_t* rot = (uint64_t*)malloc(sizeof(uint64_t) * 6 * N);
    uint64_t* plus = rot + 2 * N;
    uint64_t* aut = rot + 4 * N;
    for (uint32_t k = 0; k < len; ++k) {
        const uint32_t src = (len - 1) ^ brv(k, log_l);
        pack_l0(R, nodes + (size_t)k * 2 * N, lwes + (size_t)src * (N + 1));
    }
    for (uint32_t lvl = 1, cnt = len / 2; lvl <= log_l; ++lvl, cnt /= 2) {
        const uint32_t shift = N >> lvl;
        for (uint32_t j = 0; j < cnt; ++j) {
            uint64_t* L = nodes + (size_t)(2 * j) * 2 * N * (1u << (lvl - 1));
            uint64_t* Rn = nodes + (size_t)(2 * j + 1) * 2 * N * (1u << (lvl - 1));
            rot_poly2(rot, L, shift, N);
            vadd(plus, Rn, rot, 2 * N);
            vsub(Rn, Rn, rot, 2 * N);
            or_automorphism(R, aut, Rn, (1ULL << lvl) + 1, keys + (size_t)(lvl - 1) * d * 2 * N, b, d);
            ntt_inv(R, aut);
            ntt_inv(R, aut + N);
            vadd(L, plus, aut, 2 * N);  /* the node keeps the left slot */
        }
    }
    memcpy(out, nodes, sizeof(uint64_t) * 2 * N);
    for (uint32_t lvl = R->logN; lvl > log_l; --lvl) {
        or_automorphism(R, aut, out, (1ULL << lvl) + 1, keys + (size_t)(lvl - 1) * d * 2 * N, b, d);
        ntt_inv(R, aut);
        ntt_inv(R, aut + N);
        vadd(out, aut, out, 2 * N);
    }
    free(nodes);
    free(rot);
    return 0;
}

/* LWEtoRLWEPacker::PackConsecutively (lwe_to_rlwe_packing.cpp:260-396): the N-leaf tree walked
 * with a stack, leaf k reading input (N-1) ^ brv(k) when that index is below len and being a
 * "fake" node otherwise; two fake nodes merge into a fake node without work, any other pair
 * merges as in Pack (earlier node = left/neck, later = right/head).  Output: stack bottom once it
 * reaches level log2 N, coefficients. */
OR_EXPORT int or_pack_consecutive(const or_ring* R, uint64_t* out, const uint64_t* lwes, uint32_t len,
                                  const uint64_t* keys, uint32_t b, uint32_t d) {
    const uint32_t N = R->N, logN = R->logN;
    if (len < 1 || len > N) return -1;
    const uint32_t depth = logN + 2;
    uint64_t* st = (uint64_t*)calloc((size_t)depth * 2 * N, sizeof(uint64_t));
    uint64_t* rot = (uint64_t*)malloc(sizeof(uint64_t) * 4 * N);
    uint64_t* aut = rot + 2 * N;
    uint32_t lvl[64], fake[64];
    int top = -1;
    uint32_t next = 0;
    for (;;) {
        if (top == 0 && lvl[0] == logN) break;
        if (top >= 1 && lvl[top] == lvl[top - 1]) {
            const uint32_t o = lvl[top] + 1;
            if (fake[top] && fake[top - 1]) {
                --top;
                lvl[top] = o;
                continue;
            }
            uint64_t* neck = st + (size_t)(top - 1) * 2 * N;
            uint64_t* head = st + (size_t)top * 2 * N;
            rot_poly2(rot, neck, N >> o, N);
            vadd(neck, rot, head, 2 * N);   /* head + X^s neck */
            vsub(head, head, rot, 2 * N);   /* head - X^s neck */
            or_automorphism(R, aut, head, (1ULL << o) + 1, keys + (size_t)(o - 1) * d * 2 * N, b, d);
            ntt_inv(R, aut);
            ntt_inv(R, aut + N);
            vadd(neck, neck, aut, 2 * N);
            --top;
            lvl[top] = o;
            fake[top] = 0;
            continue;
        }
        /* push the next leaf */
        ++top;
        uint64_t* slot = st + (size_t)top * 2 * N;
        const uint32_t src = (N - 1) ^ brv(next, logN);
        if (src < len) {
            pack_l0(R, slot, lwes + (size_t)src * (N + 1));
            fake[top] = 0;
        } else {
            memset(slot, 0, sizeof(uint64_t) * 2 * N);
            fake[top] = 1;
        }
        lvl[top] = 0;
        ++next;
    }
    memcpy(out, st, sizeof(uint64_t) * 2 * N);
    free(st);
    free(rot);
    return 0;
}

/* ModulusSwitch (static/modulus_switching.cpp:36-93) for ROUND (0), FLOOR (1) and CEIL (2) in
 * the reference's own arithmetic: x87 long double (64-bit significand) on x86-64, as gcc compiles
 * the reference there.  RANDOM is unseeded in the reference and not restated. */
OR_EXPORT void or_modswitch(uint64_t* vec, uint64_t n, uint64_t source_modulus, uint64_t target_modulus, int type) {
    const long double sM = (long double)source_modulus, tM = (long double)target_modulus;
    for (uint64_t i = 0; i < n; ++i) {
        const long double x = (tM * (long double)vec[i]) / sM;
        uint64_t r;
        if (type == 0) r = (uint64_t)roundl(x);
        else if (type == 1) r = (uint64_t)floorl(x);
        else r = (uint64_t)ceill(x);
        if (r >= target_modulus) r -= target_modulus;
        vec[i] = r;
    }
}

/* LWEtoRGSWConverter (lwe_to_rgsw_conversion.cpp:116-231).  Accumulator set-up of the
 * constructor (:116-161): with block = N / digits and scale = 2^(b (D - digits)), D = ceil(64 / b),
 * t[i] = scale and t[N-1-i] = -scale for i < block/2, t[d block + j - block/2] = scale 2^(b d) for
 * d >= 1, j < block; acc = [0 | -t/2], extra = t/2.  Convert (:163-231): BR (coefficient output)
 * of acc by the input, b += extra; rows d >= 1 are X^(-d block) times row 0; trace of each row
 * (NTT out); then lev_left row i = (KS_{s^2->s}(INTT(a_i)) + (b_i, 0)) with KS in the subtract-
 * into-zero form negated, i.e. (sum_a + b_i, sum_b).  out = [lev_left: digits x 2N | lev_right:
 * digits x 2N], NTT domain; brk/method/bl/ll: the rotation key, trace keys (log2 N x dt x 2N) with
 * gadget bt, square key (ds x 2N) with gadget bs. */
OR_EXPORT int or_lwe_to_rgsw(const or_ring* R, uint64_t* out, const uint64_t* lwe, const uint64_t* brk, uint32_t n,
                             uint32_t bl, uint32_t ll, int method, const uint64_t* trace_keys, uint32_t bt, uint32_t dt,
                             const uint64_t* sq_key, uint32_t bs, uint32_t ds, uint32_t digits, uint32_t basebits) {
    const uint32_t N = R->N;
    if (digits == 0 || N % digits) return -1;
    const uint32_t block = N / digits;
    const uint32_t D = 64 / basebits + (64 % basebits ? 1 : 0);
    uint64_t* t = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    uint64_t* acc = (uint64_t*)calloc((size_t)3 * N, sizeof(uint64_t));
    uint64_t* buf = (uint64_t*)calloc((size_t)4 * N, sizeof(uint64_t));
    uint64_t scale = 1ULL << (basebits * (D - digits));
    for (uint32_t i = 0; i < block / 2; ++i) {
        t[i] = scale;
        t[N - i - 1] = OR_Q - scale;
    }
    for (uint32_t d = 1; d < digits; ++d) {
        scale <<= basebits;
        for (uint32_t j = 0; j < block; ++j) t[d * block + j - block / 2] = scale;
    }
    const uint64_t inv2 = or_inv(2);
    for (uint32_t i = 0; i < N; ++i) {
        acc[2 * N + i] = or_mul(t[i], inv2);
        acc[N + i] = or_mul(acc[2 * N + i], OR_Q - 1);
    }
    uint64_t* left = out;
    uint64_t* right = out + (size_t)2 * N * digits;
    memset(out, 0, sizeof(uint64_t) * 4 * N * digits);
    int rc = or_blind_rotate(R, right, lwe, acc, brk, n, bl, ll, method, 1);
    if (rc) goto done;
    vadd(right + N, right + N, acc + 2 * N, N);
    memcpy(buf, right, sizeof(uint64_t) * N);
    memcpy(buf + 2 * N, right + N, sizeof(uint64_t) * N);
    for (uint32_t i = 0; i < N; ++i) {
        buf[N + i] = neg(buf[i]);
        buf[3 * N + i] = neg(buf[2 * N + i]);
    }
    for (uint32_t d = 1; d < digits; ++d) {
        uint64_t* row = right + (size_t)d * 2 * N;
        memcpy(row, buf + d * block, sizeof(uint64_t) * N);
        memcpy(row + N, buf + 2 * N + d * block, sizeof(uint64_t) * N);
    }
    for (uint32_t d = 0; d < digits; ++d) {
        uint64_t* row = right + (size_t)d * 2 * N;
        memcpy(buf, row, sizeof(uint64_t) * 2 * N);
        or_trace(R, row, buf, trace_keys, bt, dt);
    }
    memset(buf, 0, sizeof(uint64_t) * 4 * N);
    for (uint32_t i = 0; i < digits; ++i) {
        uint64_t* m = right + (size_t)2 * N * i;
        uint64_t* sm = left + (size_t)2 * N * i;
        memcpy(buf, m, sizeof(uint64_t) * N);
        ntt_inv(R, buf);
        or_rlwe_keyswitch(R, sm, buf, sq_key, bs, ds);      /* sm = -sum (sm starts at 0) */
        vsub(sm, buf + 2 * N, sm, 2 * N);                     /* sm = sum */
        vadd(sm, m + N, sm, N);                               /* a += b_m */
    }
done:
    free(t);
    free(acc);
    free(buf);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* SAP digit decomposition (sap_decomposer.cpp:190-494).                    */
/* ------------------------------------------------------------------------ */

static uint32_t ilog2_ceil(uint64_t x) { /* IntLog2 (utils/math_utils.cpp:11-26) */
    uint32_t r = 0;
    if (x <= 1) return 0;
    for (x -= 1; x; x >>= 1) ++r;
    return r;
}

/* TraceEvaluator::EvalAuto(out, in, t) of the pointer overload: automorphism key log2(t-1) - 1,
 * subtracted into out (trace_evaluation.cpp:172-179, automorphism_evaluation.cpp:150-159). */
static void eval_auto_acc(const or_ring* R, uint64_t* out, const uint64_t* in, uint64_t t, const uint64_t* keys,
                          uint32_t b, uint32_t d) {
    const uint32_t N = R->N;
    const uint32_t j = ilog2_ceil(t - 1);
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * 2 * N);
    or_automorphism(R, tmp, in, t, keys + (size_t)(j - 1) * d * 2 * N, b, d);
    vadd(out, out, tmp, 2 * N);
    free(tmp);
}

/* SAPDecomposer::HomTrunc (:190-312): input RLWE (NTT) -> output RLWE (NTT) whose phase keeps
 * X^(e >> log2 radix) of a phase X^e, through a partial trace and the automorphism tree. */
OR_EXPORT void or_sap_homtrunc(const or_ring* R, uint64_t* output, const uint64_t* input, uint32_t radix,
                               const uint64_t* keys, uint32_t b, uint32_t d) {
    const uint32_t N = R->N, logN = R->logN, lr = ilog2_ceil(radix);
    /* the reference's buffer holds log2 N slots of 2N (:511); the tree reaches slot
     * log2(N/radix) + 2, beyond it for radix 2, so this one holds log2 N + 3 */
    const size_t pk_words = (size_t)2 * logN * N;
    uint64_t* pk = (uint64_t*)calloc(pk_words + (size_t)6 * N, sizeof(uint64_t));
    uint64_t* poly = (uint64_t*)calloc(N, sizeof(uint64_t));
    uint64_t* ares = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    uint64_t* sp = (uint64_t*)calloc(N, sizeof(uint64_t));
    memset(output, 0, sizeof(uint64_t) * 2 * N);
    for (uint32_t i = N - (radix - 1); i < N; ++i) poly[i] = OR_Q - 1;  /* sum_i X^-i, i < radix */
    poly[0] = 1;
    ntt_fwd(R, poly);
    for (uint32_t i = 0; i < 2 * N; ++i) pk[i] = or_mul(input[i], R->n_inv);
    vmul(pk, pk, poly, N);
    vmul(pk + N, pk + N, poly, N);
    for (uint32_t i = 0; i < lr; ++i) {
        const uint64_t t = (1ULL << (logN - i)) + 1;
        memcpy(ares, pk, sizeof(uint64_t) * 2 * N);
        ntt_inv(R, ares);
        ntt_inv(R, ares + N);
        eval_auto_acc(R, pk + 2 * N, ares, t, keys, b, d);
        vadd(pk, pk, pk + 2 * N, 2 * N);
        memset(pk + 2 * N, 0, sizeof(uint64_t) * 2 * N);
    }
    memset(pk + 2 * N, 0, sizeof(uint64_t) * (pk_words - 2 * N));
    uint32_t lvl[128];
    uint64_t ex[128];
    uint32_t head = 1;
    lvl[0] = logN - lr;
    ex[0] = 0;
    const uint32_t rev_bits = logN - lr;
    while (head != 0) {
        --head;
        const uint32_t level = lvl[head];
        const uint64_t e = ex[head];
        uint64_t* cur = pk + (size_t)head * 2 * N;
        memset(sp, 0, sizeof(uint64_t) * N);
        if (level == 0) {
            const uint64_t eb = rev_bits ? brv((uint32_t)e, rev_bits) : 0;
            sp[eb] = 1;
            ntt_fwd(R, sp);
            vmul(cur, cur, sp, N);
            vmul(cur + N, cur + N, sp, N);
            vadd(output, output, cur, 2 * N);
            memset(cur, 0, sizeof(uint64_t) * 2 * N);
            continue;
        }
        sp[N - (N >> level)] = OR_Q - 1;  /* X^(-N / 2^level) */
        ntt_fwd(R, sp);
        memcpy(cur + 2 * N, cur, sizeof(uint64_t) * 2 * N);
        ntt_inv(R, cur + 2 * N);
        ntt_inv(R, cur + 3 * N);
        eval_auto_acc(R, cur + 4 * N, cur + 2 * N, (1ULL << level) + 1, keys, b, d);
        vsub(cur + 2 * N, cur, cur + 4 * N, 2 * N);
        vmul(cur + 2 * N, cur + 2 * N, sp, N);
        vmul(cur + 3 * N, cur + 3 * N, sp, N);
        vadd(cur, cur, cur + 4 * N, 2 * N);
        memset(cur + 4 * N, 0, sizeof(uint64_t) * 2 * N);
        lvl[head] = level - 1;
        ex[head] = 2 * e;
        lvl[head + 1] = level - 1;
        ex[head + 1] = 2 * e + 1;
        head += 2;
    }
    free(pk);
    free(poly);
    free(ares);
    free(sp);
}

/* SAPDecomposer::ResetAccumulator (:314-373) with the key switch and (when Qks != Q, power of two)
 * the ROUND switch to Qks; the FLOOR switch to 2N; acc (NTT) is replaced by a fresh rotation of
 * Q/(2N) by the extracted LWE. */
OR_EXPORT int or_sap_reset(const or_ring* R, uint64_t* acc, uint32_t radix, const uint64_t* ksk, uint32_t n_out,
                           uint32_t kb, uint32_t kd, uint64_t qks, const uint64_t* brk, uint32_t n, uint32_t bl,
                           uint32_t ll, int method) {
    const uint32_t N = R->N, rl = ilog2_ceil(radix);
    uint64_t* poly = (uint64_t*)calloc(N, sizeof(uint64_t));
    uint64_t* ext = (uint64_t*)calloc((size_t)N + 1, sizeof(uint64_t));
    uint64_t* ln = (uint64_t*)calloc((size_t)n_out + 1, sizeof(uint64_t));
    uint64_t* fresh = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    int rc = 0;
    for (uint32_t i = 1; i < N; ++i) poly[i] = neg((uint64_t)(((N - i) >> rl) << rl));
    ntt_fwd(R, poly);
    vmul(acc, acc, poly, N);
    vmul(acc + N, acc + N, poly, N);
    ntt_inv(R, acc);
    ntt_inv(R, acc + N);
    or_sample_extract(ext, acc, N);
    if (qks != OR_Q) or_modswitch(ext, (uint64_t)N + 1, OR_Q, qks, 0);
    rc = or_lwe_keyswitch(ln, ext, ksk, N, n_out, kb, kd);
    if (!rc) {
        or_modswitch(ln, (uint64_t)n_out + 1, qks, 2 * (uint64_t)N, 1);
        fresh[N] = OR_Q / (2 * (uint64_t)N);
        rc = or_blind_rotate(R, acc, ln, fresh, brk, n, bl, ll, method, 0);
    }
    free(poly);
    free(ext);
    free(ln);
    free(fresh);
    return rc;
}

/* SAPDecomposer::Decompose (:375-494): digits of an LWE of power-of-two modulus, radix digits at a
 * time, each a blind rotation of the perturbed sub-phase followed by the x -> x mod radix
 * extraction and HomTrunc.  beta = Hamming weight of the LWE key.  out: digits x (N+1); returns the
 * digit count.  restart: the digit period of the accumulator reset (:480-482; UINT64_MAX = never,
 * the reference's own value, as its trace noise estimate is 0, :101-104); the reset key switches
 * N -> n with ksk (base 2^kb, kd digits). */
OR_EXPORT int or_sap_decompose(const or_ring* R, uint64_t* out, const uint64_t* input, uint32_t radix, uint32_t beta,
                               const uint64_t* brk, uint32_t n, uint32_t bl, uint32_t ll, int method,
                               const uint64_t* keys, uint32_t b, uint32_t d, uint64_t restart, const uint64_t* ksk,
                               uint32_t kb, uint32_t kd) {
    const uint32_t N = R->N, rl = ilog2_ceil(radix);
    const uint64_t mask = (uint64_t)radix - 1, lwe_mod_in = 2 * (uint64_t)N;
    const uint64_t plo = (uint64_t)radix * beta, phi = beta;
    uint64_t* acc = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    uint64_t* xp = (uint64_t*)calloc(N, sizeof(uint64_t));
    uint64_t* ext = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    uint64_t* tmp = (uint64_t*)calloc((size_t)2 * N, sizeof(uint64_t));
    uint64_t* cp = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1));
    uint64_t* sub = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1));
    int digit = 0;
    acc[N] = OR_Q / (2 * (uint64_t)N);
    ntt_fwd(R, acc + N);
    for (uint32_t i = 1; i < N; ++i) xp[N - i] = neg(i & mask);
    ntt_fwd(R, xp);
    memcpy(cp, input, sizeof(uint64_t) * (n + 1));
    /* 1 << IntLog2(a_0): the reference's x86 shift takes the count mod 64, so 1 << 64 is 1 */
    uint64_t cm = 1ULL << (ilog2_ceil(input[0]) & 63);
    for (uint32_t i = 1; i <= n; ++i)
        if (cm <= input[i]) cm <<= 1;
    while (cm > 1) {
        for (uint32_t i = 0; i <= n; ++i) {
            sub[i] = cp[i] & mask;
            cp[i] >>= rl;
        }
        cm >>= rl;
        sub[n] = (uint64_t)(((unsigned __int128)sub[n] + plo) % lwe_mod_in);
        if (cm > 1) cp[n] = (cp[n] % cm + cm - phi % cm) % cm;
        ntt_inv(R, acc);
        ntt_inv(R, acc + N);
        memcpy(tmp, acc, sizeof(uint64_t) * 2 * N);
        int rc = or_blind_rotate(R, acc, sub, tmp, brk, n, bl, ll, method, 0);
        if (rc) { digit = -1; break; }
        vmul(ext, acc, xp, N);
        vmul(ext + N, acc + N, xp, N);
        ntt_inv(R, ext);
        ntt_inv(R, ext + N);
        or_sample_extract(out + (size_t)digit * (N + 1), ext, N);
        if (restart != UINT64_MAX && (uint64_t)(digit + 1) % restart == 0 && cm != 0) {
            /* :480-482: reset with radix 0, key switch N -> n, no switch to another KS modulus */
            if (or_sap_reset(R, acc, 0, ksk, n, kb, kd, OR_Q, brk, n, bl, ll, method)) { digit = -1; break; }
        }
        or_sap_homtrunc(R, tmp, acc, radix, keys, b, d);
        memcpy(acc, tmp, sizeof(uint64_t) * 2 * N);
        ++digit;
    }
    free(acc);
    free(xp);
    free(ext);
    free(tmp);
    free(cp);
    free(sub);
    return digit;
}

/* ------------------------------------------------------------------------ */
/* Batched bootstrap for the CPU baseline: EvalFunc (negacyclic_function_    */
/* evaluator.cpp:92-101) + LWE key switch, one rotator per thread.          */
/* ------------------------------------------------------------------------ */

typedef struct {
    const or_ring* R;
    const uint64_t *lwes, *tv, *brk, *ksk;
    uint64_t* out;
    uint32_t n, b, l, method, n_out, ks_b, ks_d;
    size_t begin, end;
} bs_job;

static void* bs_worker(void* arg) {
    bs_job* J = (bs_job*)arg;
    const uint32_t N = J->R->N;
    uint64_t* acc = (uint64_t*)malloc(sizeof(uint64_t) * (3 * N + 1));
    uint64_t* ext = acc + 2 * N;
    for (size_t c = J->begin; c < J->end; ++c) {
        or_blind_rotate(J->R, acc, J->lwes + c * (J->n + 1), J->tv, J->brk, J->n, J->b, J->l, (int)J->method, 1);
        or_sample_extract(ext, acc, N);
        if (J->ksk)
            or_lwe_keyswitch(J->out + c * (J->n_out + 1), ext, J->ksk, N, J->n_out, J->ks_b, J->ks_d);
        else
            memcpy(J->out + c * (N + 1), ext, sizeof(uint64_t) * (N + 1));
    }
    free(acc);
    return NULL;
}

/* out: count x (n_out+1) if ksk != NULL, else count x (N+1).  Returns elapsed seconds. */
OR_EXPORT double or_bootstrap_batch(const or_ring* R, uint64_t* out, const uint64_t* lwes, size_t count,
                                    const uint64_t* tv, const uint64_t* brk, uint32_t n, uint32_t b, uint32_t l,
                                    int method, const uint64_t* ksk, uint32_t n_out, uint32_t ks_b, uint32_t ks_d,
                                    int threads) {
    if (threads < 1) threads = 1;
    pthread_t th[256];
    bs_job jobs[256];
    if (threads > 256) threads = 256;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; ++t) {
        bs_job J = {R, lwes, tv, brk, ksk, out, n, b, l, (uint32_t)method, n_out, ks_b, ks_d,
                    count * t / threads, count * (t + 1) / threads};
        jobs[t] = J;
        pthread_create(&th[t], NULL, bs_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
