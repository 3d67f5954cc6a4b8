// Kernel argument blocks and launchers shared by the kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tp {

struct BrArgs {
    const uint64_t* brk;   // device key layout (rows permuted lane-major, see key_pos)
    const uint64_t* lwe;   // [count][n+1], values in Z_2N, b at index n
    const uint64_t* tv;    // test vector / accumulator, coefficient domain [a|b]
    size_t tv_stride;      // 0: one shared tv, 2N: one per ciphertext
    uint64_t* out_lwe;     // [count][N+1] sample-extracted LWE, or null
    uint64_t* out_rlwe;    // [count][2N] rotated accumulator, or null
    int out_rlwe_ntt;      // 1: NTT domain (reference order), 0: coefficients
    const uint64_t* tw;    // psi_rev
    const uint64_t* itw;   // psi_inv_rev
    int n;                 // LWE dimension
    int base_log2;         // gadget base bits
    int first_shift;       // unsigned digits: b (ceil(64/b) - 1)
    int count;
    // signed digits (ternary, BMMP): x' = x + corrector mod Q, digit j = ((x' >> (64 - b(j+1)))
    // & mask) - 2^(b-1), except the digit j = max_digits - 1 (shift 0, half-range `least`)
    uint64_t corrector;
    uint64_t least;
    int max_digits;
    const uint64_t* mono;  // NTT(X^e), e < N, key-row order (ternary, BMMP)
    // chunked (continuous-batching) mode; all null / 0 for one launch over whole rotations
    const uint32_t* active;    // [count] slot processed by work-group ct (null: slot = ct)
    uint32_t* slot_step;       // [slots] next blind-rotation step of each slot (null: start at 0)
    uint64_t* state;           // [slots][2N] accumulator between launches (coefficients)
    const uint32_t* tv_index;  // [slots] test-vector row of each slot (null: tv_stride rule)
    int chunk_steps;           // steps per launch (0: run to the end)
};

struct KsArgs {
    const uint64_t* ksk;   // [n_in][digits][n_out+1] (reference layout)
    const uint64_t* corr;  // [n_out+1] sum_ij delta_j KSK[i][j][:]
    const uint64_t* in;    // [count][in_stride], n_in+1 used
    uint64_t* out;         // [count][n_out+1]
    size_t in_stride;
    uint64_t corrector;    // signed-to-unsigned corrector C mod Q
    int n_in, n_out, base_log2, digits, count;
    const uint32_t* rows;  // [count] input row of ciphertext c (null: row c); output row is c
    const int8_t* kl;      // the key as matrix-core operands (ks_limbs.hpp), or null
};

// RLWE key switching with signed digits (automorphism, trace and packing keys).
struct KeySwitchArgs {
    const uint64_t* tw;
    const uint64_t* itw;
    uint64_t corrector;   // signed-digit corrector for base_log2
    uint64_t least;
    int base_log2;
    int digits;
    int max_digits;
};

hipError_t launch_automorphism(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                               uint32_t t, const uint64_t* ksk, hipStream_t s);
hipError_t launch_trace(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                        const uint64_t* keys, hipStream_t s);
// Packing tree over len leaves; leaves whose input index is >= n_inputs are zero.
hipError_t launch_pack(int logN, const KeySwitchArgs& A, const uint64_t* lwes, const uint32_t* perm, int len,
                       int n_inputs, uint64_t* nodes, const uint64_t* keys, hipStream_t s);
// mode 0 convert ([a coef | b NTT]), 1 automorphism ([a|b] coef), 2 automorphism of a only.
hipError_t launch_rlwe_keyswitch(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                                uint32_t t, int mode, int accumulate, const uint64_t* ksk, hipStream_t s);
// Automorphism key switch of NTT-domain [a|b] rows (a inverse-transformed in registers, b permuted
// in the evaluation domain); out (+)= the mode-1 result; in may equal out.
hipError_t launch_auto_ntt(int logN, const KeySwitchArgs& A, const uint64_t* in, uint64_t* out, int count,
                           uint32_t t, int accumulate, const uint64_t* ksk, hipStream_t s);
// One level of the SAP HomTrunc tree over nodes x count NTT-domain RLWEs X[e][i] (t = 2^level + 1):
// mode 0 writes the children x + y and (x - y) polys0 to out[2e][i], out[2e + 1][i]; mode 1 writes
// the leaf contributions (x + y) polys0[2e] + (x - y) polys1[e] to out[e][i].
hipError_t launch_sap_tree(int logN, const KeySwitchArgs& A, const uint64_t* X, uint64_t* out, const uint64_t* polys0,
                           const uint64_t* polys1, int nodes, int count, uint32_t t, int mode, const uint64_t* ksk,
                           hipStream_t s);
// Modulus switch to 2^t (kind: 0 round, 1 floor, 2 ceil, 3 random with a seeded coin).
hipError_t launch_modswitch(uint64_t* dst, const uint64_t* src, size_t words, uint64_t sM, int t, int kind,
                            uint64_t seed, hipStream_t s);
// Elementwise mod-Q arithmetic (op codes TP_ELT_* of the C-ABI).
hipError_t launch_eltwise(int op, uint64_t* dst, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                          uint64_t scalar, size_t words, hipStream_t s);

hipError_t launch_ntt_batch(int logN, uint64_t* out, const uint64_t* in, size_t count, const uint64_t* tw,
                            bool inverse, hipStream_t s);
// waves: 2 (256-register variants) or 1 (512-register variants, digits 3 and 4 only).
hipError_t launch_blind_rotate(int logN, int method, int digits, const BrArgs& A, int grid, int waves,
                               hipStream_t s);
hipError_t launch_blind_rotate_w1(int logN, int method, int digits, const BrArgs& A, int grid, hipStream_t s);
// Binary CGGI at N = 1024 with the exact floating-point external product (kernels_br_fft.hip):
// fkey the transform-domain key limbs, zt the twiddle table (fft_key.hpp); waves 1 or 2 per SIMD.
hipError_t launch_blind_rotate_fft(int digits, const BrArgs& A, const void* fkey, const void* zt, int grid, int waves,
                                   hipStream_t s);
// Digit-parallel variant at N = 1024 (digits 2..4): the 2 l digit products of a step run on 2 l
// waves of one work-group per ciphertext; `grid` work-groups stride over A.count ciphertexts.
hipError_t launch_blind_rotate_latency(int method, int digits, const BrArgs& A, int grid, hipStream_t s);
// binary CGGI at N = 1024, 2 or 3 digits: the digit-parallel kernel with every transform split over
// two waves (4 * digits waves per work-group; A.tw and A.itw hold the two-wave tables 3N words in)
hipError_t launch_blind_rotate_half(int digits, const BrArgs& A, int grid, hipStream_t s);
hipError_t launch_lwe_keyswitch(const KsArgs& A, hipStream_t s);
// the matrix-core form (kernels_ks_mfma.hip); A.kl set, ks_mfma_fits(n_in, base_log2, digits)
hipError_t launch_lwe_keyswitch_mfma(const KsArgs& A, hipStream_t s);
// SAP digit decomposition (kernels_sap.hip): per-input modulus and digit count; one digit step
// of the bookkeeping; products of rows of N words with one NTT-domain polynomial; a HomTrunc leaf
// whose tree is a single leaf; the sum of the tree's leaf contributions; sample extraction into
// digit `digit` of each row's output.
hipError_t launch_sap_init(const uint64_t* lwes, uint64_t* cm, uint32_t* digits, int n, int count, int rl,
                           hipStream_t s);
hipError_t launch_sap_digit(uint64_t* cp, uint64_t* cm, uint64_t* sub, int n, int count, int rl, uint64_t mask,
                            uint64_t plo, uint64_t phi, uint64_t two_n, hipStream_t s);
hipError_t launch_sap_poly_mul(uint64_t* dst, const uint64_t* src, const uint64_t* poly, size_t rows, uint32_t N,
                               hipStream_t s);
hipError_t launch_sap_leaf(uint64_t* out, const uint64_t* x, const uint64_t* mono, size_t rows, uint32_t N,
                           int accumulate, hipStream_t s);
// out (+)= sum over `nodes` consecutive blocks of rows x N words of P.
hipError_t launch_sap_sum_nodes(uint64_t* out, const uint64_t* P, int nodes, size_t rows, uint32_t N, int accumulate,
                                hipStream_t s);
hipError_t launch_sap_extract(uint64_t* out, const uint64_t* rlwe, const uint32_t* digits, uint32_t digit,
                              size_t out_stride, int count, uint32_t N, hipStream_t s);
// LWE -> RGSW (kernels_rgsw.hip): the digit rows of the rotated accumulator, and the left Lev rows
// from the key-switched ones.
hipError_t launch_rgsw_expand(uint64_t* rows, const uint64_t* row0, const uint64_t* extra, int count, uint32_t digits,
                              uint32_t block, uint32_t N, hipStream_t s);
hipError_t launch_rgsw_left(uint64_t* left, const uint64_t* right, size_t rows, uint32_t N, hipStream_t s);
constexpr int kKsPadWords = 16;  // zero words after the device key-switching key

}  // namespace tp
