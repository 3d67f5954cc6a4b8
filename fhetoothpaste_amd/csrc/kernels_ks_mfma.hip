// LWE -> LWE key switch on the matrix cores (gfx950 v_mfma_i32_32x32x32_i8).
//
// The same sums as lwe_keyswitch_kernel (lwe_conversion.cpp:205-239): with unsigned digits
// u_ij = digit_j(x_i + C) < 2^b, out[c][col] = -(sum_ij u_ij K[i][j][col] - corr[col]) and the last
// column b - (...).  Written as a matrix product: row c of the digit matrix holds ciphertext c's
// digits, 16 slots per input word (ks_limbs.hpp); the key matrix holds each key word as eight
// biased bytes K_t - 128.  For each byte t the matrix cores accumulate S_t = sum u (K_t - 128) in
// 32 bits (exact: ks_mfma_fits), and sum_t 2^(8t) (S_t + 128 U), U = sum_ij u_ij, is the exact
// integer sum, reduced mod Q once per output word.
//
// Work-group: 4 waves, 128 ciphertexts (32 per wave) x one group of 32 output columns.  A wave's
// eight 32x32 accumulators are the eight bytes of the same 32 columns, so every lane ends with all
// eight bytes of its outputs in its own registers.  Digits are formed in registers from the input
// words (no digit matrix in memory); the key operands, stored in lane order, are staged through LDS
// a chunk of four k-steps ahead and shared by the four waves.
#include <hip/hip_runtime.h>
#include "gl64.hpp"
#include "kernels.hpp"

namespace tp {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kKmWaves = 4;
constexpr int kKmRows = 32 * kKmWaves;  // ciphertexts per work-group
constexpr int kKmChunk = 4;             // k-steps (two input words each) per staged chunk
constexpr int kKmStep = 8 * 64;         // 16-byte key operands per k-step (8 bytes x 64 lanes)
constexpr int kKmStage = kKmChunk * kKmStep / (64 * kKmWaves);  // operands each thread stages
static_assert(kKmStage * 64 * kKmWaves == kKmChunk * kKmStep, "chunk divides evenly over the work-group");

// The 16 digit slots of one word (ks_slot_digit) as four dwords, and their sum into usum.
// b = 4: the nibbles of v = (x + C) >> (64 - 4 d), even ones to slots 0..7, odd ones to 8..15.
template <int B>
__device__ __forceinline__ v4i word_digits(uint64_t x, int d, const uint32_t (&bm)[4], uint32_t& usum) {
    v4i f;
    if constexpr (B == 4) {
        const uint64_t v = x >> (64 - 4 * d);
        const uint64_t lo = v & 0x0F0F0F0F0F0F0F0FULL, hi = (v >> 4) & 0x0F0F0F0F0F0F0F0FULL;
        f[0] = (int)(uint32_t)lo;
        f[1] = (int)(uint32_t)(lo >> 32);
        f[2] = (int)(uint32_t)hi;
        f[3] = (int)(uint32_t)(hi >> 32);
        (void)bm;
    } else {
        const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
        uint32_t u[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int sh = 64 - (j + 1) * B;
            if (sh < 0) u[j] = 0;
            else if (sh >= 32) u[j] = __builtin_amdgcn_ubfe(xh, (uint32_t)(sh - 32), (uint32_t)B);
            else if (sh + B <= 32) u[j] = __builtin_amdgcn_ubfe(xl, (uint32_t)sh, (uint32_t)B);
            else u[j] = (uint32_t)(x >> sh) & ((1u << B) - 1);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            f[q] = (int)((u[4 * q] | (u[4 * q + 1] << 8) | (u[4 * q + 2] << 16) | (u[4 * q + 3] << 24)) & bm[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) usum = __builtin_amdgcn_udot4((uint32_t)f[q], 0x01010101u, usum, false);
    return f;
}

template <int B>
__global__ void __launch_bounds__(64 * kKmWaves) __attribute__((amdgpu_waves_per_eu(2)))
lwe_keyswitch_mfma_kernel(KsArgs A) {
    __shared__ v4i kst[2][kKmChunk * kKmStep];  // 2 x 32 KiB
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, half = lane >> 5;
    const int g = blockIdx.y;  // column group: columns 32 g .. 32 g + 31
    const int m0 = blockIdx.x * kKmRows + 32 * w;
    // this lane's digit row (ciphertext), padded rows read row 0 and are not written
    const int r = m0 + (lane & 31);
    const int ct = r < A.count ? (A.rows ? (int)A.rows[r] : r) : 0;
    const uint64_t* in = A.in + (size_t)ct * A.in_stride;
    const int steps = A.n_in / 2;
    const int nch = steps / kKmChunk;
    const v4i* kg = reinterpret_cast<const v4i*>(A.kl) + (size_t)g * steps * kKmStep;
    uint32_t bm[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t m = 0;
#pragma unroll
        for (int y = 0; y < 4; ++y)
            if (4 * q + y < A.digits) m |= 0xFFu << (8 * y);
        bm[q] = m;
    }
    v4i stage[kKmStage];
    uint64_t wn[kKmChunk], wc[kKmChunk];
    // chunk q: key operands of k-steps 4 q .. 4 q + 3 (contiguous, 32 KiB) and this lane's input
    // words 2 s + half of those k-steps
    auto fetch = [&](int q) __attribute__((always_inline)) {
        const v4i* src = kg + (size_t)q * kKmChunk * kKmStep;
#pragma unroll
        for (int e = 0; e < kKmStage; ++e) stage[e] = src[e * 64 * kKmWaves + tid];
#pragma unroll
        for (int c = 0; c < kKmChunk; ++c) wn[c] = in[2 * (q * kKmChunk + c) + half];
    };
    auto put = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int e = 0; e < kKmStage; ++e) kst[buf][e * 64 * kKmWaves + tid] = stage[e];
    };
    v16i acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[t][e] = 0;
    uint32_t usum = 0;
    fetch(0);
    put(0);
#pragma unroll
    for (int c = 0; c < kKmChunk; ++c) wc[c] = wn[c];
    __syncthreads();
    for (int q = 0; q < nch; ++q) {
        if (q + 1 < nch) fetch(q + 1);
        const v4i* ks = kst[q & 1];
#pragma unroll
        for (int c = 0; c < kKmChunk; ++c) {
            const v4i a = word_digits<B>(gl::add(wc[c], A.corrector), A.digits, bm, usum);
#pragma unroll
            for (int t = 0; t < 8; ++t)
                acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, ks[(c * 8 + t) * 64 + lane], acc[t], 0, 0, 0);
        }
        // the other buffer was last read in chunk q - 1, before the barrier that ended it
        if (q + 1 < nch) {
            put((q + 1) & 1);
#pragma unroll
            for (int c = 0; c < kKmChunk; ++c) wc[c] = wn[c];
        }
        __syncthreads();
    }
    // U of row l % 32 (both halves' words), then the outputs: lane l holds column 32 g + l % 32 of
    // rows 8 (e / 4) + 4 half + e % 4 (the 32x32 accumulator layout), all eight bytes
    const uint32_t U = usum + __shfl_xor(usum, 32);
    const int col = 32 * g + (lane & 31);
    const int W = A.n_out + 1;
    const uint64_t corr = col < W ? A.corr[col] : 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int mr = 8 * (e / 4) + 4 * half + (e % 4);
        const uint32_t Ur = (uint32_t)__shfl((int)U, mr);
        const int orow = m0 + mr;
        unsigned __int128 T = 0;
#pragma unroll
        for (int t = 0; t < 8; ++t)
            T += (unsigned __int128)(uint32_t)(acc[t][e] + (int)(128u * Ur)) << (8 * t);
        if (orow >= A.count || col >= W) continue;
        const uint64_t v = gl::sub(gl::reduce128((uint64_t)(T >> 64), (uint64_t)T), corr);
        uint64_t* o = A.out + (size_t)orow * W;  // compact: output row orow reads input row rows[orow]
        if (col < A.n_out) {
            o[col] = gl::neg(v);
        } else {
            const int cto = A.rows ? (int)A.rows[orow] : orow;
            o[col] = gl::sub(A.in[(size_t)cto * A.in_stride + A.n_in], v);
        }
    }
}

}  // namespace

hipError_t launch_lwe_keyswitch_mfma(const KsArgs& A, hipStream_t s) {
    const dim3 grid((A.count + kKmRows - 1) / kKmRows, (A.n_out + 1 + 31) / 32);
    switch (A.base_log2) {
#define TP_KM_CASE(B)                                                                               \
    case B:                                                                                         \
        hipLaunchKernelGGL(lwe_keyswitch_mfma_kernel<B>, grid, dim3(64 * kKmWaves), 0, s, A);        \
        return hipGetLastError();
        TP_KM_CASE(1) TP_KM_CASE(2) TP_KM_CASE(3) TP_KM_CASE(4) TP_KM_CASE(5) TP_KM_CASE(6) TP_KM_CASE(7)
#undef TP_KM_CASE
    }
    return hipErrorInvalidValue;
}

}  // namespace tp
