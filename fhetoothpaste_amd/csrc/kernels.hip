// HIP kernels for gfx950: batched NTT, the blind-rotation launcher (two-wave variants here,
// one-wave variants in kernels_br1.hip), batched LWE->LWE key switching.
#include <hip/hip_runtime.h>
#include "gl64.hpp"
#include "ntt_wave.hpp"
#include "kernels.hpp"
#include "br_kernel.hpp"

namespace tp {

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
            ntt_fwd<LOGN>(x, lds[g], tw, L);
            transpose<LOGN, Layout::kPass3, Layout::kPass1>(x, lds[g], L);  // coalesced writes
            if (live)
#pragma unroll
                for (int k = 0; k < kE; ++k) out[p * S::N + idx1<LOGN>(L, k)] = gl::canon(x[k]);
        } else {
#pragma unroll
            for (int k = 0; k < kE; ++k) x[k] = src[idx3(L, k)];
            ntt_inv<LOGN>(x, lds[g], tw, L);
            if (live)
#pragma unroll
                for (int k = 0; k < kE; ++k) out[p * S::N + idx1<LOGN>(L, k)] = gl::canon(x[k]);
        }
    }
}

// ------------------------------------------------------------------ LWE key switch

// out[c][k] = -(sum_ij d_ij(c) KSK[i][j][k]) (k < n_out), out[c][n_out] = in[c][n_in] - (...)
// (lwe_conversion.cpp:205-239).  With unsigned digits u_ij = d_ij + delta_j the sum is
// sum u_ij K - corr[k], corr[k] = sum_ij delta_j K[i][j][k] precomputed per key.  u_ij < 2^b, so
// u K_lo and u K_hi are accumulated exactly in 64 bits and reduced once at the end.
// GEMM layout: one ciphertext per lane (its digits are per-lane work), a tile of COLS key
// columns per block; the key row segment K[i][j][col0 .. col0 + COLS) is the same for the whole
// block, is staged through LDS and read as a broadcast, so each MAC is two v_mad_u64_u32.
// Tile width: 4 columns while the grid holds at most five waves per SIMD (more waves to hide the
// row loop's latency: 2048 ciphertexts at N = 1024, d = 14: 1.89 ms at 4, 2.03 ms at 8), else 8
// (fewer digit extractions per MAC: 4096 ciphertexts 3.18 ms at 8, 3.45 ms at 4; 16 columns was
// slower at every size measured, 3.88 ms at 4096).
constexpr int kKsColsMax = 16;
static_assert(kKsColsMax <= kKsPadWords, "a partial last tile reads into the zero padding");
// tile width of the row-split kernel below (its threads walk rows, not ciphertexts)
constexpr int kKsSplitCols = 16;
static_assert(kKsSplitCols <= kKsPadWords, "a partial last tile reads into the zero padding");
constexpr int kKsBlock = 256;
// Launches of at most this many ciphertexts use the row-split kernel below: the batched kernel's
// time is one lane walking all N d key rows, the same for 1 and for 256 ciphertexts (about 1.25 ms
// at N = 1024, d = 14, on 8 columns), while the row-split kernel reads the key once per ciphertext
// (about 60 us; 0.62 ms for 64, 1.51 ms for 128 against 0.80 ms on 4 columns).
constexpr int kKsRowSplitMax = 64;
constexpr int kKsMfmaMin = 16;

template <int COLS>
__global__ void __launch_bounds__(kKsBlock) lwe_keyswitch_kernel(KsArgs A) {
    static_assert(COLS <= kKsColsMax, "a partial last tile reads into the zero padding");
    const int col0 = blockIdx.x * COLS;
    const int c = blockIdx.y * kKsBlock + threadIdx.x;
    const int W = A.n_out + 1;
    const bool live = c < A.count;
    const int ct = live ? (A.rows ? (int)A.rows[c] : c) : 0;
    const uint64_t* in = A.in + (size_t)ct * A.in_stride;
    uint64_t lo[COLS], hi[COLS];
#pragma unroll
    for (int t = 0; t < COLS; ++t) lo[t] = hi[t] = 0;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    // The key rows i d + j (a segment of COLS words each) are staged through LDS, a chunk of WC input
    // words (WC d <= CH rows) at a time, loaded by the whole work-group; the next chunk's key rows and
    // input words are requested while the current chunk is used, and each wave reads a row as a
    // broadcast LDS read.  (Read directly, every row waited for its scalar loads.)  The device key
    // carries kKsPadWords words of zero padding, so a partial last tile reads in bounds; its extra
    // columns are discarded.
    constexpr int CH = 64;
    constexpr int XW = 4;                      // input words per chunk at most
    constexpr int WPT = CH * COLS / kKsBlock;  // key words each thread stages per chunk
    static_assert(WPT * kKsBlock == CH * COLS, "chunk divides evenly over the work-group");
    __shared__ uint64_t kst[2][CH * COLS];
    const uint64_t* krow = A.ksk + col0;
    const int d = A.digits;
    const int rows = A.n_in * d;
    // input words per chunk; contexts accept d <= ceil(64 / b) <= 64 = CH digits, so wc d <= CH
    const int wc = d * XW <= CH ? XW : (CH / d > 0 ? CH / d : 1);
    const int nch = (A.n_in + wc - 1) / wc;
    const int tid = threadIdx.x;
    uint64_t stage[WPT], xn[XW];
    auto fetch = [&](int q) __attribute__((always_inline)) {
#pragma unroll
        for (int w = 0; w < WPT; ++w) {
            const int idx = tid * WPT + w;
            const int r = q * wc * d + idx / COLS;  // past the last row: the last row again (unused)
            stage[w] = krow[(size_t)(r < rows ? r : rows - 1) * W + idx % COLS];
        }
#pragma unroll
        for (int w = 0; w < XW; ++w) {
            const int iw = q * wc + w;
            xn[w] = in[iw < A.n_in ? iw : A.n_in - 1];
        }
    };
    auto put = [&](int b) __attribute__((always_inline)) {
#pragma unroll
        for (int w = 0; w < WPT; ++w) kst[b][tid * WPT + w] = stage[w];
    };
    fetch(0);
    put(0);
    uint64_t xs[XW];
#pragma unroll
    for (int w = 0; w < XW; ++w) xs[w] = xn[w];
    __syncthreads();
    for (int q = 0; q < nch; ++q) {
        if (q + 1 < nch) fetch(q + 1);
        const uint64_t* ks = kst[q & 1];
#pragma unroll
        for (int w = 0; w < XW; ++w) {
            if (w >= wc || q * wc + w >= A.n_in) break;
            const uint64_t x = gl::add(xs[w], A.corrector);
            const uint64_t* kw = ks + w * d * COLS;
            for (int j = 0; j < d; ++j) {
                const uint32_t u = (uint32_t)((x >> (64 - (j + 1) * A.base_log2)) & mask);
#pragma unroll
                for (int t = 0; t < COLS; ++t) {
                    const uint64_t kv = kw[j * COLS + t];
                    lo[t] += (uint64_t)u * (uint32_t)kv;
                    hi[t] += (uint64_t)u * (uint32_t)(kv >> 32);
                }
            }
        }
        // the other buffer was last read in chunk q - 1, before the barrier that ended it
        if (q + 1 < nch) {
            put((q + 1) & 1);
#pragma unroll
            for (int w = 0; w < XW; ++w) xs[w] = xn[w];
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int t = 0; t < COLS; ++t) {
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

// Few ciphertexts: the same sums, split over the key rows instead of over ciphertexts.  Block
// (x, y) covers one tile of kKsSplitCols columns of ciphertext y; thread t takes rows t, t + kKsBlock,
// ... of the N d key rows (each a 16-word, 128-byte segment), and the partial lo / hi sums, exact
// in 64 bits like the totals they add up to, meet in LDS.  Each ciphertext reads the key once
// (the ciphertexts of a launch share it through the caches), instead of one lane per ciphertext
// walking all N d rows.
__global__ void __launch_bounds__(kKsBlock) lwe_keyswitch_one_kernel(KsArgs A) {
    __shared__ uint64_t part[2][kKsSplitCols][kKsBlock + 1];
    const int col0 = blockIdx.x * kKsSplitCols;
    const int t = threadIdx.x;
    const int W = A.n_out + 1;
    const int ct = A.rows ? (int)A.rows[blockIdx.y] : (int)blockIdx.y;
    const uint64_t* in = A.in + (size_t)ct * A.in_stride;
    uint64_t* out = A.out + (size_t)blockIdx.y * W;
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    uint64_t lo[kKsSplitCols], hi[kKsSplitCols];
#pragma unroll
    for (int c = 0; c < kKsSplitCols; ++c) lo[c] = hi[c] = 0;
    const int rows = A.n_in * A.digits;
    for (int r = t; r < rows; r += kKsBlock) {
        const int i = r / A.digits, j = r - i * A.digits;
        const uint64_t x = gl::add(in[i], A.corrector);
        const uint32_t u = (uint32_t)((x >> (64 - (j + 1) * A.base_log2)) & mask);
        const uint64_t* kr = A.ksk + (size_t)r * W + col0;
#pragma unroll
        for (int c = 0; c < kKsSplitCols; ++c) {
            const uint64_t kv = kr[c];
            lo[c] += (uint64_t)u * (uint32_t)kv;
            hi[c] += (uint64_t)u * (uint32_t)(kv >> 32);
        }
    }
#pragma unroll
    for (int c = 0; c < kKsSplitCols; ++c) {
        part[0][c][t] = lo[c];
        part[1][c][t] = hi[c];
    }
    __syncthreads();
    // thread (h, c) for h < 2, c < kKsSplitCols adds column c's partial sums of half h
    if (t < 2 * kKsSplitCols) {
        const int h = t / kKsSplitCols, c = t % kKsSplitCols;
        uint64_t sum = 0;
        for (int q = 0; q < kKsBlock; ++q) sum += part[h][c][q];
        part[h][c][kKsBlock] = sum;
    }
    __syncthreads();
    if (t < kKsSplitCols) {
        const int col = col0 + t;
        if (col >= W) return;
        const uint64_t l = part[0][t][kKsBlock], h = part[1][t][kKsBlock];
        const uint64_t t_lo = (h << 32) + l;
        const uint64_t carry = t_lo < l;
        const uint64_t t_hi = (h >> 32) + carry;
        const uint64_t acc = gl::sub(gl::reduce128(t_hi, t_lo), A.corr[col]);
        if (col < A.n_out) out[col] = gl::neg(acc);
        else out[col] = gl::sub(in[A.n_in], acc);
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

hipError_t launch_blind_rotate(int logN, int method, int digits, const BrArgs& A, int grid, int waves,
                               hipStream_t s) {
    if (waves == 1) return launch_blind_rotate_w1(logN, method, digits, A, grid, s);
    if (waves != 2) return hipErrorInvalidValue;
#define TP_BR_CASE(LG, LD, M) \
    if (logN == LG && digits == LD && method == M) { launch_br<LG, LD, M, 2>(A, grid, s); return hipGetLastError(); }
#define TP_BR_METHODS(LG, LD) TP_BR_CASE(LG, LD, 0) TP_BR_CASE(LG, LD, 1) TP_BR_CASE(LG, LD, 2)
    TP_BR_METHODS(10, 2) TP_BR_METHODS(10, 3) TP_BR_METHODS(10, 4)
    TP_BR_METHODS(11, 2) TP_BR_METHODS(11, 3) TP_BR_METHODS(11, 4)
    TP_BR_METHODS(12, 3) TP_BR_METHODS(12, 4)
#undef TP_BR_METHODS
#undef TP_BR_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_lwe_keyswitch(const KsArgs& A, hipStream_t s) {
    // the matrix-core kernel where the key is held in its operand form: its time is about 0.18 ms
    // up to a few thousand ciphertexts at 1024 -> 630, d = 14 (the k loop), the row-split kernel's
    // 0.06 ms + 9 us per ciphertext, so it takes launches of more than kKsMfmaMin
    if (A.kl && A.count > kKsMfmaMin) return launch_lwe_keyswitch_mfma(A, s);
    if (A.count <= kKsRowSplitMax) {
        hipLaunchKernelGGL(lwe_keyswitch_one_kernel, dim3((A.n_out + 1 + kKsSplitCols - 1) / kKsSplitCols, A.count),
                           dim3(kKsBlock), 0, s, A);
        return hipGetLastError();
    }
    // SIMDs of the current device (4 per CU); the runtime answers from its cached device properties
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int simds = 4 * cus;
    const int blocks_y = (A.count + kKsBlock - 1) / kKsBlock;
    const int tiles4 = (A.n_out + 1 + 3) / 4;
    const long long waves4 = (long long)blocks_y * tiles4 * (kKsBlock / 64);
    if (waves4 <= 5LL * simds) {
        hipLaunchKernelGGL(lwe_keyswitch_kernel<4>, dim3(tiles4, blocks_y), dim3(kKsBlock), 0, s, A);
    } else {
        hipLaunchKernelGGL(lwe_keyswitch_kernel<8>, dim3((A.n_out + 1 + 7) / 8, blocks_y), dim3(kKsBlock), 0, s, A);
    }
    return hipGetLastError();
}

}  // namespace tp
