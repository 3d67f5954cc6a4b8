// Kernels of the SAP digit decomposition (lib/src/operators/sap_decomposer.cpp:190-494) for a
// batch of inputs: the per-digit LWE bookkeeping, the NTT-domain products with fixed polynomials
// of HomTrunc, the sum of the tree's leaf contributions, and sample extraction into the digit
// outputs.  The blind rotations, NTTs and the tree's automorphism key switches (sap_tree_kernel)
// are in the library's other kernel files.  All HBM-bound
// grid-stride loops over canonical words; rows of N words share one polynomial (both halves of
// an RLWE [a|b] are multiplied by the same one, as the reference does at :243-246, :293-298).
#include <hip/hip_runtime.h>
#include "gl64.hpp"
#include "kernels.hpp"

namespace tp {

namespace {

constexpr int kBlock = 256;

unsigned grid_for(size_t words) {
    size_t g = (words + kBlock - 1) / kBlock;
    return (unsigned)(g > 8192 ? 8192 : (g ? g : 1));
}

// IntLog2 of the reference (utils/math_utils.cpp:11-25): ceil(log2 x), 0 for x <= 1.
__device__ __forceinline__ uint32_t ilog2_ceil(uint64_t x) { return x <= 1 ? 0 : 64 - __clzll(x - 1); }

// Per input: the modulus read off the values (sap_decomposer.cpp:427-433) and the digit count
// (iterations of :443 while the modulus exceeds 1).  1 << 64 is taken as 1, as the x86 shift the
// reference compiles to does (the count is used mod 64).
__global__ void __launch_bounds__(kBlock) sap_init_kernel(const uint64_t* __restrict__ lwes, uint64_t* __restrict__ cm,
                                                          uint32_t* __restrict__ digits, int n, int count, int rl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t* x = lwes + (size_t)i * (n + 1);
    uint64_t m = 1ULL << (ilog2_ceil(x[0]) & 63);
    for (int j = 1; j <= n; ++j)
        if (m <= x[j]) m <<= 1;
    cm[i] = m;
    uint32_t d = 0;
    while (m > 1) {
        m >>= rl;
        ++d;
    }
    digits[i] = d;
}

// One digit step over all inputs (sap_decomposer.cpp:444-455): sub = x mod radix, x >>= log radix,
// the modulus shrinks, and the perturbation adds radix * beta to the digit's b (mod 2N) and takes
// beta off the remaining b (mod the new modulus, when it exceeds 1).
__global__ void __launch_bounds__(kBlock) sap_digit_kernel(uint64_t* __restrict__ cp, uint64_t* __restrict__ cm,
                                                           uint64_t* __restrict__ sub, int n, size_t total, int rl,
                                                           uint64_t mask, uint64_t plo, uint64_t phi, uint64_t two_n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t row = (size_t)n + 1;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        uint64_t v = cp[idx];
        uint64_t s = v & mask;
        v >>= rl;
        if (idx % row == (size_t)n) {
            const size_t i = idx / row;
            const uint64_t m = cm[i] >> rl;
            cm[i] = m;
            s = (s + plo) % two_n;
            if (m > 1) {
                const uint64_t mm = m - 1;
                v = ((v & mm) + m - (phi & mm)) & mm;
            }
        }
        sub[idx] = s;
        cp[idx] = v;
    }
}

// dst[r][j] = src[r][j] * poly[j]  (rows of N words)
__global__ void __launch_bounds__(kBlock) sap_poly_mul_kernel(uint64_t* dst, const uint64_t* src,
                                                              const uint64_t* __restrict__ poly, size_t words,
                                                              uint32_t nmask) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
        dst[i] = gl::mul(src[i], poly[i & nmask]);
}

// Leaf of the tree (sap_decomposer.cpp:266-277): out (+)= x * monomial.
__global__ void __launch_bounds__(kBlock) sap_leaf_kernel(uint64_t* __restrict__ out, const uint64_t* __restrict__ x,
                                                          const uint64_t* __restrict__ mono, size_t words,
                                                          uint32_t nmask, int accumulate) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        const uint64_t p = gl::mul(x[i], mono[i & nmask]);
        out[i] = accumulate ? gl::add(out[i], p) : p;
    }
}

// Sum over the level-1 nodes of the level-synchronous tree: out[w] (+)= sum_e P[e][w], w < words
// (the per-node leaf contributions of every input, rows of 2N words).  Eight node rows are requested
// before they are added, so a thread keeps eight loads in flight instead of waiting for each.
__global__ void __launch_bounds__(kBlock) sap_sum_nodes_kernel(uint64_t* __restrict__ out,
                                                               const uint64_t* __restrict__ P, int nodes,
                                                               size_t words, int accumulate) {
    constexpr int U = 8;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        uint64_t acc = accumulate ? out[i] : 0;
        int e = 0;
        for (; e + U <= nodes; e += U) {
            uint64_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = P[(size_t)(e + u) * words + i];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = gl::add(acc, v[u]);
        }
        for (; e < nodes; ++e) acc = gl::add(acc, P[(size_t)e * words + i]);
        out[i] = acc;
    }
}

// SampleExtract at index 0 (static/sample_extraction.h:20-36) of RLWE row i ([a|b] coefficients)
// into out + i * out_stride + digit * (N + 1), for the rows with digit < digits[i] (all rows when
// digits is null).  Every word is written: -0 = 0 where the reference skips the write.
__global__ void __launch_bounds__(kBlock) sap_extract_kernel(uint64_t* __restrict__ out, const uint64_t* __restrict__ rlwe,
                                                             const uint32_t* __restrict__ digits, uint32_t digit,
                                                             size_t out_stride, int count, uint32_t N) {
    const size_t total = (size_t)count * (N + 1);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const size_t i = idx / (N + 1);
        const uint32_t j = (uint32_t)(idx - i * (N + 1));
        if (digits && digit >= digits[i]) continue;
        const uint64_t* r = rlwe + i * 2 * (size_t)N;
        uint64_t v;
        if (j == 0) v = r[0];
        else if (j < N) v = gl::neg(r[N - j]);
        else v = r[N];
        out[i * out_stride + (size_t)digit * (N + 1) + j] = v;
    }
}

}  // namespace

hipError_t launch_sap_init(const uint64_t* lwes, uint64_t* cm, uint32_t* digits, int n, int count, int rl,
                           hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(sap_init_kernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, lwes, cm, digits, n,
                       count, rl);
    return hipGetLastError();
}

hipError_t launch_sap_digit(uint64_t* cp, uint64_t* cm, uint64_t* sub, int n, int count, int rl, uint64_t mask,
                            uint64_t plo, uint64_t phi, uint64_t two_n, hipStream_t s) {
    const size_t total = (size_t)count * (n + 1);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(sap_digit_kernel, dim3(grid_for(total)), dim3(kBlock), 0, s, cp, cm, sub, n, total, rl, mask,
                       plo, phi, two_n);
    return hipGetLastError();
}

hipError_t launch_sap_poly_mul(uint64_t* dst, const uint64_t* src, const uint64_t* poly, size_t rows, uint32_t N,
                               hipStream_t s) {
    const size_t words = rows * N;
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(sap_poly_mul_kernel, dim3(grid_for(words)), dim3(kBlock), 0, s, dst, src, poly, words, N - 1);
    return hipGetLastError();
}

hipError_t launch_sap_leaf(uint64_t* out, const uint64_t* x, const uint64_t* mono, size_t rows, uint32_t N,
                           int accumulate, hipStream_t s) {
    const size_t words = rows * N;
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(sap_leaf_kernel, dim3(grid_for(words)), dim3(kBlock), 0, s, out, x, mono, words, N - 1,
                       accumulate);
    return hipGetLastError();
}

hipError_t launch_sap_sum_nodes(uint64_t* out, const uint64_t* P, int nodes, size_t rows, uint32_t N, int accumulate,
                                hipStream_t s) {
    const size_t words = rows * N;
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(sap_sum_nodes_kernel, dim3(grid_for(words)), dim3(kBlock), 0, s, out, P, nodes, words,
                       accumulate);
    return hipGetLastError();
}

hipError_t launch_sap_extract(uint64_t* out, const uint64_t* rlwe, const uint32_t* digits, uint32_t digit,
                              size_t out_stride, int count, uint32_t N, hipStream_t s) {
    const size_t total = (size_t)count * (N + 1);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(sap_extract_kernel, dim3(grid_for(total)), dim3(kBlock), 0, s, out, rlwe, digits, digit,
                       out_stride, count, N);
    return hipGetLastError();
}

}  // namespace tp
