// Elementwise steps of the LWE -> RGSW scheme switch (lib/src/operators/lwe_to_rgsw_conversion.cpp:
// 163-231) for a batch; the blind rotation, trace, inverse NTT and key switch reuse the library's
// kernels.  HBM-bound grid-stride loops over canonical words.
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

// rows[c][d] = X^(-d block) (row0[c] + (0, extra)), both halves, coefficient domain
// (:190-207: the rotated accumulator gains the constant t/2 on b, then each output digit reads the
// doubled buffer [A, -A] from offset d block).
__global__ void __launch_bounds__(kBlock) rgsw_expand_kernel(uint64_t* __restrict__ rows,
                                                             const uint64_t* __restrict__ row0,
                                                             const uint64_t* __restrict__ extra, int count,
                                                             uint32_t digits, uint32_t block, uint32_t N) {
    const size_t total = (size_t)count * digits * 2 * N;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const size_t r = idx / (2 * N);
        const uint32_t w = (uint32_t)(idx - r * 2 * N);
        const uint32_t h = w >= N, k = w - h * N;
        const size_t c = r / digits;
        const uint32_t d = (uint32_t)(r - c * digits);
        const uint32_t e = d * block + k;  // < 2N
        const uint32_t jj = e < N ? e : e - N;
        uint64_t v = row0[c * 2 * N + h * N + jj];
        if (h) v = gl::add(v, extra[jj]);
        rows[idx] = e < N ? v : gl::neg(v);
    }
}

// left = -ks, with the right row's b added to its a half (:221-229)
__global__ void __launch_bounds__(kBlock) rgsw_left_kernel(uint64_t* __restrict__ left,
                                                           const uint64_t* __restrict__ right, size_t rows,
                                                           uint32_t N) {
    const size_t total = rows * 2 * N;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const size_t r = idx / (2 * N);
        const uint32_t w = (uint32_t)(idx - r * 2 * N);
        uint64_t v = gl::neg(left[idx]);
        if (w < N) v = gl::add(v, right[r * 2 * N + N + w]);
        left[idx] = v;
    }
}

}  // namespace

hipError_t launch_rgsw_expand(uint64_t* rows, const uint64_t* row0, const uint64_t* extra, int count, uint32_t digits,
                              uint32_t block, uint32_t N, hipStream_t s) {
    const size_t total = (size_t)count * digits * 2 * N;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(rgsw_expand_kernel, dim3(grid_for(total)), dim3(kBlock), 0, s, rows, row0, extra, count,
                       digits, block, N);
    return hipGetLastError();
}

hipError_t launch_rgsw_left(uint64_t* left, const uint64_t* right, size_t rows, uint32_t N, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(rgsw_left_kernel, dim3(grid_for(rows * 2 * N)), dim3(kBlock), 0, s, left, right, rows, N);
    return hipGetLastError();
}

}  // namespace tp
