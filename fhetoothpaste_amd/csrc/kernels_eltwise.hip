// Elementwise mod-Q arithmetic over device arrays: MathWorker::{Add,Sub,Mul,FMA,MulS,Reduce}ModElw
// (lib/include/backend/backend.h:35-55; SolinasWorker, lib/src/backend/solinas.cpp:346-402).
// HBM-bound grid-stride loop, one word per lane per iteration; inputs canonical, outputs canonical.
#include <hip/hip_runtime.h>
#include "gl64.hpp"
#include "kernels.hpp"
#include "modswitch.hpp"
#include "../../include/toothpaste_amd.h"

namespace tp {

namespace {

template <int OP>
__device__ __forceinline__ uint64_t elt(uint64_t a, uint64_t b, uint64_t c, uint64_t scalar) {
    if constexpr (OP == TP_ELT_ADD) return gl::add(a, b);
    else if constexpr (OP == TP_ELT_SUB) return gl::sub(a, b);
    else if constexpr (OP == TP_ELT_MUL) return gl::mul(a, b);
    else if constexpr (OP == TP_ELT_FMA) return gl::add(gl::mul(a, b), c);
    else if constexpr (OP == TP_ELT_MULS) return gl::mul(a, scalar);
    else return a >= gl::kQ ? a - gl::kQ : a;  // TP_ELT_REDUCE
}

template <int OP>
__global__ void __launch_bounds__(256) eltwise_kernel(uint64_t* dst, const uint64_t* a, const uint64_t* b,
                                                      const uint64_t* c, uint64_t scalar, size_t words) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        const uint64_t bv = (OP == TP_ELT_ADD || OP == TP_ELT_SUB || OP == TP_ELT_MUL || OP == TP_ELT_FMA) ? b[i] : 0;
        const uint64_t cv = OP == TP_ELT_FMA ? c[i] : 0;
        dst[i] = elt<OP>(a[i], bv, cv, scalar);
    }
}

// ModulusSwitch (modulus_switching.cpp:36-93) of `words` values: canonical inputs (< sM; larger
// ones are reduced by one subtraction first), outputs in [0, 2^t).  kind 3 (RANDOM) picks floor or
// ceil by a coin from a counter-based stream (seed, index), where the reference draws an unseeded one.
__global__ void __launch_bounds__(256) modswitch_kernel(uint64_t* dst, const uint64_t* src, size_t words, uint64_t sM,
                                                        uint64_t mu, int t, int kind, uint64_t seed) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        uint64_t v = src[i];
        if (v >= sM) v -= sM;
        int k = kind;
        if (kind == 3) {
            uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            k = ((z ^ (z >> 31)) >> 63) ? kMsFloor : kMsCeil;
        }
        dst[i] = ms_switch(v, sM, mu, t, k);
    }
}

}  // namespace

hipError_t launch_modswitch(uint64_t* dst, const uint64_t* src, size_t words, uint64_t sM, int t, int kind,
                            uint64_t seed, hipStream_t s) {
    if (words == 0) return hipSuccess;
    size_t groups = (words + 255) / 256;
    if (groups > 8192) groups = 8192;
    hipLaunchKernelGGL(modswitch_kernel, dim3((unsigned)groups), dim3(256), 0, s, dst, src, words, sM,
                       ms_reciprocal(sM), t, kind, seed);
    return hipGetLastError();
}

hipError_t launch_eltwise(int op, uint64_t* dst, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                          uint64_t scalar, size_t words, hipStream_t s) {
    if (words == 0) return hipSuccess;
    // enough work-groups to fill 256 CUs several times over, then grid-stride
    size_t groups = (words + 255) / 256;
    if (groups > 8192) groups = 8192;
    const dim3 g((unsigned)groups), blk(256);
    switch (op) {
        case TP_ELT_ADD: hipLaunchKernelGGL(eltwise_kernel<TP_ELT_ADD>, g, blk, 0, s, dst, a, b, c, scalar, words); break;
        case TP_ELT_SUB: hipLaunchKernelGGL(eltwise_kernel<TP_ELT_SUB>, g, blk, 0, s, dst, a, b, c, scalar, words); break;
        case TP_ELT_MUL: hipLaunchKernelGGL(eltwise_kernel<TP_ELT_MUL>, g, blk, 0, s, dst, a, b, c, scalar, words); break;
        case TP_ELT_FMA: hipLaunchKernelGGL(eltwise_kernel<TP_ELT_FMA>, g, blk, 0, s, dst, a, b, c, scalar, words); break;
        case TP_ELT_MULS: hipLaunchKernelGGL(eltwise_kernel<TP_ELT_MULS>, g, blk, 0, s, dst, a, b, c, scalar, words); break;
        case TP_ELT_REDUCE: hipLaunchKernelGGL(eltwise_kernel<TP_ELT_REDUCE>, g, blk, 0, s, dst, a, b, c, scalar, words); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tp
