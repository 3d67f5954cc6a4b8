// Binary CGGI blind rotation at N = 1024 with an exact floating-point external product.
//
// Same result as br_kernel<10, LD, 0, *> (cggi_blind_rotation.cpp:313-386 with the binary CMux of
// mux_operator.cpp:30-152), computed differently: each step's external product
//     c_h = sum over r = (h', j) of d_r * K_{r,h}                 (negacyclic, digits unsigned)
// is evaluated over the integers.  The key polynomials K (coefficients centred in (-Q/2, Q/2])
// are split into 4 balanced 16-bit limbs whose folded complex transforms are precomputed
// (fft_key.cpp); per limb the convolution sum_r d_r * K_r^(k) is one sum of pointwise products of
// double-precision transforms (fft_wave.hpp) and one inverse transform, rounded to the nearest
// integer.  The rounding error of every limb coefficient is below the bound of
// fftk::rounding_bound (< 1/2 for the parameter sets this kernel is enabled for), so every
// rounded limb is the exact integer, and sum_k 2^(16 k) c^(k) = c_h mod Q: the accumulator is
// bit-identical to the Goldilocks-NTT formulation.
//
// Layout: one work-group of two waves per ciphertext; wave h owns accumulator half h (16
// coefficients per lane, lane L holds L + 64 k), its digits and the four limb accumulators of
// output half h.  The digit transforms are exchanged through LDS: wave h multiplies its own
// transform of digit (h, j) and the other wave's transform of digit (1 - h, j) by the key limbs of
// output half h.
#include <hip/hip_runtime.h>

#include "br_kernel.hpp"
#include "fft_limbs.hpp"
#include "fft_wave.hpp"
#include "kernels.hpp"

namespace tp {

namespace {

__device__ __forceinline__ void cmac(cf::C& acc, const cf::C& z, const cf::C& k) {
    acc.re = fma(z.re, k.re, acc.re);
    acc.re = fma(-z.im, k.im, acc.re);
    acc.im = fma(z.re, k.im, acc.im);
    acc.im = fma(z.im, k.re, acc.im);
}

// Rounded limb value x / 512 as the bit pattern of x / 512 + 1.5 2^52 (|x / 512| < 2^51: the
// pattern is 0x4338000000000000 + the nearest integer, two's complement).
__device__ __forceinline__ uint64_t round_bits(double x) {
    return (uint64_t)__double_as_longlong(fma(x, 1.0 / 512.0, 6755399441055744.0));
}

}  // namespace

template <int LD, int WPE>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(WPE)))
br_fft_kernel(BrArgs A, const cf::C* __restrict__ fkey, const cf::C* __restrict__ zt) {
    constexpr int LOGN = 10;
    using S = Shape<LOGN>;
    constexpr int N = S::N;
    constexpr size_t kStep = (size_t)2 * LD * 2 * 4 * cf::kN;  // complex values per step
    __shared__ uint64_t srcbuf[2][S::LDS_WORDS];
    __shared__ cf::C xbuf[2][cf::kSlots];
    const int h = threadIdx.x >> 6;
    const int L = threadIdx.x & 63;
    uint64_t* src = srcbuf[h];
    cf::C* xb = xbuf[h];
    const cf::C* xo = xbuf[h ^ 1];
    const uint64_t mask = (1ULL << A.base_log2) - 1;
    const cf::Tw tw = cf::load_tw(zt, L);
    for (int ct = blockIdx.x; ct < A.count; ct += gridDim.x) {
        const int slot = A.active ? (int)A.active[ct] : ct;
        const uint64_t* lwe = A.lwe + (size_t)slot * (A.n + 1);
        const uint64_t* tv = A.tv + (A.tv_index ? (size_t)A.tv_index[slot] * 2 * N : (size_t)slot * A.tv_stride);
        const int total = A.n;
        const int i0 = A.slot_step ? (int)A.slot_step[slot] : 0;
        const int i1 = (A.chunk_steps > 0 && i0 + A.chunk_steps < total) ? i0 + A.chunk_steps : total;
        uint64_t acc[kE];
        if (i0 == 0) {
            rotate_from<LOGN>(acc, tv + h * N, (uint32_t)lwe[A.n], L, false);
        } else {
            const uint64_t* st = A.state + (size_t)slot * 2 * N + h * N;
#pragma unroll
            for (int k = 0; k < kE; ++k) acc[k] = st[idx1<LOGN>(L, k)];
        }

        for (int i = i0; i < i1; ++i) {
            const uint32_t ai = (uint32_t)lwe[i];
            if (ai == 0) continue;
            const uint32_t e0 = 2 * N - ai;
            // X^e0 acc_h - acc_h, parked in LDS (every digit re-reads it)
            poly_sync<1>();
#pragma unroll
            for (int k = 0; k < kE; ++k) src[phys(idx1<LOGN>(L, k))] = acc[k];
            poly_sync<1>();
            {
                uint64_t t[kE];
                int Lh = L;
                asm volatile("" : "+v"(Lh));
                rotate_from<LOGN>(t, src, e0, Lh, true);
#pragma unroll
                for (int k = 0; k < kE; ++k) t[k] = gl::sub(t[k], acc[k]);
                poly_sync<1>();
#pragma unroll
                for (int k = 0; k < kE; ++k) src[phys(idx1<LOGN>(L, k))] = t[k];
                poly_sync<1>();
            }
            cf::C F[4][cf::kPer];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int t = 0; t < cf::kPer; ++t) F[k][t] = cf::C{0.0, 0.0};
            const cf::C* ks = fkey + (size_t)i * kStep;

#pragma unroll 1
            for (int j = 0; j < LD; ++j) {
                uint32_t d[kE];
                {
                    const int shift = A.first_shift - j * A.base_log2;
                    uint32_t Lj = (uint32_t)L;
                    asm volatile("" : "+v"(Lj));
                    const uint32_t base = Lj + (Lj >> 4);
                    constexpr uint32_t kStride = S::P + S::P / 16;
                    const uint32_t* src32 = (const uint32_t*)src;
                    if (shift >= 32) {
#pragma unroll
                        for (int k = 0; k < kE; ++k)
                            d[k] = __builtin_amdgcn_ubfe(src32[2 * (base + k * kStride) + 1], (uint32_t)(shift - 32),
                                                         (uint32_t)A.base_log2);
                    } else if (shift + A.base_log2 <= 32) {
#pragma unroll
                        for (int k = 0; k < kE; ++k)
                            d[k] = __builtin_amdgcn_ubfe(src32[2 * (base + k * kStride)], (uint32_t)shift,
                                                         (uint32_t)A.base_log2);
                    } else {
#pragma unroll
                        for (int k = 0; k < kE; ++k) d[k] = (uint32_t)((src[base + k * kStride] >> shift) & mask);
                    }
                }
                cf::C z[cf::kPer];
#pragma unroll
                for (int t = 0; t < cf::kPer; ++t) z[t] = cf::C{(double)d[t], (double)d[t + 8]};
                // key rows of this digit: own transform x limbs of half h, the other wave's likewise;
                // group t (8 values per lane) is requested while group t - 1 is used, group 0
                // during the transform's last level
                const cf::C* ko = ks + (size_t)(((h * LD + j) * 2 + h) * 4) * cf::kN;
                const cf::C* kq = ks + (size_t)((((h ^ 1) * LD + j) * 2 + h) * 4) * cf::kN;
                cf::C ka[4], kb[4];
                auto load_group = [&](cf::C (&ga)[4], cf::C (&gb)[4], int t) __attribute__((always_inline)) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        ga[k] = cf::ldc(ko, (uint32_t)(k * cf::kN + t * 64 + L));
                        gb[k] = cf::ldc(kq, (uint32_t)(k * cf::kN + t * 64 + L));
                    }
                    __builtin_amdgcn_sched_barrier(0);
                };
                cf::forward(z, xb, tw, L, [&]() __attribute__((always_inline)) { load_group(ka, kb, 0); });
                // hand the transform to the other wave
                poly_sync<1>();
#pragma unroll
                for (int t = 0; t < cf::kPer; ++t) xb[cf::phys(cf::pos3(L, t))] = z[t];
                __syncthreads();
                static_for<cf::kPer>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = T;
                    cf::C na[4], nb[4];
                    if constexpr (t + 1 < cf::kPer) load_group(na, nb, t + 1);
                    const cf::C zq = xo[cf::phys(cf::pos3(L, t))];
#pragma unroll
                    for (int k = 0; k < 4; ++k) cmac(F[k][t], z[t], ka[k]);
#pragma unroll
                    for (int k = 0; k < 4; ++k) cmac(F[k][t], zq, kb[k]);
                    if constexpr (t + 1 < cf::kPer) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            ka[k] = na[k];
                            kb[k] = nb[k];
                        }
                    }
                });
                __syncthreads();
            }

            // limb convolutions back in pass-1 layout: limb k of coefficient L + 64 m is
            // F[k][m].re (m < 8) or F[k][m - 8].im, times 512
#pragma unroll
            for (int k = 0; k < 4; ++k) cf::inverse(F[k], xb, tw, L);
            const uint64_t qm1 = glv::qm1_const();
#pragma unroll
            for (int m = 0; m < kE; ++m) {
                uint64_t bits[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) bits[k] = round_bits(m < 8 ? F[k][m].re : F[k][m - 8].im);
                acc[m] = fftl::combine(acc[m], bits, qm1);
            }
        }

        if (A.slot_step && threadIdx.x == 0) A.slot_step[slot] = (uint32_t)i1;
        if (i1 < total) {
            uint64_t* st = A.state + (size_t)slot * 2 * N + h * N;
#pragma unroll
            for (int k = 0; k < kE; ++k) st[idx1<LOGN>(L, k)] = acc[k];
            __syncthreads();
            continue;
        }
        if (A.out_lwe) {
            uint64_t* o = A.out_lwe + (size_t)slot * (N + 1);
            if (h == 0) {
#pragma unroll
                for (int k = 0; k < kE; ++k) {
                    const int j = idx1<LOGN>(L, k);
                    o[j == 0 ? 0 : N - j] = j == 0 ? acc[k] : gl::neg(acc[k]);
                }
            } else if (L == 0) {
                o[N] = acc[0];
            }
        }
        if (A.out_rlwe) {
            uint64_t* o = A.out_rlwe + (size_t)slot * 2 * N + h * N;
            if (A.out_rlwe_ntt) {
                poly_sync<1>();
                ntt_fwd<LOGN>(acc, src, A.tw, L);
#pragma unroll
                for (int k = 0; k < kE; ++k) o[idx3(L, k)] = gl::canon(acc[k]);
            } else {
#pragma unroll
                for (int k = 0; k < kE; ++k) o[idx1<LOGN>(L, k)] = acc[k];
            }
        }
        __syncthreads();
    }
}

hipError_t launch_blind_rotate_fft(int digits, const BrArgs& A, const void* fkey, const void* zt, int grid, int waves,
                                   hipStream_t s) {
    const cf::C* k = (const cf::C*)fkey;
    const cf::C* z = (const cf::C*)zt;
#define TP_FFT_CASE(LD, W)                                                                          \
    if (digits == LD && waves == W) {                                                               \
        hipLaunchKernelGGL((br_fft_kernel<LD, W>), dim3(grid), dim3(128), 0, s, A, k, z);           \
        return hipGetLastError();                                                                   \
    }
    TP_FFT_CASE(2, 2) TP_FFT_CASE(3, 2) TP_FFT_CASE(4, 2)
    TP_FFT_CASE(2, 1) TP_FFT_CASE(3, 1) TP_FFT_CASE(4, 1)
#undef TP_FFT_CASE
    return hipErrorInvalidValue;
}

}  // namespace tp
