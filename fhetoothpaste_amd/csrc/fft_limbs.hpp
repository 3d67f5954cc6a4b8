// Recombination of the four rounded limb convolutions of the floating-point external product
// (kernels_br_fft.hip) into the accumulator, mod Q.  Written with the instruction layer of
// gl_valu.hpp, so the host build (tests/cpp/fft_check.cpp) runs the device's op sequence.
#pragma once
#include "gl_valu.hpp"

namespace tp {
namespace fftl {

// Bit patterns of x + 1.5 2^52 for the rounded limb values: 0x4338000000000000 + c_k, two's
// complement, as long as |c_k| < 2^51.
constexpr uint64_t kMagicBits = 0x4338000000000000ULL;
// K0 = -(2^57 + 2^89) mod Q, and the offsets that turn b0 + 2^16 b1 (resp. b2 + 2^16 b3) into
// A + 2^57 + K0 (resp. B + 2^57), A = c0 + 2^16 c1, B = c2 + 2^16 c3.
constexpr uint64_t kK0 = 0xFBFFFFFF02000001ULL;
constexpr uint64_t kOffA = (1ULL << 57) + kK0 - kMagicBits * 0x10001ULL;
constexpr uint64_t kOffB = (1ULL << 57) - kMagicBits * 0x10001ULL;

// acc + c0 + 2^16 c1 + 2^32 c2 + 2^48 c3 mod Q, canonical, for canonical acc and |c_k| below
// 2^57 / (1 + 2^16) (then |A|, |B| < 2^57: A' = A + 2^57 + K0 < K0 + 2^58 < 2^64 and
// B' = B + 2^57 < 2^58, and acc + A' + 2^32 B' = acc + A + 2^32 B + (2^57 + 2^89 + K0), the last
// term a multiple of Q since 2^89 = 2^57 - 2^25 mod Q).  The sum is formed as a 128-bit
// (lo, hi), hi < 2^26 + 2, and reduced once (13 VALU plus the four limb adds).
GV_FN uint64_t combine(uint64_t acc, const uint64_t (&bits)[4], uint64_t qm1) {
    const uint64_t Ap = bits[0] + (bits[1] << 16) + kOffA;
    const uint64_t Bp = bits[2] + (bits[3] << 16) + kOffB;
    glv::lmask c0, c1, c2;
    const uint32_t s0 = glv::add_co(glv::lo32(acc), glv::lo32(Ap), c0);
    uint32_t s1 = glv::addc(glv::hi32(acc), glv::hi32(Ap), c0, c1);
    s1 = glv::add_co(s1, glv::lo32(Bp), c2);
    const uint32_t h0 = glv::addc0(glv::addc0(glv::hi32(Bp), c1), c2);
    uint64_t lo[1] = {glv::pk(s0, s1)}, hi[1] = {glv::pk(h0, 0u)}, r[1];
    glv::reduce_canon_g<1>(lo, hi, r, qm1);
    return r[0];
}

}  // namespace fftl
}  // namespace tp
