// One-wave-per-SIMD variants of the blind-rotation kernel (512 registers per lane): chosen per
// context by tp_autotune when they beat the two-wave variants (ternary and BMMP spill at 256).
#include "br_kernel.hpp"

namespace tp {

hipError_t launch_blind_rotate_w1(int logN, int method, int digits, const BrArgs& A, int grid, hipStream_t s) {
#define TP_BR1_CASE(LG, LD, M) \
    if (logN == LG && digits == LD && method == M) { launch_br<LG, LD, M, 1>(A, grid, s); return hipGetLastError(); }
#define TP_BR1_METHODS(LG, LD) TP_BR1_CASE(LG, LD, 0) TP_BR1_CASE(LG, LD, 1) TP_BR1_CASE(LG, LD, 2)
    TP_BR1_METHODS(10, 3) TP_BR1_METHODS(10, 4)
    TP_BR1_METHODS(11, 3) TP_BR1_METHODS(11, 4)
    TP_BR1_METHODS(12, 3) TP_BR1_METHODS(12, 4)
#undef TP_BR1_METHODS
#undef TP_BR1_CASE
    return hipErrorInvalidValue;
}

}  // namespace tp
