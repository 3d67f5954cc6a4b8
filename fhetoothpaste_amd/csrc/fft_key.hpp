// Host side of the exact floating-point external product (kernels_br_fft.hip): the twiddle
// table, the bootstrapping key in the transform domain of the folded complex FFT, and the
// rounding bound that decides whether a parameter set may use it.
//
// The device computes, per blind-rotation step and output half h,
//     c_h = sum over the 2 l digit polynomials d_r of d_r * K_{r,h}   (negacyclic, over Z)
// with every key polynomial K (coefficients taken in (-Q/2, Q/2]) split into 4 balanced limbs
// K = sum_k 2^(16 k) K^(k), |K^(k)| <= 2^15, one floating-point convolution per limb, each
// rounded to the nearest integer and recombined mod Q.  The result is exact whenever every
// limb convolution's rounding error stays below 1/2, which rounding_bound() proves.
#pragma once
#include <stdint.h>
#include <vector>

namespace tp {
namespace fftk {

constexpr int kLimbs = 4;
constexpr uint32_t kN = 1024;  // ring dimension of the device transform (512 complex points)

// zeta[k] = exp(i pi brv10(k) / 1024), k in [0, 1024), as (re, im) pairs of correctly rounded
// doubles (computed in long double).
std::vector<double> zeta_table();

// Folded transform of one real polynomial of length 1024 (coefficients as long double) in the
// device's position order: out[2p], out[2p + 1] = re, im of position p in [0, 512).
void forward_ld(const long double* a, long double* out);

// Device key of a binary CGGI bootstrapping key (reference layout [n][2 l rows][a, b][N], NTT
// domain): complex doubles [i][r = h' l + j][h][limb][t][L], position 8 L + t at index 64 t + L.
// max_abs receives max |value| over the whole key (re and im parts), which the bound uses.
std::vector<double> device_key(const uint64_t* brk, int n, int digits, double* max_abs);

// Rigorous upper bound on |computed - exact| of any limb convolution coefficient for the given
// unsigned gadget (digits of base 2^base_log2, the top one partial) when every transform-domain
// key value has modulus at most key_inf (pass kKeyInfWorst for a bound valid for any key).
double rounding_bound(int base_log2, int digits, double key_inf);
extern const double kKeyInfWorst;  // 512 sqrt(2) 2^15: the bound of any balanced-limb key

}  // namespace fftk
}  // namespace tp
