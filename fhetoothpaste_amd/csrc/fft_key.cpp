// Host side of the exact floating-point external product: twiddles, transform-domain key and
// rounding bound (declared in fft_key.hpp).
#include "fft_key.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

#include "gl64.hpp"
#include "host_ntt.hpp"

namespace tp {
namespace fftk {

namespace {

constexpr int kHalf = kN / 2;  // complex points

uint32_t brv10(uint32_t k) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((k >> b) & 1u) << (9 - b);
    return r;
}

struct ZetaLd {
    long double re[kN], im[kN];
    ZetaLd() {
        const long double pi = 3.141592653589793238462643383279502884L;
        for (uint32_t k = 0; k < kN; ++k) {
            const long double th = pi * (long double)brv10(k) / (long double)kN;
            re[k] = cosl(th);
            im[k] = sinl(th);
        }
    }
};

const ZetaLd& zeta_ld() {
    static const ZetaLd z;
    return z;
}

}  // namespace

const double kKeyInfWorst = 512.0 * 1.4142135623730951 * 32768.0;

std::vector<double> zeta_table() {
    const ZetaLd& z = zeta_ld();
    std::vector<double> t(2 * kN);
    for (uint32_t k = 0; k < kN; ++k) {
        t[2 * k] = (double)z.re[k];
        t[2 * k + 1] = (double)z.im[k];
    }
    return t;
}

// In-place Cooley-Tukey over Y^512 - i: level s, block b (positions b len .. (b + 1) len - 1,
// len = 512 >> s) pairs j and j + len / 2 with zeta[2^(s+1) + b]: the same factorisation and
// position order as the device transform (fft_wave.hpp).
void forward_ld(const long double* a, long double* out) {
    const ZetaLd& z = zeta_ld();
    long double* re = out;  // interleaved (re, im)
    for (int m = 0; m < kHalf; ++m) {
        re[2 * m] = a[m];
        re[2 * m + 1] = a[m + kHalf];
    }
    for (int s = 0; s < 9; ++s) {
        const int len = kHalf >> s, h = len / 2;
        for (int b = 0; b < (1 << s); ++b) {
            const long double wr = z.re[(2u << s) + b], wi = z.im[(2u << s) + b];
            for (int j = b * len; j < b * len + h; ++j) {
                const long double vr = out[2 * (j + h)], vi = out[2 * (j + h) + 1];
                const long double tr = vr * wr - vi * wi, ti = vr * wi + vi * wr;
                const long double ur = out[2 * j], ui = out[2 * j + 1];
                out[2 * j] = ur + tr;
                out[2 * j + 1] = ui + ti;
                out[2 * (j + h)] = ur - tr;
                out[2 * (j + h) + 1] = ui - ti;
            }
        }
    }
}

std::vector<double> device_key(const uint64_t* brk, int n, int digits, double* max_abs) {
    const size_t rows = (size_t)2 * digits;               // r = h' l + j
    const size_t per_step = rows * 2 * kLimbs * kHalf;    // complex values per step
    std::vector<double> out((size_t)n * per_step * 2);
    const HostNtt T(kN);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<double> thread_max(nt, 0.0);
    auto work = [&](unsigned id) {
        std::vector<uint64_t> poly(kN);
        std::vector<long double> limb(kN), tr(kN);
        double mx = 0.0;
        for (size_t job = id; job < (size_t)n * rows * 2; job += nt) {
            const size_t i = job / (rows * 2), r = (job / 2) % rows, h = job % 2;
            const uint64_t* src = brk + ((i * rows + r) * 2 + h) * kN;
            std::copy(src, src + kN, poly.begin());
            T.inverse(poly.data());
            // centred coefficient in (-Q/2, Q/2], then 4 balanced base-2^16 limbs
            int32_t L4[kLimbs][kN];
            for (uint32_t m = 0; m < kN; ++m) {
                int64_t v = poly[m] > gl::kQ / 2 ? (int64_t)(poly[m] - gl::kQ) : (int64_t)poly[m];
                // limbs 0..2 in [-2^15, 2^15); what remains, (v - low) / 2^48, lies in
                // [-2^15, 2^15] since |v| < 2^63 and |low| < 2^47
                for (int k = 0; k < kLimbs - 1; ++k) {
                    int64_t d = v & 0xFFFF;
                    if (d >= 0x8000) d -= 0x10000;
                    L4[k][m] = (int32_t)d;
                    v = (v - d) >> 16;
                }
                L4[kLimbs - 1][m] = (int32_t)v;
            }
            for (int k = 0; k < kLimbs; ++k) {
                for (uint32_t m = 0; m < kN; ++m) limb[m] = (long double)L4[k][m];
                forward_ld(limb.data(), tr.data());
                double* dst = out.data() + 2 * ((((i * rows + r) * 2 + h) * kLimbs + k) * kHalf);
                for (int p = 0; p < kHalf; ++p) {
                    const int L = p >> 3, t = p & 7;
                    const double re = (double)tr[2 * p], im = (double)tr[2 * p + 1];
                    dst[2 * (64 * t + L)] = re;
                    dst[2 * (64 * t + L) + 1] = im;
                    mx = std::max(mx, std::max(std::fabs(re), std::fabs(im)));
                }
            }
        }
        thread_max[id] = mx;
    };
    std::vector<std::thread> th;
    for (unsigned id = 1; id < nt; ++id) th.emplace_back(work, id);
    work(0);
    for (auto& t : th) t.join();
    if (max_abs) *max_abs = *std::max_element(thread_max.begin(), thread_max.end());
    return out;
}

// Error bound, following Higham's normwise FFT analysis (Theorem 24.2 of Accuracy and Stability
// of Numerical Algorithms): a transform of 9 levels has relative 2-norm error at most
// delta = 9 eta / (1 - 9 eta), eta = mu + gamma_4 (sqrt 2 + mu).  For one digit polynomial x
// (folded, 2-norm |x|_2) and key limb transform Y (|Y|_inf <= key_inf; each entry also within
// u/2 per part of the exact one, being correctly rounded):
//   forward error E of X:  |E|_2 <= delta sqrt(n) |x|_2, and after the product and the inverse
//                          transform (1/n, |IFFT v|_inf <= |v|_2 / sqrt n):  delta |x|_2 key_inf;
//   key rounding:          |X|_inf |E_Y|_2 / sqrt n <= u |x|_1 |y|_2;
//   product + summation:   (sqrt(2) gamma_2 + 2 l u) |x|_1 |y|_2 per term (|X|_inf <= |x|_1,
//                          |Y|_2 = sqrt(n) |y|_2);
// summed over the 2 l digit polynomials; and the inverse transform's own rounding,
// delta |c|_2 <= delta sum |x_r|_2 key_inf (Parseval, |c_r|_2 <= |x_r|_2 |Y_r|_inf).
// Worst case over digits: |x|_2 = sqrt(N) dmax, |x|_1 <= (N/2) sqrt(2) dmax; |y|_2 <= sqrt(N) 2^15.
// The recombination (fft_limbs.hpp) needs every limb value below 2^57 / (1 + 2^16) in magnitude:
// a gadget whose limb convolutions could exceed that gets an infinite bound.
double rounding_bound(int base_log2, int digits, double key_inf) {
    const double u = std::ldexp(1.0, -53);
    const double mu = u;  // twiddles correctly rounded per part: |w' - w| <= u/sqrt(2) < u
    const double g4 = 4 * u / (1 - 4 * u), g2 = 2 * u / (1 - 2 * u);
    const double eta = mu + g4 * (std::sqrt(2.0) + mu);
    const double delta = 9 * eta / (1 - 9 * eta);
    const int D = 64 / base_log2 + (64 % base_log2 ? 1 : 0);
    const int first_shift = base_log2 * (D - 1);
    const double y2 = std::sqrt((double)kN) * 32768.0;
    double sum_x2 = 0, err = 0, cmax = 0;
    for (int j = 0; j < digits; ++j) {
        const int shift = first_shift - j * base_log2;
        const int bits = std::min(base_log2, 64 - shift);
        const double dmax = std::ldexp(1.0, bits) - 1;
        cmax += 2 * kN * dmax * 32768.0;
        const double x2 = std::sqrt((double)kN) * dmax, x1 = kHalf * std::sqrt(2.0) * dmax;
        // both input halves contribute this digit
        for (int half = 0; half < 2; ++half) {
            sum_x2 += x2;
            err += delta * x2 * key_inf + (u + std::sqrt(2.0) * g2 + 2 * digits * u) * x1 * y2;
        }
    }
    if (cmax * 65537.0 >= std::ldexp(1.0, 57)) return HUGE_VAL;
    return err + delta * sum_x2 * key_inf * (1 + delta);
}

}  // namespace fftk
}  // namespace tp
