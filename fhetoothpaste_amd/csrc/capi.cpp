// C-ABI of the gfx950 bootstrapping path: contexts, key import (reference layouts ->
// device layouts), launches.  Declared in include/toothpaste_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/toothpaste_amd.h"
#include "gl64.hpp"
#include "context_internal.hpp"
#include "fft_key.hpp"
#include "host_ntt.hpp"
#include "kernels.hpp"
#include "ks_limbs.hpp"
#include "tuning_cache.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define TP_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(TP_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

uint32_t max_digits(uint32_t b) { return 64 / b + (64 % b ? 1 : 0); }

int ilog2(uint32_t x) {
    int r = 0;
    while ((1u << r) < x) ++r;
    return r;
}

}  // namespace

struct tp_context {
    tp_params p{};
    int device = 0;
    uint32_t N = 0;
    int logN = 0;
    hipStream_t stream = nullptr;
    uint64_t* d_tw = nullptr;
    uint64_t* d_itw = nullptr;
    uint64_t* d_brk = nullptr;
    size_t brk_words = 0;
    uint64_t* d_ksk = nullptr;
    uint64_t* d_corr = nullptr;
    size_t ksk_words = 0;
    uint64_t ks_corrector = 0;
    int8_t* d_kl = nullptr;      // the key-switching key as matrix-core operands (ks_limbs.hpp), where it fits
    int ks_mfma = 1;             // key switch past the row-split sizes on the matrix cores where d_kl is built
    uint64_t* d_scratch = nullptr;  // extracted LWEs of tp_bootstrap_batch_device
    size_t scratch_words = 0;
    uint64_t* d_io = nullptr;  // staging of the host-pointer entry points
    uint64_t* d_mono = nullptr;  // NTT(X^e) table (ternary, BMMP)
    uint64_t br_corrector = 0, br_least = 0;
    uint64_t* d_auto = nullptr;  // automorphism keys, log2 N x d x 2N, key-row order
    uint64_t* d_nodes = nullptr; // packing tree nodes
    size_t nodes_words = 0;
    uint32_t* d_perm = nullptr;
    size_t perm_len = 0;
    uint64_t* d_rksk = nullptr;  // RLWEtoRLWEConverter key, d x 2N, key-row order
    int br_waves = 2;            // blind-rotation variant: waves per SIMD (tp_set_option / tp_autotune)
    size_t split_max = 0;        // launches of at most this many ciphertexts use the digit-parallel kernel
    uint64_t* d_sapw = nullptr;  // SAP decomposition workspace (per-input state, accumulators, tree root)
    size_t sapw_words = 0;
    uint64_t* d_sapt = nullptr;  // SAP HomTrunc tree levels (two buffers of nodes x inputs RLWEs)
    size_t sapt_words = 0;
    uint64_t* d_sap_polys = nullptr;  // SAP NTT-domain polynomials for sap_radix
    uint32_t sap_radix = 0;
    uint64_t* d_rgsw = nullptr;       // LWE -> RGSW workspace
    size_t rgsw_words = 0;
    hipEvent_t ev[8] = {};
    // exact floating-point external product (binary CGGI at N = 1024, kernels_br_fft.hip)
    double* d_fkey = nullptr;    // transform-domain key limbs
    double* d_zeta = nullptr;    // twiddle table
    bool fft_ready = false;      // key imported and its rounding bound below 1/2
    double fft_bound = -1.0;     // that bound (-1: not computed)
    int br_engine = 0;           // 0: Goldilocks NTT (default), 1: floating-point product where ready (opt-in)
    int fft_waves = 1;           // waves per SIMD of the floating-point kernel (1 or 2)
    int br_grid = 0;             // work-groups of a br_kernel launch: 0 one per ciphertext, else at most this many
    int split_halves = 1;        // digit-parallel launches: 1 = two waves per transform where built
    int cu_count = 256;          // compute units of the context's device
    char arch[32] = {0};         // gcnArchName of the context's device
};

namespace {

// Signed-digit corrector and least-digit correction (gadget_decomp.h:16-28).
uint64_t signed_corrector(uint32_t b, uint64_t* least) {
    uint64_t C = 0;
    int64_t shift = 64 - (int64_t)b;
    while (shift > 0) { C += 1ULL << (shift + b - 1); shift -= b; }
    *least = 1ULL << (shift + b - 1);
    C += *least;
    return C % gl::kQ;
}

// Reference key rows (N words each, NTT order) -> lane-major device order: NTT index 16 L + k
// goes to ((k >> 1) * 2P + 2L + (k & 1)).
std::vector<uint64_t> to_device_rows(const uint64_t* src, size_t words, uint32_t N) {
    const uint32_t P = N / 16;
    std::vector<uint64_t> dev(words);
    for (size_t r = 0; r < words / N; ++r)
        for (uint32_t L = 0; L < P; ++L)
            for (uint32_t k = 0; k < 16; ++k)
                dev[r * N + (k >> 1) * 2 * P + 2 * L + (k & 1)] = src[r * N + 16 * L + k];
    return dev;
}

int set_device(tp_context* c) {
    TP_HIP(hipSetDevice(c->device));
    return TP_OK;
}

size_t brk_words_for(const tp_params& p) {
    const size_t rg = (size_t)4 * p.ring_dim * p.br_digits;
    switch (p.method) {
        case TP_METHOD_CGGI_BINARY: return rg * p.lwe_dim;
        case TP_METHOD_CGGI_TERNARY: return 2 * rg * p.lwe_dim;
        case TP_METHOD_BMMP: return 3 * rg * (p.lwe_dim / 2);
    }
    return 0;
}

int ensure_scratch(tp_context* c, size_t words) {
    if (c->scratch_words >= words) return TP_OK;
    if (c->d_scratch) TP_HIP(hipFree(c->d_scratch));
    c->d_scratch = nullptr;
    c->scratch_words = 0;
    if (hipMalloc(&c->d_scratch, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "scratch allocation failed");
    c->scratch_words = words;
    return TP_OK;
}

void fill_br_args(tp_context* c, tp::BrArgs& A, const uint64_t* lwes, const uint64_t* tv, size_t tv_stride,
                  uint64_t* out_lwe, uint64_t* out_rlwe, int rlwe_ntt, size_t count) {
    A.brk = c->d_brk;
    A.lwe = lwes;
    A.tv = tv;
    A.tv_stride = tv_stride;
    A.out_lwe = out_lwe;
    A.out_rlwe = out_rlwe;
    A.out_rlwe_ntt = rlwe_ntt;
    A.tw = c->d_tw;
    A.itw = c->d_itw;
    A.n = (int)c->p.lwe_dim;
    A.base_log2 = (int)c->p.br_base_log2;
    A.first_shift = (int)(c->p.br_base_log2 * (max_digits(c->p.br_base_log2) - 1));
    A.count = (int)count;
    A.corrector = c->br_corrector;
    A.least = c->br_least;
    A.max_digits = (int)max_digits(c->p.br_base_log2);
    A.mono = c->d_mono;
}

// The digit-parallel kernel (2 l waves per ciphertext) is built for N = 1024 with 2..4 digits.
bool split_built(const tp_context* c) { return c->logN == 10 && c->p.br_digits >= 2 && c->p.br_digits <= 4; }

// Parameter sets the floating-point kernel is built for: binary CGGI, N = 1024, 2..4 digits (it is
// used only when the rounding bound of the imported key is below 1/2 as well).
bool fft_built(const tp_context* c) {
    return c->logN == 10 && c->p.method == TP_METHOD_CGGI_BINARY && c->p.br_digits >= 2 && c->p.br_digits <= 4;
}
bool fft_active(const tp_context* c) { return c->br_engine == 1 && c->fft_ready; }
// digit-parallel kernel with every transform split over two waves: binary CGGI, N = 1024, 2..3 digits
bool halves_built(const tp_context* c) {
    return c->logN == 10 && c->p.method == TP_METHOD_CGGI_BINARY && (c->p.br_digits == 2 || c->p.br_digits == 3);
}

hipError_t launch_br(tp_context* c, const tp::BrArgs& A, size_t count, int grid) {
    if (split_built(c) && count <= c->split_max) {
        if (c->split_halves && halves_built(c))
            return tp::launch_blind_rotate_half((int)c->p.br_digits, A, grid, c->stream);
        return tp::launch_blind_rotate_latency((int)c->p.method, (int)c->p.br_digits, A, grid, c->stream);
    }
    if (fft_active(c))
        return tp::launch_blind_rotate_fft((int)c->p.br_digits, A, c->d_fkey, c->d_zeta, grid, c->fft_waves,
                                           c->stream);
    return tp::launch_blind_rotate(c->logN, (int)c->p.method, (int)c->p.br_digits, A, grid, c->br_waves, c->stream);
}

int launch_eval(tp_context* c, size_t count, const uint64_t* lwes, const uint64_t* tv, size_t tv_stride,
                uint64_t* out_lwe, uint64_t* out_rlwe, int rlwe_ntt, int grid) {
    if (!c->d_brk) return fail(TP_E_STATE, "blind-rotation key not imported");
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    if (tv_stride != 0 && tv_stride != 2 * (size_t)c->N) return fail(TP_E_INVALID, "tv_stride must be 0 or 2N");
    tp::BrArgs A{};
    fill_br_args(c, A, lwes, tv, tv_stride, out_lwe, out_rlwe, rlwe_ntt, count);
    if (grid <= 0) {
        grid = (int)count;
        // a fixed grid (work-groups stride over the ciphertexts) for the one-work-group-per-ciphertext kernels
        if (c->br_grid > 0 && !(split_built(c) && count <= c->split_max)) grid = std::min(grid, c->br_grid);
    }
    hipError_t e = launch_br(c, A, count, grid);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("blind rotation launch: ") + hipGetErrorString(e));
    return TP_OK;
}

}  // namespace

extern "C" {

const char* tp_last_error(void) { return g_error.c_str(); }

int tp_device_count(int* count) {
    if (!count) return fail(TP_E_INVALID, "null count");
    *count = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return TP_OK;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++*count;
    }
    return TP_OK;
}

int tp_context_create(tp_context** out, const tp_params* p, int device) {
    if (!out || !p) return fail(TP_E_INVALID, "null argument");
    *out = nullptr;
    const uint32_t N = p->ring_dim;
    if (N != 1024 && N != 2048 && N != 4096) return fail(TP_E_INVALID, "ring_dim must be 1024, 2048 or 4096");
    if (p->method > TP_METHOD_BMMP) return fail(TP_E_INVALID, "unknown method");
    if (p->br_base_log2 == 0 || p->br_base_log2 > 32 || p->br_digits == 0 ||
        p->br_digits > max_digits(p->br_base_log2))
        return fail(TP_E_INVALID, "bad gadget parameters");
    if (p->method == TP_METHOD_BMMP && (p->lwe_dim & 1)) return fail(TP_E_INVALID, "BMMP needs an even lwe_dim");
    // blind-rotation kernels are built for 2..4 digits at N = 1024 / 2048 and 3..4 at N = 4096
    if (p->br_digits < 2 || p->br_digits > 4 || (N == 4096 && p->br_digits < 3))
        return fail(TP_E_INVALID, "no blind-rotation kernel is built for this (ring_dim, digits)");
    if (p->rlwe_ks_digits) {
        const uint32_t b = p->rlwe_ks_base_log2, d = p->rlwe_ks_digits;
        if (b == 0 || b > 32 || d > max_digits(b)) return fail(TP_E_INVALID, "bad automorphism-key gadget parameters");
    }
    if (p->ks_out_dim) {
        const uint32_t b = p->ks_base_log2, d = p->ks_digits;
        if (b == 0 || b > 16 || d == 0 || d > max_digits(b) || (d == max_digits(b) && 64 % b))
            return fail(TP_E_INVALID, "bad key-switch gadget parameters");
        if ((uint64_t)N * d << b >= (1ULL << 32)) return fail(TP_E_INVALID, "key-switch accumulator bound");
    }
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("no device: ") + hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(TP_E_DEVICE, std::string("device is not gfx950: ") + prop.gcnArchName);
    tp_context* c = new tp_context;
    c->p = *p;
    c->device = device;
    c->N = N;
    c->logN = ilog2(N);
    // up to two ciphertexts per CU: two rounds of the digit-parallel kernel (one work-group per CU)
    // still beat one round of br_kernel at one wave per SIMD
    if (prop.multiProcessorCount > 0) c->cu_count = prop.multiProcessorCount;
    std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    for (char* q = c->arch; *q; ++q)
        if (*q == ':' || *q == '"' || *q == '\\') *q = '\0';  // "gfx950:sramecc+:xnack-" -> "gfx950"
    if (split_built(c)) c->split_max = 2 * (size_t)c->cu_count;
    int rc = set_device(c);
    if (rc) { delete c; return rc; }
    // device address of the no-allocation entry point's scratch array, once per device (its own
    // error is reported by that entry point; this context works without it)
    tp::noalloc_setup(device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(TP_E_DEVICE, "stream creation failed");
    }
    tp::HostNtt T(N);
    // d_tw, d_itw: psi_rev, psi_inv_rev, then the twisted transforms' tables and the two-wave
    // transforms' tables (both used at N = 1024)
    const std::vector<uint64_t> t64 = T.twisted64(), it64 = T.twisted64(true);
    const std::vector<uint64_t> h64 = T.half_tables(), ih64 = T.half_tables(true);
    if (hipMalloc(&c->d_tw, 5 * N * 8) != hipSuccess || hipMalloc(&c->d_itw, 5 * N * 8) != hipSuccess ||
        hipMalloc(&c->d_io, (size_t)8 * N * 8) != hipSuccess) {
        tp_context_destroy(c);
        return fail(TP_E_NOMEM, "table allocation failed");
    }
    bool ok = hipMemcpy(c->d_tw, T.psi_rev.data(), N * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->d_tw + N, t64.data(), 2 * N * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->d_itw, T.psi_inv_rev.data(), N * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->d_itw + N, it64.data(), 2 * N * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->d_tw + 3 * N, h64.data(), 2 * N * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->d_itw + 3 * N, ih64.data(), 2 * N * 8, hipMemcpyHostToDevice) == hipSuccess;
    for (auto& ev : c->ev) ok = ok && hipEventCreate(&ev) == hipSuccess;
    if (!ok) {
        tp_context_destroy(c);
        return fail(TP_E_DEVICE, "context set-up failed");
    }
    {
        // signed-digit corrector (gadget_decomp.h:16-28)
        const uint32_t b = p->br_base_log2;
        uint64_t C = 0;
        int64_t shift = 64 - (int64_t)b;
        while (shift > 0) { C += 1ULL << (shift + b - 1); shift -= b; }
        c->br_least = 1ULL << (shift + b - 1);
        C += c->br_least;
        c->br_corrector = C % gl::kQ;
    }
    if (p->method != TP_METHOD_CGGI_BINARY) {
        // NTT(X^e)[pos] = psi^((2 brv(pos) + 1) e), stored in the lane-major key-row order
        const uint32_t P = N / 16;
        std::vector<uint64_t> mono((size_t)N * N);
        const uint64_t psi = gl::pow(7, (gl::kQ - 1) >> (c->logN + 1));
        for (uint32_t pos = 0; pos < N; ++pos) {
            const uint64_t base = gl::mul(gl::mul(T.psi_rev[pos], T.psi_rev[pos]), psi);
            const uint32_t L = pos / 16, k = pos % 16;
            const size_t dpos = (size_t)(k >> 1) * 2 * P + 2 * L + (k & 1);
            uint64_t v = 1;
            for (uint32_t e = 0; e < N; ++e) {
                mono[(size_t)e * N + dpos] = v;
                v = gl::mul(v, base);
            }
        }
        if (hipMalloc(&c->d_mono, mono.size() * 8) != hipSuccess) {
            tp_context_destroy(c);
            return fail(TP_E_NOMEM, "monomial table allocation failed");
        }
        if (hipMemcpy(c->d_mono, mono.data(), mono.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
            tp_context_destroy(c);
            return fail(TP_E_DEVICE, "monomial table upload failed");
        }
    }
    *out = c;
    return TP_OK;
}

int tp_context_destroy(tp_context* c) {
    if (!c) return TP_OK;
    // teardown is best effort: errors here have no caller to report to
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (uint64_t* p : {c->d_tw, c->d_itw, c->d_brk, c->d_ksk, c->d_corr, c->d_scratch, c->d_io, c->d_mono,
                        c->d_auto, c->d_nodes, c->d_rksk, c->d_sapw, c->d_sapt, c->d_sap_polys, c->d_rgsw})
        if (p) (void)hipFree(p);
    for (double* p : {c->d_fkey, c->d_zeta})
        if (p) (void)hipFree(p);
    if (c->d_kl) (void)hipFree(c->d_kl);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->d_perm) (void)hipFree(c->d_perm);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TP_OK;
}

int tp_context_sync(tp_context* c) {
    if (!c) return fail(TP_E_INVALID, "null context");
    TP_HIP(hipStreamSynchronize(c->stream));
    return TP_OK;
}

int tp_context_params(tp_context* c, tp_params* out) {
    if (!c || !out) return fail(TP_E_INVALID, "null argument");
    *out = c->p;
    return TP_OK;
}

int tp_import_brk(tp_context* c, const uint64_t* brk, size_t words) {
    if (!c || !brk) return fail(TP_E_INVALID, "null argument");
    const size_t want = brk_words_for(c->p);
    if (words != want) return fail(TP_E_INVALID, "brk size mismatch: expected " + std::to_string(want));
    int rc = set_device(c);
    if (rc) return rc;
    // Permute every N-word key row into the lane-major order of the kernels:
    // NTT index 16 L + k goes to ((k >> 1) * 2P + 2L + (k & 1)).
    const uint32_t N = c->N, P = N / 16;
    std::vector<uint64_t> dev(words);
    for (size_t r = 0; r < words / N; ++r) {
        const uint64_t* src = brk + r * N;
        uint64_t* dst = dev.data() + r * N;
        for (uint32_t L = 0; L < P; ++L)
            for (uint32_t k = 0; k < 16; ++k) dst[(k >> 1) * 2 * P + 2 * L + (k & 1)] = src[16 * L + k];
    }
    if (c->d_brk) { TP_HIP(hipFree(c->d_brk)); c->d_brk = nullptr; }
    if (hipMalloc(&c->d_brk, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "brk allocation failed");
    TP_HIP(hipMemcpy(c->d_brk, dev.data(), words * 8, hipMemcpyHostToDevice));
    c->brk_words = words;
    // transform-domain key of the floating-point kernel, kept when its rounding bound certifies
    // exact results for this key (any digits of the gadget)
    c->fft_ready = false;
    c->fft_bound = -1.0;
    if (c->d_fkey) { TP_HIP(hipFree(c->d_fkey)); c->d_fkey = nullptr; }
    if (fft_built(c)) {
        double key_inf = 0.0;
        const std::vector<double> fk = tp::fftk::device_key(brk, (int)c->p.lwe_dim, (int)c->p.br_digits, &key_inf);
        c->fft_bound = tp::fftk::rounding_bound((int)c->p.br_base_log2, (int)c->p.br_digits, key_inf);
        if (c->fft_bound < 0.45) {
            if (!c->d_zeta) {
                const std::vector<double> z = tp::fftk::zeta_table();
                if (hipMalloc(&c->d_zeta, z.size() * 8) != hipSuccess) return fail(TP_E_NOMEM, "twiddle allocation failed");
                TP_HIP(hipMemcpy(c->d_zeta, z.data(), z.size() * 8, hipMemcpyHostToDevice));
            }
            if (hipMalloc(&c->d_fkey, fk.size() * 8) != hipSuccess)
                return fail(TP_E_NOMEM, "transform-domain key allocation failed");
            TP_HIP(hipMemcpy(c->d_fkey, fk.data(), fk.size() * 8, hipMemcpyHostToDevice));
            c->fft_ready = true;
        }
    }
    return TP_OK;
}

int tp_import_lwe_ksk(tp_context* c, const uint64_t* ksk, size_t words) {
    if (!c || !ksk) return fail(TP_E_INVALID, "null argument");
    if (!c->p.ks_out_dim) return fail(TP_E_INVALID, "context has no key switch (ks_out_dim = 0)");
    const uint32_t n_in = c->N, n_out = c->p.ks_out_dim, b = c->p.ks_base_log2, d = c->p.ks_digits;
    const size_t W = n_out + 1;
    if (words != (size_t)n_in * d * W) return fail(TP_E_INVALID, "ksk size mismatch");
    int rc = set_device(c);
    if (rc) return rc;
    // corr[k] = sum_ij delta_j K[i][j][k]: the signed digits are d_ij = u_ij - delta_j.
    const uint32_t D = max_digits(b);
    const uint64_t least = 1ULL << ((64 - (int)b * (int)D) + (int)b - 1);
    std::vector<uint64_t> colsum((size_t)d * W, 0), corr(W, 0);
    for (uint32_t i = 0; i < n_in; ++i)
        for (uint32_t j = 0; j < d; ++j) {
            const uint64_t* row = ksk + ((size_t)i * d + j) * W;
            uint64_t* cs = colsum.data() + (size_t)j * W;
            for (size_t k = 0; k < W; ++k) cs[k] = gl::add(cs[k], row[k]);
        }
    for (uint32_t j = 0; j < d; ++j) {
        const uint64_t delta = (j != D - 1) ? (1ULL << (b - 1)) : least;
        for (size_t k = 0; k < W; ++k) corr[k] = gl::add(corr[k], gl::mul(colsum[(size_t)j * W + k], delta));
    }
    uint64_t C = 0;
    int64_t shift = 64 - (int64_t)b;
    while (shift > 0) { C += 1ULL << (shift + b - 1); shift -= b; }
    C += 1ULL << (shift + b - 1);
    c->ks_corrector = C % gl::kQ;
    if (c->d_ksk) { TP_HIP(hipFree(c->d_ksk)); c->d_ksk = nullptr; }
    if (c->d_corr) { TP_HIP(hipFree(c->d_corr)); c->d_corr = nullptr; }
    if (hipMalloc(&c->d_ksk, (words + tp::kKsPadWords) * 8) != hipSuccess ||
        hipMalloc(&c->d_corr, W * 8) != hipSuccess)
        return fail(TP_E_NOMEM, "ksk allocation failed");
    TP_HIP(hipMemcpy(c->d_ksk, ksk, words * 8, hipMemcpyHostToDevice));
    TP_HIP(hipMemset(c->d_ksk + words, 0, tp::kKsPadWords * 8));
    TP_HIP(hipMemcpy(c->d_corr, corr.data(), W * 8, hipMemcpyHostToDevice));
    c->ksk_words = words;
    if (c->d_kl) { TP_HIP(hipFree(c->d_kl)); c->d_kl = nullptr; }
    if (tp::ks_mfma_fits(n_in, b, d)) {
        const size_t bytes = tp::ks_limb_bytes(n_in, n_out);
        std::vector<int8_t> kl(bytes);
        tp::build_ks_limbs(ksk, n_in, n_out, b, d, kl.data());
        if (hipMalloc(&c->d_kl, bytes) != hipSuccess) {
            c->d_kl = nullptr;
            return fail(TP_E_NOMEM, "ksk operand allocation failed");
        }
        TP_HIP(hipMemcpy(c->d_kl, kl.data(), bytes, hipMemcpyHostToDevice));
    }
    return TP_OK;
}

namespace {
tp::KeySwitchArgs auto_args(tp_context* c) {
    tp::KeySwitchArgs A{};
    A.tw = c->d_tw;
    A.itw = c->d_itw;
    A.base_log2 = (int)c->p.rlwe_ks_base_log2;
    A.digits = (int)c->p.rlwe_ks_digits;
    A.max_digits = (int)max_digits(c->p.rlwe_ks_base_log2);
    A.corrector = signed_corrector(c->p.rlwe_ks_base_log2, &A.least);
    return A;
}
}  // namespace

int tp_import_auto_keys(tp_context* c, const uint64_t* keys, size_t words) {
    if (!c || !keys) return fail(TP_E_INVALID, "null argument");
    if (!c->p.rlwe_ks_digits) return fail(TP_E_INVALID, "context has no automorphism gadget (rlwe_ks_digits = 0)");
    const size_t want = (size_t)c->logN * c->p.rlwe_ks_digits * 2 * c->N;
    if (words != want) return fail(TP_E_INVALID, "automorphism keys size mismatch: expected " + std::to_string(want));
    int rc = set_device(c);
    if (rc) return rc;
    std::vector<uint64_t> dev = to_device_rows(keys, words, c->N);
    if (c->d_auto) { TP_HIP(hipFree(c->d_auto)); c->d_auto = nullptr; }
    if (hipMalloc(&c->d_auto, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "automorphism key allocation failed");
    TP_HIP(hipMemcpy(c->d_auto, dev.data(), words * 8, hipMemcpyHostToDevice));
    return TP_OK;
}

int tp_automorphism_device(tp_context* c, size_t count, uint32_t t, const uint64_t* in, uint64_t* out) {
    if (!c || !in || !out) return fail(TP_E_INVALID, "null argument");
    if (!c->d_auto) return fail(TP_E_STATE, "automorphism keys not imported");
    int j = 0;
    while (j <= c->logN && ((1u << j) + 1) != t) ++j;
    if (j < 1 || j > c->logN) return fail(TP_E_INVALID, "automorphism index must be 2^j + 1, 1 <= j <= log2 N");
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    int rc = set_device(c);
    if (rc) return rc;
    const size_t key_words = (size_t)c->p.rlwe_ks_digits * 2 * c->N;
    hipError_t e = tp::launch_automorphism(c->logN, auto_args(c), in, out, (int)count, t,
                                           c->d_auto + (size_t)(j - 1) * key_words, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("automorphism launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_trace_device(tp_context* c, size_t count, const uint64_t* in, uint64_t* out) {
    if (!c || !in || !out) return fail(TP_E_INVALID, "null argument");
    if (!c->d_auto) return fail(TP_E_STATE, "automorphism keys not imported");
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    int rc = set_device(c);
    if (rc) return rc;
    hipError_t e = tp::launch_trace(c->logN, auto_args(c), in, out, (int)count, c->d_auto, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("trace launch: ") + hipGetErrorString(e));
    return TP_OK;
}

namespace {
int run_pack(tp_context* c, size_t len, size_t n_inputs, const uint64_t* lwes, const uint32_t* perm, uint64_t* out) {
    int rc = set_device(c);
    if (rc) return rc;
    const size_t words = len * 2 * c->N;
    if (c->nodes_words < words) {
        if (c->d_nodes) TP_HIP(hipFree(c->d_nodes));
        c->d_nodes = nullptr;
        c->nodes_words = 0;
        if (hipMalloc(&c->d_nodes, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "packing buffer allocation failed");
        c->nodes_words = words;
    }
    hipError_t e = tp::launch_pack(c->logN, auto_args(c), lwes, perm, (int)len, (int)n_inputs, c->d_nodes, c->d_auto,
                                   c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("packing launch: ") + hipGetErrorString(e));
    TP_HIP(hipMemcpyAsync(out, c->d_nodes, 2 * c->N * 8, hipMemcpyDeviceToDevice, c->stream));
    return TP_OK;
}
}  // namespace

int tp_pack_device(tp_context* c, size_t len, const uint64_t* lwes, const uint32_t* perm, uint64_t* out) {
    if (!c || !lwes || !out) return fail(TP_E_INVALID, "null argument");
    if (!c->d_auto) return fail(TP_E_STATE, "automorphism keys not imported");
    if (len < 2 || len > c->N || (len & (len - 1))) return fail(TP_E_INVALID, "len must be a power of two in [2, N]");
    return run_pack(c, len, len, lwes, perm, out);
}

int tp_pack_consecutive_device(tp_context* c, size_t len, const uint64_t* lwes, uint64_t* out) {
    if (!c || !lwes || !out) return fail(TP_E_INVALID, "null argument");
    if (!c->d_auto) return fail(TP_E_STATE, "automorphism keys not imported");
    if (len < 1 || len > c->N) return fail(TP_E_INVALID, "len must lie in [1, N]");
    return run_pack(c, c->N, len, lwes, nullptr, out);
}

int tp_import_rlwe_ksk(tp_context* c, const uint64_t* key, size_t words) {
    if (!c || !key) return fail(TP_E_INVALID, "null argument");
    if (!c->p.rlwe_ks_digits) return fail(TP_E_INVALID, "context has no RLWE gadget (rlwe_ks_digits = 0)");
    const size_t want = (size_t)c->p.rlwe_ks_digits * 2 * c->N;
    if (words != want) return fail(TP_E_INVALID, "RLWE key size mismatch: expected " + std::to_string(want));
    int rc = set_device(c);
    if (rc) return rc;
    std::vector<uint64_t> dev = to_device_rows(key, words, c->N);
    if (c->d_rksk) { TP_HIP(hipFree(c->d_rksk)); c->d_rksk = nullptr; }
    if (hipMalloc(&c->d_rksk, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "RLWE key allocation failed");
    TP_HIP(hipMemcpy(c->d_rksk, dev.data(), words * 8, hipMemcpyHostToDevice));
    return TP_OK;
}

int tp_rlwe_keyswitch_device(tp_context* c, size_t count, uint32_t t, int mode, int key, const uint64_t* in,
                             uint64_t* out) {
    if (!c || !in || !out) return fail(TP_E_INVALID, "null argument");
    const int base_mode = mode & 3;
    if ((mode & ~(3 | TP_RKS_ACCUMULATE)) || base_mode > TP_RKS_SWITCH_ONLY) return fail(TP_E_INVALID, "bad mode");
    if (!(t & 1) || t >= 2 * c->N) return fail(TP_E_INVALID, "automorphism index must be odd and below 2N");
    if (base_mode == TP_RKS_CONVERT && t != 1) return fail(TP_E_INVALID, "plain key switch takes t = 1");
    const uint64_t* ksk = nullptr;
    if (key == 0) {
        if (!c->d_rksk) return fail(TP_E_STATE, "RLWE key not imported");
        ksk = c->d_rksk;
    } else {
        if (key < 1 || key > c->logN) return fail(TP_E_INVALID, "automorphism key index must lie in [1, log2 N]");
        if (!c->d_auto) return fail(TP_E_STATE, "automorphism keys not imported");
        ksk = c->d_auto + (size_t)(key - 1) * c->p.rlwe_ks_digits * 2 * c->N;
    }
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    int rc = set_device(c);
    if (rc) return rc;
    hipError_t e = tp::launch_rlwe_keyswitch(c->logN, auto_args(c), in, out, (int)count, t, base_mode,
                                             (mode & TP_RKS_ACCUMULATE) ? 1 : 0, ksk, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("RLWE key switch launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_modswitch_device(tp_context* c, uint64_t* dst, const uint64_t* src, size_t words, uint64_t source_modulus,
                        uint64_t target_modulus, int type, uint64_t seed) {
    if (!c || !dst || !src) return fail(TP_E_INVALID, "null argument");
    if (source_modulus < (1ULL << 63)) return fail(TP_E_INVALID, "source modulus must be >= 2^63");
    if (target_modulus < 2 || target_modulus > (1ULL << 62) || (target_modulus & (target_modulus - 1)))
        return fail(TP_E_INVALID, "target modulus must be a power of two in [2, 2^62]");
    if (type < TP_MS_ROUND || type > TP_MS_RANDOM) return fail(TP_E_INVALID, "unknown modulus-switch type");
    int rc = set_device(c);
    if (rc) return rc;
    const int t = __builtin_ctzll(target_modulus);
    hipError_t e = tp::launch_modswitch(dst, src, words, source_modulus, t, type, seed, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("modulus switch launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_eltwise_device(tp_context* c, int op, uint64_t* dst, const uint64_t* a, const uint64_t* b, const uint64_t* cc,
                      uint64_t scalar, size_t words) {
    if (!c || !dst || !a) return fail(TP_E_INVALID, "null argument");
    if (op < TP_ELT_ADD || op > TP_ELT_REDUCE) return fail(TP_E_INVALID, "unknown elementwise op");
    if ((op <= TP_ELT_FMA && !b) || (op == TP_ELT_FMA && !cc)) return fail(TP_E_INVALID, "missing operand");
    if (op == TP_ELT_MULS && scalar >= gl::kQ) return fail(TP_E_INVALID, "scalar must be canonical");
    int rc = set_device(c);
    if (rc) return rc;
    hipError_t e = tp::launch_eltwise(op, dst, a, b, cc, scalar, words, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("elementwise launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_ntt_forward_device(tp_context* c, uint64_t* dst, const uint64_t* src, size_t count) {
    if (!c || !dst || !src) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    hipError_t e = tp::launch_ntt_batch(c->logN, dst, src, count, c->d_tw, false, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("ntt launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_ntt_inverse_device(tp_context* c, uint64_t* dst, const uint64_t* src, size_t count) {
    if (!c || !dst || !src) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    hipError_t e = tp::launch_ntt_batch(c->logN, dst, src, count, c->d_itw, true, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("ntt launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_blind_rotate(tp_context* c, uint64_t* result, const uint64_t* lwe, const uint64_t* acc, int as_coef) {
    if (!c || !result || !lwe || !acc) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    const uint32_t N = c->N, n = c->p.lwe_dim;
    for (uint32_t i = 0; i <= n; ++i)
        if (lwe[i] >= 2 * (uint64_t)N) return fail(TP_E_INVALID, "lwe entries must lie in [0, 2N)");
    rc = ensure_scratch(c, (size_t)n + 1 + 4 * N);
    if (rc) return rc;
    uint64_t* d_lwe = c->d_scratch;
    uint64_t* d_acc = d_lwe + n + 1;
    uint64_t* d_res = d_acc + 2 * N;
    TP_HIP(hipMemcpyAsync(d_lwe, lwe, (n + 1) * 8, hipMemcpyHostToDevice, c->stream));
    TP_HIP(hipMemcpyAsync(d_acc, acc, 2 * N * 8, hipMemcpyHostToDevice, c->stream));
    rc = launch_eval(c, 1, d_lwe, d_acc, 0, nullptr, d_res, as_coef ? 0 : 1, 1);
    if (rc) return rc;
    TP_HIP(hipMemcpyAsync(result, d_res, 2 * N * 8, hipMemcpyDeviceToHost, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));
    return TP_OK;
}

int tp_eval_func_batch_device(tp_context* c, size_t count, const uint64_t* lwes, const uint64_t* tv,
                              size_t tv_stride, uint64_t* out_lwe, uint64_t* out_rlwe, int rlwe_ntt) {
    if (!c || !lwes || !tv) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    return launch_eval(c, count, lwes, tv, tv_stride, out_lwe, out_rlwe, rlwe_ntt, 0);
}

int tp_lwe_keyswitch_batch_device(tp_context* c, size_t count, const uint64_t* in, size_t in_stride,
                                  uint64_t* out) {
    if (!c || !in || !out) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    return tp::launch_ks_rows(c, count, in, in_stride, nullptr, out);
}

}  // extern "C"

namespace tp {

int launch_ks_rows(tp_context* c, size_t count, const uint64_t* in, size_t in_stride, const uint32_t* rows,
                   uint64_t* out) {
    if (!c->d_ksk) return fail(TP_E_STATE, "key-switching key not imported");
    if (in_stride < (size_t)c->N + 1) return fail(TP_E_INVALID, "in_stride < N + 1");
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    tp::KsArgs A{};
    A.rows = rows;
    A.ksk = c->d_ksk;
    A.corr = c->d_corr;
    A.in = in;
    A.out = out;
    A.in_stride = in_stride;
    A.corrector = c->ks_corrector;
    A.n_in = (int)c->N;
    A.n_out = (int)c->p.ks_out_dim;
    A.base_log2 = (int)c->p.ks_base_log2;
    A.digits = (int)c->p.ks_digits;
    A.count = (int)count;
    A.kl = c->ks_mfma ? c->d_kl : nullptr;
    hipError_t e = tp::launch_lwe_keyswitch(A, c->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("key switch launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int launch_chunk(tp_context* c, size_t count, const uint32_t* active, uint32_t* step, uint64_t* state,
                 const uint64_t* lwes, const uint64_t* tv, const uint32_t* tv_index, uint64_t* ext,
                 uint32_t chunk_steps) {
    if (!c->d_brk) return fail(TP_E_STATE, "blind-rotation key not imported");
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff || chunk_steps > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    tp::BrArgs A{};
    fill_br_args(c, A, lwes, tv, 0, ext, nullptr, 0, count);
    A.active = active;
    A.slot_step = step;
    A.state = state;
    A.tv_index = tv_index;
    A.chunk_steps = (int)chunk_steps;
    hipError_t e = launch_br(c, A, count, (int)count);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("blind rotation launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int set_error(int code, const char* msg) { return fail(code, msg); }
int context_device(tp_context* c) { return c->device; }

int bootstrap_into(tp_context* c, size_t count, const uint64_t* lwes, const uint64_t* tv, size_t tv_stride,
                   uint64_t* ext, uint64_t* out) {
    if (!c->d_ksk) return fail(TP_E_STATE, "key-switching key not imported");
    int rc = launch_eval(c, count, lwes, tv, tv_stride, ext, nullptr, 0, 0);
    if (rc) return rc;
    return launch_ks_rows(c, count, ext, c->N + 1, nullptr, out);
}
hipStream_t context_stream(tp_context* c) { return c->stream; }
void context_params(tp_context* c, tp_params* p) { *p = c->p; }
int context_ready(tp_context* c) { return c->d_brk && c->d_ksk; }
int bind_device(tp_context* c) { return set_device(c); }

}  // namespace tp

extern "C" {

int tp_bootstrap_batch_device(tp_context* c, size_t count, const uint64_t* lwes, const uint64_t* tv,
                              size_t tv_stride, uint64_t* out) {
    if (!c || !lwes || !tv || !out) return fail(TP_E_INVALID, "null argument");
    if (!c->d_ksk) return fail(TP_E_STATE, "key-switching key not imported");
    int rc = set_device(c);
    if (rc) return rc;
    rc = ensure_scratch(c, count * (c->N + 1));
    if (rc) return rc;
    return tp::bootstrap_into(c, count, lwes, tv, tv_stride, c->d_scratch, out);
}

struct tp_graph {
    tp_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t* d_ext = nullptr;  // extracted LWEs between the two launches, owned by the graph
};

int tp_graph_destroy(tp_graph* g) {
    if (!g) return TP_OK;
    if (g->ctx) (void)set_device(g->ctx);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_ext) (void)hipFree(g->d_ext);
    delete g;
    return TP_OK;
}

int tp_graph_bootstrap_create(tp_context* c, size_t count, const uint64_t* lwes, const uint64_t* tv,
                              size_t tv_stride, uint64_t* out, tp_graph** graph) {
    if (!c || !lwes || !tv || !out || !graph) return fail(TP_E_INVALID, "null argument");
    *graph = nullptr;
    if (!c->d_brk || !c->d_ksk) return fail(TP_E_STATE, "blind-rotation and key-switching keys must be imported");
    if (count == 0 || count > 0x7fffffff) return fail(TP_E_INVALID, "count must lie in [1, 2^31)");
    int rc = set_device(c);
    if (rc) return rc;
    tp_graph* g = new tp_graph;
    g->ctx = c;
    if (hipMalloc(&g->d_ext, count * (c->N + 1) * 8) != hipSuccess) {
        tp_graph_destroy(g);
        return fail(TP_E_NOMEM, "graph scratch allocation failed");
    }
    // capture on the context's stream: the same launches tp_bootstrap_batch_device issues (kernels
    // only, no memset or copy nodes)
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        tp_graph_destroy(g);
        return fail(TP_E_DEVICE, "stream capture could not begin");
    }
    rc = tp::bootstrap_into(c, count, lwes, tv, tv_stride, g->d_ext, out);
    const hipError_t e = hipStreamEndCapture(c->stream, &g->graph);
    if (rc) {
        tp_graph_destroy(g);
        return rc;
    }
    if (e != hipSuccess || !g->graph) {
        tp_graph_destroy(g);
        return fail(TP_E_DEVICE, std::string("stream capture: ") + hipGetErrorString(e));
    }
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        tp_graph_destroy(g);
        return fail(TP_E_DEVICE, "graph instantiation failed");
    }
    *graph = g;
    return TP_OK;
}

int tp_graph_launch(tp_graph* g) {
    if (!g || !g->exec) return fail(TP_E_INVALID, "null graph");
    int rc = set_device(g->ctx);
    if (rc) return rc;
    const hipError_t e = hipGraphLaunch(g->exec, g->ctx->stream);
    if (e != hipSuccess) return fail(TP_E_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
    return TP_OK;
}

int tp_set_option(tp_context* c, int option, int value) {
    if (!c) return fail(TP_E_INVALID, "null context");
    if (option == TP_OPT_BR_WAVES) {
        if (value != 1 && value != 2) return fail(TP_E_INVALID, "waves per SIMD must be 1 or 2");
        if (value == 1 && c->p.br_digits != 3 && c->p.br_digits != 4)
            return fail(TP_E_INVALID, "the one-wave variant is built for 3 and 4 digits");
        c->br_waves = value;
        return TP_OK;
    }
    if (option == TP_OPT_BR_SPLIT_MAX) {
        if (value < 0) return fail(TP_E_INVALID, "the digit-parallel threshold must be >= 0");
        if (value > 0 && !split_built(c))
            return fail(TP_E_INVALID, "the digit-parallel kernel is built for N = 1024 with 2..4 digits");
        c->split_max = (size_t)value;
        return TP_OK;
    }
    if (option == TP_OPT_BR_ENGINE) {
        if (value != 0 && value != 1) return fail(TP_E_INVALID, "engine must be 0 (NTT) or 1 (floating point)");
        if (value == 1 && !fft_built(c))
            return fail(TP_E_INVALID, "the floating-point product is built for binary CGGI at N = 1024 with 2..4 digits");
        c->br_engine = value;
        return TP_OK;
    }
    if (option == TP_OPT_BR_FFT_WAVES) {
        if (value != 1 && value != 2) return fail(TP_E_INVALID, "waves per SIMD must be 1 or 2");
        c->fft_waves = value;
        return TP_OK;
    }
    if (option == TP_OPT_BR_GRID) {
        if (value < 0) return fail(TP_E_INVALID, "the work-group count must be >= 0");
        c->br_grid = value;
        return TP_OK;
    }
    if (option == TP_OPT_BR_SPLIT_HALVES) {
        if (value != 0 && value != 1) return fail(TP_E_INVALID, "split halves must be 0 or 1");
        c->split_halves = value;
        return TP_OK;
    }
    if (option == TP_OPT_KS_MFMA) {
        if (value != 0 && value != 1) return fail(TP_E_INVALID, "matrix-core key switch must be 0 or 1");
        c->ks_mfma = value;
        return TP_OK;
    }
    return fail(TP_E_INVALID, "unknown option");
}

int tp_get_option(tp_context* c, int option, int* value) {
    if (!c || !value) return fail(TP_E_INVALID, "null argument");
    if (option == TP_OPT_BR_WAVES) {
        *value = c->br_waves;
        return TP_OK;
    }
    if (option == TP_OPT_BR_SPLIT_MAX) {
        *value = (int)std::min<size_t>(c->split_max, 0x7fffffff);
        return TP_OK;
    }
    if (option == TP_OPT_BR_ENGINE) {
        *value = fft_active(c) ? 1 : 0;
        return TP_OK;
    }
    if (option == TP_OPT_BR_FFT_WAVES) {
        *value = c->fft_waves;
        return TP_OK;
    }
    if (option == TP_OPT_BR_GRID) {
        *value = c->br_grid;
        return TP_OK;
    }
    if (option == TP_OPT_BR_SPLIT_HALVES) {
        *value = c->split_halves && halves_built(c) ? 1 : 0;
        return TP_OK;
    }
    if (option == TP_OPT_KS_MFMA) {
        *value = c->ks_mfma && c->d_kl ? 1 : 0;
        return TP_OK;
    }
    return fail(TP_E_INVALID, "unknown option");
}

int tp_autotune(tp_context* c, size_t count, float* ms) {
    if (!c) return fail(TP_E_INVALID, "null context");
    if (!c->d_brk) return fail(TP_E_STATE, "blind-rotation key not imported");
    if (count == 0 || count > (1u << 20)) return fail(TP_E_INVALID, "count must lie in [1, 2^20]");
    int rc = set_device(c);
    if (rc) return rc;
    const uint32_t N = c->N, n = c->p.lwe_dim;
    // synthetic inputs: LWE entries in [0, 2N) from a fixed SplitMix64 stream, zero test vector
    std::vector<uint64_t> lwes(count * (n + 1));
    uint64_t st = 0x5DEECE66DULL;
    for (auto& x : lwes) {
        uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        x = (z ^ (z >> 31)) % (2 * (uint64_t)N);
    }
    uint64_t *d_lwe = nullptr, *d_tv = nullptr, *d_out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int saved_w = c->br_waves;
    const size_t saved_split = c->split_max;
    rc = TP_OK;
    if (hipMalloc(&d_lwe, lwes.size() * 8) != hipSuccess || hipMalloc(&d_tv, 2 * (size_t)N * 8) != hipSuccess ||
        hipMalloc(&d_out, count * (N + 1) * 8) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(TP_E_NOMEM, "autotune buffers");
    }
    if (!rc && (hipMemcpyAsync(d_lwe, lwes.data(), lwes.size() * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                hipMemsetAsync(d_tv, 0, 2 * (size_t)N * 8, c->stream) != hipSuccess))
        rc = fail(TP_E_DEVICE, "autotune upload");
    // candidates: br_kernel at two and at one wave per SIMD, the digit-parallel kernel, then the
    // floating-point kernel at one and at two waves per SIMD
    // the floating-point engine is a candidate only for a context that opted in to it
    const bool fft_candidate = c->br_engine == 1 && c->fft_ready;
    const int saved_engine = c->br_engine, saved_fw = c->fft_waves;
    const int saved_grid = c->br_grid;
    const int saved_halves = c->split_halves;
    float t_var[6] = {-1.0f, -1.0f, -1.0f, -1.0f, -1.0f, -1.0f};
    // a warm-up launch, then the faster of two timed launches (one timing flips close variants)
    auto time_launch = [&](float& t) {
        rc = launch_eval(c, count, d_lwe, d_tv, 0, d_out, nullptr, 0, 0);
        t = -1.0f;
        for (int r = 0; r < 2 && !rc; ++r) {
            float tr = -1.0f;
            if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(TP_E_DEVICE, "event record");
            if (!rc) rc = launch_eval(c, count, d_lwe, d_tv, 0, d_out, nullptr, 0, 0);
            if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&tr, e0, e1) != hipSuccess))
                rc = fail(TP_E_DEVICE, "autotune timing");
            if (!rc && (t < 0 || tr < t)) t = tr;
        }
    };
    c->br_grid = 0;
    for (int v = 0; v < 5 && !rc; ++v) {
        if (v == 1 && c->p.br_digits != 3 && c->p.br_digits != 4) continue;
        if (v == 2 && !split_built(c)) continue;
        if (v >= 3 && !fft_candidate) continue;
        c->br_waves = v == 1 ? 1 : 2;
        c->split_max = v == 2 ? count : 0;
        c->br_engine = v >= 3 ? 1 : 0;
        c->fft_waves = v == 4 ? 2 : 1;
        float t = -1.0f;
        if (v == 2 && halves_built(c)) {
            // the digit-parallel kernel in both forms: two waves per transform, then one; the
            // faster is kept (two waves win at few ciphertexts, where SIMDs would idle)
            float t_one_wave = -1.0f;
            c->split_halves = 1;
            time_launch(t);
            if (!rc) {
                c->split_halves = 0;
                time_launch(t_one_wave);
            }
            if (!rc) {
                c->split_halves = (t_one_wave >= 0 && t_one_wave < t) ? 0 : 1;
                if (c->split_halves == 0) t = t_one_wave;
            }
        } else {
            time_launch(t);
        }
        t_var[v] = t;
    }
    if (rc) {
        c->br_waves = saved_w;
        c->split_max = saved_split;
        c->br_engine = saved_engine;
        c->fft_waves = saved_fw;
        c->br_grid = saved_grid;
        c->split_halves = saved_halves;
    } else {
        c->br_waves = (t_var[1] >= 0 && t_var[1] < t_var[0]) ? 1 : 2;
        float t_one = c->br_waves == 1 ? t_var[1] : t_var[0];
        c->br_engine = 0;
        c->fft_waves = (t_var[4] >= 0 && t_var[4] < t_var[3]) ? 2 : 1;
        const float t_fft = c->fft_waves == 2 ? t_var[4] : t_var[3];
        if (t_fft >= 0 && t_fft < t_one) {
            c->br_engine = 1;
            t_one = t_fft;
        }
        // the winner on a persistent grid (work-groups = the ciphertexts resident at once, each
        // striding over the launch) when the launch holds more ciphertexts than that
        const int slots = c->cu_count * 4 * (c->br_engine ? c->fft_waves : c->br_waves);
        if (count > (size_t)slots) {
            c->split_max = 0;
            c->br_grid = slots;
            float t = -1.0f;
            time_launch(t);
            t_var[5] = t;
            c->br_grid = 0;
        }
        if (rc) {
            c->br_waves = saved_w;
            c->split_max = saved_split;
            c->br_engine = saved_engine;
            c->fft_waves = saved_fw;
            c->br_grid = saved_grid;
            c->split_halves = saved_halves;
        } else {
            if (t_var[5] >= 0 && t_var[5] < t_one) {
                c->br_grid = slots;
                t_one = t_var[5];
            }
            // the digit-parallel kernel wins at small counts: move the threshold to this count's side
            if (t_var[2] >= 0 && t_var[2] < t_one) c->split_max = std::max(saved_split, count);
            else c->split_max = std::min(saved_split, count - 1);
        }
    }
    if (ms)
        for (int v = 0; v < 6; ++v) ms[v] = t_var[v];
    (void)hipStreamSynchronize(c->stream);
    for (void* p : {(void*)d_lwe, (void*)d_tv, (void*)d_out})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

int tp_autotune_cached(tp_context* c, size_t count, const char* path, int* hit, float* ms) {
    if (!c || !path) return fail(TP_E_INVALID, "null argument");
    if (hit) *hit = 0;
    std::vector<tp::TuneEntry> entries;
    std::string err;
    if (tp::read_tuning_cache(path, entries, err)) return fail(TP_E_INVALID, err);
    tp::TuneEntry key;
    key.arch = c->arch;
    key.cus = c->cu_count;
    key.N = c->N;
    key.n = c->p.lwe_dim;
    key.b = c->p.br_base_log2;
    key.l = c->p.br_digits;
    key.method = c->p.method;
    key.count = count;
    key.fft_opt_in = (c->br_engine == 1 && c->fft_ready) ? 1 : 0;
    for (const tp::TuneEntry& e : entries) {
        if (!e.same_key(key)) continue;
        // stored settings, checked against what this build and context allow
        if ((e.waves != 1 && e.waves != 2) || (e.waves == 1 && c->p.br_digits != 3 && c->p.br_digits != 4) ||
            (e.split_max > 0 && !split_built(c)) || (e.engine == 1 && !c->fft_ready) || e.grid < 0 ||
            (e.fft_waves != 1 && e.fft_waves != 2) || (e.split_halves != 0 && e.split_halves != 1))
            return fail(TP_E_INVALID, "tuning cache entry does not fit this context");
        c->br_waves = e.waves;
        c->split_max = (size_t)e.split_max;
        c->br_engine = e.engine;
        c->fft_waves = e.fft_waves;
        c->br_grid = e.grid;
        c->split_halves = e.split_halves;
        if (ms)
            for (int v = 0; v < 6; ++v) ms[v] = v < (int)e.ms.size() ? e.ms[v] : -1.0f;
        if (hit) *hit = 1;
        return TP_OK;
    }
    float t[6];
    int rc = tp_autotune(c, count, t);
    if (rc) return rc;
    key.waves = c->br_waves;
    key.split_max = c->split_max;
    key.engine = c->br_engine;
    key.fft_waves = c->fft_waves;
    key.grid = c->br_grid;
    key.split_halves = c->split_halves;
    key.ms.assign(t, t + 6);
    if (ms)
        for (int v = 0; v < 6; ++v) ms[v] = t[v];
    entries.push_back(key);
    if (tp::write_tuning_cache(path, entries, err)) return fail(TP_E_INVALID, err);
    return TP_OK;
}

int tp_malloc(tp_context* c, uint64_t** ptr, size_t words) {
    if (!c || !ptr) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    if (hipMalloc(ptr, words ? words * 8 : 8) != hipSuccess) return fail(TP_E_NOMEM, "device allocation failed");
    return TP_OK;
}

int tp_free(tp_context* c, uint64_t* ptr) {
    if (!c) return fail(TP_E_INVALID, "null context");
    int rc = set_device(c);
    if (rc) return rc;
    TP_HIP(hipStreamSynchronize(c->stream));
    if (ptr) TP_HIP(hipFree(ptr));
    return TP_OK;
}

int tp_copy_to_device(tp_context* c, uint64_t* dst, const uint64_t* src, size_t words) {
    if (!c || !dst || !src) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    TP_HIP(hipMemcpyAsync(dst, src, words * 8, hipMemcpyHostToDevice, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));
    return TP_OK;
}

int tp_copy_to_host(tp_context* c, uint64_t* dst, const uint64_t* src, size_t words) {
    if (!c || !dst || !src) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    TP_HIP(hipMemcpyAsync(dst, src, words * 8, hipMemcpyDeviceToHost, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));
    return TP_OK;
}

int tp_copy_on_device(tp_context* c, uint64_t* dst, const uint64_t* src, size_t words) {
    if (!c || !dst || !src) return fail(TP_E_INVALID, "null argument");
    int rc = set_device(c);
    if (rc) return rc;
    TP_HIP(hipMemcpyAsync(dst, src, words * 8, hipMemcpyDeviceToDevice, c->stream));
    return TP_OK;
}

int tp_timer_mark(tp_context* c, int slot) {
    if (!c || slot < 0 || slot >= 8) return fail(TP_E_INVALID, "bad timer slot");
    TP_HIP(hipEventRecord(c->ev[slot], c->stream));
    return TP_OK;
}

int tp_timer_elapsed(tp_context* c, int a, int b, float* ms) {
    if (!c || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return fail(TP_E_INVALID, "bad timer slot");
    TP_HIP(hipEventSynchronize(c->ev[b]));
    TP_HIP(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return TP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ SAP digit decomposition
//
// SAPDecomposer (lib/src/operators/sap_decomposer.cpp:190-494) over a batch: every step of the
// reference runs for all inputs at once (one blind-rotation launch per digit, one automorphism
// key-switch launch per node of the HomTrunc tree).  RLWEs are [a|b]; a batch of them is 2 count
// rows of N words, so the per-row polynomial products broadcast one N-word polynomial.

namespace {

uint32_t int_log2_ceil(uint64_t x) {  // IntLog2 (utils/math_utils.cpp:11-25)
    if (x <= 1) return 0;
    uint32_t r = 0;
    for (x -= 1; x; x >>= 1) ++r;
    return r;
}

uint32_t bit_reverse(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < bits; ++b) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

// Polynomials that depend on (N, radix), NTT domain, in this order:
//   xp      x -> x mod radix extraction, reversed: xp[N - i] = -(i & (radix - 1))   (:412-416)
//   trunc   N^-1 sum_{i < radix} X^-i                                               (:220-233)
//   shift   [logN + 1] level l: -X^(N - N / 2^l)                                    (:280-285)
//   leaf    [N / radix] exponent e: X^brv(e) over log2(N / radix) bits              (:266-268)
//   odd     [N / (2 radix)] e: shift[1] * leaf[2e + 1], the odd leaf of a level-1 node with its shift
size_t sap_poly_words(uint32_t N, uint32_t logN, uint32_t radix) {
    return (size_t)N * (3 + logN + N / radix + N / radix / 2);
}

int sap_prepare(tp_context* c, uint32_t radix) {
    if (c->sap_radix == radix && c->d_sap_polys) return TP_OK;
    const uint32_t N = c->N, logN = (uint32_t)c->logN, lr = int_log2_ceil(radix);
    const size_t words = sap_poly_words(N, logN, radix);
    std::vector<uint64_t> h(words, 0);
    tp::HostNtt T(N);
    uint64_t* xp = h.data();
    uint64_t* trunc = xp + N;
    uint64_t* shift = trunc + N;
    uint64_t* leaf = shift + (size_t)(logN + 1) * N;
    for (uint32_t i = 1; i < N; ++i) xp[N - i] = gl::neg(i & (radix - 1));
    T.forward(xp);
    const uint64_t n_inv = gl::inv(N);
    trunc[0] = n_inv;
    for (uint32_t i = N - (radix - 1); i < N; ++i) trunc[i] = gl::neg(n_inv);
    T.forward(trunc);
    for (uint32_t l = 1; l <= logN; ++l) {
        uint64_t* p = shift + (size_t)l * N;
        p[N - (N >> l)] = gl::kQ - 1;
        T.forward(p);
    }
    const uint32_t rev_bits = logN - lr;
    for (uint32_t e = 0; e < (N >> lr); ++e) {
        uint64_t* p = leaf + (size_t)e * N;
        p[rev_bits ? bit_reverse(e, rev_bits) : 0] = 1;
        T.forward(p);
    }
    uint64_t* odd = leaf + (size_t)(N >> lr) * N;
    for (uint32_t e = 0; e < (N >> lr) / 2; ++e)
        for (uint32_t j = 0; j < N; ++j)
            odd[(size_t)e * N + j] = gl::mul(shift[N + j], leaf[(size_t)(2 * e + 1) * N + j]);
    if (c->d_sap_polys) TP_HIP(hipFree(c->d_sap_polys));
    c->d_sap_polys = nullptr;
    c->sap_radix = 0;
    if (hipMalloc(&c->d_sap_polys, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "SAP polynomial allocation failed");
    TP_HIP(hipMemcpy(c->d_sap_polys, h.data(), words * 8, hipMemcpyHostToDevice));
    c->sap_radix = radix;
    return TP_OK;
}

// Workspace carve-up for `count` inputs.
struct SapWork {
    uint64_t *cp, *sub, *cm, *acc, *tv, *ext, *slots, *rext, *rlwe_n, *rpoly, *fresh;
    uint32_t* digits;
};

int sap_workspace(tp_context* c, size_t count, uint32_t lr, SapWork& w) {
    const size_t N = c->N, n1 = (size_t)c->p.lwe_dim + 1, k1 = (size_t)c->p.ks_out_dim + 1;
    const size_t slots = 1;  // the HomTrunc root (the tree levels live in d_sapt)
    const size_t rl = 2 * N * count;
    const size_t words = 2 * n1 * count + 2 * count + 3 * rl + slots * rl + (N + 1) * count + k1 * count + 3 * N;
    if (c->sapw_words < words) {
        if (c->d_sapw) TP_HIP(hipFree(c->d_sapw));
        c->d_sapw = nullptr;
        c->sapw_words = 0;
        if (hipMalloc(&c->d_sapw, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "SAP workspace allocation failed");
        c->sapw_words = words;
    }
    uint64_t* p = c->d_sapw;
    w.cp = p; p += n1 * count;
    w.sub = p; p += n1 * count;
    w.cm = p; p += count;
    w.digits = reinterpret_cast<uint32_t*>(p); p += count;
    w.acc = p; p += rl;
    w.tv = p; p += rl;
    w.ext = p; p += rl;
    w.slots = p; p += slots * rl;
    w.rext = p; p += (N + 1) * count;
    w.rlwe_n = p; p += k1 * count;
    w.rpoly = p; p += N;
    w.fresh = p;
    return TP_OK;
}

// Q/(2N) X^0 as a coefficient-domain accumulator [0 | Q/(2N), 0, ...] (:405-407, :368).
int sap_upload_fresh(tp_context* c, const SapWork& w) {
    std::vector<uint64_t> f(2 * (size_t)c->N, 0);
    f[c->N] = gl::kQ / (2 * (uint64_t)c->N);
    TP_HIP(hipMemcpyAsync(w.fresh, f.data(), f.size() * 8, hipMemcpyHostToDevice, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));  // f is a stack buffer
    return TP_OK;
}

#define TP_LAUNCH(what, call)                                                                        \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return fail(TP_E_DEVICE, std::string(what ": ") + hipGetErrorString(e_)); \
    } while (0)

// HomTrunc (:190-312) of count RLWEs `in` (NTT) into `out` (NTT); out may equal in.
// RLWEs per HomTrunc tree buffer: inputs are taken in passes so that a pass's level-1 nodes fit
// (2 buffers of 2^16 RLWEs: 8 GiB at N = 4096, 2 GiB at N = 1024).
constexpr size_t kSapTreeRlwes = (size_t)1 << 16;

int sap_homtrunc(tp_context* c, const SapWork& w, size_t count, uint32_t radix, const uint64_t* in, uint64_t* out) {
    const uint32_t N = c->N, logN = (uint32_t)c->logN, lr = int_log2_ceil(radix);
    const size_t rows = 2 * count, rl = 2 * (size_t)N * count;
    const uint64_t* trunc = c->d_sap_polys + N;
    const uint64_t* shift = trunc + N;
    const uint64_t* leaf = shift + (size_t)(logN + 1) * N;
    const uint64_t* odd = leaf + (size_t)(N >> lr) * N;
    const size_t key_words = (size_t)c->p.rlwe_ks_digits * 2 * N;
    const tp::KeySwitchArgs A = auto_args(c);
    auto slot = [&](uint32_t h) { return w.slots + (size_t)h * rl; };
    TP_LAUNCH("SAP truncation product", tp::launch_sap_poly_mul(slot(0), in, trunc, rows, N, c->stream));
    // partial trace over t = N + 1, N/2 + 1, ..., N/2^(lr-1) + 1 (:236-250)
    // (each automorphism takes its input in the NTT domain: the reference's BackwardNTT before EvalAuto
    // happens in registers for a, and b is permuted in the evaluation domain)
    for (uint32_t i = 0; i < lr; ++i) {
        const uint32_t j = logN - i;
        TP_LAUNCH("SAP automorphism", tp::launch_auto_ntt(c->logN, A, slot(0), slot(0), (int)count, (1u << j) + 1, 1,
                                                          c->d_auto + (size_t)(j - 1) * key_words, c->stream));
    }
    // the clean-up tree (:266-311), level by level: level l holds the 2^(L0 - l) nodes of every input,
    // node e having children 2e (x + y) and 2e + 1 ((x - y) X^-(N >> l)); the leaf sum is exact in
    // Z_Q, so this order gives the reference's depth-first stack walk's result
    const uint32_t L0 = logN - lr;
    if (L0 == 0) {
        TP_LAUNCH("SAP leaf", tp::launch_sap_leaf(out, slot(0), leaf, rows, N, 0, c->stream));
        return TP_OK;
    }
    const size_t deep = (size_t)1 << (L0 - 1);  // nodes of level 1
    const size_t group = std::min(count, std::max<size_t>(1, kSapTreeRlwes / deep));  // inputs per pass
    const size_t buf_words = deep * group * 2 * N;
    if (c->sapt_words < 2 * buf_words) {
        if (c->d_sapt) TP_HIP(hipFree(c->d_sapt));
        c->d_sapt = nullptr;
        c->sapt_words = 0;
        if (hipMalloc(&c->d_sapt, 2 * buf_words * 8) != hipSuccess)
            return fail(TP_E_NOMEM, "SAP tree workspace allocation failed");
        c->sapt_words = 2 * buf_words;
    }
    for (size_t g0 = 0; g0 < count; g0 += group) {
        const int gc = (int)std::min(group, count - g0);
        const uint64_t* X = slot(0) + g0 * 2 * N;  // the root: one node per input
        int nodes = 1, which = 0;
        for (uint32_t level = L0; level >= 1; --level) {
            uint64_t* dst = c->d_sapt + (size_t)which * buf_words;
            const uint64_t* key = c->d_auto + (size_t)(level - 1) * key_words;
            if (level > 1) {
                TP_LAUNCH("SAP tree level", tp::launch_sap_tree(c->logN, A, X, dst, shift + (size_t)level * N, nullptr,
                                                                nodes, gc, (1u << level) + 1, 0, key, c->stream));
                nodes *= 2;
            } else {
                TP_LAUNCH("SAP tree leaves", tp::launch_sap_tree(c->logN, A, X, dst, leaf, odd, nodes, gc,
                                                                 (1u << level) + 1, 1, key, c->stream));
            }
            X = dst;
            which ^= 1;
        }
        TP_LAUNCH("SAP leaf sum", tp::launch_sap_sum_nodes(out + g0 * 2 * N, X, nodes, 2 * (size_t)gc, N, 0,
                                                           c->stream));
    }
    return TP_OK;
}

// ResetAccumulator (:314-373) of count accumulators (NTT, in place).  The key switch is mod Q,
// so the reference's optional ROUND switch to the key-switching modulus (:358-360) is the identity.
int sap_reset(tp_context* c, const SapWork& w, size_t count, uint64_t radix, uint64_t* acc) {
    const uint32_t N = c->N, rlog = int_log2_ceil(radix);
    const size_t rows = 2 * count;
    std::vector<uint64_t> poly(N, 0);
    for (uint32_t i = 1; i < N; ++i) poly[i] = gl::neg((uint64_t)(((N - i) >> rlog) << rlog) % gl::kQ);
    tp::HostNtt(N).forward(poly.data());
    TP_HIP(hipMemcpyAsync(w.rpoly, poly.data(), N * 8, hipMemcpyHostToDevice, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));
    TP_LAUNCH("SAP reset product", tp::launch_sap_poly_mul(w.tv, acc, w.rpoly, rows, N, c->stream));
    TP_LAUNCH("SAP inverse NTT", tp::launch_ntt_batch(c->logN, w.ext, w.tv, rows, c->d_itw, true, c->stream));
    TP_LAUNCH("SAP extraction", tp::launch_sap_extract(w.rext, w.ext, nullptr, 0, (size_t)N + 1, (int)count, N,
                                                       c->stream));
    int rc = tp::launch_ks_rows(c, count, w.rext, (size_t)N + 1, nullptr, w.rlwe_n);
    if (rc) return rc;
    const size_t k1 = (size_t)c->p.ks_out_dim + 1;
    TP_LAUNCH("SAP modulus switch", tp::launch_modswitch(w.rlwe_n, w.rlwe_n, k1 * count, gl::kQ, c->logN + 1,
                                                         TP_MS_FLOOR, 0, c->stream));
    return launch_eval(c, count, w.rlwe_n, w.fresh, 0, nullptr, acc, 1, 0);
}

int sap_check(tp_context* c, uint32_t radix, bool need_brk, bool need_auto, bool need_ksk) {
    if (radix < 2 || radix > c->N || (radix & (radix - 1))) return fail(TP_E_INVALID, "radix must be a power of two in [2, N]");
    if (need_brk && !c->d_brk) return fail(TP_E_STATE, "blind-rotation key not imported");
    if (need_auto && !c->d_auto) return fail(TP_E_STATE, "automorphism keys not imported");
    if (need_ksk) {
        if (!c->d_ksk) return fail(TP_E_STATE, "key-switching key not imported");
        if (c->p.ks_out_dim != c->p.lwe_dim) return fail(TP_E_INVALID, "reset needs ks_out_dim == lwe_dim");
    }
    return TP_OK;
}

}  // namespace

extern "C" {

int tp_sap_homtrunc_device(tp_context* c, size_t count, uint32_t radix, const uint64_t* in, uint64_t* out) {
    if (!c || !in || !out) return fail(TP_E_INVALID, "null argument");
    int rc = sap_check(c, radix, false, true, false);
    if (rc) return rc;
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    if ((rc = set_device(c)) || (rc = sap_prepare(c, radix))) return rc;
    SapWork w{};
    if ((rc = sap_workspace(c, count, int_log2_ceil(radix), w))) return rc;
    return sap_homtrunc(c, w, count, radix, in, out);
}

int tp_sap_reset_device(tp_context* c, size_t count, uint64_t radix, uint64_t* acc) {
    if (!c || !acc) return fail(TP_E_INVALID, "null argument");
    int rc = sap_check(c, 2, true, false, true);
    if (rc) return rc;
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    if ((rc = set_device(c))) return rc;
    SapWork w{};
    if ((rc = sap_workspace(c, count, 1, w)) || (rc = sap_upload_fresh(c, w))) return rc;
    return sap_reset(c, w, count, radix, acc);
}

int tp_sap_decompose_device(tp_context* c, size_t count, uint32_t radix, uint32_t beta, uint64_t restart,
                            const uint64_t* lwes, uint64_t* out, uint32_t max_digits, uint32_t* digits) {
    if (!c || !lwes || !out) return fail(TP_E_INVALID, "null argument");
    int rc = sap_check(c, radix, true, true, restart != UINT64_MAX && restart != 0);
    if (rc) return rc;
    if (restart == 0) return fail(TP_E_INVALID, "restart period must be positive (UINT64_MAX: never)");
    if (beta > c->p.lwe_dim) return fail(TP_E_INVALID, "beta (key Hamming weight) exceeds the LWE dimension");
    if (count == 0) return TP_OK;
    if (count > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    if ((rc = set_device(c)) || (rc = sap_prepare(c, radix))) return rc;
    const uint32_t N = c->N, n = c->p.lwe_dim, lr = int_log2_ceil(radix);
    SapWork w{};
    if ((rc = sap_workspace(c, count, lr, w)) || (rc = sap_upload_fresh(c, w))) return rc;
    const size_t rows = 2 * count;
    TP_HIP(hipMemcpyAsync(w.cp, lwes, (size_t)(n + 1) * count * 8, hipMemcpyDeviceToDevice, c->stream));
    TP_LAUNCH("SAP init", tp::launch_sap_init(lwes, w.cm, w.digits, (int)n, (int)count, (int)lr, c->stream));
    std::vector<uint32_t> dig(count);
    TP_HIP(hipMemcpyAsync(dig.data(), w.digits, count * 4, hipMemcpyDeviceToHost, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));
    uint32_t D = 0;
    for (uint32_t d : dig) D = d > D ? d : D;
    if (D > max_digits)
        return fail(TP_E_INVALID, "an input needs " + std::to_string(D) + " digits, more than max_digits");
    if (digits) TP_HIP(hipMemcpyAsync(digits, w.digits, count * 4, hipMemcpyDeviceToDevice, c->stream));
    const size_t out_stride = (size_t)max_digits * (N + 1);
    const uint64_t* xp = c->d_sap_polys;
    for (uint32_t k = 0; k < D; ++k) {
        TP_LAUNCH("SAP digit", tp::launch_sap_digit(w.cp, w.cm, w.sub, (int)n, (int)count, (int)lr, radix - 1,
                                                    (uint64_t)radix * beta, beta, 2 * (uint64_t)N, c->stream));
        // the accumulator enters the rotation in coefficients; the first one is Q/(2N) X^0 itself
        const uint64_t* tv = w.fresh;
        size_t tv_stride = 0;
        if (k > 0) {
            TP_LAUNCH("SAP inverse NTT", tp::launch_ntt_batch(c->logN, w.tv, w.acc, rows, c->d_itw, true, c->stream));
            tv = w.tv;
            tv_stride = 2 * (size_t)N;
        }
        if ((rc = launch_eval(c, count, w.sub, tv, tv_stride, nullptr, w.acc, 1, 0))) return rc;
        TP_LAUNCH("SAP extraction product", tp::launch_sap_poly_mul(w.ext, w.acc, xp, rows, N, c->stream));
        TP_LAUNCH("SAP inverse NTT", tp::launch_ntt_batch(c->logN, w.tv, w.ext, rows, c->d_itw, true, c->stream));
        TP_LAUNCH("SAP extraction", tp::launch_sap_extract(out, w.tv, w.digits, k, out_stride, (int)count, N,
                                                           c->stream));
        if (k + 1 == D) break;  // the accumulator is not read again
        if (restart != UINT64_MAX && (k + 1) % restart == 0) {
            // :480-482 resets with radix 0 (inputs whose modulus reached 0 have no further digits)
            if ((rc = sap_reset(c, w, count, 0, w.acc))) return rc;
        }
        if ((rc = sap_homtrunc(c, w, count, radix, w.acc, w.acc))) return rc;
    }
    return TP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ LWE -> RGSW scheme switching
//
// LWEtoRGSWConverter::Convert (lib/src/operators/lwe_to_rgsw_conversion.cpp:163-231) over a batch:
// one blind rotation launch with the staircase accumulator, the digit rows of every input in one
// pass, one trace launch over all rows, one inverse NTT and one s^2 -> s key-switch launch, and the
// left rows in one pass.

extern "C" {

int tp_lwe_to_rgsw_device(tp_context* c, size_t count, uint32_t digits, uint32_t basis_log2, const uint64_t* lwes,
                          uint64_t* out) {
    if (!c || !lwes || !out) return fail(TP_E_INVALID, "null argument");
    if (!c->d_brk) return fail(TP_E_STATE, "blind-rotation key not imported");
    if (!c->d_auto) return fail(TP_E_STATE, "automorphism (trace) keys not imported");
    if (!c->d_rksk) return fail(TP_E_STATE, "squaring key (tp_import_rlwe_ksk) not imported");
    const uint32_t N = c->N;
    if (basis_log2 == 0 || basis_log2 > 63) return fail(TP_E_INVALID, "basis_log2 must lie in [1, 63]");
    const uint32_t D = 64 / basis_log2 + (64 % basis_log2 ? 1 : 0);
    if (digits == 0 || N % digits || digits > D) return fail(TP_E_INVALID, "digits must divide N and not exceed ceil(64 / basis_log2)");
    if (count == 0) return TP_OK;
    if (count * digits > 0x7fffffff) return fail(TP_E_INVALID, "count too large");
    int rc = set_device(c);
    if (rc) return rc;
    const uint32_t block = N / digits;
    // accumulator (:116-161): staircase t of the output gadget, centred on 0; acc = [0 | -t/2], extra = t/2
    std::vector<uint64_t> host(3 * (size_t)N, 0), t(N, 0);
    uint64_t scale = 1ULL << (basis_log2 * (D - digits));
    for (uint32_t i = 0; i < block / 2; ++i) {
        t[i] = scale;
        t[N - i - 1] = gl::kQ - scale;
    }
    for (uint32_t d = 1; d < digits; ++d) {
        scale <<= basis_log2;
        for (uint32_t j = 0; j < block; ++j) t[d * block + j - block / 2] = scale;
    }
    const uint64_t inv2 = gl::inv(2);
    for (uint32_t i = 0; i < N; ++i) {
        host[2 * N + i] = gl::mul(t[i], inv2);
        host[N + i] = gl::neg(host[2 * N + i]);
    }
    const size_t rows = count * digits, rw = rows * 2 * N;
    const size_t words = 3 * (size_t)N + count * 2 * N + 3 * rw;
    if (c->rgsw_words < words) {
        if (c->d_rgsw) TP_HIP(hipFree(c->d_rgsw));
        c->d_rgsw = nullptr;
        c->rgsw_words = 0;
        if (hipMalloc(&c->d_rgsw, words * 8) != hipSuccess) return fail(TP_E_NOMEM, "LWE->RGSW workspace allocation failed");
        c->rgsw_words = words;
    }
    uint64_t* acc = c->d_rgsw;
    uint64_t* extra = acc + 2 * N;
    uint64_t* row0 = acc + 3 * N;
    uint64_t* A = row0 + count * 2 * N;  // digit rows (coefficients), then their inverse transform
    uint64_t* B = A + rw;                // right rows (NTT)
    uint64_t* C = B + rw;                // left rows
    TP_HIP(hipMemcpyAsync(acc, host.data(), 3 * (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    TP_HIP(hipStreamSynchronize(c->stream));  // host is a local buffer
    if ((rc = launch_eval(c, count, lwes, acc, 0, nullptr, row0, 0, 0))) return rc;
    TP_LAUNCH("RGSW digit rows", tp::launch_rgsw_expand(A, row0, extra, (int)count, digits, block, N, c->stream));
    TP_LAUNCH("RGSW trace", tp::launch_trace(c->logN, auto_args(c), A, B, (int)rows, c->d_auto, c->stream));
    TP_LAUNCH("RGSW inverse NTT", tp::launch_ntt_batch(c->logN, A, B, 2 * rows, c->d_itw, true, c->stream));
    // s^2 -> s switch of (a, 0): mode 2 with t = 1 ignores b (:221-224)
    TP_LAUNCH("RGSW squaring switch", tp::launch_rlwe_keyswitch(c->logN, auto_args(c), A, C, (int)rows, 1,
                                                                 TP_RKS_SWITCH_ONLY, 0, c->d_rksk, c->stream));
    TP_LAUNCH("RGSW left rows", tp::launch_rgsw_left(C, B, rows, N, c->stream));
    const size_t half = (size_t)digits * 2 * N * 8, pitch = 2 * half;
    TP_HIP(hipMemcpy2DAsync(out, pitch, C, half, half, count, hipMemcpyDeviceToDevice, c->stream));
    TP_HIP(hipMemcpy2DAsync(out + (size_t)digits * 2 * N, pitch, B, half, half, count, hipMemcpyDeviceToDevice,
                            c->stream));
    return TP_OK;
}

}  // extern "C"
