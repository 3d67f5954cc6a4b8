// C++ drop-in classes (include/toothpaste_amd.hpp) over the C-ABI: every operator owns a device
// context with its keys resident in HBM and runs its calls through the gfx950 kernels.  Host code
// here only validates, stages caller buffers in and out of device memory and prepares key
// material (seeded key generation, the automorphism of a secret key).
#include "../../include/toothpaste_amd.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>

#include "gl64.hpp"

namespace {

constexpr uint64_t kQ = gl::kQ;

[[noreturn]] void raise_status(int rc, const std::string& where) {
    const std::string msg = where + ": " + tp_last_error();
    if (rc == TP_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void check(int rc, const char* where) {
    if (rc != TP_OK) raise_status(rc, where);
}

void require(bool ok, const std::string& msg) {
    if (!ok) throw std::invalid_argument(msg);
}

// ceil(log2 x) (0 for x <= 1), the reference's IntLog2 (utils/math_utils.cpp:11-26): exact for the
// power-of-two bases and dimensions, and 64 for the Goldilocks modulus in the noise estimates.
uint64_t int_log2(uint64_t x) {
    if (x <= 1) return 0;
    uint64_t r = 0;
    for (x -= 1; x; x >>= 1) ++r;
    return r;
}

void require_goldilocks(uint64_t modulus) {
    require(modulus == kQ, "only the Goldilocks modulus 2^64 - 2^32 + 1 runs on this GPU path");
}

void require_ring(uint64_t N) {
    require(N == 1024 || N == 2048 || N == 4096, "ring dimension must be 1024, 2048 or 4096");
}

void check_lwe_entries(const uint64_t* lwe, uint64_t words, uint64_t N) {
    for (uint64_t i = 0; i < words; ++i) require(lwe[i] < 2 * N, "lwe entries must lie in [0, 2N)");
}

// Device staging buffer of a context.
struct DevBuf {
    tp_context* c;
    uint64_t* p = nullptr;
    DevBuf(tp_context* ctx, size_t words) : c(ctx) { check(tp_malloc(c, &p, words ? words : 1), "device buffer"); }
    ~DevBuf() {
        if (p) tp_free(c, p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void put(const uint64_t* src, size_t words) { check(tp_copy_to_device(c, p, src, words), "copy to device"); }
    void get(uint64_t* dst, size_t words) const { check(tp_copy_to_host(c, dst, p, words), "copy to host"); }
};

tp_params base_params(uint64_t N) {
    tp_params p{};
    p.ring_dim = (uint32_t)N;
    p.lwe_dim = 2;
    p.br_base_log2 = 16;
    p.br_digits = 4;
    p.method = TP_METHOD_CGGI_BINARY;
    return p;
}

void auto_poly(uint64_t* out, const uint64_t* in, uint64_t t, uint64_t N) {
    // EvalNegacyclicAutomorphism (utils/math_utils.cpp:70-85): out[(i t) mod N] = +-in[i]
    for (uint64_t i = 0; i < N; ++i) {
        const uint64_t e = (i * t) & (2 * N - 1);
        out[e & (N - 1)] = e >= N ? gl::neg(in[i] % kQ) : in[i] % kQ;
    }
}

// Output variance of an RGSW x RLWE product chain (cggi_blind_rotation.cpp:86-124 and
// rlwe_conversion.cpp:55-87 use the same per-product term): digits * N * basis^2 * std^2 / 12
// plus the approximation error delta^2/12 * |message|^2, delta = 2^log2(Q) / basis^digits.
long double product_variance(uint64_t modulus, uint64_t N, uint64_t basis, uint64_t digits, double std,
                             KeyDistribution dist) {
    long double delta = std::pow(2.0L, (long double)int_log2(modulus)) / std::pow((long double)basis, (long double)digits);
    if (delta <= 1) delta = 0.0L;
    delta *= delta / 12.0L;
    const long double norm2 = dist == GAUSSIAN ? 25.0L * N : (long double)N;
    long double v = (long double)digits * N;
    v *= (long double)basis * basis;
    v *= (long double)std * std;
    v /= 12.0L;
    return v + delta * norm2;
}

}  // namespace

namespace toothpaste_amd {

DeviceContext make_device_context(const tp_params& p) {
    tp_context* c = nullptr;
    check(tp_context_create(&c, &p, 0), "tp_context_create");
    return DeviceContext(c, [](tp_context* x) { tp_context_destroy(x); });
}

uint64_t fresh_seed() {
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ (uint64_t)rd();
}

}  // namespace toothpaste_amd

using toothpaste_amd::DeviceContext;
using toothpaste_amd::make_device_context;

// ------------------------------------------------------------------ MathWorker on the GPU

namespace {

class GpuSolinasWorker final : public MathWorker {
public:
    // The device context is created on first use, so contexts (and their noise estimates) can be
    // built on a host without a GPU; every transform or elementwise call needs one.
    explicit GpuSolinasWorker(uint64_t dimension) : m_N(dimension) {
        if (dimension) require_ring(dimension);
    }
    void ForwardNTT(uint64_t* result, const uint64_t* input) override { Transform(result, input, false); }
    void ForwardNTT(std::vector<uint64_t>& result, std::vector<uint64_t>& input) override {
        Transform(result.data(), input.data(), false);
    }
    void BackwardNTT(uint64_t* result, const uint64_t* input) override { Transform(result, input, true); }
    void BackwardNTT(std::vector<uint64_t>& result, std::vector<uint64_t>& input) override {
        Transform(result.data(), input.data(), true);
    }
    uint64_t AddMod(uint64_t l, uint64_t r) override { return gl::add(l, r); }
    uint64_t SubMod(uint64_t l, uint64_t r) override { return gl::sub(l, r); }
    uint64_t MulMod(uint64_t l, uint64_t r) override { return gl::mul(l, r); }
    uint64_t FMAMod(uint64_t l, uint64_t r, uint64_t s) override { return gl::add(gl::mul(l, r), s); }
    uint64_t InvMod(uint64_t x) override { return gl::inv(x % kQ); }
    uint64_t PowMod(uint64_t b, uint64_t e) override { return gl::pow(b % kQ, e); }
    void AddModElw(uint64_t* res, const uint64_t* l, const uint64_t* r, uint64_t len) override {
        Elw(TP_ELT_ADD, res, l, r, nullptr, 0, len);
    }
    void AddEqModElw(uint64_t* lo, const uint64_t* r, uint64_t len) override { Elw(TP_ELT_ADD, lo, lo, r, nullptr, 0, len); }
    void SubModElw(uint64_t* res, const uint64_t* l, const uint64_t* r, uint64_t len) override {
        Elw(TP_ELT_SUB, res, l, r, nullptr, 0, len);
    }
    void SubEqModElw(uint64_t* lo, const uint64_t* r, uint64_t len) override { Elw(TP_ELT_SUB, lo, lo, r, nullptr, 0, len); }
    void MulModElw(uint64_t* res, const uint64_t* l, const uint64_t* r, uint64_t len) override {
        Elw(TP_ELT_MUL, res, l, r, nullptr, 0, len);
    }
    void MulEqModElw(uint64_t* lo, const uint64_t* r, uint64_t len) override { Elw(TP_ELT_MUL, lo, lo, r, nullptr, 0, len); }
    void FMAModElw(uint64_t* res, const uint64_t* l, const uint64_t* r, const uint64_t* s, uint64_t len) override {
        Elw(TP_ELT_FMA, res, l, r, s, 0, len);
    }
    void FMAEqModElw(uint64_t* lo, const uint64_t* r, const uint64_t* s, uint64_t len) override {
        Elw(TP_ELT_FMA, lo, lo, r, s, 0, len);
    }
    void MulSModElw(uint64_t* res, const uint64_t* l, uint64_t r, uint64_t len) override {
        Elw(TP_ELT_MULS, res, l, nullptr, nullptr, r % kQ, len);
    }
    void MulSEqModElw(uint64_t* lo, uint64_t r, uint64_t len) override { Elw(TP_ELT_MULS, lo, lo, nullptr, nullptr, r % kQ, len); }
    void ReduceModElw(uint64_t* out, const uint64_t* in, uint64_t len) override {
        Elw(TP_ELT_REDUCE, out, in, nullptr, nullptr, 0, len);
    }
    uint64_t GetModulus() const override { return kQ; }
    uint64_t GetDimension() const override { return m_N; }

private:
    tp_context* Device() {
        if (!m_dev) m_dev = make_device_context(base_params(m_N ? m_N : 1024));
        return m_dev.get();
    }
    void Transform(uint64_t* result, const uint64_t* input, bool inverse) {
        require(m_N != 0, "this worker has no ring dimension (SelectWorkerArbitraryDimension)");
        std::lock_guard<std::mutex> lk(m_mu);
        DevBuf d(Device(), m_N);
        d.put(input, m_N);
        if (inverse) check(tp_ntt_inverse_device(Device(), d.p, d.p, 1), "BackwardNTT");
        else check(tp_ntt_forward_device(Device(), d.p, d.p, 1), "ForwardNTT");
        d.get(result, m_N);
    }
    void Elw(int op, uint64_t* res, const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t s, uint64_t len) {
        if (len == 0) return;
        std::lock_guard<std::mutex> lk(m_mu);
        tp_context* x = Device();
        DevBuf da(x, len), db(x, b ? len : 1), dc(x, c ? len : 1);
        da.put(a, len);
        if (b) db.put(b, len);
        if (c) dc.put(c, len);
        check(tp_eltwise_device(x, op, da.p, da.p, b ? db.p : nullptr, c ? dc.p : nullptr, s, len), "elementwise op");
        da.get(res, len);
    }
    uint64_t m_N;
    DeviceContext m_dev;
    std::mutex m_mu;
};

std::mutex g_worker_mu;
std::map<uint64_t, std::shared_ptr<MathWorker>> g_workers;

}  // namespace

std::shared_ptr<MathWorker> SelectWorker(uint64_t modulus, uint64_t dimension) {
    require_goldilocks(modulus);
    require_ring(dimension);
    std::lock_guard<std::mutex> lk(g_worker_mu);
    auto& w = g_workers[dimension];
    if (!w) w = std::make_shared<GpuSolinasWorker>(dimension);
    return w;
}

std::shared_ptr<MathWorker> SelectWorkerArbitraryDimension(uint64_t modulus) {
    require_goldilocks(modulus);
    std::lock_guard<std::mutex> lk(g_worker_mu);
    auto& w = g_workers[0];
    if (!w) w = std::make_shared<GpuSolinasWorker>(0);
    return w;
}

// ------------------------------------------------------------------ modulus switching

namespace {
std::mutex g_ms_mu;
DeviceContext g_ms_dev;
uint64_t g_ms_calls = 0;
}  // namespace

void ModulusSwitch(uint64_t* vec, uint64_t n, uint64_t source_modulus, uint64_t target_modulus,
                   ModulusSwitchType type) {
    if (n == 0) return;
    std::lock_guard<std::mutex> lk(g_ms_mu);
    if (!g_ms_dev) g_ms_dev = make_device_context(base_params(1024));
    tp_context* x = g_ms_dev.get();
    DevBuf d(x, n);
    d.put(vec, n);
    // RANDOM: a fresh seed per call, as the reference seeds its coin per call from random_device
    const uint64_t seed = type == ModulusSwitchType::RANDOM ? toothpaste_amd::fresh_seed() ^ ++g_ms_calls : 0;
    check(tp_modswitch_device(x, d.p, d.p, n, source_modulus, target_modulus, (int)type, seed), "ModulusSwitch");
    d.get(vec, n);
}

long double EstimateModulusSwitchingVariance(long double input_variance, uint64_t source_modulus,
                                             uint64_t target_modulus, KeyDistribution kd, uint64_t expected_l0) {
    // modulus_switching.cpp:12-34: (Qs^2 var) / Qt^2 + E[hamming weight] / 12, binary/ternary keys
    require(kd == BINARY || kd == TERNARY, "modulus-switching variance is defined for binary and ternary keys");
    const long double qs = (long double)source_modulus, qt = (long double)target_modulus;
    return (qs * qs * input_variance) / (qt * qt) + (long double)expected_l0 / 12.0L;
}

// ------------------------------------------------------------------ BlindRotator default batch

void BlindRotator::BlindRotateBatch(uint64_t* results, const uint64_t* lwes, uint64_t count,
                                    const uint64_t* accumulators, uint64_t accumulator_stride,
                                    bool output_as_coefficients) {
    auto in = std::dynamic_pointer_cast<TupleContainerImpl>(GetContext()->GetInputContainer());
    require(in != nullptr, "rotator context has no tuple input container");
    const uint64_t n = std::dynamic_pointer_cast<LWEContainerImpl>(in->GetElem(0))->GetN();
    const uint64_t N = std::dynamic_pointer_cast<RLWEContainerImpl>(in->GetElem(1))->GetN();
    std::vector<uint64_t> acc(2 * N);
    for (uint64_t c = 0; c < count; ++c) {
        const uint64_t* src = accumulators + c * accumulator_stride;
        acc.assign(src, src + 2 * N);
        BlindRotate(results + c * 2 * N, lwes + c * (n + 1), acc.data(), output_as_coefficients);
    }
}

namespace {

// Batched blind rotation on a device context (tp_eval_func_batch_device with RLWE output).
void device_rotate_batch(tp_context* dev, uint64_t N, uint64_t n, uint64_t* results, const uint64_t* lwes,
                         uint64_t count, const uint64_t* accs, uint64_t stride, bool as_coef) {
    require(stride == 0 || stride == 2 * N, "accumulator_stride must be 0 or 2N");
    if (count == 0) return;
    check_lwe_entries(lwes, count * (n + 1), N);
    const uint64_t acc_words = stride ? count * 2 * N : 2 * N;
    DevBuf d_lwe(dev, count * (n + 1)), d_acc(dev, acc_words), d_out(dev, count * 2 * N);
    d_lwe.put(lwes, count * (n + 1));
    d_acc.put(accs, acc_words);
    check(tp_eval_func_batch_device(dev, count, d_lwe.p, d_acc.p, stride, nullptr, d_out.p, as_coef ? 0 : 1),
          "BlindRotateBatch");
    d_out.get(results, count * 2 * N);
}

Container rotation_input_container(uint64_t n, uint64_t N, uint64_t modulus, long double var) {
    Container lwecont = std::make_shared<LWEContainerImpl>(n, 2 * N, 0.0);
    Container rlwecont = std::make_shared<RLWEContainerImpl>(N, modulus, var);
    return std::make_shared<TupleContainerImpl>(std::vector<Container>{lwecont, rlwecont});
}

Container rotation_output_container(const Container& input, uint64_t N, uint64_t modulus, long double var) {
    auto tp = std::dynamic_pointer_cast<TupleContainerImpl>(input);
    require(tp != nullptr, "expected the rotation input tuple");
    auto rl = std::dynamic_pointer_cast<RLWEContainerImpl>(tp->GetElem(1));
    return std::make_shared<RLWEContainerImpl>(N, modulus, var + rl->GetVariance());
}

}  // namespace

// ------------------------------------------------------------------ CGGI

CGGIBlindRotationContext::CGGIBlindRotationContext(KeyDistribution distr, uint64_t modulus, uint64_t ring_dim,
                                                   uint64_t lwe_dim, uint64_t basis, uint64_t digits, double std)
    : m_distribution(distr), m_modulus(modulus), m_N(ring_dim), m_n(lwe_dim), m_digits(digits), m_std(std),
      m_seed(TP_SEED_OS_ENTROPY) {
    SetBlindRotationBasis(basis);
    m_ntt = SelectWorker(modulus, ring_dim);
}

CGGIBlindRotationContext::CGGIBlindRotationContext(KeyDistribution distr, std::shared_ptr<MathWorker> ntt,
                                                   uint64_t lwe_dim, uint64_t basis, uint64_t digits, double std)
    : m_distribution(distr), m_modulus(ntt->GetModulus()), m_N(ntt->GetDimension()), m_n(lwe_dim), m_digits(digits),
      m_std(std), m_seed(TP_SEED_OS_ENTROPY), m_ntt(std::move(ntt)) {
    SetBlindRotationBasis(basis);
}

CGGIBlindRotationContext::CGGIBlindRotationContext(const CGGIBlindRotationContext& other)
    : OperatorContext<CGGIBlindRotator>(other), enable_shared_from_this(other), m_distribution(other.m_distribution),
      m_modulus(other.m_modulus), m_N(other.m_N), m_n(other.m_n), m_basis(other.m_basis),
      m_basis_log2(other.m_basis_log2), m_digits(other.m_digits), m_std(other.m_std), m_seed(other.m_seed),
      m_ntt(other.m_ntt) {}

void CGGIBlindRotationContext::SetBlindRotationBasis(uint64_t L) {
    m_basis = L;
    m_basis_log2 = int_log2(L);
}

long double CGGIBlindRotationContext::ComputeOutputVariance(long double) const {
    // n steps, each RGSW(s_i) [x] (X^a - 1) acc + acc: 8 x the per-product variance (cggi_blind_rotation.cpp:86-124)
    return (long double)m_n * 8.0L * product_variance(m_modulus, m_N, m_basis, m_digits, m_std, m_distribution);
}

Container CGGIBlindRotationContext::GetInputContainer() const {
    return rotation_input_container(m_n, m_N, m_modulus, ComputeOutputVariance(0.0));
}

Container CGGIBlindRotationContext::GetOutputContainer(Container input) const {
    return rotation_output_container(input, m_N, m_modulus, ComputeOutputVariance());
}

std::unique_ptr<CGGIBlindRotator> CGGIBlindRotationContext::ConstructOperator(const std::vector<GenericKey>& bundle) const {
    require(bundle.size() >= 2, "CGGI key bundle is {LWE secret, RLWE secret}");
    require(bundle[0].GetKey().size() >= m_n && bundle[1].GetKey().size() >= m_N, "secret key too short");
    auto op = std::make_unique<CGGIBlindRotator>(shared_from_this());
    op->KeyGen(bundle[0].GetKeyPtr(), bundle[1].GetKeyPtr());
    return op;
}

CGGIBlindRotator::CGGIBlindRotator(const std::shared_ptr<const CGGIBlindRotationContext>& params) : m_params(params) {
    require_goldilocks(params->m_modulus);
    require_ring(params->m_N);
    require(params->m_distribution == BINARY || params->m_distribution == TERNARY, "Unsupported key distribution");
    require(params->m_basis == (1ULL << params->m_basis_log2), "blind-rotation basis must be a power of two");
    tp_params p{};
    p.ring_dim = (uint32_t)params->m_N;
    p.lwe_dim = (uint32_t)params->m_n;
    p.br_base_log2 = (uint32_t)params->m_basis_log2;
    p.br_digits = (uint32_t)params->m_digits;
    p.method = params->m_distribution == BINARY ? TP_METHOD_CGGI_BINARY : TP_METHOD_CGGI_TERNARY;
    m_dev = make_device_context(p);
}

void CGGIBlindRotator::KeyGen(const uint64_t* lwe_key, const uint64_t* rlwe_key) {
    const auto& P = *m_params;
    const bool binary = P.m_distribution == BINARY;
    const int method = binary ? TP_METHOD_CGGI_BINARY : TP_METHOD_CGGI_TERNARY;
    m_brk.assign((size_t)(binary ? 4 : 8) * P.m_N * P.m_digits * P.m_n, 0);
    // binary keys: MakeRGSW gadget rows (the unsigned digits of BinaryCMux); ternary: signed digits
    check(tp_keygen_brk(m_brk.data(), lwe_key, rlwe_key, (uint32_t)P.m_N, (uint32_t)P.m_n, (uint32_t)P.m_basis_log2,
                        (uint32_t)P.m_digits, method, binary ? 0 : 1, P.m_std, P.m_seed),
          "blind-rotation key generation");
    check(tp_import_brk(m_dev.get(), m_brk.data(), m_brk.size()), "tp_import_brk");
    m_keys_generated = true;
}

void CGGIBlindRotator::BlindRotate(std::vector<uint64_t>& result, const std::vector<uint64_t>& lwe_vec,
                                   std::vector<uint64_t>& rlwe_acc_vec, bool output_as_coefficients) {
    BlindRotate(result.data(), lwe_vec.data(), rlwe_acc_vec.data(), output_as_coefficients);
}

void CGGIBlindRotator::BlindRotate(uint64_t* result, const uint64_t* lwe_vec, uint64_t* rlwe_acc_vec,
                                   bool output_as_coefficients) {
    require(m_keys_generated, "keys not generated");
    check(tp_blind_rotate(m_dev.get(), result ? result : rlwe_acc_vec, lwe_vec, rlwe_acc_vec, output_as_coefficients),
          "BlindRotate");
}

void CGGIBlindRotator::BlindRotateBatch(uint64_t* results, const uint64_t* lwes, uint64_t count,
                                        const uint64_t* accumulators, uint64_t accumulator_stride,
                                        bool output_as_coefficients) {
    require(m_keys_generated, "keys not generated");
    device_rotate_batch(m_dev.get(), m_params->m_N, m_params->m_n, results, lwes, count, accumulators,
                        accumulator_stride, output_as_coefficients);
}

const std::shared_ptr<const OperatorContext<BlindRotator>> CGGIBlindRotator::GetContext() const {
    // the rotator contexts are used through OperatorContext<BlindRotator> as in the reference
    // (cggi_blind_rotation.cpp:491-493)
    return std::reinterpret_pointer_cast<const OperatorContext<BlindRotator>>(m_params);
}

// ------------------------------------------------------------------ BMMP

BMMPBlindRotationContext::BMMPBlindRotationContext(KeyDistribution distr, uint64_t modulus, uint64_t ring_dim,
                                                   uint64_t lwe_dim, uint64_t basis, uint64_t digits, double std,
                                                   uint64_t step_size)
    : m_distribution(distr), m_modulus(modulus), m_N(ring_dim), m_n(lwe_dim), m_basis(basis),
      m_basis_log2(int_log2(basis)), m_digits(digits), m_step_size(step_size), m_std(std),
      m_seed(TP_SEED_OS_ENTROPY) {
    m_ntt = SelectWorker(modulus, ring_dim);
}

BMMPBlindRotationContext::BMMPBlindRotationContext(KeyDistribution distr, std::shared_ptr<MathWorker> ntt,
                                                   uint64_t lwe_dim, uint64_t basis, uint64_t digits, double std,
                                                   uint64_t step_size)
    : m_distribution(distr), m_modulus(ntt->GetModulus()), m_N(ntt->GetDimension()), m_n(lwe_dim), m_basis(basis),
      m_basis_log2(int_log2(basis)), m_digits(digits), m_step_size(step_size), m_std(std),
      m_seed(TP_SEED_OS_ENTROPY), m_ntt(std::move(ntt)) {}

long double BMMPBlindRotationContext::ComputeOutputVariance(long double) const {
    return 0.0L;  // the reference has no estimate for BMMP (bmmp_blind_rotation.cpp:122-125)
}

Container BMMPBlindRotationContext::GetInputContainer() const {
    return rotation_input_container(m_n, m_N, m_modulus, ComputeOutputVariance(0.0));
}

Container BMMPBlindRotationContext::GetOutputContainer(Container input) const {
    return rotation_output_container(input, m_N, m_modulus, ComputeOutputVariance());
}

std::unique_ptr<BMMPBlindRotator> BMMPBlindRotationContext::ConstructOperator(const std::vector<GenericKey>& bundle) const {
    require(bundle.size() >= 2, "BMMP key bundle is {LWE secret, RLWE secret}");
    require(bundle[0].GetKey().size() >= m_n && bundle[1].GetKey().size() >= m_N, "secret key too short");
    auto op = std::make_unique<BMMPBlindRotator>(shared_from_this());
    op->KeyGen(bundle[0].GetKeyPtr(), bundle[1].GetKeyPtr());
    return op;
}

BMMPBlindRotator::BMMPBlindRotator(const std::shared_ptr<const BMMPBlindRotationContext>& params) : m_params(params) {
    require_goldilocks(params->m_modulus);
    require_ring(params->m_N);
    require(params->m_distribution == BINARY, "BMMP blind rotation takes binary LWE keys");
    require(params->m_step_size == 2, "BMMP step size must be 2 (three key combinations per coefficient pair)");
    require(params->m_n % 2 == 0, "BMMP needs an even LWE dimension");
    require(params->m_basis == (1ULL << params->m_basis_log2), "blind-rotation basis must be a power of two");
    tp_params p{};
    p.ring_dim = (uint32_t)params->m_N;
    p.lwe_dim = (uint32_t)params->m_n;
    p.br_base_log2 = (uint32_t)params->m_basis_log2;
    p.br_digits = (uint32_t)params->m_digits;
    p.method = TP_METHOD_BMMP;
    m_dev = make_device_context(p);
}

void BMMPBlindRotator::KeyGen(const uint64_t* lwe_key, const uint64_t* rlwe_key) {
    const auto& P = *m_params;
    m_brk.assign((size_t)12 * P.m_N * P.m_digits * (P.m_n / 2), 0);
    check(tp_keygen_brk(m_brk.data(), lwe_key, rlwe_key, (uint32_t)P.m_N, (uint32_t)P.m_n, (uint32_t)P.m_basis_log2,
                        (uint32_t)P.m_digits, TP_METHOD_BMMP, 1, P.m_std, P.m_seed),
          "blind-rotation key generation");
    check(tp_import_brk(m_dev.get(), m_brk.data(), m_brk.size()), "tp_import_brk");
    m_keys_generated = true;
}

void BMMPBlindRotator::BlindRotate(std::vector<uint64_t>& result, const std::vector<uint64_t>& lwe_vec,
                                   std::vector<uint64_t>& rlwe_acc_vec, bool output_as_coefficients) {
    BlindRotate(result.data(), lwe_vec.data(), rlwe_acc_vec.data(), output_as_coefficients);
}

void BMMPBlindRotator::BlindRotate(uint64_t* result, const uint64_t* lwe_vec, uint64_t* rlwe_acc_vec,
                                   bool output_as_coefficients) {
    require(m_keys_generated, "keys not generated");
    check(tp_blind_rotate(m_dev.get(), result ? result : rlwe_acc_vec, lwe_vec, rlwe_acc_vec, output_as_coefficients),
          "BlindRotate");
}

void BMMPBlindRotator::BlindRotateBatch(uint64_t* results, const uint64_t* lwes, uint64_t count,
                                        const uint64_t* accumulators, uint64_t accumulator_stride,
                                        bool output_as_coefficients) {
    require(m_keys_generated, "keys not generated");
    device_rotate_batch(m_dev.get(), m_params->m_N, m_params->m_n, results, lwes, count, accumulators,
                        accumulator_stride, output_as_coefficients);
}

const std::shared_ptr<const OperatorContext<BlindRotator>> BMMPBlindRotator::GetContext() const {
    return std::reinterpret_pointer_cast<const OperatorContext<BlindRotator>>(m_params);
}

// ------------------------------------------------------------------ LWE -> LWE

LWEConversionContext::LWEConversionContext(KeyDistribution source_key_distribution, uint64_t modulus,
                                           uint64_t source_dimension, uint64_t target_dimension, uint64_t basis,
                                           uint64_t digits, double std)
    : m_source_distribution(source_key_distribution), m_modulus(modulus), m_source_dimension(source_dimension),
      m_target_dimension(target_dimension), m_basis(basis), m_basis_log2(int_log2(basis)), m_digits(digits),
      m_std(std), m_seed(TP_SEED_OS_ENTROPY) {}

long double LWEConversionContext::ComputeOutputVariance(long double input_variance) const {
    // lwe_conversion.cpp:114-146: n_src * digits * basis^2 std^2 / 12 + E|s|_0 * delta^2/12 * |s_i|^2
    long double delta = std::pow(2.0L, (long double)int_log2(m_modulus)) /
                        std::pow((long double)m_basis, (long double)m_digits);
    if (delta <= 1) delta = 0.0L;
    delta *= delta / 12.0L;
    const long double norm = m_source_distribution == GAUSSIAN ? 25.0L : 1.0L;
    uint64_t weight = m_source_dimension;
    if (m_source_distribution == BINARY) weight = m_source_dimension / 2;
    if (m_source_distribution == TERNARY) weight = 2 * m_source_dimension / 3;
    const long double per = (long double)m_digits * m_basis * m_basis * m_std * m_std / 12.0L;
    return input_variance + (long double)m_source_dimension * per + (long double)weight * delta * norm;
}

Container LWEConversionContext::GetInputContainer() const {
    return std::make_shared<LWEContainerImpl>(m_source_dimension, m_modulus, 0.0);
}

Container LWEConversionContext::GetOutputContainer(Container input) const {
    long double var = 0.0L;
    if (auto r = std::dynamic_pointer_cast<RLWEContainerImpl>(input)) var = r->GetVariance();
    else if (auto l = std::dynamic_pointer_cast<LWEContainerImpl>(input)) var = l->GetVariance();
    return std::make_shared<LWEContainerImpl>(m_target_dimension, m_modulus, ComputeOutputVariance(var));
}

std::unique_ptr<LWEtoLWEConverter> LWEConversionContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    require(keys.size() >= 2, "LWE key-switch bundle is {source key, target key}");
    require(keys[0].GetKey().size() >= m_source_dimension && keys[1].GetKey().size() >= m_target_dimension,
            "secret key too short");
    auto op = std::unique_ptr<LWEtoLWEConverter>(new LWEtoLWEConverter(shared_from_this()));
    op->KeyGen(keys[0].GetKeyPtr(), keys[1].GetKeyPtr());
    return op;
}

LWEtoLWEConverter::LWEtoLWEConverter(std::shared_ptr<const LWEConversionContext> params) : m_params(std::move(params)) {
    const auto& P = *m_params;
    require_goldilocks(P.GetModulus());
    require_ring(P.GetSourceDimension());
    require(P.GetGadgetBasis() == (1ULL << P.GetGadgetBasisLog2()), "key-switch basis must be a power of two");
    tp_params p = base_params(P.GetSourceDimension());
    p.ks_out_dim = (uint32_t)P.GetTargetDimension();
    p.ks_base_log2 = (uint32_t)P.GetGadgetBasisLog2();
    p.ks_digits = (uint32_t)P.GetGadgetDigits();
    m_dev = make_device_context(p);
}

void LWEtoLWEConverter::KeyGen(const uint64_t* source_key, const uint64_t* target_key) {
    const auto& P = *m_params;
    const uint64_t n_in = P.GetSourceDimension(), n_out = P.GetTargetDimension(), d = P.GetGadgetDigits();
    m_ksk.assign((size_t)n_in * d * (n_out + 1), 0);
    check(tp_keygen_lwe_ksk(m_ksk.data(), source_key, target_key, (uint32_t)n_in, (uint32_t)n_out,
                            (uint32_t)P.GetGadgetBasisLog2(), (uint32_t)d, P.GetStd(), P.GetSeed()),
          "LWE key-switching key generation");
    check(tp_import_lwe_ksk(m_dev.get(), m_ksk.data(), m_ksk.size()), "tp_import_lwe_ksk");
}

void LWEtoLWEConverter::Convert(uint64_t* output, const uint64_t* const input) {
    const uint64_t n_in = m_params->GetSourceDimension(), n_out = m_params->GetTargetDimension();
    tp_context* x = m_dev.get();
    DevBuf d_in(x, n_in + 1), d_res(x, n_out + 1), d_out(x, n_out + 1);
    d_in.put(input, n_in + 1);
    d_out.put(output, n_out + 1);
    check(tp_lwe_keyswitch_batch_device(x, 1, d_in.p, n_in + 1, d_res.p), "LWE key switch");
    // output[0..n_out) -= acc; output[n_out] = input[n_in] - acc[n_out]
    check(tp_eltwise_device(x, TP_ELT_ADD, d_out.p, d_out.p, d_res.p, nullptr, 0, n_out), "accumulate");
    check(tp_copy_on_device(x, d_out.p + n_out, d_res.p + n_out, 1), "copy");
    d_out.get(output, n_out + 1);
}

void LWEtoLWEConverter::Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) {
    Convert(output.data(), input.data());
}

void LWEtoLWEConverter::ConvertBatch(uint64_t* outputs, const uint64_t* inputs, uint64_t count) {
    if (count == 0) return;
    const uint64_t n_in = m_params->GetSourceDimension(), n_out = m_params->GetTargetDimension();
    tp_context* x = m_dev.get();
    DevBuf d_in(x, count * (n_in + 1)), d_out(x, count * (n_out + 1));
    d_in.put(inputs, count * (n_in + 1));
    check(tp_lwe_keyswitch_batch_device(x, count, d_in.p, n_in + 1, d_out.p), "LWE key switch");
    d_out.get(outputs, count * (n_out + 1));
}

// ------------------------------------------------------------------ RLWE -> RLWE

RLWEConversionContext::RLWEConversionContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N,
                                             uint64_t basis, uint64_t digits, double std)
    : m_source_distribution(source_key_distribution), m_modulus(modulus), m_N(N), m_basis(basis),
      m_basis_log2(int_log2(basis)), m_digits(digits), m_std(std), m_seed(TP_SEED_OS_ENTROPY) {
    m_ntt = SelectWorker(modulus, N);
}

RLWEConversionContext::RLWEConversionContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt,
                                             uint64_t basis, uint64_t digits, double std)
    : m_source_distribution(source_key_distribution), m_modulus(ntt->GetModulus()), m_N(ntt->GetDimension()),
      m_basis(basis), m_basis_log2(int_log2(basis)), m_digits(digits), m_std(std),
      m_seed(TP_SEED_OS_ENTROPY), m_ntt(std::move(ntt)) {}

long double RLWEConversionContext::ComputeOutputVariance(long double input_variance) const {
    return input_variance + product_variance(m_modulus, m_N, m_basis, m_digits, m_std, m_source_distribution);
}

Container RLWEConversionContext::GetInputContainer() const {
    // argument order as the reference writes it (rlwe_conversion.cpp:149-151)
    return std::make_shared<RLWEContainerImpl>(m_modulus, m_N, 0.0);
}

Container RLWEConversionContext::GetOutputContainer(Container input) const {
    auto in = std::dynamic_pointer_cast<RLWEContainerImpl>(input);
    return std::make_shared<RLWEContainerImpl>(m_N, m_modulus, ComputeOutputVariance(in ? in->GetVariance() : 0.0L));
}

std::unique_ptr<RLWEtoRLWEConverter> RLWEConversionContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    require(keys.size() >= 2, "RLWE key-switch bundle is {source key, target key}");
    require(keys[0].GetKey().size() >= m_N && keys[1].GetKey().size() >= m_N, "secret key too short");
    auto op = std::unique_ptr<RLWEtoRLWEConverter>(new RLWEtoRLWEConverter(shared_from_this()));
    op->KeyGen(keys[0].GetKeyPtr(), keys[1].GetKeyPtr());
    return op;
}

RLWEtoRLWEConverter::RLWEtoRLWEConverter(std::shared_ptr<const RLWEConversionContext> params) : m_params(std::move(params)) {
    const auto& P = *m_params;
    require_goldilocks(P.GetModulus());
    require_ring(P.GetDimension());
    require(P.GetGadgetBasis() == (1ULL << P.GetGadgetBasisLog2()), "gadget basis must be a power of two");
    tp_params p = base_params(P.GetDimension());
    p.rlwe_ks_base_log2 = (uint32_t)P.GetGadgetBasisLog2();
    p.rlwe_ks_digits = (uint32_t)P.GetGadgetDigits();
    m_dev = make_device_context(p);
}

void RLWEtoRLWEConverter::KeyGen(const uint64_t* source_key, const uint64_t* target_key) {
    const auto& P = *m_params;
    m_ksk.assign((size_t)P.GetGadgetDigits() * 2 * P.GetDimension(), 0);
    check(tp_keygen_rlwe_ksk(m_ksk.data(), source_key, target_key, (uint32_t)P.GetDimension(),
                             (uint32_t)P.GetGadgetBasisLog2(), (uint32_t)P.GetGadgetDigits(), P.GetStd(), P.GetSeed(), 6),
          "RLWE key-switching key generation");
    check(tp_import_rlwe_ksk(m_dev.get(), m_ksk.data(), m_ksk.size()), "tp_import_rlwe_ksk");
}

void RLWEtoRLWEConverter::Convert(uint64_t* output, const uint64_t* const input) {
    const uint64_t N = m_params->GetDimension();
    tp_context* x = m_dev.get();
    DevBuf d_in(x, 2 * N), d_out(x, 2 * N);
    d_in.put(input, 2 * N);
    d_out.put(output, 2 * N);
    check(tp_rlwe_keyswitch_device(x, 1, 1, TP_RKS_CONVERT | TP_RKS_ACCUMULATE, 0, d_in.p, d_out.p), "RLWE Convert");
    d_out.get(output, 2 * N);
}

void RLWEtoRLWEConverter::Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) {
    Convert(output.data(), input.data());
}

// ------------------------------------------------------------------ automorphism

AutomorphismContext::AutomorphismContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N,
                                         uint64_t basis, uint64_t digits, double std, uint32_t automorphism_index)
    : m_rlwe_conversion(std::make_shared<RLWEConversionContext>(source_key_distribution, modulus, N, basis, digits, std)),
      m_automorphism_index(automorphism_index) {}

AutomorphismContext::AutomorphismContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt,
                                         uint64_t basis, uint64_t digits, double std, uint32_t automorphism_index)
    : m_rlwe_conversion(std::make_shared<RLWEConversionContext>(source_key_distribution, std::move(ntt), basis, digits, std)),
      m_automorphism_index(automorphism_index) {}

AutomorphismContext::AutomorphismContext(const AutomorphismContext& other)
    : OperatorContext<AutomorphismEvaluator>(other), enable_shared_from_this(other),
      m_rlwe_conversion(std::make_shared<RLWEConversionContext>(*other.m_rlwe_conversion)),
      m_automorphism_index(other.m_automorphism_index) {}

long double AutomorphismContext::ComputeOutputVariance(long double input_variance) const {
    return m_rlwe_conversion->ComputeOutputVariance(input_variance);
}

Container AutomorphismContext::GetInputContainer() const {
    return std::make_shared<RLWEContainerImpl>(GetDimension(), GetModulus(), 0.0);
}

Container AutomorphismContext::GetOutputContainer(Container input) const {
    return m_rlwe_conversion->GetOutputContainer(input);
}

std::unique_ptr<AutomorphismEvaluator> AutomorphismContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    // automorphism_evaluation.cpp:43-63: converter from s(X^t) to s(X)
    const uint64_t N = GetDimension(), t = m_automorphism_index;
    require(!keys.empty() && keys[0].GetKey().size() >= N, "automorphism key bundle is {RLWE secret}");
    require((t & 1) && t < 2 * N, "automorphism index must be odd and below 2N");
    require_goldilocks(GetModulus());
    require_ring(N);
    std::vector<uint64_t> s(keys[0].GetKeyPtr(), keys[0].GetKeyPtr() + N), s_auto(N);
    auto_poly(s_auto.data(), s.data(), t, N);
    tp_params p = base_params(N);
    p.rlwe_ks_base_log2 = (uint32_t)GetGadgetBasisLog2();
    p.rlwe_ks_digits = (uint32_t)GetGadgetDigits();
    DeviceContext dev = make_device_context(p);
    std::vector<uint64_t> ksk((size_t)GetGadgetDigits() * 2 * N);
    check(tp_keygen_rlwe_ksk(ksk.data(), s_auto.data(), s.data(), (uint32_t)N, (uint32_t)GetGadgetBasisLog2(),
                             (uint32_t)GetGadgetDigits(), GetStd(), GetSeed(), 6),
          "automorphism key generation");
    check(tp_import_rlwe_ksk(dev.get(), ksk.data(), ksk.size()), "tp_import_rlwe_ksk");
    return std::unique_ptr<AutomorphismEvaluator>(new AutomorphismEvaluator(shared_from_this(), dev, 0));
}

AutomorphismEvaluator::AutomorphismEvaluator(std::shared_ptr<const AutomorphismContext> params, DeviceContext dev,
                                             int key_index)
    : m_params(std::move(params)), m_dev(std::move(dev)), m_key_index(key_index) {}

void AutomorphismEvaluator::Run(uint64_t* output, const uint64_t* input, int mode) {
    const uint64_t N = m_params->GetDimension();
    tp_context* x = m_dev.get();
    DevBuf d_in(x, 2 * N), d_out(x, 2 * N);
    d_in.put(input, mode == TP_RKS_SWITCH_ONLY ? N : 2 * N);
    d_out.put(output, 2 * N);
    check(tp_rlwe_keyswitch_device(x, 1, m_params->GetAutomorphismIndex(), mode | TP_RKS_ACCUMULATE, m_key_index,
                                   d_in.p, d_out.p),
          "automorphism");
    d_out.get(output, 2 * N);
}

void AutomorphismEvaluator::Eval(uint64_t* output, const uint64_t* const input) { Run(output, input, TP_RKS_AUTOMORPHISM); }

void AutomorphismEvaluator::Eval(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) {
    Run(output.data(), input.data(), TP_RKS_AUTOMORPHISM);
}

void AutomorphismEvaluator::EvalSwitchOnly(uint64_t* output, const uint64_t* const input) {
    Run(output, input, TP_RKS_SWITCH_ONLY);
}

// ------------------------------------------------------------------ trace and packing

namespace {

// The log2 N automorphism keys t = 2^j + 1 of one secret (trace_evaluation.cpp:50-62,
// lwe_to_rlwe_packing.cpp:46-58), resident on one device context.
DeviceContext auto_key_context(const AutomorphismContext& A, const std::vector<GenericKey>& keys,
                               std::vector<uint64_t>& out_keys) {
    const uint64_t N = A.GetDimension();
    require(!keys.empty() && keys[0].GetKey().size() >= N, "key bundle is {RLWE secret}");
    require_goldilocks(A.GetModulus());
    require_ring(N);
    tp_params p = base_params(N);
    p.rlwe_ks_base_log2 = (uint32_t)A.GetGadgetBasisLog2();
    p.rlwe_ks_digits = (uint32_t)A.GetGadgetDigits();
    DeviceContext dev = make_device_context(p);
    out_keys.assign((size_t)int_log2(N) * A.GetGadgetDigits() * 2 * N, 0);
    check(tp_keygen_auto_keys(out_keys.data(), keys[0].GetKeyPtr(), (uint32_t)N, (uint32_t)A.GetGadgetBasisLog2(),
                              (uint32_t)A.GetGadgetDigits(), A.GetStd(), A.GetSeed()),
          "automorphism key generation");
    check(tp_import_auto_keys(dev.get(), out_keys.data(), out_keys.size()), "tp_import_auto_keys");
    return dev;
}

}  // namespace

TraceEvaluationContext::TraceEvaluationContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N,
                                               uint64_t basis, uint64_t digits, double std)
    : m_auto_context(std::make_shared<AutomorphismContext>(source_key_distribution, modulus, N, basis, digits, std, 3)) {}

TraceEvaluationContext::TraceEvaluationContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt,
                                               uint64_t basis, uint64_t digits, double std)
    : m_auto_context(std::make_shared<AutomorphismContext>(source_key_distribution, std::move(ntt), basis, digits, std, 3)) {}

TraceEvaluationContext::TraceEvaluationContext(const TraceEvaluationContext& other)
    : OperatorContext<TraceEvaluator>(other), enable_shared_from_this(other),
      m_auto_context(std::make_shared<AutomorphismContext>(*other.m_auto_context)) {}

long double TraceEvaluationContext::ComputeOutputVariance(long double) const {
    // trace_evaluation.cpp:35-40: variance of the automorphism input container (0) x N^2 / 3
    const long double N = (long double)GetDimension();
    auto in = std::dynamic_pointer_cast<RLWEContainerImpl>(m_auto_context->GetInputContainer());
    return in->GetVariance() * (N * N / 3.0L);
}

Container TraceEvaluationContext::GetInputContainer() const { return m_auto_context->GetInputContainer(); }

Container TraceEvaluationContext::GetOutputContainer(Container input) const {
    auto in = std::dynamic_pointer_cast<RLWEContainerImpl>(input);
    return std::make_shared<RLWEContainerImpl>(GetDimension(), GetModulus(),
                                               ComputeOutputVariance(in ? in->GetVariance() : 0.0L));
}

std::unique_ptr<TraceEvaluator> TraceEvaluationContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    std::vector<uint64_t> k;
    DeviceContext dev = auto_key_context(*m_auto_context, keys, k);
    return std::unique_ptr<TraceEvaluator>(new TraceEvaluator(shared_from_this(), dev, std::move(k)));
}

TraceEvaluator::TraceEvaluator(std::shared_ptr<const TraceEvaluationContext> context, DeviceContext dev,
                               std::vector<uint64_t> keys)
    : m_context(std::move(context)), m_dev(std::move(dev)), m_keys(std::move(keys)) {
    // evaluator j-1 applies t = 2^j + 1 with automorphism key j, sharing this device context
    const uint64_t logN = int_log2(m_context->GetDimension());
    for (uint64_t j = 1; j <= logN; ++j) {
        auto ctx = std::make_shared<AutomorphismContext>(*m_context->GetAutomorphismContext());
        ctx->SetAutomorphismIndex((uint32_t)((1ULL << j) + 1));
        m_trace_evaluators.push_back(std::unique_ptr<AutomorphismEvaluator>(new AutomorphismEvaluator(ctx, m_dev, (int)j)));
    }
}

void TraceEvaluator::Eval(uint64_t* output, const uint64_t* input) {
    const uint64_t N = m_context->GetDimension();
    tp_context* x = m_dev.get();
    DevBuf d_in(x, 2 * N), d_out(x, 2 * N);
    d_in.put(input, 2 * N);
    check(tp_trace_device(x, 1, d_in.p, d_out.p), "trace");
    d_out.get(output, 2 * N);
}

void TraceEvaluator::Eval(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) {
    Eval(output.data(), input.data());
}

void TraceEvaluator::EvalAuto(uint64_t* output, const uint64_t* input, uint64_t auto_idx) {
    const uint64_t index = int_log2(auto_idx - 1);
    require(auto_idx >= 3 && ((1ULL << index) + 1) == auto_idx, "automorphism index must be 2^j + 1");
    require(index >= 1 && index <= m_trace_evaluators.size(), "automorphism index out of range");
    m_trace_evaluators[index - 1]->Eval(output, input);
}

void TraceEvaluator::EvalAuto(std::vector<uint64_t>& output, const std::vector<uint64_t>& input, uint64_t auto_idx) {
    // trace_evaluation.cpp:182-188 indexes the evaluators with log2(t - 1), not log2(t - 1) - 1
    const uint64_t index = int_log2(auto_idx - 1);
    require(auto_idx >= 3 && ((1ULL << index) + 1) == auto_idx, "automorphism index must be 2^j + 1");
    require(index < m_trace_evaluators.size(), "automorphism index out of range");
    m_trace_evaluators[index]->Eval(output, input);
}

LWEtoRLWEPackingContext::LWEtoRLWEPackingContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N,
                                                 uint64_t basis, uint64_t digits, double std)
    : m_auto_context(std::make_shared<AutomorphismContext>(source_key_distribution, modulus, N, basis, digits, std, 3)) {}

LWEtoRLWEPackingContext::LWEtoRLWEPackingContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt,
                                                 uint64_t basis, uint64_t digits, double std)
    : m_auto_context(std::make_shared<AutomorphismContext>(source_key_distribution, std::move(ntt), basis, digits, std, 3)) {}

LWEtoRLWEPackingContext::LWEtoRLWEPackingContext(const LWEtoRLWEPackingContext& other)
    : OperatorContext<LWEtoRLWEPacker>(other), enable_shared_from_this(other),
      m_auto_context(std::make_shared<AutomorphismContext>(*other.m_auto_context)) {}

long double LWEtoRLWEPackingContext::ComputeOutputVariance(long double) const {
    const long double N = (long double)GetDimension();
    auto in = std::dynamic_pointer_cast<RLWEContainerImpl>(m_auto_context->GetInputContainer());
    return in->GetVariance() * (N * N / 3.0L);
}

Container LWEtoRLWEPackingContext::GetInputContainer() const {
    return std::make_shared<VectorContainerImpl>(std::make_shared<LWEContainerImpl>(GetDimension(), GetModulus(), 0.0),
                                                 GetDimension());
}

Container LWEtoRLWEPackingContext::GetOutputContainer(Container) const {
    return std::make_shared<RLWEContainerImpl>(GetDimension(), GetModulus(), ComputeOutputVariance());
}

std::unique_ptr<LWEtoRLWEPacker> LWEtoRLWEPackingContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    std::vector<uint64_t> k;
    DeviceContext dev = auto_key_context(*m_auto_context, keys, k);
    return std::unique_ptr<LWEtoRLWEPacker>(new LWEtoRLWEPacker(shared_from_this(), dev, std::move(k)));
}

LWEtoRLWEPacker::LWEtoRLWEPacker(std::shared_ptr<const LWEtoRLWEPackingContext> context, DeviceContext dev,
                                 std::vector<uint64_t> keys)
    : m_context(std::move(context)), m_dev(std::move(dev)), m_keys(std::move(keys)) {}

void LWEtoRLWEPacker::Run(uint64_t* output, const uint64_t* input, uint64_t len, bool consecutive) {
    const uint64_t N = m_context->GetDimension();
    tp_context* x = m_dev.get();
    DevBuf d_in(x, len * (N + 1)), d_out(x, 2 * N);
    d_in.put(input, len * (N + 1));
    if (consecutive) check(tp_pack_consecutive_device(x, len, d_in.p, d_out.p), "PackConsecutively");
    else check(tp_pack_device(x, len, d_in.p, nullptr, d_out.p), "Pack");
    d_out.get(output, 2 * N);
}

void LWEtoRLWEPacker::Pack(uint64_t* output, const uint64_t* input, uint64_t len) { Run(output, input, len, false); }

void LWEtoRLWEPacker::Pack(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) {
    // lwe_to_rlwe_packing.cpp:253-258 takes len = input.size() / N
    const uint64_t N = m_context->GetDimension();
    Run(output.data(), input.data(), input.size() / N, false);
}

void LWEtoRLWEPacker::PackConsecutively(uint64_t* output, const uint64_t* input, uint64_t len) {
    Run(output, input, len, true);
}

// ------------------------------------------------------------------ function evaluation

NegacyclicFunctionEvaluationContext::NegacyclicFunctionEvaluationContext(
    std::shared_ptr<OperatorContext<BlindRotator>> blind_rotation_context)
    : m_rotation_context(std::move(blind_rotation_context)) {}

long double NegacyclicFunctionEvaluationContext::ComputeOutputVariance(long double input_variance) const {
    return m_rotation_context->ComputeOutputVariance(input_variance);
}

Container NegacyclicFunctionEvaluationContext::GetInputContainer() const {
    auto in = std::dynamic_pointer_cast<TupleContainerImpl>(m_rotation_context->GetInputContainer());
    require(in != nullptr, "rotation context has no tuple input container");
    return in->GetElem(0);
}

Container NegacyclicFunctionEvaluationContext::GetOutputContainer(Container) const {
    // as the reference (negacyclic_function_evaluator.cpp:34-41)
    return m_rotation_context->GetInputContainer();
}

std::unique_ptr<NegacyclicFunctionEvaluator> NegacyclicFunctionEvaluationContext::ConstructOperator(
    const std::vector<GenericKey>& keys) const {
    auto rotator = m_rotation_context->ConstructOperator(keys);
    return std::unique_ptr<NegacyclicFunctionEvaluator>(new NegacyclicFunctionEvaluator(shared_from_this(), std::move(rotator)));
}

NegacyclicFunctionEvaluator::NegacyclicFunctionEvaluator(std::shared_ptr<const NegacyclicFunctionEvaluationContext> ctx,
                                                         std::unique_ptr<BlindRotator> rotator)
    : m_context(std::move(ctx)), m_rotator(std::move(rotator)) {}

namespace {

void rotator_dims(BlindRotator& r, uint64_t& n, uint64_t& N) {
    if (tp_context* dev = r.DeviceContext()) {
        tp_params p{};
        check(tp_context_params(dev, &p), "tp_context_params");
        n = p.lwe_dim;
        N = p.ring_dim;
        return;
    }
    auto in = std::dynamic_pointer_cast<TupleContainerImpl>(r.GetContext()->GetInputContainer());
    require(in != nullptr, "rotator context has no tuple input container");
    n = std::dynamic_pointer_cast<LWEContainerImpl>(in->GetElem(0))->GetN();
    N = std::dynamic_pointer_cast<RLWEContainerImpl>(in->GetElem(1))->GetN();
}

}  // namespace

void NegacyclicFunctionEvaluator::EvalFunc(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED input_lwe,
                                           std::function<uint64_t(uint64_t)>& function) {
    uint64_t n = 0, N = 0;
    rotator_dims(*m_rotator, n, N);
    std::vector<uint64_t> lut(2 * N, 0);
    lut[0] = function(0);
    for (uint64_t i = 1; i < N; ++i) lut[N + i] = gl::sub(0, function(N - i) % kQ);
    EvalFunc(result, input_lwe, lut.data());
}

void NegacyclicFunctionEvaluator::EvalFunc(std::vector<uint64_t>& result, const std::vector<uint64_t>& input,
                                           std::function<uint64_t(uint64_t)>& function) {
    EvalFunc(result.data(), input.data(), function);
}

void NegacyclicFunctionEvaluator::EvalFunc(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED input_lwe,
                                           uint64_t* RESTRICTED lut_rlwe) {
    uint64_t n = 0, N = 0;
    rotator_dims(*m_rotator, n, N);
    tp_context* dev = m_rotator->DeviceContext();
    if (!dev) {
        // a rotator of another library: its own BlindRotate, then SampleExtract(0)
        std::vector<uint64_t> acc(2 * N, 0);
        m_rotator->BlindRotate(acc.data(), input_lwe, lut_rlwe, true);
        SampleExtract<uint64_t>(result, acc.data(), 0, N, kQ);
        return;
    }
    EvalFuncBatch(result, input_lwe, 1, lut_rlwe, 0);
}

void NegacyclicFunctionEvaluator::EvalFunc(std::vector<uint64_t>& result, const std::vector<uint64_t>& input,
                                           std::vector<uint64_t>& lut_rlwe) {
    EvalFunc(result.data(), input.data(), lut_rlwe.data());
}

void NegacyclicFunctionEvaluator::EvalFuncBatch(uint64_t* results, const uint64_t* input_lwes, uint64_t count,
                                                const uint64_t* lut, uint64_t lut_stride) {
    uint64_t n = 0, N = 0;
    rotator_dims(*m_rotator, n, N);
    tp_context* dev = m_rotator->DeviceContext();
    require(dev != nullptr, "EvalFuncBatch needs a rotator of this library");
    require(lut_stride == 0 || lut_stride == 2 * N, "lut_stride must be 0 or 2N");
    if (count == 0) return;
    check_lwe_entries(input_lwes, count * (n + 1), N);
    const uint64_t lut_words = lut_stride ? count * 2 * N : 2 * N;
    DevBuf d_lwe(dev, count * (n + 1)), d_lut(dev, lut_words), d_out(dev, count * (N + 1));
    d_lwe.put(input_lwes, count * (n + 1));
    d_lut.put(lut, lut_words);
    check(tp_eval_func_batch_device(dev, count, d_lwe.p, d_lut.p, lut_stride, d_out.p, nullptr, 0), "EvalFunc");
    d_out.get(results, count * (N + 1));
}

// ------------------------------------------------------------------ LWE -> RGSW

SchemeSwitchingContext::SchemeSwitchingContext(std::shared_ptr<OperatorContext<BlindRotator>> blind_rotation_context,
                                               std::shared_ptr<TraceEvaluationContext> trace_params,
                                               std::shared_ptr<RLWEConversionContext> squaring_params,
                                               uint64_t output_digits, uint64_t basis)
    : m_rotation_context(std::move(blind_rotation_context)), m_trace_params(std::move(trace_params)),
      m_squaring_params(std::move(squaring_params)), m_output_digits(output_digits), m_output_basis(basis),
      m_output_basis_log2(int_log2(basis)) {}

void SchemeSwitchingContext::SetOutputBasis(uint64_t basis) {
    m_output_basis = basis;
    m_output_basis_log2 = int_log2(basis);
}

long double SchemeSwitchingContext::ComputeOutputVariance(long double input_variance) const {
    // lwe_to_rgsw_conversion.cpp:45-51: rotation, then trace, then squaring switch
    const long double o = m_rotation_context->ComputeOutputVariance(input_variance);
    return m_squaring_params->ComputeOutputVariance(m_trace_params->ComputeOutputVariance(o));
}

Container SchemeSwitchingContext::GetInputContainer() const {
    auto br = std::dynamic_pointer_cast<TupleContainerImpl>(m_rotation_context->GetInputContainer());
    require(br != nullptr, "rotation context has no tuple input container");
    return br->GetElem(0);
}

Container SchemeSwitchingContext::GetOutputContainer(Container) const {
    return std::make_shared<RGSWContainerImpl>(m_squaring_params->GetDimension(), m_squaring_params->GetModulus(),
                                               m_output_digits, m_output_basis, ComputeOutputVariance());
}

std::unique_ptr<LWEtoRGSWConverter> SchemeSwitchingContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    // keys = {LWE secret, RLWE secret, ...}; the trace takes keys[1..], the squaring switch s^2 -> s
    require(keys.size() >= 2, "scheme-switching bundle is {LWE secret, RLWE secret}");
    std::vector<GenericKey> trace_key(keys.begin() + 1, keys.end());
    auto rotator = m_rotation_context->ConstructOperator(keys);
    auto tracer = m_trace_params->ConstructOperator(trace_key);
    const uint64_t N = m_trace_params->GetDimension();
    require(trace_key[0].GetKey().size() >= N, "RLWE secret too short");
    auto ntt = m_trace_params->GetNTT();
    std::vector<uint64_t> sq(N);
    ntt->ForwardNTT(sq.data(), trace_key[0].GetKeyPtr());
    ntt->MulEqModElw(sq.data(), sq.data(), N);  // the reference squares in place (EltwiseMultMod with aliases)
    ntt->BackwardNTT(sq.data(), sq.data());
    std::vector<GenericKey> square_keys = {{"SQUARE", sq.data(), N}, {"DEFAULT", keys[1].GetKeyPtr(), N}};
    auto squarer = m_squaring_params->ConstructOperator(square_keys);
    return std::make_unique<LWEtoRGSWConverter>(shared_from_this(), std::move(rotator), std::move(tracer),
                                                std::move(squarer));
}

LWEtoRGSWConverter::LWEtoRGSWConverter(std::shared_ptr<const SchemeSwitchingContext> context,
                                       std::unique_ptr<BlindRotator> rotator, std::unique_ptr<TraceEvaluator> trace_eval,
                                       std::unique_ptr<RLWEtoRLWEConverter> squaring_eval)
    : m_params(std::move(context)), m_rotator(std::move(rotator)), m_trace_eval(std::move(trace_eval)),
      m_square_eval(std::move(squaring_eval)) {
    // accumulator of lwe_to_rgsw_conversion.cpp:116-161: block-wise staircase of the output gadget,
    // centred on 0, halved; acc = [0 | -t/2], extra = t/2
    const uint64_t N = m_params->GetTraceContext()->GetDimension();
    const uint64_t digits = m_params->GetOutputDigits(), bb = m_params->GetOutputBasisLog2();
    require(digits >= 1 && N % digits == 0, "output digits must divide N");
    const uint64_t block = N / digits;
    const uint64_t max_digits = 64 / bb + (64 % bb ? 1 : 0);  // ceil(log2 Q / bb) for the Goldilocks Q
    require(digits <= max_digits, "more output digits than the modulus holds");
    m_acc.assign(3 * N, 0);
    m_extraction_buffer.assign(4 * N, 0);
    uint64_t scale = 1ULL << (bb * (max_digits - digits));
    for (uint64_t i = 0; i < block / 2; ++i) {
        m_acc[N + i] = scale;
        m_acc[N + N - i - 1] = kQ - scale;
    }
    for (uint64_t d = 1; d < digits; ++d) {
        scale <<= bb;
        for (uint64_t j = 0; j < block; ++j) m_acc[N + d * block + j - block / 2] = scale;
    }
    auto ntt = m_params->GetSquaringContext()->GetNTT();
    ntt->MulSEqModElw(m_acc.data() + N, gl::inv(2), N);
    std::copy(m_acc.begin() + N, m_acc.begin() + 2 * N, m_acc.begin() + 2 * N);
    ntt->MulSModElw(m_acc.data() + N, m_acc.data() + 2 * N, kQ - 1, N);
}

void LWEtoRGSWConverter::Convert(uint64_t* output, const uint64_t* const input) {
    auto sq = m_square_eval->GetContext();
    const uint64_t N = sq->GetDimension(), digits = m_params->GetOutputDigits(), block = N / digits;
    auto ntt = sq->GetNTT();
    uint64_t* buffer = m_extraction_buffer.data();
    uint64_t* const lev_left = output;
    uint64_t* const lev_right = output + 2 * N * digits;
    std::vector<uint64_t> acc(m_acc.begin(), m_acc.begin() + 2 * N);
    m_rotator->BlindRotate(lev_right, input, acc.data(), true);
    ntt->AddEqModElw(lev_right + N, m_acc.data() + 2 * N, N);
    // buffer = [A, -A, B, -B]: row d is rows 0 rotated by X^(-d block)
    std::fill(m_extraction_buffer.begin(), m_extraction_buffer.end(), 0);
    std::copy(lev_right, lev_right + N, buffer);
    std::copy(lev_right + N, lev_right + 2 * N, buffer + 2 * N);
    ntt->SubEqModElw(buffer + N, buffer, N);
    ntt->SubEqModElw(buffer + 3 * N, buffer + 2 * N, N);
    for (uint64_t d = 1; d < digits; ++d) {
        uint64_t* row = lev_right + d * 2 * N;
        std::copy(buffer + d * block, buffer + d * block + N, row);
        std::copy(buffer + 2 * N + d * block, buffer + 2 * N + d * block + N, row + N);
    }
    for (uint64_t d = 0; d < digits; ++d) {
        uint64_t* row = lev_right + d * 2 * N;
        m_trace_eval->Eval(row, row);
    }
    // RLWE(-s m) from RLWE(m): the switch from s^2 to s of (a, 0), then (a' + b, b') of its negation
    std::fill(m_extraction_buffer.begin(), m_extraction_buffer.end(), 0);
    for (uint64_t i = 0; i < digits; ++i) {
        uint64_t* m = lev_right + 2 * N * i;
        uint64_t* sm = lev_left + 2 * N * i;
        ntt->BackwardNTT(buffer, m);
        m_square_eval->Convert(sm, buffer);
        ntt->SubModElw(sm, buffer + 2 * N, sm, 2 * N);
        ntt->AddModElw(sm, m + N, sm, N);
    }
}

void LWEtoRGSWConverter::Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) {
    Convert(output.data(), input.data());
}

// ------------------------------------------------------------------ SAP digit decomposition

long double EstimateFailureProbability(long double variance, uint64_t modulus, uint64_t plaintext_modulus) {
    const long double lQ = (long double)modulus, lP = (long double)plaintext_modulus;
    return std::erfc(lQ / (2 * lP * std::sqrt(variance * 2)));
}

SAPDecompositionContext::SAPDecompositionContext(std::shared_ptr<OperatorContext<BlindRotator>> blind_rotation_context,
                                                 std::shared_ptr<TraceEvaluationContext> trace_context,
                                                 std::shared_ptr<LWEConversionContext> lwe_conversion_context,
                                                 uint64_t default_radix)
    : m_rotation_context(std::move(blind_rotation_context)), m_packing_context(std::move(trace_context)),
      m_conversion_context(std::move(lwe_conversion_context)), m_default_radix(default_radix) {
    m_restart_iter = ComputeRestartIterationForRadix(default_radix, IMPLICIT_SK_L0);
}

void SAPDecompositionContext::SetDefaultRadix(uint64_t new_radix) {
    m_default_radix = new_radix;
    m_restart_iter = ComputeRestartIterationForRadix(new_radix, IMPLICIT_SK_L0);
}

Container SAPDecompositionContext::GetInputContainer() const {
    // the conversion's output dimension, modulus unknown until the values are seen (:36-45)
    auto out = std::dynamic_pointer_cast<LWEContainerImpl>(
        m_conversion_context->GetOutputContainer(m_conversion_context->GetInputContainer()));
    require(out != nullptr, "conversion context has no LWE output container");
    return std::make_shared<LWEContainerImpl>(out->GetN(), UINT64_MAX, 0.0);
}

long double SAPDecompositionContext::ComputeOutputVariance(long double) const {
    // :47-60: per restart period, rotation -> cluster of N radix^2 / 12 -> key switch
    const long double N = (long double)m_packing_context->GetDimension();
    long double out_var = 0.0L, cluster_var = 0.0L;
    for (uint64_t i = 0; i < m_restart_iter; ++i) {
        const long double tmp_var = m_rotation_context->ComputeOutputVariance(out_var);
        const long double next = tmp_var * N * std::pow((long double)m_default_radix, 2.0L) / 12.0L;
        const long double next_out = m_conversion_context->ComputeOutputVariance(next);
        const bool fixed = i > 0 && next == cluster_var && next_out == out_var;
        cluster_var = next;
        out_var = next_out;
        if (fixed) break;
    }
    return cluster_var;
}

Container SAPDecompositionContext::GetOutputContainer(Container) const {
    auto out = std::dynamic_pointer_cast<RLWEContainerImpl>(
        m_rotation_context->GetOutputContainer(m_rotation_context->GetInputContainer()));
    require(out != nullptr, "rotation context has no RLWE output container");
    return std::make_shared<LWEContainerImpl>(out->GetN(), out->GetQ(), ComputeOutputVariance(0.0));
}

const uint64_t SAPDecompositionContext::ComputeRestartIterationForRadix(uint64_t radix, uint64_t sk_hamming) const {
    // :81-152, step for step
    const long double log_max_p_fail = -60;
    long double last_state_extraction_var = 0, current_var = 0;
    const uint64_t Q = m_packing_context->GetModulus(), N = m_packing_context->GetDimension();
    int iter = 0;
    const uint64_t worst_trunc_factor = radix, state_elem_bound = N / radix;
    while (true) {
        const long double blind_rotation_var = m_rotation_context->ComputeOutputVariance(current_var);
        const long double trunc_addition_var = worst_trunc_factor * blind_rotation_var;
        current_var = m_packing_context->ComputeOutputVariance(trunc_addition_var);
        if (current_var == 0.0L) return UINT64_MAX;
        const uint64_t factor = std::max(radix, state_elem_bound);
        const long double extraction_var = (blind_rotation_var * N * factor * factor) / 12.0L;
        const long double p_fail = EstimateFailureProbability(extraction_var, Q, factor);
        // the reference loops for as long as the estimate stays below 2^-60; a cap stands in for
        // its non-terminating case (an estimate that never grows)
        if (iter >= (1 << 20)) return UINT64_MAX;
        if (p_fail <= 0.0L) {
            iter++;
            continue;
        }
        if (std::log2(p_fail) <= log_max_p_fail) {
            iter++;
            last_state_extraction_var = extraction_var;
            continue;
        }
        break;
    }
    const long double lwe_conversion_var = m_conversion_context->ComputeOutputVariance(last_state_extraction_var);
    const long double ms_var = EstimateModulusSwitchingVariance(lwe_conversion_var, Q, 2 * N, BINARY, sk_hamming);
    const long double lwe_p_fail = EstimateFailureProbability(ms_var, 2 * N, state_elem_bound);
    if (lwe_p_fail <= 0.0L) return (uint64_t)iter;
    if (std::log2(lwe_p_fail) <= log_max_p_fail) return (uint64_t)iter;
    return (uint64_t)(iter - 1);  // the reference returns iter - 1 as uint64 (UINT64_MAX for iter 0)
}

std::unique_ptr<SAPDecomposer> SAPDecompositionContext::ConstructOperator(const std::vector<GenericKey>& keys) const {
    require(keys.size() >= 2, "SAP key bundle is {LWE secret, RLWE secret}");
    std::vector<GenericKey> conv_keys = {GenericKey(keys[1].GetKey()), GenericKey(keys[0].GetKey())};
    std::vector<GenericKey> pack_keys = {GenericKey(keys[1].GetKey())};
    uint64_t lwe_l0 = 0;
    for (uint64_t v : keys[0].GetKey()) lwe_l0 += v != 0;
    auto rotator = m_rotation_context->ConstructOperator(keys);
    auto packer = m_packing_context->ConstructOperator(pack_keys);
    auto conv = m_conversion_context->ConstructOperator(conv_keys);
    return std::unique_ptr<SAPDecomposer>(new SAPDecomposer(shared_from_this(), std::move(rotator), std::move(packer),
                                                            std::move(conv), m_default_radix, lwe_l0, m_restart_iter));
}

SAPDecomposer::SAPDecomposer(std::shared_ptr<const SAPDecompositionContext> ctx, std::unique_ptr<BlindRotator> rotator,
                             std::unique_ptr<TraceEvaluator> packer, std::unique_ptr<LWEtoLWEConverter> conv,
                             uint64_t max_radix, uint64_t lwe_sk_hamming_weight, uint64_t reset_period)
    : m_context(std::move(ctx)), m_rotator(std::move(rotator)), m_packer(std::move(packer)),
      m_lwe_converter(std::move(conv)), m_beta(lwe_sk_hamming_weight), m_max_radix(max_radix),
      m_restart_iteration(reset_period) {
    // one device context with the rotation key, the automorphism keys and the reset's switching key
    tp_context* rot = m_rotator->DeviceContext();
    require(rot != nullptr, "SAP decomposition needs a GPU blind rotator of this library");
    const std::vector<uint64_t>* brk = nullptr;
    if (auto* c = dynamic_cast<CGGIBlindRotator*>(m_rotator.get())) brk = &c->GetBlindRotationKey();
    else if (auto* b = dynamic_cast<BMMPBlindRotator*>(m_rotator.get())) brk = &b->GetBlindRotationKey();
    require(brk != nullptr, "SAP decomposition needs a CGGI or BMMP rotator");
    tp_params p{};
    check(tp_context_params(rot, &p), "tp_context_params");
    auto trace = m_context->GetTraceContext();
    auto conv_ctx = m_lwe_converter->GetContext();
    require(trace->GetDimension() == p.ring_dim, "trace and rotation ring dimensions differ");
    require(conv_ctx->GetSourceDimension() == p.ring_dim, "reset key switch must start from dimension N");
    p.rlwe_ks_base_log2 = (uint32_t)trace->GetGadgetBasisLog2();
    p.rlwe_ks_digits = (uint32_t)trace->GetGadgetDigits();
    p.ks_out_dim = (uint32_t)conv_ctx->GetTargetDimension();
    p.ks_base_log2 = (uint32_t)conv_ctx->GetGadgetBasisLog2();
    p.ks_digits = (uint32_t)conv_ctx->GetGadgetDigits();
    m_dev = make_device_context(p);
    check(tp_import_brk(m_dev.get(), brk->data(), brk->size()), "tp_import_brk");
    const auto& ak = m_packer->GetAutomorphismKeys();
    check(tp_import_auto_keys(m_dev.get(), ak.data(), ak.size()), "tp_import_auto_keys");
    const auto& ksk = m_lwe_converter->GetSwitchingKey();
    check(tp_import_lwe_ksk(m_dev.get(), ksk.data(), ksk.size()), "tp_import_lwe_ksk");
}

uint64_t SAPDecomposer::RestartFor(uint64_t radix) const {
    // :437-442
    if (radix == m_max_radix && m_beta <= SAPDecompositionContext::IMPLICIT_SK_L0) return m_restart_iteration;
    return m_context->ComputeRestartIterationForRadix(radix, m_beta);
}

void SAPDecomposer::DecomposeBatch(uint64_t* output, const uint64_t* inputs, uint64_t count, uint64_t radix,
                                   uint64_t max_digits, uint32_t* digit_counts) {
    tp_params p{};
    check(tp_context_params(m_dev.get(), &p), "tp_context_params");
    const uint64_t N = p.ring_dim, n1 = (uint64_t)p.lwe_dim + 1;
    require(max_digits >= 1 && max_digits <= 64, "max_digits must lie in [1, 64]");
    tp_context* x = m_dev.get();
    DevBuf d_in(x, count * n1), d_out(x, count * max_digits * (N + 1)), d_cnt(x, (count + 1) / 2);
    d_in.put(inputs, count * n1);
    d_out.put(output, count * max_digits * (N + 1));  // rows an input does not reach keep their words
    check(tp_sap_decompose_device(x, count, (uint32_t)radix, (uint32_t)m_beta, RestartFor(radix), d_in.p, d_out.p,
                                  (uint32_t)max_digits, reinterpret_cast<uint32_t*>(d_cnt.p)),
          "SAP Decompose");
    d_out.get(output, count * max_digits * (N + 1));
    if (digit_counts) {
        std::vector<uint64_t> c((count + 1) / 2);
        d_cnt.get(c.data(), c.size());
        std::memcpy(digit_counts, c.data(), count * sizeof(uint32_t));
    }
}

void SAPDecomposer::Decompose(uint64_t* output, const uint64_t* const input, uint64_t radix) {
    tp_params p{};
    check(tp_context_params(m_dev.get(), &p), "tp_context_params");
    const uint64_t N = p.ring_dim, W = N + 1;
    std::vector<uint64_t> digits(64 * W, 0);
    uint32_t cnt[2] = {0, 0};
    DecomposeBatch(digits.data(), input, 1, radix, 64, cnt);
    // SampleExtract leaves out[i] alone where the coefficient it negates is 0 (sample_extraction.h:29-33);
    // such words are the only zeros among out[1..N) that it would write
    for (uint32_t k = 0; k < cnt[0]; ++k) {
        const uint64_t* d = digits.data() + k * W;
        uint64_t* o = output + k * W;
        o[0] = d[0];
        for (uint64_t i = 1; i < N; ++i)
            if (d[i] != 0) o[i] = d[i];
        o[N] = d[N];
    }
}

void SAPDecomposer::Decompose(std::vector<uint64_t>& output, const std::vector<uint64_t>& input, uint64_t radix) {
    Decompose(output.data(), input.data(), radix);
}

void SAPDecomposer::HomTrunc(uint64_t* output, uint64_t* input, uint64_t radix) {
    require(radix && (radix & (radix - 1)) == 0, "radix must be a power of two");
    tp_params p{};
    check(tp_context_params(m_dev.get(), &p), "tp_context_params");
    const uint64_t N = p.ring_dim;
    tp_context* x = m_dev.get();
    DevBuf d_in(x, 2 * N), d_out(x, 2 * N);
    d_in.put(input, 2 * N);
    check(tp_sap_homtrunc_device(x, 1, (uint32_t)radix, d_in.p, d_out.p), "SAP HomTrunc");
    d_out.get(output, 2 * N);
}

void SAPDecomposer::ResetAccumulator(uint64_t* acc, uint64_t radix) {
    tp_params p{};
    check(tp_context_params(m_dev.get(), &p), "tp_context_params");
    const uint64_t N = p.ring_dim;
    tp_context* x = m_dev.get();
    DevBuf d(x, 2 * N);
    d.put(acc, 2 * N);
    check(tp_sap_reset_device(x, 1, radix, d.p), "SAP ResetAccumulator");
    d.get(acc, 2 * N);
}
