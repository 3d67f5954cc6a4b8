// C++ drop-in for the bootstrapping path of LSchild/FHEToothpaste, over the C-ABI of
// toothpaste_amd.h.  Class names, method signatures, key-bundle orders and buffer semantics follow
// the reference headers, so code written against them compiles and behaves the same:
//   OperatorContext, BlindRotator, SchemeConverter, FunctionEvaluator
//       lib/include/interfaces/{operator_context,blindrotation_operator,conversion_operator,functional_bootstrap}.h
//   MathWorker, SelectWorker                          lib/include/backend/backend.h:13-68
//   CGGIBlindRotationContext / CGGIBlindRotator       lib/include/operators/cggi_blind_rotation.h
//   BMMPBlindRotationContext / BMMPBlindRotator       lib/include/operators/bmmp_blind_rotation.h
//   LWEConversionContext / LWEtoLWEConverter,
//   RLWEConversionContext / RLWEtoRLWEConverter       lib/include/operators/endo_glwe_conversion.h
//   AutomorphismContext / AutomorphismEvaluator       lib/include/operators/automorphism_evaluation.h
//   TraceEvaluationContext / TraceEvaluator           lib/include/operators/trace_evaluation.h
//   LWEtoRLWEPackingContext / LWEtoRLWEPacker         lib/include/operators/lwe_to_rlwe_packing.h
//   NegacyclicFunctionEvaluationContext / ...Evaluator lib/include/operators/negacyclic_function_evaluator.h
//   GenericKey, containers, enums, SampleExtract      lib/include/static/{keys,container_types,enum_ids,sample_extraction}.h
// Every operator runs on the GPU (gfx950) through libtoothpaste_amd.so; there is no CPU path.
//
// Differences a caller can observe:
//   * Keys are generated inside ConstructOperator from the given secret keys with the library's
//     sampler instead of OpenFHE's: by default from ChaCha20 keyed by the operating system's
//     random source (GetSeed() returns TP_SEED_OS_ENTROPY); SetSeed(s) switches the context to a
//     reproducible stream, for tests and parity checks only (its outputs reveal its state).  Binary CGGI keys use the MakeRGSW gadget weights,
//     ternary/BMMP/RLWE keys the signed decomposer's weights 2^(64-b(j+1)).
//   * Only the Goldilocks modulus Q = 2^64 - 2^32 + 1 and ring dimensions 1024, 2048, 4096 run
//     here; other values throw std::invalid_argument (the reference falls back to HEXL).
//   * Errors throw (std::invalid_argument for bad arguments, std::runtime_error for device errors)
//     where the reference asserts.
//   * Batched entry points are added: CGGIBlindRotator/BMMPBlindRotator::BlindRotateBatch,
//     NegacyclicFunctionEvaluator::EvalFuncBatch, LWEtoLWEConverter::ConvertBatch.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "toothpaste_amd.h"

#ifndef RESTRICTED
#define RESTRICTED __restrict__
#endif

// ------------------------------------------------------------------ enums, containers, keys

enum PolyFormat { COEF, NTT };
enum KeyDistribution { BINARY, TERNARY, GAUSSIAN };
enum BlindRotationMethod { CGGI, BMMP };
enum OperatorID {
    BR_CGGI,
    BR_BMMP,
    CONV_RLWE_RLWE,
    CONV_LWE_LWE,
    CONV_LWE_RGSW,
    DECOMP_SAP,
    EVAL_AUTO,
    EVAL_TRACE,
    PACK_LWE_TO_RLWE,
    FUNC_BOOT_KS,
    NEGACYCLIC_FUNC_BOOT
};
enum ContainerLabel { RLWE, RGSW, LWE, MLWE, GLWE, NTRU, NTRUVec, Tuple, Vector, Flag, Empty };

struct ContainerImpl {
    virtual ContainerLabel GetLabel() { return Empty; }
    virtual ~ContainerImpl() = default;
};
using Container = std::shared_ptr<ContainerImpl>;

struct RLWEContainerImpl : public ContainerImpl {
    RLWEContainerImpl(uint64_t N, uint64_t Q, long double var) : m_Q(Q), m_N(N), m_var(var) {}
    ContainerLabel GetLabel() override { return RLWE; }
    uint64_t GetQ() const { return m_Q; }
    uint64_t GetN() const { return m_N; }
    long double GetVariance() const { return m_var; }

private:
    uint64_t m_Q, m_N;
    long double m_var;
};

struct LWEContainerImpl : public ContainerImpl {
    LWEContainerImpl(uint64_t n, uint64_t Q, long double var) : m_q(Q), m_n(n), m_var(var) {}
    ContainerLabel GetLabel() override { return LWE; }
    uint64_t GetQ() const { return m_q; }
    uint64_t GetN() const { return m_n; }
    long double GetVariance() const { return m_var; }

private:
    uint64_t m_q, m_n;
    long double m_var;
};

struct RGSWContainerImpl : public ContainerImpl {
    RGSWContainerImpl(uint64_t N, uint64_t Q, uint64_t digits, uint64_t basis, long double var)
        : m_Q(Q), m_N(N), m_digits(digits), m_basis(basis), m_var(var) {}
    ContainerLabel GetLabel() override { return RGSW; }
    uint64_t GetQ() const { return m_Q; }
    uint64_t GetN() const { return m_N; }
    uint64_t GetDigits() const { return m_digits; }
    uint64_t GetBasis() const { return m_basis; }
    long double GetVariance() const { return m_var; }

private:
    uint64_t m_Q, m_N, m_digits, m_basis;
    long double m_var;
};

struct TupleContainerImpl : public ContainerImpl {
    explicit TupleContainerImpl(std::vector<Container> elems) : m_elems(std::move(elems)) {}
    ContainerLabel GetLabel() override { return Tuple; }
    Container GetElem(uint64_t i) const { return m_elems.at(i); }
    uint64_t GetSize() const { return m_elems.size(); }

private:
    std::vector<Container> m_elems;
};

struct VectorContainerImpl : public ContainerImpl {
    VectorContainerImpl(Container elem, uint64_t count) : m_elem(std::move(elem)), m_count(count) {}
    ContainerLabel GetLabel() override { return Vector; }
    Container GetElemType() const { return m_elem; }
    uint64_t GetCount() const { return m_count; }

private:
    Container m_elem;
    uint64_t m_count;
};

struct GenericKey {
    GenericKey(const std::string& label, const std::vector<uint64_t>& key_vals) : m_label(label), m_key(key_vals) {}
    GenericKey(const std::string& label, const uint64_t* key_ptr, uint64_t n) : m_label(label), m_key(key_ptr, key_ptr + n) {}
    GenericKey(const std::vector<uint64_t>& key_vals) : m_key(key_vals) {}  // NOLINT(google-explicit-constructor)
    GenericKey(const uint64_t* key_ptr, uint64_t n) : m_key(key_ptr, key_ptr + n) {}
    const uint64_t* GetKeyPtr() const { return m_key.data(); }
    const std::vector<uint64_t>& GetKey() const { return m_key; }
    const std::string& GetLabel() const { return m_label; }

private:
    std::string m_label;
    std::vector<uint64_t> m_key;
};

// SampleExtract (sample_extraction.h:20-58), with the reference's exact behaviour: for
// i <= idx out[i] = in[idx - i]; for i > idx out[i] = Q - in[N - i], written only when that
// coefficient is non-zero (so pass a zeroed output); out[N] = b_idx.  The vector overload does
// not write b.
template <typename IntType>
void SampleExtract(IntType* RESTRICTED out, IntType* RESTRICTED in, uint64_t idx, uint64_t N, IntType Q) {
    for (uint64_t i = 0; i <= idx; ++i) out[i] = in[idx - i];
    for (uint64_t i = idx + 1; i < N; ++i) {
        const IntType v = in[N - i];
        if (v != 0) out[i] = Q - v;
    }
    out[N] = in[N + idx];
}

template <typename IntType, typename Allocator>
void SampleExtract(std::vector<IntType, Allocator>& out, std::vector<IntType, Allocator>& in, uint64_t idx, uint64_t N,
                   IntType Q) {
    for (uint64_t i = 0; i <= idx; ++i) out[i] = in[idx - i];
    for (uint64_t i = idx + 1; i < N; ++i) {
        const IntType v = in[N - i];
        if (v != 0) out[i] = Q - v;
    }
}

// Modulus switching (lib/include/static/modulus_switching.h): vec[i] = R(target * vec[i] / source),
// computed on the GPU with the reference's x87 extended-precision rounding reproduced exactly.
// source >= 2^63, target a power of two; RANDOM draws its coin from a seeded stream.
enum struct ModulusSwitchType { ROUND, FLOOR, CEIL, RANDOM };
void ModulusSwitch(uint64_t* vec, uint64_t n, uint64_t source_modulus, uint64_t target_modulus,
                   ModulusSwitchType type);
long double EstimateModulusSwitchingVariance(long double input_variance, uint64_t source_modulus,
                                             uint64_t target_modulus, KeyDistribution kd, uint64_t expected_l0);

// ------------------------------------------------------------------ backend

struct MathWorker {
    virtual void ForwardNTT(uint64_t* result, const uint64_t* input) = 0;
    virtual void ForwardNTT(std::vector<uint64_t>& result, std::vector<uint64_t>& input) = 0;
    virtual void BackwardNTT(uint64_t* result, const uint64_t* input) = 0;
    virtual void BackwardNTT(std::vector<uint64_t>& result, std::vector<uint64_t>& input) = 0;
    virtual uint64_t AddMod(uint64_t left, uint64_t right) = 0;
    virtual uint64_t SubMod(uint64_t left, uint64_t right) = 0;
    virtual uint64_t MulMod(uint64_t left, uint64_t right) = 0;
    virtual uint64_t FMAMod(uint64_t left, uint64_t right, uint64_t summand) = 0;
    virtual uint64_t InvMod(uint64_t left) = 0;
    virtual uint64_t PowMod(uint64_t base, uint64_t exponent) = 0;
    virtual void AddModElw(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED left,
                           const uint64_t* RESTRICTED right, uint64_t len) = 0;
    virtual void AddEqModElw(uint64_t* RESTRICTED left_out, const uint64_t* RESTRICTED right, uint64_t len) = 0;
    virtual void SubModElw(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED left,
                           const uint64_t* RESTRICTED right, uint64_t len) = 0;
    virtual void SubEqModElw(uint64_t* RESTRICTED left_out, const uint64_t* RESTRICTED right, uint64_t len) = 0;
    virtual void MulModElw(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED left,
                           const uint64_t* RESTRICTED right, uint64_t len) = 0;
    virtual void MulEqModElw(uint64_t* RESTRICTED left_out, const uint64_t* RESTRICTED right, uint64_t len) = 0;
    virtual void FMAModElw(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED left,
                           const uint64_t* RESTRICTED right, const uint64_t* RESTRICTED summand, uint64_t len) = 0;
    virtual void FMAEqModElw(uint64_t* RESTRICTED left_out, const uint64_t* RESTRICTED right,
                             const uint64_t* RESTRICTED summand, uint64_t len) = 0;
    virtual void MulSModElw(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED left, uint64_t right,
                            uint64_t len) = 0;
    virtual void MulSEqModElw(uint64_t* RESTRICTED left_out, uint64_t right, uint64_t len) = 0;
    virtual void ReduceModElw(uint64_t* out, const uint64_t* in, uint64_t len) = 0;
    virtual uint64_t GetModulus() const = 0;
    virtual uint64_t GetDimension() const = 0;
    virtual ~MathWorker() = default;
};

// The GPU Goldilocks worker for (modulus, dimension); one cached instance per dimension
// (backend.cpp:12-34).  Throws std::invalid_argument for any other modulus.
std::shared_ptr<MathWorker> SelectWorker(uint64_t modulus, uint64_t dimension);
// Scalar and elementwise operations only (dimension 0): NTT calls throw.
std::shared_ptr<MathWorker> SelectWorkerArbitraryDimension(uint64_t modulus);

// ------------------------------------------------------------------ interfaces

template <typename OperatorType>
struct OperatorContext {
    [[nodiscard]] virtual long double ComputeOutputVariance(long double input_variance = 0.0) const = 0;
    [[nodiscard]] virtual std::unique_ptr<OperatorType> ConstructOperator(const std::vector<GenericKey>& keys) const = 0;
    [[nodiscard]] virtual Container GetInputContainer() const = 0;
    [[nodiscard]] virtual Container GetOutputContainer(Container input) const = 0;
    [[nodiscard]] virtual OperatorID GetOperatorID() const = 0;
    virtual ~OperatorContext() = default;
};

struct BlindRotator {
    virtual void BlindRotate(std::vector<uint64_t>& result, const std::vector<uint64_t>& lwe_ciphertext,
                             std::vector<uint64_t>& accumulator, bool output_as_coefficients = false) = 0;
    virtual void BlindRotate(uint64_t* result, const uint64_t* lwe_vec, uint64_t* accumulator,
                             bool output_as_coefficients = false) = 0;
    [[nodiscard]] virtual const std::shared_ptr<const OperatorContext<BlindRotator>> GetContext() const = 0;
    // Batched blind rotation of `count` LWEs ([count][n+1] in Z_2N) against one accumulator
    // (accumulator_stride 0) or one per input (stride 2N), results [count][2N].  The rotators of
    // this library run the batch in one launch; the default runs BlindRotate per input.
    virtual void BlindRotateBatch(uint64_t* results, const uint64_t* lwes, uint64_t count, const uint64_t* accumulators,
                                  uint64_t accumulator_stride, bool output_as_coefficients = false);
    // The device context a GPU rotator runs on (nullptr for other rotators).
    virtual tp_context* DeviceContext() const { return nullptr; }
    virtual ~BlindRotator() = default;
};

template <typename ContextType>
struct SchemeConverter {
    virtual void Convert(uint64_t* output, const uint64_t* const input) = 0;
    virtual void Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) = 0;
    virtual const std::shared_ptr<const ContextType> GetContext() const = 0;
    virtual ~SchemeConverter() = default;
};

template <typename ContextType>
struct FunctionEvaluator {
    virtual void EvalFunc(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED input_lwe,
                          std::function<uint64_t(uint64_t)>& function) = 0;
    virtual void EvalFunc(std::vector<uint64_t>& result, const std::vector<uint64_t>& input_lwe,
                          std::function<uint64_t(uint64_t)>& function) = 0;
    virtual void EvalFunc(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED input_lwe,
                          uint64_t* RESTRICTED lut_rlwe) = 0;
    virtual void EvalFunc(std::vector<uint64_t>& result, const std::vector<uint64_t>& input_lwe,
                          std::vector<uint64_t>& lut_rlwe) = 0;
    virtual const std::shared_ptr<const ContextType> GetContext() const = 0;
    virtual ~FunctionEvaluator() = default;
};

namespace toothpaste_amd {
// Owning handle of a device context (tp_context_destroy on release).
using DeviceContext = std::shared_ptr<tp_context>;
DeviceContext make_device_context(const tp_params& p);
// Seed for contexts that were not given one (std::random_device).
uint64_t fresh_seed();
}  // namespace toothpaste_amd

// ------------------------------------------------------------------ blind rotation

struct CGGIBlindRotator;
struct CGGIBlindRotationContext : public OperatorContext<CGGIBlindRotator>,
                                  public std::enable_shared_from_this<CGGIBlindRotationContext> {
    friend struct CGGIBlindRotator;
    explicit CGGIBlindRotationContext(KeyDistribution distr, uint64_t modulus, uint64_t ring_dim, uint64_t lwe_dim,
                                      uint64_t m_basis, uint64_t m_digits, double std);
    explicit CGGIBlindRotationContext(KeyDistribution distr, std::shared_ptr<MathWorker> ntt, uint64_t lwe_dim,
                                      uint64_t m_basis, uint64_t m_digits, double std);
    CGGIBlindRotationContext(const CGGIBlindRotationContext& other);
    void SetKeyDistribution(KeyDistribution dist) { m_distribution = dist; }
    void SetModulus(uint64_t modulus) { m_modulus = modulus; }
    void SetRingDimension(uint64_t ring_dim) { m_N = ring_dim; }
    void SetNTT(std::shared_ptr<MathWorker> ntt) { m_ntt = std::move(ntt); }
    void SetLWEDimension(uint64_t lwe_dim) { m_n = lwe_dim; }
    void SetBlindRotationBasis(uint64_t L);
    void SetBlindRotationRGSWDigits(uint64_t digits) { m_digits = digits; }
    void SetStd(double std) { m_std = std; }
    void SetSeed(uint64_t seed) { m_seed = seed; }
    [[nodiscard]] KeyDistribution GetKeyDistribution() const { return m_distribution; }
    [[nodiscard]] uint64_t GetModulus() const { return m_modulus; }
    [[nodiscard]] uint64_t GetRingDimension() const { return m_N; }
    [[nodiscard]] std::shared_ptr<MathWorker> GetNTT() const { return m_ntt; }
    [[nodiscard]] uint64_t GetLWEDimension() const { return m_n; }
    [[nodiscard]] uint64_t GetBlindRotationBasis() const { return m_basis; }
    [[nodiscard]] uint64_t GetBlindRotationBasisLog2() const { return m_basis_log2; }
    [[nodiscard]] uint64_t GetBlindRotationRGSWDigits() const { return m_digits; }
    [[nodiscard]] double GetStd() const { return m_std; }
    [[nodiscard]] uint64_t GetSeed() const { return m_seed; }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<CGGIBlindRotator> ConstructOperator(const std::vector<GenericKey>& bundle) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return BR_CGGI; }

protected:
    KeyDistribution m_distribution;
    uint64_t m_modulus, m_N, m_n, m_basis = 0, m_basis_log2 = 0, m_digits;
    double m_std;
    uint64_t m_seed;
    std::shared_ptr<MathWorker> m_ntt;
};

struct CGGIBlindRotator : public BlindRotator {
    friend struct CGGIBlindRotationContext;
    explicit CGGIBlindRotator(const std::shared_ptr<const CGGIBlindRotationContext>& params);
    void BlindRotate(std::vector<uint64_t>& result, const std::vector<uint64_t>& lwe_vec,
                     std::vector<uint64_t>& rlwe_acc_vec, bool output_as_coefficients = false) override;
    void BlindRotate(uint64_t* result, const uint64_t* lwe_vec, uint64_t* rlwe_acc_vec,
                     bool output_as_coefficients = false) override;
    void BlindRotateBatch(uint64_t* results, const uint64_t* lwes, uint64_t count, const uint64_t* accumulators,
                          uint64_t accumulator_stride, bool output_as_coefficients = false) override;
    [[nodiscard]] const std::shared_ptr<const OperatorContext<BlindRotator>> GetContext() const override;
    tp_context* DeviceContext() const override { return m_dev.get(); }
    // Blind-rotation key in the reference layout (n x RGSW, or 2n x RGSW for ternary keys).
    [[nodiscard]] const std::vector<uint64_t>& GetBlindRotationKey() const { return m_brk; }
    ~CGGIBlindRotator() override = default;

private:
    void KeyGen(const uint64_t* lwe_key, const uint64_t* rlwe_key);
    std::shared_ptr<const CGGIBlindRotationContext> m_params;
    toothpaste_amd::DeviceContext m_dev;
    std::vector<uint64_t> m_brk;
    bool m_keys_generated = false;
};

struct BMMPBlindRotator;
struct BMMPBlindRotationContext : public OperatorContext<BMMPBlindRotator>,
                                  public std::enable_shared_from_this<BMMPBlindRotationContext> {
    friend struct BMMPBlindRotator;
    explicit BMMPBlindRotationContext(KeyDistribution distr, uint64_t modulus, uint64_t ring_dim, uint64_t lwe_dim,
                                      uint64_t m_basis, uint64_t m_digits, double std, uint64_t step_size);
    explicit BMMPBlindRotationContext(KeyDistribution distr, std::shared_ptr<MathWorker> ntt, uint64_t lwe_dim,
                                      uint64_t m_basis, uint64_t m_digits, double std, uint64_t step_size);
    BMMPBlindRotationContext(const BMMPBlindRotationContext& other) = default;
    void SetSeed(uint64_t seed) { m_seed = seed; }
    [[nodiscard]] KeyDistribution GetKeyDistribution() const { return m_distribution; }
    [[nodiscard]] uint64_t GetModulus() const { return m_modulus; }
    [[nodiscard]] uint64_t GetRingDimension() const { return m_N; }
    [[nodiscard]] std::shared_ptr<MathWorker> GetNTT() const { return m_ntt; }
    [[nodiscard]] uint64_t GetLWEDimension() const { return m_n; }
    [[nodiscard]] uint64_t GetBlindRotationBasis() const { return m_basis; }
    [[nodiscard]] uint64_t GetBlindRotationBasisLog2() const { return m_basis_log2; }
    [[nodiscard]] uint64_t GetBlindRotationRGSWDigits() const { return m_digits; }
    [[nodiscard]] uint64_t GetStepSize() const { return m_step_size; }
    [[nodiscard]] double GetStd() const { return m_std; }
    [[nodiscard]] uint64_t GetSeed() const { return m_seed; }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<BMMPBlindRotator> ConstructOperator(const std::vector<GenericKey>& bundle) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return BR_BMMP; }

protected:
    KeyDistribution m_distribution;
    uint64_t m_modulus, m_N, m_n, m_basis, m_basis_log2, m_digits, m_step_size;
    double m_std;
    uint64_t m_seed;
    std::shared_ptr<MathWorker> m_ntt;
};

struct BMMPBlindRotator : public BlindRotator {
    friend struct BMMPBlindRotationContext;
    explicit BMMPBlindRotator(const std::shared_ptr<const BMMPBlindRotationContext>& params);
    void BlindRotate(std::vector<uint64_t>& result, const std::vector<uint64_t>& lwe_vec,
                     std::vector<uint64_t>& rlwe_acc_vec, bool output_as_coefficients = false) override;
    void BlindRotate(uint64_t* result, const uint64_t* lwe_vec, uint64_t* rlwe_acc_vec,
                     bool output_as_coefficients = false) override;
    void BlindRotateBatch(uint64_t* results, const uint64_t* lwes, uint64_t count, const uint64_t* accumulators,
                          uint64_t accumulator_stride, bool output_as_coefficients = false) override;
    [[nodiscard]] const std::shared_ptr<const OperatorContext<BlindRotator>> GetContext() const override;
    tp_context* DeviceContext() const override { return m_dev.get(); }
    [[nodiscard]] const std::vector<uint64_t>& GetBlindRotationKey() const { return m_brk; }
    ~BMMPBlindRotator() override = default;

private:
    void KeyGen(const uint64_t* lwe_key, const uint64_t* rlwe_key);
    std::shared_ptr<const BMMPBlindRotationContext> m_params;
    toothpaste_amd::DeviceContext m_dev;
    std::vector<uint64_t> m_brk;
    bool m_keys_generated = false;
};

// ------------------------------------------------------------------ key switching

struct LWEtoLWEConverter;
struct LWEConversionContext : public OperatorContext<LWEtoLWEConverter>,
                              public std::enable_shared_from_this<LWEConversionContext> {
    LWEConversionContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t source_dimension,
                         uint64_t target_dimension, uint64_t basis, uint64_t digits, double std);
    LWEConversionContext(LWEConversionContext& other) = default;
    void SetSeed(uint64_t seed) { m_seed = seed; }
    [[nodiscard]] KeyDistribution GetSourceKeyDistribution() const { return m_source_distribution; }
    [[nodiscard]] uint64_t GetModulus() const { return m_modulus; }
    [[nodiscard]] uint64_t GetSourceDimension() const { return m_source_dimension; }
    [[nodiscard]] uint64_t GetTargetDimension() const { return m_target_dimension; }
    [[nodiscard]] uint64_t GetGadgetBasis() const { return m_basis; }
    [[nodiscard]] uint64_t GetGadgetBasisLog2() const { return m_basis_log2; }
    [[nodiscard]] uint64_t GetGadgetDigits() const { return m_digits; }
    [[nodiscard]] double GetStd() const { return m_std; }
    [[nodiscard]] uint64_t GetSeed() const { return m_seed; }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<LWEtoLWEConverter> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return CONV_LWE_LWE; }

private:
    KeyDistribution m_source_distribution;
    uint64_t m_modulus, m_source_dimension, m_target_dimension, m_basis, m_basis_log2, m_digits;
    double m_std;
    uint64_t m_seed;
};

struct LWEtoLWEConverter : public SchemeConverter<LWEConversionContext> {
    friend struct LWEConversionContext;
    // output[0..n_out) -= acc, output[n_out] = input[n_in] - acc[n_out]  (lwe_conversion.cpp:205-239)
    void Convert(uint64_t* output, const uint64_t* const input) override;
    void Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) override;
    // Batched, overwrite semantics: outputs[c] = (0.., inputs[c][n_in]) - acc_c.
    void ConvertBatch(uint64_t* outputs, const uint64_t* inputs, uint64_t count);
    [[nodiscard]] const std::shared_ptr<const LWEConversionContext> GetContext() const override { return m_params; }
    [[nodiscard]] const std::vector<uint64_t>& GetSwitchingKey() const { return m_ksk; }
    tp_context* DeviceContext() const { return m_dev.get(); }

private:
    explicit LWEtoLWEConverter(std::shared_ptr<const LWEConversionContext> params);
    void KeyGen(const uint64_t* source_key, const uint64_t* target_key);
    std::shared_ptr<const LWEConversionContext> m_params;
    toothpaste_amd::DeviceContext m_dev;
    std::vector<uint64_t> m_ksk;
};

struct RLWEtoRLWEConverter;
struct RLWEConversionContext : public OperatorContext<RLWEtoRLWEConverter>,
                               public std::enable_shared_from_this<RLWEConversionContext> {
    RLWEConversionContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N, uint64_t basis,
                          uint64_t digits, double std);
    RLWEConversionContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt, uint64_t basis,
                          uint64_t digits, double std);
    RLWEConversionContext(const RLWEConversionContext& other) = default;
    void SetSeed(uint64_t seed) { m_seed = seed; }
    void SetStd(double std) { m_std = std; }
    [[nodiscard]] KeyDistribution GetSourceKeyDistribution() const { return m_source_distribution; }
    [[nodiscard]] uint64_t GetModulus() const { return m_modulus; }
    [[nodiscard]] uint64_t GetDimension() const { return m_N; }
    [[nodiscard]] std::shared_ptr<MathWorker> GetNTT() const { return m_ntt; }
    [[nodiscard]] uint64_t GetGadgetBasis() const { return m_basis; }
    [[nodiscard]] uint64_t GetGadgetBasisLog2() const { return m_basis_log2; }
    [[nodiscard]] uint64_t GetGadgetDigits() const { return m_digits; }
    [[nodiscard]] double GetStd() const { return m_std; }
    [[nodiscard]] uint64_t GetSeed() const { return m_seed; }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<RLWEtoRLWEConverter> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return CONV_RLWE_RLWE; }

private:
    KeyDistribution m_source_distribution;
    uint64_t m_modulus, m_N, m_basis, m_basis_log2, m_digits;
    double m_std;
    uint64_t m_seed;
    std::shared_ptr<MathWorker> m_ntt;
};

struct RLWEtoRLWEConverter : public SchemeConverter<RLWEConversionContext> {
    friend struct RLWEConversionContext;
    // input [a coef | b NTT]; output (NTT) -= sum_j NTT(dec_j(a)) KSK_j, output_b += b
    // (rlwe_conversion.cpp:224-249)
    void Convert(uint64_t* output, const uint64_t* const input) override;
    void Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) override;
    const std::shared_ptr<const RLWEConversionContext> GetContext() const override { return m_params; }
    [[nodiscard]] const std::vector<uint64_t>& GetSwitchingKey() const { return m_ksk; }

private:
    explicit RLWEtoRLWEConverter(std::shared_ptr<const RLWEConversionContext> params);
    void KeyGen(const uint64_t* source_key, const uint64_t* target_key);
    std::shared_ptr<const RLWEConversionContext> m_params;
    toothpaste_amd::DeviceContext m_dev;
    std::vector<uint64_t> m_ksk;
};

// ------------------------------------------------------------------ automorphism, trace, packing

struct AutomorphismEvaluator;
struct AutomorphismContext : public OperatorContext<AutomorphismEvaluator>,
                             public std::enable_shared_from_this<AutomorphismContext> {
    AutomorphismContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N, uint64_t basis,
                        uint64_t digits, double std, uint32_t automorphism_index);
    AutomorphismContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt, uint64_t basis,
                        uint64_t digits, double std, uint32_t automorphism_index);
    AutomorphismContext(const AutomorphismContext& other);
    void SetAutomorphismIndex(uint32_t t) { m_automorphism_index = t; }
    void SetSeed(uint64_t seed) { m_rlwe_conversion->SetSeed(seed); }
    void SetStd(double std) { m_rlwe_conversion->SetStd(std); }
    [[nodiscard]] uint32_t GetAutomorphismIndex() const { return m_automorphism_index; }
    [[nodiscard]] uint64_t GetDimension() const { return m_rlwe_conversion->GetDimension(); }
    [[nodiscard]] uint64_t GetModulus() const { return m_rlwe_conversion->GetModulus(); }
    [[nodiscard]] std::shared_ptr<MathWorker> GetNTT() const { return m_rlwe_conversion->GetNTT(); }
    [[nodiscard]] uint64_t GetGadgetBasisLog2() const { return m_rlwe_conversion->GetGadgetBasisLog2(); }
    [[nodiscard]] uint64_t GetGadgetDigits() const { return m_rlwe_conversion->GetGadgetDigits(); }
    [[nodiscard]] KeyDistribution GetSourceKeyDistribution() const { return m_rlwe_conversion->GetSourceKeyDistribution(); }
    [[nodiscard]] double GetStd() const { return m_rlwe_conversion->GetStd(); }
    [[nodiscard]] uint64_t GetSeed() const { return m_rlwe_conversion->GetSeed(); }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<AutomorphismEvaluator> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return EVAL_AUTO; }

private:
    std::shared_ptr<RLWEConversionContext> m_rlwe_conversion;
    uint32_t m_automorphism_index;
};

struct TraceEvaluator;
struct AutomorphismEvaluator {
    friend struct AutomorphismContext;
    friend struct TraceEvaluator;
    // input [a coef | b coef]; output (NTT) -= KS(tau_t(a)), output_b += NTT(tau_t(b))
    // (automorphism_evaluation.cpp:150-159)
    void Eval(uint64_t* output, const uint64_t* const input);
    void Eval(std::vector<uint64_t>& output, const std::vector<uint64_t>& input);
    // as Eval with b treated as zero (automorphism_evaluation.cpp:172-177)
    void EvalSwitchOnly(uint64_t* output, const uint64_t* const input);
    [[nodiscard]] const std::shared_ptr<const AutomorphismContext> GetContext() const { return m_params; }

private:
    AutomorphismEvaluator(std::shared_ptr<const AutomorphismContext> params, toothpaste_amd::DeviceContext dev,
                          int key_index);
    void Run(uint64_t* output, const uint64_t* input, int mode);
    std::shared_ptr<const AutomorphismContext> m_params;
    toothpaste_amd::DeviceContext m_dev;
    int m_key_index;  // 0: the evaluator's own key; j: automorphism key j of a shared trace/packing context
};

struct TraceEvaluator;
struct TraceEvaluationContext : public OperatorContext<TraceEvaluator>,
                                public std::enable_shared_from_this<TraceEvaluationContext> {
    TraceEvaluationContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N, uint64_t basis,
                           uint64_t digits, double std);
    TraceEvaluationContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt, uint64_t basis,
                           uint64_t digits, double std);
    TraceEvaluationContext(const TraceEvaluationContext& other);
    void SetSeed(uint64_t seed) { m_auto_context->SetSeed(seed); }
    [[nodiscard]] uint64_t GetDimension() const { return m_auto_context->GetDimension(); }
    [[nodiscard]] uint64_t GetModulus() const { return m_auto_context->GetModulus(); }
    [[nodiscard]] std::shared_ptr<MathWorker> GetNTT() const { return m_auto_context->GetNTT(); }
    [[nodiscard]] uint64_t GetGadgetBasisLog2() const { return m_auto_context->GetGadgetBasisLog2(); }
    [[nodiscard]] uint64_t GetGadgetDigits() const { return m_auto_context->GetGadgetDigits(); }
    [[nodiscard]] KeyDistribution GetSourceKeyDistribution() const { return m_auto_context->GetSourceKeyDistribution(); }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<TraceEvaluator> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return EVAL_TRACE; }
    [[nodiscard]] std::shared_ptr<AutomorphismContext> GetAutomorphismContext() const { return m_auto_context; }

private:
    std::shared_ptr<AutomorphismContext> m_auto_context;
};

struct TraceEvaluator {
    friend struct TraceEvaluationContext;
    // input [a|b] coefficients, output NTT domain (overwritten) (trace_evaluation.cpp:135-162)
    void Eval(uint64_t* output, const uint64_t* input);
    void Eval(std::vector<uint64_t>& output, const std::vector<uint64_t>& input);
    // One automorphism t = 2^j + 1, accumulated into output.  Like the reference, the pointer
    // overload uses evaluator j-1 and the vector overload evaluator j (trace_evaluation.cpp:172-188).
    void EvalAuto(uint64_t* output, const uint64_t* input, uint64_t auto_idx);
    void EvalAuto(std::vector<uint64_t>& output, const std::vector<uint64_t>& input, uint64_t auto_idx);
    [[nodiscard]] std::shared_ptr<const TraceEvaluationContext> GetContext() const { return m_context; }
    [[nodiscard]] const std::vector<uint64_t>& GetAutomorphismKeys() const { return m_keys; }

private:
    TraceEvaluator(std::shared_ptr<const TraceEvaluationContext> context, toothpaste_amd::DeviceContext dev,
                   std::vector<uint64_t> keys);
    std::shared_ptr<const TraceEvaluationContext> m_context;
    toothpaste_amd::DeviceContext m_dev;
    std::vector<uint64_t> m_keys;
    std::vector<std::unique_ptr<AutomorphismEvaluator>> m_trace_evaluators;
};

struct LWEtoRLWEPacker;
struct LWEtoRLWEPackingContext : public OperatorContext<LWEtoRLWEPacker>,
                                 public std::enable_shared_from_this<LWEtoRLWEPackingContext> {
    LWEtoRLWEPackingContext(KeyDistribution source_key_distribution, uint64_t modulus, uint64_t N, uint64_t basis,
                            uint64_t digits, double std);
    LWEtoRLWEPackingContext(KeyDistribution source_key_distribution, std::shared_ptr<MathWorker> ntt, uint64_t basis,
                            uint64_t digits, double std);
    LWEtoRLWEPackingContext(const LWEtoRLWEPackingContext& other);
    void SetSeed(uint64_t seed) { m_auto_context->SetSeed(seed); }
    [[nodiscard]] uint64_t GetDimension() const { return m_auto_context->GetDimension(); }
    [[nodiscard]] uint64_t GetModulus() const { return m_auto_context->GetModulus(); }
    [[nodiscard]] std::shared_ptr<MathWorker> GetNTT() const { return m_auto_context->GetNTT(); }
    [[nodiscard]] uint64_t GetGadgetBasisLog2() const { return m_auto_context->GetGadgetBasisLog2(); }
    [[nodiscard]] KeyDistribution GetSourceKeyDistribution() const { return m_auto_context->GetSourceKeyDistribution(); }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<LWEtoRLWEPacker> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return PACK_LWE_TO_RLWE; }

private:
    std::shared_ptr<AutomorphismContext> m_auto_context;
};

struct LWEtoRLWEPacker {
    friend struct LWEtoRLWEPackingContext;
    // len LWEs of dimension N ([len][N+1]) -> RLWE [a|b] in coefficients (lwe_to_rlwe_packing.cpp:129-236)
    void Pack(uint64_t* output, const uint64_t* input, uint64_t len);
    void Pack(std::vector<uint64_t>& output, const std::vector<uint64_t>& input);
    // coefficient i = message i for i < len (lwe_to_rlwe_packing.cpp:260-396)
    void PackConsecutively(uint64_t* output, const uint64_t* input, uint64_t len);
    [[nodiscard]] std::shared_ptr<const LWEtoRLWEPackingContext> GetContext() const { return m_context; }
    [[nodiscard]] const std::vector<uint64_t>& GetAutomorphismKeys() const { return m_keys; }

private:
    LWEtoRLWEPacker(std::shared_ptr<const LWEtoRLWEPackingContext> context, toothpaste_amd::DeviceContext dev,
                    std::vector<uint64_t> keys);
    void Run(uint64_t* output, const uint64_t* input, uint64_t len, bool consecutive);
    std::shared_ptr<const LWEtoRLWEPackingContext> m_context;
    toothpaste_amd::DeviceContext m_dev;
    std::vector<uint64_t> m_keys;
};

// ------------------------------------------------------------------ LWE -> RGSW (scheme switching)

struct LWEtoRGSWConverter;
struct SchemeSwitchingContext : OperatorContext<LWEtoRGSWConverter>,
                                public std::enable_shared_from_this<SchemeSwitchingContext> {
    SchemeSwitchingContext(std::shared_ptr<OperatorContext<BlindRotator>> blind_rotation_context,
                           std::shared_ptr<TraceEvaluationContext> trace_params,
                           std::shared_ptr<RLWEConversionContext> squaring_params, uint64_t output_digits,
                           uint64_t basis);
    long double ComputeOutputVariance(long double input_variance = 0) const override;
    [[nodiscard]] const std::shared_ptr<OperatorContext<BlindRotator>> GetBlindRotationContext() const {
        return m_rotation_context;
    }
    [[nodiscard]] const std::shared_ptr<TraceEvaluationContext> GetTraceContext() const { return m_trace_params; }
    [[nodiscard]] const std::shared_ptr<RLWEConversionContext> GetSquaringContext() const { return m_squaring_params; }
    [[nodiscard]] uint64_t GetOutputDigits() const { return m_output_digits; }
    [[nodiscard]] uint64_t GetOutputBasis() const { return m_output_basis; }
    [[nodiscard]] uint64_t GetOutputBasisLog2() const { return m_output_basis_log2; }
    void SetBlindRotationContext(std::shared_ptr<OperatorContext<BlindRotator>> c) { m_rotation_context = std::move(c); }
    void SetTraceContext(std::shared_ptr<TraceEvaluationContext> c) { m_trace_params = std::move(c); }
    void SetSquaringParameters(std::shared_ptr<RLWEConversionContext> c) { m_squaring_params = std::move(c); }
    void SetOutputDigits(uint64_t d) { m_output_digits = d; }
    void SetOutputBasis(uint64_t basis);
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container container) const override;
    std::unique_ptr<LWEtoRGSWConverter> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return CONV_LWE_RGSW; }

private:
    std::shared_ptr<OperatorContext<BlindRotator>> m_rotation_context;
    std::shared_ptr<TraceEvaluationContext> m_trace_params;
    std::shared_ptr<RLWEConversionContext> m_squaring_params;
    uint64_t m_output_digits, m_output_basis, m_output_basis_log2;
};

struct LWEtoRGSWConverter : SchemeConverter<SchemeSwitchingContext> {
    friend struct SchemeSwitchingContext;
    LWEtoRGSWConverter(std::shared_ptr<const SchemeSwitchingContext> context, std::unique_ptr<BlindRotator> rotator,
                       std::unique_ptr<TraceEvaluator> trace_eval, std::unique_ptr<RLWEtoRLWEConverter> squaring_eval);
    [[nodiscard]] const std::shared_ptr<const SchemeSwitchingContext> GetContext() const override { return m_params; }
    // input LWE (n + 1 in Z_2N) -> output [RLWE(-s m L^i) rows | RLWE(m L^i) rows], digits x 2N each, NTT
    // domain (lwe_to_rgsw_conversion.cpp:163-231); the first half is subtracted into, as the reference
    // does, so pass it zeroed.
    void Convert(uint64_t* output, const uint64_t* const input) override;
    void Convert(std::vector<uint64_t>& output, const std::vector<uint64_t>& input) override;

private:
    std::shared_ptr<const SchemeSwitchingContext> m_params;
    std::unique_ptr<BlindRotator> m_rotator;
    std::unique_ptr<TraceEvaluator> m_trace_eval;
    std::unique_ptr<RLWEtoRLWEConverter> m_square_eval;
    std::vector<uint64_t> m_acc, m_extraction_buffer;
};

// ------------------------------------------------------------------ SAP digit decomposition

// A radix decomposer maps an encryption of a datum to encryptions of its digits
// (lib/include/interfaces/decomposition_operator.h:13-55).
template <typename ContextType>
struct RadixDecomposer {
    virtual void Decompose(uint64_t* output, const uint64_t* const input, uint64_t radix) = 0;
    virtual void Decompose(std::vector<uint64_t>& output, const std::vector<uint64_t>& input, uint64_t radix) = 0;
    virtual const std::shared_ptr<const ContextType> GetContext() const = 0;
    virtual ~RadixDecomposer() = default;
};

// erfc(Q / (2 P sqrt(2 var))) (lib/src/utils/math_utils.cpp:87-96)
long double EstimateFailureProbability(long double variance, uint64_t modulus, uint64_t plaintext_modulus);

struct SAPDecomposer;
// lib/include/operators/sap_decomposition.h:13-134, lib/src/operators/sap_decomposer.cpp:14-188
struct SAPDecompositionContext : OperatorContext<SAPDecomposer>,
                                 public std::enable_shared_from_this<SAPDecompositionContext> {
    static const uint64_t IMPLICIT_SK_L0 = 32;
    SAPDecompositionContext(std::shared_ptr<OperatorContext<BlindRotator>> blind_rotation_context,
                            std::shared_ptr<TraceEvaluationContext> trace_context,
                            std::shared_ptr<LWEConversionContext> lwe_conversion_context, uint64_t default_radix);
    // Variance of a digit after the restart period.  The reference's loop counter is 32-bit
    // (sap_decomposer.cpp:55) and never reaches a restart period of UINT64_MAX; this one stops at
    // the fixed point, which is the reference's value whenever its loop ends.
    long double ComputeOutputVariance(long double input_variance = 0) const override;
    [[nodiscard]] std::shared_ptr<OperatorContext<BlindRotator>> GetBlindRotationContext() const {
        return m_rotation_context;
    }
    [[nodiscard]] std::shared_ptr<TraceEvaluationContext> GetTraceContext() const { return m_packing_context; }
    [[nodiscard]] std::shared_ptr<LWEConversionContext> GetLWEConversionContext() const { return m_conversion_context; }
    [[nodiscard]] const uint64_t GetDefaultRadix() const { return m_default_radix; }
    [[nodiscard]] const uint64_t GetRestartIteration() const { return m_restart_iter; }
    // Digits after which the accumulator is refreshed (sap_decomposer.cpp:81-152); UINT64_MAX when
    // the trace's noise estimate is 0, as it is for noise-free automorphism inputs.
    [[nodiscard]] const uint64_t ComputeRestartIterationForRadix(uint64_t radix, uint64_t sk_hamming) const;
    void SetBlindRotationContext(std::shared_ptr<OperatorContext<BlindRotator>> c) { m_rotation_context = std::move(c); }
    void SetTraceContext(std::shared_ptr<TraceEvaluationContext> c) { m_packing_context = std::move(c); }
    void SetLWEConversionContext(std::shared_ptr<LWEConversionContext> c) { m_conversion_context = std::move(c); }
    void SetDefaultRadix(uint64_t new_radix);
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container container) const override;
    // keys = {LWE secret, RLWE secret}: rotation keys from both, automorphism keys from keys[1],
    // the reset's key switch from keys[1] to keys[0] (sap_decomposer.cpp:167-186)
    std::unique_ptr<SAPDecomposer> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return DECOMP_SAP; }

private:
    std::shared_ptr<OperatorContext<BlindRotator>> m_rotation_context;
    std::shared_ptr<TraceEvaluationContext> m_packing_context;
    std::shared_ptr<LWEConversionContext> m_conversion_context;
    uint64_t m_default_radix, m_restart_iter;
};

// lib/include/operators/sap_decomposition.h:137-229.  The three operators' keys are gathered in
// one device context, where the batched pipeline runs.
struct SAPDecomposer : RadixDecomposer<SAPDecompositionContext> {
    friend struct SAPDecompositionContext;
    // digits of input (n + 1 words, power-of-two modulus read off the values) in radix, each an LWE
    // of dimension N; writes ceil(log2 q / log2 radix) x (N + 1) words (sap_decomposer.cpp:375-494)
    void Decompose(uint64_t* output, const uint64_t* const input, uint64_t radix) override;
    void Decompose(std::vector<uint64_t>& output, const std::vector<uint64_t>& input, uint64_t radix) override;
    // Batched: inputs [count][n+1] -> output [count][max_digits][N+1], digit_counts[count] (may be
    // null).  No counterpart in the reference; one launch per step for the whole batch.
    void DecomposeBatch(uint64_t* output, const uint64_t* inputs, uint64_t count, uint64_t radix, uint64_t max_digits,
                        uint32_t* digit_counts = nullptr);
    [[nodiscard]] const std::shared_ptr<const SAPDecompositionContext> GetContext() const override { return m_context; }
    ~SAPDecomposer() override = default;
    // RLWE(X^c) -> RLWE(X^floor(c / radix)), NTT domain; output must differ from input (:190-312)
    void HomTrunc(uint64_t* output, uint64_t* input, uint64_t radix);
    // refresh of an accumulator RLWE(X^c) (NTT domain, in place) (:314-373)
    void ResetAccumulator(uint64_t* acc, uint64_t radix);
    [[nodiscard]] tp_context* DeviceContext() const { return m_dev.get(); }

private:
    SAPDecomposer(std::shared_ptr<const SAPDecompositionContext> ctx, std::unique_ptr<BlindRotator> rotator,
                  std::unique_ptr<TraceEvaluator> packer, std::unique_ptr<LWEtoLWEConverter> conv, uint64_t max_radix,
                  uint64_t lwe_sk_hamming_weight, uint64_t reset_period);
    uint64_t RestartFor(uint64_t radix) const;
    std::shared_ptr<const SAPDecompositionContext> m_context;
    std::unique_ptr<BlindRotator> m_rotator;
    std::unique_ptr<TraceEvaluator> m_packer;
    std::unique_ptr<LWEtoLWEConverter> m_lwe_converter;
    uint64_t m_beta, m_max_radix, m_restart_iteration;
    toothpaste_amd::DeviceContext m_dev;
};

// ------------------------------------------------------------------ function evaluation

struct NegacyclicFunctionEvaluator;
struct NegacyclicFunctionEvaluationContext : OperatorContext<NegacyclicFunctionEvaluator>,
                                             public std::enable_shared_from_this<NegacyclicFunctionEvaluationContext> {
    explicit NegacyclicFunctionEvaluationContext(std::shared_ptr<OperatorContext<BlindRotator>> blind_rotation_context);
    [[nodiscard]] std::shared_ptr<OperatorContext<BlindRotator>> GetBlindRotationContext() const { return m_rotation_context; }
    [[nodiscard]] long double ComputeOutputVariance(long double input_variance = 0.0) const override;
    [[nodiscard]] std::unique_ptr<NegacyclicFunctionEvaluator> ConstructOperator(const std::vector<GenericKey>& keys) const override;
    [[nodiscard]] Container GetInputContainer() const override;
    [[nodiscard]] Container GetOutputContainer(Container input) const override;
    [[nodiscard]] OperatorID GetOperatorID() const override { return NEGACYCLIC_FUNC_BOOT; }

private:
    std::shared_ptr<OperatorContext<BlindRotator>> m_rotation_context;
};

struct NegacyclicFunctionEvaluator : FunctionEvaluator<NegacyclicFunctionEvaluationContext> {
    friend struct NegacyclicFunctionEvaluationContext;
    // Test vector from f as the reference builds it (negacyclic_function_evaluator.cpp:70-89):
    // word 0 = f(0) (the a half), word N + i = -f(N - i) for i >= 1; then as the lut overload.
    void EvalFunc(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED input_lwe,
                  std::function<uint64_t(uint64_t)>& function) override;
    void EvalFunc(std::vector<uint64_t>& result, const std::vector<uint64_t>& input,
                  std::function<uint64_t(uint64_t)>& function) override;
    // Blind rotation of lut_rlwe (2N coefficients) by the input, then SampleExtract(0): result N+1
    // (negacyclic_function_evaluator.cpp:92-101).
    void EvalFunc(uint64_t* RESTRICTED result, const uint64_t* RESTRICTED input_lwe,
                  uint64_t* RESTRICTED lut_rlwe) override;
    void EvalFunc(std::vector<uint64_t>& result, const std::vector<uint64_t>& input,
                  std::vector<uint64_t>& lut_rlwe) override;
    // Batched: results [count][N+1]; lut [2N] shared (lut_stride 0) or [count][2N] (stride 2N).
    void EvalFuncBatch(uint64_t* results, const uint64_t* input_lwes, uint64_t count, const uint64_t* lut,
                       uint64_t lut_stride = 0);
    [[nodiscard]] const std::shared_ptr<const NegacyclicFunctionEvaluationContext> GetContext() const override {
        return m_context;
    }
    [[nodiscard]] BlindRotator& GetRotator() const { return *m_rotator; }
    ~NegacyclicFunctionEvaluator() override = default;

private:
    NegacyclicFunctionEvaluator(std::shared_ptr<const NegacyclicFunctionEvaluationContext> ctx,
                                std::unique_ptr<BlindRotator> rotator);
    std::shared_ptr<const NegacyclicFunctionEvaluationContext> m_context;
    std::unique_ptr<BlindRotator> m_rotator;
};
