/*
 * toothpaste_amd: C-ABI of the MI355X (gfx950) CGGI/BMMP bootstrapping path.
 *
 * This is the drop-in boundary for the hot path of LSchild/FHEToothpaste.  Every
 * entry point names the reference interface it replaces (file:line in that
 * repository).  Conventions:
 *   - every function returns 0 (TP_OK) or a negative TP_E* status; the message of
 *     the last failure on the calling thread is returned by tp_last_error();
 *   - ring elements are uint64_t values in [0, Q), Q = 2^64 - 2^32 + 1 (Goldilocks);
 *   - NTT-domain data uses the reference order: out[k] = a(psi^(2 brv(k) + 1)),
 *     psi = 7^((Q-1)/2N) (lib/src/backend/solinas.cpp:228-261);
 *   - key bytes use the reference layouts (lib/src/base_crypto.cpp:60-97,
 *     lib/src/operators/cggi_blind_rotation.cpp:264-311,
 *     lib/src/operators/bmmp_blind_rotation.cpp:229-262,
 *     lib/src/operators/lwe_conversion.cpp:174-203) and are converted on import;
 *   - "_device" entry points take device pointers and run asynchronously on the
 *     context's stream (finish with tp_context_sync); the others take host
 *     pointers and return when the result is in host memory;
 *   - calls on one context are serialised by the caller; contexts are independent.
 */
#ifndef TOOTHPASTE_AMD_H
#define TOOTHPASTE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TP_OK 0
#define TP_E_INVALID (-1)   /* bad argument (reference: assert / std::invalid_argument) */
#define TP_E_DEVICE (-2)    /* HIP runtime error or no gfx950 device */
#define TP_E_STATE (-3)     /* keys not imported yet (reference: assert(m_keys_generated)) */
#define TP_E_NOMEM (-4)     /* device allocation failed */
#define TP_E_TIMEOUT (-5)   /* tp_server_wait: request not complete in time */
#define TP_E_ENTROPY (-6)   /* key generation: the operating system's random source failed */

#define TP_Q 0xFFFFFFFF00000001ULL

/* Blind-rotation method (lib/include/static/enum_ids.h KeyDistribution + operator kind). */
#define TP_METHOD_CGGI_BINARY 0   /* CGGIBlindRotator, BINARY  (cggi_blind_rotation.cpp:313-386) */
#define TP_METHOD_CGGI_TERNARY 1  /* CGGIBlindRotator, TERNARY (cggi_blind_rotation.cpp:388-472) */
#define TP_METHOD_BMMP 2          /* BMMPBlindRotator, unroll 2 (bmmp_blind_rotation.cpp:354-419) */

/* Parameters of one blind-rotation context; mirrors the arguments of
 * CGGIBlindRotationContext(distribution, Q, N, n, basis, digits, std)
 * (lib/src/operators/cggi_blind_rotation.cpp:14-25) and of
 * LWEConversionContext (lib/include/operators/endo_glwe_conversion.h:15-60). */
typedef struct tp_params {
    uint32_t ring_dim;      /* N: 1024, 2048 or 4096 */
    uint32_t lwe_dim;       /* n */
    uint32_t br_base_log2;  /* b: gadget base 2^b */
    uint32_t br_digits;     /* l */
    uint32_t method;        /* TP_METHOD_* */
    uint32_t ks_out_dim;    /* LWE key switch N -> ks_out_dim; 0 = no key switch */
    uint32_t ks_base_log2;
    uint32_t ks_digits;
    uint32_t rlwe_ks_base_log2;  /* automorphism / trace / packing keys (RLWEtoRLWEConverter gadget) */
    uint32_t rlwe_ks_digits;     /* 0 = no automorphism keys */
} tp_params;

typedef struct tp_context tp_context;

/* Last error message of the calling thread ("" if none). */
const char* tp_last_error(void);

/* Number of usable gfx950 devices. */
int tp_device_count(int* count);

/* Create / destroy a context on a device.  Replaces OperatorContext<Op>::ConstructOperator
 * (lib/include/interfaces/operator_context.h:22-63) minus the key generation, which the
 * reference does inside ConstructOperator and which is tp_import_* here. */
int tp_context_create(tp_context** out, const tp_params* params, int device);
int tp_context_destroy(tp_context* ctx);
int tp_context_sync(tp_context* ctx);
/* Parameters a context was created with. */
int tp_context_params(tp_context* ctx, tp_params* out);

/* Import the blind-rotation key in the reference byte layout (host memory).  Replaces the
 * key material built by CGGIBlindRotator::KeyGen (cggi_blind_rotation.cpp:215-311) and
 * BMMPBlindRotator::KeyGen (bmmp_blind_rotation.cpp:222-262).  words must equal
 * n*4Nl (binary), n*8Nl (ternary) or (n/2)*12Nl (BMMP). */
int tp_import_brk(tp_context* ctx, const uint64_t* brk, size_t words);

/* Import the LWE->LWE key-switching key [N][d][n_out+1] (lwe_conversion.cpp:174-203). */
int tp_import_lwe_ksk(tp_context* ctx, const uint64_t* ksk, size_t words);

/* Import the automorphism keys for t = 2^j + 1, j = 1..log2 N, in the layout the reference
 * builds them (AutomorphismContext::ConstructOperator, automorphism_evaluation.cpp:43-63, one
 * RLWEtoRLWEConverter key each, rlwe_conversion.cpp:187-218): [log2 N][d][2N], NTT domain,
 * key j-1 switching from s(X^t) to s(X).  These are the keys TraceEvaluationContext and
 * LWEtoRLWEPackingContext construct (trace_evaluation.cpp:50-62, lwe_to_rlwe_packing.cpp:46-58). */
int tp_import_auto_keys(tp_context* ctx, const uint64_t* keys, size_t words);

/* Batched AutomorphismEvaluator::Eval (automorphism_evaluation.cpp:150-159): in [count][2N]
 * coefficients [a|b], out [count][2N] NTT domain, t = 2^j + 1 (1 <= j <= log2 N).  Overwrite
 * semantics (the reference subtracts into a zeroed output). */
int tp_automorphism_device(tp_context* ctx, size_t count, uint32_t t, const uint64_t* in, uint64_t* out);

/* Batched TraceEvaluator::Eval (trace_evaluation.cpp:135-162): in [count][2N] coefficients,
 * out [count][2N] NTT domain. */
int tp_trace_device(tp_context* ctx, size_t count, const uint64_t* in, uint64_t* out);

/* LWEtoRLWEPacker::Pack (lwe_to_rlwe_packing.cpp:129-236): len LWEs of dimension N
 * ([len][N+1], mod Q; len a power of two, 2 <= len <= N) -> one RLWE in coefficients, out [2N].
 * perm (may be NULL) reorders the inputs: input i is lwes[perm[i]].  Uses len*2N words of
 * context scratch; the tree levels are launched in order on the context stream. */
int tp_pack_device(tp_context* ctx, size_t len, const uint64_t* lwes, const uint32_t* perm, uint64_t* out);

/* RLWE->RLWE key switching with an imported key, optionally after the automorphism X -> X^t.
 * tp_import_rlwe_ksk: RLWEtoRLWEConverter key [d][2N], NTT domain, d = rlwe_ks_digits, built by
 * RLWEtoRLWEConverter::KeyGen (rlwe_conversion.cpp:187-218) or tp_keygen_rlwe_ksk.
 * tp_rlwe_keyswitch_device, count ciphertexts [count][2N] -> [count][2N] NTT domain:
 *   TP_RKS_CONVERT      in = [a coef | b NTT] (t must be 1): RLWEtoRLWEConverter::Convert
 *                       (rlwe_conversion.cpp:224-249)
 *   TP_RKS_AUTOMORPHISM in = [a|b] coefficients: AutomorphismEvaluator::Eval
 *                       (automorphism_evaluation.cpp:150-159)
 *   TP_RKS_SWITCH_ONLY  in = [a|.] coefficients, b ignored: AutomorphismEvaluator::EvalSwitchOnly
 *                       (automorphism_evaluation.cpp:172-177)
 *   | TP_RKS_ACCUMULATE out += result, the reference's subtract-into-output semantics;
 *                       without it out = result.
 * key: 0 = the imported RLWE key, j in [1, log2 N] = automorphism key j (t = 2^j + 1) of
 * tp_import_auto_keys.  t odd, 1 <= t < 2N. */
#define TP_RKS_CONVERT 0
#define TP_RKS_AUTOMORPHISM 1
#define TP_RKS_SWITCH_ONLY 2
#define TP_RKS_ACCUMULATE 4
int tp_import_rlwe_ksk(tp_context* ctx, const uint64_t* key, size_t words);
int tp_rlwe_keyswitch_device(tp_context* ctx, size_t count, uint32_t t, int mode, int key, const uint64_t* in,
                             uint64_t* out);

/* LWEtoRLWEPacker::PackConsecutively (lwe_to_rlwe_packing.cpp:260-396): len LWEs (dimension N,
 * 1 <= len <= N, [len][N+1]) -> one RLWE in coefficients whose coefficient i holds message i for
 * i < len.  The reference walks the full N-leaf tree with a stack and prunes subtrees without
 * inputs; here the tree runs level by level with zero leaves for them (merging zeros is exact).
 * Uses N*2N words of context scratch. */
int tp_pack_consecutive_device(tp_context* ctx, size_t len, const uint64_t* lwes, uint64_t* out);

/* Elementwise mod-Q arithmetic over `words` words (device pointers, inputs canonical, outputs
 * canonical, dst may alias an input): MathWorker::{AddModElw, SubModElw, MulModElw, FMAModElw,
 * MulSModElw, ReduceModElw} and their Eq forms (lib/include/backend/backend.h:35-55).
 *   TP_ELT_ADD dst = a + b    TP_ELT_SUB dst = a - b     TP_ELT_MUL dst = a * b
 *   TP_ELT_FMA dst = a*b + c  TP_ELT_MULS dst = a * scalar   TP_ELT_REDUCE dst = a mod Q (any a) */
#define TP_ELT_ADD 0
#define TP_ELT_SUB 1
#define TP_ELT_MUL 2
#define TP_ELT_FMA 3
#define TP_ELT_MULS 4
#define TP_ELT_REDUCE 5
int tp_eltwise_device(tp_context* ctx, int op, uint64_t* dst, const uint64_t* a, const uint64_t* b,
                      const uint64_t* c, uint64_t scalar, size_t words);

/* ModulusSwitch (lib/src/static/modulus_switching.cpp:36-93): dst[i] = R(tM src[i] / sM), then
 * minus tM if >= tM, for `words` values (device pointers, dst may equal src).  The quotient is
 * rounded to a 64-bit significand exactly as the reference's x87 long double does, so results
 * are bit-identical to it, near ties included.  sM >= 2^63 (e.g. the Goldilocks Q), tM = 2^t with
 * 1 <= t <= 62, inputs < sM.  type: TP_MS_ROUND (roundl), TP_MS_FLOOR, TP_MS_CEIL, TP_MS_RANDOM
 * (floor or ceil by a fair coin from the stream (seed, index); the reference's coin is unseeded). */
#define TP_MS_ROUND 0
#define TP_MS_FLOOR 1
#define TP_MS_CEIL 2
#define TP_MS_RANDOM 3
int tp_modswitch_device(tp_context* ctx, uint64_t* dst, const uint64_t* src, size_t words, uint64_t source_modulus,
                        uint64_t target_modulus, int type, uint64_t seed);

/* SAP digit decomposition, batched (lib/include/operators/sap_decomposition.h:137-228,
 * lib/src/operators/sap_decomposer.cpp:190-494).  One context holds the blind-rotation key
 * (SAPDecomposer::m_rotator), the automorphism keys (m_packer, a TraceEvaluator) and, for the
 * reset, the LWE key-switching key N -> lwe_dim (m_lwe_converter).  radix: power of two in [2, N].
 *
 * tp_sap_decompose_device replaces a loop of SAPDecomposer::Decompose(output, input, radix)
 * (:375-494) over `count` inputs.  lwes [count][lwe_dim+1] of a power-of-two modulus read off the
 * values as the reference does (:427-433); beta = Hamming weight of the LWE key (ConstructOperator,
 * :172-176); restart = digit period of the accumulator reset (the reference's
 * restart iteration; UINT64_MAX = never, its value when the trace noise estimate is 0).
 * out [count][max_digits][N+1]: digit k of input i at out + (i max_digits + k)(N+1), written for
 * k < digits[i]; digits (device, may be NULL) [count] receives each input's digit count.  Fails
 * with TP_E_INVALID, before any launch, when an input needs more than max_digits digits.
 *
 * tp_sap_homtrunc_device: SAPDecomposer::HomTrunc (:190-312), in/out [count][2N] NTT domain
 * (out may equal in).  tp_sap_reset_device: SAPDecomposer::ResetAccumulator (:314-373), acc
 * [count][2N] NTT domain in place; any radix (the reference resets with 0 inside Decompose). */
int tp_sap_decompose_device(tp_context* ctx, size_t count, uint32_t radix, uint32_t beta, uint64_t restart,
                            const uint64_t* lwes, uint64_t* out, uint32_t max_digits, uint32_t* digits);
int tp_sap_homtrunc_device(tp_context* ctx, size_t count, uint32_t radix, const uint64_t* in, uint64_t* out);
int tp_sap_reset_device(tp_context* ctx, size_t count, uint64_t radix, uint64_t* acc);

/* LWE -> RGSW scheme switching, batched: LWEtoRGSWConverter::Convert(output, input)
 * (lib/include/operators/lwe_to_rgsw_conversion.h:85-104, lwe_to_rgsw_conversion.cpp:163-231) over
 * `count` inputs.  The context holds the blind-rotation key, the trace's automorphism keys
 * (tp_import_auto_keys) and the s^2 -> s key (tp_import_rlwe_ksk), the last two with the context's
 * rlwe_ks gadget.  lwes [count][lwe_dim+1] in Z_2N; out [count][4 N digits]: per input the left Lev
 * (RLWE(-s m B^i), digits x 2N) then the right Lev (RLWE(m B^i)), NTT domain, written in full.
 * digits must divide N and not exceed ceil(64 / basis_log2). */
int tp_lwe_to_rgsw_device(tp_context* ctx, size_t count, uint32_t digits, uint32_t basis_log2, const uint64_t* lwes,
                          uint64_t* out);

/* Negacyclic NTT over Z_Q[X]/(X^N+1) of `count` polynomials (device pointers).
 * Replace SolinasWorker::ForwardNTT / BackwardNTT (lib/src/backend/solinas.cpp:228-313);
 * dst may equal src. */
int tp_ntt_forward_device(tp_context* ctx, uint64_t* dst, const uint64_t* src, size_t count);
int tp_ntt_inverse_device(tp_context* ctx, uint64_t* dst, const uint64_t* src, size_t count);

/* One blind rotation, host buffers: BlindRotator::BlindRotate(result, lwe_vec, accumulator,
 * output_as_coefficients) (lib/include/interfaces/blindrotation_operator.h:29-40).
 * lwe: n+1 values in Z_2N; acc: 2N coefficients [a|b]; result: 2N (NTT domain unless
 * as_coef); result may alias acc. */
int tp_blind_rotate(tp_context* ctx, uint64_t* result, const uint64_t* lwe, const uint64_t* acc, int as_coef);

/* Batched function evaluation: for each of `count` LWE inputs, blind rotation of the test
 * vector followed by SampleExtract at index 0.  Replaces a loop over
 * NegacyclicFunctionEvaluator::EvalFunc(result, input, lut_rlwe)
 * (negacyclic_function_evaluator.cpp:92-101).  Device pointers:
 *   lwes     [count][n+1] in Z_2N
 *   tv       [2N] coefficients, shared (tv_stride 0) or [count][2N] (tv_stride 2N)
 *   out_lwe  [count][N+1] extracted LWE (may be NULL)
 *   out_rlwe [count][2N] rotated accumulator (may be NULL), NTT domain if rlwe_ntt. */
int tp_eval_func_batch_device(tp_context* ctx, size_t count, const uint64_t* lwes, const uint64_t* tv,
                              size_t tv_stride, uint64_t* out_lwe, uint64_t* out_rlwe, int rlwe_ntt);

/* Batched LWE->LWE key switch N -> n_out (lwe_conversion.cpp:205-239), overwrite semantics:
 * out = (0, in_b) - sum d_ij KSK[i][j].  in [count][in_stride] (first N+1 used), out
 * [count][n_out+1], device pointers. */
int tp_lwe_keyswitch_batch_device(tp_context* ctx, size_t count, const uint64_t* in, size_t in_stride,
                                  uint64_t* out);

/* Full bootstrap = tp_eval_func_batch_device + tp_lwe_keyswitch_batch_device, using a device
 * scratch buffer owned by the context.  out [count][n_out+1]. */
int tp_bootstrap_batch_device(tp_context* ctx, size_t count, const uint64_t* lwes, const uint64_t* tv,
                              size_t tv_stride, uint64_t* out);

/* tp_bootstrap_batch_device without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, allocation-free serving
 * loops).  The extracted LWEs between the blind rotation and the key switch live in a 36 MiB array
 * that every GPU the library runs on holds from load until the process ends; count * (N + 1) must
 * not exceed 4,718,592 words (4,603 ciphertexts at N = 1024, 2,302 at 2048, 1,151 at 4096), else
 * TP_E_INVALID.  All calls on one GPU share the array: calls that overlap on one GPU race.
 * Output identical to tp_bootstrap_batch_device. */
int tp_bootstrap_batch_device_noalloc(tp_context* ctx, size_t count, const uint64_t* lwes, const uint64_t* tv,
                                      size_t tv_stride, uint64_t* out);

/* A full bootstrap (tp_bootstrap_batch_device's launches: blind rotation + sample extraction, then
 * the LWE key switch) captured once into a HIP graph and replayed with one call.  No counterpart in
 * the reference (its bootstraps are host loops): for callers that repeat one batch shape on fixed
 * device buffers (a serving loop with staging buffers, a benchmark), where a replay issues the
 * whole launch sequence as one submission.  The graph holds the pointers it was created with and
 * its own scratch for the extracted LWEs (count * (N + 1) words, allocated at creation): every
 * replay bootstraps the current contents of lwes/tv into out, asynchronously on the context's
 * stream; replays and other calls on the same context are ordered by that stream.  The context's
 * settings at creation (kernel variants, thresholds) are baked in.  Destroy the graph before its
 * context. */
typedef struct tp_graph tp_graph;
int tp_graph_bootstrap_create(tp_context* ctx, size_t count, const uint64_t* lwes, const uint64_t* tv,
                              size_t tv_stride, uint64_t* out, tp_graph** graph);
int tp_graph_launch(tp_graph* graph);
int tp_graph_destroy(tp_graph* graph);

/* Tuning (no counterpart in the reference).
 * TP_OPT_BR_WAVES selects the one-wave-per-ciphertext blind-rotation kernel: 2 (default; 256
 * registers per lane, two ciphertexts' waves share a SIMD) or 1 (512 registers, no spills in the
 * ternary and BMMP key combination; built for 3 and 4 digits).
 * TP_OPT_BR_SPLIT_MAX: launches of at most this many ciphertexts use the digit-parallel kernel
 * instead (2 l waves per ciphertext, one work-group of them per CU; N = 1024 with 2..4 digits
 * only, 0 switches it off).  Default: twice the device's CU count for contexts it is built for,
 * else 0.  Both kernels give bit-identical results.
 * TP_OPT_BR_ENGINE (binary CGGI at N = 1024 with 2..4 digits): 0 (default) computes each step's
 * external product with the Goldilocks NTT; 1 (opt-in) as an exact integer convolution in
 * double-precision complex arithmetic (the key split into four 16-bit limbs, kernels_br_fft.hip).
 * Engine 1 is used only when the rounding bound of the imported key certifies exact results
 * (below 1/2 for any digits of the gadget; otherwise tp_get_option reports 0); both engines give
 * bit-identical results.  TP_OPT_BR_FFT_WAVES: waves per SIMD of engine 1 (1: 512 registers,
 * default; 2: 256).
 * tp_autotune times every variant built for the context on `count` synthetic ciphertexts with
 * the imported key (the engine-1 variants only when the context opted in to engine 1), keeps
 * the fastest one-wave-per-ciphertext variant (engine and waves), times that variant on a
 * persistent grid (TP_OPT_BR_GRID = the ciphertexts resident at once) when `count` exceeds it and
 * keeps the faster grid, times the digit-parallel kernel in both forms where the two-wave form
 * is built (TP_OPT_BR_SPLIT_HALVES 1 and 0) and keeps the faster, moves the digit-parallel
 * threshold to the side of `count` that won, and reports ms[0] (NTT engine, two waves per SIMD),
 * ms[1] (one wave), ms[2] (digit-parallel, the faster form), ms[3]
 * (floating-point engine, one wave), ms[4] (two waves) and ms[5] (the winner on the persistent
 * grid), -1 where a variant is not built, not usable or not timed; ms must hold 6 floats.  The
 * context's later launches, including the server's chunks, follow these settings.
 * tp_autotune_cached(ctx, count, path, &hit, ms) keeps those choices in a JSON file at `path`,
 * keyed by device architecture, CU count, (N, n, b, l, method), count and the engine-1 opt-in: a
 * hit applies the stored settings (hit = 1, ms = the stored times) without launching anything, a
 * miss runs tp_autotune and adds its result to the file (created if absent; rewritten through
 * path.tmp and a rename, so processes tuning into one file at the same time may lose each other's
 * entries).  A malformed file or an entry this context cannot use gives TP_E_INVALID. */
#define TP_OPT_BR_WAVES 1
#define TP_OPT_BR_SPLIT_MAX 2
#define TP_OPT_BR_ENGINE 3
#define TP_OPT_BR_FFT_WAVES 4
/* TP_OPT_BR_GRID: work-groups per launch of the one-work-group-per-ciphertext blind-rotation
 * kernels; 0 (default) launches one per ciphertext, a positive value at most that many, each
 * striding over the launch's ciphertexts (results are the same either way). */
#define TP_OPT_BR_GRID 5
/* TP_OPT_BR_SPLIT_HALVES: for binary CGGI at N = 1024 with 2 or 3 digits, 1 (default) runs the
 * digit-parallel launches with every transform split over two waves (4 l waves per ciphertext,
 * kernels_br_half.hip), 0 with one wave per digit transform (2 l waves); tp_get_option reports 1
 * only where the split kernel is built.  Results are the same either way. */
#define TP_OPT_BR_SPLIT_HALVES 6
/* TP_OPT_KS_MFMA: LWE key switches of more than 16 ciphertexts run on the matrix cores (1, default)
 * where the key fits that form (base 2^b with b <= 7, at most 16 digits, n_in d (2^b - 1) 255 < 2^31:
 * the key is then also held as int8 operands, about 1.2x its size), else (0) on the vector units;
 * tp_get_option reports 1 only where the operands are built.  Results are the same either way. */
#define TP_OPT_KS_MFMA 7
int tp_set_option(tp_context* ctx, int option, int value);
int tp_get_option(tp_context* ctx, int option, int* value);
int tp_autotune(tp_context* ctx, size_t count, float* ms);
int tp_autotune_cached(tp_context* ctx, size_t count, const char* path, int* hit, float* ms);

/* Device memory helpers (so that callers need no HIP headers). */
int tp_malloc(tp_context* ctx, uint64_t** ptr, size_t words);
int tp_free(tp_context* ctx, uint64_t* ptr);
int tp_copy_to_device(tp_context* ctx, uint64_t* dst, const uint64_t* src, size_t words);
int tp_copy_to_host(tp_context* ctx, uint64_t* dst, const uint64_t* src, size_t words);
int tp_copy_on_device(tp_context* ctx, uint64_t* dst, const uint64_t* src, size_t words);

/* Elapsed milliseconds between two points of the context stream, for benchmarks:
 * tp_timer_mark(ctx, slot) records an event; tp_timer_elapsed(ctx, a, b, &ms). */
int tp_timer_mark(tp_context* ctx, int slot);
int tp_timer_elapsed(tp_context* ctx, int slot_a, int slot_b, float* ms);

/* ------------------------------------------------------------------ serving (continuous batching)
 * No counterpart in the reference, which evaluates one ciphertext per call
 * (negacyclic_function_evaluator.cpp:92-101 then lwe_conversion.cpp:205-239): a server keeps
 * the keys of several parameter sets resident and bootstraps the ciphertexts of concurrent
 * requests in shared launches.  Every ciphertext's output equals what
 * tp_bootstrap_batch_device produces for it alone, bit for bit.
 *
 * Scheduler core (host only, usable without a GPU): per parameter set a pool of max_slots[s]
 * slots of slot_bytes[s] bytes of HBM each; tp_sched_admit hands out slots to the ciphertexts
 * of queued requests in FIFO order while slots are free and the admitted footprint stays
 * within budget_bytes; tp_sched_release returns them and reports completed requests. */
typedef struct tp_sched tp_sched;
int tp_sched_create(tp_sched** out, uint32_t nsets, const size_t* slot_bytes, size_t budget_bytes,
                    const uint32_t* max_slots);
int tp_sched_destroy(tp_sched* s);
int tp_sched_submit(tp_sched* s, uint32_t set, size_t count, uint64_t* id);
int tp_sched_admit(tp_sched* s, uint32_t set, uint32_t* slots, uint64_t* req_ids, uint64_t* indices, size_t cap,
                   size_t* n_admitted);
int tp_sched_release(tp_sched* s, uint32_t set, const uint32_t* slots, const uint64_t* req_ids, size_t n,
                     int* completed);
int tp_sched_cancel(tp_sched* s, uint64_t id);
int tp_sched_forget(tp_sched* s, uint64_t id);
int tp_sched_state(tp_sched* s, uint32_t set, size_t* free_slots, size_t* queued_requests, size_t* used_bytes);

/* Server: one worker thread; each parameter set (a context with brk and LWE ksk imported)
 * runs on its context's stream.  Each iteration admits queued ciphertexts, advances every
 * active slot by chunk_steps blind-rotation steps, key-switches the slots that finished and
 * copies their outputs to the requests' host buffers.
 * tp_server_submit: count LWE ciphertexts (host, count x (n+1), entries in [0, 2N), kept
 * readable until completion), one test vector (host, 2N coefficients), output buffer
 * (host, count x (ks_out_dim+1)).  tp_server_wait returns TP_E_TIMEOUT when the request is
 * still running at the deadline; tp_server_release forgets a completed request.
 * tp_server_times: submit and completion instants of a request (seconds, steady clock). */
typedef struct tp_server tp_server;
typedef struct tp_server_config {
    uint32_t chunk_steps;        /* blind-rotation steps per launch (0: 64) */
    uint32_t max_slots_per_set;  /* slot pool size per parameter set (0: 4096) */
    size_t hbm_budget_bytes;     /* slot footprint limit over all sets (0: 90% of free HBM) */
} tp_server_config;
int tp_server_create(tp_server** out, tp_context* const* ctxs, uint32_t nsets, const tp_server_config* cfg);
int tp_server_destroy(tp_server* srv);
int tp_server_submit(tp_server* srv, uint32_t set, size_t count, const uint64_t* lwes, const uint64_t* tv,
                     uint64_t* out, uint64_t* id);
int tp_server_poll(tp_server* srv, uint64_t id, int* done);
int tp_server_wait(tp_server* srv, uint64_t id, double timeout_s);
int tp_server_cancel(tp_server* srv, uint64_t id);
int tp_server_release(tp_server* srv, uint64_t id);
int tp_server_times(tp_server* srv, uint64_t id, double* submitted, double* completed);
int tp_server_stats(tp_server* srv, uint64_t* iterations, uint64_t* launches, uint64_t* ciphertexts);

/* ------------------------------------------------------------------ wire format
 * No counterpart in the reference (it never serialises keys): a versioned record around a key or
 * ciphertext array in the reference layout, for resident serving keys and shared test vectors.
 * Header: "TPW1", version 1, kind, flags (bit 0: NTT domain), the tp_params it belongs to, the
 * payload word count and an FNV-1a 64 checksum; payload little-endian uint64.  Key kinds are
 * checked against their parameters (BRK n*4Nl / 2n*4Nl / (n/2)*12Nl words, LWE KSK N*d*(n_out+1),
 * automorphism keys log2(N)*d*2N, RLWE KSK d*2N); decoding rejects a bad magic, version, size or
 * checksum with TP_E_INVALID.  tp_wire_load with data NULL returns the word count only. */
#define TP_WIRE_RAW 0
#define TP_WIRE_BRK 1
#define TP_WIRE_LWE_KSK 2
#define TP_WIRE_AUTO_KEYS 3
#define TP_WIRE_RLWE_KSK 4
#define TP_WIRE_LWE_BATCH 5
#define TP_WIRE_RLWE_BATCH 6
int tp_wire_bytes(size_t words, size_t* bytes);
int tp_wire_encode(uint8_t* buf, size_t buf_bytes, int kind, uint32_t flags, const tp_params* params,
                   const uint64_t* data, size_t words);
int tp_wire_decode(const uint8_t* buf, size_t buf_bytes, int* kind, uint32_t* flags, tp_params* params,
                   uint64_t* data, size_t cap_words, size_t* words);
int tp_wire_save(const char* path, int kind, uint32_t flags, const tp_params* params, const uint64_t* data,
                 size_t words);
int tp_wire_load(const char* path, int* kind, uint32_t* flags, tp_params* params, uint64_t* data, size_t cap_words,
                 size_t* words);

/* ------------------------------------------------------------------ key generation
 * Host-side key generation (the reference samples with OpenFHE, which has no seed API:
 * lib/src/base_crypto.cpp:34-38,139-142).  Output bytes use the reference layouts, so they can be
 * passed to tp_import_* and to any reference-layout consumer.
 * seed == TP_SEED_OS_ENTROPY: the words come from ChaCha20 (RFC 8439) keyed with 256 bits from the
 *   operating system (getrandom) on every call; use this for real keys.
 * any other seed: a reproducible SplitMix64 stream (seed, tag), for tests and parity checks only:
 *   its outputs reveal its state, and with it the noise of every key drawn from it.
 * Gaussian noise: a 63-bit inversion table for sigma <= 21 (|e| <= 12 sigma), above that the
 *   Canonne-Kamath-Steinke rejection sampler (no tail cut).
 * dist: 0 binary, 1 ternary (P(0) = 1/2, P(+-1) = 1/4, -1 stored as Q-1).
 * gadget_kind: 0 = MakeRGSW weights L^(D-1-j) (base_crypto.cpp:68,85-95),
 *              1 = signed-decomposer weights 2^(64-b(j+1)) (gadget_decomp.h:176-193). */
#define TP_SEED_OS_ENTROPY 0xFFFFFFFFFFFFFFFFULL
int tp_keygen_secret(uint64_t* out, uint32_t len, int dist, uint64_t seed, uint64_t tag);
int tp_keygen_brk(uint64_t* brk, const uint64_t* lwe_key, const uint64_t* rlwe_key, uint32_t N, uint32_t n,
                  uint32_t base_log2, uint32_t digits, int method, int gadget_kind, double sigma, uint64_t seed);
int tp_keygen_lwe_ksk(uint64_t* ksk, const uint64_t* src, const uint64_t* dst, uint32_t n_in, uint32_t n_out,
                      uint32_t base_log2, uint32_t digits, double sigma, uint64_t seed);
/* Automorphism keys for t = 2^j + 1, j = 1..log2 N, under RLWE key s (coefficients):
 * keys [log2 N][d][2N]. */
int tp_keygen_auto_keys(uint64_t* keys, const uint64_t* s, uint32_t N, uint32_t base_log2, uint32_t digits,
                        double sigma, uint64_t seed);
/* RLWEtoRLWEConverter key from src to dst (coefficients, N each): row i = RLWE_dst(g_i src),
 * g_i = 2^(64 - b(i+1)) (kept >= 1), out [d][2N] NTT domain; stream (seed, tag). */
int tp_keygen_rlwe_ksk(uint64_t* key, const uint64_t* src, const uint64_t* dst, uint32_t N, uint32_t base_log2,
                       uint32_t digits, double sigma, uint64_t seed, uint64_t tag);
/* count discrete Gaussian samples of parameter sigma (the noise of the keygen functions) from the
 * stream (seed, tag); the sampler behind OpenFHE's DiscreteGaussianGenerator in the reference
 * (base_crypto.cpp:34-38). */
int tp_sample_gaussian(int64_t* out, size_t count, double sigma, uint64_t seed, uint64_t tag);
/* The ChaCha20 block function of the TP_SEED_OS_ENTROPY stream (RFC 8439, section 2.3), exposed
 * so that its output can be checked against the RFC's test vector. */
int tp_chacha20_block(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3], uint32_t out[16]);

#ifdef __cplusplus
}
#endif

#endif /* TOOTHPASTE_AMD_H */
