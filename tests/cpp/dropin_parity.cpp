// Exercises the C++ drop-in classes (include/toothpaste_amd.hpp) the way the reference's own tests
// use the reference classes, on seeded inputs, and writes every input and output as raw
// little-endian uint64 files into the directory given as argv[1]; tests/test_gpu_dropin.py
// checks them against the oracle.  Needs a gfx950 GPU.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/toothpaste_amd.hpp"

namespace {

constexpr uint64_t Q = 0xFFFFFFFF00000001ULL;
std::string g_dir;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    std::vector<uint64_t> modq(size_t n) {
        std::vector<uint64_t> v(n);
        for (auto& x : v) do { x = next(); } while (x >= Q);
        return v;
    }
    std::vector<uint64_t> below(size_t n, uint64_t bound) {
        std::vector<uint64_t> v(n);
        for (auto& x : v) x = next() % bound;
        return v;
    }
};

void dump(const std::string& name, const std::vector<uint64_t>& v) {
    std::ofstream f(g_dir + "/" + name + ".bin", std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * 8));
    if (!f) {
        std::cerr << "cannot write " << name << std::endl;
        std::exit(1);
    }
}

std::vector<uint64_t> secret(uint32_t len, int dist, uint64_t seed, uint64_t tag) {
    std::vector<uint64_t> s(len);
    if (tp_keygen_secret(s.data(), len, dist, seed, tag) != TP_OK) throw std::runtime_error("keygen");
    return s;
}

void step(const char* what) { std::cout << "[dropin] " << what << std::endl; }

template <class Rot>
void rotator_case(const std::string& tag, Rot& rot, uint64_t N, uint64_t n, Rng& g) {
    auto lwe = g.below(n + 1, 2 * N);
    auto acc = g.modq(2 * N);
    std::vector<uint64_t> out_ntt(2 * N), out_coef(2 * N);
    rot->BlindRotate(out_ntt, lwe, acc, false);
    rot->BlindRotate(out_coef, lwe, acc, true);
    auto alias = acc;
    rot->BlindRotate(alias.data(), lwe.data(), alias.data(), true);  // result may alias the accumulator
    auto lwes = g.below(3 * (n + 1), 2 * N);
    std::vector<uint64_t> batch(3 * 2 * N);
    rot->BlindRotateBatch(batch.data(), lwes.data(), 3, acc.data(), 0, true);
    dump(tag + "_lwe", lwe);
    dump(tag + "_acc", acc);
    dump(tag + "_out_ntt", out_ntt);
    dump(tag + "_out_coef", out_coef);
    dump(tag + "_alias", alias);
    dump(tag + "_batch_lwes", lwes);
    dump(tag + "_batch", batch);
    dump(tag + "_brk", rot->GetBlindRotationKey());
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: dropin_parity OUTDIR" << std::endl;
        return 2;
    }
    g_dir = argv[1];
    try {
        Rng g(2026);
        const uint64_t N = 1024;

        step("cggi binary");
        {
            const uint64_t n = 32;
            auto lk = secret(n, 0, 101, 1), rk = secret(N, 0, 101, 2);
            auto ctx = std::make_shared<CGGIBlindRotationContext>(BINARY, Q, N, n, 1 << 12, 3, 3.2);
            ctx->SetSeed(101);
            auto rot = ctx->ConstructOperator({GenericKey("LWE", lk), GenericKey("RLWE", rk)});
            rotator_case("cggi_bin", rot, N, n, g);

            step("function evaluator");
            auto br = std::reinterpret_pointer_cast<OperatorContext<BlindRotator>>(ctx);
            auto ev_ctx = std::make_shared<NegacyclicFunctionEvaluationContext>(br);
            auto ev = ev_ctx->ConstructOperator({GenericKey("LWE", lk), GenericKey("RLWE", rk)});
            auto lwe = g.below(n + 1, 2 * N);
            auto lut = g.modq(2 * N);
            std::vector<uint64_t> r_lut(N + 1), r_fn(N + 1);
            ev->EvalFunc(r_lut, lwe, lut);
            std::function<uint64_t(uint64_t)> f = [](uint64_t x) { return ((x * 7 + 3) % 16) * (Q / 16); };
            ev->EvalFunc(r_fn, lwe, f);
            auto lwes = g.below(4 * (n + 1), 2 * N);
            std::vector<uint64_t> r_batch(4 * (N + 1));
            ev->EvalFuncBatch(r_batch.data(), lwes.data(), 4, lut.data(), 0);
            dump("fe_lwe", lwe);
            dump("fe_lut", lut);
            dump("fe_out_lut", r_lut);
            dump("fe_out_fn", r_fn);
            dump("fe_batch_lwes", lwes);
            dump("fe_batch", r_batch);
        }

        step("cggi ternary");
        {
            const uint64_t n = 16;
            auto lk = secret(n, 1, 102, 1), rk = secret(N, 0, 102, 2);
            auto ctx = std::make_shared<CGGIBlindRotationContext>(TERNARY, Q, N, n, 1 << 16, 4, 3.2);
            ctx->SetSeed(102);
            auto rot = ctx->ConstructOperator({GenericKey("LWE", lk), GenericKey("RLWE", rk)});
            rotator_case("cggi_ter", rot, N, n, g);
        }

        step("bmmp");
        {
            const uint64_t n = 16;
            auto lk = secret(n, 0, 103, 1), rk = secret(N, 0, 103, 2);
            auto ctx = std::make_shared<BMMPBlindRotationContext>(BINARY, Q, N, n, 1 << 14, 3, 3.2, 2);
            ctx->SetSeed(103);
            auto rot = ctx->ConstructOperator({GenericKey("LWE", lk), GenericKey("RLWE", rk)});
            rotator_case("bmmp", rot, N, n, g);
        }

        step("lwe key switch");
        {
            const uint64_t n_out = 64;
            auto src = secret(N, 0, 104, 2), dst = secret(n_out, 0, 104, 1);
            auto ctx = std::make_shared<LWEConversionContext>(BINARY, Q, N, n_out, 1 << 4, 14, 3.2);
            ctx->SetSeed(104);
            auto conv = ctx->ConstructOperator({GenericKey("SRC", src), GenericKey("DST", dst)});
            auto in = g.modq(N + 1);
            auto prev = g.modq(n_out + 1);
            auto out = prev;
            conv->Convert(out, in);
            auto ins = g.modq(3 * (N + 1));
            std::vector<uint64_t> outs(3 * (n_out + 1));
            conv->ConvertBatch(outs.data(), ins.data(), 3);
            dump("lks_in", in);
            dump("lks_prev", prev);
            dump("lks_out", out);
            dump("lks_batch_in", ins);
            dump("lks_batch", outs);
            dump("lks_key", conv->GetSwitchingKey());
        }

        step("rlwe key switch");
        {
            auto src = secret(N, 0, 105, 3), dst = secret(N, 0, 105, 2);
            auto ctx = std::make_shared<RLWEConversionContext>(BINARY, Q, N, 1 << 16, 4, 3.2);
            ctx->SetSeed(105);
            auto conv = ctx->ConstructOperator({GenericKey("SRC", src), GenericKey("DST", dst)});
            auto in = g.modq(2 * N);
            auto prev = g.modq(2 * N);
            auto out = prev;
            conv->Convert(out, in);
            dump("rks_in", in);
            dump("rks_prev", prev);
            dump("rks_out", out);
            dump("rks_key", conv->GetSwitchingKey());
        }

        step("automorphism");
        {
            auto s = secret(N, 0, 106, 2);
            auto ctx = std::make_shared<AutomorphismContext>(BINARY, Q, N, 1 << 16, 4, 3.2, 5);
            ctx->SetSeed(106);
            auto ev = ctx->ConstructOperator({GenericKey("S", s)});
            auto in = g.modq(2 * N);
            auto prev = g.modq(2 * N);
            auto out = prev, out_sw = prev;
            ev->Eval(out, in);
            ev->EvalSwitchOnly(out_sw.data(), in.data());
            dump("auto_in", in);
            dump("auto_prev", prev);
            dump("auto_out", out);
            dump("auto_switch_only", out_sw);
        }

        step("trace");
        {
            auto s = secret(N, 0, 107, 2);
            auto ctx = std::make_shared<TraceEvaluationContext>(BINARY, Q, N, 1 << 16, 4, 3.2);
            ctx->SetSeed(107);
            auto tr = ctx->ConstructOperator({GenericKey("S", s)});
            auto in = g.modq(2 * N);
            std::vector<uint64_t> out(2 * N), a_ptr(2 * N, 0), a_vec(2 * N, 0);
            tr->Eval(out, in);
            tr->EvalAuto(a_ptr.data(), in.data(), 9);
            tr->EvalAuto(a_vec, in, 9);
            dump("trace_in", in);
            dump("trace_out", out);
            dump("trace_auto_ptr", a_ptr);
            dump("trace_auto_vec", a_vec);
        }

        step("packing");
        {
            auto s = secret(N, 0, 108, 2);
            auto ctx = std::make_shared<LWEtoRLWEPackingContext>(BINARY, Q, N, 1 << 16, 4, 3.2);
            ctx->SetSeed(108);
            auto pk = ctx->ConstructOperator({GenericKey("S", s)});
            auto lwes = g.modq(16 * (N + 1));
            std::vector<uint64_t> out(2 * N), out_c(2 * N);
            pk->Pack(out.data(), lwes.data(), 16);
            pk->PackConsecutively(out_c.data(), lwes.data(), 5);
            dump("pack_lwes", lwes);
            dump("pack_out", out);
            dump("pack_consecutive", out_c);
        }

        step("math worker");
        {
            auto w = SelectWorker(Q, N);
            auto x = g.modq(N);
            std::vector<uint64_t> y(N), z(N);
            w->ForwardNTT(y.data(), x.data());
            w->BackwardNTT(z.data(), y.data());
            const uint64_t L = 1000;
            auto a = g.modq(L), b = g.modq(L), c = g.modq(L);
            std::vector<uint64_t> r(6 * L);
            w->AddModElw(r.data(), a.data(), b.data(), L);
            w->SubModElw(r.data() + L, a.data(), b.data(), L);
            w->MulModElw(r.data() + 2 * L, a.data(), b.data(), L);
            w->FMAModElw(r.data() + 3 * L, a.data(), b.data(), c.data(), L);
            w->MulSModElw(r.data() + 4 * L, a.data(), 0x123456789ABCDEFULL, L);
            auto raw = std::vector<uint64_t>(L);
            for (auto& v : raw) v = g.next();
            w->ReduceModElw(r.data() + 5 * L, raw.data(), L);
            auto eq = a;
            w->AddEqModElw(eq.data(), b.data(), L);
            std::vector<uint64_t> sc = {w->AddMod(a[0], b[0]), w->SubMod(a[0], b[0]), w->MulMod(a[0], b[0]),
                                        w->FMAMod(a[0], b[0], c[0]), w->InvMod(12345), w->PowMod(7, 1000),
                                        w->InvMod(0), w->GetModulus(), w->GetDimension()};
            dump("mw_x", x);
            dump("mw_ntt", y);
            dump("mw_back", z);
            dump("mw_a", a);
            dump("mw_b", b);
            dump("mw_c", c);
            dump("mw_raw", raw);
            dump("mw_elw", r);
            dump("mw_addeq", eq);
            dump("mw_scalar", sc);
            bool threw = false;
            try {
                SelectWorker(Q - 2, N);
            } catch (const std::invalid_argument&) {
                threw = true;
            }
            if (!threw) throw std::runtime_error("SelectWorker accepted a foreign modulus");
        }

        step("lwe to rgsw");
        {
            const uint64_t n = 16;
            auto lk = secret(n, 0, 109, 1), rk = secret(N, 0, 109, 2);
            auto br = std::make_shared<CGGIBlindRotationContext>(BINARY, Q, N, n, 1 << 16, 4, 3.2);
            br->SetSeed(109);
            auto ntt = br->GetNTT();
            auto trace = std::make_shared<TraceEvaluationContext>(BINARY, ntt, 1 << 16, 4, 3.2);
            trace->SetSeed(110);
            auto square = std::make_shared<RLWEConversionContext>(BINARY, ntt, 1 << 16, 4, 3.2);
            square->SetSeed(111);
            auto rot = std::reinterpret_pointer_cast<OperatorContext<BlindRotator>>(br);
            auto ss = std::make_shared<SchemeSwitchingContext>(rot, trace, square, 4, 1 << 10);
            auto conv = ss->ConstructOperator({GenericKey("LWE", lk), GenericKey("RLWE", rk)});
            auto lwe = g.below(n + 1, 2 * N);
            std::vector<uint64_t> out(4 * N * 4, 0);
            conv->Convert(out, lwe);
            dump("rgsw_lwe", lwe);
            dump("rgsw_out", out);
        }

        step("sap decomposition");
        {
            const uint64_t n = 16, qin = 1ULL << 40;
            auto lk = secret(n, 0, 112, 1), rk = secret(N, 0, 112, 2);
            auto br = std::make_shared<CGGIBlindRotationContext>(BINARY, Q, N, n, 1 << 16, 4, 3.2);
            br->SetSeed(112);
            auto trace = std::make_shared<TraceEvaluationContext>(BINARY, br->GetNTT(), 1 << 16, 4, 3.2);
            trace->SetSeed(113);
            auto lconv = std::make_shared<LWEConversionContext>(BINARY, Q, N, n, 1 << 4, 14, 3.2);
            lconv->SetSeed(114);
            auto rot = std::reinterpret_pointer_cast<OperatorContext<BlindRotator>>(br);
            auto sap = std::make_shared<SAPDecompositionContext>(rot, trace, lconv, 32);
            // the trace's noise estimate is 0, so the accumulator is never reset (sap_decomposer.cpp:101-104)
            if (sap->GetRestartIteration() != UINT64_MAX) throw std::runtime_error("unexpected SAP restart period");
            auto dec = sap->ConstructOperator({GenericKey("LWE", lk), GenericKey("RLWE", rk)});
            auto lwe = g.below(n + 1, qin);
            const uint64_t msg = g.below(1, qin)[0];
            unsigned __int128 b = msg;
            for (uint64_t i = 0; i < n; ++i) b += (unsigned __int128)lwe[i] * lk[i];
            lwe[n] = (uint64_t)(b % qin);
            std::vector<uint64_t> out((N + 1) * 8, 0);
            dec->Decompose(out.data(), lwe.data(), 32);
            dump("sap_lwe", lwe);
            dump("sap_msg", std::vector<uint64_t>{msg});
            dump("sap_out", out);
            auto x = g.modq(2 * N);
            std::vector<uint64_t> ht(2 * N);
            dec->HomTrunc(ht.data(), x.data(), 32);
            dump("sap_ht_in", x);
            dump("sap_ht", ht);
            auto r = g.modq(2 * N);
            dump("sap_reset_in", r);
            dec->ResetAccumulator(r.data(), 32);
            dump("sap_reset", r);
        }

        step("modulus switch");
        {
            auto v = g.modq(4096);
            // near-integer quotients k Q / 2048 (where the reference's long double ceiling differs
            // from the exact one) and near-half-integers
            for (uint64_t k = 1; k < 200; ++k) {
                v[k] = (uint64_t)(((unsigned __int128)(2047 - k) * Q + 2047) / 2048);
                v[200 + k] = (uint64_t)(((unsigned __int128)(2 * k + 1) * Q) / 4096);
            }
            dump("ms_in", v);
            for (int kind = 0; kind < 3; ++kind) {
                auto w = v;
                ModulusSwitch(w.data(), w.size(), Q, 2048, (ModulusSwitchType)kind);
                dump("ms_out" + std::to_string(kind), w);
            }
            auto r = v;
            ModulusSwitch(r.data(), r.size(), Q, 2048, ModulusSwitchType::RANDOM);
            dump("ms_random", r);
        }

        step("sample extract");
        {
            auto in = g.modq(2 * N);
            in[N - 1] = 0;  // a zero coefficient: its output slot is left as the caller had it
            std::vector<uint64_t> o0(N + 1, 0), o5(N + 1, 0);
            SampleExtract<uint64_t>(o0.data(), in.data(), 0, N, Q);
            SampleExtract<uint64_t>(o5.data(), in.data(), 5, N, Q);
            dump("se_in", in);
            dump("se_0", o0);
            dump("se_5", o5);
        }
    } catch (const std::exception& e) {
        std::cerr << "dropin_parity: " << e.what() << std::endl;
        return 1;
    }
    step("done");
    return 0;
}
