// Versioned wire format for keys and ciphertext batches (host code).  The reference has no
// serialisation; its key layouts (base_crypto.cpp:60-97, cggi_blind_rotation.cpp:264-311,
// lwe_conversion.cpp:174-203, rlwe_conversion.cpp:187-218) are kept word for word as the payload,
// so a file is the reference-layout array plus a header that says what it is:
//
//   offset  size  field
//   0       4     magic "TPW1"
//   4       4     version (1)
//   8       4     kind (TP_WIRE_*)
//   12      4     flags (bit 0: payload in the NTT domain)
//   16      40    tp_params (10 x uint32) the payload belongs to
//   56      8     payload words
//   64      8     FNV-1a 64 of the payload bytes
//   72      ...   payload, little-endian uint64
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/toothpaste_amd.h"
#include "context_internal.hpp"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "payload words are copied as little-endian bytes");

namespace {

constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 72;

uint64_t fnv1a(const uint8_t* p, size_t n) {
    uint64_t h = 0xCBF29CE484222325ULL;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001B3ULL;
    }
    return h;
}

void put32(uint8_t* p, uint32_t v) {
    for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
void put64(uint8_t* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get32(const uint8_t* p) {
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)p[i] << (8 * i);
    return v;
}
uint64_t get64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

// Payload size a key kind must have for its parameters (0: any size).
size_t expected_words(int kind, const tp_params& p, size_t words) {
    const size_t N = p.ring_dim;
    int logN = 0;
    while ((1ULL << logN) < N) ++logN;
    switch (kind) {
        case TP_WIRE_BRK: {
            const size_t rg = 4 * N * p.br_digits;
            if (p.method == TP_METHOD_CGGI_BINARY) return rg * p.lwe_dim;
            if (p.method == TP_METHOD_CGGI_TERNARY) return 2 * rg * p.lwe_dim;
            return 3 * rg * (p.lwe_dim / 2);
        }
        case TP_WIRE_LWE_KSK: return N * p.ks_digits * ((size_t)p.ks_out_dim + 1);
        case TP_WIRE_AUTO_KEYS: return (size_t)logN * p.rlwe_ks_digits * 2 * N;
        case TP_WIRE_RLWE_KSK: return (size_t)p.rlwe_ks_digits * 2 * N;
        case TP_WIRE_LWE_BATCH: return words % ((size_t)p.lwe_dim + 1) == 0 ? words : 0;
        case TP_WIRE_RLWE_BATCH: return N && words % (2 * N) == 0 ? words : 0;
        default: return words;
    }
}

int werr(int code, const std::string& msg) { return tp::set_error(code, msg.c_str()); }

}  // namespace

extern "C" {

int tp_wire_bytes(size_t words, size_t* bytes) {
    if (!bytes) return werr(TP_E_INVALID, "null argument");
    *bytes = kHeader + words * 8;
    return TP_OK;
}

int tp_wire_encode(uint8_t* buf, size_t buf_bytes, int kind, uint32_t flags, const tp_params* params,
                   const uint64_t* data, size_t words) {
    if (!buf || !params || (words && !data)) return werr(TP_E_INVALID, "null argument");
    if (kind < TP_WIRE_RAW || kind > TP_WIRE_RLWE_BATCH) return werr(TP_E_INVALID, "unknown wire kind");
    if (buf_bytes < kHeader + words * 8) return werr(TP_E_INVALID, "buffer too small");
    if (expected_words(kind, *params, words) != words)
        return werr(TP_E_INVALID, "payload size does not match the parameters");
    std::memcpy(buf, "TPW1", 4);
    put32(buf + 4, kVersion);
    put32(buf + 8, (uint32_t)kind);
    put32(buf + 12, flags);
    const uint32_t* pv = reinterpret_cast<const uint32_t*>(params);
    for (int i = 0; i < 10; ++i) put32(buf + 16 + 4 * i, pv[i]);
    put64(buf + 56, words);
    uint8_t* payload = buf + kHeader;
    std::memcpy(payload, data, words * 8);  // little-endian host: the payload's byte order
    put64(buf + 64, fnv1a(payload, words * 8));
    return TP_OK;
}

// Reads the header (and, when data is non-null, the payload into data[0..cap_words)).
int tp_wire_decode(const uint8_t* buf, size_t buf_bytes, int* kind, uint32_t* flags, tp_params* params,
                   uint64_t* data, size_t cap_words, size_t* words) {
    if (!buf || !words) return werr(TP_E_INVALID, "null argument");
    if (buf_bytes < kHeader || std::memcmp(buf, "TPW1", 4) != 0) return werr(TP_E_INVALID, "not a TPW1 record");
    if (get32(buf + 4) != kVersion) return werr(TP_E_INVALID, "unsupported wire version");
    const size_t n = get64(buf + 56);
    if (n > (buf_bytes - kHeader) / 8) return werr(TP_E_INVALID, "truncated record");
    const int k = (int)get32(buf + 8);
    tp_params p{};
    uint32_t* pv = reinterpret_cast<uint32_t*>(&p);
    for (int i = 0; i < 10; ++i) pv[i] = get32(buf + 16 + 4 * i);
    if (k < TP_WIRE_RAW || k > TP_WIRE_RLWE_BATCH) return werr(TP_E_INVALID, "unknown wire kind");
    if (expected_words(k, p, n) != n) return werr(TP_E_INVALID, "payload size does not match the parameters");
    const uint8_t* payload = buf + kHeader;
    if (fnv1a(payload, n * 8) != get64(buf + 64)) return werr(TP_E_INVALID, "checksum mismatch");
    *words = n;
    if (kind) *kind = k;
    if (flags) *flags = get32(buf + 12);
    if (params) *params = p;
    if (data) {
        if (cap_words < n) return werr(TP_E_INVALID, "output buffer too small");
        std::memcpy(data, payload, n * 8);
    }
    return TP_OK;
}

int tp_wire_save(const char* path, int kind, uint32_t flags, const tp_params* params, const uint64_t* data,
                 size_t words) {
    if (!path) return werr(TP_E_INVALID, "null path");
    std::vector<uint8_t> buf(kHeader + words * 8);
    int rc = tp_wire_encode(buf.data(), buf.size(), kind, flags, params, data, words);
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return werr(TP_E_INVALID, std::string("cannot open ") + path);
    const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    const int c = std::fclose(f);
    if (w != buf.size() || c != 0) return werr(TP_E_INVALID, std::string("write failed: ") + path);
    return TP_OK;
}

// With data null, only *words (and the header fields) are returned, so a caller can size the buffer.
int tp_wire_load(const char* path, int* kind, uint32_t* flags, tp_params* params, uint64_t* data, size_t cap_words,
                 size_t* words) {
    if (!path || !words) return werr(TP_E_INVALID, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return werr(TP_E_INVALID, std::string("cannot open ") + path);
    std::vector<uint8_t> buf;
    uint8_t chunk[1 << 16];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + r);
    std::fclose(f);
    return tp_wire_decode(buf.data(), buf.size(), kind, flags, params, data, cap_words, words);
}

}  // extern "C"
