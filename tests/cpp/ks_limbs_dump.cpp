// Reads a key-switching key ([n_in][d][n_out + 1] little-endian words) and writes it as the
// matrix-core operands the key-switch kernel reads (fhetoothpaste_amd/csrc/ks_limbs.hpp), followed
// by the 16 slot digits.  tests/test_ks_limbs.py models the kernel on them against the oracle.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ks_limbs.hpp"

int main(int argc, char** argv) {
    if (argc != 7) return 2;
    const uint32_t n_in = (uint32_t)std::atoi(argv[3]), n_out = (uint32_t)std::atoi(argv[4]);
    const uint32_t b = (uint32_t)std::atoi(argv[5]), d = (uint32_t)std::atoi(argv[6]);
    if (!tp::ks_mfma_fits(n_in, b, d)) return 3;
    const size_t words = (size_t)n_in * d * (n_out + 1);
    std::vector<uint64_t> ksk(words);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(ksk.data(), 8, words, f) != words) return 1;
    std::fclose(f);
    std::vector<int8_t> kl(tp::ks_limb_bytes(n_in, n_out));
    tp::build_ks_limbs(ksk.data(), n_in, n_out, b, d, kl.data());
    int32_t slots[16];
    for (int p = 0; p < 16; ++p) slots[p] = tp::ks_slot_digit((int)b, (int)d, p);
    FILE* g = std::fopen(argv[2], "wb");
    if (!g || std::fwrite(kl.data(), 1, kl.size(), g) != kl.size() || std::fwrite(slots, 4, 16, g) != 16) return 1;
    return std::fclose(g) == 0 ? 0 : 1;
}
