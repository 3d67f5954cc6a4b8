// Writes the N = 1024 transform tables the kernels read (fhetoothpaste_amd/csrc/host_ntt.hpp) as
// raw little-endian words: psi_rev, psi_inv_rev, twisted64 (forward, inverse), half_tables
// (forward, inverse).  tests/test_ntt_tables.py checks them against the negacyclic transform.
#include <cstdio>
#include <vector>

#include "host_ntt.hpp"

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const tp::HostNtt T(1024);
    std::vector<std::vector<uint64_t>> parts = {T.psi_rev, T.psi_inv_rev, T.twisted64(false), T.twisted64(true),
                                                T.half_tables(false), T.half_tables(true)};
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 1;
    for (const auto& p : parts)
        if (std::fwrite(p.data(), 8, p.size(), f) != p.size()) return 1;
    return std::fclose(f) == 0 ? 0 : 1;
}
