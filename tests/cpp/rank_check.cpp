// Host replay of the cell-order ranking of k_reorder (ranked_member in
// sim_kernels.hpp, compiled for the host by hipcc): for random cells of
// 1..kRankCap members, slot s0 + r must receive the member with the r-th
// smallest previous index, so every cell comes out sorted, whatever order
// the scatter's atomics left it in. Prints "ok" and exits 0 on success.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../weaklycompressiblesph_amd/csrc/sim_kernels.hpp"

int main() {
    std::mt19937 rng(2026);
    std::vector<uint32_t> sizes;
    for (uint32_t c = 1; c <= 20; ++c) sizes.push_back(c);
    for (uint32_t c : {31u, 64u, 100u, 255u, wcsph::kRankCap}) sizes.push_back(c);
    long checked = 0;
    for (uint32_t cnt : sizes) {
        for (int rep = 0; rep < 20; ++rep) {
            // a slot range [s0, s0 + cnt) inside a larger perm array
            const uint32_t s0 = 7 + rep;
            std::vector<uint32_t> perm(s0 + cnt + 5, 0xdeadbeefu);
            std::vector<uint32_t> members(cnt);
            for (uint32_t t = 0; t < cnt; ++t) members[t] = (uint32_t)(rng() % 5000000u);
            std::sort(members.begin(), members.end());
            members.erase(std::unique(members.begin(), members.end()), members.end());
            const uint32_t m = (uint32_t)members.size();
            std::vector<uint32_t> shuffled = members;
            std::shuffle(shuffled.begin(), shuffled.end(), rng);
            for (uint32_t t = 0; t < m; ++t) perm[s0 + t] = shuffled[t];
            for (uint32_t t = 0; t < m; ++t) {
                const uint32_t got = wcsph::ranked_member(perm.data(), s0, s0 + m, s0 + t);
                if (got != members[t]) {
                    std::printf("cell of %u: slot %u got %u, want %u\n", m, t, got, members[t]);
                    return 1;
                }
                ++checked;
            }
        }
    }
    std::printf("ok %ld slots\n", checked);
    return 0;
}
