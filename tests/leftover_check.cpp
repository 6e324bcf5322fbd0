// The leftover-node enumeration of the colour-pair launches (vsim::Leftover)
// against brute force: for every lattice of 1..21 cells a side and each row
// class, the nodes no row of the class covers, in node-id order. Built and run
// on the CPU by test_leftover_nodes.py.
#include "../visionary-sim_amd/csrc/kernels.h"
#include <cstdio>
#include <vector>
int main()
{
    long bad = 0, cases = 0;
    for (uint32_t cx = 1; cx <= 21; cx++)
        for (uint32_t cy = 1; cy <= 21; cy++)
            for (uint32_t cz = 1; cz <= 21; cz++)
                for (uint32_t cls = 0; cls < 4; cls++) {
                    const uint32_t d[3] = { cx, cy, cz };
                    const uint32_t nx = cx + 1, ny = cy + 1, nz = cz + 1, N = nx * ny * nz;
                    std::vector<char> cov(N, 0);
                    const uint32_t py = cls & 1, pz = cls >> 1;
                    // rows of the class: cells y = py + 2a, z = pz + 2b, all x; cover nodes y..y+1, z..z+1
                    for (uint32_t z = pz; z < cz; z += 2)
                        for (uint32_t y = py; y < cy; y += 2)
                            for (uint32_t dz = 0; dz < 2; dz++)
                                for (uint32_t dy = 0; dy < 2; dy++)
                                    for (uint32_t x = 0; x < nx; x++) cov[x + nx * (y + dy + ny * (z + dz))] = 1;
                    std::vector<uint32_t> want;
                    for (uint32_t k = 0; k < N; k++)
                        if (!cov[k]) want.push_back(k);
                    vsim::Leftover L(d, cls);
                    cases++;
                    bool ok = L.total == want.size();
                    for (uint32_t i = 0; ok && i < L.total; i++) ok = L.node(i) == want[i];
                    if (!ok && bad++ < 5) printf("bad %u %u %u cls %u total %u want %zu\n", cx, cy, cz, cls, L.total, want.size());
                }
    printf("%ld cases, %ld bad\n", cases, bad);
    return bad != 0;
}
