// Checks of the stratified trace order (st_common.h): the Hilbert curve
// visits every cell of the 256 x 256 grid once, consecutive positions are
// neighbouring cells, and stratified_group_of_cell inverts the keyed Feistel
// permutation stratified_cell_of_group for several sweeps, seeds and jobs.
// Prints "order cells <distinct> steps_not_adjacent <n> inverse_mismatch <n>".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "st_common.h"

using namespace st;

int main() {
    std::vector<int> seen(65536, 0);
    long distinct = 0, far = 0, mismatch = 0;
    uint32_t prev = hilbert_cell(0);
    for (uint32_t h = 0; h < 65536; ++h) {
        const uint32_t c = hilbert_cell(h);
        if (c < 65536 && !seen[c]++) ++distinct;
        if (h) {
            const int dx = (int)(c & 255u) - (int)(prev & 255u), dy = (int)(c >> 8) - (int)(prev >> 8);
            if (std::abs(dx) + std::abs(dy) != 1) ++far;
        }
        prev = c;
    }
    const uint64_t sweeps[3] = {0, 7, 123456789ull};
    const uint32_t seeds[3][2] = {{0, 0}, {12345, 0}, {0xDEADBEEF, 0x1234}};
    for (const uint64_t sw : sweeps)
        for (const auto& sd : seeds)
            for (uint32_t job = 0; job < 3; ++job)
                for (uint32_t cell = 0; cell < 65536; ++cell) {
                    const uint64_t g = stratified_group_of_cell(sw, cell, sd[0], sd[1], job);
                    if ((g >> 16) != sw || stratified_cell_of_group(g, sd[0], sd[1], job) != cell) ++mismatch;
                }
    std::printf("order cells %ld steps_not_adjacent %ld inverse_mismatch %ld\n", distinct, far, mismatch);
    return 0;
}
