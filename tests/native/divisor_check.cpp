// Checks nnk::make_divisor / nnk::divide (the network kernels' index splitting) against
// integer division for the divisors 1..4096 and a few large ones: every n below 2^22 (divisors
// up to 64) or 2^16 (the rest), 2^16 random n below 2^31 and the edges.  Prints "ok" or the
// first failure.
#include <stdio.h>

#include <random>

#include "../../boltzmann_amd/csrc/nn_kernels.hpp"

int main() {
    std::vector<uint32_t> ds;
    for (uint32_t d = 1; d <= 4096; ++d) ds.push_back(d);
    for (uint32_t d : {6144u, 65536u, 98304u, 1000003u, (1u << 30) + 7u, 2147483647u}) ds.push_back(d);
    std::mt19937 rng(0);
    for (uint32_t d : ds) {
        const nnk::Divisor v = nnk::make_divisor(d);
        const uint32_t exhaustive = d <= 64 ? (1u << 22) : (1u << 16);
        for (uint32_t n = 0; n < exhaustive; ++n)
            if (nnk::divide(n, v) != n / d) {
                printf("FAIL d=%u n=%u\n", d, n);
                return 1;
            }
        for (int k = 0; k < (1 << 16); ++k) {
            const uint32_t n = rng() & 0x7fffffffu;
            if (nnk::divide(n, v) != n / d) {
                printf("FAIL d=%u n=%u\n", d, n);
                return 1;
            }
        }
        const uint32_t edge[] = {0x7fffffffu, 0x7ffffffeu, 0x7fffffffu - d + 1, d - 1, d, d + 1};
        for (uint32_t n : edge)
            if (nnk::divide(n, v) != n / d) {
                printf("FAIL d=%u n=%u\n", d, n);
                return 1;
            }
    }
    printf("ok\n");
    return 0;
}
