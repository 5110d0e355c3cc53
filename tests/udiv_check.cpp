// Host-side check of the multiply-shift division and the tile decode of the
// one-plane-per-workgroup kernels (stencil_benchmarks_amd/csrc/sb_internal.h):
// udiv against the hardware division for dividends below 2^31 (every edge
// case around multiples of the divisor, the largest dividend, and random
// ones), and tile_of against xcd_remap followed by plain division, over
// every workgroup of several grids.  Exits non-zero on the first mismatch.
#include <cstdio>
#include <random>

#include "../stencil_benchmarks_amd/csrc/sb_internal.h"

using sb::UDiv;

static int check_div(uint32_t d, std::mt19937_64 &rng) {
    const UDiv u = sb::make_udiv(d);
    const uint32_t top = 0x7fffffffu;
    const uint32_t edges[] = {0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, top, top - 1,
                              (top / d) * d, (top / d) * d - 1, (top / d) * d + (d - 1) / 2};
    for (uint32_t n : edges) {
        if (n > top) continue;
        if (sb::udiv(n, u) != n / d) {
            std::printf("udiv(%u, %u) = %u, expected %u\n", n, d, sb::udiv(n, u), n / d);
            return 1;
        }
    }
    for (int i = 0; i < 20000; ++i) {
        const uint32_t n = uint32_t(rng()) & top;
        if (sb::udiv(n, u) != n / d) {
            std::printf("udiv(%u, %u) = %u, expected %u\n", n, d, sb::udiv(n, u), n / d);
            return 1;
        }
    }
    return 0;
}

static int check_grid(uint32_t tiles_i, uint32_t tiles_j, uint32_t nz, int xcds) {
    const uint32_t blocks = tiles_i * tiles_j * nz;
    const sb::TileGrid g = sb::make_tile_grid(blocks, tiles_i, tiles_j, xcds);
    for (uint32_t b = 0; b < blocks; ++b) {
        uint32_t ti, tj, k;
        sb::tile_of(b, g, ti, tj, k);
        const uint32_t r = sb::xcd_remap(b, blocks, uint32_t(xcds));
        const uint32_t t = r % (tiles_i * tiles_j);
        if (k != r / (tiles_i * tiles_j) || tj != t / tiles_i || ti != t % tiles_i) {
            std::printf("grid %ux%ux%u xcds %d: block %u -> (%u, %u, %u), expected (%u, %u, %u)\n",
                        tiles_i, tiles_j, nz, xcds, b, ti, tj, k, t % tiles_i, t / tiles_i,
                        r / (tiles_i * tiles_j));
            return 1;
        }
    }
    return 0;
}

int main() {
    std::mt19937_64 rng(20261203);
    long divisors = 0;
    for (uint32_t d = 1; d <= 70000; ++d, ++divisors)
        if (check_div(d, rng)) return 1;
    for (int s = 17; s <= 31; ++s)
        for (int off = -2; off <= 2; ++off) {
            const uint32_t d = uint32_t((uint64_t(1) << s) + int64_t(off));
            if (d > 0x80000000u) continue;
            ++divisors;
            if (check_div(d, rng)) return 1;
        }
    for (int i = 0; i < 2000; ++i, ++divisors)
        if (check_div(uint32_t(rng() % 0x80000000u) + 1, rng)) return 1;
    const uint32_t grids[][3] = {{8, 64, 160}, {8, 32, 160}, {3, 7, 5}, {1, 1, 9}, {5, 1, 1},
                                 {2, 3, 1}, {16, 12, 40}, {7, 11, 13}};
    long grid_count = 0;
    for (const auto &gr : grids)
        for (int xcds : {1, 2, 8}) {
            ++grid_count;
            if (check_grid(gr[0], gr[1], gr[2], xcds)) return 1;
        }
    std::printf("ok: %ld divisors, %ld grids\n", divisors, grid_count);
    return 0;
}
