// Host check for test_abi.py::test_tile_index_matches_tile_list: the closed
// form job_tile_index() (used by the device tile-list kernel) names, for
// every tile grid up to 70 x 70 tiles, exactly the position append_job_tiles
// gives each tile, rectangles and triangles; and append_job_wave_units'
// blocks (both kernel geometries) cover every cell a job computes exactly
// once, none of them empty, with each super-tile's blocks in row-band order.
// Prints "ok <grids> <unit lists>" on success.
#include "../qld_amd/csrc/tile_order.h"

#include <stdio.h>
#include <vector>

int main()
{
    const int T = 128;
    long grids = 0;
    for (int tri = 0; tri <= 1; ++tri)
        for (int ic = 1; ic <= 70; ++ic)
            for (int jc = 1; jc <= 70; ++jc) {
                if (tri && ic != jc) continue;
                const int ni = ic * T - (ic * 37) % T, nj = jc * T - (jc * 53) % T;
                std::vector<WorkItem> items(job_tile_count(ni, nj, tri, T));
                size_t n = 0;
                append_job_tiles(items.data(), &n, 0, ni, nj, tri, T);
                if (n != items.size()) {
                    printf("count %d %d %d: %zu vs %zu\n", ic, jc, tri, n, items.size());
                    return 1;
                }
                for (size_t k = 0; k < n; ++k) {
                    const int64_t idx = job_tile_index(items[k].ti, items[k].tj, ic, jc, tri);
                    if (idx != (int64_t)k) {
                        printf("grid %d x %d tri %d: tile (%d, %d) at %zu, closed form %lld\n",
                               ic, jc, tri, items[k].ti, items[k].tj, k, (long long)idx);
                        return 1;
                    }
                }
                ++grids;
            }
    long lists = 0;
    const int geo[2][3] = {{256, 128, 64}, {128, 64, 64}};
    const int sizes[] = {1, 2, 31, 33, 64, 65, 127, 128, 129, 200, 255, 256, 257, 511, 700, 2100};
    for (const auto &g : geo)
        for (int tri = 0; tri <= 1; ++tri)
            for (int ni : sizes)
                for (int nj : sizes) {
                    if (tri && ni != nj) continue;
                    const int Tg = g[0], WTM = g[1], WTN = g[2];
                    std::vector<WorkItem> units(job_tile_count(ni, nj, tri, Tg) * (Tg / WTM) *
                                                (Tg / WTN));
                    const size_t n =
                        append_job_wave_units(units.data(), 7, ni, nj, tri, Tg, WTM, WTN);
                    std::vector<unsigned char> seen((size_t)ni * nj, 0);
                    long prev_key = -1;
                    for (size_t k = 0; k < n; ++k) {
                        const WorkItem u = units[k];
                        const int a0 = u.ti * WTM, b0 = u.tj * WTN;
                        // row-band order inside a super-tile
                        const int SI = (a0 / Tg) / QLD_SUPER_TILE, SJ = (b0 / Tg) / QLD_SUPER_TILE;
                        const long key = ((long)SJ * 4096 + SI) * 65536 + u.ti;
                        if (u.desc != 7 || key < prev_key) {
                            printf("unit order %d %d %d: unit %zu\n", ni, nj, tri, k);
                            return 1;
                        }
                        prev_key = key;
                        int cells = 0;
                        for (int i = a0; i < a0 + WTM && i < ni; ++i)
                            for (int j = b0; j < b0 + WTN && j < nj; ++j) {
                                if (tri && j <= i) continue;
                                if (seen[(size_t)i * nj + j]++) {
                                    printf("cell (%d, %d) twice: %d %d %d\n", i, j, ni, nj, tri);
                                    return 1;
                                }
                                ++cells;
                            }
                        if (!cells) {
                            printf("empty unit (%d, %d): %d %d %d\n", u.ti, u.tj, ni, nj, tri);
                            return 1;
                        }
                    }
                    for (int i = 0; i < ni; ++i)
                        for (int j = 0; j < nj; ++j)
                            if ((!tri || j > i) && !seen[(size_t)i * nj + j]) {
                                printf("cell (%d, %d) missed: %d %d %d\n", i, j, ni, nj, tri);
                                return 1;
                            }
                    // ordered by work: a permutation, most blocks first, list
                    // order kept among equals, block counts those of the
                    // kernels' variants (a block count covers its cells)
                    std::vector<WorkItem> byw(units.begin(), units.begin() + n);
                    order_wave_units_by_work(byw.data(), n, ni, nj, tri, WTM, WTN);
                    std::vector<size_t> pos((size_t)(ni / WTM + 2) * (nj / WTN + 2), 0);
                    for (size_t k = 0; k < n; ++k)
                        pos[(size_t)units[k].ti * (nj / WTN + 2) + units[k].tj] = k + 1;
                    for (size_t k = 0; k < n; ++k) {
                        const int w = wave_unit_blocks(byw[k], ni, nj, tri, WTM, WTN);
                        const size_t at = pos[(size_t)byw[k].ti * (nj / WTN + 2) + byw[k].tj];
                        if (!at || w < 1 || w > (WTM / 32) * (WTN / 32)) {
                            printf("work order %d %d %d: unit %zu\n", ni, nj, tri, k);
                            return 1;
                        }
                        if (k) {
                            const int wp = wave_unit_blocks(byw[k - 1], ni, nj, tri, WTM, WTN);
                            const size_t ap =
                                pos[(size_t)byw[k - 1].ti * (nj / WTN + 2) + byw[k - 1].tj];
                            if (wp < w || (wp == w && ap >= at)) {
                                printf("work order %d %d %d: unit %zu\n", ni, nj, tri, k);
                                return 1;
                            }
                        }
                    }
                    ++lists;
                }
    printf("ok %ld %ld\n", grids, lists);
    return 0;
}
