// Host check for test_abi.py::test_tile_index_matches_tile_list: the closed
// form job_tile_index() (used by the device tile-list kernel) names, for
// every tile grid up to 70 x 70 tiles, exactly the position append_job_tiles
// gives each tile, rectangles and triangles.  Prints "ok <grids>" on success.
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
    printf("ok %ld\n", grids);
    return 0;
}
