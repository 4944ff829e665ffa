// Tile enumeration shared by the all-pairs kernels.
//
// A launch covers either the full tile grid of an n x m rectangle (row tiles
// over table B, column tiles over table A) or, for one region against itself,
// only the tiles with tj >= ti (every cell j > i that writeResults prints,
// correlate_IO.c:403-406 of the reference, lies in such a tile).
//
// MI355X dispatches consecutive workgroup ids round-robin over its 8 XCDs,
// each with a private L2.  tile_of_block() first undoes that interleave, so
// that a run of consecutive logical tiles -- which share table panels -- lands
// on one XCD, then maps the logical index to (ti, tj).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define QLD_NUM_XCD 8
#define QLD_GROUP_ROWS 8   // row tiles per swizzle group (rectangles)

struct TileMap {
    int ti_count;   // tiles along rows (table B)
    int tj_count;   // tiles along columns (table A)
    int tri;        // 1: only tj >= ti
    int64_t total;  // number of tiles = workgroups
};

static inline TileMap make_tile_map(int n, int m, int tile_i, int tile_j, int tri)
{
    TileMap t;
    t.ti_count = (n + tile_i - 1) / tile_i;
    t.tj_count = (m + tile_j - 1) / tile_j;
    t.tri = tri;
    if (tri)
        t.total = (int64_t)t.ti_count * (t.ti_count + 1) / 2;
    else
        t.total = (int64_t)t.ti_count * t.tj_count;
    return t;
}

// Bijection on [0, G): workgroup b runs on XCD b % 8; give each XCD a
// contiguous range of logical indices.
__device__ __forceinline__ int64_t xcd_logical(int64_t b, int64_t G)
{
    int64_t x = b % QLD_NUM_XCD, local = b / QLD_NUM_XCD;
    int64_t q = G / QLD_NUM_XCD, r = G % QLD_NUM_XCD;
    return x * q + (x < r ? x : r) + local;
}

__device__ __forceinline__ void tile_of_logical(const TileMap &t, int64_t L,
                                                int &ti, int &tj)
{
    if (t.tri) {
        // L = tj*(tj+1)/2 + ti, 0 <= ti <= tj
        int64_t c = (int64_t)((sqrt(8.0 * (double)L + 1.0) - 1.0) * 0.5);
        while (c * (c + 1) / 2 > L) --c;
        while ((c + 1) * (c + 2) / 2 <= L) ++c;
        tj = (int)c;
        ti = (int)(L - c * (c + 1) / 2);
    } else {
        int64_t per_group = (int64_t)QLD_GROUP_ROWS * t.tj_count;
        int64_t g = L / per_group;
        int64_t first = g * QLD_GROUP_ROWS;
        int64_t gsz = t.ti_count - first < QLD_GROUP_ROWS ? t.ti_count - first
                                                           : QLD_GROUP_ROWS;
        int64_t w = L % per_group;
        ti = (int)(first + w % gsz);
        tj = (int)(w / gsz);
    }
}

__device__ __forceinline__ void tile_of_block(const TileMap &t, int64_t b,
                                              int64_t G, int &ti, int &tj)
{
    tile_of_logical(t, xcd_logical(b, G), ti, tj);
}
