// Order of a job's tiles in a work list.
//
// A job's output tiles are listed in 8 x 8 super-tiles, column of
// super-tiles by column (super-tile (I, J) covers row tiles I*8 .. I*8+7 and
// column tiles J*8 .. J*8+7), each super-tile column by column, so that
// tiles that run at the same time share operand panels.  A single region
// (tri) lists only the tiles with ti <= tj: every cell j > i that the
// reference prints (correlate_IO.c:403-406) lies in one of them.
//
// append_job_tiles() writes that list on the host; job_tile_index() gives the
// position of one tile in it in closed form, so the list can also be written
// on the device, one thread per tile, with no host buffer and no copy.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

// One tile: rows [ti*T, ti*T + T) x columns [tj*T, tj*T + T) of job `desc`.
struct WorkItem {
    uint32_t desc;
    uint16_t ti, tj;
};

#define QLD_SUPER_TILE 8

inline size_t job_tile_count(int ni, int nj, int tri, int T)
{
    size_t a = (size_t)((ni + T - 1) / T), b = (size_t)((nj + T - 1) / T);
    return tri ? a * (a + 1) / 2 : a * b;
}

inline void append_job_tiles(WorkItem *out, size_t *count, uint32_t desc, int ni, int nj,
                             int tri, int T)
{
    const int S = QLD_SUPER_TILE;
    const int ti_count = (ni + T - 1) / T, tj_count = (nj + T - 1) / T;
    const int si = (ti_count + S - 1) / S, sj = (tj_count + S - 1) / S;
    size_t k = *count;
    for (int J = 0; J < sj; ++J)
        for (int I = 0; I < si; ++I) {
            if (tri && I > J) continue;
            for (int tj = J * S; tj < (J + 1) * S && tj < tj_count; ++tj)
                for (int ti = I * S; ti < (I + 1) * S && ti < ti_count; ++ti)
                    if (!tri || ti <= tj) out[k++] = WorkItem{desc, (uint16_t)ti, (uint16_t)tj};
        }
    *count = k;
}

// Position of tile (ti, tj) in append_job_tiles' list (ti <= tj when tri;
// for tri, ti_count == tj_count).  With S the super-tile edge,
// (I, J) = (ti / S, tj / S), (a, b) the offsets inside it and rows(I) /
// cols(J) the tile counts of super-tile row I / column J (S except at the
// last one):
//   rectangle: every super-tile column before J is full,
//     J*S*ti_count + I*S*cols(J) + b*rows(I) + a;
//   triangle: super-tile column J' < J holds J' full blocks and a diagonal
//     block of S(S+1)/2 tiles, so
//     S*S*J(J-1)/2 + J*S(S+1)/2 + I*S*cols(J) + (I < J ? b*S + a
//                                                      : b(b+1)/2 + a).
__host__ __device__ inline int64_t job_tile_index(int ti, int tj, int ti_count, int tj_count,
                                                  int tri)
{
    const int64_t S = QLD_SUPER_TILE;
    const int64_t I = ti / S, J = tj / S, a = ti - I * S, b = tj - J * S;
    const int64_t rows = ti_count - I * S < S ? ti_count - I * S : S;
    const int64_t cols = tj_count - J * S < S ? tj_count - J * S : S;
    if (!tri) return J * S * ti_count + I * S * cols + b * rows + a;
    const int64_t before = S * S * (J * (J - 1) / 2) + J * (S * (S + 1) / 2) + I * S * cols;
    return before + (I < J ? b * S + a : b * (b + 1) / 2 + a);
}

// Wave blocks of one job (WTM x WTN blocks inside tiles of edge T) that hold
// at least one cell to compute, as units (desc, bi, bj): block rows
// [bi*WTM, bi*WTM + WTM), columns [bj*WTN, bj*WTN + WTN).  Super-tiles come
// in append_job_tiles' order; inside one, row band by row band (WTM rows),
// each band left to right across the super-tile's tiles, so that runs of
// consecutive units share their rows -- the kernels hand a workgroup's
// waves such runs together -- however many blocks a band has.  A block is
// left out when it has no row or no column inside the job, or no cell of a
// triangle's (no column inside the job right of any of its rows).  Returns the
// units written; `out` needs room for job_tile_count(...) * (T/WTM) * (T/WTN).
inline size_t append_job_wave_units(WorkItem *out, uint32_t desc, int ni, int nj, int tri,
                                    int T, int WTM, int WTN)
{
    const int S = QLD_SUPER_TILE;
    const int ti_count = (ni + T - 1) / T, tj_count = (nj + T - 1) / T;
    const int si = (ti_count + S - 1) / S, sj = (tj_count + S - 1) / S;
    const int BM = T / WTM, BN = T / WTN;
    size_t k = 0;
    for (int J = 0; J < sj; ++J)
        for (int I = 0; I < si; ++I) {
            if (tri && I > J) continue;
            for (int ti = I * S; ti < (I + 1) * S && ti < ti_count; ++ti)
                for (int wm = 0; wm < BM; ++wm)
                    for (int tj = J * S; tj < (J + 1) * S && tj < tj_count; ++tj) {
                        if (tri && ti > tj) continue;
                        for (int wn = 0; wn < BN; ++wn) {
                            const int a0 = ti * T + wm * WTM, b0 = tj * T + wn * WTN;
                            if (a0 >= ni || b0 >= nj) continue;
                            // a triangle's block holds a cell j > i iff its
                            // last column inside the job lies right of its
                            // first row
                            const int jl = (b0 + WTN < nj ? b0 + WTN : nj) - 1;
                            if (tri && jl <= a0) continue;
                            out[k++] = WorkItem{desc, (uint16_t)(ti * BM + wm),
                                                (uint16_t)(tj * BN + wn)};
                        }
                    }
        }
    return k;
}

// 32 x 32 MFMA blocks the per-wave kernels compute for unit u of a job (the
// kernels' block variants, ld_mfma.hip): 128 x 64 blocks (WTM = 128) run in
// full, as their first 32 columns only on the job's ragged right edge, or as
// the 3 blocks (1 on the edge) of a triangle's diagonal; 64 x 64 blocks skip
// the 32-row / 32-column halves past the job's edge and the block left of a
// diagonal block's diagonal.
inline int wave_unit_blocks(const WorkItem &u, int ni, int nj, int tri, int WTM, int WTN)
{
    const int a0 = u.ti * WTM, b0 = u.tj * WTN;
    const bool nfull = b0 + 32 < nj;
    if (WTM == 128) {
        if (tri && b0 == a0) return nfull ? 3 : 1;
        return nfull ? 8 : 4;
    }
    const bool mfull = a0 + 32 < ni;
    if (mfull && nfull) return tri && a0 == b0 ? 3 : 4;
    return (mfull ? 2 : 1) * (nfull ? 2 : 1);
}

// Reorders one job's units (as append_job_wave_units wrote them) so that the
// full blocks come first, in their order, and the partial ones after them,
// most blocks first: the kernels hand runs of consecutive units to a
// workgroup's waves side by side, and a run that mixes a diagonal block
// with full ones leaves three of its waves idle for most of the run.
inline void order_wave_units_by_work(WorkItem *units, size_t n, int ni, int nj, int tri,
                                     int WTM, int WTN)
{
    std::stable_sort(units, units + n, [&](const WorkItem &x, const WorkItem &y) {
        return wave_unit_blocks(x, ni, nj, tri, WTM, WTN) >
               wave_unit_blocks(y, ni, nj, tri, WTM, WTN);
    });
}
