// Popcount engine: pairwise AND + v_bcnt_u32_b32 over packed uint64 words,
// LDS-staged, with the r^2 epilogue fused into the tile.
//
// Replaces the OpenCL kernel blis_like8x4 (src/correlator/src/kernel/
// blislike.cl:109-293 of the reference) and the host loop around it
// (gemm_gpu.c:470-698): one launch covers the whole m x n (or triangular)
// problem, tables stay in HBM, and r^2 is produced in the same kernel.
//
// Work decomposition
//   workgroup : 256 lanes (4 wave64), one 128 x 128 tile of pairs
//   lane      : an 8 x 8 register tile, rows ty+16r, cols tx+16s
//   k loop    : 16 words per stage; the next stage's words are fetched into
//               registers while the current stage is consumed from LDS
// LDS holds the stage word-major (sX[k][row]) so a wave reads 16 distinct
// consecutive words of A (conflict-free) and 4 broadcast words of B.
// Per word pair the lane issues 2 v_and_b32 + 2 v_bcnt_u32_b32 (the bcnt
// accumulates), the minimum for 64-bit words on CDNA.  Counts are uint32, so
// this engine also takes the jobs beyond the MFMA engines' exact range
// (N >= 2^24, rows wider than 2^18 words), with the wide-N epilogue forms.
#include "qld_internal.h"
#include "ld_epilogue.h"
#include "ld_tiles.h"

namespace {

constexpr int PT = 128;       // tile edge (pairs)
constexpr int PKW = 16;       // words per k stage
constexpr int PLD = PT + 1;   // LDS row pitch in words (bank spread)
constexpr int PTHREADS = 256;
constexpr int PLOADS = PT * PKW / PTHREADS;   // words per lane per table

struct PopArgs {
    LdJob job;
    EpiConst ec;
    TileMap map;
};

template <int MODE>
__global__ __launch_bounds__(PTHREADS) void ld_popcount_kernel(PopArgs p)
{
    __shared__ uint64_t sA[PKW][PLD];   // table A rows -> columns j
    __shared__ uint64_t sB[PKW][PLD];   // table B rows -> rows i

    const LdJob &J = p.job;
    int ti, tj;
    tile_of_block(p.map, blockIdx.x, gridDim.x, ti, tj);
    const int i0 = ti * PT, j0 = tj * PT;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;

    // Global->register staging: lane loads word (k0 + (c & 15)) of tile row
    // c >> 4 for c = tid + 256q; 16 consecutive lanes read one 128-byte row.
    uint64_t ra[PLOADS], rb[PLOADS];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PLOADS; ++q) {
            int c = tid + PTHREADS * q;
            int row = c >> 4, k = k0 + (c & 15);
            int ja = j0 + row, ib = i0 + row;
            bool kin = k < J.words;
            ra[q] = (kin && ja < J.m) ? J.A[(size_t)ja * J.words + k] : 0ull;
            rb[q] = (kin && ib < J.n) ? J.B[(size_t)ib * J.words + k] : 0ull;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < PLOADS; ++q) {
            int c = tid + PTHREADS * q;
            sA[c & 15][c >> 4] = ra[q];
            sB[c & 15][c >> 4] = rb[q];
        }
    };

    uint32_t acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int s = 0; s < 8; ++s) acc[r][s] = 0u;

    fetch(0);
    stash();
    __syncthreads();
    for (int k0 = 0; k0 < J.words; k0 += PKW) {
        const bool more = k0 + PKW < J.words;
        if (more) fetch(k0 + PKW);
#pragma unroll 4
        for (int k = 0; k < PKW; ++k) {
            uint64_t a[8], b[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) a[s] = sA[k][tx + 16 * s];
#pragma unroll
            for (int r = 0; r < 8; ++r) b[r] = sB[k][ty + 16 * r];
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    uint64_t x = a[s] & b[r];
                    acc[r][s] += __popc((uint32_t)x);
                    acc[r][s] += __popc((uint32_t)(x >> 32));
                }
        }
        __syncthreads();
        if (more) {
            stash();
            __syncthreads();
        }
    }

    // Fused epilogue.
    if (MODE == QLD_R2_REFERENCE) {
        RefTerm ta[8], tb[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            int j = j0 + tx + 16 * s;
            ta[s] = ref_term(j < J.m ? J.bcA[j] : 0u, p.ec.fN);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int i = i0 + ty + 16 * r;
            tb[r] = ref_term(i < J.n ? J.bcB[i] : 0u, p.ec.fN);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int i = i0 + ty + 16 * r;
            if (i >= J.n) continue;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                int j = j0 + tx + 16 * s;
                if (j >= J.m || (J.tri && j <= i)) continue;
                if (J.res)
                    J.res[(size_t)i * J.ldr + j] =
                        p.ec.wide ? r2_reference_wide(acc[r][s], ta[s].p, tb[r].p, p.ec.fN)
                                  : r2_reference((float)acc[r][s], ta[s], tb[r], p.ec.invN);
                if (J.cnt) J.cnt[(size_t)i * J.ldc + j] = acc[r][s];
            }
        }
    } else {
        ExactTerm ta[8], tb[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            int j = j0 + tx + 16 * s;
            ta[s] = exact_term(j < J.m ? J.bcA[j] : 0u, p.ec.dN);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int i = i0 + ty + 16 * r;
            tb[r] = exact_term(i < J.n ? J.bcB[i] : 0u, p.ec.dN);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int i = i0 + ty + 16 * r;
            if (i >= J.n) continue;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                int j = j0 + tx + 16 * s;
                if (j >= J.m || (J.tri && j <= i)) continue;
                if (J.res)
                    J.res[(size_t)i * J.ldr + j] =
                        p.ec.wide ? r2_exact_wide(acc[r][s], ta[s], tb[r], p.ec.N)
                                  : r2_exact((float)acc[r][s], ta[s], tb[r], p.ec.dN);
                if (J.cnt) J.cnt[(size_t)i * J.ldc + j] = acc[r][s];
            }
        }
    }
}

}  // namespace

hipError_t launch_popcount(const LdJob &job, int r2_mode, hipStream_t s)
{
    PopArgs p;
    p.job = job;
    p.ec = make_epi_const(job.snp_size);
    if (job.words > QLD_MAX_WORDS) p.ec.wide = 1;   // counts may pass 2^24 too
    p.map = make_tile_map(job.n, job.m, PT, PT, job.tri);
    if (p.map.total == 0) return hipSuccess;
    dim3 grid((unsigned)p.map.total), block(PTHREADS);
    if (r2_mode == QLD_R2_EXACT)
        hipLaunchKernelGGL(ld_popcount_kernel<QLD_R2_EXACT>, grid, block, 0, s, p);
    else
        hipLaunchKernelGGL(ld_popcount_kernel<QLD_R2_REFERENCE>, grid, block, 0, s, p);
    return hipGetLastError();
}
