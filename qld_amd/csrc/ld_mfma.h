// Interface of the matrix-core engines (ld_mfma.hip).
//
// Every launch of the MFMA kernel consumes a list of tile work items.  An
// item names one job descriptor and one (ti, tj) tile of that job, so a
// single launch can mix jobs: the tiles of one region (plain call), of a
// region pair (rectangle), of thousands of windows (batched call) or of
// concurrent serving requests on different resident tables.
#pragma once

#include "qld_internal.h"
#include "tile_order.h"


// One all-pairs job on expanded operands.  Row i of the job reads expanded
// row XI + i*row_bytes and bit count bcI[i]; column j reads XJ / bcJ.  Cell
// (i, j) goes to res[i*ldr + j] (and cnt[i*ldc + j]); tri != 0 keeps j > i.
struct WorkDesc {
    const uint8_t *XI;
    const uint8_t *XJ;
    const uint32_t *bcI;
    const uint32_t *bcJ;
    float *res;
    uint32_t *cnt;
    int64_t ldr;
    int64_t ldc;
    int64_t row_bytes;   // multiple of 128 bytes (whole k stages)
    int ni, nj;
    int tri;
    int snp_size;
    float fN;            // epilogue constants of snp_size
    float pad_;
    double dN;
    double invN;
    // Packed-operand launches (fp4 expanded inside the kernel), `words`
    // uint64 per row.  Packed-into-LDS launches: PI/PJ are the job's packed
    // rows, row-major.  Direct launches: PI/PJ are row-block-tiled tables
    // (see launch_tile_rows) of rowsI/rowsJ rows whose rows rowI0/rowJ0
    // onwards are the job's rows.
    const uint64_t *PI;
    const uint64_t *PJ;
    int64_t rowI0, rowJ0;
    int rowsI, rowsJ;
    int words;
    int pad2_;
};


void fill_work_desc(WorkDesc &d, const uint8_t *XI, const uint8_t *XJ,
                    const uint32_t *bcI, const uint32_t *bcJ, int ni, int nj,
                    int64_t row_bytes, int tri, int snp_size, float *res,
                    int64_t ldr, uint32_t *cnt, int64_t ldc);

// Packed-operand variant: the same job read straight from packed rows.
void set_packed_operands(WorkDesc &d, const uint64_t *PI, const uint64_t *PJ, int words);

// Direct-operand variant: rows rowI0.. of tiled table TI (rowsI rows) and
// rows rowJ0.. of tiled table TJ (rowsJ rows).
void set_tiled_operands(WorkDesc &d, const uint64_t *TI, int64_t rowI0, int rowsI,
                        const uint64_t *TJ, int64_t rowJ0, int rowsJ, int words);

// Row-block-tiled copy of a packed table, the layout the direct kernels
// read: for each block of 32 rows and each group of 4 words, the 32 rows'
// 32-byte pieces are contiguous (1 KiB), i.e. uint64 index
//   ((row/32 * G + word/4) * 32 + row%32) * 4 + word%4,   G = ceil(words/4),
// with rows past `rows` and words past `words` zero.  One wave's load of a
// 32-row MFMA block is then 1 KiB of consecutive bytes instead of 32 pieces
// 8*words bytes apart.
int64_t tiled_bytes(int rows, int words);
// zero_queue (optional): 16 device uint32 queue heads cleared by the same
// kernel, for a persistent launch queued after it (launch_mfma_items with
// queue_zeroed); rows and words must then be positive.
hipError_t launch_tile_rows(const uint64_t *P, int rows, int words, uint64_t *T,
                            hipStream_t s, uint32_t *zero_queue = nullptr);

// Packed words per row after padding to whole 128-byte k stages.
int mfma_padded_words(int engine, int words);
// Bytes per expanded row.
int64_t mfma_row_bytes(int engine, int words);

// Tile edges the kernel is built for.
#define QLD_TILE_LARGE 256
#define QLD_TILE_SMALL 128

// Write the tile list of one job (tile edge T, desc 0) into device memory
// `out` (job_tile_count entries) in append_job_tiles' order, one thread per
// tile (job_tile_index), on stream s.
hipError_t launch_job_tiles(WorkItem *out, int ni, int nj, int tri, int T, hipStream_t s);

// Expand packed rows to fp4 nibbles / int8 bytes (rows x row_bytes).
hipError_t launch_expand(int engine, const uint64_t *T, int rows, int words,
                         uint8_t *X, hipStream_t s);

// Allow every MFMA kernel variant its dynamic LDS (call once before use).
hipError_t mfma_init_attributes();

// Per-wave units of the direct fp4 kernel: `units` lists wave blocks
// (desc, bi, bj) -- rows [bi*WTM, +WTM) x columns [bj*WTN, +WTN) of the job,
// WTM x WTN = 64 x 64 for tile 128, 128 x 64 for tile 256 (wave_block_shape)
// -- and `grid` workgroups of 4 waves drain them from eight per-XCD queues
// (queue = 16 device uint32, zeroed on the stream first), each wave claiming
// its own units.  A list built by append_job_wave_units leaves out blocks
// with nothing to compute.
hipError_t launch_mfma_wave_units(int r2_mode, int tile, const WorkDesc *descs,
                                  const WorkDesc &inline_desc, const WorkItem *units,
                                  int64_t nunits, int grid, uint32_t *queue, hipStream_t s);

// Output of one window of a shared-pair plan: the window's rows
// [start, start + size) of the table and its size x size r^2 block at `out`
// (cell (i, j) of the window at out[i * size + j], j > i written).
struct WinOut {
    int32_t start, size;
    float *out;
};

// Shared-pair units: `units` lists 128 x 64 wave blocks (0, bi, bj) of one
// triangle job over a whole table (inline_desc, its r^2 pointer unused), each
// computed once; item.desc is the offset in `wlist` of the block's window
// list (a header word -- the window count in its low 16 bits, the block's
// sub-block form above them -- then that many indices into `wins`), and the
// epilogue writes the block's r^2 into every listed window that holds the
// cell.  Sub-block forms: 0 all eight 32 x 32 blocks, 1-3 all but the first
// 1-3 row blocks, QLD_SUB_HALF the first 32 columns only (the windows hold
// no cell in the others).  Tile 256 geometry, direct fp4 operands; otherwise
// as launch_mfma_wave_units.
#define QLD_SUB_HALF 4
hipError_t launch_mfma_shared_units(int r2_mode, const WorkDesc &inline_desc,
                                    const WorkItem *units, int64_t nunits,
                                    const uint32_t *wlist, const WinOut *wins, int grid,
                                    uint32_t *queue, hipStream_t s);
void wave_block_shape(int tile, int *wtm, int *wtn);

// Split-K launch (128-tiles, direct fp4 operands): each of the `nitems`
// tiles is summed by `ksplit` workgroups over equal ranges of its word
// groups; the last to finish adds the others' partial sums and writes the
// tile's r^2 (and counts).  max_groups: the most word groups of any job's
// rows (sets the partial sums' width).  kpart: splitk_part_bytes(nitems,
// ksplit) bytes of device scratch; kcnt: nitems device uint32, zero before
// the first launch and left zero by every launch.  Launches that share
// kpart or kcnt must not overlap.
#define QLD_MAX_KSPLIT 64
hipError_t launch_mfma_splitk(int r2_mode, const WorkDesc *descs, const WorkDesc &inline_desc,
                              const WorkItem *items, int64_t nitems, int ksplit, int max_groups,
                              float *kpart, uint32_t *kcnt, hipStream_t s);
size_t splitk_part_bytes(int64_t nitems, int ksplit);

// Launch the MFMA kernel over `nitems` items.
//   descs       device array of job descriptors, or NULL to use `inline_desc`
//               for every item (single-job launches need no upload)
//   persistent  0: one workgroup per item;  >0: that many workgroups pull
//               items from eight per-XCD queues (queue = 16 device uint32,
//               zeroed by this call on the stream before the kernel); queue
//               x owns items [qbeg[x], qbeg[x+1]) (qbeg NULL: equal shares)
//   packed      nonzero: fp4 engine reading PI/PJ packed rows (no expanded
//               operand), 128-tiles only
//   queue_zeroed  the queue heads were cleared on the stream already (by
//               launch_tile_rows): no reset kernel
hipError_t launch_mfma_items(int engine, int r2_mode, int tile, int packed,
                             const WorkDesc *descs, const WorkDesc &inline_desc,
                             const WorkItem *items, int64_t nitems,
                             int persistent, uint32_t *queue, const int64_t *qbeg,
                             hipStream_t s, bool queue_zeroed = false);
