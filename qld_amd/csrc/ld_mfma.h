// Interface of the matrix-core engines (ld_mfma.hip).
#pragma once

#include "qld_internal.h"

// One all-pairs job on expanded operands.
struct MfmaJob {
    LdJob job;              // shapes, bit counts, outputs
    const uint8_t *XA;      // expanded table A (m rows x row_bytes)
    const uint8_t *XB;      // expanded table B (n rows x row_bytes)
    int64_t row_bytes;
    const uint32_t *tiles;  // device tile list, (ti | tj << 16)
    int64_t ntiles;
};

// Packed words per row after padding to whole k stages.
int mfma_padded_words(int engine, int words);
// Bytes per expanded row.
int64_t mfma_row_bytes(int engine, int words);
// Tile edge (pairs) of the MFMA kernels.
int mfma_tile_edge();

// Expand packed rows to fp4 nibbles / int8 bytes (rows x row_bytes).
hipError_t launch_expand(int engine, const uint64_t *T, int rows, int words,
                         uint8_t *X, hipStream_t s);
hipError_t launch_mfma(int engine, int r2_mode, const MfmaJob &j, hipStream_t s);
