// Internal declarations shared by the HIP translation units of libqld_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/qld_gpu.h"

// Largest sample count the engine accepts: counts must stay exact in the
// fp32 accumulators of the MFMA engines and c/N must round like the
// reference's float division (see ld_epilogue.h).
#define QLD_MAX_SNP_SIZE (1 << 24)
// Largest row width in packed words: a count is at most 64 * words, and
// fp32 sums hold every integer up to 2^24 exactly.
#define QLD_MAX_WORDS (1 << 18)

// Tile edge an expanded-operand launch may use: its buffer descriptors span
// a tile's rows and address them with 32-bit offsets, so tile * row_bytes
// is kept at or below 2^31 (128-tiles always are: row_bytes <= 64 *
// QLD_MAX_WORDS = 2^24).
static inline int expanded_tile_cap(int64_t row_bytes, int tile)
{
    return (int64_t)tile * row_bytes <= ((int64_t)1 << 31) ? tile : 128;
}

// One all-pairs job over device buffers.  Row i indexes table B, column j
// indexes table A (the reference's results[i*m + j]).
struct LdJob {
    const uint64_t *A;
    const uint32_t *bcA;
    int m;
    const uint64_t *B;
    const uint32_t *bcB;
    int n;
    int words;
    int snp_size;
    int tri;
    float *res;
    int64_t ldr;
    uint32_t *cnt;
    int64_t ldc;
};

// Per-call epilogue constants.
struct EpiConst {
    float fN;       // (float)snp_size, as the reference divides by it
    double dN;      // (double)snp_size
    double invN;    // 1.0 / snp_size, correctly rounded
};

static inline EpiConst make_epi_const(int snp_size)
{
    EpiConst e;
    e.fN = (float)snp_size;
    e.dN = (double)snp_size;
    e.invN = 1.0 / (double)snp_size;
    return e;
}

// Launchers (return hipError_t of the launch).
hipError_t launch_popcount(const LdJob &job, int r2_mode, hipStream_t s);
