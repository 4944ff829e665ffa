// Internal declarations shared by the HIP translation units of libqld_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/qld_gpu.h"

// Largest sample count the MFMA engines accept: counts must stay exact in
// their fp32 accumulators and c/N must round like the reference's float
// division (see ld_epilogue.h).
#define QLD_MAX_SNP_SIZE (1 << 24)
// Largest row width of the MFMA engines in packed words: a count is at most
// 64 * words, and fp32 sums hold every integer up to 2^24 exactly.
#define QLD_MAX_WORDS (1 << 18)
// The popcount engine counts in uint32 and takes any sample count an int
// holds (the reference's snp_size is an int) and rows of up to 2^25 words
// (counts up to 2^31); jobs beyond the MFMA bounds ("wide" jobs) run there.
#define QLD_MAX_WORDS_POPCOUNT (1 << 25)

static inline bool qld_wide_job(int words, int snp_size)
{
    return words > QLD_MAX_WORDS || snp_size >= QLD_MAX_SNP_SIZE;
}

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
    int64_t N;      // snp_size
    int wide;       // snp_size >= 2^24: the popcount engine's wide-N forms
};

static inline EpiConst make_epi_const(int snp_size)
{
    EpiConst e;
    e.fN = (float)snp_size;
    e.dN = (double)snp_size;
    e.invN = 1.0 / (double)snp_size;
    e.N = snp_size;
    e.wide = snp_size >= QLD_MAX_SNP_SIZE;
    return e;
}

// Launchers (return hipError_t of the launch).
hipError_t launch_popcount(const LdJob &job, int r2_mode, hipStream_t s);
