/*
 * qld_host.h -- C host side of the MI355X LD engine (libqld_host.so).
 *
 * Plain C on top of qld_gpu.h: the report writer that keeps qLD-compute's
 * per-region output layout, the drop-in correlate call used where the
 * reference calls correlate_gpu(), and the seeded synthetic haplotype
 * generator used by tests and benchmarks.
 */
#ifndef QLD_HOST_H
#define QLD_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "qld_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- report text ---------------------------------------------------- */

/* Format `val` with `precision` decimals the way the reference's fast float
 * printer does (fast_print.c:35-103); returns the length written to buf
 * (buf needs >= 32 bytes for the values the report prints). */
int qld_format_float(char *buf, float val, int precision);

/* Write the report header line (correlate_IO.c:378-381). */
void qld_write_results_header(FILE *fp);

/* Write the report body exactly as writeResults (correlate_IO.c:384-429):
 * rows i of table B, columns j of table A (j > i when posWset2 == 0), one
 * line per r^2 >= r2limit and < 100.  Returns lines written.  ids may be
 * NULL (empty ID columns). */
long qld_write_results(FILE *fp,
                       const uint32_t *posA, char **idA, const uint32_t *bcA, int m,
                       const uint32_t *posB, char **idB, const uint32_t *bcB, int n,
                       int snp_size, const float *results, float r2limit,
                       int posWset2);

/* Drop-in for correlate_gpu() (gemm_gpu.h:56-70, gemm_gpu.c:753-906 of the
 * reference): compute r^2 on the GPU for table A (m rows) against table B
 * (n rows; same pointer and posWset2 == 0 for a single region) and write
 * header + report to fp.  Uses uint64 rows directly (the MDF word), not the
 * reference's uint32 split.  Returns QLD_OK or a negative status; never
 * exits. */
int qld_correlate_report(qld_ctx *ctx, FILE *fp,
                         const uint64_t *tableA, const uint32_t *posA, char **idA,
                         const uint32_t *bcA, int m,
                         const uint64_t *tableB, const uint32_t *posB, char **idB,
                         const uint32_t *bcB, int n,
                         int compressed_snp_size, int snp_size, int posWset2,
                         float r2limit);

/* ---- synthetic haplotype data ---------------------------------------- */

/* Words per packed row: ceil(n_samples / 64). */
int qld_words(int n_samples);

/* Haplotype-block Markov generator.  SNP s starts a new block with
 * probability 1/mean_block (s == 0 always does); each SNP draws its minor
 * allele frequency p_s uniformly from [maf_lo, maf_hi]; the first SNP of a
 * block is Bernoulli(p_s) per sample; inside a block every sample bit copies
 * the previous SNP's bit with probability copy_prob, else it is a fresh
 * Bernoulli(p_s) draw.  Rows are packed MSB-first with the last word's low
 * bits zero; bc[s] = popcount of row s.  Deterministic in (seed, s, word);
 * independent of thread count. */
void qld_synth_blocks(uint64_t seed, int n_snps, int n_samples,
                      double mean_block, double maf_lo, double maf_hi,
                      double copy_prob, uint64_t *table, uint32_t *bc,
                      int nthreads);

/* Rectangular two-region generator: MAF ~ Beta(0.5, 2) clipped to
 * [maf_lo, maf_hi]; each region has its own independent block chain; a
 * fraction `copy_frac` of region-B SNPs are replaced by a copy of a random
 * region-A SNP with each sample bit flipped with probability flip_prob. */
void qld_synth_regions(uint64_t seed, int m_a, int n_b, int n_samples,
                       double mean_block, double maf_lo, double maf_hi,
                       double copy_prob, double copy_frac, double flip_prob,
                       uint64_t *tableA, uint32_t *bcA,
                       uint64_t *tableB, uint32_t *bcB, int nthreads);

/* Recompute bit counts of `rows` packed rows. */
void qld_bitcounts(const uint64_t *table, int rows, int words, uint32_t *bc,
                   int nthreads);

#ifdef __cplusplus
}
#endif
#endif
