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
 * reference's uint32 split.  The pairs to print are selected on the GPU
 * block by block (qld_correlate_hits); only they cross PCIe, and the text of
 * one block is formatted while the next computes.  Returns QLD_OK or a
 * negative status; never exits. */
int qld_correlate_report(qld_ctx *ctx, FILE *fp,
                         const uint64_t *tableA, const uint32_t *posA, char **idA,
                         const uint32_t *bcA, int m,
                         const uint64_t *tableB, const uint32_t *posB, char **idB,
                         const uint32_t *bcB, int n,
                         int compressed_snp_size, int snp_size, int posWset2,
                         float r2limit);

/* Report lines of selected pairs (the hits qld_correlate_hits delivers or
 * qld_compact_dev writes): rows i = row0 .. row0+rows-1 of table B, row
 * i's hits at entries offsets[i-row0] .. offsets[i-row0+1]-1 of cols
 * (table-A column) and r2.  Text as writeResults (correlate_IO.c:411-426),
 * formatted on all OpenMP threads.  Returns lines written or -1. */
long qld_write_hits(FILE *fp,
                    const uint32_t *posA, char **idA, const uint32_t *bcA, int m,
                    const uint32_t *posB, char **idB, const uint32_t *bcB, int n,
                    int snp_size, int row0, int rows, const uint64_t *offsets,
                    const uint32_t *cols, const float *r2);

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

/* ---- MDF input directories -------------------------------------------- */

#define QLD_PATH_MAX 4096

/* Message of the last failed host-library call on this thread. */
const char *qld_host_last_error(void);

/* What readHeaderFile (correlate_IO.c:79-255) takes from a directory: the
 * first regular file whose name contains "header"; its name up to the first
 * '_' is the alignment id; the line after its `#CHROM` line holds
 * `id snipsPerFile snipSize totalSnips posMin posMax` (written by
 * parse_vcf.c:620-634); `samples` counts the #CHROM line's sample columns. */
typedef struct qld_mdf_header {
    char id[256];
    char header_path[QLD_PATH_MAX];
    int snips_per_file, snip_size, total_snips, pos_min, pos_max;
    int samples;
} qld_mdf_header;

int qld_mdf_read_header(const char *dir, qld_mdf_header *h);

/* Chunk files of `id` covering the window [pos_wmin, pos_wmax], in chunk
 * order, as findFiles (correlate_IO.c:294-376) selects them from names
 * `<id>_<first>_<last>_<posMin>_<posMax>.mdf`.  *files is malloc'd; release
 * with qld_mdf_free_files.  Where the reference would rescan forever (no
 * chunk continues the index sequence) this returns QLD_ERR_NOTFOUND. */
int qld_mdf_find_files(const char *dir, const char *id, int snips_per_file,
                       int total_snips, int pos_wmin, int pos_wmax,
                       char ***files, int *nfiles);
void qld_mdf_free_files(char **files, int nfiles);

/* The same selection for chunks whose names contain `suffix` (".mdf" for
 * MDF directories, ".vcf.gz" for the parser's VCF chunks). */
int qld_find_chunks(const char *dir, const char *id, int snips_per_file, int total_snips,
                    int pos_wmin, int pos_wmax, const char *suffix, char ***files,
                    int *nfiles);

/* A loaded table: one SNP per row, `words` uint64 per row, with the POS, ID
 * and BC columns of the MDF rows (table_x64, correlate_MEM.h:66-75). */
typedef struct qld_snp_table {
    int rows, words;
    uint64_t *snp;
    uint32_t *pos, *bc;
    char **id;
    char *idbuf;
} qld_snp_table;

/* Load the rows of `files` as readTable_x64 does for MDF input
 * (correlate_IOx64.c:27-122,351-461): per file, two header lines are
 * skipped, then lines are read while pos < pos_wmax (only the first line
 * when pos_wset == 0) and every line whose POS lies in [pos_wmin, pos_wmax]
 * becomes a row; the `#CHROM` line itself is such a line with POS 0.  Each
 * row must carry exactly `words` packed words (QLD_ERR_INVALID otherwise).
 * Files are read and rows parsed on `nthreads` threads. */
int qld_mdf_read_table(char *const *files, int nfiles, int pos_wmin, int pos_wmax,
                       int pos_wset, int words, int nthreads, qld_snp_table *t);
void qld_snp_table_free(qld_snp_table *t);

/* A reader that caches the chunk files it has read: the text, the line
 * index (each data line's POS) and the rows parsed so far, so that many
 * windows over the same chunks (the reference's -inputList tasks) read and
 * parse each file and row once.  qld_mdf_reader_table has the semantics of
 * qld_mdf_read_table.  After each call the cache is trimmed to its byte
 * limit (default 4 GiB) by unloading the least recently used files that the
 * call did not read; an unloaded file is read again when needed.  Not
 * thread-safe: one caller at a time. */
typedef struct qld_mdf_reader qld_mdf_reader;
qld_mdf_reader *qld_mdf_reader_create(void);
int qld_mdf_reader_table(qld_mdf_reader *r, char *const *files, int nfiles, int pos_wmin,
                         int pos_wmax, int pos_wset, int words, int nthreads,
                         qld_snp_table *t);
/* Cache limit in bytes (0: keep only the files of the latest call). */
void qld_mdf_reader_set_limit(qld_mdf_reader *r, size_t bytes);
/* Bytes the cache holds now (file text, line index and parsed rows). */
size_t qld_mdf_reader_bytes(const qld_mdf_reader *r);
void qld_mdf_reader_free(qld_mdf_reader *r);

/* ---- VCF chunk input ------------------------------------------------------ */

/* Rows from the parser's gzip VCF chunks, as the reference correlator builds
 * them without -mdf (writeToTable_x64, correlate_IOx64.c:125-349, and
 * snip_processing.c): single-base REF/ALT, FILTER PASS or ".", INFO AF not
 * 0 or 1 and VT (if present) SNP, GT in FORMAT; the selected samples' GT
 * alleles as bases (unphased alleles shuffled by rand()), monomorphic sites
 * and sites with N before their second base dropped, the two bases coded
 * '0'/'1' in order of appearance and packed MSB-first; BC = bit count.
 * The reader owns the rand() stream (glibc's default sequence, which the
 * reference draws from unseeded): use one reader per run and read tables
 * in the reference's order (task by task, table A then B). */
typedef struct qld_vcf_reader qld_vcf_reader;
qld_vcf_reader *qld_vcf_reader_create(void);
void qld_vcf_reader_free(qld_vcf_reader *r);
/* Restart the reader's stream as srand(seed) restarts rand() (the
 * reference's -seed, correlate_our_kernel.c:387-392). */
void qld_vcf_reader_seed(qld_vcf_reader *r, unsigned int seed);
const char *qld_vcf_last_error(void);

/* First line and last #CHROM line of a header file (malloc'd); each chunk
 * must start with exactly these two lines. */
int qld_vcf_header_lines(const char *header_path, char **line1, char **chrom_line);

/* Samples of the #CHROM line selected by a list file (whitespace-separated
 * names; NULL = all): mask[s] per sample column (malloc'd), *valid_count =
 * selected samples x ploidy (makeValidList, correlate_IO.c:259-290). */
int qld_vcf_sample_mask(const char *chrom_line, const char *list_file, int ploidy,
                        unsigned char **mask, int *nsamples, int *valid_count);

/* Rows of `files` in [pos_wmin, pos_wmax] with the reference's line loop
 * (as qld_mdf_read_table); `snip_size` is the header's sample size. */
int qld_vcf_read_table(qld_vcf_reader *r, char *const *files, int nfiles, int pos_wmin,
                       int pos_wmax, int pos_wset, const char *line1, const char *chrom_line,
                       const unsigned char *mask, int nsamples, int valid_count, int ploidy,
                       int snip_size, qld_snp_table *t);

/* Write one MDF chunk file exactly as the MDF tool does
 * (process_data.c:447-460): `##fileformat=MDFv0.1`, `##SNPs:<rows>`, the
 * `#CHROM POS ID BC compSNP_k...` line, then one tab-separated row per SNP,
 * no trailing newline.  id may be NULL ("."). */
int qld_mdf_write(const char *path, const char *chrom, const uint64_t *snp,
                  const uint32_t *pos, char *const *id, const uint32_t *bc,
                  int rows, int words);

#ifdef __cplusplus
}
#endif
#endif
