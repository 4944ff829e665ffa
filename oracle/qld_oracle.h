/*
 * CPU oracle for the qLD LD hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain-C restatement of the reference algorithm and exists
 * to check the HIP engine.  Only the test suite, the smoke check and the CPU
 * baseline leg of bench.py may load it; the product path (libqld_hip.so,
 * libqld_host.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Which check pins
 * each function is listed next to it ("pinned by: ...").
 */
#ifndef QLD_ORACLE_H
#define QLD_ORACLE_H

#include <stdint.h>
#include <stdio.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pack one SNP given as one char per sample ('0'/'1'; bit = char & 1) into
 * ceil(n/64) uint64 words, MSB-first, zero-padded low bits of the last word,
 * and return its bit count.
 * Restates compressSnip_x64 (src/correlator/src/snip_processing.c:783-815).
 * pinned by: known-answer vectors and the compiled reference (tests). */
uint32_t qo_compress_snip_x64(const char *chars, int n, uint64_t *row);

/* Pairwise AND+popcount counts, brute force over words.
 * out[i*m + j] = popcount(A_j & B_i) summed over `words` words (the layout of
 * the reference's results array, gemm_cpu.c:353-363 with C[j + i*M] from the
 * column-major dgemm/dsyrk at gemm_cpu.c:203-236).
 * pinned by: numpy int64 matmul of unpacked 0/1 matrices, golden fixtures. */
void qo_counts_naive(const uint64_t *A, int m, const uint64_t *B, int n,
                     int words, uint32_t *out);

/* Same counts with the BLIS binary_gemm blocking (bin_gemm_2x2 k-loop,
 * blis/config/binary_gemm/kernels/3/bin_gemm.c:4-141, KC=256 words per rank-k
 * pass, bli_kernel.h:84-89), a 4x4 register micro-tile and OpenMP over macro
 * tiles.  tri != 0 (A == B, m == n) computes only j >= i, as cblas_dsyrk
 * Lower does (gemm_cpu.c:226-236); other cells are left untouched.
 * This is the CPU baseline timed by bench.py.
 * pinned by: qo_counts_naive on ragged shapes (tests). */
void qo_counts_blocked(const uint64_t *A, int m, const uint64_t *B, int n,
                       int words, int tri, uint32_t *out, int nthreads);
/* Vector width in bits of the popcount micro-kernel qo_counts_blocked uses
 * on this host: 512 (AND + VPOPCNTQ on AVX-512 VPOPCNTDQ hosts) or 64
 * (scalar popcnt, as bin_gemm.c does). */
int qo_counts_blocked_isa(void);

/* r^2 exactly as get_pairwise_ld_score computes it (gemm_cpu.c:343-404;
 * the same arithmetic as get_pairwise_ld_score_gpu, gemm_gpu.c:700-737), as
 * compiled by the reference's own flags (gcc -O3 -std=gnu11, no contraction of
 * p12 - p1*p2 because p1*p2 is shared with the denominator):
 *   p1 = (float)bcA[j]/N, p2 = (float)bcB[i]/N, p12 = (float)C/N
 *   x  = p1*p2 (float), t = p12 - x (float), r = t*t (float)
 *   r  = (float)((double)r / (((double)x * (1.0 - p1)) * (1.0 - p2)))
 *   r > 1.0001 -> 123.456f;  bcA[j]==0 || bcB[i]==0 -> 0.
 * Writes results[i*m + j] for all i<n, j<m (tri == 0) or j > i (tri != 0).
 * pinned by: the compiled reference get_pairwise_ld_score (bit-exact), the
 * FP64 closed form within 1e-6, and closed-form KATs (tests). */
void qo_r2_reference(const uint32_t *bcA, const uint32_t *bcB,
                     const uint32_t *counts, int m, int n, int snp_size,
                     int tri, float *results);

/* r^2 from the integer-exact numerator d = c*N - a*b (int64):
 * ((d*d) * (1/(a(N-a)))) * (1/(b(N-b))) in double, rounded once to float;
 * same special cases as qo_r2_reference.  This is the engine's EXACT mode.
 * pinned by: FP64 closed form (<= 1 float ulp), KATs (tests). */
void qo_r2_exact(const uint32_t *bcA, const uint32_t *bcB,
                 const uint32_t *counts, int m, int n, int snp_size,
                 int tri, float *results);

/* FP64 closed form r^2 = (cN - ab)^2 / (a(N-a) b(N-b)); 0 when a or b is 0,
 * NaN when the denominator is 0 otherwise.  The accuracy yardstick. */
void qo_r2_fp64(const uint32_t *bcA, const uint32_t *bcB,
                const uint32_t *counts, int m, int n, int snp_size,
                int tri, double *out);

/* Format `val` the way fprint_f does (src/correlator/src/fast_print.c:35-103)
 * into buf (NUL-terminated); returns the length.
 * pinned by: the compiled reference fprint_f and KATs (tests). */
int qo_fprint_f(char *buf, float val, int precision);

/* Report text exactly as writeResultsHeader + writeResults
 * (src/correlator/src/correlate_IO.c:378-429).  ids may be NULL (empty IDs).
 * pinned by: the compiled reference correlate() report (golden fixtures). */
void qo_write_results(FILE *fp, const uint32_t *posA, char **idA,
                      const uint32_t *bcA, int m, const uint32_t *posB,
                      char **idB, const uint32_t *bcB, int n, int snp_size,
                      const float *results, float r2limit, int posWset2,
                      int with_header);

#ifdef __cplusplus
}
#endif
#endif
