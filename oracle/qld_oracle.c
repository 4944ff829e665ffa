/*
 * CPU oracle for the qLD LD hot path -- TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker / CPU baseline; never by the product libraries.
 *
 * Built with -ffp-contract=off so that every float/double operation below is
 * rounded exactly where the C source says it is.
 */
#include "qld_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* compressSnip_x64, snip_processing.c:783-815.  The reference shifts each new
 * bit into the low end of an accumulator and stores it every 64 bits, so the
 * first sample lands in bit 63; the last partial word is shifted left by the
 * number of missing bits.  (When n is a multiple of 64 the reference also
 * stores one extra zero word past the row, :814; that write lands in the
 * next row and is overwritten later, so it is not restated.) */
uint32_t qo_compress_snip_x64(const char *chars, int n, uint64_t *row)
{
    uint64_t acc = 0;
    int nbits = 0, w = 0;
    uint32_t bc = 0;
    for (int s = 0; s < n; s++) {
        uint64_t bit = (uint64_t)((unsigned char)chars[s] & 1u);
        acc = (acc << 1) | bit;
        bc += (uint32_t)bit;
        if (++nbits == 64) {
            row[w++] = acc;
            acc = 0;
            nbits = 0;
        }
    }
    if (nbits != 0)
        row[w] = acc << (64 - nbits);
    return bc;
}

void qo_counts_naive(const uint64_t *A, int m, const uint64_t *B, int n,
                     int words, uint32_t *out)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) {
            const uint64_t *a = A + (size_t)j * words;
            const uint64_t *b = B + (size_t)i * words;
            uint32_t c = 0;
            for (int k = 0; k < words; k++)
                c += (uint32_t)__builtin_popcountll(a[k] & b[k]);
            out[(size_t)i * m + j] = c;
        }
}

/* Blocking of the binary_gemm BLIS configuration (bli_kernel.h:84-89):
 * KC = 256 words per rank-k pass; macro tiles of MC x NC rows. */
enum { QO_KC = 256, QO_MC = 72, QO_NC = 240, QO_MR = 4, QO_NR = 4 };

/* One MR x NR micro-tile over kc words: the bin_gemm_2x2 k-loop
 * (bin_gemm.c:31-110) widened to 4x4 so each loaded word is reused 4 times. */
static void micro_4x4(const uint64_t *const *a, const uint64_t *const *b,
                      int kc, uint32_t acc[QO_NR][QO_MR])
{
    uint64_t t[QO_NR][QO_MR];
    memset(t, 0, sizeof t);
    for (int k = 0; k < kc; k++) {
        uint64_t a0 = a[0][k], a1 = a[1][k], a2 = a[2][k], a3 = a[3][k];
        for (int q = 0; q < QO_NR; q++) {
            uint64_t bq = b[q][k];
            t[q][0] += (uint64_t)__builtin_popcountll(a0 & bq);
            t[q][1] += (uint64_t)__builtin_popcountll(a1 & bq);
            t[q][2] += (uint64_t)__builtin_popcountll(a2 & bq);
            t[q][3] += (uint64_t)__builtin_popcountll(a3 & bq);
        }
    }
    for (int q = 0; q < QO_NR; q++)
        for (int p = 0; p < QO_MR; p++)
            acc[q][p] += (uint32_t)t[q][p];
}

/* The same micro-tile with 512-bit AND + VPOPCNTQ: eight words of a row per
 * instruction, the 16 pair counts in 16 vector accumulators (the hosts the
 * CPU baseline runs on have AVX-512 VPOPCNTDQ; others take micro_4x4). */
__attribute__((target("avx512f,avx512vpopcntdq")))
static void micro_4x4_vpop(const uint64_t *const *a, const uint64_t *const *b,
                           int kc, uint32_t acc[QO_NR][QO_MR])
{
    __m512i t[QO_NR][QO_MR];
    for (int q = 0; q < QO_NR; q++)
        for (int p = 0; p < QO_MR; p++)
            t[q][p] = _mm512_setzero_si512();
    int k = 0;
    for (; k < kc; k += 8) {
        const __mmask8 msk = kc - k >= 8 ? (__mmask8)0xFF : (__mmask8)((1u << (kc - k)) - 1u);
        __m512i av[QO_MR];
        for (int p = 0; p < QO_MR; p++) av[p] = _mm512_maskz_loadu_epi64(msk, a[p] + k);
        for (int q = 0; q < QO_NR; q++) {
            const __m512i bq = _mm512_maskz_loadu_epi64(msk, b[q] + k);
            for (int p = 0; p < QO_MR; p++)
                t[q][p] = _mm512_add_epi64(t[q][p],
                                           _mm512_popcnt_epi64(_mm512_and_si512(av[p], bq)));
        }
    }
    for (int q = 0; q < QO_NR; q++)
        for (int p = 0; p < QO_MR; p++)
            acc[q][p] += (uint32_t)_mm512_reduce_add_epi64(t[q][p]);
}

static int qo_has_vpop(void)
{
    static int has = -1;
    if (has < 0)
        has = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vpopcntdq");
    return has;
}

int qo_counts_blocked_isa(void) { return qo_has_vpop() ? 512 : 64; }

static const uint64_t qo_zero_row[QO_KC];

void qo_counts_blocked(const uint64_t *A, int m, const uint64_t *B, int n,
                       int words, int tri, uint32_t *out, int nthreads)
{
    int mt = (m + QO_MC - 1) / QO_MC, nt = (n + QO_NC - 1) / QO_NC;
    long ntiles = (long)mt * nt;
    const int vpop = qo_has_vpop();
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (long t = 0; t < ntiles; t++) {
        int ib = (int)(t / mt), jb = (int)(t % mt);
        int i0 = ib * QO_NC, i1 = i0 + QO_NC < n ? i0 + QO_NC : n;
        int j0 = jb * QO_MC, j1 = j0 + QO_MC < m ? j0 + QO_MC : m;
        if (tri && j1 <= i0) continue;          /* tile entirely above j >= i */
        for (int i = i0; i < i1; i += QO_NR)
            for (int j = j0; j < j1; j += QO_MR) {
                if (tri && j + QO_MR <= i) continue;
                uint32_t acc[QO_NR][QO_MR];
                memset(acc, 0, sizeof acc);
                for (int pc = 0; pc < words; pc += QO_KC) {
                    int kc = words - pc < QO_KC ? words - pc : QO_KC;
                    const uint64_t *ap[QO_MR], *bp[QO_NR];
                    for (int p = 0; p < QO_MR; p++)
                        ap[p] = j + p < m ? A + (size_t)(j + p) * words + pc
                                          : qo_zero_row;
                    for (int q = 0; q < QO_NR; q++)
                        bp[q] = i + q < n ? B + (size_t)(i + q) * words + pc
                                          : qo_zero_row;
                    if (vpop) micro_4x4_vpop(ap, bp, kc, acc);
                    else micro_4x4(ap, bp, kc, acc);
                }
                for (int q = 0; q < QO_NR && i + q < n; q++)
                    for (int p = 0; p < QO_MR && j + p < m; p++)
                        if (!tri || j + p >= i + q)
                            out[(size_t)(i + q) * m + (j + p)] = acc[q][p];
            }
    }
    (void)nthreads;
}

/* get_pairwise_ld_score, gemm_cpu.c:353-392. */
void qo_r2_reference(const uint32_t *bcA, const uint32_t *bcB,
                     const uint32_t *counts, int m, int n, int snp_size,
                     int tri, float *results)
{
    const float N = (float)snp_size;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < n; i++)
        for (int j = tri ? i + 1 : 0; j < m; j++) {
            size_t at = (size_t)i * m + j;
            float r = 0.0f;
            if (bcB[i] != 0 && bcA[j] != 0) {
                float p1 = (float)bcA[j] / N;
                float p2 = (float)bcB[i] / N;
                float p12 = (float)counts[at] / N;
                float x = p1 * p2;
                float t = p12 - x;
                r = t * t;
                double den = ((double)x * (1.0 - (double)p1)) * (1.0 - (double)p2);
                r = (float)((double)r / den);
                if ((double)r > 1.0001)
                    r = 123.456f;
            }
            results[at] = r;
        }
}

void qo_r2_exact(const uint32_t *bcA, const uint32_t *bcB,
                 const uint32_t *counts, int m, int n, int snp_size,
                 int tri, float *results)
{
    const int64_t N = snp_size;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < n; i++)
        for (int j = tri ? i + 1 : 0; j < m; j++) {
            size_t at = (size_t)i * m + j;
            int64_t a = bcA[j], b = bcB[i], c = counts[at];
            float r = 0.0f;
            if (a != 0 && b != 0) {
                double d = (double)(c * N - a * b);
                double ra = 1.0 / (double)(a * (N - a));
                double rb = 1.0 / (double)(b * (N - b));
                r = (float)(((d * d) * ra) * rb);
                if ((double)r > 1.0001)
                    r = 123.456f;
            }
            results[at] = r;
        }
}

void qo_r2_fp64(const uint32_t *bcA, const uint32_t *bcB,
                const uint32_t *counts, int m, int n, int snp_size,
                int tri, double *out)
{
    const double N = (double)snp_size;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < n; i++)
        for (int j = tri ? i + 1 : 0; j < m; j++) {
            size_t at = (size_t)i * m + j;
            double a = bcA[j], b = bcB[i], c = counts[at];
            double r = 0.0;
            if (a != 0 && b != 0) {
                double d = c * N - a * b;
                r = (d * d) / ((a * (N - a)) * (b * (N - b)));
            }
            out[at] = r;
        }
}

/* fprint_f, fast_print.c:35-103.  The integer part and the 9 decimals are
 * each emitted from a digit-reversed copy, which is why trailing zeros of the
 * fraction survive and an all-zero fraction prints nothing after '.'. */
static unsigned qo_reverse(long *num)
{
    unsigned digits = 0;
    long rev = 0;
    while (*num > 0) {
        digits++;
        rev = rev * 10 + (*num) % 10;
        *num /= 10;
    }
    *num = rev;
    return digits;
}

int qo_fprint_f(char *buf, float val, int precision)
{
    static const char digit[] = "0123456789";
    int o = 0;
    unsigned long power10 = 1;
    for (int j = 0; j < precision; j++)
        power10 *= 10;
    float scaled = val * (float)power10;   /* float * unsigned long -> float */
    long long_val = (long)scaled;
    long int_part = (long)val;
    long_val = (long)((unsigned long)long_val - (unsigned long)int_part * power10);
    unsigned i_digits = qo_reverse(&int_part);
    unsigned d_digits = qo_reverse(&long_val);
    if (int_part == 0) {
        buf[o++] = '0';
    } else {
        unsigned i_numbers = 0;
        while (int_part != 0) {
            buf[o++] = digit[int_part % 10];
            int_part /= 10;
            i_numbers++;
        }
        for (unsigned j = 0; j < i_digits - i_numbers; j++)
            buf[o++] = '0';
    }
    buf[o++] = '.';
    if (long_val != 0) {
        /* precision - d_digits is unsigned in the reference; it cannot wrap
         * for the values it prints (r^2 < 100, frequencies <= 1). */
        for (unsigned j = 0; j + d_digits < (unsigned)precision; j++)
            buf[o++] = '0';
        for (unsigned j = 0; j < d_digits; j++) {
            buf[o++] = digit[long_val % 10];
            long_val /= 10;
        }
    }
    buf[o] = '\0';
    return o;
}

/* writeResultsHeader + writeResults, correlate_IO.c:378-429. */
void qo_write_results(FILE *fp, const uint32_t *posA, char **idA,
                      const uint32_t *bcA, int m, const uint32_t *posB,
                      char **idB, const uint32_t *bcB, int n, int snp_size,
                      const float *results, float r2limit, int posWset2,
                      int with_header)
{
    char f1[64], f2[64], f3[64];
    if (with_header)
        fputs("POS2\tPOS1\tID2\tID1\tfreq[2]\tfreq[1]\tLD\n", fp);
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        if (posWset2 == 0) cnt++;
        for (int j = cnt; j < m; j++) {
            float r = results[(size_t)i * m + j];
            if (r >= r2limit && r < 100) {
                qo_fprint_f(f1, 1.0f * (float)bcB[i] / (float)snp_size, 9);
                qo_fprint_f(f2, 1.0f * (float)bcA[j] / (float)snp_size, 9);
                qo_fprint_f(f3, r, 9);
                fprintf(fp, "%u\t%u\t%s\t%s\t%s\t%s\t%s\n", posB[i], posA[j],
                        idB ? idB[i] : "", idA ? idA[j] : "", f1, f2, f3);
            }
        }
    }
}
