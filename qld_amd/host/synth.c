/*
 * Seeded synthetic haplotype tables (haplotype-block Markov model).
 *
 * Every random draw is a counter-based hash of (seed, SNP, word, draw), so
 * the output does not depend on thread count or order.  A Bernoulli(q)
 * 64-bit word is built from 16 uniform words by the binary expansion of q
 * (q quantised to 1/65536): processing the bits of q from the least
 * significant set bit up, w = bit ? (w | r) : (w & r).
 *
 * Packing follows the reference's table layout (snip_processing.c:783-815):
 * sample s at bit 63 - (s % 64) of word s / 64, low bits of the last word 0.
 */
#include "qld_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

static inline uint64_t rnd(uint64_t key, uint64_t ctr)
{
    return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ull);
}

static inline double unif(uint64_t r)          /* (0, 1) */
{
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static inline uint64_t snp_key(uint64_t seed, uint64_t s)
{
    return mix64(seed ^ mix64(0xD1B54A32D192ED03ull * (s + 1)));
}

static inline uint32_t q16_of(double q)
{
    if (q <= 0) return 0;
    if (q >= 1) return 65536;
    return (uint32_t)llround(q * 65536.0);
}

static inline uint64_t bern_word(uint64_t key, uint64_t base, uint32_t q16)
{
    if (q16 == 0) return 0;
    if (q16 >= 65536) return ~0ull;
    uint64_t w = 0;
    for (int b = __builtin_ctz(q16); b < 16; b++) {
        uint64_t r = rnd(key, base + (uint64_t)b);
        w = ((q16 >> b) & 1u) ? (w | r) : (w & r);
    }
    return w;
}

int qld_words(int n_samples) { return (n_samples + 63) / 64; }

static inline uint64_t tail_mask(int n_samples)
{
    int rem = n_samples % 64;
    return rem ? ~0ull << (64 - rem) : ~0ull;
}

/* Draw streams inside one SNP key. */
enum { D_BLOCK = 0, D_MAF = 1, D_MAF2 = 2, D_MAF3 = 3, D_MAF4 = 4,
       D_COPYSEL = 5, D_COPYSRC = 6, D_WORDS = 16 };
/* Per word: draws [0,16) fresh allele, [32,48) copy mask, [48,64) flips. */
#define WBASE(w, off) ((uint64_t)D_WORDS + (uint64_t)(w) * 64u + (off))

static double beta_half_two(uint64_t key)
{
    /* Beta(0.5, 2) = G1 / (G1 + G2), G1 ~ Gamma(1/2) = Z^2 / 2,
     * G2 ~ Gamma(2) = -ln(U1 U2); Z by Box-Muller. */
    double u1 = unif(rnd(key, D_MAF)), u2 = unif(rnd(key, D_MAF2));
    double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    double g1 = 0.5 * z * z;
    double g2 = -log(unif(rnd(key, D_MAF3)) * unif(rnd(key, D_MAF4)));
    return g1 / (g1 + g2);
}

/* Fill one chain of SNPs [0, n) whose per-SNP keys derive from `seed`.
 * maf_beta != 0 draws MAF from Beta(0.5, 2) (clipped), else uniform. */
static void synth_chain(uint64_t seed, int n, int n_samples, double mean_block,
                        double lo, double hi, double copy_prob, int maf_beta,
                        uint64_t *table, int nthreads)
{
    const int W = qld_words(n_samples);
    const uint64_t last = tail_mask(n_samples);
    const double pstart = mean_block > 1.0 ? 1.0 / mean_block : 1.0;
    const uint32_t qcopy = q16_of(copy_prob);
    if (n <= 0) return;

    /* Block starts, so that blocks can be generated in parallel. */
    int *starts = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int nb = 0;
    for (int s = 0; s < n; s++)
        if (s == 0 || unif(rnd(snp_key(seed, (uint64_t)s), D_BLOCK)) < pstart)
            starts[nb++] = s;
    starts[nb] = n;

#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int b = 0; b < nb; b++) {
        for (int s = starts[b]; s < starts[b + 1]; s++) {
            uint64_t key = snp_key(seed, (uint64_t)s);
            double p = maf_beta ? beta_half_two(key)
                                : lo + (hi - lo) * unif(rnd(key, D_MAF));
            if (p < lo) p = lo;
            if (p > hi) p = hi;
            uint32_t qp = q16_of(p);
            uint64_t *row = table + (size_t)s * W;
            const uint64_t *prev = row - W;
            for (int w = 0; w < W; w++) {
                uint64_t fresh = bern_word(key, WBASE(w, 0), qp);
                uint64_t v;
                if (s == starts[b]) {
                    v = fresh;
                } else {
                    uint64_t copy = bern_word(key, WBASE(w, 32), qcopy);
                    v = (prev[w] & copy) | (fresh & ~copy);
                }
                row[w] = (w == W - 1) ? (v & last) : v;
            }
        }
    }
    (void)nthreads;
    free(starts);
}

void qld_bitcounts(const uint64_t *table, int rows, int words, uint32_t *bc,
                   int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int s = 0; s < rows; s++) {
        uint32_t c = 0;
        const uint64_t *row = table + (size_t)s * words;
        for (int w = 0; w < words; w++) c += (uint32_t)__builtin_popcountll(row[w]);
        bc[s] = c;
    }
    (void)nthreads;
}

void qld_synth_blocks(uint64_t seed, int n_snps, int n_samples,
                      double mean_block, double maf_lo, double maf_hi,
                      double copy_prob, uint64_t *table, uint32_t *bc,
                      int nthreads)
{
    synth_chain(seed, n_snps, n_samples, mean_block, maf_lo, maf_hi, copy_prob,
                0, table, nthreads);
    if (bc) qld_bitcounts(table, n_snps, qld_words(n_samples), bc, nthreads);
}

void qld_synth_regions(uint64_t seed, int m_a, int n_b, int n_samples,
                       double mean_block, double maf_lo, double maf_hi,
                       double copy_prob, double copy_frac, double flip_prob,
                       uint64_t *tableA, uint32_t *bcA,
                       uint64_t *tableB, uint32_t *bcB, int nthreads)
{
    const int W = qld_words(n_samples);
    const uint64_t last = tail_mask(n_samples);
    const uint64_t seedA = mix64(seed ^ 0xA0A0A0A0A0A0A0A0ull);
    const uint64_t seedB = mix64(seed ^ 0xB0B0B0B0B0B0B0B0ull);
    const uint32_t qflip = q16_of(flip_prob);
    synth_chain(seedA, m_a, n_samples, mean_block, maf_lo, maf_hi, copy_prob, 1,
                tableA, nthreads);
    synth_chain(seedB, n_b, n_samples, mean_block, maf_lo, maf_hi, copy_prob, 1,
                tableB, nthreads);
    if (m_a > 0) {
#ifdef _OPENMP
        if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int s = 0; s < n_b; s++) {
            uint64_t key = snp_key(seedB, (uint64_t)s);
            if (unif(rnd(key, D_COPYSEL)) >= copy_frac) continue;
            int src = (int)(unif(rnd(key, D_COPYSRC)) * m_a);
            if (src >= m_a) src = m_a - 1;
            const uint64_t *a = tableA + (size_t)src * W;
            uint64_t *row = tableB + (size_t)s * W;
            for (int w = 0; w < W; w++) {
                uint64_t v = a[w] ^ bern_word(key, WBASE(w, 48), qflip);
                row[w] = (w == W - 1) ? (v & last) : v;
            }
        }
    }
    if (bcA) qld_bitcounts(tableA, m_a, W, bcA, nthreads);
    if (bcB) qld_bitcounts(tableB, n_b, W, bcB, nthreads);
}
