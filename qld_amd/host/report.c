/*
 * Report writer and the drop-in correlate call.
 *
 * The text is byte-identical to the reference's writeResultsHeader /
 * writeResults (correlate_IO.c:378-429) with its fast float printer
 * (fast_print.c:35-103); lines are assembled in a local buffer and written
 * with one fwrite per ~1 MiB instead of one stdio call per field.
 */
#include "qld_host.h"

#include <stdlib.h>
#include <string.h>

/* The float printer: the integer part and the decimals (val * 10^precision
 * computed in float, truncated) are each emitted from a digit-reversed copy,
 * so trailing zeros of the decimals are kept and a zero fraction prints
 * nothing after the point. */
static unsigned digits_reversed(long *num)
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

int qld_format_float(char *buf, float val, int precision)
{
    int o = 0;
    unsigned long power10 = 1;
    for (int j = 0; j < precision; j++) power10 *= 10;
    long frac = (long)(val * (float)power10);
    long ipart = (long)val;
    frac = (long)((unsigned long)frac - (unsigned long)ipart * power10);
    unsigned i_digits = digits_reversed(&ipart);
    unsigned d_digits = digits_reversed(&frac);
    if (ipart == 0) {
        buf[o++] = '0';
    } else {
        unsigned emitted = 0;
        while (ipart != 0) {
            buf[o++] = (char)('0' + ipart % 10);
            ipart /= 10;
            emitted++;
        }
        for (unsigned j = emitted; j < i_digits; j++) buf[o++] = '0';
    }
    buf[o++] = '.';
    if (frac != 0) {
        for (unsigned j = d_digits; j < (unsigned)precision; j++) buf[o++] = '0';
        for (unsigned j = 0; j < d_digits; j++) {
            buf[o++] = (char)('0' + frac % 10);
            frac /= 10;
        }
    }
    buf[o] = '\0';
    return o;
}

void qld_write_results_header(FILE *fp)
{
    fputs("POS2\tPOS1\tID2\tID1\tfreq[2]\tfreq[1]\tLD\n", fp);
}

static inline char *put_u32(char *p, uint32_t v)
{
    char tmp[12];
    int k = 0;
    do {
        tmp[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) *p++ = tmp[--k];
    return p;
}

static inline char *put_str(char *p, const char *s)
{
    size_t l = strlen(s);
    memcpy(p, s, l);
    return p + l;
}

enum { OUTBUF = 1 << 20, LINE_MAX_BYTES = 2 * 12 + 2 * 3 * 32 + 16 };

long qld_write_results(FILE *fp,
                       const uint32_t *posA, char **idA, const uint32_t *bcA, int m,
                       const uint32_t *posB, char **idB, const uint32_t *bcB, int n,
                       int snp_size, const float *results, float r2limit,
                       int posWset2)
{
    char *buf = (char *)malloc(OUTBUF);
    if (!buf) return -1;
    char *p = buf;
    long lines = 0;
    /* Frequencies of table A columns are reused for every row. */
    char (*fa)[32] = (char (*)[32])malloc((size_t)(m > 0 ? m : 1) * 32);
    if (!fa) {
        free(buf);
        return -1;
    }
    for (int j = 0; j < m; j++)
        qld_format_float(fa[j], 1.0f * (float)bcA[j] / (float)snp_size, 9);
    for (int i = 0; i < n; i++) {
        int j0 = posWset2 == 0 ? i + 1 : 0;
        char fb[32];
        int have_fb = 0;
        const char *ib = idB ? idB[i] : "";
        for (int j = j0; j < m; j++) {
            float r = results[(size_t)i * m + j];
            if (!(r >= r2limit && r < 100)) continue;
            if (!have_fb) {
                qld_format_float(fb, 1.0f * (float)bcB[i] / (float)snp_size, 9);
                have_fb = 1;
            }
            const char *ia = idA ? idA[j] : "";
            size_t need = LINE_MAX_BYTES + strlen(ib) + strlen(ia);
            if ((size_t)(p - buf) + need > OUTBUF) {
                fwrite(buf, 1, (size_t)(p - buf), fp);
                p = buf;
            }
            p = put_u32(p, posB[i]);
            *p++ = '\t';
            p = put_u32(p, posA[j]);
            *p++ = '\t';
            p = put_str(p, ib);
            *p++ = '\t';
            p = put_str(p, ia);
            *p++ = '\t';
            p = put_str(p, fb);
            *p++ = '\t';
            p = put_str(p, fa[j]);
            *p++ = '\t';
            p += qld_format_float(p, r, 9);
            *p++ = '\n';
            lines++;
        }
    }
    if (p != buf) fwrite(buf, 1, (size_t)(p - buf), fp);
    free(fa);
    free(buf);
    return lines;
}

int qld_correlate_report(qld_ctx *ctx, FILE *fp,
                         const uint64_t *tableA, const uint32_t *posA, char **idA,
                         const uint32_t *bcA, int m,
                         const uint64_t *tableB, const uint32_t *posB, char **idB,
                         const uint32_t *bcB, int n,
                         int compressed_snp_size, int snp_size, int posWset2,
                         float r2limit)
{
    if (!ctx || !fp || m < 0 || n < 0) return QLD_ERR_INVALID;
    size_t cells = (size_t)m * (size_t)n;
    float *results = (float *)malloc((cells ? cells : 1) * sizeof(float));
    if (!results) return QLD_ERR_NOMEM;
    const int tri = (posWset2 == 0) && tableA == tableB && m == n;
    int rc = QLD_OK;
    if (cells)
        rc = qld_gpu_correlate(ctx, tableA, bcA, m, tableB, bcB, n,
                               compressed_snp_size, snp_size, tri, results, NULL);
    if (rc == QLD_OK) {
        qld_write_results_header(fp);
        if (qld_write_results(fp, posA, idA, bcA, m, posB, idB, bcB, n, snp_size,
                              results, r2limit, posWset2) < 0)
            rc = QLD_ERR_NOMEM;
    }
    free(results);
    return rc;
}
