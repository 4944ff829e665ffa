/*
 * Report writer and the drop-in correlate call.
 *
 * The text is byte-identical to the reference's writeResultsHeader /
 * writeResults (correlate_IO.c:378-429) with its fast float printer
 * (fast_print.c:35-103); lines are assembled in a local buffer and written
 * with one fwrite per ~1 MiB instead of one stdio call per field.
 */
#include "qld_host.h"

#include <omp.h>
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

/* ---- the drop-in correlate call ------------------------------------------ */

/* Formats the hits of row blocks delivered by qld_correlate_hits: rows are cut
 * into segments of about an equal share of the block's lines per thread (at
 * least SEG_MIN, at most SEG_MAX lines), a group of segments is formatted in
 * parallel into per-thread buffers and the buffers are written in order.  (A
 * fixed 2^18-line segment left a streamed report's blocks of about 14,000
 * lines to one thread.) */
enum { SEG_MIN = 1 << 11, SEG_MAX = 1 << 18 };

typedef struct {
    FILE *fp;
    const uint32_t *posA, *bcA, *posB, *bcB;
    char **idA, **idB;
    int snp_size;
    char (*fa)[32];
    size_t line_max;
    int nthreads;
    char **buf;
    size_t *cap, *len;
    long lines;
    int err;
} hits_writer;

static size_t max_id_len(char **ids, int n)
{
    size_t mx = 0;
    for (int i = 0; ids && i < n; i++) {
        size_t l = strlen(ids[i]);
        if (l > mx) mx = l;
    }
    return mx;
}

static char *format_row_hits(char *p, const hits_writer *w, int i, uint64_t h0, uint64_t h1,
                             const uint32_t *cols, const float *r2)
{
    if (h0 == h1) return p;
    char fb[32];
    qld_format_float(fb, 1.0f * (float)w->bcB[i] / (float)w->snp_size, 9);
    const char *ib = w->idB ? w->idB[i] : "";
    for (uint64_t h = h0; h < h1; h++) {
        const uint32_t j = cols[h];
        const char *ia = w->idA ? w->idA[j] : "";
        p = put_u32(p, w->posB[i]);
        *p++ = '\t';
        p = put_u32(p, w->posA[j]);
        *p++ = '\t';
        p = put_str(p, ib);
        *p++ = '\t';
        p = put_str(p, ia);
        *p++ = '\t';
        p = put_str(p, fb);
        *p++ = '\t';
        p = put_str(p, w->fa[j]);
        *p++ = '\t';
        p += qld_format_float(p, r2[h], 9);
        *p++ = '\n';
    }
    return p;
}

static int write_hits(void *user, int row0, int rows, const uint64_t *off,
                      const uint32_t *cols, const float *r2)
{
    hits_writer *w = (hits_writer *)user;
    if (w->err || off[rows] == off[0]) return w->err;
    uint64_t seg = (off[rows] - off[0] + (uint64_t)w->nthreads - 1) / (uint64_t)w->nthreads;
    if (seg < SEG_MIN) seg = SEG_MIN;
    if (seg > SEG_MAX) seg = SEG_MAX;
    int r = 0;
    while (r < rows) {
        /* the next group: up to nthreads segments of about `seg` hits */
        int seg_lo[64], seg_hi[64], nseg = 0;
        while (r < rows && nseg < w->nthreads) {
            int r1 = r;
            while (r1 < rows && (r1 == r || off[r1 + 1] - off[r] <= seg)) r1++;
            seg_lo[nseg] = r;
            seg_hi[nseg] = r1;
            nseg++;
            r = r1;
        }
        for (int k = 0; k < nseg; k++) {
            size_t need = (size_t)(off[seg_hi[k]] - off[seg_lo[k]]) * w->line_max + 1;
            if (need > w->cap[k]) {
                free(w->buf[k]);
                w->buf[k] = (char *)malloc(need);
                w->cap[k] = w->buf[k] ? need : 0;
                if (!w->buf[k]) return w->err = -1;
            }
        }
#pragma omp parallel for schedule(dynamic, 1) num_threads(w->nthreads)
        for (int k = 0; k < nseg; k++) {
            char *p = w->buf[k];
            for (int rr = seg_lo[k]; rr < seg_hi[k]; rr++)
                p = format_row_hits(p, w, row0 + rr, off[rr], off[rr + 1], cols, r2);
            w->len[k] = (size_t)(p - w->buf[k]);
        }
        for (int k = 0; k < nseg; k++) {
            if (w->len[k] && fwrite(w->buf[k], 1, w->len[k], w->fp) != w->len[k])
                return w->err = -1;
            w->lines += (long)(off[seg_hi[k]] - off[seg_lo[k]]);
        }
    }
    return 0;
}

static int writer_init(hits_writer *w, FILE *fp, const uint32_t *posA, char **idA,
                       const uint32_t *bcA, int m, const uint32_t *posB, char **idB,
                       const uint32_t *bcB, int n, int snp_size)
{
    memset(w, 0, sizeof *w);
    w->fp = fp;
    w->posA = posA;
    w->bcA = bcA;
    w->posB = posB;
    w->bcB = bcB;
    w->idA = idA;
    w->idB = idB;
    w->snp_size = snp_size;
    w->line_max = LINE_MAX_BYTES + max_id_len(idA, m) + max_id_len(idB, n);
    w->nthreads = omp_get_max_threads();
    if (w->nthreads > 64) w->nthreads = 64;
    if (w->nthreads < 1) w->nthreads = 1;
    w->fa = (char (*)[32])malloc((size_t)(m > 0 ? m : 1) * 32);
    w->buf = (char **)calloc((size_t)w->nthreads, sizeof(char *));
    w->cap = (size_t *)calloc((size_t)w->nthreads, sizeof(size_t));
    w->len = (size_t *)calloc((size_t)w->nthreads, sizeof(size_t));
    if (!w->fa || !w->buf || !w->cap || !w->len) return QLD_ERR_NOMEM;
    /* frequency text of every table-A column, reused by every row */
#pragma omp parallel for schedule(static) num_threads(w->nthreads)
    for (int j = 0; j < m; j++)
        qld_format_float(w->fa[j], 1.0f * (float)bcA[j] / (float)snp_size, 9);
    return QLD_OK;
}

static void writer_free(hits_writer *w)
{
    for (int k = 0; w->buf && k < w->nthreads; k++) free(w->buf[k]);
    free(w->buf);
    free(w->cap);
    free(w->len);
    free(w->fa);
}

long qld_write_hits(FILE *fp,
                    const uint32_t *posA, char **idA, const uint32_t *bcA, int m,
                    const uint32_t *posB, char **idB, const uint32_t *bcB, int n,
                    int snp_size, int row0, int rows, const uint64_t *offsets,
                    const uint32_t *cols, const float *r2)
{
    if (!fp || m < 0 || n < 0 || row0 < 0 || rows < 0 || row0 + rows > n || snp_size <= 0 ||
        (rows && !offsets))
        return -1;
    if (rows == 0 || offsets[rows] == offsets[0]) return 0;
    hits_writer w;
    long lines = -1;
    if (writer_init(&w, fp, posA, idA, bcA, m, posB, idB, bcB, n, snp_size) == QLD_OK &&
        write_hits(&w, row0, rows, offsets, cols, r2) == 0)
        lines = w.lines;
    writer_free(&w);
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
    if (!ctx || !fp || m < 0 || n < 0 || (m && !posA) || (n && !posB) || snp_size <= 0)
        return QLD_ERR_INVALID;
    qld_write_results_header(fp);
    if (m == 0 || n == 0) return QLD_OK;
    hits_writer w;
    int rc = writer_init(&w, fp, posA, idA, bcA, m, posB, idB, bcB, n, snp_size);
    if (rc == QLD_OK) {
        rc = qld_correlate_hits(ctx, tableA, bcA, m, tableB, bcB, n, compressed_snp_size,
                                snp_size, posWset2 == 0, r2limit, 0, write_hits, &w);
        if (rc != QLD_OK && w.err) rc = QLD_ERR_NOMEM;   /* out of memory or a failed write */
    }
    writer_free(&w);
    return rc;
}
