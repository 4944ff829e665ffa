/*
 * mdf.c -- MDF input directories: header discovery, chunk selection, table
 * loading and MDF writing, with the reference correlator's semantics.
 *
 * An MDF directory holds `<id>_header.vcf.gz` (a VCF header whose line after
 * `#CHROM` reads `id snipsPerFile snipSize totalSnips posMin posMax`,
 * parse_vcf.c:620-634) and chunk files `<id>_<first>_<last>_<posMin>_<posMax>.mdf`
 * (process_data.c:430-460 of the MDF tool): two `##` lines, a `#CHROM` line,
 * then one row per SNP, `chrom pos ID BC w0 w1 ...` with the packed uint64
 * words in decimal.
 *
 * Loading is two passes: a sequential scan per file decides which lines are
 * rows of the window (the reference's stop rule is sequential), then every
 * selected row is parsed in parallel straight into the packed table.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <dirent.h>
#include <errno.h>
#include <limits.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <zlib.h>

#include "qld_host.h"

static __thread char g_err[1024];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *qld_host_last_error(void) { return g_err; }

/* ---- files ------------------------------------------------------------ */

/* Whole file (plain or gzip, as gzopen reads both) into a NUL-terminated
 * buffer. */
static int slurp(const char *path, char **out, size_t *len)
{
    gzFile f = gzopen(path, "rb");
    if (!f) return fail(QLD_ERR_NOTFOUND, "cannot open %s", path);
    gzbuffer(f, 1 << 20);
    size_t cap = 1 << 20, n = 0;
    char *buf = (char *)malloc(cap + 1);
    if (!buf) { gzclose(f); return fail(QLD_ERR_NOMEM, "out of host memory reading %s", path); }
    for (;;) {
        if (cap - n < (1u << 20)) {
            cap *= 2;
            char *nb = (char *)realloc(buf, cap + 1);
            if (!nb) { free(buf); gzclose(f); return fail(QLD_ERR_NOMEM, "out of host memory reading %s", path); }
            buf = nb;
        }
        int r = gzread(f, buf + n, (unsigned)(cap - n > (1u << 30) ? (1u << 30) : cap - n));
        if (r < 0) { free(buf); gzclose(f); return fail(QLD_ERR_INVALID, "read error in %s", path); }
        if (r == 0) break;
        n += (size_t)r;
    }
    gzclose(f);
    buf[n] = '\0';
    *out = buf;
    *len = n;
    return QLD_OK;
}

/* Line cursor with the reference's reader rules (read_file.c:33-74): a line
 * is the text up to '\n' (or the end of the file); it counts only when its
 * string length is >= 3; `eof` is raised when the read hits the end of the
 * file (a last line without '\n' is still delivered). */
typedef struct {
    char *p, *end;
} cursor;

/* Returns 1 for a counted line (*line NUL-terminated in place), 0 for a
 * short line; *eof as the reference's readEOF. */
static int next_line(cursor *c, char **line, int *eof)
{
    *eof = 0;
    if (c->p >= c->end) { *eof = 1; *line = NULL; return 0; }
    char *s = c->p;
    char *nl = (char *)memchr(s, '\n', (size_t)(c->end - s));
    if (nl) {
        *nl = '\0';
        c->p = nl + 1;
    } else {
        c->p = c->end;
        *eof = 1;
    }
    *line = s;
    return strlen(s) >= 3;
}

/* Whitespace-separated words as getWordFromString (read_file.c:76-112):
 * separators are ' ' and '\t' only. */
static char *next_word(char **pp, size_t *len)
{
    char *p = *pp;
    while (*p == ' ' || *p == '\t') p++;
    if (!*p) { *pp = p; return NULL; }
    char *w = p;
    while (*p && *p != ' ' && *p != '\t') p++;
    *len = (size_t)(p - w);
    *pp = p;
    return w;
}

static int atoi_word(const char *w) { return (int)strtol(w, NULL, 10); }

/* strtoul(w, 10) for a word: plain digit strings of up to 19 digits take a
 * fast path; anything else goes through strtoul, as the reference parses. */
static inline uint64_t u64_word(const char *w, size_t len)
{
    if (len <= 19) {
        uint64_t v = 0;
        size_t i = 0;
        for (; i < len; i++) {
            unsigned d = (unsigned)(w[i] - '0');
            if (d > 9) break;
            v = v * 10 + d;
        }
        if (i == len) return v;
    }
    return (uint64_t)strtoul(w, NULL, 10);
}

/* ---- header (readHeaderFile, correlate_IO.c:79-255) -------------------- */

static void dir_with_slash(char *out, size_t cap, const char *dir)
{
    snprintf(out, cap, "%s", dir);
    size_t n = strlen(out);
    if (n == 0 || out[n - 1] != '/') {
        if (n + 1 < cap) { out[n] = '/'; out[n + 1] = '\0'; }
    }
}

static int is_regular(const char *path)
{
    struct stat st;
    return stat(path, &st) == 0 && S_ISREG(st.st_mode);
}

int qld_mdf_read_header(const char *dir, qld_mdf_header *h)
{
    if (!dir || !h) return fail(QLD_ERR_INVALID, "null argument");
    memset(h, 0, sizeof *h);
    char d[QLD_PATH_MAX];
    dir_with_slash(d, sizeof d, dir);
    DIR *dp = opendir(d);
    if (!dp) return fail(QLD_ERR_NOTFOUND, "cannot open directory: %s", d);
    struct dirent *e;
    int found = 0;
    while ((e = readdir(dp)) != NULL) {
        char path[QLD_PATH_MAX];
        snprintf(path, sizeof path, "%s%s", d, e->d_name);
        if (!is_regular(path) || !strstr(e->d_name, "header")) continue;
        if (found) continue;        /* the first header file found is used */
        found = 1;
        snprintf(h->header_path, sizeof h->header_path, "%s", path);
        /* alignment id: the name up to the first '_' or '\'' */
        size_t k = 0;
        while (e->d_name[k] && e->d_name[k] != '_' && e->d_name[k] != '\'' &&
               k + 1 < sizeof h->id) {
            h->id[k] = e->d_name[k];
            k++;
        }
        h->id[k] = '\0';
    }
    closedir(dp);
    if (!found) return fail(QLD_ERR_NOTFOUND, "could not find a header file in %s", d);

    char *buf;
    size_t len;
    int rc = slurp(h->header_path, &buf, &len);
    if (rc) return rc;
    cursor c = {buf, buf + len};
    char *line;
    int eof;
    int st = next_line(&c, &line, &eof);
    if (!st) {
        free(buf);
        return fail(QLD_ERR_INVALID, "empty header file %s", h->header_path);
    }
    static const char *fields[9] = {"#CHROM", "POS", "ID", "REF", "ALT",
                                    "QUAL", "FILTER", "INFO", "FORMAT"};
    int have_info = 0;
    while (!eof) {
        st = next_line(&c, &line, &eof);
        if (!st || strlen(line) < 6 || strncmp(line, "#CHROM", 6) != 0) continue;
        char *p = line;
        size_t wl;
        for (int f = 0; f < 9; f++) {
            char *w = next_word(&p, &wl);
            if (!w || wl != strlen(fields[f]) || strncmp(w, fields[f], wl) != 0) {
                free(buf);
                return fail(QLD_ERR_INVALID, "VCF header field %s is missing in %s",
                            fields[f], h->header_path);
            }
        }
        int samples = 0;
        while (next_word(&p, &wl)) samples++;
        if (samples == 0) {
            free(buf);
            return fail(QLD_ERR_INVALID, "no samples on the #CHROM line of %s", h->header_path);
        }
        h->samples = samples;
        st = next_line(&c, &line, &eof);
        if (!st) {
            free(buf);
            return fail(QLD_ERR_INVALID, "wrong header format in %s", h->header_path);
        }
        char idt[256];
        if (sscanf(line, "%255s %d %d %d %d %d", idt, &h->snips_per_file, &h->snip_size,
                   &h->total_snips, &h->pos_min, &h->pos_max) < 1 || strcmp(idt, h->id) != 0) {
            free(buf);
            return fail(QLD_ERR_INVALID, "different internal and external ID in %s: %s != %s",
                        h->header_path, idt, h->id);
        }
        have_info = 1;
    }
    free(buf);
    if (!have_info)
        return fail(QLD_ERR_INVALID, "no #CHROM line in header file %s", h->header_path);
    return QLD_OK;
}

/* ---- chunk selection (findFiles, correlate_IO.c:294-376) --------------- */

void qld_mdf_free_files(char **files, int nfiles)
{
    if (!files) return;
    for (int i = 0; i < nfiles; i++) free(files[i]);
    free(files);
}

int qld_mdf_find_files(const char *dir, const char *id, int snips_per_file,
                       int total_snips, int pos_wmin, int pos_wmax,
                       char ***files_out, int *nfiles_out)
{
    if (!dir || !id || !files_out || !nfiles_out) return fail(QLD_ERR_INVALID, "null argument");
    char d[QLD_PATH_MAX];
    dir_with_slash(d, sizeof d, dir);
    char **files = NULL;
    int nfiles = 0;
    int pos_min_index = 0, prev_pos_max = 0, prev_max = 1, eop = 0;
    /* The reference rescans the directory until the window is closed; each
     * scan takes every chunk that continues the index sequence.  A scan that
     * moves nothing forward would repeat forever, so it ends with an error. */
    while (!eop) {
        int before = prev_max;
        DIR *dp = opendir(d);
        if (!dp) { qld_mdf_free_files(files, nfiles); return fail(QLD_ERR_NOTFOUND, "cannot open directory: %s", d); }
        struct dirent *e;
        while ((e = readdir(dp)) != NULL) {
            char path[QLD_PATH_MAX];
            snprintf(path, sizeof path, "%s%s", d, e->d_name);
            if (!is_regular(path)) continue;
            if (!strstr(e->d_name, id) || strstr(e->d_name, "header") || !strstr(e->d_name, ".mdf"))
                continue;
            char tid[256];
            int tmin, tmax, tpmin, tpmax;
            if (sscanf(e->d_name, "%255[^_]_%d_%d_%d_%d.", tid, &tmin, &tmax, &tpmin, &tpmax) != 5)
                continue;
            if (prev_pos_max == 0) prev_pos_max = tpmax;
            if (tmin != prev_max) continue;
            int starts = (tpmin <= pos_wmin && pos_wmin <= tpmax && pos_min_index == 0) ||
                         (prev_pos_max < pos_wmin && pos_wmin < tpmin && pos_min_index == 0) ||
                         (tmin <= pos_min_index && pos_min_index <= tmax);
            if (starts && !eop && ((tmax - tmin + 1) == snips_per_file || tmax == total_snips)) {
                char **nf = (char **)realloc(files, (size_t)(nfiles + 1) * sizeof(char *));
                if (!nf) { closedir(dp); qld_mdf_free_files(files, nfiles); return fail(QLD_ERR_NOMEM, "out of host memory"); }
                files = nf;
                files[nfiles++] = strdup(path);
                if (tpmax < pos_wmax) pos_min_index = tmax + 1;
                else eop = 1;
            }
            prev_pos_max = tpmax;
            prev_max = tmax + 1;
        }
        closedir(dp);
        if (!eop && prev_max == before) {
            qld_mdf_free_files(files, nfiles);
            return fail(QLD_ERR_NOTFOUND,
                        "%s: no chunk of %s continues at SNP index %d before the window "
                        "[%d, %d] closes", d, id, prev_max, pos_wmin, pos_wmax);
        }
    }
    *files_out = files;
    *nfiles_out = nfiles;
    return QLD_OK;
}

/* ---- table loading (readTable_x64 + writeMDF_x64, correlate_IOx64.c) --- */

void qld_snp_table_free(qld_snp_table *t)
{
    if (!t) return;
    free(t->snp);
    free(t->pos);
    free(t->bc);
    free(t->id);
    free(t->idbuf);
    memset(t, 0, sizeof *t);
}

typedef struct {
    char *buf;
    size_t len;
    char **rows;   /* selected row lines (NUL-terminated in buf) */
    int nrows, cap;
    int rc;
} file_scan;

/* Pass 1 for one file: the reference's line loop (correlate_IOx64.c:371-444),
 * recording the lines whose POS lies in [wmin, wmax]. */
static int scan_file(const char *path, file_scan *fs, int wmin, int wmax, int wset)
{
    int rc = slurp(path, &fs->buf, &fs->len);
    if (rc) return rc;
    cursor c = {fs->buf, fs->buf + fs->len};
    char *line;
    int eof;
    if (!next_line(&c, &line, &eof))
        return fail(QLD_ERR_INVALID, "file %s empty", path);
    if (!next_line(&c, &line, &eof))
        return QLD_OK;          /* no second line: the file contributes nothing */
    int pos;
    do {
        if (next_line(&c, &line, &eof)) {
            char *p = line;
            size_t wl;
            char *w = next_word(&p, &wl) ? next_word(&p, &wl) : NULL;
            if (!w)
                return fail(QLD_ERR_INVALID, "%s: line without CHROM and POS fields", path);
            pos = atoi_word(w);
            if (pos >= wmin && pos <= wmax) {
                if (fs->nrows == fs->cap) {
                    int nc = fs->cap ? fs->cap * 2 : 1024;
                    char **nr = (char **)realloc(fs->rows, (size_t)nc * sizeof(char *));
                    if (!nr) return fail(QLD_ERR_NOMEM, "out of host memory");
                    fs->rows = nr;
                    fs->cap = nc;
                }
                fs->rows[fs->nrows++] = line;
            }
        } else {
            pos = 0;
        }
    } while (!eof && wset && pos < wmax);
    return QLD_OK;
}

/* Pass 2 for one row line: POS, ID, BC, then exactly `words` decimal words. */
static int parse_row(char *line, int words, uint64_t *snp, uint32_t *pos,
                     uint32_t *bc, char **id, size_t *idlen)
{
    char *p = line;
    size_t wl;
    *pos = 0;
    char *w = next_word(&p, &wl);                 /* CHROM */
    char *wp = next_word(&p, &wl);                /* POS */
    if (!w || !wp) return -1;
    *pos = (uint32_t)atoi_word(wp);
    char *wid = next_word(&p, &wl);               /* ID */
    if (!wid) return -1;
    *id = wid;
    *idlen = wl;
    char *wbc = next_word(&p, &wl);               /* BC */
    if (!wbc) return -1;
    *bc = (uint32_t)atoi_word(wbc);
    int k = 0;
    while ((w = next_word(&p, &wl)) != NULL) {
        if (k == words) return -2;
        snp[k++] = u64_word(w, wl);
    }
    return k == words ? 0 : -2;
}

int qld_mdf_read_table(char *const *files, int nfiles, int pos_wmin, int pos_wmax,
                       int pos_wset, int words, int nthreads, qld_snp_table *t)
{
    if (!t || nfiles < 0 || words <= 0 || (nfiles && !files))
        return fail(QLD_ERR_INVALID, "invalid argument");
    memset(t, 0, sizeof *t);
    t->words = words;
    file_scan *fs = (file_scan *)calloc((size_t)(nfiles ? nfiles : 1), sizeof(file_scan));
    if (!fs) return fail(QLD_ERR_NOMEM, "out of host memory");
    int rc = QLD_OK;
    char errs[1024] = "";
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int f = 0; f < nfiles; f++) {
        fs[f].rc = scan_file(files[f], &fs[f], pos_wmin, pos_wmax, pos_wset);
        if (fs[f].rc) {
#pragma omp critical
            {
                if (!rc) { rc = fs[f].rc; snprintf(errs, sizeof errs, "%s", g_err); }
            }
        }
    }
    long total = 0;
    for (int f = 0; f < nfiles && !rc; f++) total += fs[f].nrows;
    if (!rc && total > INT_MAX) rc = fail(QLD_ERR_INVALID, "too many rows (%ld)", total);
    if (!rc) {
        t->rows = (int)total;
        size_t nr = total ? (size_t)total : 1;
        t->snp = (uint64_t *)malloc(nr * (size_t)words * sizeof(uint64_t));
        t->pos = (uint32_t *)malloc(nr * sizeof(uint32_t));
        t->bc = (uint32_t *)malloc(nr * sizeof(uint32_t));
        t->id = (char **)malloc(nr * sizeof(char *));
        if (!t->snp || !t->pos || !t->bc || !t->id)
            rc = fail(QLD_ERR_NOMEM, "out of host memory for %ld rows", total);
    }
    if (!rc && total) {
        char **lines = (char **)malloc((size_t)total * sizeof(char *));
        size_t *idl = (size_t *)malloc((size_t)total * sizeof(size_t));
        if (!lines || !idl) rc = fail(QLD_ERR_NOMEM, "out of host memory");
        long r0 = 0;
        for (int f = 0; f < nfiles && !rc; f++) {
            memcpy(lines + r0, fs[f].rows, (size_t)fs[f].nrows * sizeof(char *));
            r0 += fs[f].nrows;
        }
        long bad = -1;
        int badk = 0;
        if (!rc) {
#pragma omp parallel for schedule(static, 256) num_threads(nthreads)
            for (long r = 0; r < total; r++) {
                int e = parse_row(lines[r], words, t->snp + (size_t)r * words, &t->pos[r],
                                  &t->bc[r], &t->id[r], &idl[r]);
                if (e) {
#pragma omp critical
                    {
                        if (bad < 0 || r < bad) { bad = r; badk = e; }
                    }
                }
            }
            if (bad >= 0)
                rc = fail(QLD_ERR_INVALID,
                          badk == -1 ? "row %ld (POS %u): fewer than 4 fields"
                                     : "row %ld (POS %u): the number of packed words differs "
                                       "from the %d the header's samples imply",
                          bad, t->pos[bad], words);
        }
        if (!rc) {
            /* IDs into one arena (the text buffers are released below). */
            size_t tot = 0;
            for (long r = 0; r < total; r++) tot += idl[r] + 1;
            t->idbuf = (char *)malloc(tot ? tot : 1);
            if (!t->idbuf) rc = fail(QLD_ERR_NOMEM, "out of host memory");
            size_t o = 0;
            for (long r = 0; r < total && !rc; r++) {
                memcpy(t->idbuf + o, t->id[r], idl[r]);
                t->idbuf[o + idl[r]] = '\0';
                t->id[r] = t->idbuf + o;
                o += idl[r] + 1;
            }
        }
        free(lines);
        free(idl);
    }
    for (int f = 0; f < nfiles; f++) {
        free(fs[f].buf);
        free(fs[f].rows);
    }
    free(fs);
    if (rc) {
        if (errs[0]) snprintf(g_err, sizeof g_err, "%s", errs);
        qld_snp_table_free(t);
    }
    return rc;
}

/* ---- MDF writer (process_data.c:447-460 of the MDF tool) -------------- */

int qld_mdf_write(const char *path, const char *chrom, const uint64_t *snp,
                  const uint32_t *pos, char *const *id, const uint32_t *bc,
                  int rows, int words)
{
    if (!path || !chrom || rows < 0 || words <= 0 || (rows && (!snp || !pos || !bc)))
        return fail(QLD_ERR_INVALID, "invalid argument");
    FILE *fp = fopen(path, "w");
    if (!fp) return fail(QLD_ERR_NOTFOUND, "cannot create %s: %s", path, strerror(errno));
    setvbuf(fp, NULL, _IOFBF, 1 << 20);
    fprintf(fp, "##fileformat=MDFv0.1\n##SNPs:%d\n#CHROM\tPOS\tID\tBC", rows);
    for (int k = 0; k < words; k++) fprintf(fp, "\tcompSNP_%d", k);
    for (int r = 0; r < rows; r++) {
        fprintf(fp, "\n%s\t%d\t%s\t%u", chrom, (int)pos[r], id ? id[r] : ".", bc[r]);
        const uint64_t *w = snp + (size_t)r * words;
        for (int k = 0; k < words; k++) fprintf(fp, "\t%lu", (unsigned long)w[k]);
    }
    int bad = ferror(fp);
    if (fclose(fp) != 0 || bad) return fail(QLD_ERR_INVALID, "write error on %s", path);
    return QLD_OK;
}
