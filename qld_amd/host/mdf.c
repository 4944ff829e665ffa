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
        if ((size_t)snprintf(path, sizeof path, "%s%s", d, e->d_name) >= sizeof path)
            continue;               /* longer than any path this reader opens */
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
    return qld_find_chunks(dir, id, snips_per_file, total_snips, pos_wmin, pos_wmax, ".mdf",
                           files_out, nfiles_out);
}

int qld_find_chunks(const char *dir, const char *id, int snips_per_file, int total_snips,
                    int pos_wmin, int pos_wmax, const char *suffix, char ***files_out,
                    int *nfiles_out)
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
            if ((size_t)snprintf(path, sizeof path, "%s%s", d, e->d_name) >= sizeof path)
                continue;           /* longer than any path this reader opens */
            if (!is_regular(path)) continue;
            if (!strstr(e->d_name, id) || strstr(e->d_name, "header") || !strstr(e->d_name, suffix))
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

/* One chunk file in a reader: its text (lines NUL-terminated in place), the
 * reference's view of each data line (the lines after the first two), and
 * the rows parsed so far.  kind: 0 = short line (the loop sees POS 0),
 * 1 = a line with POS, 2 = a counted line without a POS field. */
typedef struct {
    char *path;
    char *buf;
    size_t len;
    int rc;                /* load status */
    int has_data;          /* a second header line exists: data lines follow */
    int nlines;
    char **line;
    int *pos;
    unsigned char *kind;
    /* parsed rows, per data line */
    int words;             /* words per row of the parse store (0: none yet) */
    uint64_t *snp;         /* nlines x words */
    uint32_t *bc, *ppos;
    char **id;
    unsigned short *idlen;
    unsigned char *parsed; /* 0 no, 1 ok, 2 too few fields, 3 wrong word count */
    unsigned long long used;   /* reader call that last touched the file */
} mdf_file;

/* Files are allocated one by one and the registry holds pointers to them, so
 * a file's address stays valid while later calls register more files.  The
 * cache keeps at most `limit` bytes of file text, line index and parsed rows
 * after each call: least recently used files go first, never one the call
 * itself read (a single window larger than the limit is still served). */
struct qld_mdf_reader {
    mdf_file **f;
    int n, cap;
    unsigned long long calls;
    size_t limit;
};

#define QLD_MDF_CACHE_DEFAULT ((size_t)4 << 30)

qld_mdf_reader *qld_mdf_reader_create(void)
{
    qld_mdf_reader *r = (qld_mdf_reader *)calloc(1, sizeof(qld_mdf_reader));
    if (r) r->limit = QLD_MDF_CACHE_DEFAULT;
    return r;
}

void qld_mdf_reader_set_limit(qld_mdf_reader *r, size_t bytes)
{
    if (r) r->limit = bytes;
}

static void file_free_rows(mdf_file *f)
{
    free(f->snp);
    free(f->bc);
    free(f->ppos);
    free(f->id);
    free(f->idlen);
    free(f->parsed);
    f->snp = NULL;
    f->bc = f->ppos = NULL;
    f->id = NULL;
    f->idlen = NULL;
    f->parsed = NULL;
    f->words = 0;
}

/* Drop a file's text, index and rows; it stays registered and is read again
 * when a later call needs it. */
static void file_unload(mdf_file *f)
{
    free(f->buf);
    free(f->line);
    free(f->pos);
    free(f->kind);
    file_free_rows(f);
    f->buf = NULL;
    f->len = 0;
    f->line = NULL;
    f->pos = NULL;
    f->kind = NULL;
    f->nlines = 0;
    f->has_data = 0;
    f->rc = -1;
}

static size_t file_bytes(const mdf_file *f)
{
    if (f->rc != 0) return 0;
    size_t n = (size_t)f->nlines;
    size_t b = f->len + 1 + n * (sizeof(char *) + sizeof(int) + 1);
    if (f->words)
        b += n * ((size_t)f->words * 8 + 8 + sizeof(char *) + sizeof(unsigned short) + 1);
    return b;
}

size_t qld_mdf_reader_bytes(const qld_mdf_reader *r)
{
    size_t b = 0;
    if (r)
        for (int i = 0; i < r->n; i++) b += file_bytes(r->f[i]);
    return b;
}

/* Unload least recently used files, none touched by call `now`, until the
 * cache holds at most r->limit bytes. */
static void reader_trim(qld_mdf_reader *r, unsigned long long now)
{
    size_t total = qld_mdf_reader_bytes(r);
    while (total > r->limit) {
        mdf_file *victim = NULL;
        for (int i = 0; i < r->n; i++) {
            mdf_file *f = r->f[i];
            if (f->used == now || !file_bytes(f)) continue;
            if (!victim || f->used < victim->used) victim = f;
        }
        if (!victim) break;
        total -= file_bytes(victim);
        file_unload(victim);
    }
}

void qld_mdf_reader_free(qld_mdf_reader *r)
{
    if (!r) return;
    for (int i = 0; i < r->n; i++) {
        mdf_file *f = r->f[i];
        file_unload(f);
        free(f->path);
        free(f);
    }
    free(r->f);
    free(r);
}

/* Read and index one file: the reference's reader sees line 1 (must count,
 * else "file empty"), line 2 (if it does not count, no data line is read),
 * then the data lines (correlate_IOx64.c:371-444). */
static int load_file(mdf_file *f)
{
    int rc = slurp(f->path, &f->buf, &f->len);
    if (rc) return rc;
    cursor c = {f->buf, f->buf + f->len};
    char *line;
    int eof;
    if (!next_line(&c, &line, &eof)) return fail(QLD_ERR_INVALID, "file %s empty", f->path);
    if (eof || !next_line(&c, &line, &eof)) return QLD_OK;
    f->has_data = 1;
    int cap = 0;
    while (!eof) {
        int st = next_line(&c, &line, &eof);
        if (!line) break;       /* the read after a final newline */
        if (f->nlines == cap) {
            cap = cap ? cap * 2 : 4096;
            char **nl = (char **)realloc(f->line, (size_t)cap * sizeof(char *));
            int *np = (int *)realloc(f->pos, (size_t)cap * sizeof(int));
            unsigned char *nk = (unsigned char *)realloc(f->kind, (size_t)cap);
            if (nl) f->line = nl;
            if (np) f->pos = np;
            if (nk) f->kind = nk;
            if (!nl || !np || !nk) return fail(QLD_ERR_NOMEM, "out of host memory");
        }
        int k = f->nlines++;
        f->line[k] = line;
        f->pos[k] = 0;
        f->kind[k] = 0;
        if (st) {
            char *p = line;
            size_t wl;
            char *w = next_word(&p, &wl) ? next_word(&p, &wl) : NULL;
            f->kind[k] = w ? 1 : 2;
            if (w) f->pos[k] = atoi_word(w);
        }
    }
    return QLD_OK;
}

static mdf_file *reader_file(qld_mdf_reader *r, const char *path)
{
    for (int i = 0; i < r->n; i++)
        if (!strcmp(r->f[i]->path, path)) return r->f[i];
    if (r->n == r->cap) {
        int nc = r->cap ? r->cap * 2 : 64;
        mdf_file **nf = (mdf_file **)realloc(r->f, (size_t)nc * sizeof(mdf_file *));
        if (!nf) return NULL;
        r->f = nf;
        r->cap = nc;
    }
    mdf_file *f = (mdf_file *)calloc(1, sizeof *f);
    if (!f) return NULL;
    f->path = strdup(path);
    if (!f->path) {
        free(f);
        return NULL;
    }
    f->rc = -1;             /* not loaded yet */
    r->f[r->n++] = f;
    return f;
}

/* Parse data line k of f into its row slot: POS, ID, BC and the packed words
 * (strtoul per word, as writeMDF_x64 does, with a fast path for plain digit
 * strings).  A row may carry more than W words -- the MDF tool writes
 * ploidy-2 widths unless told otherwise, and the reference's own smoke run
 * then reads haploid rows from it: writeMDF_x64 (correlate_IOx64.c:102-115)
 * stores every word from the row's first slot on, the surplus lands in the
 * next row's slots, which that row overwrites, so the table keeps each row's
 * first W words.  Fewer than W words leave slots the reference never writes:
 * an error here. */
static void parse_line(mdf_file *f, int k)
{
    char *p = f->line[k];
    size_t wl;
    const int W = f->words;
    f->ppos[k] = 0;
    char *w = next_word(&p, &wl);                 /* CHROM */
    char *wp = w ? next_word(&p, &wl) : NULL;     /* POS */
    char *wid = wp ? next_word(&p, &wl) : NULL;   /* ID */
    size_t idl = wl;
    char *wbc = wid ? next_word(&p, &wl) : NULL;  /* BC */
    if (!wbc) {
        f->parsed[k] = 2;
        return;
    }
    f->ppos[k] = (uint32_t)atoi_word(wp);
    f->bc[k] = (uint32_t)atoi_word(wbc);
    uint64_t *row = f->snp + (size_t)k * W;
    int n = 0;
    while (n < W && (w = next_word(&p, &wl)) != NULL) row[n++] = u64_word(w, wl);
    if (n != W) {
        f->parsed[k] = 3;
        return;
    }
    f->id[k] = wid;
    f->idlen[k] = (unsigned short)(idl < 65535 ? idl : 65535);
    f->parsed[k] = 1;
}

static int ensure_rows(mdf_file *f, int words)
{
    if (f->words == words) return QLD_OK;
    file_free_rows(f);
    size_t n = f->nlines ? (size_t)f->nlines : 1;
    f->snp = (uint64_t *)malloc(n * (size_t)words * 8);
    f->bc = (uint32_t *)malloc(n * 4);
    f->ppos = (uint32_t *)malloc(n * 4);
    f->id = (char **)malloc(n * sizeof(char *));
    f->idlen = (unsigned short *)malloc(n * sizeof(unsigned short));
    f->parsed = (unsigned char *)calloc(n, 1);
    if (!f->snp || !f->bc || !f->ppos || !f->id || !f->idlen || !f->parsed) {
        file_free_rows(f);
        return fail(QLD_ERR_NOMEM, "out of host memory");
    }
    f->words = words;
    return QLD_OK;
}

typedef struct {
    int file, line;
} row_ref;

int qld_mdf_reader_table(qld_mdf_reader *r, char *const *files, int nfiles, int pos_wmin,
                         int pos_wmax, int pos_wset, int words, int nthreads,
                         qld_snp_table *t)
{
    if (!r || !t || nfiles < 0 || words <= 0 || (nfiles && !files))
        return fail(QLD_ERR_INVALID, "invalid argument");
    memset(t, 0, sizeof *t);
    t->words = words;
    if (nthreads <= 0) nthreads = 1;
    mdf_file **fs = (mdf_file **)malloc((size_t)(nfiles ? nfiles : 1) * sizeof(mdf_file *));
    if (!fs) return fail(QLD_ERR_NOMEM, "out of host memory");
    int rc = QLD_OK;
    const unsigned long long now = ++r->calls;
    for (int i = 0; i < nfiles && !rc; i++) {
        if (!(fs[i] = reader_file(r, files[i]))) rc = fail(QLD_ERR_NOMEM, "out of host memory");
        else fs[i]->used = now;
    }
    /* load the files this reader has not seen (each listed once) */
    char errs[1024] = "";
    if (!rc) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (int i = 0; i < nfiles; i++) {
            mdf_file *f = fs[i];
            int first = 1;
            for (int j = 0; j < i; j++)
                if (fs[j] == f) first = 0;
            if (!first || f->rc != -1) continue;
            /* g_err is thread-local: a failing load leaves its message in
             * this worker's own copy, taken into errs under the lock */
            f->rc = load_file(f);
            if (f->rc) {
#pragma omp critical
                {
                    if (!errs[0]) snprintf(errs, sizeof errs, "%s", g_err);
                }
            }
        }
        for (int i = 0; i < nfiles && !rc; i++)
            if (fs[i]->rc) {
                rc = fs[i]->rc;
                snprintf(g_err, sizeof g_err, "%s", errs[0] ? errs : "load failed");
            }
    }
    /* pass 1: the reference's per-file line loop over the index */
    row_ref *sel = NULL;
    long nsel = 0, cap = 0;
    for (int i = 0; i < nfiles && !rc; i++) {
        mdf_file *f = fs[i];
        if (!f->has_data) continue;
        for (int k = 0; k < f->nlines; k++) {
            int pos = 0;
            if (f->kind[k] == 2) {
                rc = fail(QLD_ERR_INVALID, "%s: line without CHROM and POS fields", f->path);
                break;
            }
            if (f->kind[k] == 1) {
                pos = f->pos[k];
                if (pos >= pos_wmin && pos <= pos_wmax) {
                    if (nsel == cap) {
                        cap = cap ? cap * 2 : 4096;
                        row_ref *ns = (row_ref *)realloc(sel, (size_t)cap * sizeof(row_ref));
                        if (!ns) {
                            rc = fail(QLD_ERR_NOMEM, "out of host memory");
                            break;
                        }
                        sel = ns;
                    }
                    sel[nsel].file = i;
                    sel[nsel].line = k;
                    nsel++;
                }
            }
            if (!pos_wset || pos >= pos_wmax) break;
        }
    }
    if (!rc && nsel > INT_MAX) rc = fail(QLD_ERR_INVALID, "too many rows (%ld)", nsel);
    for (int i = 0; i < nfiles && !rc; i++) rc = ensure_rows(fs[i], words);
    /* pass 2: parse the selected lines not parsed before (distinct lines) */
    if (!rc) {
#pragma omp parallel for schedule(static, 64) num_threads(nthreads)
        for (long q = 0; q < nsel; q++) {
            mdf_file *f = fs[sel[q].file];
            if (!f->parsed[sel[q].line]) {
                int dup = 0;     /* a file listed twice selects its lines twice */
                for (int j = 0; j < sel[q].file; j++)
                    if (fs[j] == f) dup = 1;
                if (!dup) parse_line(f, sel[q].line);
            }
        }
        for (long q = 0; q < nsel && !rc; q++) {
            mdf_file *f = fs[sel[q].file];
            int k = sel[q].line;
            if (f->parsed[k] == 2)
                rc = fail(QLD_ERR_INVALID, "%s: row %ld (POS %d): fewer than 4 fields", f->path,
                          q, f->pos[k]);
            else if (f->parsed[k] == 3)
                rc = fail(QLD_ERR_INVALID,
                          "%s: row %ld (POS %d): fewer packed words than the %d the "
                          "header's samples imply", f->path, q, f->pos[k], words);
        }
    }
    /* assemble */
    if (!rc) {
        size_t nr = nsel ? (size_t)nsel : 1;
        t->rows = (int)nsel;
        t->snp = (uint64_t *)malloc(nr * (size_t)words * 8);
        t->pos = (uint32_t *)malloc(nr * 4);
        t->bc = (uint32_t *)malloc(nr * 4);
        t->id = (char **)malloc(nr * sizeof(char *));
        size_t idtot = 0;
        for (long q = 0; q < nsel; q++) idtot += fs[sel[q].file]->idlen[sel[q].line] + 1u;
        t->idbuf = (char *)malloc(idtot ? idtot : 1);
        if (!t->snp || !t->pos || !t->bc || !t->id || !t->idbuf) {
            rc = fail(QLD_ERR_NOMEM, "out of host memory for %ld rows", nsel);
        } else {
            size_t o = 0;
            for (long q = 0; q < nsel; q++) {
                const mdf_file *f = fs[sel[q].file];
                const int k = sel[q].line;
                t->id[q] = t->idbuf + o;
                o += f->idlen[k] + 1u;
            }
#pragma omp parallel for schedule(static, 256) num_threads(nthreads)
            for (long q = 0; q < nsel; q++) {
                const mdf_file *f = fs[sel[q].file];
                const int k = sel[q].line;
                memcpy(t->snp + (size_t)q * words, f->snp + (size_t)k * words,
                       (size_t)words * 8);
                t->pos[q] = f->ppos[k];
                t->bc[q] = f->bc[k];
                memcpy(t->id[q], f->id[k], f->idlen[k]);
                t->id[q][f->idlen[k]] = '\0';
            }
        }
    }
    free(sel);
    free(fs);
    if (rc) qld_snp_table_free(t);
    reader_trim(r, now);
    return rc;
}

int qld_mdf_read_table(char *const *files, int nfiles, int pos_wmin, int pos_wmax,
                       int pos_wset, int words, int nthreads, qld_snp_table *t)
{
    qld_mdf_reader *r = qld_mdf_reader_create();
    if (!r) return fail(QLD_ERR_NOMEM, "out of host memory");
    int rc = qld_mdf_reader_table(r, files, nfiles, pos_wmin, pos_wmax, pos_wset, words,
                                  nthreads, t);
    qld_mdf_reader_free(r);
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
