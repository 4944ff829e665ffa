/*
 * vcf.c -- VCF chunk input: the rows the reference correlator builds from
 * the parser's gzip VCF chunks when it runs without -mdf.
 *
 * Per data line (correlate_IOx64.c:125-349, writeToTable_x64): REF and ALT
 * must be single bases, FILTER PASS or ".", INFO's AF not 0 or 1 and its VT
 * (if any) SNP, FORMAT must carry GT; each selected sample's GT alleles are
 * mapped to bases (snip_processing.c:306-419), unphased alleles shuffled with
 * the C library's rand(); sites that are not polymorphic or that hold N
 * before their second base are dropped (:469-600); the two bases become
 * '0' / '1' in order of first appearance (:602-747) and are packed MSB-first
 * (compressSnip_x64, :783-815) with the bit count as BC.
 *
 * The reference draws from the process-wide rand() stream, never seeded, in
 * row, sample and task order.  A reader here owns a private generator with
 * glibc's default state (random_r over a 128-byte TYPE_3 state seeded 1,
 * the generator rand() uses), advanced in the same order, so the shuffles
 * match without touching the process's own rand().
 */
#define _GNU_SOURCE
#include <limits.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "qld_host.h"

static __thread char v_err[1024];

static int vfail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(v_err, sizeof v_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *qld_vcf_last_error(void) { return v_err; }

struct qld_vcf_reader {
    struct random_data rd;
    char state[128];
};

qld_vcf_reader *qld_vcf_reader_create(void)
{
    qld_vcf_reader *r = (qld_vcf_reader *)calloc(1, sizeof *r);
    if (!r) return NULL;
    initstate_r(1, r->state, sizeof r->state, &r->rd);
    srandom_r(1, &r->rd);
    return r;
}

void qld_vcf_reader_free(qld_vcf_reader *r) { free(r); }

void qld_vcf_reader_seed(qld_vcf_reader *r, unsigned int seed)
{
    if (r) srandom_r(seed, &r->rd);
}

static int vrand(qld_vcf_reader *r)
{
    int32_t v;
    random_r(&r->rd, &v);
    return (int)v;
}

/* Returned by the line parser when a line cannot be handled in the current
 * draw mode (below): it is parsed again, in order, with the reader's
 * stream. */
#define VDEFER (-1000)

/* rand() values drawn ahead of the lines that use them, at most */
#ifndef QLD_VCF_AHEAD
#define QLD_VCF_AHEAD (1 << 24)
#endif

/* Where a line's rand() values come from.  STREAM: the reader's stream (a
 * sequential read).  COUNT: nothing is drawn or shuffled, the draws the line
 * would make are counted (its row is exact only when it makes none).
 * REPLAY: the line's own values, drawn from the stream beforehand in line
 * order. */
enum { VR_STREAM, VR_COUNT, VR_REPLAY };
typedef struct {
    int mode;
    qld_vcf_reader *r;
    const int32_t *vals;
    long n, used;
} vrng;

static int vdraw(vrng *g, int *v)
{
    if (g->mode == VR_STREAM) {
        *v = vrand(g->r);
        g->used++;
        return QLD_OK;
    }
    if (g->mode == VR_REPLAY && g->used < g->n) {
        *v = g->vals[g->used++];
        return QLD_OK;
    }
    return VDEFER;
}

/* ---- text helpers (read_file.c semantics) ----------------------------- */

typedef struct {
    char *p, *end;
} vcursor;

static int vnext_line(vcursor *c, char **line, int *eof)
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

/* Next word separated by ' ' or '\t'; copies it into w (NUL-terminated);
 * *eol as getWordFromString. */
static int vword(char **pp, char *w, size_t cap, int *eol)
{
    char *p = *pp;
    *eol = 0;
    while (*p == ' ' || *p == '\t') p++;
    if (!*p) { *pp = p; *eol = 1; w[0] = '\0'; return 0; }
    size_t n = 0;
    while (*p && *p != ' ' && *p != '\t') {
        if (n + 1 < cap) w[n++] = *p;
        p++;
    }
    w[n] = '\0';
    if (!*p) *eol = 1;
    *pp = p;
    return 1;
}

static int vslurp(const char *path, char **out, size_t *len)
{
    gzFile f = gzopen(path, "rb");
    if (!f) return vfail(QLD_ERR_NOTFOUND, "Failed to open file %s", path);
    size_t cap = 1 << 20, n = 0;
    char *buf = (char *)malloc(cap + 1);
    if (!buf) { gzclose(f); return vfail(QLD_ERR_NOMEM, "out of host memory"); }
    for (;;) {
        if (cap - n < (1u << 20)) {
            cap *= 2;
            char *nb = (char *)realloc(buf, cap + 1);
            if (!nb) { free(buf); gzclose(f); return vfail(QLD_ERR_NOMEM, "out of host memory"); }
            buf = nb;
        }
        int rr = gzread(f, buf + n, (unsigned)(cap - n > (1u << 30) ? (1u << 30) : cap - n));
        if (rr < 0) { free(buf); gzclose(f); return vfail(QLD_ERR_INVALID, "read error in %s", path); }
        if (rr == 0) break;
        n += (size_t)rr;
    }
    gzclose(f);
    buf[n] = '\0';
    *out = buf;
    *len = n;
    return QLD_OK;
}

/* ---- header lines and sample selection --------------------------------- */

int qld_vcf_header_lines(const char *header_path, char **line1, char **chrom_line)
{
    char *buf;
    size_t len;
    int rc = vslurp(header_path, &buf, &len);
    if (rc) return rc;
    vcursor c = {buf, buf + len};
    char *line;
    int eof;
    *line1 = *chrom_line = NULL;
    if (!vnext_line(&c, &line, &eof)) { free(buf); return vfail(QLD_ERR_INVALID, "empty header file"); }
    *line1 = strdup(line);
    while (!eof) {
        if (vnext_line(&c, &line, &eof) && strlen(line) >= 6 && !strncmp(line, "#CHROM", 6)) {
            free(*chrom_line);
            *chrom_line = strdup(line);
        }
    }
    free(buf);
    if (!*line1 || !*chrom_line) {
        free(*line1);
        free(*chrom_line);
        *line1 = *chrom_line = NULL;
        return vfail(QLD_ERR_INVALID, "no #CHROM line in %s", header_path);
    }
    return QLD_OK;
}

/* makeValidList / getFilter / sample_isValid (correlate_IO.c:35-76,259-290):
 * sample columns of the #CHROM line named in the list file (all samples
 * without one); valid_count = selected samples x ploidy.  A list whose names
 * match more columns than it has names (a name the header repeats) is an
 * error, where the reference fails its assertion (correlate_IO.c:287-288). */
int qld_vcf_sample_mask(const char *chrom_line, const char *list_file, int ploidy,
                        unsigned char **mask, int *nsamples, int *valid_count)
{
    char **names = NULL;
    int nnames = 0;
    if (list_file) {
        FILE *fp = fopen(list_file, "r");
        if (!fp) return vfail(QLD_ERR_NOTFOUND, "cannot open sample list %s", list_file);
        char w[1024];
        int oom = 0;
        while (!oom && fscanf(fp, "%1023s", w) == 1) {
            char **nn = (char **)realloc(names, (size_t)(nnames + 1) * sizeof(char *));
            if (nn) names = nn;
            char *d = nn ? strdup(w) : NULL;
            if (d) names[nnames++] = d;
            else oom = 1;
        }
        fclose(fp);
        if (oom) {
            for (int k = 0; k < nnames; k++) free(names[k]);
            free(names);
            return vfail(QLD_ERR_NOMEM, "out of host memory");
        }
    }
    size_t cap = strlen(chrom_line) + 1;
    char *w = (char *)malloc(cap);
    unsigned char *m = NULL;
    int n = 0, valid = 0, eol, fields = 0, oom = !w;
    char *p = (char *)chrom_line;
    while (!oom && vword(&p, w, cap, &eol))
        if (++fields >= 9) break;
    while (!oom && vword(&p, w, cap, &eol)) {
        unsigned char *nm = (unsigned char *)realloc(m, (size_t)(n + 1));
        if (!nm) { oom = 1; break; }
        m = nm;
        int ok = nnames == 0;
        for (int k = 0; k < nnames && !ok; k++) ok = !strcmp(names[k], w);
        m[n++] = (unsigned char)ok;
        if (ok) valid += ploidy;
        if (eol) break;
    }
    free(w);
    for (int k = 0; k < nnames; k++) free(names[k]);
    free(names);
    if (oom) {
        free(m);
        return vfail(QLD_ERR_NOMEM, "out of host memory");
    }
    if (nnames > 0 && (int64_t)valid > (int64_t)ploidy * nnames) {
        free(m);
        return vfail(QLD_ERR_INVALID, "sample list %s: %d names select %d sample columns "
                     "(a listed name appears more than once in the header)",
                     list_file, nnames, valid / (ploidy > 0 ? ploidy : 1));
    }
    *mask = m;
    *nsamples = n;
    *valid_count = valid;
    return QLD_OK;
}

/* ---- one data line (writeToTable_x64) ---------------------------------- */

enum { MAXST = 5, STATESALL_ = 8, MAXAF = 9 };

static int valid_base(char c)
{
    return c == 'A' || c == 'C' || c == 'G' || c == 'T' || c == 'N' || c == 'a' || c == 'c' ||
           c == 'g' || c == 't' || c == 'n' || c == '.';
}

static int scan_states(const char *sv, char x)
{
    for (int i = 0; i < MAXST; i++)
        if (sv[i] == x) return 1;
    return 0;
}

static int num_states(const char *sv)
{
    int n = 0;
    for (int i = 0; i < MAXST; i++) {
        if (sv[i] == 'X') return n;
        n++;
    }
    return n;
}

/* getStatesREF / getStatesALT (snip_processing.c:43-161): 1 ok, 0 no
 * states, <0 error */
static int get_states(const char *s, char *sv, int alt, int pos)
{
    int index = 0, elen = 0, slen = (int)strlen(s);
    char store = 'X';
    if (!alt) {
        for (int j = 0; j < MAXST; j++) sv[j] = 'X';
    } else {
        for (int j = 0; j < MAXST; j++)
            if (sv[j] == 'X') { index = j; break; }
    }
    for (int i = 0; i < slen; i++) {
        if (s[i] != ',') {
            if (s[i] == '<') {
                for (int j = 0; j < MAXST; j++) sv[j] = 'X';
                return 0;
            }
            if (!valid_base(s[i]))
                return vfail(QLD_ERR_INVALID, "invalid base %c (field %s) in position %d", s[i],
                             alt ? "ALT" : "REF", pos);
            elen++;
            store = s[i];
        }
        if (s[i] == ',' || i == slen - 1) {
            if (elen == 1) {
                if (scan_states(sv, store))
                    return vfail(QLD_ERR_INVALID, "Nucleotide %c (field %s) in position %d "
                                 "appears twice", store, alt ? "ALT" : "REF", pos);
                if (index >= MAXST) return vfail(QLD_ERR_INVALID, "too many states in position %d", pos);
                sv[index++] = store;
                elen = 0;
            } else {
                for (int j = 0; j < MAXST; j++) sv[j] = 'X';
                return 0;
            }
        }
    }
    return 1;
}

/* getValueAF (snip_processing.c:163-241); *af = -1 when INFO has no AF */
static int value_af(const char *s, int pos, float *af)
{
    int len = (int)strlen(s), i, j, k = 0;
    char afs[MAXAF + 1];
    float freqs[1000];
    afs[0] = 0;
    *af = -1.0f;
    if (!strcmp(".", s) || len < 4) return QLD_OK;
    for (i = 0; i < len - 3; i++)
        if (s[i] == 'A' && s[i + 1] == 'F' && s[i + 2] == '=' && (i == 0 || s[i - 1] == ';'))
            break;
    if (i == len - 3) return QLD_OK;
    i += 3;
    for (j = 0; j < MAXAF; j++) {
        if (s[i] == ',') {
            if (afs[0] == 0) return vfail(QLD_ERR_INVALID, "AF in field INFO in position %d has an empty value", pos);
            if (k >= 999) return vfail(QLD_ERR_INVALID, "too many AF values in position %d", pos);
            freqs[k] = (float)atof(afs);
            j = 0;
            ++i;
            ++k;
        }
        if (s[i] == ';' || i == len) break;
        if (s[i] == '.' || (s[i] >= '0' && s[i] <= '9') || s[i] == 'e' || s[i] == '-' ||
            s[i] == '+') {
            afs[j] = s[i];
            i++;
            afs[j + 1] = 0;
        } else {
            return vfail(QLD_ERR_INVALID, "Invalid character (%c) in AF of field INFO in position %d",
                         s[i], pos);
        }
    }
    if (strlen(afs) == 0)
        return vfail(QLD_ERR_INVALID, "AF in field INFO in position %d has no value", pos);
    float a;
    if (k == 0) {
        a = (float)atof(afs);
    } else {
        freqs[k] = (float)atof(afs);
        ++k;
        a = 0.0f;
        for (i = 0; i < k; ++i)
            if (freqs[i] > 0.0f && freqs[i] < 1.0f) { a = freqs[i]; break; }
    }
    if (a < 0.0f || a > 1.0f)
        return vfail(QLD_ERR_INVALID, "AF (%f) in line %d should be >= 0.0 and <= 1.0", a, pos);
    *af = a;
    return QLD_OK;
}

/* checkVTisSNP (:243-268): 1 SNP, 0 other, -1 no VT; -2 malformed */
static int vt_is_snp(const char *s)
{
    int len = (int)strlen(s), i;
    if (!strcmp(".", s) || len < 4) return -1;
    for (i = 0; i < len - 3; i++)
        if (s[i] == 'V' && s[i + 1] == 'T' && s[i + 2] == '=' && (i == 0 || s[i - 1] == ';'))
            break;
    if (i == len - 3) return -1;
    i += 3;
    if (i > len - 3) return -2;
    return s[i] == 'S' && s[i + 1] == 'N' && s[i + 2] == 'P';
}

/* getGTpos (:270-304) */
static int gt_pos(const char *s)
{
    int len = (int)strlen(s), n = 0;
    for (int i = 0; i < len - 1; i++) {
        if (s[i] == ':') n++;
        if (s[i] == 'G' && s[i + 1] == 'T' && (i == 0 || s[i - 1] == ':') &&
            (i + 2 == len || s[i + 2] == ':'))
            return n;
    }
    return -1;
}

/* getGTfield (:306-335), in place; returns separators + 1 */
static int gt_field(char *s, int pos)
{
    int len = (int)strlen(s), j = 0, counter = 0;
    for (int i = 0; i < len; i++) {
        if (s[i] != ':') {
            if (pos == 0) {
                s[j++] = s[i];
                if (s[i] == '|' || s[i] == '/') counter++;
            }
        } else {
            pos--;
            if (pos == -1) {
                s[j] = 0;
                break;
            }
        }
    }
    return counter + 1;
}

/* dataShuffleKnuth (:337-353): end - start draws */
static int shuffle(vrng *g, char *d, int start, int end)
{
    if (start == end) return QLD_OK;
    if (g->mode == VR_COUNT) {
        g->used += end - start;
        return QLD_OK;
    }
    for (int i = end; i > start; i--) {
        int v;
        if (vdraw(g, &v)) return VDEFER;
        int idx = start + (v % (i - start + 1));
        char t = d[idx];
        d[idx] = d[i];
        d[i] = t;
    }
    return QLD_OK;
}

/* getGTdata (:355-393) */
static int gt_data(vrng *g, const char *s, const char *sv, int total, char *out)
{
    int j = 0, start = 0, end = 0, len = (int)strlen(s);
    for (int i = 0; i < len; i++) {
        if (s[i] >= '0' && s[i] <= '9') {
            int index = s[i] - '0';
            if (index >= total || j + 1 >= 4)
                return vfail(QLD_ERR_INVALID, "genotype %s does not match REF/ALT", s);
            out[j++] = sv[index];
        } else {
            if (s[i] == '.') {
                if (j >= 4) return vfail(QLD_ERR_INVALID, "genotype %s too long", s);
                out[j++] = 'N';
            }
            if (s[i] == '/') end++;
            if (s[i] == '|') {
                if (end >= 4) return vfail(QLD_ERR_INVALID, "genotype %s too long", s);
                if (shuffle(g, out, start, end)) return VDEFER;
                start = j;
                end = j;
            }
        }
    }
    if (end >= 4) return vfail(QLD_ERR_INVALID, "genotype %s too long", s);
    if (shuffle(g, out, start, end)) return VDEFER;
    return j;
}

static int map_char(char a)
{
    switch (a) {
    case '0': return 0;
    case '1': return 1;
    case '-': case 'N': return 2;
    case 'A': return 3;
    case 'C': return 4;
    case 'G': return 5;
    case 'T': return 6;
    default: return 7;
    }
}

/* getDataType (:434-459) */
static int data_type(const int *st)
{
    if (st[STATESALL_ - 1] != 0) return -1;
    int acc = 0;
    for (int i = 3; i < STATESALL_; i++) acc += st[i];
    if ((st[0] || st[1]) && acc == 0) return st[2] ? 3 : 2;
    acc = st[0] + st[1];
    if ((st[3] || st[4] || st[5] || st[6]) && acc == 0) return st[2] ? 5 : 4;
    return 0;
}

/* removeNonPolymorphicSite with filterOut == 0 (:469-600): the scan stops
 * at the second distinct state, so an N after it does not drop the site. */
static int non_polymorphic(const char *line, int n, int type)
{
    int set[STATESALL_] = {0}, rm = 1;
    if (type == 2 || type == 3) {
        for (int i = 0; i < n; i++) {
            set[map_char(line[i])]++;
            if ((set[0] != 0) + (set[1] != 0) == 2) { rm = 0; break; }
        }
    } else if (type == 4 || type == 5) {
        for (int i = 0; i < n; i++) {
            set[map_char(line[i])]++;
            if ((set[3] != 0) + (set[4] != 0) + (set[5] != 0) + (set[6] != 0) >= 2) { rm = 0; break; }
        }
    } else {
        return 0;
    }
    if (set[2] != 0) rm = 1;
    return rm;
}

/* binaryDeduction (:602-747) */
static int binary_deduction(vrng *g, int type, int n, char *line)
{
    if (type == 2 || type == 3) return QLD_OK;
    int st[STATESALL_] = {0};
    for (int i = 0; i < n; i++) {
        int m = map_char(line[i]);
        if (m >= 3 && m <= 6) st[m - 3]++;
    }
    /* minorToMajor: sort A,C,G,T counts (a fixed 5-compare network), and
     * with more than two bases send the two minor ones to random majors */
    char alpha[4] = {'A', 'C', 'G', 'T'};
    int pairs[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
    for (int q = 0; q < 5; q++) {
        int a = pairs[q][0], b = pairs[q][1];
        if (st[b] > st[a]) {
            int t = st[a];
            char tc = alpha[a];
            st[a] = st[b];
            alpha[a] = alpha[b];
            st[b] = t;
            alpha[b] = tc;
        }
    }
    if (st[2] + st[3] != 0) {
        if (g->mode == VR_COUNT) {
            g->used += 2;
            return QLD_OK;
        }
        char rep[4];
        int total = st[0] + st[1];
        for (int i = 0; i < 2; ++i) {
            int v;
            if (vdraw(g, &v)) return VDEFER;
            int rv = v % total + 1;
            char major = rv <= st[0] ? alpha[0] : alpha[1];
            rep[i] = alpha[i + 2];
            rep[i + 2] = major;
        }
        for (int i = 0; i < n; i++) {
            if (line[i] == rep[0]) line[i] = rep[2];
            if (line[i] == rep[1]) line[i] = rep[3];
        }
    }
    int cnt[STATESALL_] = {0}, states = 0;
    for (int i = 0; i < n; i++) cnt[map_char(line[i])]++;
    cnt[2] = 0;
    for (int i = 0; i < STATESALL_; i++) states += cnt[i] != 0;
    if (states <= 0) return vfail(QLD_ERR_INVALID, "a site without any base");
    if (states == 2) {
        /* convertAlignmentBIN: first base seen -> '0', second -> '1' */
        char s0 = 'x', s1 = 'x';
        for (int i = 0; i < n; i++) {
            char t = line[i];
            if (s0 == 'x' && t != '-' && t != 'N') s0 = t;
            if (s0 != 'x' && s1 == 'x' && t != s0 && t != '-' && t != 'N') s1 = t;
            line[i] = t == '-' ? '-' : t == 'N' ? 'N' : t == s0 ? '0' : t == s1 ? '1' : 'X';
        }
    }
    return QLD_OK;
}

typedef struct {
    int rows, cap, words;
    uint64_t *snp;
    uint32_t *pos, *bc;
    char **id;
} vtab;

/* compressSnip_x64 (:783-815): '1' -> bit 63 - i%64 of word i/64; returns BC */
static uint32_t pack_row(const char *snip, int n, uint64_t *w, int words)
{
    memset(w, 0, (size_t)words * 8);
    uint32_t bc = 0;
    for (int i = 0; i < n; i++) {
        uint64_t bit = (uint64_t)((unsigned char)snip[i] & 1u);
        w[i / 64] |= bit << (63 - i % 64);
        bc += (uint32_t)bit;
    }
    return bc;
}

static int vtab_push(vtab *t, int pos, const char *id, const uint64_t *row, uint32_t bc)
{
    if (t->rows == t->cap) {
        int nc = t->cap ? t->cap * 2 : 1024;
        uint64_t *ns = (uint64_t *)realloc(t->snp, (size_t)nc * t->words * 8);
        uint32_t *np = (uint32_t *)realloc(t->pos, (size_t)nc * 4);
        uint32_t *nb = (uint32_t *)realloc(t->bc, (size_t)nc * 4);
        char **ni = (char **)realloc(t->id, (size_t)nc * sizeof(char *));
        if (ns) t->snp = ns;
        if (np) t->pos = np;
        if (nb) t->bc = nb;
        if (ni) t->id = ni;
        if (!ns || !np || !nb || !ni) return vfail(QLD_ERR_NOMEM, "out of host memory");
        t->cap = nc;
    }
    memcpy(t->snp + (size_t)t->rows * t->words, row, (size_t)t->words * 8);
    char *idc = strdup(id);
    if (!idc) return vfail(QLD_ERR_NOMEM, "out of host memory");
    t->pos[t->rows] = (uint32_t)pos;
    t->bc[t->rows] = bc;
    t->id[t->rows] = idc;
    t->rows++;
    return QLD_OK;
}

/* One data line: 1 row (the site's valid_count coded bases in snip, POS in
 * *pos_out, ID in id[256]), 0 skipped, <0 error; VDEFER when the draw mode
 * cannot handle it.  The line text is only read. */
static int vcf_line(vrng *g, char *line, const unsigned char *mask, int nmask,
                    int valid_count, int ploidy, int snip_size, char *word, size_t wcap,
                    char *snip, int *pos_out, char *id)
{
    char *p = line;
    int eol = 0, el = -1, skipped = 0, position = -1, gtpos = -1, sref = 0, salt = 0;
    char sv[MAXST];
    id[0] = '\0';
    for (int j = 0; j < MAXST; j++) sv[j] = 'X';
    while (vword(&p, word, wcap, &eol)) {
        el++;
        switch (el) {
        case 1: position = atoi(word); break;
        case 2: snprintf(id, 256, "%s", word); break;
        case 3: {
            if (strlen(word) > 1) { skipped = 1; break; }
            int s = get_states(word, sv, 0, position);
            if (s < 0) return s;
            sref = num_states(sv);
            if (sref == 0) skipped = 1;
            break;
        }
        case 4: {
            if (strlen(word) > 1) { skipped = 1; break; }
            int s = get_states(word, sv, 1, position);
            if (s < 0) return s;
            salt = num_states(sv);
            if (salt != 0 && sv[salt - 1] == '.')
                for (int q = salt - 1; q < salt - 1 + sref; ++q) sv[q] = sv[q - sref];
            if (salt == 0) skipped = 1;
            break;
        }
        case 6:
            if (strcmp("PASS", word) && strcmp(".", word)) skipped = 1;
            break;
        case 7: {
            if (!strcmp(".", word)) break;
            float af;
            int rc = value_af(word, position, &af);
            if (rc) return rc;
            if (af == 0.0f || af == 1.0f) { skipped = 1; break; }
            int vt = vt_is_snp(word);
            if (vt == -2) return vfail(QLD_ERR_INVALID, "VT in position %d has no value", position);
            if (vt == 0) skipped = 1;
            break;
        }
        case 8:
            gtpos = gt_pos(word);
            if (gtpos == -1) skipped = 1;
            break;
        default: break;
        }
        if (el == 8 || eol || skipped) break;
    }
    if (position == -1) return 0;
    if (skipped) return 0;
    int sample = 0, n = 0;
    char part[8];
    while (vword(&p, word, wcap, &eol)) {
        if (sample >= snip_size)
            return vfail(QLD_ERR_INVALID, "more samples than the header's %d in position %d",
                         snip_size, position);
        if (sample < nmask && mask[sample]) {
            int size = gt_field(word, gtpos);
            if (size > ploidy)
                return vfail(QLD_ERR_INVALID, "Detected data bigger than the ploidy of input "
                             "(position %d)", position);
            int got = gt_data(g, word, sv, salt, part);
            if (got < 0) return got;
            if (got != size)
                return vfail(QLD_ERR_INVALID, "genotype %s in position %d is malformed", word,
                             position);
            if (n + size > valid_count) {
                n += size;
                break;
            }
            memcpy(snip + n, part, (size_t)size);
            n += size;
        }
        sample++;
        if (eol) break;
    }
    if (n != valid_count)
        return vfail(QLD_ERR_INVALID, "There are %d nucleotides in position %d. Expected %d.", n,
                     position, valid_count);
    int st[STATESALL_] = {0};
    for (int i = 0; i < n; i++) st[map_char(snip[i])]++;
    int type = data_type(st);
    if (type == -1) return vfail(QLD_ERR_INVALID, "Empty alignment (no states found) in position %d", position);
    /* With an N among three or more bases, whether removeNonPolymorphicSite
     * keeps the site (and binaryDeduction then draws twice) depends on the
     * shuffled order: such a line is not counted ahead. */
    if (g->mode == VR_COUNT && type == 5 &&
        (st[3] != 0) + (st[4] != 0) + (st[5] != 0) + (st[6] != 0) > 2)
        return VDEFER;
    if (non_polymorphic(snip, n, type)) return 0;
    int rc = binary_deduction(g, type, n, snip);
    if (rc) return rc;
    *pos_out = position;
    return 1;
}

int qld_vcf_read_table(qld_vcf_reader *r, char *const *files, int nfiles, int pos_wmin,
                       int pos_wmax, int pos_wset, const char *line1, const char *chrom_line,
                       const unsigned char *mask, int nsamples, int valid_count, int ploidy,
                       int snip_size, qld_snp_table *out)
{
    if (!r || !out || nfiles < 0 || (nfiles && !files) || !line1 || !chrom_line ||
        valid_count <= 0 || ploidy <= 0)
        return vfail(QLD_ERR_INVALID, "invalid argument");
    memset(out, 0, sizeof *out);
    vtab t = {0, 0, (valid_count + 63) / 64, NULL, NULL, NULL, NULL};
    const int W = t.words;
    char *snip = (char *)malloc((size_t)valid_count + 8);
    uint64_t *row = (uint64_t *)malloc((size_t)W * 8);
    char id[256];
    int rc = snip && row ? QLD_OK : vfail(QLD_ERR_NOMEM, "out of host memory");
    /* files are inflated VBATCH at a time on separate threads, then parsed
     * in order; a file that fails to load is loaded again here so that the
     * error (and its message) is the calling thread's */
    enum { VBATCH = 8 };
    char *bufs[VBATCH] = {0};
    size_t lens[VBATCH] = {0};
    int brc[VBATCH] = {0};
    for (int f = 0; f < nfiles && !rc; f++) {
        const int slot = f % VBATCH;
        if (slot == 0) {
            const int nb = nfiles - f < VBATCH ? nfiles - f : VBATCH;
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < nb; b++) brc[b] = vslurp(files[f + b], &bufs[b], &lens[b]);
        }
        char *buf = bufs[slot];
        size_t len = lens[slot];
        bufs[slot] = NULL;
        if (brc[slot] && (rc = vslurp(files[f], &buf, &len))) break;
        size_t wcap = len + 1;
        char *word = (char *)malloc(wcap);
        if (!word) { free(buf); rc = vfail(QLD_ERR_NOMEM, "out of host memory"); break; }
        vcursor c = {buf, buf + len};
        char *line;
        int eof;
        /* the lines readTable_x64's loop hands to writeToTable_x64, in order */
        char **sel = NULL;
        int nsel = 0, cap = 0;
        size_t maxlen = 0;
        if (!vnext_line(&c, &line, &eof)) {
            rc = vfail(QLD_ERR_INVALID, "File %s empty", files[f]);
        } else if (strcmp(line, line1)) {
            rc = vfail(QLD_ERR_INVALID, "Header 1 different in %s", files[f]);
        } else if (!eof && vnext_line(&c, &line, &eof)) {
            if (strcmp(line, chrom_line)) {
                rc = vfail(QLD_ERR_INVALID, "Header 2 different in %s", files[f]);
            } else {
                int pos;
                do {
                    if (vnext_line(&c, &line, &eof)) {
                        char *p = line;
                        int e;
                        if (!vword(&p, word, wcap, &e) || !vword(&p, word, wcap, &e)) {
                            rc = vfail(QLD_ERR_INVALID, "%s: line without CHROM and POS fields",
                                       files[f]);
                            break;
                        }
                        pos = atoi(word);
                        if (pos >= pos_wmin && pos <= pos_wmax) {
                            if (nsel == cap) {
                                int nc = cap ? 2 * cap : 1024;
                                char **ns = (char **)realloc(sel, (size_t)nc * sizeof *sel);
                                if (!ns) { rc = vfail(QLD_ERR_NOMEM, "out of host memory"); break; }
                                sel = ns;
                                cap = nc;
                            }
                            sel[nsel++] = line;
                            size_t l = strlen(line);
                            if (l > maxlen) maxlen = l;
                        }
                    } else {
                        pos = 0;
                    }
                } while (!eof && pos_wset && pos < pos_wmax);
            }
        }
        /* Lines are parsed on every core.  A first pass counts the rand()
         * draws of each line (its row is already exact when it draws none:
         * haploid and phased genotypes at most two bases); the draws of the
         * other lines are then taken from the reader's stream in line order
         * and a second pass replays each line with its own values.  A line
         * the count cannot settle (an N among three or more bases) or that
         * fails stops the batch there and is parsed in order with the stream
         * itself, so the draws, the rows and the first error are those of a
         * sequential read. */
        signed char *st = NULL;
        long *cnt = NULL, *off = NULL;
        uint64_t *rows = NULL;
        uint32_t *rbc = NULL, *rpos = NULL;
        char *rid = NULL;
        if (!rc && nsel) {
            st = (signed char *)malloc((size_t)nsel);
            cnt = (long *)malloc((size_t)nsel * sizeof(long));
            off = (long *)malloc((size_t)nsel * sizeof(long));
            rows = (uint64_t *)malloc((size_t)nsel * W * 8);
            rbc = (uint32_t *)malloc((size_t)nsel * 4);
            rpos = (uint32_t *)malloc((size_t)nsel * 4);
            rid = (char *)malloc((size_t)nsel * 256);
            if (!st || !cnt || !off || !rows || !rbc || !rpos || !rid)
                rc = vfail(QLD_ERR_NOMEM, "out of host memory");
        }
        /* lines per count pass: at most QLD_VCF_AHEAD values drawn ahead */
        const int BL = valid_count >= QLD_VCF_AHEAD / 16 ? 16 : QLD_VCF_AHEAD / valid_count;
        int counted = 0;
        for (int k0 = 0; !rc && k0 < nsel;) {
            if (counted == k0) {
                const int k_end = nsel - k0 < BL ? nsel : k0 + BL;
                /* pass 1: status 0 skipped, 1 row, 2 replay (cnt draws), 3 in order */
#pragma omp parallel
                {
                    char *wb = (char *)malloc(maxlen + 1);
                    char *sb = (char *)malloc((size_t)valid_count + 8);
#pragma omp for schedule(dynamic, 8)
                    for (int k = k0; k < k_end; k++) {
                        vrng g = {VR_COUNT, NULL, NULL, 0, 0};
                        int pk = 0, v = VDEFER;
                        if (wb && sb)
                            v = vcf_line(&g, sel[k], mask, nsamples, valid_count, ploidy,
                                         snip_size, wb, maxlen + 1, sb, &pk,
                                         rid + (size_t)k * 256);
                        cnt[k] = g.used;
                        if (v < 0) {
                            st[k] = 3;
                        } else if (g.used > 0) {
                            st[k] = 2;
                        } else {
                            st[k] = (signed char)v;
                            if (v == 1) {
                                rbc[k] = pack_row(sb, valid_count, rows + (size_t)k * W, W);
                                rpos[k] = (uint32_t)pk;
                            }
                        }
                    }
                    free(wb);
                    free(sb);
                }
                counted = k_end;
            }
            /* lines k0 .. k1-1 need no in-order parse */
            int k1 = k0;
            long total = 0;
            while (k1 < counted && st[k1] != 3) {
                off[k1] = total;
                if (st[k1] == 2) total += cnt[k1];
                k1++;
            }
            if (total > 0) {
                int32_t *vals = (int32_t *)malloc((size_t)total * sizeof(int32_t));
                if (!vals) { rc = vfail(QLD_ERR_NOMEM, "out of host memory"); break; }
                for (long q = 0; q < total; q++) vals[q] = (int32_t)vrand(r);
                int bad = 0;
                /* pass 2: replay the lines that draw */
#pragma omp parallel
                {
                    char *wb = (char *)malloc(maxlen + 1);
                    char *sb = (char *)malloc((size_t)valid_count + 8);
#pragma omp for schedule(dynamic, 8)
                    for (int k = k0; k < k1; k++) {
                        if (st[k] != 2) continue;
                        vrng g = {VR_REPLAY, NULL, vals + off[k], cnt[k], 0};
                        int pk = 0, v = VDEFER;
                        if (wb && sb)
                            v = vcf_line(&g, sel[k], mask, nsamples, valid_count, ploidy,
                                         snip_size, wb, maxlen + 1, sb, &pk,
                                         rid + (size_t)k * 256);
                        if (v < 0 || g.used != g.n) {
                            const int why = (wb && sb) ? 2 : 1;
#pragma omp atomic write
                            bad = why;
                        } else {
                            st[k] = (signed char)v;
                            if (v == 1) {
                                rbc[k] = pack_row(sb, valid_count, rows + (size_t)k * W, W);
                                rpos[k] = (uint32_t)pk;
                            }
                        }
                    }
                    free(wb);
                    free(sb);
                }
                free(vals);
                if (bad) {
                    rc = bad == 1 ? vfail(QLD_ERR_NOMEM, "out of host memory")
                                  : vfail(QLD_ERR_INVALID, "%s: a line drew other rand() values "
                                          "than counted", files[f]);
                    break;
                }
            }
            for (int k = k0; k < k1 && !rc; k++)
                if (st[k] == 1)
                    rc = vtab_push(&t, (int)rpos[k], rid + (size_t)k * 256,
                                   rows + (size_t)k * W, rbc[k]);
            if (!rc && k1 < counted) {
                /* the line the count could not settle, with the stream */
                vrng g = {VR_STREAM, r, NULL, 0, 0};
                int pk = 0;
                int v = vcf_line(&g, sel[k1], mask, nsamples, valid_count, ploidy, snip_size,
                                 word, wcap, snip, &pk, id);
                if (v < 0) rc = v;
                else if (v == 1)
                    rc = vtab_push(&t, pk, id, row, pack_row(snip, valid_count, row, W));
                k1++;
            }
            k0 = k1;
        }
        free(cnt);
        free(off);
        free(st);
        free(rows);
        free(rbc);
        free(rpos);
        free(rid);
        free(sel);
        free(word);
        free(buf);
    }
    for (int b = 0; b < VBATCH; b++) free(bufs[b]);
    free(row);
    free(snip);
    if (!rc) {
        size_t nr = t.rows ? (size_t)t.rows : 1;
        size_t idtot = 0;
        for (int i = 0; i < t.rows; i++) idtot += strlen(t.id[i]) + 1;
        out->rows = t.rows;
        out->words = t.words;
        /* shrink (or, for an empty table, allocate) to the final size; a
         * failed realloc leaves the block with t, which is freed below */
        uint64_t *snp = (uint64_t *)realloc(t.snp, nr * (size_t)t.words * 8);
        if (snp) { out->snp = snp; t.snp = NULL; }
        uint32_t *pos = (uint32_t *)realloc(t.pos, nr * 4);
        if (pos) { out->pos = pos; t.pos = NULL; }
        uint32_t *bc = (uint32_t *)realloc(t.bc, nr * 4);
        if (bc) { out->bc = bc; t.bc = NULL; }
        out->id = (char **)malloc(nr * sizeof(char *));
        out->idbuf = (char *)malloc(idtot ? idtot : 1);
        if (!out->snp || !out->pos || !out->bc || !out->id || !out->idbuf) {
            rc = vfail(QLD_ERR_NOMEM, "out of host memory");
        } else {
            size_t o = 0;
            for (int i = 0; i < t.rows; i++) {
                size_t l = strlen(t.id[i]) + 1;
                memcpy(out->idbuf + o, t.id[i], l);
                out->id[i] = out->idbuf + o;
                o += l;
            }
        }
    }
    for (int i = 0; i < t.rows; i++) free(t.id[i]);
    free(t.id);
    free(t.snp);
    free(t.pos);
    free(t.bc);
    if (rc) qld_snp_table_free(out);
    return rc;
}
