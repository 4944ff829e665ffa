/*
 * qld_compute.c -- `qLD-compute` for the MI355X engine.
 *
 * Same command line, input directories and report files as the reference
 * correlator (correlate_our_kernel.c:146-472 parses, load_balancer.c:122-357
 * prepares a task, input_list.c:205-323 expands -inputList, and
 * correlate_function.c:60-290 loads the tables and correlates), with the
 * pairwise counts and r^2 computed on the GPU through qld_correlate_report().
 * Input is the parser's VCF chunks or, with -mdf, MDF chunks; every task is
 * computed (the reference's thread split
 * can drop tasks, correlate_function.c:53-57; here there is no split).
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "qld_host.h"

typedef struct {
    char input[QLD_PATH_MAX], input2[QLD_PATH_MAX], output[QLD_PATH_MAX];
    int wmin1, wmax1, wmin2, wmax2;
    int in2set, wset1, wset2;
    qld_mdf_header h1, h2;
    char **files1, **files2;
    int nfiles1, nfiles2;
} task;

typedef struct {
    char input[QLD_PATH_MAX], input2[QLD_PATH_MAX], output[QLD_PATH_MAX];
    char list[QLD_PATH_MAX];
    char samples1[QLD_PATH_MAX], samples2[QLD_PATH_MAX];
    int has_samples1, has_samples2;
    int in2set, wset1, wset2, listset, outset, pathset;
    int wmin1, wmax1, wmin2, wmax2;
    int ploidy, threads, mdf, gpu, device, dump, plain_tasks;
    int64_t batch_cells;      /* r^2 cells of one -inputList batch */
    unsigned seed;            /* -seed: srand() seed of the unphased shuffles */
    int seedset;
    int sorted, competing;    /* -sorted -competing: -inputList tasks largest first */
    float r2limit;
    int engine, r2mode;
} options;

static void die(const char *fmt, const char *a)
{
    fprintf(stderr, fmt, a);
    exit(1);
}

static double now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

static void with_slash(char *out, const char *in)
{
    snprintf(out, QLD_PATH_MAX - 1, "%s", in);
    size_t n = strlen(out);
    if (n && out[n - 1] != '/') { out[n] = '/'; out[n + 1] = '\0'; }
}

static void usage(void)
{
    printf("qLD-compute (MI355X engine)\n"
           "  -input DIR          MDF directory of the first region\n"
           "  -input2 DIR         MDF directory of the second region (optional)\n"
           "  -output FILE        report file (with -inputList: FILE stem, one\n"
           "                      report per list line, FILE_<line>.txt)\n"
           "  -ploidy P           haploid | phased_diploid | unphased_diploid\n"
           "  -posWmin1/-posWmax1 POS   window of the first region\n"
           "  -posWmin2/-posWmax2 POS   window of the second region\n"
           "  -inputList FILE     lines input,posWmin1,posWmax1,input2,posWmin2,posWmax2\n"
           "  -r2limit R          report pairs with r2 >= R (default 0.2)\n"
           "  -mdf                MDF input (default: the parser's VCF chunks)\n"
           "  -sampleList FILE    samples of the first input (VCF input)\n"
           "  -sampleList2 FILE   samples of the second input (default: -sampleList)\n"
           "  -gpu                accepted; the engine always runs on the GPU\n"
           "  -seed S             srand() seed of the unphased-genotype shuffles\n"
           "  -sorted -competing  run -inputList tasks largest first (their reports\n"
           "                      keep their line numbers); either alone: no effect\n"
           "  -blis, -rTable, -threads T\n"
           "                      accepted, no effect (MDF files are read on all\n"
           "                      host cores, up to 32)\n"
           "  -engine E           auto | popcount | mfma_fp4 | mfma_i8 (default auto)\n"
           "  -r2mode M           reference (bit-identical to the reference's float\n"
           "                      r2, default) | exact (within 1 ulp of FP64)\n"
           "  -device D           GPU ordinal (default 0)\n"
           "  -dump-tables        write the tables each task would correlate to\n"
           "                      <output>.tables instead of computing (no GPU)\n"
           "  -plain-tasks        run -inputList tasks one by one (no batching)\n"
           "  -batch-cells C      r2 cells of one -inputList batch (default 2^29; a\n"
           "                      task above C/4 runs alone through the streamed call)\n");
}

static int need_arg(int i, int argc) { return i != argc - 1; }

static void parse_args(int argc, char **argv, options *o)
{
    memset(o, 0, sizeof *o);
    o->r2limit = 0.2f;
    o->threads = 1;
    o->batch_cells = (int64_t)1 << 29;
    o->engine = QLD_ENGINE_AUTO;
    o->r2mode = QLD_R2_REFERENCE;
    if (argc < 2) die("qLD-compute needs some arguments to run. Run \"qLD-compute -help\" for more info%s\n", "");
    if (strcmp(argv[1], "-help") && strcmp(argv[1], "-h") && strcmp(argv[1], "--help"))
        printf("qLD-compute running with arguments:\n");
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (!strcmp(a, "-help") || !strcmp(a, "-h") || !strcmp(a, "--help")) {
            usage();
            exit(0);
        }
        if (!strcmp(a, "-input") || !strcmp(a, "-input2")) {
            if (need_arg(i, argc)) {
                printf("\t%s %s\n", a, argv[i + 1]);
                char *dst = !strcmp(a, "-input") ? o->input : o->input2;
                with_slash(dst, argv[++i]);
                DIR *d = opendir(dst);
                if (!d) die("\n ERROR: Directory %s does not exist.\n\n", dst);
                closedir(d);
                if (dst == o->input) o->pathset = 1;
                else o->in2set = 1;
            }
            continue;
        }
        if (!strcmp(a, "-inputList")) {
            if (!need_arg(i, argc)) die("\n ERROR: -inputList needs a file%s\n\n", "");
            printf("\t%s %s\n", a, argv[i + 1]);
            snprintf(o->list, sizeof o->list, "%s", argv[++i]);
            o->listset = 1;
            continue;
        }
        if (!strcmp(a, "-output")) {
            if (need_arg(i, argc)) {
                printf("\t%s %s\n", a, argv[i + 1]);
                snprintf(o->output, sizeof o->output, "%s", argv[++i]);
                o->outset = 1;
            }
            continue;
        }
        if (!strcmp(a, "-sampleList") || !strcmp(a, "-sampleList2")) {
            const int second = !strcmp(a, "-sampleList2");
            if (need_arg(i, argc)) {
                snprintf(second ? o->samples2 : o->samples1, QLD_PATH_MAX, "%s", argv[++i]);
                if (second) o->has_samples2 = 1;
                else o->has_samples1 = 1;
            }
            continue;
        }
        if (!strcmp(a, "-ploidy")) {
            if (!need_arg(i, argc)) die("Please insert supported ploidy (see --help for more info)%s\n", "");
            const char *p = argv[++i];
            if (!strcmp(p, "haploid")) o->ploidy = 1;
            else if (!strcmp(p, "phased_diploid") || !strcmp(p, "unphased_diploid")) o->ploidy = 2;
            else die("Please insert supported ploidy (see --help for more info)%s\n", "");
            continue;
        }
        int *iv = NULL, *set = NULL;
        if (!strcmp(a, "-posWmin1")) { iv = &o->wmin1; set = &o->wset1; }
        else if (!strcmp(a, "-posWmax1")) { iv = &o->wmax1; set = &o->wset1; }
        else if (!strcmp(a, "-posWmin2")) { iv = &o->wmin2; set = &o->wset2; }
        else if (!strcmp(a, "-posWmax2")) { iv = &o->wmax2; set = &o->wset2; }
        if (iv) {
            if (need_arg(i, argc)) { *iv = atoi(argv[++i]); *set = 1; }
            continue;
        }
        if (!strcmp(a, "-r2limit")) {
            if (need_arg(i, argc)) o->r2limit = (float)atof(argv[++i]);
            continue;
        }
        if (!strcmp(a, "-threads")) {
            if (need_arg(i, argc)) ++i;
            continue;
        }
        if (!strcmp(a, "-device")) {
            if (need_arg(i, argc)) o->device = atoi(argv[++i]);
            continue;
        }
        if (!strcmp(a, "-batch-cells")) {
            if (!need_arg(i, argc)) die("\nERROR: -batch-cells needs a value%s\n\n", "");
            long long v = atoll(argv[++i]);
            if (v < 4 || v > ((long long)1 << 40))
                die("\nERROR: -batch-cells %s is out of range (4 .. 2^40)\n\n", argv[i]);
            o->batch_cells = (int64_t)v;
            continue;
        }
        if (!strcmp(a, "-seed")) {
            /* the reference seeds rand() with it (and then stops at its own
             * "not a valid command line parameter", correlate_our_kernel.c:
             * 387-392,433-435); here the seed restarts the VCF reader's
             * stream of unphased shuffles */
            if (!need_arg(i, argc)) die("\nERROR: -seed needs a value%s\n\n", "");
            o->seed = (unsigned)atoi(argv[++i]);
            o->seedset = 1;
            continue;
        }
        if (!strcmp(a, "-engine") && need_arg(i, argc)) {
            const char *e = argv[++i];
            if (!strcmp(e, "auto")) o->engine = QLD_ENGINE_AUTO;
            else if (!strcmp(e, "popcount")) o->engine = QLD_ENGINE_POPCOUNT;
            else if (!strcmp(e, "mfma_fp4")) o->engine = QLD_ENGINE_MFMA_FP4;
            else if (!strcmp(e, "mfma_i8")) o->engine = QLD_ENGINE_MFMA_I8;
            else die("\nERROR: unknown engine %s\n\n", e);
            continue;
        }
        if (!strcmp(a, "-r2mode") && need_arg(i, argc)) {
            const char *e = argv[++i];
            if (!strcmp(e, "reference")) o->r2mode = QLD_R2_REFERENCE;
            else if (!strcmp(e, "exact")) o->r2mode = QLD_R2_EXACT;
            else die("\nERROR: unknown r2 mode %s\n\n", e);
            continue;
        }
        if (!strcmp(a, "-mdf")) { o->mdf = 1; continue; }
        if (!strcmp(a, "-dump-tables")) { o->dump = 1; continue; }
        if (!strcmp(a, "-plain-tasks")) { o->plain_tasks = 1; continue; }
        if (!strcmp(a, "-gpu")) { o->gpu = 1; continue; }
        if (!strcmp(a, "-sorted")) { o->sorted = 1; continue; }
        if (!strcmp(a, "-competing")) { o->competing = 1; continue; }
        if (!strcmp(a, "-blis") || !strcmp(a, "-rTable")) continue;
        die("\nERROR: %s is not a valid command line parameter\n\n", a);
    }
    if (!o->ploidy) die("Please insert supported ploidy (see --help for more info)%s\n", "");
    if (o->listset) {
        o->pathset = 1;
        o->wset1 = 1;
        o->in2set = 1;
        o->wset2 = 1;
    }
    if (!o->pathset) die("\n ERROR: Please specify a path for the input with -input%s\n\n", "");
    if (!o->outset) die("\n ERROR: Please specify an alignment folder with -output%s\n\n", "");
    if (o->has_samples2 && !o->has_samples1)
        die("\n ERROR: Please set sampleList before setting sampleList2%s\n\n", "");
    if (o->has_samples1 && !o->has_samples2) {
        memcpy(o->samples2, o->samples1, sizeof o->samples2);
        o->has_samples2 = 1;
    }
    if (o->mdf && o->has_samples1)
        die("\n ERROR: -sampleList applies to VCF input; MDF rows are already packed for "
            "their samples%s\n\n", "");
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    o->threads = ncpu < 1 ? 1 : ncpu > 32 ? 32 : (int)ncpu;
}

static void check(int rc, const char *what)
{
    if (rc != QLD_OK) {
        fprintf(stderr, "\n ERROR: %s: %s\n\n", what, qld_host_last_error());
        exit(1);
    }
}

/* The header of an input directory, read once per directory: the reference
 * reads it again for every task (a gzip'd file whose #CHROM line names every
 * sample; 2,048 reads took a second for a 1,024-line list), and the files
 * do not change during a run. */
static int read_header(const char *dir, qld_mdf_header *h)
{
    static struct {
        char dir[QLD_PATH_MAX];
        qld_mdf_header h;
    } cache[16];
    static int n;
    for (int i = 0; i < n; i++)
        if (!strcmp(cache[i].dir, dir)) {
            *h = cache[i].h;
            return QLD_OK;
        }
    const int rc = qld_mdf_read_header(dir, h);
    if (rc == QLD_OK && n < 16 && strlen(dir) < sizeof cache[0].dir) {
        strcpy(cache[n].dir, dir);
        cache[n++].h = *h;
    }
    return rc;
}

/* Window and task set-up of preprocess_data (load_balancer.c:122-357). */
static void prepare_task(const options *o, task *t, FILE *rep, int id)
{
    const char *suffix = o->mdf ? ".mdf" : ".vcf.gz";
    check(read_header(t->input, &t->h1), "header of -input");
    if (t->in2set) {
        check(read_header(t->input2, &t->h2), "header of -input2");
        if (t->h1.snip_size != t->h2.snip_size) {
            fprintf(stderr, "\n ERROR: The two inputs have different snip sizes.\n\t%d != %d\n",
                    t->h1.snip_size, t->h2.snip_size);
            exit(1);
        }
    } else {
        t->h2 = t->h1;
    }
    const int pmin = t->h1.pos_min, pmax = t->h1.pos_max;
    const int pmin2 = t->h2.pos_min, pmax2 = t->h2.pos_max;
    if (t->wset1) {
        if (t->wmin1 < pmin || t->wmin1 > pmax) t->wmin1 = pmin;
        if (t->wmax1 < pmin || t->wmax1 > pmax) t->wmax1 = pmax;
        if (t->wset2 && t->in2set) {
            if (t->wmin2 < pmin2 || t->wmin2 > pmax2) t->wmin2 = pmin2;
            if (t->wmax2 < pmin2 || t->wmax2 > pmax2) t->wmax2 = pmax2;
        } else if (t->wset2) {
            if (t->wmin2 < pmin || t->wmin2 > pmax) t->wmin2 = pmin;
            if (t->wmax2 < pmin || t->wmax2 > pmax) t->wmax2 = pmax;
        } else if (t->in2set) {
            t->wset2 = 1;
            t->wmin2 = (t->wmin1 < pmin2 || t->wmin1 > pmax2) ? pmin2 : t->wmin1;
            t->wmax2 = (t->wmax1 < pmin2 || t->wmax1 > pmax2) ? pmax2 : t->wmax1;
        } else {
            if (t->wmin2 < pmin || t->wmin2 > pmax) t->wmin2 = pmin;
            if (t->wmax2 < pmin || t->wmax2 > pmax) t->wmax2 = pmax;
        }
    } else {
        t->wset1 = 1;
        t->wmin1 = pmin;
        t->wmax1 = pmax;
        t->wmin2 = pmin;
        t->wmax2 = pmax;
    }
    if (t->in2set && t->wset1 && t->wset2 && t->wmin1 == t->wmin2 && t->wmax1 == t->wmax2)
        t->wset2 = 0;
    int spf = t->h1.snips_per_file, spf2 = t->h2.snips_per_file;
    if (spf < 0 || spf > t->h1.total_snips) spf = t->h1.total_snips;
    if (t->in2set && (spf2 < 0 || spf2 > t->h2.total_snips)) spf2 = t->h2.total_snips;
    check(qld_find_chunks(t->input, t->h1.id, spf, t->h1.total_snips, t->wmin1, t->wmax1,
                          suffix, &t->files1, &t->nfiles1), "chunks of -input");
    check(qld_find_chunks(t->in2set ? t->input2 : t->input, t->h2.id, spf2, t->h2.total_snips,
                          t->wmin2, t->wmax2, suffix, &t->files2, &t->nfiles2),
          "chunks of the second region");
    if (rep) {
        /* the task line of input_report.txt (input_list.c:43-59) */
        fprintf(rep, "%d)\n\tSamples:  %10d\t%10d\n\tSites:    %10d\t%10d\n", id + 1,
                t->h1.snip_size, t->h2.snip_size, t->wmax1 - t->wmin1, t->wmax2 - t->wmin2);
        fflush(rep);
    }
}

static void strip_ext(char *f)
{
    char *end = f + strlen(f);
    while (end > f && *end != '.' && *end != '\\' && *end != '/') --end;
    if (end > f && *end == '.' && end[-1] != '\\' && end[-1] != '/') *end = '\0';
}

/* -inputList expansion (input_list.c:205-323): one task per valid line. */
static task *list_tasks(options *o, int *ntask, FILE *rep)
{
    FILE *fp = fopen(o->list, "r");
    if (!fp) die("ERROR: Input list file not existing!%s\n", "");
    strip_ext(o->output);
    task *ts = NULL;
    int n = 0, line_no = 0;
    char line[4096];
    while (++line_no) {
        if (!fgets(line, sizeof line, fp)) break;
        char in1[1024], a[1024], b[1024], in2[1024], c[1024], d[1024];
        if (sscanf(line, "%1023[^,],%1023[^,],%1023[^,],%1023[^,],%1023[^,],%1023[^\n]",
                   in1, a, b, in2, c, d) != 6) {
            printf("ABORTING: Line %d ---Invalid structure---\n", line_no);
            continue;
        }
        task t;
        memset(&t, 0, sizeof t);
        with_slash(t.input, in1);
        with_slash(t.input2, in2);
        t.wmin1 = atoi(a);
        t.wmax1 = atoi(b);
        if (t.wmin1 >= t.wmax1) {
            printf("ABORTING: Line %d ---posWmin greater than posWmax---\n", line_no);
            continue;
        }
        t.wmin2 = atoi(c);
        t.wmax2 = atoi(d);
        if (t.wmin2 >= t.wmax2) {
            printf("ABORTING: Line %d ---posWmin2 greater than posWmax2---\n", line_no);
            continue;
        }
        if ((size_t)snprintf(t.output, sizeof t.output, "%s_%d.txt", o->output, line_no) >=
            sizeof t.output) {
            printf("ABORTING: Line %d ---output path too long---\n", line_no);
            continue;
        }
        if (access(t.input, F_OK) == -1 || access(t.input2, F_OK) == -1) {
            printf("ABORTING: Line %d ---Input files do not exist---\n", line_no);
            continue;
        }
        t.in2set = o->in2set;
        t.wset1 = o->wset1;
        t.wset2 = o->wset2;
        prepare_task(o, &t, rep, n);
        task *nt = (task *)realloc(ts, (size_t)(n + 1) * sizeof(task));
        if (!nt) die("ERROR: out of host memory%s\n", "");
        ts = nt;
        ts[n++] = t;
    }
    fclose(fp);
    if (!n) die("ERROR: All files in the list are invalid!%s\n", "");
    *ntask = n;
    return ts;
}

/* -sorted with -inputList sorts the task queue (correlate_our_kernel.c:
 * 589-590, MergeSort and SortedMerge, load_balancer.c:22-120): decreasing
 * chunk-file count over both regions, then decreasing span of the first
 * window, ties in list order -- a stable sort, which the reference's list
 * merge sort is.  Its worker takes tasks from the queue head only with
 * -competing; otherwise it fetches them by id, in list order
 * (correlate_function.c:52-80).  The order matters beyond scheduling: VCF
 * reads draw the unphased shuffles from one rand() stream, task by task. */
static int task_before(const task *a, const task *b)
{
    const int fa = a->nfiles1 + a->nfiles2, fb = b->nfiles1 + b->nfiles2;
    if (fa != fb) return fa > fb;
    return (a->wmax1 - a->wmin1) > (b->wmax1 - b->wmin1);
}

static void sort_tasks(task *ts, int n)
{
    task *tmp = (task *)malloc((size_t)(n > 0 ? n : 1) * sizeof(task));
    if (!tmp) die("ERROR: out of host memory%s\n", "");
    for (int w = 1; w < n; w *= 2) {
        for (int lo = 0; lo < n; lo += 2 * w) {
            int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = task_before(&ts[j], &ts[i]) ? ts[j++] : ts[i++];
            while (i < mid) tmp[k++] = ts[i++];
            while (j < hi) tmp[k++] = ts[j++];
        }
        memcpy(ts, tmp, (size_t)n * sizeof(task));
    }
    free(tmp);
}

static void dump_table(FILE *f, const qld_snp_table *t)
{
    int32_t hdr[2] = {t->rows, t->words};
    fwrite(hdr, sizeof hdr, 1, f);
    fwrite(t->pos, sizeof(uint32_t), (size_t)t->rows, f);
    fwrite(t->bc, sizeof(uint32_t), (size_t)t->rows, f);
    fwrite(t->snp, sizeof(uint64_t), (size_t)t->rows * (size_t)t->words, f);
    for (int r = 0; r < t->rows; r++) fwrite(t->id[r], 1, strlen(t->id[r]) + 1, f);
}

/* Binary dump of one task's tables: "QLDT", int32 single, posWset2,
 * snp_size, then table A and (unless single) table B as int32 rows, words,
 * uint32 pos[rows], uint32 bc[rows], uint64 words[rows*words] and the IDs as
 * NUL-terminated strings. */
static void dump_tables(const task *t, int single, const qld_snp_table *A,
                        const qld_snp_table *B)
{
    char path[QLD_PATH_MAX + 8];
    snprintf(path, sizeof path, "%s.tables", t->output);
    FILE *f = fopen(path, "wb");
    if (!f) die("\n ERROR: cannot create %s\n\n", path);
    int32_t hdr[3] = {single, t->wset2, t->h1.snip_size};
    fwrite("QLDT", 4, 1, f);
    fwrite(hdr, sizeof hdr, 1, f);
    dump_table(f, A);
    if (!single) dump_table(f, B);
    if (fclose(f) != 0) die("\n ERROR: write error on %s\n\n", path);
}

/* Every task reads through one reader: chunk files shared by tasks are read
 * and parsed once. */
static qld_mdf_reader *g_reader;

/* The tables of one task, loaded as correlate_function.c:60-240 does. */
typedef struct {
    qld_snp_table A, B;
    int single;     /* one region: B is A, only j > i */
    int upper;      /* the report prints j > i only (second window unset) */
} loaded;

static void vcheck(int rc, const char *what)
{
    if (rc != QLD_OK) {
        fprintf(stderr, "\n ERROR: %s: %s\n\n", what, qld_vcf_last_error());
        exit(1);
    }
}

/* The VCF reader: one rand() stream for the whole run, drawn in the
 * reference's order (tasks in order, table A before table B). */
static qld_vcf_reader *g_vcf;

static void load_vcf(const options *o, const task *t, int in2set, loaded *L)
{
    char *l1a, *cla, *l1b, *clb;
    vcheck(qld_vcf_header_lines(t->h1.header_path, &l1a, &cla), "header of -input");
    vcheck(qld_vcf_header_lines((in2set ? t->h2 : t->h1).header_path, &l1b, &clb),
           "header of the second region");
    unsigned char *m1, *m2;
    int n1, n2, v1, v2;
    vcheck(qld_vcf_sample_mask(cla, o->has_samples1 ? o->samples1 : NULL, o->ploidy, &m1, &n1,
                               &v1), "samples of -input");
    vcheck(qld_vcf_sample_mask(clb, o->has_samples2 ? o->samples2 : NULL, o->ploidy, &m2, &n2,
                               &v2), "samples of the second region");
    if (v1 <= 0 || v2 <= 0) die("\n ERROR: no sample selected%s\n\n", "");
    vcheck(qld_vcf_read_table(g_vcf, t->files1, t->nfiles1, t->wmin1, t->wmax1, t->wset1, l1a,
                              cla, m1, n1, v1, o->ploidy, t->h1.snip_size, &L->A),
           "reading the first region");
    if (in2set || t->wset2)
        vcheck(qld_vcf_read_table(g_vcf, t->files2, t->nfiles2, t->wmin2, t->wmax2, t->wset2,
                                  l1b, clb, m2, n2, v2, o->ploidy,
                                  (in2set ? t->h2 : t->h1).snip_size, &L->B),
               "reading the second region");
    free(l1a);
    free(cla);
    free(l1b);
    free(clb);
    free(m1);
    free(m2);
}

static void load_task(const options *o, const task *t, loaded *L)
{
    memset(L, 0, sizeof *L);
    int in2set = t->in2set && strcmp(t->input, t->input2) != 0;
    L->single = !t->wset2 && !in2set;
    L->upper = !t->wset2;
    if (!o->mdf) {
        load_vcf(o, t, in2set, L);
    } else {
        const int words1 = (t->h1.samples * o->ploidy + 63) / 64;
        const int words2 = (t->h2.samples * o->ploidy + 63) / 64;
        check(qld_mdf_reader_table(g_reader, t->files1, t->nfiles1, t->wmin1, t->wmax1,
                                   t->wset1, words1, o->threads, &L->A),
              "reading the first region");
        if (in2set)
            check(qld_mdf_reader_table(g_reader, t->files2, t->nfiles2, t->wmin2, t->wmax2,
                                       t->wset2, words2, o->threads, &L->B),
                  "reading the second region");
        else if (t->wset2)
            check(qld_mdf_reader_table(g_reader, t->files2, t->nfiles2, t->wmin2, t->wmax2,
                                       t->wset2, words1, o->threads, &L->B),
                  "reading the second region");
    }
    if (!L->single && L->B.words != L->A.words) {
        fprintf(stderr, "\n ERROR: the two regions pack %d and %d words per SNP\n\n",
                L->A.words, L->B.words);
        exit(1);
    }
}

static const qld_snp_table *table_b(const loaded *L) { return L->single ? &L->A : &L->B; }

static void free_loaded(loaded *L)
{
    qld_snp_table_free(&L->A);
    qld_snp_table_free(&L->B);
}

static void gpu_check(qld_ctx *ctx, int rc, const char *what)
{
    if (rc != QLD_OK) {
        fprintf(stderr, "\n ERROR: %s failed (%d): %s\n\n", what, rc, qld_gpu_last_error(ctx));
        exit(1);
    }
}

/* One loaded task through the drop-in correlate call
 * (correlate_function.c:239-290); frees its tables. */
static void run_loaded(qld_ctx *ctx, const options *o, task *t, loaded *L)
{
    const qld_snp_table *b = table_b(L);
    if (o->dump) {
        dump_tables(t, L->single, &L->A, b);
        free_loaded(L);
        return;
    }
    FILE *fp = fopen(t->output, "wb");
    if (!fp) die("\n ERROR: cannot create output file %s\n\n", t->output);
    gpu_check(ctx, qld_correlate_report(ctx, fp, L->A.snp, L->A.pos, L->A.id, L->A.bc,
                                        L->A.rows, b->snp, b->pos, b->id, b->bc, b->rows,
                                        L->A.words, t->h1.snip_size, t->wset2, o->r2limit),
              "GPU correlate");
    if (fclose(fp) != 0) die("\n ERROR: write error on %s\n\n", t->output);
    printf("%s: %d x %d SNPs\n", t->output, L->A.rows, b->rows);
    free_loaded(L);
}

static void run_task(qld_ctx *ctx, const options *o, task *t)
{
    loaded L;
    load_task(o, t, &L);
    run_loaded(ctx, o, t, &L);
}

/* ---- -inputList tasks as one batched launch ------------------------------
 * The reference runs every list line as an independent correlate() call
 * (input_list.c:205-323 queues them; correlate_function.c runs one GEMM per
 * task).  Here the tables of a batch of tasks are stacked into one resident
 * table, every task becomes one pair window of a persistent plan (one kernel
 * launch per tile shape for the whole batch), and each task's r^2 block is
 * compacted on the device before its report is written. */


static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) die("ERROR: out of host memory%s\n", "");
    return p;
}

static int64_t task_cells(const loaded *L)
{
    return (int64_t)L->A.rows * table_b(L)->rows;
}

static void run_batch(qld_ctx *ctx, const options *o, task *const *bt, loaded *L, int nb)
{
    const int words = L[0].A.words, snp_size = bt[0]->h1.snip_size;
    int64_t rows = 0, cells = 0, max_rows_b = 0, max_cells = 0;
    for (int k = 0; k < nb; k++) {
        rows += L[k].A.rows + (L[k].single ? 0 : L[k].B.rows);
        cells += task_cells(&L[k]);
        if (table_b(&L[k])->rows > max_rows_b) max_rows_b = table_b(&L[k])->rows;
        if (task_cells(&L[k]) > max_cells) max_cells = task_cells(&L[k]);
    }
    qld_table *tab = NULL;
    qld_plan *plan = NULL;
    void *dout = NULL, *dcnt = NULL, *doff = NULL, *dcols = NULL, *dr2 = NULL;
    qld_pair_window *w = (qld_pair_window *)xmalloc((size_t)nb * sizeof *w);
    if (rows > 0) {
        uint64_t *snp = (uint64_t *)xmalloc((size_t)rows * words * 8);
        uint32_t *bc = (uint32_t *)xmalloc((size_t)rows * 4);
        int64_t r = 0, off = 0;
        for (int k = 0; k < nb; k++) {
            const qld_snp_table *B = table_b(&L[k]);
            memset(&w[k], 0, sizeof w[k]);
            w[k].a_start = (int32_t)r;
            w[k].a_size = L[k].A.rows;
            r += L[k].A.rows;
            if (L[k].single) {
                w[k].b_start = w[k].a_start;
                w[k].tri = 1;
            } else {
                w[k].b_start = (int32_t)r;
                r += B->rows;
            }
            w[k].b_size = B->rows;
            w[k].out_offset = off;
            off += task_cells(&L[k]);
        }
        /* the tasks' rows side by side (overlapping windows repeat theirs:
         * about a gigabyte for the benchmark's list), copied in parallel */
#pragma omp parallel for schedule(dynamic, 4)
        for (int k = 0; k < nb; k++) {
            const qld_snp_table *A = &L[k].A, *B = table_b(&L[k]);
            memcpy(snp + (int64_t)w[k].a_start * words, A->snp, (size_t)A->rows * words * 8);
            memcpy(bc + w[k].a_start, A->bc, (size_t)A->rows * 4);
            if (!L[k].single) {
                memcpy(snp + (int64_t)w[k].b_start * words, B->snp, (size_t)B->rows * words * 8);
                memcpy(bc + w[k].b_start, B->bc, (size_t)B->rows * 4);
            }
        }
        gpu_check(ctx, qld_table_create(ctx, snp, bc, (int)rows, words, snp_size, &tab),
                  "table upload");
        free(snp);
        free(bc);
        gpu_check(ctx, qld_dev_alloc(ctx, (size_t)(cells ? cells : 1) * 4, &dout), "allocation");
        gpu_check(ctx, qld_plan_pairs(ctx, tab, w, nb, (float *)dout, QLD_BATCH_PERSISTENT,
                                      &plan), "batch plan");
        gpu_check(ctx, qld_plan_run(ctx, plan), "batch run");
        gpu_check(ctx, qld_sync(ctx), "batch run");
        gpu_check(ctx, qld_dev_alloc(ctx, (size_t)(max_rows_b + 1) * 4, &dcnt), "allocation");
    }
    /* Selection: every task's hits go into one device buffer (offsets per
     * task), flushed when the next task's would not fit; a flush copies the
     * buffer to the host in one transfer and formats and writes the pending
     * tasks' reports in parallel, one task per thread (tasks are a few
     * thousand lines each: too few to split one over the threads). */
    int64_t off_rows = 0;
    int64_t *obase = (int64_t *)xmalloc((size_t)(nb ? nb : 1) * sizeof(int64_t));
    for (int k = 0; k < nb; k++) {
        obase[k] = off_rows;
        off_rows += table_b(&L[k])->rows + 1;
    }
    const uint64_t cap = (uint64_t)(max_cells > (1 << 26) ? max_cells : (1 << 26));
    if (rows > 0) {
        gpu_check(ctx, qld_dev_alloc(ctx, (size_t)off_rows * 8, &doff), "allocation");
        gpu_check(ctx, qld_dev_alloc(ctx, (size_t)cap * 4, &dcols), "allocation");
        gpu_check(ctx, qld_dev_alloc(ctx, (size_t)cap * 4, &dr2), "allocation");
    }
    uint64_t *hoff = (uint64_t *)xmalloc((size_t)(off_rows ? off_rows : 1) * 8);
    uint64_t *hbase = (uint64_t *)xmalloc((size_t)(nb ? nb : 1) * sizeof(uint64_t));
    uint64_t hcap = 0, used = 0;
    uint32_t *hcols = NULL;
    float *hr2 = NULL;
    int first = 0;      /* the pending tasks are first .. k-1 */
    for (int k = 0; k <= nb; k++) {
        uint64_t total = 0;
        int rc = QLD_OK;
        if (k < nb) {
            const qld_snp_table *A = &L[k].A, *B = table_b(&L[k]);
            hbase[k] = used;
            if (task_cells(&L[k]) > 0)
                rc = qld_compact_dev(ctx, (float *)dout + w[k].out_offset, A->rows, B->rows,
                                     A->rows, 0, 0, L[k].upper, o->r2limit, (uint32_t *)dcnt,
                                     (uint64_t *)doff + obase[k], (uint32_t *)dcols + used,
                                     (float *)dr2 + used, cap - used, &total);
            if (rc != QLD_OK && rc != QLD_ERR_BUDGET) gpu_check(ctx, rc, "report selection");
        }
        if (k == nb || rc == QLD_ERR_BUDGET) {
            /* flush tasks first .. k-1 */
            if (k > first) {
                const int64_t r0 = obase[first];
                const int64_t r1 = k < nb ? obase[k] : off_rows;
                if (doff)
                    gpu_check(ctx, qld_d2h(ctx, hoff + r0, (uint64_t *)doff + r0,
                                           (size_t)(r1 - r0) * 8), "D2H");
                if (used > hcap) {
                    free(hcols);
                    free(hr2);
                    hcap = used + used / 4;
                    hcols = (uint32_t *)xmalloc(hcap * 4);
                    hr2 = (float *)xmalloc(hcap * 4);
                }
                if (used) {
                    gpu_check(ctx, qld_d2h(ctx, hcols, dcols, used * 4), "D2H");
                    gpu_check(ctx, qld_d2h(ctx, hr2, dr2, used * 4), "D2H");
                }
                int failed = -1;
#pragma omp parallel for schedule(dynamic, 1)
                for (int t = first; t < k; t++) {
                    const qld_snp_table *A = &L[t].A, *B = table_b(&L[t]);
                    FILE *fp = fopen(bt[t]->output, "wb");
                    int ok = fp != NULL;
                    if (ok) qld_write_results_header(fp);
                    if (ok && task_cells(&L[t]) > 0 && hoff[obase[t] + B->rows] > 0)
                        ok = qld_write_hits(fp, A->pos, A->id, A->bc, A->rows, B->pos, B->id,
                                            B->bc, B->rows, snp_size, 0, B->rows,
                                            hoff + obase[t], hcols + hbase[t],
                                            hr2 + hbase[t]) >= 0;
                    if (fp && fclose(fp) != 0) ok = 0;
                    if (!ok) {
#pragma omp critical
                        if (failed < 0 || t < failed) failed = t;
                    }
                }
                if (failed >= 0)
                    die("\n ERROR: cannot write output file %s\n\n", bt[failed]->output);
            }
            if (k == nb) break;
            first = k;
            used = 0;
            k--;        /* the task that did not fit, again at the buffer's start */
            continue;
        }
        used += total;
    }
    free(hoff);
    free(hbase);
    free(obase);
    free(hcols);
    free(hr2);
    free(w);
    if (plan) qld_plan_free(ctx, plan);
    if (tab) qld_table_free(ctx, tab);
    void *bufs[5] = {dout, dcnt, doff, dcols, dr2};
    for (int k = 0; k < 5; k++)
        if (bufs[k]) qld_dev_free(ctx, bufs[k]);
}

/* Tasks in list order; consecutive tasks with the same packing go into one
 * batch until -batch-cells r^2 cells (2^29 = 2 GiB by default); a task above
 * a quarter of that, or with 2^24 samples or more, runs alone through the
 * streamed drop-in call. */
static void run_list(qld_ctx *ctx, const options *o, task *tasks, int ntask)
{
    task **bt = (task **)xmalloc((size_t)ntask * sizeof(task *));
    loaded *L = (loaded *)xmalloc((size_t)ntask * sizeof(loaded));
    int nb = 0;
    int64_t cells = 0;
    double t0 = now();
    for (int i = 0; i <= ntask; i++) {
        loaded cur;
        int flush = i == ntask;
        if (!flush) {
            load_task(o, &tasks[i], &cur);
            /* resident tables feed the MFMA engines: a sample count of 2^24
             * or more (or rows of more than 2^18 words) runs alone, through
             * the streamed call and the popcount engine */
            if (task_cells(&cur) > o->batch_cells / 4 || tasks[i].h1.snip_size >= (1 << 24) ||
                cur.A.words > (1 << 18)) {
                /* every task's tables are read once and in list order: a
                 * VCF read draws the unphased shuffles from the run's
                 * rand() stream */
                if (nb) {
                    run_batch(ctx, o, bt, L, nb);
                    for (int k = 0; k < nb; k++) free_loaded(&L[k]);
                    nb = 0;
                    cells = 0;
                }
                run_loaded(ctx, o, &tasks[i], &cur);
                continue;
            }
            flush = nb && (cur.A.words != L[0].A.words ||
                           tasks[i].h1.snip_size != bt[0]->h1.snip_size ||
                           cells + task_cells(&cur) > o->batch_cells);
        }
        if (flush && nb) {
            run_batch(ctx, o, bt, L, nb);
            for (int k = 0; k < nb; k++) free_loaded(&L[k]);
            nb = 0;
            cells = 0;
        }
        if (i < ntask) {
            bt[nb] = &tasks[i];
            L[nb++] = cur;
            cells += task_cells(&cur);
        }
    }
    printf("%d tasks in %.3fs\n", ntask, now() - t0);
    free(bt);
    free(L);
}

int main(int argc, char **argv)
{
    options o;
    parse_args(argc, argv, &o);
    g_reader = qld_mdf_reader_create();
    g_vcf = qld_vcf_reader_create();
    if (!g_reader || !g_vcf) die("ERROR: out of host memory%s\n", "");
    if (o.seedset) qld_vcf_reader_seed(g_vcf, o.seed);
    if (o.gpu) printf("GPU kernel active\n");
    if (o.mdf) printf("MDF input active\n");

    FILE *rep = fopen("input_report.txt", "w");
    task *tasks;
    int ntask = 0;
    if (o.listset) {
        tasks = list_tasks(&o, &ntask, rep);
        if (o.sorted && o.competing) sort_tasks(tasks, ntask);
    } else {
        tasks = (task *)calloc(1, sizeof(task));
        if (!tasks) die("ERROR: out of host memory%s\n", "");
        snprintf(tasks[0].input, QLD_PATH_MAX, "%s", o.input);
        snprintf(tasks[0].input2, QLD_PATH_MAX, "%s", o.input2);
        snprintf(tasks[0].output, QLD_PATH_MAX, "%s", o.output);
        tasks[0].in2set = o.in2set;
        tasks[0].wset1 = o.wset1;
        tasks[0].wset2 = o.wset2;
        tasks[0].wmin1 = o.wmin1;
        tasks[0].wmax1 = o.wmax1;
        tasks[0].wmin2 = o.wmin2;
        tasks[0].wmax2 = o.wmax2;
        prepare_task(&o, &tasks[0], rep, 0);
        ntask = 1;
    }
    if (rep) fclose(rep);

    qld_cfg cfg;
    qld_cfg_default(&cfg);
    cfg.device = o.device;
    cfg.engine = o.engine;
    cfg.r2_mode = o.r2mode;
    qld_ctx *ctx = NULL;
    if (!o.dump && qld_gpu_init(&ctx, &cfg) != QLD_OK) {
        fprintf(stderr, "\n ERROR: GPU init failed: %s\n\n", qld_gpu_last_error(NULL));
        return 1;
    }
    double t0 = now();
    if (o.listset && !o.dump && !o.plain_tasks) {
        run_list(ctx, &o, tasks, ntask);
    } else {
        for (int i = 0; i < ntask; i++) run_task(ctx, &o, &tasks[i]);
    }
    for (int i = 0; i < ntask; i++) {
        qld_mdf_free_files(tasks[i].files1, tasks[i].nfiles1);
        qld_mdf_free_files(tasks[i].files2, tasks[i].nfiles2);
    }
    if (ctx) qld_gpu_release(ctx);
    printf("Finished in: %.3fs\n", now() - t0);
    printf("Output saved in: %s\n", dirname(o.output));
    free(tasks);
    qld_mdf_reader_free(g_reader);
    qld_vcf_reader_free(g_vcf);
    return 0;
}
