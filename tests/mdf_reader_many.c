/*
 * Driver for test_cli.py::test_mdf_reader_many_files_asan: reads the MDF
 * files named on the command line through one caching reader -- all of them
 * in one call, then overlapping groups across calls under a small cache
 * limit -- and prints a checksum of every table (rows with POS >= 1: the
 * #CHROM line of each file, read as a row with POS 0, stays out), so the
 * test can compare a build with AddressSanitizer against one without it.
 *
 *   mdf_reader_many WORDS LIMIT FILE...
 */
#include <stdio.h>
#include <stdlib.h>

#include "qld_host.h"

static unsigned long long table_sum(const qld_snp_table *t)
{
    unsigned long long h = 1469598103934665603ull;
    for (long r = 0; r < t->rows; r++) {
        h = (h ^ t->pos[r]) * 1099511628211ull;
        h = (h ^ t->bc[r]) * 1099511628211ull;
        for (int w = 0; w < t->words; w++)
            h = (h ^ t->snp[(size_t)r * t->words + w]) * 1099511628211ull;
    }
    return h;
}

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    const int words = atoi(argv[1]);
    const size_t limit = (size_t)strtoull(argv[2], NULL, 10);
    char **files = argv + 3;
    const int n = argc - 3;
    qld_mdf_reader *r = qld_mdf_reader_create();
    if (!r) return 3;
    qld_snp_table t;
    /* every file in one call: the registry grows past 64 and 128 entries */
    if (qld_mdf_reader_table(r, files, n, 1, 2000000000, 1, words, 4, &t)) {
        fprintf(stderr, "%s\n", qld_host_last_error());
        return 4;
    }
    printf("all %d %llu\n", t.rows, table_sum(&t));
    qld_snp_table_free(&t);
    /* overlapping groups of 70 files, twice over, under the cache limit */
    qld_mdf_reader_set_limit(r, limit);
    for (int pass = 0; pass < 2; pass++)
        for (int a = 0; a < n; a += 35) {
            const int k = a + 70 <= n ? 70 : n - a;
            if (qld_mdf_reader_table(r, files + a, k, 1, 2000000000, 1, words, 4, &t)) {
                fprintf(stderr, "%s\n", qld_host_last_error());
                return 5;
            }
            printf("group %d %d %d %llu %zu\n", a, k, t.rows, table_sum(&t),
                   qld_mdf_reader_bytes(r));
            qld_snp_table_free(&t);
        }
    qld_mdf_reader_free(r);
    return 0;
}
