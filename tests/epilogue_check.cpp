// Host check of the REFERENCE r^2 epilogue (qld_amd/csrc/ld_epilogue.h, the
// header the kernels compile) against the oracle's qo_r2_reference, bit for
// bit: every (a, b, c) for N <= 64, then random triples for large N -- the
// benchmark's sample counts, the largest accepted one and random N -- with
// c drawn uniformly, at the independence point c ~ ab/N (cancellation in
// p12 - p1 p2) and at the extremes of its range.  Prints the counts; the
// test asserts no mismatch.  Built by tests/test_epilogue_host.py with
// -ffp-contract=off, so host float and double operations round like the
// device code's.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "ld_epilogue.h"

extern "C" void qo_r2_reference(const uint32_t *bcA, const uint32_t *bcB,
                                const uint32_t *counts, int m, int n, int snp_size,
                                int tri, float *results);

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64()
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t below(uint64_t n) { return (uint32_t)(next_u64() % n); }

struct Tally {
    long long pairs = 0, mismatches = 0, slow = 0, needed = 0;
};

// One batch: columns a[j], rows b[i], counts c[i*m + j] (valid for N).
static void check_batch(int N, const std::vector<uint32_t> &a, const std::vector<uint32_t> &b,
                        const std::vector<uint32_t> &c, Tally &t)
{
    const int m = (int)a.size(), n = (int)b.size();
    std::vector<float> want((size_t)m * n);
    qo_r2_reference(a.data(), b.data(), c.data(), m, n, N, 0, want.data());
    const float fN = (float)N;
    const double invN = 1.0 / (double)N;
    std::vector<RefTerm> ta(m), tb(n);
    for (int j = 0; j < m; ++j) ta[j] = ref_term(a[j], fN);
    for (int i = 0; i < n; ++i) tb[i] = ref_term(b[i], fN);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            const size_t at = (size_t)i * m + j;
            const float cf = (float)c[at];
            bool slow = false;
            const float fast = r2_reference_fast(cf, ta[j], tb[i], invN, slow);
            const float got = r2_reference(cf, ta[j], tb[i], invN);
            uint32_t gb, wb, fb;
            memcpy(&gb, &got, 4);
            memcpy(&wb, &want[at], 4);
            memcpy(&fb, &fast, 4);
            t.pairs++;
            t.slow += slow;
            t.needed += (fb != wb);
            if (gb != wb) {
                if (t.mismatches < 10)
                    fprintf(stderr, "mismatch N=%d a=%u b=%u c=%u got %08x want %08x\n", N, a[j],
                            b[i], c[at], gb, wb);
                t.mismatches++;
            }
        }
}

static uint32_t draw_c(int N, uint32_t a, uint32_t b, int kind)
{
    const uint32_t lo = a + b > (uint32_t)N ? a + b - (uint32_t)N : 0u;
    const uint32_t hi = a < b ? a : b;
    if (kind == 1) {   // near independence: c ~ ab/N
        const double e = (double)a * (double)b / (double)N;
        int64_t c = (int64_t)e + (int64_t)below(5) - 2;
        if (c < (int64_t)lo) c = lo;
        if (c > (int64_t)hi) c = hi;
        return (uint32_t)c;
    }
    if (kind == 2) return below(2) ? lo + below(3 < hi - lo + 1 ? 3 : hi - lo + 1)
                                   : hi - below(3 < hi - lo + 1 ? 3 : hi - lo + 1);
    return lo + below((uint64_t)(hi - lo) + 1);
}

static void random_batches(int N, int batches, Tally &t)
{
    const int m = 64, n = 64;
    std::vector<uint32_t> a(m), b(n), c((size_t)m * n);
    for (int k = 0; k < batches; ++k) {
        for (int j = 0; j < m; ++j) a[j] = below((uint64_t)N + 1);
        for (int i = 0; i < n; ++i) b[i] = below((uint64_t)N + 1);
        if (k % 7 == 0) { a[0] = 0; a[1] = (uint32_t)N; b[0] = 0; b[1] = (uint32_t)N; }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) c[(size_t)i * m + j] = draw_c(N, a[j], b[i], k % 3);
        check_batch(N, a, b, c, t);
    }
}

int main(int argc, char **argv)
{
    const int batches = argc > 1 ? atoi(argv[1]) : 200;
    Tally t;
    // every (a, b, c) for N <= 64
    for (int N = 1; N <= 64; ++N) {
        std::vector<uint32_t> a(N + 1), b(1), c(N + 1);
        for (int j = 0; j <= N; ++j) a[j] = (uint32_t)j;
        for (int bb = 0; bb <= N; ++bb) {
            b[0] = (uint32_t)bb;
            const int cmax = N;   // counts above min(a, b) kept out below
            for (int cc = 0; cc <= cmax; ++cc) {
                std::vector<uint32_t> aa, cv;
                for (int j = 0; j <= N; ++j) {
                    const int lo = j + bb > N ? j + bb - N : 0, hi = j < bb ? j : bb;
                    if (cc >= lo && cc <= hi) {
                        aa.push_back((uint32_t)j);
                        cv.push_back((uint32_t)cc);
                    }
                }
                if (!aa.empty()) check_batch(N, aa, b, cv, t);
            }
        }
    }
    printf("exhaustive pairs %lld mismatches %lld slow %lld needed %lld\n", t.pairs,
           t.mismatches, t.slow, t.needed);
    Tally r;
    static const int sizes[] = {100, 1000, 16622, 22554, 100000, 200000, (1 << 24) - 1};
    for (int N : sizes) random_batches(N, batches, r);
    for (int k = 0; k < 16; ++k) random_batches(2 + (int)below((1u << 24) - 3), batches / 4, r);
    printf("random pairs %lld mismatches %lld slow %lld needed %lld\n", r.pairs, r.mismatches,
           r.slow, r.needed);
    return t.mismatches || r.mismatches ? 1 : 0;
}
