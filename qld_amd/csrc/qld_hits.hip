// Report selection on the device: r^2 blocks are compacted to the pairs the
// report prints, and only those cross PCIe.
//
// The reference writes every pair of the result matrix through one loop on
// the host (writeResults, correlate_IO.c:384-429): rows i of table B,
// columns j of table A (j > i when the second window is not set), a line
// when r^2 >= r2limit and r^2 < 100.  Here the matrix is produced block by
// block of B rows in HBM; three small kernels select the same pairs in the
// same order (count per row, exclusive scan, ordered write of (j, r^2)), the
// hits of a block are copied to pinned host memory on a second stream, and
// the caller's callback formats block b while the GPU computes block b+1.
#include "qld_internal.h"
#include "qld_ctx.h"

#include <algorithm>

namespace {

constexpr int HIT_WG = 256;

__device__ __forceinline__ bool selected(float v, float lim)
{
    return v >= lim && v < 100.0f;      // NaN and the 123.456 marker fail
}

// First column of row r that the report may print.
__device__ __forceinline__ int first_col(int r, int cols, int row0, int col0, int upper)
{
    if (!upper) return 0;
    const long long c = (long long)row0 + r + 1 - col0;
    return (int)(c < 0 ? 0 : (c > cols ? cols : c));
}

__global__ __launch_bounds__(HIT_WG) void hits_count_kernel(
    const float *__restrict__ res, int64_t ldr, int cols, int row0, int col0, int upper,
    float lim, uint32_t *__restrict__ counts)
{
    const int r = blockIdx.x;
    const float *row = res + (int64_t)r * ldr;
    uint32_t n = 0;
    for (int c = first_col(r, cols, row0, col0, upper) + threadIdx.x; c < cols; c += HIT_WG)
        n += selected(row[c], lim) ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o, 64);
    __shared__ uint32_t part[HIT_WG / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < HIT_WG / 64; w++) t += part[w];
        counts[r] = t;
    }
}

// Exclusive scan of `rows` counts into offsets[0..rows] (one workgroup).
__global__ __launch_bounds__(1024) void hits_scan_kernel(const uint32_t *__restrict__ counts,
                                                         int rows, uint64_t *__restrict__ offsets)
{
    __shared__ unsigned long long wsum[16];
    __shared__ unsigned long long carry_s;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < rows; base += 1024) {
        const int i = base + threadIdx.x;
        const unsigned long long v = i < rows ? counts[i] : 0ull;
        unsigned long long x = v;
        for (int o = 1; o < 64; o <<= 1) {
            unsigned long long y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        unsigned long long pre = carry_s;
        for (int w = 0; w < wv; w++) pre += wsum[w];
        if (i < rows) offsets[i] = pre + x - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = pre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[rows] = carry_s;
}

// Ordered write: the row is walked in chunks of 4 x 256 columns; column
// c = base + e*256 + t belongs to thread t, so lane order, wave order and e
// order together are column order, and ballot prefix counts place each hit.
__global__ __launch_bounds__(HIT_WG) void hits_write_kernel(
    const float *__restrict__ res, int64_t ldr, int cols, int row0, int col0, int upper,
    float lim, const uint64_t *__restrict__ offsets, uint32_t *__restrict__ out_cols,
    float *__restrict__ out_r2)
{
    constexpr int NW = HIT_WG / 64;
    const int r = blockIdx.x;
    const float *row = res + (int64_t)r * ldr;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __shared__ uint32_t tot[2][4][NW];
    uint64_t pos = offsets[r];
    int buf = 0;
    for (int base = first_col(r, cols, row0, col0, upper); base < cols; base += 4 * HIT_WG) {
        float v[4];
        bool p[4];
        uint32_t pre[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int c = base + e * HIT_WG + threadIdx.x;
            v[e] = c < cols ? row[c] : 0.0f;
            p[e] = c < cols && selected(v[e], lim);
            const unsigned long long mask = __ballot(p[e]);
            pre[e] = (uint32_t)__popcll(mask & below);
            if (lane == 0) tot[buf][e][wv] = (uint32_t)__popcll(mask);
        }
        __syncthreads();
        uint32_t run = 0;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            uint32_t before = run;
            for (int w = 0; w < wv; w++) before += tot[buf][e][w];
            if (p[e]) {
                const uint64_t o = pos + before + pre[e];
                out_cols[o] = (uint32_t)(col0 + base + e * HIT_WG + threadIdx.x);
                out_r2[o] = v[e];
            }
            for (int w = 0; w < NW; w++) run += tot[buf][e][w];
        }
        pos += run;
        buf ^= 1;       // the next chunk's totals go to the other buffer
    }
}


}  // namespace

// Enqueue count + scan for a block on `s`.
static hipError_t enqueue_count_scan(const float *res, int64_t ldr, int rows, int cols, int row0,
                                     int col0, int upper, float lim, uint32_t *counts,
                                     uint64_t *offsets, hipStream_t s)
{
    // Rows with no columns still get their zero count from the kernel.
    if (rows > 0)
        hipLaunchKernelGGL(hits_count_kernel, dim3(rows), dim3(HIT_WG), 0, s, res, ldr, cols,
                           row0, col0, upper, lim, counts);
    hipLaunchKernelGGL(hits_scan_kernel, dim3(1), dim3(1024), 0, s, counts, rows, offsets);
    return hipGetLastError();
}

extern "C" int qld_compact_dev(qld_ctx *ctx, const float *dres, int64_t ldr, int rows, int cols,
                               int row0, int col0, int upper, float r2limit,
                               uint32_t *dcounts, uint64_t *doffsets, uint32_t *dcols,
                               float *dr2, uint64_t capacity, uint64_t *total)
{
    if (!ctx || !total || rows < 0 || cols < 0 || (rows && (!dres || !dcounts || !doffsets)) ||
        ldr < cols) {
        if (ctx) qld_set_error(ctx, "qld_compact_dev: bad arguments");
        return QLD_ERR_INVALID;
    }
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    hipError_t e = enqueue_count_scan(dres, ldr, rows, cols, row0, col0, upper, r2limit,
                                      dcounts, doffsets, s);
    if (e != hipSuccess) return qld_hip_fail(ctx, e, "hits count/scan");
    uint64_t t = 0;
    if ((e = hipMemcpyAsync(&t, doffsets + rows, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return qld_hip_fail(ctx, e, "hits total");
    *total = t;
    if (t > capacity) {
        qld_set_error(ctx, "qld_compact_dev: %llu hits exceed capacity %llu",
                      (unsigned long long)t, (unsigned long long)capacity);
        return QLD_ERR_BUDGET;
    }
    if (t && (!dcols || !dr2)) {
        qld_set_error(ctx, "qld_compact_dev: NULL hit buffers");
        return QLD_ERR_INVALID;
    }
    if (t) {
        hipLaunchKernelGGL(hits_write_kernel, dim3(rows), dim3(HIT_WG), 0, s, dres, ldr, cols,
                           row0, col0, upper, r2limit, doffsets, dcols, dr2);
        if ((e = hipGetLastError()) != hipSuccess) return qld_hip_fail(ctx, e, "hits write");
    }
    return QLD_OK;
}

namespace {

// Pinned host buffer, grow-only.
struct PinBuf {
    void *p = nullptr;
    size_t size = 0;
    int reserve(qld_ctx *ctx, size_t bytes)
    {
        if (bytes <= size) return QLD_OK;
        release();
        hipError_t e = hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return qld_hip_fail(ctx, e, "pinned host buffer");
        }
        size = bytes ? bytes : 8;
        return QLD_OK;
    }
    void release()
    {
        if (p) hipHostFree(p);
        p = nullptr;
        size = 0;
    }
};

struct Slot {
    DevBuf d_off, d_cols, d_r2;     // rows + 1 offsets; the block's hits
    PinBuf h_off, h_cols, h_r2;     // their host copies
    PinBuf h_total;                 // the block's hit count
    hipEvent_t done = nullptr;      // hits of this slot are in host memory
    hipEvent_t ev_total = nullptr;  // h_total has arrived
    int row0 = 0, rows = 0;
};

}  // namespace

// Scratch of the streamed selection, kept by the context between calls (the
// buffers only grow): allocating and page-locking it anew cost every call a
// few milliseconds.  Calls on one context do not overlap.
struct qld_hits_cache {
    DevBuf tabA, bcA, tabB, bcB;    // the tables on the device
    DevBuf res[2], counts;          // two r^2 blocks, row hit counts
    Slot slot[3];                   // blocks b, b-1, b-2 in flight
    hipStream_t copy = nullptr;
    hipEvent_t ev_written = nullptr;
};

void qld_hits_release(qld_ctx *ctx)
{
    qld_hits_cache *c = ctx->hits;
    if (!c) return;
    if (c->copy) hipStreamSynchronize(c->copy);
    hipStreamSynchronize(ctx->stream);
    for (DevBuf *d : {&c->tabA, &c->bcA, &c->tabB, &c->bcB, &c->res[0], &c->res[1], &c->counts})
        d->release();
    for (Slot &sl : c->slot) {
        sl.d_off.release();
        sl.d_cols.release();
        sl.d_r2.release();
        sl.h_off.release();
        sl.h_cols.release();
        sl.h_r2.release();
        sl.h_total.release();
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.ev_total) hipEventDestroy(sl.ev_total);
    }
    if (c->ev_written) hipEventDestroy(c->ev_written);
    if (c->copy) hipStreamDestroy(c->copy);
    delete c;
    ctx->hits = nullptr;
}

namespace {

int hits_cache(qld_ctx *ctx, qld_hits_cache **out)
{
    if (!ctx->hits) {
        qld_hits_cache *c = new (std::nothrow) qld_hits_cache();
        if (!c) return QLD_ERR_NOMEM;
        ctx->hits = c;
        hipError_t e;
        for (Slot &sl : c->slot)
            if ((e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&sl.ev_total, hipEventDisableTiming)) != hipSuccess) {
                qld_hits_release(ctx);
                return qld_hip_fail(ctx, e, "event");
            }
        if ((e = hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&c->ev_written, hipEventDisableTiming)) != hipSuccess) {
            qld_hits_release(ctx);
            return qld_hip_fail(ctx, e, "stream/event");
        }
    }
    *out = ctx->hits;
    return QLD_OK;
}

// Device and pinned hit capacity of a slot for `hits` entries (the slot's
// previous copy has completed: its callback already ran).
int slot_reserve(qld_ctx *ctx, Slot &sl, uint64_t hits)
{
    const uint64_t want = std::max<uint64_t>(hits + hits / 4, 1u << 20) * 4;
    int rc;
    if (hits * 4 > sl.d_cols.size &&
        ((rc = sl.d_cols.reserve(ctx, want)) || (rc = sl.d_r2.reserve(ctx, want))))
        return rc;
    if (hits * 4 > sl.h_cols.size &&
        ((rc = sl.h_cols.reserve(ctx, want)) || (rc = sl.h_r2.reserve(ctx, want))))
        return rc;
    return QLD_OK;
}

}  // namespace

int qld_check_job(qld_ctx *ctx, const LdJob &j);

extern "C" int qld_correlate_hits(qld_ctx *ctx,
                                  const uint64_t *A, const uint32_t *bcA, int m,
                                  const uint64_t *B, const uint32_t *bcB, int n,
                                  int words, int snp_size, int upper, float r2limit,
                                  int64_t block_pairs, qld_hits_fn fn, void *user)
{
    if (!ctx || !fn) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    LdJob probe{A, bcA, m, B, bcB, n, words, snp_size, 0, nullptr, m, nullptr, m};
    int rc = qld_check_job(ctx, probe);
    if (rc) return rc;
    if (m == 0 || n == 0) return QLD_OK;
    const bool same = A == B && bcA == bcB && m == n;
    qld_hits_cache *run;
    if ((rc = hits_cache(ctx, &run))) return rc;
    hipError_t e;
    hipStream_t s = ctx->stream;

    // Tables on the device.
    const size_t ta = (size_t)m * words * 8, tb = (size_t)n * words * 8;
    if ((rc = run->tabA.reserve(ctx, ta)) || (rc = run->bcA.reserve(ctx, (size_t)m * 4)))
        return rc;
    if (!same && ((rc = run->tabB.reserve(ctx, tb)) || (rc = run->bcB.reserve(ctx, (size_t)n * 4))))
        return rc;
    uint64_t *dA = (uint64_t *)run->tabA.ptr;
    uint32_t *dbcA = (uint32_t *)run->bcA.ptr;
    uint64_t *dB = same ? dA : (uint64_t *)run->tabB.ptr;
    uint32_t *dbcB = same ? dbcA : (uint32_t *)run->bcB.ptr;
    if ((e = hipMemcpyAsync(dA, A, ta, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(dbcA, bcA, (size_t)m * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return qld_hip_fail(ctx, e, "H2D");
    if (!same && ((e = hipMemcpyAsync(dB, B, tb, hipMemcpyHostToDevice, s)) != hipSuccess ||
                  (e = hipMemcpyAsync(dbcB, bcB, (size_t)n * 4, hipMemcpyHostToDevice, s)) !=
                      hipSuccess))
        return qld_hip_fail(ctx, e, "H2D");

    // Blocks of B rows: about block_pairs r^2 cells each, a multiple of 256.
    if (block_pairs <= 0) block_pairs = (int64_t)1 << 26;
    int64_t R = block_pairs / std::max(m, 1);
    R = std::max<int64_t>(256, (R + 255) / 256 * 256);
    R = std::min<int64_t>(R, n);
    // Pipeline over blocks, so that the GPU always has the next block queued
    // while the host waits on it: iteration b queues block b's r^2, hit
    // counts and total; then waits for block b-1's total (block b keeps the
    // GPU busy meanwhile), sizes its buffers and queues its hit writes and
    // copies; then hands block b-2's hits to the callback.  Two r^2 blocks
    // (block b is computed while block b-1's hits are written from the
    // other) and three hit slots (b, b-1, b-2 in flight).
    if ((rc = run->res[0].reserve(ctx, (size_t)R * m * 4)) ||
        (rc = run->res[1].reserve(ctx, (size_t)R * m * 4)) ||
        (rc = run->counts.reserve(ctx, (size_t)R * 4)))
        return rc;
    for (Slot &sl : run->slot)
        if ((rc = sl.d_off.reserve(ctx, (size_t)(R + 1) * 8)) ||
            (rc = sl.h_off.reserve(ctx, (size_t)(R + 1) * 8)) ||
            (rc = sl.h_total.reserve(ctx, 8)))
            return rc;
    float *dres[2] = {(float *)run->res[0].ptr, (float *)run->res[1].ptr};
    uint32_t *dcounts = (uint32_t *)run->counts.ptr;

    struct Block {
        int64_t i0;
        int rows, col0, cols;
        int64_t ldr;
    };
    auto block = [&](int64_t b) {
        Block k;
        k.i0 = b * R;
        k.rows = (int)std::min<int64_t>(R, n - k.i0);
        k.col0 = upper ? (int)std::min<int64_t>(k.i0, m) : 0;
        k.cols = m - k.col0;
        k.ldr = std::max(k.cols, 1);
        return k;
    };
    // r^2 of block rows [i0, i0 + rows) x columns [col0, m), its hit counts
    // and their total (to the slot's pinned word).
    auto compute = [&](int64_t b) -> int {
        const Block k = block(b);
        Slot &sl = run->slot[b % 3];
        float *res = dres[b & 1];
        if (k.cols > 0) {
            if (upper && same) {
                LdJob diag{dA + (size_t)k.i0 * words, dbcA + k.i0, k.rows,
                           dA + (size_t)k.i0 * words, dbcA + k.i0, k.rows, words, snp_size, 1,
                           res, k.ldr, nullptr, k.rows};
                if ((rc = qld_run_job(ctx, diag))) return rc;
                if (k.i0 + k.rows < m) {
                    const int rest = (int)(m - k.i0 - k.rows);
                    LdJob rect{dA + (size_t)(k.i0 + k.rows) * words, dbcA + k.i0 + k.rows, rest,
                               dA + (size_t)k.i0 * words, dbcA + k.i0, k.rows, words, snp_size,
                               0, res + k.rows, k.ldr, nullptr, rest};
                    if ((rc = qld_run_job(ctx, rect))) return rc;
                }
            } else {
                LdJob job{dA + (size_t)k.col0 * words, dbcA + k.col0, k.cols,
                          dB + (size_t)k.i0 * words, dbcB + k.i0, k.rows, words, snp_size, 0,
                          res, k.ldr, nullptr, k.cols};
                if ((rc = qld_run_job(ctx, job))) return rc;
            }
        }
        if ((e = enqueue_count_scan(res, k.ldr, k.rows, k.cols, (int)k.i0, k.col0, upper,
                                    r2limit, dcounts, (uint64_t *)sl.d_off.ptr, s)) != hipSuccess ||
            (e = hipMemcpyAsync(sl.h_total.p, (uint64_t *)sl.d_off.ptr + k.rows, 8,
                                hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipEventRecord(sl.ev_total, s)) != hipSuccess)
            return qld_hip_fail(ctx, e, "hits count");
        return QLD_OK;
    };
    // Block b's hits into its slot's buffers, then copied to the host.
    auto write = [&](int64_t b) -> int {
        const Block k = block(b);
        Slot &sl = run->slot[b % 3];
        if ((e = hipEventSynchronize(sl.ev_total)) != hipSuccess)
            return qld_hip_fail(ctx, e, "hits total");
        const uint64_t hits = *(const uint64_t *)sl.h_total.p;
        if ((rc = slot_reserve(ctx, sl, hits))) return rc;
        uint64_t *d_off = (uint64_t *)sl.d_off.ptr;
        uint32_t *d_cols = (uint32_t *)sl.d_cols.ptr;
        float *d_r2 = (float *)sl.d_r2.ptr;
        if (hits) {
            hipLaunchKernelGGL(hits_write_kernel, dim3(k.rows), dim3(HIT_WG), 0, s, dres[b & 1],
                               k.ldr, k.cols, (int)k.i0, k.col0, upper, r2limit, d_off, d_cols,
                               d_r2);
            if ((e = hipGetLastError()) != hipSuccess) return qld_hip_fail(ctx, e, "hits write");
        }
        if ((e = hipEventRecord(run->ev_written, s)) != hipSuccess ||
            (e = hipStreamWaitEvent(run->copy, run->ev_written, 0)) != hipSuccess ||
            (e = hipMemcpyAsync(sl.h_off.p, d_off, (size_t)(k.rows + 1) * 8,
                                hipMemcpyDeviceToHost, run->copy)) != hipSuccess ||
            (hits && ((e = hipMemcpyAsync(sl.h_cols.p, d_cols, hits * 4, hipMemcpyDeviceToHost,
                                          run->copy)) != hipSuccess ||
                      (e = hipMemcpyAsync(sl.h_r2.p, d_r2, hits * 4, hipMemcpyDeviceToHost,
                                          run->copy)) != hipSuccess)) ||
            (e = hipEventRecord(sl.done, run->copy)) != hipSuccess)
            return qld_hip_fail(ctx, e, "hits D2H");
        sl.row0 = (int)k.i0;
        sl.rows = k.rows;
        return QLD_OK;
    };
    auto deliver = [&](int64_t b) -> int {
        Slot &sl = run->slot[b % 3];
        if ((e = hipEventSynchronize(sl.done)) != hipSuccess) return qld_hip_fail(ctx, e, "D2H");
        int u = fn(user, sl.row0, sl.rows, (const uint64_t *)sl.h_off.p,
                   (const uint32_t *)sl.h_cols.p, (const float *)sl.h_r2.p);
        if (u) {
            qld_set_error(ctx, "hits callback returned %d", u);
            return QLD_ERR_INVALID;
        }
        return QLD_OK;
    };

    const int64_t nb = (n + R - 1) / R;
    for (int64_t b = 0; b < nb; ++b) {
        if ((rc = compute(b))) return rc;
        if (b >= 1 && (rc = write(b - 1))) return rc;
        if (b >= 2 && (rc = deliver(b - 2))) return rc;
    }
    if ((rc = write(nb - 1))) return rc;
    if (nb >= 2 && (rc = deliver(nb - 2))) return rc;
    return deliver(nb - 1);
}
