// C-ABI runtime of the LD engine: contexts, device memory, dispatch.
//
// Replaces the process-global OpenCL state of the reference
// (gemm_gpu.h:36-50, gemm_gpu.c:298-447) with a per-caller context that owns
// one HIP stream, its allocations and an error string.  Nothing here calls
// exit(); every failure becomes a negative status.
#include "qld_internal.h"
#include "qld_ctx.h"
#include "ld_mfma.h"

#include <algorithm>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static char g_init_error[256];

// ---- scratch of the allocation-free entry point ----------------------------
// qld_correlate_dev_noalloc takes its scratch (queue heads, tile list and the
// row-block-tiled operand copies) from this array instead of allocating.  It
// lives in the code object: each device the library is loaded on holds its
// own copy, at its own address, until the process ends.  In host code the
// array's name is not a device address; qld_gpu_init looks the address up
// once per device (scratch_setup) and only that address is used.
#define QLD_SCRATCH_BYTES ((size_t)64 << 20)
#define QLD_SCRATCH_DEVICES 64
static __device__ __align__(256) uint8_t scratch[QLD_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[QLD_SCRATCH_DEVICES];

// With `device` current: the device address of its scratch array.
static hipError_t scratch_setup(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= QLD_SCRATCH_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

void qld_set_error(qld_ctx *ctx, const char *fmt, ...)
{
    char *dst = ctx ? ctx->err : g_init_error;
    size_t cap = ctx ? sizeof(ctx->err) : sizeof(g_init_error);
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, cap, fmt, ap);
    va_end(ap);
}

int qld_hip_fail(qld_ctx *ctx, hipError_t e, const char *what)
{
    qld_set_error(ctx, "%s: %s", what, hipGetErrorString(e));
    return QLD_ERR_DEVICE;
}

extern "C" void qld_cfg_default(qld_cfg *cfg)
{
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->engine = QLD_ENGINE_AUTO;
    cfg->r2_mode = QLD_R2_REFERENCE;
    cfg->timing = 0;
    cfg->hbm_budget = 0;
}

extern "C" int qld_gpu_init(qld_ctx **out, const qld_cfg *cfg)
{
    if (!out) {
        qld_set_error(nullptr, "qld_gpu_init: out is NULL");
        return QLD_ERR_INVALID;
    }
    *out = nullptr;
    qld_cfg c;
    if (cfg) c = *cfg; else qld_cfg_default(&c);
    if (c.engine < QLD_ENGINE_AUTO || c.engine > QLD_ENGINE_MFMA_FP4 ||
        (c.r2_mode != QLD_R2_REFERENCE && c.r2_mode != QLD_R2_EXACT)) {
        qld_set_error(nullptr, "qld_gpu_init: bad engine %d or r2_mode %d",
                      c.engine, c.r2_mode);
        return QLD_ERR_INVALID;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        qld_set_error(nullptr, "qld_gpu_init: no HIP device (%s)",
                      hipGetErrorString(e));
        return QLD_ERR_DEVICE;
    }
    if (c.device < 0 || c.device >= ndev) {
        qld_set_error(nullptr, "qld_gpu_init: device %d of %d", c.device, ndev);
        return QLD_ERR_INVALID;
    }
    if ((e = hipSetDevice(c.device)) != hipSuccess) {
        qld_set_error(nullptr, "hipSetDevice: %s", hipGetErrorString(e));
        return QLD_ERR_DEVICE;
    }
    qld_ctx *ctx = new (std::nothrow) qld_ctx();
    if (!ctx) {
        qld_set_error(nullptr, "qld_gpu_init: out of host memory");
        return QLD_ERR_NOMEM;
    }
    ctx->cfg = c;
    ctx->device = c.device;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
        qld_set_error(nullptr, "hipStreamCreate: %s", hipGetErrorString(e));
        delete ctx;
        return QLD_ERR_DEVICE;
    }
    if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess ||
        (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
        (e = hipEventCreate(&ctx->kev0)) != hipSuccess ||
        (e = hipEventCreate(&ctx->kev1)) != hipSuccess ||
        (e = hipEventCreate(&ctx->tev0)) != hipSuccess ||
        (e = hipEventCreate(&ctx->tev1)) != hipSuccess) {
        qld_set_error(nullptr, "hipEventCreate: %s", hipGetErrorString(e));
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return QLD_ERR_DEVICE;
    }
    static hipError_t attr_status = mfma_init_attributes();
    if (attr_status != hipSuccess) {
        qld_set_error(nullptr, "kernel attributes: %s", hipGetErrorString(attr_status));
        qld_gpu_release(ctx);
        return QLD_ERR_DEVICE;
    }
    if ((e = scratch_setup(c.device, &ctx->scratch)) != hipSuccess) {
        qld_set_error(nullptr, "scratch address of device %d: %s", c.device,
                      hipGetErrorString(e));
        qld_gpu_release(ctx);
        return QLD_ERR_DEVICE;
    }
    size_t freeb = 0, totalb = 0;
    hipMemGetInfo(&freeb, &totalb);
    ctx->budget = c.hbm_budget ? c.hbm_budget : (size_t)(0.9 * (double)freeb);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c.device) == hipSuccess)
        ctx->num_cu = prop.multiProcessorCount;
    *out = ctx;
    return QLD_OK;
}

extern "C" void qld_gpu_release(qld_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    qld_hits_release(ctx);
    for (auto &kv : ctx->allocs) hipFree(kv.first);
    ctx->allocs.clear();
    ctx->ws.release();
    ctx->tiles.release();
    ctx->queue.release();
    ctx->kpart.release();
    ctx->kcnt.release();
    ctx->host_stage.release();
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipEventDestroy(ctx->kev0);
    hipEventDestroy(ctx->kev1);
    hipEventDestroy(ctx->tev0);
    hipEventDestroy(ctx->tev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *qld_gpu_last_error(const qld_ctx *ctx)
{
    return ctx ? ctx->err : g_init_error;
}

// ---- memory ---------------------------------------------------------------

int qld_ctx_malloc(qld_ctx *ctx, size_t bytes, void **p)
{
    std::lock_guard<std::mutex> lk(ctx->alloc_mu);
    if (ctx->used + bytes > ctx->budget) {
        qld_set_error(ctx, "allocation of %zu bytes exceeds HBM budget "
                      "(%zu of %zu used)", bytes, ctx->used, ctx->budget);
        return QLD_ERR_BUDGET;
    }
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        qld_set_error(ctx, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return QLD_ERR_NOMEM;
    }
    ctx->used += bytes;
    return QLD_OK;
}

void qld_ctx_mfree(qld_ctx *ctx, void *p, size_t bytes)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(ctx->alloc_mu);
    hipFree(p);
    ctx->used -= bytes;
}

int DevBuf::reserve(qld_ctx *ctx, size_t bytes)
{
    if (bytes <= size) return QLD_OK;
    if (ptr) {
        hipStreamSynchronize(ctx->stream);
        qld_ctx_mfree(ctx, ptr, size);
        ptr = nullptr;
        size = 0;
    }
    owner = ctx;
    int rc = qld_ctx_malloc(ctx, bytes, &ptr);
    if (rc == QLD_OK) size = bytes;
    return rc;
}

void DevBuf::release()
{
    if (ptr && owner) qld_ctx_mfree(owner, ptr, size);
    ptr = nullptr;
    size = 0;
}

extern "C" int qld_dev_alloc(qld_ctx *ctx, size_t bytes, void **dptr)
{
    if (!ctx || !dptr) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    int rc = qld_ctx_malloc(ctx, bytes, dptr);
    if (rc == QLD_OK) ctx->allocs[*dptr] = bytes;
    return rc;
}

extern "C" int qld_dev_free(qld_ctx *ctx, void *dptr)
{
    if (!ctx) return QLD_ERR_INVALID;
    if (!dptr) return QLD_OK;
    auto it = ctx->allocs.find(dptr);
    if (it == ctx->allocs.end()) {
        qld_set_error(ctx, "qld_dev_free: %p was not allocated by this context", dptr);
        return QLD_ERR_NOTFOUND;
    }
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    qld_ctx_mfree(ctx, dptr, it->second);
    ctx->allocs.erase(it);
    return QLD_OK;
}

extern "C" int qld_h2d(qld_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_h2d");
}

extern "C" int qld_d2h(qld_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_d2h");
}

__global__ void qld_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int qld_fill_on(qld_ctx *ctx, hipStream_t s, void *dptr, uint32_t value, size_t n_u32)
{
    if (n_u32 == 0) return QLD_OK;
    size_t blocks = (n_u32 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(qld_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s,
                       (uint32_t *)dptr, value, n_u32);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "fill kernel");
}

int qld_fill_async(qld_ctx *ctx, void *dptr, uint32_t value, size_t n_u32)
{
    return qld_fill_on(ctx, ctx->stream, dptr, value, n_u32);
}

extern "C" int qld_dev_fill(qld_ctx *ctx, void *dptr, uint32_t value, size_t n_u32)
{
    if (!ctx || (!dptr && n_u32)) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    return qld_fill_async(ctx, dptr, value, n_u32);
}

__global__ void qld_copy_u128(uint4 *dst, const uint4 *src, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" int qld_dev_copy(qld_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes & 15) || (bytes && (!dst || !src))) return QLD_ERR_INVALID;
    if (!bytes) return QLD_OK;
    hipSetDevice(ctx->device);
    size_t n = bytes / 16;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(qld_copy_u128, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (uint4 *)dst, (const uint4 *)src, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_dev_copy");
}

extern "C" int qld_sync(qld_ctx *ctx)
{
    if (!ctx) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_sync");
}

// ---- all-pairs dispatch ------------------------------------------------------

static int check_job(qld_ctx *ctx, const LdJob &j)
{
    if (j.m < 0 || j.n < 0 || j.words <= 0 || j.words > QLD_MAX_WORDS_POPCOUNT) {
        qld_set_error(ctx, "bad shape m=%d n=%d words=%d (words 1 .. %d)", j.m, j.n, j.words,
                      QLD_MAX_WORDS_POPCOUNT);
        return QLD_ERR_INVALID;
    }
    // snp_size is the N of the frequencies (the header's sample count, as
    // the reference passes it: with a sample list the rows hold fewer bits
    // than N), any positive int as in the reference.
    if (j.snp_size <= 0) {
        qld_set_error(ctx, "bad snp_size %d", j.snp_size);
        return QLD_ERR_INVALID;
    }
    const int eng = ctx->cfg.engine;
    if (qld_wide_job(j.words, j.snp_size) && eng != QLD_ENGINE_AUTO &&
        eng != QLD_ENGINE_POPCOUNT) {
        qld_set_error(ctx, "snp_size %d with %d-word rows: counts can reach 2^24, beyond what "
                      "the MFMA engines' fp32 sums hold exactly (snp_size below %d and at "
                      "most %d words); use the popcount engine or AUTO", j.snp_size, j.words,
                      QLD_MAX_SNP_SIZE, QLD_MAX_WORDS);
        return QLD_ERR_INVALID;
    }
    if (j.tri && j.m != j.n) {
        qld_set_error(ctx, "triangular mode needs m == n (m=%d n=%d)", j.m, j.n);
        return QLD_ERR_INVALID;
    }
    if ((j.m && (!j.A || !j.bcA)) || (j.n && (!j.B || !j.bcB))) {
        qld_set_error(ctx, "NULL table or bit-count pointer");
        return QLD_ERR_INVALID;
    }
    if ((j.res && j.ldr < j.m) || (j.cnt && j.ldc < j.m)) {
        qld_set_error(ctx, "output pitch smaller than m");
        return QLD_ERR_INVALID;
    }
    return QLD_OK;
}

int qld_check_job(qld_ctx *ctx, const LdJob &j) { return check_job(ctx, j); }

// AUTO policy from a grid of kernel times on MI355X (single-region
// triangles of 64 to 30,000 SNPs at 256 to 100,000 samples, every engine,
// exact r^2): the fp4 engine is fastest almost everywhere -- at 16,622
// samples 5x faster than the popcount engine already at 64 SNPs, whose tile
// latency grows with the row width -- except on narrow rows (at most 16
// words, N <= 1,024) with 0.5 to 8 million pairs, where the popcount engine
// is about 10% faster (2,000 SNPs at 256 / 1,000 samples: 0.025 vs 0.028 ms).
// The int8 engine is never the fastest.  Jobs beyond the MFMA engines' exact
// range (N >= 2^24 or rows wider than 2^18 words) go to the popcount engine.
int qld_pick_engine(const qld_ctx *ctx, const LdJob &j)
{
    if (ctx->cfg.engine != QLD_ENGINE_AUTO) return ctx->cfg.engine;
    if (qld_wide_job(j.words, j.snp_size)) return QLD_ENGINE_POPCOUNT;
    const double pairs = j.tri ? 0.5 * (double)j.m * (double)j.n : (double)j.m * (double)j.n;
    if (j.words <= 16 && pairs >= 5e5 && pairs <= 8e6) return QLD_ENGINE_POPCOUNT;
    return QLD_ENGINE_MFMA_FP4;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Tile edge for one plain job.  Expanded operands: 128 x 128 tiles (two
// workgroups per CU) measured faster than 256 x 256 (3.47 vs 4.03 ms on the
// 30,000-SNP triangle).  Direct fp4 operands: 256-tiles (split over two
// workgroups) once the job has at least 4 of them per CU, 128 below: smaller
// jobs lose more to the coarser tiles' quantization than they gain per tile
// (measured triangles at 16,622 samples: 8,000 SNPs 0.223 vs 0.231 ms,
// 12,000 SNPs 0.465 vs 0.419 ms, 30,000 SNPs 2.77 vs 2.64 ms for 128 vs 256).
static int plain_tile_edge(const qld_ctx *ctx, const LdJob &job, int pmode)
{
    if (ctx->plain_tile) return ctx->plain_tile;
    if (pmode == QLD_FP4_DIRECT &&
        job_tile_count(job.n, job.m, job.tri, QLD_TILE_LARGE) >= (size_t)4 * ctx->num_cu)
        return QLD_TILE_LARGE;
    return QLD_TILE_SMALL;
}

// Split-K chunks per tile of a plain call on 128-tiles (1: none).  A tile's
// k loop is a chain of dependent loads and MFMAs over every word group, so
// a job with few tiles takes one tile's whole chain however few tiles it
// has.  Splitting each tile's groups over S workgroups shortens the chain
// S-fold and adds a fixed cost per unit (the partial sums written and, by
// the last unit, read back: S - 1 slices of 8 or 16 KiB per wave).  Model, in
// microseconds, fitted to 128-tile triangles at 16,622 and 200,000 samples:
//   time(S) = rounds(S) * (0.6 * groups / S + 8 + 1.5 * (S - 1)),
//   rounds(S) = ceil(tiles * S / (2 * CUs))   (two workgroups per CU)
// (measured 2,000 SNPs: 44.8 us unsplit, 32.0 at S = 3; 500 SNPs: 62.5 / 22.2
// at S = 4; 4,000 SNPs: 89.6 / 75.3 at S = 2; 1,000 SNPs at 200,000
// samples: 495 / 70 at S = 14).  S is the model's minimum, with at least 4
// groups per chunk.  `groups`: the word groups of the job's rows (the most
// of any job in a mixed launch).
int qld_ksplit_for(const qld_ctx *ctx, int64_t ntiles, int groups)
{
    const int64_t max_units = 16384;   // 1 GiB of partial sums
    int best = 1;
    if (ctx->ksplit) {
        best = ctx->ksplit;
    } else if (ntiles > 0) {
        const int64_t slots = 2 * (int64_t)ctx->num_cu;
        const int smax = std::min(QLD_MAX_KSPLIT, std::max(1, groups / 4));
        double best_t = 0;
        for (int s = 1; s <= smax; ++s) {
            const double rounds = (double)((ntiles * s + slots - 1) / slots);
            const double t = rounds * (0.6 * groups / s + 8.0 + 1.5 * (s - 1));
            if (s == 1 || t < best_t) {
                best_t = t;
                best = s;
            }
        }
    }
    if (ntiles > 0 && ntiles * best > max_units) best = (int)(max_units / ntiles);
    return best >= 2 ? best : 1;
}

// The device tile list of a plain job (desc 0, append_job_tiles' order),
// written by a kernel on the stream when the shape differs from the last
// call's: no host copy, so a call never waits for earlier work on the stream
// (an upload from host memory did, whenever the shape changed -- every
// block of a streamed report alternates two shapes).  Returns the count.
static int plain_tiles(qld_ctx *ctx, const LdJob &job, int T, int64_t *ntiles)
{
    const size_t cnt = job_tile_count(job.n, job.m, job.tri, T);
    *ntiles = (int64_t)cnt;
    if (ctx->tiles_key[0] == job.n && ctx->tiles_key[1] == job.m &&
        ctx->tiles_key[2] == job.tri && ctx->tiles_key[3] == T)
        return QLD_OK;
    ctx->tiles_key[0] = -1;
    const size_t tb = cnt * sizeof(WorkItem);
    int rc = ctx->tiles.reserve(ctx, tb ? tb : 8);
    if (rc) return rc;
    hipError_t e = launch_job_tiles((WorkItem *)ctx->tiles.ptr, job.n, job.m, job.tri, T,
                                    ctx->stream);
    if (e != hipSuccess) return qld_hip_fail(ctx, e, "tile list kernel");
    ctx->tiles_key[0] = job.n;
    ctx->tiles_key[1] = job.m;
    ctx->tiles_key[2] = job.tri;
    ctx->tiles_key[3] = T;
    return QLD_OK;
}

static int run_mfma(qld_ctx *ctx, int engine, const LdJob &job)
{
    // The fp4 engine reads a row-block-tiled copy of the packed rows and
    // splits it into bit planes in registers (DIRECT), unless the context
    // asks for row-major packed operands (PACKED) or an expanded copy.
    const int pmode = engine == QLD_ENGINE_MFMA_FP4 ? ctx->fp4_packed : QLD_FP4_EXPANDED;
    const bool packed = pmode != QLD_FP4_EXPANDED;
    int T = pmode == QLD_FP4_PACKED ? QLD_TILE_SMALL : plain_tile_edge(ctx, job, pmode);
    if (!packed) T = expanded_tile_cap(mfma_row_bytes(engine, job.words), T);
    if ((job.n + T - 1) / T > 65535 || (job.m + T - 1) / T > 65535) {
        qld_set_error(ctx, "too many rows for one launch (m=%d n=%d)", job.m, job.n);
        return QLD_ERR_INVALID;
    }
    int64_t ntiles = 0;
    int rc = plain_tiles(ctx, job, T, &ntiles);
    if (rc) return rc;
    const int ks = pmode == QLD_FP4_DIRECT && T == QLD_TILE_SMALL
                       ? qld_ksplit_for(ctx, ntiles, (job.words + 3) / 4)
                       : 1;
    ctx->last_ksplit = ks;
    // Persistent launches take their queue heads from the context; the
    // direct path clears them in its tiling kernel (no reset kernel).
    if (ks == 1 && (rc = ctx->queue.reserve(ctx, 16 * sizeof(uint32_t)))) return rc;
    bool queue_zeroed = false;
    hipError_t e = hipSuccess;
    WorkDesc d;
    if (pmode == QLD_FP4_DIRECT) {
        // Job rows i come from table B, columns j from table A, both read
        // from row-block-tiled copies in the workspace (made per call: the
        // caller's rows may change between calls).
        const bool same = job.A == job.B && job.m == job.n;
        const size_t bytesA = align256((size_t)tiled_bytes(job.m, job.words));
        const size_t bytesB = same ? 0 : align256((size_t)tiled_bytes(job.n, job.words));
        if ((rc = ctx->ws.reserve(ctx, bytesA + bytesB))) return rc;
        uint64_t *TA = (uint64_t *)ctx->ws.ptr;
        uint64_t *TB = same ? TA : (uint64_t *)((char *)ctx->ws.ptr + bytesA);
        queue_zeroed = ks == 1;
        e = launch_tile_rows(job.A, job.m, job.words, TA, ctx->stream,
                             queue_zeroed ? (uint32_t *)ctx->queue.ptr : nullptr);
        if (e == hipSuccess && !same) e = launch_tile_rows(job.B, job.n, job.words, TB, ctx->stream);
        if (e != hipSuccess) return qld_hip_fail(ctx, e, "row tiling kernel");
        fill_work_desc(d, nullptr, nullptr, job.bcB, job.bcA, job.n, job.m, 0, job.tri,
                       job.snp_size, job.res, job.ldr, job.cnt, job.ldc);
        set_tiled_operands(d, TB, 0, job.n, TA, 0, job.m, job.words);
    } else if (packed) {
        // Job rows i come from table B, columns j from table A.
        fill_work_desc(d, nullptr, nullptr, job.bcB, job.bcA, job.n, job.m, 0, job.tri,
                       job.snp_size, job.res, job.ldr, job.cnt, job.ldc);
        set_packed_operands(d, job.B, job.A, job.words);
    } else {
        const int64_t rb = mfma_row_bytes(engine, job.words);
        const bool same = job.A == job.B && job.m == job.n;
        const size_t bytesA = align256((size_t)job.m * rb);
        const size_t bytesB = same ? 0 : align256((size_t)job.n * rb);
        if ((rc = ctx->ws.reserve(ctx, bytesA + bytesB))) return rc;
        uint8_t *XA = (uint8_t *)ctx->ws.ptr;
        uint8_t *XB = same ? XA : XA + bytesA;
        e = launch_expand(engine, job.A, job.m, job.words, XA, ctx->stream);
        if (e == hipSuccess && !same)
            e = launch_expand(engine, job.B, job.n, job.words, XB, ctx->stream);
        if (e != hipSuccess) return qld_hip_fail(ctx, e, "expand kernel");
        fill_work_desc(d, XB, XA, job.bcB, job.bcA, job.n, job.m, rb, job.tri, job.snp_size,
                       job.res, job.ldr, job.cnt, job.ldc);
    }
    if (ks > 1) {
        // counters start zero (a fill kernel when the buffer grows) and every
        // launch leaves them zero
        const size_t had = ctx->kcnt.size;
        if ((rc = ctx->kpart.reserve(ctx, splitk_part_bytes(ntiles, ks))) ||
            (rc = ctx->kcnt.reserve(ctx, (size_t)ntiles * sizeof(uint32_t))))
            return rc;
        if (ctx->kcnt.size != had &&
            (rc = qld_fill_async(ctx, ctx->kcnt.ptr, 0u, ctx->kcnt.size / sizeof(uint32_t))))
            return rc;
        if (ctx->cfg.timing) hipEventRecord(ctx->kev0, ctx->stream);
        e = launch_mfma_splitk(ctx->cfg.r2_mode, nullptr, d, (const WorkItem *)ctx->tiles.ptr,
                               ntiles, ks, (job.words + 3) / 4, (float *)ctx->kpart.ptr,
                               (uint32_t *)ctx->kcnt.ptr, ctx->stream);
        if (ctx->cfg.timing) hipEventRecord(ctx->kev1, ctx->stream);
        return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "split-K MFMA kernel launch");
    }
    // Persistent launch: two workgroups per CU (128-tiles) or one per CU
    // per tile share (256; the launcher doubles it for split tiles) drain
    // per-XCD queues over equal shares of the tile list (expanded operands,
    // 30,000-SNP triangle: 3.47 ms vs 4.11 ms for one workgroup per tile).
    int64_t grid = (int64_t)ctx->num_cu * (T == QLD_TILE_SMALL ? 2 : 1);
    if (grid > ntiles) grid = ntiles;
    if (ctx->cfg.timing) hipEventRecord(ctx->kev0, ctx->stream);
    e = launch_mfma_items(engine, ctx->cfg.r2_mode, T, pmode, nullptr, d,
                          (const WorkItem *)ctx->tiles.ptr, ntiles, (int)grid,
                          (uint32_t *)ctx->queue.ptr, nullptr, ctx->stream, queue_zeroed);
    if (ctx->cfg.timing) hipEventRecord(ctx->kev1, ctx->stream);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "MFMA kernel launch");
}

int qld_run_job(qld_ctx *ctx, const LdJob &job)
{
    int rc = check_job(ctx, job);
    if (rc) return rc;
    if (job.m == 0 || job.n == 0 || (!job.res && !job.cnt)) return QLD_OK;
    int engine = qld_pick_engine(ctx, job);
    if (ctx->cfg.timing) hipEventRecord(ctx->ev0, ctx->stream);
    switch (engine) {
    case QLD_ENGINE_POPCOUNT: {
        if (ctx->cfg.timing) hipEventRecord(ctx->kev0, ctx->stream);
        hipError_t e = launch_popcount(job, ctx->cfg.r2_mode, ctx->stream);
        if (e != hipSuccess) return qld_hip_fail(ctx, e, "popcount kernel launch");
        if (ctx->cfg.timing) hipEventRecord(ctx->kev1, ctx->stream);
        break;
    }
    case QLD_ENGINE_MFMA_I8:
    case QLD_ENGINE_MFMA_FP4:
        if ((rc = run_mfma(ctx, engine, job))) return rc;
        break;
    default:
        qld_set_error(ctx, "engine %d not available", engine);
        return QLD_ERR_INVALID;
    }
    if (ctx->cfg.timing) hipEventRecord(ctx->ev1, ctx->stream);
    ctx->last_engine = engine;
    ctx->timed = ctx->cfg.timing != 0;
    return QLD_OK;
}

extern "C" int qld_correlate_dev(qld_ctx *ctx,
                                 const uint64_t *dA, const uint32_t *dbcA, int m,
                                 const uint64_t *dB, const uint32_t *dbcB, int n,
                                 int words, int snp_size, int triangular,
                                 float *dresults, int64_t ldr,
                                 uint32_t *dcounts, int64_t ldc)
{
    if (!ctx) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    LdJob j{dA, dbcA, m, dB, dbcB, n, words, snp_size, triangular ? 1 : 0,
            dresults, ldr, dcounts, ldc};
    return qld_run_job(ctx, j);
}

// ---- allocation-free variant of qld_correlate_dev -----------------------------
//
// The same computation with the same kernels, for callers that may not
// allocate device memory (code under HIP graph capture, tight serving loops):
// the popcount engine needs no scratch; the fp4 engine with direct operands
// (the default) takes the queue heads, the tile list -- written on the device
// in closed form instead of uploaded -- and the row-block-tiled operand
// copies from the static scratch array.  No hipMalloc, hipFree, host
// allocation or memset: the queue heads are cleared by the kernel launcher.
// Every call on a device shares that device's scratch, so calls that overlap
// on one device (two contexts, two threads) race.
extern "C" int qld_correlate_dev_noalloc(qld_ctx *ctx,
                                         const uint64_t *dA, const uint32_t *dbcA, int m,
                                         const uint64_t *dB, const uint32_t *dbcB, int n,
                                         int words, int snp_size, int triangular,
                                         float *dresults, int64_t ldr,
                                         uint32_t *dcounts, int64_t ldc)
{
    if (!ctx) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    const LdJob job{dA, dbcA, m, dB, dbcB, n, words, snp_size, triangular ? 1 : 0,
                    dresults, ldr, dcounts, ldc};
    int rc = check_job(ctx, job);
    if (rc) return rc;
    if (job.m == 0 || job.n == 0 || (!job.res && !job.cnt)) return QLD_OK;
    const int engine = qld_pick_engine(ctx, job);
    hipError_t e = hipSuccess;
    if (ctx->cfg.timing) hipEventRecord(ctx->ev0, ctx->stream);
    if (engine == QLD_ENGINE_POPCOUNT) {
        if (ctx->cfg.timing) hipEventRecord(ctx->kev0, ctx->stream);
        e = launch_popcount(job, ctx->cfg.r2_mode, ctx->stream);
        if (e != hipSuccess) return qld_hip_fail(ctx, e, "popcount kernel launch");
        if (ctx->cfg.timing) hipEventRecord(ctx->kev1, ctx->stream);
    } else {
        if (engine != QLD_ENGINE_MFMA_FP4 || ctx->fp4_packed != QLD_FP4_DIRECT) {
            qld_set_error(ctx, "qld_correlate_dev_noalloc: engine %d with fp4 operand mode %d "
                          "needs an allocated workspace (only popcount and fp4 direct "
                          "operands run without one)", engine, ctx->fp4_packed);
            return QLD_ERR_INVALID;
        }
        if (!ctx->scratch) {
            qld_set_error(ctx, "qld_correlate_dev_noalloc: no scratch on device %d",
                          ctx->device);
            return QLD_ERR_DEVICE;
        }
        const int T = plain_tile_edge(ctx, job, QLD_FP4_DIRECT);
        if ((job.n + T - 1) / T > 65535 || (job.m + T - 1) / T > 65535) {
            qld_set_error(ctx, "too many rows for one launch (m=%d n=%d)", job.m, job.n);
            return QLD_ERR_INVALID;
        }
        // scratch: queue heads | tile list | tiled A | tiled B (256-aligned)
        const bool same = job.A == job.B && job.m == job.n;
        const size_t ntiles = job_tile_count(job.n, job.m, job.tri, T);
        const size_t off_items = 256;
        const size_t off_a = off_items + align256(ntiles * sizeof(WorkItem));
        const size_t off_b = off_a + align256((size_t)tiled_bytes(job.m, job.words));
        const size_t need = off_b + (same ? 0 : align256((size_t)tiled_bytes(job.n, job.words)));
        if (need > QLD_SCRATCH_BYTES) {
            qld_set_error(ctx, "qld_correlate_dev_noalloc: %d x %d rows of %d words need %zu "
                          "bytes of scratch, more than its %zu", job.n, job.m, job.words, need,
                          (size_t)QLD_SCRATCH_BYTES);
            return QLD_ERR_BUDGET;
        }
        uint8_t *base = (uint8_t *)ctx->scratch;
        uint32_t *queue = (uint32_t *)base;
        WorkItem *items = (WorkItem *)(base + off_items);
        uint64_t *TA = (uint64_t *)(base + off_a);
        uint64_t *TB = same ? TA : (uint64_t *)(base + off_b);
        e = launch_job_tiles(items, job.n, job.m, job.tri, T, ctx->stream);
        if (e == hipSuccess) e = launch_tile_rows(job.A, job.m, job.words, TA, ctx->stream);
        if (e == hipSuccess && !same)
            e = launch_tile_rows(job.B, job.n, job.words, TB, ctx->stream);
        if (e != hipSuccess) return qld_hip_fail(ctx, e, "tile list / row tiling kernel");
        WorkDesc d;
        fill_work_desc(d, nullptr, nullptr, job.bcB, job.bcA, job.n, job.m, 0, job.tri,
                       job.snp_size, job.res, job.ldr, job.cnt, job.ldc);
        set_tiled_operands(d, TB, 0, job.n, TA, 0, job.m, job.words);
        int64_t grid = (int64_t)ctx->num_cu * (T == QLD_TILE_SMALL ? 2 : 1);
        if (grid > (int64_t)ntiles) grid = (int64_t)ntiles;
        if (ctx->cfg.timing) hipEventRecord(ctx->kev0, ctx->stream);
        e = launch_mfma_items(engine, ctx->cfg.r2_mode, T, QLD_FP4_DIRECT, nullptr, d, items,
                              (int64_t)ntiles, (int)grid, queue, nullptr, ctx->stream);
        if (ctx->cfg.timing) hipEventRecord(ctx->kev1, ctx->stream);
        if (e != hipSuccess) return qld_hip_fail(ctx, e, "MFMA kernel launch");
    }
    if (ctx->cfg.timing) hipEventRecord(ctx->ev1, ctx->stream);
    ctx->last_engine = engine;
    ctx->timed = ctx->cfg.timing != 0;
    return QLD_OK;
}

extern "C" int qld_gpu_correlate(qld_ctx *ctx,
                                 const uint64_t *A, const uint32_t *bcA, int m,
                                 const uint64_t *B, const uint32_t *bcB, int n,
                                 int words, int snp_size, int triangular,
                                 float *results, uint32_t *counts)
{
    if (!ctx) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    LdJob probe{A, bcA, m, B, bcB, n, words, snp_size, triangular ? 1 : 0,
                results, m, counts, m};
    int rc = check_job(ctx, probe);
    if (rc) return rc;
    if (m == 0 || n == 0 || (!results && !counts)) return QLD_OK;

    const bool same = (A == B) && (bcA == bcB) && (m == n);
    const size_t ta = (size_t)m * words * 8, tb = same ? 0 : (size_t)n * words * 8;
    const size_t cells = (size_t)n * m;
    const size_t align = 256;
    auto up = [&](size_t x) { return (x + align - 1) / align * align; };
    size_t off_a = 0, off_b = off_a + up(ta), off_bca = off_b + up(tb),
           off_bcb = off_bca + up((size_t)m * 4),
           off_res = off_bcb + up(same ? 0 : (size_t)n * 4),
           off_cnt = off_res + up(results ? cells * 4 : 0),
           total = off_cnt + up(counts ? cells * 4 : 0);
    if ((rc = ctx->host_stage.reserve(ctx, total))) return rc;
    char *base = (char *)ctx->host_stage.ptr;
    uint64_t *dA = (uint64_t *)(base + off_a);
    uint64_t *dB = same ? dA : (uint64_t *)(base + off_b);
    uint32_t *dbcA = (uint32_t *)(base + off_bca);
    uint32_t *dbcB = same ? dbcA : (uint32_t *)(base + off_bcb);
    float *dres = results ? (float *)(base + off_res) : nullptr;
    uint32_t *dcnt = counts ? (uint32_t *)(base + off_cnt) : nullptr;

    hipError_t e;
    hipStream_t s = ctx->stream;
    if ((e = hipMemcpyAsync(dA, A, ta, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(dbcA, bcA, (size_t)m * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return qld_hip_fail(ctx, e, "H2D");
    if (!same &&
        ((e = hipMemcpyAsync(dB, B, tb, hipMemcpyHostToDevice, s)) != hipSuccess ||
         (e = hipMemcpyAsync(dbcB, bcB, (size_t)n * 4, hipMemcpyHostToDevice, s)) != hipSuccess))
        return qld_hip_fail(ctx, e, "H2D");
    // Cells the triangular launch does not define come back as 0, as the
    // reference zeroes its results before filling them (gemm_gpu.c:808-812).
    if (triangular) {
        if (dres && (rc = qld_fill_async(ctx, dres, 0u, cells))) return rc;
        if (dcnt && (rc = qld_fill_async(ctx, dcnt, 0u, cells))) return rc;
    }
    LdJob j{dA, dbcA, m, dB, dbcB, n, words, snp_size, triangular ? 1 : 0,
            dres, m, dcnt, m};
    if ((rc = qld_run_job(ctx, j))) return rc;
    if (results &&
        (e = hipMemcpyAsync(results, dres, cells * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return qld_hip_fail(ctx, e, "D2H");
    if (counts &&
        (e = hipMemcpyAsync(counts, dcnt, cells * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return qld_hip_fail(ctx, e, "D2H");
    if ((e = hipStreamSynchronize(s)) != hipSuccess)
        return qld_hip_fail(ctx, e, "qld_gpu_correlate");
    return QLD_OK;
}

extern "C" int qld_last_kernel_ms(qld_ctx *ctx, float *ms)
{
    if (!ctx || !ms) return QLD_ERR_INVALID;
    if (!ctx->timed) {
        qld_set_error(ctx, "no timed call (set cfg.timing)");
        return QLD_ERR_NOTREADY;
    }
    hipSetDevice(ctx->device);
    hipError_t e = hipEventSynchronize(ctx->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_last_kernel_ms");
}

extern "C" int qld_last_contract_ms(qld_ctx *ctx, float *ms)
{
    if (!ctx || !ms) return QLD_ERR_INVALID;
    if (!ctx->timed) {
        qld_set_error(ctx, "no timed call (set cfg.timing)");
        return QLD_ERR_NOTREADY;
    }
    hipSetDevice(ctx->device);
    hipError_t e = hipEventSynchronize(ctx->kev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->kev0, ctx->kev1);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_last_contract_ms");
}

extern "C" int qld_timer_start(qld_ctx *ctx)
{
    if (!ctx) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipEventRecord(ctx->tev0, ctx->stream);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_timer_start");
}

extern "C" int qld_timer_stop(qld_ctx *ctx, float *ms)
{
    if (!ctx || !ms) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipEventRecord(ctx->tev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->tev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->tev0, ctx->tev1);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_timer_stop");
}

extern "C" int qld_set_fp4_operands(qld_ctx *ctx, int mode)
{
    if (!ctx || (mode != QLD_FP4_PACKED && mode != QLD_FP4_EXPANDED && mode != QLD_FP4_DIRECT))
        return QLD_ERR_INVALID;
    ctx->fp4_packed = mode;
    return QLD_OK;
}

extern "C" int qld_set_timing(qld_ctx *ctx, int on)
{
    if (!ctx) return QLD_ERR_INVALID;
    ctx->cfg.timing = on ? 1 : 0;
    return QLD_OK;
}

extern "C" int qld_set_ksplit(qld_ctx *ctx, int chunks)
{
    if (!ctx || chunks < 0 || chunks > QLD_MAX_KSPLIT) return QLD_ERR_INVALID;
    ctx->ksplit = chunks;
    return QLD_OK;
}

extern "C" int qld_last_ksplit(const qld_ctx *ctx)
{
    return ctx ? ctx->last_ksplit : QLD_ERR_INVALID;
}

extern "C" int qld_set_plain_tile(qld_ctx *ctx, int tile)
{
    if (!ctx || (tile != 0 && tile != QLD_TILE_LARGE && tile != QLD_TILE_SMALL))
        return QLD_ERR_INVALID;
    ctx->plain_tile = tile;
    return QLD_OK;
}

extern "C" int qld_last_engine(const qld_ctx *ctx)
{
    return ctx ? ctx->last_engine : QLD_ENGINE_AUTO;
}
