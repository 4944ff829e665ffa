// The context object behind the opaque qld_ctx* of the C ABI.
#pragma once

#include "qld_internal.h"
#include "ld_mfma.h"

#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

struct qld_ctx;
struct qld_hits_cache;    // scratch of the streamed selection (qld_hits.hip)

// Grow-only device buffer owned by a context.
struct DevBuf {
    void *ptr = nullptr;
    size_t size = 0;
    qld_ctx *owner = nullptr;
    int reserve(qld_ctx *ctx, size_t bytes);
    void release();
};

struct qld_ctx {
    qld_cfg cfg;
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;     // last call's kernels
    hipEvent_t kev0 = nullptr, kev1 = nullptr;   // last call's contraction kernel
    hipEvent_t tev0 = nullptr, tev1 = nullptr;   // qld_timer_start/stop
    bool timed = false;
    int last_engine = QLD_ENGINE_AUTO;
    size_t budget = 0, used = 0;
    std::mutex alloc_mu;     // guards `used`: a server's dispatcher allocates too
    std::unordered_map<void *, size_t> allocs;   // qld_dev_alloc'd blocks
    DevBuf host_stage;   // staging for the host-buffer entry point
    DevBuf ws;           // engine workspace (expanded operands)
    DevBuf tiles;        // device tile list of the last plain MFMA shape
    DevBuf queue;        // per-XCD queue heads of persistent plain launches
    DevBuf kpart;        // split-K partial sums of plain launches
    DevBuf kcnt;         // split-K arrival counters (zero between launches)
    int tiles_key[4] = {-1, -1, -1, -1};         // (ni, nj, tri, tile edge)
    int tile_cutoff = 1024;   // windows of at least this many SNPs use 256-tiles
    int plain_tile = 0;       // tile edge of plain MFMA calls; 0 = by shape
    int fp4_packed = QLD_FP4_DIRECT;   // fp4 operands: QLD_FP4_DIRECT, _PACKED or _EXPANDED
    int ksplit = 0;           // split-K chunks of plain 128-tile calls: 0 = by shape, 1 = none
    int last_ksplit = 1;      // chunks per tile of the last plain MFMA call
    void *scratch = nullptr;  // device address of this device's static scratch array
    int plan_sched = QLD_SCHED_WAVES;  // persistent plans: per-wave units or whole tiles
    bool plan_graphs = true;  // capture each plan's launches in a HIP graph
    bool plan_shared = false; // persistent plans of single-region windows count
                              // each pair once (QLD_BATCH_SHARED form)
    qld_hits_cache *hits = nullptr;   // kept between streamed-selection calls
    char err[512] = {0};
};

// A resident table: packed rows, bit counts and the expanded MFMA operand.
struct qld_table {
    int rows = 0, words = 0, snp_size = 0, engine = QLD_ENGINE_MFMA_FP4;
    int direct = 0;           // `packed` is the row-block-tiled copy the fp4
                              // kernels read into registers; no X
    int64_t row_bytes = 0;
    uint64_t *packed = nullptr;
    uint32_t *bc = nullptr;
    uint8_t *X = nullptr;
    size_t bytes_packed = 0, bytes_bc = 0, bytes_x = 0;
};

void qld_set_error(qld_ctx *ctx, const char *fmt, ...);
int qld_hip_fail(qld_ctx *ctx, hipError_t e, const char *what);
int qld_ctx_malloc(qld_ctx *ctx, size_t bytes, void **p);
void qld_ctx_mfree(qld_ctx *ctx, void *p, size_t bytes);
int qld_fill_async(qld_ctx *ctx, void *dptr, uint32_t value, size_t n_u32);
int qld_fill_on(qld_ctx *ctx, hipStream_t s, void *dptr, uint32_t value, size_t n_u32);
void qld_hits_release(qld_ctx *ctx);
// Split-K chunks per 128-tile for a direct fp4 launch of `ntiles` tiles whose
// rows have `groups` word groups (1: none); the context's forced setting
// (qld_set_ksplit) if any.
int qld_ksplit_for(const qld_ctx *ctx, int64_t ntiles, int groups);
int qld_run_job(qld_ctx *ctx, const LdJob &job);
int qld_pick_engine(const qld_ctx *ctx, const LdJob &job);
int qld_window_tile(const qld_ctx *ctx, int size);
