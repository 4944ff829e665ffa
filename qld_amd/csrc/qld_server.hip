// Serving: continuous batching of LD queries against resident tables.
//
// Requests are windows of resident tables.  A dispatcher thread owns all
// device work: it admits pending requests in arrival order while their
// outputs fit a device arena (the HBM budget), turns each admitted group into
// one MFMA launch (one job descriptor per request, 128 x 128 tiles) on one
// of several streams, and retires a group when its event completes --
// optionally after copying results to host memory on the same stream.  New
// requests are admitted whenever a stream slot and arena space are free, so
// batches overlap on the device.  Each tile depends only on its own
// request's rows, so a result does not depend on what it was batched with.
#include "qld_internal.h"
#include "qld_ctx.h"
#include "ld_mfma.h"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

enum ReqState { PENDING = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

struct Request {
    const qld_table *t = nullptr;
    int start = 0, size = 0, flags = 0;
    float *host_out = nullptr;
    int state = PENDING;
    double t_submit = 0, t_done = 0;
    int64_t out_off = -1;
    size_t out_bytes = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void *h_stage = nullptr;      // pinned: descriptors + items
    void *d_stage = nullptr;
    size_t cap = 0;
    uint32_t *d_kcnt = nullptr;   // split-K arrival counters (zero between launches)
    size_t kcnt_cap = 0;
    void *d_kpart = nullptr;      // split-K partial sums
    size_t kpart_cap = 0;
    bool busy = false;
    std::vector<int64_t> reqs;
};

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) of the output arena.
struct Arena {
    std::map<int64_t, int64_t> free_;   // offset -> length
    void init(int64_t size) { free_.clear(); free_[0] = size; }
    int64_t alloc(int64_t len)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= len) {
                int64_t off = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest) free_[off + len] = rest;
                return off;
            }
        return -1;
    }
    void release(int64_t off, int64_t len)
    {
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return;
            }
        }
        free_[off] = len;
    }
};

constexpr int64_t ARENA_ALIGN = 256;

}  // namespace

struct qld_server {
    qld_ctx *ctx = nullptr;
    qld_server_cfg cfg;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Request> reqs;
    std::vector<Slot> slots;
    uint8_t *arena_mem = nullptr;
    Arena arena;
    bool stop = false;
    std::atomic<uint64_t> gen{0};   // bumped by every submit and release: either
                                    // may let a blocked admission proceed
    int64_t next_id = 1;
    int64_t completed = 0, batches = 0, pairs = 0;
    int error = QLD_OK;
};

static size_t out_bytes_of(int size)
{
    size_t b = (size_t)size * (size_t)size * sizeof(float);
    return (b + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
}

static int slot_reserve(qld_server *s, Slot &sl, size_t bytes)
{
    if (bytes <= sl.cap) return QLD_OK;
    size_t cap = sl.cap ? sl.cap : (1u << 20);
    while (cap < bytes) cap *= 2;
    if (sl.h_stage) hipHostFree(sl.h_stage);
    if (sl.d_stage) qld_ctx_mfree(s->ctx, sl.d_stage, sl.cap);
    sl.h_stage = nullptr;
    sl.d_stage = nullptr;
    sl.cap = 0;
    if (hipHostMalloc(&sl.h_stage, cap, hipHostMallocDefault) != hipSuccess) return QLD_ERR_NOMEM;
    int rc = qld_ctx_malloc(s->ctx, cap, &sl.d_stage);
    if (rc) return rc;
    sl.cap = cap;
    return QLD_OK;
}

// Split-K scratch of slot `sl` for `items` tiles in `ks` chunks: counters
// zeroed on the slot's stream whenever their buffer is (re)allocated, left
// zero by every launch.  Only the dispatcher touches a slot, and only while
// no batch of it is in flight.
static int slot_split_reserve(qld_server *s, Slot &sl, int64_t items, int ks)
{
    const size_t cnt = (size_t)items * sizeof(uint32_t), part = splitk_part_bytes(items, ks);
    int rc;
    if (cnt > sl.kcnt_cap) {
        if (sl.d_kcnt) qld_ctx_mfree(s->ctx, sl.d_kcnt, sl.kcnt_cap);
        sl.d_kcnt = nullptr;
        sl.kcnt_cap = 0;
        const size_t cap = std::max(cnt, (size_t)4096);
        if ((rc = qld_ctx_malloc(s->ctx, cap, (void **)&sl.d_kcnt))) return rc;
        sl.kcnt_cap = cap;
        if ((rc = qld_fill_on(s->ctx, sl.stream, sl.d_kcnt, 0u, cap / sizeof(uint32_t))))
            return rc;
    }
    if (part > sl.kpart_cap) {
        if (sl.d_kpart) qld_ctx_mfree(s->ctx, sl.d_kpart, sl.kpart_cap);
        sl.d_kpart = nullptr;
        sl.kpart_cap = 0;
        if ((rc = qld_ctx_malloc(s->ctx, part, &sl.d_kpart))) return rc;
        sl.kpart_cap = part;
    }
    return QLD_OK;
}

// Launch one batch on slot `sl`: `batch` holds copies of its requests, taken
// under the server lock, so this runs without it (the HIP calls must not
// stall submit, wait or release); only the dispatcher thread touches slots.
static int launch_batch(qld_server *s, Slot &sl, const std::vector<Request> &batch)
{
    qld_ctx *ctx = s->ctx;
    const int T = QLD_TILE_SMALL;
    size_t nitems = 0;
    for (const Request &r : batch) nitems += job_tile_count(r.size, r.size, 1, T);
    const size_t nd = batch.size();
    const size_t desc_bytes = (nd * sizeof(WorkDesc) + 255) / 256 * 256;
    const size_t bytes = desc_bytes + nitems * sizeof(WorkItem);
    int rc = slot_reserve(s, sl, bytes);
    if (rc) return rc;
    WorkDesc *hd = (WorkDesc *)sl.h_stage;
    WorkItem *hi = (WorkItem *)((char *)sl.h_stage + desc_bytes);
    // Requests on direct-mode tables (packed rows into registers) and on
    // tables with an expanded operand run as two launches; the items of each
    // group are contiguous.
    size_t cnt = 0, split = 0;
    int engine = QLD_ENGINE_MFMA_FP4;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) split = cnt;
        for (size_t k = 0; k < nd; ++k) {
            const Request &r = batch[k];
            const qld_table *t = r.t;
            if (t->direct != (pass == 0)) continue;
            engine = t->engine;
            const uint32_t *bc = t->bc + r.start;
            float *out = (float *)(s->arena_mem + r.out_off);
            if (t->direct) {
                fill_work_desc(hd[k], nullptr, nullptr, bc, bc, r.size, r.size, 0, 1,
                               t->snp_size, out, r.size, nullptr, r.size);
                set_tiled_operands(hd[k], t->packed, r.start, t->rows, t->packed, r.start,
                                   t->rows, t->words);
            } else {
                const uint8_t *X = t->X + (int64_t)r.start * t->row_bytes;
                fill_work_desc(hd[k], X, X, bc, bc, r.size, r.size, t->row_bytes, 1,
                               t->snp_size, out, r.size, nullptr, r.size);
            }
            append_job_tiles(hi, &cnt, (uint32_t)k, r.size, r.size, 1, T);
        }
    }
    hipError_t e = hipMemcpyAsync(sl.d_stage, sl.h_stage, bytes, hipMemcpyHostToDevice,
                                  sl.stream);
    WorkDesc none;
    memset(&none, 0, sizeof(none));
    const WorkItem *d_items = (const WorkItem *)((char *)sl.d_stage + desc_bytes);
    // Direct-mode tiles of small batches on long rows split their k loops
    // over several workgroups (qld_ksplit_for); every chunk count gives the
    // same exact counts, so a result does not depend on its batch.
    int ks = 1, groups = 0;
    if (split > 0) {
        for (const Request &r : batch)
            if (r.t->direct) groups = std::max(groups, (r.t->words + 3) / 4);
        ks = qld_ksplit_for(ctx, (int64_t)split, groups);
    }
    if (e == hipSuccess && split > 0 && ks > 1) {
        if ((rc = slot_split_reserve(s, sl, (int64_t)split, ks))) return rc;
        e = launch_mfma_splitk(ctx->cfg.r2_mode, (const WorkDesc *)sl.d_stage, none, d_items,
                               (int64_t)split, ks, groups, (float *)sl.d_kpart, sl.d_kcnt,
                               sl.stream);
    } else if (e == hipSuccess && split > 0) {
        e = launch_mfma_items(QLD_ENGINE_MFMA_FP4, ctx->cfg.r2_mode, T, QLD_FP4_DIRECT,
                              (const WorkDesc *)sl.d_stage, none, d_items, (int64_t)split, 0,
                              nullptr, nullptr, sl.stream);
    }
    if (e == hipSuccess && cnt > split)
        e = launch_mfma_items(engine, ctx->cfg.r2_mode, T, 0, (const WorkDesc *)sl.d_stage, none,
                              d_items + split, (int64_t)(cnt - split), 0, nullptr, nullptr,
                              sl.stream);
    for (const Request &r : batch) {
        if (e == hipSuccess && (r.flags & QLD_REQ_COPY) && r.host_out)
            e = hipMemcpyAsync(r.host_out, s->arena_mem + r.out_off,
                               (size_t)r.size * r.size * sizeof(float),
                               hipMemcpyDeviceToHost, sl.stream);
    }
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    if (e != hipSuccess) return qld_hip_fail(ctx, e, "serving batch");
    return QLD_OK;
}

static void dispatcher(qld_server *s)
{
    hipSetDevice(s->ctx->device);
    const int max_batch = s->cfg.max_batch;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progress = false;
        // Retire finished batches.
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            const double t = now_s();
            for (int64_t id : sl.reqs) {
                Request &r = s->reqs[id];
                r.state = q == hipSuccess ? DONE : FAILED;
                r.t_done = t;
                if (r.flags & (QLD_REQ_COPY | QLD_REQ_DISCARD)) {
                    s->arena.release(r.out_off, (int64_t)r.out_bytes);
                    r.out_off = -1;
                }
                s->completed++;
            }
            if (q != hipSuccess) s->error = QLD_ERR_DEVICE;
            sl.reqs.clear();
            sl.busy = false;
            progress = true;
        }
        if (progress) s->cv_done.notify_all();
        // Admit pending requests into free slots (arrival order).
        for (Slot &sl : s->slots) {
            if (sl.busy || s->pending.empty()) continue;
            while (!s->pending.empty() && (int)sl.reqs.size() < max_batch) {
                Request &r = s->reqs[s->pending.front()];
                int64_t off = s->arena.alloc((int64_t)r.out_bytes);
                if (off < 0) break;      // wait for memory to free up
                r.out_off = off;
                r.state = RUNNING;
                sl.reqs.push_back(s->pending.front());
                s->pending.pop_front();
            }
            if (sl.reqs.empty()) break;
            std::vector<Request> batch;
            batch.reserve(sl.reqs.size());
            int64_t batch_pairs = 0;
            for (int64_t id : sl.reqs) {
                batch.push_back(s->reqs[id]);
                batch_pairs += (int64_t)batch.back().size * (batch.back().size - 1) / 2;
            }
            lk.unlock();
            const int launched = launch_batch(s, sl, batch);
            lk.lock();
            if (launched == QLD_OK) {
                sl.busy = true;
                s->batches++;
                s->pairs += batch_pairs;
            } else {
                for (int64_t id : sl.reqs) {
                    Request &r = s->reqs[id];
                    r.state = FAILED;
                    r.t_done = now_s();
                    s->arena.release(r.out_off, (int64_t)r.out_bytes);
                    r.out_off = -1;
                }
                sl.reqs.clear();
                s->error = QLD_ERR_DEVICE;
                s->cv_done.notify_all();
            }
            progress = true;
        }
        bool busy = false;
        for (Slot &sl : s->slots) busy |= sl.busy;
        if (s->stop && !busy) {
            // Shutting down: requests that could not be admitted fail.
            for (int64_t id : s->pending) {
                Request &r = s->reqs[id];
                r.state = FAILED;
                r.t_done = now_s();
            }
            s->pending.clear();
            s->cv_done.notify_all();
            break;
        }
        if (!progress) {
            // While work is in flight, spin on the batches' completion events
            // with the lock released (slots belong to this thread), for up
            // to 20 us at a time, and go round again as soon as a batch
            // completes or a submit or release arrives: a request's latency
            // includes the poll that notices it (a 20 us timed sleep between
            // polls, which Linux's 50 us default timer slack stretched to
            // about 70 us, put 0.09-0.11 ms p50 latencies at the benchmark's
            // rates; 1 us of slack, 0.07-0.08 ms).  The thread keeps a core
            // busy only while batches are in flight.  Otherwise sleep until a
            // submit or a release changes what can be admitted (waiting on
            // `gen`, not on `pending`, which may be non-empty but blocked on
            // memory).
            if (busy) {
                const uint64_t g = s->gen.load();
                lk.unlock();
                const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(20);
                bool done = false;
                while (!done && s->gen.load(std::memory_order_relaxed) == g &&
                       std::chrono::steady_clock::now() < until)
                    for (Slot &sl : s->slots)
                        if (sl.busy && hipEventQuery(sl.done) != hipErrorNotReady) {
                            done = true;
                            break;
                        }
                lk.lock();
            } else {
                const uint64_t g = s->gen;
                s->cv_work.wait(lk, [&] { return s->stop || s->gen != g; });
            }
        }
    }
}

extern "C" int qld_server_create(qld_ctx *ctx, const qld_server_cfg *cfg, qld_server **out)
{
    if (!ctx || !out) return QLD_ERR_INVALID;
    *out = nullptr;
    qld_server_cfg c;
    memset(&c, 0, sizeof(c));
    if (cfg) c = *cfg;
    if (!c.out_budget) c.out_budget = (size_t)4 << 30;
    if (c.max_batch <= 0) c.max_batch = 256;
    if (c.streams <= 0) c.streams = 3;
    if (c.streams > 8) c.streams = 8;
    hipSetDevice(ctx->device);
    qld_server *s = new (std::nothrow) qld_server();
    if (!s) return QLD_ERR_NOMEM;
    s->ctx = ctx;
    s->cfg = c;
    int rc = qld_ctx_malloc(ctx, c.out_budget, (void **)&s->arena_mem);
    if (rc) {
        delete s;
        return rc;
    }
    s->arena.init((int64_t)c.out_budget);
    s->slots.resize(c.streams);
    for (Slot &sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            slot_reserve(s, sl, 1u << 20) != QLD_OK) {
            qld_set_error(ctx, "qld_server_create: stream/staging setup failed");
            s->stop = true;
            qld_server_destroy(s);
            return QLD_ERR_DEVICE;
        }
    }
    s->th = std::thread(dispatcher, s);
    *out = s;
    return QLD_OK;
}

extern "C" int qld_server_destroy(qld_server *s)
{
    if (!s) return QLD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slots) {
        if (sl.stream) hipStreamSynchronize(sl.stream);
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.stream) hipStreamDestroy(sl.stream);
        if (sl.h_stage) hipHostFree(sl.h_stage);
        if (sl.d_stage) qld_ctx_mfree(s->ctx, sl.d_stage, sl.cap);
        if (sl.d_kcnt) qld_ctx_mfree(s->ctx, sl.d_kcnt, sl.kcnt_cap);
        if (sl.d_kpart) qld_ctx_mfree(s->ctx, sl.d_kpart, sl.kpart_cap);
    }
    qld_ctx_mfree(s->ctx, s->arena_mem, s->cfg.out_budget);
    delete s;
    return QLD_OK;
}

extern "C" int qld_server_submit(qld_server *s, const qld_table *t, int start, int size,
                                 int flags, float *host_out, int64_t *req_id)
{
    if (!s || !t || !req_id || start < 0 || size < 1 || (int64_t)start + size > t->rows ||
        (flags != QLD_REQ_KEEP && flags != QLD_REQ_COPY && flags != QLD_REQ_DISCARD) ||
        (flags == QLD_REQ_COPY && !host_out))
        return QLD_ERR_INVALID;
    Request r;
    r.t = t;
    r.start = start;
    r.size = size;
    r.flags = flags;
    r.host_out = host_out;
    r.out_bytes = out_bytes_of(size);
    if (r.out_bytes > s->cfg.out_budget) return QLD_ERR_BUDGET;
    r.t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return QLD_ERR_INVALID;
        const int64_t id = s->next_id++;
        s->reqs.emplace(id, r);
        s->pending.push_back(id);
        s->gen++;
        *req_id = id;
    }
    s->cv_work.notify_one();
    return QLD_OK;
}

extern "C" int qld_server_wait(qld_server *s, int64_t id, double timeout_ms, double *latency_ms)
{
    if (!s) return QLD_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return QLD_ERR_NOTFOUND;
    auto finished = [&] {
        const Request &r = s->reqs[id];
        return r.state == DONE || r.state == FAILED;
    };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, finished);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                  finished))
        return QLD_ERR_NOTREADY;
    const Request &r = s->reqs[id];
    if (latency_ms) *latency_ms = (r.t_done - r.t_submit) * 1e3;
    return r.state == DONE ? QLD_OK : QLD_ERR_DEVICE;
}

extern "C" int qld_server_fetch(qld_server *s, int64_t id, float *dst)
{
    if (!s || !dst) return QLD_ERR_INVALID;
    int64_t off;
    size_t n;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return QLD_ERR_NOTFOUND;
        Request &r = it->second;
        if (r.state != DONE) return QLD_ERR_NOTREADY;
        if (r.out_off < 0) return QLD_ERR_NOTFOUND;
        off = r.out_off;
        n = (size_t)r.size * r.size;
    }
    hipSetDevice(s->ctx->device);
    hipError_t e = hipMemcpy(dst, s->arena_mem + off, n * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return qld_hip_fail(s->ctx, e, "qld_server_fetch");
    return qld_server_release(s, id);
}

extern "C" int qld_server_release(qld_server *s, int64_t id)
{
    if (!s) return QLD_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return QLD_ERR_NOTFOUND;
        Request &r = it->second;
        if (r.state != DONE && r.state != FAILED) return QLD_ERR_NOTREADY;
        if (r.out_off >= 0) s->arena.release(r.out_off, (int64_t)r.out_bytes);
        s->reqs.erase(it);
        s->gen++;
    }
    s->cv_work.notify_one();    // freed memory may admit waiting requests
    return QLD_OK;
}

extern "C" int qld_server_stats(const qld_server *s, int64_t *completed, int64_t *batches,
                                int64_t *pairs)
{
    if (!s) return QLD_ERR_INVALID;
    qld_server *m = const_cast<qld_server *>(s);
    std::lock_guard<std::mutex> lk(m->mu);
    if (completed) *completed = s->completed;
    if (batches) *batches = s->batches;
    if (pairs) *pairs = s->pairs;
    return s->error;
}

extern "C" int qld_host_alloc(qld_ctx *ctx, size_t bytes, void **ptr)
{
    if (!ctx || !ptr) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "qld_host_alloc");
}

extern "C" int qld_host_free(qld_ctx *ctx, void *ptr)
{
    if (!ctx) return QLD_ERR_INVALID;
    if (ptr) hipHostFree(ptr);
    return QLD_OK;
}
