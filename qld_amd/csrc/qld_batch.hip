// Resident tables and window plans: the batched path.
//
// The reference correlates many windows as independent tasks, each a full
// correlate() call that reloads its table and runs one GEMM
// (input_list.c:205-323, correlate_function.c:112-290).  Here a chromosome
// is uploaded and expanded once (qld_table), and a window set becomes a
// plan: one job descriptor per window, the tile work list of every window,
// and either one persistent MFMA launch per tile shape (draining per-XCD
// work queues) or -- as the baseline -- one launch per window.  Either
// launch sequence is captured in a HIP graph and replayed.
#include "qld_internal.h"
#include "qld_ctx.h"
#include "ld_mfma.h"

#include <algorithm>
#include <numeric>
#include <string.h>
#include <vector>

struct PlanClass {
    int tile = QLD_TILE_LARGE;
    bool wave_units = false;        // items are wave blocks claimed per wave
    bool shared = false;            // shared-pair units over one table triangle
    int grid = 0;                   // persistent workgroups
    size_t ndesc = 0, nitems = 0;
    WorkDesc *d_descs = nullptr;
    WorkItem *d_items = nullptr;
    uint32_t *d_queue = nullptr;
    WorkDesc gdesc;                 // shared: the table's triangle job
    uint32_t *d_wlist = nullptr;    // shared: window lists per unit
    WinOut *d_wins = nullptr;       // shared: window outputs
};

struct PlainLaunch {
    WorkDesc desc;
    int64_t item_off, item_cnt;
    int tile;
    int ksplit = 1;                 // split-K chunks per tile (direct 128-tiles)
};

struct qld_plan {
    int mode = QLD_BATCH_PERSISTENT;
    int direct = 0;                       // packed rows into registers (fp4)
    int engine = QLD_ENGINE_MFMA_FP4;
    int r2_mode = QLD_R2_REFERENCE;
    std::vector<PlanClass> classes;       // persistent mode
    std::vector<PlainLaunch> plain;       // plain mode
    WorkItem *d_items_all = nullptr;      // plain mode item array
    float *d_kpart = nullptr;             // plain mode split-K partial sums (launches in turn)
    uint32_t *d_kcnt = nullptr;           // plain mode split-K counters, zero between launches
    std::vector<std::pair<void *, size_t>> owned;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    int64_t tiles = 0, pairs = 0;
    int launches = 0;
};

static int table_alloc(qld_ctx *ctx, void **p, size_t bytes)
{
    return qld_ctx_malloc(ctx, bytes ? bytes : 1, p);
}

static void table_release(qld_ctx *ctx, qld_table *t)
{
    qld_ctx_mfree(ctx, t->packed, t->bytes_packed);
    qld_ctx_mfree(ctx, t->bc, t->bytes_bc);
    if (t->X) qld_ctx_mfree(ctx, t->X, t->bytes_x);
    delete t;
}

static int table_create(qld_ctx *ctx, const void *rows_src, const void *bc_src,
                        hipMemcpyKind kind, int rows, int words, int snp_size,
                        qld_table **out)
{
    if (!ctx || !out || !rows_src || !bc_src || rows <= 0 || words <= 0 ||
        words > QLD_MAX_WORDS || snp_size <= 0 || snp_size >= QLD_MAX_SNP_SIZE) {
        if (ctx) qld_set_error(ctx, "qld_table_create: bad arguments (rows=%d words=%d N=%d; "
                               "resident tables feed the MFMA engines: words 1 .. %d, N 1 .. %d)",
                               rows, words, snp_size, QLD_MAX_WORDS, QLD_MAX_SNP_SIZE - 1);
        return QLD_ERR_INVALID;
    }
    *out = nullptr;
    hipSetDevice(ctx->device);
    qld_table *t = new (std::nothrow) qld_table();
    if (!t) return QLD_ERR_NOMEM;
    t->rows = rows;
    t->words = words;
    t->snp_size = snp_size;
    t->engine = ctx->cfg.engine == QLD_ENGINE_MFMA_I8 ? QLD_ENGINE_MFMA_I8 : QLD_ENGINE_MFMA_FP4;
    // fp4 tables in direct mode keep only a row-block-tiled copy of the
    // packed rows (the kernels read it straight into registers); otherwise
    // the packed rows and an expanded operand copy.
    t->direct = t->engine == QLD_ENGINE_MFMA_FP4 && ctx->fp4_packed == QLD_FP4_DIRECT;
    t->row_bytes = t->direct ? 0 : mfma_row_bytes(t->engine, words);
    t->bytes_packed = t->direct ? (size_t)tiled_bytes(rows, words) : (size_t)rows * words * 8;
    t->bytes_bc = (size_t)rows * 4;
    t->bytes_x = (size_t)rows * t->row_bytes;
    int rc;
    if ((rc = table_alloc(ctx, (void **)&t->packed, t->bytes_packed)) ||
        (rc = table_alloc(ctx, (void **)&t->bc, t->bytes_bc)) ||
        (!t->direct && (rc = table_alloc(ctx, (void **)&t->X, t->bytes_x)))) {
        table_release(ctx, t);
        return rc;
    }
    hipError_t e = hipSuccess;
    uint64_t *staged = nullptr;           // host rows staged on the device for tiling
    const size_t row_major = (size_t)rows * words * 8;
    if (t->direct) {
        const uint64_t *src = (const uint64_t *)rows_src;
        if (kind == hipMemcpyHostToDevice) {
            if ((rc = qld_ctx_malloc(ctx, row_major, (void **)&staged))) {
                table_release(ctx, t);
                return rc;
            }
            e = hipMemcpyAsync(staged, rows_src, row_major, kind, ctx->stream);
            src = staged;
        }
        if (e == hipSuccess) e = launch_tile_rows(src, rows, words, t->packed, ctx->stream);
    } else {
        e = hipMemcpyAsync(t->packed, rows_src, t->bytes_packed, kind, ctx->stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->bc, bc_src, t->bytes_bc, kind, ctx->stream);
    if (e == hipSuccess && !t->direct)
        e = launch_expand(t->engine, t->packed, rows, words, t->X, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (staged) qld_ctx_mfree(ctx, staged, row_major);
    if (e != hipSuccess) {
        table_release(ctx, t);
        return qld_hip_fail(ctx, e, "qld_table_create");
    }
    *out = t;
    return QLD_OK;
}

extern "C" int qld_table_create(qld_ctx *ctx, const uint64_t *rows_host,
                                const uint32_t *bc_host, int rows, int words,
                                int snp_size, qld_table **out)
{
    return table_create(ctx, rows_host, bc_host, hipMemcpyHostToDevice, rows, words,
                        snp_size, out);
}

extern "C" int qld_table_create_dev(qld_ctx *ctx, const uint64_t *rows_dev,
                                    const uint32_t *bc_dev, int rows, int words,
                                    int snp_size, qld_table **out)
{
    return table_create(ctx, rows_dev, bc_dev, hipMemcpyDeviceToDevice, rows, words,
                        snp_size, out);
}

extern "C" int qld_table_free(qld_ctx *ctx, qld_table *t)
{
    if (!ctx) return QLD_ERR_INVALID;
    if (!t) return QLD_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    table_release(ctx, t);
    return QLD_OK;
}

extern "C" int qld_table_rows(const qld_table *t) { return t ? t->rows : 0; }

extern "C" size_t qld_table_bytes(const qld_table *t)
{
    return t ? t->bytes_packed + t->bytes_bc + t->bytes_x : 0;
}

// ---- plans ------------------------------------------------------------------

int qld_window_tile(const qld_ctx *ctx, int size)
{
    return size >= ctx->tile_cutoff ? QLD_TILE_LARGE : QLD_TILE_SMALL;
}

static int plan_upload(qld_ctx *ctx, qld_plan *p, void **dst, const void *src, size_t bytes)
{
    int rc = qld_ctx_malloc(ctx, bytes ? bytes : 8, dst);
    if (rc) return rc;
    p->owned.push_back({*dst, bytes ? bytes : 8});
    if (!bytes) return QLD_OK;
    hipError_t e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? QLD_OK : qld_hip_fail(ctx, e, "plan upload");
}

static int plan_alloc(qld_ctx *ctx, qld_plan *p, void **dst, size_t bytes)
{
    int rc = qld_ctx_malloc(ctx, bytes, dst);
    if (rc == QLD_OK) p->owned.push_back({*dst, bytes});
    return rc;
}

static void plan_release(qld_ctx *ctx, qld_plan *p)
{
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    for (auto &o : p->owned) qld_ctx_mfree(ctx, o.first, o.second);
    delete p;
}

static int persistent_grid(const qld_ctx *ctx, int tile)
{
    // 256-tiles use 128 KiB of LDS (one workgroup per CU); 128-tiles use
    // 64 KiB and ~216 registers per lane (two per CU).
    return ctx->num_cu * (tile == QLD_TILE_LARGE ? 1 : 2);
}

// Queue the plan's kernels on the context stream (direct, or under capture).
static hipError_t plan_issue(qld_ctx *ctx, qld_plan *p)
{
    hipError_t e = hipSuccess;
    WorkDesc none;
    memset(&none, 0, sizeof(none));
    if (p->mode == QLD_BATCH_PERSISTENT || p->mode == QLD_BATCH_SHARED) {
        for (auto &c : p->classes) {
            if (!c.nitems) continue;
            if (c.shared)
                e = launch_mfma_shared_units(p->r2_mode, c.gdesc, c.d_items, (int64_t)c.nitems,
                                             c.d_wlist, c.d_wins, c.grid, c.d_queue,
                                             ctx->stream);
            else if (c.wave_units)
                e = launch_mfma_wave_units(p->r2_mode, c.tile, c.d_descs, none, c.d_items,
                                           (int64_t)c.nitems, c.grid, c.d_queue, ctx->stream);
            else
                e = launch_mfma_items(p->engine, p->r2_mode, c.tile,
                                      p->direct ? QLD_FP4_DIRECT : 0, c.d_descs, none, c.d_items,
                                      (int64_t)c.nitems, c.grid, c.d_queue, nullptr, ctx->stream);
            if (e != hipSuccess) return e;
        }
    } else {
        for (auto &l : p->plain) {
            if (l.ksplit > 1)
                e = launch_mfma_splitk(p->r2_mode, nullptr, l.desc, p->d_items_all + l.item_off,
                                       l.item_cnt, l.ksplit, (l.desc.words + 3) / 4,
                                       p->d_kpart, p->d_kcnt, ctx->stream);
            else
                e = launch_mfma_items(p->engine, p->r2_mode, l.tile,
                                      p->direct ? QLD_FP4_DIRECT : 0, nullptr, l.desc,
                                      p->d_items_all + l.item_off, l.item_cnt, 0, nullptr,
                                      nullptr, ctx->stream);
            if (e != hipSuccess) return e;
        }
    }
    return e;
}

// One unit of a plan: rows i of [b_start, b_start + b_size) against columns
// j of [a_start, a_start + a_size) of the same table; r^2 of (i, j) at
// out[out_offset + i * a_size + j]; tri: a single region, j > i only.
struct PairWin {
    int a_start, a_size, b_start, b_size;
    int64_t out_offset;
    int tri;
};

static int64_t pair_count(const PairWin &w)
{
    return w.tri ? (int64_t)w.a_size * (w.a_size - 1) / 2 : (int64_t)w.a_size * w.b_size;
}

static bool pair_empty(const PairWin &w)
{
    return w.tri ? w.a_size < 2 : (w.a_size == 0 || w.b_size == 0);
}

static int pair_tile(const qld_ctx *ctx, const qld_table *t, const PairWin &w)
{
    const int tile = qld_window_tile(ctx, std::min(w.a_size, w.b_size));
    return t->direct ? tile : expanded_tile_cap(t->row_bytes, tile);
}

// Shared-pair plan: the 128 x 64 blocks of the table's triangle that hold a
// cell of some window, each listed once with the windows it serves, in row
// band order (a workgroup's four waves claim four consecutive blocks, mostly
// of one band, so they share their rows), each 8 bands' blocks by work.  A
// block at the ragged edge of the windows' union, where the windows hold
// cells only in its lower row blocks or its first 32 columns, computes only
// those (its sub-block form, in its window list's header word).  A window [s, e) holds cells of
// block (bi, bj) iff some row r in [s, e - 1) of the band has a column in
// (r, e) inside the block: bi from s/128 to (e-2)/128, and for each, bj from
// (r0 + 1)/64 to (e - 1)/64 with r0 the band's first row inside the window.
static bool shared_plan_fits(const qld_table *t, const std::vector<PairWin> &wins)
{
    int wtm = 0, wtn = 0;
    wave_block_shape(QLD_TILE_LARGE, &wtm, &wtn);
    if (!t->direct || (int64_t)t->rows > 65535LL * wtn) return false;
    for (const PairWin &w : wins)
        if (!w.tri) return false;
    return true;
}

static int build_shared_plan(qld_ctx *ctx, const qld_table *t, const std::vector<PairWin> &wins,
                             float *out, qld_plan *p)
{
    int wtm = 0, wtn = 0;
    wave_block_shape(QLD_TILE_LARGE, &wtm, &wtn);
    if (!t->direct || (int64_t)t->rows > 65535LL * wtn) {
        qld_set_error(ctx, "shared-pair plans need a direct fp4 table of at most %lld rows",
                      65535LL * wtn);
        return QLD_ERR_INVALID;
    }
    struct Ref {
        uint32_t bi, bj, w;
    };
    std::vector<Ref> refs;
    std::vector<WinOut> outs(wins.size());
    for (size_t k = 0; k < wins.size(); ++k) {
        const PairWin &w = wins[k];
        if (!w.tri) {
            qld_set_error(ctx, "shared-pair plans take single-region windows only (window %zu)",
                          k);
            return QLD_ERR_INVALID;
        }
        outs[k] = WinOut{w.a_start, w.a_size, out + w.out_offset};
        p->pairs += pair_count(w);
        if (pair_empty(w)) continue;
        const int s = w.a_start, e = w.a_start + w.a_size;
        for (int bi = s / wtm; bi <= (e - 2) / wtm; ++bi) {
            const int r0 = std::max(s, bi * wtm);
            for (int bj = (r0 + 1) / wtn; bj <= (e - 1) / wtn; ++bj)
                refs.push_back(Ref{(uint32_t)bi, (uint32_t)bj, (uint32_t)k});
        }
    }
    std::sort(refs.begin(), refs.end(), [](const Ref &x, const Ref &y) {
        return x.bi != y.bi ? x.bi < y.bi : x.bj != y.bj ? x.bj < y.bj : x.w < y.w;
    });
    // Which 32 x 32 blocks of unit (bi, bj) hold a cell of window [ws, we):
    // rows R of the block inside the window, columns C before its end, and
    // some j in C right of some i in R (min R < max C).
    auto needed = [&](int bi, int bj, int ws, int we) {
        unsigned mask = 0;
        for (int sa = 0; sa < wtm / 32; ++sa) {
            const int rlo = std::max(bi * wtm + 32 * sa, ws);
            const int rhi = std::min(bi * wtm + 32 * sa + 32, we) - 1;
            if (rlo > rhi) continue;
            for (int sb = 0; sb < wtn / 32; ++sb) {
                const int ca = bj * wtn + 32 * sb, cmax = std::min(ca + 32, we) - 1;
                if (ca <= cmax && rlo < cmax) mask |= 1u << (sa * 2 + sb);
            }
        }
        return mask;
    };
    // MFMA blocks the kernel computes for a unit in sub-block form `sub`
    // (the kernel's variants: diagonal blocks 3 (1 on the table's last
    // columns), blocks on the table's last columns 4, forms 1-3 and
    // QLD_SUB_HALF as named).
    auto work = [&](const WorkItem &u, unsigned sub) {
        const bool nfull = (int)u.tj * wtn + 32 < t->rows;
        if ((int)u.tj * wtn == (int)u.ti * wtm) return nfull ? 3 : 1;
        if (!nfull || sub == QLD_SUB_HALF) return 4;
        return sub ? (4 - (int)sub) * 2 : 8;
    };
    std::vector<WorkItem> items;
    std::vector<uint32_t> wl;
    std::vector<int> unit_work;
    for (size_t a = 0; a < refs.size();) {
        size_t b = a;
        unsigned mask = 0;
        while (b < refs.size() && refs[b].bi == refs[a].bi && refs[b].bj == refs[a].bj) {
            const PairWin &w = wins[refs[b].w];
            mask |= needed((int)refs[a].bi, (int)refs[a].bj, w.a_start, w.a_start + w.a_size);
            ++b;
        }
        const WorkItem u{(uint32_t)wl.size(), (uint16_t)refs[a].bi, (uint16_t)refs[a].bj};
        // the cheapest form that computes every needed block: skip leading
        // row blocks without one, or the second column half
        unsigned sub = 0;
        if ((int)u.tj * wtn != (int)u.ti * wtm) {
            unsigned ro = 0;
            while (ro < 3 && !((mask >> (2 * ro)) & 3u)) ++ro;
            if (ro && work(u, ro) < work(u, sub)) sub = ro;
            if (!(mask & 0xAAu) && work(u, QLD_SUB_HALF) < work(u, sub)) sub = QLD_SUB_HALF;
        }
        items.push_back(u);
        unit_work.push_back(work(u, sub));
        wl.push_back((uint32_t)(b - a) | (sub << 16));
        for (size_t k = a; k < b; ++k) wl.push_back(refs[k].w);
        a = b;
    }
    if (items.empty()) return QLD_OK;
    // A workgroup's four waves wait for the slowest block of their strip, so
    // within every run of 8 bands the blocks go in order of work, most first
    // (full blocks, then the partial forms, the diagonal blocks together).
    {
        std::vector<size_t> idx(items.size());
        std::iota(idx.begin(), idx.end(), 0);
        for (size_t a = 0; a < idx.size();) {
            size_t b = a;
            while (b < idx.size() && items[idx[b]].ti / 8 == items[idx[a]].ti / 8) ++b;
            std::stable_sort(idx.begin() + a, idx.begin() + b, [&](size_t x, size_t y) {
                return unit_work[x] > unit_work[y];
            });
            a = b;
        }
        std::vector<WorkItem> sorted(items.size());
        for (size_t k = 0; k < idx.size(); ++k) sorted[k] = items[idx[k]];
        items.swap(sorted);
    }
    PlanClass c;
    c.tile = QLD_TILE_LARGE;
    c.wave_units = true;
    c.shared = true;
    c.nitems = items.size();
    c.grid = (int)std::min<int64_t>(2LL * ctx->num_cu, ((int64_t)items.size() + 3) / 4);
    fill_work_desc(c.gdesc, nullptr, nullptr, t->bc, t->bc, t->rows, t->rows, 0, 1, t->snp_size,
                   nullptr, 0, nullptr, 0);
    set_tiled_operands(c.gdesc, t->packed, 0, t->rows, t->packed, 0, t->rows, t->words);
    int rc;
    if ((rc = plan_upload(ctx, p, (void **)&c.d_items, items.data(),
                          items.size() * sizeof(WorkItem))) ||
        (rc = plan_upload(ctx, p, (void **)&c.d_wlist, wl.data(), wl.size() * sizeof(uint32_t))) ||
        (rc = plan_upload(ctx, p, (void **)&c.d_wins, outs.data(),
                          outs.size() * sizeof(WinOut))) ||
        (rc = plan_alloc(ctx, p, (void **)&c.d_queue, 16 * sizeof(uint32_t))))
        return rc;
    p->tiles = (int64_t)items.size();
    p->launches = 1;
    p->classes.push_back(c);
    return QLD_OK;
}

static int build_plan(qld_ctx *ctx, const qld_table *t, const std::vector<PairWin> &wins,
                      float *out, int mode, qld_plan **plan)
{
    const int nwin = (int)wins.size();
    for (int k = 0; k < nwin; ++k) {
        const PairWin &w = wins[k];
        // tile and wave-block indices are 16-bit (64-row wave blocks)
        const int64_t lim = 65535LL * 64;
        if (w.a_start < 0 || w.a_size < 0 || w.b_start < 0 || w.b_size < 0 ||
            (int64_t)w.a_start + w.a_size > t->rows || (int64_t)w.b_start + w.b_size > t->rows ||
            w.out_offset < 0 || w.a_size > lim || w.b_size > lim ||
            (w.tri && (w.a_start != w.b_start || w.a_size != w.b_size))) {
            qld_set_error(ctx, "window %d (A %d+%d, B %d+%d, tri %d) invalid for a table of "
                          "%d rows", k, w.a_start, w.a_size, w.b_start, w.b_size, w.tri, t->rows);
            return QLD_ERR_INVALID;
        }
    }
    hipSetDevice(ctx->device);
    qld_plan *p = new (std::nothrow) qld_plan();
    if (!p) return QLD_ERR_NOMEM;
    p->mode = mode;
    p->engine = t->engine;
    p->direct = t->direct;
    p->r2_mode = ctx->cfg.r2_mode;

    // Windows in table order, so that each XCD's contiguous share of the
    // work list covers one stretch of the chromosome.
    std::vector<int> order(nwin);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return wins[a].b_start != wins[b].b_start ? wins[a].b_start < wins[b].b_start
                                                  : wins[a].a_start < wins[b].a_start;
    });

    auto desc_of = [&](const PairWin &w) {
        WorkDesc d;
        if (t->direct) {
            fill_work_desc(d, nullptr, nullptr, t->bc + w.b_start, t->bc + w.a_start, w.b_size,
                           w.a_size, 0, w.tri, t->snp_size, out + w.out_offset, w.a_size,
                           nullptr, w.a_size);
            set_tiled_operands(d, t->packed, w.b_start, t->rows, t->packed, w.a_start,
                               t->rows, t->words);
        } else {
            const uint8_t *XI = t->X + (int64_t)w.b_start * t->row_bytes;
            const uint8_t *XJ = t->X + (int64_t)w.a_start * t->row_bytes;
            fill_work_desc(d, XI, XJ, t->bc + w.b_start, t->bc + w.a_start, w.b_size, w.a_size,
                           t->row_bytes, w.tri, t->snp_size, out + w.out_offset, w.a_size,
                           nullptr, w.a_size);
        }
        return d;
    };

    int rc = QLD_OK;
    if (mode == QLD_BATCH_PERSISTENT && ctx->plan_shared && shared_plan_fits(t, wins))
        mode = QLD_BATCH_SHARED;
    p->mode = mode;
    if (mode == QLD_BATCH_SHARED) {
        if ((rc = build_shared_plan(ctx, t, wins, out, p))) {
            plan_release(ctx, p);
            return rc;
        }
    } else if (mode == QLD_BATCH_PERSISTENT) {
        // Direct fp4 operands: every wave claims wave blocks of its own
        // (unless the context asks for whole tiles per workgroup).
        const bool waves = p->direct && ctx->plan_sched == QLD_SCHED_WAVES;
        for (int tile : {QLD_TILE_LARGE, QLD_TILE_SMALL}) {
            std::vector<WorkDesc> descs;
            std::vector<WorkItem> items;
            int64_t ntiles = 0;
            int wtm = 0, wtn = 0;
            wave_block_shape(tile, &wtm, &wtn);
            for (int k : order) {
                const PairWin &w = wins[k];
                if (pair_empty(w) || pair_tile(ctx, t, w) != tile) continue;
                const size_t tiles = job_tile_count(w.b_size, w.a_size, w.tri, tile);
                size_t cnt = items.size();
                if (waves) {
                    items.resize(cnt + tiles * (size_t)(tile / wtm) * (size_t)(tile / wtn));
                    const size_t n = append_job_wave_units(items.data() + cnt,
                                                           (uint32_t)descs.size(), w.b_size,
                                                           w.a_size, w.tri, tile, wtm, wtn);
                    order_wave_units_by_work(items.data() + cnt, n, w.b_size, w.a_size, w.tri,
                                             wtm, wtn);
                    cnt += n;
                    items.resize(cnt);
                } else {
                    items.resize(cnt + tiles);
                    append_job_tiles(items.data(), &cnt, (uint32_t)descs.size(), w.b_size,
                                     w.a_size, w.tri, tile);
                }
                ntiles += (int64_t)tiles;
                descs.push_back(desc_of(w));
                p->pairs += pair_count(w);
            }
            if (items.empty()) continue;
            PlanClass c;
            c.tile = tile;
            c.wave_units = waves;
            c.ndesc = descs.size();
            c.nitems = items.size();
            // wave units: two 4-wave workgroups per CU for both tile edges
            c.grid = waves ? (int)std::min<int64_t>(2LL * ctx->num_cu,
                                                    ((int64_t)items.size() + 3) / 4)
                           : (int)std::min<int64_t>(persistent_grid(ctx, tile),
                                                    (int64_t)items.size());
            if ((rc = plan_upload(ctx, p, (void **)&c.d_descs, descs.data(),
                                  descs.size() * sizeof(WorkDesc))) ||
                (rc = plan_upload(ctx, p, (void **)&c.d_items, items.data(),
                                  items.size() * sizeof(WorkItem))) ||
                (rc = plan_alloc(ctx, p, (void **)&c.d_queue, 16 * sizeof(uint32_t)))) {
                plan_release(ctx, p);
                return rc;
            }
            p->tiles += ntiles;
            p->launches += 1;
            p->classes.push_back(c);
        }
    } else {
        // One launch per window, each as a plain call on the same shape
        // would run it: direct 128-tile windows split their k loops when
        // the split model says so (qld_ksplit_for); the launches run in
        // turn, so they share one split-K scratch.
        std::vector<WorkItem> items;
        size_t part_max = 0;
        int64_t cnt_max = 0;
        for (int k : order) {
            const PairWin &w = wins[k];
            if (pair_empty(w)) continue;
            const int tile = pair_tile(ctx, t, w);
            PlainLaunch l;
            l.tile = tile;
            l.item_off = (int64_t)items.size();
            size_t cnt = items.size();
            items.resize(cnt + job_tile_count(w.b_size, w.a_size, w.tri, tile));
            append_job_tiles(items.data(), &cnt, 0, w.b_size, w.a_size, w.tri, tile);
            l.item_cnt = (int64_t)cnt - l.item_off;
            l.desc = desc_of(w);
            if (t->direct && tile == QLD_TILE_SMALL)
                l.ksplit = qld_ksplit_for(ctx, l.item_cnt, (t->words + 3) / 4);
            if (l.ksplit > 1) {
                part_max = std::max(part_max, splitk_part_bytes(l.item_cnt, l.ksplit));
                cnt_max = std::max(cnt_max, l.item_cnt);
            }
            p->plain.push_back(l);
            p->pairs += pair_count(w);
        }
        p->tiles = (int64_t)items.size();
        p->launches = (int)p->plain.size();
        if ((rc = plan_upload(ctx, p, (void **)&p->d_items_all, items.data(),
                              items.size() * sizeof(WorkItem)))) {
            plan_release(ctx, p);
            return rc;
        }
        if (cnt_max > 0) {
            // counters zeroed once here, before any capture; every launch
            // leaves them zero
            if ((rc = plan_alloc(ctx, p, (void **)&p->d_kpart, part_max)) ||
                (rc = plan_alloc(ctx, p, (void **)&p->d_kcnt, (size_t)cnt_max * 4)) ||
                (rc = qld_fill_async(ctx, p->d_kcnt, 0u, (size_t)cnt_max))) {
                plan_release(ctx, p);
                return rc;
            }
        }
    }

    // Capture the launch sequence once; replays cost one graph launch.
    if (!ctx->plan_graphs) {
        p->use_graph = false;
        *plan = p;
        return QLD_OK;
    }
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t ei = plan_issue(ctx, p);
        e = hipStreamEndCapture(ctx->stream, &p->graph);
        if (ei != hipSuccess) e = ei;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p->use_graph = false;     // direct launches still work
        if (p->exec) hipGraphExecDestroy(p->exec);
        if (p->graph) hipGraphDestroy(p->graph);
        p->exec = nullptr;
        p->graph = nullptr;
    }
    *plan = p;
    return QLD_OK;
}

extern "C" int qld_plan_windows(qld_ctx *ctx, const qld_table *t, const qld_window *wins,
                                int nwin, float *out, int mode, qld_plan **plan)
{
    if (!ctx || !t || !plan || nwin < 0 || (nwin && (!wins || !out)) ||
        (mode != QLD_BATCH_PLAIN && mode != QLD_BATCH_PERSISTENT && mode != QLD_BATCH_SHARED)) {
        if (ctx) qld_set_error(ctx, "qld_plan_windows: bad arguments");
        return QLD_ERR_INVALID;
    }
    *plan = nullptr;
    std::vector<PairWin> pw((size_t)nwin);
    for (int k = 0; k < nwin; ++k)
        pw[k] = PairWin{wins[k].start, wins[k].size, wins[k].start, wins[k].size,
                        wins[k].out_offset, 1};
    return build_plan(ctx, t, pw, out, mode, plan);
}

extern "C" int qld_plan_pairs(qld_ctx *ctx, const qld_table *t, const qld_pair_window *wins,
                              int nwin, float *out, int mode, qld_plan **plan)
{
    if (!ctx || !t || !plan || nwin < 0 || (nwin && (!wins || !out)) ||
        (mode != QLD_BATCH_PLAIN && mode != QLD_BATCH_PERSISTENT && mode != QLD_BATCH_SHARED)) {
        if (ctx) qld_set_error(ctx, "qld_plan_pairs: bad arguments");
        return QLD_ERR_INVALID;
    }
    *plan = nullptr;
    std::vector<PairWin> pw((size_t)nwin);
    for (int k = 0; k < nwin; ++k)
        pw[k] = PairWin{wins[k].a_start, wins[k].a_size, wins[k].b_start, wins[k].b_size,
                        wins[k].out_offset, wins[k].tri ? 1 : 0};
    return build_plan(ctx, t, pw, out, mode, plan);
}

extern "C" int qld_plan_run(qld_ctx *ctx, qld_plan *p)
{
    if (!ctx || !p) return QLD_ERR_INVALID;
    hipSetDevice(ctx->device);
    if (ctx->cfg.timing) hipEventRecord(ctx->ev0, ctx->stream);
    if (ctx->cfg.timing) hipEventRecord(ctx->kev0, ctx->stream);
    hipError_t e = (p->use_graph && p->exec) ? hipGraphLaunch(p->exec, ctx->stream)
                                             : plan_issue(ctx, p);
    if (e != hipSuccess) return qld_hip_fail(ctx, e, "qld_plan_run");
    if (ctx->cfg.timing) {
        hipEventRecord(ctx->kev1, ctx->stream);
        hipEventRecord(ctx->ev1, ctx->stream);
        ctx->timed = true;
    }
    ctx->last_engine = p->engine;
    return QLD_OK;
}

extern "C" int qld_plan_use_graph(qld_plan *p, int use_graph)
{
    if (!p) return QLD_ERR_INVALID;
    if (use_graph && !p->exec) return QLD_ERR_NOTFOUND;
    p->use_graph = use_graph != 0;
    return QLD_OK;
}

extern "C" int qld_plan_stats(const qld_plan *p, int64_t *tiles, int64_t *pairs,
                              int *launches)
{
    if (!p) return QLD_ERR_INVALID;
    if (tiles) *tiles = p->tiles;
    if (pairs) *pairs = p->pairs;
    if (launches) *launches = p->launches;
    return QLD_OK;
}

extern "C" int qld_plan_free(qld_ctx *ctx, qld_plan *p)
{
    if (!ctx) return QLD_ERR_INVALID;
    if (!p) return QLD_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    plan_release(ctx, p);
    return QLD_OK;
}

// ---- tile-shape autotuner ---------------------------------------------------

extern "C" int qld_set_plan_schedule(qld_ctx *ctx, int sched)
{
    if (!ctx || (sched != QLD_SCHED_TILES && sched != QLD_SCHED_WAVES)) return QLD_ERR_INVALID;
    ctx->plan_sched = sched;
    return QLD_OK;
}

extern "C" int qld_set_plan_graphs(qld_ctx *ctx, int capture)
{
    if (!ctx) return QLD_ERR_INVALID;
    ctx->plan_graphs = capture != 0;
    return QLD_OK;
}

extern "C" int qld_set_plan_shared(qld_ctx *ctx, int shared)
{
    if (!ctx) return QLD_ERR_INVALID;
    ctx->plan_shared = shared != 0;
    return QLD_OK;
}

extern "C" int qld_get_plan_shared(const qld_ctx *ctx)
{
    return ctx ? (int)ctx->plan_shared : -1;
}

extern "C" int qld_get_plan_schedule(const qld_ctx *ctx)
{
    return ctx ? ctx->plan_sched : -1;
}

extern "C" int qld_set_tile_cutoff(qld_ctx *ctx, int cutoff)
{
    if (!ctx || cutoff < 0) return QLD_ERR_INVALID;
    ctx->tile_cutoff = cutoff;
    return QLD_OK;
}

extern "C" int qld_get_tile_cutoff(const qld_ctx *ctx)
{
    return ctx ? ctx->tile_cutoff : -1;
}

// Time the window set as persistent plans under each candidate cutoff
// (windows of at least `cutoff` SNPs on 256-tiles, the rest on 128-tiles)
// and, for direct fp4 tables, each work schedule (wave blocks claimed per
// wave, or whole tiles per workgroup), and keep the fastest pair for later
// plans of this context.
extern "C" int qld_autotune_windows(qld_ctx *ctx, const qld_table *t, const qld_window *wins,
                                    int nwin, float *out, int reps, int *best_cutoff,
                                    float *best_ms)
{
    if (!ctx || !t || nwin <= 0 || !wins || !out) return QLD_ERR_INVALID;
    if (reps <= 0) reps = 3;
    static const int candidates[] = {0, 256, 384, 512, 768, 1024, 1536, 2048, 1 << 30};
    const int saved = ctx->tile_cutoff, saved_sched = ctx->plan_sched;
    int best = saved, best_sched = saved_sched;
    float best_t = 3.4e38f;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return QLD_ERR_DEVICE;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return QLD_ERR_DEVICE;
    }
    // Every candidate plan is built first and then timed in two rounds
    // (fastest round kept), after a warm-up: the GPU's clocks ramp up over
    // the first launches after idle, which would otherwise penalize
    // whichever candidate runs first.
    struct Cand {
        int sched, cut;
        qld_plan *p;
        float ms;
    };
    std::vector<Cand> cands;
    int rc = QLD_OK;
    const bool saved_shared = ctx->plan_shared;
    ctx->plan_shared = false;          // the per-window candidates first
    const int nsched = t->direct ? 2 : 1;
    for (int si = 0; si < nsched && rc == QLD_OK; ++si) {
        ctx->plan_sched = t->direct ? (si ? QLD_SCHED_TILES : QLD_SCHED_WAVES) : saved_sched;
        for (int cut : candidates) {
            ctx->tile_cutoff = cut;
            qld_plan *p = nullptr;
            if ((rc = qld_plan_windows(ctx, t, wins, nwin, out, QLD_BATCH_PERSISTENT, &p))) break;
            cands.push_back(Cand{ctx->plan_sched, cut, p, 3.4e38f});
        }
    }
    // and the shared-pair form, where it applies (its tile shape is fixed)
    ctx->plan_sched = saved_sched;
    ctx->tile_cutoff = saved;
    if (rc == QLD_OK) {
        std::vector<PairWin> pw((size_t)nwin);
        for (int k = 0; k < nwin; ++k)
            pw[k] = PairWin{wins[k].start, wins[k].size, wins[k].start, wins[k].size,
                            wins[k].out_offset, 1};
        if (shared_plan_fits(t, pw)) {
            qld_plan *p = nullptr;
            rc = qld_plan_windows(ctx, t, wins, nwin, out, QLD_BATCH_SHARED, &p);
            if (rc == QLD_OK) cands.push_back(Cand{-1, -1, p, 3.4e38f});
        }
    }
    for (int w = 0; w < 3 * reps && rc == QLD_OK && !cands.empty(); ++w)
        rc = qld_plan_run(ctx, cands[0].p);
    for (int round = 0; round < 2 && rc == QLD_OK; ++round)
        for (auto &c : cands) {
            if ((rc = qld_plan_run(ctx, c.p))) break;           // warm-up
            hipEventRecord(e0, ctx->stream);
            for (int r = 0; r < reps && rc == QLD_OK; ++r) rc = qld_plan_run(ctx, c.p);
            hipEventRecord(e1, ctx->stream);
            float ms = 0.0f;
            if (rc == QLD_OK && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                c.ms = std::min(c.ms, ms / reps);
            if (rc) break;
        }
    bool best_shared = saved_shared;
    for (auto &c : cands) {
        if (rc == QLD_OK && c.ms < best_t) {
            best_t = c.ms;
            best_shared = c.cut < 0;
            if (!best_shared) {
                best = c.cut;
                best_sched = c.sched;
            }
        }
        qld_plan_free(ctx, c.p);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // the per-window settings stay those of the fastest per-window candidate
    // (they serve window sets the shared form cannot take)
    if (rc == QLD_OK && best_shared) {
        float bt = 3.4e38f;
        for (auto &c : cands)
            if (c.cut >= 0 && c.ms < bt) {
                bt = c.ms;
                best = c.cut;
                best_sched = c.sched;
            }
    }
    ctx->tile_cutoff = rc == QLD_OK ? best : saved;
    ctx->plan_sched = rc == QLD_OK ? best_sched : saved_sched;
    ctx->plan_shared = rc == QLD_OK ? best_shared : saved_shared;
    if (best_cutoff) *best_cutoff = ctx->tile_cutoff;
    if (best_ms) *best_ms = best_t;
    return rc;
}
