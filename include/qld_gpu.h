/*
 * qld_gpu.h -- C ABI of the MI355X LD engine (libqld_hip.so).
 *
 * The engine replaces the OpenCL GPU path of qLD-compute: the pairwise
 * AND+popcount "GEMM" over packed haplotype rows and the host-side r^2
 * epilogue.  All entry points are plain C: pointers, sizes and int status
 * codes; no HIP or C++ types cross this boundary.  No entry point calls
 * exit(); every failure is reported as a negative status and a message
 * retrievable with qld_gpu_last_error().
 *
 * Data layout shared with the reference (include/header.h:60-62,
 * include/correlate_MEM.h:66-75 of the reference):
 *   table    row-major uint64 words, one SNP per row, `words` = ceil(N/64)
 *            words per row, sample s at bit 63-(s%64) of word s/64, the
 *            low bits of the last word zero (snip_processing.c:783-815)
 *   bc       uint32 popcount of each row (the MDF "BC" column)
 *   results  float, results[i*ldr + j] = r^2(A_j, B_i): row i indexes table
 *            B, column j indexes table A (gemm_cpu.c:353-363)
 *   counts   uint32, same indexing as results: popcount(A_j & B_i)
 * Triangular mode (A == B, posWset2 == 0 in the reference) defines the cells
 * j > i only -- exactly the cells writeResults prints
 * (correlate_IO.c:403-406); other cells are left untouched.
 */
#ifndef QLD_GPU_H
#define QLD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define QLD_OK            0
#define QLD_ERR_INVALID  (-1)   /* bad argument or shape                  */
#define QLD_ERR_NOMEM    (-2)   /* device or host allocation failed       */
#define QLD_ERR_DEVICE   (-3)   /* HIP runtime / launch error             */
#define QLD_ERR_BUDGET   (-4)   /* request exceeds the HBM budget         */
#define QLD_ERR_NOTREADY (-5)   /* asynchronous result not ready yet      */
#define QLD_ERR_NOTFOUND (-6)   /* unknown handle / request id            */

/* ---- contraction engines --------------------------------------------- */
#define QLD_ENGINE_AUTO      0  /* pick per shape (autotuned table)       */
#define QLD_ENGINE_POPCOUNT  1  /* VALU AND + v_bcnt over packed words    */
#define QLD_ENGINE_MFMA_I8   2  /* int8 MFMA over unpacked 0/1 bytes      */
#define QLD_ENGINE_MFMA_FP4  3  /* fp4 (e2m1) MFMA over unpacked 0/1      */

/* ---- r^2 epilogue modes ------------------------------------------------ */
/* REFERENCE: bit-identical to get_pairwise_ld_score (gemm_cpu.c:343-404)
 *            as the reference's own build compiles it (float p's, float
 *            numerator, double denominator and division).
 * EXACT:     integer-exact D*N^2 = c*N - a*b, double ratio, rounded once
 *            to float (<= ~1 float ulp from the FP64 closed form).
 * Both keep the reference's special cases: a or b == 0 -> 0; a or b == N ->
 * NaN (never printed); r^2 > 1.0001 -> 123.456. */
#define QLD_R2_REFERENCE 0
#define QLD_R2_EXACT     1

typedef struct qld_ctx qld_ctx;

typedef struct qld_cfg {
    int    device;      /* HIP device ordinal                              */
    int    engine;      /* QLD_ENGINE_*                                    */
    int    r2_mode;     /* QLD_R2_*                                        */
    int    timing;      /* nonzero: record HIP events around each kernel   */
    size_t hbm_budget;  /* bytes the context may allocate; 0 = 90% of free */
} qld_cfg;

/* Fill *cfg with defaults (device 0, AUTO engine, REFERENCE r^2). */
void qld_cfg_default(qld_cfg *cfg);

/* Replaces gpu_init() (gemm_gpu.h:52, gemm_gpu.c:298-447): binds the
 * device, creates the context's stream and workspace.  Thread-safe; one
 * context per calling thread is the intended use. */
int qld_gpu_init(qld_ctx **ctx, const qld_cfg *cfg);

/* Replaces gpu_release() (gemm_gpu.h:54, gemm_gpu.c:886-906 frees). */
void qld_gpu_release(qld_ctx *ctx);

/* Message of the last failure on this context ("" if none); ctx may be
 * NULL for failures of qld_gpu_init itself. */
const char *qld_gpu_last_error(const qld_ctx *ctx);

/* Replaces the compute part of correlate_gpu() (gemm_gpu.h:56-70,
 * gemm_gpu.c:753-845: gpu_gemm + get_pairwise_ld_score_gpu).  Host
 * buffers, synchronous.  A: m x words, B: n x words (B may equal A);
 * triangular requires A == B content and m == n.  results (n*m floats) and
 * counts (n*m uint32) are each optional (NULL = not wanted).  snp_size is
 * the N of the frequencies, any positive int as in the reference: the
 * header's sample count, as correlate_function_gpu passes it
 * (correlate_function_gpu.c:240-275), which exceeds the rows' bits when a
 * sample list selects fewer samples.  The MFMA engines take N < 2^24 and
 * rows of at most 2^18 words (their fp32 sums are exact below 2^24); larger
 * jobs run on the popcount engine (uint32 counts, rows of up to 2^25 words),
 * which AUTO picks for them, and an explicit MFMA engine rejects them. */
int qld_gpu_correlate(qld_ctx *ctx,
                      const uint64_t *A, const uint32_t *bcA, int m,
                      const uint64_t *B, const uint32_t *bcB, int n,
                      int words, int snp_size, int triangular,
                      float *results, uint32_t *counts);

/* ---- device-resident interface ----------------------------------------- */
/* Device memory owned by the context (counted against hbm_budget). */
int qld_dev_alloc(qld_ctx *ctx, size_t bytes, void **dptr);
int qld_dev_free(qld_ctx *ctx, void *dptr);
int qld_h2d(qld_ctx *ctx, void *dst, const void *src, size_t bytes);
int qld_d2h(qld_ctx *ctx, void *dst, const void *src, size_t bytes);
int qld_dev_fill(qld_ctx *ctx, void *dptr, uint32_t value, size_t n_u32);
/* Device-to-device copy by a plain streaming kernel on the context's stream
 * (asynchronous); `bytes` must be a multiple of 16.  Also the known-size
 * reference when calibrating HBM traffic counters. */
int qld_dev_copy(qld_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Wait for all work queued on the context's stream. */
int qld_sync(qld_ctx *ctx);

/* Asynchronous on the context's stream: same computation as
 * qld_gpu_correlate on device buffers.  ldr / ldc are the row pitches (in
 * elements) of results / counts; either output may be NULL. */
int qld_correlate_dev(qld_ctx *ctx,
                      const uint64_t *dA, const uint32_t *dbcA, int m,
                      const uint64_t *dB, const uint32_t *dbcB, int n,
                      int words, int snp_size, int triangular,
                      float *dresults, int64_t ldr,
                      uint32_t *dcounts, int64_t ldc);

/* Allocation-free variant of qld_correlate_dev: same arguments, same results
 * bit for bit, and it allocates nothing (no hipMalloc/hipFree, no host
 * allocation, no memset node).  No counterpart in the reference: it is for
 * callers that cannot allocate -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation.  Its scratch (queue heads, tile
 * list, row-block-tiled operand copies) is a static 64 MiB device array per
 * device, held from library load to process end; shapes whose scratch would
 * exceed it return QLD_ERR_BUDGET (at 16,622 samples: a triangle of up to
 * about 32,000 rows, or about 32,000 rows of A and B together; the scratch
 * holds 2,080 bytes per row there).  Runs the popcount
 * engine, or the fp4 engine with direct operands (the defaults); other
 * engine / operand settings return QLD_ERR_INVALID.  All calls on one device
 * share its scratch: calls that overlap on a device (several contexts or
 * threads) race and must be serialized by the caller. */
int qld_correlate_dev_noalloc(qld_ctx *ctx,
                              const uint64_t *dA, const uint32_t *dbcA, int m,
                              const uint64_t *dB, const uint32_t *dbcB, int n,
                              int words, int snp_size, int triangular,
                              float *dresults, int64_t ldr,
                              uint32_t *dcounts, int64_t ldc);

/* Milliseconds of device time of the kernels queued by the last
 * qld_correlate_dev / qld_gpu_correlate call (needs cfg.timing); waits for
 * them.  Mirrors the reference's per-phase GEMM timer (gemm_gpu.c:823-829). */
int qld_last_kernel_ms(qld_ctx *ctx, float *ms);

/* Milliseconds of the contraction kernel alone (popcount or MFMA, with
 * its fused epilogue) of the last call; needs cfg.timing; waits for it. */
int qld_last_contract_ms(qld_ctx *ctx, float *ms);

/* Turn cfg.timing on or off for later calls.  Each timed call records four
 * HIP events on the stream, about 15-30 us per call on the 30,000-SNP
 * region; leave it off where only throughput matters. */
int qld_set_timing(qld_ctx *ctx, int on);

/* Device-time stopwatch over a region of the context's stream: start
 * records a HIP event on the stream, stop records a second one, waits for
 * it and returns the elapsed milliseconds. */
int qld_timer_start(qld_ctx *ctx);
int qld_timer_stop(qld_ctx *ctx, float *ms);

/* Operand source of the fp4 engine (plain calls, and tables created after
 * the call): DIRECT (default) reads the packed rows, from a row-block-tiled
 * copy made per call (or held by the table), straight into registers and
 * splits bits into e2m1 planes there (no LDS ring, no expanded copy, 1 bit
 * of memory traffic per sample); EXPANDED first writes an fp4
 * copy of the tables (4 bits per sample) that the kernel streams into LDS by
 * DMA; PACKED reads the packed rows into LDS and expands them there
 * (128-tiles). */
#define QLD_FP4_DIRECT   2
#define QLD_FP4_PACKED   1
#define QLD_FP4_EXPANDED 0
int qld_set_fp4_operands(qld_ctx *ctx, int mode);

/* Tile edge of the MFMA engines for plain calls: 256, 128, or 0 for the
 * default (direct fp4: 256 once the job has at least 4 such tiles per CU,
 * else 128; expanded operands: 128). */
int qld_set_plain_tile(qld_ctx *ctx, int tile);

/* Split-K for plain calls on 128-tiles of the direct fp4 engine (and for the
 * serving batches and one-launch-per-window plans of the same context): each tile's
 * word groups are summed by `chunks` workgroups and the last to finish adds
 * the partial sums (exact integers) and writes the tile.  0 (default): by
 * shape -- the chunk count that a cost model of the tiles' k-loop chains and
 * the per-chunk overhead puts fastest, with at least 4 word groups per chunk
 * (jobs of a few hundred tiles or fewer are split; large ones are not); 1: never;
 * 2 .. 64: always that many (fewer where the partial sums would pass 1 GiB).
 * The no-allocation entry point never splits.  Returns QLD_ERR_INVALID
 * outside 0 .. 64.  qld_last_ksplit gives the chunks per tile of the last
 * plain MFMA call (1: not split). */
int qld_set_ksplit(qld_ctx *ctx, int chunks);
int qld_last_ksplit(const qld_ctx *ctx);

/* Engine the last call used (QLD_ENGINE_*), after AUTO resolution. */
int qld_last_engine(const qld_ctx *ctx);

/* ---- resident tables ------------------------------------------------------ */
/* A packed table kept in HBM together with its expanded matrix-core operand
 * form (fp4 or int8, per the context's engine; fp4 for AUTO/POPCOUNT too, so
 * windows and requests can always run on the matrix cores).  Replaces the
 * reference's per-task table reload (correlate_function.c:155-228) for
 * callers that correlate many windows of one chromosome. */
typedef struct qld_table qld_table;

/* Copy `rows` packed rows and their bit counts from host memory into device
 * memory, in the layout the context's engine reads: a row-block-tiled copy
 * of the same size for the default fp4 direct operands, the packed rows plus
 * an expanded fp4 / int8 copy otherwise.  The caller's rows are not kept.
 * Tables feed the MFMA engines: words 1 .. 2^18, snp_size 1 .. 2^24 - 1. */
int qld_table_create(qld_ctx *ctx, const uint64_t *rows_host,
                     const uint32_t *bc_host, int rows, int words,
                     int snp_size, qld_table **out);
/* Same from device buffers (copied). */
int qld_table_create_dev(qld_ctx *ctx, const uint64_t *rows_dev,
                         const uint32_t *bc_dev, int rows, int words,
                         int snp_size, qld_table **out);
int qld_table_free(qld_ctx *ctx, qld_table *t);
/* Rows, words and device bytes held by a table. */
int qld_table_rows(const qld_table *t);
size_t qld_table_bytes(const qld_table *t);

/* ---- batched windows ------------------------------------------------------ */
/* One window = rows [start, start + size) of a resident table, correlated
 * with itself; its r^2 triangle (cells j > i of a size x size row-major
 * block, pitch `size`) goes to out + out_offset (floats).  This is the
 * reference's many-task mode (input_list.c:205-323, split_region:17-39),
 * one correlate() per window. */
typedef struct qld_window {
    int32_t start;
    int32_t size;
    int64_t out_offset;
} qld_window;

#define QLD_BATCH_PLAIN      0  /* one kernel launch per window             */
#define QLD_BATCH_PERSISTENT 1  /* every window's tiles in one persistent
                                   launch with per-XCD work queues         */
#define QLD_BATCH_SHARED     2  /* one persistent launch over the table's
                                   triangle: each pair that any window holds
                                   is counted once and its r^2 written into
                                   every window holding it (overlapping
                                   windows share work); single-region windows
                                   of a direct fp4 table only               */

typedef struct qld_plan qld_plan;

/* Prepare a window set: validates windows, picks tile shapes (autotuned
 * table), uploads job descriptors and the tile work list, and captures the
 * launch sequence in a HIP graph.  `out` is a device buffer of at least
 * max(out_offset + size*size) floats. */
int qld_plan_windows(qld_ctx *ctx, const qld_table *t, const qld_window *wins,
                     int nwin, float *out, int mode, qld_plan **plan);
/* A general unit of a plan, for the reference's two-window and two-input
 * tasks (load_balancer.c:122-357, correlate_function.c:239-290): rows i of
 * table rows [b_start, b_start + b_size) against columns j of
 * [a_start, a_start + a_size) of the same table; r^2 of (i, j) is written to
 * out[out_offset + i * a_size + j].  tri = 1 is a single region (a == b):
 * only j > i is defined. */
typedef struct qld_pair_window {
    int32_t a_start, a_size;
    int32_t b_start, b_size;
    int64_t out_offset;
    int32_t tri;
    int32_t reserved;
} qld_pair_window;

/* As qld_plan_windows, for pair windows. */
int qld_plan_pairs(qld_ctx *ctx, const qld_table *t, const qld_pair_window *wins,
                   int nwin, float *out, int mode, qld_plan **plan);

/* Queue one execution of the plan on the context's stream (asynchronous). */
int qld_plan_run(qld_ctx *ctx, qld_plan *plan);
/* Replay through the captured HIP graph (nonzero) or by direct launches. */
int qld_plan_use_graph(qld_plan *plan, int use_graph);
/* Whether plans built afterwards capture their launches in a HIP graph
 * (default 1); without one, qld_plan_run launches the kernels directly. */
int qld_set_plan_graphs(qld_ctx *ctx, int capture);
/* Tiles and pairs the plan computes; kernel launches per run.  Pairs count
 * every window's pairs (the outputs); tiles are work units (256 x 256 tiles,
 * or for QLD_BATCH_SHARED the 128 x 64 blocks counted once). */
int qld_plan_stats(const qld_plan *plan, int64_t *tiles, int64_t *pairs,
                   int *launches);
int qld_plan_free(qld_ctx *ctx, qld_plan *plan);

/* ---- serving ----------------------------------------------------------------- */
/* A request loop for LD queries against resident tables.  Each request is a
 * window (rows [start, start + size) of one table) whose r^2 triangle is
 * computed on the device.  A dispatcher thread admits pending requests in
 * arrival order while their outputs fit the output budget, packs them into
 * one launch (one job descriptor per request), and keeps up to `streams`
 * such batches in flight on separate HIP streams, so requests that arrive
 * while others run are admitted as slots and memory free up.  Each result
 * is bit-identical whatever it was batched with.  No counterpart in the
 * reference (its tasks run one by one, correlate_function_gpu.c:55-81). */
typedef struct qld_server qld_server;

typedef struct qld_server_cfg {
    size_t out_budget;   /* device bytes for outputs in flight (0 = 4 GiB)   */
    int    max_batch;    /* requests per launch (0 = 256)                    */
    int    streams;      /* batches in flight (0 = 3)                        */
    int    reserved;
} qld_server_cfg;

/* Request flags. */
#define QLD_REQ_KEEP    0  /* result stays on the device until fetched      */
#define QLD_REQ_COPY    1  /* copy into host_out (pinned memory from
                              qld_host_alloc is fastest), then release     */
#define QLD_REQ_DISCARD 2  /* release at completion (latency measurement)  */

int qld_server_create(qld_ctx *ctx, const qld_server_cfg *cfg, qld_server **out);
/* Stop accepting, finish everything in flight, free all results. */
int qld_server_destroy(qld_server *s);
/* Queue a request (thread-safe, non-blocking).  QLD_ERR_BUDGET if its
 * output (size*size floats) can never fit the budget. */
int qld_server_submit(qld_server *s, const qld_table *t, int start, int size,
                      int flags, float *host_out, int64_t *req_id);
/* Wait up to timeout_ms (< 0: forever) for completion; QLD_ERR_NOTREADY on
 * timeout.  latency_ms (optional) = submit to result ready. */
int qld_server_wait(qld_server *s, int64_t req_id, double timeout_ms,
                    double *latency_ms);
/* Copy a completed QLD_REQ_KEEP result (size*size floats, cells j > i
 * defined) to host memory and release it. */
int qld_server_fetch(qld_server *s, int64_t req_id, float *dst);
/* Forget a completed request (and release a kept result). */
int qld_server_release(qld_server *s, int64_t req_id);
/* Totals since creation: requests completed, batches launched, pairs. */
int qld_server_stats(const qld_server *s, int64_t *completed, int64_t *batches,
                     int64_t *pairs);

/* Pinned host memory for QLD_REQ_COPY destinations and fast copies. */
int qld_host_alloc(qld_ctx *ctx, size_t bytes, void **ptr);
int qld_host_free(qld_ctx *ctx, void *ptr);

/* Work scheduling of persistent plans with direct fp4 operands.  WAVES
 * (default): the work list holds wave blocks (64 x 64 in 128-tiles,
 * 128 x 64 in 256-tiles) with something to compute, and every wave claims
 * its own, so no wave waits for the others of its workgroup on a diagonal
 * or ragged-edge tile.  TILES: a workgroup claims a whole tile.  Plans built
 * afterwards use the setting. */
#define QLD_SCHED_TILES 0
#define QLD_SCHED_WAVES 1
int qld_set_plan_schedule(qld_ctx *ctx, int sched);
int qld_get_plan_schedule(const qld_ctx *ctx);

/* Tile-shape policy of window plans: windows of at least `cutoff` SNPs use
 * 256 x 256 tiles, smaller ones 128 x 128 (0 = all large). */
int qld_set_tile_cutoff(qld_ctx *ctx, int cutoff);
int qld_get_tile_cutoff(const qld_ctx *ctx);
/* Whether persistent plans (QLD_BATCH_PERSISTENT) of single-region windows
 * on a direct fp4 table take the shared-pair form of QLD_BATCH_SHARED
 * (default 0; the autotuner sets it when that form is faster). */
int qld_set_plan_shared(qld_ctx *ctx, int shared);
int qld_get_plan_shared(const qld_ctx *ctx);
/* Autotuner: time the given window set as persistent plans under each
 * candidate cutoff -- and, for tables with direct fp4 operands, under both
 * work schedules and in the shared-pair form -- (`reps` runs each after a
 * warm-up), keep the fastest cutoff and schedule of the per-window forms
 * and whether the shared-pair form beat them in the context, and report the
 * cutoff (qld_get_plan_schedule and qld_get_plan_shared give the rest).
 * Writes r^2 into `out` as a plan run does. */
int qld_autotune_windows(qld_ctx *ctx, const qld_table *t, const qld_window *wins,
                         int nwin, float *out, int reps, int *best_cutoff,
                         float *best_ms);

/* ---- report selection on the device ------------------------------------
 *
 * The reference prints pair (i, j) -- i a row of table B, j a column of
 * table A, j > i when its second window is unset -- when r^2 >= r2limit and
 * r^2 < 100 (writeResults, correlate_IO.c:403-410), after copying the whole
 * r^2 matrix to the host.  These calls select the same pairs on the GPU. */

/* Hits of rows [row0, row0 + rows) of table B: row r's hits are entries
 * offsets[r] .. offsets[r+1]-1 of cols (column j of table A, ascending) and
 * r2.  Pointers are valid during the call only.  Return 0 to continue. */
typedef int (*qld_hits_fn)(void *user, int row0, int rows, const uint64_t *offsets,
                           const uint32_t *cols, const float *r2);

/* Host tables in, selected pairs out: r^2 of B (n rows) against A (m rows)
 * is computed in blocks of about `block_pairs` cells (<= 0: 2^26) of B
 * rows; each block is compacted on the device (`upper`: only j > i) and its
 * hits are copied to pinned host memory and handed to `fn` on the calling
 * thread, in row order, while the next block computes.  A single region
 * passes the same table pointers as A and B with upper = 1 (only the upper
 * triangle is computed).  Replaces the full-matrix D2H of correlate_gpu
 * (gemm_gpu.c:849-861) for report output. */
int qld_correlate_hits(qld_ctx *ctx,
                       const uint64_t *A, const uint32_t *bcA, int m,
                       const uint64_t *B, const uint32_t *bcB, int n,
                       int words, int snp_size, int upper, float r2limit,
                       int64_t block_pairs, qld_hits_fn fn, void *user);

/* Device-resident variant for one block already in HBM: results
 * dres[r * ldr + c] = r^2 of B row row0 + r and A column col0 + c, c < cols.
 * Counts hits per row into dcounts[rows], their exclusive prefix into
 * doffsets[rows + 1], sets *total, and -- when *total <= capacity -- writes
 * the hits (column index, r^2) into dcols / dr2 in row-major order.
 * QLD_ERR_BUDGET when capacity is too small (*total says how much is
 * needed).  Synchronizes the context stream to read the total. */
int qld_compact_dev(qld_ctx *ctx, const float *dres, int64_t ldr, int rows, int cols,
                    int row0, int col0, int upper, float r2limit,
                    uint32_t *dcounts, uint64_t *doffsets, uint32_t *dcols, float *dr2,
                    uint64_t capacity, uint64_t *total);

#ifdef __cplusplus
}
#endif
#endif
