// Matrix-core engines: the pair counts as a 0/1 matrix product on MFMA.
//
// Each packed haplotype bit is first expanded to one matrix element:
//   FP4 engine : e2m1 nibble 0x2 (= 1.0) or 0x0, two samples per byte,
//                v_mfma_scale_f32_16x16x128_f8f6f4 (scale 2^0), fp32 sums
//   I8 engine  : byte 0x01 or 0x00, v_mfma_i32_16x16x64_i8, int32 sums
// A count is a sum of at most N products 1*1, so fp32 accumulators are exact
// for N < 2^24 and the counts equal the popcount engine's bit for bit.
// The pairing of samples to K positions only has to be the same for every
// row; the expansion keeps each 64-bit word's samples together.
//
// Work decomposition (both engines), tile edge T = 256 or 128
//   workgroup : 4 wave64 (2 x 2), one T x T tile of pairs of one job
//   wave      : T/2 x T/2 = (T/32)^2 MFMA blocks of 16 x 16 (256 or 64
//               fp32/int32 accumulators per lane)
//   k stage   : 128 bytes of every tile row (256 fp4 / 128 int8 samples),
//               two MFMA k-steps; stages are double-buffered in LDS and
//               filled with buffer_load_dwordx4 ... lds (no register
//               staging; out-of-range rows read as 0 through the buffer
//               descriptor's bound)
// LDS rows are 8 chunks of 16 bytes stored at chunk ^ (row & 7) so that the
// ds_read_b128 of 16 consecutive rows by an MFMA operand fetch spreads over
// all banks.
#include "qld_internal.h"
#include "ld_epilogue.h"
#include "ld_tiles.h"
#include "ld_mfma.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int SB = 128;          // bytes per row per k stage
constexpr int THREADS = 256;     // 4 wave64 in a 2 x 2 arrangement

// Geometry of a TILE x TILE workgroup tile.
template <int TILE>
struct Geo {
    static constexpr int WT = TILE / 2;              // wave tile edge
    static constexpr int NB = WT / 16;               // 16 x 16 MFMA blocks per wave edge
    static constexpr int ROWS_PER_WAVE = TILE / 4;   // rows each wave stages per operand
    static constexpr int LOADS = ROWS_PER_WAVE / 8;  // 1 KiB wave loads per operand
    static constexpr int TILE_BYTES = TILE * SB;     // one operand, one stage
    static constexpr int LDS_BYTES = 4 * TILE_BYTES; // 2 operands x 2 stages
};

struct FP4 {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c)
    {
        v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
        v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a8, b8, c, 4, 4, 0,
                                                                 127, 0, 127);
    }
    static __device__ __forceinline__ float count(float v) { return v; }
};

struct I8 {
    typedef v4i acc_t;
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c)
    {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ float count(int v) { return (float)v; }
};

struct KArgs {
    WorkDesc inline_desc;     // used when descs == nullptr
    const WorkDesc *descs;
    const WorkItem *items;
    int64_t nitems;
    uint32_t *queue;          // persistent launches: 8 per-XCD queue heads
};

typedef __attribute__((address_space(3))) void lds_void;

template <int TILE>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t panel_rsrc(const uint8_t *base, int row0,
                                                            int rows, int64_t row_bytes)
{
    int64_t have = (int64_t)(rows - row0) * row_bytes;
    int64_t lim = (int64_t)TILE * row_bytes;
    int64_t nrec = have < lim ? have : lim;
    if (nrec < 0) nrec = 0;
    return __builtin_amdgcn_make_buffer_rsrc((void *)(base + (int64_t)row0 * row_bytes),
                                             0, (int)nrec, 0x00020000);
}

// Queue one k stage of one operand: wave w fills tile rows
// [w*R, w*R + R), R = TILE/4.  Lane L of load q lands in LDS slot
// (w*R + 8q)*8 + L, i.e. row w*R + 8q + L/8, physical chunk L%8, which holds
// logical chunk (L%8) ^ (L/8).
template <int TILE>
__device__ __forceinline__ void stage_load(__amdgpu_buffer_rsrc_t rs, uint8_t *lds_tile,
                                           int64_t row_bytes, int k0, int wave, int lane)
{
    typedef Geo<TILE> G;
    const int r_in = lane >> 3, pc = lane & 7;
    const int c = pc ^ r_in;
#pragma unroll
    for (int q = 0; q < G::LOADS; ++q) {
        int row = wave * G::ROWS_PER_WAVE + q * 8 + r_in;
        uint32_t voff = (uint32_t)((int64_t)row * row_bytes + k0 + c * 16);
        uint8_t *dst = lds_tile + (wave * G::ROWS_PER_WAVE + q * 8) * SB;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void *)dst, 16, voff, 0, 0, 0);
    }
}

template <int N>
__device__ __forceinline__ void wait_vm()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Workgroup barrier that also orders LDS traffic: this wave's outstanding
// ds_reads complete first, and the compiler may not move memory accesses
// across it (the bare s_barrier builtin orders neither).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One tile of one job: k loop over double-buffered LDS stages, then the
// fused epilogue straight to the job's result rows.
template <class OP, int MODE, int TILE>
__device__ __forceinline__ void mfma_tile(const WorkDesc &d, int ti, int tj, uint8_t *smem)
{
    typedef Geo<TILE> G;
    constexpr int NB = G::NB, WT = G::WT;
    auto sA = [&](int b) { return smem + (2 * b) * G::TILE_BYTES; };
    auto sB = [&](int b) { return smem + (2 * b + 1) * G::TILE_BYTES; };

    const int i0 = ti * TILE, j0 = tj * TILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t rb = d.row_bytes;
    const int stages = (int)(rb / SB);

    // MFMA operand "A" = job rows i, operand "B" = job columns j.
    const __amdgpu_buffer_rsrc_t rsI = panel_rsrc<TILE>(d.XI, i0, d.ni, rb);
    const __amdgpu_buffer_rsrc_t rsJ = panel_rsrc<TILE>(d.XJ, j0, d.nj, rb);

    typename OP::acc_t acc[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = typename OP::acc_t{0, 0, 0, 0};

    stage_load<TILE>(rsI, sA(0), rb, 0, wave, lane);
    stage_load<TILE>(rsJ, sB(0), rb, 0, wave, lane);

    const int r16 = lane & 15, g = lane >> 4;
    for (int s = 0; s < stages; ++s) {
        const int cur = s & 1;
        if (s + 1 < stages) {
            stage_load<TILE>(rsI, sA(cur ^ 1), rb, (s + 1) * SB, wave, lane);
            stage_load<TILE>(rsJ, sB(cur ^ 1), rb, (s + 1) * SB, wave, lane);
            wait_vm<2 * G::LOADS>();
        } else {
            wait_vm<0>();
        }
        lds_barrier();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int c = ks * 4 + g;
            v4i bv[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                int row = wn * WT + b * 16 + r16;
                bv[b] = *(const v4i *)(sB(cur) + row * SB + ((c ^ (row & 7)) << 4));
            }
#pragma unroll
            for (int a = 0; a < NB; ++a) {
                int row = wm * WT + a * 16 + r16;
                v4i av = *(const v4i *)(sA(cur) + row * SB + ((c ^ (row & 7)) << 4));
#pragma unroll
                for (int b = 0; b < NB; ++b) acc[a][b] = OP::mma(av, bv[b], acc[a][b]);
            }
        }
        lds_barrier();
    }

    // Fused epilogue.  Block (a, b) register v holds the pair
    // i = i0 + wm*WT + a*16 + 4*(lane/16) + v,  j = j0 + wn*WT + b*16 + lane%16.
    const int jb = j0 + wn * WT + r16;
    if (MODE == QLD_R2_REFERENCE) {
        RefTerm tj_[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            int j = jb + b * 16;
            tj_[b] = ref_term(j < d.nj ? d.bcJ[j] : 0u, d.fN);
        }
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int i = i0 + wm * WT + a * 16 + 4 * g + v;
                if (i >= d.ni) continue;
                RefTerm ti_ = ref_term(d.bcI[i], d.fN);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    int j = jb + b * 16;
                    if (j >= d.nj || (d.tri && j <= i)) continue;
                    float c = OP::count(acc[a][b][v]);
                    if (d.res) d.res[(int64_t)i * d.ldr + j] = r2_reference(c, tj_[b], ti_, d.invN);
                    if (d.cnt) d.cnt[(int64_t)i * d.ldc + j] = (uint32_t)c;
                }
            }
    } else {
        ExactTerm tj_[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            int j = jb + b * 16;
            tj_[b] = exact_term(j < d.nj ? d.bcJ[j] : 0u, d.dN);
        }
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int i = i0 + wm * WT + a * 16 + 4 * g + v;
                if (i >= d.ni) continue;
                ExactTerm ti_ = exact_term(d.bcI[i], d.dN);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    int j = jb + b * 16;
                    if (j >= d.nj || (d.tri && j <= i)) continue;
                    float c = OP::count(acc[a][b][v]);
                    if (d.res) d.res[(int64_t)i * d.ldr + j] = r2_exact(c, tj_[b], ti_, d.dN);
                    if (d.cnt) d.cnt[(int64_t)i * d.ldc + j] = (uint32_t)c;
                }
            }
    }
}

template <class OP, int MODE, int TILE>
__device__ __forceinline__ void run_item(const KArgs &p, int64_t idx, uint8_t *smem)
{
    const WorkItem it = p.items[idx];
    if (p.descs) {
        const WorkDesc d = p.descs[it.desc];
        mfma_tile<OP, MODE, TILE>(d, it.ti, it.tj, smem);
    } else {
        mfma_tile<OP, MODE, TILE>(p.inline_desc, it.ti, it.tj, smem);
    }
}

// One workgroup per item (PERSIST == 0) or a fixed grid draining eight
// per-XCD queues: workgroup b starts on queue b % 8 -- the XCD it runs on --
// whose items are a contiguous range of the list, then steals from the
// other queues.  Every claim advances a head, so the loop ends once all
// eight ranges are exhausted.
template <class OP, int MODE, int TILE, int PERSIST>
__global__ __launch_bounds__(THREADS, 1) void ld_mfma_kernel(KArgs p)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    if (!PERSIST) {
        run_item<OP, MODE, TILE>(p, xcd_logical(blockIdx.x, gridDim.x), smem);
        return;
    }
    __shared__ int64_t claimed;
    const int x0 = (int)(blockIdx.x % QLD_NUM_XCD);
    for (int q = 0; q < QLD_NUM_XCD; ++q) {
        const int x = (x0 + q) % QLD_NUM_XCD;
        const int64_t begin = p.nitems * x / QLD_NUM_XCD;
        const int64_t end = p.nitems * (x + 1) / QLD_NUM_XCD;
        for (;;) {
            if (threadIdx.x == 0) {
                int64_t k = (int64_t)atomicAdd(&p.queue[x], 1u);
                claimed = begin + k < end ? begin + k : -1;
            }
            __syncthreads();
            const int64_t idx = claimed;
            __syncthreads();
            if (idx < 0) break;
            run_item<OP, MODE, TILE>(p, idx, smem);
        }
    }
}

// Bit expansion.  Word w of a packed row (64 samples) becomes 32 bytes (fp4)
// or 64 bytes (int8) at byte 32w / 64w of the expanded row; bit e of the
// word (LSB first) becomes element e.  Words W..Wp-1 are written as zeros.
__device__ __forceinline__ uint32_t nibbles_of_byte(uint32_t x)
{
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x << 1;                     // e2m1 1.0 = 0b0010
}

__device__ __forceinline__ uint64_t bytes_of_byte(uint32_t x)
{
    uint64_t t = x & 0xFFu;
    t = (t | (t << 28)) & 0x0000000F0000000Full;
    t = (t | (t << 14)) & 0x0003000300030003ull;
    t = (t | (t << 7)) & 0x0101010101010101ull;
    return t;
}

__global__ void expand_fp4_kernel(const uint64_t *T, int rows, int W, int Wp,
                                  uint8_t *X, int64_t row_bytes)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * Wp) return;
    int r = (int)(idx / Wp), w = (int)(idx % Wp);
    uint64_t v = w < W ? T[(int64_t)r * W + w] : 0ull;
    uint4 lo, hi;
    lo.x = nibbles_of_byte((uint32_t)(v >> 0) & 0xFF);
    lo.y = nibbles_of_byte((uint32_t)(v >> 8) & 0xFF);
    lo.z = nibbles_of_byte((uint32_t)(v >> 16) & 0xFF);
    lo.w = nibbles_of_byte((uint32_t)(v >> 24) & 0xFF);
    hi.x = nibbles_of_byte((uint32_t)(v >> 32) & 0xFF);
    hi.y = nibbles_of_byte((uint32_t)(v >> 40) & 0xFF);
    hi.z = nibbles_of_byte((uint32_t)(v >> 48) & 0xFF);
    hi.w = nibbles_of_byte((uint32_t)(v >> 56) & 0xFF);
    uint4 *dst = (uint4 *)(X + (int64_t)r * row_bytes + (int64_t)w * 32);
    dst[0] = lo;
    dst[1] = hi;
}

__global__ void expand_i8_kernel(const uint64_t *T, int rows, int W, int Wp,
                                 uint8_t *X, int64_t row_bytes)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * Wp) return;
    int r = (int)(idx / Wp), w = (int)(idx % Wp);
    uint64_t v = w < W ? T[(int64_t)r * W + w] : 0ull;
    uint64_t *dst = (uint64_t *)(X + (int64_t)r * row_bytes + (int64_t)w * 64);
#pragma unroll
    for (int b = 0; b < 8; ++b) dst[b] = bytes_of_byte((uint32_t)(v >> (8 * b)));
}

}  // namespace

int mfma_padded_words(int engine, int words)
{
    // A k stage is 128 bytes of a row: 4 packed words as fp4, 2 as int8.
    int q = engine == QLD_ENGINE_MFMA_FP4 ? 4 : 2;
    return (words + q - 1) / q * q;
}

int64_t mfma_row_bytes(int engine, int words)
{
    int64_t per_word = engine == QLD_ENGINE_MFMA_FP4 ? 32 : 64;
    return per_word * mfma_padded_words(engine, words);
}

void fill_work_desc(WorkDesc &d, const uint8_t *XI, const uint8_t *XJ,
                    const uint32_t *bcI, const uint32_t *bcJ, int ni, int nj,
                    int64_t row_bytes, int tri, int snp_size, float *res,
                    int64_t ldr, uint32_t *cnt, int64_t ldc)
{
    EpiConst e = make_epi_const(snp_size);
    d.XI = XI;
    d.XJ = XJ;
    d.bcI = bcI;
    d.bcJ = bcJ;
    d.res = res;
    d.cnt = cnt;
    d.ldr = ldr;
    d.ldc = ldc;
    d.row_bytes = row_bytes;
    d.ni = ni;
    d.nj = nj;
    d.tri = tri;
    d.snp_size = snp_size;
    d.fN = e.fN;
    d.pad_ = 0.0f;
    d.dN = e.dN;
    d.invN = e.invN;
}

size_t job_tile_count(int ni, int nj, int tri, int T)
{
    size_t a = (size_t)((ni + T - 1) / T), b = (size_t)((nj + T - 1) / T);
    return tri ? a * (a + 1) / 2 : a * b;
}

void append_job_tiles(WorkItem *out, size_t *count, uint32_t desc, int ni, int nj,
                      int tri, int T)
{
    const int S = 8;
    const int ti_count = (ni + T - 1) / T, tj_count = (nj + T - 1) / T;
    const int si = (ti_count + S - 1) / S, sj = (tj_count + S - 1) / S;
    size_t k = *count;
    for (int J = 0; J < sj; ++J)
        for (int I = 0; I < si; ++I) {
            if (tri && I > J) continue;
            for (int tj = J * S; tj < (J + 1) * S && tj < tj_count; ++tj)
                for (int ti = I * S; ti < (I + 1) * S && ti < ti_count; ++ti)
                    if (!tri || ti <= tj) out[k++] = WorkItem{desc, (uint16_t)ti, (uint16_t)tj};
        }
    *count = k;
}

hipError_t launch_expand(int engine, const uint64_t *T, int rows, int words,
                         uint8_t *X, hipStream_t s)
{
    if (rows <= 0) return hipSuccess;
    int Wp = mfma_padded_words(engine, words);
    int64_t total = (int64_t)rows * Wp;
    unsigned blocks = (unsigned)((total + 255) / 256);
    int64_t rb = mfma_row_bytes(engine, words);
    if (engine == QLD_ENGINE_MFMA_FP4)
        hipLaunchKernelGGL(expand_fp4_kernel, dim3(blocks), dim3(256), 0, s, T, rows,
                           words, Wp, X, rb);
    else
        hipLaunchKernelGGL(expand_i8_kernel, dim3(blocks), dim3(256), 0, s, T, rows,
                           words, Wp, X, rb);
    return hipGetLastError();
}

__global__ void zero_queue_kernel(uint32_t *q)
{
    if (threadIdx.x < 16) q[threadIdx.x] = 0u;
}

template <class OP, int MODE, int TILE, int PERSIST>
static hipError_t set_lds_limit()
{
    return hipFuncSetAttribute((const void *)ld_mfma_kernel<OP, MODE, TILE, PERSIST>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)Geo<TILE>::LDS_BYTES);
}

template <class OP, int MODE>
static hipError_t set_lds_limits_mode()
{
    hipError_t e;
    if ((e = set_lds_limit<OP, MODE, QLD_TILE_LARGE, 0>()) != hipSuccess) return e;
    if ((e = set_lds_limit<OP, MODE, QLD_TILE_LARGE, 1>()) != hipSuccess) return e;
    if ((e = set_lds_limit<OP, MODE, QLD_TILE_SMALL, 0>()) != hipSuccess) return e;
    return set_lds_limit<OP, MODE, QLD_TILE_SMALL, 1>();
}

// Raise the dynamic LDS limit of every kernel variant (done once per
// process, before any launch -- and so before any graph capture).
hipError_t mfma_init_attributes()
{
    hipError_t e;
    if ((e = set_lds_limits_mode<FP4, QLD_R2_REFERENCE>()) != hipSuccess) return e;
    if ((e = set_lds_limits_mode<FP4, QLD_R2_EXACT>()) != hipSuccess) return e;
    if ((e = set_lds_limits_mode<I8, QLD_R2_REFERENCE>()) != hipSuccess) return e;
    return set_lds_limits_mode<I8, QLD_R2_EXACT>();
}

template <class OP, int MODE, int TILE, int PERSIST>
static hipError_t launch_one(const KArgs &a, unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL((ld_mfma_kernel<OP, MODE, TILE, PERSIST>), dim3(grid), dim3(THREADS),
                       (size_t)Geo<TILE>::LDS_BYTES, s, a);
    return hipGetLastError();
}

template <class OP, int MODE, int TILE>
static hipError_t launch_mode(const KArgs &a, int persistent, hipStream_t s)
{
    if (persistent > 0)
        return launch_one<OP, MODE, TILE, 1>(a, (unsigned)persistent, s);
    return launch_one<OP, MODE, TILE, 0>(a, (unsigned)a.nitems, s);
}

template <class OP, int MODE>
static hipError_t launch_tile(const KArgs &a, int tile, int persistent, hipStream_t s)
{
    if (tile == QLD_TILE_SMALL) return launch_mode<OP, MODE, QLD_TILE_SMALL>(a, persistent, s);
    return launch_mode<OP, MODE, QLD_TILE_LARGE>(a, persistent, s);
}

hipError_t launch_mfma_items(int engine, int r2_mode, int tile,
                             const WorkDesc *descs, const WorkDesc &inline_desc,
                             const WorkItem *items, int64_t nitems,
                             int persistent, uint32_t *queue, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    KArgs a;
    a.inline_desc = inline_desc;
    a.descs = descs;
    a.items = items;
    a.nitems = nitems;
    a.queue = queue;
    if (persistent > 0) {
        if (!queue) return hipErrorInvalidValue;
        hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(64), 0, s, queue);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (engine == QLD_ENGINE_MFMA_FP4)
        return r2_mode == QLD_R2_EXACT ? launch_tile<FP4, QLD_R2_EXACT>(a, tile, persistent, s)
                                       : launch_tile<FP4, QLD_R2_REFERENCE>(a, tile, persistent, s);
    return r2_mode == QLD_R2_EXACT ? launch_tile<I8, QLD_R2_EXACT>(a, tile, persistent, s)
                                   : launch_tile<I8, QLD_R2_REFERENCE>(a, tile, persistent, s);
}
