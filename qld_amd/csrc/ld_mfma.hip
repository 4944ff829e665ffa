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
// Work decomposition (both engines)
//   workgroup : 4 wave64 (2 x 2), one 256 x 256 tile of pairs
//   wave      : 128 x 128 = 8 x 8 MFMA blocks of 16 x 16 (256 fp32/int32
//               accumulators per lane)
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

constexpr int MT = 256;          // tile edge
constexpr int WT = 128;          // wave tile edge
constexpr int NB = WT / 16;      // MFMA blocks per wave edge
constexpr int SB = 128;          // bytes per row per k stage
constexpr int TILE_BYTES = MT * SB;            // one operand, one stage
constexpr int THREADS = 256;
constexpr int LOADS_PER_WAVE = MT * SB / 1024 / 4;   // 1 KiB wave loads per operand

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

struct MfmaArgs {
    const uint8_t *XA;      // expanded table A (columns j)
    const uint8_t *XB;      // expanded table B (rows i)
    int64_t row_bytes;      // expanded row pitch (multiple of SB)
    int stages;             // row_bytes / SB
    const uint32_t *tiles;  // (ti | tj << 16) per logical tile
    int64_t ntiles;
    const uint32_t *bcA;
    const uint32_t *bcB;
    int m, n, tri;
    float *res;
    int64_t ldr;
    uint32_t *cnt;
    int64_t ldc;
    EpiConst ec;
};

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t panel_rsrc(const uint8_t *base,
                                                            int row0, int rows,
                                                            int64_t row_bytes)
{
    int64_t have = (int64_t)(rows - row0) * row_bytes;
    int64_t lim = (int64_t)MT * row_bytes;
    int64_t nrec = have < lim ? have : lim;
    if (nrec < 0) nrec = 0;
    return __builtin_amdgcn_make_buffer_rsrc((void *)(base + (int64_t)row0 * row_bytes),
                                             0, (int)nrec, 0x00020000);
}

// Queue one k stage of one operand: wave w fills tile rows [64w, 64w + 64).
// Lane L of load q lands in LDS slot (64w + 8q)*8 + L, i.e. row 64w+8q+L/8,
// physical chunk L%8, which holds logical chunk (L%8) ^ (L/8).
__device__ __forceinline__ void stage_load(__amdgpu_buffer_rsrc_t rs, uint8_t *lds_tile,
                                           int64_t row_bytes, int k0, int wave, int lane)
{
    const int r_in = lane >> 3, pc = lane & 7;
    const int c = pc ^ r_in;
#pragma unroll
    for (int q = 0; q < LOADS_PER_WAVE; ++q) {
        int row = wave * 64 + q * 8 + r_in;
        uint32_t voff = (uint32_t)((int64_t)row * row_bytes + k0 + c * 16);
        uint8_t *dst = lds_tile + (wave * 64 + q * 8) * SB;
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

template <class OP, int MODE>
__global__ __launch_bounds__(THREADS, 1) void ld_mfma_kernel(MfmaArgs p)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    // Stage buffer b: operand A (rows i) at 2b*TILE_BYTES, B (cols j) after it.
    auto sA = [&](int b) { return smem + (2 * b) * TILE_BYTES; };
    auto sB = [&](int b) { return smem + (2 * b + 1) * TILE_BYTES; };

    const int64_t L = xcd_logical(blockIdx.x, gridDim.x);
    const uint32_t t = p.tiles[L];
    const int ti = (int)(t & 0xFFFF), tj = (int)(t >> 16);
    const int i0 = ti * MT, j0 = tj * MT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    // Operand "A" of the MFMA = rows i (table B); operand "B" = columns j.
    const __amdgpu_buffer_rsrc_t rsI = panel_rsrc(p.XB, i0, p.n, p.row_bytes);
    const __amdgpu_buffer_rsrc_t rsJ = panel_rsrc(p.XA, j0, p.m, p.row_bytes);

    typename OP::acc_t acc[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = typename OP::acc_t{0, 0, 0, 0};

    stage_load(rsI, sA(0), p.row_bytes, 0, wave, lane);
    stage_load(rsJ, sB(0), p.row_bytes, 0, wave, lane);

    const int r16 = lane & 15, g = lane >> 4;
    for (int s = 0; s < p.stages; ++s) {
        const int cur = s & 1;
        if (s + 1 < p.stages) {
            stage_load(rsI, sA(cur ^ 1), p.row_bytes, (s + 1) * SB, wave, lane);
            stage_load(rsJ, sB(cur ^ 1), p.row_bytes, (s + 1) * SB, wave, lane);
            wait_vm<2 * LOADS_PER_WAVE>();
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
    // i = i0 + wm*128 + a*16 + 4*(lane/16) + v,  j = j0 + wn*128 + b*16 + lane%16.
    const int jb = j0 + wn * WT + r16;
    if (MODE == QLD_R2_REFERENCE) {
        RefTerm tj_[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            int j = jb + b * 16;
            tj_[b] = ref_term(j < p.m ? p.bcA[j] : 0u, p.ec.fN);
        }
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int i = i0 + wm * WT + a * 16 + 4 * g + v;
                if (i >= p.n) continue;
                RefTerm ti_ = ref_term(p.bcB[i], p.ec.fN);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    int j = jb + b * 16;
                    if (j >= p.m || (p.tri && j <= i)) continue;
                    float c = OP::count(acc[a][b][v]);
                    if (p.res) p.res[(int64_t)i * p.ldr + j] = r2_reference(c, tj_[b], ti_, p.ec.invN);
                    if (p.cnt) p.cnt[(int64_t)i * p.ldc + j] = (uint32_t)c;
                }
            }
    } else {
        ExactTerm tj_[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            int j = jb + b * 16;
            tj_[b] = exact_term(j < p.m ? p.bcA[j] : 0u, p.ec.dN);
        }
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int i = i0 + wm * WT + a * 16 + 4 * g + v;
                if (i >= p.n) continue;
                ExactTerm ti_ = exact_term(p.bcB[i], p.ec.dN);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    int j = jb + b * 16;
                    if (j >= p.m || (p.tri && j <= i)) continue;
                    float c = OP::count(acc[a][b][v]);
                    if (p.res) p.res[(int64_t)i * p.ldr + j] = r2_exact(c, tj_[b], ti_, p.ec.dN);
                    if (p.cnt) p.cnt[(int64_t)i * p.ldc + j] = (uint32_t)c;
                }
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

int mfma_tile_edge() { return MT; }

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

template <class OP, int MODE>
static hipError_t launch_one(const MfmaArgs &a, hipStream_t s)
{
    const size_t lds = 4 * (size_t)TILE_BYTES;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)ld_mfma_kernel<OP, MODE>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((ld_mfma_kernel<OP, MODE>), dim3((unsigned)a.ntiles), dim3(THREADS),
                       lds, s, a);
    return hipGetLastError();
}

hipError_t launch_mfma(int engine, int r2_mode, const MfmaJob &j, hipStream_t s)
{
    MfmaArgs a;
    a.XA = j.XA;
    a.XB = j.XB;
    a.row_bytes = j.row_bytes;
    a.stages = (int)(j.row_bytes / SB);
    a.tiles = j.tiles;
    a.ntiles = j.ntiles;
    a.bcA = j.job.bcA;
    a.bcB = j.job.bcB;
    a.m = j.job.m;
    a.n = j.job.n;
    a.tri = j.job.tri;
    a.res = j.job.res;
    a.ldr = j.job.ldr;
    a.cnt = j.job.cnt;
    a.ldc = j.job.ldc;
    a.ec = make_epi_const(j.job.snp_size);
    if (a.ntiles == 0) return hipSuccess;
    if (engine == QLD_ENGINE_MFMA_FP4)
        return r2_mode == QLD_R2_EXACT ? launch_one<FP4, QLD_R2_EXACT>(a, s)
                                       : launch_one<FP4, QLD_R2_REFERENCE>(a, s);
    return r2_mode == QLD_R2_EXACT ? launch_one<I8, QLD_R2_EXACT>(a, s)
                                   : launch_one<I8, QLD_R2_REFERENCE>(a, s);
}
