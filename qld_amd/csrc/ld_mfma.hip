// Matrix-core engines: the pair counts as a 0/1 matrix product on MFMA.
//
// Each packed haplotype bit becomes one matrix element:
//   FP4 engine : e2m1 nibble 0x2 (= 1.0) or 0x0, two samples per byte,
//                v_mfma_scale_f32_32x32x64_f8f6f4 (scale 2^0), fp32 sums;
//                with direct operands the packed dwords are split into bit
//                planes in registers instead (fp4_plane: nibbles 0.5, 1 or
//                2, brought back to products of 1.0 by the MFMA's scale)
//   I8 engine  : byte 0x01 or 0x00, v_mfma_i32_32x32x32_i8, int32 sums
// (On gfx950 the 32x32 forms retire more work per cycle than the 16x16
// ones: measured 8.1 vs 4.8 PFLOP/s for fp4, 3.8 vs 2.9 POP/s for int8.)
// A count is a sum of at most N products 1*1, so fp32 accumulators are exact
// for N < 2^24 and the counts equal the popcount engine's bit for bit.
// The pairing of samples to K positions only has to be the same for every
// row; the expansion keeps each 64-bit word's samples together.
//
// Work decomposition, tile edge T = 256 or 128 (configurations below)
//   direct fp4 (default): each wave reads its own 32-row operand blocks
//               from the row-block-tiled table into registers (kloop_regs,
//               no LDS ring, no barrier); 256-tiles are two workgroups of 4
//               waves of 128 x 64, 128-tiles one workgroup of 4 waves of
//               64 x 64; LDS only stages the r^2 epilogue
//   operands in LDS (expanded fp4 / int8, or packed rows expanded in LDS):
//     workgroup : one T x T tile (8 waves of 128 x 64 for 256-tiles, 4 of
//                 64 x 64 for 128-tiles)
//     k stage   : SB bytes of every tile row (SB/32 MFMA k-steps) in a ring
//                 of NSTAGE LDS buffers, filled with buffer_load_dwordx4 ...
//                 lds (no register staging; out-of-range rows read as 0
//                 through the buffer descriptor's bound); one barrier per
//                 stage.  LDS rows are stored as 16-byte chunks permuted per
//                 row (Geo::swz) so that the ds_read_b128 of consecutive rows
//                 by an MFMA operand fetch spreads over all banks.
#include "qld_internal.h"
#include "ld_epilogue.h"
#include "ld_tiles.h"
#include "ld_mfma.h"

#include <stdlib.h>
#include <type_traits>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Geometry of a TILE x TILE tile computed by WGM x WGN wave64s (each a
// WTM x WTN block of 32 x 32 MFMA tiles) whose k stages carry SBY bytes of
// every row (SBY/32 MFMA k-steps) in a ring of NST LDS buffers.  SPLIT > 1
// (direct operands only) hands the tile to SPLIT workgroups, each taking
// WGM/SPLIT rows of waves, sized so that SPLIT workgroups share a CU: while
// one is in its r^2 epilogue, the other keeps the matrix cores busy.
template <int TILE, int SBY, int NST, int WGM, int WGN, int SPLIT_ = 1>
struct Geo {
    static constexpr int EDGE = TILE;
    static constexpr int SB = SBY;
    static constexpr int NSTAGE = NST;
    static constexpr int SPLIT = SPLIT_;
    static constexpr int WAVES_M = WGM / SPLIT;      // wave rows per workgroup
    static constexpr int WAVES_N = WGN;
    static constexpr int WAVES = WAVES_M * WGN;
    static constexpr int THREADS = 64 * WAVES;
    static constexpr int KSTEPS = SB / 32;
    static constexpr int CHUNKS = SB / 16;           // 16-byte chunks per LDS row
    static constexpr int ROWS_PER_LOAD = 1024 / SB;  // rows per 1 KiB wave load
    static constexpr int WTM = TILE / WGM;           // wave tile rows (operand A)
    static constexpr int WTN = TILE / WGN;           // wave tile columns (operand B)
    static constexpr int NBM = WTM / 32;             // 32 x 32 MFMA blocks per wave
    static constexpr int NBN = WTN / 32;
    static constexpr int ROWS_PER_WAVE = TILE / WAVES;   // rows each wave stages per operand
    static constexpr int LOADS = ROWS_PER_WAVE / ROWS_PER_LOAD;
    static constexpr int TILE_BYTES = TILE * SB;     // one operand, one stage
    static constexpr int STAGE_BYTES = 2 * TILE_BYTES;
    static constexpr int LDS_BYTES = NSTAGE * STAGE_BYTES;
    // Epilogue: each wave stages 32 rows of its counts at a time, plus the
    // r^2 terms of its WTM rows.
    static constexpr int EP_PITCH = WTN + 8;         // words: rows r and r+4 land 32 banks apart
    static constexpr int EP_WAVE_BYTES = 32 * EP_PITCH * 4;
    static constexpr int EP_TERMS_OFF = WAVES * EP_WAVE_BYTES;
    static constexpr int LANES_PER_ROW = WTN / 4;    // each lane owns 4 columns
    static constexpr int ROWS_PER_IT = 64 / LANES_PER_ROW;
    static constexpr int RPL = WTM / 64 > 0 ? WTM / 64 : 1;   // epilogue rows per lane
    static constexpr int EP_BYTES = EP_TERMS_OFF + WAVES * WTM * 16;
    static_assert(EP_BYTES <= LDS_BYTES, "epilogue LDS");
    static_assert(LOADS >= 1 && NSTAGE >= 2, "stage geometry");
    static_assert(WGM % SPLIT == 0, "split geometry");
    static_assert(ROWS_PER_IT >= 1 && 32 % ROWS_PER_IT == 0, "epilogue geometry");

    // Physical 16-byte chunk of logical chunk c of an LDS row.  The
    // ds_read_b128 of an MFMA operand fetch reads one logical chunk of 16
    // consecutive rows per 64-bank pass (8 rows per 32-bank pass); the row
    // function spreads every such group over distinct 16-byte bank groups
    // (checked exhaustively for both pass shapes).  Its own inverse.
    static __device__ __forceinline__ int swz(int row, int c)
    {
        if (CHUNKS == 8) return c ^ ((row ^ (row >> 3)) & 7);
        return c ^ (((row >> 2) & 1) | ((((row >> 1) ^ (row >> 3)) & 1) << 1));
    }
};

// Kernel configurations: (tile, stage bytes, ring depth, waves down, waves
// across[, workgroups per tile]).  256-tiles with operands in LDS: 8 waves
// of 128 x 64 (two per SIMD, 0.75 KiB of LDS operand reads per MFMA) over a
// 4-deep ring of 64-byte stages (measured faster than 4 waves of 128 x 128
// over a 2-deep ring of 128-byte stages).  Direct operands: the same 8 waves
// of 128 x 64 as two workgroups of 4 (L256W8H; no ring, the stage fields
// only size the unused ring geometry).  128-tiles: 4 waves of 64 x 64 over
// a 2-deep ring of 128-byte stages (a 4-deep ring of 64-byte stages
// measured slower).
#define QLD_CFG_L256W8 256, 64, 4, 2, 4
#define QLD_CFG_L256W8H 256, 64, 4, 2, 4, 2
#define QLD_CFG_S128 128, 128, 2, 2, 2
typedef Geo<QLD_CFG_L256W8> GeoL8;
typedef Geo<QLD_CFG_L256W8H> GeoL8H;
typedef Geo<QLD_CFG_S128> GeoS;

// v_mfma_scale_f32_32x32x64_f8f6f4 with e2m1 operands (scale 2^0): on gfx950
// it retires twice the flops per cycle of the 16x16x128 form.
struct FP4 {
    typedef v16f acc_t;
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c)
    {
        v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
        v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, c, 4, 4, 0, 127,
                                                                0, 127);
    }
    // The same product with operand A scaled by 2^(EA-127) (E8M0 exponent).
    template <int EA>
    static __device__ __forceinline__ acc_t mma_scaled(v4i a, v4i b, acc_t c)
    {
        v8i a8 = {a[0], a[1], a[2], a[3], 0, 0, 0, 0};
        v8i b8 = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, c, 4, 4, 0, EA,
                                                                0, 127);
    }
    // Counts travel through LDS as raw 32-bit words.
    static __device__ __forceinline__ uint32_t bits(float v) { return __float_as_uint(v); }
    static __device__ __forceinline__ float count(uint32_t w) { return __uint_as_float(w); }
};

struct I8 {
    typedef v16i acc_t;
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c)
    {
        return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ uint32_t bits(int v) { return (uint32_t)v; }
    static __device__ __forceinline__ float count(uint32_t w) { return (float)(int)w; }
};

struct KArgs {
    WorkDesc inline_desc;     // used when descs == nullptr
    const WorkDesc *descs;
    const WorkItem *items;
    int64_t nitems;
    uint32_t *queue;          // persistent launches: 8 per-XCD queue heads
    int64_t qbeg[QLD_NUM_XCD + 1];   // queue x owns items [qbeg[x], qbeg[x+1])
    const uint32_t *wlist;    // shared-pair units: window lists (item.desc = offset)
    const WinOut *wins;       // shared-pair units: window outputs
    int ksplit;               // split-K units: chunks per item
    float *kpart;             // split-K units: partial sums, 64 KiB per unit
    uint32_t *kcnt;           // split-K units: arrivals per item (zero between launches)
};

typedef __attribute__((address_space(3))) void lds_void;
// Global-memory views of the job's pointers: accesses through them compile
// to global_* instructions instead of flat_* ones (a pending flat access
// makes the compiler wait for every outstanding load at each later wait).
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) float g_f32;
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) const uint32_t g_cu32;
typedef __attribute__((address_space(1))) v4f g_v4f;
typedef __attribute__((address_space(1))) v4u g_v4u;

template <int TILE>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t panel_rsrc(const uint8_t *base, int row0,
                                                            int rows, int64_t row_bytes)
{
    int64_t have = (int64_t)(rows - row0) * row_bytes;
    int64_t lim = (int64_t)TILE * row_bytes;
    int64_t nrec = have < lim ? have : lim;
    if (nrec < 0) nrec = 0;
    // num_records is an unsigned 32-bit field; launches keep TILE * row_bytes
    // at or below 2^31 (expanded_tile_cap)
    if (nrec > 0xFFFFFFFFll) nrec = 0xFFFFFFFFll;
    return __builtin_amdgcn_make_buffer_rsrc((void *)(base + (int64_t)row0 * row_bytes),
                                             0, (int)(uint32_t)nrec, 0x00020000);
}

// Queue one k stage of one operand: wave w fills tile rows
// [w*R, w*R + R), R = TILE/WAVES.  A wave load covers 1 KiB = ROWS_PER_LOAD
// rows x CHUNKS chunks; lane L lands at row L/CHUNKS of the load, physical
// chunk L%CHUNKS, which holds logical chunk swz(row, L%CHUNKS).
template <class G>
__device__ __forceinline__ void stage_load(__amdgpu_buffer_rsrc_t rs, uint8_t *lds_tile,
                                           int64_t row_bytes, int k0, int wave, int lane)
{
    const int r_in = lane / G::CHUNKS, pc = lane % G::CHUNKS;
#pragma unroll
    for (int q = 0; q < G::LOADS; ++q) {
        int row = wave * G::ROWS_PER_WAVE + q * G::ROWS_PER_LOAD + r_in;
        const int c = G::swz(row, pc);
        uint32_t voff = (uint32_t)((int64_t)row * row_bytes + k0 + c * 16);
        uint8_t *dst = lds_tile + (wave * G::ROWS_PER_WAVE + q * G::ROWS_PER_LOAD) * G::SB;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void *)dst, 16, voff, 0, 0, 0);
    }
}

// The wave's index in the workgroup as a scalar: buffer descriptors, LDS
// DMA destinations and tile origins derived from it stay in SGPRs (a value
// the compiler cannot prove wave-uniform costs a readfirstlane loop around
// every buffer load).
__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// A value loaded identically by every lane (a work item, a job descriptor),
// moved to SGPRs dword by dword.  Without it the compiler keeps such loads in
// VGPRs (it cannot prove the memory unclobbered) and everything derived from
// them -- buffer descriptors above all -- needs a readfirstlane loop per use.
template <class T>
__device__ __forceinline__ T wave_uniform(const T &v)
{
    static_assert(sizeof(T) % 4 == 0, "dword-sized");
    T out;
    const int *src = reinterpret_cast<const int *>(&v);
    int *dst = reinterpret_cast<int *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

// Element idx of an array that no lane writes while the kernel runs (work
// items, job descriptors), read through the constant address space: with a
// wave-uniform index the loads are scalar (s_load, counted by lgkmcnt), so
// they never wait behind the vector-memory counter -- which the r^2 stores
// of the unit before keep busy.
template <class T>
__device__ __forceinline__ T load_uniform(const T *base, int64_t idx)
{
    static_assert(sizeof(T) % 4 == 0, "dword-sized");
    typedef const __attribute__((address_space(4))) uint32_t cu32;
    const cu32 *src = (const cu32 *)(uintptr_t)(base + idx);
    T out;
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
    return out;
}

// Claim the next unit of queue q: the old head, incremented.  An atomic
// increment, not an add: the compiler rewrites a lane's atomic add into a
// wave-wide form whose result it broadcasts at once -- a wait, right where
// the claim is issued, for every memory operation pending before it, the
// r^2 stores included -- and leaves increments alone.
__device__ __forceinline__ uint32_t claim_head(uint32_t *q)
{
    return __builtin_amdgcn_atomic_inc32(q, 0xFFFFFFFFu, __ATOMIC_RELAXED, "agent");
}

// What a claim loop does inside a unit: begin() once the k loop's first
// operands have arrived (the k loop drains every memory operation there, the
// unit before's r^2 stores among them, so a claim issued then waits behind
// nothing), end() after the k loop, before the epilogue's stores (where the
// claim, its result long back, is read without a wait).
struct NoHook {
    __device__ __forceinline__ void begin() const {}
    __device__ __forceinline__ void end() const {}
};

template <class B, class E>
struct Hooks {
    B b;
    E e;
    __device__ __forceinline__ void begin() const { b(); }
    __device__ __forceinline__ void end() const { e(); }
};

template <class B, class E>
__device__ __forceinline__ Hooks<B, E> make_hooks(const B &b, const E &e)
{
    return Hooks<B, E>{b, e};
}

template <int N>
__device__ __forceinline__ void wait_vm()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Wait until at most `ahead` later stages (2*LOADS loads each) are pending.
template <int PER_STAGE>
__device__ __forceinline__ void wait_stage(int ahead)
{
    if (ahead >= 2) wait_vm<2 * PER_STAGE>();
    else if (ahead == 1) wait_vm<PER_STAGE>();
    else wait_vm<0>();
}

// Workgroup barrier that also orders LDS traffic: this wave's outstanding
// ds_reads complete first, and the compiler may not move memory accesses
// across it (the bare s_barrier builtin orders neither).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// r^2 of 4 consecutive cells of one row.
template <int MODE>
struct Epi;

// Row terms travel through LDS as 16 bytes.
union TermBits {
    uint4 u;
    struct {
        float f;
        uint32_t z;
        double d;
    } s;
};

template <>
struct Epi<QLD_R2_REFERENCE> {
    typedef RefTerm term;
    static constexpr int ROW_UNROLL = 1;   // rows in flight (register budget)
    static __device__ __forceinline__ term make(uint32_t bc, const WorkDesc &d)
    {
        return ref_term(bc, d.fN);
    }
    static __device__ __forceinline__ uint4 pack(const term &t)
    {
        TermBits b;
        b.s.f = t.p;
        b.s.z = __float_as_uint(t.ip);
        b.s.d = t.A;
        return b.u;
    }
    static __device__ __forceinline__ term unpack(uint4 u)
    {
        TermBits b;
        b.u = u;
        term t;
        t.p = b.s.f;
        t.ip = __uint_as_float(b.s.z);
        t.A = b.s.d;
        return t;
    }
    // Four pairs without a division; a lane whose pair may round otherwise
    // (about one pair in 2^19) redoes that pair with the reference's
    // division, under one branch for the four.
    static __device__ __forceinline__ void r2x4(const float (&c)[4], const term (&col)[4],
                                                const term &row, const WorkDesc &d,
                                                float (&out)[4])
    {
        bool s0, s1, s2, s3;
        out[0] = r2_reference_fast(c[0], col[0], row, d.invN, s0);
        out[1] = r2_reference_fast(c[1], col[1], row, d.invN, s1);
        out[2] = r2_reference_fast(c[2], col[2], row, d.invN, s2);
        out[3] = r2_reference_fast(c[3], col[3], row, d.invN, s3);
        if (s0 | s1 | s2 | s3) {
            if (s0) out[0] = r2_reference_div(c[0], col[0].p, row.p, d.invN);
            if (s1) out[1] = r2_reference_div(c[1], col[1].p, row.p, d.invN);
            if (s2) out[2] = r2_reference_div(c[2], col[2].p, row.p, d.invN);
            if (s3) out[3] = r2_reference_div(c[3], col[3].p, row.p, d.invN);
        }
    }
    // Four rows of one column.
    static __device__ __forceinline__ void r2x4_rows(const float (&c)[4], const term &col,
                                                     const term (&row)[4], const WorkDesc &d,
                                                     float (&out)[4])
    {
        bool s0, s1, s2, s3;
        out[0] = r2_reference_fast(c[0], col, row[0], d.invN, s0);
        out[1] = r2_reference_fast(c[1], col, row[1], d.invN, s1);
        out[2] = r2_reference_fast(c[2], col, row[2], d.invN, s2);
        out[3] = r2_reference_fast(c[3], col, row[3], d.invN, s3);
        if (s0 | s1 | s2 | s3) {
            if (s0) out[0] = r2_reference_div(c[0], col.p, row[0].p, d.invN);
            if (s1) out[1] = r2_reference_div(c[1], col.p, row[1].p, d.invN);
            if (s2) out[2] = r2_reference_div(c[2], col.p, row[2].p, d.invN);
            if (s3) out[3] = r2_reference_div(c[3], col.p, row[3].p, d.invN);
        }
    }
};

template <>
struct Epi<QLD_R2_EXACT> {
    typedef ExactTerm term;
    static constexpr int ROW_UNROLL = 2;
    static __device__ __forceinline__ term make(uint32_t bc, const WorkDesc &d)
    {
        return exact_term(bc, d.dN);
    }
    // bc fits a float exactly (< 2^24).
    static __device__ __forceinline__ uint4 pack(const term &t)
    {
        TermBits b;
        b.s.f = (float)t.bc;
        b.s.z = t.zero;
        b.s.d = t.rv;
        return b.u;
    }
    static __device__ __forceinline__ term unpack(uint4 u)
    {
        TermBits b;
        b.u = u;
        term t;
        t.bc = (double)b.s.f;
        t.zero = b.s.z != 0;
        t.rv = b.s.d;
        return t;
    }
    static __device__ __forceinline__ void r2x4(const float (&c)[4], const term (&col)[4],
                                                const term &row, const WorkDesc &d,
                                                float (&out)[4])
    {
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = r2_exact(c[q], col[q], row, d.dN);
    }
    static __device__ __forceinline__ void r2x4_rows(const float (&c)[4], const term &col,
                                                     const term (&row)[4], const WorkDesc &d,
                                                     float (&out)[4])
    {
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = r2_exact(c[q], col, row[q], d.dN);
    }
};

// k loop over expanded operands: a ring of LDS stages filled by DMA.
template <class OP, class G>
__device__ __forceinline__ void kloop_dma(const WorkDesc &d, int i0, int j0,
                                          typename OP::acc_t (&acc)[G::NBM][G::NBN],
                                          uint8_t *smem)
{
    constexpr int TILE = G::EDGE, SB = G::SB, NSTAGE = G::NSTAGE, KSTEPS = G::KSTEPS;
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN;
    auto sA = [&](int b) { return smem + b * G::STAGE_BYTES; };
    auto sB = [&](int b) { return smem + b * G::STAGE_BYTES + G::TILE_BYTES; };
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int wm = wave / G::WAVES_N, wn = wave % G::WAVES_N;
    const int64_t rb = d.row_bytes;
    const int stages = (int)(rb / SB);

    // MFMA operand "A" = job rows i, operand "B" = job columns j.
    const __amdgpu_buffer_rsrc_t rsI = panel_rsrc<TILE>(d.XI, i0, d.ni, rb);
    const __amdgpu_buffer_rsrc_t rsJ = panel_rsrc<TILE>(d.XJ, j0, d.nj, rb);

#pragma unroll
    for (int p = 0; p < NSTAGE - 1; ++p)
        if (p < stages) {
            stage_load<G>(rsI, sA(p), rb, p * SB, wave, lane);
            stage_load<G>(rsJ, sB(p), rb, p * SB, wave, lane);
        }

    // 32x32 operand fetch: lane l reads 16 bytes of row l%32, chunk
    // 2*ks + l/32 of the stage (the same chunk for both operands).
    const int r32 = lane & 31, half = lane >> 5;
    for (int s = 0; s < stages; ++s) {
        const int cur = s % NSTAGE;
        int ahead = stages - 1 - s;
        wait_stage<2 * G::LOADS>(ahead < NSTAGE - 2 ? ahead : NSTAGE - 2);
        lds_barrier();
        if (s + NSTAGE - 1 < stages) {
            const int nxt = (s + NSTAGE - 1) % NSTAGE;
            stage_load<G>(rsI, sA(nxt), rb, (s + NSTAGE - 1) * SB, wave, lane);
            stage_load<G>(rsJ, sB(nxt), rb, (s + NSTAGE - 1) * SB, wave, lane);
        }
        v4i av[KSTEPS][NBM], bv[KSTEPS][NBN];
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            const int c = 2 * ks + half;
#pragma unroll
            for (int b = 0; b < NBN; ++b) {
                int row = wn * WTN + b * 32 + r32;
                bv[ks][b] = *(const v4i *)(sB(cur) + row * SB + (G::swz(row, c) << 4));
            }
#pragma unroll
            for (int a = 0; a < NBM; ++a) {
                int row = wm * WTM + a * 32 + r32;
                av[ks][a] = *(const v4i *)(sA(cur) + row * SB + (G::swz(row, c) << 4));
            }
        }
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks)
#pragma unroll
            for (int a = 0; a < NBM; ++a)
#pragma unroll
                for (int b = 0; b < NBN; ++b)
                    acc[a][b] = OP::mma(av[ks][a], bv[ks][b], acc[a][b]);
    }
}

// e2m1 expansion of 32 packed samples into one 16-byte MFMA chunk: dword k
// holds, in nibble n, 1.0 (0x2) iff bit 4n+k of x is set.  The sample -> K
// mapping only has to agree between the two operands, which it does.
__device__ __forceinline__ v4i fp4_chunk(uint32_t x)
{
    const uint32_t M = 0x22222222u;
    v4i c;
    c[0] = (int)((x << 1) & M);
    c[1] = (int)(x & M);
    c[2] = (int)((x >> 1) & M);
    c[3] = (int)((x >> 2) & M);
    return c;
}

// Bit-plane expansion of 128 packed samples (4 dwords) for the direct k
// loop: plane P keeps bit P of every nibble of every dword in place, so a
// plane costs one AND (plane 3, whose bit would be the e2m1 sign, is first
// moved down to bit 2).  The nibble values are 0.5 (bit 0), 1.0 (bit 1) and
// 2.0 (bit 2); the A operand's E8M0 scale brings every product of two set
// bits back to exactly 1.0 (scale 4 for plane 0, 1 for plane 1, 1/4 for
// planes 2 and 3).  The four planes of a lane's 128 bits are the k values
// of four consecutive MFMA k steps; the sample -> K mapping is the same for
// both operands.
template <int P>
__device__ __forceinline__ v4i fp4_plane(const v4i &x)
{
    v4i c;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        c[i] = P < 3 ? (int)((uint32_t)x[i] & (0x11111111u << P))
                     : (int)(((uint32_t)x[i] >> 1) & 0x44444444u);
    return c;
}
template <int P> struct PlaneScale { static constexpr int EA = P == 0 ? 129 : P == 1 ? 127 : 125; };

// k loop over packed operands (fp4 only): lane t stages row t%TILE of
// operand t/TILE.  Each 128-byte LDS stage row is the expansion of 4 packed
// words (chunk c = packed dword c); packed words are fetched three stages
// ahead into registers, expanded and written to the LDS stage the MFMAs
// read next.  Memory traffic is 1 bit per sample instead of 4.
template <class G>
__device__ __forceinline__ void kloop_packed(const WorkDesc &d, int i0, int j0,
                                             v16f (&acc)[G::NBM][G::NBN], uint8_t *smem)
{
    constexpr int TILE = G::EDGE, SB = G::SB, KSTEPS = G::KSTEPS;
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN;
    static_assert(G::NSTAGE == 2 && SB == 128 && 2 * TILE == G::THREADS, "packed geometry");
    auto sA = [&](int b) { return smem + b * G::STAGE_BYTES; };
    auto sB = [&](int b) { return smem + b * G::STAGE_BYTES + G::TILE_BYTES; };
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int wm = wave / G::WAVES_N, wn = wave % G::WAVES_N;
    const int W = d.words;
    const int stages = (W + 3) / 4;

    // Staging role of this lane.
    const int opnd = threadIdx.x / TILE, srow = threadIdx.x % TILE;
    const int grow = (opnd ? j0 : i0) + srow;
    const bool rvalid = grow < (opnd ? d.nj : d.ni);
    const uint64_t *src = (opnd ? d.PJ : d.PI) + (int64_t)(rvalid ? grow : 0) * W;
    uint8_t *dst_row = (opnd ? TILE * SB : 0) + srow * SB + smem;

    auto fetch = [&](uint64_t (&r)[4], int s) {
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int k = 4 * s + w;
            r[w] = (rvalid && k < W) ? src[k] : 0ull;
        }
    };
    auto expand = [&](const uint64_t (&r)[4], int buf) {
        uint8_t *row = dst_row + buf * G::STAGE_BYTES;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t x = (uint32_t)(r[c >> 1] >> (32 * (c & 1)));
            *(v4i *)(row + (G::swz(srow, c) << 4)) = fp4_chunk(x);
        }
    };

    const int r32 = lane & 31, half = lane >> 5;
    auto compute = [&](int buf) {
        v4i av[KSTEPS][NBM], bv[KSTEPS][NBN];
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            const int c = 2 * ks + half;
#pragma unroll
            for (int b = 0; b < NBN; ++b) {
                int row = wn * WTN + b * 32 + r32;
                bv[ks][b] = *(const v4i *)(sB(buf) + row * SB + (G::swz(row, c) << 4));
            }
#pragma unroll
            for (int a = 0; a < NBM; ++a) {
                int row = wm * WTM + a * 32 + r32;
                av[ks][a] = *(const v4i *)(sA(buf) + row * SB + (G::swz(row, c) << 4));
            }
        }
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks)
#pragma unroll
            for (int a = 0; a < NBM; ++a)
#pragma unroll
                for (int b = 0; b < NBN; ++b)
                    acc[a][b] = FP4::mma(av[ks][a], bv[ks][b], acc[a][b]);
    };

    // Three register sets in rotation (unrolled so each set is static):
    // stage s+1 is expanded while stage s+2 and s+3 are in flight.
    uint64_t R0[4], R1[4], R2[4];
    fetch(R0, 0);
    fetch(R1, 1);
    fetch(R2, 2);
    expand(R0, 0);
    fetch(R0, 3);
    // One step: barrier, expand stage s+1 from Rn, refill Rn with stage s+4
    // (Rn is free once expanded), compute stage s.
    auto step = [&](int s, uint64_t (&Rn)[4]) {
        lds_barrier();
        if (s + 1 < stages) {
            expand(Rn, (s + 1) & 1);
            fetch(Rn, s + 4);
        }
        compute(s & 1);
    };
    int s = 0;
    for (; s + 3 <= stages; s += 3) {
        step(s, R1);
        step(s + 1, R2);
        step(s + 2, R0);
    }
    if (s < stages) step(s, R1);
    if (s + 1 < stages) step(s + 1, R2);
}

// k loop straight from the packed rows (fp4, no LDS, no barrier), read from
// the row-block-tiled copy (launch_tile_rows).  For each group of 4 k-steps
// (4 packed words = 256 samples of a row) lane l of a 32 x 32 operand fetch
// loads 16 bytes -- words 2h, 2h+1 of the group, h = l/32 -- of row l%32
// with one buffer_load_dwordx4 (the wave's 64 loads are 1 KiB of
// consecutive bytes), split into four bit planes in registers (fp4_plane).
// Both operands use the same sample -> K mapping, so the products pair
// samples correctly.  The next group's words are in flight while this
// group's MFMAs run.  Rows past the table and words past the row end are
// zero in the tiled copy; rows past the job but inside the table are read
// and their products never stored.
// DOFF: the wave's block straddles a triangle's diagonal, its first column
// DOFF 32-blocks right of its first row: a 32 x 32 block (a, b) with
// a - b > DOFF lies wholly below the diagonal (no pair j > i) and is not
// computed (DOFF = QLD_NO_DIAG: every block is).
// MA, NA: only the first MA row and NA column blocks of the wave's block
// hold rows / columns of the job (a ragged edge): the others are neither
// loaded nor computed.  RO: nor are its first RO row blocks (shared-pair
// units whose windows hold only the block's lower rows).
// gb, ge: the word groups [gb, ge) to sum over (ge < 0: to the row end);
// split-K units take one such range of a tile.
#define QLD_NO_DIAG 99
template <class G, int DOFF = QLD_NO_DIAG, int MA = G::NBM, int NA = G::NBN,
          class Hook = NoHook, int RO = 0>
__device__ __forceinline__ void kloop_regs(const WorkDesc &d, int i0, int j0, int wm, int wn,
                                           v16f (&acc)[G::NBM][G::NBN],
                                           const Hook &hook = Hook(), int gb = 0, int ge = -1)
{
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN;
    const int lane = threadIdx.x & 63;
    const int r32 = lane & 31, h = lane >> 5;
    const int W = d.words;
    const int groups = (W + 3) / 4;
    const uint32_t blk = (uint32_t)groups * 1024u;     // bytes per 32-row block
    const int a0 = i0 + wm * WTM, b0 = j0 + wn * WTN;
    // Operand rows [r0, r0 + rows) of a tiled table of tab_rows rows: a
    // buffer descriptor over the 32-row blocks they touch (never past the
    // table: anything beyond reads as zero), and each lane's byte offset in
    // it for its row of 32-row MFMA block q (rows need not start on a block).
    auto rsrc = [&](const uint64_t *T, int64_t r0, int tab_rows, int rows) {
        const int64_t first = r0 >> 5;
        int64_t last = (r0 + rows + 31) >> 5;
        const int64_t tab_blocks = ((int64_t)tab_rows + 31) >> 5;
        if (last > tab_blocks) last = tab_blocks;
        const int64_t nb = last > first ? last - first : 0;
        const uint64_t base =
            wave_uniform((uint64_t)(uintptr_t)((const char *)T + first * (int64_t)blk));
        return __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)base, 0,
                                                 wave_uniform((int)(nb * blk)), 0x00020000);
    };
    auto lane_off = [&](int64_t r0, int q) {
        const int rel = (int)(r0 & 31) + 32 * q + r32;
        return (uint32_t)(rel >> 5) * blk + (uint32_t)(rel & 31) * 32u + 16u * (uint32_t)h;
    };
    const int64_t rA = d.rowI0 + a0, rB = d.rowJ0 + b0;
    const __amdgpu_buffer_rsrc_t rsA = rsrc(d.PI, rA, d.rowsI, WTM);
    const __amdgpu_buffer_rsrc_t rsB = rsrc(d.PJ, rB, d.rowsJ, WTN);
    uint32_t offA[NBM], offB[NBN];
#pragma unroll
    for (int a = 0; a < NBM; ++a) offA[a] = lane_off(rA, a);
#pragma unroll
    for (int b = 0; b < NBN; ++b) offB[b] = lane_off(rB, b);

    // Group g (4 words = 256 samples of every row) of the wave's blocks;
    // the tiled table is zero past each row's last word, so no masking.
    auto fetch = [&](v4i (&xa)[NBM], v4i (&xb)[NBN], int g) {
        const uint32_t k = (uint32_t)g * 1024u;
#pragma unroll
        for (int a = RO; a < MA; ++a)
            if (a - DOFF <= NA - 1)     // row block a meets some computed column block
                xa[a] = (v4i)__builtin_amdgcn_raw_buffer_load_b128(rsA, offA[a] + k, 0, 0);
#pragma unroll
        for (int b = 0; b < NA; ++b)
            xb[b] = (v4i)__builtin_amdgcn_raw_buffer_load_b128(rsB, offB[b] + k, 0, 0);
    };
    auto plane = [&](auto pc, const v4i (&xa)[NBM], const v4i (&xb)[NBN]) {
        constexpr int P = decltype(pc)::value;
        v4i bv[NBN];
#pragma unroll
        for (int b = 0; b < NA; ++b) bv[b] = fp4_plane<P>(xb[b]);
#pragma unroll
        for (int a = RO; a < MA; ++a) {
            if (a - DOFF > NA - 1) continue;
            const v4i av = fp4_plane<P>(xa[a]);
#pragma unroll
            for (int b = 0; b < NA; ++b)
                if (a - b <= DOFF)
                    acc[a][b] = FP4::mma_scaled<PlaneScale<P>::EA>(av, bv[b], acc[a][b]);
        }
    };
    auto compute = [&](const v4i (&xa)[NBM], const v4i (&xb)[NBN]) {
        plane(std::integral_constant<int, 0>(), xa, xb);
        plane(std::integral_constant<int, 1>(), xa, xb);
        plane(std::integral_constant<int, 2>(), xa, xb);
        plane(std::integral_constant<int, 3>(), xa, xb);
    };
    // Two register sets: group g+1 in flight while group g computes.  The
    // loads in the loop are unconditional, so it carries no register copies
    // (the group read past the last is beyond the rows' blocks or the
    // descriptor, and is never computed).
    const int full = ge < 0 ? groups : ge;
    v4i pa[NBM], pb[NBN], qa[NBM], qb[NBN];
    fetch(pa, pb, gb);
    // Drain every outstanding memory operation once here (the previous
    // tile's r^2 stores among them).  The compiler treats loads and stores
    // pending together as completing out of order and would otherwise wait
    // for everything (vmcnt 0) at every k step of the loop instead of for
    // the older group only.
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    hook.begin();
    int g = gb;
    for (; g + 2 <= full; g += 2) {
        // Scheduling barriers keep each group's loads a whole compute ahead
        // of their use (left alone, the scheduler sinks them next to the
        // first use; measured on the 30,000-SNP triangle: 2.58 vs 2.79 ms
        // with split 256-tiles, 2.76 vs 3.06 ms with 128-tiles).
        fetch(qa, qb, g + 1);
        __builtin_amdgcn_sched_barrier(0);
        compute(pa, pb);
        __builtin_amdgcn_sched_barrier(0);
        fetch(pa, pb, g + 2);
        __builtin_amdgcn_sched_barrier(0);
        compute(qa, qb);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (g < full) compute(pa, pb);
}

// k loop of one 64 x 64 wave block of a 128-tile (direct operands): a block
// that reaches past the job's last row or column computes only its first
// 32-row / 32-column half there; a square block on a triangle's diagonal
// skips its 32 x 32 block below the diagonal.  Word groups [gb, ge).
template <class G, class Hook = NoHook>
__device__ __forceinline__ void kloop_direct128(const WorkDesc &d, int i0, int j0, int wm,
                                                int wn, v16f (&acc)[G::NBM][G::NBN],
                                                const Hook &hook = Hook(), int gb = 0,
                                                int ge = -1)
{
    static_assert(G::WTM == 64 && G::WTN == 64, "128-tile wave blocks");
    const int a0 = i0 + wm * G::WTM, b0 = j0 + wn * G::WTN;
    const bool mfull = a0 + 32 < d.ni, nfull = b0 + 32 < d.nj;
    if (mfull && nfull) {
        if (d.tri && a0 == b0)
            kloop_regs<G, 0, G::NBM, G::NBN, Hook>(d, i0, j0, wm, wn, acc, hook, gb, ge);
        else
            kloop_regs<G, QLD_NO_DIAG, G::NBM, G::NBN, Hook>(d, i0, j0, wm, wn, acc, hook, gb,
                                                             ge);
    } else if (mfull) {
        kloop_regs<G, QLD_NO_DIAG, 2, 1, Hook>(d, i0, j0, wm, wn, acc, hook, gb, ge);
    } else if (nfull) {
        kloop_regs<G, QLD_NO_DIAG, 1, 2, Hook>(d, i0, j0, wm, wn, acc, hook, gb, ge);
    } else {
        kloop_regs<G, QLD_NO_DIAG, 1, 1, Hook>(d, i0, j0, wm, wn, acc, hook, gb, ge);
    }
}

template <class OP, int MODE, class G, bool INTERIOR>
__device__ __forceinline__ void tile_epilogue_impl(const WorkDesc &d, int i0, int j0, int wm, int wn,
                                                   int wave,
                                                   typename OP::acc_t (&acc)[G::NBM][G::NBN],
                                                   const uint32_t *bc_row, const uint32_t *bc_col,
                                                   uint8_t *smem)
{
    typedef Epi<MODE> E;
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN, RPL = G::RPL;
    const int lane = threadIdx.x & 63;
    const int r32 = lane & 31, half = lane >> 5;
    const int cq = lane % G::LANES_PER_ROW, rq = lane / G::LANES_PER_ROW;
    const int jc = j0 + wn * WTN + 4 * cq;              // first of this lane's 4 columns
    uint32_t *ep = (uint32_t *)(smem + wave * G::EP_WAVE_BYTES);
    uint4 *terms = (uint4 *)(smem + G::EP_TERMS_OFF) + wave * WTM;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
        if (lane + q * 64 < WTM) terms[q * 64 + lane] = E::pack(E::make(bc_row[q], d));
    typename E::term col[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) col[q] = E::make(bc_col[q], d);
#pragma unroll
    for (int a = 0; a < NBM; ++a) {
#pragma unroll
        for (int b = 0; b < NBN; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int row = 8 * (r >> 2) + 4 * half + (r & 3);
                ep[row * G::EP_PITCH + b * 32 + r32] = OP::bits(acc[a][b][r]);
            }
        const int ibase = i0 + wm * WTM + a * 32;
        // this lane's first row of the block; row steps are uniform offsets
        // (scalar arithmetic), so a row costs one vector address add
        g_f32 *const res0 = (g_f32 *)d.res + (int64_t)(ibase + rq) * d.ldr + jc;
#pragma unroll E::ROW_UNROLL
        for (int it = 0; it < 32 / G::ROWS_PER_IT; ++it) {
            const int rl = it * G::ROWS_PER_IT + rq;
            const int i = ibase + rl;
            const uint4 c4 = *(const uint4 *)(ep + rl * G::EP_PITCH + 4 * cq);
            const typename E::term rowt = E::unpack(terms[a * 32 + rl]);
            if (!INTERIOR && i >= d.ni) continue;
            const float cv[4] = {OP::count(c4.x), OP::count(c4.y), OP::count(c4.z),
                                 OP::count(c4.w)};
            float rv[4];
            E::r2x4(cv, col, rowt, d, rv);
            g_f32 *const rdst = res0 + (int64_t)(it * G::ROWS_PER_IT) * d.ldr;
            g_u32 *const cdst = (g_u32 *)d.cnt + (int64_t)i * d.ldc + jc;   // rarely asked for
            if constexpr (INTERIOR) {
                // Every column of the tile is inside the job (and above the
                // diagonal for a triangle) and rows are 16-byte aligned.
                // Non-temporal: the kernel never reads them back (1% faster
                // on the 30,000-SNP triangle).
                if (d.res)
                    __builtin_nontemporal_store(v4f{rv[0], rv[1], rv[2], rv[3]}, (g_v4f *)rdst);
                if (d.cnt)
                    __builtin_nontemporal_store(v4u{(uint32_t)cv[0], (uint32_t)cv[1],
                                                    (uint32_t)cv[2], (uint32_t)cv[3]},
                                                (g_v4u *)cdst);
            } else {
                if (d.res) {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (jc + q < d.nj && (!d.tri || jc + q > i)) rdst[q] = rv[q];
                }
                if (d.cnt) {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (jc + q < d.nj && (!d.tri || jc + q > i)) cdst[q] = (uint32_t)cv[q];
                }
            }
        }
    }
}

// Fused epilogue.  Block (a, b) register r holds the pair
//   row wm*WT + a*32 + 8*(r/4) + 4*(lane/32) + r%4, col wn*WT + b*32 + lane%32
// of the tile.  The r^2 terms of the wave's WT rows go to LDS once; then
// for each a the wave stages those 32 rows of counts in its own LDS region
// and every lane turns 4 consecutive counts of a row into r^2.  A tile
// wholly inside the job (and, for a triangle, wholly above the diagonal)
// with 16-byte aligned rows writes them with one 16-byte store per lane and
// row (rows of WT floats coalesce); any other tile checks every element.
template <class OP, int MODE, class G>
__device__ __forceinline__ void tile_epilogue(const WorkDesc &d, int i0, int j0, int wm, int wn,
                                              int wave, typename OP::acc_t (&acc)[G::NBM][G::NBN],
                                              const uint32_t *bc_row, const uint32_t *bc_col,
                                              uint8_t *smem)
{
    // the wave's own block decides: the direct kernels' waves work alone
    const int a0 = i0 + wm * G::WTM, b0 = j0 + wn * G::WTN;
    const bool inside = a0 + G::WTM <= d.ni && b0 + G::WTN <= d.nj &&
                        (!d.tri || b0 >= a0 + G::WTM);
    const bool aligned =
        (!d.res || ((d.ldr & 3) == 0 && (((uintptr_t)d.res) & 15) == 0)) &&
        (!d.cnt || ((d.ldc & 3) == 0 && (((uintptr_t)d.cnt) & 15) == 0));
    if (inside && aligned)
        tile_epilogue_impl<OP, MODE, G, true>(d, i0, j0, wm, wn, wave, acc, bc_row, bc_col, smem);
    else
        tile_epilogue_impl<OP, MODE, G, false>(d, i0, j0, wm, wn, wave, acc, bc_row, bc_col, smem);
}

// Shared-pair epilogue: one wave block of a triangle over a whole table,
// written into every window of its list that holds some of its cells.  Lane
// l owns column jb + l of the block (its bit count came in bc_col[0]): for
// each 32-row block the wave stages the counts in LDS, reads row by row its
// column's count (rows are uniform, row terms a broadcast read) and keeps
// the 32 r^2 values in registers; then, window by window, the rows inside
// the window (a uniform range) are stored at the window's
// out[(i - start) * size + (j - start)] for columns j in (i, window end):
// one store per row of up to 64 consecutive floats, its row address scalar,
// marked non-temporal (the kernel never reads them back; 1% faster on the
// benchmark's 4,096 windows).
template <class OP, int MODE, class G>
__device__ __forceinline__ void tile_epilogue_shared(const WorkDesc &d, int i0, int j0, int wm,
                                                     int wn, int wave,
                                                     typename OP::acc_t (&acc)[G::NBM][G::NBN],
                                                     const uint32_t *bc_row,
                                                     const uint32_t *bc_col, uint8_t *smem,
                                                     const uint32_t *wl, const WinOut *wins)
{
    typedef Epi<MODE> E;
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN, RPL = G::RPL;
    static_assert(WTN == 64, "one column per lane");
    const int lane = threadIdx.x & 63;
    const int r32 = lane & 31, half = lane >> 5;
    const int jb = j0 + wn * WTN, j = jb + lane;
    uint32_t *ep = (uint32_t *)(smem + wave * G::EP_WAVE_BYTES);
    uint4 *terms = (uint4 *)(smem + G::EP_TERMS_OFF) + wave * WTM;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
        if (lane + q * 64 < WTM) terms[q * 64 + lane] = E::pack(E::make(bc_row[q], d));
    const typename E::term colt = E::make(bc_col[0], d);
    const int nw = (int)(load_uniform(wl, 0) & 0xFFFFu);
#pragma unroll
    for (int a = 0; a < NBM; ++a) {
#pragma unroll
        for (int b = 0; b < NBN; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int row = 8 * (r >> 2) + 4 * half + (r & 3);
                ep[row * G::EP_PITCH + b * 32 + r32] = OP::bits(acc[a][b][r]);
            }
        const int ibase = i0 + wm * WTM + a * 32;
        if (ibase >= d.ni) continue;
        float r2v[32];
#pragma unroll
        for (int g = 0; g < 8; ++g) {
            float cv[4];
            typename E::term rowt[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cv[u] = OP::count(ep[(4 * g + u) * G::EP_PITCH + lane]);
                rowt[u] = E::unpack(terms[a * 32 + 4 * g + u]);
            }
            float rv[4];
            E::r2x4_rows(cv, colt, rowt, d, rv);
#pragma unroll
            for (int u = 0; u < 4; ++u) r2v[4 * g + u] = rv[u];
        }
        for (int k = 0; k < nw; ++k) {
            const WinOut w = load_uniform(wins, (int64_t)load_uniform(wl, 1 + k));
            const int ws = w.start, we = w.start + w.size;
            const int r0 = (ws > ibase ? ws : ibase) - ibase;
            const int r1 = (we < ibase + 32 ? we : ibase + 32) - ibase;
            if (r0 >= r1) continue;
            g_f32 *row = (g_f32 *)w.out + (int64_t)(ibase - ws) * w.size + (jb - ws);
            const bool colok = j < we;
            if (r0 == 0 && r1 == 32 && jb >= ibase + 32) {
                // all 32 rows inside the window and every column right of
                // them: one lane mask for the block, scalar row addresses
                if (colok) {
#pragma unroll
                    for (int rl = 0; rl < 32; ++rl, row += w.size)
                        __builtin_nontemporal_store(r2v[rl], &row[lane]);
                }
            } else {
#pragma unroll
                for (int rl = 0; rl < 32; ++rl, row += w.size)
                    if (rl >= r0 && rl < r1 && colok && j > ibase + rl)
                        __builtin_nontemporal_store(r2v[rl], &row[lane]);
            }
        }
    }
}

// One tile of one job: k loop (expanded operands by DMA, or packed operands
// expanded in the kernel), then the fused epilogue.
// wm, wn: the wave's block of the tile; wave: the wave's LDS epilogue
// region.  Cooperative k loops (LDS ring) need wm, wn = this wave's own.
template <class OP, int MODE, class G, int PACKED, bool WAVE_UNITS = false, class Hook = NoHook,
          bool SHARED = false>
__device__ __forceinline__ void mfma_tile(const WorkDesc &d, int ti, int tj, int wm, int wn,
                                          int wave, uint8_t *smem, const Hook &hook = Hook(),
                                          const uint32_t *wl = nullptr,
                                          const WinOut *wins = nullptr)
{
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN, TILE = G::EDGE;
    constexpr int RPL = G::RPL;                         // epilogue rows per lane
    const int i0 = ti * TILE, j0 = tj * TILE;
    const int lane = threadIdx.x & 63;
    // Direct operands: a wave whose block of a triangle's diagonal tile lies
    // wholly below the diagonal (every column < every row: no j > i), or
    // whose block has no row or no column inside the job (ragged edges of
    // short windows), has nothing to compute or write; its waves do not
    // share a k loop with the others, so it simply skips the tile.
    if constexpr (PACKED == 2)
        if ((d.tri && j0 + wn * WTN + WTN <= i0 + wm * WTM + 1) || i0 + wm * WTM >= d.ni ||
            j0 + wn * WTN >= d.nj) {
            hook.begin();
            hook.end();
            return;
        }

    // Bit counts of this lane's epilogue rows/columns, fetched now so their
    // latency hides under the k loop.
    uint32_t bc_row[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        int i = i0 + wm * WTM + q * 64 + lane;
        bc_row[q] = (lane + q * 64 < WTM && i < d.ni) ? ((g_cu32 *)d.bcI)[i] : 0u;
    }
    const int jc = j0 + wn * WTN + 4 * (lane % G::LANES_PER_ROW);
    uint32_t bc_col[4];
    if constexpr (SHARED) {
        // the shared-pair epilogue gives each lane one column
        const int jl = j0 + wn * WTN + lane;
        bc_col[0] = jl < d.nj ? ((g_cu32 *)d.bcJ)[jl] : 0u;
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) bc_col[q] = jc + q < d.nj ? ((g_cu32 *)d.bcJ)[jc + q] : 0u;
    }

    typename OP::acc_t acc[NBM][NBN];
#pragma unroll
    for (int a = 0; a < NBM; ++a)
#pragma unroll
        for (int b = 0; b < NBN; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;

    static_assert(G::SPLIT == 1 || PACKED == 2, "split tiles need direct operands");
    if constexpr (PACKED == 2) {
        // no LDS ring: epilogue regions are per wave
        const int a0 = i0 + wm * WTM, b0 = j0 + wn * WTN;
        if constexpr (WTM == WTN && NBM == 2 && NBN == 2) {
            // 128-tiles (the short windows of plans)
            kloop_direct128<G, Hook>(d, i0, j0, wm, wn, acc, hook);
        } else if constexpr (WAVE_UNITS && WTM == 128 && WTN == 64) {
            // 128 x 64 wave blocks of 256-tiles: a block whose columns start
            // on its first row (a triangle's diagonal) has pairs j > i only
            // in its first 64 rows, and only in their 32-blocks on or right
            // of the block diagonal: 3 of its 8 MFMA blocks; a block whose
            // second 32 columns lie past the job's last column (a window's
            // ragged right edge) computes its first 32 columns only.  (The
            // block starting 64 columns right of its rows lacks only its
            // last 32-block; it runs in full: that variant pushes the kernel
            // past its register budget.)  Only the per-wave kernels of plans
            // carry the variants: in a single region such blocks are rare
            // and the extra code would only grow the kernel.  On the
            // benchmark's 4,096 windows the variants leave 7.6% of the
            // MFMA blocks outside the triangles, against 10.4% without the
            // ragged-edge ones (the 32 x 32 granularity's floor: 6.2%).
            const bool nfull = b0 + 32 < d.nj;
            // shared-pair units: the host marks blocks whose windows hold
            // cells only below their first 1-3 row blocks, or only in their
            // first 32 columns (the union of the windows' triangles ends
            // inside the block); those blocks are not computed
            int sub = 0;
            if constexpr (SHARED) sub = (int)(load_uniform(wl, 0) >> 16);
            if (d.tri && b0 == a0) {
                if (nfull)
                    kloop_regs<G, 0, 2, 2, Hook>(d, i0, j0, wm, wn, acc, hook);
                else
                    kloop_regs<G, 0, 1, 1, Hook>(d, i0, j0, wm, wn, acc, hook);
            } else if (!nfull || (SHARED && sub == QLD_SUB_HALF)) {
                kloop_regs<G, QLD_NO_DIAG, G::NBM, 1, Hook>(d, i0, j0, wm, wn, acc, hook);
            } else if (SHARED && sub == 1) {
                kloop_regs<G, QLD_NO_DIAG, G::NBM, G::NBN, Hook, 1>(d, i0, j0, wm, wn, acc, hook);
            } else if (SHARED && sub == 2) {
                kloop_regs<G, QLD_NO_DIAG, G::NBM, G::NBN, Hook, 2>(d, i0, j0, wm, wn, acc, hook);
            } else if (SHARED && sub == 3) {
                kloop_regs<G, QLD_NO_DIAG, G::NBM, G::NBN, Hook, 3>(d, i0, j0, wm, wn, acc, hook);
            } else {
                kloop_regs<G, QLD_NO_DIAG, G::NBM, G::NBN, Hook>(d, i0, j0, wm, wn, acc, hook);
            }
        } else {
            kloop_regs<G, QLD_NO_DIAG, G::NBM, G::NBN, Hook>(d, i0, j0, wm, wn, acc, hook);
        }
    } else {
        hook.begin();
        if constexpr (PACKED == 1) kloop_packed<G>(d, i0, j0, acc, smem);
        else kloop_dma<OP, G>(d, i0, j0, acc, smem);
        lds_barrier();   // every wave is done with the ring before the epilogue reuses it
    }
    hook.end();
    if constexpr (SHARED)
        tile_epilogue_shared<OP, MODE, G>(d, i0, j0, wm, wn, wave, acc, bc_row, bc_col, smem, wl,
                                          wins);
    else
        tile_epilogue<OP, MODE, G>(d, i0, j0, wm, wn, wave, acc, bc_row, bc_col, smem);
}

template <class OP, int MODE, class G, int PACKED, class Hook = NoHook>
__device__ __forceinline__ void run_item(const KArgs &p, int64_t idx, int wm, int wn, int wave,
                                         uint8_t *smem, const Hook &hook = Hook())
{
    const WorkItem it = load_uniform(p.items, idx);
    if (p.descs) {
        const WorkDesc d = load_uniform(p.descs, it.desc);
        mfma_tile<OP, MODE, G, PACKED, false, Hook>(d, it.ti, it.tj, wm, wn, wave, smem, hook);
    } else {
        mfma_tile<OP, MODE, G, PACKED, false, Hook>(p.inline_desc, it.ti, it.tj, wm, wn, wave,
                                                    smem, hook);
    }
}

// One wave block of a per-wave unit list: item (desc, bi, bj) names rows
// [bi*WTM, bi*WTM + WTM) and columns [bj*WTN, bj*WTN + WTN) of the job.
template <class OP, int MODE, class G, int PACKED, bool SHARED, class Hook>
__device__ __forceinline__ void run_wave_unit(const KArgs &p, int64_t idx, int wave,
                                              uint8_t *smem, const Hook &hook)
{
    constexpr int BM = G::EDGE / G::WTM, BN = G::EDGE / G::WTN;   // wave blocks per tile side
    const WorkItem it = load_uniform(p.items, idx);
    const int bi = it.ti, bj = it.tj;
    if constexpr (SHARED) {
        mfma_tile<OP, MODE, G, PACKED, true, Hook, true>(p.inline_desc, bi / BM, bj / BN, bi % BM,
                                                         bj % BN, wave, smem, hook,
                                                         p.wlist + it.desc, p.wins);
    } else if (p.descs) {
        const WorkDesc d = load_uniform(p.descs, it.desc);
        mfma_tile<OP, MODE, G, PACKED, true, Hook>(d, bi / BM, bj / BN, bi % BM, bj % BN, wave,
                                                   smem, hook);
    } else {
        mfma_tile<OP, MODE, G, PACKED, true, Hook>(p.inline_desc, bi / BM, bj / BN, bi % BM,
                                                   bj % BN, wave, smem, hook);
    }
}

// Persistent launches: a fixed grid drains eight per-XCD queues; workgroup
// b starts on queue b % 8 -- the XCD it runs on -- whose units come from a
// contiguous range of the list, then steals from the other queues.  Every
// claim advances a head, so the loop ends once all eight ranges are
// exhausted.  Thread 0 claims for the workgroup.
// PERSIST == 1: a claim is one unit, item u / SPLIT, wave rows
//   (u % SPLIT) * WAVES_M onwards (the queue heads count units).
// PERSIST == 2 (direct operands only): the list holds wave blocks, only those
//   with something to compute, and a claim is a strip of WAVES consecutive
//   ones, one per wave (the queue heads count strips).  A strip's blocks
//   mostly share their rows, and the workgroup's waves run them side by side
//   (one barrier per strip), so each row block is fetched once for the four
//   of them; waves claiming blocks on their own ran 8% slower on a 30,000-SNP
//   region.
// Claims run one unit ahead: the next claim is issued once the leader's k
// loop has its first operands (its initial drain leaves nothing pending, the
// previous unit's r^2 stores included) and published in LDS after that k
// loop, so no claim waits behind stores.  The two LDS slots alternate, so
// the barrier that ends a unit is the only one.
template <class OP, int MODE, class G, int PERSIST, int PACKED>
__device__ __forceinline__ void run_queues(const KArgs &p, int wave, uint8_t *smem)
{
    constexpr int64_t PER = PERSIST >= 2 ? G::WAVES : 1;   // list entries per claim
    constexpr int64_t SUB = PERSIST >= 2 ? 1 : G::SPLIT;   // claims per list entry
    const int wr = wave / G::WAVES_N, wn = wave % G::WAVES_N;
    __shared__ uint32_t claimed[2];
    __shared__ int claimed_x[2];      // -1 once every queue is drained
    int x = (int)(blockIdx.x % QLD_NUM_XCD);
    int left = QLD_NUM_XCD;           // queues thread 0 has not found empty
    const bool leader = threadIdx.x == 0;
    // thread 0: publish head k of queue x in slot s, or move on to the next
    // queue (a blocking claim there) if queue x is drained
    auto settle = [&](int s, uint32_t k) {
        for (;;) {
            if ((int64_t)k * PER < (p.qbeg[x + 1] - p.qbeg[x]) * SUB) {
                claimed[s] = k;
                claimed_x[s] = x;
                return;
            }
            if (--left == 0) {
                claimed_x[s] = -1;
                return;
            }
            x = (x + 1) % QLD_NUM_XCD;
            k = claim_head(&p.queue[x]);
        }
    };
    if (leader) settle(0, claim_head(&p.queue[x]));
    __syncthreads();
    for (int slot = 0;; slot ^= 1) {
        const int cx = __builtin_amdgcn_readfirstlane(claimed_x[slot]);
        if (cx < 0) break;
        const int64_t k = (uint32_t)__builtin_amdgcn_readfirstlane((int)claimed[slot]);
        uint32_t pend = 0;
        const auto ahead = make_hooks(
            [&]() {
                if (leader) pend = claim_head(&p.queue[x]);
            },
            [&]() {
                // read in every wave, not only under the leader's branch:
                // the compiler then knows the claim complete on every path
                const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)pend);
                if (leader) settle(slot ^ 1, c);
            });
        if constexpr (PERSIST >= 2) {
            const int64_t idx = p.qbeg[cx] + k * PER + wave;
            // the last strip of a queue may be short (wave 0, the leader's,
            // always has a block)
            if (idx < p.qbeg[cx + 1])
                run_wave_unit<OP, MODE, G, PACKED, PERSIST == 3>(p, idx, wave, smem, ahead);
        } else {
            const int64_t u = p.qbeg[cx] * SUB + k;
            run_item<OP, MODE, G, PACKED>(p, u / SUB, (int)(u % SUB) * G::WAVES_M + wr, wn, wave,
                                          smem, ahead);
        }
        __syncthreads();
    }
}

template <class OP, int MODE, class G, int PERSIST, int PACKED>
__global__ __launch_bounds__(G::THREADS, G::SPLIT) void ld_mfma_kernel(KArgs p)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    const int wave = wave_id();
    const int wr = wave / G::WAVES_N, wn = wave % G::WAVES_N;
    static_assert(PERSIST < 2 || PACKED == 2, "per-wave units need direct operands");
    if constexpr (PERSIST == 0) {
        // one workgroup per unit: workgroup b takes unit b
        const int64_t u = (int64_t)blockIdx.x;
        run_item<OP, MODE, G, PACKED>(p, u / G::SPLIT, (int)(u % G::SPLIT) * G::WAVES_M + wr,
                                      wn, wave, smem);
    } else {
        run_queues<OP, MODE, G, PERSIST, PACKED>(p, wave, smem);
    }
}

// Split-K units (128-tiles, direct fp4 operands), for jobs too small to
// fill the GPU with whole tiles: a tile's k loop is long (260 word groups
// of dependent loads at 16,622 samples) and one workgroup per tile leaves
// most CUs idle.  Unit u of the launch is chunk c = u % S of item u / S:
// the word groups [G*c/S, G*(c+1)/S) of its tile (G = the job's groups of 4
// words).  Every wave writes its partial sums -- integers, exact in fp32 --
// to its own slice of `kpart`, then the workgroup counts itself in at
// kcnt[item]; the last of the item's S workgroups to arrive adds the other
// S-1 slices to its own sums (exact, in any order), resets the counter for
// the next launch and runs the r^2 epilogue.  No workgroup waits for
// another.
// Visibility across XCDs (each has its own L2): the slices are written and
// read as agent-scope relaxed atomics (stores and loads coherent at the
// device's coherence point, past the XCD's L2), every wave waits for its
// stores to complete before the workgroup barrier that precedes the arrival
// atomic, and the last workgroup reads the slices only after that atomic
// returned.  C++ release/acquire fences would do the same by writing back
// and invalidating the whole L2 of the XCD at every arrival, which evicts
// the operand rows of every other workgroup there (the split launch ran 2-7x
// slower than the unsplit one with them).
// P16: every chunk's counts are below 2^16 (at most 255 word groups per
// chunk), and the slices hold them as 16-bit integers: half the bytes of
// fp32 slices, which are about half the kernel's HBM traffic.
constexpr int SPLITK_SLICE = GeoS::NBM * GeoS::NBN * 16 * 64;   // floats per wave

template <int MODE, bool P16>
__global__ __launch_bounds__(GeoS::THREADS) void ld_splitk_kernel(KArgs p)
{
    typedef GeoS G;
    constexpr int NBM = G::NBM, NBN = G::NBN, WTM = G::WTM, WTN = G::WTN, TILE = G::EDGE;
    constexpr int RPL = G::RPL;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int last;
    const int S = p.ksplit;
    const int64_t item = (int64_t)blockIdx.x / S;
    const int c = (int)((int64_t)blockIdx.x % S);
    const WorkItem it = load_uniform(p.items, item);
    const WorkDesc d = p.descs ? load_uniform(p.descs, (int64_t)it.desc) : p.inline_desc;
    const int wave = wave_id();
    const int wm = wave / G::WAVES_N, wn = wave % G::WAVES_N;
    const int lane = threadIdx.x & 63;
    const int i0 = it.ti * TILE, j0 = it.tj * TILE;
    const int a0 = i0 + wm * WTM, b0 = j0 + wn * WTN;
    // as in mfma_tile: a block with no pair j > i, or no row or column in
    // the job, computes and writes nothing (the same in every chunk)
    const bool skip = (d.tri && b0 + WTN <= a0 + 1) || a0 >= d.ni || b0 >= d.nj;
    const int groups = (d.words + 3) / 4;
    const int gb = (int)((int64_t)groups * c / S), ge = (int)((int64_t)groups * (c + 1) / S);

    uint32_t bc_row[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
        const int i = a0 + q * 64 + lane;
        bc_row[q] = (lane + q * 64 < WTM && i < d.ni) ? ((g_cu32 *)d.bcI)[i] : 0u;
    }
    const int jc = b0 + 4 * (lane % G::LANES_PER_ROW);
    uint32_t bc_col[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bc_col[q] = jc + q < d.nj ? ((g_cu32 *)d.bcJ)[jc + q] : 0u;

    v16f acc[NBM][NBN];
#pragma unroll
    for (int a = 0; a < NBM; ++a)
#pragma unroll
        for (int b = 0; b < NBN; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;
    if (!skip) kloop_direct128<G>(d, i0, j0, wm, wn, acc, NoHook(), gb, ge);

    // slice (item, chunk, wave), in 8-byte words: word 64 * (PW * block + k)
    // + lane holds registers RW * k .. RW * k + RW - 1 of MFMA block
    // a * NBN + b (RW = 2 fp32 values, or 4 16-bit integers with P16)
    constexpr int WORDS = SPLITK_SLICE / 2;        // slice stride in 8-byte words
    constexpr int RW = P16 ? 4 : 2, PW = 16 / RW;  // registers per word, words per block
    uint64_t *const slices = (uint64_t *)p.kpart + (item * S * G::WAVES + wave) * WORDS + lane;
    auto word = [&](int a, int b, int k) {
        uint64_t w = 0;
#pragma unroll
        for (int e = 0; e < RW; ++e)
            w |= P16 ? (uint64_t)(uint32_t)acc[a][b][RW * k + e] << (16 * e)
                     : (uint64_t)__float_as_uint(acc[a][b][RW * k + e]) << (32 * e);
        return w;
    };
    if (!skip) {
        uint64_t *const mine = slices + (int64_t)c * G::WAVES * WORDS;
#pragma unroll
        for (int a = 0; a < NBM; ++a)
#pragma unroll
            for (int b = 0; b < NBN; ++b)
#pragma unroll
                for (int k = 0; k < PW; ++k)
                    __hip_atomic_store(mine + 64 * (PW * (a * NBN + b) + k), word(a, b, k),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wave's slice is written
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev =
            __hip_atomic_fetch_add(p.kcnt + item, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = prev == (uint32_t)(S - 1);
        // every unit of the item has counted itself in: clear for the next launch
        if (last) __hip_atomic_store(p.kcnt + item, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!last || skip) return;
    for (int s2 = 0; s2 < S; ++s2) {
        if (s2 == c) continue;
        uint64_t *const other = slices + (int64_t)s2 * G::WAVES * WORDS;
        uint64_t v[NBM * NBN * PW];
#pragma unroll
        for (int k = 0; k < NBM * NBN * PW; ++k)
            v[k] = __hip_atomic_load(other + 64 * k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int a = 0; a < NBM; ++a)
#pragma unroll
            for (int b = 0; b < NBN; ++b)
#pragma unroll
                for (int k = 0; k < PW; ++k) {
                    const uint64_t x = v[PW * (a * NBN + b) + k];
#pragma unroll
                    for (int e = 0; e < RW; ++e)
                        acc[a][b][RW * k + e] +=
                            P16 ? (float)(uint32_t)((x >> (16 * e)) & 0xFFFFu)
                                : __uint_as_float((uint32_t)(x >> (32 * e)));
                }
    }
    tile_epilogue<FP4, MODE, G>(d, i0, j0, wm, wn, wave, acc, bc_row, bc_col, smem);
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
    // Rows hold whole 128-byte k stages: 4 packed words as fp4, 2 as int8.
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
    d.PI = nullptr;
    d.PJ = nullptr;
    d.rowI0 = d.rowJ0 = 0;
    d.rowsI = d.rowsJ = 0;
    d.words = 0;
    d.pad2_ = 0;
}

void set_packed_operands(WorkDesc &d, const uint64_t *PI, const uint64_t *PJ, int words)
{
    d.PI = PI;
    d.PJ = PJ;
    d.rowI0 = d.rowJ0 = 0;
    d.rowsI = d.ni;
    d.rowsJ = d.nj;
    d.words = words;
}

void set_tiled_operands(WorkDesc &d, const uint64_t *TI, int64_t rowI0, int rowsI,
                        const uint64_t *TJ, int64_t rowJ0, int rowsJ, int words)
{
    d.PI = TI;
    d.PJ = TJ;
    d.rowI0 = rowI0;
    d.rowJ0 = rowJ0;
    d.rowsI = rowsI;
    d.rowsJ = rowsJ;
    d.words = words;
}

int64_t tiled_bytes(int rows, int words)
{
    return (int64_t)((rows + 31) / 32) * ((words + 3) / 4) * 1024;
}

// One thread per cell of the job's tile grid; a triangle's cells below the
// diagonal write nothing.
__global__ void job_tiles_kernel(WorkItem *out, int ti_count, int tj_count, int tri)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ti_count * tj_count) return;
    const int ti = (int)(t % ti_count), tj = (int)(t / ti_count);
    if (tri && ti > tj) return;
    out[job_tile_index(ti, tj, ti_count, tj_count, tri)] =
        WorkItem{0u, (uint16_t)ti, (uint16_t)tj};
}

hipError_t launch_job_tiles(WorkItem *out, int ni, int nj, int tri, int T, hipStream_t s)
{
    const int ti_count = (ni + T - 1) / T, tj_count = (nj + T - 1) / T;
    const int64_t cells = (int64_t)ti_count * tj_count;
    if (cells <= 0) return hipSuccess;
    hipLaunchKernelGGL(job_tiles_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s,
                       out, ti_count, tj_count, tri);
    return hipGetLastError();
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

// One thread per 32-byte piece (row r, word group g) of the tiled table:
// consecutive threads write consecutive pieces; each reads its row's four
// words (zero past the row or the table).  zq (optional): 16 queue heads the
// first threads clear for the persistent launch that follows, which then
// needs no reset kernel of its own.
__global__ void tile_rows_kernel(const uint64_t *P, int rows, int W, int groups,
                                 int64_t pieces, uint64_t *T, uint32_t *zq)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (zq && t < 16) zq[t] = 0u;
    if (t >= pieces) return;
    const int r = (int)(t & 31);
    const int64_t bg = t >> 5;
    const int g = (int)(bg % groups);
    const int row = (int)(bg / groups) * 32 + r;
    uint64_t v[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int word = 4 * g + w;
        v[w] = (row < rows && word < W) ? P[(int64_t)row * W + word] : 0ull;
    }
    ulonglong2 *dst = (ulonglong2 *)(T + t * 4);
    dst[0] = make_ulonglong2(v[0], v[1]);
    dst[1] = make_ulonglong2(v[2], v[3]);
}

hipError_t launch_tile_rows(const uint64_t *P, int rows, int words, uint64_t *T,
                            hipStream_t s, uint32_t *zero_queue)
{
    if (rows <= 0 || words <= 0) return zero_queue ? hipErrorInvalidValue : hipSuccess;
    const int groups = (words + 3) / 4;
    const int64_t pieces = (int64_t)((rows + 31) / 32) * groups * 32;
    hipLaunchKernelGGL(tile_rows_kernel, dim3((unsigned)((pieces + 255) / 256)), dim3(256), 0, s,
                       P, rows, words, groups, pieces, T, zero_queue);
    return hipGetLastError();
}

__global__ void zero_queue_kernel(uint32_t *q)
{
    if (threadIdx.x < 16) q[threadIdx.x] = 0u;
}

// Dynamic LDS of a kernel variant: the k-stage ring (which the epilogue
// reuses), or only the epilogue regions when operands go straight to
// registers.
template <class G, int PACKED>
constexpr int lds_bytes() { return PACKED == 2 ? G::EP_BYTES : G::LDS_BYTES; }

template <class OP, int MODE, class G, int PERSIST, int PACKED = 0>
static hipError_t set_lds_limit()
{
    return hipFuncSetAttribute((const void *)ld_mfma_kernel<OP, MODE, G, PERSIST, PACKED>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               lds_bytes<G, PACKED>());
}

template <class OP, int MODE>
static hipError_t set_lds_limits_mode()
{
    hipError_t e;
    if ((e = set_lds_limit<OP, MODE, GeoL8, 0>()) != hipSuccess) return e;
    if ((e = set_lds_limit<OP, MODE, GeoL8, 1>()) != hipSuccess) return e;
    if ((e = set_lds_limit<OP, MODE, GeoS, 0>()) != hipSuccess) return e;
    return set_lds_limit<OP, MODE, GeoS, 1>();
}

template <int MODE>
static hipError_t set_lds_limits_packed()
{
    hipError_t e;
    if ((e = set_lds_limit<FP4, MODE, GeoS, 0, 1>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoS, 1, 1>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoS, 0, 2>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoS, 1, 2>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoL8H, 0, 2>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoS, 2, 2>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoL8H, 2, 2>()) != hipSuccess) return e;
    if ((e = set_lds_limit<FP4, MODE, GeoL8H, 3, 2>()) != hipSuccess) return e;
    return set_lds_limit<FP4, MODE, GeoL8H, 1, 2>();
}

// Raise the dynamic LDS limit of every kernel variant (done once per
// process, before any launch -- and so before any graph capture).
hipError_t mfma_init_attributes()
{
    hipError_t e;
    if ((e = set_lds_limits_mode<FP4, QLD_R2_REFERENCE>()) != hipSuccess) return e;
    if ((e = set_lds_limits_mode<FP4, QLD_R2_EXACT>()) != hipSuccess) return e;
    if ((e = set_lds_limits_mode<I8, QLD_R2_REFERENCE>()) != hipSuccess) return e;
    if ((e = set_lds_limits_mode<I8, QLD_R2_EXACT>()) != hipSuccess) return e;
    if ((e = set_lds_limits_packed<QLD_R2_REFERENCE>()) != hipSuccess) return e;
    return set_lds_limits_packed<QLD_R2_EXACT>();
}

template <class OP, int MODE, class G, int PERSIST, int PACKED = 0>
static hipError_t launch_one(const KArgs &a, unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL((ld_mfma_kernel<OP, MODE, G, PERSIST, PACKED>), dim3(grid),
                       dim3(G::THREADS), (size_t)(lds_bytes<G, PACKED>()), s, a);
    return hipGetLastError();
}

template <class OP, int MODE, class G, int PACKED = 0>
static hipError_t launch_mode(const KArgs &a, int persistent, hipStream_t s)
{
    if (persistent > 0)
        return launch_one<OP, MODE, G, 1, PACKED>(a, (unsigned)persistent * G::SPLIT, s);
    return launch_one<OP, MODE, G, 0, PACKED>(a, (unsigned)(a.nitems * G::SPLIT), s);
}

template <class OP, int MODE>
static hipError_t launch_tile(const KArgs &a, int tile, int persistent, hipStream_t s)
{
    if (tile == QLD_TILE_SMALL) return launch_mode<OP, MODE, GeoS>(a, persistent, s);
    return launch_mode<OP, MODE, GeoL8>(a, persistent, s);
}

hipError_t launch_mfma_wave_units(int r2_mode, int tile, const WorkDesc *descs,
                                  const WorkDesc &inline_desc, const WorkItem *units,
                                  int64_t nunits, int grid, uint32_t *queue, hipStream_t s)
{
    if (nunits <= 0) return hipSuccess;
    if (!queue || grid <= 0) return hipErrorInvalidValue;
    KArgs a = {};
    a.inline_desc = inline_desc;
    a.descs = descs;
    a.items = units;
    a.nitems = nunits;
    a.queue = queue;
    for (int x = 0; x <= QLD_NUM_XCD; ++x) a.qbeg[x] = nunits * x / QLD_NUM_XCD;
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(64), 0, s, queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (tile == QLD_TILE_SMALL)
        return r2_mode == QLD_R2_EXACT
                   ? launch_one<FP4, QLD_R2_EXACT, GeoS, 2, 2>(a, (unsigned)grid, s)
                   : launch_one<FP4, QLD_R2_REFERENCE, GeoS, 2, 2>(a, (unsigned)grid, s);
    return r2_mode == QLD_R2_EXACT
               ? launch_one<FP4, QLD_R2_EXACT, GeoL8H, 2, 2>(a, (unsigned)grid, s)
               : launch_one<FP4, QLD_R2_REFERENCE, GeoL8H, 2, 2>(a, (unsigned)grid, s);
}

hipError_t launch_mfma_shared_units(int r2_mode, const WorkDesc &inline_desc,
                                    const WorkItem *units, int64_t nunits,
                                    const uint32_t *wlist, const WinOut *wins, int grid,
                                    uint32_t *queue, hipStream_t s)
{
    if (nunits <= 0) return hipSuccess;
    if (!queue || !wlist || !wins || grid <= 0) return hipErrorInvalidValue;
    KArgs a = {};
    a.inline_desc = inline_desc;
    a.descs = nullptr;
    a.items = units;
    a.nitems = nunits;
    a.queue = queue;
    a.wlist = wlist;
    a.wins = wins;
    for (int x = 0; x <= QLD_NUM_XCD; ++x) a.qbeg[x] = nunits * x / QLD_NUM_XCD;
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(64), 0, s, queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return r2_mode == QLD_R2_EXACT
               ? launch_one<FP4, QLD_R2_EXACT, GeoL8H, 3, 2>(a, (unsigned)grid, s)
               : launch_one<FP4, QLD_R2_REFERENCE, GeoL8H, 3, 2>(a, (unsigned)grid, s);
}

hipError_t launch_mfma_splitk(int r2_mode, const WorkDesc *descs, const WorkDesc &inline_desc,
                              const WorkItem *items, int64_t nitems, int ksplit, int max_groups,
                              float *kpart, uint32_t *kcnt, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    if (ksplit < 1 || ksplit > QLD_MAX_KSPLIT || !kpart || !kcnt || max_groups < 1 ||
        nitems * ksplit > 0x7FFFFFFF)
        return hipErrorInvalidValue;
    // a chunk of g word groups counts at most 256 g samples per pair
    const bool p16 = (max_groups + ksplit - 1) / ksplit <= 255;
    KArgs a = {};
    a.inline_desc = inline_desc;
    a.descs = descs;
    a.items = items;
    a.nitems = nitems;
    a.ksplit = ksplit;
    a.kpart = kpart;
    a.kcnt = kcnt;
    const unsigned grid = (unsigned)(nitems * ksplit);
    constexpr size_t lds = lds_bytes<GeoS, 2>();
    const dim3 g(grid), b(GeoS::THREADS);
    if (r2_mode == QLD_R2_EXACT) {
        if (p16) hipLaunchKernelGGL((ld_splitk_kernel<QLD_R2_EXACT, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((ld_splitk_kernel<QLD_R2_EXACT, false>), g, b, lds, s, a);
    } else {
        if (p16) hipLaunchKernelGGL((ld_splitk_kernel<QLD_R2_REFERENCE, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((ld_splitk_kernel<QLD_R2_REFERENCE, false>), g, b, lds, s, a);
    }
    return hipGetLastError();
}

size_t splitk_part_bytes(int64_t nitems, int ksplit)
{
    return (size_t)nitems * (size_t)ksplit * GeoS::WAVES * SPLITK_SLICE * sizeof(float);
}

void wave_block_shape(int tile, int *wtm, int *wtn)
{
    *wtm = tile == QLD_TILE_SMALL ? GeoS::WTM : GeoL8H::WTM;
    *wtn = tile == QLD_TILE_SMALL ? GeoS::WTN : GeoL8H::WTN;
}

hipError_t launch_mfma_items(int engine, int r2_mode, int tile, int packed,
                             const WorkDesc *descs, const WorkDesc &inline_desc,
                             const WorkItem *items, int64_t nitems,
                             int persistent, uint32_t *queue, const int64_t *qbeg,
                             hipStream_t s, bool queue_zeroed)
{
    if (nitems <= 0) return hipSuccess;
    KArgs a = {};
    a.inline_desc = inline_desc;
    a.descs = descs;
    a.items = items;
    a.nitems = nitems;
    a.queue = queue;
    for (int x = 0; x <= QLD_NUM_XCD; ++x)
        a.qbeg[x] = qbeg ? qbeg[x] : nitems * x / QLD_NUM_XCD;
    if (persistent > 0) {
        if (!queue) return hipErrorInvalidValue;
        if (!queue_zeroed) {
            hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(64), 0, s, queue);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    if (packed == 2) {      // packed rows straight into registers
        if (engine != QLD_ENGINE_MFMA_FP4) return hipErrorInvalidValue;
        if (tile == QLD_TILE_SMALL)
            return r2_mode == QLD_R2_EXACT
                       ? launch_mode<FP4, QLD_R2_EXACT, GeoS, 2>(a, persistent, s)
                       : launch_mode<FP4, QLD_R2_REFERENCE, GeoS, 2>(a, persistent, s);
        return r2_mode == QLD_R2_EXACT
                   ? launch_mode<FP4, QLD_R2_EXACT, GeoL8H, 2>(a, persistent, s)
                   : launch_mode<FP4, QLD_R2_REFERENCE, GeoL8H, 2>(a, persistent, s);
    }
    if (packed) {
        if (engine != QLD_ENGINE_MFMA_FP4 || tile != QLD_TILE_SMALL) return hipErrorInvalidValue;
        return r2_mode == QLD_R2_EXACT
                   ? launch_mode<FP4, QLD_R2_EXACT, GeoS, 1>(a, persistent, s)
                   : launch_mode<FP4, QLD_R2_REFERENCE, GeoS, 1>(a, persistent, s);
    }
    if (engine == QLD_ENGINE_MFMA_FP4)
        return r2_mode == QLD_R2_EXACT ? launch_tile<FP4, QLD_R2_EXACT>(a, tile, persistent, s)
                                       : launch_tile<FP4, QLD_R2_REFERENCE>(a, tile, persistent, s);
    return r2_mode == QLD_R2_EXACT ? launch_tile<I8, QLD_R2_EXACT>(a, tile, persistent, s)
                                   : launch_tile<I8, QLD_R2_REFERENCE>(a, tile, persistent, s);
}
