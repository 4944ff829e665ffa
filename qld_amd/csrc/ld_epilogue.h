// Fused r^2 epilogue: pair count -> r^2, shared by every contraction engine.
//
// REFERENCE mode restates get_pairwise_ld_score (gemm_cpu.c:353-392 of the
// reference; identical arithmetic in get_pairwise_ld_score_gpu,
// gemm_gpu.c:708-733) operation for operation, as its own build
// (gcc -O3 -std=gnu11 -mfpmath=sse) compiles it: p1*p2 is computed once in
// float and reused by the denominator, so p12 - p1*p2 is NOT fused into an
// FMA.  This file is compiled with -ffp-contract=off so the device code
// rounds at the same points.
//
// p12 = (float)c / (float)N is formed as (float)((double)c * (1.0/N)).  That
// equals the correctly rounded float quotient for c <= N < 2^24: c/N is never
// a float midpoint (that would need a 25-bit odd numerator, i.e. c >= 2^24)
// and is at least 2^-25/N (relative) away from every midpoint, while the
// double product is within 2^-52 of c/N.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Largest float r with (double)r <= 1.0001, so (double)r > 1.0001 <=> r > it.
#define QLD_R2_CLAMP_F 1.0000998973846435546875f
#define QLD_R2_FLAG_F 123.456f

// Per-SNP terms of the REFERENCE epilogue.
struct RefTerm {
    float p;      // (float)bc / (float)N
    double omp;   // 1.0 - (double)p  (exact in double)
    bool zero;    // bc == 0 -> r^2 forced to 0
};

__device__ __forceinline__ RefTerm ref_term(uint32_t bc, float fN)
{
    RefTerm t;
    t.p = __fdiv_rn((float)bc, fN);
    t.omp = 1.0 - (double)t.p;
    t.zero = (bc == 0u);
    return t;
}

// a = column term (table A, val_1 in the reference), b = row term (table B,
// val_2); the denominator multiplies (1 - val_1) before (1 - val_2).
__device__ __forceinline__ float r2_reference(float c, const RefTerm &a,
                                              const RefTerm &b, double invN)
{
    float p12 = (float)((double)c * invN);
    float x = __fmul_rn(a.p, b.p);
    float t = __fsub_rn(p12, x);
    float num = __fmul_rn(t, t);
    double den = __dmul_rn(__dmul_rn((double)x, a.omp), b.omp);
    float r = (float)__ddiv_rn((double)num, den);
    r = (r > QLD_R2_CLAMP_F) ? QLD_R2_FLAG_F : r;
    return (a.zero || b.zero) ? 0.0f : r;
}

// Per-SNP terms of the EXACT epilogue.
struct ExactTerm {
    double bc;    // bit count
    double rv;    // 1.0 / (bc * (N - bc)), correctly rounded (inf if 0)
    bool zero;
};

__device__ __forceinline__ ExactTerm exact_term(uint32_t bc, double dN)
{
    ExactTerm t;
    t.bc = (double)bc;
    t.rv = __ddiv_rn(1.0, t.bc * (dN - t.bc));
    t.zero = (bc == 0u);
    return t;
}

// d = c*N - a*b is exact in double (every term < 2^48);
// r^2 = ((d*d) * ra) * rb rounded once to float.
__device__ __forceinline__ float r2_exact(float c, const ExactTerm &a,
                                          const ExactTerm &b, double dN)
{
    double d = __fma_rn((double)c, dN, -(a.bc * b.bc));
    double q = __dmul_rn(__dmul_rn(__dmul_rn(d, d), a.rv), b.rv);
    float r = (float)q;
    r = (r > QLD_R2_CLAMP_F) ? QLD_R2_FLAG_F : r;
    return (a.zero || b.zero) ? 0.0f : r;
}
