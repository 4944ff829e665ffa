// Fused r^2 epilogue: pair count -> r^2, shared by every contraction engine.
//
// REFERENCE mode returns exactly what get_pairwise_ld_score (gemm_cpu.c:353-392
// of the reference; identical arithmetic in get_pairwise_ld_score_gpu,
// gemm_gpu.c:708-733) returns as its own build (gcc -O3 -std=gnu11
// -mfpmath=sse) compiles it:
//   p1 = a/N, p2 = b/N, p12 = c/N, x = p1*p2, t = p12 - x, num = t*t   (float)
//   den = ((double)x * (1 - p1)) * (1 - p2)                            (double)
//   r = (float)((double)num / den); r > 1.0001 -> 123.456; a or b == 0 -> 0
// p1*p2 is computed once in float and reused by the denominator, so
// p12 - p1*p2 is NOT fused into an FMA.  Device code is compiled with
// -ffp-contract=off so every float and double operation here rounds once,
// where the source says; the header also compiles as plain C++ so the same
// functions can be checked on the host.
//
// p12 = (float)c / (float)N is formed as (float)((double)c * (1.0/N)).  That
// equals the correctly rounded float quotient for c <= N < 2^24: c/N is never
// a float midpoint (that would need a 25-bit odd numerator, i.e. c >= 2^24)
// and is at least 2^-25/N (relative) away from every midpoint, while the
// double product is within 2^-52 of c/N.
//
// The quotient (float)(num / den) is not divided per pair.  With per-SNP
// terms A = 1/(p(1-p)) (double) and ip = 1/p (float):
//   x = p1 p2 - e, e = fma(p1, p2, -x) exactly, so
//   num / (x (1-p1)(1-p2)) = num A1 A2 (1 + d + d^2 + ...), d = e / (p1 p2),
// and Q = fma(P, dl, P), P = (num A1) A2, dl = e (ip1 ip2), is within
// 2^-45.5 (relative) of num / den: dl is within 2^-46 of d (|d| <= 2^-24,
// four float roundings), d^2 <= 2^-48, and P, den and the fma add nine
// double roundings.  That is at most 171 double ulps of Q, counting the
// half ulp of den's quotient.  Only float midpoints change what
// (float)Q rounds to, so when Q is more than 512 ulps from the nearest one
// (its low 29 significand bits, the bits float drops, more than 512 from
// 0x10000000) the correctly rounded double quotient rounds to the same
// float as Q.  Otherwise -- and for Q >= 1.0000896 (the clamp), infinities
// and NaNs (p = 1 gives A = inf) -- the pair takes the reference's own
// division.  Q is 0 or at least 2^-92, so float rounding never reaches the
// subnormal range, where midpoints sit elsewhere: a nonzero |t| below 2^-25
// needs p12 and x above 2^-25, whose ulps are multiples of 2^-48, so
// num >= 2^-96, and den <= 1/16.  A pair leaves the fast path with
// probability about 2^-19.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define QLD_EPI __host__ __device__ __forceinline__
#else
#define QLD_EPI inline
#endif

// Largest float r with (double)r <= 1.0001, so (double)r > 1.0001 <=> r > it.
#define QLD_R2_CLAMP_F 1.0000998973846435546875f
#define QLD_R2_FLAG_F 123.456f
// High word of the double 1 + 94/2^20 = 1.0000896: quotients at or above it
// (and infinities, NaNs, negative values) take the division.
#define QLD_R2_FAST_HI 0x3FF0005Eu
// (low word << 3) + this <= 2 * 8 * 512 <=> within 512 ulps of a float midpoint.
#define QLD_R2_MID_BIAS 0x80001000u
#define QLD_R2_MID_SPAN 0x2000u

// Per-SNP terms of the REFERENCE epilogue.
struct RefTerm {
    float p;      // (float)bc / (float)N
    float ip;     // 1 / p rounded to float; 0 when bc == 0
    double A;     // 1 / (p (1 - p)) rounded to double; 0 when bc == 0, inf when p == 1
};

QLD_EPI RefTerm ref_term(uint32_t bc, float fN)
{
    RefTerm t;
    t.p = (float)bc / fN;
    t.ip = bc == 0u ? 0.0f : 1.0f / t.p;
    t.A = bc == 0u ? 0.0 : 1.0 / ((double)t.p * (1.0 - (double)t.p));
    return t;
}

// The reference's operations one by one (p1 = column term, table A, val_1
// in the reference; p2 = row term, table B, val_2; the denominator
// multiplies (1 - val_1) before (1 - val_2)).
QLD_EPI float r2_reference_div(float c, float p1, float p2, double invN)
{
    const float p12 = (float)((double)c * invN);
    const float x = p1 * p2;
    const float t = p12 - x;
    const float num = t * t;
    const double den = ((double)x * (1.0 - (double)p1)) * (1.0 - (double)p2);
    float r = (float)((double)num / den);
    r = (r > QLD_R2_CLAMP_F) ? QLD_R2_FLAG_F : r;
    return (p1 == 0.0f || p2 == 0.0f) ? 0.0f : r;
}

// The same value without a division; slow is set when it may differ, and
// the caller then takes r2_reference_div.
QLD_EPI float r2_reference_fast(float c, const RefTerm &a, const RefTerm &b, double invN,
                                bool &slow)
{
    const float p12 = (float)((double)c * invN);
    const float x = a.p * b.p;
    const float t = p12 - x;
    const float num = t * t;
    const float e = __builtin_fmaf(a.p, b.p, -x);
    const float dl = e * (a.ip * b.ip);
    const double P = ((double)num * a.A) * b.A;
    const double Q = __builtin_fma(P, (double)dl, P);
    const uint64_t bits = __builtin_bit_cast(uint64_t, Q);
    const uint32_t lo = (uint32_t)bits, hi = (uint32_t)(bits >> 32);
    slow = (uint32_t)((lo << 3) + QLD_R2_MID_BIAS) <= QLD_R2_MID_SPAN || hi >= QLD_R2_FAST_HI;
    return (float)Q;
}

QLD_EPI float r2_reference(float c, const RefTerm &a, const RefTerm &b, double invN)
{
    bool slow;
    const float r = r2_reference_fast(c, a, b, invN, slow);
    return slow ? r2_reference_div(c, a.p, b.p, invN) : r;
}

// N >= 2^24 (the popcount engine): the reference's operations with p12
// formed as it forms it, the count rounded to float and divided by the
// rounded N (c/N may then be a float midpoint, and c and N inexact).
QLD_EPI float r2_reference_wide(uint32_t c, float p1, float p2, float fN)
{
    const float p12 = (float)c / fN;
    const float x = p1 * p2;
    const float t = p12 - x;
    const float num = t * t;
    const double den = ((double)x * (1.0 - (double)p1)) * (1.0 - (double)p2);
    float r = (float)((double)num / den);
    r = (r > QLD_R2_CLAMP_F) ? QLD_R2_FLAG_F : r;
    return (p1 == 0.0f || p2 == 0.0f) ? 0.0f : r;
}

// Per-SNP terms of the EXACT epilogue.
struct ExactTerm {
    double bc;    // bit count
    double rv;    // 1.0 / (bc * (N - bc)), correctly rounded (inf if 0)
    bool zero;
};

QLD_EPI ExactTerm exact_term(uint32_t bc, double dN)
{
    ExactTerm t;
    t.bc = (double)bc;
    t.rv = 1.0 / (t.bc * (dN - t.bc));
    t.zero = (bc == 0u);
    return t;
}

// d = c*N - a*b is exact in double (every term < 2^48);
// r^2 = ((d*d) * ra) * rb rounded once to float.
QLD_EPI float r2_exact(float c, const ExactTerm &a, const ExactTerm &b, double dN)
{
    const double d = __builtin_fma((double)c, dN, -(a.bc * b.bc));
    const double q = ((d * d) * a.rv) * b.rv;
    float r = (float)q;
    r = (r > QLD_R2_CLAMP_F) ? QLD_R2_FLAG_F : r;
    return (a.zero || b.zero) ? 0.0f : r;
}

// N >= 2^24: d = c*N - a*b in 64-bit integers (|d| < 2^62), rounded once to
// double; ra and rb are then within one double rounding of 1/(a(N-a)) and
// 1/(b(N-b)), so r^2 stays within one float ulp of the FP64 closed form.
QLD_EPI float r2_exact_wide(uint32_t c, const ExactTerm &a, const ExactTerm &b, int64_t N)
{
    const int64_t d = (int64_t)c * N - (int64_t)a.bc * (int64_t)b.bc;
    const double dd = (double)d;
    const double q = ((dd * dd) * a.rv) * b.rv;
    float r = (float)q;
    r = (r > QLD_R2_CLAMP_F) ? QLD_R2_FLAG_F : r;
    return (a.zero || b.zero) ? 0.0f : r;
}
