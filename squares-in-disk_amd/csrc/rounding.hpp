// Directed-rounding binary64 arithmetic for gfx950.
//
// Every kernel of the prover core switches the wave's MODE register to
// round-toward-minus-infinity for f32 and f64 (enter_round_down) and computes
//   lower bounds as RD(x op y)      (one VALU instruction)
//   upper bounds as -RD((-x) op y)  (= RU(x op y); the negation is a free source modifier)
// exactly like the reference's SSE code (ivarp/include/ivarp/number/float_interval_addsub.hpp,
// float_interval_mul.hpp).  The add/mul/fma instructions are written as inline
// assembly so the compiler, which assumes round-to-nearest, can neither fold the
// negations away nor move them across the mode switch.
//
// Division and square root have no single correctly rounded instruction on
// CDNA.  They start from the compiler's sequence (accurate to a few ulp in any
// rounding mode) and are then corrected to the exact directed result with the
// exact residual a - q*b (resp. a - s*s) computed by one fma (Markstein /
// Tuckerman correction, Muller et al., Handbook of Floating-Point Arithmetic,
// ch. 4 and 5).  The residual is exact whenever q is within a few ulp of the
// quotient and no underflow occurs, so the sign test decides the rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqd {
namespace dev {

__device__ __forceinline__ void enter_round_down() {
    // MODE.FP_ROUND: bits [1:0] f32, bits [3:2] f64/f16; 2 = toward -infinity.
    __asm__ volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10" ::: "memory");
}

__device__ __forceinline__ double rd_add(double a, double b) {
    double r;
    __asm__ volatile("v_add_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// RU(a + b) = -RD(-a - b)
__device__ __forceinline__ double ru_add(double a, double b) {
    double r;
    __asm__ volatile("v_add_f64 %0, -%1, -%2" : "=v"(r) : "v"(a), "v"(b));
    return -r;
}
__device__ __forceinline__ double rd_sub(double a, double b) {
    double r;
    __asm__ volatile("v_add_f64 %0, %1, -%2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// RU(a - b) = -RD(-a + b)
__device__ __forceinline__ double ru_sub(double a, double b) {
    double r;
    __asm__ volatile("v_add_f64 %0, -%1, %2" : "=v"(r) : "v"(a), "v"(b));
    return -r;
}
__device__ __forceinline__ double rd_mul(double a, double b) {
    double r;
    __asm__ volatile("v_mul_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// RU(a * b) = -RD((-a) * b)
__device__ __forceinline__ double ru_mul(double a, double b) {
    double r;
    __asm__ volatile("v_mul_f64 %0, -%1, %2" : "=v"(r) : "v"(a), "v"(b));
    return -r;
}
// a - b*c, rounded down; exact whenever the true value is representable.
__device__ __forceinline__ double fms_rd(double a, double b, double c) {
    double r;
    __asm__ volatile("v_fma_f64 %0, -%1, %2, %3" : "=v"(r) : "v"(b), "v"(c), "v"(a));
    return r;
}

__device__ __forceinline__ std::uint64_t bits(double x) { return __builtin_bit_cast(std::uint64_t, x); }
__device__ __forceinline__ double from_bits(std::uint64_t u) { return __builtin_bit_cast(double, u); }
__device__ __forceinline__ bool is_finite(double x) { return (bits(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull; }
// next double toward -infinity / +infinity (finite x)
__device__ __forceinline__ double next_down(double x) {
    std::uint64_t u = bits(x);
    if((u << 1) == 0) return from_bits(0x8000000000000001ull);
    return from_bits((u >> 63) ? u + 1 : u - 1);
}
__device__ __forceinline__ double next_up(double x) {
    std::uint64_t u = bits(x);
    if((u << 1) == 0) return from_bits(0x0000000000000001ull);
    return from_bits((u >> 63) ? u - 1 : u + 1);
}

// Opaque initial approximations (kept out of constant folding).
__device__ __forceinline__ double approx_div(double a, double b) {
    __asm__ volatile("" : "+v"(a), "+v"(b));
    return a / b;
}
__device__ __forceinline__ double approx_sqrt(double a) {
    __asm__ volatile("" : "+v"(a));
    return __builtin_sqrt(a);
}

// RD(a / b) for normal operands whose quotient is normal (|a|, |b| in [0.5, 1)).
__device__ inline double rd_div_core(double a, double b) {
    double q = approx_div(a, b);
    const bool bneg = b < 0.0;
    // q too large while (a - q b)/b < 0
    #pragma unroll 1
    for(int i = 0; i < 4; ++i) {
        double r = fms_rd(a, q, b);
        if(r != 0.0 && ((r < 0.0) != bneg)) q = next_down(q); else break;
    }
    #pragma unroll 1
    for(int i = 0; i < 4; ++i) {
        double qn = next_up(q);
        double r = fms_rd(a, qn, b);
        if(r == 0.0 || ((r < 0.0) == bneg)) q = qn; else break;
    }
    return q;
}
// RD(a / b).  The operands are normalised to mantissas in [0.5, 1) so that the
// residual of the correction is exact even for subnormal inputs; the exponent is
// restored with ldexp, which rounds in the current (downward) mode when the
// result is subnormal or overflows.  The subnormal grid is a subset of the grid
// of the scaled quotient, so this second downward rounding equals one rounding.
__device__ inline double rd_div(double a, double b) {
    if(!is_finite(a) || !is_finite(b) || a == 0.0 || b == 0.0) return approx_div(a, b);  // exact IEEE specials
    int ea, eb;
    const double ma = __builtin_frexp(a, &ea);
    const double mb = __builtin_frexp(b, &eb);
    return __builtin_ldexp(rd_div_core(ma, mb), ea - eb);
}
// RU(a / b) = -RD(-a / b)
__device__ inline double ru_div(double a, double b) { return -rd_div(-a, b); }

// RD(sqrt(a)) for a in [0.5, 2).
__device__ inline double rd_sqrt_core(double a) {
    double s = approx_sqrt(a);
    #pragma unroll 1
    for(int i = 0; i < 4; ++i) {
        if(fms_rd(a, s, s) < 0.0) s = next_down(s); else break;
    }
    #pragma unroll 1
    for(int i = 0; i < 4; ++i) {
        double t = next_up(s);
        if(fms_rd(a, t, t) >= 0.0) s = t; else break;
    }
    return s;
}
// RD(sqrt(a)): a = m 2^(2k) with m in [0.5, 2), sqrt(a) = sqrt(m) 2^k exactly.
__device__ inline double rd_sqrt(double a) {
    if(!(a > 0.0) || !is_finite(a)) return approx_sqrt(a);  // 0, -0, +inf, negative, NaN
    int e;
    double m = __builtin_frexp(a, &e);
    if(e & 1) { m = __builtin_ldexp(m, 1); e -= 1; }
    return __builtin_ldexp(rd_sqrt_core(m), e / 2);
}
// RU(sqrt(a))
__device__ inline double ru_sqrt(double a) {
    if(!(a > 0.0) || !is_finite(a)) return approx_sqrt(a);
    int e;
    double m = __builtin_frexp(a, &e);
    if(e & 1) { m = __builtin_ldexp(m, 1); e -= 1; }
    double s = rd_sqrt_core(m);
    if(fms_rd(m, s, s) > 0.0) s = next_up(s);
    return __builtin_ldexp(s, e / 2);
}

} // namespace dev
} // namespace sqd
