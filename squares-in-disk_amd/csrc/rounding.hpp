// Directed-rounding binary64 arithmetic for gfx950.
//
// Every kernel of the prover core switches the wave's MODE register to
// round-toward-minus-infinity for f32 and f64 (enter_round_down) and computes
//   lower bounds as RD(x op y)      (one VALU instruction)
//   upper bounds as -RD((-x) op y)  (= RU(x op y); the negation is a free source modifier)
// like the reference's SSE code (ivarp/include/ivarp/number/float_interval_addsub.hpp,
// float_interval_mul.hpp).  The mode switch is a volatile asm statement with a
// memory clobber at kernel entry, before any load: every floating-point value of
// the kernel derives from memory read after it.
//
// Division and square root have no single correctly rounded instruction on
// CDNA.  They start from an estimate accurate to a few ulp in any rounding mode
// (reciprocal and reciprocal-square-root estimates refined by Newton /
// Goldschmidt steps) and are then corrected to the exact directed
// result from the sign of the residual a - q*b (resp. a - s*s), computed by one
// fma (Markstein / Tuckerman correction, Muller et al., Handbook of
// Floating-Point Arithmetic, ch. 4 and 5).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqd {
namespace dev {

__device__ __forceinline__ void enter_round_down() {
    // MODE.FP_ROUND: bits [1:0] f32, bits [3:2] f64/f16; 2 = toward -infinity.
    __asm__ volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 10" ::: "memory");
}

// The operations below are plain IEEE operations evaluated in the wave's
// round-down mode.  The compiler assumes round-to-nearest, so they are written
// in forms that no transformation valid under round-to-nearest can change:
// moving a negation across an add, a sub or an fma changes the sign of an exact
// zero and is only done by the compiler with no-signed-zeros (never enabled),
// while a negation *is* moved across a plain multiply.  Products are therefore
// fma(a, b, +0.0), which equals RD(a * b) in round-down mode including the sign
// of zero (-0 + +0 = -0 when rounding down) and cannot be rewritten as a
// multiply without no-signed-zeros.  The negations end up as free source
// modifiers or sign-bit flips.
__device__ __forceinline__ double rd_add(double a, double b) { return a + b; }
// RU(a + b) = -RD(-a - b)
__device__ __forceinline__ double ru_add(double a, double b) { return -((-a) - b); }
__device__ __forceinline__ double rd_sub(double a, double b) { return a - b; }
// RU(a - b) = -RD(b - a)
__device__ __forceinline__ double ru_sub(double a, double b) { return -(b - a); }
__device__ __forceinline__ double rd_mul(double a, double b) { return __builtin_fma(a, b, 0.0); }
// RU(a * b) = -RD(-a * b)
__device__ __forceinline__ double ru_mul(double a, double b) { return -__builtin_fma(-a, b, 0.0); }
// a - b*c, rounded down; exact whenever the true value is representable.
__device__ __forceinline__ double fms_rd(double a, double b, double c) { return __builtin_fma(-b, c, a); }
// RU(a*b + c) = -RD(-a*b - c)
__device__ __forceinline__ double fma_ru(double a, double b, double c) { return -__builtin_fma(-a, b, -c); }

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

// IEEE a / b for the special operands (zeros, infinities, NaN).
__device__ __forceinline__ double special_div(double a, double b) {
    __asm__ volatile("" : "+v"(a), "+v"(b));
    return a / b;
}
// a / b for |a|, |b| in [2^-501, 2^501]: reciprocal estimate, two Newton steps,
// and a residual correction whose last rounding is upward, so that the result is
// RD(a / b) or the next double above it (up to the estimate's tiny error).
__device__ __forceinline__ double div_estimate(double a, double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return fma_ru(r, y, q);
}
// IEEE sqrt for the special inputs (zeros, negatives, infinities, NaN).
__device__ __forceinline__ double special_sqrt(double a) {
    __asm__ volatile("" : "+v"(a));
    return __builtin_sqrt(a);
}

// The directed results below are decided by the sign of an exact residual.
// For a candidate s of sqrt(m): fma gives RD(m - s*s); a nonzero residual keeps
// its sign under downward rounding as long as it does not underflow to zero,
// which the operand ranges guarantee (see the callers).  The candidate q of a / b
// is judged the same way from RD(a - q*b).

// Slow path: walk q to RD(a / b) (used when the approximation is more than one
// step off, which the fast path detects).
__device__ __noinline__ double rd_div_walk(double a, double b, double q) {
    const bool bneg = b < 0.0;
    for(int i = 0; i < 64; ++i) {  // q too large: (a - q b) has the sign opposite to b
        double r = fms_rd(a, q, b);
        if(r != 0.0 && ((r < 0.0) != bneg)) q = next_down(q); else break;
    }
    for(int i = 0; i < 64; ++i) {
        double qn = next_up(q);
        double r = fms_rd(a, qn, b);
        if(r == 0.0 || ((r < 0.0) == bneg)) q = qn; else break;
    }
    return q;
}
// RD(a / b) from a candidate q: one step in the direction the residual points,
// confirmed by the residual of the neighbour; otherwise the slow walk.
__device__ __forceinline__ double rd_div_fix(double a, double b, double q) {
    const bool bneg = b < 0.0;
    const double r0 = fms_rd(a, q, b);                 // sign(a - q b) = sign(b) * sign(x - q)
    const bool above = r0 != 0.0 && ((r0 < 0.0) != bneg);   // q > x
    const double s1 = above ? next_down(q) : next_up(q);
    const double r1 = fms_rd(a, s1, b);
    const bool s1_above = r1 != 0.0 && ((r1 < 0.0) != bneg);  // s1 > x
    // above: s1 < q, RD = s1 when s1 <= x.   below/equal: RD = q when s1 > x, s1 when s1 == x.
    const bool ok = above ? !s1_above : (s1_above || r1 == 0.0);
    const double res = above ? s1 : (r1 == 0.0 ? s1 : q);
    if(__builtin_expect(!ok, 0)) return rd_div_walk(a, b, q);
    return res;
}
// RD(a / b).  Operands with magnitudes in [2^-500, 2^500] keep the quotient and
// the residual far from underflow and overflow; others are first normalised to
// mantissas in [0.5, 1) with frexp and the exponent restored with ldexp, which
// rounds downward in the current mode when the result is subnormal (the
// subnormal grid is a subset of the normalised quotient's grid, so that second
// downward rounding equals a single one).
__device__ inline double rd_div(double a, double b) {
    if(!is_finite(a) || !is_finite(b) || a == 0.0 || b == 0.0) return special_div(a, b);  // exact IEEE specials
    const unsigned ea = static_cast<unsigned>(bits(a) >> 52) & 0x7ffu;
    const unsigned eb = static_cast<unsigned>(bits(b) >> 52) & 0x7ffu;
    if(ea - 523u <= 1000u && eb - 523u <= 1000u)  // biased exponents in [523, 1523]
        return rd_div_fix(a, b, div_estimate(a, b));
    int xa, xb;
    const double ma = __builtin_frexp(a, &xa);
    const double mb = __builtin_frexp(b, &xb);
    return __builtin_ldexp(rd_div_fix(ma, mb, div_estimate(ma, mb)), xa - xb);
}
// RU(a / b) = -RD(-a / b)
__device__ inline double ru_div(double a, double b) { return -rd_div(-a, b); }

// sqrt(m) for m in [2^-1000, 2^1024): reciprocal square root estimate, one
// Goldschmidt step and two residual corrections, the last rounded upward, so
// that the result is RD(sqrt(m)) or the next double above it.
__device__ __forceinline__ double sqrt_estimate(double m) {
    const double y = __builtin_amdgcn_rsq(m);
    double g = m * y;
    double h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, m);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, m);
    return fma_ru(d, h, g);
}
__device__ __noinline__ double rd_sqrt_walk(double m, double s) {
    for(int i = 0; i < 64; ++i) {
        if(fms_rd(m, s, s) < 0.0) s = next_down(s); else break;
    }
    for(int i = 0; i < 64; ++i) {
        const double t = next_up(s);
        if(fms_rd(m, t, t) >= 0.0) s = t; else break;
    }
    return s;
}
// RD(sqrt(m)) from a positive candidate g (as rd_div_fix).
__device__ __forceinline__ double rd_sqrt_fix(double m, double g) {
    const double r0 = fms_rd(m, g, g);
    const bool above = r0 < 0.0;                        // g > sqrt(m)
    const double s1 = from_bits(bits(g) + (above ? ~0ull : 1ull));   // g > 0: neighbour below / above
    const double r1 = fms_rd(m, s1, s1);
    const bool ok = above ? (r1 >= 0.0) : (r1 <= 0.0);
    const double res = above ? s1 : (r1 == 0.0 ? s1 : g);
    if(__builtin_expect(!ok, 0)) return rd_sqrt_walk(m, g);
    return res;
}
// Positive finite m: inputs below 2^-900 are scaled by 2^1000 (an even power,
// so the root scales exactly by 2^500) to keep the residual from underflowing.
__device__ __forceinline__ double scale_for_sqrt(double a, bool& tiny) {
    tiny = a < 0x1p-900;
    return tiny ? __builtin_ldexp(a, 1000) : a;
}
// RD(sqrt(a))
__device__ inline double rd_sqrt(double a) {
    if(!(a > 0.0) || !is_finite(a)) return special_sqrt(a);  // 0, -0, +inf, negative, NaN
    bool tiny;
    const double m = scale_for_sqrt(a, tiny);
    const double s = rd_sqrt_fix(m, sqrt_estimate(m));
    return tiny ? __builtin_ldexp(s, -500) : s;
}
// RU(sqrt(a)): RD result, one step up unless the root is exact.
__device__ inline double ru_sqrt(double a) {
    if(!(a > 0.0) || !is_finite(a)) return special_sqrt(a);
    bool tiny;
    const double m = scale_for_sqrt(a, tiny);
    double s = rd_sqrt_fix(m, sqrt_estimate(m));
    if(fms_rd(m, s, s) != 0.0) s = from_bits(bits(s) + 1ull);
    return tiny ? __builtin_ldexp(s, -500) : s;
}

// ---------------------------------------------------------------- binary32
// The same operations on floats (the FP32 prefilter).  The MODE write above sets
// round-down for f32 too; the kernels keep f32 denormals (IEEE mode), so
// directed results near zero are exact.
__device__ __forceinline__ float rd_add(float a, float b) { return a + b; }
__device__ __forceinline__ float ru_add(float a, float b) { return -((-a) - b); }
__device__ __forceinline__ float rd_sub(float a, float b) { return a - b; }
__device__ __forceinline__ float ru_sub(float a, float b) { return -(b - a); }
__device__ __forceinline__ float rd_mul(float a, float b) { return __builtin_fmaf(a, b, 0.0f); }
__device__ __forceinline__ float ru_mul(float a, float b) { return -__builtin_fmaf(-a, b, 0.0f); }
__device__ __forceinline__ float fms_rd(float a, float b, float c) { return __builtin_fmaf(-b, c, a); }

__device__ __forceinline__ std::uint32_t bits(float x) { return __builtin_bit_cast(std::uint32_t, x); }
__device__ __forceinline__ float from_bits32(std::uint32_t u) { return __builtin_bit_cast(float, u); }
__device__ __forceinline__ bool is_finite(float x) { return (bits(x) & 0x7f800000u) != 0x7f800000u; }
__device__ __forceinline__ float next_down(float x) {
    const std::uint32_t u = bits(x);
    if((u << 1) == 0) return from_bits32(0x80000001u);
    if(u == 0x7f800000u) return from_bits32(0x7f7fffffu);     // +inf -> largest float
    return from_bits32((u >> 31) ? u + 1 : u - 1);
}
__device__ __forceinline__ float next_up(float x) {
    const std::uint32_t u = bits(x);
    if((u << 1) == 0) return from_bits32(0x00000001u);
    if(u == 0xff800000u) return from_bits32(0xff7fffffu);     // -inf -> most negative float
    return from_bits32((u >> 31) ? u - 1 : u + 1);
}

// Outward conversion of a double: the largest float <= d / the smallest float
// >= d (NaN stays NaN).  The hardware conversion rounds in the current mode;
// the comparison with the exact widening makes the result independent of it.
__device__ __forceinline__ float f32_down(double d) {
    float f = static_cast<float>(d);
    if(static_cast<double>(f) > d) f = next_down(f);
    return f;
}
__device__ __forceinline__ float f32_up(double d) {
    float f = static_cast<float>(d);
    if(static_cast<double>(f) < d) f = next_up(f);
    return f;
}

// Division: a float quotient rounded down equals the binary64 quotient rounded
// down and then to float (every float is a double, so the largest float below
// the exact quotient is also the largest float below its binary64 round-down).
// Divisions are rare in the programs; this reuses the binary64 algorithm.
__device__ inline float rd_div(float a, float b) {
    return f32_down(rd_div(static_cast<double>(a), static_cast<double>(b)));
}
__device__ inline float ru_div(float a, float b) {
    return f32_up(ru_div(static_cast<double>(a), static_cast<double>(b)));
}

// Square root of a float in [2^-100, 2^128): the hardware estimate
// (v_sqrt_f32, a few ulp at worst in any rounding mode) corrected by the sign of
// the exact residual m - s*s.  For m >= 2^-100 the root is >= 2^-50 and a
// nonzero residual is at least 2^-96 in magnitude, so it never rounds to zero.
__device__ __noinline__ float rd_sqrt_walk(float m, float s) {
    for(int i = 0; i < 64; ++i) {
        if(fms_rd(m, s, s) < 0.0f) s = next_down(s); else break;
    }
    for(int i = 0; i < 64; ++i) {
        const float t = next_up(s);
        if(fms_rd(m, t, t) >= 0.0f) s = t; else break;
    }
    return s;
}
__device__ __forceinline__ float rd_sqrt_fix(float m, float g) {
    const float r0 = fms_rd(m, g, g);
    const bool above = r0 < 0.0f;
    const float s1 = from_bits32(bits(g) + (above ? ~0u : 1u));
    const float r1 = fms_rd(m, s1, s1);
    const bool ok = above ? (r1 >= 0.0f) : (r1 <= 0.0f);
    const float res = above ? s1 : (r1 == 0.0f ? s1 : g);
    if(__builtin_expect(!ok, 0)) return rd_sqrt_walk(m, g);
    return res;
}
__device__ __forceinline__ float special_sqrt(float a) {
    __asm__ volatile("" : "+v"(a));
    return __builtin_sqrtf(a);
}
// Positive finite inputs below 2^-100 are scaled by 2^100 (the root by 2^50).
__device__ __forceinline__ float scale_for_sqrt(float a, bool& tiny) {
    tiny = a < 0x1p-100f;
    return tiny ? a * 0x1p100f : a;   // exact: a power of two, no overflow
}
__device__ inline float rd_sqrt(float a) {
    if(!(a > 0.0f) || !is_finite(a)) return special_sqrt(a);
    bool tiny;
    const float m = scale_for_sqrt(a, tiny);
    const float s = rd_sqrt_fix(m, __builtin_amdgcn_sqrtf(m));
    return tiny ? __builtin_ldexpf(s, -50) : s;
}
__device__ inline float ru_sqrt(float a) {
    if(!(a > 0.0f) || !is_finite(a)) return special_sqrt(a);
    bool tiny;
    const float m = scale_for_sqrt(a, tiny);
    float s = rd_sqrt_fix(m, __builtin_amdgcn_sqrtf(m));
    if(fms_rd(m, s, s) != 0.0f) s = from_bits32(bits(s) + 1u);
    return tiny ? __builtin_ldexpf(s, -50) : s;
}

} // namespace dev
} // namespace sqd
