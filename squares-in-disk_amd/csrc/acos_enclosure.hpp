// An enclosure of acos(t) in binary64 interval arithmetic, for the binary32
// evaluation of the FP32 prefilter.
//
// The binary64 evaluation needs acos rounded correctly in each direction, as the
// reference has it from MPFR (acos_fixed.hpp); the binary32 evaluation only needs
// floats below and above acos(t), which this gives at a small fraction of the
// cost: the arcsine series evaluated on intervals with every operation rounded
// outward (the kernels run in round-down mode; upper bounds are -RD(-x)), plus a
// bound on the series' tail:
//   |t| <= 1/2:  acos(t) = pi/2 - asin(t),         asin(a) = a P(a^2)
//   |t| >  1/2:  acos(|t|) = 2 asin(sqrt(u)) = 2 sqrt(u) P(u),  u = (1 - |t|) / 2
//                (1 - |t| and the halving are exact), acos(t) = pi - acos(|t|) for t < 0
//   P(z) = sum_k c_k z^k,  c_k = binom(2k, k) / (4^k (2k + 1)),  0 <= z <= 1/4.
// The c_k are positive and decrease, so the first K terms are below P(z) and
// the rest are at most c_K z^K / (1 - z) <= c_K 4^-K (4/3), about 2^-39.5 for K =
// 16: the enclosure is about that wide relative to its value, and converting
// it outward to floats gives the correctly rounded floats except in about one
// case in 2^16 (then one float wider).  Every operation here is an IEEE
// operation with directed rounding, so the CPU restatement reproduces the bits.
#pragma once
#include "rounding.hpp"

namespace sqd {
namespace dev {
namespace acosenc {

constexpr int kTerms = 16;

// c_k rounded down (up = false) or up, exactly: floor / ceil of c_k 2^s with
// 2^52 <= c_k 2^s < 2^53, scaled back by a power of two.
__host__ __device__ constexpr double series_coeff(int k, bool up) {
    unsigned __int128 num = 1, den = 1;
    for(int i = 1; i <= k; ++i) num = num * static_cast<unsigned __int128>(k + i) / static_cast<unsigned __int128>(i);
    for(int i = 0; i < k; ++i) den *= 4;
    den *= static_cast<unsigned __int128>(2 * k + 1);
    int s = 0;
    while((num << s) < (den << 52)) ++s;             // c_k 2^s >= 2^52
    const unsigned __int128 sc = num << s;
    unsigned long long q = static_cast<unsigned long long>(sc / den);
    if(up && sc % den != 0) ++q;
    double r = static_cast<double>(q);               // q <= 2^53: exact
    for(int i = 0; i < s; ++i) r *= 0.5;             // exact (no underflow)
    return r;
}
// Upper bound of the tail sum_{k >= kTerms} c_k z^k for z <= 1/4, which is at
// most c_K 4^-K (4/3): 2 c_K 4^-K with c_K rounded up (the scalings are exact).
__host__ __device__ constexpr double tail_bound() {
    double r = series_coeff(kTerms, true);
    for(int i = 0; i < kTerms; ++i) r *= 0.25;
    return r * 2.0;
}

struct Coeffs {
    double lo[kTerms], hi[kTerms];
};
__host__ __device__ constexpr Coeffs coeffs() {
    Coeffs c{};
    for(int k = 0; k < kTerms; ++k) { c.lo[k] = series_coeff(k, false); c.hi[k] = series_coeff(k, true); }
    return c;
}

// [lo, hi] of P(z) for 0 <= zl <= z <= zh <= 1/4 (Horner, lower bounds rounded
// down and upper bounds up; all terms are non-negative).
__device__ __forceinline__ void series(double zl, double zh, double& pl, double& ph) {
    constexpr Coeffs C = coeffs();
    constexpr double tail = tail_bound();
    pl = C.lo[kTerms - 1];
    ph = C.hi[kTerms - 1];
#pragma unroll
    for(int k = kTerms - 2; k >= 0; --k) {
        pl = __builtin_fma(zl, pl, C.lo[k]);          // RD(zl pl + c)
        ph = fma_ru(zh, ph, C.hi[k]);                 // RU(zh ph + c)
    }
    ph = ru_add(ph, tail);
}

// Lower and upper bound of acos(t) for t in [-1, 1]; NaN outside.  One series
// evaluation serves both ranges (z = t^2 or u), so that lanes of a wave on
// either side of 1/2 share it.
__device__ __forceinline__ void acos_enclose(double t, double& lo, double& hi) {
    constexpr double kHalfPiLo = 0x1.921fb54442d18p+0, kHalfPiHi = 0x1.921fb54442d19p+0;
    constexpr double kPiLo = 0x1.921fb54442d18p+1, kPiHi = 0x1.921fb54442d19p+1;
    if(!(t >= -1.0 && t <= 1.0)) { lo = hi = __builtin_nan(""); return; }
    const double a = t < 0.0 ? -t : t;
    const bool small = a <= 0.5;
    // u = (1 - a) / 2, exact, written -(a - 1) / 2 so that it is +0 for a = 1
    // (x - x is -0 when rounding down; a select on a == 1 would be folded away
    // by a compiler that assumes round-to-nearest, where 1 - 1 is +0)
    const double u = -(a - 1.0) * 0.5;
    double pl, ph;
    series(small ? rd_mul(a, a) : u, small ? ru_mul(a, a) : u, pl, ph);
    double ml = a, mh = a;                                        // asin(a) = a P(a^2)
    if(!small) {                                                  // acos(a) = 2 sqrt(u) P(u) (doubling exact)
        ml = 2.0 * rd_sqrt(u);
        mh = 2.0 * ru_sqrt(u);
    }
    const double rl = rd_mul(ml, pl), rh = ru_mul(mh, ph);
    if(small) {
        lo = t >= 0.0 ? rd_sub(kHalfPiLo, rh) : rd_add(kHalfPiLo, rl);
        hi = t >= 0.0 ? ru_sub(kHalfPiHi, rl) : ru_add(kHalfPiHi, rh);
    } else {
        lo = t >= 0.0 ? rl : rd_sub(kPiLo, rh);
        hi = t >= 0.0 ? rh : ru_sub(kPiHi, rl);
    }
}

} // namespace acosenc
} // namespace dev
} // namespace sqd
