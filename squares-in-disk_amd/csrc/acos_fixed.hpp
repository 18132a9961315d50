// Directed-rounding arc cosine for binary64, in exact 128-bit fixed-point
// integer arithmetic (independent of the floating-point rounding mode).
//
// Fast path (acos_directed): RD(acos(x)) is the largest double y with
// cos(y) >= x; starting from an approximation, neighbouring doubles are tested
// with a fixed-point cos (see cos_ge).  Fallback (acos_series), below: acos(x)
// itself in fixed point.
//
// The reference rounds acos correctly in the required direction with MPFR
// (ivarp/src/interval_acos.cpp:29-66): acos([l, u]) = [RD(acos(u)), RU(acos(l))].
// Here acos(x) is evaluated as an integer Y = acos(x) * 2^124 with a proven error
// bound, and the directed double results are read off the interval
// [Y - E, Y + E]:
//   |x| <= 1/2:  acos(x) = pi/2 - asin(x)
//   |x| >  1/2:  acos(|x|) = 2 asin(sqrt(u)) = 2 sqrt(u) sum_k c_k u^k,  u = (1 - |x|) / 2
//                (exact integer square root of the normalised u)
//   x < 0:       acos(x) = pi - acos(-x)
//   asin(s) = sum_k c_k s^(2k+1),  c_{k+1} = c_k (2k+1)^2 / ((2k+2)(2k+3)),  s <= 1/2.
// Every truncation rounds toward zero, so a computed series is below the true one
// by at most 7 units of 2^-124 per term (64 terms) plus a geometric tail of less
// than 10 units, i.e. under 512 units; E = 2048 units bounds every branch (see
// acos_directed for the scales).
//
// When [Y - E, Y + E] contains a double the direction cannot be decided at this
// precision; the result is then widened by that double (still an enclosure of
// acos(x)) and the caller's hard-case flag is set.  For |x| < 1 acos(x) is
// irrational, so this is the only way the result can differ from the correctly
// rounded one.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqd {
namespace dev {
namespace acosfx {

using u128 = unsigned __int128;
constexpr int kF = 124;                  // fractional bits
constexpr u128 kUnit = 1;
constexpr u128 kErr = 2048;              // error bound E in units of 2^-124

// pi * 2^124 and pi/2 * 2^124, rounded down (Machin's formula in integer
// arithmetic; the hex expansion of pi is 3.243F6A8885A308D313198A2E0370734...)
__host__ __device__ constexpr u128 pi_fx() { return (u128(0x3243f6a8885a308dull) << 64) | 0x313198a2e0370734ull; }
__host__ __device__ constexpr u128 half_pi_fx() { return (u128(0x1921fb54442d1846ull) << 64) | 0x9898cc51701b839aull; }

// floor(a * b / 2^124) for a, b < 2^127 with a * b < 2^252.
__host__ __device__ inline u128 mulq(u128 a, u128 b) {
    const std::uint64_t a0 = static_cast<std::uint64_t>(a), a1 = static_cast<std::uint64_t>(a >> 64);
    const std::uint64_t b0 = static_cast<std::uint64_t>(b), b1 = static_cast<std::uint64_t>(b >> 64);
    const u128 p00 = u128(a0) * b0, p01 = u128(a0) * b1, p10 = u128(a1) * b0, p11 = u128(a1) * b1;
    const u128 mid = (p00 >> 64) + static_cast<std::uint64_t>(p01) + static_cast<std::uint64_t>(p10);
    const u128 hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
    return (hi << (128 - kF)) | (static_cast<std::uint64_t>(mid) >> (kF - 64));
}

// floor((2k+1)^2 / ((2k+2)(2k+3)) * 2^124), by binary long division.
__host__ __device__ constexpr u128 series_ratio(int k) {
    const std::uint64_t num = static_cast<std::uint64_t>(2 * k + 1) * static_cast<std::uint64_t>(2 * k + 1);
    const std::uint64_t den = static_cast<std::uint64_t>(2 * k + 2) * static_cast<std::uint64_t>(2 * k + 3);
    std::uint64_t rem = num;   // num < den
    u128 q = 0;
    for(int i = 0; i < kF; ++i) {
        rem <<= 1;
        q <<= 1;
        if(rem >= den) { rem -= den; q |= 1; }
    }
    return q;
}

// asin(S / 2^124) * 2^124 rounded down, for S <= 2^123 (s <= 1/2).
__host__ __device__ inline u128 asin_fx(u128 s) {
    const u128 t2 = mulq(s, s);
    u128 sum = s, term = s;
    for(int k = 0; k < 64 && term != 0; ++k) {
        term = mulq(mulq(term, t2), series_ratio(k));
        sum += term;
    }
    return sum;
}

// sum_k c_k u^k * 2^124 rounded down, for U = u * 2^124 <= 2^122 (u <= 1/4):
// asin(s) = s * series(s^2).
__host__ __device__ inline u128 asin_ratio_fx(u128 u) {
    u128 sum = kUnit << kF, term = kUnit << kF;
    for(int k = 0; k < 64 && term != 0; ++k) {
        term = mulq(mulq(term, u), series_ratio(k));
        sum += term;
    }
    return sum;
}

// floor(sqrt(U * 2^124)) for U < 2^122, digit by digit.
__host__ __device__ inline u128 sqrt_fx(u128 u) {
    u128 res = 0, rem = 0;
    for(int i = 122; i >= 0; --i) {
        const int j = 2 * i;
        const unsigned two = j >= kF ? static_cast<unsigned>((u >> (j - kF)) & 3u) : 0u;
        rem = (rem << 2) | two;
        const u128 trial = (res << 2) | 1u;
        if(rem >= trial) { rem -= trial; res = (res << 1) | 1u; }
        else res <<= 1;
    }
    return res;
}

__host__ __device__ inline int msb128(u128 v) {
    const std::uint64_t h = static_cast<std::uint64_t>(v >> 64);
    return h ? 127 - __builtin_clzll(h) : 63 - __builtin_clzll(static_cast<std::uint64_t>(v));
}
__host__ __device__ inline double from_bits_h(std::uint64_t u) { return __builtin_bit_cast(double, u); }
__host__ __device__ inline std::uint64_t to_bits_h(double d) { return __builtin_bit_cast(std::uint64_t, d); }

// Largest double <= V / 2^(124 + z) (V > 0); *inexact reports dropped bits.
__host__ __device__ inline double fx_to_double_down(u128 v, int z, bool* inexact) {
    const int p = msb128(v);
    std::uint64_t mant;
    bool lost = false;
    if(p >= 52) {
        mant = static_cast<std::uint64_t>(v >> (p - 52));
        lost = (v & ((kUnit << (p - 52)) - 1)) != 0;
    } else {
        mant = static_cast<std::uint64_t>(v) << (52 - p);
    }
    // value = mant * 2^(p - 52 - 124 - z), mant in [2^52, 2^53): biased exponent
    // p - 176 - z + 1075 (the values here are >= 2^-200, normal doubles)
    const std::uint64_t e = static_cast<std::uint64_t>(p - 52 - kF - z + 1075);
    *inexact = lost;
    return from_bits_h((e << 52) | (mant & 0x000fffffffffffffull));
}
__host__ __device__ inline double fx_down(u128 v, int z) {
    bool ix;
    return v == 0 ? 0.0 : fx_to_double_down(v, z, &ix);
}
__host__ __device__ inline double fx_up(u128 v, int z) {
    if(v == 0) return 0.0;
    bool ix;
    const double d = fx_to_double_down(v, z, &ix);
    return ix ? from_bits_h(to_bits_h(d) + 1) : d;
}

// ---------------------------------------------------------------- cos test
// cos is strictly decreasing on [0, pi], so RD(acos(x)) is the largest double
// y <= pi with cos(y) >= x.  cos(y) of a double y in [0, pi] is evaluated here
// in the same fixed point: r = |y - pi/2| (exact but for the rounding of pi/2,
// < 1 unit), cos(y) = -+sin(r), and sin(r) = sin(a) for a = r <= pi/4 or
// cos(pi/2 - r) otherwise, so that every series argument is at most pi/4 and
// z = argument^2 < 0.62.  The alternating series are summed from the inside,
//   sin(a) = a (c0 - z (c1 - z (c2 - ...))),  c_k = 1/(2k+1)!
//   cos(b) =    d0 - z (d1 - z (d2 - ...)),   d_k = 1/(2k)!,
// where every partial sum lies in (0, c_k] (the terms decrease for z < 6), so
// unsigned arithmetic suffices.  17 terms leave a tail below 2^-129; each step
// truncates by less than 1 unit and the coefficients are floor(2^124 / n!)
// exactly, and earlier errors shrink by z < 1, so the result is within
// kCosErr units of the true value.
constexpr u128 kCosErr = 64;
constexpr int kCosTerms = 17;

// floor(2^124 / n!) (nested floor divisions of an integer equal one floor division)
__host__ __device__ constexpr u128 inv_factorial(int n) {
    u128 v = kUnit << kF;
    for(int i = 2; i <= n; ++i) v /= static_cast<u128>(i);
    return v;
}
struct TrigCoeffs {
    u128 s[kCosTerms];   // floor(2^124 / (2k+1)!)
    u128 c[kCosTerms];   // floor(2^124 / (2k)!)
};
__host__ __device__ constexpr TrigCoeffs trig_coeffs() {
    TrigCoeffs t{};
    for(int k = 0; k < kCosTerms; ++k) {
        t.s[k] = inv_factorial(2 * k + 1);
        t.c[k] = inv_factorial(2 * k);
    }
    return t;
}

// sin(A / 2^124) * 2^124 for A <= pi/4 * 2^124, within kCosErr units
__host__ __device__ inline u128 sin_fx(u128 a) {
    constexpr TrigCoeffs T = trig_coeffs();
    const u128 z = mulq(a, a);
    u128 s = T.s[kCosTerms - 1];
    for(int k = kCosTerms - 2; k >= 0; --k) s = T.s[k] - mulq(z, s);
    return mulq(a, s);
}
// cos(B / 2^124) * 2^124 for B <= pi/4 * 2^124, within kCosErr units
__host__ __device__ inline u128 cos_fx(u128 b) {
    constexpr TrigCoeffs T = trig_coeffs();
    const u128 z = mulq(b, b);
    u128 s = T.c[kCosTerms - 1];
    for(int k = kCosTerms - 2; k >= 0; --k) s = T.c[k] - mulq(z, s);
    return s;
}

using i128 = __int128;

// x as a signed fixed-point value, truncated toward zero (exact for |x| >= 2^-71)
__host__ __device__ inline i128 to_fx_signed(double x) {
    const std::uint64_t b = to_bits_h(x);
    const int ex = static_cast<int>((b >> 52) & 0x7ff);
    const std::uint64_t m = ex ? ((b & 0x000fffffffffffffull) | 0x0010000000000000ull) : (b & 0x000fffffffffffffull);
    const int sh = (ex ? ex : 1) - 1075 + kF;
    const u128 X = sh >= 0 ? (u128(m) << sh) : (sh > -64 ? u128(m >> (-sh)) : u128(0));
    return (b >> 63) ? -static_cast<i128>(X) : static_cast<i128>(X);
}

// 1 if cos(y) >= x, 0 if cos(y) < x, -1 if undecided at this precision
// (y a double in [0, pi] with y = 0 or y >= 2^-71, x in [-1, 1]).
__host__ __device__ inline int cos_ge(double y, double x) {
    const u128 Y = static_cast<u128>(to_fx_signed(y));
    const u128 H = half_pi_fx();
    const bool neg = Y >= H;                        // y >= pi/2 (up to the rounding of pi/2)
    const u128 r = neg ? Y - H : H - Y;
    const u128 Q = H >> 1;                          // pi/4
    u128 s;
    if(r <= Q) s = sin_fx(r);
    else s = cos_fx(r <= H ? H - r : r - H);        // cos is even: |pi/2 - r|
    const i128 c = neg ? -static_cast<i128>(s) : static_cast<i128>(s);
    const i128 X = to_fx_signed(x);
    const i128 e = static_cast<i128>(kCosErr + 2);  // series, pi/2 rounding, truncation of x
    if(c - e > X) return 1;
    if(c + e < X) return 0;
    return -1;
}

// A cheaper first try of cos_ge, decisive in all but rare cases: the same
// fixed point with two savings.
//  * mulq_low drops the limb products below 2^96 (with 32-bit limbs a_i, b_j,
//    those with i + j <= 2): they add less than 2^64 (1 + 2 2^32 + 3 2^64) <
//    3.01 2^128 before the shift by 124, so the result is at most 49 units
//    below floor(a b / 2^124) and never above it.  10 of the 16 limb products.
//  * 12 terms of each series: the alternating tails are at most
//    z^12 / 24! < 2^36.6 units (cos) and a z^12 / 25! < 2^31.6 units (sin) for
//    z < 0.617, a < 0.786.
// Rounding errors (units of 2^-124): a product is less than 50 below the exact
// one (49 dropped, 1 floor) and z less than 50 below z; each Horner step adds
// less than 1 (coefficient) + 50 (product) + 50 (z's error times a partial sum
// <= 1) and scales the earlier error by z < 0.617, so the sum is within
// 101 / 0.383 < 264 units, and sin's final product keeps it below 50 + 0.786 x
// 264 < 264.  With the tails, pi/2 and x (under 3 units) the evaluation is
// within 2^37 units; kCosErrFast allows twice that.  The cosines of neighbouring
// doubles y differ by about 2^-53 sin y, so away from y = 0 and pi the test is
// undecided in about one comparison in 2^32 (cos_ge then decides); near them
// (x near +-1) cos_ge decides more often.
constexpr u128 kCosErrFast = kUnit << 38;
constexpr int kCosTermsFast = 12;

// floor(a b / 2^124) minus 0 to 49, for a, b < 2^125 (limb products of weight
// >= 2^96 only; the sum of weight 2^128 and above fits 128 bits).
__host__ __device__ inline u128 mulq_low(u128 a, u128 b) {
    std::uint32_t x[4], w[4];
    for(int i = 0; i < 4; ++i) {
        x[i] = static_cast<std::uint32_t>(a >> (32 * i));
        w[i] = static_cast<std::uint32_t>(b >> (32 * i));
    }
    auto p = [&](int i, int j) { return u128(static_cast<std::uint64_t>(x[i]) * w[j]); };
    const u128 lo = p(0, 3) + p(1, 2) + p(2, 1) + p(3, 0);                       // weight 2^96
    const u128 hi = p(1, 3) + p(2, 2) + p(3, 1) + ((p(2, 3) + p(3, 2)) << 32) + (p(3, 3) << 64);   // weight 2^128
    return (hi << 4) + (lo >> 28);
}
__host__ __device__ inline u128 sin_fx_fast(u128 a) {
    constexpr TrigCoeffs T = trig_coeffs();
    const u128 z = mulq_low(a, a);
    u128 s = T.s[kCosTermsFast - 1];
    for(int k = kCosTermsFast - 2; k >= 0; --k) s = T.s[k] - mulq_low(z, s);
    return mulq_low(a, s);
}
__host__ __device__ inline u128 cos_fx_fast(u128 b) {
    constexpr TrigCoeffs T = trig_coeffs();
    const u128 z = mulq_low(b, b);
    u128 s = T.c[kCosTermsFast - 1];
    for(int k = kCosTermsFast - 2; k >= 0; --k) s = T.c[k] - mulq_low(z, s);
    return s;
}
// cos_ge at the cheaper precision; -1 also when that does not decide.  *d
// receives cos(y) - x at scale 2^124 (within kCosErrFast units).
__host__ __device__ inline int cos_ge_fast(double y, double x, i128* d = nullptr) {
    const u128 Y = static_cast<u128>(to_fx_signed(y));
    const u128 H = half_pi_fx();
    const bool neg = Y >= H;
    const u128 r = neg ? Y - H : H - Y;
    const u128 Q = H >> 1;
    const u128 s = r <= Q ? sin_fx_fast(r) : cos_fx_fast(r <= H ? H - r : r - H);
    const i128 c = neg ? -static_cast<i128>(s) : static_cast<i128>(s);
    const i128 X = to_fx_signed(x);
    const i128 e = static_cast<i128>(kCosErrFast);
    if(d) *d = c - X;
    if(c - e > X) return 1;
    if(c + e < X) return 0;
    return -1;
}
// cos_ge at full precision, out of line: the walk needs it rarely (see
// cos_ge_fast), and inline its 128-bit temporaries would raise the register
// count, and so lower the occupancy, of every kernel that evaluates acos.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __noinline__ int cos_ge_full(double y, double x) { return cos_ge(y, x); }
#else
inline int cos_ge_full(double y, double x) { return cos_ge(y, x); }
#endif
// cos_ge, tried at the cheaper precision first where that usually decides
// (y at least 2^-10 away from 0 and pi: neighbouring cosines at least about
// 2^-63 apart).
// *d (when the cheaper test decided; *fast tells) as in cos_ge_fast.
__host__ __device__ inline int cos_ge_2(double y, double x, i128* d = nullptr, bool* fast = nullptr) {
    if(y >= 0x1p-10 && y <= 0x1.91ffb54442d18p+1) {       // pi - 2^-10, rounded down
        const int t = cos_ge_fast(y, x, d);
        if(t >= 0) {
            if(fast) *fast = true;
            return t;
        }
    }
    if(fast) *fast = false;
    return cos_ge_full(y, x);
}

// The walk's second step from its first one's value, for y0 in [2^-10, pi -
// 2^-10] where the first step's cheaper test gives d = cos(y0) - x within
// 2^-86.  The neighbour's value follows from the slope:
//   cos(y0) - cos(y0 + u) = 2 sin(y0 + u/2) sin(u/2) = u sin(y0) (1 + eps)
// for the spacing u of the doubles at y0 (u <= 2^-51), |eps| < 2^-40 since
// sin(y0) >= sin(2^-10) > 2^-11; the same for y0 - u'.  With |d| <= 2^-45,
// sin(y0) = sqrt(1 - cos^2 y0) is sqrt((1 - x)(1 + x)) within a relative 2^-22
// (|cos^2 y0 - x^2| <= 2^-44 against sin^2 y0 > 2^-22), plus the few roundings
// of the double evaluation, in any rounding mode.  Each side is widened by a
// relative 2^-20 (and d by 2^-80 absolute).  True when the walk ends at the
// neighbour:
//   up (cos(y0) >= x, t = 1): cos(y0 + u) < x, i.e. d < u sin(y0) (1 + eps);
//   down (cos(y0) < x, t = 0): cos(y0 - u') >= x, i.e. -d <= u' sin(y0) (1 + eps).
// False inside the band and whenever the walk goes on past the neighbour.
__host__ __device__ inline bool acos_slope_ends(double x, double y0, double y1, i128 d) {
    const u128 m = static_cast<u128>(d < 0 ? -d : d);
    if(m >> 79) return false;                                        // |d| >= 2^-45
    const double md = (static_cast<double>(static_cast<std::uint64_t>(m >> 32)) * 0x1p32 +
                       static_cast<double>(static_cast<std::uint64_t>(m & 0xffffffffull))) * 0x1p-124;
    const double m_hi = md * (1.0 + 0x1p-20) + 0x1p-80;              // |d|, rounded up generously
    const double sn = __builtin_sqrt((1.0 - x) * (1.0 + x));        // sin(y0) within a relative 2^-22
    const double u = y1 > y0 ? y1 - y0 : y0 - y1;                    // exact: a power of two
    return m_hi < u * sn * (1.0 - 0x1p-20);
}

// RD(acos(x)) for x in (-1, 1) by walking from an approximation y0 with cos_ge;
// false when the walk does not settle within a few steps or a comparison is
// undecided (the caller then uses the series below).  One comparison per step,
// at one place in the code: from y0 upward while cos(y) >= x (the last such y
// is the result), or downward until cos(y) >= x.
__host__ __device__ inline bool acos_walk(double x, double y0, double* lo) {
    constexpr std::uint64_t kPiDown = 0x400921fb54442d18ull;   // largest double below pi
    const double pi_down = from_bits_h(kPiDown);
    double y = y0;
    if(!(y >= 0x1p-60)) y = 0x1p-60;                 // also catches NaN
    if(y > pi_down) y = pi_down;
    int dir = 0;                                      // +1 up (cos(y0) >= x), -1 down
    double prev = y;
    for(int i = 0; i < 9; ++i) {
        i128 d = 0;
        bool fast = false;
        const int t = cos_ge_2(y, x, &d, &fast);
        if(t < 0) return false;
        if(i == 0) {
            dir = t == 1 ? 1 : -1;
            // the second step from the slope, when it decides (acos_slope_ends)
            if(fast) {
                const double y1 = from_bits_h(to_bits_h(y) + (t == 1 ? 1 : -1));
                if(acos_slope_ends(x, y, y1, d)) {
                    *lo = t == 1 ? y : y1;
                    return true;
                }
            }
        }
        if(dir > 0) {
            if(t == 0) { *lo = prev; return true; }
            if(y == pi_down) { *lo = y; return true; }
            prev = y;
            y = from_bits_h(to_bits_h(y) + 1);
        } else {
            if(t == 1) { *lo = y; return true; }
            y = from_bits_h(to_bits_h(y) - 1);
            if(!(y >= 0x1p-60)) return false;
        }
    }
    return false;
}

// x in [-1, 1] (finite); returns RD(acos(x)) and RU(acos(x)) from the series
// alone; *hard is set when the direction could not be decided (results widened).
//   |x| <= 1/2: Y = (pi/2 -+ asin|x|) * 2^124, absolute error < E.
//   |x| >  1/2: u = (1 - |x|) / 2 (exact), scaled by 4^z so that the square root
//       carries 122 significant bits; Y = 2 sqrt(u) series(u) * 2^(124 + z) is
//       below the true value by less than E relative to its leading bit
//       (relative error < 2^-114, whatever the magnitude of acos), then for x < 0
//       acos(x) = pi - acos(|x|) at scale 2^124.
__host__ __device__ inline void acos_series(double x, double* lo, double* hi, bool* hard) {
    *hard = false;
    if(!(x >= -1.0 && x <= 1.0)) { *lo = *hi = __builtin_nan(""); return; }   // outside the domain, NaN
    if(x == 1.0) { *lo = 0.0; *hi = 0.0; return; }
    const bool neg = x < 0.0;
    const double ax = neg ? -x : x;
    // |x| as a fixed-point integer (exact unless |x| < 2^-71: then truncated by < 1 unit)
    const std::uint64_t b = to_bits_h(ax);
    const int ex = static_cast<int>((b >> 52) & 0x7ff);
    const std::uint64_t m = ex ? ((b & 0x000fffffffffffffull) | 0x0010000000000000ull) : (b & 0x000fffffffffffffull);
    const int sh = (ex ? ex : 1) - 1075 + kF;
    const u128 X = sh >= 0 ? (u128(m) << sh) : (sh > -64 ? u128(m >> (-sh)) : u128(0));
    u128 ylo, yhi;
    int z = 0;
    if(X <= (kUnit << (kF - 1))) {
        const u128 as = asin_fx(X);
        const u128 Y = neg ? half_pi_fx() + as : half_pi_fx() - as;
        ylo = Y > kErr ? Y - kErr : 0;
        yhi = Y + kErr;
    } else {
        const u128 U = ((kUnit << kF) - X) >> 1;      // (1 - |x|) / 2, exact, < 2^122
        const int p = msb128(U);
        const int zz = p < 120 ? (121 - p) / 2 : 0;
        const u128 S = sqrt_fx(U << (2 * zz));        // sqrt(u) * 2^(124 + zz), >= 2^122 (rounded down)
        const u128 Yz = mulq(S, asin_ratio_fx(U)) << 1;
        if(neg) {
            const u128 Y = pi_fx() - (Yz >> zz);      // acos(|x|) at scale 2^124 is below its true value by < E + 1
            ylo = Y - kErr - 1;
            yhi = Y + 1;
        } else {
            z = zz;
            ylo = Yz;
            yhi = Yz + kErr;
        }
    }
    const double dl = fx_down(ylo, z), dh = fx_down(yhi, z);
    const double ul = fx_up(ylo, z), uh = fx_up(yhi, z);
    *lo = dl;
    *hi = uh;
    *hard = (dl != dh) || (ul != uh);
}

// acos_series out of line (the rare fallback of acos_directed).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __noinline__ void acos_series_call(double x, double* lo, double* hi, bool* hard) { acos_series(x, lo, hi, hard); }
#else
inline void acos_series_call(double x, double* lo, double* hi, bool* hard) { acos_series(x, lo, hi, hard); }
#endif

// RD(acos(x)) and RU(acos(x)) for x in [-1, 1]: the cos walk from the math
// library's acos (any approximation works: the walk decides), else the series.
// acos(x) is irrational for x != 1, so RU is the next double above RD.
__host__ __device__ inline void acos_directed(double x, double* lo, double* hi, bool* hard) {
    *hard = false;
    if(x > -1.0 && x < 1.0) {
        double l;
        if(acos_walk(x, ::acos(x), &l)) {
            *lo = l;
            *hi = from_bits_h(to_bits_h(l) + 1);
            return;
        }
    }
    acos_series_call(x, lo, hi, hard);
}

} // namespace acosfx
} // namespace dev
} // namespace sqd
