// Directed-rounding arc cosine for binary64, in exact 128-bit fixed-point
// integer arithmetic (independent of the floating-point rounding mode).
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

// x in [-1, 1] (finite); returns RD(acos(x)) and RU(acos(x)); *hard is set when
// the direction could not be decided (results widened).
//   |x| <= 1/2: Y = (pi/2 -+ asin|x|) * 2^124, absolute error < E.
//   |x| >  1/2: u = (1 - |x|) / 2 (exact), scaled by 4^z so that the square root
//       carries 122 significant bits; Y = 2 sqrt(u) series(u) * 2^(124 + z) is
//       below the true value by less than E relative to its leading bit
//       (relative error < 2^-114, whatever the magnitude of acos), then for x < 0
//       acos(x) = pi - acos(|x|) at scale 2^124.
__host__ __device__ inline void acos_directed(double x, double* lo, double* hi, bool* hard) {
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

} // namespace acosfx
} // namespace dev
} // namespace sqd
