// Fixed-point (denominator 10^6) compile-time bound arithmetic used by the lemma
// lowering.  Every function here restates the integer algorithm of
// ivarp/include/ivarp/number/fixed_point_bounds.hpp, fixed_point_bounds_impl.hpp
// and fixed_point_bounds_sqrt.hpp, including their saturation rules and the
// exact digit-by-digit long division, because these bounds decide runtime
// behaviour (multiplication sign cases, sqrt clamping, if_then_else folding,
// filtered constraints, propagation dependencies).
#pragma once
#include <cstdint>
#include <algorithm>

namespace ivarp {
namespace fixed_point_bounds {

constexpr std::int64_t denom() noexcept { return 1000000; }
constexpr std::int64_t max_bound() noexcept { return INT64_C(9223372036854775807); }
constexpr std::int64_t min_bound() noexcept { return -max_bound(); }

constexpr bool is_lb(std::int64_t v) noexcept { return v > min_bound(); }
constexpr bool is_ub(std::int64_t v) noexcept { return v < max_bound(); }
constexpr bool is_finite(std::int64_t l, std::int64_t u) noexcept { return l > min_bound() && u < max_bound(); }
constexpr bool nonzero(std::int64_t l, std::int64_t u) noexcept { return l > 0 || u < 0; }

namespace detail {
inline std::int64_t uadd(std::uint64_t a, std::uint64_t b) noexcept {
    std::uint64_t s = a + b;
    return s >= static_cast<std::uint64_t>(max_bound()) ? max_bound() : static_cast<std::int64_t>(s);
}
inline std::int64_t sadd(std::int64_t a, std::int64_t b) noexcept {
    if(a < 0 && b < 0) return -uadd(static_cast<std::uint64_t>(-a), static_cast<std::uint64_t>(-b));
    if(a > 0 && b > 0) return uadd(static_cast<std::uint64_t>(a), static_cast<std::uint64_t>(b));
    return a + b;
}
inline std::int64_t sat_umul(std::uint64_t a, std::uint64_t b) noexcept {
    if(b == 0) return 0;
    if(a > static_cast<std::uint64_t>(max_bound()) / b) return max_bound();
    return static_cast<std::int64_t>(a * b);
}
// a, b >= 0; product rounded toward zero (rd) or away from zero (ru) at 1e-6.
inline std::int64_t mag_mul(std::int64_t a, std::int64_t b, bool up) noexcept {
    const std::int64_t D = denom();
    std::int64_t ia = a / D, ib = b / D, fa = a % D, fb = b % D;
    std::int64_t whole = sat_umul(static_cast<std::uint64_t>(sat_umul(ia, ib)), D);
    std::int64_t cross = uadd(static_cast<std::uint64_t>(sat_umul(ia, fb)),
                              static_cast<std::uint64_t>(sat_umul(ib, fa)));
    std::int64_t ff = fa * fb;
    std::int64_t frac = ff / D + ((up && ff % D != 0) ? 1 : 0);
    return uadd(static_cast<std::uint64_t>(uadd(whole, cross)), frac);
}
inline unsigned lead_zeros(std::uint64_t x) noexcept { return x == 0 ? 64u : static_cast<unsigned>(__builtin_clzll(x)); }
// (a * 10^6) / b with 128-bit intermediate, following the two-digit (base 2^32)
// schoolbook division of the reference, including its choice of comparison digit
// in the second quotient-digit correction loop, so that results agree bit-for-bit.
inline std::int64_t long_div(std::uint64_t a, std::uint64_t b, std::uint64_t round_up) noexcept {
    const std::uint64_t D = static_cast<std::uint64_t>(denom());
    const std::uint64_t M32 = 0xffffffffull;
    std::uint64_t lo_prod = (a & M32) * D;
    std::uint64_t hi_prod = (a >> 32) * D + (lo_prod >> 32);
    std::uint64_t nh = hi_prod >> 32;
    std::uint64_t nl = ((hi_prod & M32) << 32) | (lo_prod & M32);
    if(nh >= b) return max_bound();
    unsigned sh = lead_zeros(b);
    if(sh > 0) {
        b <<= sh;
        nh = (nh << sh) | (nl >> (64u - sh));
        nl <<= sh;
    }
    const std::uint64_t bh = b >> 32, bl = b & M32;
    std::uint64_t qhat = nh / bh, rhat = nh % bh;
    while((qhat >> 32) != 0 || (qhat & M32) * bl > ((rhat << 32) | (nl >> 32))) {
        qhat -= 1;
        rhat += bh;
        if((rhat >> 32) != 0) break;
    }
    std::uint32_t q_hi = static_cast<std::uint32_t>(qhat);
    std::uint64_t u = ((nh << 32) | (nl >> 32)) - static_cast<std::uint64_t>(q_hi) * b;
    qhat = u / bh;
    rhat = u % bh;
    while((qhat >> 32) != 0 || (qhat & M32) * bl > ((rhat << 32) | (nl >> 32))) {
        qhat -= 1;
        rhat += bh;
        if((rhat >> 32) != 0) break;
    }
    std::uint32_t q_lo = static_cast<std::uint32_t>(qhat);
    std::uint64_t quot = (static_cast<std::uint64_t>(q_hi) << 32) | q_lo;
    std::uint64_t rem = (((u << 32) | (nl & M32)) - static_cast<std::uint64_t>(q_lo) * b) >> sh;
    if(quot >= static_cast<std::uint64_t>(max_bound())) return max_bound();
    return static_cast<std::int64_t>(quot + round_up * (rem != 0 ? 1u : 0u));
}
inline std::int64_t mag_div_rd(std::int64_t a, std::int64_t b) noexcept {
    return b >= max_bound() ? 0 : long_div(static_cast<std::uint64_t>(a), static_cast<std::uint64_t>(b), 0);
}
inline std::int64_t mag_div_ru(std::int64_t a, std::int64_t b) noexcept {
    return (b == 0 || a >= max_bound()) ? max_bound()
                                        : long_div(static_cast<std::uint64_t>(a), static_cast<std::uint64_t>(b), 1);
}
inline std::int64_t iabs(std::int64_t v) noexcept { return v < 0 ? -v : v; }
} // namespace detail

inline std::int64_t int_to_fp(std::int64_t v) noexcept {
    return v < 0 ? -detail::sat_umul(static_cast<std::uint64_t>(-v), denom())
                 : detail::sat_umul(static_cast<std::uint64_t>(v), denom());
}
inline std::int64_t add_rd(std::int64_t a, std::int64_t b) noexcept {
    return (a <= min_bound() || b <= min_bound()) ? min_bound() : detail::sadd(a, b);
}
inline std::int64_t add_ru(std::int64_t a, std::int64_t b) noexcept {
    return (a >= max_bound() || b >= max_bound()) ? max_bound() : detail::sadd(a, b);
}
inline std::int64_t mul_rd(std::int64_t a, std::int64_t b) noexcept {
    using namespace detail;
    return ((a < 0) != (b < 0)) ? -mag_mul(iabs(a), iabs(b), true) : mag_mul(iabs(a), iabs(b), false);
}
inline std::int64_t mul_ru(std::int64_t a, std::int64_t b) noexcept {
    using namespace detail;
    return ((a < 0) != (b < 0)) ? -mag_mul(iabs(a), iabs(b), false) : mag_mul(iabs(a), iabs(b), true);
}
inline std::int64_t div_rd(std::int64_t a, std::int64_t b) noexcept {
    using namespace detail;
    return ((a < 0) != (b < 0)) ? -mag_div_ru(iabs(a), iabs(b)) : mag_div_rd(iabs(a), iabs(b));
}
inline std::int64_t div_ru(std::int64_t a, std::int64_t b) noexcept {
    using namespace detail;
    return ((a < 0) != (b < 0)) ? -mag_div_rd(iabs(a), iabs(b)) : mag_div_ru(iabs(a), iabs(b));
}

struct Bounds { std::int64_t lb, ub; };

inline std::int64_t iv_mul_lb(std::int64_t la, std::int64_t ua, std::int64_t lb, std::int64_t ub) noexcept {
    if((!is_lb(la) && ub > 0) || (!is_lb(lb) && ua > 0)) return min_bound();
    if((!is_ub(ua) && lb < 0) || (!is_ub(ub) && la < 0)) return min_bound();
    return std::min(std::min(mul_rd(la, lb), mul_rd(la, ub)), std::min(mul_rd(ua, lb), mul_rd(ua, ub)));
}
inline std::int64_t iv_mul_ub(std::int64_t la, std::int64_t ua, std::int64_t lb, std::int64_t ub) noexcept {
    if((!is_ub(ua) && ub > 0) || (!is_ub(ub) && ua > 0)) return max_bound();
    if((!is_lb(la) && lb < 0) || (!is_lb(lb) && la < 0)) return max_bound();
    return std::max(std::max(mul_ru(la, lb), mul_ru(la, ub)), std::max(mul_ru(ua, lb), mul_ru(ua, ub)));
}
inline std::int64_t iv_div_lb(std::int64_t la, std::int64_t ua, std::int64_t lb, std::int64_t ub) noexcept {
    if((lb < 0 && ub > 0) || (lb == 0 && ub == 0)) return min_bound();
    if(la == 0 && ua == 0) return 0;
    if(lb == 0) return la < 0 ? min_bound() : div_rd(la, ub);
    if(ub == 0) return ua > 0 ? min_bound() : detail::mag_div_rd(-ua, -lb);
    if(lb > 0) return div_rd(la, la < 0 ? lb : ub);
    return div_rd(ua, ua > 0 ? ub : lb);
}
inline std::int64_t iv_div_ub(std::int64_t la, std::int64_t ua, std::int64_t lb, std::int64_t ub) noexcept {
    if((lb < 0 && ub > 0) || (lb == 0 && ub == 0)) return max_bound();
    if(la == 0 && ua == 0) return 0;
    if(lb == 0) return ua > 0 ? max_bound() : div_ru(ua, ub);
    if(ub == 0) return la < 0 ? max_bound() : -detail::mag_div_ru(la, -lb);
    if(lb > 0) return div_ru(ua, ua < 0 ? ub : lb);
    return div_ru(la, la > 0 ? lb : ub);
}
inline std::int64_t pow_lb(std::int64_t lb, std::int64_t ub, unsigned p) noexcept {
    if(!is_lb(lb) && !is_ub(ub)) return min_bound();
    if(p == 0) return int_to_fp(1);
    if(p % 2 == 1 && !is_lb(lb)) return min_bound();
    int sign = (p % 2 == 0) ? 1 : (lb < 0 ? -1 : 1);
    std::int64_t op = (p % 2 == 0 && ub < 0) ? ub : ((p % 2 == 0 && lb < 0 && ub > 0) ? 0 : lb);
    op = detail::iabs(op);
    std::int64_t r = op;
    for(unsigned i = 1; i < p; ++i) r = sign > 0 ? mul_rd(r, op) : mul_ru(r, op);
    return r * sign;
}
inline std::int64_t pow_ub(std::int64_t lb, std::int64_t ub, unsigned p) noexcept {
    if(!is_lb(lb) && !is_ub(ub)) return max_bound();
    if(p == 0) return int_to_fp(1);
    int sign = (p % 2 == 0) ? 1 : (ub < 0 ? -1 : 1);
    std::int64_t op = (p % 2 == 0) ? std::max(-lb, ub) : ub;
    op = detail::iabs(op);
    std::int64_t r = op;
    for(unsigned i = 1; i < p; ++i) r = sign > 0 ? mul_ru(r, op) : mul_rd(r, op);
    return r * sign;
}
namespace detail {
inline std::int64_t bisect_sqrt_rd(std::int64_t x, std::int64_t lo, std::int64_t hi) noexcept {
    while(hi - lo > 1) {
        std::int64_t mid = (lo + hi) / 2;
        if(mul_ru(mid, mid) >= x) hi = mid; else lo = mid;
    }
    return mul_ru(hi, hi) == x ? hi : lo;
}
inline std::int64_t bisect_sqrt_ru(std::int64_t x, std::int64_t lo, std::int64_t hi) noexcept {
    while(hi - lo > 1) {
        std::int64_t mid = (lo + hi) / 2;
        if(mul_rd(mid, mid) > x) hi = mid; else lo = mid;
    }
    return mul_rd(lo, lo) == x ? lo : hi;
}
} // namespace detail
inline std::int64_t sqrt_rd(std::int64_t x) noexcept {
    if(x == 0) return 0;
    if(x >= max_bound()) return INT64_C(3037000499976);
    if(x < denom()) return detail::bisect_sqrt_rd(x, x, denom());
    if(x <= int_to_fp(16)) return detail::bisect_sqrt_rd(x, x / 4, x);
    return detail::bisect_sqrt_rd(x, 0, x / 4 + 1);
}
inline std::int64_t sqrt_ru(std::int64_t x) noexcept {
    if(x == 0) return 0;
    if(x >= max_bound()) return max_bound();
    if(x < denom()) return detail::bisect_sqrt_ru(x, x, denom());
    if(x <= int_to_fp(16)) return detail::bisect_sqrt_ru(x, x / 4, x);
    return detail::bisect_sqrt_ru(x, 0, x / 4 + 1);
}
inline Bounds iv_sqrt(std::int64_t lb, std::int64_t ub) noexcept {
    if(lb < 0) return {min_bound(), max_bound()};
    return {sqrt_rd(lb), sqrt_ru(ub)};
}

enum class Order { LT, LE, EQ, GE, GT, UNKNOWN };
inline Order iv_order(std::int64_t la, std::int64_t ua, std::int64_t lb, std::int64_t ub) noexcept {
    if(is_ub(ua) && is_lb(lb)) {
        if(ua < lb) return Order::LT;
        if(ua == lb) return (is_ub(ub) && is_lb(la) && ub == la) ? Order::EQ : Order::LE;
    }
    if(is_lb(la) && is_ub(ub)) {
        if(ub < la) return Order::GT;
        if(ub == la) return Order::GE;
    }
    return Order::UNKNOWN;
}

} // namespace fixed_point_bounds
} // namespace ivarp
