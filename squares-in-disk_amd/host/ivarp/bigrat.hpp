// Exact rational arithmetic for constant folding in the lemma lowering.
//
// The reference folds argument-free sub-expressions (sqrt(1.6), 2/sqrt(5), ...)
// with GMP rationals and 256-bit MPFR intervals and then rounds the enclosure
// outward to double (ivarp/include/ivarp/constant_folding/constant_folding_expr.hpp:45-50,
// ivarp/src/convert_number.cpp:72-93).  This file provides the same capability
// without GMP: arbitrary-precision integers, exact rationals, rational intervals
// whose irrational operations (sqrt) are enclosed to 2^-320, and the
// truncate-then-step conversion to a double interval.
#pragma once
#include <cstdint>
#include <vector>
#include <string>
#include <stdexcept>
#include <cmath>
#include <limits>
#include <utility>

namespace ivarp {
namespace exact {

// Unsigned magnitude, little-endian base-2^32 limbs, no leading zero limbs.
class Nat {
public:
    Nat() = default;
    explicit Nat(std::uint64_t v) { while(v) { d.push_back(static_cast<std::uint32_t>(v)); v >>= 32; } }
    bool zero() const noexcept { return d.empty(); }
    std::size_t bits() const noexcept {
        if(d.empty()) return 0;
        return 32 * (d.size() - 1) + (32 - static_cast<std::size_t>(__builtin_clz(d.back())));
    }
    static int cmp(const Nat& a, const Nat& b) noexcept {
        if(a.d.size() != b.d.size()) return a.d.size() < b.d.size() ? -1 : 1;
        for(std::size_t i = a.d.size(); i-- > 0;) {
            if(a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
        }
        return 0;
    }
    static Nat add(const Nat& a, const Nat& b) {
        Nat r; r.d.resize(std::max(a.d.size(), b.d.size()) + 1, 0);
        std::uint64_t c = 0;
        for(std::size_t i = 0; i < r.d.size(); ++i) {
            std::uint64_t s = c + (i < a.d.size() ? a.d[i] : 0) + (i < b.d.size() ? b.d[i] : 0);
            r.d[i] = static_cast<std::uint32_t>(s); c = s >> 32;
        }
        r.trim(); return r;
    }
    // requires a >= b
    static Nat sub(const Nat& a, const Nat& b) {
        Nat r; r.d.resize(a.d.size(), 0);
        std::int64_t br = 0;
        for(std::size_t i = 0; i < a.d.size(); ++i) {
            std::int64_t s = static_cast<std::int64_t>(a.d[i]) - br - (i < b.d.size() ? b.d[i] : 0);
            br = s < 0 ? 1 : 0;
            r.d[i] = static_cast<std::uint32_t>(s + (br ? (INT64_C(1) << 32) : 0));
        }
        if(br) throw std::logic_error("Nat::sub underflow");
        r.trim(); return r;
    }
    static Nat mul(const Nat& a, const Nat& b) {
        if(a.zero() || b.zero()) return Nat{};
        Nat r; r.d.assign(a.d.size() + b.d.size(), 0);
        for(std::size_t i = 0; i < a.d.size(); ++i) {
            std::uint64_t c = 0;
            for(std::size_t j = 0; j < b.d.size(); ++j) {
                std::uint64_t t = static_cast<std::uint64_t>(a.d[i]) * b.d[j] + r.d[i + j] + c;
                r.d[i + j] = static_cast<std::uint32_t>(t); c = t >> 32;
            }
            std::size_t k = i + b.d.size();
            while(c) { std::uint64_t t = static_cast<std::uint64_t>(r.d[k]) + c; r.d[k] = static_cast<std::uint32_t>(t); c = t >> 32; ++k; }
        }
        r.trim(); return r;
    }
    Nat shl(std::size_t s) const {
        if(zero()) return *this;
        Nat r; std::size_t limbs = s / 32, b = s % 32;
        r.d.assign(limbs, 0);
        std::uint32_t carry = 0;
        for(std::uint32_t v : d) {
            r.d.push_back(b ? ((v << b) | carry) : v);
            carry = b ? (v >> (32 - b)) : 0;
        }
        if(carry) r.d.push_back(carry);
        r.trim(); return r;
    }
    Nat shr(std::size_t s) const {
        std::size_t limbs = s / 32, b = s % 32;
        if(limbs >= d.size()) return Nat{};
        Nat r; r.d.assign(d.begin() + static_cast<std::ptrdiff_t>(limbs), d.end());
        if(b) {
            for(std::size_t i = 0; i < r.d.size(); ++i) {
                std::uint32_t hi = (i + 1 < r.d.size()) ? r.d[i + 1] : 0;
                r.d[i] = (r.d[i] >> b) | (hi << (32 - b));
            }
        }
        r.trim(); return r;
    }
    // Binary long division: returns quotient, sets rem.
    static Nat divmod(const Nat& a, const Nat& b, Nat& rem) {
        if(b.zero()) throw std::domain_error("Nat division by zero");
        if(cmp(a, b) < 0) { rem = a; return Nat{}; }
        std::size_t n = a.bits();
        Nat q; q.d.assign((n + 31) / 32, 0);
        Nat r;
        for(std::size_t i = n; i-- > 0;) {
            r = r.shl(1);
            if((a.d[i / 32] >> (i % 32)) & 1u) { if(r.d.empty()) r.d.push_back(1); else r.d[0] |= 1u; }
            if(cmp(r, b) >= 0) { r = sub(r, b); q.d[i / 32] |= (1u << (i % 32)); }
        }
        q.trim(); rem = r; return q;
    }
    static Nat gcd(Nat a, Nat b) {
        while(!b.zero()) { Nat r; divmod(a, b, r); a = std::move(b); b = std::move(r); }
        return a;
    }
    // floor(sqrt(a))
    static Nat isqrt(const Nat& a) {
        if(a.zero()) return Nat{};
        Nat x = Nat(1).shl((a.bits() + 1) / 2 + 1);
        for(;;) {
            Nat r; Nat y = add(x, divmod(a, x, r)).shr(1);
            if(cmp(y, x) >= 0) return x;
            x = std::move(y);
        }
    }
    bool is_one() const noexcept { return d.size() == 1 && d[0] == 1; }
    bool bit(std::size_t i) const noexcept { return i / 32 < d.size() && ((d[i / 32] >> (i % 32)) & 1u); }
    std::uint64_t low64() const noexcept {
        std::uint64_t v = 0;
        if(d.size() > 0) v |= d[0];
        if(d.size() > 1) v |= static_cast<std::uint64_t>(d[1]) << 32;
        return v;
    }
private:
    void trim() { while(!d.empty() && d.back() == 0) d.pop_back(); }
    std::vector<std::uint32_t> d;
};

// Exact rational number in lowest terms; den > 0.
class Rat {
public:
    Rat() : den(1) {}
    Rat(std::int64_t v) : neg(v < 0), num(static_cast<std::uint64_t>(v < 0 ? -v : v)), den(1) {} // NOLINT
    Rat(bool negative, Nat n, Nat d) : neg(negative), num(std::move(n)), den(std::move(d)) { norm(); }
    static Rat from_decimal(const std::string& s) {
        Nat n, d(1);
        bool seen_point = false;
        for(char c : s) {
            if(c == '.') { seen_point = true; continue; }
            if(c == '\'') continue;
            if(c < '0' || c > '9') throw std::invalid_argument("bad decimal literal: " + s);
            n = Nat::add(Nat::mul(n, Nat(10)), Nat(static_cast<std::uint64_t>(c - '0')));
            if(seen_point) d = Nat::mul(d, Nat(10));
        }
        return Rat(false, n, d);
    }
    bool is_zero() const noexcept { return num.zero(); }
    bool negative() const noexcept { return neg && !num.zero(); }
    int sign() const noexcept { return num.zero() ? 0 : (neg ? -1 : 1); }
    const Nat& numer() const noexcept { return num; }
    const Nat& denom() const noexcept { return den; }
    friend Rat operator-(const Rat& a) { Rat r = a; r.neg = !r.neg; r.norm(); return r; }
    friend Rat operator+(const Rat& a, const Rat& b) {
        Nat x = Nat::mul(a.num, b.den), y = Nat::mul(b.num, a.den), dd = Nat::mul(a.den, b.den);
        if(a.neg == b.neg) return Rat(a.neg, Nat::add(x, y), dd);
        if(Nat::cmp(x, y) >= 0) return Rat(a.neg, Nat::sub(x, y), dd);
        return Rat(b.neg, Nat::sub(y, x), dd);
    }
    friend Rat operator-(const Rat& a, const Rat& b) { return a + (-b); }
    friend Rat operator*(const Rat& a, const Rat& b) {
        return Rat(a.neg != b.neg, Nat::mul(a.num, b.num), Nat::mul(a.den, b.den));
    }
    friend Rat operator/(const Rat& a, const Rat& b) {
        if(b.num.zero()) throw std::domain_error("rational division by zero");
        return Rat(a.neg != b.neg, Nat::mul(a.num, b.den), Nat::mul(a.den, b.num));
    }
    static int cmp(const Rat& a, const Rat& b) {
        int sa = a.sign(), sb = b.sign();
        if(sa != sb) return sa < sb ? -1 : 1;
        int c = Nat::cmp(Nat::mul(a.num, b.den), Nat::mul(b.num, a.den));
        return sa < 0 ? -c : c;
    }
    friend bool operator<(const Rat& a, const Rat& b) { return cmp(a, b) < 0; }
    friend bool operator==(const Rat& a, const Rat& b) { return cmp(a, b) == 0; }
    // Value truncated toward zero to a double, plus whether that is exact.
    // Restricted to the normal double range, which is all constant folding needs.
    double to_double_trunc(bool& exact) const {
        if(num.zero()) { exact = true; return 0.0; }
        long k = static_cast<long>(num.bits()) - static_cast<long>(den.bits());
        // want q = floor(|r| * 2^(52-k)) in [2^52, 2^53)
        auto scaled = [&](long e, Nat& rem) {
            Nat n = num, dd = den;
            if(e >= 0) n = n.shl(static_cast<std::size_t>(e)); else dd = dd.shl(static_cast<std::size_t>(-e));
            Nat q = Nat::divmod(n, dd, rem);
            return q;
        };
        Nat rem;
        long e = 52 - k;
        Nat q = scaled(e, rem);
        if(q.bits() < 53) { e += 1; q = scaled(e, rem); }
        if(q.bits() > 53) { e -= 1; q = scaled(e, rem); }
        exact = rem.zero();
        double v = std::ldexp(static_cast<double>(q.low64()), static_cast<int>(-e));
        if(!std::isfinite(v) || (v != 0.0 && std::fabs(v) < std::numeric_limits<double>::min()))
            throw std::range_error("constant outside the normal double range");
        return neg ? -v : v;
    }
    std::string str() const;
private:
    void norm() {
        if(num.zero()) { neg = false; den = Nat(1); return; }
        if(den.is_one()) return;
        Nat g = Nat::gcd(num, den);
        if(!g.is_one()) { Nat r; num = Nat::divmod(num, g, r); den = Nat::divmod(den, g, r); }
    }
    bool neg = false;
    Nat num, den;
};

// Outward double enclosure of an exact rational, with the reference's
// truncate-then-step rule (ivarp/src/convert_number.cpp:72-93).
inline std::pair<double, double> outward_double(const Rat& r) {
    bool exact = false;
    double d = r.to_double_trunc(exact);
    if(exact) return {d, d};
    if(d == 0) {
        if(r.negative()) return {-std::numeric_limits<double>::min(), 0.0};
        return {0.0, std::numeric_limits<double>::min()};
    }
    if(d > 0) return {d, std::nextafter(d, std::numeric_limits<double>::infinity())};
    return {std::nextafter(d, -std::numeric_limits<double>::infinity()), d};
}

// Closed interval with rational end points; used to evaluate argument-free
// sub-expressions.  Operations that are exact on rationals stay exact; sqrt
// encloses the true root within 2^-320.
struct RInterval {
    Rat lo, hi;
    bool undefined = false;
    static RInterval point(const Rat& r) { return {r, r, false}; }
    bool is_point() const { return lo == hi; }
};

inline RInterval iv_add(const RInterval& a, const RInterval& b) { return {a.lo + b.lo, a.hi + b.hi, a.undefined || b.undefined}; }
inline RInterval iv_sub(const RInterval& a, const RInterval& b) { return {a.lo - b.hi, a.hi - b.lo, a.undefined || b.undefined}; }
inline RInterval iv_neg(const RInterval& a) { return {-a.hi, -a.lo, a.undefined}; }
inline RInterval iv_mul(const RInterval& a, const RInterval& b) {
    Rat p[4] = {a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi};
    Rat mn = p[0], mx = p[0];
    for(const Rat& x : p) { if(x < mn) mn = x; if(mx < x) mx = x; }
    return {mn, mx, a.undefined || b.undefined};
}
inline RInterval iv_div(const RInterval& a, const RInterval& b) {
    if(b.lo.sign() <= 0 && b.hi.sign() >= 0)
        throw std::domain_error("constant folding: division by an interval containing zero");
    RInterval inv{Rat(1) / b.hi, Rat(1) / b.lo, b.undefined};
    return iv_mul(a, inv);
}
inline Rat sqrt_bound(const Rat& x, bool upper) {
    constexpr std::size_t P = 320;
    // sqrt(n/d) = sqrt(n*d)/d; scale by 2^P
    Nat nd = Nat::mul(x.numer(), x.denom()).shl(2 * P);
    Nat s = Nat::isqrt(nd);
    if(upper && Nat::cmp(Nat::mul(s, s), nd) != 0) s = Nat::add(s, Nat(1));
    return Rat(false, s, x.denom().shl(P));
}
inline RInterval iv_sqrt(const RInterval& a) {
    if(a.hi.sign() < 0) throw std::domain_error("constant folding: sqrt of a negative constant");
    bool und = a.undefined;
    Rat lo = a.lo;
    if(lo.sign() < 0) { lo = Rat(0); und = true; }
    return {sqrt_bound(lo, false), sqrt_bound(a.hi, true), und};
}
inline RInterval iv_square(const RInterval& a) {
    RInterval r = iv_mul(a, a);
    if(r.lo.sign() < 0) r.lo = Rat(0);
    return r;
}
inline RInterval iv_max(const RInterval& a, const RInterval& b) {
    return {a.lo < b.lo ? b.lo : a.lo, a.hi < b.hi ? b.hi : a.hi, a.undefined || b.undefined};
}
inline RInterval iv_min(const RInterval& a, const RInterval& b) {
    return {b.lo < a.lo ? b.lo : a.lo, b.hi < a.hi ? b.hi : a.hi, a.undefined || b.undefined};
}
inline std::pair<double, double> outward_double(const RInterval& r) {
    return {outward_double(r.lo).first, outward_double(r.hi).second};
}

} // namespace exact
} // namespace ivarp
