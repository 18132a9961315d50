// Interval operations of the GPU prover core (outward rounded; binary64, and
// binary32 for the FP32 prefilter: every operation is a template over the bound type).
//
// Semantics follow the reference's IDouble operations bit for bit:
//   + -            ivarp/include/ivarp/number/float_interval_addsub.hpp:53-102
//   *              float_interval_mul.hpp:33-166 (sign cases, NaN -> 0 unless finite)
//   /              float_interval_div.hpp:45-107, interval_div.hpp:57-288
//   sqrt           math_fn/interval_sqrt.hpp:99-116 (clamp + undefined flag)
//   square         math_fn/fixed_pow.hpp:183-187
//   min / max      number/minmax.hpp:59-87
//   comparisons    number/interval_comparisons.hpp:32-176, IBool bool.hpp:70-149
//   if_then_else   math_fn/eval/if_then_else.hpp:64-82
// Decisions the reference takes from compile-time bounds (sign of an operand,
// finiteness) arrive here as template arguments: Tri::T / Tri::F when the
// compile-time bounds decide, Tri::R when the runtime value decides.
#pragma once
#include "rounding.hpp"
#include "acos_fixed.hpp"
#include "acos_enclosure.hpp"

namespace sqd {
namespace dev {

template<class R> struct IVT { R lb, ub; bool und; };
using IV = IVT<double>;
struct IB { bool def, poss; };

enum Tri : int { F = 0, T = 1, R = 2 };

template<class X> __device__ __forceinline__ X inf_v() { return static_cast<X>(__builtin_huge_val()); }

template<class X> __device__ __forceinline__ IVT<X> iv_add(IVT<X> a, IVT<X> b) { return IVT<X>{rd_add(a.lb, b.lb), ru_add(a.ub, b.ub), a.und || b.und}; }
template<class X> __device__ __forceinline__ IVT<X> iv_sub(IVT<X> a, IVT<X> b) { return IVT<X>{rd_sub(a.lb, b.ub), ru_sub(a.ub, b.lb), a.und || b.und}; }
template<class X> __device__ __forceinline__ IVT<X> iv_neg(IVT<X> a) { return IVT<X>{-a.ub, -a.lb, a.und}; }

template<class X> __device__ __forceinline__ X nan_to_zero(X v) { return v != v ? X(0) : v; }

template<bool Finite, class X>
__device__ __forceinline__ void do_mul(X l1, X l2, X u1, X u2, X& lo, X& hi) {
    lo = rd_mul(l1, l2);
    hi = ru_mul(u1, u2);
    if(!Finite) { lo = nan_to_zero(lo); hi = nan_to_zero(hi); }
}

template<int C> __device__ __forceinline__ bool tri(bool runtime) { return C == T ? true : (C == F ? false : runtime); }

// Multiplication with the reference's nine sign cases.  NN/NP: "non-negative" /
// "non-positive" decisions for each operand.
template<int NN1, int NP1, int NN2, int NP2, bool Finite, class X>
__device__ inline IVT<X> iv_mul(IVT<X> a, IVT<X> b) {
    X lo, hi;
    const bool a_nn = tri<NN1>(a.lb >= X(0)), a_np = tri<NP1>(a.ub <= X(0));
    const bool b_nn = tri<NN2>(b.lb >= X(0)), b_np = tri<NP2>(b.ub <= X(0));
    if(a_nn) {
        if(b_nn) do_mul<Finite>(a.lb, b.lb, a.ub, b.ub, lo, hi);
        else if(b_np) do_mul<Finite>(a.ub, b.lb, a.lb, b.ub, lo, hi);
        else do_mul<Finite>(a.ub, b.lb, a.ub, b.ub, lo, hi);
    } else if(a_np) {
        if(b_nn) do_mul<Finite>(a.lb, b.ub, a.ub, b.lb, lo, hi);
        else if(b_np) do_mul<Finite>(a.ub, b.ub, a.lb, b.lb, lo, hi);
        else do_mul<Finite>(a.lb, b.ub, a.lb, b.lb, lo, hi);
    } else if(b_nn) {
        do_mul<Finite>(a.lb, b.ub, a.ub, b.ub, lo, hi);
    } else if(b_np) {
        do_mul<Finite>(a.ub, b.lb, a.lb, b.lb, lo, hi);
    } else {
        X l1, h1, l2, h2;
        do_mul<Finite>(a.lb, b.ub, a.lb, b.lb, l1, h1);
        do_mul<Finite>(b.lb, a.ub, a.ub, b.ub, l2, h2);
        lo = (l2 <= l1) ? l2 : l1;
        hi = (h1 <= h2) ? h2 : h1;
    }
    return IVT<X>{lo, hi, a.und || b.und};
}

// square(x) = x * x with run-time sign checks, then bounded below by 0
// (math_fn/fixed_pow.hpp:183-187).  The product's sign cases with both operands
// equal reduce to: lb >= 0: [lb^2, ub^2]; ub <= 0: [ub^2, lb^2]; otherwise the
// lower bound lb*ub < 0 is raised to +0 and the upper bound is the larger of
// lb^2 and ub^2 (same comparison as the general product, so the same bits).
template<class X>
__device__ inline IVT<X> iv_square(IVT<X> a) {
    X lo, hi;
    if(a.lb >= X(0)) {
        lo = rd_mul(a.lb, a.lb);
        hi = ru_mul(a.ub, a.ub);
    } else if(a.ub <= X(0)) {
        lo = rd_mul(a.ub, a.ub);
        hi = ru_mul(a.lb, a.lb);
    } else {
        const X lh = nan_to_zero(ru_mul(a.lb, a.lb)), uh = nan_to_zero(ru_mul(a.ub, a.ub));
        lo = X(0);
        hi = (lh <= uh) ? uh : lh;
    }
    lo = nan_to_zero(lo);
    hi = nan_to_zero(hi);
    if(lo < X(0)) lo = X(0);
    return IVT<X>{lo, hi, a.und};
}

template<class X> __device__ __forceinline__ int sgn(X v) { return (v > X(0)) - (v < X(0)); }

template<class X>
__device__ inline void divide_by_zero(IVT<X>& num, IVT<X> den, int sl, int su, int dsum) {
    const X inf = inf_v<X>();
    num.und = true;
    if(dsum == 0 || su - sl == 2) { num.lb = -inf; num.ub = inf; return; }
    if(dsum > 0) {
        if(sl >= 0) { num.lb = rd_div(num.lb, den.ub); num.ub = inf; }
        else { num.ub = ru_div(num.ub, den.ub); num.lb = -inf; }
    } else {
        if(sl >= 0) { num.ub = ru_div(num.lb, den.lb); num.lb = -inf; }
        else { num.lb = rd_div(num.ub, den.lb); num.ub = inf; }
    }
}

template<class X>
__device__ __forceinline__ X div_lb_inf(X n, int ni, X d, int di) {
    if(!ni && !di) return rd_div(n, d);
    if(di != 0) return X(0);
    return -inf_v<X>();
}
template<class X>
__device__ __forceinline__ X div_ub_inf(X n, int ni, X d, int di) {
    if(!ni && !di) return ru_div(n, d);
    if(di != 0) return X(0);
    return inf_v<X>();
}

template<class X>
__device__ inline void divide_by_infinite(IVT<X>& num, IVT<X> den) {
    num.und = num.und || den.und;
    const int snl = sgn(num.lb), snu = sgn(num.ub), sdl = sgn(den.lb), sdu = sgn(den.ub);
    const int inl = is_finite(num.lb) ? 0 : -1, inu = is_finite(num.ub) ? 0 : 1;
    const int idl = is_finite(den.lb) ? 0 : -1, idu = is_finite(den.ub) ? 0 : 1;
    X l, u;
    if(sdl > 0) {
        if(snl >= 0) { l = div_lb_inf(num.lb, inl, den.ub, idu); u = div_ub_inf(num.ub, inu, den.lb, idl); }
        else if(snu <= 0) { l = div_lb_inf(num.lb, inl, den.lb, idl); u = div_ub_inf(num.ub, inu, den.ub, idu); }
        else { l = div_lb_inf(num.lb, inl, den.lb, idl); u = div_ub_inf(num.ub, inu, den.lb, idl); }
    } else if(sdu < 0) {
        if(snl >= 0) { l = div_lb_inf(num.ub, inu, den.ub, idu); u = div_ub_inf(num.lb, inl, den.lb, idl); }
        else if(snu <= 0) { l = div_lb_inf(num.ub, inu, den.lb, idl); u = div_ub_inf(num.lb, inl, den.ub, idu); }
        else { l = div_lb_inf(num.ub, inu, den.ub, idu); u = div_ub_inf(num.lb, inl, den.ub, idu); }
    } else {
        divide_by_zero(num, den, snl, snu, sdl + sdu);
        return;
    }
    num.lb = l;
    num.ub = u;
}

// Division; POS/NEG: "denominator positive/negative"; NN/NP for the numerator;
// CTFinite: both operands have finite compile-time bounds.
template<int NN, int NP, int POS, int NEG, bool CTFinite, class X>
__device__ inline IVT<X> iv_div(IVT<X> num, IVT<X> den) {
    if(!CTFinite && !(is_finite(num.lb) && is_finite(num.ub) && is_finite(den.lb) && is_finite(den.ub))) {
        divide_by_infinite(num, den);
        return num;
    }
    num.und = num.und || den.und;
    X l, u;
    if(tri<POS>(den.lb > X(0))) {
        if(tri<NN>(num.lb >= X(0))) { l = rd_div(num.lb, den.ub); u = ru_div(num.ub, den.lb); }
        else if(tri<NP>(num.ub <= X(0))) { l = rd_div(num.lb, den.lb); u = ru_div(num.ub, den.ub); }
        else { l = rd_div(num.lb, den.lb); u = ru_div(num.ub, den.lb); }
    } else if(tri<NEG>(den.ub < X(0))) {
        if(tri<NN>(num.lb >= X(0))) { l = rd_div(num.ub, den.ub); u = ru_div(num.lb, den.lb); }
        else if(tri<NP>(num.ub <= X(0))) { l = rd_div(num.ub, den.lb); u = ru_div(num.lb, den.ub); }
        else { l = rd_div(num.ub, den.ub); u = ru_div(num.lb, den.ub); }
    } else {
        divide_by_zero(num, den, sgn(num.lb), sgn(num.ub), sgn(den.lb) + sgn(den.ub));
        return num;
    }
    num.lb = l;
    num.ub = u;
    return num;
}

template<int NN, class X>
__device__ inline IVT<X> iv_sqrt(IVT<X> x) {
    X lb = x.lb, ub = x.ub;
    bool und = x.und;
    if(!tri<NN>(x.lb >= X(0))) {
        if(ub < X(0)) return IVT<X>{-inf_v<X>(), inf_v<X>(), true};
        lb = X(0);
        und = true;
    }
    return IVT<X>{rd_sqrt(lb), ru_sqrt(ub), und};
}

// Evaluations of acos whose rounding direction could not be decided (see
// acos_fixed.hpp); one counter per compiled program, read by the host.
static __device__ unsigned long long g_acos_hard = 0;

// acos with the reference's domain handling (math_fn/interval_asin.hpp:13-27,
// interval_acos.hpp:19-29): where the argument's compile-time bounds allow values
// outside [-1, 1] the interval is clamped and marked possibly undefined.
template<bool CheckLb, bool CheckUb>
__device__ inline IV iv_acos(IV x) {
    bool und = x.und;
    if(CheckLb && x.lb < -1.0) { x.lb = -1.0; und = true; }
    if(CheckUb && x.ub > 1.0) { x.ub = 1.0; und = true; }
    double l1, h1, l2, h2;
    bool hard1, hard2;
    acosfx::acos_directed(x.ub, &l1, &h1, &hard1);
    acosfx::acos_directed(x.lb, &l2, &h2, &hard2);
    if(hard1 || hard2) atomicAdd(&g_acos_hard, 1ull);
    return IV{l1, h2, und};
}
// The binary32 version: the same domain handling, and the binary64 enclosure of
// acosenc::acos_enclose (about 2^-39 wide) converted outward, which contains
// acos and is the correctly rounded pair of floats but in rare cases (then one
// float wider); the binary32 pass only has to enclose.
template<bool CheckLb, bool CheckUb>
__device__ inline IVT<float> iv_acos(IVT<float> x) {
    bool und = x.und;
    double l = static_cast<double>(x.lb), u = static_cast<double>(x.ub);
    if(CheckLb && l < -1.0) { l = -1.0; und = true; }
    if(CheckUb && u > 1.0) { u = 1.0; und = true; }
    double lo, hi, dummy;
    acosenc::acos_enclose(u, lo, dummy);
    acosenc::acos_enclose(l, dummy, hi);
    return IVT<float>{f32_down(lo), f32_up(hi), und};
}

template<class X> __device__ __forceinline__ X dmax(X n1, X n2) { return n2 < n1 ? n1 : n2; }
template<class X> __device__ __forceinline__ X dmin(X n1, X n2) { return n1 < n2 ? n1 : n2; }
template<class X> __device__ __forceinline__ IVT<X> iv_max(IVT<X> a, IVT<X> b) { return IVT<X>{dmax(a.lb, b.lb), dmax(a.ub, b.ub), a.und || b.und}; }
template<class X> __device__ __forceinline__ IVT<X> iv_min(IVT<X> a, IVT<X> b) { return IVT<X>{dmin(a.lb, b.lb), dmin(a.ub, b.ub), a.und || b.und}; }

template<class X>
__device__ __forceinline__ IVT<X> iv_ite(IB c, IVT<X> t, IVT<X> e) {
    if(c.def) return t;
    if(!c.poss) return e;
    IVT<X> r = t;
    if(e.lb < r.lb) r.lb = e.lb;
    if(e.ub > r.ub) r.ub = e.ub;
    r.und = r.und || e.und;
    return r;
}

template<class X>
__device__ __forceinline__ IB ib_less(IVT<X> a, IVT<X> b) {
    return IB{!a.und && !b.und && a.ub < b.lb, a.und || b.und || a.lb < b.ub};
}
__device__ __forceinline__ IB ib_not(IB x) { return IB{!x.poss, !x.def}; }
template<class X> __device__ __forceinline__ IB ib_lt(IVT<X> a, IVT<X> b) { return ib_less(a, b); }
template<class X> __device__ __forceinline__ IB ib_gt(IVT<X> a, IVT<X> b) { return ib_less(b, a); }
template<class X> __device__ __forceinline__ IB ib_ge(IVT<X> a, IVT<X> b) { return ib_not(ib_less(a, b)); }
template<class X> __device__ __forceinline__ IB ib_le(IVT<X> a, IVT<X> b) { return ib_not(ib_less(b, a)); }
template<class X> __device__ __forceinline__ IB ib_eq(IVT<X> a, IVT<X> b) {
    return IB{!a.und && !b.und && a.lb == a.ub && b.lb == b.ub && a.lb == b.lb,
              a.und || b.und || (!(a.ub < b.lb) && !(b.ub < a.lb))};
}
template<class X> __device__ __forceinline__ IB ib_neq(IVT<X> a, IVT<X> b) { return ib_not(ib_eq(a, b)); }
__device__ __forceinline__ IB ib_or(IB a, IB b) { return IB{a.def || b.def, a.poss || b.poss}; }
__device__ __forceinline__ IB ib_and(IB a, IB b) { return IB{a.def && b.def, a.poss && b.poss}; }

} // namespace dev
} // namespace sqd
