// Per-operation compile-time bound rules (the EvalBounds of each reference tag).
#pragma once
#include <vector>
#include <utility>
#include <stdexcept>
#include "expr.hpp"
#include "fixed_point.hpp"

namespace ivarp {
namespace rules {
namespace fpb = fixed_point_bounds;

// ivarp/include/ivarp/math_fn/ternary_ops.hpp:40-51
inline fpb::Bounds ite_bounds(bool clb, bool cub, fpb::Bounds t, fpb::Bounds e) {
    fpb::Bounds b3 = clb ? t : e;
    fpb::Bounds b2 = !cub ? e : t;
    return {std::min(b2.lb, b3.lb), std::max(b2.ub, b3.ub)};
}

inline fpb::Bounds expr_bounds(Op op, const std::vector<fpb::Bounds>& k) {
    switch(op) {
        case Op::Neg: return {-k[0].ub, -k[0].lb};
        case Op::Add: return {fpb::add_rd(k[0].lb, k[1].lb), fpb::add_ru(k[0].ub, k[1].ub)};
        case Op::Sub: return {fpb::add_rd(k[0].lb, -k[1].ub), fpb::add_ru(k[0].ub, -k[1].lb)};
        case Op::Mul: return {fpb::iv_mul_lb(k[0].lb, k[0].ub, k[1].lb, k[1].ub),
                              fpb::iv_mul_ub(k[0].lb, k[0].ub, k[1].lb, k[1].ub)};
        case Op::Div: return {fpb::iv_div_lb(k[0].lb, k[0].ub, k[1].lb, k[1].ub),
                              fpb::iv_div_ub(k[0].lb, k[0].ub, k[1].lb, k[1].ub)};
        case Op::Sqrt: return fpb::iv_sqrt(k[0].lb, k[0].ub);
        case Op::Square: return {fpb::pow_lb(k[0].lb, k[0].ub, 2), fpb::pow_ub(k[0].lb, k[0].ub, 2)};
        case Op::Max: {
            // fixed_point_bounds::maximum folds from the right with a strict >
            std::int64_t l = k.back().lb, u = k.back().ub;
            for(std::size_t i = k.size() - 1; i-- > 0;) {
                l = k[i].lb > l ? k[i].lb : l;
                u = k[i].ub > u ? k[i].ub : u;
            }
            return {l, u};
        }
        case Op::Min: {
            std::int64_t l = k.back().lb, u = k.back().ub;
            for(std::size_t i = k.size() - 1; i-- > 0;) {
                l = k[i].lb < l ? k[i].lb : l;
                u = k[i].ub < u ? k[i].ub : u;
            }
            return {l, u};
        }
        case Op::Acos: return {fpb::min_bound(), fpb::max_bound()}; // tag without EvalBounds
        default: throw std::logic_error("expr_bounds: not an expression operation");
    }
}

inline void cmp_bounds(Op op, fpb::Bounds a, fpb::Bounds b, bool& lb, bool& ub) {
    using O = fpb::Order;
    O o = fpb::iv_order(a.lb, a.ub, b.lb, b.ub);
    switch(op) {
        case Op::Le: lb = (o == O::LE || o == O::LT || o == O::EQ); ub = (o != O::GT); return;
        case Op::Lt: lb = (o == O::LT); ub = (o == O::UNKNOWN || o == O::LE || o == O::LT); return;
        case Op::Ge: lb = (o == O::GE || o == O::GT || o == O::EQ); ub = (o != O::LT); return;
        case Op::Gt: lb = (o == O::GT); ub = (o == O::UNKNOWN || o == O::GE || o == O::GT); return;
        case Op::Eq: lb = (o == O::EQ); ub = (o != O::GT && o != O::LT); return;
        case Op::Neq: lb = (o == O::GT || o == O::LT); ub = (o != O::EQ); return;
        default: throw std::logic_error("cmp_bounds: not a comparison");
    }
}

inline void junctor_bounds(Op op, const std::vector<std::pair<bool, bool>>& k, bool& lb, bool& ub) {
    switch(op) {
        case Op::Not: lb = !k[0].second; ub = !k[0].first; return;
        case Op::OrSeq: case Op::Or:
            lb = false; ub = false;
            for(auto& p : k) { lb = lb || p.first; ub = ub || p.second; }
            return;
        case Op::AndSeq: case Op::And:
            lb = true; ub = true;
            for(auto& p : k) { lb = lb && p.first; ub = ub && p.second; }
            return;
        default: throw std::logic_error("junctor_bounds: unsupported predicate operation");
    }
}

} // namespace rules
} // namespace ivarp
