// <ivarp/critical_printer.hpp> of the reference's lemma sources: a critical
// handler that prints every variable of a critical cuboid and further named
// expressions evaluated on it (ivarp/include/ivarp/critical_printer.hpp:30-120).
// The expressions are evaluated on the host in binary64 intervals widened
// outward by one unit in the last place after every operation (two for acos),
// an enclosure of their range that is meant for printing, not for proofs.
#pragma once
#include <cmath>
#include <limits>
#include <mutex>
#include <ostream>
#include <sstream>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "ivarp/math_fn.hpp"
#include "ivarp/run_prover.hpp"

namespace ivarp {

template<typename ExpressionType> struct PrintExpr {
    std::string name;
    ExpressionType expr;
};

template<typename StringArg, typename EType> inline auto printable_expression(StringArg&& name, EType&& expr) {
    return PrintExpr<std::decay_t<EType>>{std::string(std::forward<StringArg>(name)), std::forward<EType>(expr)};
}

namespace impl {

struct HostIv {
    double lb, ub;
    bool und;
};
struct HostIb {
    bool def, poss;
};

inline double down(double x, int ulps = 1) {
    for(int i = 0; i < ulps; ++i) x = std::nextafter(x, -std::numeric_limits<double>::infinity());
    return x;
}
inline double up(double x, int ulps = 1) {
    for(int i = 0; i < ulps; ++i) x = std::nextafter(x, std::numeric_limits<double>::infinity());
    return x;
}

inline HostIb host_pred(const NodeP& n, const std::vector<IntervalValue>& box);

inline HostIv host_eval(const NodeP& n, const std::vector<IntervalValue>& box) {
    auto k = [&](std::size_t i) { return host_eval(n->kids[i], box); };
    const double inf = std::numeric_limits<double>::infinity();
    switch(n->op) {
        case Op::Const: return HostIv{n->c_lb, n->c_ub, n->c_undef};
        case Op::Arg: {
            const IntervalValue& v = box.at(static_cast<std::size_t>(n->arg));
            return HostIv{v.lb, v.ub, false};
        }
        case Op::Neg: { HostIv a = k(0); return HostIv{-a.ub, -a.lb, a.und}; }
        case Op::Add: { HostIv a = k(0), b = k(1); return HostIv{down(a.lb + b.lb), up(a.ub + b.ub), a.und || b.und}; }
        case Op::Sub: { HostIv a = k(0), b = k(1); return HostIv{down(a.lb - b.ub), up(a.ub - b.lb), a.und || b.und}; }
        case Op::Mul: case Op::Square: {
            HostIv a = k(0), b = n->op == Op::Square ? a : k(1);
            double p[4] = {a.lb * b.lb, a.lb * b.ub, a.ub * b.lb, a.ub * b.ub};
            double lo = inf, hi = -inf;
            for(double v : p) { v = std::isnan(v) ? 0.0 : v; lo = std::fmin(lo, v); hi = std::fmax(hi, v); }
            lo = down(lo);
            if(n->op == Op::Square && lo < 0) lo = 0;
            return HostIv{lo, up(hi), a.und || b.und};
        }
        case Op::Div: {
            HostIv a = k(0), b = k(1);
            if(b.lb <= 0 && b.ub >= 0) return HostIv{-inf, inf, true};
            double p[4] = {a.lb / b.lb, a.lb / b.ub, a.ub / b.lb, a.ub / b.ub};
            double lo = inf, hi = -inf;
            for(double v : p) { lo = std::fmin(lo, v); hi = std::fmax(hi, v); }
            return HostIv{down(lo), up(hi), a.und || b.und};
        }
        case Op::Sqrt: {
            HostIv a = k(0);
            if(a.ub < 0) return HostIv{-inf, inf, true};
            const bool und = a.und || a.lb < 0;
            return HostIv{a.lb <= 0 ? 0.0 : std::fmax(0.0, down(std::sqrt(a.lb))), up(std::sqrt(a.ub)), und};
        }
        case Op::Acos: {
            HostIv a = k(0);
            const bool und = a.und || a.lb < -1 || a.ub > 1;
            const double l = std::fmax(-1.0, a.lb), u = std::fmin(1.0, a.ub);
            return HostIv{std::fmax(0.0, down(std::acos(u), 2)), up(std::acos(l), 2), und};
        }
        case Op::Max: case Op::Min: {
            HostIv r = k(0);
            for(std::size_t i = 1; i < n->kids.size(); ++i) {
                HostIv b = k(i);
                r = n->op == Op::Max ? HostIv{std::fmax(r.lb, b.lb), std::fmax(r.ub, b.ub), r.und || b.und}
                                     : HostIv{std::fmin(r.lb, b.lb), std::fmin(r.ub, b.ub), r.und || b.und};
            }
            return r;
        }
        case Op::Ite: {
            const HostIb c = host_pred(n->kids[0], box);
            if(c.def) return k(1);
            if(!c.poss) return k(2);
            HostIv t = k(1), e = k(2);
            return HostIv{std::fmin(t.lb, e.lb), std::fmax(t.ub, e.ub), t.und || e.und};
        }
        default: return HostIv{-inf, inf, true};
    }
}

inline HostIb host_pred(const NodeP& n, const std::vector<IntervalValue>& box) {
    auto less = [](HostIv a, HostIv b) { return HostIb{!a.und && !b.und && a.ub < b.lb, a.und || b.und || a.lb < b.ub}; };
    auto neg = [](HostIb x) { return HostIb{!x.poss, !x.def}; };
    switch(n->op) {
        case Op::BoolConst: return HostIb{n->bval_def, n->bval_poss};
        case Op::Lt: return less(host_eval(n->kids[0], box), host_eval(n->kids[1], box));
        case Op::Gt: return less(host_eval(n->kids[1], box), host_eval(n->kids[0], box));
        case Op::Le: return neg(less(host_eval(n->kids[1], box), host_eval(n->kids[0], box)));
        case Op::Ge: return neg(less(host_eval(n->kids[0], box), host_eval(n->kids[1], box)));
        case Op::Eq: case Op::Neq: {
            HostIv a = host_eval(n->kids[0], box), b = host_eval(n->kids[1], box);
            HostIb e{!a.und && !b.und && a.lb == a.ub && b.lb == b.ub && a.lb == b.lb,
                     a.und || b.und || (!(a.ub < b.lb) && !(b.ub < a.lb))};
            return n->op == Op::Eq ? e : neg(e);
        }
        case Op::Not: return neg(host_pred(n->kids[0], box));
        case Op::Or: case Op::OrSeq: case Op::And: case Op::AndSeq: {
            const bool is_or = n->op == Op::Or || n->op == Op::OrSeq;
            HostIb r{!is_or, !is_or};
            for(const auto& c : n->kids) {
                HostIb v = host_pred(c, box);
                r = is_or ? HostIb{r.def || v.def, r.poss || v.poss} : HostIb{r.def && v.def, r.poss && v.poss};
            }
            return r;
        }
        default: return HostIb{false, true};
    }
}

} // namespace impl

template<typename... PrintExprs> class CriticalPrinter {
public:
    CriticalPrinter(const ConstraintSystem* lookup, std::ostream* output, PrintExprs... exprs)
        : name_lookup(lookup), output(output), exprs(std::move(exprs)...) {}

    template<typename Context, typename ValueArray> void operator()(const Context&, const ValueArray& values) const {
        std::ostringstream out;
        out << "Critical:\n";
        std::vector<IntervalValue> box;
        for(std::size_t i = 0; i < name_lookup->num_args(); ++i) {
            out << '\t' << name_lookup->arg_name(i) << ": " << values[i] << std::endl;
            box.push_back(IntervalValue{values[i].lb, values[i].ub});
        }
        std::apply([&](const auto&... e) { (print_one(out, e, box), ...); }, exprs);
        std::lock_guard<std::mutex> guard(*lock);
        *output << out.str();
    }

private:
    template<typename E> static void print_one(std::ostringstream& out, const PrintExpr<E>& e, const std::vector<IntervalValue>& box) {
        const impl::HostIv v = impl::host_eval(Expr(e.expr).node, box);
        out << '\t' << e.name << ": " << IntervalValue{v.lb, v.ub} << (v.und ? " (possibly undefined)" : "") << std::endl;
    }
    const ConstraintSystem* name_lookup;
    std::ostream* output;
    std::tuple<PrintExprs...> exprs;
    std::shared_ptr<std::mutex> lock = std::make_shared<std::mutex>();
};

template<typename... PrintExprs>
inline auto critical_printer(std::ostream& output, const ConstraintSystem& system, PrintExprs&&... exprs) {
    return CriticalPrinter<std::decay_t<PrintExprs>...>(&system, &output, std::forward<PrintExprs>(exprs)...);
}

} // namespace ivarp
