// Construction of the runtime expression DAG and compile-time bound evaluation.
// See expr.hpp.  Bound rules per node kind restate the EvalBounds members of the
// reference's tags: ivarp/include/ivarp/math_fn/binary_ops.hpp:32-104,
// unary_ops.hpp:33-38, unary_fns.hpp:144-147 (sqrt), fixed_pow.hpp:36-41,
// n_ary_ops.hpp:30-78, ternary_ops.hpp:40-51, math_pred/binary_ops.hpp:126-326,
// math_pred/n_ary_ops.hpp:30-105, math_pred/unary_ops.hpp:34-40.
#include "expr.hpp"
#include "bounds_rules.hpp"
#include <unordered_map>
#include <algorithm>
#include <limits>

namespace ivarp {
namespace fpb = fixed_point_bounds;

const char* op_name(Op o) noexcept {
    switch(o) {
        case Op::Const: return "const"; case Op::Arg: return "arg"; case Op::Neg: return "neg";
        case Op::Add: return "add"; case Op::Sub: return "sub"; case Op::Mul: return "mul";
        case Op::Div: return "div"; case Op::Sqrt: return "sqrt"; case Op::Square: return "square";
        case Op::Max: return "max"; case Op::Min: return "min"; case Op::Ite: return "ite";
        case Op::Acos: return "acos"; case Op::BoolConst: return "bconst"; case Op::Lt: return "lt";
        case Op::Le: return "le"; case Op::Gt: return "gt"; case Op::Ge: return "ge"; case Op::Eq: return "eq";
        case Op::Neq: return "neq"; case Op::Not: return "not"; case Op::OrSeq: return "orseq";
        case Op::AndSeq: return "andseq"; case Op::Or: return "or"; case Op::And: return "and";
    }
    return "?";
}

NodeP make_node(Op op, std::vector<NodeP> kids) {
    auto n = std::make_shared<Node>();
    n->op = op;
    n->kids = std::move(kids);
    return n;
}
NodeP make_arg(int index) {
    auto n = std::make_shared<Node>();
    n->op = Op::Arg;
    n->arg = index;
    return n;
}

// ---------------- literals (BoundedRational) ----------------
Literal make_int_literal(const char* digits) {
    std::string s;
    for(const char* p = digits; *p; ++p) if(*p != '\'') s.push_back(*p);
    std::size_t nz = s.find_first_not_of('0');
    s = (nz == std::string::npos) ? std::string("0") : s.substr(nz);
    exact::Rat v = exact::Rat::from_decimal(s);
    // Literals with 19 or more digits (or 19 digits starting with 9) do not fit;
    // their bounds become [8999999999999999999, max].
    bool fits = s.size() < 19 || (s.size() == 19 && s[0] < '9');
    if(!fits) return Literal{v, INT64_C(8999999999999999999), fpb::max_bound()};
    std::int64_t iv = std::stoll(s);
    std::int64_t f = fpb::int_to_fp(iv);
    return Literal{v, f, f};
}
Literal make_decimal_literal(const char* digits) {
    std::string s;
    for(const char* p = digits; *p; ++p) if(*p != '\'') s.push_back(*p);
    std::size_t pt = s.find('.');
    std::string ip = pt == std::string::npos ? s : s.substr(0, pt);
    std::string fp = pt == std::string::npos ? std::string() : s.substr(pt + 1);
    std::size_t nz = ip.find_first_not_of('0');
    ip = (nz == std::string::npos) ? std::string() : ip.substr(nz);
    std::int64_t ival = ip.empty() ? 0 : std::stoll(ip);
    std::int64_t base = fpb::int_to_fp(ival);
    std::int64_t frac = 0, mult = fpb::denom() / 10;
    bool tail_nonzero = false;
    for(std::size_t i = 0; i < fp.size(); ++i) {
        int dgt = fp[i] - '0';
        if(i < 6) { frac += mult * dgt; mult /= 10; }
        else if(dgt != 0) tail_nonzero = true;
    }
    std::int64_t lb = fpb::detail::uadd(static_cast<std::uint64_t>(base), static_cast<std::uint64_t>(frac));
    std::int64_t ub = fpb::detail::uadd(static_cast<std::uint64_t>(base),
                                        static_cast<std::uint64_t>(frac + (tail_nonzero ? 1 : 0)));
    return Literal{exact::Rat::from_decimal(s), lb, ub};
}
Literal operator-(const Literal& a) { return Literal{-a.value, -a.ub, -a.lb}; }
Literal operator+(const Literal& a, const Literal& b) {
    return Literal{a.value + b.value, fpb::add_rd(a.lb, b.lb), fpb::add_ru(a.ub, b.ub)};
}
Literal operator-(const Literal& a, const Literal& b) {
    return Literal{a.value - b.value, fpb::add_rd(a.lb, -b.ub), fpb::add_ru(a.ub, -b.lb)};
}
Literal operator*(const Literal& a, const Literal& b) {
    return Literal{a.value * b.value, fpb::iv_mul_lb(a.lb, a.ub, b.lb, b.ub), fpb::iv_mul_ub(a.lb, a.ub, b.lb, b.ub)};
}
Literal operator/(const Literal& a, const Literal& b) {
    return Literal{a.value / b.value, fpb::iv_div_lb(a.lb, a.ub, b.lb, b.ub), fpb::iv_div_ub(a.lb, a.ub, b.lb, b.ub)};
}

NodeP literal_node(const Literal& l) {
    auto n = std::make_shared<Node>();
    n->op = Op::Const;
    n->exact_value = std::make_shared<exact::RInterval>(exact::RInterval::point(l.value));
    auto d = exact::outward_double(l.value);
    n->c_lb = d.first;
    n->c_ub = d.second;
    n->bounded = true;
    n->lb = l.lb;
    n->ub = l.ub;
    return n;
}

Expr::Expr(int v) {
    auto n = std::make_shared<Node>();
    n->op = Op::Const;
    n->exact_value = std::make_shared<exact::RInterval>(exact::RInterval::point(exact::Rat(v)));
    n->c_lb = n->c_ub = static_cast<double>(v);
    n->bounded = true;
    n->lb = fpb::int_to_fp(std::numeric_limits<int>::min());
    n->ub = fpb::int_to_fp(std::numeric_limits<int>::max());
    node = n;
}

Pred::Pred(bool b) {
    auto n = std::make_shared<Node>();
    n->op = Op::BoolConst;
    n->bounded = true;
    n->plb = false;
    n->pub = true;
    n->bval_def = n->bval_poss = b;
    node = n;
}

// ---------------- expression construction ----------------
Expr operator-(const Expr& a) {
    // -(-x) is x (Negate specialisation, ivarp/include/ivarp/math_fn/unary_ops.hpp:52-64).
    if(a.node->op == Op::Neg) return Expr(a.node->kids[0]);
    return Expr(make_node(Op::Neg, {a.node}));
}
Expr operator+(const Expr& a, const Expr& b) { return Expr(make_node(Op::Add, {a.node, b.node})); }
Expr operator-(const Expr& a, const Expr& b) { return Expr(make_node(Op::Sub, {a.node, b.node})); }
Expr operator*(const Expr& a, const Expr& b) { return Expr(make_node(Op::Mul, {a.node, b.node})); }
Expr operator/(const Expr& a, const Expr& b) { return Expr(make_node(Op::Div, {a.node, b.node})); }
Expr sqrt(const Expr& a) { return Expr(make_node(Op::Sqrt, {a.node})); }
Expr acos(const Expr& a) { return Expr(make_node(Op::Acos, {a.node})); }
Expr square(const Expr& a) { return Expr(make_node(Op::Square, {a.node})); }
Expr maximum_v(std::vector<NodeP> kids) {
    if(kids.size() == 1) return Expr(kids[0]);
    return Expr(make_node(Op::Max, std::move(kids)));
}
Expr minimum_v(std::vector<NodeP> kids) {
    if(kids.size() == 1) return Expr(kids[0]);
    return Expr(make_node(Op::Min, std::move(kids)));
}
Expr if_then_else_impl(const Pred& c, const Expr& t, const Expr& e) {
    return Expr(make_node(Op::Ite, {c.node, t.node, e.node}));
}
Pred compare(Op op, const Expr& a, const Expr& b) { return Pred(make_node(op, {a.node, b.node})); }

// n-ary junctors merge with operands carrying the same junctor
// (ivarp/include/ivarp/math_fn/math_pred/n_ary_ops.hpp:120-205).
static Pred junctor(Op op, const Pred& a, const Pred& b) {
    std::vector<NodeP> k;
    if(a.node->op == op) k = a.node->kids; else k.push_back(a.node);
    if(b.node->op == op) k.insert(k.end(), b.node->kids.begin(), b.node->kids.end()); else k.push_back(b.node);
    return Pred(make_node(op, std::move(k)));
}
Pred operator||(const Pred& a, const Pred& b) { return junctor(Op::OrSeq, a, b); }
Pred operator&&(const Pred& a, const Pred& b) { return junctor(Op::AndSeq, a, b); }
Pred operator|(const Pred& a, const Pred& b) { return junctor(Op::Or, a, b); }
Pred operator&(const Pred& a, const Pred& b) { return junctor(Op::And, a, b); }
Pred operator!(const Pred& a) { return Pred(make_node(Op::Not, {a.node})); }

// Symbolic call: replace argument i by replacements[i] throughout the tree.
Expr Expr::call(const std::vector<NodeP>& r) const {
    std::unordered_map<const Node*, NodeP> memo;
    int nargs = num_args(node);
    if(static_cast<std::size_t>(nargs) != r.size())
        throw std::invalid_argument("symbolic call with the wrong number of arguments");
    struct Sub {
        const std::vector<NodeP>& r;
        std::unordered_map<const Node*, NodeP>& memo;
        NodeP go(const NodeP& n) {
            auto it = memo.find(n.get());
            if(it != memo.end()) return it->second;
            NodeP out;
            if(n->op == Op::Arg) out = r[static_cast<std::size_t>(n->arg)];
            else if(n->kids.empty()) out = n;
            else {
                auto m = std::make_shared<Node>(*n);
                bool changed = false;
                for(auto& k : m->kids) { NodeP nk = go(k); changed |= (nk != k); k = nk; }
                out = changed ? NodeP(m) : n;
            }
            memo[n.get()] = out;
            return out;
        }
    } s{r, memo};
    return Expr(s.go(node));
}

int num_args(const NodeP& n) {
    if(n->op == Op::Arg) return n->arg + 1;
    int m = 0;
    for(const auto& k : n->kids) m = std::max(m, num_args(k));
    return m;
}

// ---------------- compile-time bounds (CompileTimeBounds) ----------------
fpb::Bounds ct_bounds(const NodeP& n, const std::vector<fpb::Bounds>& a) {
    if(n->bounded) return {n->lb, n->ub};
    if(n->op == Op::Arg) {
        if(static_cast<std::size_t>(n->arg) >= a.size())
            throw std::invalid_argument("bound expression uses a variable of higher index");
        return a[static_cast<std::size_t>(n->arg)];
    }
    if(n->op == Op::Ite) {
        bool plb, pub;
        ct_pred_bounds(n->kids[0], a, plb, pub);
        return rules::ite_bounds(plb, pub, ct_bounds(n->kids[1], a), ct_bounds(n->kids[2], a));
    }
    std::vector<fpb::Bounds> kb;
    for(const auto& k : n->kids) kb.push_back(ct_bounds(k, a));
    return rules::expr_bounds(n->op, kb);
}
void ct_pred_bounds(const NodeP& n, const std::vector<fpb::Bounds>& a, bool& plb, bool& pub) {
    if(n->bounded) { plb = n->plb; pub = n->pub; return; }
    switch(n->op) {
        case Op::Lt: case Op::Le: case Op::Gt: case Op::Ge: case Op::Eq: case Op::Neq:
            rules::cmp_bounds(n->op, ct_bounds(n->kids[0], a), ct_bounds(n->kids[1], a), plb, pub);
            return;
        default: break;
    }
    std::vector<std::pair<bool, bool>> kb;
    for(const auto& k : n->kids) { bool l, u; ct_pred_bounds(k, a, l, u); kb.emplace_back(l, u); }
    rules::junctor_bounds(n->op, kb, plb, pub);
}

namespace impl {
ConstraintSystem finish_constraint_system(CSBuilder&& b) {
    ConstraintSystem cs;
    std::size_t n = b.vars.size();
    cs.descriptions_by_index.resize(n);
    std::vector<bool> seen(n, false);
    for(const auto& v : b.vars) {
        if(v.index < 0 || static_cast<std::size_t>(v.index) >= n || seen[static_cast<std::size_t>(v.index)])
            throw std::invalid_argument("there must be exactly one description per argument index");
        seen[static_cast<std::size_t>(v.index)] = true;
        cs.descriptions_by_index[static_cast<std::size_t>(v.index)] = v;
    }
    // bound constraints first, in the order the variables were given
    for(const auto& v : b.vars) {
        Expr x(make_arg(v.index));
        cs.constraints.push_back((x >= v.lb).node);
        cs.constraints.push_back((x <= v.ub).node);
    }
    cs.constraints.insert(cs.constraints.end(), b.cons.begin(), b.cons.end());
    for(std::size_t i = 0; i < n; ++i) {
        const auto& d = cs.descriptions_by_index[i];
        if(num_args(d.lb.node) > static_cast<int>(i) || num_args(d.ub.node) > static_cast<int>(i))
            throw std::invalid_argument("variable bound depends on a variable of higher index");
        std::int64_t lb = ct_bounds(d.lb.node, cs.ct_bounds).lb;
        std::int64_t ub = ct_bounds(d.ub.node, cs.ct_bounds).ub;
        cs.ct_bounds.push_back({lb, ub});
    }
    return cs;
}
}

} // namespace ivarp
