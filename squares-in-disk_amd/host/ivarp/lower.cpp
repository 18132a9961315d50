// Lowering passes, restating the reference's compile-time refactoring on the
// runtime DAG:
//  * constant folding of argument-free sub-expressions
//      ivarp/include/ivarp/constant_folding/fold_constants.hpp:32-110, cuda_transformer.hpp:37-76
//  * bound-and-simplify with compile-time (fixed point) bounds
//      ivarp/include/ivarp/refactor_constraint_system/bound_and_simplify*.hpp, math_fn/ternary_ops.hpp:54-79
//  * rewriting of relational constraints into bounds on one variable
//      refactor_constraint_system/constraints_to_bounds/rewriters/relational_ops.hpp:31-129
//  * 5N iterations of bound tightening, then filtering of always-true constraints
//      refactor_constraint_system.hpp:187-194, refactor_constraints_impl.hpp:31-58, filter_constraints.hpp:35-61
//  * runtime bound table with dependency lists, runtime constraint table
//      runtime_bound_table.hpp:54-219, runtime_constraint_table.hpp:31-115,
//      bound_dependency_analysis/cbd_*.hpp
//  * split policy processing   prover_input/process_var_splitting.hpp:30-139
#include "lower.hpp"
#include "lower_passes.hpp"
#include "bounds_rules.hpp"
#include <unordered_map>
#include <map>
#include <tuple>
#include <algorithm>
#include <cstring>
#include <limits>

namespace ivarp {
namespace {
namespace fpb = fixed_point_bounds;
using fpb::Bounds;

NodeP with_bounds(const NodeP& n, std::int64_t lb, std::int64_t ub) {
    auto m = std::make_shared<Node>(*n);
    m->bounded = true;
    m->lb = lb;
    m->ub = ub;
    return m;
}
NodeP bool_const(bool v) {
    auto m = std::make_shared<Node>();
    m->op = Op::BoolConst;
    m->bounded = true;
    m->plb = m->pub = v;
    m->bval_def = m->bval_poss = v;
    return m;
}

// ---------------- constant folding ----------------
exact::RInterval exact_eval(const NodeP& n) {
    using namespace exact;
    switch(n->op) {
        case Op::Const:
            if(!n->exact_value) throw std::logic_error("constant without exact value inside a foldable tree");
            return *n->exact_value;
        case Op::Neg: return iv_neg(exact_eval(n->kids[0]));
        case Op::Add: return iv_add(exact_eval(n->kids[0]), exact_eval(n->kids[1]));
        case Op::Sub: return iv_sub(exact_eval(n->kids[0]), exact_eval(n->kids[1]));
        case Op::Mul: return iv_mul(exact_eval(n->kids[0]), exact_eval(n->kids[1]));
        case Op::Div: return iv_div(exact_eval(n->kids[0]), exact_eval(n->kids[1]));
        case Op::Sqrt: return iv_sqrt(exact_eval(n->kids[0]));
        case Op::Square: return iv_square(exact_eval(n->kids[0]));
        case Op::Max: case Op::Min: {
            RInterval r = exact_eval(n->kids.back());
            for(std::size_t i = n->kids.size() - 1; i-- > 0;) {
                RInterval k = exact_eval(n->kids[i]);
                r = n->op == Op::Max ? iv_max(k, r) : iv_min(k, r);
            }
            return r;
        }
        default: throw std::logic_error(std::string("constant folding of '") + op_name(n->op) + "' is not supported");
    }
}

struct Folder {
    std::unordered_map<const Node*, NodeP> memo;
    NodeP go(const NodeP& n) {
        auto it = memo.find(n.get());
        if(it != memo.end()) return it->second;
        NodeP out;
        if(n->op == Op::Const || n->op == Op::BoolConst) out = n;
        else if(num_args(n) == 0) {
            if(is_pred_op(n->op)) throw std::logic_error("argument-free predicates are not supported");
            Bounds b = ct_bounds(n, {});
            exact::RInterval v = exact_eval(n);
            auto d = exact::outward_double(v);
            auto m = std::make_shared<Node>();
            m->op = Op::Const;
            m->bounded = true;
            m->lb = b.lb;
            m->ub = b.ub;
            m->c_lb = d.first;
            m->c_ub = d.second;
            m->c_undef = v.undefined;
            out = m;
        } else {
            auto m = std::make_shared<Node>(*n);
            for(auto& k : m->kids) k = go(k);
            out = m;
        }
        memo[n.get()] = out;
        return out;
    }
};

// ---------------- bound and simplify ----------------
struct BoundAndSimplify {
    const std::vector<Bounds>& ab;
    std::unordered_map<const Node*, NodeP> memo;

    NodeP go(const NodeP& n) {
        auto it = memo.find(n.get());
        if(it != memo.end()) return it->second;
        NodeP out = apply(n);
        memo[n.get()] = out;
        return out;
    }

    NodeP apply(const NodeP& n) {
        switch(n->op) {
            case Op::Const: case Op::BoolConst: return n;
            case Op::Arg: {
                Bounds b = ab.at(static_cast<std::size_t>(n->arg));
                if(n->bounded) { b.lb = std::max(b.lb, n->lb); b.ub = std::min(b.ub, n->ub); }
                return with_bounds(n, b.lb, b.ub);
            }
            case Op::Ite: {
                NodeP c = go(n->kids[0]), t = go(n->kids[1]), e = go(n->kids[2]);
                NodeP r;
                if(c->plb && c->pub) r = t;
                else if(!c->plb && !c->pub) r = e;
                else {
                    auto m = std::make_shared<Node>();
                    m->op = Op::Ite;
                    m->kids = {c, t, e};
                    m->bounded = true;
                    m->lb = std::min(t->lb, e->lb);
                    m->ub = std::max(t->ub, e->ub);
                    r = m;
                }
                if(n->bounded) {
                    if(r->op == Op::Const)
                        throw std::logic_error("if_then_else folded to a constant inside a bounded expression");
                    r = with_bounds(r, std::max(r->lb, n->lb), std::min(r->ub, n->ub));
                }
                return r;
            }
            default: break;
        }
        auto m = std::make_shared<Node>();
        m->op = n->op;
        m->arg = n->arg;
        m->bounded = true;
        for(const auto& k : n->kids) m->kids.push_back(go(k));
        if(!is_pred_op(n->op)) {
            std::vector<Bounds> kb;
            for(const auto& k : m->kids) kb.push_back({k->lb, k->ub});
            Bounds eb = rules::expr_bounds(n->op, kb);
            if(n->bounded) { eb.lb = std::max(eb.lb, n->lb); eb.ub = std::min(eb.ub, n->ub); }
            m->lb = eb.lb;
            m->ub = eb.ub;
            return m;
        }
        bool plb, pub;
        switch(n->op) {
            case Op::Lt: case Op::Le: case Op::Gt: case Op::Ge: case Op::Eq: case Op::Neq:
                rules::cmp_bounds(n->op, {m->kids[0]->lb, m->kids[0]->ub}, {m->kids[1]->lb, m->kids[1]->ub}, plb, pub);
                break;
            default: {
                std::vector<std::pair<bool, bool>> kb;
                for(const auto& k : m->kids) kb.emplace_back(k->plb, k->pub);
                rules::junctor_bounds(n->op, kb, plb, pub);
            }
        }
        if(n->bounded) {
            // BoundedPredicate: simplify by the combined bounds
            bool l = plb || n->plb, u = pub && n->pub;
            if(l && u) return bool_const(true);
            if(!l && !u) return bool_const(false);
            m->plb = false;
            m->pub = true;
            return m;
        }
        m->plb = plb;
        m->pub = pub;
        return m;
    }
};

// ---------------- rewriting to bounds ----------------
struct Rewrite { int target; flat::Dir dir; NodeP expr; };

bool is_arg(int t, const NodeP& a) { return a->op == Op::Arg && a->arg == t; }
bool is_arg_or_neg(int t, const NodeP& a) {
    if(a->op == Op::Neg) return is_arg(t, a->kids[0]);
    return is_arg(t, a);
}
bool depends_on(const NodeP& n, int t) {
    if(n->op == Op::Arg) return n->arg == t;
    for(const auto& k : n->kids) if(depends_on(k, t)) return true;
    return false;
}
Op swap_sides(Op o) {
    switch(o) {
        case Op::Lt: return Op::Gt; case Op::Gt: return Op::Lt;
        case Op::Le: return Op::Ge; case Op::Ge: return Op::Le;
        default: return o;
    }
}
std::vector<Rewrite> rewrite(int t, const NodeP& c) {
    switch(c->op) {
        case Op::Le: case Op::Lt: case Op::Ge: case Op::Gt: case Op::Eq: {
            const NodeP& a1 = c->kids[0];
            const NodeP& a2 = c->kids[1];
            bool l = is_arg_or_neg(t, a1), r = is_arg_or_neg(t, a2);
            if(l == r) return {};
            const NodeP& a = l ? a1 : a2;
            const NodeP& na = l ? a2 : a1;
            Op st = l ? c->op : swap_sides(c->op);
            if(a->op == Op::Neg) throw std::logic_error("constraints on a negated variable are not supported");
            flat::Dir d = (st == Op::Le || st == Op::Lt) ? flat::Dir::LEQ
                        : (st == Op::Ge || st == Op::Gt) ? flat::Dir::GEQ : flat::Dir::BOTH;
            return {Rewrite{t, d, na}};
        }
        case Op::Not:
            if(depends_on(c, t)) throw std::logic_error("negated constraints are not supported by the rewriter");
            return {};
        default: return {};
    }
}

// ---------------- bound dependency analysis ----------------
struct Deps { bool ll = false, lu = false, ul = false, uu = false; };
enum Uses { U_LB = 1, U_UB = 2, U_BOTH = 3 };
bool dep_on_lb(int used, const Deps& d) { return ((used & U_LB) && d.ll) || ((used & U_UB) && d.ul); }
bool dep_on_ub(int used, const Deps& d) { return ((used & U_LB) && d.lu) || ((used & U_UB) && d.uu); }
Deps computation_uses(int lb1, int ub1, const Deps& d1, int lb2, int ub2, const Deps& d2) {
    return Deps{dep_on_lb(lb1, d1) || dep_on_lb(lb2, d2), dep_on_ub(lb1, d1) || dep_on_ub(lb2, d2),
                dep_on_lb(ub1, d1) || dep_on_lb(ub2, d2), dep_on_ub(ub1, d1) || dep_on_ub(ub2, d2)};
}
Deps deps(const NodeP& n, int a) {
    switch(n->op) {
        case Op::Arg: return n->arg == a ? Deps{true, false, false, true} : Deps{};
        case Op::Const: case Op::BoolConst: return Deps{};
        case Op::Neg: { Deps c = deps(n->kids[0], a); return Deps{c.ul, c.uu, c.ll, c.lu}; }
        case Op::Add: {
            Deps x = deps(n->kids[0], a), y = deps(n->kids[1], a);
            return Deps{x.ll || y.ll, x.lu || y.lu, x.ul || y.ul, x.uu || y.uu};
        }
        case Op::Sub: {
            Deps x = deps(n->kids[0], a), y = deps(n->kids[1], a);
            return Deps{x.ll || y.ul, x.lu || y.uu, x.ul || y.ll, x.uu || y.lu};
        }
        case Op::Mul: {
            const NodeP& c1 = n->kids[0];
            const NodeP& c2 = n->kids[1];
            int s1 = c1->lb >= 0 ? 1 : (c1->ub <= 0 ? -1 : 0);
            int s2 = c2->lb >= 0 ? 1 : (c2->ub <= 0 ? -1 : 0);
            int lb1 = s2 == 1 ? U_LB : s2 == -1 ? U_UB : s1 == 1 ? U_UB : s1 == -1 ? U_LB : U_BOTH;
            int ub1 = s2 == 1 ? U_UB : s2 == -1 ? U_LB : s1 == 1 ? U_UB : s1 == -1 ? U_LB : U_BOTH;
            int lb2 = s1 == 1 ? U_LB : s1 == -1 ? U_UB : s2 == 1 ? U_UB : s2 == -1 ? U_LB : U_BOTH;
            int ub2 = s1 == 1 ? U_UB : s1 == -1 ? U_LB : s2 == 1 ? U_UB : s2 == -1 ? U_LB : U_BOTH;
            return computation_uses(lb1, ub1, deps(c1, a), lb2, ub2, deps(c2, a));
        }
        case Op::Div: {
            const NodeP& c1 = n->kids[0];
            const NodeP& c2 = n->kids[1];
            int s1 = c1->lb >= 0 ? 1 : (c1->ub <= 0 ? -1 : 0);
            int s2 = c2->lb > 0 ? 1 : (c2->ub < 0 ? -1 : 0);
            int lb1 = s2 == 1 ? U_LB : s2 == -1 ? U_UB : U_BOTH;
            int ub1 = s2 == 1 ? U_UB : s2 == -1 ? U_LB : U_BOTH;
            int lb2 = s2 == 0 ? U_BOTH : s1 == 1 ? U_UB : s1 == -1 ? U_LB : s2 == 1 ? U_LB : U_UB;
            int ub2 = s2 == 0 ? U_BOTH : s1 == 1 ? U_LB : s1 == -1 ? U_UB : s2 == 1 ? U_LB : U_UB;
            return computation_uses(lb1, ub1, deps(c1, a), lb2, ub2, deps(c2, a));
        }
        default: {
            bool any = false;
            for(const auto& k : n->kids) { Deps d = deps(k, a); any = any || d.ll || d.lu || d.ul || d.uu; }
            return Deps{any, any, any, any};
        }
    }
}

// ---------------- flattening with hash-consing ----------------
struct Flattener {
    flat::Program& prog;
    std::map<std::tuple<int, std::vector<int>, int, std::uint64_t, std::uint64_t, int, std::int64_t, std::int64_t, int>, int> ids;
    std::unordered_map<const Node*, int> memo;

    static std::uint64_t bits(double d) { std::uint64_t u; std::memcpy(&u, &d, 8); return u; }

    int go(const NodeP& n) {
        auto it = memo.find(n.get());
        if(it != memo.end()) return it->second;
        flat::FNode f;
        f.op = to_fop(n->op);
        for(const auto& k : n->kids) f.kids.push_back(go(k));
        f.arg = n->arg;
        if(n->op == Op::Const) { f.c_lb = n->c_lb; f.c_ub = n->c_ub; f.c_undef = n->c_undef; }
        if(n->op == Op::BoolConst) { f.b_def = n->bval_def; f.b_poss = n->bval_poss; }
        if(!n->bounded) throw std::logic_error("unbounded node after lowering");
        if(is_pred_op(n->op)) { f.plb = n->plb; f.pub = n->pub; f.lb = 0; f.ub = 0; }
        else { f.lb = n->lb; f.ub = n->ub; }
        int flags = (f.c_undef ? 1 : 0) | (f.b_def ? 2 : 0) | (f.b_poss ? 4 : 0) | (f.plb ? 8 : 0) | (f.pub ? 16 : 0);
        auto key = std::make_tuple(static_cast<int>(f.op), f.kids, f.arg, bits(f.c_lb), bits(f.c_ub), flags, f.lb, f.ub, 0);
        auto jt = ids.find(key);
        int id;
        if(jt != ids.end()) id = jt->second;
        else {
            id = static_cast<int>(prog.nodes.size());
            prog.nodes.push_back(f);
            ids.emplace(key, id);
        }
        memo[n.get()] = id;
        return id;
    }
    static flat::FOp to_fop(Op o) {
        using flat::FOp;
        switch(o) {
            case Op::Const: return FOp::Const; case Op::Arg: return FOp::Arg; case Op::Neg: return FOp::Neg;
            case Op::Add: return FOp::Add; case Op::Sub: return FOp::Sub; case Op::Mul: return FOp::Mul;
            case Op::Div: return FOp::Div; case Op::Sqrt: return FOp::Sqrt; case Op::Square: return FOp::Square;
            case Op::Max: return FOp::Max; case Op::Min: return FOp::Min; case Op::Ite: return FOp::Ite;
            case Op::Acos: return FOp::Acos; case Op::BoolConst: return FOp::BoolConst; case Op::Lt: return FOp::Lt;
            case Op::Le: return FOp::Le; case Op::Gt: return FOp::Gt; case Op::Ge: return FOp::Ge;
            case Op::Eq: return FOp::Eq; case Op::Neq: return FOp::Neq; case Op::Not: return FOp::Not;
            // sequencing and non-sequencing junctors evaluate to the same IBool
            case Op::OrSeq: case Op::Or: return FOp::Or;
            case Op::AndSeq: case Op::And: return FOp::And;
        }
        throw std::logic_error("unknown op");
    }
};

} // namespace

flat::Program lower_constraint_system(const ConstraintSystem& cs, const std::vector<std::uint64_t>& splitting,
                                      const std::string& name)
{
    const int N = static_cast<int>(cs.num_args());
    if(splitting.size() != static_cast<std::size_t>(N))
        throw std::invalid_argument("wrong number of elements in the variable splitting");
    std::vector<Bounds> ab = cs.ct_bounds;

    // constant folding (transform_for_cuda)
    Folder folder;
    std::vector<NodeP> cons;
    for(const auto& c : cs.constraints) cons.push_back(folder.go(c));

    // 5*N iterations of bound-and-simplify plus compile-time bound tightening
    const int iterations = 5 * N;
    for(int it = 0; it < iterations; ++it) {
        BoundAndSimplify bs{ab, {}};
        for(auto& c : cons) c = bs.go(c);
        std::vector<Bounds> nb = ab;
        for(int t = 0; t < N; ++t) {
            for(const auto& c : cons) {
                for(const Rewrite& rw : rewrite(t, c)) {
                    Bounds r = ct_bounds(rw.expr, ab);
                    auto& tb = nb[static_cast<std::size_t>(t)];
                    if(static_cast<unsigned>(rw.dir) & static_cast<unsigned>(flat::Dir::GEQ)) tb.lb = std::max(tb.lb, r.lb);
                    if(static_cast<unsigned>(rw.dir) & static_cast<unsigned>(flat::Dir::LEQ)) tb.ub = std::min(tb.ub, r.ub);
                }
            }
            if(nb[static_cast<std::size_t>(t)].lb > nb[static_cast<std::size_t>(t)].ub)
                throw std::logic_error("empty compile-time bounds (proof complete at compile time)");
        }
        ab = nb;
    }
    {
        BoundAndSimplify bs{ab, {}};
        for(auto& c : cons) c = bs.go(c);
        std::vector<NodeP> kept;
        for(const auto& c : cons) if(!c->plb) kept.push_back(c);
        cons.swap(kept);
    }

    flat::Program p;
    p.name = name;
    p.nargs = N;
    for(int i = 0; i < N; ++i) p.arg_names.push_back(cs.arg_name(static_cast<std::size_t>(i)));

    // split policy: the first variable is always dynamic; dynamic ones form a prefix
    bool seen_static = false;
    p.initial_queue_size = 1;
    for(int i = 0; i < N; ++i) {
        std::uint64_t s = splitting[static_cast<std::size_t>(i)];
        if(i == 0 && !is_dynamic_subdivision(s)) s = dynamic_subdivision(static_cast<std::uint32_t>(s), 4);
        flat::FSplit sp;
        sp.arg = i;
        sp.dynamic = is_dynamic_subdivision(s);
        sp.initial = static_cast<std::uint32_t>(s >> 32u);
        sp.subs = static_cast<std::uint32_t>(s & 0xffffffffu);
        if(sp.dynamic) {
            if(seen_static) throw std::invalid_argument("dynamically split variables must be a prefix");
            p.initial_queue_size *= sp.initial;
        } else {
            seen_static = true;
        }
        p.splits.push_back(sp);
    }

    Flattener fl{p, {}, {}};
    // RBT in argument-major order
    for(int t = 0; t < N; ++t) {
        for(const auto& c : cons) {
            for(const Rewrite& rw : rewrite(t, c)) {
                flat::FBound b;
                b.target = t;
                b.dir = rw.dir;
                b.expr = fl.go(rw.expr);
                b.honor_undef = !(fpb::is_lb(rw.expr->lb) || fpb::is_ub(rw.expr->ub));
                p.bounds.push_back(b);
                (void)b;
            }
        }
    }
    // dependency lists (NeedsUpdate / UpdateBounds)
    std::vector<NodeP> bound_exprs;
    for(int t = 0; t < N; ++t)
        for(const auto& c : cons)
            for(const Rewrite& rw : rewrite(t, c)) bound_exprs.push_back(rw.expr);
    p.update.assign(static_cast<std::size_t>(N), std::vector<std::vector<int>>(3));
    for(int a = 0; a < N; ++a) {
        for(std::size_t k = 0; k < p.bounds.size(); ++k) {
            Deps d = deps(bound_exprs[k], a);
            unsigned dir = static_cast<unsigned>(p.bounds[k].dir);
            bool need_lb = dir & static_cast<unsigned>(flat::Dir::GEQ);
            bool need_ub = dir & static_cast<unsigned>(flat::Dir::LEQ);
            bool on_lb = (need_lb && d.ll) || (need_ub && d.ul);
            bool on_ub = (need_lb && d.lu) || (need_ub && d.uu);
            auto& u = p.update[static_cast<std::size_t>(a)];
            if(on_lb) u[0].push_back(static_cast<int>(k));
            if(on_ub) u[1].push_back(static_cast<int>(k));
            if(on_lb || on_ub) u[2].push_back(static_cast<int>(k));
        }
    }
    // RCT rows
    p.rows.assign(static_cast<std::size_t>(N), {});
    {
        std::vector<int> ids;   // flattened in constraint order
        for(const auto& c : cons) ids.push_back(fl.go(c));
        const auto rows = passes::constraint_rows(cons, N);
        for(std::size_t a = 0; a < rows.size(); ++a)
            for(int i : rows[a]) p.rows[a].push_back(ids[static_cast<std::size_t>(i)]);
    }
    // compile-time box and its outward double enclosure
    for(int a = 0; a < N; ++a) {
        const Bounds& b = ab[static_cast<std::size_t>(a)];
        p.ct_lb.push_back(b.lb);
        p.ct_ub.push_back(b.ub);
        double lo = -std::numeric_limits<double>::infinity(), hi = std::numeric_limits<double>::infinity();
        if(fpb::is_lb(b.lb)) lo = exact::outward_double(exact::Rat(b.lb) / exact::Rat(fpb::denom())).first;
        if(fpb::is_ub(b.ub)) hi = exact::outward_double(exact::Rat(b.ub) / exact::Rat(fpb::denom())).second;
        p.box_lb.push_back(lo);
        p.box_ub.push_back(hi);
    }
    return p;
}

namespace passes {

NodeP bound_and_simplify(const NodeP& n, const std::vector<fixed_point_bounds::Bounds>& arg_bounds) {
    BoundAndSimplify bs{arg_bounds, {}};
    return bs.go(n);
}

std::vector<BoundRewrite> rewrite_to_bounds(int target, const NodeP& constraint) {
    std::vector<BoundRewrite> out;
    for(const Rewrite& r : rewrite(target, constraint)) out.push_back(BoundRewrite{r.target, r.dir, r.expr});
    return out;
}

BoundDependencies bound_dependencies(const NodeP& bounded_expr, int arg) {
    const Deps d = deps(bounded_expr, arg);
    return BoundDependencies{d.ll, d.lu, d.ul, d.uu};
}

std::vector<std::vector<int>> constraint_rows(const std::vector<NodeP>& constraints, int nargs) {
    std::vector<std::vector<int>> rows(static_cast<std::size_t>(nargs));
    for(std::size_t i = 0; i < constraints.size(); ++i) {
        const NodeP& c = constraints[i];
        const int na = num_args(c);
        bool unrewritten = true;
        for(int a = 0; a < nargs && unrewritten; ++a) if(!rewrite(a, c).empty()) unrewritten = false;
        if(na >= 1 && na < nargs && unrewritten) rows[static_cast<std::size_t>(na - 1)].push_back(static_cast<int>(i));
        rows[static_cast<std::size_t>(nargs - 1)].push_back(static_cast<int>(i));
    }
    return rows;
}

} // namespace passes

} // namespace ivarp
