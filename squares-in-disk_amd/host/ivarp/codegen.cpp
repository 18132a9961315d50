// Emits the HIP program struct of one lowered program (see csrc/core.hpp for
// the interface).  Compile-time decisions that the reference takes from its
// fixed-point bounds (operand signs, finiteness, whether an undefined bound value
// disables a bound) are resolved here and become template arguments or are
// dropped from the generated code.
#include "codegen.hpp"
#include "fixed_point.hpp"
#include <algorithm>
#include <sstream>
#include <stdexcept>
#include <functional>
#include <map>
#include <set>
#include <utility>
#include <vector>
#include <cmath>
#include <cstdio>

namespace ivarp {
namespace codegen {
namespace {
using flat::FOp;
using flat::Program;
namespace fpb = fixed_point_bounds;

std::string dlit(double d) {
    if(d == __builtin_huge_val()) return "__builtin_huge_val()";
    if(d == -__builtin_huge_val()) return "(-__builtin_huge_val())";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", d);
    return buf;
}
// The largest float <= d (up: the smallest float >= d), as the kernels'
// outward conversion gives it; the conversion here rounds to nearest and the
// comparison with the exact widening corrects it, zero signs included.
float f32_outward(double d, bool up) {
    float f = static_cast<float>(d);
    if(!up && static_cast<double>(f) > d) f = std::nextafter(f, -__builtin_huge_valf());
    if(up && static_cast<double>(f) < d) f = std::nextafter(f, __builtin_huge_valf());
    return f;
}
std::string flit(float f) {
    if(f == __builtin_huge_valf()) return "__builtin_huge_valf()";
    if(f == -__builtin_huge_valf()) return "(-__builtin_huge_valf())";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%af", static_cast<double>(f));
    return buf;
}
int tri_nn(const flat::FNode& b) { return b.lb >= 0 ? 1 : (b.ub <= 0 ? 0 : 2); }
int tri_np(const flat::FNode& b) { return b.ub <= 0 ? 1 : (b.lb >= 0 ? 0 : 2); }
int tri_pos(const flat::FNode& b) { return b.lb > 0 ? 1 : (b.ub <= 0 ? 0 : 2); }
int tri_neg(const flat::FNode& b) { return b.ub < 0 ? 1 : (b.lb >= 0 ? 0 : 2); }
bool ct_finite(const flat::FNode& b) { return b.lb > fpb::min_bound() && b.ub < fpb::max_bound(); }

struct Emitter {
    const Program& p;
    std::map<int, int> const_index;
    std::vector<double> consts;

    int cidx(int node) {
        auto it = const_index.find(node);
        if(it != const_index.end()) return it->second;
        int i = static_cast<int>(consts.size() / 2);
        consts.push_back(p.nodes[static_cast<std::size_t>(node)].c_lb);
        consts.push_back(p.nodes[static_cast<std::size_t>(node)].c_ub);
        const_index[node] = i;
        return i;
    }
    // nodes whose values are read from the hoisted-term struct `h` instead of
    // being computed (set while emitting a level's hoisted row check)
    std::set<int> subst;

    void collect(int id, std::vector<int>& order, std::vector<char>& seen) {
        if(seen[static_cast<std::size_t>(id)]) return;
        seen[static_cast<std::size_t>(id)] = 1;
        if(!subst.count(id))
            for(int k : p.nodes[static_cast<std::size_t>(id)].kids) collect(k, order, seen);
        order.push_back(id);
    }
    bool is_pred(int id) const {
        switch(p.nodes[static_cast<std::size_t>(id)].op) {
            case FOp::BoolConst: case FOp::Lt: case FOp::Le: case FOp::Gt: case FOp::Ge: case FOp::Eq: case FOp::Neq:
            case FOp::Not: case FOp::Or: case FOp::And: return true;
            default: return false;
        }
    }
    std::string v(int id) { return "v" + std::to_string(id); }
    std::string b(int id) { return "p" + std::to_string(id); }
    // Evaluation of the sub-DAG below `root`, lazy like the reference's
    // evaluation: an if_then_else evaluates its then-branch only where the
    // condition is possibly true and its else-branch only where it is possibly
    // false (math_fn/eval/if_then_else.hpp:64-82), and || / && stop at the first
    // definitely true / not possibly true operand (predicate_eval/n_ary.hpp:63-135).
    // Every node is computed once, in the innermost scope that contains all of
    // its uses: the common prefix of the (conditional node, operand) paths from
    // the root to it.  Results are those of the full evaluation, since all nodes
    // are pure.
    using Ctx = std::vector<std::pair<int, int>>;
    std::string eval(int root, const std::string& indent) { return eval(std::vector<int>{root}, indent); }
    // The same for several roots, all at the top scope.
    std::string eval(const std::vector<int>& roots, const std::string& indent) {
        std::vector<int> order;
        std::vector<char> seen(p.nodes.size(), 0);
        for(int r : roots) collect(r, order, seen);
        std::map<int, Ctx> ctx;
        for(int r : roots) ctx[r] = {};
        for(auto it = order.rbegin(); it != order.rend(); ++it) {
            const int id = *it;
            const flat::FNode& n = p.nodes[static_cast<std::size_t>(id)];
            const Ctx& c = ctx.at(id);
            const bool lazy = n.op == FOp::Ite || n.op == FOp::Or || n.op == FOp::And;
            if(subst.count(id)) continue;
            for(std::size_t i = 0; i < n.kids.size(); ++i) {
                Ctx ck = c;
                if(lazy && i > 0) ck.push_back({id, static_cast<int>(i)});
                const int k = n.kids[i];
                auto f = ctx.find(k);
                if(f == ctx.end()) { ctx[k] = ck; continue; }
                Ctx& cur = f->second;
                std::size_t m = 0;
                while(m < cur.size() && m < ck.size() && cur[m] == ck[m]) ++m;
                cur.resize(m);
            }
        }
        std::ostringstream o;
        emit_scope(order, ctx, {}, indent, o);
        return o.str();
    }
    // Weighted size of the code emitted in `scope` (and the scopes nested in it):
    // roughly the instructions it costs; square roots, divisions and acos weigh
    // more than the rest.
    int scope_cost(const std::vector<int>& order, const std::map<int, Ctx>& ctx, const Ctx& scope) const {
        int cost = 0;
        for(int id : order) {
            const Ctx& c = ctx.at(id);
            if(c.size() < scope.size() || !std::equal(scope.begin(), scope.end(), c.begin())) continue;
            if(subst.count(id)) continue;
            switch(p.nodes[static_cast<std::size_t>(id)].op) {
                case FOp::Const: case FOp::Arg: case FOp::BoolConst: break;
                case FOp::Sqrt: cost += 12; break;
                case FOp::Div: cost += 16; break;
                case FOp::Acos: cost += 200; break;
                case FOp::Mul: case FOp::Square: cost += 3; break;
                default: cost += 1; break;
            }
        }
        return cost;
    }
    // Lazy nodes whose skipped operands cost less than this are evaluated in
    // full instead: where the lanes of a wave disagree on the condition both
    // branches run anyway, and the branch itself (exec-mask bookkeeping) costs
    // more than a few interval operations.  All terms are pure, so the value is
    // the same either way.
    static constexpr int kEagerCost = 12;

    void emit_scope(const std::vector<int>& order, const std::map<int, Ctx>& ctx, const Ctx& scope,
                    const std::string& indent, std::ostringstream& o) {
        for(int id : order) {
            if(ctx.at(id) != scope) continue;
            const flat::FNode& n = p.nodes[static_cast<std::size_t>(id)];
            auto sub = [&](int i) { Ctx c = scope; c.push_back({id, i}); return c; };
            if(subst.count(id)) {
                if(is_pred(id)) o << indent << "const IB " << b(id) << " = h." << b(id) << ";\n";
                else o << indent << "const IV " << v(id) << " = h." << v(id) << ";\n";
                continue;
            }
            if(n.op == FOp::Ite && scope_cost(order, ctx, sub(1)) + scope_cost(order, ctx, sub(2)) < kEagerCost) {
                emit_scope(order, ctx, sub(1), indent, o);
                emit_scope(order, ctx, sub(2), indent, o);
                o << indent << "const IV " << v(id) << " = iv_ite(" << b(n.kids[0]) << ", " << v(n.kids[1]) << ", "
                  << v(n.kids[2]) << ");\n";
                continue;
            }
            if(n.op == FOp::Ite) {
                const std::string c = b(n.kids[0]), t = "t" + std::to_string(id), e = "e" + std::to_string(id);
                o << indent << "IV " << t << "{}, " << e << "{};\n";
                o << indent << "if(" << c << ".poss) {\n";
                emit_scope(order, ctx, sub(1), indent + "    ", o);
                o << indent << "    " << t << " = " << v(n.kids[1]) << ";\n" << indent << "}\n";
                o << indent << "if(!" << c << ".def) {\n";
                emit_scope(order, ctx, sub(2), indent + "    ", o);
                o << indent << "    " << e << " = " << v(n.kids[2]) << ";\n" << indent << "}\n";
                o << indent << "const IV " << v(id) << " = iv_ite(" << c << ", " << t << ", " << e << ");\n";
                continue;
            }
            if(n.op == FOp::Or || n.op == FOp::And) {
                const bool is_or = n.op == FOp::Or;
                int skipped = 0;
                for(std::size_t i = 1; i < n.kids.size(); ++i) skipped += scope_cost(order, ctx, sub(static_cast<int>(i)));
                if(skipped < kEagerCost) {
                    for(std::size_t i = 1; i < n.kids.size(); ++i) emit_scope(order, ctx, sub(static_cast<int>(i)), indent, o);
                    o << indent << "IB " << b(id) << " = " << b(n.kids[0]) << ";\n";
                    for(std::size_t i = 1; i < n.kids.size(); ++i)
                        o << indent << b(id) << " = " << (is_or ? "ib_or(" : "ib_and(") << b(id) << ", " << b(n.kids[i]) << ");\n";
                    continue;
                }
                o << indent << "IB " << b(id) << " = " << b(n.kids[0]) << ";\n";
                for(std::size_t i = 1; i < n.kids.size(); ++i) {
                    o << indent << "if(" << (is_or ? "!" : "") << b(id) << (is_or ? ".def" : ".poss") << ") {\n";
                    emit_scope(order, ctx, sub(static_cast<int>(i)), indent + "    ", o);
                    o << indent << "    " << b(id) << " = " << (is_or ? "ib_or(" : "ib_and(") << b(id) << ", "
                      << b(n.kids[i]) << ");\n" << indent << "}\n";
                }
                continue;
            }
            o << indent << node_code(id) << "\n";
        }
    }
    std::string node_code(int id) {
        const flat::FNode& n = p.nodes[static_cast<std::size_t>(id)];
        auto K = [&](std::size_t i) { return n.kids[i]; };
        auto KN = [&](std::size_t i) -> const flat::FNode& { return p.nodes[static_cast<std::size_t>(n.kids[i])]; };
        // Binary64 constants are literals, and the compiler folds operations on
        // literals alone in round-to-nearest, not in the directed rounding the
        // kernels run in: such terms must have been folded exactly by lowering.
        if(n.op != FOp::Const && n.op != FOp::Arg && !n.kids.empty()) {
            bool all_const = true;
            for(std::size_t i = 0; i < n.kids.size(); ++i) all_const = all_const && KN(i).op == FOp::Const;
            if(all_const)
                throw std::runtime_error("codegen: node " + std::to_string(id) + " operates on constants only");
        }
        std::ostringstream o;
        switch(n.op) {
            case FOp::Const: {
                int c = cidx(id);
                o << "const IV " << v(id) << "{kc<Num, " << 2 * c << ">(C), kc<Num, " << 2 * c + 1 << ">(C), "
                  << (n.c_undef ? "true" : "false") << "};";
                break;
            }
            case FOp::Arg: o << "const IV " << v(id) << "{lb[" << n.arg << "], ub[" << n.arg << "], false};"; break;
            case FOp::Neg: o << "const IV " << v(id) << " = iv_neg(" << v(K(0)) << ");"; break;
            case FOp::Add: o << "const IV " << v(id) << " = iv_add(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Sub: o << "const IV " << v(id) << " = iv_sub(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Mul:
                o << "const IV " << v(id) << " = iv_mul<" << tri_nn(KN(0)) << ", " << tri_np(KN(0)) << ", "
                  << tri_nn(KN(1)) << ", " << tri_np(KN(1)) << ", "
                  << ((ct_finite(KN(0)) && ct_finite(KN(1))) ? "true" : "false") << ">(" << v(K(0)) << ", " << v(K(1)) << ");";
                break;
            case FOp::Div:
                o << "const IV " << v(id) << " = iv_div<" << tri_nn(KN(0)) << ", " << tri_np(KN(0)) << ", "
                  << tri_pos(KN(1)) << ", " << tri_neg(KN(1)) << ", "
                  << ((ct_finite(KN(0)) && ct_finite(KN(1))) ? "true" : "false") << ">(" << v(K(0)) << ", " << v(K(1)) << ");";
                break;
            case FOp::Sqrt: o << "const IV " << v(id) << " = iv_sqrt<" << tri_nn(KN(0)) << ">(" << v(K(0)) << ");"; break;
            case FOp::Square: o << "const IV " << v(id) << " = iv_square(" << v(K(0)) << ");"; break;
            case FOp::Max: case FOp::Min: {
                const char* f = n.op == FOp::Max ? "iv_max" : "iv_min";
                std::string acc = v(K(n.kids.size() - 1));
                for(std::size_t i = n.kids.size() - 1; i-- > 0;) acc = std::string(f) + "(" + v(K(i)) + ", " + acc + ")";
                o << "const IV " << v(id) << " = " << acc << ";";
                break;
            }
            case FOp::Ite: case FOp::Or: case FOp::And: throw std::logic_error("lazy node emitted as plain code");
            case FOp::Acos: {
                // domain checks only where the compile-time bounds leave [-1, 1]
                const bool chk_lb = KN(0).lb < fpb::int_to_fp(-1), chk_ub = KN(0).ub > fpb::int_to_fp(1);
                o << "const IV " << v(id) << " = iv_acos<" << (chk_lb ? "true" : "false") << ", "
                  << (chk_ub ? "true" : "false") << ">(" << v(K(0)) << ");";
                break;
            }
            case FOp::BoolConst: o << "const IB " << b(id) << "{" << (n.b_def ? "true" : "false") << ", " << (n.b_poss ? "true" : "false") << "};"; break;
            case FOp::Lt: o << "const IB " << b(id) << " = ib_lt(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Le: o << "const IB " << b(id) << " = ib_le(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Gt: o << "const IB " << b(id) << " = ib_gt(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Ge: o << "const IB " << b(id) << " = ib_ge(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Eq: o << "const IB " << b(id) << " = ib_eq(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Neq: o << "const IB " << b(id) << " = ib_neq(" << v(K(0)) << ", " << v(K(1)) << ");"; break;
            case FOp::Not: o << "const IB " << b(id) << " = ib_not(" << b(K(0)) << ");"; break;
        }
        return o.str();
    }
};

} // namespace

bool supported(const flat::Program&) { return true; }

std::string generate(const flat::Program& p, const std::string& key) {
    Emitter e{p, {}, {}};
    const std::string ns = "sqd_gen_" + key;
    std::vector<int> dyn, sta;
    for(std::size_t i = 0; i < p.splits.size(); ++i) (p.splits[i].dynamic ? dyn : sta).push_back(static_cast<int>(i));
    std::ostringstream body;

    // constraint predicates, one function per constraint root
    std::map<int, bool> roots;
    for(const auto& r : p.rows) for(int c : r) roots[c] = true;
    for(const auto& kv : roots) {
        body << "    template<class Num> __device__ __forceinline__ static IB c" << kv.first
             << "(const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
        body << "        using IV = IVT<Num>;\n        (void)lb; (void)ub; (void)C;\n";
        body << e.eval(kv.first, "        ");
        body << "        return " << e.b(kv.first) << ";\n    }\n";
    }
    // per constraint: the interval of (left - right) of a top-level comparison,
    // for the FP32-versus-FP64 width statistics (other roots: an empty interval)
    std::vector<int> cons;
    for(const auto& kv : roots) cons.push_back(kv.first);
    for(int c : cons) {
        const flat::FNode& n = p.nodes[static_cast<std::size_t>(c)];
        const bool cmp = n.op == FOp::Lt || n.op == FOp::Le || n.op == FOp::Gt || n.op == FOp::Ge;
        body << "    template<class Num> __device__ static IVT<Num> d" << c
             << "(const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
        body << "        using IV = IVT<Num>;\n        (void)lb; (void)ub; (void)C;\n";
        if(cmp) {
            body << e.eval(c, "        ");
            body << "        (void)" << e.b(c) << ";\n";
            body << "        return iv_sub(" << e.v(n.kids[0]) << ", " << e.v(n.kids[1]) << ");\n    }\n";
        } else {
            body << "        return IV{Num(1), Num(0), true};\n    }\n";
        }
    }
    body << "    static constexpr int NCONS = " << cons.size() << ";\n";
    body << "    static constexpr int cons_id(int j) { switch(j) {";
    for(std::size_t j = 0; j < cons.size(); ++j) body << " case " << j << ": return " << cons[j] << ";";
    body << " default: return -1; } }\n";
    body << "    template<class Num> __device__ static IVT<Num> cons_diff(int j, const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
    body << "        switch(j) {\n";
    for(std::size_t j = 0; j < cons.size(); ++j) body << "        case " << j << ": return d" << cons[j] << "(lb, ub, C);\n";
    body << "        default: return IVT<Num>{Num(1), Num(0), true};\n        }\n    }\n";
    // values for the enclosure checks: the interval of bound k's expression and
    // the IBool of constraint j (in cons_id order) on a box, evaluated exactly as
    // apply and the row checks evaluate them
    body << "    template<class Num> __device__ static IVT<Num> bound_value(int k, const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
    body << "        using IV = IVT<Num>;\n        (void)lb; (void)ub; (void)C;\n        switch(k) {\n";
    for(std::size_t k = 0; k < p.bounds.size(); ++k) {
        body << "        case " << k << ": {\n" << e.eval(p.bounds[k].expr, "            ");
        body << "            return " << e.v(p.bounds[k].expr) << ";\n        }\n";
    }
    body << "        default: return IV{Num(1), Num(0), true};\n        }\n    }\n";
    body << "    template<class Num> __device__ static IB cons_value(int j, const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
    body << "        switch(j) {\n";
    for(std::size_t j = 0; j < cons.size(); ++j) body << "        case " << j << ": return c" << cons[j] << "(lb, ub, C);\n";
    body << "        default: return IB{false, true};\n        }\n    }\n";
    // bound application
    body << "    template<class Num> __device__ __forceinline__ static int apply(int k, Num (&lb)[N], Num (&ub)[N], const Num* C) {\n";
    body << "        using IV = IVT<Num>;\n        (void)lb; (void)ub; (void)C;\n        switch(k) {\n";
    for(std::size_t k = 0; k < p.bounds.size(); ++k) {
        const flat::FBound& bd = p.bounds[k];
        const int t = bd.target;
        body << "        case " << k << ": {\n" << e.eval(bd.expr, "            ");
        const std::string f = e.v(bd.expr);
        if(bd.honor_undef) body << "            if(" << f << ".und) return 0;\n";
        if(bd.dir == flat::Dir::LEQ) {
            body << "            if(" << f << ".ub < ub[" << t << "]) { ub[" << t << "] = " << f << ".ub; return " << f
                 << ".ub < lb[" << t << "] ? -1 : " << 1 + t * 3 + 1 << "; }\n            return 0;\n";
        } else if(bd.dir == flat::Dir::GEQ) {
            body << "            if(" << f << ".lb > lb[" << t << "]) { lb[" << t << "] = " << f << ".lb; return " << f
                 << ".lb > ub[" << t << "] ? -1 : " << 1 + t * 3 + 0 << "; }\n            return 0;\n";
        } else {
            body << "            int ch = 0;\n"
                 << "            if(" << f << ".ub < ub[" << t << "]) { ub[" << t << "] = " << f << ".ub; ch |= 2; }\n"
                 << "            if(" << f << ".lb > lb[" << t << "]) { lb[" << t << "] = " << f << ".lb; ch |= 1; }\n"
                 << "            if(ub[" << t << "] < lb[" << t << "]) return -1;\n"
                 << "            return ch ? " << 1 + t * 3 << " + (ch - 1) : 0;\n";
        }
        body << "        }\n";
    }
    body << "        default: return 0;\n        }\n    }\n";
    // dependency lists
    body << "    template<class Fn> __device__ static void deps(int list, Fn&& f) {\n        switch(list) {\n";
    for(int a = 0; a < p.nargs; ++a) {
        for(int w = 0; w < 3; ++w) {
            const auto& l = p.update[static_cast<std::size_t>(a)][static_cast<std::size_t>(w)];
            if(l.empty()) continue;
            // a bound listed twice is queued once (the queue skips members), so
            // only its first occurrence is emitted: propagate() applies the
            // initial list in order before it builds the queue
            std::vector<int> seen_once;
            body << "        case " << a * 3 + w << ":";
            for(int x : l) {
                if(std::find(seen_once.begin(), seen_once.end(), x) != seen_once.end()) continue;
                seen_once.push_back(x);
                body << " f(" << x << ");";
            }
            body << " break;\n";
        }
    }
    body << "        default: break;\n        }\n    }\n";
    // constraint rows per static level, tested in row order (cpu_prover_core.hpp:179-191)
    body << "    template<int L, class Num> __device__ __forceinline__ static bool rows_violated(const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
    body << "        (void)lb; (void)ub; (void)C;\n";
    for(std::size_t L = 0; L < sta.size(); ++L) {
        int from = p.splits[static_cast<std::size_t>(sta[L])].arg;
        int to = L + 1 < sta.size() ? p.splits[static_cast<std::size_t>(sta[L + 1])].arg : p.nargs;
        std::vector<int> level;
        for(int r = from; r < to; ++r)
            for(int c : p.rows[static_cast<std::size_t>(r)]) level.push_back(c);
        body << "        if constexpr(L == " << L << ") {\n";
        for(int c : level) body << "            if(!c" << c << "(lb, ub, C).poss) return true;\n";
        body << "        }\n";
    }
    body << "        return false;\n    }\n";
    body << "    template<class Num> __device__ __forceinline__ static bool all_satisfied(const Num (&lb)[N], const Num (&ub)[N], const Num* C) {\n";
    for(int c : p.rows[static_cast<std::size_t>(p.nargs - 1)]) body << "        if(!c" << c << "(lb, ub, C).poss) return false;\n";
    body << "        (void)lb; (void)ub; (void)C;\n        return true;\n    }\n";

    // Hoisted row checks for the FP32 prefilter.  At static level L the rows
    // are checked on a box whose variables other than the split variable A are
    // the parent's for every child, so each term free of A has the same value
    // for all children of a parent: hoist<L> computes the terms the other terms
    // read (the frontier) once per parent into Hoist<L>, and
    // rows_violated_hoisted<L> evaluates the rest per child, with the same
    // results as rows_violated<L> on that box.
    std::ostringstream hoist_types, hoist_fns;
    hoist_fns << "    template<int L, class Num> __device__ __forceinline__ static void hoist(const Num (&lb)[N], const Num (&ub)[N], "
                 "const Num* C, Hoist<L, Num>& h) {\n        using IV = IVT<Num>;\n        (void)lb; (void)ub; (void)C; (void)h;\n";
    std::ostringstream rows_h;
    rows_h << "    template<int L, class Num> __device__ __forceinline__ static bool rows_violated_hoisted(const Num (&lb)[N], "
              "const Num (&ub)[N], const Num* C, const Hoist<L, Num>& h) {\n        using IV = IVT<Num>;\n"
              "        (void)lb; (void)ub; (void)C; (void)h;\n";
    std::string hoist_sel = "NoHoist";
    for(std::size_t L = sta.size(); L-- > 0;) {
        const int A = p.splits[static_cast<std::size_t>(sta[L])].arg;
        int from = A;
        int to = L + 1 < sta.size() ? p.splits[static_cast<std::size_t>(sta[L + 1])].arg : p.nargs;
        std::vector<int> level;
        for(int r = from; r < to; ++r)
            for(int c : p.rows[static_cast<std::size_t>(r)]) level.push_back(c);
        // dependence on A
        std::vector<int> dep(p.nodes.size(), -1);
        std::function<bool(int)> depends = [&](int id) -> bool {
            int& d = dep[static_cast<std::size_t>(id)];
            if(d >= 0) return d != 0;
            const flat::FNode& n = p.nodes[static_cast<std::size_t>(id)];
            bool r = n.op == FOp::Arg && n.arg == A;
            for(int k : n.kids) r = depends(k) || r;
            d = r ? 1 : 0;
            return r;
        };
        // frontier: A-free terms (not constants or variables) read by terms that depend on A, or whole A-free constraints
        std::set<int> frontier;
        std::set<int> visited;
        std::function<void(int)> walk = [&](int id) {
            if(!visited.insert(id).second) return;
            const flat::FNode& n = p.nodes[static_cast<std::size_t>(id)];
            for(int k : n.kids) {
                const flat::FNode& kn = p.nodes[static_cast<std::size_t>(k)];
                if(!depends(k)) {
                    if(kn.op != FOp::Const && kn.op != FOp::Arg && kn.op != FOp::BoolConst) frontier.insert(k);
                } else {
                    walk(k);
                }
            }
        };
        for(int c : level) {
            if(depends(c)) walk(c);
            else if(p.nodes[static_cast<std::size_t>(c)].op != FOp::BoolConst) frontier.insert(c);
        }
        const std::string hname = "HoistL" + std::to_string(L);
        hoist_types << "template<class Num> struct " << hname << " {";
        for(int f : frontier) {
            if(e.is_pred(f)) hoist_types << " IB " << e.b(f) << ";";
            else hoist_types << " IVT<Num> " << e.v(f) << ";";
        }
        hoist_types << " };\n";
        hoist_sel = "std::conditional_t<L == " + std::to_string(L) + ", " + hname + "<Num>, " + hoist_sel + ">";
        // hoist<L>: the frontier terms
        hoist_fns << "        if constexpr(L == " << L << ") {\n";
        if(!frontier.empty()) {
            std::vector<int> roots(frontier.begin(), frontier.end());
            hoist_fns << e.eval(roots, "            ");
            for(int f : frontier)
                hoist_fns << "            h." << (e.is_pred(f) ? e.b(f) : e.v(f)) << " = " << (e.is_pred(f) ? e.b(f) : e.v(f)) << ";\n";
        }
        hoist_fns << "        }\n";
        // rows_violated_hoisted<L>
        rows_h << "        if constexpr(L == " << L << ") {\n";
        e.subst = frontier;
        for(int c : level) {
            if(frontier.count(c)) {
                rows_h << "            if(!h." << e.b(c) << ".poss) return true;\n";
            } else {
                rows_h << "            {\n" << e.eval(c, "                ");
                rows_h << "                if(!" << e.b(c) << ".poss) return true;\n            }\n";
            }
        }
        e.subst.clear();
        rows_h << "        }\n";
    }
    hoist_fns << "    }\n";
    rows_h << "        return false;\n    }\n";
    body << "    template<int L, class Num> using Hoist = " << hoist_sel << ";\n";
    body << hoist_fns.str() << rows_h.str();

    std::ostringstream o;
    o << "// Generated from the lowered program '" << key << "' (" << p.name << "); do not edit.\n";
    o << "#include \"kernel_set.hpp\"\n\n";
    o << "namespace " << ns << " {\nusing namespace sqd::dev;\n\n";
    o << hoist_types.str() << "\n";
    o << "struct P {\n";
    o << "    static constexpr int N = " << p.nargs << ";\n";
    o << "    static constexpr int NB = " << p.bounds.size() << ";\n";
    o << "    static constexpr int NC = " << e.consts.size() / 2 << ";\n";
    o << "    static constexpr int NDYN = " << dyn.size() << ";\n";
    o << "    static constexpr int NSTAT = " << sta.size() << ";\n";
    // constants as literals (folded into the instructions that use them): the
    // binary64 intervals, and the same converted outward to binary32
    o << "    static constexpr double K[] = {";
    for(std::size_t i = 0; i < e.consts.size(); ++i) o << (i ? ", " : "") << dlit(e.consts[i]);
    if(e.consts.empty()) o << "0.0, 0.0";
    o << "};\n";
    o << "    static constexpr float K32[] = {";
    for(std::size_t i = 0; i < e.consts.size(); ++i) o << (i ? ", " : "") << flit(f32_outward(e.consts[i], i & 1));
    if(e.consts.empty()) o << "0.0f, 0.0f";
    o << "};\n";
    o << "    template<class Num, int I> __device__ static __forceinline__ Num kc(const Num* C) {\n"
         "        (void)C;\n"
         "        if constexpr(sizeof(Num) == sizeof(double)) { constexpr double k = K[I]; return k; }\n"
         "        else { constexpr float k = K32[I]; return k; }\n"
         "    }\n";
    auto table = [&](const char* name, const std::vector<int>& idx, auto get) {
        o << "    static constexpr int " << name << "(int j) { switch(j) {";
        for(std::size_t j = 0; j < idx.size(); ++j) o << " case " << j << ": return " << get(p.splits[static_cast<std::size_t>(idx[j])]) << ";";
        o << " default: return 1; } }\n";
    };
    table("dyn_arg", dyn, [](const flat::FSplit& s) { return s.arg; });
    table("dyn_initial", dyn, [](const flat::FSplit& s) { return static_cast<int>(s.initial); });
    table("dyn_subs", dyn, [](const flat::FSplit& s) { return static_cast<int>(s.subs); });
    table("st_arg", sta, [](const flat::FSplit& s) { return s.arg; });
    table("st_subs", sta, [](const flat::FSplit& s) { return static_cast<int>(s.subs); });
    // Register budget of the binary64 static level kernels as resident waves
    // per SIMD (0: the compiler's choice), measured per kernel (per proof, with
    // the constraint functions inlined): One's and Two's level 1 at 4 (128
    // VGPRs, a few spilled; 3 was slower), Three's level 1 at 4 (274 against
    // 292 us), More1's level 1 at 3 (264 against 301 ms at the compiler's 176
    // VGPRs), Four's level 2 at 3 (3.46 against 3.92 ms) and More2's level 2 at
    // 3 (19.8 ms; 4 spilled 426 VGPRs and took 38.7 ms).
    int waves_l1 = 0, waves_l2 = 0;
    if(key == "one" || key == "two" || key == "three") waves_l1 = 4;
    if(key == "more1") waves_l1 = 3;
    if(key == "four" || key == "more2") waves_l2 = 3;
    o << "    static constexpr int level_waves(int l) { return l == 1 ? " << waves_l1 << " : l == 2 ? " << waves_l2
      << " : 0; }\n";
    // Whether the static levels' propagation applies the initial bounds inline
    // before it builds its queue (same results; fewer instructions, more
    // registers).  Measured per kernel: faster for every statement's dominant
    // level but Four's level 2 (193 instead of 171 VGPRs: 2 waves per SIMD
    // instead of 3, 4.72 against 4.37 ms per proof).
    int slow_prop_level = -1;
    if(key == "four") slow_prop_level = 2;
    o << "    static constexpr bool prop_fast(int l) { return l != " << slow_prop_level << "; }\n";
    o << body.str();
    o << "};\n\n";
    o << "static const double k_consts[] = {";
    for(std::size_t i = 0; i < e.consts.size(); ++i) o << (i ? ", " : "") << dlit(e.consts[i]);
    if(e.consts.empty()) o << "0.0, 0.0";
    o << "};\n\n";
    char fp[32];
    std::snprintf(fp, sizeof fp, "0x%016llxull", static_cast<unsigned long long>(p.fingerprint()));
    o << "static const sqd::KernelSet k_set = sqd::make_kernel_set<P>(" << fp << ", \"" << key << "\", k_consts);\n";
    o << "static const int k_registered = (sqd::register_kernel_set(&k_set), 0);\n";
    o << "} // namespace " << ns << "\n";
    return o.str();
}

} // namespace codegen
} // namespace ivarp
