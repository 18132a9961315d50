// Test driver of the lowering's passes (host/ivarp): reads one command per line
// on stdin and prints its result, for tests/test_lowering_kat.py.  Test
// infrastructure only.
//
//   fp <fn> <int>...                 fixed-point bound function (mul_rd, div_ru, sqrt_rd,
//                                    iv_order, iv_sqrt, pow_lb, ...): prints its result
//   bs <n> <lb ub>*n <expr>          bound_and_simplify: "E lb ub" or "P plb pub"
//   ctb <n> <lb ub>*n <expr>         compile-time bounds of the raw tree: as bs
//   dep <n> <lb ub>*n <arg> <expr>   dependencies of the bounded expression on arg:
//                                    lb_on_lb lb_on_ub ub_on_lb ub_on_ub (0/1)
//   rw <target> <pred>               bound rewriting: the count, then per bound
//                                    "dir c_lb c_ub ct_lb ct_ub" of its expression
//   rows <n> <m> <pred>*m            rows of the runtime constraint table
//   node <expr>                      op, compile-time bounds and constant value of the node
//
// Expressions are s-expressions: (arg i), (z v) an integer literal v_Z, (i v) a
// C++ int, (neg e), (add a b), (sub a b), (mul a b), (div a b), (sqrt a),
// (square a), (max a b), (min a b), (ite p a b), (lt a b), (le a b), (gt a b),
// (ge a b), (eq a b), (neq a b), (not p), (or p q), (and p q).
#include <cstdio>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "expr.hpp"
#include "fixed_point.hpp"
#include "lower_passes.hpp"

using namespace ivarp;
namespace fpb = ivarp::fixed_point_bounds;

namespace {

struct Parser {
    std::string s;
    std::size_t i = 0;
    void ws() { while(i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i; }
    std::string word() {
        ws();
        std::size_t j = i;
        while(i < s.size() && s[i] != ' ' && s[i] != ')' && s[i] != '(') ++i;
        return s.substr(j, i - j);
    }
    void expect(char c) {
        ws();
        if(i >= s.size() || s[i] != c) throw std::runtime_error(std::string("expected ") + c + " in " + s);
        ++i;
    }
    // an expression or predicate node
    NodeP node() {
        expect('(');
        const std::string op = word();
        NodeP r;
        if(op == "arg") r = Expr(make_arg(std::stoi(word()))).node;
        else if(op == "z") r = Expr(make_int_literal(word().c_str())).node;
        else if(op == "i") r = Expr(std::stoi(word())).node;
        else {
            std::vector<NodeP> k;
            for(;;) {
                ws();
                if(i < s.size() && s[i] == ')') break;
                k.push_back(node());
            }
            auto E = [&](std::size_t j) { return Expr(k.at(j)); };
            auto P = [&](std::size_t j) { return Pred(k.at(j)); };
            if(op == "neg") r = (-E(0)).node;
            else if(op == "add") r = (E(0) + E(1)).node;
            else if(op == "sub") r = (E(0) - E(1)).node;
            else if(op == "mul") r = (E(0) * E(1)).node;
            else if(op == "div") r = (E(0) / E(1)).node;
            else if(op == "sqrt") r = sqrt(E(0)).node;
            else if(op == "square") r = square(E(0)).node;
            else if(op == "max") r = maximum(E(0), E(1)).node;
            else if(op == "min") r = minimum(E(0), E(1)).node;
            else if(op == "ite") r = if_then_else(P(0), E(1), E(2)).node;
            else if(op == "lt") r = (E(0) < E(1)).node;
            else if(op == "le") r = (E(0) <= E(1)).node;
            else if(op == "gt") r = (E(0) > E(1)).node;
            else if(op == "ge") r = (E(0) >= E(1)).node;
            else if(op == "eq") r = (E(0) == E(1)).node;
            else if(op == "neq") r = (E(0) != E(1)).node;
            else if(op == "not") r = (!P(0)).node;
            else if(op == "or") r = (P(0) || P(1)).node;
            else if(op == "and") r = (P(0) && P(1)).node;
            else throw std::runtime_error("unknown operation " + op);
        }
        expect(')');
        return r;
    }
};

const char* order_name(fpb::Order o) {
    switch(o) {
        case fpb::Order::LT: return "LT";
        case fpb::Order::LE: return "LE";
        case fpb::Order::EQ: return "EQ";
        case fpb::Order::GE: return "GE";
        case fpb::Order::GT: return "GT";
        default: return "UNKNOWN";
    }
}

long long num(std::istringstream& in) {
    long long v;
    if(!(in >> v)) throw std::runtime_error("expected an integer");
    return v;
}

std::string fp(const std::string& fn, std::istringstream& in) {
    std::vector<long long> a;
    long long v;
    while(in >> v) a.push_back(v);
    auto need = [&](std::size_t n) { if(a.size() != n) throw std::runtime_error(fn + ": wrong argument count"); };
    auto one = [](long long x) { return std::to_string(x); };
    if(fn == "mul_rd") { need(2); return one(fpb::mul_rd(a[0], a[1])); }
    if(fn == "mul_ru") { need(2); return one(fpb::mul_ru(a[0], a[1])); }
    if(fn == "div_rd") { need(2); return one(fpb::div_rd(a[0], a[1])); }
    if(fn == "div_ru") { need(2); return one(fpb::div_ru(a[0], a[1])); }
    if(fn == "add_rd") { need(2); return one(fpb::add_rd(a[0], a[1])); }
    if(fn == "add_ru") { need(2); return one(fpb::add_ru(a[0], a[1])); }
    if(fn == "sqrt_rd") { need(1); return one(fpb::sqrt_rd(a[0])); }
    if(fn == "sqrt_ru") { need(1); return one(fpb::sqrt_ru(a[0])); }
    if(fn == "int_to_fp") { need(1); return one(fpb::int_to_fp(a[0])); }
    if(fn == "iv_sqrt") { need(2); auto b = fpb::iv_sqrt(a[0], a[1]); return one(b.lb) + " " + one(b.ub); }
    if(fn == "pow_lb") { need(3); return one(fpb::pow_lb(a[0], a[1], static_cast<unsigned>(a[2]))); }
    if(fn == "pow_ub") { need(3); return one(fpb::pow_ub(a[0], a[1], static_cast<unsigned>(a[2]))); }
    if(fn == "iv_mul_lb") { need(4); return one(fpb::iv_mul_lb(a[0], a[1], a[2], a[3])); }
    if(fn == "iv_mul_ub") { need(4); return one(fpb::iv_mul_ub(a[0], a[1], a[2], a[3])); }
    if(fn == "iv_div_lb") { need(4); return one(fpb::iv_div_lb(a[0], a[1], a[2], a[3])); }
    if(fn == "iv_div_ub") { need(4); return one(fpb::iv_div_ub(a[0], a[1], a[2], a[3])); }
    if(fn == "iv_order") { need(4); return order_name(fpb::iv_order(a[0], a[1], a[2], a[3])); }
    throw std::runtime_error("unknown fixed-point function " + fn);
}

std::vector<fpb::Bounds> arg_bounds(std::istringstream& in) {
    const long long n = num(in);
    std::vector<fpb::Bounds> b;
    for(long long k = 0; k < n; ++k) {
        const long long lo = num(in), hi = num(in);
        b.push_back({lo, hi});
    }
    return b;
}

std::string rest(std::istringstream& in) {
    std::string r;
    std::getline(in, r);
    return r;
}

std::string describe(const NodeP& n) {
    std::ostringstream o;
    if(is_pred_op(n->op)) o << "P " << (n->plb ? 1 : 0) << " " << (n->pub ? 1 : 0);
    else o << "E " << n->lb << " " << n->ub;
    return o.str();
}

std::string run(const std::string& line) {
    std::istringstream in(line);
    std::string cmd;
    in >> cmd;
    if(cmd == "fp") {
        std::string fn;
        in >> fn;
        return fp(fn, in);
    }
    if(cmd == "bs" || cmd == "ctb") {
        auto ab = arg_bounds(in);
        Parser p{rest(in)};
        NodeP n = p.node();
        if(cmd == "bs") return describe(passes::bound_and_simplify(n, ab));
        if(is_pred_op(n->op)) {
            bool plb = false, pub = true;
            ct_pred_bounds(n, ab, plb, pub);
            return std::string("P ") + (plb ? "1" : "0") + " " + (pub ? "1" : "0");
        }
        auto b = ct_bounds(n, ab);
        return "E " + std::to_string(b.lb) + " " + std::to_string(b.ub);
    }
    if(cmd == "dep") {
        auto ab = arg_bounds(in);
        const int a = static_cast<int>(num(in));
        Parser p{rest(in)};
        auto d = passes::bound_dependencies(passes::bound_and_simplify(p.node(), ab), a);
        std::ostringstream o;
        o << d.lb_depends_on_lb << " " << d.lb_depends_on_ub << " " << d.ub_depends_on_lb << " " << d.ub_depends_on_ub;
        return o.str();
    }
    if(cmd == "rw") {
        const int t = static_cast<int>(num(in));
        Parser p{rest(in)};
        auto rws = passes::rewrite_to_bounds(t, p.node());
        std::ostringstream o;
        o << rws.size();
        for(const auto& r : rws) {
            const char* dir = r.dir == flat::Dir::LEQ ? "LEQ" : r.dir == flat::Dir::GEQ ? "GEQ" : "BOTH";
            o.precision(17);
            o << " " << dir << " " << r.expr->c_lb << " " << r.expr->c_ub << " " << r.expr->lb << " " << r.expr->ub;
        }
        return o.str();
    }
    if(cmd == "rows") {
        const int n = static_cast<int>(num(in));
        const long long m = num(in);
        Parser p{rest(in)};
        std::vector<NodeP> cons;
        for(long long k = 0; k < m; ++k) cons.push_back(p.node());
        auto rows = passes::constraint_rows(cons, n);
        std::ostringstream o;
        for(std::size_t a = 0; a < rows.size(); ++a) {
            o << (a ? " |" : "|");
            for(int i : rows[a]) o << " " << i;
        }
        return o.str();
    }
    if(cmd == "node") {
        Parser p{rest(in)};
        NodeP n = p.node();
        std::ostringstream o;
        o << op_name(n->op) << " " << n->lb << " " << n->ub << " " << n->c_lb << " " << n->c_ub;
        return o.str();
    }
    throw std::runtime_error("unknown command " + cmd);
}

} // namespace

int main() {
    std::string line;
    while(std::getline(std::cin, line)) {
        if(line.empty()) continue;
        try {
            std::cout << run(line) << "\n";
        } catch(const std::exception& e) {
            std::cout << "ERROR " << e.what() << "\n";
        }
    }
    return 0;
}
