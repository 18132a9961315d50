// Runtime expression DAG with the lemma-facing API of ivarp's expression
// templates (ivarp/include/ivarp/math_fn/*.hpp, math_fn/math_pred/*.hpp,
// number/literals.hpp, variable_description.hpp, constraint_system.hpp).
//
// The reference encodes every expression in its C++ type and lowers it at
// compile time.  Here the same user-facing names build a DAG of shared nodes at
// run time; the lowering in lower.hpp then reproduces the reference's compile-time
// passes on that DAG.  Lemma sources therefore read exactly like the reference's.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>
#include <stdexcept>
#include <initializer_list>
#include <type_traits>
#include "fixed_point.hpp"
#include "bigrat.hpp"

namespace ivarp {

enum class Op : std::uint8_t {
    // expressions
    Const, Arg, Neg, Add, Sub, Mul, Div, Sqrt, Square, Max, Min, Ite, Acos,
    // predicates
    BoolConst, Lt, Le, Gt, Ge, Eq, Neq, Not, OrSeq, AndSeq, Or, And
};
inline bool is_pred_op(Op o) noexcept { return static_cast<unsigned>(o) >= static_cast<unsigned>(Op::BoolConst); }
const char* op_name(Op o) noexcept;

struct Node {
    Op op = Op::Const;
    std::vector<std::shared_ptr<const Node>> kids;
    int arg = -1;
    // Exact value of a literal constant (raw trees only).
    std::shared_ptr<const exact::RInterval> exact_value;
    // Interval value of a constant (outward double enclosure) and undefinedness.
    double c_lb = 0.0, c_ub = 0.0;
    bool c_undef = false;
    // Runtime value of a boolean constant (IBool {definitely, possibly}).
    bool bval_def = false, bval_poss = false;
    // Compile-time bounds: present on constants and on every node produced by
    // bound-and-simplify (the BoundedMathExpr / BoundedPredicate wrappers).
    bool bounded = false;
    std::int64_t lb = fixed_point_bounds::min_bound(), ub = fixed_point_bounds::max_bound();
    bool plb = false, pub = true;
};
using NodeP = std::shared_ptr<const Node>;

NodeP make_node(Op op, std::vector<NodeP> kids);
NodeP make_arg(int index);

// A decimal or integer literal with fixed-point compile-time bounds (BoundedRational).
struct Literal {
    exact::Rat value;
    std::int64_t lb, ub;
};
Literal make_int_literal(const char* digits);
Literal make_decimal_literal(const char* digits);
Literal operator-(const Literal& a);
Literal operator+(const Literal& a, const Literal& b);
Literal operator-(const Literal& a, const Literal& b);
Literal operator*(const Literal& a, const Literal& b);
Literal operator/(const Literal& a, const Literal& b);

inline namespace literals {
inline Literal operator""_Z(const char* s) { return make_int_literal(s); }
inline Literal operator""_X(const char* s) { return make_decimal_literal(s); }
}

NodeP literal_node(const Literal& l);

class Expr;
template<typename T> struct IsExprLike : std::integral_constant<bool,
    std::is_same<std::decay_t<T>, Expr>::value || std::is_same<std::decay_t<T>, Literal>::value> {};

class Expr {
public:
    Expr() = default;
    explicit Expr(NodeP n) : node(std::move(n)) {}
    Expr(const Literal& l) : node(literal_node(l)) {} // NOLINT: implicit promotion like MathConstant
    // Built-in integers become constants whose compile-time bounds are the range
    // of `int` (NumberToConstant / BoundsFromIntType, math_fn/constant/number_to_constant.hpp:28-34).
    Expr(int v); // NOLINT
    // Symbolic call: substitute x0, x1, ... by the given expressions.
    template<typename... A> Expr operator()(const A&... a) const {
        std::vector<NodeP> r{Expr(a).node...};
        return call(r);
    }
    Expr call(const std::vector<NodeP>& replacements) const;
    NodeP node;
};

class Pred {
public:
    Pred() = default;
    explicit Pred(NodeP n) : node(std::move(n)) {}
    Pred(bool b); // NOLINT: implicit promotion like MathBoolConstant
    NodeP node;
};

namespace args {
// Inline variables: initialised before any later definition in a translation
// unit that includes this header, so lemma statements may use them at namespace scope.
inline const Expr x0{make_arg(0)}, x1{make_arg(1)}, x2{make_arg(2)}, x3{make_arg(3)}, x4{make_arg(4)},
                  x5{make_arg(5)}, x6{make_arg(6)}, x7{make_arg(7)}, x8{make_arg(8)}, x9{make_arg(9)};
}

inline Expr ensure_expr(const Expr& e) { return e; }
inline Expr ensure_expr(const Literal& l) { return Expr(l); }

Expr operator-(const Expr& a);
Expr operator+(const Expr& a, const Expr& b);
Expr operator-(const Expr& a, const Expr& b);
Expr operator*(const Expr& a, const Expr& b);
Expr operator/(const Expr& a, const Expr& b);
#define IVARP_MIXED_OP(op) \
    inline Expr operator op(const Expr& a, const Literal& b) { return a op Expr(b); } \
    inline Expr operator op(const Literal& a, const Expr& b) { return Expr(a) op b; }
IVARP_MIXED_OP(+) IVARP_MIXED_OP(-) IVARP_MIXED_OP(*) IVARP_MIXED_OP(/)
#undef IVARP_MIXED_OP

Expr sqrt(const Expr& a);
Expr acos(const Expr& a);
Expr square(const Expr& a);
inline Expr square(const Literal& a) { return square(Expr(a)); }
Expr maximum_v(std::vector<NodeP> kids);
Expr minimum_v(std::vector<NodeP> kids);
template<typename... A> Expr maximum(const A&... a) { return maximum_v({Expr(a).node...}); }
template<typename... A> Expr minimum(const A&... a) { return minimum_v({Expr(a).node...}); }
Expr if_then_else_impl(const Pred& c, const Expr& t, const Expr& e);
template<typename T, typename E> Expr if_then_else(const Pred& c, const T& t, const E& e) {
    return if_then_else_impl(c, Expr(t), Expr(e));
}

Pred compare(Op op, const Expr& a, const Expr& b);
#define IVARP_CMP(op, code) \
    inline Pred operator op(const Expr& a, const Expr& b) { return compare(Op::code, a, b); } \
    inline Pred operator op(const Expr& a, const Literal& b) { return compare(Op::code, a, Expr(b)); } \
    inline Pred operator op(const Literal& a, const Expr& b) { return compare(Op::code, Expr(a), b); }
IVARP_CMP(<, Lt) IVARP_CMP(<=, Le) IVARP_CMP(>, Gt) IVARP_CMP(>=, Ge) IVARP_CMP(==, Eq) IVARP_CMP(!=, Neq)
#undef IVARP_CMP

Pred operator||(const Pred& a, const Pred& b);
Pred operator&&(const Pred& a, const Pred& b);
Pred operator|(const Pred& a, const Pred& b);
Pred operator&(const Pred& a, const Pred& b);
Pred operator!(const Pred& a);

// Variable description: name and lower/upper bound expressions (may use
// variables of smaller index).
struct VariableDescription {
    int index;
    std::string name;
    Expr lb, ub;
};
template<typename L, typename U>
VariableDescription variable(const Expr& arg, std::string name, const L& lb, const U& ub) {
    if(!arg.node || arg.node->op != Op::Arg) throw std::invalid_argument("variable() needs an argument x_i");
    return VariableDescription{arg.node->arg, std::move(name), Expr(lb), Expr(ub)};
}

class ConstraintSystem {
public:
    std::size_t num_args() const noexcept { return descriptions_by_index.size(); }
    std::string arg_name(std::size_t i) const { return descriptions_by_index.at(i).name; }
    // Constraints in reference order: the bound constraints x >= lb, x <= ub of each
    // variable (in the order the variables were given), then the user constraints.
    std::vector<NodeP> constraints;
    std::vector<VariableDescription> descriptions_by_index;
    // Compile-time bounds of the variables, computed in index order from the
    // variables' bound expressions (ivarp/include/ivarp/constraint_system/extract_ct_bounds.hpp:65-85).
    std::vector<fixed_point_bounds::Bounds> ct_bounds;
};

namespace impl {
struct CSBuilder {
    std::vector<VariableDescription> vars;
    std::vector<NodeP> cons;
    void add(const VariableDescription& v) { vars.push_back(v); }
    void add(const Pred& p) { cons.push_back(p.node); }
};
ConstraintSystem finish_constraint_system(CSBuilder&& b);
}
template<typename... A> ConstraintSystem constraint_system(const A&... a) {
    impl::CSBuilder b;
    int dummy[] = {0, (b.add(a), 0)...};
    (void)dummy;
    return impl::finish_constraint_system(std::move(b));
}

// Compile-time bounds of an expression/predicate tree (CompileTimeBounds metafunction).
fixed_point_bounds::Bounds ct_bounds(const NodeP& n, const std::vector<fixed_point_bounds::Bounds>& args);
void ct_pred_bounds(const NodeP& n, const std::vector<fixed_point_bounds::Bounds>& args, bool& plb, bool& pub);
int num_args(const NodeP& n);

} // namespace ivarp
