// The passes of the lowering (lower.cpp) one by one, as the reference's own
// lowering tests exercise them: bound-and-simplify with compile-time bounds
// (refactor_constraint_system/bound_and_simplify.hpp), the rewriting of a
// constraint into bounds on one variable (constraints_to_bounds/rewriters/
// relational_ops.hpp:31-129), the bound dependency analysis
// (bound_dependency_analysis.hpp:65-80) and the runtime constraint table's rows
// (runtime_constraint_table.hpp:31-115).
#pragma once
#include <vector>
#include "expr.hpp"
#include "program.hpp"

namespace ivarp {
namespace passes {

// The tree with compile-time bounds on every node (and predicates decided at
// compile time folded), for arguments with the given bounds.
NodeP bound_and_simplify(const NodeP& n, const std::vector<fixed_point_bounds::Bounds>& arg_bounds);

// The bounds a constraint imposes on argument `target` (none if it is not a
// comparison with the argument alone on one side).
struct BoundRewrite {
    int target;
    flat::Dir dir;
    NodeP expr;
};
std::vector<BoundRewrite> rewrite_to_bounds(int target, const NodeP& constraint);

// Which bound of a bounded expression may depend on which bound of `arg`.
struct BoundDependencies {
    bool lb_depends_on_lb, lb_depends_on_ub, ub_depends_on_lb, ub_depends_on_ub;
};
BoundDependencies bound_dependencies(const NodeP& bounded_expr, int arg);

// Rows of the runtime constraint table over `nargs` arguments: row a < nargs - 1
// holds the constraints no argument's bound rewriting takes whose highest
// argument is a, the last row every constraint (indices into `constraints`).
std::vector<std::vector<int>> constraint_rows(const std::vector<NodeP>& constraints, int nargs);

} // namespace passes
} // namespace ivarp
