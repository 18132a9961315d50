// <ivarp/math_fn.hpp> of the reference's lemma sources (src/*.cpp, src/*.hpp),
// over this project's run-time expression DSL: args::x0.., the _Z / _X
// literals, sqrt, square, maximum, minimum, if_then_else, acos, ensure_expr,
// comparisons and junctors, variable and constraint_system
// (host/ivarp/expr.hpp), and the prover-input surface DefaultContextWithNumberType,
// IDouble, U64Pack, dynamic_subdivision and prover_input (host/ivarp/lower.hpp).
// With this directory on the include path the reference's lemma files compile
// unchanged and lower to the programs the GPU core proves.
#pragma once
#include "expr.hpp"
#include "lower.hpp"
