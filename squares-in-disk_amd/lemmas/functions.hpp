// Auxiliary functions of the "Packing Squares into a Disk" proofs, written
// against the ivarp-compatible DSL.  These are the paper's formulas; they are the
// workload and must be stated exactly as in the reference
// (src/fn_aux.hpp:27-55, src/XY.hpp:27-47, src/B1.hpp-B6.hpp).
#pragma once
#include "ivarp/expr.hpp"

namespace sqd_lemmas {
using namespace ivarp;
using namespace ivarp::args;

// T(u), the top boundary of the disk as a function of the height u (src/fn_aux.hpp)
inline const Expr T = (2_Z / sqrt(ensure_expr(5_Z))) * sqrt(maximum(0_Z, 1_Z - square(x0) / 5_Z)) - 0.8_X * x0;
// T^{-1}(s)
inline const Expr T_inv = sqrt(maximum(1_Z - 0.25_X * square(x0), 0_Z)) - x0;
// w(y_t, h): width available at height y_t for a strip of height h
inline const Expr w = 2_Z * sqrt(maximum(0_Z, minimum(1_Z - square(x0), 1_Z - square(x0 - x1))));
// sigma(s_1)
inline const Expr sigma = if_then_else(x0 <= sqrt((2_Z + sqrt(ensure_expr(2_Z))) / 3_Z),
    0.25_X * (-x0 - 2_Z * T_inv(x0) + sqrt(8_Z - square(x0 - 2_Z * T_inv(x0)))),
    0.2_X * (sqrt(20_Z - square(x0)) - 2_Z * x0));
inline const Expr s1_star = sqrt((1_Z / 3_Z) * (2_Z + sqrt(ensure_expr(2_Z))));
// 0 if s_n > sigma, else 0.83 sigma^2
inline const Expr last_square = if_then_else(sigma(x0) >= x1, 0.83_X * square(sigma(x0)), 0_Z);

// X(a_i, b_i, u) and Y(a, h_i, w_i, h_{i+1}) (src/XY.hpp)
inline const Expr X = if_then_else(
    x1 >= 0_Z,
    sqrt(1_Z - square(x1 + x2)) - x2,
    if_then_else(
        x0 < 0_Z,
        sqrt(maximum(0_Z, 1_Z - square(x0 + x2))) - x2,
        if_then_else(
            x2 <= 2_Z * minimum(x0, -x1),
            T_inv(x2),
            sqrt(maximum(0_Z, 1_Z - square(x2 - minimum(x0, -x1)))) - x2)));
inline const Expr Y = 0.5_X * x2 - x1 + X(x0, x0 - x1, x3);

// B_1(h_i, w_i, h_{i+1})
inline const Expr B_1 = maximum(0.5_X * x0 * x1 + 0.25_X * square(x0),
                                square(x0) + (x1 - x0 - x2) * x2,
                                0.5_X * x0 * (x1 + x0) - square(x2));
// B_2(h_i, w_i, h_{i+1})
inline const Expr B_2 = if_then_else(x1 < x0 + x2, square(x0),
                                     if_then_else(x1 < 2_Z * x0, square(x0) + square(x2), B_1(x0, x1, x2)));
// B_3(a, h_i, w_i, h_{i+1})
inline const Expr B_3 = square(x1) + maximum(maximum(0_Z, Y(x0, x1, x2, x3) * x3),
                                             minimum(square(maximum(0_Z, Y(x0, x1, x2, x3))), 2_Z * square(x3)));
// B_4(a, h_i, w_i, h_{i+1})
inline const Expr B_4 = maximum(B_2(x1, x2, x3), B_3(x0, x1, x2, x3));
// B_5(h_j, H_j)
inline const Expr B5_d = minimum(1_Z, maximum(0_Z, 1_Z - x1 + x0));
inline const Expr B5_A = acos(B5_d) - B5_d * sqrt(1_Z - square(B5_d));
inline const Expr B_5 = B5_A + square(x0) - x1 * x0;
// B_6(H_R, W_R, h_{j+1})
inline const Expr B_6 = (x0 * x1) / 2_Z + (x2 * x0) / 4_Z;

} // namespace sqd_lemmas
