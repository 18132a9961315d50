// The eleven lemma statements of "Packing Squares into a Disk with Optimal
// Worst-Case Density", stated with the ivarp-compatible DSL.  Each system and
// split policy is the workload the reference proves (file:line of the reference
// statement in the comment above each one); the formulas must be identical.
#include "lemmas.hpp"
#include "functions.hpp"
#include <stdexcept>

namespace sqd_lemmas {
namespace {

// src/top_packing.hpp:40-51 (statement 1) and :53-60 (statement 2)
ConstraintSystem top_packing(int which) {
    const Expr s1 = x0;
    const Expr l1 = sqrt(1_Z - square(T_inv(s1))) - 0.5_X * s1;
    const Expr sigma_2sq2 = (1_Z / (2_Z * sqrt(ensure_expr(2_Z)))) * sigma(s1);
    const auto s1d = variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X)));
    if(which == 1) {
        const Expr xp1 = 0.5_X * s1 + sigma_2sq2;
        const Expr yp1 = if_then_else(s1 <= s1_star, T_inv(s1) + sigma(s1) + sigma_2sq2, 0.5_X * sigma + sigma_2sq2);
        const Expr F_TP1 = square(xp1) + square(yp1);
        return constraint_system(s1d, s1 >= l1, F_TP1 > 1_Z);
    }
    const Expr xp2 = 0.5_X * s1 + 0.645_X * sigma(s1);
    const Expr yp2 = if_then_else(s1 <= s1_star, T_inv(s1) + 2_Z * 0.645_X * sigma(s1), 0.79_X * sigma(s1));
    const Expr F_TP2 = square(xp2) + square(yp2);
    return constraint_system(s1d, s1 >= l1, F_TP2 > 1_Z);
}

// src/one_subcontainer.cpp:38-61
ConstraintSystem one_subcontainer() {
    const Expr s1 = x0, h1 = x1, sn = x2;
    const Expr z = T(-T_inv(s1) + h1);
    const Expr w1 = w(T_inv(s1), h1);
    const Expr F_OC = square(s1) + B_4(T_inv(s1), h1, w1, sn) + square(sn) + last_square(s1, sn);
    const Pred not_case_2 = (s1 > 1_Z / sqrt(ensure_expr(2_Z))) || w1 < 2_Z * h1 ||
                            square(s1) + square(h1) + 2_Z * square(sn) <= 1.56_X;
    return constraint_system(
        variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))),
        variable(h1, "h_1", 0_Z, s1),
        variable(sn, "s_n", 0_Z, h1),
        F_OC <= 1.6_X,
        s1 >= h1, h1 >= sn, z >= 0_Z,
        sn >= z,
        h1 <= T_inv(s1) + 1_Z,
        not_case_2);
}

// src/two_subcontainer.cpp:38-67
ConstraintSystem two_subcontainers() {
    const Expr s1 = x0, h1 = x1, h2 = x2, sn = x3;
    const Expr z = T(-T_inv(s1) + h1 + h2);
    const Expr w1 = w(T_inv(s1), h1);
    const Expr w2 = w(T_inv(s1) - h1, h2);
    const Expr F_2C_1 = square(s1) + B_4(T_inv(s1), h1, w1, h2) + B_4(T_inv(s1) - h1, h2, w2, sn) +
                        square(sn) + last_square(s1, sn);
    return constraint_system(
        variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))),
        variable(h1, "h_1", 0_Z, s1),
        variable(h2, "h_2", 0_Z, h1),
        variable(sn, "s_n", maximum(0_Z, z), h2),
        z >= 0_Z,
        F_2C_1 <= 1.6_X,
        h1 <= s1, h2 <= h1,
        sn <= h2, h1 <= T_inv(s1) + 1_Z,
        h2 <= T_inv(s1) + 1_Z - h1);
}

// src/three_subcontainer.cpp:37-68
ConstraintSystem three_subcontainers() {
    const Expr s1 = x0, h1 = x1, h2 = x2, h3 = x3, sn = x4;
    const Expr z = T(-T_inv(s1) + h1 + h2 + h3);
    const Expr w1 = w(T_inv(s1), h1);
    const Expr w2 = w(T_inv(s1) - h1, h2);
    const Expr w3 = w(T_inv(s1) - h1 - h2, h3);
    const Expr F_3C_1 = square(s1) + B_4(T_inv(s1), h1, w1, h2) + B_4(T_inv(s1) - h1, h2, w2, h3) +
                        B_4(T_inv(s1) - h1 - h2, h3, w3, sn) + square(sn) + last_square(s1, sn);
    return constraint_system(
        variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))),
        variable(h1, "h_1", 0_Z, s1),
        variable(h2, "h_2", 0_Z, h1),
        variable(h3, "h_3", 0_Z, h2),
        variable(sn, "s_n", maximum(0_Z, z), h3),
        z >= 0_Z,
        h1 <= s1, h2 <= h1,
        sn <= h2, h1 <= T_inv(s1) + 1_Z,
        h2 <= T_inv(s1) + 1_Z - h1,
        h3 <= T_inv(s1) + 1_Z - h1 - h2,
        F_3C_1 <= 1.6_X);
}

// src/four_subcontainer.cpp:37-74
ConstraintSystem four_subcontainers() {
    const Expr s1 = x0, h1 = x1, h2 = x2, h3 = x3, h4 = x4, sn = x5;
    const Expr z = T(-T_inv(s1) + h1 + h2 + h3 + h4);
    const Expr w1 = w(T_inv(s1), h1);
    const Expr w2 = w(T_inv(s1) - h1, h2);
    const Expr w3 = w(T_inv(s1) - h1 - h2, h3);
    const Expr w4 = w(T_inv(s1) - h1 - h2 - h3, h4);
    const Expr F_4C_1 = square(s1) + B_4(T_inv(s1), h1, w1, h2) + B_4(T_inv(s1) - h1, h2, w2, h3) +
                        B_4(T_inv(s1) - h1 - h2, h3, w3, h4) + B_4(T_inv(s1) - h1 - h2 - h3, h4, w4, sn) +
                        square(sn) + last_square(s1, sn);
    return constraint_system(
        variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))),
        variable(h1, "h_1", 0_Z, s1),
        variable(h2, "h_2", 0_Z, h1),
        variable(h3, "h_3", 0_Z, h2),
        variable(h4, "h_4", 0_Z, h3),
        variable(sn, "s_n", maximum(0_Z, z), h4),
        z >= 0_Z,
        h1 <= T_inv(s1) + 1_Z,
        h2 <= T_inv(s1) + 1_Z - h1,
        h3 <= T_inv(s1) + 1_Z - h1 - h2,
        h4 <= T_inv(s1) + 1_Z - h1 - h2 - h3,
        F_4C_1 <= 1.6_X);
}

// Five, six and seven subcontainers with s_n > sigma share one shape:
// src/five_subcontainer.cpp:38-80, src/six_subcontainer.cpp:38-86, src/seven_subcontainer.cpp:38-91.
ConstraintSystem k_subcontainers(int k) {
    const Expr s1 = x0, sn = x1;
    std::vector<Expr> h;
    for(int i = 0; i < k; ++i) h.push_back(Expr(make_arg(2 + i)));
    Expr hsum = -T_inv(s1);
    for(int i = 0; i < k; ++i) hsum = hsum + h[static_cast<std::size_t>(i)];
    const Expr z = T(hsum);
    std::vector<Expr> ws;
    Expr a = T_inv(s1);
    std::vector<Expr> as;
    for(int i = 0; i < k; ++i) {
        as.push_back(a);
        ws.push_back(w(a, h[static_cast<std::size_t>(i)]));
        a = a - h[static_cast<std::size_t>(i)];
    }
    Expr F = square(s1) + square(sn);
    for(int i = 0; i < k; ++i) {
        const Expr& next = (i + 1 < k) ? h[static_cast<std::size_t>(i + 1)] : sn;
        F = F + B_4(as[static_cast<std::size_t>(i)], h[static_cast<std::size_t>(i)], ws[static_cast<std::size_t>(i)], next);
    }
    impl::CSBuilder b;
    b.add(variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))));
    b.add(variable(sn, "s_n", sigma(s1), s1));
    for(int i = 0; i < k; ++i) {
        const Expr& upper = i == 0 ? s1 : h[static_cast<std::size_t>(i - 1)];
        b.add(variable(h[static_cast<std::size_t>(i)], "h_" + std::to_string(i + 1), sn, upper));
    }
    b.add(sn > z);
    b.add(z >= 0_Z);
    b.add(F <= 1.6_X);
    // h_i <= T_inv(s1) + 1 - h_1 - ... - h_{i-1} - (k-i) sigma(s1)
    for(int i = 0; i < k; ++i) {
        Expr rhs = T_inv(s1) + 1_Z;
        for(int j = 0; j < i; ++j) rhs = rhs - h[static_cast<std::size_t>(j)];
        int mult = k - 1 - i;
        if(mult > 1) rhs = rhs - make_int_literal(std::to_string(mult).c_str()) * sigma(s1);
        else if(mult == 1) rhs = rhs - sigma(s1);
        b.add(h[static_cast<std::size_t>(i)] <= rhs);
    }
    return impl::finish_constraint_system(std::move(b));
}

// src/more_subcontainers.cpp:31-79
ConstraintSystem more_subcontainers_1() {
    const Expr s1 = x0, h1 = x1, h2 = x2, h3 = x3, h4 = x4;
    const Expr H4 = 1_Z + T_inv(s1) - h1 - h2 - h3;
    const Expr w1 = w(T_inv(s1), h1);
    const Expr w2 = w(T_inv(s1) - h1, h2);
    const Expr w3 = w(T_inv(s1) - h1 - h2, h3);
    const Expr least_each = T_inv(s1) * (1_Z / 3_Z);
    const Expr F_MSC_1 = square(s1) + 0.83_X * square(sigma(s1)) + B_5(h4, H4) + B_4(T_inv(s1), h1, w1, h2) +
                         B_4(T_inv(s1) - h1, h2, w2, h3) + B_4(T_inv(s1) - h1 - h2, h3, w3, h4);
    const Expr F_MSC_1p1 = square(s1) + 0.83_X * square(sigma(s1));
    const Expr F_MSC_1p3 = square(s1) + 0.83_X * square(sigma(s1)) + B_4(T_inv(s1), h1, w1, h2) +
                           B_4(T_inv(s1) - h1, h2, w2, h3);
    return constraint_system(
        variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))),
        variable(h1, "h_1", least_each, s1),
        variable(h2, "h_2", least_each, h1),
        variable(h3, "h_3", least_each, h2),
        variable(h4, "h_4", 0_Z, h3),
        0_Z <= H4, H4 <= 1_Z,
        F_MSC_1p1 <= 1.6_X,
        F_MSC_1p3 <= 1.6_X,
        F_MSC_1 <= 1.6_X);
}

// src/more_subcontainers.cpp:82-128
ConstraintSystem more_subcontainers_2() {
    const Expr s1 = x0, h1 = x1, h2 = x2, h3 = x3, dely = x4, h_jp1 = x5;
    const Expr w1 = w(T_inv(s1), h1);
    const Expr w2 = w(T_inv(s1) - h1, h2);
    const Expr H_R = maximum(0_Z, T_inv(s1) - h1 - h2) + dely;
    const Expr W_R = w(maximum(0_Z, T_inv(s1) - h1 - h2), H_R);
    const Expr H_jp1 = 1_Z - dely;
    const Expr F_MSC_2 = square(s1) + 0.83_X * square(sigma) + B_4(T_inv(s1), h1, w1, h2) +
                         B_4(T_inv(s1) - h1, h2, w2, h3) + B_5(h_jp1, H_jp1) + B_6(H_R, W_R, h_jp1);
    const Expr F_MSC_2p1 = square(s1) + 0.83_X * square(sigma) + B_4(T_inv(s1), h1, w1, h2) +
                           B_4(T_inv(s1) - h1, h2, w2, h3);
    return constraint_system(
        variable(s1, "s_1", 0.295_X, sqrt(ensure_expr(1.6_X))),
        variable(h1, "h_1", 0_Z, s1),
        variable(h2, "h_2", 0_Z, h1),
        variable(h3, "h_3", 0_Z, h2),
        variable(dely, "Delta_y", 0_Z, h3),
        variable(h_jp1, "h_{j+1}", 0_Z, h3),
        h1 <= T_inv(s1),
        h2 <= T_inv(s1) - h1,
        h3 <= T_inv(s1) - h1 - h2,
        F_MSC_2 <= 1.6_X,
        F_MSC_2p1 <= 1.6_X);
}

// Small known-answer systems of the reference's prover tests.
// ivarp/test/expected_criticals.cpp:54-61 (run with generation_count = 2)
ConstraintSystem expected_criticals_2d() {
    const Expr x = x0, y = x1;
    return constraint_system(variable(x, "x", 0_Z, 2_Z), variable(y, "y", 0_Z, 2_Z),
                             x >= y, x <= 3_Z - y, y <= 3_Z - x);
}
// ivarp/test/expected_criticals2.cpp:30-40 (run with generation_count 0, 1, 2)
ConstraintSystem expected_criticals_4d() {
    const Expr w = x0, x = x1, y = x2, z = x3;
    return constraint_system(variable(w, "w", 0_Z, 1_Z), variable(x, "x", 0_Z, 1_Z),
                             variable(y, "y", 0_Z, 1_Z), variable(z, "z", 0_Z, 1_Z),
                             w + x + y + z >= 0);
}

// A known-answer system of this implementation's own (no reference
// counterpart), shaped for the per-parent paths of the static levels: the last
// static level splits z 128 ways, its rows hold terms free of z (hoisted once
// per parent: sqrt(x + y)), z <= y lets a child's z range raise y's lower bound
// (a child that changes another variable is decided apart from its siblings),
// and the final generation leaves critical boxes.  Its counts and critical
// boxes are checked against the CPU oracle.
ConstraintSystem expected_criticals_hoisted() {
    const Expr x = x0, y = x1, z = x2;
    return constraint_system(variable(x, "x", 0_Z, 1_Z), variable(y, "y", 0_Z, 1_Z), variable(z, "z", 0_Z, 1_Z),
                             z <= y, sqrt(x + y) + z >= 1_Z);
}

constexpr std::uint64_t dyn(std::uint32_t i, std::uint32_t s) { return ivarp::dynamic_subdivision(i, s); }

} // namespace

const std::vector<LemmaDef>& all_lemmas() {
    static const std::vector<LemmaDef> defs = {
        {"top1", "Top Packing Lemma, statement (1)", [] { return top_packing(1); }, {dyn(128, 8)}, 128, 128},
        {"top2", "Top Packing Lemma, statement (2)", [] { return top_packing(2); }, {dyn(128, 8)}, 168, 168},
        {"one", "One Subcontainer Lemma", one_subcontainer, {dyn(128, 8), 256, 256}, 2869888, 2858678},
        {"two", "Two Subcontainer Lemma", two_subcontainers, {dyn(128, 8), dyn(128, 8), 64, 64}, 7181696, 7069587},
        {"three", "Three Subcontainer Lemma", three_subcontainers,
         {dyn(64, 8), dyn(64, 8), dyn(16, 4), 64, 32}, 3549856, 3440786},
        {"four", "Four Subcontainer Lemma", four_subcontainers,
         {dyn(64, 8), dyn(64, 8), dyn(16, 4), 32, 32, 16}, 37245328, 35026303},
        {"five", "Five Subcontainers, s_n > sigma", [] { return k_subcontainers(5); },
         {dyn(128, 4), dyn(16, 4), dyn(16, 4), 16, 16, 16, 4}, 3392, 3188},
        {"six", "Six Subcontainers, s_n > sigma", [] { return k_subcontainers(6); },
         {dyn(128, 4), dyn(16, 4), dyn(16, 4), dyn(16, 4), 16, 16, 16, 4}, 14848, 13928},
        {"seven", "Seven Subcontainers, s_n > sigma", [] { return k_subcontainers(7); },
         {dyn(128, 4), dyn(16, 4), dyn(16, 4), dyn(16, 4), 16, 16, 16, 16, 4}, 1808, 1703},
        {"more1", ">= 5 Subcontainers, s_n <= sigma, y_3 <= 0", more_subcontainers_1,
         {dyn(128, 4), dyn(64, 4), dyn(64, 4), 32, 32}, 1056154896, 1022752013},
        {"more2", ">= 5 Subcontainers, s_n <= sigma, y_3 > 0", more_subcontainers_2,
         {dyn(128, 4), dyn(64, 4), dyn(64, 4), 32, 32, 16}, 123127968, 115680431},
        // known-answer systems of the reference's tests; totals are for generation_count = 2
        {"ec1", "expected criticals, 2 variables", expected_criticals_2d, {dyn(8, 2), 4}, 408, 352},
        {"ec2", "expected criticals, 4 variables", expected_criticals_4d, {dyn(4, 4), dyn(4, 4), 8, 8}, 614868, 541696},
        {"ec3", "expected criticals, 3 variables, a 128-way last static level", expected_criticals_hoisted,
         {dyn(4, 2), 8, 128}, 45436, 45056},
    };
    return defs;
}

const LemmaDef& lemma(const std::string& key) {
    for(const auto& d : all_lemmas()) if(d.key == key) return d;
    throw std::invalid_argument("unknown lemma: " + key);
}

ivarp::flat::Program lower_lemma(const LemmaDef& def) {
    return ivarp::lower_constraint_system(def.system(), def.splitting, def.title);
}

} // namespace sqd_lemmas
