// Registry of the eleven lemma statements proven by the reference
// (src/main.cpp:33-43 runs them in this order).
#pragma once
#include <functional>
#include <string>
#include <vector>
#include "ivarp/lower.hpp"

namespace sqd_lemmas {

struct LemmaDef {
    std::string key;     // short identifier, e.g. "one"
    std::string title;   // the name printed by the reference's run_proof
    std::function<ivarp::ConstraintSystem()> system;
    std::vector<std::uint64_t> splitting;
    // Published totals (prover_output_23_11_2020.txt) for num_cuboids / num_leaf_cuboids.
    std::uint64_t ref_cuboids, ref_leaves;
};

const std::vector<LemmaDef>& all_lemmas();
const LemmaDef& lemma(const std::string& key);
ivarp::flat::Program lower_lemma(const LemmaDef& def);

} // namespace sqd_lemmas
