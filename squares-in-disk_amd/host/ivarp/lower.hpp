// Lowering of a constraint system into the flat program (the "refactoring" of
// ivarp/include/ivarp/refactor_constraint_system.hpp and prover_input.hpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "expr.hpp"
#include "program.hpp"

namespace ivarp {

// Number-type tags and contexts of the reference's API; only IDouble contexts
// are proven on the GPU (the FP32 prefilter is an internal pass of the core).
struct IDouble {};
struct IFloat {};
template<typename NumberType_> struct DefaultContextWithNumberType { using NumberType = NumberType_; };

template<std::uint64_t... S> struct U64Pack {};
constexpr std::uint64_t dynamic_subdivision(std::uint32_t initial, std::uint32_t subdivision) {
    return (static_cast<std::uint64_t>(initial) << 32u) | subdivision;
}
constexpr bool is_dynamic_subdivision(std::uint64_t s) { return (s >> 32u) != 0; }

// Run the reference's lowering on a constraint system with the given split
// policy (one 64-bit split code per variable, in argument order).
flat::Program lower_constraint_system(const ConstraintSystem& cs, const std::vector<std::uint64_t>& splitting,
                                      const std::string& name = std::string());

struct ProverInput {
    flat::Program program;
    std::size_t num_args() const noexcept { return static_cast<std::size_t>(program.nargs); }
};

namespace impl {
template<typename Pack> struct PackValues;
template<std::uint64_t... S> struct PackValues<U64Pack<S...>> {
    static std::vector<std::uint64_t> get() { return {S...}; }
};
}

// prover_input<Context, U64Pack<split policy>>(system) as in
// ivarp/include/ivarp/prover_input.hpp:118-122.
template<typename Context, typename VariableSplitting>
ProverInput prover_input(const ConstraintSystem& cs) {
    static_assert(std::is_same<typename Context::NumberType, IDouble>::value,
                  "the GPU prover core proves IDouble contexts");
    return ProverInput{lower_constraint_system(cs, impl::PackValues<VariableSplitting>::get())};
}

} // namespace ivarp
