// <ivarp/prover.hpp> of the reference's lemma sources: the expression DSL and
// the prover entry point.
#pragma once
#include "ivarp/math_fn.hpp"
#include "ivarp/run_prover.hpp"
