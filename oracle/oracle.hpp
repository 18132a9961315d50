// CPU oracle: test infrastructure only.
//
// A plain C++ restatement of ivarp's CPU prover core (the reference's parity
// oracle), run on the lowered program text.  It is used by the tests, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker and
// the CPU baseline; nothing in the product links, loads or calls it.
//
// Restated reference functions (file:line):
//   interval + - * / sqrt acos, min/max, comparisons, IBool
//        math_fn/interval_acos.hpp:19-29 with ivarp/src/interval_acos.cpp:29-66 (MPFR),
//        ivarp/include/ivarp/number/float_interval_addsub.hpp:53-102,
//        float_interval_mul.hpp:33-166, float_interval_div.hpp:45-107, interval_div.hpp:57-288,
//        math_fn/interval_sqrt.hpp:72-116, number/minmax.hpp:59-177,
//        number/interval_comparisons.hpp:32-290, bool.hpp:70-149,
//        math_fn/fixed_pow.hpp:173-200, math_fn/eval/if_then_else.hpp:64-82
//   bound application      refactor_constraint_system/constraints_to_bounds/compile_time_bound.hpp:42-85
//   propagation            bound_propagation/iterated_recursive.hpp:40-131,
//                          bound_propagation/dynamic_bound_application.hpp:59-153, index_queue_set.hpp:38-101
//   splitting              splitter.hpp:36-121
//   static enumeration,    cpu_prover/cpu_prover_core.hpp:52-256
//   hull and criticals
//   driver                 prover/proof_driver.hpp:300-438,472-509, prover/pd_impl_split_into.hpp:29-75
//   initial box            refactor_constraint_system/refactor_constraints_impl.hpp:61-101
//   binary32 intervals     the same operations on float (ivarp's IFloat number type,
//                          number/interval.hpp:346-347), for the FP32 prefilter
#pragma once
#include <cstdint>
#include <cstddef>

extern "C" {

struct OracleCounts {
    std::uint64_t num_cuboids;
    std::uint64_t num_leaf_cuboids;
    std::uint64_t num_critical_cuboids;
    std::uint64_t num_repeated_nodes;
};

// Parse a lowered program (text form).  Returns 0 on success.
int oracle_load(const char* program_text, void** handle);
void oracle_free(void* handle);
int oracle_nargs(void* handle);
const char* oracle_last_error(void);

// Initial runtime box: outward compile-time box plus 50*N rounds of all bounds.
int oracle_initial_box(void* handle, double* lb, double* ub);

// Whole proof with `threads` worker threads and the given generation count.
// Criticals (up to max_criticals) are written as [lb0,ub0,lb1,ub1,...] per cuboid.
int oracle_run(void* handle, int generations, int threads, OracleCounts* counts,
               double* criticals, std::uint64_t max_criticals, std::uint64_t* n_criticals);

// Same, restricted to the initial-queue index range [first, last) of the
// compacted initial queue (queue order of the reference's proof initialisation).
int oracle_run_slice(void* handle, int generations, int threads, std::uint64_t first, std::uint64_t last,
                     OracleCounts* counts, std::uint64_t* initial_queue_len);

// The initial queue (compacted), n entries of N intervals; returns the length.
int oracle_initial_queue(void* handle, double* lb, double* ub, std::uint64_t cap, std::uint64_t* n, OracleCounts* counts);

// One non-final pass of the core over n boxes (SoA: lb[a*n+i], ub[a*n+i]).
// For every input: has_hull[i] and the joined hull of the surviving static leaves.
int oracle_nonfinal(void* handle, std::uint64_t n, const double* lb, const double* ub,
                    double* hull_lb, double* hull_ub, std::uint8_t* has_hull, OracleCounts* counts);
// One final pass: criticals are appended in enumeration order (SoA per critical, stride N).
int oracle_final(void* handle, std::uint64_t n, const double* lb, const double* ub,
                 double* crit_lb, double* crit_ub, std::uint64_t cap, std::uint64_t* n_crit, OracleCounts* counts);
// Dynamic split of n boxes (depth d -> d+1): surviving children appended (stride N).
int oracle_split(void* handle, std::uint64_t n, const double* lb, const double* ub,
                 double* out_lb, double* out_ub, std::uint64_t cap, std::uint64_t* n_out, OracleCounts* counts);

// Elementwise directed-rounding primitives for arithmetic parity tests.
// op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt(a), 5 acos(a), 6 acos(a) (the same; the
// device evaluates it by another method).  out_lo = RD(a op b), out_hi = RU(a op b).
// Ops 8..13: the same in binary32 (operands must be floats; results widened
// exactly), 14: outward conversion of a to binary32.
int oracle_rounded_op(int op, std::uint64_t n, const double* a, const double* b, double* out_lo, double* out_hi);

// Interval operations on n operand pairs, 3 doubles per interval (lb, ub,
// undefined flag 0/1), without compile-time bound information (run-time sign
// decisions only).  op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt (of a), 5 square (of
// a), 6 negation (of a), 7 max, 8 min, 9 splitter point: a = (lb, ub, -),
// b = (i, n, -) -> out lb = point i of n, 10 acos (of a) clamped to [-1, 1].
int oracle_interval_op(int op, std::uint64_t n, const double* a, const double* b, double* out);

// Binary64 arithmetic operations (add/sub/mul/div/sqrt) executed since the last reset.
std::uint64_t oracle_flop_count(void);
void oracle_flop_reset(void);
void oracle_flop_enable(int on);
// The same count for one phase, and the number of child cuboids the phase
// processed: mode 0 non-final static level, 1 final static level, 2 dynamic split
// level, 3 initial-queue level, 4 everything else (level ignored).
int oracle_phase_counts(int mode, int level, std::uint64_t* flops, std::uint64_t* units);
// Binary32 operations of the phase (FP32 prefilter).
int oracle_phase_counts32(int mode, int level, std::uint64_t* flops);

// FP32 prefilter (no counterpart in the reference): with it on, every child of a
// static level is first propagated and row-checked in outward-rounded binary32;
// a child binary32 proves empty or infeasible is a pruned leaf, the others are
// decided in binary64 as usual.  Applies to oracle_run, oracle_run_slice,
// oracle_nonfinal and oracle_final.
int oracle_set_fp32_prefilter(void* handle, int on);
// oracle_run from the initial subdivisions [first, last) of the first dynamic
// variable only (the proof service's unit of work; no counterpart in the reference).
int oracle_set_init_range(void* handle, std::uint64_t first, std::uint64_t last);
// Static-level children of the last computing call that binary32 evaluated / pruned.
int oracle_fp32_stats(void* handle, std::uint64_t* checked, std::uint64_t* decided);
// Bound and constraint values on n boxes (SoA, stride n), as the device's
// sqd_eval_program gives them: bound k's expression on box i to
// bound_values[(i * n_bounds + k) * 3 + {lb, ub, undefined}], the IBool of
// constraint j (the constraint-row roots in ascending node order) to
// cons_values[(i * n_cons + j) * 2 + {definitely, possibly}]; precision 32
// evaluates on the box converted outward to binary32, as the prefilter does.
// The evaluation is the one the bound applications and row checks use
// (EvaluatorT::eval_bound / possibly).
int oracle_eval_program(void* handle, int precision, std::uint64_t n, const double* lb, const double* ub,
                        double* bound_values, std::uint8_t* cons_values);
}
