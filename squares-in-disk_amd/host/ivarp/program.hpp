// The lowered ("flat") form of one lemma: the input of the GPU prover core and
// of the CPU oracle.  It carries everything the reference's ProverInput type
// encodes (ivarp/include/ivarp/prover_input.hpp:52-85): the runtime bound table
// (RBT) with its dependency lists, the runtime constraint table (RCT), the split
// policy and the compile-time box, as a hash-consed DAG of interval operations.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include <ostream>
#include <istream>

namespace ivarp {
namespace flat {

enum class FOp : std::uint8_t {
    Const = 0, Arg, Neg, Add, Sub, Mul, Div, Sqrt, Square, Max, Min, Ite, Acos,
    BoolConst, Lt, Le, Gt, Ge, Eq, Neq, Not, Or, And
};
const char* fop_name(FOp o) noexcept;
bool fop_from_name(const std::string& s, FOp& out) noexcept;
inline bool fop_is_pred(FOp o) noexcept { return static_cast<unsigned>(o) >= static_cast<unsigned>(FOp::BoolConst); }

struct FNode {
    FOp op = FOp::Const;
    std::vector<int> kids;       // node ids (always smaller than this node's id)
    int arg = -1;                // Arg
    double c_lb = 0, c_ub = 0;   // Const: outward double enclosure
    bool c_undef = false;        // Const: possibly undefined
    bool b_def = false, b_poss = false; // BoolConst: IBool value
    // Compile-time bounds (fixed point, 1e-6) of an expression node, or the
    // compile-time truth bounds of a predicate node.
    std::int64_t lb = 0, ub = 0;
    bool plb = false, pub = true;
};

enum class Dir : std::uint8_t { LEQ = 1, GEQ = 2, BOTH = 3 };

struct FBound {
    int target;      // argument the bound tightens
    Dir dir;         // target <= f (LEQ), target >= f (GEQ), target == f (BOTH)
    int expr;        // node id of f
    bool honor_undef;// f's compile-time bounds are both infinite: an undefined value disables the bound
};

struct FSplit {
    int arg;
    std::uint32_t initial;   // subdivisions of the initial queue (dynamic variables)
    std::uint32_t subs;      // subdivisions per generation / per static level
    bool dynamic;
};

struct Program {
    std::string name;
    int nargs = 0;
    std::vector<std::string> arg_names;
    std::vector<FSplit> splits;           // one per argument, dynamic ones first
    std::vector<FNode> nodes;             // topologically ordered DAG
    std::vector<FBound> bounds;           // RBT, argument-major order
    // update[arg][k]: bounds to re-apply after lb (k=0), ub (k=1) or both (k=2) of arg changed
    std::vector<std::vector<std::vector<int>>> update;
    std::vector<std::vector<int>> rows;   // RCT: rows[arg] = predicate node ids; rows[nargs-1] = all
    std::vector<std::int64_t> ct_lb, ct_ub; // compile-time box (fixed point)
    std::vector<double> box_lb, box_ub;   // outward double enclosure of the compile-time box
    std::uint64_t initial_queue_size = 1;

    int num_dynamic() const noexcept {
        int n = 0;
        for(const auto& s : splits) if(s.dynamic) ++n;
        return n;
    }
    // 64-bit FNV-1a over the canonical text form; identifies compiled kernels.
    std::uint64_t fingerprint() const;
    void write(std::ostream& out) const;
    static Program read(std::istream& in);
};

} // namespace flat
} // namespace ivarp
