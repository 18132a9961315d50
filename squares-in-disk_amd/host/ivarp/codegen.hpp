// Generation of the per-program HIP kernel source from a lowered program.
#pragma once
#include <string>
#include "program.hpp"

namespace ivarp {
namespace codegen {

// True if every operation of the program has a GPU implementation.
bool supported(const flat::Program& p);
// HIP source defining the program struct and registering its kernel set.
std::string generate(const flat::Program& p, const std::string& key);

} // namespace codegen
} // namespace ivarp
