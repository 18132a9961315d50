// <ivarp/progress_printer.hpp> of the reference's lemma sources: a progress
// observer printing the frontier size, the cuboid and critical counts and the
// frontier head's first variables (ivarp/include/ivarp/progress_printer.hpp:29-100).
#pragma once
#include <algorithm>
#include <exception>
#include <ostream>
#include <sstream>

#include "ivarp/run_prover.hpp"

namespace ivarp {

class ProgressPrinter {
public:
    ProgressPrinter(std::ostream& output, std::size_t max_num_vars, const ConstraintSystem* arg_names)
        : output(&output), max_num_vars(std::min(max_num_vars, arg_names->num_args())), arg_names(arg_names) {}

    template<typename Context, typename EntryType>
    void observe_progress(const Context&, const EntryType& queue_head, const ProgressInfo& info) {
        std::ostringstream buffer;
        buffer << "Queue size: " << info.queue_size << ", cuboids: " << info.cuboid_count << "(" << info.critical_count
               << ')';
        for(std::size_t i = 0; i < max_num_vars && i < queue_head.elements.size(); ++i)
            buffer << ", " << arg_names->arg_name(i) << ": " << queue_head.elements[i];
        *output << buffer.str() << '\n';
    }
    void observe_done() const noexcept {}
    void observe_error(std::exception_ptr, std::size_t thread_id) {
        *output << "Proof aborting with error thrown by prover thread #" << thread_id << std::endl;
    }

private:
    std::ostream* output;
    std::size_t max_num_vars;
    const ConstraintSystem* arg_names;
};

} // namespace ivarp
