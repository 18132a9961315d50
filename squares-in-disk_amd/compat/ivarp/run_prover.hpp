// <ivarp/run_prover.hpp> of the reference's lemma sources: ProverSettings,
// ProofInformation and run_prover (ivarp/include/ivarp/run_prover.hpp:57-74,
// prover_input.hpp:124-159, prover/proof_information.hpp:29-44), running the
// proof on the GPU core through the C ABI (include/sqd.h): the lowered program is
// bound to its compiled kernels, proved with the frontier resident in HBM
// (sqd_prove), every critical cuboid goes to the handler (the core's critical
// sink) and progress reports go to the observer.
#pragma once
#include <chrono>
#include <cstddef>
#include <functional>
#include <iomanip>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ivarp/math_fn.hpp"
#include "sqd.h"

namespace ivarp {

struct ProverSettings {
    /// Devices to prove on (the first one is used); {-1}: device 0.
    std::vector<int> cuda_device_ids = {-1};
    /// Generations of dynamic subdivisions.
    unsigned generation_count = 8;
    /// Threads of the CPU prover; unused by the GPU core.
    int thread_count = -1;
    /// Dynamic cuboids per round (-1: chosen by the core).
    int dequeue_buffer_size = -1;
    /// Unused by the GPU core (its blocks have 256 threads).
    int cuda_threads_per_block = 256;
    /// Delay before the first progress report, and between two reports.
    std::chrono::milliseconds start_progress_after{8000};
    std::chrono::milliseconds progress_every{2000};
    /// The GPU core never repeats a node at the same depth (as with 1 here, the
    /// value the reference's lemma driver sets); other values are refused.
    int max_iterations_per_node = -1;
    float iteration_volume_criterion = -1.f;
    float iteration_single_dimension_criterion = -1.f;
};

struct ProofInformation {
    std::size_t num_cuboids{0};
    std::size_t num_leaf_cuboids{0};
    std::size_t num_critical_cuboids{0};
    std::size_t num_repeated_nodes{0};
};

struct ProgressInfo {
    std::size_t queue_size{0};
    std::size_t cuboid_count{0};
    std::size_t leaf_count{0};
    std::size_t critical_count{0};
};

/// An interval of a critical cuboid or of the frontier's head, printed as [lb, ub].
struct IntervalValue {
    double lb, ub;
};
inline std::ostream& operator<<(std::ostream& o, const IntervalValue& v) {
    std::ostringstream s;
    s << std::setprecision(17) << '[' << v.lb << ", " << v.ub << ']';
    return o << s.str();
}

/// The frontier's head handed to a progress observer (its `elements`).
struct QueueHead {
    std::vector<IntervalValue> elements;
};

struct DefaultProgressObserver {
    template<typename Context, typename EntryType>
    void observe_progress(const Context&, const EntryType&, const ProgressInfo&) {}
    void observe_done() const noexcept {}
    void observe_error(std::exception_ptr, std::size_t) {}
};

namespace impl {

inline void sqd_check(int rc) {
    if(rc != SQD_OK) throw std::runtime_error(sqd_last_error());
}

struct ProverHooks {
    std::size_t nargs = 0;
    std::function<void(const std::vector<IntervalValue>&)> critical;
    std::function<void(const sqd_progress&)> progress;

    static void on_critical(void* user, const double* lb, const double* ub, size_t n, size_t stride) {
        auto* h = static_cast<ProverHooks*>(user);
        std::vector<IntervalValue> box(h->nargs);
        for(size_t i = 0; i < n; ++i) {
            for(std::size_t a = 0; a < h->nargs; ++a) box[a] = IntervalValue{lb[a * stride + i], ub[a * stride + i]};
            h->critical(box);
        }
    }
    static void on_progress(void* user, const sqd_progress* p) { static_cast<ProverHooks*>(user)->progress(*p); }
};

// Owns a program and a core for one proof.
struct ProofRun {
    sqd_program* program = nullptr;
    sqd_core* core = nullptr;
    ~ProofRun() {
        if(core) sqd_core_destroy(core);
        if(program) sqd_program_destroy(program);
    }
};

} // namespace impl

/// Proves the lowered statement on the GPU core; true when it ends without
/// critical cuboids.  The handler is called as handler(context, values) for
/// every critical cuboid (values[i]: the interval of variable i), the observer
/// with observe_progress(context, queue_head, info) between rounds after
/// start_progress_after, then observe_done() or, when the proof fails,
/// observe_error() before the exception propagates.
template<typename ProverInputType, typename ProgressObserver, typename OnCritical>
inline bool run_prover(const ProverInputType& input, const OnCritical& handler, ProofInformation* info,
                       const ProverSettings& s, ProgressObserver progress) {
    using Context = DefaultContextWithNumberType<IDouble>;
    if(s.max_iterations_per_node > 1)
        throw std::invalid_argument("the GPU prover core does not repeat nodes (max_iterations_per_node > 1)");
    impl::ProofRun run;
    std::ostringstream text;
    input.program.write(text);
    impl::sqd_check(sqd_program_from_text(text.str().c_str(), &run.program));
    sqd_settings settings;
    sqd_settings_default(&settings);
    settings.generation_count = s.generation_count;
    if(s.dequeue_buffer_size > 0) settings.dequeue_buffer_size = static_cast<uint32_t>(s.dequeue_buffer_size);
    const int device = (!s.cuda_device_ids.empty() && s.cuda_device_ids.front() >= 0) ? s.cuda_device_ids.front() : 0;
    impl::sqd_check(sqd_core_create(device, run.program, &settings, &run.core));
    impl::ProverHooks hooks;
    hooks.nargs = input.num_args();
    hooks.critical = [&](const std::vector<IntervalValue>& box) { handler(Context{}, box); };
    hooks.progress = [&](const sqd_progress& p) {
        if(p.done) return;   // the end is reported below, after the counts
        ProgressInfo pi;
        pi.queue_size = p.queue_size;
        pi.cuboid_count = p.cuboid_count;
        pi.leaf_count = p.leaf_count;
        pi.critical_count = p.critical_count;
        QueueHead head;
        if(p.head_lb)
            for(std::size_t a = 0; a < hooks.nargs; ++a) head.elements.push_back(IntervalValue{p.head_lb[a], p.head_ub[a]});
        progress.observe_progress(Context{}, head, pi);
    };
    impl::sqd_check(sqd_core_set_critical_sink(run.core, &impl::ProverHooks::on_critical, &hooks));
    impl::sqd_check(sqd_core_set_progress(run.core, &impl::ProverHooks::on_progress, &hooks,
                                          static_cast<double>(s.start_progress_after.count()) / 1000.0,
                                          static_cast<double>(s.progress_every.count()) / 1000.0));
    sqd_counts counts{};
    size_t n_critical = 0;
    const int rc = sqd_prove(run.core, &counts, nullptr, nullptr, 0, &n_critical, nullptr);
    if(rc != SQD_OK) {
        progress.observe_error(std::make_exception_ptr(std::runtime_error(sqd_last_error())), 0);
        impl::sqd_check(rc);
    }
    progress.observe_done();
    ProofInformation out;
    out.num_cuboids = counts.num_cuboids;
    out.num_leaf_cuboids = counts.num_leaf_cuboids;
    out.num_critical_cuboids = counts.num_critical_cuboids;
    out.num_repeated_nodes = counts.num_repeated_nodes;
    if(info) *info = out;
    return out.num_critical_cuboids == 0;
}

template<typename ProverInputType, typename OnCritical>
inline bool run_prover(const ProverInputType& input, const OnCritical& handler, ProofInformation* info = nullptr,
                       const ProverSettings& s = {}) {
    return run_prover(input, handler, info, s, DefaultProgressObserver{});
}

template<typename ProverInputType, typename OnCritical>
inline bool run_prover(const ProverInputType& input, const OnCritical& handler, const ProverSettings& s) {
    return run_prover(input, handler, nullptr, s, DefaultProgressObserver{});
}

} // namespace ivarp
