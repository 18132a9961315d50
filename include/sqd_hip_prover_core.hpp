/*
 * sqd_hip_prover_core.hpp - the reference-side binding of the C ABI (sqd.h):
 * a prover core for ivarp's ProofDriver that runs on the MI355X.
 *
 * It implements the duck-typed core concept that ProofDriver drives
 * (ivarp/include/ivarp/prover/proof_driver.hpp:218-627; the CPU core is
 * cpu_prover/cpu_prover_core.hpp:31-260, the CUDA core
 * with_cuda/ivarp_cuda/gpu_prover_core.cuh:199-475) over the C ABI only.  It is
 * included after ivarp's own headers, which supply BasicProverCore,
 * ProofDriverQueueEntry, CuboidCounts and ProverSettings; it would sit next to
 * CPUProverCore, selected in run_prover.hpp:57-74 when a HIP device is present,
 * with one driver thread per device.
 *
 *   handle_cuboids_nonfinal  -> sqd_nonfinal  (hull of the surviving static leaves)
 *   handle_cuboids_final     -> sqd_final     (every critical to on_critical, through
 *                                              the core's critical sink: none is lost
 *                                              however many a call finds)
 *   initialize_per_thread    -> sqd_core_create on device thread_id
 *
 * The statement's tables (runtime bounds, constraints, dynamic bound
 * application) are compiled into the program's kernels: the program is given by
 * its lowered text (sqd_program_from_text) or a built-in key.  Errors become
 * std::runtime_error with sqd_last_error(), which the driver rethrows on the
 * caller as it does for the CUDA core's errors.
 */
#ifndef SQD_HIP_PROVER_CORE_HPP
#define SQD_HIP_PROVER_CORE_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sqd.h"

namespace ivarp {

template<typename ProverInputType, typename OnCritical>
class HipProverCore : public BasicProverCore<ProverInputType, OnCritical> {
public:
    using Context = typename ProverInputType::Context;
    using NumberType = typename Context::NumberType;   // IDouble
    static constexpr std::size_t num_args = ProverInputType::num_args;
    using QueueElement = ProofDriverQueueEntry<NumberType, num_args>;

    struct FromText {};
    // A built-in statement ("one", "four", ...).
    explicit HipProverCore(const char* lemma_key) { check(sqd_program_from_lemma(lemma_key, &program_)); }
    // A lowered program text (the lowering's output for any statement whose
    // kernels are compiled into the library).
    HipProverCore(FromText, const std::string& program_text) {
        check(sqd_program_from_text(program_text.c_str(), &program_));
    }
    HipProverCore(const HipProverCore&) = delete;
    HipProverCore& operator=(const HipProverCore&) = delete;
    ~HipProverCore() {
        for(sqd_core* c : cores_) sqd_core_destroy(c);
        sqd_program_destroy(program_);
    }

    static void replace_default_settings(ProverSettings& s) { s.dequeue_buffer_size = 8192; }

    // set_runtime_bounds / set_runtime_constraints / set_dynamic_bound_application
    // are inherited: the tables are compiled into the program's kernels.

    void initialize(std::size_t num_threads) {
        cores_.assign(num_threads, nullptr);
        sinks_.assign(num_threads, Sink{});
    }
    void initialize_per_thread(std::size_t thread_id, std::size_t /*num_threads*/) {
        sqd_settings s;
        sqd_settings_default(&s);
        s.generation_count = static_cast<uint32_t>(this->settings->generation_count);
        check(sqd_core_create(static_cast<int>(thread_id), program_, &s, &cores_[thread_id]));
        sinks_[thread_id].self = this;
        check(sqd_core_set_critical_sink(cores_[thread_id], &HipProverCore::on_chunk, &sinks_[thread_id]));
    }

    CuboidCounts handle_cuboids_nonfinal(std::size_t t, const std::vector<QueueElement>& in,
                                         std::vector<QueueElement>* out) {
        const std::size_t n = in.size();
        std::vector<double> lb(num_args * n), ub(num_args * n), hl(num_args * n), hu(num_args * n);
        std::vector<uint8_t> has(n);
        to_soa(in, lb, ub);
        sqd_counts c{};
        check(sqd_nonfinal(cores_[t], lb.data(), ub.data(), n, hl.data(), hu.data(), has.data(), &c));
        for(std::size_t i = 0; i < n; ++i) {
            if(!has[i]) continue;
            QueueElement e;
            for(std::size_t a = 0; a < num_args; ++a) e.elements[a] = NumberType{hl[a * n + i], hu[a * n + i]};
            e.depth = in[i].depth;
            out->push_back(e);
        }
        return counts(c);
    }

    // Every critical of the call reaches on_critical (the sink receives them in
    // chunks as the device finds them), whatever sqd_settings.critical_capacity is.
    CuboidCounts handle_cuboids_final(std::size_t t, const std::vector<QueueElement>& in) {
        const std::size_t n = in.size();
        std::vector<double> lb(num_args * n), ub(num_args * n);
        to_soa(in, lb, ub);
        sqd_counts c{};
        size_t k = 0;
        check(sqd_final(cores_[t], lb.data(), ub.data(), n, nullptr, nullptr, 0, &k, &c));
        return counts(c);
    }

private:
    struct Sink {
        HipProverCore* self = nullptr;
    };
    // sqd_critical_fn: one chunk of criticals in SoA layout -> on_critical per box
    static void on_chunk(void* user, const double* lb, const double* ub, size_t n, size_t stride) {
        HipProverCore* self = static_cast<Sink*>(user)->self;
        for(size_t i = 0; i < n; ++i) {
            NumberType box[num_args];
            for(std::size_t a = 0; a < num_args; ++a) box[a] = NumberType{lb[a * stride + i], ub[a * stride + i]};
            (*self->on_critical)(Context{}, box);
        }
    }
    static void check(int rc) {
        if(rc != SQD_OK) throw std::runtime_error(sqd_last_error());
    }
    static CuboidCounts counts(const sqd_counts& c) {
        CuboidCounts r;
        r.num_cuboids = c.num_cuboids;
        r.num_leaf_cuboids = c.num_leaf_cuboids;
        r.num_critical_cuboids = c.num_critical_cuboids;
        r.num_repeated_nodes = c.num_repeated_nodes;
        return r;
    }
    static void to_soa(const std::vector<QueueElement>& in, std::vector<double>& lb, std::vector<double>& ub) {
        const std::size_t n = in.size();
        for(std::size_t i = 0; i < n; ++i)
            for(std::size_t a = 0; a < num_args; ++a) {
                lb[a * n + i] = in[i].elements[a].lb();
                ub[a * n + i] = in[i].elements[a].ub();
            }
    }
    sqd_program* program_ = nullptr;
    std::vector<sqd_core*> cores_;
    std::vector<Sink> sinks_;
};

} // namespace ivarp

#endif /* SQD_HIP_PROVER_CORE_HPP */
