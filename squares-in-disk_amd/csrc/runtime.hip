// Host runtime of the GPU prover core: program binding, device buffers and the
// round loop of the proof driver, exported through the C ABI of include/sqd.h.
//
// The proof runs as rounds over a LIFO frontier resident in HBM
// (prover/proof_driver.hpp:393-438 pops up to dequeue_buffer_size cuboids per
// batch from a LIFO queue): each round pops K dynamic cuboids from the top of
// the device stack, runs the static levels (non-final pass, hull per cuboid),
// sorts the hulls into final (depth >= generation_count) and split lists, runs
// the final pass and the dynamic split, and pushes the children back onto the
// stack.  The count totals do not depend on the processing order.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sqd.h"
#include "kernel_set.hpp"
#include "ivarp/program.hpp"
#include "lemmas.hpp"

namespace sqd {

static std::vector<const KernelSet*>& registry() {
    static std::vector<const KernelSet*> r;
    return r;
}
void register_kernel_set(const KernelSet* ks) { registry().push_back(ks); }
const KernelSet* find_kernel_set(std::uint64_t fp) {
    for(const KernelSet* k : registry()) if(k->fingerprint == fp) return k;
    return nullptr;
}

namespace {

thread_local std::string t_error;

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if(e != hipSuccess) throw Failure(SQD_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template<typename F> int guarded(F&& f) {
    try { f(); return SQD_OK; }
    catch(const Failure& e) { t_error = e.what(); return e.code; }
    catch(const std::invalid_argument& e) { t_error = e.what(); return SQD_ERR_ARGUMENT; }
    catch(const std::exception& e) { t_error = e.what(); return SQD_ERR_INTERNAL; }
}

// Device-resident control block of a core: the counters, every list counter
// and the round's pop range, in one allocation so that one copy brings the
// whole state to the host.
struct DeviceState {
    dev::DeviceCounters counters;
    unsigned long long stack_count, fin_count, spl_count, crit_count, scratch_count;
    unsigned long long round_k, round_base;   // the round pops [round_base, round_base + round_k)
    unsigned long long host_n;                // input count of host-driven calls
    unsigned long long frontier_full;         // nonzero: a round could not pop anything
    unsigned long long level_count[64];       // output counter of (mode, level): mode * 16 + level
    unsigned long long f32_decided;           // static-level children pruned by the FP32 prefilter
    unsigned long long f32_checked;           // static-level children the FP32 prefilter evaluated
    unsigned long long undec_count[32];       // FP32 prefilter: undecided children of (mode, level), mode * 16 + level
    unsigned long long part_n;                // inputs of one part of a final pass repeated in parts
};

static_assert(sizeof(DeviceState) % sizeof(unsigned long long) == 0, "state is a block of 64-bit words");

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
    if(threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
// Start of a proof: the job's control block to zero and its initial runtime
// box (computed once when the core was created, as the reference computes it
// when it builds the prover input, refactor_constraints_impl.hpp:114) as the
// single entry of the initial-queue input list, at depth 0.  Row blockIdx.y
// works on job blockIdx.y.  With `start_flag` (a word per job) only the jobs
// whose flag is set start; the others keep their state and get an empty
// initial-queue input.
__global__ void k_proof_start(unsigned long long* words, unsigned n_words, const double* box, int n_args,
                              double* list_lb, double* list_ub, unsigned* list_depth, unsigned long long list_stride,
                              unsigned long long count_word, const unsigned long long* start_flag,
                              unsigned long long stride) {
    const unsigned long long off = dev::job_offset(stride);
    words = dev::job_ptr(words, off);
    box = dev::job_ptr(box, off);
    list_lb = dev::job_ptr(list_lb, off);
    list_ub = dev::job_ptr(list_ub, off);
    list_depth = dev::job_ptr(list_depth, off);
    start_flag = dev::job_ptr(start_flag, off);
    if(blockIdx.x != 0) return;
    if(start_flag && *start_flag == 0ull) {
        if(threadIdx.x == 0) words[count_word] = 0ull;
        return;
    }
    for(unsigned i = threadIdx.x; i < n_words; i += blockDim.x) words[i] = 0ull;
    for(int a = static_cast<int>(threadIdx.x); a < n_args; a += static_cast<int>(blockDim.x)) {
        list_lb[static_cast<unsigned long long>(a) * list_stride] = box[a];
        list_ub[static_cast<unsigned long long>(a) * list_stride] = box[n_args + a];
    }
    if(threadIdx.x == 0 && list_depth) list_depth[0] = 0u;
    __syncthreads();
    if(threadIdx.x == 0) words[count_word] = 1ull;
}
// Most jobs of one batch (sqd_prove_batch).
constexpr unsigned kMaxJobs = 64;
// The initial-queue range of each job of a batch: [lo[j], hi[j]), and whether
// the job starts with the next proof start (continuous batches start only the
// rows that were admitted since the last round).
struct JobRanges {
    unsigned lo[kMaxJobs];
    unsigned hi[kMaxJobs];
    unsigned start[kMaxJobs];
};
// Kernels below with a `stride` argument run one grid row per job of a batch:
// row blockIdx.y works on job blockIdx.y's memory (job 0's pointers shifted by
// blockIdx.y * stride bytes).
__global__ void k_set_ranges(unsigned long long* p, unsigned long long stride, JobRanges r) {
    p = dev::job_ptr(p, dev::job_offset(stride));
    if(threadIdx.x == 0 && blockIdx.x == 0) {
        p[0] = r.lo[blockIdx.y];
        p[1] = r.hi[blockIdx.y];
        p[2] = r.start[blockIdx.y];
    }
}
// Start of a proof (or of a host-driven call): the whole control block (state
// and per-block count slots) to zero.
__global__ void k_state_reset(unsigned long long* words, unsigned long long n_words, unsigned long long stride) {
    words = dev::job_ptr(words, dev::job_offset(stride));
    for(unsigned long long i = threadIdx.x; blockIdx.x == 0 && i < n_words; i += blockDim.x) words[i] = 0ull;
}
// The control blocks of a batch's jobs side by side in `out` (one copy to the host).
__global__ void k_gather_states(const unsigned long long* st, unsigned long long stride, unsigned long long* out,
                                unsigned words) {
    st = dev::job_ptr(st, dev::job_offset(stride));
    for(unsigned i = threadIdx.x; blockIdx.x == 0 && i < words; i += blockDim.x)
        out[static_cast<unsigned long long>(blockIdx.y) * words + i] = st[i];
}
// Start of a round: pop K = min(k_max, S, room / (fanout - 1)) cuboids from
// the top of the stack (their slots are reused by the children the round
// pushes), and clear the round's list counters.
__global__ void k_round_begin(DeviceState* st, unsigned long long k_max, unsigned long long stack_cap,
                              unsigned long long fanout, unsigned long long stride) {
    st = dev::job_ptr(st, dev::job_offset(stride));
    const unsigned t = threadIdx.x;
    if(blockIdx.x != 0) return;
    if(t == 0) {
        const unsigned long long S = st->stack_count;
        unsigned long long K = S < k_max ? S : k_max;
        if(fanout > 1) {
            const unsigned long long room = stack_cap > S ? stack_cap - S : 0;
            const unsigned long long lim = room / (fanout - 1);
            K = K < lim ? K : lim;
        }
        if(K == 0 && S > 0) st->frontier_full = 1;
        st->round_k = K;
        st->round_base = S - K;
        st->stack_count = S - K;
        st->fin_count = 0;
        st->spl_count = 0;
    }
    if(t < 64) st->level_count[t] = 0;
    if(t < 32) st->undec_count[t] = 0;
}
// Adds the per-block count slots into the state's counters and clears them
// (run before the state is read).  Each block folds kFoldThreads slots of its
// job (one per thread), reduces them across its waves and adds the sums to the
// counters with atomics, so that the slots are read by many compute units at
// once instead of one.
constexpr unsigned kFoldThreads = 256;
__global__ __launch_bounds__(kFoldThreads) void k_fold_counts(DeviceState* st, unsigned long long* slots,
                                                              unsigned long long n_slots, unsigned long long stride) {
    st = dev::job_ptr(st, dev::job_offset(stride));
    slots = dev::job_ptr(slots, dev::job_offset(stride));
    constexpr int W = 5;   // used words of a slot
    __shared__ unsigned long long part[W][kFoldThreads / 64];
    unsigned long long acc[W] = {0, 0, 0, 0, 0};
    for(unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_slots;
        i += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        unsigned long long* s = slots + static_cast<unsigned long long>(dev::kSlotWords) * i;
        unsigned long long v[W];
#pragma unroll
        for(int j = 0; j < W; ++j) v[j] = s[j];
#pragma unroll
        for(int j = 0; j < W; ++j) { acc[j] += v[j]; s[j] = 0ull; }
    }
#pragma unroll
    for(int j = 0; j < W; ++j) acc[j] = dev::wave_sum(acc[j]);
    if(__lane_id() == 0)
        for(int j = 0; j < W; ++j) part[j][threadIdx.x / 64] = acc[j];
    __syncthreads();
    if(threadIdx.x == 0) {
        unsigned long long sum[W] = {0, 0, 0, 0, 0};
        for(unsigned w = 0; w < blockDim.x / 64; ++w)
            for(int j = 0; j < W; ++j) sum[j] += part[j][w];
        if(sum[0]) atomicAdd(&st->counters.cuboids, sum[0]);
        if(sum[1]) atomicAdd(&st->counters.leaves, sum[1]);
        if(sum[2]) atomicAdd(&st->counters.criticals, sum[2]);
        if(sum[3]) atomicAdd(&st->f32_decided, sum[3]);
        if(sum[4]) atomicAdd(&st->f32_checked, sum[4]);
    }
}
// Level counters to zero (between chunks of a host-driven call).
__global__ void k_levels_reset(DeviceState* st) {
    if(blockIdx.x == 0 && threadIdx.x < 64) st->level_count[threadIdx.x] = 0;
    if(blockIdx.x == 0 && threadIdx.x < 32) st->undec_count[threadIdx.x] = 0;
}
// hull record initial values: min-key, max-key, zero-sign lb, zero-sign ub
__global__ void k_init_hull(unsigned long long* p, unsigned long long n_records, unsigned long long stride) {
    p = dev::job_ptr(p, dev::job_offset(stride));
    for(unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_records;
        i += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        p[4 * i + 0] = ~0ull;
        p[4 * i + 1] = 0ull;
        p[4 * i + 2] = ~0ull;
        p[4 * i + 3] = ~0ull;
    }
}

// Serialises graph capture with device allocation and freeing: the runtime may
// order those against the legacy stream, which a capture in progress on another
// thread's stream does not allow.
std::mutex g_capture_mutex;

// Device bytes allocated by dalloc on this thread (a core records the bytes its
// creation allocated).
thread_local std::uint64_t t_dalloc_bytes = 0;
template<typename T> T* dalloc(std::size_t n) {
    void* p = nullptr;
    if(n == 0) n = 1;
    hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    t_dalloc_bytes += n * sizeof(T);
    return static_cast<T*>(p);
}

// Device memory of one job, laid out before it is allocated: take() reserves an
// aligned range and records where its address goes; place() sets every
// recorded pointer once the allocation exists.
struct Layout {
    std::size_t size = 0;
    std::vector<std::pair<std::size_t, std::function<void(char*)>>> fix;
    template<typename T> void take(T*& p, std::size_t n) {
        size = (size + 255) & ~static_cast<std::size_t>(255);
        const std::size_t off = size;
        T** where = &p;
        fix.push_back({off, [where](char* at) { *where = reinterpret_cast<T*>(at); }});
        size += (n ? n : 1) * sizeof(T);
    }
    std::size_t stride() const { return (size + 255) & ~static_cast<std::size_t>(255); }
    void place(char* base) const {
        for(const auto& f : fix) f.second(base + f.first);
    }
};

// A device box list (storage in the core's job memory); its counter lives in
// the DeviceState.
struct DevList {
    int N = 0;
    unsigned long long cap = 0;
    double* lb = nullptr;
    double* ub = nullptr;
    unsigned* depth = nullptr;
    unsigned* origin = nullptr;
    unsigned* leaf = nullptr;
    unsigned long long* count = nullptr;
    void lay(Layout& l, int n_args, unsigned long long capacity, bool with_depth, bool with_origin) {
        N = n_args;
        cap = capacity;
        l.take(lb, static_cast<std::size_t>(cap) * N);
        l.take(ub, static_cast<std::size_t>(cap) * N);
        if(with_depth) l.take(depth, static_cast<std::size_t>(cap));
        if(with_origin) { l.take(origin, static_cast<std::size_t>(cap)); l.take(leaf, static_cast<std::size_t>(cap)); }
    }
    dev::BoxList view() const { return view(count); }
    dev::BoxList view(unsigned long long* counter) const {
        dev::BoxList v;
        v.lb = lb; v.ub = ub; v.depth = depth; v.origin = origin; v.leaf = leaf;
        v.stride = cap; v.cap = cap; v.count = counter;
        return v;
    }
};

std::uint64_t product(const std::vector<std::uint32_t>& v, std::size_t upto) {
    std::uint64_t p = 1;
    for(std::size_t i = 0; i < upto && i < v.size(); ++i) p *= v[i];
    return p;
}

} // namespace
} // namespace sqd

struct sqd_program {
    ivarp::flat::Program prog;
    const sqd::KernelSet* ks = nullptr;
    std::string text;
};

struct sqd_core {
    int device = 0;
    const sqd_program* program = nullptr;
    const sqd::KernelSet* ks = nullptr;
    sqd_settings settings{};
    hipStream_t stream = nullptr;
    int N = 0;
    unsigned cus = 0;            // compute units
    unsigned grid_override = 0;  // level grid from the settings (0: resident capacity)
    double* d_consts = nullptr;
    double* d_box_ct = nullptr;     // 2N: outward compile-time box
    double* d_box_init = nullptr;   // 2N: initial runtime box
    sqd::DeviceState* d_state = nullptr;   // followed by n_slots count slots of dev::kSlotWords words
    sqd::DeviceState* h_state = nullptr;   // pinned host copies of the jobs' blocks (jobs of them)
    // Batches: every pointer above and below marked (job) is job 0's; job j's
    // memory is the same layout job_stride bytes further on, all in d_jobs.
    unsigned jobs = 1;                      // jobs per batch (sqd_settings.jobs)
    unsigned ny = 1;                        // grid rows of the launches being enqueued: 1 or jobs
    unsigned states = 1;                    // jobs whose state the last copy to h_state holds
    unsigned long long job_stride = 0;      // bytes of one job's part of d_jobs
    char* d_jobs = nullptr;                 // the jobs' memory, jobs parts
    unsigned long long* d_gather = nullptr; // jobs control blocks side by side (k_gather_states)
    unsigned long long* d_slots = nullptr;
    unsigned long long n_slots = 0;        // >= the largest grid of a counting kernel
    unsigned long long* d_discard = nullptr;   // count slots that are never folded (final passes repeated in parts)
    unsigned long long* level_slots = nullptr; // the slots the level kernels add to: d_slots, or d_discard
    std::size_t state_bytes = 0;
    unsigned long long* d_hull = nullptr;
    unsigned clean = 0;               // jobs (from 0) whose hull records and count slots are known reset
    unsigned char* d_has = nullptr;
    sqd::DevList stack, lvl[2], fin, spl, crit, scratch;
    std::uint64_t device_bytes = 0;  // device memory allocated by sqd_core_create
    sqd_critical_fn crit_fn = nullptr;       // critical sink (sqd_core_set_critical_sink)
    void* crit_user = nullptr;
    std::vector<double> crit_host;           // one part of the criticals on its way to the sink
    unsigned long long crit_delivered = 0;   // criticals handed to the sink by the current call
    sqd_progress_fn progress_fn = nullptr;   // progress reports of sqd_prove
    void* progress_user = nullptr;
    double progress_start = 0, progress_every = 0;
    std::vector<double> head;        // 2N: the frontier's top for progress reports
    bool fp32 = false;               // some static level runs the FP32-prefiltered kernels
    // sqd_prove as HIP graphs: the initial queue with the first round, and one later round
    hipGraphExec_t g_first = nullptr, g_round = nullptr;
    // continuous batches (sqd_batch_admit / sqd_batch_step): per row 0 free, 1 admitted
    // (starts with the next step), 2 running; the rows' ranges; the most cuboids each
    // row's frontier held after a round; graphs of a step with and without starting rows
    std::vector<unsigned char> row_state;
    std::vector<unsigned> row_first, row_last;
    std::vector<unsigned long long> row_peak;
    hipGraphExec_t g_cont_start = nullptr, g_cont_round = nullptr;
    unsigned proofs = 0;             // sqd_prove calls so far (graphs are captured from the second)
    unsigned long long* d_range = nullptr;   // initial-queue range of the first variable's subdivisions
    unsigned long long* d_undec = nullptr;   // FP32 prefilter: undecided child indices of one level
    unsigned long long undec_cap = 0;
    unsigned long long k_max = 0;    // dynamic cuboids per round
    unsigned long long split_fanout = 1;
    std::vector<std::uint32_t> st_subs, dyn_subs, dyn_initial;
    std::vector<double> box_init;                  // host copy of the compile-time box (uploaded per job)
    unsigned long long range_init[3] = {0, 1, 1};  // host copy of the full initial range (and start flag)
    // profiling: events around each launch, resolved at the next stream sync
    struct Pending { int kind; hipEvent_t a, b; };
    bool profiling = false;
    int profile_only = -1;           // kind to time, or -1 for every kind
    std::vector<hipEvent_t> ev_pool;
    std::size_t ev_used = 0;
    std::vector<Pending> pending;
    std::map<int, sqd_kernel_stat> stats;
};

namespace sqd {
namespace {

// Kernel kinds for the timing statistics: level kernels are mode * 16 + level.
enum : int { KIND_HULL_FINALIZE = 100, KIND_FINAL_NOSTATIC, KIND_INIT_HULL, KIND_HULL_DECODE, KIND_INITIAL_BOX,
             KIND_SET_COUNT };

// A static level's kernels are named after how the core decides it: the
// binary32 prefilter with binary64 for the undecided children, the binary64
// group test, or the plain binary64 pass.
sqd_kernel_stat make_stat(int kind, const sqd_settings& settings) {
    sqd_kernel_stat s{};
    static const char* modes[] = {"NONFINAL", "FINAL", "SPLIT", "INIT"};
    if(kind < 64) {
        s.mode = kind / 16;
        s.level = kind % 16;
        const char* kernel = "dynamic_level";
        if(s.mode <= 1)
            kernel = ((settings.fp32_prefilter >> s.level) & 1u) ? "static_level_fp32"
                   : ((settings.group_levels >> s.level) & 1u)  ? "static_level_group64"
                                                                 : "static_level";
        std::snprintf(s.name, sizeof s.name, "%s<%s,%d>", kernel, modes[s.mode], s.level);
    } else {
        s.mode = -1;
        s.level = -1;
        const char* n = kind == KIND_HULL_FINALIZE ? "hull_finalize" : kind == KIND_FINAL_NOSTATIC ? "final_nostatic"
                      : kind == KIND_INIT_HULL ? "init_hull" : kind == KIND_HULL_DECODE ? "hull_decode"
                      : kind == KIND_INITIAL_BOX ? "initial_box" : "set_count";
        std::snprintf(s.name, sizeof s.name, "%s", n);
    }
    return s;
}

hipEvent_t next_event(sqd_core* c) {
    if(c->ev_used == c->ev_pool.size()) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Launches through `f` (returning hipError_t); with profiling on, brackets the
// launch with two events on the core's stream.
template<typename F> void launch(sqd_core* c, int kind, const char* what, F&& f) {
    hipEvent_t a = nullptr;
    const bool timed = c->profiling && (c->profile_only < 0 || c->profile_only == kind);
    if(timed) {
        a = next_event(c);
        hip_check(hipEventRecord(a, c->stream), "hipEventRecord");
    }
    hip_check(f(), what);
    if(timed) {
        hipEvent_t b = next_event(c);
        hip_check(hipEventRecord(b, c->stream), "hipEventRecord");
        c->pending.push_back({kind, a, b});
    }
}

// Waits for the core's stream and folds the finished launches into the statistics.
void sync(sqd_core* c) {
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for(const auto& p : c->pending) {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        auto it = c->stats.find(p.kind);
        if(it == c->stats.end()) it = c->stats.emplace(p.kind, make_stat(p.kind, c->settings)).first;
        it->second.launches += 1;
        it->second.milliseconds += ms;
    }
    c->pending.clear();
    c->ev_used = 0;
}

void set_count(sqd_core* c, unsigned long long* p, unsigned long long v) {
    launch(c, KIND_SET_COUNT, "k_set_u64", [&] {
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, c->stream, p, v);
        return hipGetLastError();
    });
}
void reset_state(sqd_core* c) {
    launch(c, KIND_SET_COUNT, "k_state_reset", [&] {
        hipLaunchKernelGGL(k_state_reset, dim3(1, c->ny), dim3(256), 0, c->stream, reinterpret_cast<unsigned long long*>(c->d_state),
                           static_cast<unsigned long long>(c->state_bytes / sizeof(unsigned long long)), c->job_stride);
        return hipGetLastError();
    });
}
// Enqueues the fold of the count slots into the counters and the copy of the
// device state (of every job of the launches' rows) to its pinned host copy.
void enqueue_state_copy(sqd_core* c) {
    launch(c, KIND_SET_COUNT, "k_fold_counts", [&] {
        const unsigned blocks = static_cast<unsigned>((c->n_slots + kFoldThreads - 1) / kFoldThreads);
        hipLaunchKernelGGL(k_fold_counts, dim3(blocks, c->ny), dim3(kFoldThreads), 0, c->stream, c->d_state, c->d_slots, c->n_slots,
                           c->job_stride);
        return hipGetLastError();
    });
    if(c->ny == 1) {
        hip_check(hipMemcpyAsync(c->h_state, c->d_state, sizeof(DeviceState), hipMemcpyDeviceToHost, c->stream),
                  "copy state");
        return;
    }
    constexpr unsigned words = sizeof(DeviceState) / sizeof(unsigned long long);
    launch(c, KIND_SET_COUNT, "k_gather_states", [&] {
        hipLaunchKernelGGL(k_gather_states, dim3(1, c->ny), dim3(256), 0, c->stream,
                           reinterpret_cast<const unsigned long long*>(c->d_state), c->job_stride, c->d_gather, words);
        return hipGetLastError();
    });
    hip_check(hipMemcpyAsync(c->h_state, c->d_gather, sizeof(DeviceState) * c->ny, hipMemcpyDeviceToHost, c->stream),
              "copy states");
}
// Waits for the state copy enqueued last and checks the overflow flags (of
// every job: one failing job fails the whole call).
const DeviceState& finish_state(sqd_core* c) {
    sync(c);
    c->states = c->ny;
    for(unsigned j = 0; j < c->ny; ++j) {
        if(c->h_state[j].counters.overflow)
            throw Failure(SQD_ERR_CAPACITY, "a device level buffer overflowed; raise sqd_settings.level_bytes");
        if(c->h_state[j].frontier_full)
            throw Failure(SQD_ERR_CAPACITY, "frontier capacity exhausted; raise sqd_settings.frontier_capacity");
        if(c->h_state[j].crit_count > c->crit.cap && !(c->crit_fn && c->ny == 1))
            throw Failure(SQD_ERR_CAPACITY, "more critical cuboids (" + std::to_string(c->h_state[j].crit_count) +
                                                ") than sqd_settings.critical_capacity (" + std::to_string(c->crit.cap) +
                                                "); set a critical sink (sqd_core_set_critical_sink) to receive every "
                                                "one, or raise the capacity");
    }
    return *c->h_state;
}
// With profiling on: adds the child cuboids of the round whose state was copied
// last to the statistics of the level kernels that timed them.  A static level's
// children are its inputs times its subdivisions; the inputs of level 0 are the
// round's popped cuboids (non-final) or its final list, those of level L the
// survivors of level L - 1; the split levels alike from the split list.
void count_children(sqd_core* c) {
    if(!c->profiling) return;
    auto add = [&](int kind, unsigned long long n) {
        auto it = c->stats.find(kind);
        if(it != c->stats.end()) it->second.children += n;
    };
    const int S = c->ks->NSTAT, D = c->ks->NDYN;
    for(unsigned j = 0; j < c->ny; ++j) {
        const DeviceState& h = c->h_state[j];
        for(int mode : {dev::NONFINAL, dev::FINAL}) {
            unsigned long long in = mode == dev::NONFINAL ? h.round_k : h.fin_count;
            for(int L = 0; L < S; ++L) {
                add(mode * 16 + L, in * c->st_subs[static_cast<std::size_t>(L)]);
                in = h.level_count[mode * 16 + L];
            }
        }
        unsigned long long in = h.spl_count;
        for(int J = 0; J < D; ++J) {
            add(dev::SPLIT * 16 + J, in * c->dyn_subs[static_cast<std::size_t>(J)]);
            in = h.level_count[dev::SPLIT * 16 + J];
        }
    }
}

// Whether some job of the last state copy still has cuboids on its frontier.
bool frontier_left(const sqd_core* c) {
    for(unsigned j = 0; j < c->ny; ++j)
        if(c->h_state[j].stack_count > 0) return true;
    return false;
}
// Folds the count slots and brings the device state to the host (one copy, one sync).
const DeviceState& read_state(sqd_core* c) {
    enqueue_state_copy(c);
    return finish_state(c);
}
sqd_counts counts_of(const DeviceState& h) {
    sqd_counts r;
    r.num_cuboids = h.counters.cuboids;
    r.num_leaf_cuboids = h.counters.leaves;
    r.num_critical_cuboids = h.counters.criticals;
    r.num_repeated_nodes = 0;
    return r;
}

dev::LevelArgs level_args(sqd_core* c, const dev::BoxList& in, const unsigned long long* base,
                          const unsigned long long* count, const dev::BoxList& out) {
    dev::LevelArgs a;
    a.in = in;
    a.out = out;
    a.in_base = base;
    a.in_count = count;
    a.hull = c->d_hull;
    a.counters = &c->d_state->counters;
    a.block_counts = c->level_slots;
    a.consts = c->d_consts;
    a.generations = c->settings.generation_count;
    a.undec = c->d_undec;
    a.undec_count = nullptr;
    a.undec_cap = c->undec_cap;
    a.init_range = c->d_range;
    a.children_per_lane = static_cast<int>(std::min<uint32_t>(c->settings.prefilter_group, 64u));
    a.job_stride = c->job_stride;
    a.jobs = c->ny;
    return a;
}

// Static level L of `mode` over the inputs [*base, *base + *count) of `in`
// (base may be null: 0), at most `max_items` of them, writing `out`.  Its
// undecided-list counter undec_count[mode * 16 + L] must be zero on entry.
void run_static_level(sqd_core* c, int mode, int L, const dev::BoxList& in, const unsigned long long* base,
                      const unsigned long long* count, unsigned long long max_items, const dev::BoxList& out) {
    dev::BoxList src = in;
    if(L == 0) { src.origin = nullptr; src.leaf = nullptr; }  // the input index is the origin
    dev::LevelArgs a = level_args(c, src, base, count, out);
    a.undec_count = &c->d_state->undec_count[mode * 16 + L];
    const unsigned long long children = max_items * c->st_subs[static_cast<std::size_t>(L)];
    launch(c, mode * 16 + L, "static level launch",
           [&] {
               const bool f32 = (c->settings.fp32_prefilter >> L) & 1u;
               const bool grp = (c->settings.group_levels >> L) & 1u;
               const int m = mode | (f32 ? dev::kModeF32 : grp ? dev::kModeGroup : 0);
               return c->ks->static_level(m, L, a, children, c->cus, c->grid_override, c->stream);
           });
}

// Static levels over the inputs [*base, *base + *count) of `in` (base may be
// null: 0), at most `max_items` of them.  NONFINAL joins the leaves into d_hull;
// FINAL appends critical leaves to `crit_out`.  Level L writes lvl[L % 2] with
// the counter level_count[mode * 16 + L], which must be zero on entry.  With
// `levels` >= 0 only the first `levels` levels run.
void run_static(sqd_core* c, int mode, const dev::BoxList& in, const unsigned long long* base,
                const unsigned long long* count, unsigned long long max_items, const dev::BoxList& crit_out,
                int levels = -1) {
    const int S = c->ks->NSTAT;
    const int upto = levels < 0 ? S : std::min(levels, S);
    dev::BoxList prev = in;
    const unsigned long long* prev_base = base;
    const unsigned long long* prev_count = count;
    unsigned long long items = max_items;
    for(int L = 0; L < upto; ++L) {
        const bool last = (L == S - 1);
        dev::BoxList out = last ? crit_out : c->lvl[L % 2].view(&c->d_state->level_count[mode * 16 + L]);
        run_static_level(c, mode, L, prev, prev_base, prev_count, items, out);
        const unsigned long long children = items * c->st_subs[static_cast<std::size_t>(L)];
        prev = out;
        prev_base = nullptr;
        prev_count = out.count;
        items = std::min<unsigned long long>(children, c->lvl[L % 2].cap);
    }
}

// Dynamic levels (SPLIT or INIT) from the inputs [0, *count) of `in` (at most
// `max_items`) to `out`, appending to its counter.
void run_dynamic(sqd_core* c, int mode, const dev::BoxList& in, const unsigned long long* count,
                 unsigned long long max_items, const dev::BoxList& out) {
    const int D = c->ks->NDYN;
    dev::BoxList prev = in;
    const unsigned long long* prev_count = count;
    unsigned long long items = max_items;
    for(int J = 0; J < D; ++J) {
        const bool last = (J == D - 1);
        dev::BoxList o = last ? out : c->lvl[J % 2].view(&c->d_state->level_count[mode * 16 + J]);
        o.origin = nullptr;
        o.leaf = nullptr;
        dev::BoxList src = prev;
        src.origin = nullptr;
        src.leaf = nullptr;
        dev::LevelArgs a = level_args(c, src, nullptr, prev_count, o);
        const unsigned long long children =
            items * (mode == dev::INIT ? c->dyn_initial[static_cast<std::size_t>(J)] : c->dyn_subs[static_cast<std::size_t>(J)]);
        launch(c, mode * 16 + J, "dynamic level launch",
               [&] { return c->ks->dynamic_level(mode, J, a, children, c->cus, c->grid_override, c->stream); });
        prev = o;
        prev_count = o.count;
        items = std::min<unsigned long long>(children, c->lvl[J % 2].cap);
    }
}

void upload_soa(sqd_core* c, DevList& l, const double* lb, const double* ub, unsigned long long n, unsigned depth) {
    if(n > l.cap) throw Failure(SQD_ERR_CAPACITY, "input larger than the device list");
    for(int a = 0; a < c->N; ++a) {
        hip_check(hipMemcpyAsync(l.lb + static_cast<std::size_t>(a) * l.cap, lb + static_cast<std::size_t>(a) * n,
                                 n * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload lb");
        hip_check(hipMemcpyAsync(l.ub + static_cast<std::size_t>(a) * l.cap, ub + static_cast<std::size_t>(a) * n,
                                 n * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload ub");
    }
    std::vector<unsigned> d;
    if(l.depth) {
        d.assign(n, depth);
        hip_check(hipMemcpyAsync(l.depth, d.data(), n * sizeof(unsigned), hipMemcpyHostToDevice, c->stream), "upload depth");
    }
    set_count(c, l.count, n);
    sync(c);
}
void download_soa(sqd_core* c, const DevList& l, unsigned long long n, double* lb, double* ub, unsigned long long stride) {
    for(int a = 0; a < c->N; ++a) {
        hip_check(hipMemcpyAsync(lb + static_cast<std::size_t>(a) * stride, l.lb + static_cast<std::size_t>(a) * l.cap,
                                 n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download lb");
        hip_check(hipMemcpyAsync(ub + static_cast<std::size_t>(a) * stride, l.ub + static_cast<std::size_t>(a) * l.cap,
                                 n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download ub");
    }
    sync(c);
}

// Hull records and count slots are cleared by the kernels that consume them
// (hull finalize / decode, the count fold).  Every computing entry point starts
// with begin_call and marks the core clean when it completes; after a failed
// call both are reset in full before the next use.
void begin_call(sqd_core* c) {
    // any other computing call ends a continuous batch: its rows' jobs are dropped
    std::fill(c->row_state.begin(), c->row_state.end(), static_cast<unsigned char>(0));
    if(c->clean < c->ny) {
        if(c->ks->NSTAT > 0)
            launch(c, KIND_INIT_HULL, "k_init_hull", [&] {
                hipLaunchKernelGGL(k_init_hull, dim3(256, c->ny), dim3(256), 0, c->stream, c->d_hull,
                                   c->k_max * static_cast<unsigned long long>(c->N), c->job_stride);
                return hipGetLastError();
            });
        reset_state(c);
    }
    c->clean = 0;   // until the call completes
}
// A call that completed leaves the hull records and count slots of its jobs reset.
void end_call(sqd_core* c) { c->clean = c->ny; }
void levels_reset(sqd_core* c) {
    launch(c, KIND_SET_COUNT, "k_levels_reset", [&] {
        hipLaunchKernelGGL(k_levels_reset, dim3(1), dim3(64), 0, c->stream, c->d_state);
        return hipGetLastError();
    });
}

// Initial runtime box into d_box_init and, if `to_scratch`, as the single
// entry of the scratch list.  With `reset`, the kernel first clears the whole
// device state (as k_state_reset does).
void launch_initial_box(sqd_core* c, bool to_scratch, bool reset) {
    dev::BoxList list{};
    if(to_scratch) list = c->scratch.view();
    unsigned long long* zero = reset ? reinterpret_cast<unsigned long long*>(c->d_state) : nullptr;
    const unsigned words = reset ? static_cast<unsigned>(sizeof(DeviceState) / sizeof(unsigned long long)) : 0u;
    launch(c, KIND_INITIAL_BOX, "initial box", [&] {
        return c->ks->initial_box(c->d_box_ct, c->d_box_ct + c->N, c->d_consts, c->d_box_init, list, zero, words,
                                  c->ny, c->job_stride, c->stream);
    });
}

// The ranges of the first dynamic variable's initial subdivisions that the next
// initial queues (one per job of the launches' rows) start from.
void set_ranges(sqd_core* c, const JobRanges& r) {
    launch(c, KIND_SET_COUNT, "k_set_ranges", [&] {
        hipLaunchKernelGGL(k_set_ranges, dim3(1, c->ny), dim3(64), 0, c->stream, c->d_range, c->job_stride, r);
        return hipGetLastError();
    });
}
void set_range(sqd_core* c, unsigned long long first, unsigned long long last) {
    JobRanges r{};
    r.lo[0] = static_cast<unsigned>(first);
    r.hi[0] = static_cast<unsigned>(last);
    set_ranges(c, r);
}
unsigned long long full_range(const sqd_core* c) { return c->dyn_initial.empty() ? 1ull : c->dyn_initial[0]; }

// Enqueues a cleared device state and the initial queue on the (empty) stack;
// nothing is waited for.  The initial box is the one computed when the core
// was created (d_box_init).
void enqueue_initial_queue(sqd_core* c, bool only_flagged = false) {
    launch(c, KIND_INITIAL_BOX, "proof start", [&] {
        const unsigned words = static_cast<unsigned>(sizeof(DeviceState) / sizeof(unsigned long long));
        const unsigned long long count_word =
            static_cast<unsigned long long>(reinterpret_cast<char*>(c->scratch.count) - reinterpret_cast<char*>(c->d_state)) /
            sizeof(unsigned long long);
        hipLaunchKernelGGL(k_proof_start, dim3(1, c->ny), dim3(64), 0, c->stream,
                           reinterpret_cast<unsigned long long*>(c->d_state), words, c->d_box_init, c->N, c->scratch.lb,
                           c->scratch.ub, c->scratch.depth, c->scratch.cap, count_word,
                           only_flagged ? c->d_range + 2 : nullptr, c->job_stride);
        return hipGetLastError();
    });
    run_dynamic(c, dev::INIT, c->scratch.view(), c->scratch.count, 1, c->stack.view());
}

// Calls the core's progress callback with the last state brought to the host.
void report_progress(sqd_core* c, unsigned long long rounds, bool done, int status) {
    if(!c->progress_fn) return;
    const DeviceState& h = *c->h_state;
    sqd_progress pr{};
    pr.queue_size = done && status != SQD_OK ? 0 : h.stack_count;
    pr.cuboid_count = h.counters.cuboids;
    pr.leaf_count = h.counters.leaves;
    pr.critical_count = h.counters.criticals;
    pr.rounds = rounds;
    pr.done = done ? 1 : 0;
    pr.status = status;
    if(!done && h.stack_count > 0) {
        c->head.assign(2 * static_cast<std::size_t>(c->N), 0.0);
        const unsigned long long top = h.stack_count - 1;
        for(int a = 0; a < c->N; ++a) {
            hip_check(hipMemcpyAsync(&c->head[static_cast<std::size_t>(a)], c->stack.lb + static_cast<std::size_t>(a) * c->stack.cap + top,
                                     sizeof(double), hipMemcpyDeviceToHost, c->stream), "progress head");
            hip_check(hipMemcpyAsync(&c->head[static_cast<std::size_t>(c->N + a)], c->stack.ub + static_cast<std::size_t>(a) * c->stack.cap + top,
                                     sizeof(double), hipMemcpyDeviceToHost, c->stream), "progress head");
        }
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        pr.head_lb = c->head.data();
        pr.head_ub = c->head.data() + c->N;
    }
    c->progress_fn(c->progress_user, &pr);
}

// Captures the launches and copies that `enqueue` makes on the core's stream
// into an executable HIP graph (kernels and one device-to-host copy; nothing
// that needs a memset node).
template<typename F> hipGraphExec_t capture_graph(sqd_core* c, F&& enqueue) {
    std::lock_guard<std::mutex> lock(g_capture_mutex);
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        enqueue();
    } catch(...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(c->stream, &tmp);
        if(tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    hip_check(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(err, "hipGraphInstantiate");
    return e;
}

// Enqueues one round over the top of the stack: pop (on the device), non-final
// static levels with hull join, hull sorting into final and split lists, final
// pass, dynamic split pushing the children.  Nothing is waited for.
void enqueue_round(sqd_core* c) {
    DeviceState* st = c->d_state;
    launch(c, KIND_SET_COUNT, "k_round_begin", [&] {
        hipLaunchKernelGGL(k_round_begin, dim3(1, c->ny), dim3(64), 0, c->stream, st, c->k_max, c->stack.cap,
                           c->split_fanout, c->job_stride);
        return hipGetLastError();
    });
    const unsigned long long K = c->k_max;   // bound on the round's size
    dev::BoxList items = c->stack.view();
    if(c->ks->NSTAT > 0) run_static(c, dev::NONFINAL, items, &st->round_base, &st->round_k, K, c->crit.view());
    launch(c, KIND_HULL_FINALIZE, "hull finalize", [&] {
        return c->ks->hull_finalize(items, &st->round_base, &st->round_k, K, c->d_hull, c->fin.view(), c->spl.view(),
                                    &st->counters, c->d_slots, c->settings.generation_count, c->ny, c->job_stride,
                                    c->stream);
    });
    if(c->ks->NSTAT > 0) run_static(c, dev::FINAL, c->fin.view(), nullptr, c->fin.count, K, c->crit.view());
    else launch(c, KIND_FINAL_NOSTATIC, "final", [&] {
        return c->ks->final_nostatic(c->fin.view(), c->crit.view(), c->d_slots,
                                     static_cast<unsigned>((K + 255) / 256), c->ny, c->job_stride, c->stream);
    });
    run_dynamic(c, dev::SPLIT, c->spl.view(), c->spl.count, K, c->stack.view());
}

// Hands the criticals [0, k) of the critical list (k <= its capacity) to the sink.
void deliver_criticals(sqd_core* c, unsigned long long k) {
    if(k == 0) return;
    c->crit_host.assign(2 * static_cast<std::size_t>(c->N) * k, 0.0);
    double* lb = c->crit_host.data();
    double* ub = lb + static_cast<std::size_t>(c->N) * k;
    download_soa(c, c->crit, k, lb, ub, k);
    c->crit_delivered += k;
    c->crit_fn(c->crit_user, lb, ub, static_cast<size_t>(k), static_cast<size_t>(k));
}

// Entries [off, cap) of a list as a list of their own, counted by `count`.
dev::BoxList list_from(const dev::BoxList& l, unsigned long long off, unsigned long long* count) {
    dev::BoxList v = l;
    v.lb = l.lb + off;
    v.ub = l.ub + off;
    if(l.depth) v.depth = l.depth + off;
    if(l.origin) v.origin = l.origin + off;
    if(l.leaf) v.leaf = l.leaf + off;
    v.cap = l.cap > off ? l.cap - off : 0;
    v.count = count;
    return v;
}

// Hands the k criticals of the last final pass (over the n_fin boxes of the fin
// list, job 0) to the sink, then empties the critical list.  When the pass found
// more criticals than the list holds, the pass is repeated in parts whose
// criticals fit: the levels before the last over the whole fin list, then the
// last level over slices of its inputs (an input yields at most that level's
// subdivisions in criticals), with the count slots of the repetition discarded so
// that nothing is counted twice.  The repetition runs the same kernels on the
// same boxes, so its criticals are the ones the pass counted.
void drain_criticals(sqd_core* c, unsigned long long k, unsigned long long n_fin) {
    DeviceState* st = c->d_state;
    if(k <= c->crit.cap) {
        deliver_criticals(c, k);
    } else {
        struct Discard {
            sqd_core* c;
            explicit Discard(sqd_core* core) : c(core) { c->level_slots = c->d_discard; }
            ~Discard() { c->level_slots = c->d_slots; }
        } discard(c);
        const int S = c->ks->NSTAT;
        const unsigned long long cap = c->crit.cap;
        dev::BoxList src = c->fin.view();
        unsigned long long n_in = std::min<unsigned long long>(n_fin, c->fin.cap);
        if(S >= 2) {
            levels_reset(c);
            run_static(c, dev::FINAL, c->fin.view(), nullptr, c->fin.count, n_in, c->crit.view(), S - 1);
            n_in = read_state(c).level_count[dev::FINAL * 16 + S - 2];
            src = c->lvl[(S - 2) % 2].view();
        }
        n_in = std::min(n_in, src.cap);
        // a part's children must fit the critical list and, on a prefiltered or
        // grouped last level, the undecided list; without static levels a part
        // is at most a round's final list
        const unsigned long long subs = S > 0 ? c->st_subs[static_cast<std::size_t>(S - 1)] : 1ull;
        const unsigned long long room = c->undec_cap ? std::min(cap, c->undec_cap) : cap;
        unsigned long long per = std::max<unsigned long long>(1, room / subs);
        if(S == 0) per = std::min(per, c->k_max);
        for(unsigned long long off = 0; off < n_in; off += per) {
            const unsigned long long m = std::min(per, n_in - off);
            set_count(c, &st->part_n, m);
            set_count(c, c->crit.count, 0);
            const dev::BoxList part = list_from(src, off, &st->part_n);
            if(S == 0) {
                launch(c, KIND_FINAL_NOSTATIC, "final", [&] {
                    return c->ks->final_nostatic(part, c->crit.view(), c->d_discard,
                                                 static_cast<unsigned>((m + 255) / 256), 1u, 0ull, c->stream);
                });
            } else {
                set_count(c, &st->undec_count[dev::FINAL * 16 + S - 1], 0);
                run_static_level(c, dev::FINAL, S - 1, part, nullptr, &st->part_n, m, c->crit.view());
            }
            deliver_criticals(c, std::min(read_state(c).crit_count, cap));
        }
    }
    set_count(c, c->crit.count, 0);
}

// The most cuboids each job's frontier held after a round (rows of the last state copy).
void track_peaks(sqd_core* c) {
    if(c->row_peak.size() < c->ny) c->row_peak.resize(c->ny, 0ull);
    for(unsigned j = 0; j < c->ny; ++j) c->row_peak[j] = std::max(c->row_peak[j], c->h_state[j].stack_count);
}

// After a round of a proof on a core of one job: hands the round's criticals to
// the sink, if one is set.
void drain_round(sqd_core* c) {
    if(!c->crit_fn || c->ny != 1) return;
    const DeviceState& h = c->h_state[0];
    if(h.crit_count) drain_criticals(c, h.crit_count, h.fin_count);
}

// Proves the initial ranges [first[j], last[j]) of n jobs, leaving each job's
// final state in h_state[j].  A core of one job runs one grid row; a batch core
// always runs all its rows (the jobs beyond n get empty ranges and do no work),
// so that its captured graphs fit every call.  One host round trip per round:
// the round's size is decided on the device, the host only learns whether some
// frontier is left.  From the second proof on a core (the first one fills the
// launch statistics), the rounds are replayed as captured graphs unless
// profiling is on.  Progress reports (single jobs only) follow job 0.
void prove_jobs(sqd_core* c, unsigned n, const uint32_t* first, const uint32_t* last, double* seconds,
                bool progress_reports) {
    const uint32_t all = c->dyn_initial.empty() ? 1u : static_cast<uint32_t>(c->dyn_initial[0]);
    JobRanges r{};
    for(unsigned j = 0; j < n; ++j) {
        if(first[j] >= last[j] || last[j] > all)
            throw std::invalid_argument("initial range must satisfy first < last <= the first variable's subdivisions");
        r.lo[j] = first[j];
        r.hi[j] = last[j];
    }
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    auto t0 = std::chrono::steady_clock::now();
    struct Rows {           // every launch of the call runs the core's rows
        sqd_core* c;
        explicit Rows(sqd_core* core) : c(core) { c->ny = c->jobs; }
        ~Rows() { c->ny = 1; }
    } rows(c);
    set_ranges(c, r);
    begin_call(c);
    c->crit_delivered = 0;
    c->row_peak.assign(c->ny, 0ull);
    const bool graphs = c->settings.graphs && !c->profiling && c->proofs > 0;
    if(graphs && !c->g_first) {
        c->g_first = capture_graph(c, [&] { enqueue_initial_queue(c); enqueue_round(c); enqueue_state_copy(c); });
        c->g_round = capture_graph(c, [&] { enqueue_round(c); enqueue_state_copy(c); });
    }
    unsigned long long rounds = 0;
    auto last_report = std::chrono::steady_clock::now();
    bool reported = false;
    auto progress = [&] {                      // between rounds, when due
        if(!progress_reports || !c->progress_fn) return;
        const auto now = std::chrono::steady_clock::now();
        if(std::chrono::duration<double>(now - t0).count() < c->progress_start) return;
        if(reported && std::chrono::duration<double>(now - last_report).count() < c->progress_every) return;
        reported = true;
        last_report = now;
        report_progress(c, rounds, false, SQD_OK);
    };
    if(graphs) {
        hip_check(hipGraphLaunch(c->g_first, c->stream), "hipGraphLaunch");
        finish_state(c);
        track_peaks(c);
        drain_round(c);
        ++rounds;
        while(frontier_left(c)) {
            progress();
            hip_check(hipGraphLaunch(c->g_round, c->stream), "hipGraphLaunch");
            finish_state(c);
            track_peaks(c);
            drain_round(c);
            ++rounds;
        }
    } else {
        enqueue_initial_queue(c);
        do {
            if(rounds > 0) progress();
            enqueue_round(c);
            read_state(c);
            track_peaks(c);
            count_children(c);
            drain_round(c);
            ++rounds;
        } while(frontier_left(c));
    }
    if(c->crit_fn && c->ny == 1 && c->crit_delivered != c->h_state[0].counters.criticals)
        throw Failure(SQD_ERR_INTERNAL, "criticals handed to the sink differ from the criticals counted");
    ++c->proofs;
    end_call(c);
    if(progress_reports) report_progress(c, rounds, true, SQD_OK);
    auto t1 = std::chrono::steady_clock::now();
    if(seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

// Settings, launch shape and the layout of one job's device memory for a core
// of `program` on `device` (nothing is allocated): the sizes follow from the
// settings, and the rounds are bounded so that no level list and no undecided
// list can overflow.  `state_words` receives the place of the control block in
// the layout.
void configure_core(sqd_core* c, int device, const sqd_program* program, const sqd_settings* settings, Layout& lay,
                    unsigned long long*& state_words) {
    if(!program->ks)
        throw Failure(SQD_ERR_NO_KERNELS, "no kernels are compiled for this program (fingerprint not registered)");
    c->device = device;
    c->program = program;
    c->ks = program->ks;
    c->N = program->prog.nargs;
    if(c->N < 1 || c->N > 64) throw std::invalid_argument("programs need 1 to 64 variables");
    if(settings) c->settings = *settings;
    else sqd_settings_default(&c->settings);
    sqd_settings& s = c->settings;
    if(s.frontier_capacity == 0) s.frontier_capacity = 1ull << 23;
    if(s.level_bytes == 0) s.level_bytes = 1ull << 30;
    if(s.critical_capacity == 0) s.critical_capacity = 1ull << 20;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    c->cus = static_cast<unsigned>(prop.multiProcessorCount);
    c->grid_override = s.grid_blocks;
    const auto& prog = program->prog;
    for(const auto& sp : prog.splits) {
        if(sp.dynamic) { c->dyn_subs.push_back(sp.subs); c->dyn_initial.push_back(sp.initial); }
        else c->st_subs.push_back(sp.subs);
    }
    unsigned used = 0;   // the levels the masks can select
    for(std::size_t L = 0; L < c->st_subs.size() && L < 32; ++L) used |= 1u << L;
    c->fp32 = (s.fp32_prefilter & used) != 0;

    // level buffers hold the survivors of one level; size K for the worst case
    const unsigned long long per_entry = 16ull * static_cast<unsigned long long>(c->N) + 12ull;
    const unsigned long long lvl_cap = std::max<unsigned long long>(1024, s.level_bytes / per_entry);
    unsigned long long worst_static = 1, worst_split = 1;
    for(std::size_t L = 0; L + 1 < c->st_subs.size(); ++L) worst_static = std::max<unsigned long long>(worst_static, product(c->st_subs, L + 1));
    for(std::size_t J = 0; J + 1 < c->dyn_subs.size(); ++J) worst_split = std::max<unsigned long long>(worst_split, product(c->dyn_subs, J + 1));
    unsigned long long worst_init = product(c->dyn_initial, c->dyn_initial.size() > 0 ? c->dyn_initial.size() - 1 : 0);
    if(worst_init > lvl_cap) throw Failure(SQD_ERR_CAPACITY, "initial queue levels do not fit the level buffers");
    c->split_fanout = product(c->dyn_subs, c->dyn_subs.size());
    unsigned long long kmax = s.dequeue_buffer_size ? s.dequeue_buffer_size : 8192;
    kmax = std::min(kmax, lvl_cap / worst_static);
    kmax = std::min(kmax, lvl_cap / worst_split);
    // the FP32 prefilter and the binary64 group test list the children they
    // leave undecided (8 bytes each, level_bytes of them), up to all children
    // of the level, so a round is bounded by the largest such level's children
    // per dynamic cuboid
    const unsigned long long undec_cap = std::max<unsigned long long>(1024, s.level_bytes / 8);
    unsigned long long worst_undec = 1;
    for(std::size_t L = 0; L < c->st_subs.size() && L < 32; ++L)
        if(((s.group_levels | s.fp32_prefilter) >> L) & 1u)
            worst_undec = std::max<unsigned long long>(worst_undec, product(c->st_subs, L + 1));
    kmax = std::min(kmax, undec_cap / worst_undec);
    if(kmax == 0) throw Failure(SQD_ERR_CAPACITY, "level buffers too small for one cuboid");
    c->k_max = kmax;
    // a critical sink receives a final pass in parts of at least one input of
    // the last static level, so the critical list holds at least its children
    if(!c->st_subs.empty()) s.critical_capacity = std::max<unsigned long long>(s.critical_capacity, c->st_subs.back());
    // count slots: one per block of the largest counting grid (level kernels
    // run at most 8 blocks of 256 threads per CU; the small per-round
    // kernels need at most k_max / 256 + 1 blocks)
    c->n_slots = std::max<unsigned long long>(8ull * c->cus, kmax / 256 + 2);
    if(c->grid_override > c->n_slots) c->grid_override = static_cast<unsigned>(c->n_slots);
    c->state_bytes = sizeof(DeviceState) + static_cast<std::size_t>(c->n_slots) * dev::kSlotWords * sizeof(unsigned long long);
    // one job's memory, laid out once; a batch holds `jobs` copies of it
    lay.take(state_words, c->state_bytes / sizeof(unsigned long long));
    lay.take(c->d_discard, static_cast<std::size_t>(c->n_slots) * dev::kSlotWords);
    const unsigned long long stack_cap = std::max<unsigned long long>(s.frontier_capacity, prog.initial_queue_size + 1);
    c->stack.lay(lay, c->N, stack_cap, true, false);
    for(auto& l : c->lvl) l.lay(lay, c->N, lvl_cap, true, true);
    c->fin.lay(lay, c->N, kmax, true, false);
    c->spl.lay(lay, c->N, kmax, true, false);
    c->crit.lay(lay, c->N, s.critical_capacity, false, false);
    c->scratch.lay(lay, c->N, std::max<unsigned long long>(kmax, 1024), true, false);
    lay.take(c->d_hull, static_cast<std::size_t>(kmax) * c->N * 4);
    lay.take(c->d_range, 3);   // first, last, start flag
    if(c->fp32 || (s.group_levels & used)) {
        c->undec_cap = undec_cap;
        lay.take(c->d_undec, static_cast<std::size_t>(undec_cap));
    }
    lay.take(c->d_has, static_cast<std::size_t>(kmax));
    const std::size_t n_consts = 2 * static_cast<std::size_t>(std::max(1, c->ks->NC));
    lay.take(c->d_consts, n_consts);
    lay.take(c->d_box_ct, 2 * static_cast<std::size_t>(c->N));
    lay.take(c->d_box_init, 2 * static_cast<std::size_t>(c->N));
    c->jobs = s.jobs ? s.jobs : 1u;
    if(c->jobs > kMaxJobs) throw std::invalid_argument("sqd_settings.jobs: at most 64 jobs per batch");
    c->job_stride = lay.stride();
    // the sizing rules the level kernels rely on (they also clamp their input
    // counts to the list capacities)
    if(kmax > c->fin.cap || kmax > c->spl.cap || kmax * worst_static > lvl_cap || kmax * worst_split > lvl_cap ||
       (c->undec_cap && kmax * worst_undec > c->undec_cap))
        throw Failure(SQD_ERR_INTERNAL, "core sizing rule violated");
}

// The initial runtime box of every job (refactor_constraints_impl.hpp:61-101),
// once per core: it depends on the program alone.
void compute_initial_boxes(sqd_core* c) {
    const unsigned rows = c->ny;
    c->ny = c->jobs;
    try {
        launch_initial_box(c, false, false);
    } catch(...) {
        c->ny = rows;
        throw;
    }
    c->ny = rows;
}

// Device pointers into the control block once the job memory is placed.
void bind_state(sqd_core* c, unsigned long long* state_words) {
    c->d_state = reinterpret_cast<DeviceState*>(state_words);
    c->d_slots = reinterpret_cast<unsigned long long*>(c->d_state + 1);
    c->level_slots = c->d_slots;
    DeviceState* st = c->d_state;
    c->stack.count = &st->stack_count;
    c->fin.count = &st->fin_count;
    c->spl.count = &st->spl_count;
    c->crit.count = &st->crit_count;
    c->scratch.count = &st->scratch_count;
}

// Every job's copy of the constants, the compile-time box and the full range
// (enqueued on the core's stream from host memory).
void upload_job_constants(sqd_core* c) {
    const auto& prog = c->program->prog;
    c->range_init[0] = 0ull;
    c->range_init[1] = c->dyn_initial.empty() ? 1ull : c->dyn_initial[0];
    c->range_init[2] = 1ull;
    c->box_init.assign(prog.box_lb.begin(), prog.box_lb.end());
    c->box_init.insert(c->box_init.end(), prog.box_ub.begin(), prog.box_ub.end());
    const std::size_t n_consts = 2 * static_cast<std::size_t>(std::max(1, c->ks->NC));
    for(unsigned j = 0; j < c->jobs; ++j) {
        const std::size_t off = static_cast<std::size_t>(c->job_stride) * j;
        auto at = [&](auto* q) { return reinterpret_cast<char*>(q) + off; };
        hip_check(hipMemcpyAsync(at(c->d_range), c->range_init, sizeof c->range_init, hipMemcpyHostToDevice, c->stream),
                  "initial range");
        hip_check(hipMemcpyAsync(at(c->d_consts), c->ks->consts, sizeof(double) * n_consts, hipMemcpyHostToDevice,
                                 c->stream), "upload constants");
        hip_check(hipMemcpyAsync(at(c->d_box_ct), c->box_init.data(), c->box_init.size() * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream), "upload box");
    }
}

} // namespace
} // namespace sqd

namespace sqd {
void set_last_error(const std::string& message) { t_error = message; }
}

// ---------------------------------------------------------------- scratch
// Device memory of sqd_prove_range_noalloc: one job's layout (control block,
// frontier, level buffers, lists, hull records, constants) placed in this
// array instead of an allocation.  Every device the library runs on holds its
// own copy until the process ends.  Host code never passes the array's name to
// a kernel or a copy: its device address is read once per device
// (hipGetSymbolAddress) and kept in g_scratch.
constexpr std::size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace sqd {
namespace {
constexpr int kScratchDevices = 64;
// Per device: the lock that serialises the calls sharing its scratch array,
// the array's device address (set once) and the pageable host copy of the
// control block.
struct ScratchDevice {
    std::mutex m;
    char* ptr = nullptr;
    DeviceState host_state;
};
ScratchDevice g_scratch[kScratchDevices];
} // namespace
} // namespace sqd

using namespace sqd;

extern "C" {

const char* sqd_last_error(void) { return t_error.c_str(); }

void sqd_settings_default(sqd_settings* s) {
    if(!s) return;
    *s = sqd_settings{};
    s->generation_count = 8;
    s->graphs = 1;
}

int sqd_lemma_count(void) { return static_cast<int>(sqd_lemmas::all_lemmas().size()); }
const char* sqd_lemma_key(int i) {
    const auto& l = sqd_lemmas::all_lemmas();
    return (i >= 0 && static_cast<std::size_t>(i) < l.size()) ? l[static_cast<std::size_t>(i)].key.c_str() : nullptr;
}
const char* sqd_lemma_title(int i) {
    const auto& l = sqd_lemmas::all_lemmas();
    return (i >= 0 && static_cast<std::size_t>(i) < l.size()) ? l[static_cast<std::size_t>(i)].title.c_str() : nullptr;
}

static void bind(sqd_program* p) {
    std::ostringstream o;
    p->prog.write(o);
    p->text = o.str();
    p->ks = find_kernel_set(p->prog.fingerprint());
}

int sqd_program_from_text(const char* text, sqd_program** out) {
    return guarded([&] {
        if(!text || !out) throw std::invalid_argument("null argument");
        std::istringstream in(text);
        auto p = std::make_unique<sqd_program>();
        try { p->prog = ivarp::flat::Program::read(in); }
        catch(const std::runtime_error& e) { throw std::invalid_argument(e.what()); }
        bind(p.get());
        *out = p.release();
    });
}

int sqd_program_from_lemma(const char* key, sqd_program** out) {
    return guarded([&] {
        if(!key || !out) throw std::invalid_argument("null argument");
        const auto& d = sqd_lemmas::lemma(key);
        auto p = std::make_unique<sqd_program>();
        p->prog = sqd_lemmas::lower_lemma(d);
        bind(p.get());
        *out = p.release();
    });
}

void sqd_program_destroy(sqd_program* p) { delete p; }

int sqd_program_text(const sqd_program* p, char* buffer, size_t capacity, size_t* len) {
    return guarded([&] {
        if(!p) throw std::invalid_argument("null program");
        if(len) *len = p->text.size() + 1;
        if(buffer && capacity > 0) {
            std::size_t n = std::min(capacity - 1, p->text.size());
            std::memcpy(buffer, p->text.data(), n);
            buffer[n] = 0;
        }
    });
}
int sqd_program_nargs(const sqd_program* p) { return p ? p->prog.nargs : -1; }
int sqd_program_has_kernels(const sqd_program* p) { return (p && p->ks) ? 1 : 0; }

int sqd_core_create(int device, const sqd_program* program, const sqd_settings* settings, sqd_core** out) {
    return guarded([&] {
        if(!program || !out) throw std::invalid_argument("null argument");
        std::unique_lock<std::mutex> alloc_lock(g_capture_mutex);
        struct Destroy {
            std::unique_lock<std::mutex>* lock;
            void operator()(sqd_core* p) const { lock->unlock(); sqd_core_destroy(p); }
        };
        std::unique_ptr<sqd_core, Destroy> c(new sqd_core, Destroy{&alloc_lock});   // frees what was allocated if creation fails
        const std::uint64_t bytes0 = t_dalloc_bytes;
        Layout lay;
        unsigned long long* state_words = nullptr;
        configure_core(c.get(), device, program, settings, lay, state_words);
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        c->d_jobs = dalloc<char>(static_cast<std::size_t>(c->job_stride) * c->jobs);
        lay.place(c->d_jobs);
        bind_state(c.get(), state_words);
        if(c->jobs > 1) c->d_gather = dalloc<unsigned long long>(sizeof(DeviceState) / sizeof(unsigned long long) * c->jobs);
        void* hs = nullptr;
        hip_check(hipHostMalloc(&hs, sizeof(DeviceState) * c->jobs, hipHostMallocDefault), "hipHostMalloc");
        c->h_state = static_cast<DeviceState*>(hs);
        std::memset(c->h_state, 0, sizeof(DeviceState) * c->jobs);
        upload_job_constants(c.get());
        compute_initial_boxes(c.get());
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        c->device_bytes = t_dalloc_bytes - bytes0;
        *out = c.release();
    });
}

void sqd_core_destroy(sqd_core* c) {
    if(!c) return;
    std::lock_guard<std::mutex> lock(g_capture_mutex);
    (void)hipSetDevice(c->device);
    if(c->stream) (void)hipStreamSynchronize(c->stream);
    for(void* p : {static_cast<void*>(c->d_jobs), static_cast<void*>(c->d_gather)})
        if(p) (void)hipFree(p);
    if(c->h_state) (void)hipHostFree(c->h_state);
    if(c->g_first) (void)hipGraphExecDestroy(c->g_first);
    if(c->g_round) (void)hipGraphExecDestroy(c->g_round);
    if(c->g_cont_start) (void)hipGraphExecDestroy(c->g_cont_start);
    if(c->g_cont_round) (void)hipGraphExecDestroy(c->g_cont_round);
    for(hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if(c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int sqd_initial_box(sqd_core* c, double* lb, double* ub) {
    return guarded([&] {
        if(!c || !lb || !ub) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        hip_check(hipMemcpyAsync(lb, c->d_box_init, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipMemcpyAsync(ub, c->d_box_init + c->N, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream), "copy");
        sync(c);
    });
}

int sqd_initial_queue(sqd_core* c, double* lb, double* ub, size_t capacity, size_t* n, sqd_counts* counts) {
    return guarded([&] {
        if(!c || !n) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        set_range(c, 0, full_range(c));
        begin_call(c);
        enqueue_initial_queue(c);
        const DeviceState& h = read_state(c);
        const unsigned long long len = h.stack_count;
        const sqd_counts r = counts_of(h);
        end_call(c);
        *n = len;
        if(lb && ub) download_soa(c, c->stack, std::min<unsigned long long>(len, capacity), lb, ub, capacity);
        if(counts) *counts = r;
    });
}

int sqd_nonfinal(sqd_core* c, const double* lb, const double* ub, size_t n,
                 double* hull_lb, double* hull_ub, uint8_t* has_hull, sqd_counts* counts) {
    return guarded([&] {
        if(!c || (n && (!lb || !ub || !hull_lb || !hull_ub || !has_hull))) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        begin_call(c);
        reset_state(c);
        std::vector<double> clb, cub, hlb, hub;
        for(std::size_t off = 0; off < n; off += c->k_max) {
            const std::size_t m = std::min<std::size_t>(c->k_max, n - off);
            clb.assign(static_cast<std::size_t>(c->N) * m, 0);
            cub.assign(static_cast<std::size_t>(c->N) * m, 0);
            for(int a = 0; a < c->N; ++a)
                for(std::size_t i = 0; i < m; ++i) {
                    clb[static_cast<std::size_t>(a) * m + i] = lb[static_cast<std::size_t>(a) * n + off + i];
                    cub[static_cast<std::size_t>(a) * m + i] = ub[static_cast<std::size_t>(a) * n + off + i];
                }
            upload_soa(c, c->stack, clb.data(), cub.data(), m, 0);
            levels_reset(c);
            if(c->ks->NSTAT > 0) run_static(c, dev::NONFINAL, c->stack.view(), nullptr, c->stack.count, m, c->crit.view());
            launch(c, KIND_HULL_DECODE, "decode", [&] {
                return c->ks->hull_decode(c->stack.view(), m, c->d_hull, c->scratch.view(), c->d_has, c->stream);
            });
            hlb.assign(static_cast<std::size_t>(c->N) * m, 0);
            hub.assign(static_cast<std::size_t>(c->N) * m, 0);
            download_soa(c, c->scratch, m, hlb.data(), hub.data(), m);
            hip_check(hipMemcpyAsync(has_hull + off, c->d_has, m, hipMemcpyDeviceToHost, c->stream), "copy has");
            sync(c);
            for(int a = 0; a < c->N; ++a)
                for(std::size_t i = 0; i < m; ++i) {
                    hull_lb[static_cast<std::size_t>(a) * n + off + i] = hlb[static_cast<std::size_t>(a) * m + i];
                    hull_ub[static_cast<std::size_t>(a) * n + off + i] = hub[static_cast<std::size_t>(a) * m + i];
                }
        }
        sqd_counts r = counts_of(read_state(c));
        end_call(c);
        if(c->ks->NSTAT == 0) r.num_leaf_cuboids += n;  // without static variables each box is its own leaf
        if(counts) *counts = r;
    });
}

int sqd_final(sqd_core* c, const double* lb, const double* ub, size_t n,
              double* crit_lb, double* crit_ub, size_t capacity, size_t* n_critical, sqd_counts* counts) {
    return guarded([&] {
        if(!c || (n && (!lb || !ub))) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        if(n > c->fin.cap) throw Failure(SQD_ERR_CAPACITY, "too many boxes for one final pass");
        begin_call(c);
        reset_state(c);
        upload_soa(c, c->fin, lb, ub, n, 0);
        if(c->ks->NSTAT > 0) run_static(c, dev::FINAL, c->fin.view(), nullptr, c->fin.count, n, c->crit.view());
        else launch(c, KIND_FINAL_NOSTATIC, "final", [&] {
            return c->ks->final_nostatic(c->fin.view(), c->crit.view(), c->d_slots,
                                         static_cast<unsigned>((n + 255) / 256 + 1), 1u, 0ull, c->stream);
        });
        c->crit_delivered = 0;
        const DeviceState& h = read_state(c);
        const unsigned long long k = h.crit_count;
        const sqd_counts r = counts_of(h);
        if(c->crit_fn) {                  // every critical to the sink, none to the output arrays
            drain_criticals(c, k, n);
            sync(c);
            if(c->crit_delivered != r.num_critical_cuboids)
                throw Failure(SQD_ERR_INTERNAL, "criticals handed to the sink differ from the criticals counted");
        }
        end_call(c);
        if(n_critical) *n_critical = c->crit_fn ? r.num_critical_cuboids : k;
        if(crit_lb && crit_ub && !c->crit_fn)
            download_soa(c, c->crit, std::min<unsigned long long>({k, capacity, c->crit.cap}), crit_lb, crit_ub, capacity);
        if(counts) *counts = r;
    });
}

int sqd_split(sqd_core* c, const double* lb, const double* ub, size_t n,
              double* out_lb, double* out_ub, size_t capacity, size_t* n_out, sqd_counts* counts) {
    return guarded([&] {
        if(!c || (n && (!lb || !ub))) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        if(n > c->spl.cap) throw Failure(SQD_ERR_CAPACITY, "too many boxes for one split");
        begin_call(c);
        reset_state(c);
        upload_soa(c, c->spl, lb, ub, n, 1);
        run_dynamic(c, dev::SPLIT, c->spl.view(), c->spl.count, n, c->stack.view());
        const DeviceState& h = read_state(c);
        const unsigned long long k = h.stack_count;
        const sqd_counts r = counts_of(h);
        end_call(c);
        if(n_out) *n_out = k;
        if(out_lb && out_ub) download_soa(c, c->stack, std::min<unsigned long long>(k, capacity), out_lb, out_ub, capacity);
        if(counts) *counts = r;
    });
}

int sqd_prove(sqd_core* c, sqd_counts* counts, double* crit_lb, double* crit_ub,
              size_t capacity, size_t* n_critical, double* seconds) {
    if(!c) { t_error = "null core"; return SQD_ERR_ARGUMENT; }
    const uint32_t all = c->dyn_initial.empty() ? 1u : static_cast<uint32_t>(c->dyn_initial[0]);
    return sqd_prove_range(c, 0, all, counts, crit_lb, crit_ub, capacity, n_critical, seconds);
}

int sqd_prove_range(sqd_core* c, uint32_t first, uint32_t last, sqd_counts* counts, double* crit_lb,
                    double* crit_ub, size_t capacity, size_t* n_critical, double* seconds) {
    const int rc = guarded([&] {
        if(!c) throw std::invalid_argument("null core");
        prove_jobs(c, 1, &first, &last, seconds, true);
        const DeviceState& h = c->h_state[0];
        const bool sink = c->crit_fn && c->jobs == 1;   // the criticals went to the sink
        const unsigned long long k = sink ? h.counters.criticals : h.crit_count;
        if(counts) *counts = counts_of(h);
        if(n_critical) *n_critical = k;
        if(crit_lb && crit_ub && k && !sink)
            download_soa(c, c->crit, std::min<unsigned long long>({k, capacity, c->crit.cap}), crit_lb, crit_ub, capacity);
    });
    if(rc != SQD_OK && c && c->progress_fn) {     // the reference's observe_error
        const std::string msg = t_error;
        try { report_progress(c, 0, true, rc); } catch(...) {}
        t_error = msg;
    }
    return rc;
}

int sqd_prove_batch(sqd_core* c, size_t n_jobs, const uint32_t* first, const uint32_t* last,
                    sqd_job_counts* results, double* seconds) {
    return guarded([&] {
        if(!c || !first || !last || !results) throw std::invalid_argument("null argument");
        if(n_jobs < 1 || n_jobs > c->jobs) throw std::invalid_argument("a batch holds 1 to sqd_settings.jobs jobs");
        prove_jobs(c, static_cast<unsigned>(n_jobs), first, last, seconds, false);
        for(size_t j = 0; j < n_jobs; ++j) {
            results[j].counts = counts_of(c->h_state[j]);
            results[j].n_critical = c->h_state[j].crit_count;
            results[j].fp32_checked = c->h_state[j].f32_checked;
            results[j].fp32_decided = c->h_state[j].f32_decided;
            results[j].frontier_peak = j < c->row_peak.size() ? c->row_peak[j] : 0ull;
        }
    });
}

int sqd_batch_admit(sqd_core* c, uint32_t row, uint32_t first, uint32_t last) {
    return guarded([&] {
        if(!c) throw std::invalid_argument("null core");
        if(row >= c->jobs) throw std::invalid_argument("row beyond the core's batch size");
        if(first >= last || last > full_range(c))
            throw std::invalid_argument("initial range must satisfy first < last <= the first variable's subdivisions");
        if(c->row_state.size() != c->jobs) {
            c->row_state.assign(c->jobs, 0);
            c->row_first.assign(c->jobs, 0u);
            c->row_last.assign(c->jobs, 0u);
        }
        if(c->row_state[row] != 0) throw std::invalid_argument("the row holds a job");
        c->row_state[row] = 1;
        c->row_first[row] = first;
        c->row_last[row] = last;
    });
}

int sqd_batch_step(sqd_core* c, sqd_job_counts* results, uint64_t* finished, uint32_t* active) {
    return guarded([&] {
        if(!c || !results || !finished || !active) throw std::invalid_argument("null argument");
        *finished = 0;
        *active = 0;
        if(c->row_state.size() != c->jobs) return;   // nothing admitted yet
        bool running = false, starting = false;
        for(unsigned char st : c->row_state) { running = running || st == 2; starting = starting || st == 1; }
        if(!running && !starting) return;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        struct Rows {
            sqd_core* c;
            explicit Rows(sqd_core* core) : c(core) { c->ny = c->jobs; }
            ~Rows() { c->ny = 1; }
        } rows(c);
        std::vector<unsigned char> state = c->row_state;   // begin_call below drops the rows' jobs
        try {
            if(!running) {
                begin_call(c);                            // a fresh batch: reset what a failed call left
                c->row_state = state;
                c->row_peak.assign(c->jobs, 0ull);
            } else if(c->clean < c->ny) {
                throw Failure(SQD_ERR_INTERNAL, "continuous batch interrupted");
            }
            c->clean = 0;
            if(starting) {
                JobRanges r{};
                for(unsigned j = 0; j < c->jobs; ++j) {
                    r.lo[j] = c->row_first[j];
                    r.hi[j] = c->row_last[j];
                    r.start[j] = c->row_state[j] == 1 ? 1u : 0u;
                    if(c->row_state[j] == 1) c->row_peak[j] = 0;
                }
                set_ranges(c, r);
            }
            if(c->settings.graphs && !c->profiling) {
                hipGraphExec_t& g = starting ? c->g_cont_start : c->g_cont_round;
                if(!g) g = capture_graph(c, [&] {
                    if(starting) enqueue_initial_queue(c, true);
                    enqueue_round(c);
                    enqueue_state_copy(c);
                });
                hip_check(hipGraphLaunch(g, c->stream), "hipGraphLaunch");
                finish_state(c);
            } else {
                if(starting) enqueue_initial_queue(c, true);
                enqueue_round(c);
                read_state(c);
            }
        } catch(...) {
            std::fill(c->row_state.begin(), c->row_state.end(), static_cast<unsigned char>(0));
            throw;
        }
        track_peaks(c);
        for(unsigned j = 0; j < c->jobs; ++j) {
            if(c->row_state[j] == 0) continue;
            c->row_state[j] = 2;
            const DeviceState& h = c->h_state[j];
            if(h.stack_count > 0) { ++*active; continue; }
            results[j].counts = counts_of(h);
            results[j].n_critical = h.crit_count;
            results[j].fp32_checked = h.f32_checked;
            results[j].fp32_decided = h.f32_decided;
            results[j].frontier_peak = c->row_peak[j];
            *finished |= 1ull << j;
            c->row_state[j] = 0;
        }
        c->clean = c->ny;   // a completed round leaves hull records and count slots reset
    });
}

int sqd_batch_frontier(sqd_core* c, uint64_t* counts, size_t n) {
    return guarded([&] {
        if(!c || (!counts && n)) throw std::invalid_argument("null argument");
        for(size_t j = 0; j < n; ++j) {
            const bool held = j < c->row_state.size() && c->row_state[j] != 0 && j < c->states;
            counts[j] = held ? c->h_state[j].stack_count : 0ull;
        }
    });
}

int sqd_core_job_criticals(sqd_core* c, size_t job, double* crit_lb, double* crit_ub, size_t capacity, size_t* n) {
    return guarded([&] {
        if(!c || !n) throw std::invalid_argument("null argument");
        if(job >= c->jobs) throw std::invalid_argument("job index beyond the core's batch size");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        // the last call's state copy holds the job's count when the call was a batch
        const unsigned long long k = job < c->states ? c->h_state[job].crit_count : 0ull;
        *n = k;
        if(crit_lb && crit_ub && k) {
            DevList l = c->crit;
            const std::size_t off = static_cast<std::size_t>(c->job_stride) * job;
            l.lb = reinterpret_cast<double*>(reinterpret_cast<char*>(l.lb) + off);
            l.ub = reinterpret_cast<double*>(reinterpret_cast<char*>(l.ub) + off);
            download_soa(c, l, std::min<unsigned long long>({k, capacity, c->crit.cap}), crit_lb, crit_ub, capacity);
        }
    });
}

int sqd_core_set_critical_sink(sqd_core* c, sqd_critical_fn fn, void* user) {
    return guarded([&] {
        if(!c) throw std::invalid_argument("null core");
        c->crit_fn = fn;
        c->crit_user = fn ? user : nullptr;
    });
}

int sqd_core_set_progress(sqd_core* c, sqd_progress_fn fn, void* user, double start_after_seconds,
                          double every_seconds) {
    return guarded([&] {
        if(!c) throw std::invalid_argument("null core");
        c->progress_fn = fn;
        c->progress_user = user;
        c->progress_start = start_after_seconds;
        c->progress_every = every_seconds;
    });
}

int sqd_core_criticals(sqd_core* c, double* crit_lb, double* crit_ub, size_t capacity, size_t* n) {
    return guarded([&] {
        if(!c || !n) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        const unsigned long long k = read_state(c).crit_count;
        *n = k;
        if(crit_lb && crit_ub && k)
            download_soa(c, c->crit, std::min<unsigned long long>({k, capacity, c->crit.cap}), crit_lb, crit_ub, capacity);
    });
}

int sqd_core_shape(const sqd_core* c, uint32_t* initial_subdivisions, uint32_t* static_levels) {
    return guarded([&] {
        if(!c || !initial_subdivisions || !static_levels) throw std::invalid_argument("null argument");
        *initial_subdivisions = static_cast<uint32_t>(full_range(c));
        *static_levels = static_cast<uint32_t>(c->st_subs.size());
    });
}

int sqd_core_memory(const sqd_core* c, uint64_t* device_bytes) {
    return guarded([&] {
        if(!c || !device_bytes) throw std::invalid_argument("null argument");
        *device_bytes = c->device_bytes;
    });
}

int sqd_core_acos_fallbacks(sqd_core* c, uint64_t* n) {
    return guarded([&] {
        if(!c || !n) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        unsigned long long v = 0;
        hip_check(c->ks->acos_hard(&v, c->stream), "acos counter");
        sync(c);
        *n = v;
    });
}

int sqd_core_fp32_stats(sqd_core* c, uint64_t* checked, uint64_t* decided) {
    return guarded([&] {
        if(!c || !checked || !decided) throw std::invalid_argument("null argument");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        *checked = c->h_state->f32_checked;
        *decided = c->h_state->f32_decided;
    });
}

int sqd_width_stats(sqd_core* c, const double* lb, const double* ub, size_t n, sqd_width_stat* out,
                    size_t capacity, size_t* n_constraints) {
    return guarded([&] {
        if(!c || !n_constraints || (n && (!lb || !ub))) throw std::invalid_argument("null argument");
        const int nc = c->ks->NCONS;
        *n_constraints = static_cast<size_t>(nc);
        if(!out || nc == 0) return;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        const std::size_t box_bytes = sizeof(double) * static_cast<std::size_t>(c->N) * n;
        double *d_lb = nullptr, *d_ub = nullptr, *d_out = nullptr;
        std::vector<double> h(3 * static_cast<std::size_t>(nc), 0.0);
        std::unique_lock<std::mutex> lock(g_capture_mutex);
        auto release = [&] { for(double* q : {d_lb, d_ub, d_out}) if(q) (void)hipFree(q); };
        try {
            d_lb = dalloc<double>(std::max<std::size_t>(1, static_cast<std::size_t>(c->N) * n));
            d_ub = dalloc<double>(std::max<std::size_t>(1, static_cast<std::size_t>(c->N) * n));
            d_out = dalloc<double>(3 * static_cast<std::size_t>(nc));
            hip_check(hipMemcpyAsync(d_lb, lb, box_bytes, hipMemcpyHostToDevice, c->stream), "upload boxes");
            hip_check(hipMemcpyAsync(d_ub, ub, box_bytes, hipMemcpyHostToDevice, c->stream), "upload boxes");
            hip_check(hipMemcpyAsync(d_out, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream), "clear");
            hip_check(c->ks->width_stats(d_lb, d_ub, n, c->d_consts, d_out, c->stream), "width statistics launch");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            hip_check(hipMemcpyAsync(h.data(), d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "download");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        } catch(...) { release(); throw; }
        release();
        for(int j = 0; j < nc && static_cast<size_t>(j) < capacity; ++j) {
            sqd_width_stat& w = out[j];
            w.constraint = c->ks->cons_id(j);
            w.reserved = 0;
            w.samples = static_cast<uint64_t>(h[3 * j + 1]);
            w.mean_ratio = w.samples ? h[3 * j] / h[3 * j + 1] : 0.0;
            std::uint64_t mb;
            std::memcpy(&mb, &h[3 * j + 2], sizeof mb);
            double mx;
            std::memcpy(&mx, &mb, sizeof mx);
            w.max_ratio = mx;
        }
    });
}

int sqd_eval_program(sqd_core* c, int precision, const double* lb, const double* ub, size_t n, double* bound_values,
                     uint8_t* cons_values, size_t* n_bounds, size_t* n_constraints) {
    return guarded([&] {
        if(!c || !n_bounds || !n_constraints || (n && (!lb || !ub))) throw std::invalid_argument("null argument");
        if(precision != 64 && precision != 32) throw std::invalid_argument("precision must be 64 or 32");
        const std::size_t nb = static_cast<std::size_t>(c->ks->NB), nc = static_cast<std::size_t>(c->ks->NCONS);
        *n_bounds = nb;
        *n_constraints = nc;
        if(!bound_values || !cons_values || n == 0) return;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        const std::size_t box_bytes = sizeof(double) * static_cast<std::size_t>(c->N) * n;
        double *d_lb = nullptr, *d_ub = nullptr, *d_b = nullptr;
        unsigned char* d_c = nullptr;
        std::unique_lock<std::mutex> lock(g_capture_mutex);
        auto release = [&] {
            for(void* q : {static_cast<void*>(d_lb), static_cast<void*>(d_ub), static_cast<void*>(d_b),
                           static_cast<void*>(d_c)})
                if(q) (void)hipFree(q);
        };
        try {
            d_lb = dalloc<double>(static_cast<std::size_t>(c->N) * n);
            d_ub = dalloc<double>(static_cast<std::size_t>(c->N) * n);
            d_b = dalloc<double>(std::max<std::size_t>(1, n * nb * 3));
            d_c = dalloc<unsigned char>(std::max<std::size_t>(1, n * nc * 2));
            hip_check(hipMemcpyAsync(d_lb, lb, box_bytes, hipMemcpyHostToDevice, c->stream), "upload boxes");
            hip_check(hipMemcpyAsync(d_ub, ub, box_bytes, hipMemcpyHostToDevice, c->stream), "upload boxes");
            hip_check(c->ks->eval_program(precision, d_lb, d_ub, n, d_b, d_c, c->stream), "evaluation launch");
            if(nb)
                hip_check(hipMemcpyAsync(bound_values, d_b, n * nb * 3 * sizeof(double), hipMemcpyDeviceToHost,
                                         c->stream), "download bound values");
            if(nc)
                hip_check(hipMemcpyAsync(cons_values, d_c, n * nc * 2, hipMemcpyDeviceToHost, c->stream),
                          "download constraint values");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        } catch(...) { release(); throw; }
        release();
    });
}

int sqd_core_set_profiling(sqd_core* c, int enable, int only_mode, int only_level) {
    return guarded([&] {
        if(!c) throw std::invalid_argument("null core");
        if(only_mode > 3 || (only_mode >= 0 && (only_level < 0 || only_level > 15)))
            throw std::invalid_argument("profiling filter: mode 0..3 and level 0..15");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        sync(c);
        c->profiling = enable != 0;
        c->profile_only = only_mode >= 0 ? only_mode * 16 + only_level : -1;
        if(c->profiling) c->stats.clear();
    });
}

int sqd_core_kernel_stats(const sqd_core* c, sqd_kernel_stat* out, size_t capacity, size_t* count) {
    return guarded([&] {
        if(!c || !count) throw std::invalid_argument("null argument");
        *count = c->stats.size();
        std::size_t i = 0;
        for(const auto& kv : c->stats) {
            if(!out || i >= capacity) break;
            out[i++] = kv.second;
        }
    });
}

int sqd_prove_range_noalloc(int device, const sqd_program* program, const sqd_settings* settings, uint32_t first,
                            uint32_t last, sqd_counts* counts, double* crit_lb, double* crit_ub, size_t capacity,
                            size_t* n_critical) {
    return guarded([&] {
        if(!program) throw std::invalid_argument("null program");
        int n_dev = 0;
        hip_check(hipGetDeviceCount(&n_dev), "hipGetDeviceCount");
        if(device < 0 || device >= n_dev || device >= kScratchDevices) throw std::invalid_argument("no such device");
        ScratchDevice& sd = g_scratch[device];
        std::lock_guard<std::mutex> lock(sd.m);
        hip_check(hipSetDevice(device), "hipSetDevice");
        if(!sd.ptr) {   // set-up, once per device, with the device current
            void* p = nullptr;
            hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
            if(!p) throw Failure(SQD_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave no address");
            sd.ptr = static_cast<char*>(p);
        }
        sqd_settings s;
        if(settings) s = *settings;
        else sqd_settings_default(&s);
        s.jobs = 1;     // one job's layout
        s.graphs = 0;   // a graph would be an allocation
        sqd_core c;     // host-side state only; its stream is the device's null stream
        Layout lay;
        unsigned long long* state_words = nullptr;
        configure_core(&c, device, program, &s, lay, state_words);
        if(c.job_stride > kScratchBytes)
            throw Failure(SQD_ERR_CAPACITY, "these settings need " + std::to_string(c.job_stride) +
                                                " bytes of device memory, the scratch array of sqd_prove_range_noalloc "
                                                "holds " + std::to_string(kScratchBytes) +
                                                "; lower level_bytes, frontier_capacity or critical_capacity");
        lay.place(sd.ptr);
        bind_state(&c, state_words);
        c.h_state = &sd.host_state;
        upload_job_constants(&c);
        compute_initial_boxes(&c);
        // begin_call clears the hull records and the control block (kernels, as
        // the array keeps the last call's contents)
        prove_jobs(&c, 1, &first, &last, nullptr, false);
        const DeviceState& h = c.h_state[0];
        const unsigned long long k = h.crit_count;
        if(counts) *counts = counts_of(h);
        if(n_critical) *n_critical = k;
        if(crit_lb && crit_ub && k)
            download_soa(&c, c.crit, std::min<unsigned long long>({k, capacity, c.crit.cap}), crit_lb, crit_ub, capacity);
    });
}

} // extern "C"
