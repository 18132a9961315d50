// Proof service and autotuner, host code over the C ABI of the prover core.
//
// Service: proof jobs (a program plus a range of its first dynamic variable's
// initial subdivisions, see sqd_prove_range) are queued as they arrive and run
// by worker threads, each on a prover core of the job's program with its own
// HIP stream, so that the kernels of several jobs overlap on the device (the
// streams map onto the device's hardware queues) while each job's frontier stays
// in HBM.  Cores are pooled per program and created on demand within a device
// memory budget; an idle core of another program is destroyed to make room.  A
// worker runs a continuous batch of one program on a core: queued jobs of the
// program take its free rows, and a job that arrives while the batch runs joins
// at the next round boundary (sqd_batch_step).
//
// Autotuner: times proofs of a slice of the program's first variable (batches
// of sqd_settings.jobs copies of it on a batch core) and picks, by coordinate
// descent, the FP32-prefilter level mask, the binary64 group-test level mask,
// the children per group, the dynamic cuboids per round (dequeue_buffer_size)
// and the blocks per level launch.
// These settings change the schedule, not the counts (counts do not depend on
// the order or the batching of the frontier); the prefilter could prune a child
// binary64 keeps, so candidates whose counts differ from the plain run's are
// rejected.
//
// Neither has a counterpart in the reference (its ProofDriver runs one proof
// with one core per device, prover/proof_driver.hpp:393-438).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include "sqd.h"

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0, Clock::time_point t1) { return std::chrono::duration<double>(t1 - t0).count(); }

}  // namespace

namespace sqd {
void set_last_error(const std::string& message);   // the message sqd_last_error() returns (runtime.hip)
}

namespace {

int fail(int code, const std::string& msg) {
    sqd::set_last_error(msg);
    return code;
}

// ---------------------------------------------------------------- autotuner
struct Trial {
    sqd_settings s;
    double seconds = 0;
    sqd_counts counts{};
    int status = SQD_OK;
};

bool same_counts(const sqd_counts& a, const sqd_counts& b) {
    return a.num_cuboids == b.num_cuboids && a.num_leaf_cuboids == b.num_leaf_cuboids &&
           a.num_critical_cuboids == b.num_critical_cuboids;
}

// Median time of the proofs of [lo, hi) on a fresh core with settings s: two
// untimed proofs first (the second captures the graphs), then at least `reps`
// and as many as fit in about `budget` seconds (at most 200).  With s.jobs > 1
// each proof is a batch of s.jobs copies of the range (the way the service runs
// them), so that the kernels' share of the time is the batch's.
Trial time_settings(int device, const sqd_program* prog, const sqd_settings& s, uint32_t lo, uint32_t hi, int reps,
                    double budget) {
    Trial t;
    t.s = s;
    sqd_core* core = nullptr;
    t.status = sqd_core_create(device, prog, &s, &core);
    if(t.status != SQD_OK) return t;
    const size_t jobs = s.jobs > 1 ? s.jobs : 1;
    std::vector<uint32_t> first(jobs, lo), last(jobs, hi);
    std::vector<sqd_job_counts> jc(jobs);
    std::vector<double> times;
    int want = reps;
    for(int r = 0; r < want + 2 && t.status == SQD_OK; ++r) {
        const auto t0 = Clock::now();
        t.status = sqd_prove_batch(core, jobs, first.data(), last.data(), jc.data(), nullptr);
        const double dt = seconds_since(t0, Clock::now());
        t.counts = jc[0].counts;
        for(size_t j = 1; j < jobs && t.status == SQD_OK; ++j)
            if(!same_counts(jc[j].counts, jc[0].counts)) t.status = SQD_ERR_INTERNAL;
        if(r == 1 && dt > 0) want = std::max(reps, std::min(200, static_cast<int>(budget / dt)));
        if(r >= 2) times.push_back(dt);
    }
    sqd_core_destroy(core);
    if(t.status == SQD_OK) {
        std::sort(times.begin(), times.end());
        t.seconds = times[times.size() / 2];
    }
    return t;
}

}  // namespace

extern "C" {

void sqd_autotune_options_default(sqd_autotune_options* o) {
    if(!o) return;
    *o = sqd_autotune_options{};
    o->target_seconds = 0.02;
    o->repetitions = 3;
}

int sqd_autotune(int device, const sqd_program* prog, const sqd_settings* base, const sqd_autotune_options* opt,
                 sqd_settings* best, sqd_autotune_report* report) {
    try {
        if(!prog || !best) return fail(SQD_ERR_ARGUMENT, "null argument");
        sqd_autotune_options o;
        sqd_autotune_options_default(&o);
        if(opt) o = *opt;
        if(o.repetitions < 1) o.repetitions = 1;
        sqd_settings s0;
        if(base) s0 = *base;
        else sqd_settings_default(&s0);
        s0.fp32_prefilter = 0;
        s0.group_levels = 0;
        // the tuning workload: the whole range, or a centred slice of it whose
        // proof takes about target_seconds (sized on a core of one job)
        sqd_settings s_probe = s0;
        s_probe.jobs = 1;
        s_probe.graphs = 0;   // a graph capture would make one of the timed proofs slow
        sqd_core* probe = nullptr;
        int rc = sqd_core_create(device, prog, &s_probe, &probe);
        if(rc != SQD_OK) return rc;
        uint32_t all = 0, nstat = 0;
        rc = sqd_core_shape(probe, &all, &nstat);
        // start from the middle sixteenth, then halve or double the slice
        uint32_t w = std::max<uint32_t>(1, all / 16);
        auto centre = [&](uint32_t width, uint32_t& a, uint32_t& b) {
            a = (all - width) / 2;
            b = a + width;
        };
        uint32_t lo = 0, hi = all;
        centre(w, lo, hi);
        if(rc == SQD_OK) {   // one untimed proof first: the first launches of a process load code
            sqd_counts c{};
            size_t nc = 0;
            rc = sqd_prove_range(probe, lo, hi, &c, nullptr, nullptr, 0, &nc, nullptr);
        }
        if(rc == SQD_OK) {
            for(int step = 0; step < 12; ++step) {
                double dt = 0;
                for(int r = 0; r < 2 && rc == SQD_OK; ++r) {   // the faster of two proofs
                    sqd_counts c{};
                    size_t nc = 0;
                    const auto t0 = Clock::now();
                    rc = sqd_prove_range(probe, lo, hi, &c, nullptr, nullptr, 0, &nc, nullptr);
                    const double d = seconds_since(t0, Clock::now());
                    dt = r == 0 ? d : std::min(dt, d);
                }
                if(rc != SQD_OK) break;
                if(dt > o.target_seconds && w > 1) w = std::max<uint32_t>(1, w / 2);
                else if(dt < o.target_seconds / 4 && w < all) w = std::min(all, w * 2);
                else break;
                centre(w, lo, hi);
            }
        }
        unsigned cus = 0;
        {
            hipDeviceProp_t prop;
            if(hipGetDeviceProperties(&prop, device) == hipSuccess) cus = static_cast<unsigned>(prop.multiProcessorCount);
        }
        sqd_core_destroy(probe);
        if(rc != SQD_OK) return rc;

        int trials = 0;
        Trial plain = time_settings(device, prog, s0, lo, hi, o.repetitions, o.target_seconds);
        ++trials;
        if(plain.status != SQD_OK) return plain.status;
        Trial cur = plain;
        auto consider = [&](sqd_settings s) {
            Trial t = time_settings(device, prog, s, lo, hi, o.repetitions, o.target_seconds);
            ++trials;
            if(t.status == SQD_OK && same_counts(t.counts, plain.counts) && t.seconds < cur.seconds) cur = t;
        };
        // 1. FP32 prefilter: every subset of the static levels
        const uint32_t masks = nstat >= 8 ? 256u : (1u << nstat);
        const sqd_settings s_mask = cur.s;
        for(uint32_t m = 1; m < masks; ++m) {
            sqd_settings s = s_mask;
            s.fp32_prefilter = m;
            consider(s);
        }
        // 2. binary64 group test on the levels the prefilter does not take, and
        //    (unless fp32_first) on its own, without the prefilter
        const sqd_settings s_grp = cur.s;
        for(uint32_t m = 1; m < masks; ++m) {
            if(s_grp.fp32_prefilter && !(m & s_grp.fp32_prefilter)) {
                sqd_settings s = s_grp;
                s.group_levels = m;
                consider(s);
            }
            if(!o.fp32_first || !s_grp.fp32_prefilter) {
                sqd_settings s = s_mask;
                s.group_levels = m;
                consider(s);
            }
        }
        // 3. children per group (when a level groups them)
        if(cur.s.fp32_prefilter || cur.s.group_levels) {
            const sqd_settings s_p = cur.s;
            for(uint32_t g : {1u, 2u, 4u, 8u, 16u, 32u, 64u}) {
                sqd_settings s = s_p;
                s.prefilter_group = g;
                consider(s);
            }
        }
        // 4. dynamic cuboids per round
        const sqd_settings s_k = cur.s;
        for(uint32_t k : {1024u, 4096u, 16384u, 65536u}) {
            sqd_settings s = s_k;
            s.dequeue_buffer_size = k;
            consider(s);
        }
        // 5. blocks per level launch (0: the kernel's resident capacity)
        const sqd_settings s_g = cur.s;
        for(uint32_t g : {cus, 2 * cus, 4 * cus}) {
            if(g == 0) continue;
            sqd_settings s = s_g;
            s.grid_blocks = g;
            consider(s);
        }
        *best = cur.s;
        if(report) {
            report->first = lo;
            report->last = hi;
            report->trials = static_cast<uint32_t>(trials);
            report->plain_seconds = plain.seconds;
            report->best_seconds = cur.seconds;
            report->counts = cur.counts;
        }
        return SQD_OK;
    } catch(const std::exception& e) {
        return fail(SQD_ERR_INTERNAL, e.what());
    }
}

}  // extern "C"

// ---------------------------------------------------------------- service
struct sqd_service {
    sqd_service_options opt{};
    std::mutex m;
    std::condition_variable cv;        // queue, cores and job state changes
    struct Job {
        const sqd_program* prog = nullptr;
        uint32_t first = 0, last = 0;
        Clock::time_point submitted, started, finished;
        bool done = false;
        bool joined = false;   // admitted into a batch that was already running
        sqd_job_result result{};
        std::string error;
    };
    std::map<uint64_t, Job> jobs;
    std::deque<uint64_t> queue;
    uint64_t next_id = 1;
    struct Slot {
        uint64_t id;
        const sqd_program* prog;
        sqd_core* core;       // nullptr while the worker that reserved the slot creates it
        uint64_t bytes;
        bool busy;
    };
    std::vector<Slot> pool;
    uint64_t next_slot = 1;
    uint64_t bytes_in_use = 0;
    std::map<const sqd_program*, uint64_t> core_bytes;     // measured size of a core per program
    std::map<const sqd_program*, sqd_settings> settings;   // per program (autotuned or the defaults)
    std::map<const sqd_program*, bool> tuning;             // autotune in progress
    // continuous batches running per program: their program's queued jobs join
    // them as rows free up, so another worker starts a further batch (and a
    // further core) of the program only when a whole batch of its jobs waits
    std::map<const sqd_program*, int> running_progs;
    // admission by frontier bytes (opt.frontier_budget): the frontier bytes of
    // the running jobs (each job's cuboids after its last round, or the estimate
    // it was admitted with before its first), and per program the most frontier
    // bytes per initial subdivision that a finished job held (the estimate)
    uint64_t frontier_in_flight = 0;
    std::map<const sqd_program*, double> frontier_per_unit;
    sqd_service_stats stats{};
    bool stop = false;
    std::vector<std::thread> workers;

    sqd_settings base_settings() const {
        sqd_settings s;
        sqd_settings_default(&s);
        s.level_bytes = opt.level_bytes;
        s.frontier_capacity = opt.frontier_capacity;
        s.fp32_prefilter = opt.fp32_prefilter;
        return s;
    }

    Slot* find_slot(uint64_t id) {
        for(auto& s : pool)
            if(s.id == id) return &s;
        return nullptr;
    }

    // With the lock held: reserves a free core of `prog` (core set) or room for
    // a new one within the memory budget (core == nullptr: the caller creates
    // it), destroying idle cores of other programs or waiting as needed.
    // Returns the slot id, 0 when the service stops.
    uint64_t acquire(std::unique_lock<std::mutex>& lk, const sqd_program* prog) {
        for(;;) {
            if(stop) return 0;
            for(auto& s : pool)
                if(s.prog == prog && !s.busy && s.core) {
                    s.busy = true;
                    return s.id;
                }
            const uint64_t need = core_bytes.count(prog) ? core_bytes[prog]
                                                         : opt.core_bytes_estimate * opt.batch_jobs;
            if(bytes_in_use + need <= opt.memory_budget || pool.empty()) {
                pool.push_back(Slot{next_slot++, prog, nullptr, need, true});
                bytes_in_use += need;
                return pool.back().id;
            }
            // make room: destroy an idle core of another program
            bool freed = false;
            for(size_t i = 0; i < pool.size(); ++i)
                if(!pool[i].busy && pool[i].core) {
                    sqd_core* c = pool[i].core;
                    bytes_in_use -= pool[i].bytes;
                    pool.erase(pool.begin() + static_cast<std::ptrdiff_t>(i));
                    lk.unlock();
                    sqd_core_destroy(c);
                    lk.lock();
                    ++stats.cores_destroyed;
                    freed = true;
                    break;
                }
            if(!freed) cv.wait(lk);
        }
    }

    // Frontier bytes of one cuboid of `prog` (its bounds and its depth).
    static uint64_t cuboid_bytes(const sqd_program* prog) {
        return 16ull * static_cast<uint64_t>(std::max(1, sqd_program_nargs(prog))) + 4ull;
    }
    // With the lock held: the frontier bytes job `id` is admitted with (at
    // least one cuboid per initial subdivision of its range).
    uint64_t frontier_estimate(uint64_t id) {
        const Job& j = jobs[id];
        const double range = static_cast<double>(j.last - j.first);
        double per = static_cast<double>(cuboid_bytes(j.prog));
        auto f = frontier_per_unit.find(j.prog);
        if(f != frontier_per_unit.end()) per = std::max(per, f->second);
        return static_cast<uint64_t>(per * range);
    }
    // With the lock held: whether job `id` fits the frontier budget now (a job
    // is always admitted when no job runs, so that every job ends).
    bool frontier_fits(uint64_t id) {
        return opt.frontier_budget == 0 || stats.rows_running == 0 ||
               frontier_in_flight + frontier_estimate(id) <= opt.frontier_budget;
    }
    void frontier_add(int64_t delta) {
        frontier_in_flight = static_cast<uint64_t>(static_cast<int64_t>(frontier_in_flight) + delta);
        stats.frontier_bytes = frontier_in_flight;
        stats.frontier_bytes_peak = std::max(stats.frontier_bytes_peak, frontier_in_flight);
    }
    // With the lock held: the oldest queued job of `prog` (0: none), left queued.
    uint64_t peek_queued(const sqd_program* prog) {
        for(uint64_t id : queue)
            if(jobs[id].prog == prog) return id;
        return 0;
    }

    // With the lock held: the oldest queued job of `prog`, taken off the queue (0: none).
    uint64_t take_queued(const sqd_program* prog) {
        for(auto it = queue.begin(); it != queue.end(); ++it)
            if(jobs[*it].prog == prog) {
                const uint64_t id = *it;
                queue.erase(it);
                return id;
            }
        return 0;
    }

    // With the lock held: completes job `id` with its result.
    void complete(uint64_t id, const sqd_job_counts& jc, int status, const std::string& error) {
        Job& j = jobs[id];
        j.finished = Clock::now();
        sqd_job_result r{};
        r.counts = jc.counts;
        r.n_critical = jc.n_critical;
        r.frontier_peak = jc.frontier_peak;
        r.status = status;
        r.joined_running = j.joined ? 1u : 0u;
        r.queued_seconds = seconds_since(j.submitted, j.started);
        r.run_seconds = seconds_since(j.started, j.finished);
        r.latency_seconds = seconds_since(j.submitted, j.finished);
        j.result = r;
        j.error = error;
        j.done = true;
        ++stats.jobs_completed;
    }

    // A worker takes the program of the oldest queued job, a core of that
    // program, and runs a continuous batch on it: queued jobs of the program
    // fill the core's free rows, the batch advances a round at a time
    // (sqd_batch_step), a job whose frontier empties completes and frees its
    // row, and jobs that arrived meanwhile join at the next round boundary.  The
    // batch ends when no row holds a job.
    void run_worker(int device) {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(m);
        for(;;) {
            // the program of the oldest queued job that no batch runs, or that has
            // at least a batch of jobs queued and an idle core (a new core costs
            // its allocation and first launches: only for a backlog of four
            // batches)
            const sqd_program* prog = nullptr;
            cv.wait(lk, [&] {
                if(stop) return true;
                std::map<const sqd_program*, uint32_t> waiting;
                for(uint64_t id : queue) ++waiting[jobs[id].prog];
                for(uint64_t id : queue) {
                    const sqd_program* p = jobs[id].prog;
                    if(!frontier_fits(id)) continue;
                    auto r = running_progs.find(p);
                    if(r == running_progs.end() || r->second == 0) { prog = p; return true; }
                    if(waiting[p] < opt.batch_jobs) continue;
                    bool idle = false;
                    for(const Slot& sl : pool) idle = idle || (sl.prog == p && !sl.busy && sl.core);
                    if(idle || waiting[p] >= 4 * opt.batch_jobs) { prog = p; return true; }
                }
                return false;
            });
            if(stop) return;
            ++running_progs[prog];
            struct Release {                   // every exit below holds the lock
                sqd_service* s;
                const sqd_program* p;
                ~Release() {
                    if(--s->running_progs[p] == 0) s->running_progs.erase(p);
                    s->cv.notify_all();
                }
            } release{this, prog};
            // settings of the program (tuned once, by the first worker that needs them)
            if(!settings.count(prog)) {
                if(opt.autotune && !tuning[prog]) {
                    tuning[prog] = true;
                    lk.unlock();
                    sqd_settings best = base_settings(), base = base_settings();
                    base.jobs = opt.batch_jobs;       // tuned on batches as the workers run them
                    if(sqd_autotune(device, prog, &base, nullptr, &best, nullptr) != SQD_OK) best = base_settings();
                    best.level_bytes = opt.level_bytes;
                    best.frontier_capacity = opt.frontier_capacity;
                    lk.lock();
                    settings[prog] = best;
                    tuning[prog] = false;
                    cv.notify_all();
                } else if(opt.autotune) {
                    cv.wait(lk, [&] { return stop || settings.count(prog) != 0; });
                    if(stop) return;
                } else {
                    settings[prog] = base_settings();
                }
            }
            bool queued = false;   // another worker may have taken the program's jobs meanwhile
            for(uint64_t id : queue) queued = queued || jobs[id].prog == prog;
            if(!queued) continue;
            const uint64_t slot = acquire(lk, prog);
            if(slot == 0) return;
            sqd_settings s = settings[prog];
            s.jobs = opt.batch_jobs;
            sqd_core* core = find_slot(slot)->core;
            if(!core) {
                lk.unlock();
                const int rc = sqd_core_create(device, prog, &s, &core);
                const std::string err = rc != SQD_OK ? sqd_last_error() : std::string();
                lk.lock();
                if(rc == SQD_OK) ++stats.cores_created;
                if(rc != SQD_OK) {                  // creation failed: give the room back, fail one job
                    Slot* sl = find_slot(slot);
                    bytes_in_use -= sl->bytes;
                    pool.erase(pool.begin() + (sl - pool.data()));
                    const uint64_t id = take_queued(prog);
                    if(id) {
                        jobs[id].started = Clock::now();
                        complete(id, sqd_job_counts{}, rc, err);
                    }
                    cv.notify_all();
                    continue;
                }
            }
            const uint32_t rows = opt.batch_jobs;
            std::vector<uint64_t> row_job(rows, 0);
            std::vector<uint32_t> row_first(rows, 0), row_last(rows, 0);   // read without the lock
            std::vector<uint64_t> row_bytes(rows, 0);                       // frontier bytes counted
            std::vector<uint64_t> row_cuboids(rows, 0);
            const uint64_t bpc = cuboid_bytes(prog);
            uint32_t running = 0;
            auto admit = [&] {   // with the lock held: queued jobs of prog into free rows
                const bool joining = running > 0;
                for(uint32_t r = 0; r < rows; ++r) {
                    if(row_job[r]) continue;
                    const uint64_t next = peek_queued(prog);
                    if(!next || !frontier_fits(next)) break;
                    const uint64_t id = take_queued(prog);
                    Job& j = jobs[id];
                    j.started = Clock::now();
                    j.joined = joining;
                    if(sqd_batch_admit(core, r, j.first, j.last) != SQD_OK) {
                        complete(id, sqd_job_counts{}, SQD_ERR_ARGUMENT, sqd_last_error());
                        continue;
                    }
                    row_job[r] = id;
                    row_first[r] = j.first;
                    row_last[r] = j.last;
                    row_bytes[r] = frontier_estimate(id);
                    frontier_add(static_cast<int64_t>(row_bytes[r]));
                    ++running;
                    ++stats.rows_running;
                    stats.rows_peak = std::max(stats.rows_peak, stats.rows_running);
                }
            };
            admit();
            std::vector<sqd_job_counts> jc(rows);
            while(running > 0) {
                lk.unlock();
                uint64_t finished = 0;
                uint32_t active = 0;
                const int rc = sqd_batch_step(core, jc.data(), &finished, &active);
                if(rc == SQD_OK) sqd_batch_frontier(core, row_cuboids.data(), rows);
                std::vector<int> status(rows, SQD_OK);
                std::vector<std::string> errors(rows);
                if(rc != SQD_OK) {
                    // one job's buffers overflowed and the step dropped every row's job:
                    // prove each alone to tell which one failed
                    finished = 0;
                    for(uint32_t r = 0; r < rows; ++r) {
                        if(!row_job[r]) continue;
                        const uint32_t first = row_first[r], last = row_last[r];
                        jc[r] = sqd_job_counts{};
                        status[r] = sqd_prove_batch(core, 1, &first, &last, &jc[r], nullptr);
                        if(status[r] != SQD_OK) errors[r] = sqd_last_error();
                        finished |= 1ull << r;
                    }
                }
                lk.lock();
                for(uint32_t r = 0; r < rows; ++r) {
                    if(!row_job[r]) continue;
                    if(!((finished >> r) & 1ull)) {          // still running: its frontier now
                        const uint64_t now = row_cuboids[r] * bpc;
                        frontier_add(static_cast<int64_t>(now) - static_cast<int64_t>(row_bytes[r]));
                        row_bytes[r] = now;
                        continue;
                    }
                    if(status[r] == SQD_OK && row_last[r] > row_first[r]) {
                        double& u = frontier_per_unit[prog];
                        u = std::max(u, static_cast<double>(jc[r].frontier_peak * bpc) / (row_last[r] - row_first[r]));
                    }
                    frontier_add(-static_cast<int64_t>(row_bytes[r]));
                    row_bytes[r] = 0;
                    complete(row_job[r], jc[r], status[r], errors[r]);
                    row_job[r] = 0;
                    --running;
                    --stats.rows_running;
                }
                cv.notify_all();
                admit();   // jobs that arrived while the round ran join at this boundary
            }
            uint64_t bytes = 0;
            sqd_core_memory(core, &bytes);
            Slot* sl = find_slot(slot);          // busy slots are never removed by others
            if(bytes) {
                bytes_in_use = bytes_in_use - sl->bytes + bytes;
                sl->bytes = bytes;
                core_bytes[prog] = bytes;
            }
            sl->core = core;
            sl->busy = false;
            cv.notify_all();
        }
    }
};

extern "C" {

void sqd_service_options_default(sqd_service_options* o) {
    if(!o) return;
    *o = sqd_service_options{};
    o->device = 0;
    o->workers = 8;
    o->memory_budget = 128ull << 30;
    o->fp32_prefilter = SQD_FP32_ALL_LEVELS;
    o->autotune = 0;
    o->level_bytes = 256ull << 20;
    o->frontier_capacity = 1ull << 22;
    o->core_bytes_estimate = 2ull << 30;
    o->batch_jobs = 8;
}

int sqd_service_create(const sqd_service_options* opt, sqd_service** out) {
    try {
        if(!out) return fail(SQD_ERR_ARGUMENT, "null argument");
        auto* s = new sqd_service;
        sqd_service_options_default(&s->opt);
        if(opt) s->opt = *opt;
        if(s->opt.workers < 1) s->opt.workers = 1;
        if(s->opt.workers > 64) s->opt.workers = 64;
        if(s->opt.core_bytes_estimate == 0) s->opt.core_bytes_estimate = 2ull << 30;
        if(s->opt.batch_jobs < 1) s->opt.batch_jobs = 1;
        if(s->opt.batch_jobs > 64) s->opt.batch_jobs = 64;
        const int device = s->opt.device;
        for(uint32_t i = 0; i < s->opt.workers; ++i) s->workers.emplace_back([s, device] { s->run_worker(device); });
        *out = s;
        return SQD_OK;
    } catch(const std::exception& e) {
        return fail(SQD_ERR_INTERNAL, e.what());
    }
}

int sqd_service_get_settings(sqd_service* s, const sqd_program* prog, int* known, sqd_settings* out) {
    if(!s || !prog || !known || !out) return fail(SQD_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->settings.find(prog);
    *known = it != s->settings.end() ? 1 : 0;
    if(*known) {
        *out = it->second;
        out->jobs = s->opt.batch_jobs;
    }
    return SQD_OK;
}

int sqd_service_set_settings(sqd_service* s, const sqd_program* prog, const sqd_settings* settings) {
    if(!s || !prog || !settings) return fail(SQD_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    if(s->tuning[prog]) return fail(SQD_ERR_ARGUMENT, "the program's settings are being tuned");
    for(const auto& sl : s->pool)
        if(sl.prog == prog) return fail(SQD_ERR_ARGUMENT, "the program already has cores; set its settings before its first job");
    sqd_settings v = *settings;
    v.level_bytes = s->opt.level_bytes;
    v.frontier_capacity = s->opt.frontier_capacity;
    s->settings[prog] = v;
    s->cv.notify_all();
    return SQD_OK;
}

int sqd_service_submit(sqd_service* s, const sqd_program* prog, uint32_t first, uint32_t last, uint64_t* job) {
    if(!s || !prog || !job) return fail(SQD_ERR_ARGUMENT, "null argument");
    if(first >= last) return fail(SQD_ERR_ARGUMENT, "empty initial range");
    std::lock_guard<std::mutex> lk(s->m);
    const uint64_t id = s->next_id++;
    sqd_service::Job& j = s->jobs[id];
    j.prog = prog;
    j.first = first;
    j.last = last;
    j.submitted = Clock::now();
    s->queue.push_back(id);
    *job = id;
    s->cv.notify_all();
    return SQD_OK;
}

int sqd_service_submit_many(sqd_service* s, const sqd_program* prog, const uint32_t* first, const uint32_t* last,
                            size_t n, uint64_t* jobs) {
    if(!s || !prog || (n && (!first || !last || !jobs))) return fail(SQD_ERR_ARGUMENT, "null argument");
    for(size_t i = 0; i < n; ++i)
        if(first[i] >= last[i]) return fail(SQD_ERR_ARGUMENT, "empty initial range (job " + std::to_string(i) + ")");
    if(n == 0) return SQD_OK;
    std::lock_guard<std::mutex> lk(s->m);
    const Clock::time_point now = Clock::now();
    for(size_t i = 0; i < n; ++i) {
        const uint64_t id = s->next_id++;
        sqd_service::Job& j = s->jobs[id];
        j.prog = prog;
        j.first = first[i];
        j.last = last[i];
        j.submitted = now;
        s->queue.push_back(id);
        jobs[i] = id;
    }
    s->cv.notify_all();
    return SQD_OK;
}

int sqd_service_poll(sqd_service* s, uint64_t job, int* done, sqd_job_result* out) {
    if(!s || !done) return fail(SQD_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(job);
    if(it == s->jobs.end()) return fail(SQD_ERR_ARGUMENT, "unknown job");
    *done = it->second.done ? 1 : 0;
    if(it->second.done && out) *out = it->second.result;
    return SQD_OK;
}

int sqd_service_wait(sqd_service* s, uint64_t job, sqd_job_result* out) {
    if(!s) return fail(SQD_ERR_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(job);
    if(it == s->jobs.end()) return fail(SQD_ERR_ARGUMENT, "unknown job");
    s->cv.wait(lk, [&] { return it->second.done || s->stop; });
    if(!it->second.done) return fail(SQD_ERR_INTERNAL, "service stopped");
    if(out) *out = it->second.result;
    const sqd_service::Job j = it->second;
    s->jobs.erase(it);
    if(j.result.status != SQD_OK) sqd::set_last_error(j.error);
    return j.result.status;
}

int sqd_service_get_stats(sqd_service* s, sqd_service_stats* out) {
    if(!s || !out) return fail(SQD_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    return SQD_OK;
}

void sqd_service_destroy(sqd_service* s) {
    if(!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
        s->cv.notify_all();
    }
    for(auto& t : s->workers) t.join();
    for(auto& slot : s->pool)
        if(slot.core) sqd_core_destroy(slot.core);
    delete s;
}

}  // extern "C"
