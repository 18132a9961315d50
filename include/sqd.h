/*
 * sqd.h - C ABI of the MI355X (gfx950) branch-and-bound prover core.
 *
 * The reference's drop-in boundary is the duck-typed "prover core" consumed by
 * ProofDriver (ivarp/include/ivarp/prover/proof_driver.hpp:218-627), implemented
 * by CPUProverCore (cpu_prover/cpu_prover_core.hpp:31-260) and CUDAProverCore
 * (with_cuda/ivarp_cuda/gpu_prover_core.cuh:199-475) and selected by run_prover
 * (run_prover.hpp:57-74).  Each entry point below names the member it replaces.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative sqd_status; the message
 *    of the last failure on the calling thread is available from sqd_last_error().
 *    No C++ exception crosses this boundary.
 *  - Boxes are passed in structure-of-arrays layout: for n boxes of N intervals,
 *    lb[a * n + i] and ub[a * n + i] are the bounds of argument a of box i.
 *  - Host buffers are caller-owned; device buffers are owned by the library.
 *  - One sqd_core per device; a core is thread-compatible, not thread-safe
 *    (like the reference, calls for one core are never concurrent).
 *  - Counts follow the reference's CuboidCounts semantics
 *    (prover/proof_driver.hpp:48-70, prover/proof_information.hpp:32-44).
 */
#ifndef SQD_H
#define SQD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sqd_status {
    SQD_OK = 0,
    SQD_ERR_ARGUMENT = -1,   /* invalid argument or malformed program text */
    SQD_ERR_NO_KERNELS = -2, /* no compiled kernels for this program */
    SQD_ERR_DEVICE = -3,     /* HIP runtime error (launch, memory, device) */
    SQD_ERR_CAPACITY = -4,   /* a device buffer was too small for the workload */
    SQD_ERR_INTERNAL = -5
} sqd_status;

typedef struct sqd_program sqd_program;
typedef struct sqd_core sqd_core;

/* ProofInformation / CuboidCounts (prover/proof_information.hpp:32-44) */
typedef struct sqd_counts {
    uint64_t num_cuboids;
    uint64_t num_leaf_cuboids;
    uint64_t num_critical_cuboids;
    uint64_t num_repeated_nodes;
} sqd_counts;

/* The subset of ProverSettings (prover_input.hpp:124-159) the GPU core uses,
 * plus its own sizing knobs.  Fill with sqd_settings_default(); zero sizing
 * fields take the defaults, generation_count is used as given. */
typedef struct sqd_settings {
    uint32_t generation_count;     /* generations of dynamic splits (reference default 8) */
    uint32_t dequeue_buffer_size;  /* dynamic cuboids per round; default chosen per program */
    uint64_t frontier_capacity;    /* cuboids the device frontier can hold; default 1<<23 */
    uint64_t level_bytes;          /* bytes per level buffer; default 1<<30 */
    uint64_t critical_capacity;    /* criticals the device list holds (default 1<<20; at least the
                                      last static level's subdivisions): more fail a call unless a
                                      critical sink receives them (sqd_core_set_critical_sink) */
    uint32_t grid_blocks;          /* blocks per level launch; default: the kernel's resident
                                      capacity (occupancy x CUs), capped by the work */
    uint32_t fp32_prefilter;       /* bit mask of the static levels (bit L: level L) that first
                                      decide each child in outward-rounded binary32 and
                                      re-evaluate only the undecided ones in binary64;
                                      0 = off, SQD_FP32_ALL_LEVELS = every level (no
                                      counterpart in the reference, which evaluates
                                      everything in binary64) */
    uint32_t graphs;               /* nonzero: sqd_prove replays its rounds as captured HIP
                                      graphs from the second proof on a core (default 1) */
    uint32_t prefilter_group;      /* FP32 prefilter: consecutive children of one parent a
                                      lane first tests as one group (a power of two up to
                                      64; 0 = chosen per launch, 1 = no groups) */
    uint32_t jobs;                 /* jobs per batch (sqd_prove_batch), at most 64; the core
                                      holds that many copies of its device memory and runs
                                      every launch with one grid row per job; 0 or 1: one */
    uint32_t group_levels;         /* bit mask of the static levels (bit L: level L) decided
                                      in binary64 with the group test: the rows are first
                                      checked once per group of prefilter_group consecutive
                                      children, the children of the groups left one by one;
                                      exactly the results of the plain pass (no counterpart in
                                      the reference; levels in fp32_prefilter take that) */
} sqd_settings;

#define SQD_FP32_ALL_LEVELS 0xffffffffu

/* Defaults: generation_count = 8, graphs = 1, every other field 0 (sizes chosen
 * by the core, no FP32 prefilter). */
void sqd_settings_default(sqd_settings* settings);

/* ---------------------------------------------------------------- programs */

/* Parse a lowered program (the text produced by the lemma lowering) and bind it
 * to the kernels compiled for it.  Replaces prover_input<...>(system) plus the
 * template instantiation of the core (prover_input.hpp:118-122). */
int sqd_program_from_text(const char* text, sqd_program** out);
/* Lower one of the built-in lemma statements ("top1", "top2", "one", "two",
 * "three", "four", "five", "six", "seven", "more1", "more2"). */
int sqd_program_from_lemma(const char* key, sqd_program** out);
void sqd_program_destroy(sqd_program* program);
/* Canonical text of the program; *len receives the size including the NUL. */
int sqd_program_text(const sqd_program* program, char* buffer, size_t capacity, size_t* len);
int sqd_program_nargs(const sqd_program* program);
/* 1 if kernels are compiled into this library for the program, else 0. */
int sqd_program_has_kernels(const sqd_program* program);
/* Number of built-in lemmas and the key of lemma i. */
int sqd_lemma_count(void);
const char* sqd_lemma_key(int i);
const char* sqd_lemma_title(int i);

/* ---------------------------------------------------------------- cores */

/* initialize / initialize_per_thread / set_* (prover/basic_prover_core.hpp:30-67,
 * gpu_prover_core.cuh:241-263,427-464): allocate device state on `device`. */
int sqd_core_create(int device, const sqd_program* program, const sqd_settings* settings, sqd_core** out);
void sqd_core_destroy(sqd_core* core);
/* Device memory the core allocated (frontier, level buffers, lists, state).
 * No counterpart in the reference. */
int sqd_core_memory(const sqd_core* core, uint64_t* device_bytes);

/* Initial runtime box (refactor_constraints_impl.hpp:61-101): N lb then N ub. */
int sqd_initial_box(sqd_core* core, double* lb, double* ub);

/* Initial queue (ProofDriver::proof_init_*, proof_driver.hpp:320-391): up to
 * `capacity` boxes in SoA with stride `capacity`; *n receives the queue length. */
int sqd_initial_queue(sqd_core* core, double* lb, double* ub, size_t capacity, size_t* n, sqd_counts* counts);

/* handle_cuboids_nonfinal (cpu_prover_core.hpp:52-60): for each of n boxes the
 * hull of its surviving static leaves; has_hull[i] = 0 if none survived. */
int sqd_nonfinal(sqd_core* core, const double* lb, const double* ub, size_t n,
                 double* hull_lb, double* hull_ub, uint8_t* has_hull, sqd_counts* counts);

/* handle_cuboids_final (cpu_prover_core.hpp:62-71): critical leaves of n boxes,
 * up to `capacity` (SoA, stride capacity); *n_critical receives the total (the
 * rest can be fetched with sqd_core_criticals).  With a critical sink set, every
 * critical goes to the sink instead and none to crit_lb/crit_ub. */
int sqd_final(sqd_core* core, const double* lb, const double* ub, size_t n,
              double* crit_lb, double* crit_ub, size_t capacity, size_t* n_critical, sqd_counts* counts);

/* split_into (pd_impl_split_into.hpp:29-75): children of n boxes that survive
 * propagation and every constraint (SoA, stride capacity). */
int sqd_split(sqd_core* core, const double* lb, const double* ub, size_t n,
              double* out_lb, double* out_ub, size_t capacity, size_t* n_out, sqd_counts* counts);

/* ProofDriver::run (proof_driver.hpp:472-489) with the frontier resident in HBM:
 * returns the counts; criticals (up to capacity, SoA stride capacity) are copied
 * out when crit_lb/crit_ub are non-NULL.  *seconds (optional) receives the time
 * spent on the device loop. */
int sqd_prove(sqd_core* core, sqd_counts* counts, double* crit_lb, double* crit_ub,
              size_t capacity, size_t* n_critical, double* seconds);

/* The same proof restricted to the initial subdivisions [first, last) of the
 * first dynamic variable (proof_driver.hpp:320-391 splits it into `initial`
 * parts first): the exact union of those subtrees of the full proof, so the
 * counts of a partition of [0, initial) add up to sqd_prove's.  The unit of work
 * of the proof service.  No counterpart in the reference. */
int sqd_prove_range(sqd_core* core, uint32_t first, uint32_t last, sqd_counts* counts, double* crit_lb,
                    double* crit_ub, size_t capacity, size_t* n_critical, double* seconds);

/* Several proof jobs of one program in the same launches (no counterpart in the
 * reference, which proves one statement at a time): job j proves the initial
 * range [first[j], last[j]) exactly as sqd_prove_range would, with its own
 * frontier, level buffers, counters and criticals, in the core's job memory j.
 * Every kernel of a round runs one grid row per job, so the per-launch cost and
 * the small kernels of a round (initial queue, round control, count fold) are
 * shared by the batch, and one job's idle rows leave the machine to the others.
 * n_jobs <= sqd_settings.jobs of the core; the call ends when every job has.
 * A job whose buffers overflow fails the whole call (SQD_ERR_CAPACITY).
 * sqd_prove and sqd_prove_range on such a core run a batch of one. */
typedef struct sqd_job_counts {
    sqd_counts counts;
    uint64_t n_critical;      /* criticals the job found (sqd_core_job_criticals) */
    uint64_t fp32_checked;    /* static-level children the FP32 prefilter evaluated */
    uint64_t fp32_decided;    /* ... and pruned */
    uint64_t frontier_peak;   /* most cuboids the job's frontier held after a round */
} sqd_job_counts;
int sqd_prove_batch(sqd_core* core, size_t n_jobs, const uint32_t* first, const uint32_t* last,
                    sqd_job_counts* results, double* seconds);
/* Continuous batches on a core of sqd_settings.jobs rows (no counterpart in the
 * reference, which proves one statement at a time): jobs join free rows at
 * round boundaries and leave when their frontier is empty, so that a job
 * arriving while others run starts with the next round instead of waiting for
 * the whole batch.  sqd_batch_admit puts the job [first, last) (as in
 * sqd_prove_range) into the free row `row`; it starts with the next step.
 * sqd_batch_step runs one round on every row that holds a job (after the initial
 * queues of the rows admitted since the last step, only theirs): *finished
 * receives the mask of rows whose job ended in this step (their counts in
 * results[row], which needs sqd_settings.jobs entries; their criticals stay
 * available to sqd_core_job_criticals until the row starts another job), *active
 * the number of rows still running.  Each job's counts and criticals are exactly
 * those of its range proved alone.  A failing step (a buffer too small for one
 * job) drops every row's job; any other computing call on the core ends the
 * continuous batch the same way. */
int sqd_batch_admit(sqd_core* core, uint32_t row, uint32_t first, uint32_t last);
int sqd_batch_step(sqd_core* core, sqd_job_counts* results, uint64_t* finished, uint32_t* active);
/* The cuboids on each row's frontier after the last step (counts[j] for rows
 * j < n and j < sqd_settings.jobs; rows without a job hold 0). */
int sqd_batch_frontier(sqd_core* core, uint64_t* counts, size_t n);

/* The critical cuboids of job `job` of the last sqd_prove_batch (as
 * sqd_core_criticals: up to capacity boxes, SoA stride capacity; *n the total). */
int sqd_core_job_criticals(sqd_core* core, size_t job, double* crit_lb, double* crit_ub, size_t capacity,
                           size_t* n);

/* The critical cuboids found by the last sqd_prove on this core (they stay in
 * device memory until the next call): up to `capacity` boxes (SoA, stride
 * capacity); *n receives the total found.  The reference hands each critical
 * to its on_critical callback (prover/proof_driver.hpp:499-509). */
int sqd_core_criticals(sqd_core* core, double* crit_lb, double* crit_ub, size_t capacity, size_t* n);

/* on_critical (set_on_critical, prover/basic_prover_core.hpp:30-67; called per
 * critical in cpu_prover/cpu_prover_core.hpp:137-149; the reference's GPU core
 * re-launches until its output buffer holds every critical,
 * with_cuda/ivarp_cuda/gpu_prover_core.cuh:293-316,
 * cuda_output_buffer.cuh:68-107): with a sink set, sqd_prove, sqd_prove_range
 * and sqd_final on a core of one job hand every critical cuboid to `fn` as the
 * proof finds them, in chunks of n boxes in SoA layout (lb[a * stride + i]),
 * valid during the call, on the calling thread.  The device critical list is
 * drained after every round; a round with more criticals than
 * critical_capacity repeats its final pass in parts that fit (counting nothing
 * twice), so no critical is lost.  Those calls then copy no criticals to their
 * own output arrays and report the total in *n_critical.  Without a sink, more
 * criticals than critical_capacity fail the call with SQD_ERR_CAPACITY.  Batches
 * of several jobs keep up to critical_capacity criticals per job.  fn = NULL
 * removes the sink. */
typedef void (*sqd_critical_fn)(void* user, const double* lb, const double* ub, size_t n, size_t stride);
int sqd_core_set_critical_sink(sqd_core* core, sqd_critical_fn fn, void* user);

/* Progress of sqd_prove / sqd_prove_range: the reference's ProgressInfo and
 * observer calls (prover/proof_driver.hpp:551-603, progress_printer.hpp:47-90):
 * after `start_after_seconds`, then at most every `every_seconds`, between two
 * rounds, and once more when the proof ends (done = 1, with its status).  The
 * callback runs on the calling thread; head_lb/head_ub (N values each) hold the
 * next cuboid of the frontier and are valid during the call only. */
typedef struct sqd_progress {
    uint64_t queue_size;          /* cuboids on the device frontier */
    uint64_t cuboid_count;
    uint64_t leaf_count;
    uint64_t critical_count;
    uint64_t rounds;              /* rounds completed */
    int32_t done;                 /* 1 on the last call of a proof */
    int32_t status;               /* SQD_OK, or the error that ended the proof */
    const double* head_lb;        /* top of the frontier; NULL when it is empty */
    const double* head_ub;
} sqd_progress;
typedef void (*sqd_progress_fn)(void* user, const sqd_progress* progress);
/* fn = NULL turns progress reports off. */
int sqd_core_set_progress(sqd_core* core, sqd_progress_fn fn, void* user, double start_after_seconds,
                          double every_seconds);

/* Evaluations of acos by this core's program whose rounding direction could not
 * be decided at 2^-113 precision, so that the interval was widened by one double
 * (still an enclosure); counted since the library was loaded.  Zero means every
 * acos was rounded exactly like the reference's MPFR evaluation.  No counterpart
 * in the reference. */
int sqd_core_acos_fallbacks(sqd_core* core, uint64_t* n);

/* ---------------------------------------------------------------- FP32 prefilter */

/* With sqd_settings.fp32_prefilter set: the static-level children of the last
 * computing call (sqd_prove, sqd_nonfinal, sqd_final) that the binary32 pass
 * evaluated, and those it decided (pruned) without binary64.  No counterpart in
 * the reference. */
int sqd_core_fp32_stats(sqd_core* core, uint64_t* checked, uint64_t* decided);

/* Binary32-versus-binary64 overestimation per constraint: over n boxes (SoA,
 * stride n) the width of the interval of (left side - right side) of each
 * constraint that is a plain comparison, evaluated on the outward binary32 box
 * with outward binary32 constants, divided by the same width in binary64.
 * Entries are in the order of the program's constraint rows' roots; other
 * constraints (boolean combinations) report 0 samples.  No counterpart in the
 * reference. */
typedef struct sqd_width_stat {
    int32_t constraint;   /* node index of the constraint in the lowered program */
    uint32_t reserved;
    uint64_t samples;     /* boxes with finite, defined widths and a positive binary64 width */
    double mean_ratio;    /* mean of width32 / width64 */
    double max_ratio;
} sqd_width_stat;
int sqd_width_stats(sqd_core* core, const double* lb, const double* ub, size_t n, sqd_width_stat* out,
                    size_t capacity, size_t* n_constraints);

/* Values of the program's expressions on n boxes (SoA, stride n), in binary64
 * (precision 64) or, on each box converted outward to binary32, in binary32
 * (precision 32): for the enclosure checks of tests against point evaluations
 * in exact arithmetic (the reference's own tests evaluate its functions on
 * random points and intervals, ivarp/test/test_util.hpp:113-235,
 * test/cuda/cuda_test_fn.cuh:29-86).  bound_values[(i * n_bounds + k) * 3 + {0, 1,
 * 2}] receives lb, ub and the possibly-undefined flag (0/1) of the expression of
 * runtime bound k (the program's bound order); cons_values[(i * n_constraints +
 * j) * 2 + {0, 1}] the IBool (definitely, possibly) of constraint j (the order of
 * sqd_width_stats).  *n_bounds and *n_constraints receive the counts (NULL
 * outputs: only those).  No counterpart in the reference. */
int sqd_eval_program(sqd_core* core, int precision, const double* lb, const double* ub, size_t n,
                     double* bound_values, uint8_t* cons_values, size_t* n_bounds, size_t* n_constraints);

/* The number of initial subdivisions of the first dynamic variable (the range
 * sqd_prove_range selects from) and the number of static levels (the bits of
 * sqd_settings.fp32_prefilter that apply).  No counterpart in the reference. */
int sqd_core_shape(const sqd_core* core, uint32_t* initial_subdivisions, uint32_t* static_levels);

/* ---------------------------------------------------------------- autotuner */

/* Chooses, by coordinate descent over timed proofs of a slice of the program's
 * first variable, the FP32-prefilter level mask, the binary64 group-test level
 * mask (group_levels), the children per group (prefilter_group), the dynamic
 * cuboids per round (dequeue_buffer_size) and the
 * blocks per level launch (grid_blocks), starting from `base` (NULL: the
 * defaults).  With base->jobs > 1 every timed proof is a batch of that many
 * copies of the slice (sqd_prove_batch), as the proof service runs them, and the
 * report's times are per batch.  Candidates whose counts differ from the plain
 * binary64 run are rejected.  No counterpart in the reference. */
typedef struct sqd_autotune_options {
    double target_seconds;    /* slice of the first variable to time: about this long, and about
                                 this much timing per candidate (default 0.02) */
    int32_t repetitions;      /* at least this many timed proofs per candidate, median taken (default 3) */
    uint32_t fp32_first;      /* nonzero: the binary64 group test is tried only on the levels the
                                 best prefilter mask leaves; 0 (default): also on its own, without
                                 the prefilter (exact binary64 results) */
} sqd_autotune_options;
typedef struct sqd_autotune_report {
    uint32_t first, last;     /* the slice [first, last) that was timed */
    uint32_t trials;          /* candidates timed */
    uint32_t reserved;
    double plain_seconds;     /* the base settings without the prefilter */
    double best_seconds;      /* the chosen settings */
    sqd_counts counts;        /* counts of the slice (equal for every accepted candidate) */
} sqd_autotune_report;
void sqd_autotune_options_default(sqd_autotune_options* options);
int sqd_autotune(int device, const sqd_program* program, const sqd_settings* base, const sqd_autotune_options* options,
                 sqd_settings* best, sqd_autotune_report* report);

/* ---------------------------------------------------------------- proof service */

/* A queue of proof jobs (a program and a range [first, last) of its first
 * variable's initial subdivisions, as in sqd_prove_range) run by worker threads,
 * each on a pooled core of the job's program with its own stream.  A worker
 * takes the program of the oldest queued job and runs a continuous batch of it
 * (sqd_batch_admit / sqd_batch_step): queued jobs of the program fill the core's
 * batch_jobs rows, a job leaves its row when its frontier is empty, and jobs
 * that arrive meanwhile join at the next round boundary; the batches of up to
 * `workers` workers overlap on the device.  Cores are created on
 * demand within `memory_budget` bytes of device memory; idle cores of other
 * programs are destroyed to make room.  Jobs may be submitted from any thread
 * while others run.  Programs must outlive the service.  No counterpart in the
 * reference (ProofDriver runs one proof at a time, prover/proof_driver.hpp:393-438). */
typedef struct sqd_service sqd_service;
typedef struct sqd_service_options {
    int32_t device;
    uint32_t workers;             /* batches in flight (default 8) */
    uint64_t memory_budget;       /* device bytes for the pooled cores (default 128 GiB) */
    uint32_t fp32_prefilter;      /* prefilter level mask of every core (default SQD_FP32_ALL_LEVELS) */
    uint32_t autotune;            /* nonzero: sqd_autotune each program on its first job */
    uint64_t level_bytes;         /* per core (default 256 MiB) */
    uint64_t frontier_capacity;   /* per core (default 1 << 22 cuboids) */
    uint64_t core_bytes_estimate; /* size assumed per job of a program's first core, before one
                                     has been measured (default 2 GiB) */
    uint32_t batch_jobs;          /* queued jobs of one program a worker proves together
                                     (sqd_prove_batch; each core holds that many; default 8) */
    uint32_t reserved;
    uint64_t frontier_budget;     /* bytes of frontier the running jobs may hold together
                                     (0: no limit, the default); see sqd_service_get_stats */
} sqd_service_options;
typedef struct sqd_job_result {
    sqd_counts counts;
    uint64_t n_critical;
    double queued_seconds;        /* submission to start */
    double run_seconds;           /* start to completion */
    double latency_seconds;       /* submission to completion */
    int32_t status;               /* SQD_OK or the job's error */
    uint32_t joined_running;      /* 1: the job joined a batch that was already running */
    uint64_t frontier_peak;       /* most cuboids its frontier held after a round */
} sqd_job_result;
void sqd_service_options_default(sqd_service_options* options);
int sqd_service_create(const sqd_service_options* options, sqd_service** out);
int sqd_service_submit(sqd_service* service, const sqd_program* program, uint32_t first, uint32_t last, uint64_t* job);
/* n jobs of `program` at once, job i over [first[i], last[i]); their ids in
 * jobs[0..n).  One lock and one wake-up of the workers for all of them, for
 * callers that hand over many jobs together (a job stream's burst, a whole
 * proof's ranges).  No counterpart in the reference (which proves one
 * statement at a time, prover/proof_driver.hpp:393-438).  All ranges are
 * checked before any job is queued: on an error none is. */
int sqd_service_submit_many(sqd_service* service, const sqd_program* program, const uint32_t* first,
                            const uint32_t* last, size_t n, uint64_t* jobs);
/* The settings the service's cores of `program` use (*known = 0 before its first
 * job has chosen them), with jobs = batch_jobs. */
int sqd_service_get_settings(sqd_service* service, const sqd_program* program, int* known, sqd_settings* out);
/* Fixes the settings of `program`'s cores (instead of the defaults or autotuning);
 * only before its first job.  level_bytes and frontier_capacity stay the
 * service's, jobs is batch_jobs. */
int sqd_service_set_settings(sqd_service* service, const sqd_program* program, const sqd_settings* settings);
/* *done = 1 once the job has finished (its result is then copied to *out). */
int sqd_service_poll(sqd_service* service, uint64_t job, int* done, sqd_job_result* out);
/* Blocks until the job finishes, copies its result and forgets the job; returns its status. */
int sqd_service_wait(sqd_service* service, uint64_t job, sqd_job_result* out);
/* Stops the workers (running jobs finish, queued ones are dropped) and frees the cores. */
void sqd_service_destroy(sqd_service* service);

/* Admission by frontier bytes.  A running job's frontier (the cuboids on its
 * device stack, 16 bytes per variable plus 4 each) is counted after every
 * round; a queued job is admitted (into a free row of a running batch, or a new
 * batch) only while the running jobs' frontier bytes plus its own estimate stay
 * within frontier_budget, or when no job runs at all.  A job's estimate is its
 * program's largest frontier peak per initial subdivision among the finished
 * jobs (at least one cuboid), times its range.  No counterpart in the
 * reference; for callers sharing one device among many jobs. */
typedef struct sqd_service_stats {
    uint64_t frontier_bytes;      /* running jobs' frontier bytes now */
    uint64_t frontier_bytes_peak; /* most of them at once */
    uint32_t rows_running;        /* jobs running now (rows of continuous batches) */
    uint32_t rows_peak;           /* most at once */
    uint64_t cores_created;
    uint64_t cores_destroyed;     /* idle cores destroyed to make room */
    uint64_t jobs_completed;
} sqd_service_stats;
int sqd_service_get_stats(sqd_service* service, sqd_service_stats* out);

/* ---------------------------------------------------------------- timing */

/* Device time per kernel kind, from HIP events recorded on the core's stream
 * around every launch while profiling is on.  No counterpart in the reference
 * (its ProofInformation, ivarp/include/ivarp/prover/proof_information.hpp,
 * holds counts only); used by the throughput measurement. */
typedef struct sqd_kernel_stat {
    char name[48];        /* "static_level<NONFINAL,1>", "hull_finalize", ... */
    int32_t mode;         /* 0 non-final, 1 final, 2 split, 3 initial queue, -1 other kernels */
    int32_t level;        /* static or dynamic level; -1 for other kernels */
    uint64_t launches;
    double milliseconds;  /* summed event time over the launches */
    uint64_t children;    /* child cuboids those launches processed (static and split levels of
                             proofs; 0 for other kernels) */
} sqd_kernel_stat;

/* enable != 0 turns profiling on and clears the statistics; 0 turns it off.
 * With only_mode >= 0 just the level kernel (only_mode, only_level) is timed,
 * which keeps the event overhead out of the other launches. */
int sqd_core_set_profiling(sqd_core* core, int enable, int only_mode, int only_level);
/* Copies up to `capacity` entries (ordered by kind); *count receives the number of kinds. */
int sqd_core_kernel_stats(const sqd_core* core, sqd_kernel_stat* out, size_t capacity, size_t* count);

/* ---------------------------------------------------------------- tests */

/* Elementwise directed rounding on the device (op: 0 add, 1 sub, 2 mul, 3 div,
 * 4 sqrt(a), 5 acos(a), 6 acos(a) by the fixed-point series alone, the fallback
 * of 5): lo = RD(a op b), hi = RU(a op b).  Ops 8..13 are the
 * same in binary32 (add, sub, mul, div, sqrt, acos; operands must be floats,
 * results are floats widened exactly), op 14 the outward conversion of a to
 * binary32.  Arithmetic parity checks (acos: ivarp/src/interval_acos.cpp:29-66
 * rounds with MPFR). */
int sqd_rounded_op(int device, int op, size_t n, const double* a, const double* b, double* lo, double* hi);

/* Interval operations on the device (run-time sign decisions, as for operands
 * without compile-time bounds): n operand pairs of 3 doubles (lb, ub, undefined
 * flag 0/1), op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt(a), 5 square(a), 6 -a,
 * 7 max, 8 min, 9 splitter point (a = (lb, ub, -), b = (i, n, -) -> out lb),
 * 10 acos(a) with domain clamping to [-1, 1].
 * Interval parity checks (ivarp/test/interval_*.cpp, splitter.cpp). */
int sqd_interval_op(int device, int op, size_t n, const double* a, const double* b, double* out);

/* Evaluations of acos in the test kernels whose rounding direction could not be
 * decided at 2^-113 precision (results widened instead); 0 in every run so far. */
int sqd_test_acos_fallbacks(int device, uint64_t* n);

/* ---------------------------------------------------------------- without allocation */

/* sqd_prove_range without a core and without allocating device memory: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  One job
 * of `program` proves the initial range [first, last) with `settings` (NULL:
 * the defaults; jobs and graphs are ignored) exactly as sqd_prove_range does,
 * its device memory laid out in a static 64 MiB array that each device holds
 * until the process ends (SQD_ERR_CAPACITY, before any work, when the settings
 * need more: lower level_bytes, frontier_capacity, critical_capacity).  The
 * call allocates no device or pinned memory, no stream and no event; it runs
 * on the device's null stream.  Calls on one device share the array and are
 * serialised by a lock (overlapping calls would race on it); each call clears
 * what must start at zero.  Outputs as sqd_prove_range (criticals up to
 * capacity, SoA stride capacity; *n_critical the total). */
int sqd_prove_range_noalloc(int device, const sqd_program* program, const sqd_settings* settings, uint32_t first,
                            uint32_t last, sqd_counts* counts, double* crit_lb, double* crit_ub, size_t capacity,
                            size_t* n_critical);

const char* sqd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SQD_H */
