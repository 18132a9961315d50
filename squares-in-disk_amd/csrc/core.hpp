// Generic GPU prover-core kernels, instantiated once per lowered program.
//
// One lane handles one child cuboid ("one cuboid per lane").  A static level
// kernel takes a list of parent boxes (structure-of-arrays in HBM), splits the
// level's variable, runs interval constraint propagation and the constraint
// rows, and stream-compacts the survivors with a wave ballot + prefix count into
// the next level's list.  The last static level joins the surviving leaves into
// the per-dynamic-cuboid hull with 64-bit atomics on order-preserving keys (the
// critical-cuboid reduction).  The dynamic split (split_into) and the initial
// queue are level kernels of the same shape over the dynamic variables.
//
// Reference semantics restated here:
//   static enumeration / counts   ivarp/include/ivarp/cpu_prover/cpu_prover_core.hpp:179-256
//   hull (AddToRange)             cpu_prover_core.hpp:97-134, number/interval.hpp:77-84
//   dynamic split                 prover/pd_impl_split_into.hpp:29-75
//   initial queue                 prover/proof_driver.hpp:320-391
//   propagation                   bound_propagation/iterated_recursive.hpp:40-131,
//                                 dynamic_bound_application.hpp:104-123, index_queue_set.hpp:38-101
//   splitter                      splitter.hpp:44-59
//   initial box                   refactor_constraint_system/refactor_constraints_impl.hpp:61-101
//
// A program struct P (generated from the lowered program) provides:
//   N, NB (bounds), NC (constants), NDYN, NSTAT,
//   dyn_arg(j), dyn_initial(j), dyn_subs(j), st_arg(j), st_subs(j)  (constexpr)
//   apply(k, lb, ub, C)      -> -1 empty, 0 unchanged, else 1 + update-list id
//   deps(list, f)            calls f(a) for each bound index a in update list `list`
//   rows_violated<L>(lb, ub, C)   constraint rows of static level L
//   all_satisfied(lb, ub, C)      every remaining constraint possibly holds
//   level_waves(L), prop_fast(L)  per-level register budget and propagation form
#pragma once
#include <type_traits>
#include "interval.hpp"

namespace sqd {
namespace dev {

constexpr int kBlock = 256;

// Register budget of the level kernels, as a minimum of resident waves per SIMD
// (0: the compiler's choice).
#ifndef SQD_LEVEL_WAVES_PER_EU
#define SQD_LEVEL_WAVES_PER_EU 0
#endif
#if SQD_LEVEL_WAVES_PER_EU > 0
#define SQD_LEVEL_ATTR __attribute__((amdgpu_waves_per_eu(SQD_LEVEL_WAVES_PER_EU)))
#else
#define SQD_LEVEL_ATTR
#endif
// The binary64 static level kernels take their program's budget for the level
// (P::level_waves(L); 0 or 1: no minimum) unless SQD_LEVEL_WAVES_PER_EU is set.
template<class P, int L>
constexpr int static_level_waves() {
    return SQD_LEVEL_WAVES_PER_EU > 0 ? SQD_LEVEL_WAVES_PER_EU : (P::level_waves(L) > 0 ? P::level_waves(L) : 1);
}

enum Mode : int { NONFINAL = 0, FINAL = 1, SPLIT = 2, INIT = 3 };
// Hoisted-term struct of a program without static levels.
struct NoHoist {};
// Flag or'ed into a static-level mode at launch: the FP32-prefiltered kernel.
constexpr int kModeF32 = 4;
// Static-level mode bit: binary64 with the group test (k_static_group64).
constexpr int kModeGroup = 8;

struct DeviceCounters {
    unsigned long long cuboids;
    unsigned long long leaves;
    unsigned long long criticals;
    unsigned long long overflow;
};

// A list of boxes in structure-of-arrays layout: lb[a * stride + i].
struct BoxList {
    double* lb;
    double* ub;
    unsigned* depth;     // SPLIT/INIT/stack lists
    unsigned* origin;    // static levels: index of the dynamic cuboid within the round
    unsigned* leaf;      // static levels: enumeration index within that cuboid's grid
    unsigned long long stride;
    unsigned long long cap;
    unsigned long long* count;  // device counter of valid entries
};

// Arguments of a level kernel.  The input range lives in device memory (the
// round's pop position and size, or a list's counter), so a whole round can be
// enqueued without the host knowing its size.
struct LevelArgs {
    BoxList in;
    BoxList out;
    const unsigned long long* in_base;   // first input entry (nullptr: 0)
    const unsigned long long* in_count;  // number of inputs
    unsigned long long* hull;            // NONFINAL last level: [origin][arg][4] keys
    DeviceCounters* counters;            // overflow flag
    unsigned long long* block_counts;    // per-block count slots (block_counts_add)
    const double* consts;
    unsigned generations;
    // FP32 prefilter: children (static-level child indices) binary32 left undecided
    unsigned long long* undec;
    unsigned long long* undec_count;
    unsigned long long undec_cap;
    // initial queue: the first dynamic variable's initial subdivisions [range[0], range[1])
    // (nullptr: all of them)
    const unsigned long long* init_range;
    // FP32 prefilter: most children of one parent per lane (0: chosen from the launch size)
    int children_per_lane;
    // batched launches: the grid's row blockIdx.y works on job blockIdx.y, whose
    // memory is job 0's shifted by blockIdx.y * job_stride bytes (see job_args);
    // `jobs` is the number of rows the host launches
    unsigned long long job_stride;
    unsigned jobs;
};

// ---------------------------------------------------------------- batched jobs
// A batch of jobs keeps every job's device memory (lists, control block, count
// slots, hull records, constants) in one allocation of equal parts, so job j's
// pointers are job 0's plus j times the part's size.  A kernel launched over a
// grid of several rows shifts its pointer arguments by its row's offset; null
// pointers (an absent list column, a missing input base) stay null.
template<class T>
__device__ __forceinline__ T* job_ptr(T* p, unsigned long long off) {
    return p ? reinterpret_cast<T*>(reinterpret_cast<unsigned long long>(p) + off) : p;
}
__device__ __forceinline__ unsigned long long job_offset(unsigned long long stride) {
    return static_cast<unsigned long long>(blockIdx.y) * stride;
}
__device__ __forceinline__ BoxList job_list(BoxList b, unsigned long long off) {
    b.lb = job_ptr(b.lb, off);
    b.ub = job_ptr(b.ub, off);
    b.depth = job_ptr(b.depth, off);
    b.origin = job_ptr(b.origin, off);
    b.leaf = job_ptr(b.leaf, off);
    b.count = job_ptr(b.count, off);
    return b;
}
__device__ __forceinline__ LevelArgs job_args(LevelArgs a) {
    const unsigned long long off = job_offset(a.job_stride);
    a.in = job_list(a.in, off);
    a.out = job_list(a.out, off);
    a.in_base = job_ptr(a.in_base, off);
    a.in_count = job_ptr(a.in_count, off);
    a.hull = job_ptr(a.hull, off);
    a.counters = job_ptr(a.counters, off);
    a.block_counts = job_ptr(a.block_counts, off);
    a.consts = job_ptr(a.consts, off);
    a.undec = job_ptr(a.undec, off);
    a.undec_count = job_ptr(a.undec_count, off);
    a.init_range = job_ptr(a.init_range, off);
    return a;
}

// ---------------------------------------------------------------- hull keys
__device__ __forceinline__ unsigned long long order_key(double x) {
    unsigned long long u = bits(x == 0.0 ? 0.0 : x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_order_key(unsigned long long k) {
    return from_bits((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
}

// ---------------------------------------------------------------- splitter
__device__ __forceinline__ double split_width(double lb, double ub, int n) {
    return rd_div(rd_sub(ub, lb), static_cast<double>(n));
}
__device__ __forceinline__ double split_point(double lb, double ub, double w, int i, int n) {
    return i >= n ? ub : rd_add(lb, rd_mul(static_cast<double>(i), w));
}

// ---------------------------------------------------------------- propagation
// Per-lane queue state: the ring of pending bound indices and the list of
// "unstable" bounds live in LDS (one byte column per lane); membership masks and
// the 4-bit recursion counters live in registers.
template<class P>
struct PropScratch {
    unsigned char ring[P::NB > 0 ? P::NB : 1][kBlock];
    unsigned char ulist[P::NB > 0 ? P::NB : 1][kBlock];
};

__device__ __forceinline__ unsigned rec_get(const unsigned long long (&rec)[3], int a) {
    const unsigned long long w = a < 16 ? rec[0] : (a < 32 ? rec[1] : rec[2]);
    return static_cast<unsigned>((w >> ((a & 15) * 4)) & 15ull);
}
__device__ __forceinline__ void rec_inc(unsigned long long (&rec)[3], int a) {
    const unsigned long long inc = 1ull << ((a & 15) * 4);
    if(a < 16) rec[0] += inc; else if(a < 32) rec[1] += inc; else rec[2] += inc;
}

// FAST (the codegen's P::prop_fast for a level; the results are the same
// either way): the queue starts with the bounds that depend on `arg` and is
// worked off in order.  Every call site passes a compile-time `arg`, so the
// initial bounds are known here: they are applied first, in queue order,
// without the queue.  While none of them changes the box nothing else is
// queued, and when all leave it unchanged the propagation is over (the common
// case).  From the first change on, the queue below takes over in exactly the
// state the plain algorithm reaches there: the initial bounds after it still
// queued, and the dependents of its change queued after them.  The inlined
// initial bounds cost registers, which lowers the occupancy of some kernels.
template<class P, bool FAST = true, class Num>
__device__ __forceinline__ bool propagate(Num (&lb)[P::N], Num (&ub)[P::N], int arg, int it_limit, unsigned rec_limit,
                          const Num* C, PropScratch<P>& s)
{
    constexpr int NB = P::NB;
    static_assert(NB <= 48, "at most 48 bounds (4-bit recursion counters in three words)");
    if(NB == 0 || it_limit <= 0) return false;
    int done = 0, r_first = 0;
    if constexpr(FAST) {
        // a queue that receives all NB bounds at once reads as empty (begin ==
        // end, see the ring below), and then nothing is applied
        int m = 0;
        P::deps(arg * 3 + 2, [&](int) { ++m; });
        if(m >= NB) return false;
        P::deps(arg * 3 + 2, [&](int k) {
            if(r_first != 0) return;
            r_first = P::apply(k, lb, ub, C);
            ++done;
        });
        if(r_first < 0) return true;
        if(r_first == 0) return false;
    }
    const int lane = threadIdx.x;
    int qb = 0, qe = 0;
    unsigned long long present = 0, unstable = 0;
    unsigned long long rec[3] = {0, 0, 0};
    int nun = 0;
    // A full ring reads as empty (begin == end), as in the reference's IndexQueueSet.
    auto enqueue = [&](int e) -> bool {
        if((present >> e) & 1ull) return false;
        present |= 1ull << e;
        s.ring[qe][lane] = static_cast<unsigned char>(e);
        if(++qe >= NB) qe = 0;
        return true;
    };
    auto handle = [&](int r) {
        P::deps(r - 1, [&](int a) {
            if(rec_get(rec, a) >= rec_limit) {
                if(!((present >> a) & 1ull) && !((unstable >> a) & 1ull)) {
                    s.ulist[nun][lane] = static_cast<unsigned char>(a);
                    ++nun;
                    unstable |= 1ull << a;
                }
            } else if(enqueue(a)) {
                rec_inc(rec, a);
            }
        });
    };
    int seen = 0;
    P::deps(arg * 3 + 2, [&](int a) {
        if(seen++ >= done) enqueue(a);   // the initial bounds not yet applied
    });
    if(FAST) handle(r_first);
    for(int it = 0; it < it_limit; ++it) {
        while(qb != qe) {
            const int k = s.ring[qb][lane];
            present &= ~(1ull << k);
            if(++qb >= NB) qb = 0;
            const int r = P::apply(k, lb, ub, C);
            if(r < 0) return true;
            if(r > 0) handle(r);
        }
        if(nun > 0) {
            for(int j = 0; j < nun; ++j) enqueue(s.ulist[j][lane]);
            rec[0] = rec[1] = rec[2] = 0;
            unstable = 0;
            nun = 0;
        } else {
            break;
        }
    }
    return false;
}

// The inline first step of propagate<P, true> alone: the initial bounds of a
// change to `arg`, applied in queue order until the first one that changes the
// box.  -1: a bound emptied the box (propagate returns true); 0: none changed it
// and the propagation is over with this box (as in propagate); 1: one changed
// it, and only propagate itself, from the start, gives the propagated box.
template<class P, class Num>
__device__ __forceinline__ int propagate_initial(Num (&lb)[P::N], Num (&ub)[P::N], int arg, const Num* C) {
    constexpr int NB = P::NB;
    if(NB == 0) return 0;
    int m = 0;
    P::deps(arg * 3 + 2, [&](int) { ++m; });
    if(m >= NB) return 0;   // a full queue reads as empty: nothing is applied
    int r_first = 0;
    P::deps(arg * 3 + 2, [&](int k) {
        if(r_first == 0) r_first = P::apply(k, lb, ub, C);
    });
    return r_first < 0 ? -1 : (r_first > 0 ? 1 : 0);
}

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ unsigned long long wave_append(bool flag, unsigned long long* counter) {
    const unsigned long long mask = __ballot(flag);
    const unsigned long long active = __ballot(1);
    const int lane = __lane_id();
    const int first = __ffsll(static_cast<long long>(active)) - 1;
    unsigned long long base = 0;
    if(lane == first && mask) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, first);
    return base + static_cast<unsigned long long>(__popcll(mask & ((1ull << lane) - 1ull)));
}
// The same with n entries for every flagged lane (n equal on all lanes).
__device__ __forceinline__ unsigned long long wave_append_n(bool flag, unsigned n, unsigned long long* counter) {
    const unsigned long long mask = __ballot(flag);
    const unsigned long long active = __ballot(1);
    const int lane = __lane_id();
    const int first = __ffsll(static_cast<long long>(active)) - 1;
    unsigned long long base = 0;
    if(lane == first && mask) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)) * n);
    base = __shfl(base, first);
    return base + static_cast<unsigned long long>(__popcll(mask & ((1ull << lane) - 1ull))) * n;
}
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
// Words of one per-block count slot: cuboids, leaves, criticals, children
// decided by the FP32 prefilter, children it checked.
constexpr int kSlotWords = 8;

// Adds the block's counts to the block's own slot of `slots` (kSlotWords words
// per block index).  Device-scope atomics on one address from
// all XCDs serialise, so the counts are not accumulated in one place on the
// device: blocks of one launch own distinct slots and launches on a stream do
// not overlap, so a plain read-modify-write is exact, and the host sums the
// slots.  Every thread of the block calls this once, at the end of the kernel.
__device__ inline void block_counts_add(unsigned long long* slots, unsigned long long cub, unsigned long long leaf,
                                        unsigned long long crit, unsigned long long f32_dec = 0,
                                        unsigned long long f32_chk = 0) {
    __shared__ unsigned long long part[5][kBlock / 64];
    unsigned long long v[5] = {cub, leaf, crit, f32_dec, f32_chk};
    const unsigned w = threadIdx.x / 64u;
#pragma unroll
    for(int j = 0; j < 5; ++j) {
        v[j] = wave_sum(v[j]);
        if(__lane_id() == 0) part[j][w] = v[j];
    }
    __syncthreads();
    if(threadIdx.x < 5) {
        unsigned long long a = 0;
        for(unsigned i = 0; i < blockDim.x / 64u; ++i) a += part[threadIdx.x][i];
        unsigned long long* slot = slots + static_cast<unsigned long long>(kSlotWords) * blockIdx.x;
        if(a) slot[threadIdx.x] += a;
    }
}

template<class P>
__device__ __forceinline__ void load_consts(double* dst, const double* src) {
    for(int i = threadIdx.x; i < 2 * P::NC; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}
// Binary32 constants: each binary64 constant interval converted outward.
template<class P>
__device__ __forceinline__ void load_consts_f32(float* dst, const double* src) {
    for(int i = threadIdx.x; i < 2 * P::NC; i += blockDim.x) dst[i] = (i & 1) ? f32_up(src[i]) : f32_down(src[i]);
    __syncthreads();
}

// A value equal on every lane of the wave, held in scalar registers from here
// on (the compiler cannot tell that a value loaded from memory or derived from
// threadIdx.x / 64 is the same on all lanes, and keeps it in vector registers).
__device__ __forceinline__ unsigned long long wave_uniform(unsigned long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}
__device__ __forceinline__ unsigned long long input_base(const LevelArgs& a) { return a.in_base ? *a.in_base : 0ull; }
// Inputs [base, base + n) of a level, n clamped to the input list's capacity (a
// count past it was raised by a pass that could not store its survivors, which
// set the overflow flag and fails the call).
__device__ __forceinline__ unsigned long long input_count(const LevelArgs& a, unsigned long long base) {
    const unsigned long long n = *a.in_count;
    const unsigned long long room = a.in.cap > base ? a.in.cap - base : 0ull;
    return n < room ? n : room;
}

template<class P>
__device__ __forceinline__ void load_box(const BoxList& l, unsigned long long i, double (&lb)[P::N], double (&ub)[P::N]) {
#pragma unroll
    for(int a = 0; a < P::N; ++a) {
        lb[a] = l.lb[static_cast<unsigned long long>(a) * l.stride + i];
        ub[a] = l.ub[static_cast<unsigned long long>(a) * l.stride + i];
    }
}
template<class P>
__device__ __forceinline__ void store_box(const BoxList& l, unsigned long long i, const double (&lb)[P::N], const double (&ub)[P::N]) {
#pragma unroll
    for(int a = 0; a < P::N; ++a) {
        l.lb[static_cast<unsigned long long>(a) * l.stride + i] = lb[a];
        l.ub[static_cast<unsigned long long>(a) * l.stride + i] = ub[a];
    }
}

// ---------------------------------------------------------------- hull join
// Joins the leaves of lanes with `join` set into their dynamic cuboid's hull
// record [origin][arg][4]: min order key of lb, max order key of ub, and for
// bounds equal to zero the first leaf index with its sign bit (the reference
// keeps the first of equal bounds, so the sign of a zero bound is that of the
// first leaf in enumeration order).  Called by every lane of the wave.  When all
// joining lanes share one origin (the usual case: a wave covers consecutive
// children of one parent) the min/max are reduced across the wave first and one
// lane issues the atomics.
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}
template<int N>
__device__ __forceinline__ void hull_join(unsigned long long* hull, bool join, unsigned origin, unsigned leaf,
                                          const double (&lb)[N], const double (&ub)[N]) {
    const unsigned long long jm = __ballot(join);
    if(!jm) return;
    const int first = __ffsll(static_cast<long long>(jm)) - 1;
    const unsigned o0 = __shfl(origin, first);
    bool odd = join && origin != o0;
#pragma unroll
    for(int a = 0; a < N; ++a) odd = odd || (join && (lb[a] != lb[a] || ub[a] != ub[a]));  // NaN: keep per lane
    unsigned long long* rec = hull + static_cast<unsigned long long>(join ? origin : o0) * N * 4;
    if(__ballot(odd) == 0) {
#pragma unroll
        for(int a = 0; a < N; ++a) {
            const double mn = wave_min(join ? lb[a] : __builtin_inf());
            const double mx = wave_max(join ? ub[a] : -__builtin_inf());
            if(__lane_id() == first) {
                atomicMin(rec + a * 4 + 0, order_key(mn));
                atomicMax(rec + a * 4 + 1, order_key(mx));
            }
        }
    } else if(join) {
#pragma unroll
        for(int a = 0; a < N; ++a) {
            atomicMin(rec + a * 4 + 0, order_key(lb[a]));
            atomicMax(rec + a * 4 + 1, order_key(ub[a]));
        }
    }
    if(join) {
#pragma unroll
        for(int a = 0; a < N; ++a) {
            if(lb[a] == 0.0) atomicMin(rec + a * 4 + 2, (static_cast<unsigned long long>(leaf) << 1) | (bits(lb[a]) >> 63));
            if(ub[a] == 0.0) atomicMin(rec + a * 4 + 3, (static_cast<unsigned long long>(leaf) << 1) | (bits(ub[a]) >> 63));
        }
    }
}

// ---------------------------------------------------------------- static levels
// Largest group of consecutive children of one parent that one lane tests at
// once (a power of two dividing SUBS, at most 64).
template<int SUBS> constexpr int children_per_lane() {
    int c = 64;
    while(c > 1 && SUBS % c != 0) c >>= 1;
    return c;
}
// Children per group of a launch: as args.children_per_lane asks, or the most up
// to 16 that still leave two groups per thread of the grid (small launches need
// the parallelism more).
template<int SUBS, int CPL>
__device__ __forceinline__ int group_size(const LevelArgs& args, unsigned long long n_items, unsigned long long stride) {
    for(int c = CPL; c > 1; c >>= 1) {
        if(SUBS % c != 0) continue;
        if(args.children_per_lane > 0 ? c <= args.children_per_lane
                                      : c <= 16 && n_items * static_cast<unsigned long long>(SUBS / c) >= 2 * stride)
            return c;
    }
    return 1;
}

// Child w of a static level: the parent box with the level's variable replaced
// by subrange w % SUBS of the splitter (splitter.hpp:44-59), its dynamic cuboid
// (origin) and its enumeration index within that cuboid's grid (leaf).
template<class P, int L>
__device__ __forceinline__ void static_child(const LevelArgs& args, unsigned long long in_base, unsigned long long w,
                                             double (&lb)[P::N], double (&ub)[P::N], unsigned& origin, unsigned& leaf) {
    constexpr int ARG = P::st_arg(L);
    constexpr int SUBS = P::st_subs(L);
    const unsigned long long item = w / SUBS;
    const int i = static_cast<int>(w % SUBS);
    const unsigned long long src = in_base + item;
    load_box<P>(args.in, src, lb, ub);
    if(args.in.origin) { origin = args.in.origin[src]; leaf = args.in.leaf[src]; }
    else { origin = static_cast<unsigned>(item); leaf = 0; }
    leaf = leaf * SUBS + static_cast<unsigned>(i);
    const double pl = lb[ARG], pu = ub[ARG];
    const double wd = split_width(pl, pu, SUBS);
    lb[ARG] = split_point(pl, pu, wd, i, SUBS);
    ub[ARG] = split_point(pl, pu, wd, i + 1, SUBS);
}

// The binary64 decision for child w (cpu_prover_core.hpp:221-256): propagate,
// check the level's constraint rows, then count the leaf, join the hull (last
// non-final level), record a critical (last final level) or append the survivor
// to the next level's list.  Called by every lane of the wave (the hull join and
// the append are wave-wide); `valid` marks the lanes that hold a child.
template<class P, int MODE, int L>
__device__ __forceinline__ void static_decide(const LevelArgs& args, unsigned long long in_base, bool valid,
                                              unsigned long long w, const double* s_c, PropScratch<P>& s_prop,
                                              unsigned long long& c_leaf, unsigned long long& c_crit) {
    constexpr int N = P::N;
    constexpr int ARG = P::st_arg(L);
    constexpr bool LAST = (L == P::NSTAT - 1);
    bool survive = false, join = false;
    double lb[N], ub[N];
    unsigned origin = 0, leaf = 0;
    if(valid) {
        static_child<P, L>(args, in_base, w, lb, ub, origin, leaf);
        if(propagate<P, P::prop_fast(L)>(lb, ub, ARG, N, 4u, s_c, s_prop)) {
            ++c_leaf;
        } else if(P::template rows_violated<L>(lb, ub, s_c)) {
            ++c_leaf;
        } else if(LAST) {
            ++c_leaf;
            if(MODE == FINAL) { ++c_crit; survive = true; }
            else join = true;
        } else {
            survive = true;
        }
    }
    if(LAST && MODE == NONFINAL) hull_join<N>(args.hull, join, origin, leaf, lb, ub);
    if(!LAST || MODE == FINAL) {
        const unsigned long long pos = wave_append(survive, args.out.count);
        if(survive) {
            if(pos < args.out.cap) {
                store_box<P>(args.out, pos, lb, ub);
                if(args.out.origin) { args.out.origin[pos] = origin; args.out.leaf[pos] = leaf; }
            } else if(!LAST) {
                // a level list is sized so that this cannot happen; the critical
                // list's count keeps going past its capacity, which the host
                // compares with it (and then repeats the final pass in parts)
                atomicAdd(&args.counters->overflow, 1ull);
            }
        }
    }
}

// The binary64 decision of child i of the parent (plb, pub), whose terms free
// of the level's variable were hoisted into h (in LDS): as static_decide, but
// the rows are checked with the hoisted terms when propagation left every other
// variable bit for bit as in the parent (the child's box is then, in every
// variable those terms read, the box h was computed for), and in full otherwise.
// DEFER: the child is decided here only when the level's initial bounds leave
// the box unchanged (then the hoisted terms apply) or empty it; otherwise
// `defer` is set and nothing is counted or written for it: the caller decides
// it later with static_decide, together with other such children.
// Called by every lane of the wave.
template<class P, int MODE, int L, bool DEFER>
__device__ __forceinline__ void static_decide_hoisted(const LevelArgs& args, bool valid, const double (&plb)[P::N],
                                                      const double (&pub)[P::N], double pl, double pu, double wd, int i,
                                                      unsigned origin, unsigned leaf,
                                                      const typename P::template Hoist<L, double>& h,
                                                      const double* s_c, PropScratch<P>& s_prop,
                                                      unsigned long long& c_leaf, unsigned long long& c_crit,
                                                      bool& defer) {
    constexpr int N = P::N;
    constexpr int ARG = P::st_arg(L);
    constexpr int SUBS = P::st_subs(L);
    constexpr bool LAST = (L == P::NSTAT - 1);
    bool survive = false, join = false;
    double lb[N], ub[N];
#pragma unroll
    for(int a = 0; a < N; ++a) { lb[a] = plb[a]; ub[a] = pub[a]; }
    defer = false;
    if(valid) {
        lb[ARG] = split_point(pl, pu, wd, i, SUBS);
        ub[ARG] = split_point(pl, pu, wd, i + 1, SUBS);
        bool pruned;
        if constexpr(DEFER) {
            const int r = propagate_initial<P>(lb, ub, ARG, s_c);
            pruned = r < 0;
            defer = r > 0;
        } else {
            pruned = propagate<P, P::prop_fast(L)>(lb, ub, ARG, N, 4u, s_c, s_prop);
        }
        if(defer) {
        } else if(pruned) {
            ++c_leaf;
        } else {
            // with DEFER the box is the parent's apart from the level's variable
            bool same = true;
#pragma unroll
            for(int a = 0; a < N; ++a)
                if(!DEFER && a != ARG) same = same && bits(lb[a]) == bits(plb[a]) && bits(ub[a]) == bits(pub[a]);
            const bool violated = same ? P::template rows_violated_hoisted<L>(lb, ub, s_c, h)
                                       : P::template rows_violated<L>(lb, ub, s_c);
            if(violated) {
                ++c_leaf;
            } else if(LAST) {
                ++c_leaf;
                if(MODE == FINAL) { ++c_crit; survive = true; }
                else join = true;
            } else {
                survive = true;
            }
        }
    }
    if(LAST && MODE == NONFINAL) hull_join<N>(args.hull, join, origin, leaf, lb, ub);
    if(!LAST || MODE == FINAL) {
        const unsigned long long pos = wave_append(survive, args.out.count);
        if(survive) {
            if(pos < args.out.cap) {
                store_box<P>(args.out, pos, lb, ub);
                if(args.out.origin) { args.out.origin[pos] = origin; args.out.leaf[pos] = leaf; }
            } else if(!LAST) {
                // a level list is sized so that this cannot happen; the critical
                // list's count keeps going past its capacity, which the host
                // compares with it (and then repeats the final pass in parts)
                atomicAdd(&args.counters->overflow, 1ull);
            }
        }
    }
}

// Whether the binary64 static level L hoists its split-free terms per wave:
// when a wave can take all children of a parent in several trips of 64, the
// level has such terms, and it is not the first static level (whose parents,
// the round's dynamic cuboids, are usually too few to give every wave one).
template<class P, int L>
constexpr bool wave_hoisting() {
    return L >= 1 && P::st_subs(L) % 64 == 0 && P::st_subs(L) >= 128 &&
           sizeof(typename P::template Hoist<L, double>) > 1;
}

// Level L of the static grid (modes NONFINAL and FINAL).
//
// F32 = false: every child w of the level's parents is decided in binary64.
// Where wave_hoisting<P, L>() holds, each wave takes whole parents: it computes
// the parent's split-free terms once into its LDS slot, then decides the
// SUBS children in trips of 64 (static_decide_hoisted).
// F32 = true: the children listed in args.undec (by k_static_prefilter, those
// binary32 left undecided; by k_static_group64, those of the groups that passed
// the group test) are decided in binary64, exactly as above.
template<class P, int MODE, int L, bool F32>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(static_level_waves<P, L>())))
void k_static_level(LevelArgs job0) {
    const LevelArgs args = job_args(job0);
    constexpr int SUBS = P::st_subs(L);
    const unsigned long long in_base = input_base(args);
    const unsigned long long n_items = input_count(args, in_base);
    // the list holds at most undec_cap entries; a pass that found more counted
    // them (and raised the overflow flag, which fails the call) but stored only
    // those that fit
    const unsigned long long listed = F32 ? *args.undec_count : 0ull;
    const unsigned long long total =
        F32 ? (listed < args.undec_cap ? listed : args.undec_cap) : n_items * static_cast<unsigned long long>(SUBS);
    const unsigned long long first = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    // a block without inputs leaves at once (block-uniform: no barrier is skipped
    // by part of a block, and it has no counts to add); the hoisting path hands
    // out whole parents per wave
    if(!F32 && wave_hoisting<P, L>() ? static_cast<unsigned long long>(blockIdx.x) * (blockDim.x / 64) >= n_items
                                     : static_cast<unsigned long long>(blockIdx.x) * blockDim.x >= total)
        return;
    // the generated code takes the constants as literals (P::kc) and does not
    // read the array pointer its functions take
    const double* const s_c = nullptr;
    __shared__ PropScratch<P> s_prop;
    enter_round_down();
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    unsigned long long c_cub = 0, c_leaf = 0, c_crit = 0;
    // every lane of a wave runs the same number of loop trips, so the wave-wide
    // ballots always see all lanes of the wave
    if constexpr(!F32 && wave_hoisting<P, L>()) {
        constexpr int N = P::N;
        constexpr int ARG = P::st_arg(L);
        using H = typename P::template Hoist<L, double>;
        __shared__ H s_h[kBlock / 64];
        // the wave's loop state is the same on all its lanes: in scalar
        // registers it leaves the vector registers to the deferred children's
        // full decision, which spills otherwise
        const unsigned wib = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);   // wave in block
        H& h = s_h[wib];
        // Children whose initial bounds change another variable (a few percent;
        // they need the full propagation and the full rows) are listed per wave
        // and decided 64 at a time, so that a wave does not run that path for
        // one or two of its lanes.  The list holds fewer than 64 children
        // between parents, so 64 + SUBS entries suffice.
        constexpr bool DEFER = P::prop_fast(L);
        __shared__ unsigned long long s_def[kBlock / 64][DEFER ? 64 + SUBS : 1];
        unsigned long long* const dq = s_def[wib];
        unsigned dcount = 0;   // wave-uniform
        const int lane = static_cast<int>(__lane_id());
        const unsigned long long below = (1ull << lane) - 1ull;
        const unsigned long long waves = static_cast<unsigned long long>(gridDim.x) * (blockDim.x / 64);
        const unsigned long long wave = static_cast<unsigned long long>(blockIdx.x) * (blockDim.x / 64) + wib;
        const unsigned long long items = wave_uniform(n_items);
        const unsigned long long base = wave_uniform(in_base);
        const unsigned long long wtrips = (items + waves - 1) / waves;   // the same for every wave
        for(unsigned long long trip = 0; trip < wtrips; ++trip) {
            const unsigned long long item = wave + trip * waves;
            const bool valid = item < items;                              // wave-uniform
            double plb[N], pub[N];
            double pl = 0, pu = 0, wd = 0;
            unsigned p_origin = 0, p_leaf = 0;
#pragma unroll
            for(int a = 0; a < N; ++a) { plb[a] = 0.0; pub[a] = 0.0; }
            if(valid) {
                const unsigned long long src = base + item;
                load_box<P>(args.in, src, plb, pub);
                if(args.in.origin) { p_origin = args.in.origin[src]; p_leaf = args.in.leaf[src]; }
                else { p_origin = static_cast<unsigned>(item); p_leaf = 0; }
                pl = plb[ARG];
                pu = pub[ARG];
                wd = split_width(pl, pu, SUBS);
                H hv;
                P::template hoist<L>(plb, pub, s_c, hv);
                if(lane == 0) { h = hv; c_cub += SUBS; }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll 1
            for(int k = 0; k < SUBS / 64; ++k) {
                const int i = lane + 64 * k;
                bool defer;
                static_decide_hoisted<P, MODE, L, DEFER>(args, valid, plb, pub, pl, pu, wd, i, p_origin,
                                                         p_leaf * SUBS + static_cast<unsigned>(i), h, s_c, s_prop,
                                                         c_leaf, c_crit, defer);
                if constexpr(DEFER) {
                    const unsigned long long m = __ballot(defer);
                    if(defer) dq[dcount + __popcll(m & below)] = item * SUBS + static_cast<unsigned>(i);
                    dcount += static_cast<unsigned>(__popcll(m));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if constexpr(DEFER) {
                // full trips of 64 while the list holds them; after the last
                // parent, the rest
                const bool last = trip + 1 == wtrips;
                while(dcount >= 64 || (last && dcount > 0)) {
                    const unsigned take = dcount >= 64 ? 64u : dcount;
                    dcount -= take;
                    const bool v = static_cast<unsigned>(lane) < take;
                    const unsigned long long w = v ? dq[dcount + lane] : 0ull;
                    static_decide<P, MODE, L>(args, base, v, w, s_c, s_prop, c_leaf, c_crit);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
    } else {
        const unsigned long long trips = (total + stride - 1) / stride;
        for(unsigned long long trip = 0; trip < trips; ++trip) {
            const unsigned long long t = first + trip * stride;
            const bool valid = t < total;
            unsigned long long w = t;
            if(F32) w = valid ? args.undec[t] : 0ull;
            else if(valid && w % SUBS == 0) c_cub += SUBS;
            static_decide<P, MODE, L>(args, in_base, valid, w, s_c, s_prop, c_leaf, c_crit);
        }
    }
    block_counts_add(args.block_counts, c_cub, c_leaf, MODE == FINAL ? c_crit : 0ull);
}

// The FP32 prefilter of static level L.  Each child (the parent box with the
// level's variable A split) is tested in binary32 on outward-rounded boxes with
// outward binary32 constants: it is pruned when propagation empties it, or when
// a constraint of its rows fails on the hybrid box made of the parent's other
// variables and the propagated child's A.  Both tests are rigorous (the hybrid
// box contains the propagated child, since propagation only narrows).  The
// other children are appended to args.undec for k_static_level<..., true>,
// which decides them in binary64; survivors, hulls and criticals therefore
// always come from the binary64 evaluation.
//
// Groups: each lane first takes a group of cpl consecutive children of one
// parent (cpl from args.children_per_lane or the launch size, at most CPL) and
// checks the rows once on the parent's box with A over the whole group, the
// terms free of A evaluated once (P::hoist).  The outward-rounded operations are
// inclusion-isotone where defined (each bound is the directed rounding of the
// exact bound, a monotone function of the operands' bounds; a sub-box is
// possibly undefined only where the box is, and a comparison with a possibly
// undefined operand is never definite), and every child's hybrid box lies in
// the group's box, so a group that fails a row holds only children that the
// per-child test prunes too: they are counted as pruned by it.  The children of
// the other groups are gathered in LDS and spread over the block's lanes for
// the per-child test (propagation, then the rows on the hybrid box); the
// results, and the counts of checked and decided children, are those of the
// per-child test alone.  Propagation is not part of the group test: its
// iteration depends on which bounds changed, so it is not isotone.
// The binary32 pass runs in its own kernel so that its register budget (and
// with it the occupancy) is not that of the binary64 code.
template<class P, int MODE, int L>
__global__ __launch_bounds__(kBlock) SQD_LEVEL_ATTR void k_static_prefilter(LevelArgs job0) {
    const LevelArgs args = job_args(job0);
    constexpr int N = P::N;
    constexpr int ARG = P::st_arg(L);
    constexpr int SUBS = P::st_subs(L);
    constexpr int CPL = children_per_lane<SUBS>();   // most children per lane
    // constants are literals in the generated code (P::kc, binary32 converted
    // outward); the array pointer its functions take is not read
    const float* const s_c32 = nullptr;
    __shared__ PropScratch<P> s_prop;
    enter_round_down();
    const unsigned long long in_base = input_base(args);
    const unsigned long long n_items = input_count(args, in_base);
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    const int cpl = group_size<SUBS, CPL>(args, n_items, stride);
    const int step = SUBS / cpl;                                                       // groups per parent
    const unsigned long long total = n_items * static_cast<unsigned long long>(step);  // groups
    const unsigned long long first = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if(static_cast<unsigned long long>(blockIdx.x) * blockDim.x >= total) return;   // no groups in any trip (block-uniform)
    const unsigned long long trips = (total + stride - 1) / stride;
    __shared__ unsigned long long s_groups[kBlock];   // the trip's groups that the group test left
    __shared__ unsigned long long s_ngroups;
    // The hoisted terms of the parents of one block trip (at most MAXP of them:
    // a trip's kBlock consecutive groups, at least SUBS / CPL per parent), kept
    // for the per-child test when the table fits in 16 KiB.
    using H = typename P::template Hoist<L, float>;
    constexpr int MAXP = kBlock / (SUBS / CPL) + 2;
    constexpr bool SHARE = sizeof(H) > 1 && MAXP * sizeof(H) <= 16384;
    __shared__ H s_h[SHARE ? MAXP : 1];
    unsigned long long c_cub = 0, c_leaf = 0, c_dec = 0, c_chk = 0;
    for(unsigned long long trip = 0; trip < trips; ++trip) {
        if(threadIdx.x == 0) s_ngroups = 0;
        __syncthreads();
        // 1. group test: the lane's group is children [g * cpl, (g + 1) * cpl) of
        //    one parent; the rows are checked once on the parent's box with the
        //    level's variable over the whole group
        const unsigned long long s0 = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + trip * stride;
        const unsigned long long item0 = s0 / static_cast<unsigned long long>(step);   // the trip's first parent
        const unsigned long long slot = first + trip * stride;
        bool keep = false;
        if(slot < total) {
            const unsigned long long item = slot / static_cast<unsigned long long>(step);
            const int g = static_cast<int>(slot % static_cast<unsigned long long>(step));
            if(g == 0) c_cub += SUBS;
            keep = true;
            if(cpl > 1 || SHARE) {
                double lb[N], ub[N];
                load_box<P>(args.in, in_base + item, lb, ub);
                float hl[N], hu[N];
#pragma unroll
                for(int a = 0; a < N; ++a) { hl[a] = f32_down(lb[a]); hu[a] = f32_up(ub[a]); }
                H h;
                P::template hoist<L>(hl, hu, s_c32, h);
                if constexpr(SHARE) {
                    if(g == 0 || slot == s0) s_h[item - item0] = h;   // one lane per parent
                }
                if(cpl > 1) {
                    const double wd = split_width(lb[ARG], ub[ARG], SUBS);
                    hl[ARG] = f32_down(split_point(lb[ARG], ub[ARG], wd, g * cpl, SUBS));
                    hu[ARG] = f32_up(split_point(lb[ARG], ub[ARG], wd, (g + 1) * cpl, SUBS));
                    if(P::template rows_violated_hoisted<L>(hl, hu, s_c32, h)) {
                        keep = false;
                        c_chk += static_cast<unsigned long long>(cpl);
                        c_leaf += static_cast<unsigned long long>(cpl);
                        c_dec += static_cast<unsigned long long>(cpl);
                    }
                }
            }
        }
        const unsigned long long at = wave_append(keep, &s_ngroups);
        if(keep) s_groups[at] = slot;
        __syncthreads();
        // 2. the children of the groups left, spread over the block's lanes
        const unsigned long long m = s_ngroups * static_cast<unsigned long long>(cpl);
        for(unsigned long long base = 0; base < m; base += blockDim.x) {
            const unsigned long long q = base + threadIdx.x;
            bool undecided = false;
            unsigned long long w = 0;
            if(q < m) {
                const unsigned long long gs = s_groups[q / static_cast<unsigned long long>(cpl)];
                const unsigned long long item = gs / static_cast<unsigned long long>(step);
                const int i = static_cast<int>(gs % static_cast<unsigned long long>(step)) * cpl +
                              static_cast<int>(q % static_cast<unsigned long long>(cpl));
                w = item * SUBS + static_cast<unsigned long long>(i);
                double lb[N], ub[N];
                load_box<P>(args.in, in_base + item, lb, ub);
                float plb[N], pub[N];
#pragma unroll
                for(int a = 0; a < N; ++a) { plb[a] = f32_down(lb[a]); pub[a] = f32_up(ub[a]); }
                const double wd = split_width(lb[ARG], ub[ARG], SUBS);
                float lb32[N], ub32[N];
#pragma unroll
                for(int a = 0; a < N; ++a) { lb32[a] = plb[a]; ub32[a] = pub[a]; }
                lb32[ARG] = f32_down(split_point(lb[ARG], ub[ARG], wd, i, SUBS));
                ub32[ARG] = f32_up(split_point(lb[ARG], ub[ARG], wd, i + 1, SUBS));
                ++c_chk;
                bool pruned = propagate<P, P::prop_fast(L)>(lb32, ub32, ARG, N, 4u, s_c32, s_prop);
                if(!pruned) {
                    if constexpr(SHARE) {
                        plb[ARG] = lb32[ARG];
                        pub[ARG] = ub32[ARG];
                        pruned = P::template rows_violated_hoisted<L>(plb, pub, s_c32, s_h[item - item0]);
                    } else {
                        H h;
                        P::template hoist<L>(plb, pub, s_c32, h);
                        plb[ARG] = lb32[ARG];
                        pub[ARG] = ub32[ARG];
                        pruned = P::template rows_violated_hoisted<L>(plb, pub, s_c32, h);
                    }
                }
                if(pruned) {
                    ++c_leaf;
                    ++c_dec;
                } else {
                    undecided = true;
                }
            }
            const unsigned long long pos = wave_append(undecided, args.undec_count);
            if(undecided) {
                if(pos < args.undec_cap) args.undec[pos] = w;
                else atomicAdd(&args.counters->overflow, 1ull);
            }
        }
        __syncthreads();
    }
    block_counts_add(args.block_counts, c_cub, c_leaf, 0ull, c_dec, c_chk);
}

// The binary64 group test of static level L (the levels of
// sqd_settings.group_levels that the FP32 prefilter does not take): each lane
// checks the rows once, in binary64, on the parent's box with A over a group of
// cpl consecutive children (the terms free of A hoisted), and appends the
// children of a group that passes to args.undec; k_static_level<..., true> then
// decides them one by one exactly as the plain pass does.  The binary64
// operations are inclusion-isotone where defined (as for the FP32 prefilter's
// group test) and every child's propagated box lies in the group's box, so a group that fails a row holds only children that the plain
// pass prunes too (by propagation or by a row): they are counted as its pruned
// leaves, and survivors, hulls, criticals and counts are exactly the plain
// pass's.  The test runs in its own kernel so that neither pass carries the
// other's registers.
template<class P, int MODE, int L>
__global__ __launch_bounds__(kBlock) SQD_LEVEL_ATTR void k_static_group64(LevelArgs job0) {
    const LevelArgs args = job_args(job0);
    constexpr int N = P::N;
    constexpr int ARG = P::st_arg(L);
    constexpr int SUBS = P::st_subs(L);
    constexpr int CPL = children_per_lane<SUBS>();
    const double* const s_c = nullptr;   // constants are literals in the generated code (P::kc)
    enter_round_down();
    const unsigned long long in_base = input_base(args);
    const unsigned long long n_items = input_count(args, in_base);
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    const int cpl = group_size<SUBS, CPL>(args, n_items, stride);
    const int step = SUBS / cpl;                                                       // groups per parent
    const unsigned long long total = n_items * static_cast<unsigned long long>(step);  // groups
    const unsigned long long first = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if(static_cast<unsigned long long>(blockIdx.x) * blockDim.x >= total) return;   // no groups (block-uniform)
    const unsigned long long trips = (total + stride - 1) / stride;
    unsigned long long c_cub = 0, c_leaf = 0;
    for(unsigned long long trip = 0; trip < trips; ++trip) {
        const unsigned long long slot = first + trip * stride;
        bool keep = false;
        unsigned long long w0 = 0;
        if(slot < total) {
            const unsigned long long item = slot / static_cast<unsigned long long>(step);
            const int g = static_cast<int>(slot % static_cast<unsigned long long>(step));
            if(g == 0) c_cub += SUBS;
            w0 = item * SUBS + static_cast<unsigned long long>(g * cpl);
            keep = true;
            if(cpl > 1) {
                double lb[N], ub[N];
                load_box<P>(args.in, in_base + item, lb, ub);
                typename P::template Hoist<L, double> h;
                P::template hoist<L>(lb, ub, s_c, h);
                const double pl = lb[ARG], pu = ub[ARG];
                const double wd = split_width(pl, pu, SUBS);
                lb[ARG] = split_point(pl, pu, wd, g * cpl, SUBS);
                ub[ARG] = split_point(pl, pu, wd, (g + 1) * cpl, SUBS);
                if(P::template rows_violated_hoisted<L>(lb, ub, s_c, h)) {
                    keep = false;
                    c_leaf += static_cast<unsigned long long>(cpl);
                }
            }
        }
        const unsigned long long pos = wave_append_n(keep, static_cast<unsigned>(cpl), args.undec_count);
        if(keep) {
            if(pos + static_cast<unsigned long long>(cpl) <= args.undec_cap) {
                for(int k = 0; k < cpl; ++k) args.undec[pos + k] = w0 + static_cast<unsigned long long>(k);
            } else {
                atomicAdd(&args.counters->overflow, 1ull);
            }
        }
    }
    block_counts_add(args.block_counts, c_cub, c_leaf, 0ull);
}

// ---------------------------------------------------------------- dynamic levels
// Level J of the dynamic split (SPLIT: per-generation subdivisions, INIT: the
// initial queue).  The last level keeps children that possibly satisfy every
// constraint and appends them to the output (the cuboid stack).
template<class P, int MODE, int J>
__global__ __launch_bounds__(kBlock) SQD_LEVEL_ATTR void k_dynamic_level(LevelArgs job0) {
    const LevelArgs args = job_args(job0);
    constexpr int N = P::N;
    constexpr int ARG = P::dyn_arg(J);
    constexpr int SUBS = MODE == INIT ? P::dyn_initial(J) : P::dyn_subs(J);
    constexpr bool LAST = (J == P::NDYN - 1);
    constexpr unsigned REC = MODE == INIT ? (N < 255 ? N : 255) : (N < 32 ? N : 32);
    const double* const s_c = nullptr;   // constants are literals in the generated code (P::kc)
    __shared__ PropScratch<P> s_prop;
    enter_round_down();
    const unsigned long long in_base = input_base(args);
    const unsigned long long n_items = input_count(args, in_base);
    // the initial queue may start from a range of the first variable's subdivisions
    constexpr bool RANGED = (MODE == INIT && J == 0);
    const unsigned long long r_lo = RANGED && args.init_range ? args.init_range[0] : 0ull;
    const unsigned long long r_n = RANGED && args.init_range ? args.init_range[1] - r_lo : static_cast<unsigned long long>(SUBS);
    const unsigned long long total = n_items * r_n;
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    const unsigned long long first = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if(static_cast<unsigned long long>(blockIdx.x) * blockDim.x >= total) return;   // no children (block-uniform)
    const unsigned long long trips = (total + stride - 1) / stride;
    unsigned long long c_cub = 0, c_leaf = 0;
    for(unsigned long long trip = 0; trip < trips; ++trip) {
        const unsigned long long w = first + trip * stride;
        bool survive = false;
        double lb[N], ub[N];
        unsigned depth = 0;
        if(w < total) {
            const unsigned long long item = w / r_n;
            const int i = static_cast<int>(r_lo + w % r_n);
            const unsigned long long src = in_base + item;
            load_box<P>(args.in, src, lb, ub);
            depth = args.in.depth ? args.in.depth[src] : 0u;
            ++c_cub;
            const double pl = lb[ARG], pu = ub[ARG];
            const double wd = split_width(pl, pu, SUBS);
            lb[ARG] = split_point(pl, pu, wd, i, SUBS);
            ub[ARG] = split_point(pl, pu, wd, i + 1, SUBS);
            if(propagate<P>(lb, ub, ARG, 2 * N, REC, s_c, s_prop)) {
                ++c_leaf;
            } else if(LAST && !P::all_satisfied(lb, ub, s_c)) {
                ++c_leaf;
            } else {
                survive = true;
            }
        }
        const unsigned long long pos = wave_append(survive, args.out.count);
        if(survive) {
            if(pos < args.out.cap) {
                store_box<P>(args.out, pos, lb, ub);
                if(args.out.depth) args.out.depth[pos] = depth;
            } else {
                atomicAdd(&args.counters->overflow, 1ull);
            }
        }
    }
    block_counts_add(args.block_counts, c_cub, c_leaf, 0ull);
}

// ---------------------------------------------------------------- round glue
// Turn the hull records of one round into boxes.  Hulls at depth >= generations
// go to `fin` (final pass), the others to `spl` with depth + 1 (dynamic split).
// Without static variables the popped cuboid itself is the only leaf.
// Hull records start as (min key, max key, zero-sign keys) = (~0, 0, ~0, ~0)
// and are put back to that state by the kernel that consumes them.
__device__ __forceinline__ void clear_hull_record(unsigned long long* rec, int n_args) {
    for(int a = 0; a < n_args; ++a) {
        rec[a * 4 + 0] = ~0ull;
        rec[a * 4 + 1] = 0ull;
        rec[a * 4 + 2] = ~0ull;
        rec[a * 4 + 3] = ~0ull;
    }
}

// Turn the hull records of one round into boxes.  Hulls at depth >= generations
// go to `fin` (final pass), the others to `spl` with depth + 1 (dynamic split).
// Without static variables the popped cuboid itself is the only leaf.  The
// round's items are [*base, *base + *count) of `items`.  Row blockIdx.y of the
// grid works on job blockIdx.y.
template<class P>
__global__ __launch_bounds__(kBlock) void k_hull_finalize(BoxList items, const unsigned long long* base_ptr,
                                                          const unsigned long long* count_ptr, unsigned long long* hull,
                                                          BoxList fin, BoxList spl, DeviceCounters* counters,
                                                          unsigned long long* block_counts, unsigned generations,
                                                          unsigned long long job_stride) {
    constexpr int N = P::N;
    const unsigned long long off = job_offset(job_stride);
    items = job_list(items, off);
    fin = job_list(fin, off);
    spl = job_list(spl, off);
    base_ptr = job_ptr(base_ptr, off);
    count_ptr = job_ptr(count_ptr, off);
    hull = job_ptr(hull, off);
    counters = job_ptr(counters, off);
    block_counts = job_ptr(block_counts, off);
    const unsigned long long n = *count_ptr, base = *base_ptr;
    const unsigned long long d = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool have = false;
    double lb[N], ub[N];
    unsigned depth = 0;
    if(d < n) {
        depth = items.depth[base + d];
        if(P::NSTAT == 0) {
            load_box<P>(items, base + d, lb, ub);
            have = true;
        } else {
            unsigned long long* rec = hull + d * N * 4;
            have = rec[0] != ~0ull;
            if(have) {
#pragma unroll
                for(int a = 0; a < N; ++a) {
                    lb[a] = from_order_key(rec[a * 4 + 0]);
                    ub[a] = from_order_key(rec[a * 4 + 1]);
                    if(lb[a] == 0.0) lb[a] = (rec[a * 4 + 2] & 1ull) ? -0.0 : 0.0;
                    if(ub[a] == 0.0) ub[a] = (rec[a * 4 + 3] & 1ull) ? -0.0 : 0.0;
                }
                clear_hull_record(rec, N);
            }
        }
    }
    if(P::NSTAT == 0) block_counts_add(block_counts, 0ull, have ? 1ull : 0ull, 0ull);
    const bool to_final = have && depth >= generations;
    const bool to_split = have && !to_final;
    const unsigned long long pf = wave_append(to_final, fin.count);
    const unsigned long long ps = wave_append(to_split, spl.count);
    if(to_final) {
        if(pf < fin.cap) { store_box<P>(fin, pf, lb, ub); fin.depth[pf] = depth; }
        else atomicAdd(&counters->overflow, 1ull);
    }
    if(to_split) {
        if(ps < spl.cap) { store_box<P>(spl, ps, lb, ub); spl.depth[ps] = depth + 1; }
        else atomicAdd(&counters->overflow, 1ull);
    }
}

// Decode hull records in input order (sqd_nonfinal): entry d of `out` holds the
// hull of input d, has[d] = 0 when no leaf survived.
template<class P>
__global__ __launch_bounds__(kBlock) void k_hull_decode(BoxList items, unsigned long long n, unsigned long long* hull,
                                                        BoxList out, unsigned char* has) {
    constexpr int N = P::N;
    const unsigned long long d = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if(d >= n) return;
    double lb[N], ub[N];
    bool ok;
    if(P::NSTAT == 0) {
        load_box<P>(items, d, lb, ub);
        ok = true;
    } else {
        unsigned long long* rec = hull + d * N * 4;
        ok = rec[0] != ~0ull;
#pragma unroll
        for(int a = 0; a < N; ++a) {
            lb[a] = ok ? from_order_key(rec[a * 4 + 0]) : 0.0;
            ub[a] = ok ? from_order_key(rec[a * 4 + 1]) : 0.0;
            if(ok && lb[a] == 0.0) lb[a] = (rec[a * 4 + 2] & 1ull) ? -0.0 : 0.0;
            if(ok && ub[a] == 0.0) ub[a] = (rec[a * 4 + 3] & 1ull) ? -0.0 : 0.0;
        }
        if(ok) clear_hull_record(rec, N);
    }
    store_box<P>(out, d, lb, ub);
    has[d] = ok ? 1 : 0;
}

// Final pass without static variables: every cuboid is one critical leaf.  Row
// blockIdx.y of the grid works on job blockIdx.y.
template<class P>
__global__ __launch_bounds__(kBlock) void k_final_nostatic(BoxList fin, BoxList out, unsigned long long* block_counts,
                                                           unsigned long long job_stride) {
    const unsigned long long off = job_offset(job_stride);
    fin = job_list(fin, off);
    out = job_list(out, off);
    block_counts = job_ptr(block_counts, off);
    const unsigned long long n = *fin.count < fin.cap ? *fin.count : fin.cap;
    const unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool v = i < n;
    double lb[P::N], ub[P::N];
    if(v) load_box<P>(fin, i, lb, ub);
    const unsigned long long pos = wave_append(v, out.count);
    if(v && pos < out.cap) store_box<P>(out, pos, lb, ub);
    block_counts_add(block_counts, 0ull, v ? 1ull : 0ull, v ? 1ull : 0ull);
}

// Initial runtime box: outward compile-time box, then 50*N rounds of all bounds.
// First `zero_words` words at `zero` are cleared (a fresh control block).  The
// box goes to `out` (N lb, N ub) and, when list.lb is set, to entry 0 of `list`
// (depth 0) with the list's count set to 1.  Row blockIdx.y of the grid works on
// job blockIdx.y (every pointer shifted by its offset, as job_args does).
template<class P>
__global__ void k_initial_box(const double* box_lb, const double* box_ub, const double* consts, double* out,
                              BoxList list, unsigned long long* zero, unsigned zero_words,
                              unsigned long long job_stride) {
    const unsigned long long off = job_offset(job_stride);
    box_lb = job_ptr(box_lb, off);
    box_ub = job_ptr(box_ub, off);
    consts = job_ptr(consts, off);
    out = job_ptr(out, off);
    list = job_list(list, off);
    zero = job_ptr(zero, off);
    __shared__ double s_c[2 * (P::NC > 0 ? P::NC : 1)];
    for(unsigned i = threadIdx.x; zero && i < zero_words; i += blockDim.x) zero[i] = 0ull;
    __threadfence_block();
    load_consts<P>(s_c, consts);
    enter_round_down();
    if(threadIdx.x != 0 || blockIdx.x != 0) return;
    double lb[P::N], ub[P::N];
#pragma unroll
    for(int a = 0; a < P::N; ++a) { lb[a] = box_lb[a]; ub[a] = box_ub[a]; }
    // Up to 50 N sweeps over all bounds, each sweep stopping at an empty result.
    // A sweep that changes nothing leaves every later sweep unchanged too (each
    // bound application is a function of the box alone), so the loop stops there
    // with the same box the full count of sweeps would give.
    for(int it = 0; it < 50 * P::N; ++it) {
        bool changed = false;
        for(int k = 0; k < P::NB; ++k) {
            const int r = P::apply(k, lb, ub, s_c);
            if(r != 0) changed = true;
            if(r < 0) break;
        }
        if(!changed) break;
    }
#pragma unroll
    for(int a = 0; a < P::N; ++a) { out[a] = lb[a]; out[P::N + a] = ub[a]; }
    if(list.lb) {
        store_box<P>(list, 0, lb, ub);
        if(list.depth) list.depth[0] = 0;
        *list.count = 1;
    }
}

// FP32-versus-FP64 width of each constraint's compared difference (left minus
// right side of a top-level comparison) over n boxes given in SoA form (stride n).
// Per constraint j: out[3j] += sum of width32 / width64, out[3j + 1] += number of
// boxes where both widths are finite, defined and width64 > 0, out[3j + 2] = the
// largest ratio (bit pattern of a positive double, maximised as an integer).
// Statistics only: the quotient is not directionally rounded.
template<class P>
__global__ __launch_bounds__(kBlock) void k_width_stats(const double* lb, const double* ub, unsigned long long n,
                                                        const double* consts, double* out) {
    constexpr int N = P::N;
    __shared__ double s_c[2 * (P::NC > 0 ? P::NC : 1)];
    __shared__ float s_c32[2 * (P::NC > 0 ? P::NC : 1)];
    load_consts<P>(s_c, consts);
    load_consts_f32<P>(s_c32, s_c);
    enter_round_down();
    const unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    double bl[N], bu[N];
    float fl[N], fu[N];
#pragma unroll
    for(int a = 0; a < N; ++a) {
        bl[a] = valid ? lb[static_cast<unsigned long long>(a) * n + i] : 0.0;
        bu[a] = valid ? ub[static_cast<unsigned long long>(a) * n + i] : 0.0;
        fl[a] = f32_down(bl[a]);
        fu[a] = f32_up(bu[a]);
    }
    for(int j = 0; j < P::NCONS; ++j) {
        double r = 0.0;
        bool ok = false;
        if(valid) {
            const IV d = P::cons_diff(j, bl, bu, s_c);
            const IVT<float> f = P::cons_diff(j, fl, fu, s_c32);
            const double w64 = d.ub - d.lb;
            const double w32 = static_cast<double>(f.ub) - static_cast<double>(f.lb);
            ok = !d.und && !f.und && is_finite(w64) && is_finite(w32) && w64 > 0.0;
            r = ok ? w32 / w64 : 0.0;
        }
        double s = r;
#pragma unroll
        for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        const unsigned long long cnt = wave_sum(ok ? 1ull : 0ull);
        const double mx = wave_max(ok ? r : 0.0);
        if(__lane_id() == 0 && cnt) {
            atomicAdd(out + 3 * j, s);
            atomicAdd(out + 3 * j + 1, static_cast<double>(cnt));
            atomicMax(reinterpret_cast<unsigned long long*>(out + 3 * j + 2), bits(mx));
        }
    }
}

// Bound and constraint values on n boxes (SoA, stride n), for the enclosure
// checks against point evaluations: the interval of bound k's expression on box
// i goes to out_b[(i * NB + k) * 3 + {0: lb, 1: ub, 2: possibly undefined}], the
// IBool of constraint j (cons_id order) to out_c[(i * NCONS + j) * 2 + {0:
// definitely, 1: possibly}].  With Num = float the box is first converted
// outward to binary32, as the prefilter converts its boxes, and the results are
// widened exactly.
template<class P, class Num>
__global__ __launch_bounds__(kBlock) void k_eval_program(const double* lb, const double* ub, unsigned long long n,
                                                         double* out_b, unsigned char* out_c) {
    constexpr int N = P::N;
    enter_round_down();
    const unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i >= n) return;
    Num bl[N], bu[N];
#pragma unroll
    for(int a = 0; a < N; ++a) {
        const double l = lb[static_cast<unsigned long long>(a) * n + i];
        const double u = ub[static_cast<unsigned long long>(a) * n + i];
        if constexpr(std::is_same<Num, float>::value) { bl[a] = f32_down(l); bu[a] = f32_up(u); }
        else { bl[a] = l; bu[a] = u; }
    }
    const Num* const C = nullptr;   // constants are literals in the generated code
#pragma unroll 1
    for(int k = 0; k < P::NB; ++k) {
        const IVT<Num> v = P::bound_value(k, bl, bu, C);
        double* o = out_b + (i * P::NB + static_cast<unsigned long long>(k)) * 3;
        o[0] = static_cast<double>(v.lb);
        o[1] = static_cast<double>(v.ub);
        o[2] = v.und ? 1.0 : 0.0;
    }
#pragma unroll 1
    for(int j = 0; j < P::NCONS; ++j) {
        const IB c = P::cons_value(j, bl, bu, C);
        unsigned char* o = out_c + (i * P::NCONS + static_cast<unsigned long long>(j)) * 2;
        o[0] = c.def ? 1 : 0;
        o[1] = c.poss ? 1 : 0;
    }
}

} // namespace dev
} // namespace sqd
