// Host-side launch table for the kernels of one compiled program.  Each
// generated program file instantiates make_kernel_set<P>() and registers it under
// the program's fingerprint; the runtime looks the table up when a program is
// bound (sqd_program_from_text / sqd_program_from_lemma).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <utility>
#include "core.hpp"

namespace sqd {

struct KernelSet {
    std::uint64_t fingerprint;
    const char* key;
    int N, NB, NC, NDYN, NSTAT;
    const double* consts;   // 2*NC doubles: [lb, ub] per constant
    // Level launches.  `max_children` bounds the number of children of the
    // launch (host-known upper bound); `cus` is the device's compute-unit count.
    // The grid is the kernel's resident capacity (occupancy x CUs), capped by the
    // blocks the bound needs; a nonzero `grid` overrides it.  A static-level mode
    // or'ed with dev::kModeF32 selects the FP32-prefiltered variant, with
    // dev::kModeGroup the binary64 group test.
    hipError_t (*static_level)(int mode, int level, const dev::LevelArgs& a, unsigned long long max_children,
                               unsigned cus, unsigned grid, hipStream_t s);
    hipError_t (*dynamic_level)(int mode, int level, const dev::LevelArgs& a, unsigned long long max_children,
                                unsigned cus, unsigned grid, hipStream_t s);
    // Hulls of the round's items [*base, *base + *count) (at most max_n of them).
    // These three launch `jobs` grid rows, row j on job j's memory (job 0's
    // pointers plus j * job_stride bytes); the level launches take the same from
    // LevelArgs.
    hipError_t (*hull_finalize)(dev::BoxList items, const unsigned long long* base, const unsigned long long* count,
                                unsigned long long max_n, unsigned long long* hull, dev::BoxList fin,
                                dev::BoxList spl, dev::DeviceCounters* c, unsigned long long* block_counts,
                                unsigned gens, unsigned jobs, unsigned long long job_stride, hipStream_t s);
    hipError_t (*final_nostatic)(dev::BoxList fin, dev::BoxList out, unsigned long long* block_counts, unsigned grid,
                                 unsigned jobs, unsigned long long job_stride, hipStream_t s);
    hipError_t (*initial_box)(const double* lb, const double* ub, const double* consts, double* out,
                              dev::BoxList list, unsigned long long* zero, unsigned zero_words, unsigned jobs,
                              unsigned long long job_stride, hipStream_t s);
    hipError_t (*hull_decode)(dev::BoxList items, unsigned long long n, unsigned long long* hull,
                              dev::BoxList out, unsigned char* has, hipStream_t s);
    // Undecided acos roundings of this program so far (device-wide counter).
    hipError_t (*acos_hard)(unsigned long long* out, hipStream_t s);
    // Constraints (roots of the constraint rows) and the FP32-versus-FP64 width
    // statistics of their compared differences over n boxes (k_width_stats).
    int NCONS;
    int (*cons_id)(int j);
    hipError_t (*width_stats)(const double* lb, const double* ub, unsigned long long n, const double* consts,
                              double* out, hipStream_t s);
    // Bound and constraint values on n boxes (k_eval_program), in binary64
    // (precision 64) or binary32 (32).
    hipError_t (*eval_program)(int precision, const double* lb, const double* ub, unsigned long long n,
                               double* out_b, unsigned char* out_c, hipStream_t s);
};

void register_kernel_set(const KernelSet* ks);
const KernelSet* find_kernel_set(std::uint64_t fingerprint);

namespace detail {

// Resident blocks per CU of a kernel (queried once per instantiation).
template<typename K>
unsigned blocks_per_cu(K kernel) {
    int n = 0;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, dev::kBlock, 0) != hipSuccess || n < 1) n = 1;
    return static_cast<unsigned>(n);
}
// Rows of a level launch: one per job of the batch (job_args).
inline unsigned jobs_of(const dev::LevelArgs& a) { return a.jobs ? a.jobs : 1u; }
inline unsigned level_grid(unsigned resident, unsigned long long max_children, unsigned grid) {
    if(grid) return grid;
    const unsigned long long need = (max_children + dev::kBlock - 1) / dev::kBlock;
    return static_cast<unsigned>(need < 1 ? 1 : (need < resident ? need : resident));
}

// One static level: the binary64 kernel over every child, or (F32) the binary32
// prefilter followed by the binary64 kernel over the children it left undecided.
template<class P, int MODE, int L, bool F32>
hipError_t launch_one_static(const dev::LevelArgs& a, unsigned long long max_children, unsigned cus, unsigned grid,
                             hipStream_t s) {
    if(F32) {
        static const unsigned pre_cu = blocks_per_cu(dev::k_static_prefilter<P, MODE, L>);
        hipLaunchKernelGGL((dev::k_static_prefilter<P, MODE, L>), dim3(level_grid(pre_cu * cus, max_children, grid), jobs_of(a)),
                           dim3(dev::kBlock), 0, s, a);
        const hipError_t e = hipGetLastError();
        if(e != hipSuccess) return e;
    }
    static const unsigned per_cu = blocks_per_cu(dev::k_static_level<P, MODE, L, F32>);
    hipLaunchKernelGGL((dev::k_static_level<P, MODE, L, F32>),
                       dim3(level_grid(per_cu * cus, max_children, grid), jobs_of(a)), dim3(dev::kBlock), 0, s, a);
    return hipGetLastError();
}
// One static level with the binary64 group test: the test, then the binary64
// kernel over the children of the groups it left.
template<class P, int MODE, int L>
hipError_t launch_one_grouped(const dev::LevelArgs& a, unsigned long long max_children, unsigned cus, unsigned grid,
                              hipStream_t s) {
    static const unsigned pre_cu = blocks_per_cu(dev::k_static_group64<P, MODE, L>);
    hipLaunchKernelGGL((dev::k_static_group64<P, MODE, L>),
                       dim3(level_grid(pre_cu * cus, max_children, grid), jobs_of(a)), dim3(dev::kBlock), 0, s, a);
    const hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    static const unsigned per_cu = blocks_per_cu(dev::k_static_level<P, MODE, L, true>);
    hipLaunchKernelGGL((dev::k_static_level<P, MODE, L, true>),
                       dim3(level_grid(per_cu * cus, max_children, grid), jobs_of(a)), dim3(dev::kBlock), 0, s, a);
    return hipGetLastError();
}
template<class P, int MODE, int... L>
hipError_t launch_grouped(int level, const dev::LevelArgs& a, unsigned long long mc, unsigned cus, unsigned grid,
                          hipStream_t s, std::integer_sequence<int, L...>) {
    hipError_t e = hipErrorInvalidValue;
    (void)a; (void)mc; (void)cus; (void)grid; (void)s; (void)level;
    ((level == L ? (e = launch_one_grouped<P, MODE, L>(a, mc, cus, grid, s), 0) : 0), ...);
    return e;
}
template<class P, int MODE, int J>
hipError_t launch_one_dynamic(const dev::LevelArgs& a, unsigned long long max_children, unsigned cus, unsigned grid,
                              hipStream_t s) {
    static const unsigned per_cu = blocks_per_cu(dev::k_dynamic_level<P, MODE, J>);
    hipLaunchKernelGGL((dev::k_dynamic_level<P, MODE, J>),
                       dim3(level_grid(per_cu * cus, max_children, grid), jobs_of(a)), dim3(dev::kBlock), 0, s, a);
    return hipGetLastError();
}
template<class P, int MODE, bool F32, int... L>
hipError_t launch_static(int level, const dev::LevelArgs& a, unsigned long long mc, unsigned cus, unsigned grid,
                         hipStream_t s, std::integer_sequence<int, L...>) {
    hipError_t e = hipErrorInvalidValue;
    (void)a; (void)mc; (void)cus; (void)grid; (void)s; (void)level;
    ((level == L ? (e = launch_one_static<P, MODE, L, F32>(a, mc, cus, grid, s), 0) : 0), ...);
    return e;
}
template<class P, int MODE, int... J>
hipError_t launch_dynamic(int level, const dev::LevelArgs& a, unsigned long long mc, unsigned cus, unsigned grid,
                          hipStream_t s, std::integer_sequence<int, J...>) {
    hipError_t e = hipErrorInvalidValue;
    (void)a; (void)mc; (void)cus; (void)grid; (void)s; (void)level;
    ((level == J ? (e = launch_one_dynamic<P, MODE, J>(a, mc, cus, grid, s), 0) : 0), ...);
    return e;
}

template<class P>
hipError_t static_level(int mode, int level, const dev::LevelArgs& a, unsigned long long mc, unsigned cus,
                        unsigned grid, hipStream_t s) {
    auto seq = std::make_integer_sequence<int, P::NSTAT>{};
    if(mode == dev::NONFINAL) return launch_static<P, dev::NONFINAL, false>(level, a, mc, cus, grid, s, seq);
    if(mode == dev::FINAL) return launch_static<P, dev::FINAL, false>(level, a, mc, cus, grid, s, seq);
    if(mode == (dev::NONFINAL | dev::kModeF32))
        return launch_static<P, dev::NONFINAL, true>(level, a, mc, cus, grid, s, seq);
    if(mode == (dev::FINAL | dev::kModeF32)) return launch_static<P, dev::FINAL, true>(level, a, mc, cus, grid, s, seq);
    if(mode == (dev::NONFINAL | dev::kModeGroup)) return launch_grouped<P, dev::NONFINAL>(level, a, mc, cus, grid, s, seq);
    if(mode == (dev::FINAL | dev::kModeGroup)) return launch_grouped<P, dev::FINAL>(level, a, mc, cus, grid, s, seq);
    return hipErrorInvalidValue;
}
template<class P>
hipError_t dynamic_level(int mode, int level, const dev::LevelArgs& a, unsigned long long mc, unsigned cus,
                         unsigned grid, hipStream_t s) {
    auto seq = std::make_integer_sequence<int, P::NDYN>{};
    if(mode == dev::SPLIT) return launch_dynamic<P, dev::SPLIT>(level, a, mc, cus, grid, s, seq);
    if(mode == dev::INIT) return launch_dynamic<P, dev::INIT>(level, a, mc, cus, grid, s, seq);
    return hipErrorInvalidValue;
}
template<class P>
hipError_t hull_finalize(dev::BoxList items, const unsigned long long* base, const unsigned long long* count,
                         unsigned long long max_n, unsigned long long* hull, dev::BoxList fin, dev::BoxList spl,
                         dev::DeviceCounters* c, unsigned long long* block_counts, unsigned gens, unsigned jobs,
                         unsigned long long job_stride, hipStream_t s) {
    if(max_n == 0) return hipSuccess;
    unsigned grid = static_cast<unsigned>((max_n + dev::kBlock - 1) / dev::kBlock);
    hipLaunchKernelGGL((dev::k_hull_finalize<P>), dim3(grid, jobs ? jobs : 1u), dim3(dev::kBlock), 0, s, items, base,
                       count, hull, fin, spl, c, block_counts, gens, job_stride);
    return hipGetLastError();
}
template<class P>
hipError_t final_nostatic(dev::BoxList fin, dev::BoxList out, unsigned long long* block_counts, unsigned grid,
                          unsigned jobs, unsigned long long job_stride, hipStream_t s) {
    hipLaunchKernelGGL((dev::k_final_nostatic<P>), dim3(grid, jobs ? jobs : 1u), dim3(dev::kBlock), 0, s, fin, out,
                       block_counts, job_stride);
    return hipGetLastError();
}
template<class P>
hipError_t initial_box(const double* lb, const double* ub, const double* consts, double* out, dev::BoxList list,
                       unsigned long long* zero, unsigned zero_words, unsigned jobs, unsigned long long job_stride,
                       hipStream_t s) {
    hipLaunchKernelGGL((dev::k_initial_box<P>), dim3(1, jobs ? jobs : 1u), dim3(dev::kBlock), 0, s, lb, ub, consts,
                       out, list, zero, zero_words, job_stride);
    return hipGetLastError();
}

template<class P>
hipError_t hull_decode(dev::BoxList items, unsigned long long n, unsigned long long* hull, dev::BoxList out,
                       unsigned char* has, hipStream_t s) {
    if(n == 0) return hipSuccess;
    unsigned grid = static_cast<unsigned>((n + dev::kBlock - 1) / dev::kBlock);
    hipLaunchKernelGGL((dev::k_hull_decode<P>), dim3(grid), dim3(dev::kBlock), 0, s, items, n, hull, out, has);
    return hipGetLastError();
}

template<class P>
int cons_id(int j) { return P::cons_id(j); }
template<class P>
hipError_t width_stats(const double* lb, const double* ub, unsigned long long n, const double* consts, double* out,
                       hipStream_t s) {
    if(n == 0) return hipSuccess;
    const unsigned grid = static_cast<unsigned>((n + dev::kBlock - 1) / dev::kBlock);
    hipLaunchKernelGGL((dev::k_width_stats<P>), dim3(grid), dim3(dev::kBlock), 0, s, lb, ub, n, consts, out);
    return hipGetLastError();
}

template<class P>
hipError_t eval_program(int precision, const double* lb, const double* ub, unsigned long long n, double* out_b,
                        unsigned char* out_c, hipStream_t s) {
    if(n == 0) return hipSuccess;
    const unsigned grid = static_cast<unsigned>((n + dev::kBlock - 1) / dev::kBlock);
    if(precision == 32)
        hipLaunchKernelGGL((dev::k_eval_program<P, float>), dim3(grid), dim3(dev::kBlock), 0, s, lb, ub, n, out_b, out_c);
    else
        hipLaunchKernelGGL((dev::k_eval_program<P, double>), dim3(grid), dim3(dev::kBlock), 0, s, lb, ub, n, out_b, out_c);
    return hipGetLastError();
}

inline hipError_t acos_hard(unsigned long long* out, hipStream_t s) {
    return hipMemcpyFromSymbolAsync(out, HIP_SYMBOL(dev::g_acos_hard), sizeof(unsigned long long), 0,
                                    hipMemcpyDeviceToHost, s);
}

} // namespace detail

template<class P>
KernelSet make_kernel_set(std::uint64_t fingerprint, const char* key, const double* consts) {
    KernelSet k;
    k.fingerprint = fingerprint;
    k.key = key;
    k.N = P::N;
    k.NB = P::NB;
    k.NC = P::NC;
    k.NDYN = P::NDYN;
    k.NSTAT = P::NSTAT;
    k.consts = consts;
    k.static_level = &detail::static_level<P>;
    k.dynamic_level = &detail::dynamic_level<P>;
    k.hull_finalize = &detail::hull_finalize<P>;
    k.final_nostatic = &detail::final_nostatic<P>;
    k.initial_box = &detail::initial_box<P>;
    k.hull_decode = &detail::hull_decode<P>;
    k.acos_hard = &detail::acos_hard;
    k.NCONS = P::NCONS;
    k.cons_id = &detail::cons_id<P>;
    k.width_stats = &detail::width_stats<P>;
    k.eval_program = &detail::eval_program<P>;
    return k;
}

} // namespace sqd
