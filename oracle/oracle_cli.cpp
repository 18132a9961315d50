// Command-line driver of the CPU oracle (test infrastructure only).
//   sqd_oracle <program-file> [--threads T] [--generations G] [--flops] [--slice FIRST LAST]
//              [--range FIRST LAST] [--fp32]
// --slice proves the initial-queue entries [FIRST, LAST); --range proves from the
// first dynamic variable's initial subdivisions [FIRST, LAST) only (the unit of
// sqd_prove_range).
// Prints one JSON line with the counts, the wall time, the thread count,
// with --fp32 the FP32 prefilter's checked/decided children and, with --flops,
// the binary64 (and binary32) operation counts (total and per phase).
#include "oracle.hpp"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>

int main(int argc, char** argv) {
    if(argc < 2) { std::fprintf(stderr, "usage: sqd_oracle <program> [--threads T] [--generations G] [--flops] [--slice A B] [--fp32]\n"); return 2; }
    int threads = static_cast<int>(std::thread::hardware_concurrency());
    int gens = 8;
    bool flops = false;
    bool slice = false;
    bool fp32 = false;
    bool range = false;
    unsigned long long first = 0, last = 0, r_first = 0, r_last = 0;
    for(int i = 2; i < argc; ++i) {
        if(!std::strcmp(argv[i], "--threads") && i + 1 < argc) threads = std::atoi(argv[++i]);
        else if(!std::strcmp(argv[i], "--generations") && i + 1 < argc) gens = std::atoi(argv[++i]);
        else if(!std::strcmp(argv[i], "--flops")) flops = true;
        else if(!std::strcmp(argv[i], "--fp32")) fp32 = true;
        else if(!std::strcmp(argv[i], "--slice") && i + 2 < argc) { slice = true; first = std::strtoull(argv[++i], nullptr, 10); last = std::strtoull(argv[++i], nullptr, 10); }
        else if(!std::strcmp(argv[i], "--range") && i + 2 < argc) { range = true; r_first = std::strtoull(argv[++i], nullptr, 10); r_last = std::strtoull(argv[++i], nullptr, 10); }
    }
    std::ifstream f(argv[1]);
    std::stringstream ss;
    ss << f.rdbuf();
    void* h = nullptr;
    if(oracle_load(ss.str().c_str(), &h) != 0) { std::fprintf(stderr, "load: %s\n", oracle_last_error()); return 1; }
    oracle_set_fp32_prefilter(h, fp32 ? 1 : 0);
    if(range && oracle_set_init_range(h, r_first, r_last) != 0) { std::fprintf(stderr, "bad --range\n"); return 2; }
    oracle_flop_enable(flops ? 1 : 0);
    oracle_flop_reset();
    OracleCounts c{};
    std::uint64_t ncrit = 0, iql = 0;
    auto t0 = std::chrono::steady_clock::now();
    int rc = slice ? oracle_run_slice(h, gens, threads, first, last, &c, &iql)
                   : oracle_run(h, gens, threads, &c, nullptr, 0, &ncrit);
    auto t1 = std::chrono::steady_clock::now();
    if(rc != 0) { std::fprintf(stderr, "run: %s\n", oracle_last_error()); return 1; }
    double s = std::chrono::duration<double>(t1 - t0).count();
    // per-phase operation and child-cuboid counts: "mode.level": [flops, cuboids, flops32]
    std::string phases;
    for(int m = 0; m < 5 && flops; ++m)
        for(int l = 0; l < (m == 4 ? 1 : 16); ++l) {
            std::uint64_t pf = 0, pu = 0, pf32 = 0;
            if(oracle_phase_counts(m, l, &pf, &pu) != 0 || oracle_phase_counts32(m, l, &pf32) != 0) continue;
            if(pf == 0 && pu == 0 && pf32 == 0) continue;
            char buf[128];
            std::snprintf(buf, sizeof buf, "%s\"%d.%d\": [%llu, %llu, %llu]", phases.empty() ? "" : ", ", m, l,
                          static_cast<unsigned long long>(pf), static_cast<unsigned long long>(pu),
                          static_cast<unsigned long long>(pf32));
            phases += buf;
        }
    std::uint64_t f32c = 0, f32d = 0;
    oracle_fp32_stats(h, &f32c, &f32d);
    std::printf("{\"num_cuboids\": %llu, \"num_leaf_cuboids\": %llu, \"num_critical_cuboids\": %llu, "
                "\"seconds\": %.6f, \"threads\": %d, \"flops\": %llu, \"initial_queue\": %llu, "
                "\"fp32_checked\": %llu, \"fp32_decided\": %llu, \"phases\": {%s}}\n",
                static_cast<unsigned long long>(c.num_cuboids), static_cast<unsigned long long>(c.num_leaf_cuboids),
                static_cast<unsigned long long>(c.num_critical_cuboids), s, threads,
                static_cast<unsigned long long>(oracle_flop_count()), static_cast<unsigned long long>(iql),
                static_cast<unsigned long long>(f32c), static_cast<unsigned long long>(f32d), phases.c_str());
    oracle_free(h);
    return 0;
}
