// Command-line front end of the lemma lowering.
//   sqd_lower list                    list the registered lemmas
//   sqd_lower dump <key> [file]       write the lowered program of one lemma
//   sqd_lower dump-all <dir>          write <dir>/<key>.prog for every lemma
//   sqd_lower summary <key>           print the table sizes of one lemma
#include <fstream>
#include <iostream>
#include <iomanip>
#include "lemmas.hpp"
#include "ivarp/codegen.hpp"

using namespace sqd_lemmas;

static void summary(const LemmaDef& d, const ivarp::flat::Program& p) {
    std::cout << d.key << ": N=" << p.nargs << " nodes=" << p.nodes.size() << " bounds=" << p.bounds.size()
              << " rows=[";
    for(int a = 0; a < p.nargs; ++a) std::cout << (a ? "," : "") << p.rows[static_cast<std::size_t>(a)].size();
    std::cout << "] iqs=" << p.initial_queue_size << " box=";
    std::cout << std::setprecision(17);
    for(int a = 0; a < p.nargs; ++a) std::cout << " [" << p.box_lb[static_cast<std::size_t>(a)] << ","
                                              << p.box_ub[static_cast<std::size_t>(a)] << "]";
    std::cout << " fp=" << std::hex << p.fingerprint() << std::dec << "\n";
}

int main(int argc, char** argv) {
    try {
        std::string cmd = argc > 1 ? argv[1] : "list";
        if(cmd == "list") {
            for(const auto& d : all_lemmas()) std::cout << d.key << "\t" << d.title << "\n";
            return 0;
        }
        if(cmd == "summary" && argc > 2) {
            const auto& d = lemma(argv[2]);
            summary(d, lower_lemma(d));
            return 0;
        }
        if(cmd == "dump" && argc > 2) {
            const auto& d = lemma(argv[2]);
            auto p = lower_lemma(d);
            if(argc > 3) { std::ofstream f(argv[3]); p.write(f); }
            else p.write(std::cout);
            return 0;
        }
        if(cmd == "dump-all" && argc > 2) {
            for(const auto& d : all_lemmas()) {
                auto p = lower_lemma(d);
                std::ofstream f(std::string(argv[2]) + "/" + d.key + ".prog");
                p.write(f);
                summary(d, p);
            }
            return 0;
        }
        if(cmd == "gen" && argc > 2) {
            // lowered program text and generated HIP source for every lemma the GPU supports
            for(const auto& d : all_lemmas()) {
                auto p = lower_lemma(d);
                { std::ofstream f(std::string(argv[2]) + "/" + d.key + ".prog"); p.write(f); }
                if(!ivarp::codegen::supported(p)) { std::cout << "skip " << d.key << "\n"; continue; }
                std::ofstream f(std::string(argv[2]) + "/" + d.key + ".hip");
                f << ivarp::codegen::generate(p, d.key);
                std::cout << "gen " << d.key << "\n";
            }
            return 0;
        }
        std::cerr << "usage: sqd_lower list | summary <key> | dump <key> [file] | dump-all <dir> | gen <dir>\n";
        return 2;
    } catch(const std::exception& e) {
        std::cerr << "sqd_lower: " << e.what() << "\n";
        return 1;
    }
}
