// Host build of the device's directed binary64 acos (acos_fixed.hpp, compiled
// here for the CPU), for tests/test_acos_host.py: RD and RU of acos(x) for n
// inputs, evaluated with the floating-point environment rounding to nearest
// (round_down = 0) or downward (round_down = 1, as the level kernels run).
#include <cfenv>
#include "acos_fixed.hpp"

extern "C" int acos_host_directed(const double* x, double* lo, double* hi, long n, int round_down) {
    const int saved = std::fegetround();
    std::fesetround(round_down ? FE_DOWNWARD : FE_TONEAREST);
    int hard = 0;
    for(long i = 0; i < n; ++i) {
        bool h = false;
        sqd::dev::acosfx::acos_directed(x[i], &lo[i], &hi[i], &h);
        hard += h ? 1 : 0;
    }
    std::fesetround(saved);
    return hard;
}
