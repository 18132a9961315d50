// CPU oracle (test infrastructure only).  See oracle.hpp for the restated
// reference functions.  All arithmetic runs with the FPU in round-toward-minus-
// infinity, upper bounds are computed as -RD(-x op y) like the reference, and
// the upper bound of sqrt switches the rounding mode to upward around one
// sqrtsd.  The value barrier `opq` keeps the compiler from folding the
// negations (the build also uses -frounding-math).
#include "oracle.hpp"
#include <atomic>
#include <cfenv>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <set>
#include <vector>

#include <mpfr.h>

namespace {

// ---------------------------------------------------------------- arithmetic
// Operation counting for the roofline figures: every binary64 add, sub, mul,
// div and sqrt is counted once, attributed to the phase that executes it.  A
// phase is (mode, level) with mode 0 non-final static level, 1 final static
// level, 2 dynamic split level, 3 initial-queue level (the kernel kinds of the
// device path); kPhaseOther collects everything else (initial box).
constexpr int kPhaseLevels = 16;
constexpr int kPhaseOther = 4 * kPhaseLevels;
constexpr int kPhases = kPhaseOther + 1;
static std::atomic<std::uint64_t> g_flops{0};
static std::atomic<std::uint64_t> g_phase_flops[kPhases];
static std::atomic<std::uint64_t> g_phase_units[kPhases];
static bool g_count_flops = false;
thread_local std::uint64_t t_phase_flops[kPhases];
thread_local std::uint64_t t_phase_units[kPhases];
thread_local int t_phase = kPhaseOther;

inline double opq(double x) { __asm__ volatile("" : "+x"(x)); return x; }
inline float opq(float x) { __asm__ volatile("" : "+x"(x)); return x; }
inline void count(int k) { if(g_count_flops) t_phase_flops[t_phase] += static_cast<std::uint64_t>(k); }
// binary32 operations (the FP32 prefilter) are counted apart
thread_local std::uint64_t t_phase_flops32[kPhases];
static std::atomic<std::uint64_t> g_phase_flops32[kPhases];
inline void count32(int k) { if(g_count_flops) t_phase_flops32[t_phase] += static_cast<std::uint64_t>(k); }
inline void enter_phase(int mode, std::size_t level, std::uint64_t units) {
    t_phase = level < static_cast<std::size_t>(kPhaseLevels) ? mode * kPhaseLevels + static_cast<int>(level) : kPhaseOther;
    if(g_count_flops) t_phase_units[t_phase] += units;
}
// Moves this thread's counts into the global totals.
void flush_thread_counts() {
    for(int i = 0; i < kPhases; ++i) {
        g_flops += t_phase_flops[i];
        g_phase_flops[i] += t_phase_flops[i];
        g_phase_units[i] += t_phase_units[i];
        g_phase_flops32[i] += t_phase_flops32[i];
        t_phase_flops[i] = 0;
        t_phase_units[i] = 0;
        t_phase_flops32[i] = 0;
    }
    t_phase = kPhaseOther;
}

// The whole oracle runs with the FPU in round-toward-minus-infinity (SSE for
// binary64 and binary32 alike), so each operation is one correctly rounded
// downward result; upper bounds use RU(x) = -RD(-x), square roots toggle the
// mode (math_fn/interval_sqrt.hpp:72-116).
inline double rd_add(double a, double b) { count(1); return opq(opq(a) + opq(b)); }
inline double rd_sub(double a, double b) { count(1); return opq(opq(a) - opq(b)); }
inline double rd_mul(double a, double b) { count(1); return opq(opq(a) * opq(b)); }
inline double rd_div(double a, double b) { count(1); return opq(opq(a) / opq(b)); }
inline double rd_sqrt(double a) { count(1); return opq(__builtin_sqrt(opq(a))); }
inline float rd_add(float a, float b) { count32(1); return opq(opq(a) + opq(b)); }
inline float rd_sub(float a, float b) { count32(1); return opq(opq(a) - opq(b)); }
inline float rd_mul(float a, float b) { count32(1); return opq(opq(a) * opq(b)); }
inline float rd_div(float a, float b) { count32(1); return opq(opq(a) / opq(b)); }
inline float rd_sqrt(float a) { count32(1); return opq(__builtin_sqrtf(opq(a))); }
template<class X> inline X ru_add(X a, X b) { return -rd_add(-a, -b); }
template<class X> inline X ru_sub(X a, X b) { return -rd_add(-a, b); }
template<class X> inline X ru_mul(X a, X b) { return -rd_mul(-a, b); }
template<class X> inline X ru_div(X a, X b) { return -rd_div(-a, b); }
inline double sqrt_up(double a) { return __builtin_sqrt(a); }
inline float sqrt_up(float a) { return __builtin_sqrtf(a); }
template<class X> inline X ru_sqrt(X a) {
    if(sizeof(X) == 4) count32(1); else count(1);
    __asm__ volatile("" ::: "memory");
    std::fesetround(FE_UPWARD);
    __asm__ volatile("" ::: "memory");
    X r = opq(sqrt_up(opq(a)));
    __asm__ volatile("" ::: "memory");
    std::fesetround(FE_DOWNWARD);
    __asm__ volatile("" ::: "memory");
    return r;
}
// Outward conversion of a double to float (rounding down, resp. up).
inline float f32_down(double d) { return static_cast<float>(opq(d)); }
inline float f32_up(double d) { return -static_cast<float>(opq(-d)); }

struct RoundDownGuard {
    int old;
    RoundDownGuard() : old(std::fegetround()) { std::fesetround(FE_DOWNWARD); }
    ~RoundDownGuard() { std::fesetround(old); }
};

constexpr double kInf = std::numeric_limits<double>::infinity();
template<class X> constexpr X inf_v() { return std::numeric_limits<X>::infinity(); }
constexpr std::int64_t kMaxB = INT64_C(9223372036854775807);
constexpr std::int64_t kMinB = -kMaxB;

template<class X> struct IVT { X lb, ub; bool und; };
using IV = IVT<double>;
struct IB { bool def, poss; };

// compile-time bound classification helpers (fixed_point_bounds.hpp:72-180)
struct CT { std::int64_t lb, ub; };
inline bool ct_finite(CT b) { return b.lb > kMinB && b.ub < kMaxB; }
template<class X> inline bool nonneg(CT b, const IVT<X>& f) { if(b.lb >= 0) return true; if(b.ub <= 0) return false; return f.lb >= 0; }
template<class X> inline bool nonpos(CT b, const IVT<X>& f) { if(b.ub <= 0) return true; if(b.lb >= 0) return false; return f.ub <= 0; }
template<class X> inline bool positive(CT b, const IVT<X>& f) { if(b.lb > 0) return true; if(b.ub <= 0) return false; return f.lb > 0; }
template<class X> inline bool negative(CT b, const IVT<X>& f) { if(b.ub < 0) return true; if(b.lb >= 0) return false; return f.ub < 0; }
const CT kUnbounded{kMinB, kMaxB};

template<class X> inline X nan_to_zero(X v) { return v != v ? X(0) : v; }
template<class X> inline void do_mul(X l1, X l2, X u1, X u2, bool finite, X& lo, X& hi) {
    lo = rd_mul(l1, l2);
    hi = ru_mul(u1, u2);
    if(!finite) { lo = nan_to_zero(lo); hi = nan_to_zero(hi); }
}
template<class X> IVT<X> mul_check_signs(const IVT<X>& a, const IVT<X>& b, CT B1, CT B2) {
    const bool fin = ct_finite(B1) && ct_finite(B2);
    X lo, hi;
    if(nonneg(B1, a)) {
        if(nonneg(B2, b)) do_mul(a.lb, b.lb, a.ub, b.ub, fin, lo, hi);
        else if(nonpos(B2, b)) do_mul(a.ub, b.lb, a.lb, b.ub, fin, lo, hi);
        else do_mul(a.ub, b.lb, a.ub, b.ub, fin, lo, hi);
    } else if(nonpos(B1, a)) {
        if(nonneg(B2, b)) do_mul(a.lb, b.ub, a.ub, b.lb, fin, lo, hi);
        else if(nonpos(B2, b)) do_mul(a.ub, b.ub, a.lb, b.lb, fin, lo, hi);
        else do_mul(a.lb, b.ub, a.lb, b.lb, fin, lo, hi);
    } else if(nonneg(B2, b)) {
        do_mul(a.lb, b.ub, a.ub, b.ub, fin, lo, hi);
    } else if(nonpos(B2, b)) {
        do_mul(a.ub, b.lb, a.lb, b.lb, fin, lo, hi);
    } else {
        X l1, h1, l2, h2;
        do_mul(a.lb, b.ub, a.lb, b.lb, fin, l1, h1);
        do_mul(b.lb, a.ub, a.ub, b.ub, fin, l2, h2);
        lo = (l2 <= l1) ? l2 : l1;
        hi = (h1 <= h2) ? h2 : h1;
    }
    return IVT<X>{lo, hi, a.und || b.und};
}

template<class X> inline int sgn(X v) { return (v > 0) - (v < 0); }
template<class X> void divide_by_zero(IVT<X>& num, const IVT<X>& den, int sl, int su, int dsum) {
    const X inf = inf_v<X>();
    num.und = true;
    if(dsum == 0 || su - sl == 2) { num.lb = -inf; num.ub = inf; return; }
    if(dsum > 0) {
        if(sl >= 0) { num.lb = rd_div(num.lb, den.ub); num.ub = inf; }
        else { num.ub = ru_div(num.ub, den.ub); num.lb = -inf; }
    } else {
        if(sl >= 0) { num.ub = ru_div(num.lb, den.lb); num.lb = -inf; }
        else { num.lb = rd_div(num.ub, den.lb); num.ub = inf; }
    }
}
template<class X> struct WInf { X v; int inf_sign; };
template<class X> inline X div_lb_inf(WInf<X> n, WInf<X> d) {
    if(!n.inf_sign && !d.inf_sign) return rd_div(n.v, d.v);
    if(d.inf_sign != 0) return X(0);
    return -inf_v<X>();
}
template<class X> inline X div_ub_inf(WInf<X> n, WInf<X> d) {
    if(!n.inf_sign && !d.inf_sign) return ru_div(n.v, d.v);
    if(d.inf_sign != 0) return X(0);
    return inf_v<X>();
}
template<class X> void divide_by_infinite(IVT<X>& num, const IVT<X>& den) {
    num.und = num.und || den.und;
    int snl = sgn(num.lb), snu = sgn(num.ub), sdl = sgn(den.lb), sdu = sgn(den.ub);
    WInf<X> nl{num.lb, std::isfinite(num.lb) ? 0 : -1}, nu{num.ub, std::isfinite(num.ub) ? 0 : 1};
    WInf<X> dl{den.lb, std::isfinite(den.lb) ? 0 : -1}, du{den.ub, std::isfinite(den.ub) ? 0 : 1};
    auto set = [&](WInf<X> a, WInf<X> b, WInf<X> c, WInf<X> d) { X l = div_lb_inf(a, b), u = div_ub_inf(c, d); num.lb = l; num.ub = u; };
    if(sdl > 0) {
        if(snl >= 0) set(nl, du, nu, dl);
        else if(snu <= 0) set(nl, dl, nu, du);
        else set(nl, dl, nu, dl);
    } else if(sdu < 0) {
        if(snl >= 0) set(nu, du, nl, dl);
        else if(snu <= 0) set(nu, dl, nl, du);
        else set(nu, du, nl, du);
    } else {
        divide_by_zero(num, den, snl, snu, sdl + sdu);
    }
}
template<class X> IVT<X> interval_div(IVT<X> num, const IVT<X>& den, CT BN, CT BD) {
    bool fin_ct = ct_finite(BN) && ct_finite(BD);
    bool fin_rt = std::isfinite(num.lb) && std::isfinite(num.ub) && std::isfinite(den.lb) && std::isfinite(den.ub);
    if(!(fin_ct || fin_rt)) { divide_by_infinite(num, den); return num; }
    num.und = num.und || den.und;
    auto ia = [&](X a, X b, X c, X d) { X l = rd_div(a, b), u = ru_div(c, d); num.lb = l; num.ub = u; };
    if(positive(BD, den)) {
        if(nonneg(BN, num)) ia(num.lb, den.ub, num.ub, den.lb);
        else if(nonpos(BN, num)) ia(num.lb, den.lb, num.ub, den.ub);
        else ia(num.lb, den.lb, num.ub, den.lb);
    } else if(negative(BD, den)) {
        if(nonneg(BN, num)) ia(num.ub, den.ub, num.lb, den.lb);
        else if(nonpos(BN, num)) ia(num.ub, den.lb, num.lb, den.ub);
        else ia(num.ub, den.ub, num.lb, den.ub);
    } else {
        divide_by_zero(num, den, sgn(num.lb), sgn(num.ub), sgn(den.lb) + sgn(den.ub));
    }
    return num;
}
template<class X> IVT<X> interval_sqrt(const IVT<X>& x, CT B) {
    X lb = x.lb, ub = x.ub;
    bool und = x.und;
    if(!nonneg(B, x)) {
        if(ub < 0) return IVT<X>{-inf_v<X>(), inf_v<X>(), true};
        lb = 0;
        und = true;
    }
    return IVT<X>{rd_sqrt(lb), ru_sqrt(ub), und};
}
// acos rounded down / up, correctly (MPFR at 53 bits), like the reference's
// cpu_acos_rd / cpu_acos_ru (ivarp/src/interval_acos.cpp:29-66).  NaN outside [-1, 1].
double acos_dir(double x, bool up) {
    count(1);
    mpfr_t m;
    mpfr_init2(m, 53);
    mpfr_set_d(m, x, MPFR_RNDN);
    mpfr_acos(m, m, up ? MPFR_RNDU : MPFR_RNDD);
    const double r = mpfr_get_d(m, up ? MPFR_RNDU : MPFR_RNDD);
    mpfr_clear(m);
    return r;
}
// fn_domain_handle_bounds (math_fn/interval_asin.hpp:13-27) + interval_acos (interval_acos.hpp:19-29)
IV interval_acos(IV x, CT B) {
    constexpr std::int64_t kOne = 1000000;
    bool und = x.und;
    if(B.lb < -kOne && x.lb < -1.0) { x.lb = -1.0; und = true; }
    if(B.ub > kOne && x.ub > 1.0) { x.ub = 1.0; und = true; }
    return IV{acos_dir(x.ub, false), acos_dir(x.lb, true), und};
}
// binary32 (the FP32 prefilter's own definition, not the reference's): an
// enclosure of acos in binary64 interval arithmetic, converted outward.  The
// arcsine series P(z) = sum_k c_k z^k, c_k = binom(2k, k) / (4^k (2k + 1)), on
// z = t^2 (|t| <= 1/2: acos t = pi/2 - t P(t^2)) or on z = u = (1 - |t|) / 2
// (acos|t| = 2 sqrt(u) P(u), acos t = pi - acos|t| for t < 0), 16 terms with
// every operation rounded outward, plus 2 c_16 4^-16 for the tail (z <= 1/4).
// The same operations in the same order as the device (csrc/acos_enclosure.hpp).
constexpr int kAcosTerms = 16;
constexpr double acos_coeff(int k, bool up) {
    unsigned __int128 num = 1, den = 1;
    for(int i = 1; i <= k; ++i) num = num * static_cast<unsigned __int128>(k + i) / static_cast<unsigned __int128>(i);
    for(int i = 0; i < k; ++i) den *= 4;
    den *= static_cast<unsigned __int128>(2 * k + 1);
    int sh = 0;
    while((num << sh) < (den << 52)) ++sh;
    const unsigned __int128 sc = num << sh;
    unsigned long long q = static_cast<unsigned long long>(sc / den);
    if(up && sc % den != 0) ++q;
    double r = static_cast<double>(q);
    for(int i = 0; i < sh; ++i) r *= 0.5;
    return r;
}
inline double rd_fma(double a, double b, double c) { count(1); return opq(std::fma(opq(a), opq(b), opq(c))); }
inline double ru_fma(double a, double b, double c) { return -rd_fma(-a, b, -c); }
void acos_enclose(double t, double& lo, double& hi) {
    static const double kHalfPiLo = 0x1.921fb54442d18p+0, kHalfPiHi = 0x1.921fb54442d19p+0;
    static const double kPiLo = 0x1.921fb54442d18p+1, kPiHi = 0x1.921fb54442d19p+1;
    static double clo[kAcosTerms], chi[kAcosTerms], tail;
    static const bool init = [] {
        for(int k = 0; k < kAcosTerms; ++k) { clo[k] = acos_coeff(k, false); chi[k] = acos_coeff(k, true); }
        double r = acos_coeff(kAcosTerms, true);
        for(int i = 0; i < kAcosTerms; ++i) r *= 0.25;
        tail = r * 2.0;
        return true;
    }();
    (void)init;
    if(!(t >= -1.0 && t <= 1.0)) { lo = hi = std::numeric_limits<double>::quiet_NaN(); return; }
    const double a = t < 0.0 ? -t : t;
    const bool small = a <= 0.5;
    const double u = opq(-opq(a - 1.0) * 0.5);   // exact; +0 for a = 1 (a - 1 is -0 rounding down)
    const double zl = small ? rd_mul(a, a) : u, zh = small ? ru_mul(a, a) : u;
    double pl = clo[kAcosTerms - 1], ph = chi[kAcosTerms - 1];
    for(int k = kAcosTerms - 2; k >= 0; --k) {
        pl = rd_fma(zl, pl, clo[k]);
        ph = ru_fma(zh, ph, chi[k]);
    }
    ph = ru_add(ph, tail);
    double ml = a, mh = a;
    if(!small) { ml = 2.0 * rd_sqrt(u); mh = 2.0 * ru_sqrt(u); }
    const double rl = rd_mul(ml, pl), rh = ru_mul(mh, ph);
    if(small) {
        lo = t >= 0.0 ? rd_sub(kHalfPiLo, rh) : rd_add(kHalfPiLo, rl);
        hi = t >= 0.0 ? ru_sub(kHalfPiHi, rl) : ru_add(kHalfPiHi, rh);
    } else {
        lo = t >= 0.0 ? rl : rd_sub(kPiLo, rh);
        hi = t >= 0.0 ? rh : ru_sub(kPiHi, rl);
    }
}
IVT<float> interval_acos(IVT<float> x, CT B) {
    constexpr std::int64_t kOne = 1000000;
    bool und = x.und;
    double l = x.lb, u = x.ub;
    if(B.lb < -kOne && l < -1.0) { l = -1.0; und = true; }
    if(B.ub > kOne && u > 1.0) { u = 1.0; und = true; }
    double lo, hi, dummy;
    acos_enclose(u, lo, dummy);
    acos_enclose(l, dummy, hi);
    return IVT<float>{f32_down(lo), f32_up(hi), und};
}

template<class X> inline X dmax(X n1, X n2) { return n2 < n1 ? n1 : n2; }
template<class X> inline X dmin(X n1, X n2) { return n1 < n2 ? n1 : n2; }

template<class X> inline IB less(const IVT<X>& a, const IVT<X>& b) {
    return IB{!a.und && !b.und && a.ub < b.lb, a.und || b.und || a.lb < b.ub};
}
inline IB inot(IB x) { return IB{!x.poss, !x.def}; }

// ---------------------------------------------------------------- program
enum Op { C_CONST, C_ARG, C_NEG, C_ADD, C_SUB, C_MUL, C_DIV, C_SQRT, C_SQUARE, C_MAX, C_MIN, C_ITE, C_ACOS,
          C_BCONST, C_LT, C_LE, C_GT, C_GE, C_EQ, C_NEQ, C_NOT, C_OR, C_AND };
const char* const kOpNames[] = {"const", "arg", "neg", "add", "sub", "mul", "div", "sqrt", "square", "max", "min",
                                "ite", "acos", "bconst", "lt", "le", "gt", "ge", "eq", "neq", "not", "or", "and"};

struct PNode {
    int op;
    std::vector<int> kids;
    int arg;
    double c_lb, c_ub;
    bool c_undef, b_def, b_poss;
    CT ct;
};
struct PBound { int target, dir, expr; bool honor; };
struct PSplit { int arg; std::uint32_t initial, subs; bool dynamic; };

struct Program {
    int N = 0;
    std::vector<PSplit> splits;
    std::vector<PNode> nodes;
    std::vector<PBound> bounds;
    std::vector<std::vector<std::vector<int>>> update;
    std::vector<std::vector<int>> rows;
    std::vector<double> box_lb, box_ub;
    std::vector<std::vector<int>> bound_order;  // nodes needed per bound, topological
    std::vector<std::vector<int>> pred_order;   // nodes needed per node (for constraint roots)
};

double from_bits(std::uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }

void collect(const Program& p, int root, std::vector<int>& out, std::vector<char>& seen) {
    if(seen[static_cast<std::size_t>(root)]) return;
    seen[static_cast<std::size_t>(root)] = 1;
    for(int k : p.nodes[static_cast<std::size_t>(root)].kids) collect(p, k, out, seen);
    out.push_back(root);
}

Program parse(const std::string& text) {
    Program p;
    std::istringstream in(text);
    std::string line;
    if(!std::getline(in, line) || line != "sqd-program 1") throw std::runtime_error("bad program header");
    bool ended = false;
    while(std::getline(in, line)) {
        std::istringstream s(line);
        std::string key;
        s >> key;
        if(key == "end") { ended = true; break; }
        if(key == "nargs") {
            s >> p.N;
            p.update.assign(static_cast<std::size_t>(p.N), std::vector<std::vector<int>>(3));
            p.rows.assign(static_cast<std::size_t>(p.N), {});
            p.box_lb.assign(static_cast<std::size_t>(p.N), 0);
            p.box_ub.assign(static_cast<std::size_t>(p.N), 0);
        } else if(key == "split") {
            PSplit sp; int d; s >> sp.arg >> sp.initial >> sp.subs >> d; sp.dynamic = d != 0; p.splits.push_back(sp);
        } else if(key == "node") {
            PNode n; std::size_t id, nk; std::string op; std::uint64_t cl, cu; int cund, bd, bp, pl, pu;
            s >> id >> op >> n.arg >> nk;
            n.op = -1;
            for(int i = 0; i < 23; ++i) if(op == kOpNames[i]) n.op = i;
            if(n.op < 0) throw std::runtime_error("unknown op " + op);
            n.kids.resize(nk);
            for(auto& k : n.kids) s >> k;
            s >> cl >> cu >> cund >> bd >> bp >> n.ct.lb >> n.ct.ub >> pl >> pu;
            n.c_lb = from_bits(cl); n.c_ub = from_bits(cu);
            n.c_undef = cund != 0; n.b_def = bd != 0; n.b_poss = bp != 0;
            if(id != p.nodes.size()) throw std::runtime_error("node order");
            p.nodes.push_back(n);
        } else if(key == "bound") {
            PBound b; std::size_t k; int hon; s >> k >> b.target >> b.dir >> b.expr >> hon; b.honor = hon != 0;
            p.bounds.push_back(b);
        } else if(key == "update") {
            int a, w; std::size_t n; s >> a >> w >> n;
            auto& l = p.update.at(static_cast<std::size_t>(a)).at(static_cast<std::size_t>(w));
            l.resize(n); for(auto& x : l) s >> x;
        } else if(key == "row") {
            int a; std::size_t n; s >> a >> n;
            auto& r = p.rows.at(static_cast<std::size_t>(a));
            r.resize(n); for(auto& x : r) s >> x;
        } else if(key == "box") {
            int a; std::uint64_t l, u; s >> a >> l >> u;
            p.box_lb.at(static_cast<std::size_t>(a)) = from_bits(l);
            p.box_ub.at(static_cast<std::size_t>(a)) = from_bits(u);
        }
        if(s.fail()) throw std::runtime_error("malformed line: " + line);
    }
    if(!ended) throw std::runtime_error("program text is truncated");
    for(const auto& b : p.bounds) {
        std::vector<int> o; std::vector<char> seen(p.nodes.size(), 0);
        collect(p, b.expr, o, seen);
        p.bound_order.push_back(o);
    }
    p.pred_order.resize(p.nodes.size());
    for(const auto& r : p.rows) for(int c : r) {
        if(!p.pred_order[static_cast<std::size_t>(c)].empty()) continue;
        std::vector<int> o; std::vector<char> seen(p.nodes.size(), 0);
        collect(p, c, o, seen);
        p.pred_order[static_cast<std::size_t>(c)] = o;
    }
    return p;
}

// ---------------------------------------------------------------- evaluation
template<class X> struct CuboidT {
    std::vector<X> lb, ub;
    std::uint64_t depth = 0;
};
using Cuboid = CuboidT<double>;

// Interval evaluation of the program's nodes on one box, in binary64 or (the
// FP32 prefilter) in binary32 with every constant converted outward.
template<class X> struct EvaluatorT {
    using IV = IVT<X>;
    using Cuboid = CuboidT<X>;
    const Program& p;
    std::vector<IV> v;
    std::vector<IB> pb;
    std::vector<X> c_lb, c_ub;
    explicit EvaluatorT(const Program& prog) : p(prog), v(prog.nodes.size()), pb(prog.nodes.size()) {
        for(const PNode& n : prog.nodes) {
            if(sizeof(X) == 4) { c_lb.push_back(static_cast<X>(f32_down(n.c_lb))); c_ub.push_back(static_cast<X>(f32_up(n.c_ub))); }
            else { c_lb.push_back(static_cast<X>(n.c_lb)); c_ub.push_back(static_cast<X>(n.c_ub)); }
        }
    }

    void eval_node(int id, const Cuboid& c) {
        const PNode& n = p.nodes[static_cast<std::size_t>(id)];
        auto K = [&](int i) -> const IV& { return v[static_cast<std::size_t>(n.kids[static_cast<std::size_t>(i)])]; };
        auto KP = [&](int i) -> IB { return pb[static_cast<std::size_t>(n.kids[static_cast<std::size_t>(i)])]; };
        auto KCT = [&](int i) -> CT { return p.nodes[static_cast<std::size_t>(n.kids[static_cast<std::size_t>(i)])].ct; };
        IV& r = v[static_cast<std::size_t>(id)];
        switch(n.op) {
            case C_CONST: r = IV{c_lb[static_cast<std::size_t>(id)], c_ub[static_cast<std::size_t>(id)], n.c_undef}; return;
            case C_ARG: r = IV{c.lb[static_cast<std::size_t>(n.arg)], c.ub[static_cast<std::size_t>(n.arg)], false}; return;
            case C_NEG: r = IV{-K(0).ub, -K(0).lb, K(0).und}; return;
            case C_ADD: r = IV{rd_add(K(0).lb, K(1).lb), ru_add(K(0).ub, K(1).ub), K(0).und || K(1).und}; return;
            case C_SUB: r = IV{rd_sub(K(0).lb, K(1).ub), ru_sub(K(0).ub, K(1).lb), K(0).und || K(1).und}; return;
            case C_MUL: r = mul_check_signs(K(0), K(1), KCT(0), KCT(1)); return;
            case C_DIV: r = interval_div(K(0), K(1), KCT(0), KCT(1)); return;
            case C_SQRT: r = interval_sqrt(K(0), KCT(0)); return;
            case C_ACOS: r = interval_acos(K(0), KCT(0)); return;
            case C_SQUARE: {
                IV x = mul_check_signs(K(0), K(0), kUnbounded, kUnbounded);
                if(x.lb < 0) x.lb = 0;
                r = x;
                return;
            }
            case C_MAX: case C_MIN: {
                std::size_t m = n.kids.size();
                IV acc = v[static_cast<std::size_t>(n.kids[m - 1])];
                for(std::size_t i = m - 1; i-- > 0;) {
                    const IV& k = v[static_cast<std::size_t>(n.kids[i])];
                    if(n.op == C_MAX) acc = IV{dmax(k.lb, acc.lb), dmax(k.ub, acc.ub), k.und || acc.und};
                    else acc = IV{dmin(k.lb, acc.lb), dmin(k.ub, acc.ub), k.und || acc.und};
                }
                r = acc;
                return;
            }
            case C_ITE: {
                IB cnd = KP(0);
                if(cnd.def) { r = K(1); return; }
                if(!cnd.poss) { r = K(2); return; }
                IV t = K(1);
                const IV& e = K(2);
                if(e.lb < t.lb) t.lb = e.lb;
                if(e.ub > t.ub) t.ub = e.ub;
                t.und = t.und || e.und;
                r = t;
                return;
            }
            case C_BCONST: pb[static_cast<std::size_t>(id)] = IB{n.b_def, n.b_poss}; return;
            case C_LT: pb[static_cast<std::size_t>(id)] = less(K(0), K(1)); return;
            case C_GT: pb[static_cast<std::size_t>(id)] = less(K(1), K(0)); return;
            case C_GE: pb[static_cast<std::size_t>(id)] = inot(less(K(0), K(1))); return;
            case C_LE: pb[static_cast<std::size_t>(id)] = inot(less(K(1), K(0))); return;
            case C_EQ: case C_NEQ: {
                const IV& a = K(0); const IV& b = K(1);
                IB e{!a.und && !b.und && a.lb == a.ub && b.lb == b.ub && a.lb == b.lb,
                     a.und || b.und || (!(a.ub < b.lb) && !(b.ub < a.lb))};
                pb[static_cast<std::size_t>(id)] = n.op == C_EQ ? e : inot(e);
                return;
            }
            case C_NOT: pb[static_cast<std::size_t>(id)] = inot(KP(0)); return;
            case C_OR: case C_AND: {
                IB acc = n.op == C_OR ? IB{false, false} : IB{true, true};
                for(int k : n.kids) {
                    IB x = pb[static_cast<std::size_t>(k)];
                    if(n.op == C_OR) { acc.def = acc.def || x.def; acc.poss = acc.poss || x.poss; }
                    else { acc.def = acc.def && x.def; acc.poss = acc.poss && x.poss; }
                }
                pb[static_cast<std::size_t>(id)] = acc;
                return;
            }
            default: throw std::runtime_error("unsupported node");
        }
    }
    // Demand-driven evaluation, memoized per evaluation (one bound or one
    // constraint on one box).  if_then_else evaluates only the branches its
    // condition allows (math_fn/eval/if_then_else.hpp:64-82); || and && stop at a
    // definitely true / not possibly true operand (predicate_eval/n_ary.hpp:63-135).
    // The values equal those of evaluating every node; the operation count is
    // that of the lazy evaluation.
    std::vector<std::uint32_t> stamp = std::vector<std::uint32_t>(p.nodes.size(), 0);
    std::uint32_t epoch = 0;
    void begin() {
        if(++epoch == 0) { std::fill(stamp.begin(), stamp.end(), 0u); epoch = 1; }
    }
    void need(int id, const Cuboid& c) {
        const std::size_t i = static_cast<std::size_t>(id);
        if(stamp[i] == epoch) return;
        const PNode& n = p.nodes[i];
        auto kid = [&](std::size_t j) { return n.kids[j]; };
        switch(n.op) {
            case C_ITE: {
                need(kid(0), c);
                const IB cnd = pb[static_cast<std::size_t>(kid(0))];
                if(cnd.poss) need(kid(1), c);
                if(!cnd.def) need(kid(2), c);
                eval_node(id, c);   // reads only the branches the condition selects
                break;
            }
            case C_OR: case C_AND: {
                const bool is_or = n.op == C_OR;
                need(kid(0), c);
                IB acc = pb[static_cast<std::size_t>(kid(0))];
                for(std::size_t j = 1; j < n.kids.size(); ++j) {
                    if(is_or ? acc.def : !acc.poss) break;
                    need(kid(j), c);
                    const IB x = pb[static_cast<std::size_t>(kid(j))];
                    if(is_or) { acc.def = acc.def || x.def; acc.poss = acc.poss || x.poss; }
                    else { acc.def = acc.def && x.def; acc.poss = acc.poss && x.poss; }
                }
                pb[i] = acc;
                break;
            }
            default:
                for(int k : n.kids) need(k, c);
                eval_node(id, c);
                break;
        }
        stamp[i] = epoch;
    }
    const IV& eval_bound(int k, const Cuboid& c) {
        begin();
        const int e = p.bounds[static_cast<std::size_t>(k)].expr;
        need(e, c);
        return v[static_cast<std::size_t>(e)];
    }
    bool possibly(int pred, const Cuboid& c) {
        begin();
        need(pred, c);
        return pb[static_cast<std::size_t>(pred)].poss;
    }
};

using Evaluator = EvaluatorT<double>;
using Evaluator32 = EvaluatorT<float>;

enum BoundID { B_NONE = 0, B_LB = 1, B_UB = 2, B_BOTH = 3, B_EMPTY = 4 };

template<class X>
int apply_bound(EvaluatorT<X>& ev, int k, CuboidT<X>& c) {
    const PBound& b = ev.p.bounds[static_cast<std::size_t>(k)];
    IVT<X> iv = ev.eval_bound(k, c);
    if(b.honor && iv.und) return B_NONE;
    X& tl = c.lb[static_cast<std::size_t>(b.target)];
    X& tu = c.ub[static_cast<std::size_t>(b.target)];
    if(b.dir == 1) {
        if(iv.ub < tu) { tu = iv.ub; return iv.ub < tl ? B_EMPTY : B_UB; }
    } else if(b.dir == 2) {
        if(iv.lb > tl) { tl = iv.lb; return iv.lb > tu ? B_EMPTY : B_LB; }
    } else {
        int lbc = B_NONE, ubc = B_NONE;
        if(iv.ub < tu) { tu = iv.ub; ubc = B_UB; }
        if(iv.lb > tl) { tl = iv.lb; lbc = B_LB; }
        if(tu < tl) return B_EMPTY;
        return lbc | ubc;
    }
    return B_NONE;
}

// Fixed-capacity ring queue with membership set; empty iff begin == end, so a
// queue holding every bound index at once reads as empty (index_queue_set.hpp:62-91).
struct IndexQueue {
    std::size_t cap;
    std::size_t begin = 0, end = 0;
    std::vector<int> elems;
    std::vector<char> present;
    explicit IndexQueue(std::size_t n) : cap(n), elems(n), present(n, 0) {}
    bool empty() const { return cap == 0 || begin == end; }
    bool enqueue(int e) {
        if(cap == 0) return false;
        if(present[static_cast<std::size_t>(e)]) return false;
        present[static_cast<std::size_t>(e)] = 1;
        elems[end] = e;
        if(++end >= cap) end = 0;
        return true;
    }
    int dequeue() {
        int e = elems[begin];
        present[static_cast<std::size_t>(e)] = 0;
        if(++begin >= cap) begin = 0;
        return e;
    }
};

template<class X>
bool propagate(EvaluatorT<X>& ev, CuboidT<X>& c, int arg, std::size_t it_limit, unsigned rec_limit) {
    const Program& p = ev.p;
    const std::size_t nb = p.bounds.size();
    IndexQueue q(nb);
    std::vector<unsigned> rec(nb, 0);
    std::vector<char> unstable(nb, 0);
    std::vector<int> unstable_list;
    for(int k : p.update[static_cast<std::size_t>(arg)][2]) q.enqueue(k);
    auto handle = [&](const std::vector<int>& lst) {
        for(int a : lst) {
            if(rec[static_cast<std::size_t>(a)] >= rec_limit) {
                if(!q.present[static_cast<std::size_t>(a)] && !unstable[static_cast<std::size_t>(a)]) {
                    unstable_list.push_back(a);
                    unstable[static_cast<std::size_t>(a)] = 1;
                }
            } else if(q.enqueue(a)) {
                rec[static_cast<std::size_t>(a)] += 1;
            }
        }
    };
    for(std::size_t i = 0; i < it_limit; ++i) {
        while(!q.empty()) {
            int k = q.dequeue();
            int r = apply_bound(ev, k, c);
            int t = p.bounds[static_cast<std::size_t>(k)].target;
            switch(r) {
                case B_EMPTY: return true;
                case B_LB: handle(p.update[static_cast<std::size_t>(t)][0]); break;
                case B_UB: handle(p.update[static_cast<std::size_t>(t)][1]); break;
                case B_BOTH: handle(p.update[static_cast<std::size_t>(t)][2]); break;
                default: break;
            }
        }
        if(!unstable_list.empty()) {
            for(int a : unstable_list) q.enqueue(a);
            std::fill(rec.begin(), rec.end(), 0u);
            std::fill(unstable.begin(), unstable.end(), 0);
            unstable_list.clear();
        } else {
            break;
        }
    }
    return false;
}

struct Splitter {
    double lb, ub, w;
    int n;
    Splitter(double l, double u, int cnt) : lb(l), ub(u), n(cnt) { w = rd_div(rd_sub(u, l), static_cast<double>(cnt)); }
    double point(int i) const { return i >= n ? ub : rd_add(lb, rd_mul(static_cast<double>(i), w)); }
};

struct Counts {
    std::uint64_t cuboids = 0, leaves = 0, criticals = 0, repeated = 0;
    std::uint64_t f32_checked = 0, f32_decided = 0;   // FP32 prefilter (static-level children)
    Counts& operator+=(const Counts& o) {
        cuboids += o.cuboids; leaves += o.leaves; criticals += o.criticals; repeated += o.repeated;
        f32_checked += o.f32_checked; f32_decided += o.f32_decided;
        return *this;
    }
};

struct Core {
    const Program& p;
    std::vector<int> static_splits;   // indices into p.splits
    std::vector<int> dynamic_splits;
    explicit Core(const Program& prog) : p(prog) {
        for(std::size_t i = 0; i < prog.splits.size(); ++i) {
            if(prog.splits[i].dynamic) dynamic_splits.push_back(static_cast<int>(i));
            else static_splits.push_back(static_cast<int>(i));
        }
    }
    bool satisfies_all(Evaluator& ev, const Cuboid& c) const {
        for(int cn : p.rows[static_cast<std::size_t>(p.N - 1)]) if(!ev.possibly(cn, c)) return false;
        return true;
    }
    template<class X>
    bool violates_rows(EvaluatorT<X>& ev, const CuboidT<X>& c, int from, int to) const {
        for(int r = from; r < to; ++r)
            for(int cn : p.rows[static_cast<std::size_t>(r)]) if(!ev.possibly(cn, c)) return true;
        return false;
    }
    // FP32 prefilter of one static-level child q (the parent box with the split
    // variable `arg` replaced by the child's subrange), in binary32 on the
    // outward-converted box: the child is pruned when propagation empties it, or
    // when a constraint of its rows fails on the hybrid box made of the parent's
    // other variables and the propagated child's split variable.  The hybrid box
    // contains the propagated child (propagation only narrows), so both tests
    // are rigorous; the hybrid box makes every term free of `arg` equal for all
    // children of a parent.
    bool f32_prunes(Evaluator32& ev32, const Cuboid& q, int arg, int row_end) const {
        CuboidT<float> f, h;
        for(std::size_t a = 0; a < q.lb.size(); ++a) { f.lb.push_back(f32_down(q.lb[a])); f.ub.push_back(f32_up(q.ub[a])); }
        h = f;
        if(propagate(ev32, f, arg, static_cast<std::size_t>(p.N), 4)) return true;
        h.lb[static_cast<std::size_t>(arg)] = f.lb[static_cast<std::size_t>(arg)];
        h.ub[static_cast<std::size_t>(arg)] = f.ub[static_cast<std::size_t>(arg)];
        return violates_rows(ev32, h, arg, row_end);
    }
    // Static enumeration (cpu_prover_core.hpp:221-256); `leaf` is called for each
    // surviving leaf.  With ev32 set, each child is first tried in binary32
    // (a child pruned there is a leaf; the others are decided in binary64).
    template<typename Leaf>
    void enumerate(Evaluator& ev, const Cuboid& e, std::size_t level, int mode, Counts& cnt, Leaf&& leaf,
                   Evaluator32* ev32 = nullptr) const {
        if(level == static_splits.size()) { cnt.leaves++; leaf(e); return; }
        const PSplit& s = p.splits[static_cast<std::size_t>(static_splits[level])];
        const int arg = s.arg;
        const int subs = static_cast<int>(s.subs);
        const int row_end = level + 1 < static_splits.size()
                            ? p.splits[static_cast<std::size_t>(static_splits[level + 1])].arg : p.N;
        enter_phase(mode, level, static_cast<std::uint64_t>(subs));
        Splitter sp(e.lb[static_cast<std::size_t>(arg)], e.ub[static_cast<std::size_t>(arg)], subs);
        for(int i = 0; i < subs; ++i) {
            enter_phase(mode, level, 0);
            Cuboid q = e;
            q.lb[static_cast<std::size_t>(arg)] = sp.point(i);
            q.ub[static_cast<std::size_t>(arg)] = sp.point(i + 1);
            if(ev32) {
                cnt.f32_checked++;
                if(f32_prunes(*ev32, q, arg, row_end)) { cnt.f32_decided++; cnt.leaves++; continue; }
            }
            if(propagate(ev, q, arg, static_cast<std::size_t>(p.N), 4)) { cnt.leaves++; continue; }
            if(violates_rows(ev, q, arg, row_end)) { cnt.leaves++; continue; }
            enumerate(ev, q, level + 1, mode, cnt, leaf, ev32);
        }
        cnt.cuboids += static_cast<std::uint64_t>(subs);
    }
    // Non-final pass with max_iterations_per_node = 1: the hull of surviving leaves.
    bool nonfinal(Evaluator& ev, const Cuboid& e, Cuboid& hull, Counts& cnt, Evaluator32* ev32 = nullptr) const {
        bool empty = true;
        enumerate(ev, e, 0, 0, cnt, [&](const Cuboid& leaf) {
            if(empty) { hull = leaf; empty = false; return; }
            for(int a = 0; a < p.N; ++a) {
                std::size_t i = static_cast<std::size_t>(a);
                if(leaf.lb[i] < hull.lb[i]) hull.lb[i] = leaf.lb[i];
                if(leaf.ub[i] > hull.ub[i]) hull.ub[i] = leaf.ub[i];
            }
        }, ev32);
        if(!empty) hull.depth = e.depth;
        return !empty;
    }
    template<typename Crit>
    void final_pass(Evaluator& ev, const Cuboid& e, Counts& cnt, Crit&& crit, Evaluator32* ev32 = nullptr) const {
        enumerate(ev, e, 0, 1, cnt, [&](const Cuboid& leaf) { cnt.criticals++; crit(leaf); }, ev32);
    }
    // Dynamic split (pd_impl_split_into.hpp:29-75) or initial split (proof_driver.hpp:320-391).
    // initial subdivisions of the first dynamic variable the initial split enumerates
    std::uint64_t init_lo = 0, init_hi = ~0ull;
    template<typename Out>
    void split(Evaluator& ev, const Cuboid& e, std::size_t level, bool initial, Counts& cnt, Out&& out) const {
        const std::size_t itlim = 2 * static_cast<std::size_t>(p.N);
        const unsigned reclim = initial ? static_cast<unsigned>(std::min(p.N, 255)) : static_cast<unsigned>(std::min(p.N, 32));
        if(level == dynamic_splits.size()) {
            if(satisfies_all(ev, e)) out(e); else cnt.leaves++;
            return;
        }
        const PSplit& s = p.splits[static_cast<std::size_t>(dynamic_splits[level])];
        const int n = static_cast<int>(initial ? s.initial : s.subs);
        const int mode = initial ? 3 : 2;
        enter_phase(mode, level, static_cast<std::uint64_t>(n));
        Splitter sp(e.lb[static_cast<std::size_t>(s.arg)], e.ub[static_cast<std::size_t>(s.arg)], n);
        const bool ranged = initial && level == 0;
        const int i0 = ranged ? static_cast<int>(std::min<std::uint64_t>(init_lo, static_cast<std::uint64_t>(n))) : 0;
        const int i1 = ranged ? static_cast<int>(std::min<std::uint64_t>(init_hi, static_cast<std::uint64_t>(n))) : n;
        for(int i = i0; i < i1; ++i) {
            enter_phase(mode, level, 0);
            cnt.cuboids++;
            Cuboid q = e;
            q.lb[static_cast<std::size_t>(s.arg)] = sp.point(i);
            q.ub[static_cast<std::size_t>(s.arg)] = sp.point(i + 1);
            if(propagate(ev, q, s.arg, itlim, reclim)) { cnt.leaves++; continue; }
            split(ev, q, level + 1, initial, cnt, out);
        }
    }
};

Cuboid initial_box(const Program& p) {
    Cuboid c;
    c.lb = p.box_lb;
    c.ub = p.box_ub;
    Evaluator ev(p);
    for(int it = 0; it < 50 * p.N; ++it) {
        for(std::size_t k = 0; k < p.bounds.size(); ++k)
            if(apply_bound(ev, static_cast<int>(k), c) == B_EMPTY) break;
    }
    return c;
}

std::vector<Cuboid> initial_queue(const Program& p, Counts& cnt, std::uint64_t lo = 0, std::uint64_t hi = ~0ull) {
    Core core(p);
    core.init_lo = lo;
    core.init_hi = hi;
    Evaluator ev(p);
    std::vector<Cuboid> q;
    Cuboid root = initial_box(p);
    core.split(ev, root, 0, true, cnt, [&](const Cuboid& c) { q.push_back(c); });
    return q;
}

// LIFO frontier shared by worker threads; counts do not depend on the schedule.
struct Frontier {
    std::mutex m;
    std::condition_variable cv;
    std::vector<Cuboid> stack;
    int busy = 0;
    bool pop(std::vector<Cuboid>& out, std::size_t batch) {
        std::unique_lock<std::mutex> l(m);
        cv.wait(l, [&] { return !stack.empty() || busy == 0; });
        if(stack.empty()) return false;
        for(std::size_t i = 0; i < batch && !stack.empty(); ++i) { out.push_back(std::move(stack.back())); stack.pop_back(); }
        ++busy;
        return true;
    }
    void push_done(std::vector<Cuboid>& items) {
        std::unique_lock<std::mutex> l(m);
        for(auto& c : items) stack.push_back(std::move(c));
        items.clear();
        --busy;
        cv.notify_all();
    }
};

struct RunResult {
    Counts counts;
    std::vector<Cuboid> criticals;
};

void run_from_queue(const Program& p, std::vector<Cuboid> start, int generations, int threads, RunResult& res,
                    bool fp32 = false) {
    Core core(p);
    Frontier f;
    f.stack = std::move(start);
    std::mutex res_m;
    std::vector<std::thread> ts;
    std::vector<Counts> counts(static_cast<std::size_t>(threads));
    for(int t = 0; t < threads; ++t) {
        ts.emplace_back([&, t] {
            RoundDownGuard g;
            Evaluator ev(p);
            Evaluator32 ev32(p);
            Evaluator32* e32 = fp32 ? &ev32 : nullptr;
            Counts& cnt = counts[static_cast<std::size_t>(t)];
            std::vector<Cuboid> batch, produced;
            while(f.pop(batch, 32)) {
                for(const Cuboid& e : batch) {
                    Cuboid hull;
                    if(!core.nonfinal(ev, e, hull, cnt, e32)) continue;
                    if(hull.depth >= static_cast<std::uint64_t>(generations)) {
                        std::vector<Cuboid> crit;
                        core.final_pass(ev, hull, cnt, [&](const Cuboid& c) { crit.push_back(c); }, e32);
                        if(!crit.empty()) {
                            std::lock_guard<std::mutex> l(res_m);
                            for(auto& c : crit) res.criticals.push_back(c);
                        }
                    } else {
                        hull.depth += 1;
                        core.split(ev, hull, 0, false, cnt, [&](const Cuboid& c) { produced.push_back(c); });
                    }
                }
                batch.clear();
                f.push_done(produced);
            }
            flush_thread_counts();
        });
    }
    for(auto& t : ts) t.join();
    for(const auto& c : counts) res.counts += c;
}

thread_local std::string t_error;

struct Handle {
    Program prog;
    std::uint64_t init_lo = 0, init_hi = ~0ull;   // initial range of the first dynamic variable
    bool fp32 = false;       // FP32 prefilter in the static levels
    Counts last;             // counts of the last computing call
};

template<typename F> int guarded(F&& f) {
    try { f(); flush_thread_counts(); return 0; }
    catch(const std::exception& e) { flush_thread_counts(); t_error = e.what(); return -1; }
}

void to_out(const Counts& c, OracleCounts* o) {
    if(!o) return;
    o->num_cuboids = c.cuboids;
    o->num_leaf_cuboids = c.leaves;
    o->num_critical_cuboids = c.criticals;
    o->num_repeated_nodes = c.repeated;
}

Cuboid soa_get(const Program& p, std::uint64_t n, const double* lb, const double* ub, std::uint64_t i) {
    Cuboid c;
    c.lb.resize(static_cast<std::size_t>(p.N));
    c.ub.resize(static_cast<std::size_t>(p.N));
    for(int a = 0; a < p.N; ++a) {
        c.lb[static_cast<std::size_t>(a)] = lb[static_cast<std::uint64_t>(a) * n + i];
        c.ub[static_cast<std::size_t>(a)] = ub[static_cast<std::uint64_t>(a) * n + i];
    }
    return c;
}

} // namespace

namespace {
template<class X>
void eval_program_boxes(const Program& p, std::uint64_t n, const double* lb, const double* ub, double* bv,
                        std::uint8_t* cv) {
    std::set<int> roots;
    for(const auto& r : p.rows) for(int c : r) roots.insert(c);
    const std::vector<int> cons(roots.begin(), roots.end());
    EvaluatorT<X> ev(p);
    CuboidT<X> c;
    c.lb.resize(static_cast<std::size_t>(p.N));
    c.ub.resize(static_cast<std::size_t>(p.N));
    const std::size_t nb = p.bounds.size(), nc = cons.size();
    for(std::uint64_t i = 0; i < n; ++i) {
        for(int a = 0; a < p.N; ++a) {
            const double l = lb[static_cast<std::uint64_t>(a) * n + i], u = ub[static_cast<std::uint64_t>(a) * n + i];
            if(sizeof(X) == 4) { c.lb[static_cast<std::size_t>(a)] = static_cast<X>(f32_down(l)); c.ub[static_cast<std::size_t>(a)] = static_cast<X>(f32_up(u)); }
            else { c.lb[static_cast<std::size_t>(a)] = static_cast<X>(l); c.ub[static_cast<std::size_t>(a)] = static_cast<X>(u); }
        }
        for(std::size_t k = 0; k < nb; ++k) {
            const IVT<X> v = ev.eval_bound(static_cast<int>(k), c);
            double* o = bv + (i * nb + k) * 3;
            o[0] = static_cast<double>(v.lb);
            o[1] = static_cast<double>(v.ub);
            o[2] = v.und ? 1.0 : 0.0;
        }
        for(std::size_t j = 0; j < nc; ++j) {
            ev.begin();
            ev.need(cons[j], c);
            const IB b = ev.pb[static_cast<std::size_t>(cons[j])];
            cv[(i * nc + j) * 2] = b.def ? 1 : 0;
            cv[(i * nc + j) * 2 + 1] = b.poss ? 1 : 0;
        }
    }
}
} // namespace

extern "C" {

int oracle_load(const char* text, void** handle) {
    return guarded([&] { auto h = std::make_unique<Handle>(); h->prog = parse(text); *handle = h.release(); });
}
void oracle_free(void* handle) { delete static_cast<Handle*>(handle); }
int oracle_nargs(void* handle) { return static_cast<Handle*>(handle)->prog.N; }
const char* oracle_last_error(void) { return t_error.c_str(); }

int oracle_initial_box(void* handle, double* lb, double* ub) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Cuboid c = initial_box(p);
        for(int a = 0; a < p.N; ++a) { lb[a] = c.lb[static_cast<std::size_t>(a)]; ub[a] = c.ub[static_cast<std::size_t>(a)]; }
    });
}

int oracle_initial_queue(void* handle, double* lb, double* ub, std::uint64_t cap, std::uint64_t* n, OracleCounts* counts) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Counts cnt;
        std::vector<Cuboid> q = initial_queue(p, cnt);
        *n = q.size();
        for(std::uint64_t i = 0; i < q.size() && i < cap; ++i)
            for(int a = 0; a < p.N; ++a) {
                lb[i * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)] = q[i].lb[static_cast<std::size_t>(a)];
                ub[i * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)] = q[i].ub[static_cast<std::size_t>(a)];
            }
        to_out(cnt, counts);
    });
}

int oracle_run(void* handle, int generations, int threads, OracleCounts* counts,
               double* criticals, std::uint64_t max_criticals, std::uint64_t* n_criticals) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Counts init;
        Handle* hr = static_cast<Handle*>(handle);
        std::vector<Cuboid> q = initial_queue(p, init, hr->init_lo, hr->init_hi);
        // the reference pops the last element first; reverse-free order does not change counts
        RunResult r;
        Handle* h = static_cast<Handle*>(handle);
        run_from_queue(p, std::move(q), generations, threads < 1 ? 1 : threads, r, h->fp32);
        r.counts += init;
        h->last = r.counts;
        to_out(r.counts, counts);
        if(n_criticals) *n_criticals = r.criticals.size();
        for(std::uint64_t i = 0; criticals && i < r.criticals.size() && i < max_criticals; ++i)
            for(int a = 0; a < p.N; ++a) {
                criticals[(i * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)) * 2] = r.criticals[i].lb[static_cast<std::size_t>(a)];
                criticals[(i * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)) * 2 + 1] = r.criticals[i].ub[static_cast<std::size_t>(a)];
            }
    });
}

int oracle_run_slice(void* handle, int generations, int threads, std::uint64_t first, std::uint64_t last,
                     OracleCounts* counts, std::uint64_t* initial_queue_len) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Counts init;
        std::vector<Cuboid> q = initial_queue(p, init);
        if(initial_queue_len) *initial_queue_len = q.size();
        if(last > q.size()) last = q.size();
        if(first > last) first = last;
        std::vector<Cuboid> part(q.begin() + static_cast<std::ptrdiff_t>(first), q.begin() + static_cast<std::ptrdiff_t>(last));
        RunResult r;
        Handle* h = static_cast<Handle*>(handle);
        run_from_queue(p, std::move(part), generations, threads < 1 ? 1 : threads, r, h->fp32);
        h->last = r.counts;
        to_out(r.counts, counts);
    });
}

int oracle_nonfinal(void* handle, std::uint64_t n, const double* lb, const double* ub,
                    double* hull_lb, double* hull_ub, std::uint8_t* has_hull, OracleCounts* counts) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Core core(p);
        Evaluator ev(p);
        Evaluator32 ev32(p);
        Handle* hd = static_cast<Handle*>(handle);
        Evaluator32* e32 = hd->fp32 ? &ev32 : nullptr;
        Counts cnt;
        for(std::uint64_t i = 0; i < n; ++i) {
            Cuboid e = soa_get(p, n, lb, ub, i), h;
            bool ok = core.nonfinal(ev, e, h, cnt, e32);
            has_hull[i] = ok ? 1 : 0;
            for(int a = 0; a < p.N; ++a) {
                hull_lb[static_cast<std::uint64_t>(a) * n + i] = ok ? h.lb[static_cast<std::size_t>(a)] : 0.0;
                hull_ub[static_cast<std::uint64_t>(a) * n + i] = ok ? h.ub[static_cast<std::size_t>(a)] : 0.0;
            }
        }
        hd->last = cnt;
        to_out(cnt, counts);
    });
}

int oracle_final(void* handle, std::uint64_t n, const double* lb, const double* ub,
                 double* crit_lb, double* crit_ub, std::uint64_t cap, std::uint64_t* n_crit, OracleCounts* counts) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Core core(p);
        Evaluator ev(p);
        Evaluator32 ev32(p);
        Handle* hd = static_cast<Handle*>(handle);
        Evaluator32* e32 = hd->fp32 ? &ev32 : nullptr;
        Counts cnt;
        std::uint64_t k = 0;
        for(std::uint64_t i = 0; i < n; ++i) {
            Cuboid e = soa_get(p, n, lb, ub, i);
            core.final_pass(ev, e, cnt, [&](const Cuboid& c) {
                if(k < cap)
                    for(int a = 0; a < p.N; ++a) {
                        crit_lb[k * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)] = c.lb[static_cast<std::size_t>(a)];
                        crit_ub[k * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)] = c.ub[static_cast<std::size_t>(a)];
                    }
                ++k;
            }, e32);
        }
        *n_crit = k;
        hd->last = cnt;
        to_out(cnt, counts);
    });
}

int oracle_split(void* handle, std::uint64_t n, const double* lb, const double* ub,
                 double* out_lb, double* out_ub, std::uint64_t cap, std::uint64_t* n_out, OracleCounts* counts) {
    return guarded([&] {
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        Core core(p);
        Evaluator ev(p);
        Counts cnt;
        std::uint64_t k = 0;
        for(std::uint64_t i = 0; i < n; ++i) {
            Cuboid e = soa_get(p, n, lb, ub, i);
            core.split(ev, e, 0, false, cnt, [&](const Cuboid& c) {
                if(k < cap)
                    for(int a = 0; a < p.N; ++a) {
                        out_lb[k * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)] = c.lb[static_cast<std::size_t>(a)];
                        out_ub[k * static_cast<std::uint64_t>(p.N) + static_cast<std::uint64_t>(a)] = c.ub[static_cast<std::size_t>(a)];
                    }
                ++k;
            });
        }
        *n_out = k;
        to_out(cnt, counts);
    });
}

int oracle_rounded_op(int op, std::uint64_t n, const double* a, const double* b, double* lo, double* hi) {
    return guarded([&] {
        RoundDownGuard g;
        for(std::uint64_t i = 0; i < n; ++i) {
            switch(op) {
                case 0: lo[i] = rd_add(a[i], b[i]); hi[i] = ru_add(a[i], b[i]); break;
                case 1: lo[i] = rd_sub(a[i], b[i]); hi[i] = ru_sub(a[i], b[i]); break;
                case 2: lo[i] = rd_mul(a[i], b[i]); hi[i] = ru_mul(a[i], b[i]); break;
                case 3: lo[i] = rd_div(a[i], b[i]); hi[i] = ru_div(a[i], b[i]); break;
                case 4: lo[i] = rd_sqrt(a[i]); hi[i] = ru_sqrt(a[i]); break;
                case 5: case 6: lo[i] = acos_dir(a[i], false); hi[i] = acos_dir(a[i], true); break;
                case 8: case 9: case 10: case 11: case 12: {
                    const float x = static_cast<float>(a[i]), y = static_cast<float>(b[i]);
                    if(static_cast<double>(x) != a[i] || static_cast<double>(y) != b[i])
                        if(!(a[i] != a[i]) && !(b[i] != b[i])) throw std::invalid_argument("binary32 operands must be floats");
                    float l = 0, h = 0;
                    switch(op) {
                        case 8: l = rd_add(x, y); h = ru_add(x, y); break;
                        case 9: l = rd_sub(x, y); h = ru_sub(x, y); break;
                        case 10: l = rd_mul(x, y); h = ru_mul(x, y); break;
                        case 11: l = rd_div(x, y); h = ru_div(x, y); break;
                        default: l = rd_sqrt(x); h = ru_sqrt(x); break;
                    }
                    lo[i] = l;
                    hi[i] = h;
                    break;
                }
                case 13: {
                    const float x = static_cast<float>(a[i]);
                    const IVT<float> r = interval_acos(IVT<float>{x, x, false}, CT{-1000000, 1000000});   // no clamping
                    lo[i] = r.lb;
                    hi[i] = r.ub;
                    break;
                }
                case 14: lo[i] = f32_down(a[i]); hi[i] = f32_up(a[i]); break;
                default: throw std::invalid_argument("unknown op");
            }
        }
    });
}

int oracle_interval_op(int op, std::uint64_t n, const double* a, const double* b, double* out) {
    return guarded([&] {
        if(op < 0 || op > 10) throw std::invalid_argument("unknown interval op");
        RoundDownGuard g;
        for(std::uint64_t i = 0; i < n; ++i) {
            const IV x{a[3 * i], a[3 * i + 1], a[3 * i + 2] != 0.0};
            const IV y{b[3 * i], b[3 * i + 1], b[3 * i + 2] != 0.0};
            IV r{0.0, 0.0, false};
            switch(op) {
                case 0: r = IV{rd_add(x.lb, y.lb), ru_add(x.ub, y.ub), x.und || y.und}; break;
                case 1: r = IV{rd_sub(x.lb, y.ub), ru_sub(x.ub, y.lb), x.und || y.und}; break;
                case 2: r = mul_check_signs(x, y, kUnbounded, kUnbounded); break;
                case 3: r = interval_div(x, y, kUnbounded, kUnbounded); break;
                case 4: r = interval_sqrt(x, kUnbounded); break;
                case 5: r = mul_check_signs(x, x, kUnbounded, kUnbounded); if(r.lb < 0) r.lb = 0; break;
                case 6: r = IV{-x.ub, -x.lb, x.und}; break;
                case 7: r = IV{dmax(x.lb, y.lb), dmax(x.ub, y.ub), x.und || y.und}; break;
                case 8: r = IV{dmin(x.lb, y.lb), dmin(x.ub, y.ub), x.und || y.und}; break;
                case 9: {
                    const Splitter sp(x.lb, x.ub, static_cast<int>(y.ub));
                    r.lb = sp.point(static_cast<int>(y.lb));
                    break;
                }
                case 10: r = interval_acos(x, kUnbounded); break;
                default: break;
            }
            out[3 * i] = r.lb;
            out[3 * i + 1] = r.ub;
            out[3 * i + 2] = r.und ? 1.0 : 0.0;
        }
    });
}

std::uint64_t oracle_flop_count(void) { return g_flops.load(); }
void oracle_flop_reset(void) {
    g_flops = 0;
    for(int i = 0; i < kPhases; ++i) { g_phase_flops[i] = 0; g_phase_units[i] = 0; g_phase_flops32[i] = 0; }
}
int oracle_phase_counts(int mode, int level, std::uint64_t* flops, std::uint64_t* units) {
    if(mode < 0 || mode > 4 || level < 0 || level >= kPhaseLevels || !flops || !units) return -1;
    const int ph = mode == 4 ? kPhaseOther : mode * kPhaseLevels + level;
    *flops = g_phase_flops[ph].load();
    *units = g_phase_units[ph].load();
    return 0;
}
void oracle_flop_enable(int on) { g_count_flops = on != 0; }
int oracle_phase_counts32(int mode, int level, std::uint64_t* flops) {
    if(mode < 0 || mode > 4 || level < 0 || level >= kPhaseLevels || !flops) return -1;
    const int ph = mode == 4 ? kPhaseOther : mode * kPhaseLevels + level;
    *flops = g_phase_flops32[ph].load();
    return 0;
}

int oracle_eval_program(void* handle, int precision, std::uint64_t n, const double* lb, const double* ub,
                        double* bound_values, std::uint8_t* cons_values) {
    return guarded([&] {
        if(precision != 64 && precision != 32) throw std::invalid_argument("precision must be 64 or 32");
        RoundDownGuard g;
        const Program& p = static_cast<Handle*>(handle)->prog;
        if(precision == 32) eval_program_boxes<float>(p, n, lb, ub, bound_values, cons_values);
        else eval_program_boxes<double>(p, n, lb, ub, bound_values, cons_values);
    });
}

int oracle_set_init_range(void* handle, std::uint64_t first, std::uint64_t last) {
    if(!handle || first >= last) return -1;
    static_cast<Handle*>(handle)->init_lo = first;
    static_cast<Handle*>(handle)->init_hi = last;
    return 0;
}

int oracle_set_fp32_prefilter(void* handle, int on) {
    if(!handle) return -1;
    static_cast<Handle*>(handle)->fp32 = on != 0;
    return 0;
}
int oracle_fp32_stats(void* handle, std::uint64_t* checked, std::uint64_t* decided) {
    if(!handle || !checked || !decided) return -1;
    *checked = static_cast<Handle*>(handle)->last.f32_checked;
    *decided = static_cast<Handle*>(handle)->last.f32_decided;
    return 0;
}

} // extern "C"
