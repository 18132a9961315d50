// Canonical text form of a lowered program.  Doubles are written as their IEEE
// bit patterns so that the form round-trips exactly.
#include "program.hpp"
#include <sstream>
#include <stdexcept>
#include <cstring>

namespace ivarp {
namespace flat {

static const char* const kNames[] = {
    "const", "arg", "neg", "add", "sub", "mul", "div", "sqrt", "square", "max", "min", "ite", "acos",
    "bconst", "lt", "le", "gt", "ge", "eq", "neq", "not", "or", "and"
};

const char* fop_name(FOp o) noexcept { return kNames[static_cast<unsigned>(o)]; }
bool fop_from_name(const std::string& s, FOp& out) noexcept {
    for(unsigned i = 0; i < sizeof(kNames) / sizeof(kNames[0]); ++i) {
        if(s == kNames[i]) { out = static_cast<FOp>(i); return true; }
    }
    return false;
}

static std::uint64_t bits_of(double d) { std::uint64_t u; std::memcpy(&u, &d, 8); return u; }
static double double_of(std::uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }

void Program::write(std::ostream& o) const {
    o << "sqd-program 1\n";
    o << "name " << name << "\n";
    o << "nargs " << nargs << "\n";
    for(int i = 0; i < nargs; ++i) o << "arg " << i << " " << arg_names[static_cast<std::size_t>(i)] << "\n";
    for(const auto& s : splits) o << "split " << s.arg << " " << s.initial << " " << s.subs << " " << (s.dynamic ? 1 : 0) << "\n";
    o << "iqs " << initial_queue_size << "\n";
    for(std::size_t i = 0; i < nodes.size(); ++i) {
        const FNode& n = nodes[i];
        o << "node " << i << " " << fop_name(n.op) << " " << n.arg << " " << n.kids.size();
        for(int k : n.kids) o << " " << k;
        o << " " << bits_of(n.c_lb) << " " << bits_of(n.c_ub) << " " << (n.c_undef ? 1 : 0)
          << " " << (n.b_def ? 1 : 0) << " " << (n.b_poss ? 1 : 0)
          << " " << n.lb << " " << n.ub << " " << (n.plb ? 1 : 0) << " " << (n.pub ? 1 : 0) << "\n";
    }
    for(std::size_t k = 0; k < bounds.size(); ++k) {
        const FBound& b = bounds[k];
        o << "bound " << k << " " << b.target << " " << static_cast<int>(b.dir) << " " << b.expr << " "
          << (b.honor_undef ? 1 : 0) << "\n";
    }
    for(int a = 0; a < nargs; ++a) {
        for(int w = 0; w < 3; ++w) {
            const auto& l = update[static_cast<std::size_t>(a)][static_cast<std::size_t>(w)];
            o << "update " << a << " " << w << " " << l.size();
            for(int x : l) o << " " << x;
            o << "\n";
        }
    }
    for(int a = 0; a < nargs; ++a) {
        const auto& r = rows[static_cast<std::size_t>(a)];
        o << "row " << a << " " << r.size();
        for(int x : r) o << " " << x;
        o << "\n";
    }
    for(int a = 0; a < nargs; ++a) {
        o << "ct " << a << " " << ct_lb[static_cast<std::size_t>(a)] << " " << ct_ub[static_cast<std::size_t>(a)] << "\n";
    }
    for(int a = 0; a < nargs; ++a) {
        o << "box " << a << " " << bits_of(box_lb[static_cast<std::size_t>(a)]) << " "
          << bits_of(box_ub[static_cast<std::size_t>(a)]) << "\n";
    }
    o << "end\n";
}

Program Program::read(std::istream& in) {
    Program p;
    std::string line;
    auto fail = [](const std::string& why) { throw std::runtime_error("malformed program: " + why); };
    if(!std::getline(in, line) || line != "sqd-program 1") fail("header");
    while(std::getline(in, line)) {
        std::istringstream s(line);
        std::string key;
        s >> key;
        if(key == "end") return p;
        if(key == "name") { std::getline(s >> std::ws, p.name); }
        else if(key == "nargs") {
            s >> p.nargs;
            p.arg_names.assign(static_cast<std::size_t>(p.nargs), "");
            p.update.assign(static_cast<std::size_t>(p.nargs), std::vector<std::vector<int>>(3));
            p.rows.assign(static_cast<std::size_t>(p.nargs), {});
            p.ct_lb.assign(static_cast<std::size_t>(p.nargs), 0);
            p.ct_ub.assign(static_cast<std::size_t>(p.nargs), 0);
            p.box_lb.assign(static_cast<std::size_t>(p.nargs), 0);
            p.box_ub.assign(static_cast<std::size_t>(p.nargs), 0);
        }
        else if(key == "arg") { int i; s >> i; std::getline(s >> std::ws, p.arg_names.at(static_cast<std::size_t>(i))); }
        else if(key == "split") {
            FSplit sp; int dyn; s >> sp.arg >> sp.initial >> sp.subs >> dyn; sp.dynamic = dyn != 0; p.splits.push_back(sp);
        }
        else if(key == "iqs") s >> p.initial_queue_size;
        else if(key == "node") {
            FNode n; std::size_t id, nk; std::string op; std::uint64_t cl, cu; int cund, bd, bp, pl, pu;
            s >> id >> op >> n.arg >> nk;
            if(id != p.nodes.size()) fail("node order");
            if(!fop_from_name(op, n.op)) fail("op " + op);
            n.kids.resize(nk);
            for(auto& k : n.kids) { s >> k; if(k < 0 || static_cast<std::size_t>(k) >= id) fail("kid order"); }
            s >> cl >> cu >> cund >> bd >> bp >> n.lb >> n.ub >> pl >> pu;
            n.c_lb = double_of(cl); n.c_ub = double_of(cu); n.c_undef = cund != 0;
            n.b_def = bd != 0; n.b_poss = bp != 0; n.plb = pl != 0; n.pub = pu != 0;
            p.nodes.push_back(n);
        }
        else if(key == "bound") {
            FBound b; std::size_t k; int dir, hon; s >> k >> b.target >> dir >> b.expr >> hon;
            if(k != p.bounds.size()) fail("bound order");
            b.dir = static_cast<Dir>(dir); b.honor_undef = hon != 0; p.bounds.push_back(b);
        }
        else if(key == "update") {
            int a, w; std::size_t n; s >> a >> w >> n;
            auto& l = p.update.at(static_cast<std::size_t>(a)).at(static_cast<std::size_t>(w));
            l.resize(n); for(auto& x : l) s >> x;
        }
        else if(key == "row") {
            int a; std::size_t n; s >> a >> n;
            auto& r = p.rows.at(static_cast<std::size_t>(a));
            r.resize(n); for(auto& x : r) s >> x;
        }
        else if(key == "ct") { int a; s >> a; s >> p.ct_lb.at(static_cast<std::size_t>(a)) >> p.ct_ub.at(static_cast<std::size_t>(a)); }
        else if(key == "box") {
            int a; std::uint64_t l, u; s >> a >> l >> u;
            p.box_lb.at(static_cast<std::size_t>(a)) = double_of(l); p.box_ub.at(static_cast<std::size_t>(a)) = double_of(u);
        }
        else if(!key.empty()) fail("key " + key);
        if(s.fail()) fail("line: " + line);
    }
    fail("missing end");
    return p;
}

std::uint64_t Program::fingerprint() const {
    // names are cosmetic: identical lowered systems share one fingerprint
    Program anon = *this;
    anon.name.clear();
    for(auto& a : anon.arg_names) a = "_";
    std::ostringstream o;
    anon.write(o);
    const std::string s = o.str();
    std::uint64_t h = 1469598103934665603ull;
    for(unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
    return h;
}

} // namespace flat
} // namespace ivarp
