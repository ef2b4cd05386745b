// Operation count of one RKF45 trial of the reference integrator (test /
// measurement infrastructure, part of liboracle.so).  The oracle's
// rkf45_step and right-hand sides are instantiated on a counting scalar:
// every +, -, *, / between non-constant operands, every sqrt and every pow is
// tallied; operations between compile-time constants (the tableau quotients,
// folded by any compiler) are not.  The totals are the algorithmic FLOPs per
// trial used for the roofline of the integration kernel.
#include <cstdint>

namespace oracle {
struct Cnt;
}
namespace {
struct Tally { uint64_t add = 0, mul = 0, div = 0, sqrt = 0, pow = 0; };
thread_local Tally g_tally;
}  // namespace

namespace oracle {
struct Cnt {
    double v = 0;
    bool k = false;   // a constant
    Cnt() = default;
    Cnt(double x) : v(x), k(true) {}
    static Cnt var(double x) { Cnt c; c.v = x; c.k = false; return c; }
};
inline Cnt mk(double v, bool k) { Cnt c; c.v = v; c.k = k; return c; }
inline Cnt operator+(Cnt a, Cnt b) { if (!(a.k && b.k)) g_tally.add++; return mk(a.v + b.v, a.k && b.k); }
inline Cnt operator-(Cnt a, Cnt b) { if (!(a.k && b.k)) g_tally.add++; return mk(a.v - b.v, a.k && b.k); }
inline Cnt operator*(Cnt a, Cnt b) { if (!(a.k && b.k)) g_tally.mul++; return mk(a.v * b.v, a.k && b.k); }
inline Cnt operator/(Cnt a, Cnt b) { if (!(a.k && b.k)) g_tally.div++; return mk(a.v / b.v, a.k && b.k); }
inline Cnt operator-(Cnt a) { return mk(-a.v, a.k); }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
}  // namespace oracle

#include "oracle.hpp"

namespace oracle {
template <> inline Cnt sqrt_<Cnt>(Cnt v) { g_tally.sqrt++; return mk(std::sqrt(v.v), false); }
template <> inline Cnt pow_<Cnt>(Cnt a, Cnt b) { g_tally.pow++; return mk(std::pow(a.v, b.v), false); }
template <> inline Cnt abs_<Cnt>(Cnt v) { return mk(std::fabs(v.v), v.k); }
template <> inline Cnt max_<Cnt>(Cnt a, Cnt b) { return a.v > b.v ? a : b; }
}  // namespace oracle

extern "C" {

// counts[5] = {add/sub, mul, div, sqrt, pow} of one trial (metric: 1 = Schwarzschild, 2 = Kerr).
void orc_trial_opcount(int metric, uint64_t counts[5]) {
    using namespace oracle;
    g_tally = Tally{};
    State<Cnt> y0{{Cnt::var(3.0), Cnt::var(4.0), Cnt::var(20.0), Cnt::var(6e8)},
                  {Cnt::var(0.3), Cnt::var(-0.2), Cnt::var(-1.5), Cnt::var(-1.6)}};
    Cnt h = Cnt::var(1.0), ht = Cnt::var(0.0);
    Cnt a = Cnt::var(-0.9), ms = Cnt::var(1.0);
    auto f = [&](const State<Cnt>& s) { return metric == kKerr ? kerr_rhs(s, a) : schwarzschild_rhs(s, ms); };
    rkf45_step<Cnt>(f, y0, &h, &ht, Cnt::var(1000.0), 1u, Cnt::var(1e-7), Cnt::var(1e-7));
    counts[0] = g_tally.add; counts[1] = g_tally.mul; counts[2] = g_tally.div;
    counts[3] = g_tally.sqrt; counts[4] = g_tally.pow;
}

// counts of one right-hand-side evaluation
void orc_rhs_opcount(int metric, uint64_t counts[5]) {
    using namespace oracle;
    g_tally = Tally{};
    State<Cnt> y0{{Cnt::var(3.0), Cnt::var(4.0), Cnt::var(20.0), Cnt::var(6e8)},
                  {Cnt::var(0.3), Cnt::var(-0.2), Cnt::var(-1.5), Cnt::var(-1.6)}};
    if (metric == kKerr) kerr_rhs(y0, Cnt::var(-0.9));
    else schwarzschild_rhs(y0, Cnt::var(1.0));
    counts[0] = g_tally.add; counts[1] = g_tally.mul; counts[2] = g_tally.div;
    counts[3] = g_tally.sqrt; counts[4] = g_tally.pow;
}

}  // extern "C"
