// flopcount.cpp -- the oracle compiled over an operation-counting scalar.
//
// TEST INFRASTRUCTURE ONLY (like the oracle it includes): it gives the
// algorithmic FP64 operation count of the reference's pair bodies as the
// oracle restates them (equations.py:14-567, kernels as defined in the
// oracle), branch-exact on the inputs it is given.  Every `double` of the
// oracle becomes a Counted, whose + - * / and sqrt (and cbrt, exp) each count
// one operation while a pair of a pass is being evaluated (the oracle's
// ORC_PAIR_BEGIN/BODY/END hooks): the pair's symbols (make_sym: r, rho_ij,
// W_i, grad W_i, grad W_j, dW_i/dh, each apart: ORC_SYM_PART) count apart
// from the pass's loop body.
// Comparisons, min/max, abs, floor/ceil, negation and conversions count
// nothing.  The arithmetic itself is the oracle's, value for value.
// Counters are plain globals: run it on one thread (orc_set_threads(1)).
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orcflop {
int on = 0, slot = 0;
int64_t ops[16], pairs[5];  // ops[pass]: bodies, ops[10 + part]: symbols (all passes)
inline void tick() {
    if (on) ++ops[slot];
}
}  // namespace orcflop

struct Counted {
    double v;
    Counted() = default;
    Counted(double x) : v(x) {}
    Counted(int x) : v((double)x) {}
    Counted(long x) : v((double)x) {}
    Counted(long long x) : v((double)x) {}
    Counted(unsigned x) : v((double)x) {}
    Counted(unsigned long x) : v((double)x) {}
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator unsigned() const { return (unsigned)v; }
    explicit operator unsigned long() const { return (unsigned long)v; }
    explicit operator double() const { return v; }
    Counted& operator+=(Counted b) { orcflop::tick(); v += b.v; return *this; }
    Counted& operator-=(Counted b) { orcflop::tick(); v -= b.v; return *this; }
    Counted& operator*=(Counted b) { orcflop::tick(); v *= b.v; return *this; }
    Counted& operator/=(Counted b) { orcflop::tick(); v /= b.v; return *this; }
};
static_assert(sizeof(Counted) == sizeof(double), "Counted must keep the layout of double");

inline Counted operator+(Counted a, Counted b) { orcflop::tick(); return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { orcflop::tick(); return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { orcflop::tick(); return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { orcflop::tick(); return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted operator+(Counted a) { return a; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
inline bool operator!(Counted a) { return !a.v; }
inline Counted sqrt(Counted a) { orcflop::tick(); return Counted(std::sqrt(a.v)); }
inline Counted cbrt(Counted a) { orcflop::tick(); return Counted(std::cbrt(a.v)); }
inline Counted exp(Counted a) { orcflop::tick(); return Counted(std::exp(a.v)); }
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted floor(Counted a) { return Counted(std::floor(a.v)); }
inline Counted ceil(Counted a) { return Counted(std::ceil(a.v)); }
inline Counted fmax(Counted a, Counted b) { return Counted(std::fmax(a.v, b.v)); }
inline Counted fmin(Counted a, Counted b) { return Counted(std::fmin(a.v, b.v)); }
inline Counted ldexp(Counted a, int e) { return Counted(std::ldexp(a.v, e)); }
inline bool isfinite(Counted a) { return std::isfinite(a.v); }

#define ORC_PAIR_BEGIN(p)        \
    do {                         \
        orcflop::on = 1;         \
        orcflop::slot = 10;      \
        ++orcflop::pairs[(p)];   \
    } while (0)
#define ORC_PAIR_BODY(p) (orcflop::slot = (p))
#define ORC_SYM_PART(k) (orcflop::slot = 10 + (k))
#define ORC_PAIR_END() (orcflop::on = 0)

#define double Counted
extern "C" {
#include "sphmhd_oracle.c"
}
#undef double

// counts since the last reset: body[pass] and pairs[pass] for the five
// passes (0 density, 1 Cullen-Dehnen, 2 MHD accelerations, 3 magnetic
// stress, 4 magnetic state rates), sym[part] over all passes for the six
// symbols (0 r, 1 rho_ij, 2 W_i, 3 grad W_i, 4 grad W_j, 5 dW_i/dh)
extern "C" void orcflop_reset(void) {
    std::memset(orcflop::ops, 0, sizeof orcflop::ops);
    std::memset(orcflop::pairs, 0, sizeof orcflop::pairs);
}
extern "C" void orcflop_get(int64_t* body, int64_t* sym, int64_t* pairs) {
    for (int k = 0; k < 5; ++k) {
        body[k] = orcflop::ops[k];
        pairs[k] = orcflop::pairs[k];
    }
    for (int k = 0; k < 6; ++k) sym[k] = orcflop::ops[10 + k];
}
