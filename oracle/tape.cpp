// tape.cpp -- the oracle compiled over a taped reverse-mode scalar.
//
// TEST INFRASTRUCTURE ONLY (like the oracle it includes): it gives the exact
// gradient of a rollout of the oracle's discrete map -- the warm start of
// application.py:110-147 and PEC steps of integrator.py:12-82 around the
// group plan of scheme.py:70-81 -- with respect to every property of the
// initial state and the per-instance parameters, by reverse accumulation over
// a tape of the FP64 operations the oracle performs.  It is the CPU reference
// gradient the HIP adjoint step is checked against.
//
// Every `double` of the oracle becomes a Taped: its value (the oracle's own
// arithmetic, value for value, run on one thread) and the index of the tape
// node that produced it (0: a constant).  + - * / sqrt cbrt exp record one
// node with the local partial derivatives of their operands; comparisons,
// min/max and branches act on the value, so the tape is the derivative of
// the path the oracle took (neighbour sets, cells, converged flags and the
// particle attaining the minimum time step fixed by the values).  fmax/fmin
// pass on the operand they return (ties: the first, as the HIP kernels'
// duals do); fabs has derivative 0 at 0, sqrt and cbrt at 0 too; floor, ceil
// and ldexp of a value are constants.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

struct Taped;
namespace tape {
struct Node {
    int32_t a, b;  // operand nodes (0: none)
    double da, db; // d(result)/d(operand)
};
std::vector<Node> nodes(1, Node{0, 0, 0.0, 0.0});  // node 0: the constant
bool overflow = false;
constexpr size_t kMaxNodes = 1u << 30;
inline int32_t push(int32_t a, double da, int32_t b, double db) {
    if (!a && !b) return 0;
    if (nodes.size() >= kMaxNodes) {
        overflow = true;
        return 0;
    }
    nodes.push_back(Node{a, b, da, db});
    return (int32_t)(nodes.size() - 1);
}
}  // namespace tape

struct Taped {
    double v;
    int32_t i;
    int32_t pad;
    Taped() = default;
    constexpr Taped(double x) : v(x), i(0), pad(0) {}
    constexpr Taped(int x) : v((double)x), i(0), pad(0) {}
    constexpr Taped(long x) : v((double)x), i(0), pad(0) {}
    constexpr Taped(long long x) : v((double)x), i(0), pad(0) {}
    constexpr Taped(unsigned x) : v((double)x), i(0), pad(0) {}
    constexpr Taped(unsigned long x) : v((double)x), i(0), pad(0) {}
    Taped(double x, int32_t idx) : v(x), i(idx), pad(0) {}
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator unsigned() const { return (unsigned)v; }
    explicit operator unsigned long() const { return (unsigned long)v; }
    explicit operator double() const { return v; }
};

inline Taped operator+(Taped a, Taped b) { return Taped(a.v + b.v, tape::push(a.i, 1.0, b.i, 1.0)); }
inline Taped operator-(Taped a, Taped b) { return Taped(a.v - b.v, tape::push(a.i, 1.0, b.i, -1.0)); }
inline Taped operator*(Taped a, Taped b) { return Taped(a.v * b.v, tape::push(a.i, b.v, b.i, a.v)); }
inline Taped operator/(Taped a, Taped b) {
    const double q = a.v / b.v;
    return Taped(q, tape::push(a.i, 1.0 / b.v, b.i, -q / b.v));
}
inline Taped operator-(Taped a) { return Taped(-a.v, tape::push(a.i, -1.0, 0, 0.0)); }
inline Taped operator+(Taped a) { return a; }
inline Taped& operator+=(Taped& a, Taped b) { return a = a + b; }
inline Taped& operator-=(Taped& a, Taped b) { return a = a - b; }
inline Taped& operator*=(Taped& a, Taped b) { return a = a * b; }
inline Taped& operator/=(Taped& a, Taped b) { return a = a / b; }
inline bool operator<(Taped a, Taped b) { return a.v < b.v; }
inline bool operator>(Taped a, Taped b) { return a.v > b.v; }
inline bool operator<=(Taped a, Taped b) { return a.v <= b.v; }
inline bool operator>=(Taped a, Taped b) { return a.v >= b.v; }
inline bool operator==(Taped a, Taped b) { return a.v == b.v; }
inline bool operator!=(Taped a, Taped b) { return a.v != b.v; }
inline bool operator!(Taped a) { return !a.v; }
inline Taped sqrt(Taped a) {
    const double s = std::sqrt(a.v);
    return Taped(s, tape::push(a.i, s > 0.0 ? 0.5 / s : 0.0, 0, 0.0));
}
inline Taped cbrt(Taped a) {
    const double s = std::cbrt(a.v);
    return Taped(s, tape::push(a.i, s != 0.0 ? 1.0 / (3.0 * s * s) : 0.0, 0, 0.0));
}
inline Taped exp(Taped a) {
    const double e = std::exp(a.v);
    return Taped(e, tape::push(a.i, e, 0, 0.0));
}
inline Taped fabs(Taped a) { return a.v < 0.0 ? -a : a; }
inline Taped floor(Taped a) { return Taped(std::floor(a.v)); }
inline Taped ceil(Taped a) { return Taped(std::ceil(a.v)); }
inline Taped fmax(Taped a, Taped b) { return a.v >= b.v ? a : b; }
inline Taped fmin(Taped a, Taped b) { return a.v <= b.v ? a : b; }
inline Taped ldexp(Taped a, int e) { return Taped(std::ldexp(a.v, e)); }
inline bool isfinite(Taped a) { return std::isfinite(a.v); }

#define double Taped
extern "C" {
#include "sphmhd_oracle.c"
}
#undef double

// The parameter block as the library's ABI lays it out (sphmhd_params of
// include/sphmhd.h, whose doubles became Taped in this translation unit).
struct PlainParams {
    int32_t dim, kernel, density_iterate, min_iterations, max_iterations, periodic, adaptive_timestep, reserved;
    double lo[3], hi[3];
    double gamma, mu0, hfact, artificial_viscosity, artificial_thermal_conductivity,
        artificial_magnetic_dissipation, cleaning_speed_factor, cleaning_damping_factor, timestep_factor, alpha_min,
        alpha_max, tf;
};
static_assert(sizeof(PlainParams) == 176, "sphmhd_params layout");

// the differentiated parameters, in the order of sphmhd_tparam
enum { TP_AV = 0, TP_AC, TP_AR, TP_CHF, TP_SIGMA, TP_ALPHA_MAX, NTP };

static Taped input(double x) {
    tape::nodes.push_back(tape::Node{0, 0, 0.0, 0.0});
    return Taped(x, (int32_t)(tape::nodes.size() - 1));
}

extern "C" {

int tape_nprops(void) { return NPROP; }

// A rollout of n_instances instances (params: PlainParams each, offsets:
// n_instances + 1): fields[q] (NPROP host arrays in the oracle's property
// order, null = zero) and itype set the initial state; then the warm start
// (warm != 0) and nsteps PEC steps; the objective of instance k is
// sum_q weights[q] * total_q of instance k at the end (SPHMHD_NTOTALS
// weights).  Returns the objectives in value[k], the gradient of their sum
// with respect to every initial property in grads[q] (NPROP arrays or null:
// not wanted) and with respect to the parameters of sphmhd_tparam in
// grad_params[k * NTP + p]; tape_len receives the tape's node count.
// Returns 0, or the oracle's error code, or -100 when the tape overflowed.
int tape_rollout(const void* params, int32_t n_instances, const int64_t* offsets, const double* const* fields,
                 const int32_t* itype, int64_t nsteps, int32_t warm, const double* weights, double* value,
                 double* const* grads, double* grad_params, int64_t* tape_len) {
    tape::nodes.assign(1, tape::Node{0, 0, 0.0, 0.0});
    tape::overflow = false;
    const PlainParams* pp = (const PlainParams*)params;
    std::vector<sphmhd_params> prm(n_instances);
    std::vector<int32_t> pidx((size_t)n_instances * NTP);
    for (int k = 0; k < n_instances; ++k) {
        const PlainParams& a = pp[k];
        sphmhd_params& b = prm[k];
        std::memset((void*)&b, 0, sizeof b);
        b.dim = a.dim;
        b.kernel = a.kernel;
        b.density_iterate = a.density_iterate;
        b.min_iterations = a.min_iterations;
        b.max_iterations = a.max_iterations;
        b.periodic = a.periodic;
        b.adaptive_timestep = a.adaptive_timestep;
        b.reserved = a.reserved;
        for (int q = 0; q < 3; ++q) {
            b.lo[q] = a.lo[q];
            b.hi[q] = a.hi[q];
        }
        b.gamma = a.gamma;
        b.mu0 = a.mu0;
        b.hfact = a.hfact;
        b.artificial_viscosity = input(a.artificial_viscosity);
        b.artificial_thermal_conductivity = input(a.artificial_thermal_conductivity);
        b.artificial_magnetic_dissipation = input(a.artificial_magnetic_dissipation);
        b.cleaning_speed_factor = input(a.cleaning_speed_factor);
        b.cleaning_damping_factor = input(a.cleaning_damping_factor);
        b.alpha_max = input(a.alpha_max);
        b.timestep_factor = a.timestep_factor;
        b.alpha_min = a.alpha_min;
        b.tf = a.tf;
        pidx[(size_t)k * NTP + TP_AV] = b.artificial_viscosity.i;
        pidx[(size_t)k * NTP + TP_AC] = b.artificial_thermal_conductivity.i;
        pidx[(size_t)k * NTP + TP_AR] = b.artificial_magnetic_dissipation.i;
        pidx[(size_t)k * NTP + TP_CHF] = b.cleaning_speed_factor.i;
        pidx[(size_t)k * NTP + TP_SIGMA] = b.cleaning_damping_factor.i;
        pidx[(size_t)k * NTP + TP_ALPHA_MAX] = b.alpha_max.i;
    }
    orc_set_threads(1);
    orc_t* o = orc_create(prm.data(), n_instances, offsets);
    if (!o) return SPHMHD_EINVAL;
    const int64_t n = o->n;
    std::vector<int32_t> fidx((size_t)NPROP * n, 0);
    for (int q = 0; q < NPROP; ++q)
        for (int64_t i = 0; i < n; ++i) {
            Taped x = input(fields[q] ? fields[q][i] : 0.0);
            o->p[q][i] = x;
            fidx[(size_t)q * n + i] = x.i;
        }
    if (itype) std::memcpy(o->itype, itype, sizeof(int32_t) * n);
    int rc = 0;
    if (warm) rc = orc_warm_start(o);
    if (rc == 0 && nsteps > 0) orc_step(o, nsteps);
    std::vector<Taped> tot((size_t)n_instances * SPHMHD_NTOTALS);
    orc_totals(o, tot.data());
    std::vector<Taped> obj(n_instances, Taped(0.0));
    for (int k = 0; k < n_instances; ++k) {
        for (int q = 0; q < SPHMHD_NTOTALS; ++q)
            if (weights[q] != 0.0) obj[k] += weights[q] * tot[(size_t)k * SPHMHD_NTOTALS + q];
        value[k] = obj[k].v;
    }
    if (tape_len) *tape_len = (int64_t)tape::nodes.size();
    if (tape::overflow) rc = -100;
    if (rc == 0) {
        std::vector<double> adj(tape::nodes.size(), 0.0);
        for (int k = 0; k < n_instances; ++k)
            if (obj[k].i) adj[obj[k].i] += 1.0;
        for (size_t j = tape::nodes.size() - 1; j > 0; --j) {
            const double g = adj[j];
            if (g == 0.0) continue;
            const tape::Node& nd = tape::nodes[j];
            if (nd.a) adj[nd.a] += g * nd.da;
            if (nd.b) adj[nd.b] += g * nd.db;
        }
        for (int q = 0; q < NPROP; ++q)
            if (grads && grads[q])
                for (int64_t i = 0; i < n; ++i) grads[q][i] = adj[fidx[(size_t)q * n + i]];
        if (grad_params)
            for (size_t k = 0; k < pidx.size(); ++k) grad_params[k] = adj[pidx[k]];
    }
    orc_destroy(o);
    tape::nodes.clear();
    tape::nodes.shrink_to_fit();
    return rc;
}

}  // extern "C"
