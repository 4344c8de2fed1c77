// sph_internal.h -- shared definitions of the gfx950 SPH-MHD library.
//
// Layout in HBM
//   * Bound particle properties stay in the caller's storage order (one FP64
//     array per property, the PySPH ParticleArray layout).
//   * Each evaluation bins particles into cells, radix-sorts the cell keys and
//     then writes cell-ordered AoS "packs" of exactly what the pair loops read
//     from a neighbour: a 32-byte record for the density pass and a 144-byte
//     (2-D) or 160-byte (3-D) record for the force pass.  Pair kernels run
//     one thread per destination in cell order, so a wave's neighbours sit in
//     a handful of contiguous pack runs.
//   * Per-instance scalars (grid, clock, reductions) live in one device array
//     of Inst, so an ensemble of uneven instances runs in the same launches as
//     a single instance.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <stdint.h>

#include "../../include/sphmhd.h"

namespace sph {

constexpr int kBlock = 256;  // workgroup size (the largest shape of the tuned launches)
constexpr int kWave = 64;
// launch classes whose workgroup size is tuned per context (64, 128 or 256)
enum { LC_DENSITY = 0, LC_LISTS, LC_FORCE, kLaunchClasses };
constexpr int kScanItems = 4;    // cells per thread of the cell-count scan
constexpr int kScanCells = 256 * kScanItems;  // cells per scan block
constexpr int kCellsPerParticle = 8;  // a grid has at most this many cells per particle (max(1, 8 N))
constexpr int kSmallBucket = 128; // cells up to this size are ordered by rank, one thread per particle
constexpr int kBigLds = 4096;    // larger cells up to this size are sorted in LDS

// Force-pack record: the FP64 values the force pair loop reads from a
// neighbour, 18 in 2-D (144 B) and 20 in 3-D (160 B), packed densely so that
// a record is a whole number of 16-byte units.  (A unit-major layout, one
// array per unit, was measured too: it gathers no faster and is slower on the
// 40-neighbour quintic workloads.)  The pair loop is bound by the
// bytes it gathers, so whatever the loop can rebuild exactly from the record
// is left out: p = (gamma - 1) rho e and 1/rho, P/rho^2, B/rho are recomputed
// per pair with the same operations the prologue uses.  au, av, aw are the
// accelerations of the previous evaluation; PSICH is psipred * c_h.
template <int DIM>
struct FRec {
    static constexpr int X = 0, Y = 1, Z = 2;
    static constexpr int H1 = DIM == 3 ? 3 : 2;  // 1 / h
    static constexpr int M = H1 + 1, RHO = M + 1, E = RHO + 1;
    static constexpr int U = E + 1, V = U + 1, W = V + 1, OM = W + 1;
    static constexpr int BX = OM + 1, BY = BX + 1, BZ = BY + 1, AL = BZ + 1;
    static constexpr int AU = AL + 1, AV = AU + 1, AW = AV + 1;  // AW only in 3-D
    static constexpr int PSICH = DIM == 3 ? AW + 1 : AV + 1;
    static constexpr int VW = PSICH + 1;  // wave speed
    static constexpr int N = VW + 1;
    static constexpr int UNITS = N / 2;  // 16-byte units per record
};
static_assert(FRec<2>::N == 18 && FRec<3>::N == 20, "force record sizes");
constexpr int kFRecMax = 20;  // doubles allocated per particle

// One instance's parameters, grid and clock, then the targets of its
// reductions.  Fields that kernels update with device-scope atomics while
// other workgroups of the same launch read the instance sit on 128-byte lines
// of their own, one line per launch family: an atomic on a line that also
// holds fields every workgroup loads made those loads miss the cache (the
// force pass ran 20-25% slower with dt_adapt's target beside the fields its
// epilogue reads).
struct Inst {
    // static parameters (filled at create)
    int64_t off0, off1;
    int32_t dim, kernel, density_iterate, min_iter;
    int32_t max_iter, periodic, adaptive;
    int32_t range_shift;  // the stencil's per-range image shifts equal the minimum image (see stencil_from_hmax)
    int32_t split;        // cells per R h_max along an axis (grid_setup); above 1 the lists prune by cell h
    double lo[3], hi[3], L[3], halfL[3];
    double gamma, gamma1, mu0, hfact;
    double av, ac, ar, chf;
    double sigma, tsf, alpha_min, alpha_max;
    double tf, cleaning_active, c_cour, c_force;
    // grid of the current evaluation
    int32_t n[3];
    int32_t srad[3];
    double ceff[3];
    int64_t cell_base, ncell;
    // clock
    double t, dt;
    // density-iteration state (the group control's; read by the passes)
    unsigned long long hmax_bits;       // max h for the current pass
    int32_t iters, done;
    int32_t spread;  // lanes per particle of this instance in the next pass (set with its count)
    // reduction targets (bit patterns of positive doubles order like integers)
    // force pass: min dt_adapt and the warning bits (1: density hit
    // max_iterations, 2: NaN)
    alignas(128) unsigned long long dtp_bits;  // min dt_adapt of the last force pass
    int32_t status;
    // density passes
    alignas(128) unsigned long long hmax_next_bits;  // max h of still-unconverged particles after the running pass
    unsigned long long hmax_conv_bits;               // max h of converged particles (frozen for the evaluation)
    int32_t unconv;
    int32_t passes;  // most passes a particle took (the passes iterated within one launch)
    // the evaluation's start: h_max, h_min and the particles' box, reduced as
    // maxima of kRed bit patterns (see red_values); red_stage accumulates
    // them in the fused stage kernel (zero between steps)
    alignas(128) unsigned long long red[8];
    unsigned long long red_stage[8];
};

// The evaluation's control block.  The first line holds what every
// workgroup reads (written between launches or by a launch's last
// workgroup); each family of device-scope atomics has a 128-byte line of its
// own, so they do not make those reads miss.
struct Ctl {
    int64_t total_cells;
    int32_t all_done;
    int32_t status;
    int32_t iters_max;
    int32_t any_split;  // some instance's grid has split cells (grid_setup)
    // Steps are enqueued one ahead of the host's wait.  An evaluation that
    // needs more density passes or longer lists sets abort when it ends, and
    // every kernel of the step enqueued after it then exits at entry, leaving
    // the state as that evaluation left it for the host to finish.
    int32_t abort;
    int32_t seq;  // evaluations ended so far (the host mirror's slot is seq % 2)
    int32_t at_tf;  // set by an evaluation's end: its step brings every clock to tf
    // 1 while a multi-step run is before its last step: the force pass then
    // skips the diagnostic rows no later kernel reads (set by the host)
    int32_t lean;
    // the next density pass maps spread lanes to each particle (the largest
    // Inst::spread of the instances still iterating)
    int32_t spread;
    // list build
    alignas(128) int64_t pairs;
    int32_t max_nb;
    int32_t overflow;  // a neighbour list did not fit its slots
    // lengths of the two active-particle lists of the density passes (a pass
    // reads one and appends to the other)
    alignas(128) int32_t act0;
    alignas(128) int32_t act1;
    alignas(128) int32_t nbig;   // cells with more than kSmallBucket particles in this sort
    alignas(128) int64_t count;  // particles counted by k_count
    // histogram of h for the 3-D cell size (a context of one 3-D instance,
    // else null): kHistCopies rings of kHistSlots counts, filled by the
    // launch that reduces h_max and emptied by its grid set-up
    alignas(128) uint32_t* hist;
};

__host__ __device__ inline int32_t* act_count(Ctl* c, int k) { return k ? &c->act1 : &c->act0; }

// the host's copy of the control block (pinned, written by the device)
struct CtlView {
    int64_t total_cells, pairs, count;
    int32_t all_done, status, max_nb, iters_max, overflow, nbig, any_split, abort, seq;
    int32_t at_tf;  // the step of this evaluation brings every instance's clock to its tf
};

__device__ inline CtlView ctl_view(const Ctl& c) {
    CtlView v;
    v.total_cells = c.total_cells;
    v.pairs = c.pairs;
    v.count = c.count;
    v.all_done = c.all_done;
    v.status = c.status;
    v.max_nb = c.max_nb;
    v.iters_max = c.iters_max;
    v.overflow = c.overflow;
    v.nbig = c.nbig;
    v.any_split = c.any_split;
    v.abort = c.abort;
    v.seq = c.seq;
    v.at_tf = c.at_tf;
    return v;
}

struct Fields {
    double* f[SPHMHD_NFIELDS];
    int32_t* itype;
};

// the rows StrainDiagnostics fills (dvdx_xx, _xy, _xz, _yx, ..., _zz), all
// null unless bound by sphmhd_bind_strain
struct Strain {
    double* d[9];
};

// ---------------------------------------------------------------------------
// exact helpers (must round exactly like the CPU oracle)
// ---------------------------------------------------------------------------

__host__ __device__ inline double kernel_radius(int k) {
    return (k == SPHMHD_QUINTIC || k == SPHMHD_GAUSSIAN || k == SPHMHD_SUPER_GAUSSIAN) ? 3.0 : 2.0;
}

__device__ __forceinline__ double min_image(double d, double L, double halfL, int periodic) {
    if (periodic) {
        if (d > halfL) d -= L;
        else if (d < -halfL) d += L;
    }
    return d;
}

__device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    double s = dx * dx + dy * dy;
    return s + dz * dz;
}

// sqrt of a squared distance: the Goldschmidt step and the two Newton
// corrections of the library's correctly rounded sqrt, without its rescaling
// of inputs below 2^-767 and its class test (zero, infinity, NaN).  The
// input is r2 + 2^-600, which is r2 itself for every r2 >= 2^-546 (2^-600 is
// below half its ulp), so the result equals sqrt(r2) bit for bit there; a
// distance of zero (a particle with itself) gives 2^-300, a q whose square
// vanishes beside 1 and whose product with a kernel derivative vanishes
// beside W(0), so the density sums keep their bits.
__device__ __forceinline__ double sqrt_r2(double r2) {
    const double x = r2 + 0x1p-600;
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;
    const double e = ::fma(-h, s, 0.5);
    s = ::fma(s, e, s);
    h = ::fma(h, e, h);
    double d = ::fma(-s, s, x);
    s = ::fma(d, h, s);
    d = ::fma(-s, s, x);
    return ::fma(d, h, s);
}

// sqrt of a non-negative x by sqrt_r2's steps, exactly 0 at 0: equal to
// sqrt(x) bit for bit for x = 0 and every x >= 2^-546 (NaN stays NaN)
__device__ __forceinline__ double sqrt_nn(double x) {
    const double s = sqrt_r2(x);
    return x == 0.0 ? 0.0 : s;
}

__device__ __forceinline__ int cell_coord(double x, double lo, double ceff, int n) {
    double f = floor((x - lo) / ceff);
    if (f < 0.0) return 0;
    if (f > (double)(n - 1)) return n - 1;
    return (int)f;
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

// True in exactly one workgroup of a launch: the last to arrive here.  Every
// workgroup of the grid must call it (no early exit before it); the counter
// is back at zero when the launch ends.  This runs a launch's small serial
// tail (group control, grid set-up, a top-level scan) inside the launch
// instead of as a kernel of its own.
//
// Visibility: what a tail reads of its own launch's results is written with
// device-scope atomics or st_dev and read with ld_dev, which are coherent
// across the XCDs; plain stores may still sit in an XCD's L2 (a device-scope
// fence would write back and invalidate the L2 in every workgroup: measured,
// it doubled the step).  Ordering: on gfx950 a workgroup-scope release does
// not wait for outstanding vector memory operations, so every wave first
// waits for all of its own (stores and atomics without return count in
// vmcnt) with s_waitcnt 0; only then does the barrier let thread 0 issue the
// arrival atomic, so the last workgroup never reads a slot another workgroup
// has not finished writing.
//
// Same-address atomics serialise (about 10 ns each), so a thousand
// workgroups do not share one counter: they arrive in kArriveGroups groups
// (blockIdx mod kArriveGroups, one 128-byte line per group counter) and the
// last of each group arrives at the top counter.
constexpr int kArriveGroups = 32;
constexpr int kArriveStride = 32;  // int32 per counter line
constexpr int kArriveInts = (kArriveGroups + 1) * kArriveStride;

__device__ __forceinline__ bool last_to_arrive(int32_t* arrive) {
    __shared__ int s_last;
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt, expcnt, lgkmcnt all zero
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = (int)(gridDim.x * gridDim.y * gridDim.z);
        const int b = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
        const int g = b % kArriveGroups;
        const int members = total / kArriveGroups + (g < total % kArriveGroups ? 1 : 0);
        const int groups = total < kArriveGroups ? total : kArriveGroups;
        int last = 0;
        int32_t* gc = arrive + (1 + g) * kArriveStride;
        if (atomicAdd(gc, 1) == members - 1) {
            atomicExch(gc, 0);
            if (atomicAdd(arrive, 1) == groups - 1) {
                atomicExch(arrive, 0);
                last = 1;
            }
        }
        s_last = last;
    }
    __syncthreads();
    return s_last;
}

// device-coherent load and store (bypass the non-coherent caches)
template <class T>
__device__ __forceinline__ T ld_dev(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dev(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// start of an evaluation: control block and per-instance reductions cleared
// (reset_dt: the pending dt too); run by one workgroup
__device__ inline void eval_begin_body(Inst* inst, int ninst, Ctl* ctl, int reset_dt) {
    if (threadIdx.x == 0) {
        ctl->all_done = 0;
        ctl->status = 0;
        ctl->max_nb = 0;
        ctl->iters_max = 0;
        ctl->pairs = 0;
        ctl->count = 0;
        ctl->act0 = 0;
        ctl->act1 = 0;
        ctl->overflow = 0;
        ctl->at_tf = 0;
    }
    for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
        Inst& s = inst[k];
        s.hmax_bits = 0ull;
        for (int q = 0; q < 8; ++q) s.red[q] = 0ull;
        s.hmax_next_bits = 0ull;
        s.hmax_conv_bits = 0ull;
        if (reset_dt) s.dtp_bits = 0x7FF0000000000000ull;  // +inf
        s.iters = 0;
        s.unconv = 0;
        s.passes = 0;
        s.done = 0;
        s.status = 0;
    }
}

// stencil half-width per axis from the instance's current h_max
// range_shift: a candidate of stencil cell offset k along an axis lies within
// (|k| + 1) cell widths of the particle, so when every axis has more than
// 2 (s + 1) cells (s the stencil half-width) no unwrapped candidate is more
// than half a box away and every wrapped one is: the minimum image of a whole
// stencil range is then one shift (0 or +-L), and adding it gives the very
// operation min_image performs (d + (-L) == d - L in IEEE arithmetic)
__device__ inline void stencil_from_hmax(Inst& s) {
    double R = kernel_radius(s.kernel);
    double hm = bitsd(s.hmax_bits);
    int shift_ok = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < s.dim) {
            double r = ceil(R * hm / s.ceff[a]);
            s.srad[a] = !(r < 1.0e6) ? 1000000 : (r < 1.0 ? 1 : (int)r);
            if (s.periodic && !((int64_t)s.n[a] > 2 * ((int64_t)s.srad[a] + 1))) shift_ok = 0;
        } else {
            s.srad[a] = 0;
        }
    }
    s.range_shift = shift_ok;
}

// Reductions at an evaluation's start, all as maxima of 64-bit patterns (zero
// = nothing yet): 0 h_max (positive doubles order like their bits), 1 the
// complement of h_min's bits, 2-4 the ordered bits of max x, y, z, 5-7 the
// complements of the ordered bits of min x, y, z.
constexpr int kRed = 8;

// a double's bits mapped so that unsigned order is numeric order
__host__ __device__ inline unsigned long long ord_bits(double x) {
    unsigned long long b;
    memcpy(&b, &x, sizeof b);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double ord_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    double x;
    memcpy(&x, &b, sizeof x);
    return x;
}
__device__ __forceinline__ void red_values(unsigned long long (&v)[kRed], double h, double x, double y, double z) {
    v[0] = dbits(h);
    v[1] = ~dbits(h);
    v[2] = ord_bits(x);
    v[3] = ord_bits(y);
    v[4] = ord_bits(z);
    v[5] = ~ord_bits(x);
    v[6] = ~ord_bits(y);
    v[7] = ~ord_bits(z);
}

// margin of an open domain's grid box around the particles' extent [mn, mx]:
// 1% of the extent, at least 1e-9 of its largest coordinate (and of 1)
__host__ __device__ inline double box_pad(double mn, double mx) {
    return ::fmax(0.01 * (mx - mn), 1.0e-9 * ::fmax(::fmax(::fabs(mn), ::fabs(mx)), 1.0));
}

// cells per R h_max along an axis: min(floor(h_max / h_min), 4) when h_max
// >= 2 h_min, else 1
__host__ __device__ inline double cell_split(double hmax, double hmin) {
    if (hmin > 0.0 && hmax / hmin >= 2.0) return ::fmin(::floor(hmax / hmin), 4.0);
    return 1.0;
}

// Levels of h for the 3-D cell size: the biased exponent and top four
// mantissa bits of a positive h (16 levels per octave, increasing with h),
// counted in a ring of kHistSlots slots; level_value(l) is the smallest h of
// level l, so level l holds [level_value(l), level_value(l + 1)).
constexpr int kHistSlots = 64;
constexpr int kHistCopies = 8;  // rings, one per group of workgroups (same-address atomics serialise)
__host__ __device__ inline int h_level(double h) {
    unsigned long long b;
    memcpy(&b, &h, sizeof b);
    return (int)(b >> 48);
}
__host__ __device__ inline double level_value(int l) { return ldexp(1.0 + (double)(l & 15) / 16.0, (l >> 4) - 1023); }

// A workgroup's part of the h histogram: counts per slot in LDS (lh), then
// one atomic per occupied slot into one of the rings (callers put a barrier
// between the three)
__device__ __forceinline__ void hist_clear(uint32_t* lh) {
    for (int t = threadIdx.x; t < kHistSlots; t += blockDim.x) lh[t] = 0u;
}
__device__ __forceinline__ void hist_add(uint32_t* lh, double h) {
    if (h > 0.0) atomicAdd(&lh[h_level(h) & (kHistSlots - 1)], 1u);
}
__device__ __forceinline__ void hist_flush(uint32_t* hist, const uint32_t* lh) {
    uint32_t* dst = hist + (blockIdx.x % kHistCopies) * kHistSlots;
    for (int t = threadIdx.x; t < kHistSlots; t += blockDim.x)
        if (lh[t]) atomicAdd(&dst[t], lh[t]);
}

// Cell size of a 3-D instance from its h histogram (ring[(l0 + j) % slots],
// j < nl: the counts of levels l0 .. l0 + nl - 1, l0 the level of h_min):
// the candidate among c0 (the h_max rule) and half of R times the upper edge
// of each occupied level that minimises a model of the scan cost, the
// largest such cell on ties.  Per particle of level upper edge u (capped at
// h_max): s = max(1, ceil(R u / c)) stencil cells per axis, (2s+1)^3 cells of
// (c/u)^3 h-volumes holding hfact^3 particles each, plus kRowCost
// candidates' worth per stencil row ((2s+1)^2); a cell whose h_max stencil
// spans more than two cells makes the lists prune and bound each particle's
// reach, which costs kSplitCost times as much per modelled candidate (the
// factor that times best on the 128^3 blast: its grid splits once h_max
// passes about 1.2 times the bulk h).  Contraction off: the oracle makes the
// same choice from the same counts.
constexpr double kRowCost = 6.0;
constexpr double kSplitCost = 1.15;

// candidate cand (-1: c0) and its modelled cost
__host__ __device__ inline double cell_candidate(double R, double hm, int l0, int cand, double c0) {
#pragma clang fp contract(off)
    if (cand < 0) return c0;
    const double u = level_value(l0 + cand + 1);
    return 0.5 * (R * (u < hm ? u : hm));
}
__host__ __device__ inline double cell_cost_3d(double R, double hfact, double hm, int l0, int nl,
                                               const uint32_t* ring, double c) {
#pragma clang fp contract(off)
    const double k3 = hfact * hfact * hfact;
    double cost = 0.0;
    for (int j = 0; j < nl; ++j) {
        const uint32_t nj = ring[(l0 + j) & (kHistSlots - 1)];
        if (!nj) continue;
        const double v = level_value(l0 + j + 1);
        const double u = v < hm ? v : hm;
        double sj = ceil(R * u / c);
        if (sj < 1.0) sj = 1.0;
        const double w = 2.0 * sj + 1.0;
        const double x = w * c / u;
        cost += (double)nj * (k3 * (x * x * x) + kRowCost * (w * w));
    }
    return cost;
}

// cells per R h_max along an axis as the grid is built: in 3-D at least 2
// (twice the h split, at most 4): a 5^3 stencil of half-width cells holds
// 125/8 = 15.6 cell volumes of candidates against 27 for 3^3 full cells
__host__ __device__ inline double cell_split_dim(double hmax, double hmin, int dim) {
    const double k = cell_split(hmax, hmin);
    return dim == 3 ? ::fmin(2.0 * k, 4.0) : k;
}

// Grid set-up (one workgroup): one thread per instance, then a serial prefix
// of cell bases.
__device__ inline void grid_setup(Inst* inst, int ninst, Ctl* ctl) {
    // a single 3-D instance: its h histogram, summed over the rings (which
    // are emptied for the next launch) by the whole workgroup
    __shared__ uint32_t hring[kHistSlots];
    __shared__ double cand_c[kHistSlots + 1], cand_cost[kHistSlots + 1];
    __shared__ double model_c;
    uint32_t* hist = ctl->hist;
    if (hist) {
        for (int t = threadIdx.x; t < kHistSlots; t += blockDim.x) hring[t] = 0u;
        __syncthreads();
        for (int t = threadIdx.x; t < kHistSlots * kHistCopies; t += blockDim.x) {
            const uint32_t v = ld_dev(&hist[t]);
            if (v) {
                st_dev(&hist[t], 0u);
                atomicAdd(&hring[t % kHistSlots], v);
            }
        }
        __syncthreads();
        // the model's candidates costed in parallel (one per thread), then
        // the cheapest, the largest cell on ties
        const Inst& s0 = inst[0];
        const double hm = bitsd(ld_dev(&s0.red[0]));
        const unsigned long long r1 = ld_dev(&s0.red[1]);
        const double hn = r1 ? bitsd(~r1) : 0.0;
        const double R = kernel_radius(s0.kernel);
        const bool on = s0.dim == 3 && hn > 0.0 && hm >= hn && h_level(hm) - h_level(hn) < kHistSlots;
        const int l0 = on ? h_level(hn) : 0, nl = on ? h_level(hm) - l0 + 1 : 0;
        const double c0 = R * hm / cell_split_dim(hm, hn, 3);
        for (int t = threadIdx.x; t <= nl; t += blockDim.x) {
            const int cand = t - 1;
            const bool use = cand < 0 || hring[(l0 + cand) & (kHistSlots - 1)] != 0u;
            const double c = cell_candidate(R, hm, l0, cand, c0);
            cand_c[t] = c;
            double q = use ? cell_cost_3d(R, s0.hfact, hm, l0, nl, hring, c) : -1.0;
            if (use && ceil(R * hm / c) > 2.0) q = kSplitCost * q;
            cand_cost[t] = q;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double best = cand_c[0], bcost = cand_cost[0];
            for (int t = 1; t <= nl; ++t) {
                const double q = cand_cost[t];
                if (q >= 0.0 && (q < bcost || (q == bcost && cand_c[t] > best))) {
                    best = cand_c[t];
                    bcost = q;
                }
            }
            model_c = on ? best : -1.0;
        }
        __syncthreads();
    }
    int k = threadIdx.x;
    for (; k < ninst; k += blockDim.x) {
        Inst& s = inst[k];
        int64_t nk = s.off1 - s.off0;
        int64_t cap = kCellsPerParticle * nk > 1 ? kCellsPerParticle * nk : 1;
        double R = kernel_radius(s.kernel);
        // h_max, h_min and the box were reduced by this launch's atomics
        unsigned long long rv[kRed];
        for (int q = 0; q < kRed; ++q) rv[q] = ld_dev(&s.red[q]);
        s.hmax_bits = rv[0];
        const double hm = bitsd(rv[0]);
        const double hn = rv[1] ? bitsd(~rv[1]) : 0.0;
        // an open domain's grid spans the particles' box at this evaluation
        if (!s.periodic && nk > 0) {
            for (int a = 0; a < s.dim; ++a) {
                const double mn = ord_value(~rv[5 + a]), mx = ord_value(rv[2 + a]);
                const double pad = box_pad(mn, mx);
                s.hi[a] = mx + pad;
                s.lo[a] = mn - pad;
                s.L[a] = s.hi[a] - s.lo[a];
                s.halfL[a] = 0.5 * s.L[a];
            }
        }
        // where h varies by 2x or more the cells are split below R h_max (at
        // most 4 per axis), so the small-h particles do not scan cells sized
        // for the largest h
        double split = cell_split_dim(hm, hn, s.dim);
        double c = R * hm / split;
        // a single 3-D instance: cells sized by the scan-cost model over its
        // h levels (while h spans less than the ring), at most 8 per R h_max
        const bool model = hist && s.dim == 3 && hn > 0.0 && hm >= hn && h_level(hm) - h_level(hn) < kHistSlots;
        if (model) {
            c = ::fmax(model_c, R * hm / 8.0);
            split = R * hm / c;
        }
        double lmax = 0.0;
        for (int a = 0; a < s.dim; ++a) lmax = ::fmax(lmax, s.L[a]);
        if (!(c > 0.0) || !(c <= lmax)) {  // h_max of 0, NaN or above the box: one cell
            c = lmax;
            split = 1.0;
        }
        for (int guard = 0; guard < 80; ++guard) {
            int64_t prod = 1;
            for (int a = 0; a < 3; ++a) {
                if (a < s.dim) {
                    double nf = floor(s.L[a] / c);
                    int na = nf < 1.0 ? 1 : (nf > 1.0e9 ? 1000000000 : (int)nf);
                    s.n[a] = na;
                    prod *= na;
                    if (prod > cap) break;
                } else {
                    s.n[a] = 1;
                }
            }
            if (prod <= cap) break;
            // too many cells: first undo the split, then coarsen
            if (model) {
                c = 2.0 * c;
            } else if (split > 1.0) {
                split -= 1.0;
                c = R * hm / split;
            } else {
                c = 2.0 * c;
            }
        }
        for (int a = 0; a < 3; ++a) s.ceff[a] = a < s.dim ? s.L[a] / s.n[a] : 1.0;
        s.ncell = (int64_t)s.n[0] * s.n[1] * s.n[2];
        // above 1 only where h varies (the lists then prune by cell h): with
        // the model, where the h_max stencil is wider than half cells' 2
        s.split = cell_split(hm, hn) > 1.0 ? (int)split : 1;
        if (model) s.split = ceil(R * hm / c) > 2.0 ? 2 : 1;
        stencil_from_hmax(s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t base = 0;
        for (int q = 0; q < ninst; ++q) {
            inst[q].cell_base = base;
            base += inst[q].ncell;
        }
        ctl->total_cells = base;
        ctl->nbig = 0;
        int split = 0;
        for (int q = 0; q < ninst; ++q) split |= inst[q].split > 1;
        ctl->any_split = split;
    }
}

// Maxima of the first NR of the kRed patterns v over a block of kBlock
// threads, per run of equal instances: waves reduce by shuffles, then thread
// 0 calls emit(k, m, whole) once per run (m: the maxima; entries from NR on
// are zero), whole telling that the run is the whole block.  kcur < 0: the
// thread holds nothing (v all zero).  Periodic domains need only h_max and
// h_min (NR = 2), open ones the box too (NR = kRed).
template <int NR, class Emit>
__device__ __forceinline__ void block_max_runs(int kcur, unsigned long long (&v)[kRed], Emit emit) {
    constexpr int NW = kBlock / kWave;
    __shared__ unsigned long long wv[NW][kRed];
    __shared__ int wk[NW];
    __shared__ int wmixed;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (threadIdx.x == 0) wmixed = 0;
    int k0 = kcur;
    for (int off = 32; off > 0; off >>= 1) k0 = max(k0, __shfl_xor(k0, off));
    const bool uniform = __all(kcur == k0 || kcur < 0);
    __syncthreads();
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                const unsigned long long o = __shfl_xor(v[q], off);
                v[q] = o > v[q] ? o : v[q];
            }
        }
    } else {
        if (kcur >= 0) emit(kcur, v, false);  // rare: a wave straddling instances
        if (lane == 0) wmixed = 1;
    }
    if (lane == 0) {
        for (int q = 0; q < NR; ++q) wv[wave][q] = v[q];
        wk[wave] = uniform ? k0 : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < nw;) {
            const int k = wk[w];
            unsigned long long m[kRed] = {};
            for (int q = 0; q < NR; ++q) m[q] = wv[w][q];
            int r = w + 1;
            for (; r < nw && wk[r] == k; ++r)
                for (int q = 0; q < NR; ++q) m[q] = wv[r][q] > m[q] ? wv[r][q] : m[q];
            if (k >= 0) emit(k, m, w == 0 && r == nw && !wmixed);
            w = r;
        }
    }
}

__device__ __forceinline__ unsigned long long run_heads(uint32_t key, bool valid) {
    const uint32_t prev = __shfl_up(key, 1);
    const int lane = threadIdx.x & (kWave - 1);
    return __ballot(valid && (lane == 0 || prev != key));
}

// one past the last lane of the run that holds lane `lane` (the next head
// above it, else one past the last valid lane)
__device__ __forceinline__ int run_end(unsigned long long heads, unsigned long long vmask, int lane) {
    const unsigned long long above = lane == kWave - 1 ? 0ull : heads & (~0ull << (lane + 1));
    return above ? __ffsll((long long)above) - 1 : kWave - __clzll(vmask);
}

// per-cell largest h of cell-ordered particles (lanes of a wave in sorted
// order, key their cell): each run of equal keys issues one atomicMax of its
// largest h's bits (positive doubles order like their bits; h that is not
// positive counts as 0, as fmax(h_i, h) then never exceeds h_i either).
// Lanes with use == false contribute nothing.
__device__ __forceinline__ void cell_hmax(unsigned long long* hcell, uint32_t key, double h, bool use) {
    if (!__ballot(use)) return;
    const unsigned long long heads = run_heads(key, use);
    const unsigned long long vmask = __ballot(use);
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long v = use && h > 0.0 ? dbits(h) : 0ull;
    const int end = use ? run_end(heads, vmask, lane) : lane;
    for (int off = 1; off < kWave; off <<= 1) {
        const unsigned long long o = __shfl_down(v, off);
        if (lane + off < end && o > v) v = o;
    }
    if (use && ((heads >> lane) & 1ull) && v) atomicMax(&hcell[key], v);
}

// ---------------------------------------------------------------------------
// XCD-aware tiles
// ---------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Gather kernels launch 8 * ceil(tiles / 8) blocks and map block b to tile
// (b % 8) * ceil(tiles / 8) + b / 8, so every XCD walks one contiguous range of
// cell-ordered particles and its L2 keeps the neighbours they share.  Blocks
// mapped beyond the last tile have nothing to do.
constexpr int kXcds = 8;

__host__ __device__ inline unsigned xcd_grid(int64_t tiles) {
    return (unsigned)(((tiles + kXcds - 1) / kXcds) * kXcds);
}

__device__ __forceinline__ int64_t xcd_tile(int64_t tiles) {
    const int64_t per = (tiles + kXcds - 1) / kXcds;
    const int64_t b = blockIdx.x;
    return (b % kXcds) * per + b / kXcds;
}

// ---------------------------------------------------------------------------
// Persistent launches of the pair kernels
// ---------------------------------------------------------------------------
// A pair kernel covers a sequence of tiles, tile q being the B consecutive
// sorted particles (or active-list entries) a block of B threads owns; a
// tile may hold the end of one instance and the start of the next, and every
// particle is evaluated with its own instance's parameters, so the uneven
// instances of an ensemble share the launches.  A plain launch has one block
// per tile (xcd_tile).  A persistent launch has as many blocks as the GPU
// holds resident at once, and each pulls tiles from work queues until all are
// empty: first the queue of its XCD (the contiguous eighth of the tiles
// xcd_tile would give that XCD, so neighbouring tiles still share an L2),
// then the others in turn.  The queue heads (kXcds counters on lines of their
// own) are zeroed by the launch's last workgroup, so a persistent launch
// needs no clearing launch before it and replays from a captured graph.
constexpr int kQueueInts = kXcds * kArriveStride;

// the block's next tile, the same in every thread; -1 once every queue is
// empty.  polls (thread 0's, start 0): queues found empty so far.
__device__ __forceinline__ int64_t pull_tile(int32_t* queue, int64_t tiles, int& polls) {
    __shared__ int64_t s_tile;
    if (threadIdx.x == 0) {
        const int64_t per = (tiles + kXcds - 1) / kXcds;
        const int home = (int)(blockIdx.x % kXcds);
        int64_t got = -1;
        while (polls < kXcds) {
            const int x = (home + polls) % kXcds;
            const int64_t lo = (int64_t)x * per, hi = tiles < lo + per ? tiles : lo + per;
            if (lo < hi) {
                const int64_t k = atomicAdd(&queue[x * kArriveStride], 1);
                if (lo + k < hi) {
                    got = lo + k;
                    break;
                }
            }
            ++polls;  // this queue is empty
        }
        s_tile = got;
    }
    __syncthreads();
    const int64_t t = s_tile;
    __syncthreads();  // s_tile is written again by the next call
    return t;
}

// by the last workgroup of a persistent launch, every block having pulled
// its last tile
__device__ __forceinline__ void reset_queues(int32_t* queue) {
    if (threadIdx.x < kXcds) atomicExch(&queue[threadIdx.x * kArriveStride], 0);
}

// ---------------------------------------------------------------------------
// stencil traversal
// ---------------------------------------------------------------------------
// cstart[c] / cend[c] hold, for every cell c (empty cells included), the
// sorted positions [first key >= c, first key > c), so the x-adjacent cells of
// one stencil row form ONE contiguous range [cstart[first], cend[last]).
// for_each_row_range calls f(a, b, sx, sy, sz) for each such range in cell
// order (z, y outer, x inner), splitting a row in two where it wraps a
// periodic boundary; (sx, sy, sz) is the range's image shift (0 or +-L per
// axis: +L where the stencil wrapped below 0, -L above n), which is the
// minimum image of each of its candidates when s.range_shift is set.
// stencil span of the cells c0..c1 of one axis; for c0 == c1 exactly the span
// for_each_row_range forms
__device__ __forceinline__ void axis_span(int c0, int c1, int sr, int na, int periodic, int& lo, int& hi) {
    if (periodic && 2 * sr + 1 + (c1 - c0) >= na) { lo = 0; hi = na - 1; }
    else if (periodic) { lo = c0 - sr; hi = c1 + sr; }
    else { lo = max(c0 - sr, 0); hi = min(c1 + sr, na - 1); }
}

// the x pieces of [lo, hi] (-n0 <= lo <= hi < 2 n0) after periodic wrapping
// (at most two), with their image shifts: +L for cells wrapped from below 0,
// -L for cells wrapped from n0 and above
__device__ __forceinline__ int x_pieces(int lo, int hi, int n0, int* xa, int* xb) {
    if (hi < 0) { xa[0] = lo + n0; xb[0] = hi + n0; return 1; }
    if (lo >= n0) { xa[0] = lo - n0; xb[0] = hi - n0; return 1; }
    if (lo < 0) { xa[0] = lo + n0; xb[0] = n0 - 1; xa[1] = 0; xb[1] = hi; return 2; }
    if (hi >= n0) { xa[0] = lo; xb[0] = n0 - 1; xa[1] = 0; xb[1] = hi - n0; return 2; }
    xa[0] = lo; xb[0] = hi; return 1;
}

__device__ __forceinline__ void x_shifts(int lo, int hi, int n0, double L, double* sx) {
    sx[0] = lo < 0 ? L : (lo >= n0 ? -L : 0.0);
    sx[1] = hi >= n0 ? -L : 0.0;
}

// Geometric pruning of one particle's stencil.  A cell is passed over when
// its box, in the frame of the stencil's image shifts, lies farther from the
// particle than R max(h_i, h_c), h_c the cell's largest h where per-cell
// values are given (hcell) and otherwise a bound hb on every h.  Such a cell
// holds no pair of the particle, so every sum and list comes out exactly as
// from the full scan.  Whole stencil rows are dropped (by the row's largest h
// where per-cell values are given, before any cell of it is read) and row
// ends trimmed, so a row stays one contiguous range.  The first and last cells of a
// periodic axis can hold particles binned there by clamping (outside the box
// before the first wrap), so their gap along that axis counts as zero; gaps
// are shrunk by a margin far above the rounding of the binning and of the
// distance arithmetic.  Only for grids with range shifts (the unwrapped
// stencil frame is then the minimum image).
struct Prune {
    double x[3], lo[3], c[3], m[3];
    int ci[3], n[3];
    int per;
    double R, hi, hb;
    const unsigned long long* hcell;
    const unsigned long long* hrow;  // with hcell: the largest h of each row, at the row's first cell

    // gap along axis a between the particle and stencil cell q (unwrapped)
    __device__ __forceinline__ double gap(int a, int q) const {
        if (q == ci[a]) return 0.0;
        if (per) {
            const int w = q < 0 ? q + n[a] : (q >= n[a] ? q - n[a] : q);
            if (w == 0 || w == n[a] - 1) return 0.0;
        }
        const double g = q > ci[a] ? (lo[a] + q * c[a]) - x[a] : x[a] - (lo[a] + (q + 1) * c[a]);
        return g > m[a] ? g - m[a] : 0.0;
    }
    __device__ __forceinline__ bool far(double g2, double h) const {
        const double r = R * ::fmax(hi, h);
        return g2 > r * r * (1.0 + 1.0e-12);
    }
    // trims [xl, xh] of the stencil row starting at global cell `row`;
    // false when nothing is left
    __device__ __forceinline__ bool trim(double g2yz, int64_t row, int& xl, int& xh) const {
        // without per-cell values a row's largest h (when given) bounds the
        // h of every neighbour in it together with hb
        const double hr = hrow ? bitsd(hrow[row]) : hb;
        const double hrb = hcell ? hr : ::fmin(hr, hb);
        if (far(g2yz, hrb)) return false;
        if (!hcell) {
            // start from the span the reach leaves along x, widened by a
            // cell each side (the gap tests below trim it exactly: a cell's
            // gap grows away from the particle), keeping the edge cells of
            // a periodic axis, whose gap counts as zero
            const double r = R * ::fmax(hi, hrb);
            const double rx = ::sqrt(::fmax(r * r * (1.0 + 1.0e-12) - g2yz, 0.0));
            const int a = (int)::floor((x[0] - rx - lo[0]) / c[0]) - 1;
            const int b = (int)::floor((x[0] + rx - lo[0]) / c[0]) + 1;
            int nl = xl > a ? xl : a, nh = xh < b ? xh : b;
            if (per) {
                for (int q : {-1, 0, n[0] - 1, n[0]})
                    if (q >= xl && q <= xh) {
                        nl = q < nl ? q : nl;
                        nh = q > nh ? q : nh;
                    }
            }
            xl = nl;
            xh = nh;
        }
        auto skip = [&](int q) {
            const double g = gap(0, q);
            if (!hcell) return far(g2yz + g * g, hrb);
            const int w = q < 0 ? q + n[0] : (q >= n[0] ? q - n[0] : q);
            return far(g2yz + g * g, bitsd(hcell[row + w]));
        };
        while (xl <= xh && skip(xl)) ++xl;
        while (xh >= xl && skip(xh)) --xh;
        return xl <= xh;
    }
};

// hb: the bound on every h for cells without values (0 for a reach of R h_i);
// hrow (the largest h of each row, at the row's first cell) drops and trims
// rows by min(the row's largest h, hb) where no per-cell values are given
template <int DIM>
__device__ __forceinline__ void prune_init(Prune& p, const Inst& s, int64_t key, double x, double y, double z,
                                           double R, double hi, double hb, const unsigned long long* hcell,
                                           const unsigned long long* hrow) {
    const int64_t local = key - s.cell_base;
    p.ci[0] = (int)(local % s.n[0]);
    p.ci[1] = (int)((local / s.n[0]) % s.n[1]);
    p.ci[2] = DIM == 3 ? (int)(local / ((int64_t)s.n[0] * s.n[1])) : 0;
    p.x[0] = x;
    p.x[1] = y;
    p.x[2] = z;
    for (int a = 0; a < 3; ++a) {
        p.lo[a] = s.lo[a];
        p.c[a] = s.ceff[a];
        p.n[a] = s.n[a];
        p.m[a] = 1.0e-9 * s.ceff[a] + 1.0e-12 * (::fabs(s.lo[a]) + ::fabs(s.L[a]));
    }
    p.per = s.periodic;
    p.R = R;
    p.hi = hi;
    p.hb = hb;
    p.hcell = hcell;
    p.hrow = hrow;
}

template <int DIM, class F>
__device__ __forceinline__ void for_each_row_range(const Inst& s, int64_t key, const int32_t* __restrict__ cstart,
                                                   const int32_t* __restrict__ cend, F&& f,
                                                   const int* srad = nullptr, const Prune* pr = nullptr) {
    const int n0 = s.n[0], n1 = s.n[1], n2 = s.n[2];
    const int64_t local = key - s.cell_base;
    const int cx = (int)(local % n0);
    const int cy = (int)((local / n0) % n1);
    const int cz = DIM == 3 ? (int)(local / ((int64_t)n0 * n1)) : 0;
    int lo[3], hi[3];
    const int c[3] = {cx, cy, cz};
    for (int a = 0; a < 3; ++a) {
        const int na = s.n[a], sr = srad ? srad[a] : s.srad[a];
        if (a >= DIM) { lo[a] = 0; hi[a] = 0; }
        else if (s.periodic && 2 * sr + 1 >= na) { lo[a] = 0; hi[a] = na - 1; }
        else if (s.periodic) { lo[a] = c[a] - sr; hi[a] = c[a] + sr; }
        else { lo[a] = max(c[a] - sr, 0); hi[a] = min(c[a] + sr, na - 1); }
    }
    // x pieces of the full row; a pruned row recomputes its own
    int xa0[2], xb0[2];
    double sx0[2];
    const int nseg0 = x_pieces(lo[0], hi[0], n0, xa0, xb0);
    x_shifts(lo[0], hi[0], n0, s.L[0], sx0);
    if (!pr) {
        // every row has the same x pieces: the next row's bounds are loaded
        // while the current row's candidates are scanned, so a row costs no
        // extra memory round trip before its first candidate
        auto row_of = [&](int qy, int qz) {
            const int wy = ((qy % n1) + n1) % n1;
            const int wz = DIM == 3 ? ((qz % n2) + n2) % n2 : 0;
            return s.cell_base + (int64_t)n0 * (wy + (int64_t)n1 * wz);
        };
        int qy = lo[1], qz = lo[2];
        int64_t row = row_of(qy, qz);
        int a0 = cstart[row + xa0[0]], b0 = cend[row + xb0[0]];
        int a1 = nseg0 == 2 ? cstart[row + xa0[1]] : 0, b1 = nseg0 == 2 ? cend[row + xb0[1]] : 0;
        while (true) {
            int ny = qy + 1, nz = qz;
            if (ny > hi[1]) {
                ny = lo[1];
                ++nz;
            }
            const bool more = nz <= hi[2];
            int na0 = 0, nb0 = 0, na1 = 0, nb1 = 0;
            if (more) {
                const int64_t nrow = row_of(ny, nz);
                na0 = cstart[nrow + xa0[0]];
                nb0 = cend[nrow + xb0[0]];
                if (nseg0 == 2) {
                    na1 = cstart[nrow + xa0[1]];
                    nb1 = cend[nrow + xb0[1]];
                }
            }
            const double sz = DIM == 3 ? (qz < 0 ? s.L[2] : (qz >= n2 ? -s.L[2] : 0.0)) : 0.0;
            const double sy = qy < 0 ? s.L[1] : (qy >= n1 ? -s.L[1] : 0.0);
            f(a0, b0, sx0[0], sy, sz);
            if (nseg0 == 2) f(a1, b1, sx0[1], sy, sz);
            if (!more) break;
            qy = ny;
            qz = nz;
            a0 = na0;
            b0 = nb0;
            a1 = na1;
            b1 = nb1;
        }
        return;
    }
    for (int qz = lo[2]; qz <= hi[2]; ++qz) {
        const int wz = DIM == 3 ? ((qz % n2) + n2) % n2 : 0;
        const double sz = DIM == 3 ? (qz < 0 ? s.L[2] : (qz >= n2 ? -s.L[2] : 0.0)) : 0.0;
        const double gz = pr && DIM == 3 ? pr->gap(2, qz) : 0.0;
        for (int qy = lo[1]; qy <= hi[1]; ++qy) {
            const int wy = ((qy % n1) + n1) % n1;
            const double sy = qy < 0 ? s.L[1] : (qy >= n1 ? -s.L[1] : 0.0);
            const int64_t row = s.cell_base + (int64_t)n0 * (wy + (int64_t)n1 * wz);
            int xa[2] = {xa0[0], xa0[1]}, xb[2] = {xb0[0], xb0[1]};
            double sx[2] = {sx0[0], sx0[1]};
            int nseg = nseg0;
            if (pr) {
                int xl = lo[0], xh = hi[0];
                const double gy = pr->gap(1, qy);
                if (!pr->trim(gy * gy + gz * gz, row, xl, xh)) continue;
                nseg = x_pieces(xl, xh, n0, xa, xb);
                x_shifts(xl, xh, n0, s.L[0], sx);
            }
            f(cstart[row + xa[0]], cend[row + xb[0]], sx[0], sy, sz);
            if (nseg == 2) f(cstart[row + xa[1]], cend[row + xb[1]], sx[1], sy, sz);
        }
    }
}

// for_each_row_range for a particle at (x, y, z) that needs only the
// candidates within sqrt(rt2) of it (with range shifts): rows farther than
// that in y, z are passed over and the others trimmed along x to the cells
// the radius reaches, by a margin far above the rounding of the binning and
// of the arithmetic, keeping the edge cells of a periodic axis (they can hold
// particles binned by clamping).  Every candidate within the radius is
// visited, in the same order as by the full stencil.
template <int DIM, class F>
__device__ __forceinline__ void for_each_row_range_trim(const Inst& s, int64_t key, double x, double y, double z,
                                                        double rt2, const int32_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ cend, F&& f, const int* srad) {
    const int n0 = s.n[0], n1 = s.n[1], n2 = s.n[2];
    const int64_t local = key - s.cell_base;
    const int cx = (int)(local % n0);
    const int cy = (int)((local / n0) % n1);
    const int cz = DIM == 3 ? (int)(local / ((int64_t)n0 * n1)) : 0;
    int lo[3], hi[3];
    const int c[3] = {cx, cy, cz};
    for (int a = 0; a < 3; ++a) {
        const int na = s.n[a], sr = srad[a];
        if (a >= DIM) { lo[a] = 0; hi[a] = 0; }
        else if (s.periodic && 2 * sr + 1 >= na) { lo[a] = 0; hi[a] = na - 1; }
        else if (s.periodic) { lo[a] = c[a] - sr; hi[a] = c[a] + sr; }
        else { lo[a] = max(c[a] - sr, 0); hi[a] = min(c[a] + sr, na - 1); }
    }
    const int per = s.periodic;
    // gap between the particle and unwrapped cell q of axis a
    auto gap = [&](int a, int q, double xa) {
        if (q == c[a]) return 0.0;
        if (per) {
            const int w = q < 0 ? q + s.n[a] : (q >= s.n[a] ? q - s.n[a] : q);
            if (w == 0 || w == s.n[a] - 1) return 0.0;
        }
        const double g = q > c[a] ? (s.lo[a] + q * s.ceff[a]) - xa : xa - (s.lo[a] + (q + 1) * s.ceff[a]);
        const double m = 1.0e-9 * s.ceff[a] + 1.0e-12 * (::fabs(s.lo[a]) + ::fabs(s.L[a]));
        return g > m ? g - m : 0.0;
    };
    const double r2 = rt2 * (1.0 + 1.0e-12);
    const double inv0 = 1.0 / s.ceff[0];
    const double xr = (x - s.lo[0]) * inv0;  // x in cells
    for (int qz = lo[2]; qz <= hi[2]; ++qz) {
        const int wz = DIM == 3 ? ((qz % n2) + n2) % n2 : 0;
        const double sz = DIM == 3 ? (qz < 0 ? s.L[2] : (qz >= n2 ? -s.L[2] : 0.0)) : 0.0;
        const double gz = DIM == 3 ? gap(2, qz, z) : 0.0;
        for (int qy = lo[1]; qy <= hi[1]; ++qy) {
            const double gy = gap(1, qy, y);
            const double g2 = gy * gy + gz * gz;
            if (g2 > r2) continue;
            const double wc = ::sqrt(r2 - g2) * inv0 + 1.0e-9;  // half-width in cells, widened
            int xl = (int)::floor(xr - wc), xh = (int)::floor(xr + wc);
            xl = xl > lo[0] ? xl : lo[0];
            xh = xh < hi[0] ? xh : hi[0];
            if (per) {
                for (int q : {-1, 0, n0 - 1, n0})
                    if (q >= lo[0] && q <= hi[0]) {
                        xl = q < xl ? q : xl;
                        xh = q > xh ? q : xh;
                    }
            }
            if (xl > xh) continue;
            const int wy = ((qy % n1) + n1) % n1;
            const double sy = qy < 0 ? s.L[1] : (qy >= n1 ? -s.L[1] : 0.0);
            const int64_t row = s.cell_base + (int64_t)n0 * (wy + (int64_t)n1 * wz);
            int xa[2], xb[2];
            double sx[2];
            const int nseg = x_pieces(xl, xh, n0, xa, xb);
            x_shifts(xl, xh, n0, s.L[0], sx);
            f(cstart[row + xa[0]], cend[row + xb[0]], sx[0], sy, sz);
            if (nseg == 2) f(cstart[row + xa[1]], cend[row + xb[1]], sx[1], sy, sz);
        }
    }
}

// f(j, rec) for j in [a, b), four independent loads in flight per thread;
// the last one to three candidates load together too (a stencil row holds a
// few candidates in 3-D, so one at a time cost a memory round trip each)
template <class T, class F>
__device__ __forceinline__ void scan_range4(const T* __restrict__ rec, int a, int b, F&& f) {
    int j = a;
    for (; j + 4 <= b; j += 4) {
        const T o0 = rec[j], o1 = rec[j + 1], o2 = rec[j + 2], o3 = rec[j + 3];
        f(j, o0);
        f(j + 1, o1);
        f(j + 2, o2);
        f(j + 3, o3);
    }
    if (j < b) {
        const int r = b - j;  // 1 to 3; the loads past the end repeat rec[j]
        const T o0 = rec[j], o1 = rec[r > 1 ? j + 1 : j], o2 = rec[r > 2 ? j + 2 : j];
        f(j, o0);
        if (r > 1) f(j + 1, o1);
        if (r > 2) f(j + 2, o2);
    }
}

// f(j, rec) for j = a + g, a + g + G, ... < b: lane g of a group of G lanes
// sharing one destination, four loads in flight per lane
template <class T, class F>
__device__ __forceinline__ void scan_range4_strided(const T* __restrict__ rec, int a, int b, int g, int G, F&& f) {
    int j = a + g;
    for (; j + 3 * G < b; j += 4 * G) {
        const T o0 = rec[j], o1 = rec[j + G], o2 = rec[j + 2 * G], o3 = rec[j + 3 * G];
        f(j, o0);
        f(j + G, o1);
        f(j + 2 * G, o2);
        f(j + 3 * G, o3);
    }
    for (; j < b; j += G) f(j, rec[j]);
}

// ---------------------------------------------------------------------------
// block staging of candidate records
// ---------------------------------------------------------------------------
// A block's particles are consecutive in cell order, so they sit in one or two
// cell rows (y, z fixed), and every stencil range of theirs lies inside a few
// contiguous runs of sorted records: for each of the block's rows, each
// stencil row offset and each periodic x piece, the run from the row's first
// x - s to its last x + s.  Wave 0 lists the runs (one per lane), the block
// copies them into LDS with coalesced loads, and each thread then scans its
// own ranges out of LDS.  A range is read from LDS only when a staged run
// contains it in sorted-index space, so the copy always holds exactly the
// records the range names; anything else (more rows, wide stencils, an
// instance boundary, too many records) is read from global memory as before.
constexpr int kStageCap = 1024;     // staged 32-byte records per block of kBlock threads (32 KB)
constexpr int kStageSegs = 2;       // cell rows a staged group may span

// runs staged for one group of particles: a[k] <= j < b[k] sits at off[k] + j - a[k]
template <int SLOTS>
struct StageRuns {
    int on;
    int nslot;   // runs in use: segments x nz x ny x 2 pieces
    int ny, nz;  // stencil rows per group row (y, z)
    int64_t r0;  // first cell row (y + n1 z) of the group
    int a[SLOTS], b[SLOTS], off[SLOTS + 1];
};
using StageHdr = StageRuns<kWave>;

// Plans the runs of the group of particles [t0, t0 + count) (clipped to n).
// Called by all 64 lanes of one wave; lane k (< SLOTS) computes run k.  The
// plan is off when the group straddles instances, spans more than kStageSegs
// cell rows, needs more than SLOTS runs or more than cap records.
template <int DIM, int SLOTS>
__device__ void stage_plan(StageRuns<SLOTS>& H, const uint32_t* __restrict__ skey,
                           const int32_t* __restrict__ cstart, const int32_t* __restrict__ cend,
                           const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of, int64_t t0,
                           int64_t count, int64_t n, int cap) {
    static_assert(SLOTS <= kWave, "one run per lane");
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t tl = min(t0 + count, n) - 1;
    const int k = inst_of[t0];
    const Inst& s = inst[k];
    const int n0 = s.n[0], n1 = s.n[1], n2 = s.n[2];
    const int64_t l0 = (int64_t)skey[t0] - s.cell_base, l1 = (int64_t)skey[tl] - s.cell_base;
    const int64_t r0 = l0 / n0, r1 = l1 / n0;
    const int ny = min(2 * s.srad[1] + 1, n1);
    const int nz = DIM == 3 ? min(2 * s.srad[2] + 1, n2) : 1;
    const int64_t nseg = r1 - r0 + 1;
    const bool on = inst_of[tl] == k && nseg <= kStageSegs && ny <= SLOTS && nz <= SLOTS &&
                    nseg * ny * nz * 2 <= SLOTS;
    const int nslot = on ? (int)(nseg * ny * nz * 2) : 0;
    int a = 0, len = 0;
    if (lane < nslot) {
        const int p = lane & 1;
        int rest = lane >> 1;
        const int iy = rest % ny;
        rest /= ny;
        const int iz = rest % nz;
        const int seg = rest / nz;
        const int64_t row = r0 + seg;
        const int cy = (int)(row % n1), cz = DIM == 3 ? (int)(row / n1) : 0;
        const int xa = seg == 0 ? (int)(l0 % n0) : 0;
        const int xb = seg == nseg - 1 ? (int)(l1 % n0) : n0 - 1;
        int lo0, hi0, lo1, hi1, lo2 = 0, hi2 = 0;
        axis_span(xa, xb, s.srad[0], n0, s.periodic, lo0, hi0);
        axis_span(cy, cy, s.srad[1], n1, s.periodic, lo1, hi1);
        if (DIM == 3) axis_span(cz, cz, s.srad[2], n2, s.periodic, lo2, hi2);
        int pa[2], pb[2];
        const int np = x_pieces(lo0, hi0, n0, pa, pb);
        const int qy = lo1 + iy, qz = lo2 + iz;
        if (qy <= hi1 && qz <= hi2 && p < np) {
            const int wy = ((qy % n1) + n1) % n1;
            const int wz = DIM == 3 ? ((qz % n2) + n2) % n2 : 0;
            const int64_t rowc = s.cell_base + (int64_t)n0 * (wy + (int64_t)n1 * wz);
            a = cstart[rowc + pa[p]];
            len = cend[rowc + pb[p]] - a;
        }
    }
    int incl = len;
    for (int d = 1; d < kWave; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    const int total = __shfl(incl, kWave - 1);
    if (lane < SLOTS) {
        H.a[lane] = a;
        H.b[lane] = a + len;
        H.off[lane] = incl - len;
    }
    if (lane == 0) {
        H.on = on && total <= cap;
        H.nslot = nslot;
        H.ny = ny;
        H.nz = nz;
        H.r0 = r0;
        H.off[SLOTS] = total;
    }
}

// all threads of the block: copy the planned runs into LDS
template <class T, int SLOTS>
__device__ __forceinline__ void stage_fill(const StageRuns<SLOTS>& H, T* stage, const T* __restrict__ src) {
    if (!H.on) return;
    const int total = H.off[SLOTS];
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        int lo = 0, hi = SLOTS;  // off[lo] <= idx < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (H.off[mid] <= idx) lo = mid;
            else hi = mid;
        }
        stage[idx] = src[H.a[lo] + (idx - H.off[lo])];
    }
}

// for_each_row_range over the staged copy: f(j, rec_j, sx, sy, sz) for every
// candidate j (with its range's image shift, as for_each_row_range), from LDS
// where a staged run holds the range (pruned ranges lie inside the full ones,
// so a run that holds a full range holds its pruned part too)
template <int DIM, class T, class F>
__device__ __forceinline__ void for_each_candidate(const Inst& s, int64_t key, const int32_t* __restrict__ cstart,
                                                   const int32_t* __restrict__ cend, const StageHdr& H,
                                                   const T* stage, const T* __restrict__ glob, F&& f,
                                                   const Prune* pr = nullptr, const int* srad = nullptr) {
    const int n0 = s.n[0], n1 = s.n[1], n2 = s.n[2];
    const int64_t local = key - s.cell_base;
    const int cx = (int)(local % n0);
    const int cy = (int)((local / n0) % n1);
    const int cz = DIM == 3 ? (int)(local / ((int64_t)n0 * n1)) : 0;
    int lo0, hi0, lo1, hi1, lo2 = 0, hi2 = 0;
    const int* sr = srad ? srad : s.srad;  // the instance's stencil, or a particle's narrower one
    axis_span(cx, cx, sr[0], n0, s.periodic, lo0, hi0);
    axis_span(cy, cy, sr[1], n1, s.periodic, lo1, hi1);
    if (DIM == 3) axis_span(cz, cz, sr[2], n2, s.periodic, lo2, hi2);
    int xa0[2], xb0[2];
    double sx0[2];
    const int nseg0 = x_pieces(lo0, hi0, n0, xa0, xb0);
    x_shifts(lo0, hi0, n0, s.L[0], sx0);
    const int64_t seg = local / n0 - H.r0;
    const bool staged = H.on && seg >= 0 && seg < kStageSegs;
    for (int qz = lo2; qz <= hi2; ++qz) {
        const int wz = DIM == 3 ? ((qz % n2) + n2) % n2 : 0;
        const double sz = DIM == 3 ? (qz < 0 ? s.L[2] : (qz >= n2 ? -s.L[2] : 0.0)) : 0.0;
        const double gz = pr && DIM == 3 ? pr->gap(2, qz) : 0.0;
        for (int qy = lo1; qy <= hi1; ++qy) {
            const int wy = ((qy % n1) + n1) % n1;
            const double sy = qy < 0 ? s.L[1] : (qy >= n1 ? -s.L[1] : 0.0);
            const int64_t row = s.cell_base + (int64_t)n0 * (wy + (int64_t)n1 * wz);
            int xa[2] = {xa0[0], xa0[1]}, xb[2] = {xb0[0], xb0[1]};
            double sxg[2] = {sx0[0], sx0[1]};
            int nseg = nseg0;
            if (pr) {
                int xl = lo0, xh = hi0;
                const double gy = pr->gap(1, qy);
                if (!pr->trim(gy * gy + gz * gz, row, xl, xh)) continue;
                nseg = x_pieces(xl, xh, n0, xa, xb);
                x_shifts(xl, xh, n0, s.L[0], sxg);
            }
            const int base = (((int)seg * H.nz + (qz - lo2)) * H.ny + (qy - lo1)) * 2;
            for (int g = 0; g < nseg; ++g) {
                const int a = cstart[row + xa[g]], b = cend[row + xb[g]];
                int k = -1;
                if (staged && base + 1 < H.nslot) {
                    if (H.a[base] <= a && b <= H.b[base]) k = base;
                    else if (H.a[base + 1] <= a && b <= H.b[base + 1]) k = base + 1;
                }
                const double sx = sxg[g];
                auto fs = [&](int j, const T& o) { f(j, o, sx, sy, sz); };
                if (k >= 0) {
                    const int delta = H.off[k] - H.a[k];
                    int j = a;
                    for (; j + 4 <= b; j += 4) {
                        const T o0 = stage[j + delta], o1 = stage[j + 1 + delta];
                        const T o2 = stage[j + 2 + delta], o3 = stage[j + 3 + delta];
                        fs(j, o0);
                        fs(j + 1, o1);
                        fs(j + 2, o2);
                        fs(j + 3, o3);
                    }
                    for (; j < b; ++j) fs(j, stage[j + delta]);
                } else {
                    scan_range4(glob, a, b, fs);
                }
            }
        }
    }
}

// time step after a step: the pending min dt_adapt (adaptive) or the current
// dt, clipped to tf - t (solver.py:61)
__host__ __device__ inline double next_dt(const Inst& s, double t, double dt_now, double pending) {
    double dt = s.adaptive ? pending : dt_now;
    if (t + dt > s.tf) dt = s.tf - t;
    return dt;
}

// ---------------------------------------------------------------------------
// SPH kernel shapes w(q), w'(q) (normalisation applied by the caller)
// ---------------------------------------------------------------------------

__host__ __device__ inline double kernel_sigma(int k, int dim) {
    const double pi = 3.14159265358979323846;
    switch (k) {
    case SPHMHD_CUBIC: return dim == 2 ? 10.0 / (7.0 * pi) : 1.0 / pi;
    case SPHMHD_QUINTIC: return dim == 2 ? 7.0 / (478.0 * pi) : 1.0 / (120.0 * pi);
    case SPHMHD_WENDLAND_C2: return dim == 2 ? 7.0 / (4.0 * pi) : 21.0 / (16.0 * pi);
    case SPHMHD_WENDLAND_C4: return dim == 2 ? 9.0 / (4.0 * pi) : 495.0 / (256.0 * pi);
    case SPHMHD_WENDLAND_C6: return dim == 2 ? 78.0 / (28.0 * pi) : 1365.0 / (512.0 * pi);
    case SPHMHD_GAUSSIAN:
    case SPHMHD_SUPER_GAUSSIAN: return dim == 2 ? 1.0 / pi : 1.0 / (pi * 1.7724538509055160273);
    }
    return 0.0;
}

// W and dW/dq of the kernel shapes (w(q) scaled by sigma / h^d elsewhere);
// generic in the scalar type so the tangent step differentiates the same
// arithmetic
template <int K, int DIM, class T>
__host__ __device__ __forceinline__ void shape(T q, T& w, T& dw) {
    if constexpr (K == SPHMHD_CUBIC) {
        T t = fmax(2.0 - q, 0.0);
        if (q > 1.0) {
            w = 0.25 * t * t * t;
            dw = -0.75 * t * t;
        } else {
            w = 1.0 - 1.5 * q * q * (1.0 - 0.5 * q);
            dw = -3.0 * q * (1.0 - 0.75 * q);
        }
    } else if constexpr (K == SPHMHD_QUINTIC) {
        T t3 = fmax(3.0 - q, 0.0), t2 = fmax(2.0 - q, 0.0), t1 = fmax(1.0 - q, 0.0);
        T a3 = t3 * t3, a2 = t2 * t2, a1 = t1 * t1;
        T b3 = a3 * a3, b2 = a2 * a2, b1 = a1 * a1;
        w = b3 * t3 - 6.0 * (b2 * t2) + 15.0 * (b1 * t1);
        dw = -5.0 * b3 + 30.0 * b2 - 75.0 * b1;
    } else if constexpr (K == SPHMHD_WENDLAND_C2) {
        T t = fmax(1.0 - 0.5 * q, 0.0);
        T t3 = t * t * t;
        w = t3 * t * (2.0 * q + 1.0);
        dw = -5.0 * q * t3;
    } else if constexpr (K == SPHMHD_WENDLAND_C4) {
        T t = fmax(1.0 - 0.5 * q, 0.0);
        T t2 = t * t;
        T t5 = t2 * t2 * t;
        w = t5 * t * ((35.0 / 12.0) * q * q + 3.0 * q + 1.0);
        dw = -(14.0 / 3.0) * q * (1.0 + 2.5 * q) * t5;
    } else if constexpr (K == SPHMHD_WENDLAND_C6) {
        T t = fmax(1.0 - 0.5 * q, 0.0);
        T t2 = t * t;
        T t4 = t2 * t2;
        T t7 = t4 * t2 * t;
        w = t7 * t * (4.0 * q * q * q + 6.25 * q * q + 4.0 * q + 1.0);
        dw = -5.5 * q * t7 * (1.0 + 3.5 * q + 4.0 * q * q);
    } else if constexpr (K == SPHMHD_GAUSSIAN) {
        T g = q < 3.0 ? exp(-q * q) : T(0.0);
        w = g;
        dw = -2.0 * q * g;
    } else {
        T g = q < 3.0 ? exp(-q * q) : T(0.0);
        const double d2 = 0.5 * DIM;
        w = g * (d2 + 1.0 - q * q);
        dw = g * (-2.0 * q * (d2 + 1.0 - q * q) - 2.0 * q);
    }
}

template <int K, int DIM, class T>
__host__ __device__ __forceinline__ T shape_dw(T q) {
    T w, dw;
    shape<K, DIM>(q, w, dw);
    return dw;
}

// ---------------------------------------------------------------------------
// host-side launchers (defined in the .hip files)
// ---------------------------------------------------------------------------

struct Dev {
    // sizes
    int64_t n;
    int32_t ninst;
    int32_t dim, kernel;
    int32_t periodic;       // every instance's domain is periodic (they agree)
    int64_t cell_cap;
    int32_t nb_stride;      // neighbour-list capacity per particle
    int32_t pass;           // density pass index within the evaluation
    int32_t dens_trim = 1;  // 3-D first density pass trims its stencil rows to R h_i (same results)
    int32_t dens_fused = 1;  // the first density pass iterates its unconverged particles to the end (2-D)
    bool prune_lists;       // list builds use the pruned kernel (some grid was split)
    hipStream_t stream;
    // device arrays
    Inst* inst;
    Ctl* ctl;
    CtlView* hctl_dev;      // device view of the pinned host mirror of ctl: 3 slots
    int32_t* arrive;        // kArriveInts arrival counters of last_to_arrive (zero between launches)
    unsigned long long* hblk;  // per block of the fused stage kernel: its kRed maxima (instance in kblk)
    int32_t* kblk;          // that block's instance, or -1 when it straddles instances
    int32_t blk[3];         // workgroup size of the density, list (prep + build) and force launches
    int32_t persist[3];     // 1: that launch class runs persistent (work queues), 0: one block per tile
    int32_t* queue;         // kQueueInts: the work-queue heads of persistent launches (zero between launches)
    int32_t* inst_of;       // storage index -> instance (== sorted position -> instance)
    uint32_t *keys, *keys_alt;
    int32_t *perm, *perm_alt;
    int32_t* ccount;        // particles per cell (cell_cap; zero between sorts)
    int32_t* cblock;        // per scan block of kScanCells cells: exclusive offset
    int32_t* bigc;          // cells with more than kSmallBucket particles
    int32_t *cstart, *cend;
    int32_t *cls, *cle;     // the scan's block-local cell starts and ends (cell_cap)
    unsigned long long* hcell;  // per cell (cell_cap): bits of its largest h, for split grids; then
                                // (cell_cap more, at a row's first cell) the largest h of the row;
                                // then two cell_cap buffers of the h maxima over each cell's stencil
    uint32_t* hist;             // the h histogram of a single 3-D instance (Ctl::hist), or null
    double* dpack;          // 4 doubles per sorted particle: {x,y,z,m} for the density
                            // passes, then {x,y,z,h} for the neighbour-list build
    int32_t* act[2];        // active-particle lists of the density passes
    double* fpack;          // FRec<dim>::N doubles per sorted particle
    int32_t* nb;            // nb_stride x n, column major
    int32_t* nbcnt;
    double* scratch3;       // 3 x n (warm-start B copy)
    double* aprev;          // 3 x n: the accelerations the force pass reads as "previous",
                            // kept by the first prep of an evaluation for a redone pass
    double* totals;         // ninst x SPHMHD_NTOTALS
    double* tot_scratch;    // this device's copy of the static totals scratch (sph_totals_static.hip)
    Fields fields;
    Strain strain;          // optional velocity-gradient rows (sphmhd_bind_strain)
};

// sph_totals_static.hip: resolves the current device's copy of its static
// scratch array (once per device); sph_capi.hip: a context's device state
// and its longest instance, for entry points defined in other files
int scratch_setup(double** out);
int ctx_view(sphmhd_ctx* c, Dev** d, int64_t* longest);

// tangent-linear step (sph_tangent.hip)
constexpr int kTanDirs = SPHMHD_NTANGENTS;

struct TFields {
    double* f[kTanDirs][SPHMHD_NFIELDS];
};

struct InstTan {
    double dt[kTanDirs];                   // tangent of the current dt
    double dtp[kTanDirs];                  // tangent of the pending min dt_adapt
    double par[SPHMHD_NTPARAMS][kTanDirs]; // parameter tangents
    int32_t argmin;                        // storage index attaining the pending min
    int32_t pad;
};

struct TanDev {
    TFields tf;
    InstTan* itan = nullptr;  // ninst
    double* aprev = nullptr;  // 3 x (1 + kTanDirs) x n: previous accelerations and tangents
    double* bsave = nullptr;  // 3 x (1 + kTanDirs) x n: warm-start B copy
    double* totals = nullptr; // ninst x SPHMHD_NTOTALS x (1 + kTanDirs)
};

void launch_t_eval_begin(Dev& d, TanDev& g);
void launch_t_density(Dev& d, TanDev& g);
void launch_t_prep(Dev& d, TanDev& g, bool keep_prev);
void launch_t_force(Dev& d, TanDev& g);
void launch_t_initialize(Dev& d, TanDev& g);
void launch_t_stage1(Dev& d, TanDev& g);
void launch_t_stage2(Dev& d, TanDev& g);
void launch_t_clock(Dev& d, TanDev& g);
void launch_t_warm_dt(Dev& d, TanDev& g);
void launch_t_sync_initial(Dev& d, TanDev& g);
void launch_t_save(Dev& d, TanDev& g, int field, double* buf, bool restore);
void launch_t_copy(Dev& d, TanDev& g, int dst, int src);
void launch_t_totals(Dev& d, TanDev& g, double* out);

// adjoint (reverse-mode) step (sph_adjoint.hip)
constexpr int kAdjBlock = 128;    // workgroup of the reverse pair and density kernels
constexpr int kAdjLanes = 8;      // lanes sharing one particle in those kernels
constexpr int kAdjAcc = 35;       // sums of the force pass (their bars: AdjDev::acc rows)
constexpr int kAdjPassMax = 32;   // density passes one evaluation's reverse sweep holds per particle

// the properties a PEC step reads from the step before it (the state a step
// is recomputed from); the rest is written by initialize or the evaluation
// before it is read
enum {
    kSX = 0, kSY, kSZ, kSU, kSV, kSW, kSE, kSH, kSALPHA1, kSBX, kSBY, kSBZ, kSPSI,
    kSAU, kSAV, kSAW, kSAE, kSAH, kSCS, kSALPHALOC, kSABX, kSABY, kSABZ, kSAPSI, kNState
};
constexpr int kStateFields[kNState] = {
    SPHMHD_F_X, SPHMHD_F_Y, SPHMHD_F_Z, SPHMHD_F_U, SPHMHD_F_V, SPHMHD_F_W, SPHMHD_F_E, SPHMHD_F_H,
    SPHMHD_F_ALPHA1, SPHMHD_F_BEVOLX, SPHMHD_F_BEVOLY, SPHMHD_F_BEVOLZ, SPHMHD_F_PSI,
    SPHMHD_F_AU, SPHMHD_F_AV, SPHMHD_F_AW, SPHMHD_F_AE, SPHMHD_F_AH, SPHMHD_F_CS, SPHMHD_F_ALPHALOC,
    SPHMHD_F_ABEVOLX, SPHMHD_F_ABEVOLY, SPHMHD_F_ABEVOLZ, SPHMHD_F_APSI};

// property rows to copy (kernel argument)
struct RowList {
    int count;
    int f[SPHMHD_NFIELDS];
};
struct Weights {
    double w[SPHMHD_NTOTALS];
};

// per instance: bars of the dt of the step being reversed and of the
// pending dt of the evaluation being reversed, the particle attaining that
// pending minimum, the largest h of the evaluation's density passes, the
// parameter bars and the status (1: the recomputed density iteration ended
// elsewhere than the forward one, 2: more passes than the sweep holds)
struct AdjInst {
    double dtbar;
    double pendbar;
    double carry;  // fixed steps: the next dt's bar carried to this step's dt
    double par[SPHMHD_NTPARAMS];
    unsigned long long hmax_bits;
    int32_t argmin;
    int32_t status;
};

struct AdjDev {
    double* b[SPHMHD_NFIELDS];  // bar rows (the caller's), null where the field is null
    double* acc;        // 35 x n: bars of the force pass's sums (by sorted position)
    int32_t* argmax;    // n: sorted position of the neighbour attaining dt_cfl's maximum (-1: none)
    double* par;        // SPHMHD_NTPARAMS x n: per-particle parameter bars
    double* dtrow;      // n: per-particle bars of the step's dt
    double* ain;        // 5 x n: h, converged, au, av, aw entering the evaluation
    double* pass;       // 5 x passmax x n: per density pass h, sw, sg (values, then bars), dsw/dh,
                        // dsg/dh (by sorted position)
    int32_t* npass;     // n: density passes the particle ran (by sorted position)
    double* bsave;      // 3 x n: warm start: the bar of the saved initial B
    AdjInst* ai;        // ninst
    const double* snap; // kNState x n: the state the step being reversed started from
    int32_t passmax;
};

void launch_a_rows(Dev& d, const RowList& L, double* buf, bool restore);
void launch_a_clock(Dev& d, double* clk, bool restore);
void launch_a_save_entry(Dev& d, AdjDev& a);
void launch_a_seed_totals(Dev& d, AdjDev& a, const Weights& w);
void launch_a_zero(Dev& d, double* p, int64_t count);
// mode 0: the step clock (adaptive: the pending minimum; fixed: the current
// dt), 1: the warm start's first dt (the pending minimum), 2: none
void launch_a_eval_begin(Dev& d, AdjDev& a, const double* clk_next, int mode);
void launch_a_stage2(Dev& d, AdjDev& a);
void launch_a_eval(Dev& d, AdjDev& a);
void launch_a_stage1_init(Dev& d, AdjDev& a);
void launch_a_reduce_dt(Dev& d, AdjDev& a);
void launch_a_reduce_par(Dev& d, AdjDev& a);
void launch_a_sync(Dev& d, AdjDev& a);
void launch_a_bsave(Dev& d, AdjDev& a, bool take);

void launch_eval_begin(Dev& d, bool reset_dt);
void launch_pack_xh(Dev& d);
void launch_reset_nb_stats(Dev& d);
void launch_reset_force_outputs(Dev& d);
void launch_count_unconverged(Dev& d);
void launch_hmax_grid(Dev& d);
void launch_keys_sort_ranges(Dev& d, bool pack);  // pack: also write the density pack
void launch_pack_density(Dev& d);
void launch_density(Dev& d);
void launch_g0_control(Dev& d);
// keep_prev: the prep reads au, av, aw from the fields and keeps a copy in
// aprev; otherwise (a redone force pass) it reads them back from aprev
void launch_prep_force(Dev& d, bool keep_prev);
void launch_nb_build(Dev& d, bool cell_h = false);
void launch_force(Dev& d);
// sets the message sphmhd_last_error() returns; returns code
int report_error(int code, const std::string& msg);

// grid of a persistent launch: the blocks of `fn` (B threads, lds bytes of
// dynamic LDS) the GPU holds at once, a multiple of kXcds, at most the
// plain grid of `tiles` tiles
unsigned persistent_grid(const void* fn, int B, size_t lds, int64_t tiles);

void launch_eval_end(Dev& d);
void launch_ctl_export(Dev& d);
void launch_clear_abort(Dev& d);
void launch_set_lean(Dev& d, int lean);

void launch_initialize(Dev& d);
void launch_stage1(Dev& d);
void launch_stage2(Dev& d);
void launch_stage2_init_stage1(Dev& d);
void launch_init_stage1(Dev& d);
void launch_stage2_clock(Dev& d);
void launch_clock(Dev& d);
void launch_sync_initial(Dev& d);
void launch_copy(Dev& d, double* dst, const double* src, int64_t n);
void launch_totals(Dev& d);
void launch_warm_dt(Dev& d);
void launch_fill_i32(Dev& d, int32_t* p, int32_t v, int64_t n);

}  // namespace sph
