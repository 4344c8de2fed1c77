// sph_pairs.hip -- the pair loops of one force evaluation.
//
//   k_density   G0  DensityIteration loop + post_loop (equations.py:38-88)
//   k_g0_control    Group(iterate=True, min 2, max 250) control (scheme.py:71)
//   k_prep_force G1 + G2 + G4 (equations.py:104-130, 476-485) and the cell-ordered
//                   force pack with per-particle hoisted terms
//   k_force     G3 + G5 + G7 + G8 pair loops (equations.py:143-567) fused into one
//               neighbour traversal, with the G3/G5/G8 post_loops and G6 + G9
//               (equations.py:133-140, 570-624) as the epilogue.
//
// Fusion is legal because G3, G5, G7 and G8 read only neighbour values fixed by
// G0..G4; the one hazard (G3 reads the neighbours' previous a while G5 writes
// a) is removed by packing the previous a into the force pack before the loop.
// One thread per destination in cell order; FP64 throughout; no atomics on
// particle data (gather form), so results are deterministic.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "sph_internal.h"

namespace sph {

__device__ __forceinline__ double ipow_dim(double x, int dim) { return dim == 2 ? x * x : x * x * x; }
__device__ __forceinline__ double iroot_dim(double x, int dim) { return dim == 2 ? sqrt(x) : cbrt(x); }

// ---------------------------------------------------------------------------
// G0: density sum + Newton step for h
// ---------------------------------------------------------------------------

// A pass either scans every particle (act_in == nullptr: the first pass of an
// evaluation) or only the sorted positions listed in act_in (count at
// *cnt_in): the particles still unconverged after the previous pass.  A
// converged particle's h, omega and density sum cannot change again within
// the evaluation (its own h is frozen and the sum only involves its h, x and
// the neighbours' m, x), so skipping it is exact.  Particles that stay
// unconverged are appended to act_out.
//
// A later pass over few particles runs G lanes per particle (SPREAD): lane g
// takes every G-th candidate of each stencil range, an xor butterfly
// combines the G partial sums (every lane then holds the same bits) and the
// group's first lane finishes the particle.  Such a pass is bound by the
// latency of one particle's scan, which G lanes cut G-fold.  G is a power of
// two up to 64 chosen per instance by the group control from that
// instance's own count and size (Inst::spread), so an instance's bits do not
// depend on the launch shape or on the other instances of an ensemble; the
// launch maps Ctl::spread lanes (the largest G) to each particle and a
// particle uses the first G of them.
constexpr int kSpreadMax = 64;            // lanes per particle at most (a wave)
constexpr int64_t kSpreadSlots = 262144;  // lanes an instance's spread pass fills at most (four waves per SIMD)

__device__ __forceinline__ int spread_lanes(int64_t count, int64_t n) {
    const int64_t cap = n < kSpreadSlots ? n : kSpreadSlots;
    int G = 1;
    while (G < kSpreadMax && count * (2 * G) <= cap) G *= 2;
    return G;
}

// DensityIteration.post_loop (equations.py:54-86) from one pass's sums:
// writes h, omega and the converged flag; returns whether the particle stays
// unconverged (hout = its next h, else its final h)
template <int K, int DIM>
__device__ __forceinline__ bool density_newton(const Fields& F, const Inst& s, int32_t i, double hi, double mi,
                                               double sw, double sg, double& hout) {
    const double h1 = 1.0 / hi;
    const double fac = DIM == 2 ? kernel_sigma(K, 2) * h1 * h1 : kernel_sigma(K, 3) * h1 * h1 * h1;
    const double rho_sum = sw * fac;
    const double dwdh = -h1 * (fac * sg);
    const double hi0 = F.f[SPHMHD_F_H0][i];
    const double rhoi = mi / ipow_dim(hi / s.hfact, DIM);
    const double dhdrhoi = -hi / (DIM * rhoi);
    const double omegai = 1.0 - dhdrhoi * dwdh;
    F.f[SPHMHD_F_OMEGA][i] = 1.0 / omegai;
    const double func = rhoi - rho_sum;
    const double dfdh1 = omegai > 0.0 ? dhdrhoi / omegai : dhdrhoi / fabs(omegai + 2.220446049250313e-16);
    double hnew = hi - func * dfdh1;
    if (hnew > 1.2 * hi) hnew = 1.2 * hi;
    else if (hnew < 0.8 * hi) hnew = 0.8 * hi;
    const double diff = fabs(hnew - hi) / hi0;
    const bool fin = ((diff < 1.0e-4) && (omegai > 0.0) && (hi > 0.0)) || !s.density_iterate;
    if (!fin) {
        hout = hnew;
        F.f[SPHMHD_F_CONVERGED][i] = 0.0;
    } else {
        hout = s.hfact * iroot_dim(mi / rho_sum, DIM);
        F.f[SPHMHD_F_CONVERGED][i] = 1.0;
    }
    F.f[SPHMHD_F_H][i] = hout;
    return !fin;
}

// Lane g of G lanes' part of one particle's density sums at h: every G-th
// candidate of each stencil range.  The stencil follows this h alone (the
// pass's stencil, sized for the h_max its pass began with, may be smaller by
// now), and so does the choice between the ranges' image shifts and a
// minimum image per candidate, which give the same r^2 bits.
template <int K, int DIM>
__device__ __forceinline__ void density_lane_sums(const Inst& s, const double4* __restrict__ dpack, int64_t key,
                                                  const int32_t* __restrict__ cstart,
                                                  const int32_t* __restrict__ cend, const double4& me, double hi,
                                                  int g, int G, double& sw, double& sg) {
    const double R = kernel_radius(K);
    const double h1 = 1.0 / hi;
    const double Rh = R * hi;
    const double Rh2 = Rh * Rh;
    const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
    const double H0 = s.halfL[0], H1 = s.halfL[1], H2 = s.halfL[2];
    const int per = s.periodic;
    int sr[3] = {0, 0, 0};
    bool shift = true;
    for (int a = 0; a < DIM; ++a) {
        const double q = ceil(Rh / s.ceff[a]);
        sr[a] = q < 1.0 ? 1 : (q < 1.0e6 ? (int)q : 1000000);
        if (per && !((int64_t)s.n[a] > 2 * ((int64_t)sr[a] + 1))) shift = false;
    }
    auto scan = [&](auto shift_c) {
        constexpr bool kShift = decltype(shift_c)::value;
        auto body = [&](int, const double4& o, double sx, double sy, double sz) {
            const double dx = kShift ? (me.x - o.x) + sx : min_image(me.x - o.x, L0, H0, per);
            const double dy = kShift ? (me.y - o.y) + sy : min_image(me.y - o.y, L1, H1, per);
            const double dz = DIM == 3 ? (kShift ? (me.z - o.z) + sz : min_image(me.z - o.z, L2, H2, per)) : 0.0;
            const double r2 = r2_exact(dx, dy, dz);
            if (r2 < Rh2) {
                const double q = sqrt_r2(r2) * h1;
                double w, dw;
                shape<K, DIM>(q, w, dw);
                sw += o.w * w;
                sg += o.w * ((double)DIM * w + q * dw);
            }
        };
        auto rf = [&](int a, int b, double sx, double sy, double sz) {
            scan_range4_strided(dpack, a, b, g, G, [&](int j, const double4& o) { body(j, o, sx, sy, sz); });
        };
        for_each_row_range<DIM>(s, key, cstart, cend, rf, sr);
    };
    if (shift) scan(std::true_type{});
    else scan(std::false_type{});
}

// Per instance, what the block's particles that finished in one round leave
// for the group control: the largest h of those converged (hc) and of those
// left unconverged (hu), how many were left unconverged (uc) and the most
// passes one took (np).  Waves reduce by shuffles and thread 0 issues one
// atomic per quantity and instance run; a wave holding several instances
// falls back to lane atomics.  Every thread of the block calls it.
__device__ __forceinline__ void reduce_finished(Inst* inst, int ki, bool fin, unsigned long long hc,
                                                unsigned long long hu, int uc, int np) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    constexpr int NW = kBlock / kWave;
    const int nw = blockDim.x / kWave;
    __shared__ int rk[NW], ruc[NW], rnp[NW];
    __shared__ unsigned long long rhc[NW], rhu[NW];
    if (!fin) {
        hc = hu = 0ull;
        uc = np = 0;
    }
    const unsigned long long fm = __ballot(fin);
    const int k0 = fm ? __shfl(ki, __ffsll((long long)fm) - 1) : -1;
    const bool uniform = __all(!fin || ki == k0);
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(hc, off);
            hc = o > hc ? o : hc;
            o = __shfl_xor(hu, off);
            hu = o > hu ? o : hu;
            uc += __shfl_xor(uc, off);
            const int op = __shfl_xor(np, off);
            np = op > np ? op : np;
        }
    } else if (fin) {
        if (hc) atomicMax(&inst[ki].hmax_conv_bits, hc);
        if (hu) atomicMax(&inst[ki].hmax_next_bits, hu);
        if (uc) atomicAdd(&inst[ki].unconv, uc);
        if (np > 1) atomicMax(&inst[ki].passes, np);
    }
    if (lane == 0) {
        rk[w] = uniform ? k0 : -1;
        rhc[w] = hc;
        rhu[w] = hu;
        ruc[w] = uc;
        rnp[w] = np;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < nw;) {
            const int k = rk[q];
            unsigned long long c = rhc[q], u = rhu[q];
            int cu = ruc[q], mp = rnp[q];
            int r = q + 1;
            for (; r < nw && rk[r] == k; ++r) {
                c = rhc[r] > c ? rhc[r] : c;
                u = rhu[r] > u ? rhu[r] : u;
                cu += ruc[r];
                mp = rnp[r] > mp ? rnp[r] : mp;
            }
            if (k >= 0) {
                if (c) atomicMax(&inst[k].hmax_conv_bits, c);
                if (u) atomicMax(&inst[k].hmax_next_bits, u);
                if (cu) atomicAdd(&inst[k].unconv, cu);
                if (mp > 1) atomicMax(&inst[k].passes, mp);
            }
            q = r;
        }
    }
    __syncthreads();
}

// The passes after the first within the first pass's workgroup.  A
// particle's Newton iteration reads only its own h and its neighbours'
// positions and masses, which no other particle's iteration changes, so the
// block iterates each of its particles still unconverged to convergence (or
// to its instance's max_iterations) at once: those particles are listed in
// LDS and each pass gives each a fixed number of lanes (their sums combined
// by an xor butterfly), as many rounds as the list needs.  A particle's bits
// then depend on nothing but that lane count: not on the launch shape, the
// block or the other instances of an ensemble.  The iteration ends within the
// launch; the group control takes the passes from the largest count a
// particle needed (at least min_iterations, as the group iteration runs).
// Lanes per particle: 4 in 2-D (the 512^2 rotor's density 61 us per step,
// against 65 with 8 lanes and 68 with 16); 2 in 3-D, where the stencil rows
// are many and short.
constexpr int kFuseLanes2 = 4, kFuseLanes3 = 2;

template <int K, int DIM>
__device__ __forceinline__ void density_fused_passes(const Fields& F, const double4* __restrict__ dpack,
                                                     const int32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ skey,
                                                     const int32_t* __restrict__ cstart,
                                                     const int32_t* __restrict__ cend, Inst* inst,
                                                     const int32_t* __restrict__ inst_of, int ki, int64_t t,
                                                     bool live, bool unconv, double hout) {
    __shared__ int32_t lt[2][kBlock];
    __shared__ double lh[2][kBlock];
    __shared__ int lcnt[2];
    if (threadIdx.x == 0) lcnt[0] = 0;
    __syncthreads();
    // the first pass's outcome: particles that go on are listed, the rest finish
    const bool go = unconv && inst[ki].max_iter > 1;
    const bool left = unconv && !go;
    if (go) {
        const int slot = atomicAdd(&lcnt[0], 1);
        lt[0][slot] = (int32_t)t;
        lh[0][slot] = hout;
    }
    reduce_finished(inst, ki, live && !go, unconv ? 0ull : dbits(hout), left ? dbits(hout) : 0ull, left ? 1 : 0, 1);
    constexpr int kFuseLanes = DIM == 2 ? kFuseLanes2 : kFuseLanes3;
    const int B = blockDim.x;
    int cur = 0;
    for (int p = 2;; ++p) {
        const int cnt = lcnt[cur];
        if (cnt == 0) break;
        if (threadIdx.x == 0) lcnt[cur ^ 1] = 0;
        __syncthreads();
        for (int base = 0; base < cnt * kFuseLanes; base += B) {
            const int idx = base + threadIdx.x;
            const int item = idx / kFuseLanes, g = idx & (kFuseLanes - 1);
            const bool v = item < cnt;
            const int64_t tt = v ? (int64_t)lt[cur][item] : 0;
            const double h = v ? lh[cur][item] : 0.0;
            const int kk = v ? inst_of[tt] : 0;
            const Inst& s = inst[kk];
            double sw = 0.0, sg = 0.0;
            double4 me = make_double4(0.0, 0.0, 0.0, 0.0);
            if (v) {
                me = dpack[tt];
                density_lane_sums<K, DIM>(s, dpack, (int64_t)skey[tt], cstart, cend, me, h, g, kFuseLanes, sw, sg);
            }
            for (int off = 1; off < kFuseLanes; off <<= 1) {
                sw += __shfl_xor(sw, off);
                sg += __shfl_xor(sg, off);
            }
            const bool lead = v && g == 0;
            bool more = false, stays = false;
            double ho = h;
            if (lead) {
                stays = density_newton<K, DIM>(F, s, perm[tt], h, me.w, sw, sg, ho);
                more = stays && p < s.max_iter;
                if (more) {
                    const int slot = atomicAdd(&lcnt[cur ^ 1], 1);
                    lt[cur ^ 1][slot] = (int32_t)tt;
                    lh[cur ^ 1][slot] = ho;
                }
            }
            const bool ends = stays && !more;
            reduce_finished(inst, kk, lead && !more, stays ? 0ull : dbits(ho), ends ? dbits(ho) : 0ull,
                            ends ? 1 : 0, p);
        }
        cur ^= 1;
    }
}

template <int K, int DIM, bool SPREAD, bool TRIM = false, bool FUSE = false>
__device__ __forceinline__ void density_block(Fields F, const double4* __restrict__ dpack,
                                              const int32_t* __restrict__ perm, const uint32_t* __restrict__ skey,
                                              const int32_t* __restrict__ cstart, const int32_t* __restrict__ cend,
                                              Inst* inst, const int32_t* __restrict__ inst_of, int64_t n,
                                              const int32_t* __restrict__ act_in, const int32_t* cnt_in,
                                              int32_t* __restrict__ act_out, int32_t* cnt_out, int64_t tile,
                                              int G, int trim = 1) {
    const int B = blockDim.x;  // 64, 128 or 256 (the launch shape is tuned)
    const int64_t idx0 = tile * B;
    // Gm lanes mapped to each item (B is a multiple of Gm, so they are
    // consecutive lanes of one wave)
    const int Gm = SPREAD ? G : 1;
    const int64_t items = (act_in ? (int64_t)*cnt_in : n) * Gm;
    if (idx0 >= items) return;  // whole block past the end (the block barriers below need every wave)
    const int64_t idx = idx0 + threadIdx.x;
    const bool valid = idx < items;
    const int g = SPREAD ? (int)(threadIdx.x & (Gm - 1)) : 0;
    const int64_t item = SPREAD ? idx / Gm : idx;
    int64_t t = valid ? (act_in ? (int64_t)act_in[item] : item) : 0;
    int ki = valid ? inst_of[t] : 0;
    const Inst& s = inst[ki];
    const int Gk = SPREAD ? min(max(s.spread, 1), Gm) : 1;  // this particle's lanes: the first Gk of its Gm
    bool live = valid && !s.done;
    int32_t i = valid ? perm[t] : 0;
    double hi = live ? F.f[SPHMHD_F_H][i] : 0.0;
    bool active = live && F.f[SPHMHD_F_CONVERGED][i] != 1.0 && g < Gk;
    double hout = hi;
    bool unconv = false;

    double sw = 0.0, sg = 0.0;
    double4 me = make_double4(0.0, 0.0, 0.0, 0.0);
    if (active) {
        const double R = kernel_radius(K);
        me = dpack[t];
        const double h1 = 1.0 / hi;
        const double Rh = R * hi;
        const double Rh2 = Rh * Rh;
        const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
        const double H0 = s.halfL[0], H1 = s.halfL[1], H2 = s.halfL[2];
        const int per = s.periodic;
        // this particle's own stencil: every j with r < R h_i lies within
        // ceil(R h_i / c_eff) cells of it (at most the pass's stencil, which
        // is sized for the largest h), so where h varies the small-h
        // particles scan fewer cells
        int sr[3] = {0, 0, 0};
        for (int a = 0; a < DIM; ++a) {
            const double q = ceil(Rh / s.ceff[a]);
            sr[a] = q < 1.0 ? 1 : (q < (double)s.srad[a] ? (int)q : s.srad[a]);
        }
        // one loop with the ranges' image shifts, one with min_image per
        // candidate (small periodic grids); both give the same r^2 bits
        auto scan = [&](auto shift_c) {
            constexpr bool kShift = decltype(shift_c)::value;
            auto body = [&](int j, const double4& o, double sx, double sy, double sz) {
                const double dx = kShift ? (me.x - o.x) + sx : min_image(me.x - o.x, L0, H0, per);
                const double dy = kShift ? (me.y - o.y) + sy : min_image(me.y - o.y, L1, H1, per);
                const double dz = DIM == 3 ? (kShift ? (me.z - o.z) + sz : min_image(me.z - o.z, L2, H2, per)) : 0.0;
                const double r2 = r2_exact(dx, dy, dz);
                if (r2 < Rh2) {
                    const double q = sqrt_r2(r2) * h1;
                    double w, dw;
                    shape<K, DIM>(q, w, dw);
                    sw += o.w * w;
                    sg += o.w * ((double)DIM * w + q * dw);
                }
            };
            auto rf = [&](int a, int b, double sx, double sy, double sz) {
                auto f = [&](int j, const double4& o) { body(j, o, sx, sy, sz); };
                if constexpr (SPREAD) scan_range4_strided(dpack, a, b, g, Gk, f);
                else scan_range4(dpack, a, b, f);
            };
            // the first pass in 3-D scans only the cells within R h_i of the
            // particle (the same hits in the same order, fewer candidates)
            if constexpr (DIM == 3 && !SPREAD && kShift && TRIM) {
                if (trim) {
                    for_each_row_range_trim<DIM>(s, (int64_t)skey[t], me.x, me.y, me.z, Rh2, cstart, cend, rf, sr);
                    return;
                }
            }
            for_each_row_range<DIM>(s, (int64_t)skey[t], cstart, cend, rf, sr);
        };
        if (s.range_shift) scan(std::true_type{});
        else scan(std::false_type{});
    }
    if constexpr (SPREAD) {
        // every lane of the wave takes part; a particle's first Gk lanes
        // exchange only among themselves
        for (int off = 1; off < Gm; off <<= 1) {
            const double ow = __shfl_xor(sw, off), og = __shfl_xor(sg, off);
            if (off < Gk) {
                sw += ow;
                sg += og;
            }
        }
        if (g != 0) live = active = false;
    }

    if (active) unconv = density_newton<K, DIM>(F, s, i, hi, me.w, sw, sg, hout);
    if constexpr (FUSE) {
        density_fused_passes<K, DIM>(F, dpack, perm, skey, cstart, cend, inst, inst_of, ki, (int64_t)t, live,
                                     unconv, hout);
        return;
    }

    // Still-unconverged particles go to the next pass's list, and per instance
    // the max h over converged particles (frozen from now on), the max h over
    // the still-unconverged ones and their count are reduced.  Same-address
    // atomics serialise, so waves combine in LDS and thread 0 issues one
    // atomic per quantity and block (per instance run when a block straddles
    // instances; a wave that straddles instances falls back to lane atomics).
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    constexpr int NW = kBlock / kWave;  // LDS slots: waves of the largest shape
    const int nw = blockDim.x / kWave;
    __shared__ int wcnt[NW], wbase[NW], wk[NW], wuc[NW];
    __shared__ unsigned long long whc[NW], whu[NW];
    const unsigned long long um = __ballot(unconv);
    const bool conv_now = live && !unconv;
    unsigned long long hc = conv_now ? dbits(hout) : 0ull;
    unsigned long long hu = unconv ? dbits(hout) : 0ull;
    int uc = unconv ? 1 : 0;
    const int k0 = __shfl(ki, 0);  // lane 0 is valid whenever any lane is
    const bool uniform = __all(ki == k0 || !valid);
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(hc, off);
            hc = o > hc ? o : hc;
            o = __shfl_xor(hu, off);
            hu = o > hu ? o : hu;
            uc += __shfl_xor(uc, off);
        }
    } else if (live) {
        if (hc) atomicMax(&inst[ki].hmax_conv_bits, hc);
        if (hu) atomicMax(&inst[ki].hmax_next_bits, hu);
        if (uc) atomicAdd(&inst[ki].unconv, uc);
    }
    if (lane == 0) {
        wcnt[w] = __popcll(um);
        wk[w] = uniform && valid ? k0 : -1;
        whc[w] = hc;
        whu[w] = hu;
        wuc[w] = uc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int q = 0; q < nw; ++q) {
            wbase[q] = tot;
            tot += wcnt[q];
        }
        if (tot) {
            const int base = atomicAdd(cnt_out, tot);
            for (int q = 0; q < nw; ++q) wbase[q] += base;
        }
        for (int q = 0; q < nw;) {
            const int k = wk[q];
            unsigned long long c = whc[q], u = whu[q];
            int cu = wuc[q];
            int r = q + 1;
            for (; r < nw && wk[r] == k; ++r) {
                c = whc[r] > c ? whc[r] : c;
                u = whu[r] > u ? whu[r] : u;
                cu += wuc[r];
            }
            if (k >= 0 && !inst[k].done) {
                if (c) atomicMax(&inst[k].hmax_conv_bits, c);
                if (u) atomicMax(&inst[k].hmax_next_bits, u);
                if (cu) atomicAdd(&inst[k].unconv, cu);
            }
            q = r;
        }
    }
    __syncthreads();
    if (unconv) {
        const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        act_out[wbase[w] + __popcll(um & lt)] = (int32_t)t;
    }
}

// Group iteration control: iterations >= min and converged, or == max.
// Also clears the counter of the active list the next pass will write
// (next_list), so passes can be launched back to back without the host.
// One workgroup.
// fused: the pass iterated every particle to the end (k_density FUSE), and
// the passes are the most any particle needed, at least min_iterations.
__device__ void g0_control_body(Inst* inst, int ninst, Ctl* ctl, int next_list, int fused = 0) {
    __shared__ int alldone, spread;
    if (threadIdx.x == 0) {
        alldone = 1;
        spread = 1;
        *act_count(ctl, next_list) = 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
        Inst& s = inst[k];
        if (!s.done) {
            // unconv and the two h maxima were reduced by the pass's atomics
            const int unconv = ld_dev(&s.unconv);
            const unsigned long long hconv = ld_dev(&s.hmax_conv_bits), hnext = ld_dev(&s.hmax_next_bits);
            if (fused) {
                const int np = max(ld_dev(&s.passes), 1);
                s.iters = !s.density_iterate ? 1 : min(max(s.min_iter, np), s.max_iter);
                s.passes = 0;
            } else {
                s.iters += 1;
            }
            bool conv = unconv == 0;
            bool stop = fused || !s.density_iterate || (s.iters >= s.min_iter && conv) || s.iters >= s.max_iter;
            if (s.density_iterate && !(s.iters >= s.min_iter && conv) && s.iters >= s.max_iter) s.status |= 1;
            s.unconv = 0;
            s.spread = spread_lanes(unconv, s.off1 - s.off0);
            s.hmax_bits = hconv > hnext ? hconv : hnext;
            s.hmax_next_bits = 0ull;
            stencil_from_hmax(s);
            if (stop) s.done = 1;
            atomicMax(&ctl->iters_max, s.iters);
        }
        if (!s.done) {
            atomicAnd(&alldone, 0);
            atomicMax(&spread, s.spread);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->all_done = alldone;
        ctl->spread = spread;
    }
}

__global__ void k_g0_control(Inst* inst, int ninst, Ctl* ctl, int next_list) {
    g0_control_body(inst, ninst, ctl, next_list);
}

// one density pass; its last workgroup runs the group control.  LATER: a
// pass after the first (act_in set), with the spread lanes; the first pass
// has an instantiation of its own that carries none of that code.  FUSE: a
// first pass whose workgroups go on to iterate their unconverged particles to
// the end (density_fused_passes), so that no later pass is needed.
template <int K, int DIM, bool LATER, bool FUSE = false>
__global__ void __launch_bounds__(kBlock, DIM == 3 ? 4 : 1) k_density(Fields F, const double4* __restrict__ dpack,
                                                    const int32_t* __restrict__ perm, const uint32_t* __restrict__ skey,
                                                    const int32_t* __restrict__ cstart, const int32_t* __restrict__ cend,
                                                    Inst* inst, const int32_t* __restrict__ inst_of, int64_t n,
                                                    const int32_t* __restrict__ act_in, const int32_t* cnt_in,
                                                    int32_t* __restrict__ act_out, int32_t* cnt_out, Ctl* ctl,
                                                    int ninst, int next_list, int32_t* arrive, int32_t* queue,
                                                    int trim) {
    if (ctl->abort) return;  // the whole launch: the step after an unfinished evaluation
    const int B = blockDim.x;
    const int G = LATER ? max(ctl->spread, 1) : 1;
    auto run = [&](int64_t tile) {
        if (!LATER)
            density_block<K, DIM, false, true, FUSE>(F, dpack, perm, skey, cstart, cend, inst, inst_of, n, act_in,
                                                     cnt_in, act_out, cnt_out, tile, 1, trim);
        else if (G == 1)
            density_block<K, DIM, false>(F, dpack, perm, skey, cstart, cend, inst, inst_of, n, act_in, cnt_in,
                                         act_out, cnt_out, tile, 1);
        else
            density_block<K, DIM, LATER>(F, dpack, perm, skey, cstart, cend, inst, inst_of, n, act_in, cnt_in,
                                         act_out, cnt_out, tile, G);
    };
    if (queue) {  // persistent: tiles of this pass's items from the work queues
        const int64_t items = (act_in ? (int64_t)*cnt_in : n) * G;
        int polls = 0;
        for (int64_t tile; (tile = pull_tile(queue, (items + B - 1) / B, polls)) >= 0;) run(tile);
    } else if (LATER) {
        // a later pass's items fill the first tiles: blocks take them in
        // dispatch order, which spreads them over every XCD (an ensemble's
        // G * count can pass the grid: then blocks take further tiles)
        const int64_t tiles = ((int64_t)*cnt_in * G + B - 1) / B;
        for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) run(tile);
    } else {
        run(xcd_tile((n + B - 1) / B));
    }
    if (last_to_arrive(arrive)) {
        g0_control_body(inst, ninst, ctl, next_list, FUSE ? 1 : 0);
        if (queue) reset_queues(queue);
    }
}

// ---------------------------------------------------------------------------
// G1 + G2 + G4 and the force pack
// ---------------------------------------------------------------------------

// A particle's inputs to k_prep_force, all loaded before its first store: a
// load after the property stores could alias them for the compiler, which
// would then wait for another round trip to memory.
struct PrepIn {
    double h, m, epred, bxp, byp, bzp, x, y, z, au, av, aw, upred, vpred, wpred, omega, alpha, psip;
};

template <int DIM>
__device__ __forceinline__ void write_force_record(double2* stage, const PrepIn& in, double rho, double p, double bx,
                                                   double by, double bz, double vwave, double chf) {
    using R = FRec<DIM>;
    const double x = in.x, y = in.y, z = in.z, h = in.h, m = in.m, e = in.epred;
    const double au = in.au, av = in.av, aw = in.aw;
    double r[R::N];
    r[R::X] = x;
    r[R::Y] = y;
    if constexpr (DIM == 3) r[R::Z] = z;
    r[R::H1] = 1.0 / h;
    r[R::M] = m;
    r[R::RHO] = rho;
    r[R::E] = e;
    r[R::U] = in.upred;
    r[R::V] = in.vpred;
    r[R::W] = in.wpred;
    r[R::OM] = in.omega;
    r[R::BX] = bx;
    r[R::BY] = by;
    r[R::BZ] = bz;
    r[R::AL] = in.alpha;
    r[R::AU] = au;
    r[R::AV] = av;
    if constexpr (DIM == 3) r[R::AW] = aw;
    r[R::PSICH] = in.psip * (chf * vwave);
    r[R::VW] = vwave;
#pragma unroll
    for (int q = 0; q < R::UNITS; ++q) stage[q] = make_double2(r[2 * q], r[2 * q + 1]);
}

// one particle of k_prep_force: the group G1/G2/G4 epilogue values into the
// property rows, its force record into stage
__device__ __forceinline__ void prep_force_one(Fields F, const int32_t* __restrict__ perm, const Inst* __restrict__ inst,
                                               const int32_t* __restrict__ inst_of, double2* stage,
                                               double4* __restrict__ xh, int64_t t, int64_t n,
                                               double* __restrict__ aprev, int keep_prev) {
    const Inst& s = inst[inst_of[t]];
    const int32_t i = perm[t];
    const int dim = s.dim;
    double** f = F.f;
    // every input first (the previous accelerations from their rows on the
    // first prep of an evaluation, which the force pass then overwrites, or
    // back from aprev on a redone pass)
    PrepIn in;
    in.h = f[SPHMHD_F_H][i];
    in.m = f[SPHMHD_F_M][i];
    in.epred = f[SPHMHD_F_EPRED][i];
    in.bxp = f[SPHMHD_F_BEVOLXPRED][i];
    in.byp = f[SPHMHD_F_BEVOLYPRED][i];
    in.bzp = f[SPHMHD_F_BEVOLZPRED][i];
    in.x = f[SPHMHD_F_X][i];
    in.y = f[SPHMHD_F_Y][i];
    in.z = dim == 3 ? f[SPHMHD_F_Z][i] : 0.0;
    const double* asrc[3] = {f[SPHMHD_F_AU], f[SPHMHD_F_AV], f[SPHMHD_F_AW]};
    if (!keep_prev) {
        asrc[0] = aprev;
        asrc[1] = aprev + n;
        asrc[2] = aprev + 2 * n;
    }
    in.au = asrc[0][i];
    in.av = asrc[1][i];
    in.aw = asrc[2][i];
    in.upred = f[SPHMHD_F_UPRED][i];
    in.vpred = f[SPHMHD_F_VPRED][i];
    in.wpred = f[SPHMHD_F_WPRED][i];
    in.omega = f[SPHMHD_F_OMEGA][i];
    in.alpha = f[SPHMHD_F_ALPHA1][i];
    in.psip = f[SPHMHD_F_PSIPRED][i];
    const double h = in.h, m = in.m, epred = in.epred;
    // DensityFromSmoothingLength (equations.py:118)
    const double rho = m / ipow_dim(h / s.hfact, dim);
    // IdealGasEOS (equations.py:105-106)
    const double p = s.gamma1 * rho * epred;
    const double cs = sqrt(s.gamma * p / rho);
    // MagneticStateReconstruction (equations.py:477-485)
    const double bx = in.bxp * rho;
    const double by = in.byp * rho;
    const double bz = in.bzp * rho;
    f[SPHMHD_F_RHO][i] = rho;
    f[SPHMHD_F_P][i] = p;
    f[SPHMHD_F_CS][i] = cs;
    f[SPHMHD_F_BX][i] = bx;
    f[SPHMHD_F_BY][i] = by;
    f[SPHMHD_F_BZ][i] = bz;
    f[SPHMHD_F_BXPRED][i] = bx;
    f[SPHMHD_F_BYPRED][i] = by;
    f[SPHMHD_F_BZPRED][i] = bz;

    const double b2 = bx * bx + by * by + bz * bz;
    const double vwave = sqrt(cs * cs + b2 / (s.mu0 * rho));
    xh[t] = make_double4(in.x, in.y, in.z, h);
    if (keep_prev) {
        aprev[i] = in.au;
        aprev[n + i] = in.av;
        aprev[2 * n + i] = in.aw;
    }
    if (dim == 2) write_force_record<2>(stage, in, rho, p, bx, by, bz, vwave, s.chf);
    else write_force_record<3>(stage, in, rho, p, bx, by, bz, vwave, s.chf);
}

__global__ void k_prep_force(Fields F, const int32_t* __restrict__ perm, const Inst* __restrict__ inst,
                             const int32_t* __restrict__ inst_of, double* __restrict__ fpack,
                             double4* __restrict__ xh, int64_t n, double* __restrict__ aprev, int keep_prev,
                             const Ctl* ctl) {
    if (ctl->abort) return;
    const int64_t t0 = xcd_tile((n + blockDim.x - 1) / blockDim.x) * blockDim.x;
    if (t0 >= n) return;  // whole block past the end
    // Records are built in LDS and stored as one contiguous run per block:
    // each lane writing its own 144/160-byte record would leave every store
    // instruction touching 64 partial lines.
    __shared__ double2 stage[kBlock * kFRecMax / 2];
    const int64_t t = t0 + threadIdx.x;
    const int dim = inst[inst_of[t0]].dim;  // the same for every instance
    const int units = dim == 3 ? FRec<3>::UNITS : FRec<2>::UNITS;
    if (t < n) prep_force_one(F, perm, inst, inst_of, stage + threadIdx.x * units, xh, t, n, aprev, keep_prev);
    __syncthreads();
    const int64_t cnt = min((int64_t)blockDim.x, n - t0) * units;
    double2* out = (double2*)fpack + t0 * units;
    for (int64_t q = threadIdx.x; q < cnt; q += blockDim.x) out[q] = stage[q];
}

// ---------------------------------------------------------------------------
// fused force pass
// ---------------------------------------------------------------------------

template <int DIM>
struct Acc {
    double ax, ay, az, ae, div, dtcfl;
    double bx, by, bz, divbs, divbd;
    // Cullen-Dehnen sums (equations.py:329-354); in 2-D only the entries that
    // can be non-zero are kept
    double rxx, rxy, rxz, ryy, ryz, rzz;
    double vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz;
    double axx, axy, axz, ayx, ayy, ayz, azx, azy, azz;
};

template <int K, int DIM>
__device__ __forceinline__ void force_block(Fields F, const double* __restrict__ fpack,
                                            const int32_t* __restrict__ perm, const int32_t* __restrict__ nb,
                                            const int32_t* __restrict__ nbcnt, Inst* inst,
                                            const int32_t* __restrict__ inst_of, int64_t n, int stride,
                                            int lean, Strain strain, int64_t tile) {
    using R = FRec<DIM>;
    constexpr int U = R::UNITS;
    const double2* __restrict__ pk = (const double2*)fpack;

    const int B = blockDim.x;  // 64, 128 or 256 (the launch shape is tuned)
    const int64_t t0 = tile * B;
    if (t0 >= n) return;  // whole block past the end (the block barrier below needs every wave)
    const int64_t t = t0 + threadIdx.x;
    const bool valid = t < n;
    const int ki = valid ? inst_of[t] : 0;
    const Inst& s = inst[ki];
    double dt_adapt = 1.0e300;
    bool bad = false;

    // destination values (rows of the pack are read once, so plain loads)
    double me[R::N];
    int cnt = 0;
    if (valid) {
#pragma unroll
        for (int c = 0; c < U; ++c) {
            const double2 v = pk[t * U + c];
            me[2 * c] = v.x;
            me[2 * c + 1] = v.y;
        }
        cnt = min(nbcnt[t], stride);  // a longer list is re-run after growing
    } else {
#pragma unroll
        for (int c = 0; c < R::N; ++c) me[c] = 1.0;
    }
    const double xi = me[R::X], yi = me[R::Y], zi = DIM == 3 ? me[R::Z] : 0.0;
    const double h1i = me[R::H1];
    const double rhoi = me[R::RHO], ei = me[R::E];
    const double pi_ = s.gamma1 * rhoi * ei;  // IdealGasEOS, as k_prep_force
    const double ui = me[R::U], vi = me[R::V], wi = me[R::W];
    const double omi = me[R::OM], ali = me[R::AL];
    const double bxi = me[R::BX], byi = me[R::BY], bzi = me[R::BZ];
    const double aui = me[R::AU], avi = me[R::AV], awi = DIM == 3 ? me[R::AW] : 0.0;
    const double psichi = me[R::PSICH], vwi = me[R::VW];
    const double irhoi = 1.0 / rhoi, hiri = 0.5 * irhoi;
    const double inv_mu0 = 1.0 / s.mu0;
    const double r2ii = irhoi * irhoi;
    const double pbi = pi_ * r2ii;
    const double brxi = bxi * irhoi, bryi = byi * irhoi, brzi = bzi * irhoi;
    const double bro2i = (brxi * brxi + bryi * bryi + brzi * brzi) * inv_mu0;
    const double sig = DIM == 2 ? kernel_sigma(K, 2) : kernel_sigma(K, 3);
    const double fac_i = (DIM == 2 ? sig * h1i * h1i : sig * h1i * h1i * h1i) * h1i;
    const double av = s.av, ac = s.ac, ar = s.ar, gamma1 = s.gamma1;
    const double Lx = s.L[0], Ly = s.L[1], Lz = s.L[2];
    const double hLx = s.halfL[0], hLy = s.halfL[1], hLz = s.halfL[2];
    const int periodic = s.periodic;

    // The pass is bound by the bytes gathered from neighbour records, so the
    // record is kept small.  In 2-D the next neighbour's record is loaded while
    // the current one is computed (the index one slot further ahead); 3-D and
    // the Gaussian kernels have no registers to spare for that (it would halve
    // their occupancy).
    constexpr bool kPrefetch = DIM == 2 && K != SPHMHD_GAUSSIAN && K != SPHMHD_SUPER_GAUSSIAN;
    int jn = cnt > (kPrefetch ? 1 : 0) ? nb[(kPrefetch ? (int64_t)n : 0) + t] : -1;
    double2 nxt[U];
    if (kPrefetch && cnt > 0) {
        const int64_t j0 = nb[t];
#pragma unroll
        for (int c = 0; c < U; ++c) nxt[c] = pk[j0 * U + c];
    }

    Acc<DIM> A = {};
    for (int q = 0; q < cnt; ++q) {
        if (!kPrefetch) {
#pragma unroll
            for (int c = 0; c < U; ++c) nxt[c] = pk[(int64_t)jn * U + c];
        }
        double o[R::N];
#pragma unroll
        for (int c = 0; c < U; ++c) {
            o[2 * c] = nxt[c].x;
            o[2 * c + 1] = nxt[c].y;
        }
        if (kPrefetch && jn >= 0) {
#pragma unroll
            for (int c = 0; c < U; ++c) nxt[c] = pk[(int64_t)jn * U + c];
        }
        const int ahead = kPrefetch ? 2 : 1;
        jn = q + ahead < cnt ? nb[(int64_t)(q + ahead) * n + t] : -1;
        {
            const double h1j = o[R::H1], mj = o[R::M], rhoj = o[R::RHO], ej = o[R::E];
            const double pj = gamma1 * rhoj * ej;
            const double uj = o[R::U], vj = o[R::V], wj = o[R::W], omj = o[R::OM];
            const double bxj = o[R::BX], byj = o[R::BY], bzj = o[R::BZ], alj = o[R::AL];
            const double auj = o[R::AU], avj = o[R::AV], awj = DIM == 3 ? o[R::AW] : 0.0;
            const double psichj = o[R::PSICH], vwj = o[R::VW];
            const double irhoj = 1.0 / rhoj, hirj = 0.5 * irhoj;
            const double r2ij = irhoj * irhoj;
            const double pbj = pj * r2ij;
            const double brxj = bxj * irhoj, bryj = byj * irhoj, brzj = bzj * irhoj;
            const double bro2j = (brxj * brxj + bryj * bryj + brzj * brzj) * inv_mu0;

            double dx = min_image(xi - o[R::X], Lx, hLx, periodic);
            double dy = min_image(yi - o[R::Y], Ly, hLy, periodic);
            double dz = DIM == 3 ? min_image(zi - o[R::Z], Lz, hLz, periodic) : 0.0;
            const double r2 = r2_exact(dx, dy, dz);
            const double r = sqrt_r2(r2);  // r2 = 0 (the particle itself) gives 2^-300: below both cut-offs

            // kernel gradients DWI, DWJ (0 for r <= 1e-12)
            double gi = 0.0, gj = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
            const bool far12 = r > 1.0e-12;
            if (far12) {
                const double ir = 1.0 / r;
                gi = shape_dw<K, DIM>(r * h1i) * fac_i * ir;
                const double facj = (DIM == 2 ? sig * h1j * h1j : sig * h1j * h1j * h1j) * h1j;
                gj = shape_dw<K, DIM>(r * h1j) * facj * ir;
                ex = dx * ir;
                ey = dy * ir;
                ez = dz * ir;
            }
            const double dwix = gi * dx, dwiy = gi * dy, dwiz = gi * dz;
            const double dwjx = gj * dx, dwjy = gj * dy, dwjz = gj * dz;

            // relative velocity, projections
            const double dvx = ui - uj, dvy = vi - vj, dvz = wi - wj;
            const double gri = omi * (ex * dwix + ey * dwiy + ez * dwiz);
            const double grj = omj * (ex * dwjx + ey * dwjy + ez * dwjz);
            const double dot = dvx * ex + dvy * ey + dvz * ez;

            // ---- MHDAccelerations (equations.py:217-275); unit vector is 0 below 1e-8
            {
                const bool near8 = r < 1.0e-8;
                const double dotA = near8 ? 0.0 : dot;
                const double griA = near8 ? 0.0 : gri;
                const double grjA = near8 ? 0.0 : grj;
                A.div += mj * dotA * griA;
                const double vsig2 = sqrt_nn(fabs(pi_ - pj) / (0.5 * (rhoi + rhoj)));
                double c = fmax(vwi - 2.0 * dotA, vwj - 2.0 * dotA);
                c = fmax(c, 0.0);
                A.dtcfl = fmax(A.dtcfl, c);
                double sx = pbi * omi, sjx = pbj * omj;  // P/rho^2 * omega
                if (dotA < 0.0) {
                    const double vsi = fmax(ali * vwi - 2.0 * dotA, 0.0);
                    const double vsj = fmax(alj * vwj - 2.0 * dotA, 0.0);
                    const double qi = -hiri * vsi * dotA * av;
                    const double qj = -hirj * vsj * dotA * av;
                    sx += qi * omi;
                    sjx += qj * omj;
                    A.ae += mj * qi * dotA * griA;
                }
                A.ax -= mj * (sx * dwix + sjx * dwjx);
                A.ay -= mj * (sx * dwiy + sjx * dwjy);
                A.az -= mj * (sx * dwiz + sjx * dwjz);
                A.ae += mj * pbi * omi * (dvx * dwix + dvy * dwiy + dvz * dwiz);
                A.ae += ac * vsig2 * (ei - ej) * mj * (hiri * griA + hirj * grjA);
            }
            // ---- MagneticStressAccelerations (equations.py:149-195)
            {
                const double gradp = mj * (0.5 * bro2i * gri + 0.5 * bro2j * grj);
                const double pi_b = (brxi * ex + bryi * ey + brzi * ez) * gri;
                const double pj_b = (brxj * ex + bryj * ey + brzj * ez) * grj;
                const double sm = mj * inv_mu0;
                A.ax += -ex * gradp + sm * (brxi * pi_b + brxj * pj_b);
                A.ay += -ey * gradp + sm * (bryi * pi_b + bryj * pj_b);
                A.az += -ez * gradp + sm * (brzi * pi_b + brzj * pj_b);
            }
            // ---- MagneticStateRates (equations.py:507-559)
            {
                const double pbi_ = bxi * ex + byi * ey + bzi * ez;
                const double pbj_ = bxj * ex + byj * ey + bzj * ez;
                const double dbx = bxi - bxj, dby = byi - byj, dbz = bzi - bzj;
                const double pdb = dbx * ex + dby * ey + dbz * ez;
                const double db2 = dbx * dbx + dby * dby + dbz * dbz;
                const double pmi = mj * r2ii * gri;
                const double pmj = mj * r2ij * grj;
                const double brt = -(pmi * pbi_);
                const double tx = dvx - dot * ex, ty = dvy - dot * ey, tz = dvz - dot * ez;
                const double vsb = sqrt_nn(tx * tx + ty * ty + tz * tz);
                const double diss = 0.5 * (pmi + pmj) * vsb * ar;
                const double dpsi = pmi * psichi + pmj * psichj;
                A.bx += brt * dvx + diss * dbx - dpsi * ex;
                A.by += brt * dvy + diss * dby - dpsi * ey;
                A.bz += brt * dvz + diss * dbz - dpsi * ez;
                A.ae += -0.5 * db2 * diss * inv_mu0;
                A.divbs += pmi * pbi_ + pmj * pbj_;
                A.divbd += -mj * pdb * gri;
            }
            // ---- CullenDehnenAlphaDiagnostics loop (equations.py:319-354)
            {
                const double rx = mj * dwix, ry = mj * dwiy, rz = mj * dwiz;
                const double dax = aui - auj, day = avi - avj;
                A.vxx += dvx * rx; A.vxy += dvx * ry;
                A.vyx += dvy * rx; A.vyy += dvy * ry;
                A.vzx += dvz * rx; A.vzy += dvz * ry;
                A.axx += dax * rx; A.ayy += day * ry;
                if constexpr (DIM == 3) {
                    const double daz = awi - awj;
                    A.rxx += -dx * rx; A.rxy += -dx * ry; A.rxz += -dx * rz;
                    A.ryy += -dy * ry; A.ryz += -dy * rz; A.rzz += -dz * rz;
                    A.vxz += dvx * rz; A.vyz += dvy * rz; A.vzz += dvz * rz;
                    A.axy += dax * ry; A.axz += dax * rz;
                    A.ayx += day * rx; A.ayz += day * rz;
                    A.azx += daz * rx; A.azy += daz * ry; A.azz += daz * rz;
                }
            }
        }
    }

    if (valid) {
        // ---- epilogue ------------------------------------------------------
        const int32_t i = perm[t];
        double** f = F.f;
        const double hi = f[SPHMHD_F_H][i];
        const double cs = f[SPHMHD_F_CS][i], psipred = f[SPHMHD_F_PSIPRED][i];
        // MHDAccelerations.post_loop (equations.py:279-280)
        const double div = -A.div / rhoi;
        // SmoothingLengthRateFromForceDivergence (equations.py:140)
        const double ah = hi * div * (1.0 / (double)DIM);

        // CullenDehnenAlphaDiagnostics.post_loop (equations.py:358-429)
        double div_a, dvxdx, dvxdy, dvxdz, dvydx, dvydy, dvydz, dvzdx, dvzdy, dvzdz;
        const double termnorm = omi / rhoi;
        bool inverted = false;
        if constexpr (DIM == 3) {
            const double rxx = A.rxx, rxy = A.rxy, rxz = A.rxz, ryy = A.ryy, ryz = A.ryz, rzz = A.rzz;
            const double denom = rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
            if (fabs(denom) > 2.2250738585072014e-308) {
                inverted = true;
                const double dd = 1.0 / denom;
                const double txx = ryy * rzz - ryz * ryz;
                const double txy = rxz * ryz - rzz * rxy;
                const double txz = rxy * ryz - rxz * ryy;
                const double tyy = rzz * rxx - rxz * rxz;
                const double tyz = rxy * rxz - rxx * ryz;
                const double tzz = rxx * ryy - rxy * rxy;
                const double gax = (A.axx * txx + A.axy * txy + A.axz * txz) * dd;
                const double gay = (A.ayx * txy + A.ayy * tyy + A.ayz * tyz) * dd;
                const double gaz = (A.azx * txz + A.azy * tyz + A.azz * tzz) * dd;
                div_a = -(gax + gay + gaz);
                dvxdx = -(A.vxx * txx + A.vxy * txy + A.vxz * txz) * dd;
                dvxdy = -(A.vxx * txy + A.vxy * tyy + A.vxz * tyz) * dd;
                dvxdz = -(A.vxx * txz + A.vxy * tyz + A.vxz * tzz) * dd;
                dvydx = -(A.vyx * txx + A.vyy * txy + A.vyz * txz) * dd;
                dvydy = -(A.vyx * txy + A.vyy * tyy + A.vyz * tyz) * dd;
                dvydz = -(A.vyx * txz + A.vyy * tyz + A.vyz * tzz) * dd;
                dvzdx = -(A.vzx * txx + A.vzy * txy + A.vzz * txz) * dd;
                dvzdy = -(A.vzx * txy + A.vzy * tyy + A.vzz * tyz) * dd;
                dvzdz = -(A.vzx * txz + A.vzy * tyz + A.vzz * tzz) * dd;
            }
        }
        if (!inverted) {
            // 2-D: the 3x3 determinant of the reference is identically zero, so
            // its fallback branch (equations.py:400-410) always applies.
            div_a = -termnorm * (A.axx + A.ayy + A.azz);
            dvxdx = -termnorm * A.vxx; dvxdy = -termnorm * A.vxy; dvxdz = -termnorm * A.vxz;
            dvydx = -termnorm * A.vyx; dvydy = -termnorm * A.vyy; dvydz = -termnorm * A.vyz;
            dvzdx = -termnorm * A.vzx; dvzdy = -termnorm * A.vzy; dvzdz = -termnorm * A.vzz;
        }
        const double divvdt = div_a - (dvxdx * dvxdx + dvydy * dvydy + dvzdz * dvzdz +
                                       2.0 * (dvxdy * dvydx + dvxdz * dvzdx + dvydz * dvzdy));
        // StrainDiagnostics (equations.py:432-471): the same velocity
        // gradient from the same sums, on request (rows bound) and with the
        // other diagnostics
        if (!lean && strain.d[0]) {
            strain.d[0][i] = dvxdx; strain.d[1][i] = dvxdy; strain.d[2][i] = dvxdz;
            strain.d[3][i] = dvydx; strain.d[4][i] = dvydy; strain.d[5][i] = dvydz;
            strain.d[6][i] = dvzdx; strain.d[7][i] = dvzdy; strain.d[8][i] = dvzdz;
        }
        const double divv = dvxdx + dvydy + dvzdz;
        const double cvx = dvzdy - dvydz, cvy = dvxdz - dvzdx, cvz = dvydx - dvxdy;
        const double comp = fmax(-divv, 0.0);
        const double fc = comp * comp;
        const double den = fc + (cvx * cvx + cvy * cvy + cvz * cvz);
        const double xil = den > 0.0 ? fc / den : 1.0;
        const double source = 10.0 * hi * hi * xil * fmax(-divvdt, 0.0);
        const double alphaloc = fmax(s.alpha_min, s.alpha_max * fmax(fmin(source / (cs * cs), 1.0), 0.0));

        // MagneticStateRates.post_loop (equations.py:563-567)
        const double ch = s.chf * vwi;
        const double apsi = -ch * A.divbd / rhoi - s.sigma * psipred * ch / hi -
                            0.5 * s.cleaning_active * psipred * div;
        double dtcfl = fmax(A.dtcfl, vwi);
        dtcfl = fmax(dtcfl, ch);

        // EnergyLimiter (equations.py:595-598)
        double ae = A.ae;
        if (dtcfl > 0.0) {
            const double dtc = s.c_cour * hi / dtcfl;
            if (ei + dtc * ae < 0.0 && ei > 0.0) ae /= 1.0 - dtc * ae / ei;
        }
        // DivBCorrection (equations.py:608-624)
        const double b2 = bxi * bxi + byi * byi + bzi * bzi;
        const double divb = A.divbs;
        double frac = 0.0;
        if (b2 > 0.0) {
            const double beta = 2.0 * s.mu0 * pi_ / b2;
            frac = fmax(0.0, fmin(1.0, (10.0 - beta) * 0.125));
        }
        const double corr = divb * frac * inv_mu0;
        const double ax = A.ax - bxi * corr, ay = A.ay - byi * corr, az = A.az - bzi * corr;
        // AdaptiveTimestep (equations.py:579-585)
        dt_adapt = 1.0e29;
        if (dtcfl > 0.0) dt_adapt = s.c_cour * hi / dtcfl;
        const double f2 = ax * ax + ay * ay + az * az;
        if (f2 > 0.0) dt_adapt = fmin(dt_adapt, s.c_force * sqrt(hi / sqrt(f2)));
        bad = !isfinite(f2) || !isfinite(ae);

        // the rates the stages and the next evaluation read
        f[SPHMHD_F_AU][i] = ax;
        f[SPHMHD_F_AV][i] = ay;
        f[SPHMHD_F_AW][i] = az;
        f[SPHMHD_F_AE][i] = ae;
        f[SPHMHD_F_AH][i] = ah;
        f[SPHMHD_F_ABEVOLX][i] = A.bx;
        f[SPHMHD_F_ABEVOLY][i] = A.by;
        f[SPHMHD_F_ABEVOLZ][i] = A.bz;
        f[SPHMHD_F_APSI][i] = apsi;
        f[SPHMHD_F_ALPHALOC][i] = alphaloc;
        // the reference's diagnostics, which no later kernel reads: skipped
        // in the evaluations of a multi-step run that are not its last
        if (!lean) {
            f[SPHMHD_F_DIV][i] = div;
            f[SPHMHD_F_DIV_FOR_PSI][i] = div;
            f[SPHMHD_F_DT_CFL][i] = dtcfl;
            f[SPHMHD_F_DT_ADAPT][i] = dt_adapt;
            f[SPHMHD_F_DIVBSYMM][i] = rhoi * divb;
            f[SPHMHD_F_DIVBDIFF][i] = A.divbd;
            f[SPHMHD_F_XILIMITER][i] = xil;
            f[SPHMHD_F_DIVVDT][i] = divvdt;
        }
    }

    // min dt_adapt and the NaN flag per instance: wave, then block, then one
    // atomic per instance run of the block (same-address atomics serialise)
    constexpr int NW = kBlock / kWave;  // LDS slots: waves of the largest shape
    const int nw = blockDim.x / kWave;
    __shared__ unsigned long long wdb[NW];
    __shared__ int wk[NW], wbad[NW];
    const int w = threadIdx.x / kWave;
    unsigned long long db = dbits(dt_adapt);
    const int k0 = __shfl(ki, 0);  // lane 0 is valid whenever any lane is
    const bool uniform = __all(ki == k0 || !valid);
    const bool anybad = __any(bad);
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(db, off);
            db = o < db ? o : db;
        }
    } else if (valid) {
        atomicMin(&inst[ki].dtp_bits, db);
        if (bad) atomicOr(&inst[ki].status, 2);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        wk[w] = uniform && valid ? k0 : -1;
        wdb[w] = db;
        wbad[w] = anybad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < nw;) {
            const int k = wk[q];
            unsigned long long m = wdb[q];
            int b = wbad[q];
            int r = q + 1;
            for (; r < nw && wk[r] == k; ++r) {
                m = wdb[r] < m ? wdb[r] : m;
                b |= wbad[r];
            }
            if (k >= 0) {
                atomicMin(&inst[k].dtp_bits, m);
                if (b) atomicOr(&inst[k].status, 2);
            }
            q = r;
        }
    }
}

// end of an evaluation (one workgroup): the instances' warning bits (1:
// density hit max_iterations, 2: NaN) into ctl->status, whether the step
// being evaluated (clock t, dt) ends at or past every instance's tf, then ctl
// into its host mirror, which the host reads after the stream's wait
__device__ void eval_end_body(const Inst* inst, int ninst, Ctl* ctl, CtlView* hout) {
    __shared__ int wst[kBlock / kWave], wfin[kBlock / kWave];
    int st = 0, fin = 1;
    for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
        st |= ld_dev(&inst[k].status);  // force pass's atomics
        if (!(inst[k].t + inst[k].dt >= inst[k].tf)) fin = 0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        st |= __shfl_xor(st, off);
        fin &= __shfl_xor(fin, off);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        wst[threadIdx.x / kWave] = st;
        wfin[threadIdx.x / kWave] = fin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)blockDim.x / kWave; ++w) {
            st |= wst[w];
            fin &= wfin[w];
        }
        CtlView v = ctl_view(*ctl);
        v.status |= st;
        v.at_tf = fin;
        ctl->at_tf = fin;
        ctl->status = v.status;
        // more passes or longer lists needed, or the run ends with this step:
        // the step enqueued next exits
        v.abort = (!v.all_done || v.overflow || fin) ? 1 : 0;
        ctl->abort = v.abort;
        ctl->seq = v.seq + 1;
        hout[v.seq & 1] = v;
    }
}

__global__ void __launch_bounds__(kBlock) k_eval_end(const Inst* __restrict__ inst, int ninst, Ctl* ctl,
                                                     CtlView* hout) {
    eval_end_body(inst, ninst, ctl, hout);
}

// the control block into the host mirror's third slot (synchronous reads)
__global__ void k_ctl_export(const Ctl* ctl, CtlView* hout) { hout[2] = ctl_view(*ctl); }

__global__ void k_clear_abort(Ctl* ctl) { ctl->abort = 0; }

__global__ void k_set_lean(Ctl* ctl, int v) { ctl->lean = v; }

// the fused force pass; its last workgroup ends the evaluation
template <int K, int DIM>
__global__ void __launch_bounds__(kBlock, DIM == 3 ? 2 : 1) k_force(Fields F, const double* __restrict__ fpack,
                                                  const int32_t* __restrict__ perm, const int32_t* __restrict__ nb,
                                                  const int32_t* __restrict__ nbcnt, Inst* inst,
                                                  const int32_t* __restrict__ inst_of, int64_t n, int stride,
                                                  int ninst, Ctl* ctl, CtlView* hout, int32_t* arrive,
                                                  Strain strain, int32_t* queue) {
    if (ctl->abort) return;  // the whole launch: the step after an unfinished evaluation
    const int B = blockDim.x;
    const int lean = ctl->lean;
    if (queue) {  // persistent: tiles from the work queues
        int polls = 0;
        for (int64_t tile; (tile = pull_tile(queue, (n + B - 1) / B, polls)) >= 0;)
            force_block<K, DIM>(F, fpack, perm, nb, nbcnt, inst, inst_of, n, stride, lean, strain, tile);
    } else {
        force_block<K, DIM>(F, fpack, perm, nb, nbcnt, inst, inst_of, n, stride, lean, strain,
                            xcd_tile((n + B - 1) / B));
    }
    if (last_to_arrive(arrive)) {
        eval_end_body(inst, ninst, ctl, hout);
        if (queue) reset_queues(queue);
    }
}

// ---------------------------------------------------------------------------
// launchers (kernel x dim dispatch)
// ---------------------------------------------------------------------------

static inline unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

unsigned persistent_grid(const void* fn, int B, size_t lds, int64_t tiles) {
    // resident blocks per CU of each (kernel, shape), queried once
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, int> per_cu;
    static int cus = 0;
    int blocks = 1;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (cus == 0) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
                cus = 256;
        }
        const auto key = std::make_tuple(fn, B, lds);
        auto it = per_cu.find(key);
        if (it == per_cu.end()) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, B, lds) != hipSuccess || nb < 1) nb = 1;
            it = per_cu.emplace(key, nb).first;
        }
        blocks = it->second;
    }
    int64_t g = (int64_t)blocks * cus;
    g = std::max<int64_t>(kXcds, g / kXcds * kXcds);
    return (unsigned)std::min<int64_t>(g, xcd_grid(tiles));
}

#define SPH_DISPATCH(KID, DIMV, CALL)                                                    \
    switch (KID) {                                                                       \
    case SPHMHD_CUBIC: if (DIMV == 2) CALL(SPHMHD_CUBIC, 2) else CALL(SPHMHD_CUBIC, 3) break; \
    case SPHMHD_QUINTIC: if (DIMV == 2) CALL(SPHMHD_QUINTIC, 2) else CALL(SPHMHD_QUINTIC, 3) break; \
    case SPHMHD_WENDLAND_C2: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C2, 2) else CALL(SPHMHD_WENDLAND_C2, 3) break; \
    case SPHMHD_WENDLAND_C4: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C4, 2) else CALL(SPHMHD_WENDLAND_C4, 3) break; \
    case SPHMHD_WENDLAND_C6: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C6, 2) else CALL(SPHMHD_WENDLAND_C6, 3) break; \
    case SPHMHD_GAUSSIAN: if (DIMV == 2) CALL(SPHMHD_GAUSSIAN, 2) else CALL(SPHMHD_GAUSSIAN, 3) break; \
    default: if (DIMV == 2) CALL(SPHMHD_SUPER_GAUSSIAN, 2) else CALL(SPHMHD_SUPER_GAUSSIAN, 3) break; \
    }

// pass 0 scans every particle and writes list 0 (with dens_fused it iterates
// them to the end and lists none); pass p > 0 reads list (p - 1) % 2 and
// writes list p % 2; the pass ends with the group control (the next pass
// writes list (p + 1) % 2)
void launch_density(Dev& d) {
    if (d.n == 0) return;
    const int p = d.pass;
    const int32_t* ain = p == 0 ? nullptr : d.act[(p - 1) & 1];
    const int32_t* cin = p == 0 ? nullptr : act_count(d.ctl, (p - 1) & 1);
    int32_t* aout = d.act[p & 1];
    int32_t* cout = act_count(d.ctl, p & 1);
    const int B = d.blk[LC_DENSITY];
    int32_t* q = d.persist[LC_DENSITY] ? d.queue : nullptr;
#define CALL_DENS_L(KK, DD, LL, FF)                                                                         \
    { const unsigned g = q ? persistent_grid((const void*)k_density<KK, DD, LL, FF>, B, 0, nblk(d.n, B))    \
                           : xcd_grid(nblk(d.n, B));                                                        \
      hipLaunchKernelGGL((k_density<KK, DD, LL, FF>), dim3(g), dim3(B), 0, d.stream, d.fields,              \
                         (const double4*)d.dpack, d.perm, d.keys, d.cstart, d.cend, d.inst, d.inst_of, d.n, \
                         ain, cin, aout, cout, d.ctl, d.ninst, (p + 1) & 1, d.arrive, q, d.dens_trim); }
#define CALL_DENS(KK, DD)                                                                                   \
    { if (p > 0) CALL_DENS_L(KK, DD, true, false)                                                           \
      else if (d.dens_fused) CALL_DENS_L(KK, DD, false, true)                                               \
      else CALL_DENS_L(KK, DD, false, false) }
    SPH_DISPATCH(d.kernel, d.dim, CALL_DENS)
#undef CALL_DENS
#undef CALL_DENS_L
}

void launch_g0_control(Dev& d) {
    // the next pass (d.pass + 1) writes list (d.pass + 1) % 2
    hipLaunchKernelGGL(k_g0_control, dim3(1), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl, (d.pass + 1) & 1);
}

void launch_prep_force(Dev& d, bool keep_prev) {
    if (d.n == 0) return;
    const int B = d.blk[LC_LISTS];
    hipLaunchKernelGGL(k_prep_force, dim3(xcd_grid(nblk(d.n, B))), dim3(B), 0, d.stream, d.fields, d.perm, d.inst,
                       d.inst_of, d.fpack, (double4*)d.dpack, d.n, d.aprev, keep_prev ? 1 : 0,
                       d.ctl);
}

void launch_eval_end(Dev& d) {
    hipLaunchKernelGGL(k_eval_end, dim3(1), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl, d.hctl_dev);
}

void launch_ctl_export(Dev& d) {
    hipLaunchKernelGGL(k_ctl_export, dim3(1), dim3(1), 0, d.stream, d.ctl, d.hctl_dev);
}

void launch_clear_abort(Dev& d) { hipLaunchKernelGGL(k_clear_abort, dim3(1), dim3(1), 0, d.stream, d.ctl); }

void launch_set_lean(Dev& d, int lean) { hipLaunchKernelGGL(k_set_lean, dim3(1), dim3(1), 0, d.stream, d.ctl, lean); }

void launch_force(Dev& d) {
    if (d.n == 0) return;
    const int B = d.blk[LC_FORCE];
    int32_t* q = d.persist[LC_FORCE] ? d.queue : nullptr;
#define CALL_FORCE(KK, DD)                                                                                  \
    { const unsigned g = q ? persistent_grid((const void*)k_force<KK, DD>, B, 0, nblk(d.n, B))              \
                           : xcd_grid(nblk(d.n, B));                                                        \
      hipLaunchKernelGGL((k_force<KK, DD>), dim3(g), dim3(B), 0, d.stream, d.fields, d.fpack,               \
                         d.perm, d.nb, d.nbcnt, d.inst, d.inst_of, d.n, d.nb_stride, d.ninst, d.ctl,          \
                         d.hctl_dev, d.arrive, d.strain, q); }
    SPH_DISPATCH(d.kernel, d.dim, CALL_FORCE)
#undef CALL_FORCE
}

}  // namespace sph
