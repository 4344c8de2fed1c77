// sph_tangent.hip -- the tangent-linear (forward-mode) step.
//
// Every quantity of the step is carried as Dual<SPHMHD_NTANGENTS>: its value
// and its derivative along three directions of the inputs (initial state and
// the parameters of sphmhd_tparam).  The kernels restate the primal bodies of
// sph_pairs.hip and sph_stage.hip expression for expression over the dual
// type; every branch, clamp and min/max acts on the value, so the derivative
// is that of the discrete map the primal step computes (density Newton
// iteration unrolled as run, neighbour sets and cells held fixed, adaptive dt
// through the particle that attains the minimum).  The neighbour search and
// list build are the primal kernels, run on the values.
//
// Storage: the value of every field is the bound primal row; its three
// tangents are rows of the caller's tangent storage (TFields).  Per-instance
// tangents (dt, pending dt, parameters) live in InstTan.
#include <climits>
#include <type_traits>

#include "sph_dual.h"
#include "sph_internal.h"

namespace sph {

using D3 = Dual<kTanDirs>;

__device__ __forceinline__ D3 ldT(const Fields& F, const TFields& G, int k, int64_t i) {
    D3 r(F.f[k][i]);
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) r.d[q] = G.f[q][k][i];
    return r;
}

__device__ __forceinline__ void stT(const Fields& F, const TFields& G, int k, int64_t i, const D3& x) {
    F.f[k][i] = x.v;
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) G.f[q][k][i] = x.d[q];
}

__device__ __forceinline__ D3 parT(double v, const InstTan& it, int p) {
    D3 r(v);
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) r.d[q] = it.par[p][q];
    return r;
}

// loads and stores of ND directions starting at direction d0
template <int ND>
__device__ __forceinline__ Dual<ND> ldN(const Fields& F, const TFields& G, int k, int64_t i, int d0) {
    Dual<ND> r(F.f[k][i]);
#pragma unroll
    for (int q = 0; q < ND; ++q) r.d[q] = G.f[d0 + q][k][i];
    return r;
}

template <int ND>
__device__ __forceinline__ void stN(const Fields& F, const TFields& G, int k, int64_t i, const Dual<ND>& x, int d0,
                                    bool value) {
    if (value) F.f[k][i] = x.v;
#pragma unroll
    for (int q = 0; q < ND; ++q) G.f[d0 + q][k][i] = x.d[q];
}

template <int ND>
__device__ __forceinline__ Dual<ND> parN(double v, const InstTan& it, int p, int d0) {
    Dual<ND> r(v);
#pragma unroll
    for (int q = 0; q < ND; ++q) r.d[q] = it.par[p][d0 + q];
    return r;
}

template <class T>
__device__ __forceinline__ T ipow_dim_t(const T& x, int dim) { return dim == 2 ? x * x : x * x * x; }
template <class T>
__device__ __forceinline__ T iroot_dim_t(const T& x, int dim) { return dim == 2 ? sqrt(x) : cbrt(x); }

#define TFLD(name) SPHMHD_F_##name

// ---------------------------------------------------------------------------
// bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_t_eval_begin(InstTan* it, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ninst) it[k].argmin = INT_MAX;
}

// ---------------------------------------------------------------------------
// G0: density sum + Newton step (k_density over duals)
// ---------------------------------------------------------------------------

// Three lanes per particle, one tangent direction each (21 particles per
// wave, the last lane idle): the lanes of a particle scan the same candidates
// in lockstep (the candidate loads broadcast), every lane reads the old h,
// omega and converged before any lane writes, and only the direction-0 lane
// stores the values and joins the primal's list and reductions.
constexpr int kTPW = kWave / kTanDirs;  // particles per wave

template <int K, int DIM>
__global__ void __launch_bounds__(kBlock) k_t_density(Fields F, TFields G, const double4* __restrict__ dpack,
                                                      const int32_t* __restrict__ perm,
                                                      const uint32_t* __restrict__ skey,
                                                      const int32_t* __restrict__ cstart,
                                                      const int32_t* __restrict__ cend, Inst* inst,
                                                      const int32_t* __restrict__ inst_of, int64_t n,
                                                      const int32_t* __restrict__ act_in, const int32_t* cnt_in,
                                                      int32_t* __restrict__ act_out, int32_t* cnt_out) {
    using D1 = Dual<1>;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const int slot = lane / kTanDirs, dir = lane - kTanDirs * slot;
    const int64_t idx0 = (int64_t)blockIdx.x * nw * kTPW;
    const int64_t items = act_in ? (int64_t)*cnt_in : n;
    if (idx0 >= items) return;  // whole block past the end (the block barriers below need every wave)
    const int64_t idx = idx0 + (int64_t)w * kTPW + slot;
    const bool valid = slot < kTPW && idx < items;
    const bool lead = dir == 0;
    const int64_t t = valid ? (act_in ? (int64_t)act_in[idx] : idx) : 0;
    const int ki = valid ? inst_of[t] : 0;
    const Inst& s = inst[ki];
    const bool live = valid && !s.done;
    const int32_t i = valid ? perm[t] : 0;
    D1 hi = live ? ldN<1>(F, G, TFLD(H), i, dir) : D1(0.0);
    const bool active = live && F.f[SPHMHD_F_CONVERGED][i] != 1.0;
    D1 hout = hi;
    bool unconv = false;

    if (active) {
        const double R = kernel_radius(K);
        const double4 me = dpack[t];
        const D1 xi = ldN<1>(F, G, TFLD(X), i, dir), yi = ldN<1>(F, G, TFLD(Y), i, dir);
        const D1 zi = DIM == 3 ? ldN<1>(F, G, TFLD(Z), i, dir) : D1(0.0);
        const D1 h1 = 1.0 / hi;
        const double Rh = R * hi.v;
        const double Rh2 = Rh * Rh;
        const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
        const double H0 = s.halfL[0], H1 = s.halfL[1], H2 = s.halfL[2];
        const int per = s.periodic;
        // this particle's own stencil, as the primal pass: every j with
        // r < R h_i lies within ceil(R h_i / c_eff) cells of it
        int sr[3] = {0, 0, 0};
        for (int a = 0; a < DIM; ++a) {
            const double q = ceil(Rh / s.ceff[a]);
            sr[a] = q < 1.0 ? 1 : (q < (double)s.srad[a] ? (int)q : s.srad[a]);
        }
        D1 sw(0.0), sg(0.0);
        // the value test as the primal pass (range shifts where exact)
        auto scan = [&](auto shift_c) {
            constexpr bool kShift = decltype(shift_c)::value;
            for_each_row_range<DIM>(s, (int64_t)skey[t], cstart, cend, [&](int a, int b, double sx, double sy,
                                                                          double sz) {
                for (int j = a; j < b; ++j) {
                    const double4 o = dpack[j];
                    const double ddx = me.x - o.x, ddy = me.y - o.y, ddz = me.z - o.z;
                    const double dxv = kShift ? ddx + sx : min_image(ddx, L0, H0, per);
                    const double dyv = kShift ? ddy + sy : min_image(ddy, L1, H1, per);
                    const double dzv = DIM == 3 ? (kShift ? ddz + sz : min_image(ddz, L2, H2, per)) : 0.0;
                    if (r2_exact(dxv, dyv, dzv) < Rh2) {
                        const int32_t jj = perm[j];
                        // the minimum image shifts by a constant: the tangent is x_i' - x_j'
                        const D1 dx = (xi - ldN<1>(F, G, TFLD(X), jj, dir)) + (dxv - ddx);
                        const D1 dy = (yi - ldN<1>(F, G, TFLD(Y), jj, dir)) + (dyv - ddy);
                        const D1 dz = DIM == 3 ? (zi - ldN<1>(F, G, TFLD(Z), jj, dir)) + (dzv - ddz) : D1(0.0);
                        const D1 q = sqrt(dx * dx + dy * dy + dz * dz) * h1;
                        D1 w1, dw;
                        shape<K, DIM>(q, w1, dw);
                        sw += o.w * w1;
                        sg += o.w * ((double)DIM * w1 + q * dw);
                    }
                }
            }, sr);
        };
        if (s.range_shift) scan(std::true_type{});
        else scan(std::false_type{});
        const D1 fac = DIM == 2 ? kernel_sigma(K, 2) * h1 * h1 : kernel_sigma(K, 3) * h1 * h1 * h1;
        const D1 rho_sum = sw * fac;
        const D1 dwdh = -h1 * (fac * sg);

        // DensityIteration.post_loop (equations.py:54-86)
        const double mi = me.w;
        const D1 hi0 = ldN<1>(F, G, TFLD(H0), i, dir);
        const D1 rhoi = mi / ipow_dim_t(hi / s.hfact, DIM);
        const D1 dhdrhoi = -hi / ((double)DIM * rhoi);
        const D1 omegai = 1.0 - dhdrhoi * dwdh;
        const D1 func = rhoi - rho_sum;
        const D1 dfdh1 = omegai > 0.0 ? dhdrhoi / omegai : dhdrhoi / fabs(omegai + 2.220446049250313e-16);
        D1 hnew = hi - func * dfdh1;
        if (hnew > 1.2 * hi) hnew = 1.2 * hi;
        else if (hnew < 0.8 * hi) hnew = 0.8 * hi;
        const double diff = ::fabs(hnew.v - hi.v) / hi0.v;
        const bool fin = ((diff < 1.0e-4) && (omegai.v > 0.0) && (hi.v > 0.0)) || !s.density_iterate;
        if (!fin) {
            hout = hnew;
            unconv = true;
        } else {
            hout = s.hfact * iroot_dim_t(mi / rho_sum, DIM);
        }
        // every lane of the wave has read the old values (lockstep): store
        __builtin_amdgcn_wave_barrier();
        stN<1>(F, G, TFLD(OMEGA), i, 1.0 / omegai, dir, lead);
        if (lead) F.f[SPHMHD_F_CONVERGED][i] = fin ? 1.0 : 0.0;
        stN<1>(F, G, TFLD(H), i, hout, dir, lead);
    }

    // the primal's list and reductions, on the values (direction-0 lanes)
    const bool mine = live && lead;
    constexpr int NW = kBlock / kWave;
    __shared__ int wcnt[NW], wbase[NW], wk[NW], wuc[NW];
    __shared__ unsigned long long whc[NW], whu[NW];
    const unsigned long long um = __ballot(unconv && lead);
    const bool conv_now = mine && !unconv;
    unsigned long long hc = conv_now ? dbits(hout.v) : 0ull;
    unsigned long long hu = unconv && lead ? dbits(hout.v) : 0ull;
    int uc = unconv && lead ? 1 : 0;
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
    } else if (mine) {
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
    if (unconv && lead) {
        const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        act_out[wbase[w] + __popcll(um & lt)] = (int32_t)t;
    }
}

// ---------------------------------------------------------------------------
// G1 + G2 + G4 (k_prep_force over duals), the {x,y,z,h} list pack and a
// snapshot of the previous accelerations the force pass reads from neighbours
// ---------------------------------------------------------------------------

__global__ void k_t_prep(Fields F, TFields G, const int32_t* __restrict__ perm, const Inst* __restrict__ inst,
                         const int32_t* __restrict__ inst_of, double4* __restrict__ xh, double* __restrict__ aprev,
                         int64_t n, int keep_prev) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Inst& s = inst[inst_of[t]];
    const int32_t i = perm[t];
    const int dim = s.dim;
    const D3 h = ldT(F, G, TFLD(H), i);
    const double m = F.f[SPHMHD_F_M][i];
    const D3 rho = m / ipow_dim_t(h / s.hfact, dim);
    const D3 epred = ldT(F, G, TFLD(EPRED), i);
    const D3 p = s.gamma1 * rho * epred;
    const D3 cs = sqrt(s.gamma * p / rho);
    const D3 bx = ldT(F, G, TFLD(BEVOLXPRED), i) * rho;
    const D3 by = ldT(F, G, TFLD(BEVOLYPRED), i) * rho;
    const D3 bz = ldT(F, G, TFLD(BEVOLZPRED), i) * rho;
    stT(F, G, TFLD(RHO), i, rho);
    stT(F, G, TFLD(P), i, p);
    stT(F, G, TFLD(CS), i, cs);
    stT(F, G, TFLD(BX), i, bx);
    stT(F, G, TFLD(BY), i, by);
    stT(F, G, TFLD(BZ), i, bz);
    stT(F, G, TFLD(BXPRED), i, bx);
    stT(F, G, TFLD(BYPRED), i, by);
    stT(F, G, TFLD(BZPRED), i, bz);
    const double z = dim == 3 ? F.f[SPHMHD_F_Z][i] : 0.0;
    xh[t] = make_double4(F.f[SPHMHD_F_X][i], F.f[SPHMHD_F_Y][i], z, h.v);
    // the previous accelerations, with tangents, as the first prep of the
    // evaluation found them (a redone force pass reads the same ones)
    if (keep_prev) return;
    const int ka[3] = {TFLD(AU), TFLD(AV), TFLD(AW)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const D3 a = ldT(F, G, ka[c], i);
        aprev[(int64_t)(c * (1 + kTanDirs)) * n + i] = a.v;
#pragma unroll
        for (int q = 0; q < kTanDirs; ++q) aprev[(int64_t)(c * (1 + kTanDirs) + 1 + q) * n + i] = a.d[q];
    }
}

__device__ __forceinline__ D3 ld_aprev(const double* __restrict__ aprev, int c, int64_t n, int64_t i) {
    D3 r(aprev[(int64_t)(c * (1 + kTanDirs)) * n + i]);
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) r.d[q] = aprev[(int64_t)(c * (1 + kTanDirs) + 1 + q) * n + i];
    return r;
}

// ---------------------------------------------------------------------------
// G3 + G5 + G7 + G8 pair loops and G3/G5/G6/G8/G9 epilogues (k_force over duals)
// ---------------------------------------------------------------------------

template <int ND>
__device__ __forceinline__ Dual<ND> ld_aprevN(const double* __restrict__ aprev, int c, int64_t n, int64_t i, int d0) {
    Dual<ND> r(aprev[(int64_t)(c * (1 + kTanDirs)) * n + i]);
#pragma unroll
    for (int q = 0; q < ND; ++q) r.d[q] = aprev[(int64_t)(c * (1 + kTanDirs) + 1 + d0 + q) * n + i];
    return r;
}

template <int DIM, class D3>
struct AccT {
    D3 ax, ay, az, ae, div, dtcfl;
    D3 bx, by, bz, divbs, divbd;
    D3 rxx, rxy, rxz, ryy, ryz, rzz;
    D3 vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz;
    D3 axx, axy, axz, ayx, ayy, ayz, azx, azy, azz;
};

// the neighbour-side values the pair loop reads (the force-pack record's
// contents, rebuilt from the fields with the prep's arithmetic)
template <class D3>
struct SideT {
    D3 x, y, z, h1, m, rho, p, e, u, v, w, om, bx, by, bz, al, au, av, aw, psich, vw, hir;
};

template <int DIM, int ND>
__device__ __forceinline__ SideT<Dual<ND>> load_side(const Fields& F, const TFields& G,
                                                     const double* __restrict__ aprev, int64_t n, int32_t i,
                                                     const Inst& s, const Dual<ND>& chf, int d0) {
    using D3 = Dual<ND>;
    SideT<D3> o;
    o.x = ldN<ND>(F, G, TFLD(X), i, d0);
    o.y = ldN<ND>(F, G, TFLD(Y), i, d0);
    o.z = DIM == 3 ? ldN<ND>(F, G, TFLD(Z), i, d0) : D3(0.0);
    const D3 h = ldN<ND>(F, G, TFLD(H), i, d0);
    o.h1 = 1.0 / h;
    o.m = D3(F.f[SPHMHD_F_M][i]);
    o.rho = ldN<ND>(F, G, TFLD(RHO), i, d0);
    o.e = ldN<ND>(F, G, TFLD(EPRED), i, d0);
    o.p = s.gamma1 * o.rho * o.e;
    o.u = ldN<ND>(F, G, TFLD(UPRED), i, d0);
    o.v = ldN<ND>(F, G, TFLD(VPRED), i, d0);
    o.w = ldN<ND>(F, G, TFLD(WPRED), i, d0);
    o.om = ldN<ND>(F, G, TFLD(OMEGA), i, d0);
    o.bx = ldN<ND>(F, G, TFLD(BX), i, d0);
    o.by = ldN<ND>(F, G, TFLD(BY), i, d0);
    o.bz = ldN<ND>(F, G, TFLD(BZ), i, d0);
    o.al = ldN<ND>(F, G, TFLD(ALPHA1), i, d0);
    o.au = ld_aprevN<ND>(aprev, 0, n, i, d0);
    o.av = ld_aprevN<ND>(aprev, 1, n, i, d0);
    o.aw = ld_aprevN<ND>(aprev, 2, n, i, d0);
    const D3 cs = ldN<ND>(F, G, TFLD(CS), i, d0);
    const D3 b2 = o.bx * o.bx + o.by * o.by + o.bz * o.bz;
    o.vw = sqrt(cs * cs + b2 / (s.mu0 * o.rho));
    o.psich = ldN<ND>(F, G, TFLD(PSIPRED), i, d0) * (chf * o.vw);
    o.hir = 0.5 * (1.0 / o.rho);
    return o;
}

// ND directions per workgroup, the directions spread over blockIdx.y: with
// one direction the dual accumulators fit in registers (three would spill),
// and small systems get three times the workgroups.  Only the direction-0
// workgroups store values and reduce the pending dt.
template <int K, int DIM, int ND>
__global__ void __launch_bounds__(kBlock) k_t_force(Fields F, TFields G, const double* __restrict__ aprev,
                                                    const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ nb,
                                                    const int32_t* __restrict__ nbcnt, Inst* inst,
                                                    const InstTan* __restrict__ itan,
                                                    const int32_t* __restrict__ inst_of, int64_t n, int stride) {
    using D3 = Dual<ND>;
    const int d0 = blockIdx.y * ND;
    const bool lead = blockIdx.y == 0;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x;  // 64, 128 or 256 (the force pass's tuned shape)
    if (t0 >= n) return;
    const int64_t t = t0 + threadIdx.x;
    const bool valid = t < n;
    const int ki = valid ? inst_of[t] : 0;
    const Inst& s = inst[ki];
    const InstTan& it = itan[ki];
    D3 dt_adapt(1.0e300);
    bool bad = false;
    if (valid) {
        const int32_t i = perm[t];
        const D3 av = parN<ND>(s.av, it, SPHMHD_TP_ARTIFICIAL_VISCOSITY, d0);
        const D3 ac = parN<ND>(s.ac, it, SPHMHD_TP_ARTIFICIAL_THERMAL_CONDUCTIVITY, d0);
        const D3 ar = parN<ND>(s.ar, it, SPHMHD_TP_ARTIFICIAL_MAGNETIC_DISSIPATION, d0);
        const D3 chf = parN<ND>(s.chf, it, SPHMHD_TP_CLEANING_SPEED_FACTOR, d0);
        const D3 sigma = parN<ND>(s.sigma, it, SPHMHD_TP_CLEANING_DAMPING_FACTOR, d0);
        const D3 alpha_max = parN<ND>(s.alpha_max, it, SPHMHD_TP_ALPHA_MAX, d0);
        const SideT<D3> me = load_side<DIM, ND>(F, G, aprev, n, i, s, chf, d0);
        const D3 xi = me.x, yi = me.y, zi = me.z, h1i = me.h1;
        const D3 rhoi = me.rho, pi_ = me.p, ei = me.e;
        const D3 ui = me.u, vi = me.v, wi = me.w, omi = me.om, ali = me.al;
        const D3 bxi = me.bx, byi = me.by, bzi = me.bz;
        const D3 aui = me.au, avi = me.av, awi = me.aw;
        const D3 psichi = me.psich, vwi = me.vw, hiri = me.hir;
        const double inv_mu0 = 1.0 / s.mu0;
        const D3 irhoi = 2.0 * hiri;
        const D3 r2ii = irhoi * irhoi;
        const D3 pbi = pi_ * r2ii;
        const D3 brxi = bxi * irhoi, bryi = byi * irhoi, brzi = bzi * irhoi;
        const D3 bro2i = (brxi * brxi + bryi * bryi + brzi * brzi) * inv_mu0;
        const double sig = DIM == 2 ? kernel_sigma(K, 2) : kernel_sigma(K, 3);
        const D3 fac_i = (DIM == 2 ? sig * h1i * h1i : sig * h1i * h1i * h1i) * h1i;
        const double Lx = s.L[0], Ly = s.L[1], Lz = s.L[2];
        const double hLx = s.halfL[0], hLy = s.halfL[1], hLz = s.halfL[2];
        const int periodic = s.periodic;

        AccT<DIM, D3> A;
        const int cnt = min(nbcnt[t], stride);
        for (int q = 0; q < cnt; ++q) {
            const int32_t jj = perm[nb[(int64_t)q * n + t]];
            const SideT<D3> o = load_side<DIM, ND>(F, G, aprev, n, jj, s, chf, d0);
            const D3 h1j = o.h1, mj = o.m, rhoj = o.rho, pj = o.p, ej = o.e;
            const D3 uj = o.u, vj = o.v, wj = o.w, omj = o.om;
            const D3 bxj = o.bx, byj = o.by, bzj = o.bz, alj = o.al;
            const D3 auj = o.au, avj = o.av, awj = o.aw;
            const D3 psichj = o.psich, vwj = o.vw, hirj = o.hir;
            const D3 irhoj = 2.0 * hirj;
            const D3 r2ij = irhoj * irhoj;
            const D3 pbj = pj * r2ij;
            const D3 brxj = bxj * irhoj, bryj = byj * irhoj, brzj = bzj * irhoj;
            const D3 bro2j = (brxj * brxj + bryj * bryj + brzj * brzj) * inv_mu0;

            const double ddx = xi.v - o.x.v, ddy = yi.v - o.y.v, ddz = zi.v - o.z.v;
            const D3 dx = (xi - o.x) + (min_image(ddx, Lx, hLx, periodic) - ddx);
            const D3 dy = (yi - o.y) + (min_image(ddy, Ly, hLy, periodic) - ddy);
            const D3 dz = DIM == 3 ? (zi - o.z) + (min_image(ddz, Lz, hLz, periodic) - ddz) : D3(0.0);
            const D3 r = sqrt(dx * dx + dy * dy + dz * dz);

            D3 gi(0.0), gj(0.0), ex(0.0), ey(0.0), ez(0.0);
            if (r > 1.0e-12) {
                const D3 ir = 1.0 / r;
                gi = shape_dw<K, DIM>(r * h1i) * fac_i * ir;
                const D3 facj = (DIM == 2 ? sig * h1j * h1j : sig * h1j * h1j * h1j) * h1j;
                gj = shape_dw<K, DIM>(r * h1j) * facj * ir;
                ex = dx * ir;
                ey = dy * ir;
                ez = dz * ir;
            }
            const D3 dwix = gi * dx, dwiy = gi * dy, dwiz = gi * dz;
            const D3 dwjx = gj * dx, dwjy = gj * dy, dwjz = gj * dz;
            const D3 dvx = ui - uj, dvy = vi - vj, dvz = wi - wj;
            const D3 gri = omi * (ex * dwix + ey * dwiy + ez * dwiz);
            const D3 grj = omj * (ex * dwjx + ey * dwjy + ez * dwjz);
            const D3 dot = dvx * ex + dvy * ey + dvz * ez;
            // MHDAccelerations (equations.py:217-275)
            {
                const bool near8 = r < 1.0e-8;
                const D3 dotA = near8 ? D3(0.0) : dot;
                const D3 griA = near8 ? D3(0.0) : gri;
                const D3 grjA = near8 ? D3(0.0) : grj;
                A.div += mj * dotA * griA;
                const D3 vsig2 = sqrt(fabs(pi_ - pj) / (0.5 * (rhoi + rhoj)));
                D3 c = fmax(vwi - 2.0 * dotA, vwj - 2.0 * dotA);
                c = fmax(c, 0.0);
                A.dtcfl = fmax(A.dtcfl, c);
                D3 sx = pbi * omi, sjx = pbj * omj;
                if (dotA < 0.0) {
                    const D3 vsi = fmax(ali * vwi - 2.0 * dotA, 0.0);
                    const D3 vsj = fmax(alj * vwj - 2.0 * dotA, 0.0);
                    const D3 qi = -hiri * vsi * dotA * av;
                    const D3 qj = -hirj * vsj * dotA * av;
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
            // MagneticStressAccelerations (equations.py:149-195)
            {
                const D3 gradp = mj * (0.5 * bro2i * gri + 0.5 * bro2j * grj);
                const D3 pi_b = (brxi * ex + bryi * ey + brzi * ez) * gri;
                const D3 pj_b = (brxj * ex + bryj * ey + brzj * ez) * grj;
                const D3 sm = mj * inv_mu0;
                A.ax += -ex * gradp + sm * (brxi * pi_b + brxj * pj_b);
                A.ay += -ey * gradp + sm * (bryi * pi_b + bryj * pj_b);
                A.az += -ez * gradp + sm * (brzi * pi_b + brzj * pj_b);
            }
            // MagneticStateRates (equations.py:507-559)
            {
                const D3 pbi_ = bxi * ex + byi * ey + bzi * ez;
                const D3 pbj_ = bxj * ex + byj * ey + bzj * ez;
                const D3 dbx = bxi - bxj, dby = byi - byj, dbz = bzi - bzj;
                const D3 pdb = dbx * ex + dby * ey + dbz * ez;
                const D3 db2 = dbx * dbx + dby * dby + dbz * dbz;
                const D3 pmi = mj * r2ii * gri;
                const D3 pmj = mj * r2ij * grj;
                const D3 brt = -(pmi * pbi_);
                const D3 tx = dvx - dot * ex, ty = dvy - dot * ey, tz = dvz - dot * ez;
                const D3 vsb = sqrt(tx * tx + ty * ty + tz * tz);
                const D3 diss = 0.5 * (pmi + pmj) * vsb * ar;
                const D3 dpsi = pmi * psichi + pmj * psichj;
                A.bx += brt * dvx + diss * dbx - dpsi * ex;
                A.by += brt * dvy + diss * dby - dpsi * ey;
                A.bz += brt * dvz + diss * dbz - dpsi * ez;
                A.ae += -0.5 * db2 * diss * inv_mu0;
                A.divbs += pmi * pbi_ + pmj * pbj_;
                A.divbd += -mj * pdb * gri;
            }
            // CullenDehnenAlphaDiagnostics loop (equations.py:319-354)
            {
                const D3 rx = mj * dwix, ry = mj * dwiy, rz = mj * dwiz;
                const D3 dax = aui - auj, day = avi - avj;
                A.vxx += dvx * rx; A.vxy += dvx * ry;
                A.vyx += dvy * rx; A.vyy += dvy * ry;
                A.vzx += dvz * rx; A.vzy += dvz * ry;
                A.axx += dax * rx; A.ayy += day * ry;
                if constexpr (DIM == 3) {
                    const D3 daz = awi - awj;
                    A.rxx += -dx * rx; A.rxy += -dx * ry; A.rxz += -dx * rz;
                    A.ryy += -dy * ry; A.ryz += -dy * rz; A.rzz += -dz * rz;
                    A.vxz += dvx * rz; A.vyz += dvy * rz; A.vzz += dvz * rz;
                    A.axy += dax * ry; A.axz += dax * rz;
                    A.ayx += day * rx; A.ayz += day * rz;
                    A.azx += daz * rx; A.azy += daz * ry; A.azz += daz * rz;
                }
            }
        }

        // ---- epilogue (as k_force) ----
        const D3 hi = ldN<ND>(F, G, TFLD(H), i, d0);
        const D3 cs = ldN<ND>(F, G, TFLD(CS), i, d0), psipred = ldN<ND>(F, G, TFLD(PSIPRED), i, d0);
        const D3 div = -A.div / rhoi;
        const D3 ah = hi * div * (1.0 / (double)DIM);
        D3 div_a, dvxdx, dvxdy, dvxdz, dvydx, dvydy, dvydz, dvzdx, dvzdy, dvzdz;
        const D3 termnorm = omi / rhoi;
        bool inverted = false;
        if constexpr (DIM == 3) {
            const D3 rxx = A.rxx, rxy = A.rxy, rxz = A.rxz, ryy = A.ryy, ryz = A.ryz, rzz = A.rzz;
            const D3 denom = rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
            if (::fabs(denom.v) > 2.2250738585072014e-308) {
                inverted = true;
                const D3 dd = 1.0 / denom;
                const D3 txx = ryy * rzz - ryz * ryz;
                const D3 txy = rxz * ryz - rzz * rxy;
                const D3 txz = rxy * ryz - rxz * ryy;
                const D3 tyy = rzz * rxx - rxz * rxz;
                const D3 tyz = rxy * rxz - rxx * ryz;
                const D3 tzz = rxx * ryy - rxy * rxy;
                const D3 gax = (A.axx * txx + A.axy * txy + A.axz * txz) * dd;
                const D3 gay = (A.ayx * txy + A.ayy * tyy + A.ayz * tyz) * dd;
                const D3 gaz = (A.azx * txz + A.azy * tyz + A.azz * tzz) * dd;
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
            div_a = -termnorm * (A.axx + A.ayy + A.azz);
            dvxdx = -termnorm * A.vxx; dvxdy = -termnorm * A.vxy; dvxdz = -termnorm * A.vxz;
            dvydx = -termnorm * A.vyx; dvydy = -termnorm * A.vyy; dvydz = -termnorm * A.vyz;
            dvzdx = -termnorm * A.vzx; dvzdy = -termnorm * A.vzy; dvzdz = -termnorm * A.vzz;
        }
        const D3 divvdt = div_a - (dvxdx * dvxdx + dvydy * dvydy + dvzdz * dvzdz +
                                   2.0 * (dvxdy * dvydx + dvxdz * dvzdx + dvydz * dvzdy));
        const D3 divv = dvxdx + dvydy + dvzdz;
        const D3 cvx = dvzdy - dvydz, cvy = dvxdz - dvzdx, cvz = dvydx - dvxdy;
        const D3 comp = fmax(-divv, 0.0);
        const D3 fc = comp * comp;
        const D3 den = fc + (cvx * cvx + cvy * cvy + cvz * cvz);
        const D3 xil = den > 0.0 ? fc / den : D3(1.0);
        const D3 source = 10.0 * hi * hi * xil * fmax(-divvdt, 0.0);
        const D3 alphaloc = fmax(s.alpha_min, alpha_max * fmax(fmin(source / (cs * cs), 1.0), 0.0));

        const D3 ch = chf * vwi;
        const D3 apsi = -ch * A.divbd / rhoi - sigma * psipred * ch / hi - 0.5 * s.cleaning_active * psipred * div;
        D3 dtcfl = fmax(A.dtcfl, vwi);
        dtcfl = fmax(dtcfl, ch);

        D3 ae = A.ae;
        if (dtcfl > 0.0) {
            const D3 dtc = s.c_cour * hi / dtcfl;
            if (ei + dtc * ae < 0.0 && ei > 0.0) ae /= 1.0 - dtc * ae / ei;
        }
        const D3 b2 = bxi * bxi + byi * byi + bzi * bzi;
        const D3 divb = A.divbs;
        D3 frac(0.0);
        if (b2 > 0.0) {
            const D3 beta = 2.0 * s.mu0 * pi_ / b2;
            frac = fmax(0.0, fmin(1.0, (10.0 - beta) * 0.125));
        }
        const D3 corr = divb * frac * inv_mu0;
        const D3 ax = A.ax - bxi * corr, ay = A.ay - byi * corr, az = A.az - bzi * corr;
        dt_adapt = D3(1.0e29);
        if (dtcfl > 0.0) dt_adapt = s.c_cour * hi / dtcfl;
        const D3 f2 = ax * ax + ay * ay + az * az;
        if (f2 > 0.0) dt_adapt = fmin(dt_adapt, s.c_force * sqrt(hi / sqrt(f2)));
        bad = !isfinite(f2) || !isfinite(ae);

        stN<ND>(F, G, TFLD(AU), i, ax, d0, lead);
        stN<ND>(F, G, TFLD(AV), i, ay, d0, lead);
        stN<ND>(F, G, TFLD(AW), i, az, d0, lead);
        stN<ND>(F, G, TFLD(AE), i, ae, d0, lead);
        stN<ND>(F, G, TFLD(AH), i, ah, d0, lead);
        stN<ND>(F, G, TFLD(DIV), i, div, d0, lead);
        stN<ND>(F, G, TFLD(DIV_FOR_PSI), i, div, d0, lead);
        stN<ND>(F, G, TFLD(DT_CFL), i, dtcfl, d0, lead);
        stN<ND>(F, G, TFLD(DT_ADAPT), i, dt_adapt, d0, lead);
        stN<ND>(F, G, TFLD(ABEVOLX), i, A.bx, d0, lead);
        stN<ND>(F, G, TFLD(ABEVOLY), i, A.by, d0, lead);
        stN<ND>(F, G, TFLD(ABEVOLZ), i, A.bz, d0, lead);
        stN<ND>(F, G, TFLD(APSI), i, apsi, d0, lead);
        stN<ND>(F, G, TFLD(DIVBSYMM), i, rhoi * divb, d0, lead);
        stN<ND>(F, G, TFLD(DIVBDIFF), i, A.divbd, d0, lead);
        stN<ND>(F, G, TFLD(ALPHALOC), i, alphaloc, d0, lead);
        stN<ND>(F, G, TFLD(XILIMITER), i, xil, d0, lead);
        stN<ND>(F, G, TFLD(DIVVDT), i, divvdt, d0, lead);
    }
    // the primal's min dt_adapt and NaN flag per instance
    if (!lead) return;  // the whole workgroup: uniform
    constexpr int NW = kBlock / kWave;  // LDS slots: waves of the largest shape
    const int nw = blockDim.x / kWave;
    __shared__ unsigned long long wdb[NW];
    __shared__ int wk[NW], wbad[NW];
    const int w = threadIdx.x / kWave;
    unsigned long long db = dbits(dt_adapt.v);
    const int k0 = __shfl(ki, 0);
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

// the tangent of the pending dt: the lowest storage index attaining the
// minimum, then its dt_adapt tangent
__global__ void k_t_dt_argmin(Fields F, const Inst* __restrict__ inst, InstTan* itan,
                              const int32_t* __restrict__ inst_of, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = inst_of[i];
    if (F.f[SPHMHD_F_DT_ADAPT][i] == bitsd(inst[k].dtp_bits)) atomicMin(&itan[k].argmin, (int)i);
}

__global__ void k_t_dt_tangent(TFields G, InstTan* itan, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    const int a = itan[k].argmin;
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) itan[k].dtp[q] = a != INT_MAX ? G.f[q][SPHMHD_F_DT_ADAPT][a] : 0.0;
}

// ---------------------------------------------------------------------------
// MHDStep (integrator.py:12-82) over duals, the clock and the warm start
// ---------------------------------------------------------------------------

__global__ void k_t_initialize(Fields F, TFields G, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto cp = [&](int dst, int src) { stT(F, G, dst, i, ldT(F, G, src, i)); };
    cp(TFLD(X0), TFLD(X));
    cp(TFLD(Y0), TFLD(Y));
    if (F.f[SPHMHD_F_Z0]) stT(F, G, TFLD(Z0), i, F.f[SPHMHD_F_Z] ? ldT(F, G, TFLD(Z), i) : D3(0.0));
    cp(TFLD(U0), TFLD(U));
    cp(TFLD(V0), TFLD(V));
    cp(TFLD(W0), TFLD(W));
    cp(TFLD(UPRED), TFLD(U));
    cp(TFLD(VPRED), TFLD(V));
    cp(TFLD(WPRED), TFLD(W));
    cp(TFLD(E0), TFLD(E));
    cp(TFLD(EPRED), TFLD(E));
    cp(TFLD(H0), TFLD(H));
    stT(F, G, TFLD(CONVERGED), i, D3(0.0));
    stT(F, G, TFLD(OMEGA), i, D3(1.0));
    cp(TFLD(ALPHA10), TFLD(ALPHA1));
    cp(TFLD(BEVOLX0), TFLD(BEVOLX));
    cp(TFLD(BEVOLY0), TFLD(BEVOLY));
    cp(TFLD(BEVOLZ0), TFLD(BEVOLZ));
    cp(TFLD(PSI0), TFLD(PSI));
    cp(TFLD(BEVOLXPRED), TFLD(BEVOLX));
    cp(TFLD(BEVOLYPRED), TFLD(BEVOLY));
    cp(TFLD(BEVOLZPRED), TFLD(BEVOLZ));
    cp(TFLD(PSIPRED), TFLD(PSI));
}

__device__ __forceinline__ D3 dtT(const Inst& s, const InstTan& it) {
    D3 r(s.dt);
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) r.d[q] = it.dt[q];
    return r;
}

__device__ __forceinline__ void t_wrap(const Fields& F, const Inst& s, int64_t i) {
    if (!s.periodic) return;
    // a constant shift: the tangents are unchanged
    double* x = F.f[SPHMHD_F_X];
    double* y = F.f[SPHMHD_F_Y];
    x[i] = x[i] < s.lo[0] ? x[i] + s.L[0] : (x[i] >= s.hi[0] ? x[i] - s.L[0] : x[i]);
    y[i] = y[i] < s.lo[1] ? y[i] + s.L[1] : (y[i] >= s.hi[1] ? y[i] - s.L[1] : y[i]);
    if (s.dim == 3) {
        double* z = F.f[SPHMHD_F_Z];
        z[i] = z[i] < s.lo[2] ? z[i] + s.L[2] : (z[i] >= s.hi[2] ? z[i] - s.L[2] : z[i]);
    }
}

__global__ void k_t_stage1(Fields F, TFields G, const Inst* __restrict__ inst, const InstTan* __restrict__ itan,
                           const int32_t* __restrict__ inst_of, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = inst_of[i];
    const Inst& s = inst[k];
    if (F.itype[i] == 1) {
        const D3 dt = dtT(s, itan[k]);
        const D3 dtb2 = 0.5 * dt;
        const D3 au = ldT(F, G, TFLD(AU), i), av = ldT(F, G, TFLD(AV), i), aw = ldT(F, G, TFLD(AW), i);
        const D3 u0 = ldT(F, G, TFLD(U0), i), v0 = ldT(F, G, TFLD(V0), i), w0 = ldT(F, G, TFLD(W0), i);
        const D3 u = u0 + dtb2 * au, v = v0 + dtb2 * av, w = w0 + dtb2 * aw;
        stT(F, G, TFLD(U), i, u);
        stT(F, G, TFLD(V), i, v);
        stT(F, G, TFLD(W), i, w);
        stT(F, G, TFLD(UPRED), i, u0 + dt * au);
        stT(F, G, TFLD(VPRED), i, v0 + dt * av);
        stT(F, G, TFLD(WPRED), i, w0 + dt * aw);
        stT(F, G, TFLD(X), i, ldT(F, G, TFLD(X0), i) + dt * u);
        stT(F, G, TFLD(Y), i, ldT(F, G, TFLD(Y0), i) + dt * v);
        if (s.dim == 3) stT(F, G, TFLD(Z), i, ldT(F, G, TFLD(Z0), i) + dt * w);
        const D3 e0 = ldT(F, G, TFLD(E0), i), ae = ldT(F, G, TFLD(AE), i);
        stT(F, G, TFLD(E), i, e0 + dtb2 * ae);
        stT(F, G, TFLD(EPRED), i, e0 + dt * ae);
        const D3 h = ldT(F, G, TFLD(H0), i) + dt * ldT(F, G, TFLD(AH), i);
        stT(F, G, TFLD(H), i, h);
        const D3 decay = 0.1 * ldT(F, G, TFLD(CS), i) / h;
        const D3 a0 = ldT(F, G, TFLD(ALPHA10), i), al = ldT(F, G, TFLD(ALPHALOC), i);
        stT(F, G, TFLD(ALPHA1), i, a0 < al ? al : (a0 + dt * al * decay) / (1.0 + dt * decay));
        const int b0[4] = {TFLD(BEVOLX0), TFLD(BEVOLY0), TFLD(BEVOLZ0), TFLD(PSI0)};
        const int ba[4] = {TFLD(ABEVOLX), TFLD(ABEVOLY), TFLD(ABEVOLZ), TFLD(APSI)};
        const int bn[4] = {TFLD(BEVOLX), TFLD(BEVOLY), TFLD(BEVOLZ), TFLD(PSI)};
        const int bp[4] = {TFLD(BEVOLXPRED), TFLD(BEVOLYPRED), TFLD(BEVOLZPRED), TFLD(PSIPRED)};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const D3 x0 = ldT(F, G, b0[c], i), a = ldT(F, G, ba[c], i);
            stT(F, G, bn[c], i, x0 + dtb2 * a);
            stT(F, G, bp[c], i, x0 + dt * a);
        }
    }
    t_wrap(F, s, i);
}

__global__ void k_t_stage2(Fields F, TFields G, const Inst* __restrict__ inst, const InstTan* __restrict__ itan,
                           const int32_t* __restrict__ inst_of, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = inst_of[i];
    const Inst& s = inst[k];
    if (F.itype[i] == 1) {
        const D3 dtb2 = 0.5 * dtT(s, itan[k]);
        stT(F, G, TFLD(U), i, ldT(F, G, TFLD(U), i) + dtb2 * ldT(F, G, TFLD(AU), i));
        stT(F, G, TFLD(V), i, ldT(F, G, TFLD(V), i) + dtb2 * ldT(F, G, TFLD(AV), i));
        stT(F, G, TFLD(W), i, ldT(F, G, TFLD(W), i) + dtb2 * ldT(F, G, TFLD(AW), i));
        const D3 e = ldT(F, G, TFLD(E), i) + dtb2 * ldT(F, G, TFLD(AE), i);
        stT(F, G, TFLD(E), i, e);
        const D3 rho = ldT(F, G, TFLD(RHO), i);
        stT(F, G, TFLD(P), i, s.gamma1 * rho * e);
        const D3 bx = ldT(F, G, TFLD(BEVOLX), i) + dtb2 * ldT(F, G, TFLD(ABEVOLX), i);
        const D3 by = ldT(F, G, TFLD(BEVOLY), i) + dtb2 * ldT(F, G, TFLD(ABEVOLY), i);
        const D3 bz = ldT(F, G, TFLD(BEVOLZ), i) + dtb2 * ldT(F, G, TFLD(ABEVOLZ), i);
        stT(F, G, TFLD(BEVOLX), i, bx);
        stT(F, G, TFLD(BEVOLY), i, by);
        stT(F, G, TFLD(BEVOLZ), i, bz);
        stT(F, G, TFLD(PSI), i, ldT(F, G, TFLD(PSI), i) + dtb2 * ldT(F, G, TFLD(APSI), i));
        stT(F, G, TFLD(BX), i, bx * rho);
        stT(F, G, TFLD(BY), i, by * rho);
        stT(F, G, TFLD(BZ), i, bz * rho);
    }
    t_wrap(F, s, i);
}

// t += dt; dt = the pending minimum (adaptive) clipped to tf - t.  The time
// itself carries no tangent (the step count, not the end time, is fixed), so
// a step clipped to tf - t has none either.
__global__ void k_t_clock(Inst* inst, InstTan* itan, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    Inst& s = inst[k];
    InstTan& it = itan[k];
    const double t = s.t + s.dt;
    const double pending = bitsd(s.dtp_bits);
    const double dt = next_dt(s, t, s.dt, pending);
    const bool clipped = t + (s.adaptive ? pending : s.dt) > s.tf;
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) it.dt[q] = clipped ? 0.0 : (s.adaptive ? it.dtp[q] : it.dt[q]);
    s.dt = dt;
    s.t = t;
}

// application.py:134-147: dt = the pending minimum, clipped to tf - t
__global__ void k_t_warm_dt_commit(Inst* inst, InstTan* itan, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    Inst& s = inst[k];
    const double pending = bitsd(s.dtp_bits);
    const bool clipped = s.t + pending > s.tf;
    s.dt = clipped ? s.tf - s.t : pending;
#pragma unroll
    for (int q = 0; q < kTanDirs; ++q) itan[k].dt[q] = clipped ? 0.0 : itan[k].dtp[q];
}

__global__ void k_t_sync_initial(Fields F, TFields G, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const D3 rho = ldT(F, G, TFLD(RHO), i);
    const D3 bx = ldT(F, G, TFLD(BX), i) / rho, by = ldT(F, G, TFLD(BY), i) / rho, bz = ldT(F, G, TFLD(BZ), i) / rho;
    stT(F, G, TFLD(BEVOLX), i, bx);
    stT(F, G, TFLD(BEVOLY), i, by);
    stT(F, G, TFLD(BEVOLZ), i, bz);
    stT(F, G, TFLD(PSI), i, D3(0.0));
    stT(F, G, TFLD(BEVOLXPRED), i, bx);
    stT(F, G, TFLD(BEVOLYPRED), i, by);
    stT(F, G, TFLD(BEVOLZPRED), i, bz);
    stT(F, G, TFLD(PSIPRED), i, D3(0.0));
    stT(F, G, TFLD(UPRED), i, ldT(F, G, TFLD(U), i));
    stT(F, G, TFLD(VPRED), i, ldT(F, G, TFLD(V), i));
    stT(F, G, TFLD(WPRED), i, ldT(F, G, TFLD(W), i));
    stT(F, G, TFLD(EPRED), i, ldT(F, G, TFLD(E), i));
}

// copies field src (value and tangents) into a scratch row block, or back
__global__ void k_t_save(Fields F, TFields G, int field, double* __restrict__ dst, int64_t n, int restore) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (restore) {
        F.f[field][i] = dst[i];
#pragma unroll
        for (int q = 0; q < kTanDirs; ++q) G.f[q][field][i] = dst[(int64_t)(q + 1) * n + i];
    } else {
        dst[i] = F.f[field][i];
#pragma unroll
        for (int q = 0; q < kTanDirs; ++q) dst[(int64_t)(q + 1) * n + i] = G.f[q][field][i];
    }
}

__global__ void k_t_copy(Fields F, TFields G, int dst, int src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) stT(F, G, dst, i, ldT(F, G, src, i));
}

// conserved totals (k_totals) with their tangents: out[k][q][0] the value,
// out[k][q][1 + d] the tangent along direction d
__global__ void k_t_totals(Fields F, TFields G, const Inst* __restrict__ inst, double* __restrict__ out) {
    const Inst& s = inst[blockIdx.x];
    D3 acc[SPHMHD_NTOTALS];
    for (int64_t i = s.off0 + threadIdx.x; i < s.off1; i += blockDim.x) {
        const double m = F.f[SPHMHD_F_M][i];
        const D3 rho = ldT(F, G, TFLD(RHO), i);
        const D3 u = ldT(F, G, TFLD(U), i), v = ldT(F, G, TFLD(V), i), w = ldT(F, G, TFLD(W), i);
        const D3 bx = ldT(F, G, TFLD(BX), i), by = ldT(F, G, TFLD(BY), i), bz = ldT(F, G, TFLD(BZ), i);
        const D3 v2 = u * u + v * v + w * w;
        const D3 b2 = bx * bx + by * by + bz * bz;
        acc[SPHMHD_T_MASS] += D3(m);
        acc[SPHMHD_T_MOMX] += m * u;
        acc[SPHMHD_T_MOMY] += m * v;
        acc[SPHMHD_T_MOMZ] += m * w;
        acc[SPHMHD_T_EKIN] += 0.5 * m * v2;
        acc[SPHMHD_T_ETHERM] += m * ldT(F, G, TFLD(E), i);
        acc[SPHMHD_T_EMAG] += 0.5 * m * b2 / (s.mu0 * rho);
        if (b2 > 0.0)
            acc[SPHMHD_T_DIVB] += (m / rho) * ldT(F, G, TFLD(H), i) * fabs(ldT(F, G, TFLD(DIVBSYMM), i)) / sqrt(b2);
        acc[SPHMHD_T_VOLUME] += m / rho;
    }
    __shared__ double red[kBlock];
    for (int q = 0; q < SPHMHD_NTOTALS; ++q) {
        for (int c = 0; c <= kTanDirs; ++c) {
            red[threadIdx.x] = c == 0 ? acc[q].v : acc[q].d[c - 1];
            __syncthreads();
            for (int off = kBlock / 2; off > 0; off >>= 1) {
                if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
                __syncthreads();
            }
            if (threadIdx.x == 0) out[((int64_t)blockIdx.x * SPHMHD_NTOTALS + q) * (1 + kTanDirs) + c] = red[0];
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned tblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

#define T_DISPATCH(KID, DIMV, CALL)                                                       \
    switch (KID) {                                                                        \
    case SPHMHD_CUBIC: if (DIMV == 2) CALL(SPHMHD_CUBIC, 2) else CALL(SPHMHD_CUBIC, 3) break; \
    case SPHMHD_QUINTIC: if (DIMV == 2) CALL(SPHMHD_QUINTIC, 2) else CALL(SPHMHD_QUINTIC, 3) break; \
    case SPHMHD_WENDLAND_C2: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C2, 2) else CALL(SPHMHD_WENDLAND_C2, 3) break; \
    case SPHMHD_WENDLAND_C4: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C4, 2) else CALL(SPHMHD_WENDLAND_C4, 3) break; \
    case SPHMHD_WENDLAND_C6: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C6, 2) else CALL(SPHMHD_WENDLAND_C6, 3) break; \
    case SPHMHD_GAUSSIAN: if (DIMV == 2) CALL(SPHMHD_GAUSSIAN, 2) else CALL(SPHMHD_GAUSSIAN, 3) break; \
    default: if (DIMV == 2) CALL(SPHMHD_SUPER_GAUSSIAN, 2) else CALL(SPHMHD_SUPER_GAUSSIAN, 3) break; \
    }

void launch_t_eval_begin(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_eval_begin, dim3(tblk(d.ninst)), dim3(kBlock), 0, d.stream, g.itan, d.ninst);
}

void launch_t_density(Dev& d, TanDev& g) {
    if (d.n == 0) return;
    const int p = d.pass;
    const int32_t* ain = p == 0 ? nullptr : d.act[(p - 1) & 1];
    const int32_t* cin = p == 0 ? nullptr : act_count(d.ctl, (p - 1) & 1);
    int32_t* aout = d.act[p & 1];
    int32_t* cout = act_count(d.ctl, p & 1);
    const int B = d.blk[LC_DENSITY];
    const int64_t per_block = (int64_t)(B / kWave) * kTPW;
    const unsigned grid = (unsigned)((d.n + per_block - 1) / per_block);
#define CALL_TD(KK, DD)                                                                                       \
    { hipLaunchKernelGGL((k_t_density<KK, DD>), dim3(grid), dim3(B), 0, d.stream, d.fields, g.tf,              \
                         (const double4*)d.dpack, d.perm, d.keys, d.cstart, d.cend, d.inst, d.inst_of, d.n, ain, \
                         cin, aout, cout); }
    T_DISPATCH(d.kernel, d.dim, CALL_TD)
#undef CALL_TD
}

void launch_t_prep(Dev& d, TanDev& g, bool keep_prev) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(k_t_prep, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, d.perm, d.inst,
                       d.inst_of, (double4*)d.dpack, g.aprev, d.n, keep_prev ? 1 : 0);
}

void launch_t_force(Dev& d, TanDev& g) {
    if (d.n == 0) return;
    const int B = d.blk[LC_FORCE];
#define CALL_TF(KK, DD)                                                                                      \
    { hipLaunchKernelGGL((k_t_force<KK, DD, 1>), dim3((d.n + B - 1) / B, kTanDirs), dim3(B), 0, d.stream, d.fields, \
                         g.tf, g.aprev, d.perm, d.nb, d.nbcnt, d.inst, g.itan, d.inst_of, d.n, d.nb_stride); }
    T_DISPATCH(d.kernel, d.dim, CALL_TF)
#undef CALL_TF
    hipLaunchKernelGGL(k_t_dt_argmin, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst, g.itan,
                       d.inst_of, d.n);
    hipLaunchKernelGGL(k_t_dt_tangent, dim3(tblk(d.ninst)), dim3(kBlock), 0, d.stream, g.tf, g.itan, d.ninst);
}

void launch_t_initialize(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_initialize, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, d.n);
}

void launch_t_stage1(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_stage1, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, d.inst, g.itan,
                       d.inst_of, d.n);
}

void launch_t_stage2(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_stage2, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, d.inst, g.itan,
                       d.inst_of, d.n);
}

void launch_t_clock(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_clock, dim3(tblk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, g.itan, d.ninst);
}

void launch_t_warm_dt(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_warm_dt_commit, dim3(tblk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, g.itan,
                       d.ninst);
}

void launch_t_sync_initial(Dev& d, TanDev& g) {
    hipLaunchKernelGGL(k_t_sync_initial, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, d.n);
}

void launch_t_save(Dev& d, TanDev& g, int field, double* buf, bool restore) {
    hipLaunchKernelGGL(k_t_save, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, field, buf, d.n,
                       restore ? 1 : 0);
}

void launch_t_copy(Dev& d, TanDev& g, int dst, int src) {
    hipLaunchKernelGGL(k_t_copy, dim3(tblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, g.tf, dst, src, d.n);
}

void launch_t_totals(Dev& d, TanDev& g, double* out) {
    hipLaunchKernelGGL(k_t_totals, dim3(d.ninst), dim3(kBlock), 0, d.stream, d.fields, g.tf, d.inst, out);
}

}  // namespace sph
