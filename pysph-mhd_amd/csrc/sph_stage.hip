// sph_stage.hip -- per-particle (HBM-bound) kernels around the pair loops.
//
//   MHDStep.initialize / stage1 / stage2 (integrator.py:12-82) with the
//   periodic wrap of PySPH's update_domain, the fused stage2(n) + initialize
//   + stage1(n+1) kernel used inside multi-step runs, the solver clock
//   (solver.py:57-61), the warm-start state sync (application.py:100-108) and
//   initial dt (application.py:134-147), and the conserved totals.
#include "sph_internal.h"

namespace sph {

#define FLD(name) F.f[SPHMHD_F_##name][i]

__device__ __forceinline__ void do_initialize(const Fields& F, int64_t i) {
    FLD(X0) = FLD(X);
    FLD(Y0) = FLD(Y);
    if (F.f[SPHMHD_F_Z0]) FLD(Z0) = F.f[SPHMHD_F_Z] ? FLD(Z) : 0.0;
    FLD(U0) = FLD(U);
    FLD(V0) = FLD(V);
    FLD(W0) = FLD(W);
    FLD(UPRED) = FLD(U);
    FLD(VPRED) = FLD(V);
    FLD(WPRED) = FLD(W);
    FLD(E0) = FLD(E);
    FLD(EPRED) = FLD(E);
    FLD(H0) = FLD(H);
    FLD(CONVERGED) = 0.0;
    FLD(OMEGA) = 1.0;
    FLD(ALPHA10) = FLD(ALPHA1);
    FLD(BEVOLX0) = FLD(BEVOLX);
    FLD(BEVOLY0) = FLD(BEVOLY);
    FLD(BEVOLZ0) = FLD(BEVOLZ);
    FLD(PSI0) = FLD(PSI);
    FLD(BEVOLXPRED) = FLD(BEVOLX);
    FLD(BEVOLYPRED) = FLD(BEVOLY);
    FLD(BEVOLZPRED) = FLD(BEVOLZ);
    FLD(PSIPRED) = FLD(PSI);
}

__device__ __forceinline__ double wrap1(double x, double lo, double hi, double L) {
    if (x < lo) x += L;
    else if (x >= hi) x -= L;
    return x;
}

__device__ __forceinline__ void do_wrap(const Fields& F, const Inst& s, int64_t i) {
    if (!s.periodic) return;
    FLD(X) = wrap1(FLD(X), s.lo[0], s.hi[0], s.L[0]);
    FLD(Y) = wrap1(FLD(Y), s.lo[1], s.hi[1], s.L[1]);
    if (s.dim == 3) FLD(Z) = wrap1(FLD(Z), s.lo[2], s.hi[2], s.L[2]);
}

__device__ __forceinline__ void do_stage1(const Fields& F, const Inst& s, int64_t i, double dt) {
    if (F.itype[i] != 1) return;
    const double dtb2 = 0.5 * dt;
    const double u = FLD(U0) + dtb2 * FLD(AU);
    const double v = FLD(V0) + dtb2 * FLD(AV);
    const double w = FLD(W0) + dtb2 * FLD(AW);
    FLD(U) = u;
    FLD(V) = v;
    FLD(W) = w;
    FLD(UPRED) = FLD(U0) + dt * FLD(AU);
    FLD(VPRED) = FLD(V0) + dt * FLD(AV);
    FLD(WPRED) = FLD(W0) + dt * FLD(AW);
    FLD(X) = FLD(X0) + dt * u;
    FLD(Y) = FLD(Y0) + dt * v;
    if (s.dim == 3) FLD(Z) = FLD(Z0) + dt * w;
    FLD(E) = FLD(E0) + dtb2 * FLD(AE);
    FLD(EPRED) = FLD(E0) + dt * FLD(AE);
    const double h = FLD(H0) + dt * FLD(AH);
    FLD(H) = h;
    const double decay = 0.1 * FLD(CS) / h;
    const double a0 = FLD(ALPHA10), al = FLD(ALPHALOC);
    FLD(ALPHA1) = a0 < al ? al : (a0 + dt * al * decay) / (1.0 + dt * decay);
    FLD(BEVOLX) = FLD(BEVOLX0) + dtb2 * FLD(ABEVOLX);
    FLD(BEVOLY) = FLD(BEVOLY0) + dtb2 * FLD(ABEVOLY);
    FLD(BEVOLZ) = FLD(BEVOLZ0) + dtb2 * FLD(ABEVOLZ);
    FLD(PSI) = FLD(PSI0) + dtb2 * FLD(APSI);
    FLD(BEVOLXPRED) = FLD(BEVOLX0) + dt * FLD(ABEVOLX);
    FLD(BEVOLYPRED) = FLD(BEVOLY0) + dt * FLD(ABEVOLY);
    FLD(BEVOLZPRED) = FLD(BEVOLZ0) + dt * FLD(ABEVOLZ);
    FLD(PSIPRED) = FLD(PSI0) + dt * FLD(APSI);
}

__device__ __forceinline__ void do_stage2(const Fields& F, const Inst& s, int64_t i, double dt) {
    if (F.itype[i] != 1) return;
    const double dtb2 = 0.5 * dt;
    FLD(U) += dtb2 * FLD(AU);
    FLD(V) += dtb2 * FLD(AV);
    FLD(W) += dtb2 * FLD(AW);
    const double e = FLD(E) + dtb2 * FLD(AE);
    FLD(E) = e;
    const double rho = FLD(RHO);
    FLD(P) = s.gamma1 * rho * e;
    const double bx = FLD(BEVOLX) + dtb2 * FLD(ABEVOLX);
    const double by = FLD(BEVOLY) + dtb2 * FLD(ABEVOLY);
    const double bz = FLD(BEVOLZ) + dtb2 * FLD(ABEVOLZ);
    FLD(BEVOLX) = bx;
    FLD(BEVOLY) = by;
    FLD(BEVOLZ) = bz;
    FLD(PSI) += dtb2 * FLD(APSI);
    FLD(BX) = bx * rho;
    FLD(BY) = by * rho;
    FLD(BZ) = bz * rho;
}

__global__ void k_initialize(Fields F, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) do_initialize(F, i);
}

__global__ void k_stage1(Fields F, const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Inst& s = inst[inst_of[i]];
    do_stage1(F, s, i, s.dt);
    do_wrap(F, s, i);
}

__global__ void k_stage2(Fields F, const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Inst& s = inst[inst_of[i]];
    do_stage2(F, s, i, s.dt);
    do_wrap(F, s, i);
}

// stage2 of step n, then initialize and stage1 of step n+1 with the step the
// clock kernel is about to commit.  The same arithmetic as do_stage2,
// do_initialize, do_stage1 and do_wrap in that order, expression for
// expression (so it rounds identically), but every field is loaded once into
// registers and every output stored once: through the property pointers the
// compiler cannot tell the fields apart and would reload and re-store them.
__device__ __forceinline__ double4 stage2_init_stage1_one(const Fields& F, const Inst* __restrict__ inst,
                                                          const int32_t* __restrict__ inst_of, int64_t i) {
    const Inst& s = inst[inst_of[i]];
    const double dt = s.dt;
    const double dtn = next_dt(s, s.t + s.dt, s.dt, bitsd(s.dtp_bits));
    const bool fluid = F.itype[i] == 1;
    const bool has_z = F.f[SPHMHD_F_Z] != nullptr, has_z0 = F.f[SPHMHD_F_Z0] != nullptr;

    double u = FLD(U), v = FLD(V), w = FLD(W), e = FLD(E);
    double bx = FLD(BEVOLX), by = FLD(BEVOLY), bz = FLD(BEVOLZ), psi = FLD(PSI);
    const double au = FLD(AU), av = FLD(AV), aw = FLD(AW), ae = FLD(AE);
    const double abx = FLD(ABEVOLX), aby = FLD(ABEVOLY), abz = FLD(ABEVOLZ), apsi = FLD(APSI);
    double x = FLD(X), y = FLD(Y), z = has_z ? FLD(Z) : 0.0;
    double h = FLD(H), alpha = FLD(ALPHA1);
    // every input is loaded before the first store: a load after a store to
    // another field could alias it for the compiler, which then waits for a
    // second round trip to memory
    const double rho = FLD(RHO), ah = FLD(AH), cs = FLD(CS), al = FLD(ALPHALOC);

    // stage2 (step n)
    if (fluid) {
        const double dtb2 = 0.5 * dt;
        u += dtb2 * au;
        v += dtb2 * av;
        w += dtb2 * aw;
        e = e + dtb2 * ae;
        FLD(P) = s.gamma1 * rho * e;
        bx = bx + dtb2 * abx;
        by = by + dtb2 * aby;
        bz = bz + dtb2 * abz;
        psi += dtb2 * apsi;
        FLD(BX) = bx * rho;
        FLD(BY) = by * rho;
        FLD(BZ) = bz * rho;
    }
    // initialize (step n+1)
    const double x0 = x, y0 = y, u0 = u, v0 = v, w0 = w, e0 = e, h0 = h, a10 = alpha;
    const double bx0 = bx, by0 = by, bz0 = bz, psi0 = psi;
    FLD(X0) = x0;
    FLD(Y0) = y0;
    if (has_z0) FLD(Z0) = z;
    FLD(U0) = u0;
    FLD(V0) = v0;
    FLD(W0) = w0;
    FLD(E0) = e0;
    FLD(H0) = h0;
    FLD(CONVERGED) = 0.0;
    FLD(OMEGA) = 1.0;
    FLD(ALPHA10) = a10;
    FLD(BEVOLX0) = bx0;
    FLD(BEVOLY0) = by0;
    FLD(BEVOLZ0) = bz0;
    FLD(PSI0) = psi0;
    double upred = u, vpred = v, wpred = w, epred = e;
    double bxp = bx, byp = by, bzp = bz, psip = psi;
    // stage1 (step n+1)
    if (fluid) {
        const double dtb2 = 0.5 * dtn;
        u = u0 + dtb2 * au;
        v = v0 + dtb2 * av;
        w = w0 + dtb2 * aw;
        upred = u0 + dtn * au;
        vpred = v0 + dtn * av;
        wpred = w0 + dtn * aw;
        x = x0 + dtn * u;
        y = y0 + dtn * v;
        if (s.dim == 3) z = z + dtn * w;  // z0 == z here
        e = e0 + dtb2 * ae;
        epred = e0 + dtn * ae;
        h = h0 + dtn * ah;
        const double decay = 0.1 * cs / h;
        alpha = a10 < al ? al : (a10 + dtn * al * decay) / (1.0 + dtn * decay);
        bx = bx0 + dtb2 * abx;
        by = by0 + dtb2 * aby;
        bz = bz0 + dtb2 * abz;
        psi = psi0 + dtb2 * apsi;
        bxp = bx0 + dtn * abx;
        byp = by0 + dtn * aby;
        bzp = bz0 + dtn * abz;
        psip = psi0 + dtn * apsi;
    }
    // wrap
    if (s.periodic) {
        x = wrap1(x, s.lo[0], s.hi[0], s.L[0]);
        y = wrap1(y, s.lo[1], s.hi[1], s.L[1]);
        if (s.dim == 3) z = wrap1(z, s.lo[2], s.hi[2], s.L[2]);
    }
    FLD(U) = u;
    FLD(V) = v;
    FLD(W) = w;
    FLD(UPRED) = upred;
    FLD(VPRED) = vpred;
    FLD(WPRED) = wpred;
    FLD(X) = x;
    FLD(Y) = y;
    if (s.dim == 3) FLD(Z) = z;
    FLD(E) = e;
    FLD(EPRED) = epred;
    FLD(H) = h;
    FLD(ALPHA1) = alpha;
    FLD(BEVOLX) = bx;
    FLD(BEVOLY) = by;
    FLD(BEVOLZ) = bz;
    FLD(PSI) = psi;
    FLD(BEVOLXPRED) = bxp;
    FLD(BEVOLYPRED) = byp;
    FLD(BEVOLZPRED) = bzp;
    FLD(PSIPRED) = psip;
    return make_double4(x, y, z, h);  // the evaluation's positions and h
}

// t += dt; dt = the pending step clipped to tf (solver.py:57-61)
__device__ __forceinline__ void advance_clock(Inst& s) {
    const double t = s.t + s.dt;
    s.dt = next_dt(s, t, s.dt, bitsd(s.dtp_bits));
    s.t = t;
}

// The end of both stage launches: each block leaves its kRed maxima of
// red_values in its own slot (hblk, kblk; a block straddling instances uses
// atomics), and the last workgroup combines the slots, commits the clock
// (advance: after a step's stage2) or clears the abort flag (the first step
// of a call, as k_eval_begin), starts the evaluation (eval_begin_body with
// the pending dt cleared) and sets up its cell grid.
template <int NR>
__device__ __forceinline__ void stage_reduce_begin(int k, unsigned long long (&v)[kRed], Inst* inst, int ninst,
                                                   Ctl* ctl, int32_t* arrive, unsigned long long* hblk,
                                                   int32_t* kblk, bool advance) {
    if (uint32_t* hist = ctl->hist) {  // h levels for the 3-D cell size
        __shared__ uint32_t lh[kHistSlots];
        hist_clear(lh);
        __syncthreads();
        if (k >= 0) hist_add(lh, bitsd(v[0]));
        __syncthreads();
        hist_flush(hist, lh);
    }
    bool slot = false;
    block_max_runs<NR>(k, v, [&](int q, const unsigned long long* m, bool whole) {
        if (whole) {
            for (int c = 0; c < NR; ++c) st_dev(&hblk[(int64_t)blockIdx.x * kRed + c], m[c]);
            st_dev(&kblk[blockIdx.x], q);
            slot = true;
        } else {  // a block straddling instances
            for (int c = 0; c < NR; ++c) atomicMax(&inst[q].red_stage[c], m[c]);
        }
    });
    if (threadIdx.x == 0 && !slot) st_dev(&kblk[blockIdx.x], -1);
    if (last_to_arrive(arrive)) {
        // the blocks' slots, each thread a contiguous chunk
        const int nb = (int)gridDim.x;
        const int per = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
        const int q0 = (int)threadIdx.x * per, q1 = min(nb, q0 + per);
        int kc = -1;
        unsigned long long mc[kRed] = {};
        auto flush = [&](int q, const unsigned long long* m) {
            for (int c = 0; c < NR; ++c) atomicMax(&inst[q].red_stage[c], m[c]);
        };
        for (int q = q0; q < q1; ++q) {
            const int kq = ld_dev(&kblk[q]);
            if (kq < 0) continue;
            if (kq != kc) {
                if (kc >= 0) flush(kc, mc);
                kc = kq;
                for (int c = 0; c < kRed; ++c) mc[c] = 0ull;
            }
            for (int c = 0; c < NR; ++c) {
                const unsigned long long x = ld_dev(&hblk[(int64_t)q * kRed + c]);
                mc[c] = x > mc[c] ? x : mc[c];
            }
        }
        block_max_runs<NR>(kc, mc, [&](int q, const unsigned long long* m, bool) { flush(q, m); });
        __syncthreads();
        if (advance) {
            for (int q = threadIdx.x; q < ninst; q += blockDim.x) advance_clock(inst[q]);
        } else if (threadIdx.x == 0) {
            ctl->abort = 0;
        }
        __syncthreads();
        eval_begin_body(inst, ninst, ctl, 1);
        __syncthreads();
        // the reductions, cleared for the next step; the same thread then
        // sets up that instance's grid
        for (int q = threadIdx.x; q < ninst; q += blockDim.x) {
            for (int c = 0; c < kRed; ++c) {
                inst[q].red[c] = ld_dev(&inst[q].red_stage[c]);
                inst[q].red_stage[c] = 0ull;
            }
        }
        grid_setup(inst, ninst, ctl);
    }
}

// The predicted h and positions of every particle are reduced to its
// instance's h_max, h_min and box on the way (the reduction k_hmax would
// otherwise launch for): each block leaves its kRed maxima in its own slot
// (hblk, kblk; same-address atomics from every block would serialise), and
// the last workgroup combines the slots, commits the clock, starts the next
// evaluation (eval_begin_body with the pending dt cleared) and sets up its
// cell grid: work that would otherwise be three more launches.
template <int NR>
__global__ void __launch_bounds__(kBlock) k_stage2_init_stage1(Fields F, Inst* inst,
                                                               const int32_t* __restrict__ inst_of, int64_t n,
                                                               int ninst, Ctl* ctl, int32_t* arrive,
                                                               unsigned long long* hblk, int32_t* kblk) {
    if (ctl->abort) return;  // the whole launch: the step after an unfinished evaluation
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1;
    unsigned long long v[kRed] = {};
    if (i < n) {
        const double4 r = stage2_init_stage1_one(F, inst, inst_of, i);
        red_values(v, r.w, r.x, r.y, r.z);
        k = inst_of[i];
    }
    stage_reduce_begin<NR>(k, v, inst, ninst, ctl, arrive, hblk, kblk, true);
}

// The last step of a call: stage2 (k_stage2's body), then the clock commit
// (k_clock) by the last workgroup, once every particle has read the step
__global__ void __launch_bounds__(kBlock) k_stage2_clock(Fields F, Inst* inst, const int32_t* __restrict__ inst_of,
                                                         int64_t n, int ninst, int32_t* arrive) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const Inst& s = inst[inst_of[i]];
        do_stage2(F, s, i, s.dt);
        do_wrap(F, s, i);
    }
    if (last_to_arrive(arrive))
        for (int q = threadIdx.x; q < ninst; q += blockDim.x) advance_clock(inst[q]);
}

// The first step of a call: initialize and stage1 (k_initialize and
// k_stage1's bodies, in that order), the h and box reductions of the
// evaluation that follows, and its start (as k_eval_begin, then k_hmax's
// tail): three launches fewer than the unfused path.
template <int NR>
__global__ void __launch_bounds__(kBlock) k_init_stage1(Fields F, Inst* inst, const int32_t* __restrict__ inst_of,
                                                        int64_t n, int ninst, Ctl* ctl, int32_t* arrive,
                                                        unsigned long long* hblk, int32_t* kblk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1;
    unsigned long long v[kRed] = {};
    if (i < n) {
        do_initialize(F, i);
        const Inst& s = inst[inst_of[i]];
        do_stage1(F, s, i, s.dt);
        do_wrap(F, s, i);
        const double* z = F.f[SPHMHD_F_Z];
        red_values(v, F.f[SPHMHD_F_H][i], F.f[SPHMHD_F_X][i], F.f[SPHMHD_F_Y][i], z ? z[i] : 0.0);
        k = inst_of[i];
    }
    stage_reduce_begin<NR>(k, v, inst, ninst, ctl, arrive, hblk, kblk, false);
}

__global__ void k_clock(Inst* inst, int ninst) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    advance_clock(inst[k]);
}

// application.py:134-147 followed by the solver's first _get_timestep
// (solver.py:40): equal to min dt_adapt when every dt_cfl > 0.
__global__ void k_warm_dt(Fields F, Inst* inst, const int32_t* __restrict__ inst_of, int64_t n, Ctl* ctl) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(F.f[SPHMHD_F_DT_CFL][i] > 0.0)) atomicOr(&ctl->status, 4);
}

__global__ void k_warm_dt_commit(Inst* inst, int ninst) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    Inst& s = inst[k];
    const double pending = bitsd(s.dtp_bits);
    double dt = pending;
    if (s.t + dt > s.tf) dt = s.tf - s.t;
    s.dt = dt;
}

__global__ void k_sync_initial(Fields F, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rho = FLD(RHO);
    const double bx = FLD(BX) / rho, by = FLD(BY) / rho, bz = FLD(BZ) / rho;
    FLD(BEVOLX) = bx;
    FLD(BEVOLY) = by;
    FLD(BEVOLZ) = bz;
    FLD(PSI) = 0.0;
    FLD(BEVOLXPRED) = bx;
    FLD(BEVOLYPRED) = by;
    FLD(BEVOLZPRED) = bz;
    FLD(PSIPRED) = 0.0;
    FLD(UPRED) = FLD(U);
    FLD(VPRED) = FLD(V);
    FLD(WPRED) = FLD(W);
    FLD(EPRED) = FLD(E);
}

__global__ void k_count_unconverged(Fields F, int64_t n, Ctl* ctl) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = i < n && F.f[SPHMHD_F_CONVERGED][i] != 1.0;
    unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd((unsigned long long*)&ctl->count, (unsigned long long)__popcll(m));
}

__global__ void k_zero_count(Ctl* ctl) { ctl->count = 0; }

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// conserved totals: one block per instance, fixed-order tree reduction
__global__ void k_totals(Fields F, const Inst* __restrict__ inst, double* __restrict__ out) {
    const Inst& s = inst[blockIdx.x];
    double acc[SPHMHD_NTOTALS];
#pragma unroll
    for (int q = 0; q < SPHMHD_NTOTALS; ++q) acc[q] = 0.0;
    for (int64_t i = s.off0 + threadIdx.x; i < s.off1; i += blockDim.x) {
        const double m = FLD(M), rho = FLD(RHO);
        const double u = FLD(U), v = FLD(V), w = FLD(W);
        const double bx = FLD(BX), by = FLD(BY), bz = FLD(BZ);
        const double v2 = u * u + v * v + w * w;
        const double b2 = bx * bx + by * by + bz * bz;
        acc[SPHMHD_T_MASS] += m;
        acc[SPHMHD_T_MOMX] += m * u;
        acc[SPHMHD_T_MOMY] += m * v;
        acc[SPHMHD_T_MOMZ] += m * w;
        acc[SPHMHD_T_EKIN] += 0.5 * m * v2;
        acc[SPHMHD_T_ETHERM] += m * FLD(E);
        acc[SPHMHD_T_EMAG] += 0.5 * m * b2 / (s.mu0 * rho);
        if (b2 > 0.0) acc[SPHMHD_T_DIVB] += (m / rho) * FLD(H) * fabs(FLD(DIVBSYMM)) / sqrt(b2);
        acc[SPHMHD_T_VOLUME] += m / rho;
    }
    __shared__ double red[kBlock];
    for (int q = 0; q < SPHMHD_NTOTALS; ++q) {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int off = kBlock / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(int64_t)blockIdx.x * SPHMHD_NTOTALS + q] = red[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void launch_initialize(Dev& d) {
    if (d.n) hipLaunchKernelGGL(k_initialize, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.n);
}
void launch_stage1(Dev& d) {
    if (d.n) hipLaunchKernelGGL(k_stage1, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst, d.inst_of, d.n);
}
void launch_stage2(Dev& d) {
    if (d.n) hipLaunchKernelGGL(k_stage2, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst, d.inst_of, d.n);
}
void launch_stage2_init_stage1(Dev& d) {
    if (d.n)
    {
        // periodic domains reduce h_max and h_min, open ones their box too
        if (d.periodic)
            hipLaunchKernelGGL(k_stage2_init_stage1<2>, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst,
                               d.inst_of, d.n, d.ninst, d.ctl, d.arrive, d.hblk, d.kblk);
        else
            hipLaunchKernelGGL(k_stage2_init_stage1<kRed>, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields,
                               d.inst, d.inst_of, d.n, d.ninst, d.ctl, d.arrive, d.hblk, d.kblk);
    }
}
void launch_stage2_clock(Dev& d) {
    if (d.n == 0) {
        launch_clock(d);
        return;
    }
    hipLaunchKernelGGL(k_stage2_clock, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst, d.inst_of, d.n,
                       d.ninst, d.arrive);
}

void launch_init_stage1(Dev& d) {
    if (d.n == 0) return;
    if (d.periodic)
        hipLaunchKernelGGL(k_init_stage1<2>, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst, d.inst_of,
                           d.n, d.ninst, d.ctl, d.arrive, d.hblk, d.kblk);
    else
        hipLaunchKernelGGL(k_init_stage1<kRed>, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst,
                           d.inst_of, d.n, d.ninst, d.ctl, d.arrive, d.hblk, d.kblk);
}
void launch_clock(Dev& d) {
    hipLaunchKernelGGL(k_clock, dim3(nblk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, d.ninst);
}
void launch_sync_initial(Dev& d) {
    if (d.n) hipLaunchKernelGGL(k_sync_initial, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.n);
}
void launch_copy(Dev& d, double* dst, const double* src, int64_t n) {
    if (n) hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(kBlock), 0, d.stream, dst, src, n);
}
void launch_fill_i32(Dev& d, int32_t* p, int32_t v, int64_t n) {
    if (n) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(kBlock), 0, d.stream, p, v, n);
}
void launch_warm_dt(Dev& d) {
    if (d.n) hipLaunchKernelGGL(k_warm_dt, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst, d.inst_of, d.n, d.ctl);
    hipLaunchKernelGGL(k_warm_dt_commit, dim3(nblk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, d.ninst);
}
void launch_count_unconverged(Dev& d) {
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, d.stream, d.ctl);
    if (d.n) hipLaunchKernelGGL(k_count_unconverged, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.n, d.ctl);
}
void launch_totals(Dev& d) {
    hipLaunchKernelGGL(k_totals, dim3(d.ninst), dim3(kBlock), 0, d.stream, d.fields, d.inst, d.totals);
}

}  // namespace sph
