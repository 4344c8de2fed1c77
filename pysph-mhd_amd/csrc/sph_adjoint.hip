// sph_adjoint.hip -- the adjoint (reverse-mode) step: kernels.
//
// The reverse sweep of one PEC step (integrator.py:12-82 around the group
// plan of scheme.py:70-81) runs after the step's forward part (initialize,
// stage1 and the evaluation) has been recomputed by the primal kernels, so
// the property rows hold the evaluation's inputs and outputs.  Every bar row
// b[f] is the derivative of the objective with respect to the current
// version of property f; the kernels walk the step backwards:
//
//   k_a_stage2        MHDStep.stage2 (B = Bevol rho, the half-step updates)
//   k_a_force_epi     per destination: the force pass's sums recomputed over
//                     its neighbour list, then the epilogue reversed: bars of
//                     the 35 sums (acc rows) and of the particle's own values
//   k_a_force_pair    per particle k, over its list: the pair bodies reversed
//                     with k as destination (its sums' bars) and with k as
//                     source of each neighbour j (j's sums' bars), gathered so
//                     that a thread only writes its own particle's bars
//   k_a_prep          G1/G2/G4 (rho from h, EOS, B = Bevolpred rho)
//   k_a_density       G0 per particle: the Newton passes recomputed from the h
//                     entering the evaluation, then reversed pass by pass
//                     (post_loop over Dual<3>); per pass the bars of the two
//                     sums are kept for the positions
//   k_a_density_pos   the positions' share of every pass's sums, gathered over
//                     a stencil covering the largest h of any pass
//   k_a_stage1_init   MHDStep.stage1 and initialize
// and the per-instance dt: the pending minimum's bar enters at the particle
// attaining it (lowest storage index on ties); the bars of a step's dt
// (summed over its particles) become the pending bar of the evaluation
// before it.  Gradients are the derivative of the discrete map the forward
// step computes (sph_adjoint.h).
#include <climits>

#include "sph_adjoint.h"

namespace sph {

using namespace adj;

#define FLD(name) F.f[SPHMHD_F_##name][i]
#define BAR(name) A.b[SPHMHD_F_##name][i]

namespace {

template <int K, int DIM>
__device__ __forceinline__ Par<double> par_of(const Inst& s) {
    Par<double> P;
    P.av = s.av;
    P.ac = s.ac;
    P.ar = s.ar;
    P.chf = s.chf;
    P.sigma = s.sigma;
    P.alpha_max = s.alpha_max;
    P.gamma = s.gamma;
    P.gamma1 = s.gamma1;
    P.mu0 = s.mu0;
    P.inv_mu0 = 1.0 / s.mu0;
    P.alpha_min = s.alpha_min;
    P.c_cour = s.c_cour;
    P.c_force = s.c_force;
    P.cleaning_active = s.cleaning_active;
    P.hfact = s.hfact;
    P.ksig = kernel_sigma(K, DIM);
    return P;
}

// the force pass's inputs of particle i: predicted v, e, B, psi, final h,
// omega, rho, cs and the previous accelerations entering the evaluation
template <int DIM>
__device__ __forceinline__ Raw<double> load_raw(const Fields& F, const AdjDev& A, int64_t i, int64_t n) {
    Raw<double> r;
    r.x = FLD(X);
    r.y = FLD(Y);
    r.z = DIM == 3 ? FLD(Z) : 0.0;
    r.h = FLD(H);
    r.rho = FLD(RHO);
    r.e = FLD(EPRED);
    r.u = FLD(UPRED);
    r.v = FLD(VPRED);
    r.w = FLD(WPRED);
    r.om = FLD(OMEGA);
    r.bx = FLD(BX);
    r.by = FLD(BY);
    r.bz = FLD(BZ);
    r.al = FLD(ALPHA1);
    r.au = A.ain[2 * n + i];
    r.av = A.ain[3 * n + i];
    r.aw = A.ain[4 * n + i];
    r.psi = FLD(PSIPRED);
    r.cs = FLD(CS);
    r.m = FLD(M);
    return r;
}

// a neighbour's side from its force-pack record (sorted position j): the
// values the primal pair loop reads, the derived ones rebuilt with side_of's
// operations
template <int DIM>
__device__ __forceinline__ Side<double> side_of_rec(const double* __restrict__ fpack, int64_t j, const Par<double>& P,
                                                    double* pos) {
    using R = FRec<DIM>;
    const double* r = fpack + (int64_t)R::N * j;
    Side<double> s;
    pos[0] = r[R::X];
    pos[1] = r[R::Y];
    pos[2] = DIM == 3 ? r[R::Z] : 0.0;
    s.h1 = r[R::H1];
    s.m = r[R::M];
    s.rho = r[R::RHO];
    s.e = r[R::E];
    s.p = P.gamma1 * s.rho * s.e;
    s.u = r[R::U];
    s.v = r[R::V];
    s.w = r[R::W];
    s.om = r[R::OM];
    s.bx = r[R::BX];
    s.by = r[R::BY];
    s.bz = r[R::BZ];
    s.al = r[R::AL];
    s.au = r[R::AU];
    s.av = r[R::AV];
    s.aw = DIM == 3 ? r[FRec<3>::AW] : 0.0;
    s.psich = r[R::PSICH];
    s.vw = r[R::VW];
    s.hir = 0.5 * (1.0 / s.rho);
    return s;
}

template <int DIM>
__device__ __forceinline__ void sep_pos(const Inst& s, const double* a, const double* b, double* d) {
    d[0] = min_image(a[0] - b[0], s.L[0], s.halfL[0], s.periodic);
    d[1] = min_image(a[1] - b[1], s.L[1], s.halfL[1], s.periodic);
    d[2] = DIM == 3 ? min_image(a[2] - b[2], s.L[2], s.halfL[2], s.periodic) : 0.0;
}

// the bars of particle t's sums (rows by sorted position)
template <int DIM>
__device__ __forceinline__ void load_acc(const AdjDev& A, int64_t t, int64_t n, Acc<double>& B) {
    double* p = acc_ptr(B);
#pragma unroll
    for (int k = 0; k < kNAcc; ++k) p[k] = acc_used(k, DIM) ? A.acc[(int64_t)k * n + t] : 0.0;
}

// A group of kAdjLanes consecutive lanes shares one particle: each lane takes
// every kAdjLanes-th neighbour (or candidate), and the lanes' partial sums
// are combined by a butterfly of shuffles (a fixed order: the results do not
// depend on the launch).
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
    for (int off = kAdjLanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ double gmax(double v) {
#pragma unroll
    for (int off = kAdjLanes / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}
template <class S>
__device__ __forceinline__ void gsum_all(S& x) {
    double* p = reinterpret_cast<double*>(&x);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(S) / sizeof(double)); ++k) p[k] = gsum(p[k]);
}

}  // namespace

// ---------------------------------------------------------------------------
// bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_a_begin_eval(AdjInst* ai, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    ai[k].argmin = INT_MAX;
    ai[k].hmax_bits = 0ull;
}

// the particle attaining the pending minimum (lowest storage index)
__global__ void k_a_argmin(Fields F, const Inst* __restrict__ inst, AdjInst* ai, const int32_t* __restrict__ inst_of,
                           int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = inst_of[i];
    if (F.f[SPHMHD_F_DT_ADAPT][i] == bitsd(inst[k].dtp_bits)) atomicMin(&ai[k].argmin, (int)i);
}

// The clock between the evaluation being reversed and the next step:
// dt_next = the pending minimum clipped to tf - t (adaptive), else the
// current dt; a clipped step carries no derivative (the end time is fixed).
// dtbar holds the bar of dt_next (from the step after); it becomes the
// pending bar, or (fixed steps) carries to the current dt.
// mode 1: the warm start's first dt (always the pending minimum, clipped);
// mode 2: no dt follows this evaluation.
__global__ void k_a_pend(const Inst* __restrict__ inst, AdjInst* ai, const double* __restrict__ clk_next, int ninst,
                         int mode) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    const Inst& s = inst[k];
    AdjInst& a = ai[k];
    a.pendbar = 0.0;
    a.carry = 0.0;
    if (mode != 2) {
        const double dt_next = clk_next[2 * k + 1];
        if (s.adaptive || mode == 1) {
            if (dt_next == bitsd(s.dtp_bits)) a.pendbar = a.dtbar;
        } else if (dt_next == s.dt) {
            a.carry = a.dtbar;
        }
    }
    a.dtbar = 0.0;
}

// sum of the per-particle dt bars of one instance (one block per instance)
__global__ void k_a_reduce_dt(const Inst* __restrict__ inst, AdjInst* ai, const double* __restrict__ row) {
    const Inst& s = inst[blockIdx.x];
    // four partial sums, so four loads are in flight per thread
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t stride = blockDim.x;
    int64_t i = s.off0 + threadIdx.x;
    for (; i + 3 * stride < s.off1; i += 4 * stride) {
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += row[i + q * stride];
    }
    for (; i < s.off1; i += stride) acc[0] += row[i];
    __shared__ double red[kBlock];
    red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) ai[blockIdx.x].dtbar = red[0] + ai[blockIdx.x].carry;
}

// parameter bars of one instance (one block per instance)
__global__ void k_a_reduce_par(const Inst* __restrict__ inst, AdjInst* ai, const double* __restrict__ par,
                               int64_t n) {
    const Inst& s = inst[blockIdx.x];
    __shared__ double red[kBlock];
    for (int p = 0; p < SPHMHD_NTPARAMS; ++p) {
        double acc = 0.0;
        for (int64_t i = s.off0 + threadIdx.x; i < s.off1; i += blockDim.x) acc += par[(int64_t)p * n + i];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int off = kBlock / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) ai[blockIdx.x].par[p] = red[0];
        __syncthreads();
    }
}

// rows of the listed properties to or from a buffer (rows in list order)
__global__ void k_a_rows(Fields F, RowList L, double* __restrict__ buf, int64_t n, int restore) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;  // one row per grid row: the row's pointer is uniform
    double* f = F.f[L.f[r]];
    if (i >= n || !f) return;
    if (restore) f[i] = buf[(int64_t)r * n + i];
    else buf[(int64_t)r * n + i] = f[i];
}

// per-instance clock {t, dt} to or from a device array
__global__ void k_a_clock(Inst* inst, double* clk, int ninst, int restore) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    if (restore) {
        inst[k].t = clk[2 * k];
        inst[k].dt = clk[2 * k + 1];
    } else {
        clk[2 * k] = inst[k].t;
        clk[2 * k + 1] = inst[k].dt;
    }
}

// what the evaluation reads before it writes it: h, converged and the
// previous accelerations
__global__ void k_a_save_entry(Fields F, double* __restrict__ ain, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ain[i] = FLD(H);
    ain[n + i] = FLD(CONVERGED);
    ain[2 * n + i] = FLD(AU);
    ain[3 * n + i] = FLD(AV);
    ain[4 * n + i] = FLD(AW);
}

// bars of the final state for the objective sum_q w_q total_q (k_totals'
// formulas; the divB total has no derivative and must have weight 0)
__global__ void k_a_seed_totals(Fields F, AdjDev A, const Inst* __restrict__ inst,
                                const int32_t* __restrict__ inst_of, int64_t n, Weights W) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Inst& s = inst[inst_of[i]];
    const double m = FLD(M), rho = FLD(RHO);
    const double u = FLD(U), v = FLD(V), w = FLD(W);
    const double bx = FLD(BX), by = FLD(BY), bz = FLD(BZ);
    const double* q = W.w;
    BAR(U) += q[SPHMHD_T_MOMX] * m + q[SPHMHD_T_EKIN] * m * u;
    BAR(V) += q[SPHMHD_T_MOMY] * m + q[SPHMHD_T_EKIN] * m * v;
    BAR(W) += q[SPHMHD_T_MOMZ] * m + q[SPHMHD_T_EKIN] * m * w;
    BAR(E) += q[SPHMHD_T_ETHERM] * m;
    const double mr = s.mu0 * rho;
    BAR(BX) += q[SPHMHD_T_EMAG] * m * bx / mr;
    BAR(BY) += q[SPHMHD_T_EMAG] * m * by / mr;
    BAR(BZ) += q[SPHMHD_T_EMAG] * m * bz / mr;
    const double b2 = bx * bx + by * by + bz * bz;
    BAR(RHO) += -q[SPHMHD_T_EMAG] * 0.5 * m * b2 / (mr * rho) - q[SPHMHD_T_VOLUME] * m / (rho * rho);
}

// ---------------------------------------------------------------------------
// MHDStep.stage2 (integrator.py:66-82), then the periodic wrap (identity)
// ---------------------------------------------------------------------------

__global__ void k_a_stage2(Fields F, AdjDev A, const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of,
                           int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dtb = 0.0;
    if (F.itype[i] == 1) {
        const Inst& s = inst[inst_of[i]];
        const double dt = s.dt, dtb2 = 0.5 * dt;
        const double rho = FLD(RHO);
        // B = Bevol'' rho, Bevol'' = Bevol' + dt/2 aBevol
        const double ex = FLD(BEVOLX) + dtb2 * FLD(ABEVOLX);
        const double ey = FLD(BEVOLY) + dtb2 * FLD(ABEVOLY);
        const double ez = FLD(BEVOLZ) + dtb2 * FLD(ABEVOLZ);
        BAR(BEVOLX) += BAR(BX) * rho;
        BAR(BEVOLY) += BAR(BY) * rho;
        BAR(BEVOLZ) += BAR(BZ) * rho;
        BAR(RHO) += BAR(BX) * ex + BAR(BY) * ey + BAR(BZ) * ez;
        BAR(BX) = 0.0;
        BAR(BY) = 0.0;
        BAR(BZ) = 0.0;
        BAR(P) = 0.0;  // p = gamma1 rho e: read by nothing after the step
        // x += dt/2 a (in place: the bar of x passes through)
        BAR(APSI) += dtb2 * BAR(PSI);
        dtb += 0.5 * FLD(APSI) * BAR(PSI);
        BAR(ABEVOLX) += dtb2 * BAR(BEVOLX);
        BAR(ABEVOLY) += dtb2 * BAR(BEVOLY);
        BAR(ABEVOLZ) += dtb2 * BAR(BEVOLZ);
        dtb += 0.5 * (FLD(ABEVOLX) * BAR(BEVOLX) + FLD(ABEVOLY) * BAR(BEVOLY) + FLD(ABEVOLZ) * BAR(BEVOLZ));
        BAR(AE) += dtb2 * BAR(E);
        dtb += 0.5 * FLD(AE) * BAR(E);
        BAR(AU) += dtb2 * BAR(U);
        BAR(AV) += dtb2 * BAR(V);
        BAR(AW) += dtb2 * BAR(W);
        dtb += 0.5 * (FLD(AU) * BAR(U) + FLD(AV) * BAR(V) + FLD(AW) * BAR(W));
    }
    A.dtrow[i] = dtb;
}

// ---------------------------------------------------------------------------
// the force pass (G3 + G5..G9), reversed
// ---------------------------------------------------------------------------

template <int K, int DIM>
__global__ void __launch_bounds__(kAdjBlock) k_a_force_epi(Fields F, AdjDev A, const double* __restrict__ fpack,
                                                          const int32_t* __restrict__ perm,
                                                          const int32_t* __restrict__ nb,
                                                          const int32_t* __restrict__ nbcnt,
                                                          const Inst* __restrict__ inst,
                                                          const int32_t* __restrict__ inst_of, int64_t n,
                                                          int stride) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = gid / kAdjLanes;
    const int lane = (int)(gid % kAdjLanes);
    if (t >= n) return;  // the whole group
    const int32_t i = perm[t];
    const int ki = inst_of[t];
    const Inst& s = inst[ki];
    const Par<double> P = par_of<K, DIM>(s);
    const Raw<double> me = load_raw<DIM>(F, A, i, n);
    const Side<double> S = side_of<DIM>(me, P);
    const double xi[3] = {me.x, me.y, me.z};
    Acc<double> acc;
    acc_zero(acc);
    // the maximum pair candidate of dt_cfl and the first list slot attaining
    // it (the forward max keeps its first operand on ties; -1: the initial 0)
    double best = 0.0;
    int qbest = -1;
    const int cnt = min(nbcnt[t], stride);
    for (int q = lane; q < cnt; q += kAdjLanes) {
        const int32_t j = nb[(int64_t)q * n + t];
        double xj[3];
        const Side<double> So = side_of_rec<DIM>(fpack, j, P, xj);
        double d[3];
        sep_pos<DIM>(s, xi, xj, d);
        const double c = pair_fwd<K, DIM>(S, So, d[0], d[1], d[2], P, acc);
        if (c > best) {
            best = c;
            qbest = q;
        }
    }
    double* ap0 = acc_ptr(acc);
#pragma unroll
    for (int k = 0; k < kNAcc; ++k)
        if (k != 5 && acc_used(k, DIM)) ap0[k] = gsum(ap0[k]);
    acc.dtcfl = gmax(acc.dtcfl);
    {
        int qb = (best == acc.dtcfl && qbest >= 0) ? qbest : INT_MAX;
#pragma unroll
        for (int off = kAdjLanes / 2; off > 0; off >>= 1) qb = min(qb, __shfl_xor(qb, off));
        qbest = qb;
    }
    if (lane != 0) return;
    const int32_t amax = qbest == INT_MAX ? -1 : nb[(int64_t)qbest * n + t];
    Epi<double> O;
    O.ax = BAR(AU);
    O.ay = BAR(AV);
    O.az = BAR(AW);
    O.ae = BAR(AE);
    O.ah = BAR(AH);
    O.abx = BAR(ABEVOLX);
    O.aby = BAR(ABEVOLY);
    O.abz = BAR(ABEVOLZ);
    O.apsi = BAR(APSI);
    O.alphaloc = BAR(ALPHALOC);
    O.dt_adapt = A.ai[ki].argmin == i ? A.ai[ki].pendbar : 0.0;
    Acc<double> Ab;
    acc_zero(Ab);
    RawBar rb{};
    double b_vw = 0.0, b_p = 0.0;
    EpiParBar pb{0.0, 0.0, 0.0};
    epi_rev<DIM>(acc, me, S, P, O, Ab, rb, &b_vw, &b_p, pb);
    SideBar sb{};
    sb.vw = b_vw;
    sb.p = b_p;
    side_rev(me, S, P, sb, rb);
    const double* ap = acc_ptr(Ab);
#pragma unroll
    for (int k = 0; k < kNAcc; ++k)
        if (acc_used(k, DIM)) A.acc[(int64_t)k * n + t] = ap[k];
    A.argmax[t] = amax;
    // the sums' outputs are consumed; what the epilogue read of the particle
    // adds to the bars of those values
    BAR(AE) = 0.0;
    BAR(AH) = 0.0;
    BAR(ABEVOLX) = 0.0;
    BAR(ABEVOLY) = 0.0;
    BAR(ABEVOLZ) = 0.0;
    BAR(APSI) = 0.0;
    BAR(ALPHALOC) = 0.0;
    BAR(H) += rb.h;
    BAR(CS) += rb.cs;
    BAR(PSIPRED) += rb.psi;
    BAR(RHO) += rb.rho;
    BAR(OMEGA) += rb.om;
    BAR(BX) += rb.bx;
    BAR(BY) += rb.by;
    BAR(BZ) += rb.bz;
    BAR(EPRED) += rb.e;
    A.par[(int64_t)SPHMHD_TP_CLEANING_SPEED_FACTOR * n + i] += pb.chf;
    A.par[(int64_t)SPHMHD_TP_CLEANING_DAMPING_FACTOR * n + i] += pb.sigma;
    A.par[(int64_t)SPHMHD_TP_ALPHA_MAX * n + i] += pb.alpha_max;
}

template <int K, int DIM>
__global__ void __launch_bounds__(kAdjBlock) k_a_force_pair(Fields F, AdjDev A, const double* __restrict__ fpack,
                                                           const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ nb,
                                                           const int32_t* __restrict__ nbcnt,
                                                           const Inst* __restrict__ inst,
                                                           const int32_t* __restrict__ inst_of, int64_t n,
                                                           int stride) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = gid / kAdjLanes;
    const int lane = (int)(gid % kAdjLanes);
    if (t >= n) return;  // the whole group
    const int32_t i = perm[t];
    const Inst& s = inst[inst_of[t]];
    const Par<double> P = par_of<K, DIM>(s);
    const Raw<double> me = load_raw<DIM>(F, A, i, n);
    const Side<double> S = side_of<DIM>(me, P);
    const double xi[3] = {me.x, me.y, me.z};
    Acc<double> Bi;
    load_acc<DIM>(A, t, n, Bi);
    const int32_t amax_i = A.argmax[t];
    struct Part {
        SideBar own;
        double xb[3];
        PairParBar pb;
    } part{};
    const int cnt = min(nbcnt[t], stride);
    for (int q = lane; q < cnt; q += kAdjLanes) {
        const int32_t j = nb[(int64_t)q * n + t];
        double xj[3];
        const Side<double> So = side_of_rec<DIM>(fpack, j, P, xj);
        double d[3];
        sep_pos<DIM>(s, xi, xj, d);
        // t as the destination: the bars of t's sums
        {
            SideBar jb{};
            double db[3] = {0.0, 0.0, 0.0};
            pair_rev<K, DIM>(S, So, d[0], d[1], d[2], P, Bi, amax_i == j, part.own, jb, db, part.pb);
            part.xb[0] += db[0];
            part.xb[1] += db[1];
            part.xb[2] += db[2];
        }
        // t as a source of j: the bars of j's sums (the separation is -d)
        {
            Acc<double> Bj;
            load_acc<DIM>(A, j, n, Bj);
            SideBar ib{};
            double db[3] = {0.0, 0.0, 0.0};
            PairParBar pdrop{0.0, 0.0, 0.0};  // counted by j's own group
            pair_rev<K, DIM>(So, S, -d[0], -d[1], -d[2], P, Bj, A.argmax[j] == t, ib, part.own, db, pdrop);
            part.xb[0] -= db[0];
            part.xb[1] -= db[1];
            part.xb[2] -= db[2];
        }
    }
    gsum_all(part);
    if (lane != 0) return;
    const SideBar& own = part.own;
    RawBar rb{};
    side_rev(me, S, P, own, rb);
    BAR(X) += part.xb[0];
    BAR(Y) += part.xb[1];
    if (DIM == 3) BAR(Z) += part.xb[2];
    BAR(H) += rb.h;
    BAR(RHO) += rb.rho;
    BAR(EPRED) += rb.e;
    BAR(UPRED) += rb.u;
    BAR(VPRED) += rb.v;
    BAR(WPRED) += rb.w;
    BAR(OMEGA) += rb.om;
    BAR(BX) += rb.bx;
    BAR(BY) += rb.by;
    BAR(BZ) += rb.bz;
    BAR(ALPHA1) += rb.al;
    // the previous accelerations entering the evaluation: their bar replaces
    // the (consumed) bar of the accelerations the evaluation produced
    BAR(AU) = rb.au;
    BAR(AV) = rb.av;
    BAR(AW) = rb.aw;
    BAR(PSIPRED) += rb.psi;
    BAR(CS) += rb.cs;
    A.par[(int64_t)SPHMHD_TP_ARTIFICIAL_VISCOSITY * n + i] += part.pb.av;
    A.par[(int64_t)SPHMHD_TP_ARTIFICIAL_THERMAL_CONDUCTIVITY * n + i] += part.pb.ac;
    A.par[(int64_t)SPHMHD_TP_ARTIFICIAL_MAGNETIC_DISSIPATION * n + i] += part.pb.ar;
    // psich = psipred chf vw: every pair's psich of this particle, in either role
    A.par[(int64_t)SPHMHD_TP_CLEANING_SPEED_FACTOR * n + i] += own.psich * me.psi * S.vw;
}

// G1 + G2 + G4 (equations.py:105-106, 117-118, 477-485): rho = m / (h/k)^d,
// p = (gamma - 1) rho epred, cs = sqrt(gamma p / rho), B = Bevolpred rho
__global__ void k_a_prep(Fields F, AdjDev A, const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of,
                         int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Inst& s = inst[inst_of[i]];
    const int dim = s.dim;
    const double h = FLD(H), m = FLD(M);
    const double rho = m / (dim == 2 ? (h / s.hfact) * (h / s.hfact) : (h / s.hfact) * (h / s.hfact) * (h / s.hfact));
    const double epred = FLD(EPRED);
    const double p = s.gamma1 * rho * epred;
    const double cs = sqrt(s.gamma * p / rho);
    double brho = BAR(RHO);
    const double bbx = BAR(BX) + BAR(BXPRED), bby = BAR(BY) + BAR(BYPRED), bbz = BAR(BZ) + BAR(BZPRED);
    BAR(BEVOLXPRED) += bbx * rho;
    BAR(BEVOLYPRED) += bby * rho;
    BAR(BEVOLZPRED) += bbz * rho;
    brho += bbx * FLD(BEVOLXPRED) + bby * FLD(BEVOLYPRED) + bbz * FLD(BEVOLZPRED);
    double bp = BAR(P);
    if (cs > 0.0) {
        const double u = BAR(CS) * 0.5 / cs;
        bp += u * s.gamma / rho;
        brho += -u * s.gamma * p / (rho * rho);
    }
    brho += bp * s.gamma1 * epred;
    BAR(EPRED) += bp * s.gamma1 * rho;
    BAR(H) += brho * (-(double)dim * rho / h);
    BAR(RHO) = 0.0;
    BAR(CS) = 0.0;
    BAR(P) = 0.0;
    BAR(BX) = 0.0;
    BAR(BY) = 0.0;
    BAR(BZ) = 0.0;
    BAR(BXPRED) = 0.0;
    BAR(BYPRED) = 0.0;
    BAR(BZPRED) = 0.0;
}

// ---------------------------------------------------------------------------
// G0: the density iteration (equations.py:14-86), reversed
// ---------------------------------------------------------------------------

// the density sums of particle t at h, with their derivative in h: the
// primal pass's scan (sph_pairs.hip density_block) over a dual h, reading
// the density pack {x, y, z, m} in sorted order
template <int K, int DIM>
__device__ __forceinline__ void density_sums(const Inst& s, const double4* __restrict__ dpack,
                                             const uint32_t* __restrict__ skey, const int32_t* __restrict__ cstart,
                                             const int32_t* __restrict__ cend, int64_t t, int lane,
                                             const double4& me, double hi, Dual<1>& sw, Dual<1>& sg) {
    const double R = kernel_radius(K);
    Dual<1> h(hi);
    h.d[0] = 1.0;
    const Dual<1> h1 = 1.0 / h;
    const double Rh = R * hi;
    const double Rh2 = Rh * Rh;
    int sr[3] = {0, 0, 0};
    for (int a = 0; a < DIM; ++a) {
        const double q = ceil(Rh / s.ceff[a]);
        sr[a] = q < 1.0 ? 1 : (q < 1.0e6 ? (int)q : 1000000);
    }
    const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
    const double H0 = s.halfL[0], H1 = s.halfL[1], H2 = s.halfL[2];
    const int per = s.periodic;
    sw = Dual<1>(0.0);
    sg = Dual<1>(0.0);
    for_each_row_range<DIM>(s, (int64_t)skey[t], cstart, cend, [&](int a, int b, double, double, double) {
        for (int j = a + lane; j < b; j += kAdjLanes) {
            const double4 o = dpack[j];
            const double dx = min_image(me.x - o.x, L0, H0, per);
            const double dy = min_image(me.y - o.y, L1, H1, per);
            const double dz = DIM == 3 ? min_image(me.z - o.z, L2, H2, per) : 0.0;
            const double r2 = r2_exact(dx, dy, dz);
            if (r2 < Rh2) {
                const Dual<1> q = sqrt(r2) * h1;
                Dual<1> w, dw;
                shape<K, DIM>(q, w, dw);
                sw += o.w * w;
                sg += o.w * ((double)DIM * w + q * dw);
            }
        }
    }, sr);
    sw.v = gsum(sw.v);
    sw.d[0] = gsum(sw.d[0]);
    sg.v = gsum(sg.v);
    sg.d[0] = gsum(sg.d[0]);
}

// per particle (sorted position t): the passes of the forward iteration from
// the h entering the evaluation (rows of A.pass by sorted position), then
// reversed; A.pass rows: 0 h, 1 sw then its bar, 2 sg then its bar, 3 dsw/dh,
// 4 dsg/dh
template <int K, int DIM>
__global__ void __launch_bounds__(kAdjBlock) k_a_density(Fields F, AdjDev A, const double4* __restrict__ dpack,
                                                        const int32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ skey,
                                                        const int32_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ cend,
                                                        const Inst* __restrict__ inst,
                                                        const int32_t* __restrict__ inst_of, int64_t n) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = gid / kAdjLanes;
    const int lane = (int)(gid % kAdjLanes);
    const bool valid = t < n;
    const int32_t i = valid ? perm[t] : 0;
    const int ki = valid ? inst_of[t] : -1;
    // converged entering the evaluation: h and omega pass through
    const bool active = valid && A.ain[n + i] != 1.0;
    if (valid && !active && lane == 0) A.npass[t] = 0;
    double hmax = 0.0;
    if (active) {
        const Inst& s = inst[ki];
        const double ksig = kernel_sigma(K, DIM);
        const double4 me = dpack[t];
        const double mi = me.w, hi0 = FLD(H0);
        const int maxp = min(s.max_iter, A.passmax);
        const int64_t pn = (int64_t)A.passmax * n;  // rows of one pass quantity
        double h = A.ain[i];
        hmax = h;
        int np = 0;
        double hout = h, omega = 1.0;
        bool fin = false;
        for (int p = 0; p < maxp; ++p) {
            Dual<1> sw, sg;
            density_sums<K, DIM>(s, dpack, skey, cstart, cend, t, lane, me, h, sw, sg);
            if (lane == 0) {
                A.pass[(int64_t)p * n + t] = h;
                A.pass[pn + (int64_t)p * n + t] = sw.v;
                A.pass[2 * pn + (int64_t)p * n + t] = sg.v;
                A.pass[3 * pn + (int64_t)p * n + t] = sw.d[0];
                A.pass[4 * pn + (int64_t)p * n + t] = sg.d[0];
            }
            newton_post<DIM>(h, sw.v, sg.v, mi, hi0, s.hfact, ksig, s.density_iterate, hout, omega, fin);
            np = p + 1;
            if (fin) break;
            h = hout;
            hmax = fmax(hmax, h);
        }
        if (lane == 0) {
            A.npass[t] = np;
            if (!fin && np == A.passmax && s.max_iter > A.passmax) atomicOr(&A.ai[ki].status, 2);
            // the recomputed iteration must end where the forward one did
            if (!(fabs(hout - FLD(H)) <= 1.0e-10 * fabs(FLD(H)))) atomicOr(&A.ai[ki].status, 1);
            // reverse, pass by pass: (h_next, omega) = post_loop(h, sw(h), sg(h))
            double hbar = BAR(H), wbar = BAR(OMEGA);
            for (int p = np - 1; p >= 0; --p) {
                const double hp = A.pass[(int64_t)p * n + t];
                Dual<3> hd(hp), swd(A.pass[pn + (int64_t)p * n + t]), sgd(A.pass[2 * pn + (int64_t)p * n + t]);
                hd.d[0] = 1.0;
                swd.d[1] = 1.0;
                sgd.d[2] = 1.0;
                Dual<3> ho, om;
                bool f;
                newton_post<DIM>(hd, swd, sgd, mi, hi0, s.hfact, ksig, s.density_iterate, ho, om, f);
                const double bsw = hbar * ho.d[1] + wbar * om.d[1];
                const double bsg = hbar * ho.d[2] + wbar * om.d[2];
                const double bh = hbar * ho.d[0] + wbar * om.d[0] + bsw * A.pass[3 * pn + (int64_t)p * n + t] +
                                  bsg * A.pass[4 * pn + (int64_t)p * n + t];
                A.pass[pn + (int64_t)p * n + t] = bsw;
                A.pass[2 * pn + (int64_t)p * n + t] = bsg;
                hbar = bh;
                wbar = 0.0;  // an earlier pass's omega is overwritten
            }
            BAR(H) = hbar;  // the bar of h entering the evaluation
            BAR(OMEGA) = 0.0;
        }
    }
    // the largest h of any pass, per instance: combined over the wave first
    // (same-address atomics serialise), one atomic per wave and instance
    unsigned long long hb = active ? dbits(hmax) : 0ull;
    const int k0 = __shfl(ki, 0);
    const bool uniform = __all(ki == k0 || ki < 0);
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(hb, off);
            hb = o > hb ? o : hb;
        }
        if ((threadIdx.x & (kWave - 1)) == 0 && k0 >= 0 && hb) atomicMax(&A.ai[k0].hmax_bits, hb);
    } else if (active && lane == 0) {
        atomicMax(&A.ai[ki].hmax_bits, hb);
    }
}

// The positions' share: every pass p of every particle a sums
// m_b w(r_ab / h_a^p) and m_b (d w + q w') over r_ab < R h_a^p; with the
// pass's bars (swbar, sgbar) their gradient in the separation x_a - x_b is
// m_b (swbar w' + sgbar ((d + 1) w' + q w'')) / h (x_a - x_b) / r.  Particle
// t gathers its own passes (destination) and every candidate's (source).
template <int K, int DIM>
__global__ void __launch_bounds__(kAdjBlock) k_a_density_pos(Fields F, AdjDev A, const double4* __restrict__ dpack,
                                                            const int32_t* __restrict__ perm,
                                                            const uint32_t* __restrict__ skey,
                                                            const int32_t* __restrict__ cstart,
                                                            const int32_t* __restrict__ cend,
                                                            const Inst* __restrict__ inst,
                                                            const int32_t* __restrict__ inst_of, int64_t n) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = gid / kAdjLanes;
    const int lane = (int)(gid % kAdjLanes);
    if (t >= n) return;  // the whole group
    const int32_t i = perm[t];
    const int ki = inst_of[t];
    const Inst& s = inst[ki];
    const double hm = bitsd(A.ai[ki].hmax_bits);
    if (!(hm > 0.0)) return;  // no particle of the instance ran a pass
    const double R = kernel_radius(K);
    const int64_t pn = (int64_t)A.passmax * n;
    int sr[3] = {0, 0, 0};
    for (int a = 0; a < DIM; ++a) {
        const double q = ceil(R * hm / s.ceff[a]);
        sr[a] = q < 1.0 ? 1 : (q < 1.0e6 ? (int)q : 1000000);
    }
    const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
    const double H0 = s.halfL[0], H1 = s.halfL[1], H2 = s.halfL[2];
    const int per = s.periodic;
    const double4 me = dpack[t];
    const int npi = A.npass[t];
    // the reach of this particle's passes
    double Ri = 0.0;
    for (int p = 0; p < npi; ++p) Ri = fmax(Ri, R * A.pass[(int64_t)p * n + t]);
    const double Rm = R * hm;
    const double Rm2 = Rm * Rm;
    double xb[3] = {0.0, 0.0, 0.0};
    // d phi / d(x_a - x_b) / (x_a - x_b) of pass p of a (sorted position) against b
    auto coef = [&](int64_t a, int p, double mb, double r) {
        const double h = A.pass[(int64_t)p * n + a];
        const double bsw = A.pass[pn + (int64_t)p * n + a], bsg = A.pass[2 * pn + (int64_t)p * n + a];
        const double q = r / h;
        double dw, d2w;
        shape_d12<K, DIM>(q, dw, d2w);
        return mb * (bsw * dw + bsg * ((double)(DIM + 1) * dw + q * d2w)) / (h * r);
    };
    for_each_row_range<DIM>(s, (int64_t)skey[t], cstart, cend, [&](int a, int b, double, double, double) {
        for (int j = a + lane; j < b; j += kAdjLanes) {
            const double4 o = dpack[j];
            const double dx = min_image(me.x - o.x, L0, H0, per);
            const double dy = min_image(me.y - o.y, L1, H1, per);
            const double dz = DIM == 3 ? min_image(me.z - o.z, L2, H2, per) : 0.0;
            const double r2 = r2_exact(dx, dy, dz);
            if (!(r2 > 0.0) || !(r2 < Rm2)) continue;
            const double r = sqrt(r2);
            double c = 0.0;
            if (r < Ri)
                for (int p = 0; p < npi; ++p) {
                    const double Rh = R * A.pass[(int64_t)p * n + t];
                    if (r2 < Rh * Rh) c += coef(t, p, o.w, r);
                }
            const int npj = A.npass[j];
            for (int p = 0; p < npj; ++p) {
                const double Rh = R * A.pass[(int64_t)p * n + j];
                if (r2 < Rh * Rh) c += coef(j, p, me.w, r);  // (x_j - x_t) = -d: the two signs cancel
            }
            xb[0] += c * dx;
            xb[1] += c * dy;
            xb[2] += c * dz;
        }
    }, sr);
    xb[0] = gsum(xb[0]);
    xb[1] = gsum(xb[1]);
    xb[2] = gsum(xb[2]);
    if (lane != 0) return;
    BAR(X) += xb[0];
    BAR(Y) += xb[1];
    if (DIM == 3) BAR(Z) += xb[2];
}

// ---------------------------------------------------------------------------
// MHDStep.stage1 and initialize (integrator.py:12-64), reversed
// ---------------------------------------------------------------------------

// snap: the state rows the step started from (the rates, cs and alpha_loc of
// the previous evaluation, in kStateFields order)
__global__ void k_a_stage1_init(Fields F, AdjDev A, const Inst* __restrict__ inst,
                                const int32_t* __restrict__ inst_of, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool has_z = F.f[SPHMHD_F_Z] != nullptr;
    double dtb = 0.0;
    if (F.itype[i] == 1) {
        const Inst& s = inst[inst_of[i]];
        const double dt = s.dt, dtb2 = 0.5 * dt;
        const double* S = A.snap;
        auto old = [&](int r) { return S[(int64_t)r * n + i]; };
        const double au = old(kSAU), av = old(kSAV), aw = old(kSAW), ae = old(kSAE), ah = old(kSAH);
        const double abx = old(kSABX), aby = old(kSABY), abz = old(kSABZ), apsi = old(kSAPSI);
        const double cs = old(kSCS), al = old(kSALPHALOC);
        double b_au = 0.0, b_av = 0.0, b_aw = 0.0, b_ae = 0.0, b_ah = 0.0;
        double b_abx = 0.0, b_aby = 0.0, b_abz = 0.0, b_apsi = 0.0, b_cs = 0.0, b_al = 0.0;
        // predicted B, psi: x = x0 + dt a
        BAR(PSI0) += BAR(PSIPRED);
        b_apsi += dt * BAR(PSIPRED);
        dtb += apsi * BAR(PSIPRED);
        BAR(PSIPRED) = 0.0;
        BAR(BEVOLX0) += BAR(BEVOLXPRED);
        BAR(BEVOLY0) += BAR(BEVOLYPRED);
        BAR(BEVOLZ0) += BAR(BEVOLZPRED);
        b_abx += dt * BAR(BEVOLXPRED);
        b_aby += dt * BAR(BEVOLYPRED);
        b_abz += dt * BAR(BEVOLZPRED);
        dtb += abx * BAR(BEVOLXPRED) + aby * BAR(BEVOLYPRED) + abz * BAR(BEVOLZPRED);
        BAR(BEVOLXPRED) = 0.0;
        BAR(BEVOLYPRED) = 0.0;
        BAR(BEVOLZPRED) = 0.0;
        // half-step B, psi: x = x0 + dt/2 a
        BAR(PSI0) += BAR(PSI);
        b_apsi += dtb2 * BAR(PSI);
        dtb += 0.5 * apsi * BAR(PSI);
        BAR(PSI) = 0.0;
        BAR(BEVOLX0) += BAR(BEVOLX);
        BAR(BEVOLY0) += BAR(BEVOLY);
        BAR(BEVOLZ0) += BAR(BEVOLZ);
        b_abx += dtb2 * BAR(BEVOLX);
        b_aby += dtb2 * BAR(BEVOLY);
        b_abz += dtb2 * BAR(BEVOLZ);
        dtb += 0.5 * (abx * BAR(BEVOLX) + aby * BAR(BEVOLY) + abz * BAR(BEVOLZ));
        BAR(BEVOLX) = 0.0;
        BAR(BEVOLY) = 0.0;
        BAR(BEVOLZ) = 0.0;
        // alpha1 (integrator.py:47-50) with h = h0 + dt ah the h entering the evaluation
        const double h = A.ain[i];
        double bh = 0.0;
        const double a0 = FLD(ALPHA10);
        const double ba = BAR(ALPHA1);
        if (a0 < al) {
            b_al += ba;
        } else {
            const double dec = 0.1 * cs / h;
            const double num = a0 + dt * al * dec, den = 1.0 + dt * dec;
            const double alpha = num / den;
            const double bnum = ba / den, bden = -ba * alpha / den;
            BAR(ALPHA10) += bnum;
            b_al += bnum * dt * dec;
            const double bdec = bnum * dt * al + bden * dt;
            dtb += bnum * al * dec + bden * dec;
            b_cs += bdec * 0.1 / h;
            bh += -bdec * dec / h;
        }
        BAR(ALPHA1) = 0.0;
        // h = h0 + dt ah
        bh += BAR(H);
        BAR(H0) += bh;
        b_ah += dt * bh;
        dtb += ah * bh;
        BAR(H) = 0.0;
        // e
        BAR(E0) += BAR(EPRED) + BAR(E);
        b_ae += dt * BAR(EPRED) + dtb2 * BAR(E);
        dtb += ae * BAR(EPRED) + 0.5 * ae * BAR(E);
        BAR(EPRED) = 0.0;
        BAR(E) = 0.0;
        // x = x0 + dt u (u after stage1)
        BAR(X0) += BAR(X);
        BAR(Y0) += BAR(Y);
        BAR(U) += dt * BAR(X);
        BAR(V) += dt * BAR(Y);
        dtb += FLD(U) * BAR(X) + FLD(V) * BAR(Y);
        BAR(X) = 0.0;
        BAR(Y) = 0.0;
        if (s.dim == 3 && has_z) {
            BAR(Z0) += BAR(Z);
            BAR(W) += dt * BAR(Z);
            dtb += FLD(W) * BAR(Z);
            BAR(Z) = 0.0;
        }
        // u: upred = u0 + dt a, u = u0 + dt/2 a
        BAR(U0) += BAR(UPRED) + BAR(U);
        BAR(V0) += BAR(VPRED) + BAR(V);
        BAR(W0) += BAR(WPRED) + BAR(W);
        b_au += dt * BAR(UPRED) + dtb2 * BAR(U);
        b_av += dt * BAR(VPRED) + dtb2 * BAR(V);
        b_aw += dt * BAR(WPRED) + dtb2 * BAR(W);
        dtb += au * BAR(UPRED) + av * BAR(VPRED) + aw * BAR(WPRED) +
               0.5 * (au * BAR(U) + av * BAR(V) + aw * BAR(W));
        BAR(UPRED) = 0.0;
        BAR(VPRED) = 0.0;
        BAR(WPRED) = 0.0;
        BAR(U) = 0.0;
        BAR(V) = 0.0;
        BAR(W) = 0.0;
        // the rates, cs and alpha_loc the step started from
        BAR(AU) += b_au;
        BAR(AV) += b_av;
        BAR(AW) += b_aw;
        BAR(AE) += b_ae;
        BAR(AH) += b_ah;
        BAR(ABEVOLX) += b_abx;
        BAR(ABEVOLY) += b_aby;
        BAR(ABEVOLZ) += b_abz;
        BAR(APSI) += b_apsi;
        BAR(CS) += b_cs;
        BAR(ALPHALOC) += b_al;
    }
    // initialize (integrator.py:12-36): copies of the state
    BAR(PSI) += BAR(PSIPRED) + BAR(PSI0);
    BAR(PSIPRED) = 0.0;
    BAR(PSI0) = 0.0;
    BAR(BEVOLX) += BAR(BEVOLXPRED) + BAR(BEVOLX0);
    BAR(BEVOLY) += BAR(BEVOLYPRED) + BAR(BEVOLY0);
    BAR(BEVOLZ) += BAR(BEVOLZPRED) + BAR(BEVOLZ0);
    BAR(BEVOLXPRED) = 0.0;
    BAR(BEVOLYPRED) = 0.0;
    BAR(BEVOLZPRED) = 0.0;
    BAR(BEVOLX0) = 0.0;
    BAR(BEVOLY0) = 0.0;
    BAR(BEVOLZ0) = 0.0;
    BAR(ALPHA1) += BAR(ALPHA10);
    BAR(ALPHA10) = 0.0;
    BAR(OMEGA) = 0.0;
    BAR(CONVERGED) = 0.0;
    BAR(H) += BAR(H0);
    BAR(H0) = 0.0;
    BAR(E) += BAR(EPRED) + BAR(E0);
    BAR(EPRED) = 0.0;
    BAR(E0) = 0.0;
    BAR(U) += BAR(UPRED) + BAR(U0);
    BAR(V) += BAR(VPRED) + BAR(V0);
    BAR(W) += BAR(WPRED) + BAR(W0);
    BAR(UPRED) = 0.0;
    BAR(VPRED) = 0.0;
    BAR(WPRED) = 0.0;
    BAR(U0) = 0.0;
    BAR(V0) = 0.0;
    BAR(W0) = 0.0;
    BAR(X) += BAR(X0);
    BAR(Y) += BAR(Y0);
    BAR(X0) = 0.0;
    BAR(Y0) = 0.0;
    if (A.b[SPHMHD_F_Z0]) {
        if (A.b[SPHMHD_F_Z]) BAR(Z) += BAR(Z0);
        BAR(Z0) = 0.0;
    }
    A.dtrow[i] += dtb;
}

// ---------------------------------------------------------------------------
// the warm start (application.py:100-131), reversed
// ---------------------------------------------------------------------------

// h0 = h, then _sync_initial_mhd_state: Bevol = B / rho, psi = 0, predicted =
// current (rho and B: the values the forward sync read)
__global__ void k_a_sync(Fields F, AdjDev A, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BAR(E) += BAR(EPRED);
    BAR(EPRED) = 0.0;
    BAR(U) += BAR(UPRED);
    BAR(V) += BAR(VPRED);
    BAR(W) += BAR(WPRED);
    BAR(UPRED) = 0.0;
    BAR(VPRED) = 0.0;
    BAR(WPRED) = 0.0;
    BAR(PSI) = 0.0;
    BAR(PSIPRED) = 0.0;
    const double rho = FLD(RHO);
    const double gx = BAR(BEVOLX) + BAR(BEVOLXPRED), gy = BAR(BEVOLY) + BAR(BEVOLYPRED),
                 gz = BAR(BEVOLZ) + BAR(BEVOLZPRED);
    BAR(BX) += gx / rho;
    BAR(BY) += gy / rho;
    BAR(BZ) += gz / rho;
    BAR(RHO) += -(gx * FLD(BX) + gy * FLD(BY) + gz * FLD(BZ)) / (rho * rho);
    BAR(BEVOLX) = 0.0;
    BAR(BEVOLY) = 0.0;
    BAR(BEVOLZ) = 0.0;
    BAR(BEVOLXPRED) = 0.0;
    BAR(BEVOLYPRED) = 0.0;
    BAR(BEVOLZPRED) = 0.0;
    BAR(H) += BAR(H0);
    BAR(H0) = 0.0;
}

// the field restored after the first evaluation: the bar of the B it wrote
// goes to the saved initial B (take = 1); then (take = 0) back onto B
__global__ void k_a_bsave(AdjDev A, int64_t n, int take) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int f[3] = {SPHMHD_F_BX, SPHMHD_F_BY, SPHMHD_F_BZ};
    for (int c = 0; c < 3; ++c) {
        if (take) {
            A.bsave[(int64_t)c * n + i] = A.b[f[c]][i];
            A.b[f[c]][i] = 0.0;
        } else {
            A.b[f[c]][i] += A.bsave[(int64_t)c * n + i];
        }
    }
}

__global__ void k_a_zero_rows(double* __restrict__ p, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned ablk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// Elementwise reverse kernels run 64-thread workgroups: a rollout the
// adjoint differentiates is small (C5: 16,296 particles), and 256-thread
// workgroups left three quarters of the CUs idle while each thread walked a
// long chain of dependent loads.
constexpr int kAdjElem = 64;

#define A_DISPATCH(KID, DIMV, CALL)                                                       \
    switch (KID) {                                                                        \
    case SPHMHD_CUBIC: if (DIMV == 2) CALL(SPHMHD_CUBIC, 2) else CALL(SPHMHD_CUBIC, 3) break; \
    case SPHMHD_QUINTIC: if (DIMV == 2) CALL(SPHMHD_QUINTIC, 2) else CALL(SPHMHD_QUINTIC, 3) break; \
    case SPHMHD_WENDLAND_C2: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C2, 2) else CALL(SPHMHD_WENDLAND_C2, 3) break; \
    case SPHMHD_WENDLAND_C4: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C4, 2) else CALL(SPHMHD_WENDLAND_C4, 3) break; \
    case SPHMHD_WENDLAND_C6: if (DIMV == 2) CALL(SPHMHD_WENDLAND_C6, 2) else CALL(SPHMHD_WENDLAND_C6, 3) break; \
    case SPHMHD_GAUSSIAN: if (DIMV == 2) CALL(SPHMHD_GAUSSIAN, 2) else CALL(SPHMHD_GAUSSIAN, 3) break; \
    default: if (DIMV == 2) CALL(SPHMHD_SUPER_GAUSSIAN, 2) else CALL(SPHMHD_SUPER_GAUSSIAN, 3) break; \
    }

void launch_a_rows(Dev& d, const RowList& L, double* buf, bool restore) {
    if (d.n && L.count)
        hipLaunchKernelGGL(k_a_rows, dim3(ablk(d.n, kAdjElem), L.count), dim3(kAdjElem), 0, d.stream, d.fields, L, buf,
                           d.n, restore ? 1 : 0);
}

void launch_a_clock(Dev& d, double* clk, bool restore) {
    hipLaunchKernelGGL(k_a_clock, dim3(ablk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, clk, d.ninst,
                       restore ? 1 : 0);
}

void launch_a_save_entry(Dev& d, AdjDev& a) {
    if (d.n) hipLaunchKernelGGL(k_a_save_entry, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, a.ain, d.n);
}

void launch_a_seed_totals(Dev& d, AdjDev& a, const Weights& w) {
    if (d.n)
        hipLaunchKernelGGL(k_a_seed_totals, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, a, d.inst,
                           d.inst_of, d.n, w);
}

void launch_a_zero(Dev& d, double* p, int64_t count) {
    if (count > 0) hipLaunchKernelGGL(k_a_zero_rows, dim3(ablk(count, kAdjElem)), dim3(kAdjElem), 0, d.stream, p, count);
}

// the pending dt's bar for the evaluation just recomputed: clk_next is the
// device clock {t, dt} per instance of the step after it
void launch_a_eval_begin(Dev& d, AdjDev& a, const double* clk_next, int mode) {
    hipLaunchKernelGGL(k_a_begin_eval, dim3(ablk(d.ninst)), dim3(kBlock), 0, d.stream, a.ai, d.ninst);
    if (d.n)
        hipLaunchKernelGGL(k_a_argmin, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, d.inst, a.ai, d.inst_of,
                           d.n);
    hipLaunchKernelGGL(k_a_pend, dim3(ablk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, a.ai, clk_next, d.ninst,
                       mode);
}

void launch_a_stage2(Dev& d, AdjDev& a) {
    if (d.n)
        hipLaunchKernelGGL(k_a_stage2, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, a, d.inst, d.inst_of,
                           d.n);
}

void launch_a_eval(Dev& d, AdjDev& a) {
    if (d.n == 0) return;
    const unsigned g = ablk(d.n * kAdjLanes, kAdjBlock);
#define CALL_A(KK, DD)                                                                                            \
    {                                                                                                             \
        hipLaunchKernelGGL((k_a_force_epi<KK, DD>), dim3(g), dim3(kAdjBlock), 0, d.stream, d.fields, a, d.fpack,  \
                           d.perm, d.nb, d.nbcnt, d.inst, d.inst_of, d.n, d.nb_stride);                           \
        hipLaunchKernelGGL((k_a_force_pair<KK, DD>), dim3(g), dim3(kAdjBlock), 0, d.stream, d.fields, a, d.fpack, \
                           d.perm, d.nb, d.nbcnt, d.inst, d.inst_of, d.n, d.nb_stride);                           \
        hipLaunchKernelGGL(k_a_prep, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, a, d.inst, d.inst_of,  \
                           d.n);                                                                                  \
        launch_pack_density(d);                                                                                   \
        hipLaunchKernelGGL((k_a_density<KK, DD>), dim3(g), dim3(kAdjBlock), 0, d.stream, d.fields, a,             \
                           (const double4*)d.dpack, d.perm, d.keys, d.cstart, d.cend, d.inst, d.inst_of, d.n);    \
        hipLaunchKernelGGL((k_a_density_pos<KK, DD>), dim3(g), dim3(kAdjBlock), 0, d.stream, d.fields, a,         \
                           (const double4*)d.dpack, d.perm, d.keys, d.cstart, d.cend, d.inst, d.inst_of, d.n);    \
    }
    A_DISPATCH(d.kernel, d.dim, CALL_A)
#undef CALL_A
}

void launch_a_stage1_init(Dev& d, AdjDev& a) {
    if (d.n)
        hipLaunchKernelGGL(k_a_stage1_init, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, a, d.inst,
                           d.inst_of, d.n);
}

void launch_a_reduce_dt(Dev& d, AdjDev& a) {
    hipLaunchKernelGGL(k_a_reduce_dt, dim3(d.ninst), dim3(kBlock), 0, d.stream, d.inst, a.ai, a.dtrow);
}

void launch_a_reduce_par(Dev& d, AdjDev& a) {
    hipLaunchKernelGGL(k_a_reduce_par, dim3(d.ninst), dim3(kBlock), 0, d.stream, d.inst, a.ai, a.par, d.n);
}

void launch_a_sync(Dev& d, AdjDev& a) {
    if (d.n) hipLaunchKernelGGL(k_a_sync, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, d.fields, a, d.n);
}

void launch_a_bsave(Dev& d, AdjDev& a, bool take) {
    if (d.n) hipLaunchKernelGGL(k_a_bsave, dim3(ablk(d.n, kAdjElem)), dim3(kAdjElem), 0, d.stream, a, d.n, take ? 1 : 0);
}

}  // namespace sph
