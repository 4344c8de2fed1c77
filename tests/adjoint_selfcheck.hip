// adjoint_selfcheck.hip -- host-only check of the hand-written reverse pair
// and epilogue functions of sph_adjoint.h against forward-mode duals.
//
// For random inputs x, a random direction v of the inputs and random output
// bars ybar, the reverse functions give g = d<ybar, f(x)>/dx and a dual
// evaluation of f along v gives <ybar, f'(x) v>; the two must agree to
// rounding: <g, v> == <ybar, f'(x) v>.  Inputs are drawn so that every
// branch of the bodies is taken in some trials (viscosity on and off, the
// pair being the maximum of dt_cfl or not, the 3-D velocity-gradient
// inversion or its fallback, the energy limiter, the divB correction's
// clamp, the Courant or force limit of dt).
//
// Prints one line per check: "<name> <trials> <max relative error>".
// Built and run by tests/test_adjoint_cpu.py (hipcc, run on the host).
#include <cmath>
#include <cstdio>
#include <random>

#include "../pysph-mhd_amd/csrc/sph_adjoint.h"

using namespace sph;
using namespace sph::adj;
using D = Dual<1>;

static std::mt19937_64 rng(12345);
static double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

template <class T>
static Par<T> make_par(const Par<double>& p, const double* dir) {
    Par<T> q;
    q.gamma = p.gamma;
    q.gamma1 = p.gamma1;
    q.mu0 = p.mu0;
    q.inv_mu0 = p.inv_mu0;
    q.alpha_min = p.alpha_min;
    q.c_cour = p.c_cour;
    q.c_force = p.c_force;
    q.cleaning_active = p.cleaning_active;
    q.hfact = p.hfact;
    q.ksig = p.ksig;
    const double* src = &p.av;
    T* dst = &q.av;
    for (int k = 0; k < 6; ++k) {
        dst[k] = T(src[k]);
        if constexpr (!std::is_same<T, double>::value) dst[k].d[0] = dir ? dir[k] : 0.0;
    }
    return q;
}

template <class T>
static Raw<T> make_raw(const Raw<double>& r, const double* dir) {
    Raw<T> q;
    const double* src = &r.x;
    T* dst = &q.x;
    for (int k = 0; k < 19; ++k) {
        dst[k] = T(src[k]);
        if constexpr (!std::is_same<T, double>::value) dst[k].d[0] = dir ? dir[k] : 0.0;
    }
    q.m = r.m;
    return q;
}

static Par<double> rand_par(int kernel, int dim) {
    Par<double> p;
    p.gamma = uni(1.3, 1.8);
    p.gamma1 = p.gamma - 1.0;
    p.mu0 = uni(0.5, 2.0);
    p.inv_mu0 = 1.0 / p.mu0;
    p.alpha_min = uni(0.0, 0.3);
    p.c_cour = uni(0.1, 0.4);
    p.c_force = uni(0.1, 0.3);
    p.cleaning_active = 1.0;
    p.hfact = 1.2;
    p.ksig = kernel_sigma(kernel, dim);
    p.av = uni(0.5, 1.5);
    p.ac = uni(0.0, 1.0);
    p.ar = uni(0.0, 1.0);
    p.chf = uni(0.5, 1.5);
    p.sigma = uni(0.2, 1.5);
    p.alpha_max = uni(0.5, 1.5);
    return p;
}

static Raw<double> rand_raw(double h, bool visc_bias) {
    Raw<double> r;
    r.x = uni(0, 1);
    r.y = uni(0, 1);
    r.z = uni(0, 1);
    r.h = h * uni(0.8, 1.2);
    r.rho = uni(0.5, 2.0);
    r.e = uni(0.5, 2.0);
    const double vs = visc_bias ? 2.0 : 0.3;
    r.u = uni(-vs, vs);
    r.v = uni(-vs, vs);
    r.w = uni(-vs, vs);
    r.om = uni(0.8, 1.2);
    r.bx = uni(-1, 1);
    r.by = uni(-1, 1);
    r.bz = uni(-1, 1);
    r.al = uni(0.0, 1.0);
    r.au = uni(-1, 1);
    r.av = uni(-1, 1);
    r.aw = uni(-1, 1);
    r.psi = uni(-0.5, 0.5);
    r.cs = uni(0.5, 1.5);
    r.m = uni(0.5, 1.5) * 1e-3;
    return r;
}

static double rel(double a, double b, double scale) { return fabs(a - b) / (fabs(a) + fabs(b) + scale); }

template <int K, int DIM>
static double check_pair(int trials, const char* name) {
    double worst = 0.0;
    const double R = kernel_radius(K);
    for (int t = 0; t < trials; ++t) {
        const Par<double> P = rand_par(K, DIM);
        const double h = 0.05;
        const Raw<double> ri = rand_raw(h, t % 2 == 0), rj = rand_raw(h, t % 2 == 0);
        double d[3];
        const double rmax = R * fmax(ri.h, rj.h);
        do {
            d[0] = uni(-rmax, rmax);
            d[1] = uni(-rmax, rmax);
            d[2] = DIM == 3 ? uni(-rmax, rmax) : 0.0;
        } while (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] >= rmax * rmax);
        Acc<double> B;
        double* bp = acc_ptr(B);
        for (int k = 0; k < kNAcc; ++k) bp[k] = acc_used(k, DIM) ? uni(-1, 1) : 0.0;
        const bool argmax = t % 3 != 0;
        // reverse
        const Side<double> si = side_of<DIM>(ri, P), sj = side_of<DIM>(rj, P);
        SideBar ib{}, jb{};
        double db[3] = {0, 0, 0};
        PairParBar pb{0, 0, 0};
        pair_rev<K, DIM>(si, sj, d[0], d[1], d[2], P, B, argmax, ib, jb, db, pb);
        RawBar rbi{}, rbj{};
        side_rev(ri, si, P, ib, rbi);
        side_rev(rj, sj, P, jb, rbj);
        const double chf_bar = ib.psich * ri.psi * si.vw + jb.psich * rj.psi * sj.vw;
        // forward along a random direction
        double vi[19], vj[19], vd[3], vp[6];
        for (double& x : vi) x = uni(-1, 1);
        for (double& x : vj) x = uni(-1, 1);
        for (double& x : vd) x = uni(-1, 1) * h;
        for (double& x : vp) x = uni(-1, 1);
        if (DIM == 2) vi[2] = vj[2] = vi[16] = vj[16] = vd[2] = 0.0;  // z and aw do not enter in 2-D
        vp[4] = vp[5] = 0.0;  // sigma, alpha_max do not enter the pair
        const Raw<D> Ri = make_raw<D>(ri, vi), Rj = make_raw<D>(rj, vj);
        const Par<D> PD = make_par<D>(P, vp);
        const Side<D> SI = side_of<DIM>(Ri, PD), SJ = side_of<DIM>(Rj, PD);
        D dx(d[0]), dy(d[1]), dz(d[2]);
        dx.d[0] = vd[0];
        dy.d[0] = vd[1];
        dz.d[0] = vd[2];
        Acc<D> A;
        acc_zero(A);
        const D cand = pair_fwd<K, DIM>(SI, SJ, dx, dy, dz, PD, A);
        double fwd = 0.0, scale = 0.0;
        const D* ap = acc_ptr(A);
        for (int k = 0; k < kNAcc; ++k) {
            if (k == 5) continue;  // dt_cfl: through the candidate
            fwd += bp[k] * ap[k].d[0];
            scale += fabs(bp[k] * ap[k].d[0]);
        }
        if (argmax && cand.v >= 0.0) fwd += B.dtcfl * cand.d[0];
        // reverse contraction
        const double* gi = &rbi.x;
        const double* gj = &rbj.x;
        double rev = 0.0;
        for (int k = 0; k < 19; ++k) {
            if (k < 3) continue;  // positions enter through d
            rev += gi[k] * vi[k] + gj[k] * vj[k];
        }
        rev += db[0] * vd[0] + db[1] * vd[1] + db[2] * vd[2];
        rev += pb.av * vp[0] + pb.ac * vp[1] + pb.ar * vp[2] + chf_bar * vp[3];
        const double e = rel(fwd, rev, 1e-12 * scale + 1e-300);
        if (e > worst) worst = e;
    }
    printf("%s %d %.3e\n", name, trials, worst);
    return worst;
}

template <int DIM>
static double check_epi(int trials, const char* name) {
    double worst = 0.0;
    int branches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < trials; ++t) {
        Par<double> P = rand_par(SPHMHD_QUINTIC, DIM);
        if (t % 2 == 0) P.alpha_min = 0.0;  // alpha_loc above its floor
        Raw<double> me = rand_raw(t % 2 == 0 ? 0.5 : 0.05, false);
        if (t % 6 == 3) {  // beta = 2 mu0 p / B^2 inside (2, 10): the divB correction's clamp is open
            me.bx *= 0.5;
            me.by *= 0.5;
            me.bz *= 0.5;
        }
        Acc<double> A;
        double* ap = acc_ptr(A);
        for (int k = 0; k < kNAcc; ++k) ap[k] = acc_used(k, DIM) ? uni(-1, 1) : 0.0;
        if (DIM == 3) {
            // a positive definite -R (the CD sums are -sum x_ij (m grad W)^T) or a singular one
            A.rxx = -uni(1, 2);
            A.ryy = -uni(1, 2);
            A.rzz = -uni(1, 2);
            A.rxy = uni(-0.3, 0.3);
            A.rxz = uni(-0.3, 0.3);
            A.ryz = uni(-0.3, 0.3);
            if (t % 5 == 0) A.rxx = A.rxy = A.rxz = 0.0;
        }
        A.dtcfl = uni(0.0, 3.0);
        if (t % 4 == 1) {  // energy limiter
            me.e = 0.01;
            A.ae = -uni(5.0, 50.0);
        }
        if (t % 3 == 2) me.cs = uni(3.0, 6.0);  // large cs: Courant limit
        // output bars
        Epi<double> O;
        double* op = epi_ptr(O);
        for (int k = 0; k < kNEpi; ++k) op[k] = uni(-1, 1);
        const Side<double> S = side_of<DIM>(me, P);
        Acc<double> Ab;
        acc_zero(Ab);
        RawBar rb{};
        double b_vw = 0.0, b_p = 0.0;
        EpiParBar pb{0, 0, 0};
        epi_rev<DIM>(A, me, S, P, O, Ab, rb, &b_vw, &b_p, pb);
        SideBar sb{};
        sb.vw = b_vw;
        sb.p = b_p;
        side_rev(me, S, P, sb, rb);
        // dual forward
        double va[kNAcc], vm[19], vp[6];
        for (int k = 0; k < kNAcc; ++k) va[k] = acc_used(k, DIM) ? uni(-1, 1) : 0.0;
        for (double& x : vm) x = uni(-1, 1);
        for (double& x : vp) x = uni(-1, 1);
        vp[0] = vp[1] = vp[2] = 0.0;  // av, ac, ar do not enter the epilogue
        // the raw inputs the epilogue reads: h, rho, e, om, bx, by, bz, psi, cs
        for (int k : {0, 1, 2, 6, 7, 8, 13, 14, 15, 16}) vm[k] = 0.0;
        Acc<D> AD;
        D* adp = acc_ptr(AD);
        for (int k = 0; k < kNAcc; ++k) {
            adp[k] = D(ap[k]);
            adp[k].d[0] = va[k];
        }
        const Raw<D> MD = make_raw<D>(me, vm);
        const Par<D> PD = make_par<D>(P, vp);
        const Side<D> SD = side_of<DIM>(MD, PD);
        const Epi<D> E = epi_fwd<DIM>(AD, MD, SD, PD);
        const D* ep = epi_ptr(E);
        double fwd = 0.0, scale = 0.0;
        for (int k = 0; k < kNEpi; ++k) {
            fwd += op[k] * ep[k].d[0];
            scale += fabs(op[k] * ep[k].d[0]);
        }
        double rev = 0.0;
        const double* abp = acc_ptr(Ab);
        for (int k = 0; k < kNAcc; ++k) rev += abp[k] * va[k];
        const double* g = &rb.x;
        for (int k = 0; k < 19; ++k) rev += g[k] * vm[k];
        rev += pb.chf * vp[3] + pb.sigma * vp[4] + pb.alpha_max * vp[5];
        const double e = rel(fwd, rev, 1e-12 * scale + 1e-300);
        if (e > worst) worst = e;
        // branch coverage
        branches[0] += E.dt_adapt.v < 1e29;
        branches[1] += (me.e == 0.01);
        branches[2] += E.alphaloc.v > P.alpha_min;
        const double b2 = me.bx * me.bx + me.by * me.by + me.bz * me.bz;
        const double beta = 2.0 * P.mu0 * S.p / b2;
        branches[3] += beta > 2.0 && beta < 10.0;
    }
    printf("%s %d %.3e\n", name, trials, worst);
    printf("%s_branches dt %d limiter %d alpha %d divb_clamp_open %d\n", name, branches[0], branches[1],
           branches[2], branches[3]);
    return worst;
}

int main() {
    check_pair<SPHMHD_CUBIC, 2>(400, "pair_cubic_2d");
    check_pair<SPHMHD_WENDLAND_C6, 2>(400, "pair_wendland_c6_2d");
    check_pair<SPHMHD_QUINTIC, 3>(400, "pair_quintic_3d");
    check_pair<SPHMHD_GAUSSIAN, 3>(200, "pair_gaussian_3d");
    check_epi<2>(600, "epi_2d");
    check_epi<3>(600, "epi_3d");
    return 0;
}
