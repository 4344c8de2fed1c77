// sph_adjoint.h -- per-pair and per-particle pieces of the adjoint (reverse-
// mode) step, host and device code.
//
// The step's derivative is taken of the discrete map the forward step
// computes: neighbour sets and cells fixed, the density Newton iteration as
// run (pass by pass), every branch, clamp and min/max as taken (ties of
// max/min: the first operand, as the duals of sph_dual.h), the adaptive dt
// through the particle attaining the minimum.  The force pass of one
// destination i sums a contribution C(i <- j) of every neighbour j into 35
// accumulators (equations.py:143-567 fused: MHDAccelerations,
// MagneticStressAccelerations, MagneticStateRates, CullenDehnenAlpha
// Diagnostics loops), then a per-particle epilogue (their post_loops and
// G6/G9: equations.py:137-140, 356-429, 559-624) turns the sums into the
// rates.  Here:
//
//   side_of   the per-particle values a pair reads (from its raw fields);
//   pair_fwd  C(i <- j), generic in the scalar type (double for the
//             recomputation, Dual for the self-check);
//   pair_rev  the gradient of <Abar_i, C(i <- j)> with respect to both
//             sides and the separation, by hand (reverse accumulation);
//   epi_fwd / epi_rev  the epilogue and its reverse;
//   newton_post  DensityIteration.post_loop (equations.py:54-86) generic in
//             the scalar type: its derivative comes from Dual<3>.
//
// The forward arithmetic restates the force kernel's expressions (the
// tangent step's k_t_force, which restates sph_pairs.hip's k_force), so the
// recomputation takes the forward step's branches.
#pragma once

#include <hip/hip_runtime.h>

// sph_internal.h first: its generic kernel shapes must see the double
// overloads of fmax/exp, the duals' are found through their namespace
#include "sph_internal.h"
#include "sph_dual.h"

#define ADJ_HD __host__ __device__ __forceinline__

namespace sph {
namespace adj {

// the double overloads beside the duals' (found through their namespace)
using ::cbrt;
using ::fabs;
using ::fmax;
using ::fmin;
using ::sqrt;

constexpr double kFloatTiny = 2.2250738585072014e-308;

// differentiated parameters (sphmhd_tparam order) and the instance's constants
template <class T>
struct Par {
    T av, ac, ar, chf, sigma, alpha_max;
    double gamma, gamma1, mu0, inv_mu0, alpha_min, c_cour, c_force, cleaning_active, hfact, ksig;
};

// a particle's raw inputs of the force pass (predicted v, e, B; previous a)
template <class T>
struct Raw {
    T x, y, z, h, rho, e, u, v, w, om, bx, by, bz, al, au, av, aw, psi, cs;
    double m;
};
struct RawBar {
    double x, y, z, h, rho, e, u, v, w, om, bx, by, bz, al, au, av, aw, psi, cs;
};

// what the pair reads of a side
template <class T>
struct Side {
    T h1, rho, hir, p, e, u, v, w, om, bx, by, bz, al, au, av, aw, psich, vw;
    double m;
};
struct SideBar {
    double h1, rho, hir, p, e, u, v, w, om, bx, by, bz, al, au, av, aw, psich, vw;
};

// the force pass's accumulators (the epilogue's inputs)
template <class T>
struct Acc {
    T ax, ay, az, ae, div, dtcfl, bx, by, bz, divbs, divbd;
    T rxx, rxy, rxz, ryy, ryz, rzz;
    T vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz;
    T axx, axy, axz, ayx, ayy, ayz, azx, azy, azz;
};
constexpr int kNAcc = kAdjAcc;
template <class T>
ADJ_HD T* acc_ptr(Acc<T>& a) { return &a.ax; }
template <class T>
ADJ_HD const T* acc_ptr(const Acc<T>& a) { return &a.ax; }
// the accumulators a DIM-dimensional pass writes (indices into Acc as an array)
ADJ_HD bool acc_used(int k, int dim) {
    if (dim == 3) return true;
    // 2-D: ax..divbd (0-10), vxx vxy vyx vyy vzx vzy, axx ayy
    return k <= 10 || k == 17 || k == 18 || k == 20 || k == 21 || k == 23 || k == 24 || k == 26 || k == 30;
}

template <class T>
ADJ_HD T zero() { return T(0.0); }

template <class T>
ADJ_HD void acc_zero(Acc<T>& a) {
    T* p = acc_ptr(a);
    for (int k = 0; k < kNAcc; ++k) p[k] = T(0.0);
}

template <int DIM, class T>
ADJ_HD Side<T> side_of(const Raw<T>& r, const Par<T>& P) {
    Side<T> s;
    s.h1 = 1.0 / r.h;
    s.rho = r.rho;
    s.e = r.e;
    s.p = P.gamma1 * r.rho * r.e;
    s.u = r.u;
    s.v = r.v;
    s.w = r.w;
    s.om = r.om;
    s.bx = r.bx;
    s.by = r.by;
    s.bz = r.bz;
    s.al = r.al;
    s.au = r.au;
    s.av = r.av;
    s.aw = r.aw;
    const T b2 = r.bx * r.bx + r.by * r.by + r.bz * r.bz;
    s.vw = sqrt(r.cs * r.cs + b2 / (P.mu0 * r.rho));
    s.psich = r.psi * (P.chf * s.vw);
    s.hir = 0.5 * (1.0 / r.rho);
    s.m = r.m;
    return s;
}

// side bars -> raw bars (psich's share of the cleaning-speed parameter is
// left to the caller: chf_bar += sb.psich * psi * vw where it counts)
ADJ_HD void side_rev(const Raw<double>& r, const Side<double>& s, const Par<double>& P, const SideBar& sb,
                     RawBar& rb) {
    rb.h += -sb.h1 * s.h1 * s.h1;
    rb.rho += sb.rho;
    rb.e += sb.e;
    rb.u += sb.u;
    rb.v += sb.v;
    rb.w += sb.w;
    rb.om += sb.om;
    rb.bx += sb.bx;
    rb.by += sb.by;
    rb.bz += sb.bz;
    rb.al += sb.al;
    rb.au += sb.au;
    rb.av += sb.av;
    rb.aw += sb.aw;
    // psich = psi * (chf * vw)
    rb.psi += sb.psich * (P.chf * s.vw);
    const double bvw = sb.vw + sb.psich * r.psi * P.chf;
    // vw = sqrt(cs^2 + b2 / (mu0 rho))
    if (s.vw > 0.0) {
        const double u = bvw * 0.5 / s.vw;
        const double b2 = r.bx * r.bx + r.by * r.by + r.bz * r.bz;
        const double mr = P.mu0 * r.rho;
        rb.cs += u * 2.0 * r.cs;
        const double bb2 = u / mr;
        rb.rho += -u * b2 / (mr * r.rho);
        rb.bx += 2.0 * r.bx * bb2;
        rb.by += 2.0 * r.by * bb2;
        rb.bz += 2.0 * r.bz * bb2;
    }
    // p = gamma1 rho e
    rb.rho += sb.p * P.gamma1 * r.e;
    rb.e += sb.p * P.gamma1 * r.rho;
    // hir = 0.5 / rho
    rb.rho += -sb.hir * 0.5 / (r.rho * r.rho);
}

// W'(q) and W''(q) of the kernel shape (w scaled by sigma/h^d elsewhere): the
// shape's own arithmetic over a dual
template <int K, int DIM>
ADJ_HD void shape_d12(double q, double& dw, double& d2w) {
    Dual<1> qd(q);
    qd.d[0] = 1.0;
    Dual<1> w, g;
    shape<K, DIM>(qd, w, g);
    dw = g.v;
    d2w = g.d[0];
}

// C(i <- j) into A; returns the pair's dt_cfl candidate (after the clamp at 0)
template <int K, int DIM, class T>
ADJ_HD T pair_fwd(const Side<T>& I, const Side<T>& J, const T& dx, const T& dy, const T& dz, const Par<T>& P,
                  Acc<T>& A) {
    const double inv_mu0 = P.inv_mu0;
    const double sig = P.ksig;
    const T irhoi = 2.0 * I.hir, irhoj = 2.0 * J.hir;
    const T r2ii = irhoi * irhoi, r2ij = irhoj * irhoj;
    const T pbi = I.p * r2ii, pbj = J.p * r2ij;
    const T brxi = I.bx * irhoi, bryi = I.by * irhoi, brzi = I.bz * irhoi;
    const T brxj = J.bx * irhoj, bryj = J.by * irhoj, brzj = J.bz * irhoj;
    const T bro2i = (brxi * brxi + bryi * bryi + brzi * brzi) * inv_mu0;
    const T bro2j = (brxj * brxj + bryj * bryj + brzj * brzj) * inv_mu0;
    const T fac_i = (DIM == 2 ? sig * I.h1 * I.h1 : sig * I.h1 * I.h1 * I.h1) * I.h1;
    const T mj(J.m);
    const T r = sqrt(dx * dx + dy * dy + dz * dz);
    T gi(0.0), gj(0.0), ex(0.0), ey(0.0), ez(0.0);
    if (r > 1.0e-12) {
        const T ir = 1.0 / r;
        gi = shape_dw<K, DIM>(r * I.h1) * fac_i * ir;
        const T facj = (DIM == 2 ? sig * J.h1 * J.h1 : sig * J.h1 * J.h1 * J.h1) * J.h1;
        gj = shape_dw<K, DIM>(r * J.h1) * facj * ir;
        ex = dx * ir;
        ey = dy * ir;
        ez = dz * ir;
    }
    const T dwix = gi * dx, dwiy = gi * dy, dwiz = gi * dz;
    const T dwjx = gj * dx, dwjy = gj * dy, dwjz = gj * dz;
    const T dvx = I.u - J.u, dvy = I.v - J.v, dvz = I.w - J.w;
    const T gri = I.om * (ex * dwix + ey * dwiy + ez * dwiz);
    const T grj = J.om * (ex * dwjx + ey * dwjy + ez * dwjz);
    const T dot = dvx * ex + dvy * ey + dvz * ez;
    T cand;
    // MHDAccelerations (equations.py:217-275)
    {
        const bool near8 = r < 1.0e-8;
        const T dotA = near8 ? T(0.0) : dot;
        const T griA = near8 ? T(0.0) : gri;
        const T grjA = near8 ? T(0.0) : grj;
        A.div += mj * dotA * griA;
        const T vsig2 = sqrt(fabs(I.p - J.p) / (0.5 * (I.rho + J.rho)));
        T c = fmax(I.vw - 2.0 * dotA, J.vw - 2.0 * dotA);
        c = fmax(c, 0.0);
        cand = c;
        A.dtcfl = fmax(A.dtcfl, c);
        T sx = pbi * I.om, sjx = pbj * J.om;
        if (dotA < 0.0) {
            const T vsi = fmax(I.al * I.vw - 2.0 * dotA, 0.0);
            const T vsj = fmax(J.al * J.vw - 2.0 * dotA, 0.0);
            const T qi = -I.hir * vsi * dotA * P.av;
            const T qj = -J.hir * vsj * dotA * P.av;
            sx += qi * I.om;
            sjx += qj * J.om;
            A.ae += mj * qi * dotA * griA;
        }
        A.ax -= mj * (sx * dwix + sjx * dwjx);
        A.ay -= mj * (sx * dwiy + sjx * dwjy);
        A.az -= mj * (sx * dwiz + sjx * dwjz);
        A.ae += mj * pbi * I.om * (dvx * dwix + dvy * dwiy + dvz * dwiz);
        A.ae += P.ac * vsig2 * (I.e - J.e) * mj * (I.hir * griA + J.hir * grjA);
    }
    // MagneticStressAccelerations (equations.py:149-195)
    {
        const T gradp = mj * (0.5 * bro2i * gri + 0.5 * bro2j * grj);
        const T pi_b = (brxi * ex + bryi * ey + brzi * ez) * gri;
        const T pj_b = (brxj * ex + bryj * ey + brzj * ez) * grj;
        const T sm = mj * inv_mu0;
        A.ax += -ex * gradp + sm * (brxi * pi_b + brxj * pj_b);
        A.ay += -ey * gradp + sm * (bryi * pi_b + bryj * pj_b);
        A.az += -ez * gradp + sm * (brzi * pi_b + brzj * pj_b);
    }
    // MagneticStateRates (equations.py:507-559)
    {
        const T pbi_ = I.bx * ex + I.by * ey + I.bz * ez;
        const T pbj_ = J.bx * ex + J.by * ey + J.bz * ez;
        const T dbx = I.bx - J.bx, dby = I.by - J.by, dbz = I.bz - J.bz;
        const T pdb = dbx * ex + dby * ey + dbz * ez;
        const T db2 = dbx * dbx + dby * dby + dbz * dbz;
        const T pmi = mj * r2ii * gri;
        const T pmj = mj * r2ij * grj;
        const T brt = -(pmi * pbi_);
        const T tx = dvx - dot * ex, ty = dvy - dot * ey, tz = dvz - dot * ez;
        const T vsb = sqrt(tx * tx + ty * ty + tz * tz);
        const T diss = 0.5 * (pmi + pmj) * vsb * P.ar;
        const T dpsi = pmi * I.psich + pmj * J.psich;
        A.bx += brt * dvx + diss * dbx - dpsi * ex;
        A.by += brt * dvy + diss * dby - dpsi * ey;
        A.bz += brt * dvz + diss * dbz - dpsi * ez;
        A.ae += -0.5 * db2 * diss * inv_mu0;
        A.divbs += pmi * pbi_ + pmj * pbj_;
        A.divbd += -mj * pdb * gri;
    }
    // CullenDehnenAlphaDiagnostics loop (equations.py:319-354)
    {
        const T rx = mj * dwix, ry = mj * dwiy, rz = mj * dwiz;
        const T dax = I.au - J.au, day = I.av - J.av;
        A.vxx += dvx * rx;
        A.vxy += dvx * ry;
        A.vyx += dvy * rx;
        A.vyy += dvy * ry;
        A.vzx += dvz * rx;
        A.vzy += dvz * ry;
        A.axx += dax * rx;
        A.ayy += day * ry;
        if constexpr (DIM == 3) {
            const T daz = I.aw - J.aw;
            A.rxx += -dx * rx;
            A.rxy += -dx * ry;
            A.rxz += -dx * rz;
            A.ryy += -dy * ry;
            A.ryz += -dy * rz;
            A.rzz += -dz * rz;
            A.vxz += dvx * rz;
            A.vyz += dvy * rz;
            A.vzz += dvz * rz;
            A.axy += dax * ry;
            A.axz += dax * rz;
            A.ayx += day * rx;
            A.ayz += day * rz;
            A.azx += daz * rx;
            A.azy += daz * ry;
            A.azz += daz * rz;
        }
    }
    return cand;
}

// parameter bars of one pair
struct PairParBar {
    double av, ac, ar;
};

// Reverse of pair_fwd: with B the bars of the destination's accumulators,
// adds d<B, C(i <- j)>/d(side i) to Ib, /d(side j) to Jb, /d(dx, dy, dz) to
// db[3] and the parameter bars to pb.  argmax: this pair's candidate is the
// destination's maximum dt_cfl (the bar B.dtcfl flows through it).
template <int K, int DIM>
ADJ_HD void pair_rev(const Side<double>& I, const Side<double>& J, double dx, double dy, double dz,
                     const Par<double>& P, const Acc<double>& B, bool argmax, SideBar& Ib, SideBar& Jb,
                     double* db, PairParBar& pb) {
    // ---- the forward values ----
    const double inv_mu0 = P.inv_mu0;
    const double sig = P.ksig;
    const double irhoi = 2.0 * I.hir, irhoj = 2.0 * J.hir;
    const double r2ii = irhoi * irhoi, r2ij = irhoj * irhoj;
    const double pbi = I.p * r2ii, pbj = J.p * r2ij;
    const double brxi = I.bx * irhoi, bryi = I.by * irhoi, brzi = I.bz * irhoi;
    const double brxj = J.bx * irhoj, bryj = J.by * irhoj, brzj = J.bz * irhoj;
    const double bro2i = (brxi * brxi + bryi * bryi + brzi * brzi) * inv_mu0;
    const double bro2j = (brxj * brxj + bryj * bryj + brzj * brzj) * inv_mu0;
    const double fac_i = (DIM == 2 ? sig * I.h1 * I.h1 : sig * I.h1 * I.h1 * I.h1) * I.h1;
    const double fac_j = (DIM == 2 ? sig * J.h1 * J.h1 : sig * J.h1 * J.h1 * J.h1) * J.h1;
    const double mj = J.m;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const bool far12 = r > 1.0e-12;
    double ir = 0.0, wpi = 0.0, wpj = 0.0, w2i = 0.0, w2j = 0.0;
    double gi = 0.0, gj = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
    if (far12) {
        ir = 1.0 / r;
        shape_d12<K, DIM>(r * I.h1, wpi, w2i);
        shape_d12<K, DIM>(r * J.h1, wpj, w2j);
        gi = wpi * fac_i * ir;
        gj = wpj * fac_j * ir;
        ex = dx * ir;
        ey = dy * ir;
        ez = dz * ir;
    }
    const double dwix = gi * dx, dwiy = gi * dy, dwiz = gi * dz;
    const double dwjx = gj * dx, dwjy = gj * dy, dwjz = gj * dz;
    const double dvx = I.u - J.u, dvy = I.v - J.v, dvz = I.w - J.w;
    const double Si = ex * dwix + ey * dwiy + ez * dwiz;
    const double Sj = ex * dwjx + ey * dwjy + ez * dwjz;
    const double gri = I.om * Si, grj = J.om * Sj;
    const double dot = dvx * ex + dvy * ey + dvz * ez;
    const bool near8 = r < 1.0e-8;
    const double dotA = near8 ? 0.0 : dot, griA = near8 ? 0.0 : gri, grjA = near8 ? 0.0 : grj;
    const double rhoij = 0.5 * (I.rho + J.rho);
    const double pdiff = I.p - J.p;
    const double Q = fabs(pdiff) / rhoij;
    const double vsig2 = sqrt(Q);
    const double ci = I.vw - 2.0 * dotA, cj = J.vw - 2.0 * dotA;
    const bool take_i = ci >= cj;
    const bool cpos = (take_i ? ci : cj) >= 0.0;
    const bool visc = dotA < 0.0;
    double vsi = 0.0, vsj = 0.0, qvi = 0.0, qvj = 0.0;
    bool vsi_on = false, vsj_on = false;
    double sx = pbi * I.om, sjx = pbj * J.om;
    if (visc) {
        const double ti = I.al * I.vw - 2.0 * dotA, tj = J.al * J.vw - 2.0 * dotA;
        vsi_on = ti >= 0.0;
        vsj_on = tj >= 0.0;
        vsi = vsi_on ? ti : 0.0;
        vsj = vsj_on ? tj : 0.0;
        qvi = -I.hir * vsi * dotA * P.av;
        qvj = -J.hir * vsj * dotA * P.av;
        sx += qvi * I.om;
        sjx += qvj * J.om;
    }
    const double gradp = mj * (0.5 * bro2i * gri + 0.5 * bro2j * grj);
    const double Pi = brxi * ex + bryi * ey + brzi * ez, Pj = brxj * ex + bryj * ey + brzj * ez;
    const double pi_b = Pi * gri, pj_b = Pj * grj;
    const double sm = mj * inv_mu0;
    const double pbi_ = I.bx * ex + I.by * ey + I.bz * ez;
    const double pbj_ = J.bx * ex + J.by * ey + J.bz * ez;
    const double dbx = I.bx - J.bx, dby = I.by - J.by, dbz = I.bz - J.bz;
    const double pdb = dbx * ex + dby * ey + dbz * ez;
    const double db2 = dbx * dbx + dby * dby + dbz * dbz;
    const double pmi = mj * r2ii * gri, pmj = mj * r2ij * grj;
    const double brt = -(pmi * pbi_);
    const double tx = dvx - dot * ex, ty = dvy - dot * ey, tz = dvz - dot * ez;
    const double vsb = sqrt(tx * tx + ty * ty + tz * tz);
    const double sdiss = 0.5 * (pmi + pmj);
    const double diss = sdiss * vsb * P.ar;
    const double dpsi = pmi * I.psich + pmj * J.psich;
    const double rx = mj * dwix, ry = mj * dwiy, rz = mj * dwiz;
    const double dax = I.au - J.au, day = I.av - J.av, daz = I.aw - J.aw;

    // ---- reverse ----
    double b_dvx = 0.0, b_dvy = 0.0, b_dvz = 0.0, b_dax = 0.0, b_day = 0.0, b_daz = 0.0;
    double b_dx = 0.0, b_dy = 0.0, b_dz = 0.0;
    double b_ex = 0.0, b_ey = 0.0, b_ez = 0.0;
    double b_gri = 0.0, b_grj = 0.0;
    // CullenDehnenAlphaDiagnostics loop
    double b_rx = B.vxx * dvx + B.vyx * dvy + B.vzx * dvz + B.axx * dax;
    double b_ry = B.vxy * dvx + B.vyy * dvy + B.vzy * dvz + B.ayy * day;
    double b_rz = 0.0;
    b_dvx += B.vxx * rx + B.vxy * ry;
    b_dvy += B.vyx * rx + B.vyy * ry;
    b_dvz += B.vzx * rx + B.vzy * ry;
    b_dax += B.axx * rx;
    b_day += B.ayy * ry;
    if constexpr (DIM == 3) {
        b_rx += B.ayx * day + B.azx * daz - B.rxx * dx;
        b_ry += B.axy * dax + B.azy * daz - B.rxy * dx - B.ryy * dy;
        b_rz = B.vxz * dvx + B.vyz * dvy + B.vzz * dvz + B.axz * dax + B.ayz * day + B.azz * daz - B.rxz * dx -
               B.ryz * dy - B.rzz * dz;
        b_dvx += B.vxz * rz;
        b_dvy += B.vyz * rz;
        b_dvz += B.vzz * rz;
        b_dax += B.axy * ry + B.axz * rz;
        b_day += B.ayx * rx + B.ayz * rz;
        b_daz = B.azx * rx + B.azy * ry + B.azz * rz;
        b_dx += -(B.rxx * rx + B.rxy * ry + B.rxz * rz);
        b_dy += -(B.ryy * ry + B.ryz * rz);
        b_dz += -(B.rzz * rz);
    }
    double b_dwix = mj * b_rx, b_dwiy = mj * b_ry, b_dwiz = mj * b_rz;
    double b_dwjx = 0.0, b_dwjy = 0.0, b_dwjz = 0.0;

    // MagneticStateRates
    const double b_brt = B.bx * dvx + B.by * dvy + B.bz * dvz;
    b_dvx += B.bx * brt;
    b_dvy += B.by * brt;
    b_dvz += B.bz * brt;
    const double b_diss = B.bx * dbx + B.by * dby + B.bz * dbz - 0.5 * db2 * inv_mu0 * B.ae;
    double b_dbx = B.bx * diss, b_dby = B.by * diss, b_dbz = B.bz * diss;
    const double b_dpsi = -(B.bx * ex + B.by * ey + B.bz * ez);
    b_ex += -B.bx * dpsi;
    b_ey += -B.by * dpsi;
    b_ez += -B.bz * dpsi;
    const double b_db2 = -0.5 * diss * inv_mu0 * B.ae;
    double b_pmi = B.divbs * pbi_, b_pmj = B.divbs * pbj_;
    double b_pbi_ = B.divbs * pmi, b_pbj_ = B.divbs * pmj;
    const double b_pdb = -mj * gri * B.divbd;
    b_gri += -mj * pdb * B.divbd;
    b_pmi += b_dpsi * I.psich;
    b_pmj += b_dpsi * J.psich;
    Ib.psich += b_dpsi * pmi;
    Jb.psich += b_dpsi * pmj;
    b_pmi += 0.5 * vsb * P.ar * b_diss;
    b_pmj += 0.5 * vsb * P.ar * b_diss;
    const double b_vsb = sdiss * P.ar * b_diss;
    pb.ar += sdiss * vsb * b_diss;
    double b_tx = 0.0, b_ty = 0.0, b_tz = 0.0;
    if (vsb > 0.0) {
        const double g = b_vsb / vsb;
        b_tx = g * tx;
        b_ty = g * ty;
        b_tz = g * tz;
    }
    b_dvx += b_tx;
    b_dvy += b_ty;
    b_dvz += b_tz;
    double b_dot = -(b_tx * ex + b_ty * ey + b_tz * ez);
    b_ex += -b_tx * dot;
    b_ey += -b_ty * dot;
    b_ez += -b_tz * dot;
    b_pmi += -b_brt * pbi_;
    b_pbi_ += -b_brt * pmi;
    double b_r2ii = b_pmi * mj * gri, b_r2ij = b_pmj * mj * grj;
    b_gri += b_pmi * mj * r2ii;
    b_grj += b_pmj * mj * r2ij;
    b_dbx += 2.0 * dbx * b_db2 + b_pdb * ex;
    b_dby += 2.0 * dby * b_db2 + b_pdb * ey;
    b_dbz += 2.0 * dbz * b_db2 + b_pdb * ez;
    b_ex += b_pdb * dbx + b_pbi_ * I.bx + b_pbj_ * J.bx;
    b_ey += b_pdb * dby + b_pbi_ * I.by + b_pbj_ * J.by;
    b_ez += b_pdb * dbz + b_pbi_ * I.bz + b_pbj_ * J.bz;
    Ib.bx += b_dbx + b_pbi_ * ex;
    Ib.by += b_dby + b_pbi_ * ey;
    Ib.bz += b_dbz + b_pbi_ * ez;
    Jb.bx += -b_dbx + b_pbj_ * ex;
    Jb.by += -b_dby + b_pbj_ * ey;
    Jb.bz += -b_dbz + b_pbj_ * ez;

    // MagneticStressAccelerations
    b_ex += -B.ax * gradp;
    b_ey += -B.ay * gradp;
    b_ez += -B.az * gradp;
    const double b_gradp = -(B.ax * ex + B.ay * ey + B.az * ez);
    const double b_pi_b = sm * (B.ax * brxi + B.ay * bryi + B.az * brzi);
    const double b_pj_b = sm * (B.ax * brxj + B.ay * bryj + B.az * brzj);
    double b_brxi = sm * B.ax * pi_b, b_bryi = sm * B.ay * pi_b, b_brzi = sm * B.az * pi_b;
    double b_brxj = sm * B.ax * pj_b, b_bryj = sm * B.ay * pj_b, b_brzj = sm * B.az * pj_b;
    const double b_Pi = b_pi_b * gri, b_Pj = b_pj_b * grj;
    b_gri += b_pi_b * Pi;
    b_grj += b_pj_b * Pj;
    b_brxi += b_Pi * ex;
    b_bryi += b_Pi * ey;
    b_brzi += b_Pi * ez;
    b_brxj += b_Pj * ex;
    b_bryj += b_Pj * ey;
    b_brzj += b_Pj * ez;
    b_ex += b_Pi * brxi + b_Pj * brxj;
    b_ey += b_Pi * bryi + b_Pj * bryj;
    b_ez += b_Pi * brzi + b_Pj * brzj;
    const double b_bro2i = b_gradp * mj * 0.5 * gri, b_bro2j = b_gradp * mj * 0.5 * grj;
    b_gri += b_gradp * mj * 0.5 * bro2i;
    b_grj += b_gradp * mj * 0.5 * bro2j;

    // MHDAccelerations
    const double wae = B.ae;
    const double Eij = I.e - J.e, G = I.hir * griA + J.hir * grjA;
    pb.ac += wae * vsig2 * Eij * mj * G;
    const double b_vsig2 = wae * P.ac * Eij * mj * G;
    const double b_E = wae * P.ac * vsig2 * mj * G;
    const double b_G = wae * P.ac * vsig2 * Eij * mj;
    Ib.e += b_E;
    Jb.e -= b_E;
    Ib.hir += b_G * griA;
    Jb.hir += b_G * grjA;
    double b_griA = b_G * I.hir, b_grjA = b_G * J.hir;
    const double V = dvx * dwix + dvy * dwiy + dvz * dwiz;
    double b_pbi = wae * mj * I.om * V;
    Ib.om += wae * mj * pbi * V;
    const double b_V = wae * mj * pbi * I.om;
    b_dvx += b_V * dwix;
    b_dvy += b_V * dwiy;
    b_dvz += b_V * dwiz;
    b_dwix += b_V * dvx;
    b_dwiy += b_V * dvy;
    b_dwiz += b_V * dvz;
    const double b_sx = -mj * (B.ax * dwix + B.ay * dwiy + B.az * dwiz);
    const double b_sjx = -mj * (B.ax * dwjx + B.ay * dwjy + B.az * dwjz);
    b_dwix += -mj * sx * B.ax;
    b_dwiy += -mj * sx * B.ay;
    b_dwiz += -mj * sx * B.az;
    b_dwjx += -mj * sjx * B.ax;
    b_dwjy += -mj * sjx * B.ay;
    b_dwjz += -mj * sjx * B.az;
    double b_dotA = 0.0;
    if (visc) {
        double b_qvi = wae * mj * dotA * griA;
        b_dotA += wae * mj * qvi * griA;
        b_griA += wae * mj * qvi * dotA;
        b_qvi += b_sx * I.om;
        Ib.om += b_sx * qvi;
        const double b_qvj = b_sjx * J.om;
        Jb.om += b_sjx * qvj;
        Ib.hir += -b_qvi * vsi * dotA * P.av;
        const double b_vsi = -b_qvi * I.hir * dotA * P.av;
        b_dotA += -b_qvi * I.hir * vsi * P.av;
        pb.av += -b_qvi * I.hir * vsi * dotA;
        Jb.hir += -b_qvj * vsj * dotA * P.av;
        const double b_vsj = -b_qvj * J.hir * dotA * P.av;
        b_dotA += -b_qvj * J.hir * vsj * P.av;
        pb.av += -b_qvj * J.hir * vsj * dotA;
        if (vsi_on) {
            Ib.al += b_vsi * I.vw;
            Ib.vw += b_vsi * I.al;
            b_dotA += -2.0 * b_vsi;
        }
        if (vsj_on) {
            Jb.al += b_vsj * J.vw;
            Jb.vw += b_vsj * J.al;
            b_dotA += -2.0 * b_vsj;
        }
    }
    b_pbi += b_sx * I.om;
    Ib.om += b_sx * pbi;
    const double b_pbj = b_sjx * J.om;
    Jb.om += b_sjx * pbj;
    if (argmax && cpos) {
        const double bc = B.dtcfl;
        if (take_i) Ib.vw += bc;
        else Jb.vw += bc;
        b_dotA += -2.0 * bc;
    }
    if (vsig2 > 0.0) {
        const double b_Q = b_vsig2 * 0.5 / vsig2;
        const double b_apd = b_Q / rhoij;
        const double b_rhoij = -b_Q * Q / rhoij;
        Ib.rho += 0.5 * b_rhoij;
        Jb.rho += 0.5 * b_rhoij;
        const double sgn = pdiff < 0.0 ? -1.0 : 1.0;
        Ib.p += sgn * b_apd;
        Jb.p -= sgn * b_apd;
    }
    b_dotA += B.div * mj * griA;
    b_griA += B.div * mj * dotA;
    if (!near8) {
        b_dot += b_dotA;
        b_gri += b_griA;
        b_grj += b_grjA;
    }

    // pair geometry
    b_dvx += b_dot * ex;
    b_dvy += b_dot * ey;
    b_dvz += b_dot * ez;
    b_ex += b_dot * dvx;
    b_ey += b_dot * dvy;
    b_ez += b_dot * dvz;
    Ib.om += b_gri * Si;
    Jb.om += b_grj * Sj;
    const double b_Si = b_gri * I.om, b_Sj = b_grj * J.om;
    b_ex += b_Si * dwix + b_Sj * dwjx;
    b_ey += b_Si * dwiy + b_Sj * dwjy;
    b_ez += b_Si * dwiz + b_Sj * dwjz;
    b_dwix += b_Si * ex;
    b_dwiy += b_Si * ey;
    b_dwiz += b_Si * ez;
    b_dwjx += b_Sj * ex;
    b_dwjy += b_Sj * ey;
    b_dwjz += b_Sj * ez;
    Ib.u += b_dvx;
    Ib.v += b_dvy;
    Ib.w += b_dvz;
    Jb.u -= b_dvx;
    Jb.v -= b_dvy;
    Jb.w -= b_dvz;
    Ib.au += b_dax;
    Ib.av += b_day;
    Jb.au -= b_dax;
    Jb.av -= b_day;
    if constexpr (DIM == 3) {
        Ib.aw += b_daz;
        Jb.aw -= b_daz;
    }
    const double b_gi = b_dwix * dx + b_dwiy * dy + b_dwiz * dz;
    const double b_gj = b_dwjx * dx + b_dwjy * dy + b_dwjz * dz;
    b_dx += b_dwix * gi + b_dwjx * gj;
    b_dy += b_dwiy * gi + b_dwjy * gj;
    b_dz += b_dwiz * gi + b_dwjz * gj;
    // the side values per pair: B/rho, P/rho^2, 1/rho from hir
    b_brxi += b_bro2i * 2.0 * brxi * inv_mu0;
    b_bryi += b_bro2i * 2.0 * bryi * inv_mu0;
    b_brzi += b_bro2i * 2.0 * brzi * inv_mu0;
    b_brxj += b_bro2j * 2.0 * brxj * inv_mu0;
    b_bryj += b_bro2j * 2.0 * bryj * inv_mu0;
    b_brzj += b_bro2j * 2.0 * brzj * inv_mu0;
    Ib.bx += b_brxi * irhoi;
    Ib.by += b_bryi * irhoi;
    Ib.bz += b_brzi * irhoi;
    Jb.bx += b_brxj * irhoj;
    Jb.by += b_bryj * irhoj;
    Jb.bz += b_brzj * irhoj;
    double b_irhoi = b_brxi * I.bx + b_bryi * I.by + b_brzi * I.bz;
    double b_irhoj = b_brxj * J.bx + b_bryj * J.by + b_brzj * J.bz;
    Ib.p += b_pbi * r2ii;
    b_r2ii += b_pbi * I.p;
    Jb.p += b_pbj * r2ij;
    b_r2ij += b_pbj * J.p;
    b_irhoi += 2.0 * irhoi * b_r2ii;
    b_irhoj += 2.0 * irhoj * b_r2ij;
    Ib.hir += 2.0 * b_irhoi;
    Jb.hir += 2.0 * b_irhoj;
    if (far12) {
        b_dx += b_ex * ir;
        b_dy += b_ey * ir;
        b_dz += b_ez * ir;
        double b_ir = b_ex * dx + b_ey * dy + b_ez * dz + b_gi * wpi * fac_i + b_gj * wpj * fac_j;
        const double b_qi = b_gi * fac_i * ir * w2i, b_qj = b_gj * fac_j * ir * w2j;
        const double b_faci = b_gi * wpi * ir, b_facj = b_gj * wpj * ir;
        double b_r = b_qi * I.h1 + b_qj * J.h1;
        Ib.h1 += b_qi * r + b_faci * (double)(DIM + 1) * fac_i / I.h1;
        Jb.h1 += b_qj * r + b_facj * (double)(DIM + 1) * fac_j / J.h1;
        b_r += -b_ir * ir * ir;
        b_dx += b_r * dx * ir;
        b_dy += b_r * dy * ir;
        b_dz += b_r * dz * ir;
    }
    db[0] += b_dx;
    db[1] += b_dy;
    db[2] += DIM == 3 ? b_dz : 0.0;
}

// ---------------------------------------------------------------------------
// the force epilogue (equations.py:137-140, 276-278, 356-429, 559-624)
// ---------------------------------------------------------------------------

template <class T>
struct Epi {
    T ax, ay, az, ae, ah, abx, aby, abz, apsi, alphaloc, dt_adapt;
};
constexpr int kNEpi = 11;
template <class T>
ADJ_HD T* epi_ptr(Epi<T>& e) { return &e.ax; }
template <class T>
ADJ_HD const T* epi_ptr(const Epi<T>& e) { return &e.ax; }

// me: the destination's raw values (h, cs, psi = psipred, rho, om, bx..bz =
// predicted B, e = epred); S = side_of(me)
template <int DIM, class T>
ADJ_HD Epi<T> epi_fwd(const Acc<T>& A, const Raw<T>& me, const Side<T>& S, const Par<T>& P) {
    Epi<T> o;
    const T hi = me.h, cs = me.cs, psipred = me.psi, rhoi = me.rho, omi = me.om, ei = me.e;
    const T vwi = S.vw, pi_ = S.p;
    const T bxi = me.bx, byi = me.by, bzi = me.bz;
    const double inv_mu0 = P.inv_mu0;
    const T div = -A.div / rhoi;
    o.ah = hi * div * (1.0 / (double)DIM);
    T div_a, dvxdx, dvxdy, dvxdz, dvydx, dvydy, dvydz, dvzdx, dvzdy, dvzdz;
    const T termnorm = omi / rhoi;
    bool inverted = false;
    if constexpr (DIM == 3) {
        const T rxx = A.rxx, rxy = A.rxy, rxz = A.rxz, ryy = A.ryy, ryz = A.ryz, rzz = A.rzz;
        const T denom = rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
        if (fabs(denom) > kFloatTiny) {
            inverted = true;
            const T dd = 1.0 / denom;
            const T txx = ryy * rzz - ryz * ryz;
            const T txy = rxz * ryz - rzz * rxy;
            const T txz = rxy * ryz - rxz * ryy;
            const T tyy = rzz * rxx - rxz * rxz;
            const T tyz = rxy * rxz - rxx * ryz;
            const T tzz = rxx * ryy - rxy * rxy;
            const T gax = (A.axx * txx + A.axy * txy + A.axz * txz) * dd;
            const T gay = (A.ayx * txy + A.ayy * tyy + A.ayz * tyz) * dd;
            const T gaz = (A.azx * txz + A.azy * tyz + A.azz * tzz) * dd;
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
        dvxdx = -termnorm * A.vxx;
        dvxdy = -termnorm * A.vxy;
        dvxdz = -termnorm * A.vxz;
        dvydx = -termnorm * A.vyx;
        dvydy = -termnorm * A.vyy;
        dvydz = -termnorm * A.vyz;
        dvzdx = -termnorm * A.vzx;
        dvzdy = -termnorm * A.vzy;
        dvzdz = -termnorm * A.vzz;
    }
    const T divvdt =
        div_a - (dvxdx * dvxdx + dvydy * dvydy + dvzdz * dvzdz + 2.0 * (dvxdy * dvydx + dvxdz * dvzdx + dvydz * dvzdy));
    const T divv = dvxdx + dvydy + dvzdz;
    const T cvx = dvzdy - dvydz, cvy = dvxdz - dvzdx, cvz = dvydx - dvxdy;
    const T comp = fmax(-divv, 0.0);
    const T fc = comp * comp;
    const T den = fc + (cvx * cvx + cvy * cvy + cvz * cvz);
    const T xil = den > 0.0 ? fc / den : T(1.0);
    const T source = 10.0 * hi * hi * xil * fmax(-divvdt, 0.0);
    o.alphaloc = fmax(P.alpha_min, P.alpha_max * fmax(fmin(source / (cs * cs), 1.0), 0.0));

    const T ch = P.chf * vwi;
    o.apsi = -ch * A.divbd / rhoi - P.sigma * psipred * ch / hi - 0.5 * P.cleaning_active * psipred * div;
    T dtcfl = fmax(A.dtcfl, vwi);
    dtcfl = fmax(dtcfl, ch);
    T ae = A.ae;
    if (dtcfl > 0.0) {
        const T dtc = P.c_cour * hi / dtcfl;
        if (ei + dtc * ae < 0.0 && ei > 0.0) ae /= 1.0 - dtc * ae / ei;
    }
    o.ae = ae;
    const T b2 = bxi * bxi + byi * byi + bzi * bzi;
    const T divb = A.divbs;
    T frac(0.0);
    if (b2 > 0.0) {
        const T beta = 2.0 * P.mu0 * pi_ / b2;
        frac = fmax(0.0, fmin(1.0, (10.0 - beta) * 0.125));
    }
    const T corr = divb * frac * inv_mu0;
    o.ax = A.ax - bxi * corr;
    o.ay = A.ay - byi * corr;
    o.az = A.az - bzi * corr;
    o.abx = A.bx;
    o.aby = A.by;
    o.abz = A.bz;
    T dt_adapt(1.0e29);
    if (dtcfl > 0.0) dt_adapt = P.c_cour * hi / dtcfl;
    const T f2 = o.ax * o.ax + o.ay * o.ay + o.az * o.az;
    if (f2 > 0.0) dt_adapt = fmin(dt_adapt, P.c_force * sqrt(hi / sqrt(f2)));
    o.dt_adapt = dt_adapt;
    return o;
}

struct EpiParBar {
    double chf, sigma, alpha_max;
};

// Reverse of epi_fwd: O = bars of its outputs; adds the bars of the
// accumulators to Ab, of the raw values to rb (h, cs, psi, rho, om, bx..bz,
// e), of S.vw and S.p to *b_vw, *b_p and of the parameters to pb.
template <int DIM>
ADJ_HD void epi_rev(const Acc<double>& A, const Raw<double>& me, const Side<double>& S, const Par<double>& P,
                    const Epi<double>& O, Acc<double>& Ab, RawBar& rb, double* b_vw, double* b_p, EpiParBar& pb) {
    // ---- the forward values ----
    const double hi = me.h, cs = me.cs, psi = me.psi, rho = me.rho, om = me.om, ei = me.e;
    const double vw = S.vw, p = S.p;
    const double bx = me.bx, by = me.by, bz = me.bz;
    const double inv_mu0 = P.inv_mu0;
    const double div = -A.div / rho;
    const double tn = om / rho;
    const double V[3][3] = {{A.vxx, A.vxy, A.vxz}, {A.vyx, A.vyy, A.vyz}, {A.vzx, A.vzy, A.vzz}};
    const double Aa[3][3] = {{A.axx, A.axy, A.axz}, {A.ayx, A.ayy, A.ayz}, {A.azx, A.azy, A.azz}};
    double M[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    bool inverted = false;
    if constexpr (DIM == 3) {
        const double rxx = A.rxx, rxy = A.rxy, rxz = A.rxz, ryy = A.ryy, ryz = A.ryz, rzz = A.rzz;
        const double denom =
            rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
        if (fabs(denom) > kFloatTiny) {
            inverted = true;
            const double dd = 1.0 / denom;
            const double txx = ryy * rzz - ryz * ryz, txy = rxz * ryz - rzz * rxy, txz = rxy * ryz - rxz * ryy;
            const double tyy = rzz * rxx - rxz * rxz, tyz = rxy * rxz - rxx * ryz, tzz = rxx * ryy - rxy * rxy;
            M[0][0] = txx * dd;
            M[0][1] = M[1][0] = txy * dd;
            M[0][2] = M[2][0] = txz * dd;
            M[1][1] = tyy * dd;
            M[1][2] = M[2][1] = tyz * dd;
            M[2][2] = tzz * dd;
        }
    }
    double dv[3][3];
    double div_a = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int b = 0; b < 3; ++b) {
            if (inverted) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += V[r][c] * M[c][b];
                dv[r][b] = -s;
            } else {
                dv[r][b] = -tn * V[r][b];
            }
        }
    if (inverted) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) div_a -= Aa[r][c] * M[c][r];
    } else {
        div_a = -tn * (A.axx + A.ayy + A.azz);
    }
    const double divvdt = div_a - (dv[0][0] * dv[0][0] + dv[1][1] * dv[1][1] + dv[2][2] * dv[2][2] +
                                   2.0 * (dv[0][1] * dv[1][0] + dv[0][2] * dv[2][0] + dv[1][2] * dv[2][1]));
    const double divv = dv[0][0] + dv[1][1] + dv[2][2];
    const double cvx = dv[2][1] - dv[1][2], cvy = dv[0][2] - dv[2][0], cvz = dv[1][0] - dv[0][1];
    const bool comp_on = -divv >= 0.0;
    const double comp = comp_on ? -divv : 0.0;
    const double fc = comp * comp;
    const double den = fc + (cvx * cvx + cvy * cvy + cvz * cvz);
    const double xil = den > 0.0 ? fc / den : 1.0;
    const bool g_on = -divvdt >= 0.0;
    const double g = g_on ? -divvdt : 0.0;
    const double source = 10.0 * hi * hi * xil * g;
    const double s1 = source / (cs * cs);
    const bool s1_lo = s1 <= 1.0;
    const double x1 = s1_lo ? s1 : 1.0;
    const bool s1_on = s1_lo && x1 >= 0.0;
    const double c1 = x1 >= 0.0 ? x1 : 0.0;
    const bool al_on = !(P.alpha_min >= P.alpha_max * c1);
    const double ch = P.chf * vw;
    const bool keep_a = A.dtcfl >= vw;
    const double dtc0 = keep_a ? A.dtcfl : vw;
    const bool keep_0 = dtc0 >= ch;
    const double dtcfl = keep_0 ? dtc0 : ch;
    bool lim = false;
    double dtc = 0.0, zl = 1.0;
    if (dtcfl > 0.0) {
        dtc = P.c_cour * hi / dtcfl;
        if (ei + dtc * A.ae < 0.0 && ei > 0.0) {
            lim = true;
            zl = 1.0 - dtc * A.ae / ei;
        }
    }
    const double ae = lim ? A.ae / zl : A.ae;
    const double b2 = bx * bx + by * by + bz * bz;
    const double divb = A.divbs;
    double frac = 0.0, beta = 0.0;
    bool frac_on = false;
    if (b2 > 0.0) {
        beta = 2.0 * P.mu0 * p / b2;
        const double y = (10.0 - beta) * 0.125;
        const double z = 1.0 <= y ? 1.0 : y;
        frac = 0.0 >= z ? 0.0 : z;
        frac_on = !(1.0 <= y) && !(0.0 >= y);
    }
    const double corr = divb * frac * inv_mu0;
    const double ax = A.ax - bx * corr, ay = A.ay - by * corr, az = A.az - bz * corr;
    const bool c_on = dtcfl > 0.0;
    const double dta_c = c_on ? P.c_cour * hi / dtcfl : 1.0e29;
    const double f2 = ax * ax + ay * ay + az * az;
    bool take_f = false;
    double sf = 0.0, gq = 0.0;
    if (f2 > 0.0) {
        sf = sqrt(f2);
        gq = hi / sf;
        const double dtf = P.c_force * sqrt(gq);
        take_f = !(dta_c <= dtf);
    }

    // ---- reverse ----
    double b_ax = O.ax, b_ay = O.ay, b_az = O.az, b_dtcfl = 0.0, b_rho = 0.0, b_div = 0.0, b_ch = 0.0;
    // dt_adapt
    if (take_f) {
        const double b_gq = O.dt_adapt * P.c_force * 0.5 / sqrt(gq);
        rb.h += b_gq / sf;
        const double b_sf = -b_gq * gq / sf;
        const double b_f2 = b_sf * 0.5 / sf;
        b_ax += 2.0 * ax * b_f2;
        b_ay += 2.0 * ay * b_f2;
        b_az += 2.0 * az * b_f2;
    } else if (c_on) {
        rb.h += O.dt_adapt * P.c_cour / dtcfl;
        b_dtcfl += -O.dt_adapt * dta_c / dtcfl;
    }
    // divB correction
    Ab.ax += b_ax;
    Ab.ay += b_ay;
    Ab.az += b_az;
    rb.bx += -b_ax * corr;
    rb.by += -b_ay * corr;
    rb.bz += -b_az * corr;
    const double b_corr = -(b_ax * bx + b_ay * by + b_az * bz);
    Ab.divbs += b_corr * frac * inv_mu0;
    if (frac_on) {
        const double b_beta = -0.125 * (b_corr * divb * inv_mu0);
        *b_p += b_beta * 2.0 * P.mu0 / b2;
        const double b_b2 = -b_beta * beta / b2;
        rb.bx += 2.0 * bx * b_b2;
        rb.by += 2.0 * by * b_b2;
        rb.bz += 2.0 * bz * b_b2;
    }
    // energy limiter
    if (lim) {
        Ab.ae += O.ae / zl;
        const double b_z = -O.ae * ae / zl;
        const double b_dtc = -b_z * A.ae / ei;
        Ab.ae += -b_z * dtc / ei;
        rb.e += b_z * dtc * A.ae / (ei * ei);
        rb.h += b_dtc * P.c_cour / dtcfl;
        b_dtcfl += -b_dtc * dtc / dtcfl;
    } else {
        Ab.ae += O.ae;
    }
    // dt_cfl = max(A.dtcfl, vw, ch)
    if (keep_0) {
        if (keep_a) Ab.dtcfl += b_dtcfl;
        else *b_vw += b_dtcfl;
    } else {
        b_ch += b_dtcfl;
    }
    // apsi
    {
        const double b = O.apsi;
        b_ch += b * (-A.divbd / rho - P.sigma * psi / hi);
        Ab.divbd += -b * ch / rho;
        b_rho += b * ch * A.divbd / (rho * rho);
        pb.sigma += -b * psi * ch / hi;
        rb.psi += b * (-P.sigma * ch / hi - 0.5 * P.cleaning_active * div);
        rb.h += b * P.sigma * psi * ch / (hi * hi);
        b_div += -b * 0.5 * P.cleaning_active * psi;
    }
    pb.chf += b_ch * vw;
    *b_vw += b_ch * P.chf;
    // alpha_loc
    double b_c1 = 0.0;
    if (al_on) {
        pb.alpha_max += O.alphaloc * c1;
        b_c1 = O.alphaloc * P.alpha_max;
    }
    const double b_s1 = s1_on ? b_c1 : 0.0;
    const double b_source = b_s1 / (cs * cs);
    rb.cs += -b_s1 * 2.0 * s1 / cs;
    rb.h += b_source * 20.0 * hi * xil * g;
    const double b_xil = b_source * 10.0 * hi * hi * g;
    const double b_g = b_source * 10.0 * hi * hi * xil;
    const double b_divvdt = g_on ? -b_g : 0.0;
    double b_fc = 0.0, b_den = 0.0;
    if (den > 0.0) {
        b_fc = b_xil / den;
        b_den = -b_xil * xil / den;
    }
    b_fc += b_den;
    const double b_cvx = 2.0 * cvx * b_den, b_cvy = 2.0 * cvy * b_den, b_cvz = 2.0 * cvz * b_den;
    const double b_comp = 2.0 * comp * b_fc;
    const double b_divv = comp_on ? -b_comp : 0.0;
    double Db[3][3];
    for (int r = 0; r < 3; ++r)
        for (int b = 0; b < 3; ++b) Db[r][b] = 0.0;
    Db[2][1] += b_cvx;
    Db[1][2] -= b_cvx;
    Db[0][2] += b_cvy;
    Db[2][0] -= b_cvy;
    Db[1][0] += b_cvz;
    Db[0][1] -= b_cvz;
    for (int a = 0; a < 3; ++a) Db[a][a] += b_divv - 2.0 * dv[a][a] * b_divvdt;
    Db[0][1] += -2.0 * dv[1][0] * b_divvdt;
    Db[1][0] += -2.0 * dv[0][1] * b_divvdt;
    Db[0][2] += -2.0 * dv[2][0] * b_divvdt;
    Db[2][0] += -2.0 * dv[0][2] * b_divvdt;
    Db[1][2] += -2.0 * dv[2][1] * b_divvdt;
    Db[2][1] += -2.0 * dv[1][2] * b_divvdt;
    const double b_diva = b_divvdt;
    double Vb[3][3], Aab[3][3];
    if (inverted) {
        double Mb[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int b = 0; b < 3; ++b) s += Db[r][b] * M[c][b];
                Vb[r][c] = -s;
                Aab[r][c] = -b_diva * M[c][r];
            }
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int r = 0; r < 3; ++r) s += V[r][c] * Db[r][b];
                Mb[c][b] = -s - b_diva * Aa[b][c];
            }
        // M = R^-1: Rbar = -M Mbar M (M symmetric), R symmetric in 6 entries
        double T1[3][3], Rb[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += M[a][c] * Mb[c][b];
                T1[a][b] = s;
            }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += T1[a][c] * M[c][b];
                Rb[a][b] = -s;
            }
        Ab.rxx += Rb[0][0];
        Ab.rxy += Rb[0][1] + Rb[1][0];
        Ab.rxz += Rb[0][2] + Rb[2][0];
        Ab.ryy += Rb[1][1];
        Ab.ryz += Rb[1][2] + Rb[2][1];
        Ab.rzz += Rb[2][2];
    } else {
        double b_tn = -b_diva * (A.axx + A.ayy + A.azz);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                Vb[r][c] = -tn * Db[r][c];
                Aab[r][c] = 0.0;
                b_tn += -V[r][c] * Db[r][c];
            }
        Aab[0][0] = Aab[1][1] = Aab[2][2] = -tn * b_diva;
        rb.om += b_tn / rho;
        b_rho += -b_tn * tn / rho;
    }
    Ab.vxx += Vb[0][0];
    Ab.vxy += Vb[0][1];
    Ab.vxz += Vb[0][2];
    Ab.vyx += Vb[1][0];
    Ab.vyy += Vb[1][1];
    Ab.vyz += Vb[1][2];
    Ab.vzx += Vb[2][0];
    Ab.vzy += Vb[2][1];
    Ab.vzz += Vb[2][2];
    Ab.axx += Aab[0][0];
    Ab.axy += Aab[0][1];
    Ab.axz += Aab[0][2];
    Ab.ayx += Aab[1][0];
    Ab.ayy += Aab[1][1];
    Ab.ayz += Aab[1][2];
    Ab.azx += Aab[2][0];
    Ab.azy += Aab[2][1];
    Ab.azz += Aab[2][2];
    // ah = h div / DIM, div = -A.div / rho
    rb.h += O.ah * div * (1.0 / (double)DIM);
    b_div += O.ah * hi * (1.0 / (double)DIM);
    Ab.div += -b_div / rho;
    b_rho += b_div * A.div / (rho * rho);
    rb.rho += b_rho;
    Ab.bx += O.abx;
    Ab.by += O.aby;
    Ab.bz += O.abz;
}

// ---------------------------------------------------------------------------
// the density iteration (DensityIteration.post_loop, equations.py:54-86)
// ---------------------------------------------------------------------------

template <class T>
ADJ_HD T ipow_d(const T& x, int dim) { return dim == 2 ? x * x : x * x * x; }
template <class T>
ADJ_HD T iroot_d(const T& x, int dim) { return dim == 2 ? sqrt(x) : cbrt(x); }

// One pass's post-loop from h and the pass's sums sw = sum m_j w(q_ij),
// sg = sum m_j (d w + q w'): the next h (converged: k (m / rho_sum)^(1/d))
// and the omega the pass stores; fin: the particle converged in this pass.
template <int DIM, class T>
ADJ_HD void newton_post(const T& hi, const T& sw, const T& sg, double mi, double hi0, double hfact, double ksig,
                        int iterate, T& hout, T& omega, bool& fin) {
    const T h1 = 1.0 / hi;
    const T fac = DIM == 2 ? ksig * h1 * h1 : ksig * h1 * h1 * h1;
    const T rho_sum = sw * fac;
    const T dwdh = -h1 * (fac * sg);
    const T rhoi = mi / ipow_d(hi / hfact, DIM);
    const T dhdrhoi = -hi / ((double)DIM * rhoi);
    const T omegai = 1.0 - dhdrhoi * dwdh;
    omega = 1.0 / omegai;
    const T func = rhoi - rho_sum;
    const T dfdh1 = omegai > 0.0 ? dhdrhoi / omegai : dhdrhoi / fabs(omegai + 2.220446049250313e-16);
    T hnew = hi - func * dfdh1;
    if (hnew > 1.2 * hi) hnew = 1.2 * hi;
    else if (hnew < 0.8 * hi) hnew = 0.8 * hi;
    const double diff = fabs(val(hnew) - val(hi)) / hi0;
    fin = ((diff < 1.0e-4) && (val(omegai) > 0.0) && (val(hi) > 0.0)) || !iterate;
    hout = fin ? T(hfact * iroot_d(mi / rho_sum, DIM)) : hnew;
}

}  // namespace adj
}  // namespace sph
