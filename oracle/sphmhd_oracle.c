/*
 * sphmhd_oracle.c -- FP64 CPU restatement of the reference SPH-MHD step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see the header).
 *
 * Every equation body is restated literally, one C function per reference
 * method, in the reference's operation order, compiled with
 * -ffp-contract=off so that no multiply-add is fused.  The group driver
 * follows the plan of src/scheme.py:70-81 group by group (no fusion), with
 * OpenMP over destination particles (every pair loop is gather-form: only
 * the destination is written).
 *
 * PySPH semantics the reference relies on but does not ship are defined here
 * (the HIP library follows the same definitions):
 *   kernels   W = sigma_d h^-d w(q), dW/dq, grad W = x (sigma h^-d w'(q)) h^-1 / r
 *             (0 when r <= 1e-12), dW/dh = -sigma h^-d h^-1 (d w + q w').
 *   pairs     XIJ = x_i - x_j with the minimum image on periodic axes
 *             (d > L/2 -> d - L, d < -L/2 -> d + L), RIJ = sqrt(XIJ.XIJ),
 *             RHOIJ = (rho_i + rho_j)/2, WI/GHI/DWI with h_i, DWJ with h_j,
 *             self pairs included.
 *   neighbours j is a neighbour of i iff r^2 < (R max(h_i, h_j))^2.
 *   cells     built once per evaluation from h_max and h_min (positions do
 *             not move during an evaluation): c = R h_max / k with
 *             k = min(floor(h_max/h_min), 4) when h_max >= 2 h_min, else 1
 *             (cells split where h varies, so the small-h particles do not
 *             scan cells sized for the largest h), and in 3-D k doubled (at
 *             most 4: half-width cells); a single 3-D instance whose h spans
 *             fewer than 64 levels (16 per octave) takes the cell size of
 *             cell_model_3d instead, at least R h_max / 8;
 *             n_a = max(1, floor(L_a/c)), while prod n_a > max(1, 8 N) the
 *             split is undone first and c then doubled (the model's c is
 *             doubled); c_eff_a = L_a / n_a over the box: the domain on
 *             periodic axes, else the particles' box at this evaluation
 *             widened by box_pad on each side;
 *             key = ix + n_x (iy + n_y iz) with i_a = floor((x_a - lo_a)/c_eff_a)
 *             clamped to [0, n_a - 1]; particles stably sorted by key.
 *   stencil   per pass, s_a = max(1, ceil(R h_max_now / c_eff_a)); an axis
 *             with 2 s_a + 1 >= n_a is scanned whole.
 *   groups    initialize (all particles) -> loop (neighbours) -> post_loop;
 *             an iterated group stops when iterations >= min and every
 *             particle converged, or iterations == max.
 *   PEC step  initialize, stage1, wrap, evaluate, stage2, wrap; then
 *             t += dt and dt = min dt_adapt (clipped to tf - t).
 *   wrap      x < lo -> x + L, x >= hi -> x - L (periodic axes).
 */
#include "sphmhd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* equations.py:6-11 */
#define DENSITY_TOLERANCE 1.0e-4
#define COURANT_FACTOR 0.3
#define FORCE_FACTOR 0.25
#define PAIR_DISTANCE_EPSILON 1.0e-12
#define INTERACTION_DISTANCE_EPSILON 1.0e-8
#define FLOAT_TINY 2.2250738585072014e-308

/* cell grid: at most this many cells per particle (max(1, 8 N)) */
#define kCellsPerParticle 8

/* particles.py:1-129 (itype is held separately as int32). */
#define ORC_PROPS(X)                                                                        \
    X(x) X(y) X(z) X(gid) X(tag) X(u) X(v) X(w) X(rho) X(rho_sum) X(h) X(m) X(cs) X(p) X(e) \
    X(au) X(av) X(aw) X(arho) X(ae) X(ah) X(x0) X(y0) X(z0) X(u0) X(v0) X(w0)               \
    X(upred) X(vpred) X(wpred) X(epred) X(e0) X(h0) X(div) X(dt_cfl) X(dt_force)            \
    X(dt_adapt) X(grhox) X(grhoy) X(grhoz) X(dwdh) X(omega) X(converged) X(alpha1)          \
    X(alpha10) X(Bx) X(By) X(Bz) X(Bevolx) X(Bevoly) X(Bevolz) X(psi)                       \
    X(Bevolxpred) X(Bevolypred) X(Bevolzpred) X(psipred) X(Bxpred) X(Bypred) X(Bzpred)      \
    X(aBevolx) X(aBevoly) X(aBevolz) X(apsi) X(Bevolx0) X(Bevoly0) X(Bevolz0) X(psi0)       \
    X(divBsymm) X(divBdiff) X(div_for_psi) X(alphaloc) X(xilimiter) X(divvdt)               \
    X(cd_rxx) X(cd_rxy) X(cd_rxz) X(cd_ryy) X(cd_ryz) X(cd_rzz)                             \
    X(cd_vxx) X(cd_vxy) X(cd_vxz) X(cd_vyx) X(cd_vyy) X(cd_vyz) X(cd_vzx) X(cd_vzy) X(cd_vzz) \
    X(cd_axx) X(cd_axy) X(cd_axz) X(cd_ayx) X(cd_ayy) X(cd_ayz) X(cd_azx) X(cd_azy) X(cd_azz) \
    X(dvdx_xx) X(dvdx_xy) X(dvdx_xz) X(dvdx_yx) X(dvdx_yy) X(dvdx_yz)                       \
    X(dvdx_zx) X(dvdx_zy) X(dvdx_zz)

#define ORC_ENUM(n) P_##n,
enum { ORC_PROPS(ORC_ENUM) NPROP };
#define ORC_NAME(n) #n,
static const char* PROP_NAMES[NPROP] = {ORC_PROPS(ORC_NAME)};

typedef struct {
    double xij[3], rij, r2ij, rhoij, wi, dwi[3], dwj[3], ghi;
} sym_t;

typedef struct {
    int n[3];
    double lo[3]; /* the grid's origin: the domain's on periodic axes, else the particles' box */
    double ceff[3], L[3], halfL[3];
    int64_t ncell;
    int64_t base; /* global key offset of this instance */
} grid_t;

struct orc {
    int64_t n;
    int ninst;
    int64_t* off;
    sphmhd_params* prm;
    double* p[NPROP];
    int32_t* itype;
    double *t, *dt, *dt_pending;
    int32_t* last_iters;
    int64_t steps;    /* PEC steps taken by orc_step */
    grid_t* grid;
    uint32_t* key;    /* per sorted position (global key) */
    int32_t* perm;    /* sorted position -> storage index */
    int32_t* cellof;  /* storage index -> instance-local cell */
    int64_t *cstart, *cend; /* per global cell: sorted range */
    int64_t cell_cap;
};

/* ------------------------------------------------------------------ */
/* Kernels (PySPH kernels, restated from the SPH literature)           */
/* ------------------------------------------------------------------ */

static const double ORC_PI = 3.14159265358979323846;

double orc_kernel_radius(int kernel) {
    switch (kernel) {
    case SPHMHD_CUBIC: return 2.0;
    case SPHMHD_QUINTIC: return 3.0;
    case SPHMHD_WENDLAND_C2: return 2.0;
    case SPHMHD_WENDLAND_C4: return 2.0;
    case SPHMHD_WENDLAND_C6: return 2.0;
    case SPHMHD_GAUSSIAN: return 3.0;
    case SPHMHD_SUPER_GAUSSIAN: return 3.0;
    }
    return 0.0;
}

static double k_sigma(int kernel, int dim) {
    const double pi = ORC_PI;
    switch (kernel) {
    case SPHMHD_CUBIC: return dim == 2 ? 10.0 / (7.0 * pi) : 1.0 / pi;
    case SPHMHD_QUINTIC: return dim == 2 ? 7.0 / (478.0 * pi) : 1.0 / (120.0 * pi);
    case SPHMHD_WENDLAND_C2: return dim == 2 ? 7.0 / (4.0 * pi) : 21.0 / (16.0 * pi);
    case SPHMHD_WENDLAND_C4: return dim == 2 ? 9.0 / (4.0 * pi) : 495.0 / (256.0 * pi);
    case SPHMHD_WENDLAND_C6: return dim == 2 ? 78.0 / (28.0 * pi) : 1365.0 / (512.0 * pi);
    case SPHMHD_GAUSSIAN: return dim == 2 ? 1.0 / pi : 1.0 / (pi * sqrt(pi));
    case SPHMHD_SUPER_GAUSSIAN: return dim == 2 ? 1.0 / pi : 1.0 / (pi * sqrt(pi));
    }
    return 0.0;
}

static double p5(double t) { double t2 = t * t; double t4 = t2 * t2; return t4 * t; }
static double p4(double t) { double t2 = t * t; return t2 * t2; }

/* shape w(q) and its derivative w'(q), without normalisation */
static void k_shape(int kernel, int dim, double q, double* w, double* dw) {
    switch (kernel) {
    case SPHMHD_CUBIC:
        if (q > 2.0) { *w = 0.0; *dw = 0.0; }
        else if (q > 1.0) { double t = 2.0 - q; *w = 0.25 * t * t * t; *dw = -0.75 * t * t; }
        else { *w = 1.0 - 1.5 * q * q * (1.0 - 0.5 * q); *dw = -3.0 * q * (1.0 - 0.75 * q); }
        return;
    case SPHMHD_QUINTIC: {
        double t3 = 3.0 - q, t2 = 2.0 - q, t1 = 1.0 - q;
        if (q > 3.0) { *w = 0.0; *dw = 0.0; }
        else if (q > 2.0) { *w = p5(t3); *dw = -5.0 * p4(t3); }
        else if (q > 1.0) { *w = p5(t3) - 6.0 * p5(t2); *dw = -5.0 * p4(t3) + 30.0 * p4(t2); }
        else {
            *w = p5(t3) - 6.0 * p5(t2) + 15.0 * p5(t1);
            *dw = -5.0 * p4(t3) + 30.0 * p4(t2) - 75.0 * p4(t1);
        }
        return;
    }
    case SPHMHD_WENDLAND_C2: {
        if (q >= 2.0) { *w = 0.0; *dw = 0.0; return; }
        double t = 1.0 - 0.5 * q;
        double t3 = t * t * t;
        *w = t3 * t * (2.0 * q + 1.0);
        *dw = -5.0 * q * t3;
        return;
    }
    case SPHMHD_WENDLAND_C4: {
        if (q >= 2.0) { *w = 0.0; *dw = 0.0; return; }
        double t = 1.0 - 0.5 * q;
        double t5 = p5(t);
        *w = t5 * t * ((35.0 / 12.0) * q * q + 3.0 * q + 1.0);
        *dw = -(14.0 / 3.0) * q * (1.0 + 2.5 * q) * t5;
        return;
    }
    case SPHMHD_WENDLAND_C6: {
        if (q >= 2.0) { *w = 0.0; *dw = 0.0; return; }
        double t = 1.0 - 0.5 * q;
        double t7 = p4(t) * t * t * t;
        *w = t7 * t * (4.0 * q * q * q + 6.25 * q * q + 4.0 * q + 1.0);
        *dw = -5.5 * q * t7 * (1.0 + 3.5 * q + 4.0 * q * q);
        return;
    }
    case SPHMHD_GAUSSIAN: {
        if (q >= 3.0) { *w = 0.0; *dw = 0.0; return; }
        double g = exp(-q * q);
        *w = g;
        *dw = -2.0 * q * g;
        return;
    }
    case SPHMHD_SUPER_GAUSSIAN: {
        if (q >= 3.0) { *w = 0.0; *dw = 0.0; return; }
        double g = exp(-q * q);
        double d2 = 0.5 * (double)dim;
        *w = g * (d2 + 1.0 - q * q);
        *dw = g * (-2.0 * q * (d2 + 1.0 - q * q) - 2.0 * q);
        return;
    }
    }
    *w = 0.0; *dw = 0.0;
}

static double k_fac(int kernel, int dim, double h1) {
    double f = k_sigma(kernel, dim) * h1 * h1;
    if (dim == 3) f = f * h1;
    return f;
}

double orc_kernel_w(int kernel, int dim, double r, double h) {
    double h1 = 1.0 / h, q = r * h1, w, dw;
    k_shape(kernel, dim, q, &w, &dw);
    return w * k_fac(kernel, dim, h1);
}

double orc_kernel_dwdq(int kernel, int dim, double r, double h) {
    double h1 = 1.0 / h, q = r * h1, w, dw;
    k_shape(kernel, dim, q, &w, &dw);
    return dw * k_fac(kernel, dim, h1);
}

void orc_kernel_grad(int kernel, int dim, const double* xij, double r, double h, double* grad) {
    double h1 = 1.0 / h, tmp;
    if (r > 1e-12) {
        double wdash = orc_kernel_dwdq(kernel, dim, r, h);
        tmp = wdash * h1 / r;
    } else {
        tmp = 0.0;
    }
    grad[0] = tmp * xij[0];
    grad[1] = tmp * xij[1];
    grad[2] = tmp * xij[2];
}

double orc_kernel_gradh(int kernel, int dim, double r, double h) {
    double h1 = 1.0 / h, q = r * h1, w, dw;
    k_shape(kernel, dim, q, &w, &dw);
    double fac = k_fac(kernel, dim, h1);
    w = w * fac;
    dw = dw * fac;
    return -h1 * ((double)dim * w + q * dw);
}

/* ------------------------------------------------------------------ */
/* Equation bodies (equations.py), literal                             */
/* ------------------------------------------------------------------ */

#define D(name) o->p[P_##name][d]
#define S(name) o->p[P_##name][s]

static double ipow(double x, int n) { return n == 2 ? x * x : x * x * x; }
static double iroot(double x, int n) { return n == 2 ? sqrt(x) : cbrt(x); }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* DensityIteration.initialize, equations.py:25-35 */
static void density_initialize(orc_t* o, int64_t d) {
    D(rho) = 0.0;
    D(div) = 0.0;
    D(grhox) = 0.0;
    D(grhoy) = 0.0;
    D(grhoz) = 0.0;
    D(arho) = 0.0;
    D(dwdh) = 0.0;
}

/* DensityIteration.loop, equations.py:38-50 */
static void density_loop(orc_t* o, int64_t d, int64_t s, const sym_t* y) {
    double mj = S(m);
    double dvx = D(upred) - S(upred);
    double dvy = D(vpred) - S(vpred);
    double dvz = D(wpred) - S(wpred);
    double vijdotdwij = dvx * y->dwi[0] + dvy * y->dwi[1] + dvz * y->dwi[2];
    D(rho) += mj * y->wi;
    D(arho) += mj * vijdotdwij;
    D(grhox) += mj * y->dwi[0];
    D(grhoy) += mj * y->dwi[1];
    D(grhoz) += mj * y->dwi[2];
    D(dwdh) += mj * y->ghi;
}

/* DensityIteration.post_loop, equations.py:53-88; returns -1 when the
 * particle sets equation_has_converged = -1 (equations.py:79). */
static int density_post_loop(orc_t* o, const sphmhd_params* P, int64_t d) {
    const int dim = P->dim;
    const double k = P->hfact;
    const int iterate_only_once = !P->density_iterate;
    int flag = 1;
    if (D(converged) != 1) {
        double mi = D(m);
        double hi = D(h);
        double hi0 = D(h0);
        double rhoi = mi / ipow(hi / k, dim);
        double dhdrhoi = -hi / (dim * rhoi);
        double dwdhi = D(dwdh);
        double omegai = 1.0 - dhdrhoi * dwdhi;
        double gradhi = 1.0 / omegai;
        D(omega) = gradhi;
        double func = rhoi - D(rho);
        double dfdh1 = omegai > 0.0 ? dhdrhoi / omegai : dhdrhoi / fabs(omegai + 2.220446049250313e-16);
        double hnew = hi - func * dfdh1;
        if (hnew > 1.2 * hi) hnew = 1.2 * hi;
        else if (hnew < 0.8 * hi) hnew = 0.8 * hi;
        double diff = fabs(hnew - hi) / hi0;
        int finalize_density = ((diff < DENSITY_TOLERANCE) && (omegai > 0.0) && (hi > 0.0)) || iterate_only_once;
        if (!finalize_density) {
            flag = -1;
            D(h) = hnew;
            D(converged) = 0;
        } else {
            D(arho) *= D(omega);
            D(ah) = D(arho) * dhdrhoi;
            D(h) = k * iroot(mi / D(rho), dim);
            D(converged) = 1;
        }
    }
    D(div) = -D(arho) / D(rho);
    return flag;
}

/* IdealGasEOS.loop, equations.py:104-106 */
static void ideal_gas_eos(orc_t* o, const sphmhd_params* P, int64_t d) {
    double gamma1 = P->gamma - 1.0;
    D(p) = gamma1 * D(rho) * D(epred);
    D(cs) = sqrt(P->gamma * D(p) / D(rho));
}

/* DensityFromSmoothingLength.initialize, equations.py:116-118 */
static void rho_from_h(orc_t* o, const sphmhd_params* P, int64_t d) {
    D(rho_sum) = D(rho);
    D(rho) = D(m) / ipow(D(h) / P->hfact, P->dim);
}

/* SmoothingLengthRateFromDensityRelation.initialize, equations.py:128-130 */
static void ah_from_rho(orc_t* o, const sphmhd_params* P, int64_t d) {
    double rhoh = D(m) / ipow(D(h) / P->hfact, P->dim);
    D(ah) *= rhoh / D(rho);
}

/* SmoothingLengthRateFromForceDivergence.initialize, equations.py:139-140 */
static void ah_from_div(orc_t* o, const sphmhd_params* P, int64_t d) {
    double dim1 = 1.0 / P->dim;
    D(ah) = D(h) * D(div) * dim1;
}

/* MagneticStressAccelerations.loop, equations.py:149-195 */
static void magstress_loop(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const sym_t* y) {
    const double mu0 = P->mu0;
    double runix, runiy, runiz;
    if (y->rij > PAIR_DISTANCE_EPSILON) {
        runix = y->xij[0] / y->rij;
        runiy = y->xij[1] / y->rij;
        runiz = y->xij[2] / y->rij;
    } else {
        runix = 0.0; runiy = 0.0; runiz = 0.0;
    }
    double grad_i = D(omega) * (runix * y->dwi[0] + runiy * y->dwi[1] + runiz * y->dwi[2]);
    double grad_j = S(omega) * (runix * y->dwj[0] + runiy * y->dwj[1] + runiz * y->dwj[2]);
    double rho_i = D(rho);
    double rho_j = S(rho);
    double brhox_i = D(Bxpred) / rho_i;
    double brhoy_i = D(Bypred) / rho_i;
    double brhoz_i = D(Bzpred) / rho_i;
    double brhox_j = S(Bxpred) / rho_j;
    double brhoy_j = S(Bypred) / rho_j;
    double brhoz_j = S(Bzpred) / rho_j;
    double bro2_i = (brhox_i * brhox_i + brhoy_i * brhoy_i + brhoz_i * brhoz_i) / mu0;
    double bro2_j = (brhox_j * brhox_j + brhoy_j * brhoy_j + brhoz_j * brhoz_j) / mu0;
    double gradp = S(m) * (0.5 * bro2_i * grad_i + 0.5 * bro2_j * grad_j);

    double sxx_i = -S(m) * brhox_i * brhox_i / mu0;
    double sxy_i = -S(m) * brhox_i * brhoy_i / mu0;
    double sxz_i = -S(m) * brhox_i * brhoz_i / mu0;
    double syy_i = -S(m) * brhoy_i * brhoy_i / mu0;
    double syz_i = -S(m) * brhoy_i * brhoz_i / mu0;
    double szz_i = -S(m) * brhoz_i * brhoz_i / mu0;
    double sxx_j = -S(m) * brhox_j * brhox_j / mu0;
    double sxy_j = -S(m) * brhox_j * brhoy_j / mu0;
    double sxz_j = -S(m) * brhox_j * brhoz_j / mu0;
    double syy_j = -S(m) * brhoy_j * brhoy_j / mu0;
    double syz_j = -S(m) * brhoy_j * brhoz_j / mu0;
    double szz_j = -S(m) * brhoz_j * brhoz_j / mu0;

    double projsx = (sxx_i * runix + sxy_i * runiy + sxz_i * runiz) * grad_i + (sxx_j * runix + sxy_j * runiy + sxz_j * runiz) * grad_j;
    double projsy = (sxy_i * runix + syy_i * runiy + syz_i * runiz) * grad_i + (sxy_j * runix + syy_j * runiy + syz_j * runiz) * grad_j;
    double projsz = (sxz_i * runix + syz_i * runiy + szz_i * runiz) * grad_i + (sxz_j * runix + syz_j * runiy + szz_j * runiz) * grad_j;
    double ax = -runix * gradp - projsx;
    double ay = -runiy * gradp - projsy;
    double az = -runiz * gradp - projsz;
    D(au) += ax;
    D(av) += ay;
    D(aw) += az;
}

/* MHDAccelerations.initialize, equations.py:207-214 */
static void mhdacc_initialize(orc_t* o, int64_t d) {
    D(au) = 0.0;
    D(av) = 0.0;
    D(aw) = 0.0;
    D(ae) = 0.0;
    D(dt_cfl) = 0.0;
    D(div) = 0.0;
    D(div_for_psi) = 0.0;
}

/* MHDAccelerations.loop, equations.py:217-275 */
static void mhdacc_loop(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const sym_t* y) {
    const double beta = 2.0;
    const double mu0 = P->mu0;
    double p_i = D(p);
    double pj = S(p);
    double rhoi = D(rho);
    double rhoj = S(rho);
    double rhoi2 = rhoi * rhoi;
    double rhoj2 = rhoj * rhoj;
    double pibrhoi2 = p_i / rhoi2;
    double pjbrhoj2 = pj / rhoj2;
    double mj = S(m);
    double bi2 = D(Bxpred) * D(Bxpred) + D(Bypred) * D(Bypred) + D(Bzpred) * D(Bzpred);
    double bj2 = S(Bxpred) * S(Bxpred) + S(Bypred) * S(Bypred) + S(Bzpred) * S(Bzpred);
    double vwave_i = sqrt(D(cs) * D(cs) + bi2 / (mu0 * rhoi));
    double vwave_j = sqrt(S(cs) * S(cs) + bj2 / (mu0 * rhoj));
    double runix, runiy, runiz;
    if (y->rij < INTERACTION_DISTANCE_EPSILON) {
        runix = 0.0; runiy = 0.0; runiz = 0.0;
    } else {
        runix = y->xij[0] / y->rij;
        runiy = y->xij[1] / y->rij;
        runiz = y->xij[2] / y->rij;
    }
    double dvx = D(upred) - S(upred);
    double dvy = D(vpred) - S(vpred);
    double dvz = D(wpred) - S(wpred);
    double dot = dvx * runix + dvy * runiy + dvz * runiz;
    double omegai = D(omega);
    double omegaj = S(omega);
    double grad_i = omegai * (runix * y->dwi[0] + runiy * y->dwi[1] + runiz * y->dwi[2]);
    double grad_j = omegaj * (runix * y->dwj[0] + runiy * y->dwj[1] + runiz * y->dwj[2]);
    D(div) += S(m) * dot * grad_i;
    double pdiff = fabs(p_i - pj);
    double vsig2 = sqrt(pdiff / y->rhoij);
    double dt_cfl = vwave_i - beta * dot;
    dt_cfl = dmax(dt_cfl, vwave_j - beta * dot);
    dt_cfl = dmax(dt_cfl, 0.0);
    D(dt_cfl) = dmax(D(dt_cfl), dt_cfl);

    if (dot < 0.0) {
        double vsigav_i = dmax(D(alpha1) * vwave_i - beta * dot, 0.0);
        double vsigav_j = dmax(S(alpha1) * vwave_j - beta * dot, 0.0);
        double qrho2_i = -0.5 / rhoi * vsigav_i * dot * P->artificial_viscosity;
        double qrho2_j = -0.5 / rhoj * vsigav_j * dot * P->artificial_viscosity;
        D(au) += -mj * (qrho2_i * omegai * y->dwi[0] + qrho2_j * omegaj * y->dwj[0]);
        D(av) += -mj * (qrho2_i * omegai * y->dwi[1] + qrho2_j * omegaj * y->dwj[1]);
        D(aw) += -mj * (qrho2_i * omegai * y->dwi[2] + qrho2_j * omegaj * y->dwj[2]);
        D(ae) += mj * qrho2_i * dot * grad_i;
    }

    D(au) += -mj * (pibrhoi2 * omegai * y->dwi[0] + pjbrhoj2 * omegaj * y->dwj[0]);
    D(av) += -mj * (pibrhoi2 * omegai * y->dwi[1] + pjbrhoj2 * omegaj * y->dwj[1]);
    D(aw) += -mj * (pibrhoi2 * omegai * y->dwi[2] + pjbrhoj2 * omegaj * y->dwj[2]);

    double vijdotdwi = dvx * y->dwi[0] + dvy * y->dwi[1] + dvz * y->dwi[2];
    D(ae) += mj * pibrhoi2 * omegai * vijdotdwi;

    double eij = D(epred) - S(epred);
    D(ae) += P->artificial_thermal_conductivity * vsig2 * eij * (0.5 * S(m) / rhoi * grad_i + 0.5 * mj / rhoj * grad_j);
}

/* MHDAccelerations.post_loop, equations.py:278-280 */
static void mhdacc_post_loop(orc_t* o, int64_t d) {
    D(div_for_psi) = -D(div) / D(rho);
    D(div) = -D(div) / D(rho);
}

/* CullenDehnenAlphaDiagnostics.initialize, equations.py:288-315 */
static void cd_initialize(orc_t* o, int64_t d) {
    D(alphaloc) = 0.0;
    D(xilimiter) = 1.0;
    D(divvdt) = 0.0;
    for (int q = P_cd_rxx; q <= P_cd_azz; ++q) o->p[q][d] = 0.0;
}

/* CullenDehnenAlphaDiagnostics.loop, equations.py:318-354 */
static void cd_loop(orc_t* o, int64_t d, int64_t s, const sym_t* y) {
    double runx = S(m) * y->dwi[0];
    double runy = S(m) * y->dwi[1];
    double runz = S(m) * y->dwi[2];
    double dvx = D(upred) - S(upred);
    double dvy = D(vpred) - S(vpred);
    double dvz = D(wpred) - S(wpred);
    double dax = D(au) - S(au);
    double day = D(av) - S(av);
    double daz = D(aw) - S(aw);

    D(cd_rxx) += -y->xij[0] * runx;
    D(cd_rxy) += -y->xij[0] * runy;
    D(cd_rxz) += -y->xij[0] * runz;
    D(cd_ryy) += -y->xij[1] * runy;
    D(cd_ryz) += -y->xij[1] * runz;
    D(cd_rzz) += -y->xij[2] * runz;

    D(cd_vxx) += dvx * runx;
    D(cd_vxy) += dvx * runy;
    D(cd_vxz) += dvx * runz;
    D(cd_vyx) += dvy * runx;
    D(cd_vyy) += dvy * runy;
    D(cd_vyz) += dvy * runz;
    D(cd_vzx) += dvz * runx;
    D(cd_vzy) += dvz * runy;
    D(cd_vzz) += dvz * runz;

    D(cd_axx) += dax * runx;
    D(cd_axy) += dax * runy;
    D(cd_axz) += dax * runz;
    D(cd_ayx) += day * runx;
    D(cd_ayy) += day * runy;
    D(cd_ayz) += day * runz;
    D(cd_azx) += daz * runx;
    D(cd_azy) += daz * runy;
    D(cd_azz) += daz * runz;
}

/* CullenDehnenAlphaDiagnostics.post_loop, equations.py:357-429.  alpha_min
 * and alpha_max generalise the clamp of :428; (0, 1) reproduce it. */
static void cd_post_loop(orc_t* o, const sphmhd_params* P, int64_t d) {
    double rxx = D(cd_rxx), rxy = D(cd_rxy), rxz = D(cd_rxz);
    double ryy = D(cd_ryy), ryz = D(cd_ryz), rzz = D(cd_rzz);
    double denom = rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
    double termnorm = D(omega) / D(rho);
    double div_a, dvxdx, dvxdy, dvxdz, dvydx, dvydy, dvydz, dvzdx, dvzdy, dvzdz;

    if (fabs(denom) > FLOAT_TINY) {
        double ddenom = 1.0 / denom;
        double term_xx = ryy * rzz - ryz * ryz;
        double term_xy = rxz * ryz - rzz * rxy;
        double term_xz = rxy * ryz - rxz * ryy;
        double term_yy = rzz * rxx - rxz * rxz;
        double term_yz = rxy * rxz - rxx * ryz;
        double term_zz = rxx * ryy - rxy * rxy;

        double gradaxdx = (D(cd_axx) * term_xx + D(cd_axy) * term_xy + D(cd_axz) * term_xz) * ddenom;
        double gradaydy = (D(cd_ayx) * term_xy + D(cd_ayy) * term_yy + D(cd_ayz) * term_yz) * ddenom;
        double gradazdz = (D(cd_azx) * term_xz + D(cd_azy) * term_yz + D(cd_azz) * term_zz) * ddenom;
        div_a = -(gradaxdx + gradaydy + gradazdz);

        double gradvxdx = (D(cd_vxx) * term_xx + D(cd_vxy) * term_xy + D(cd_vxz) * term_xz) * ddenom;
        double gradvxdy = (D(cd_vxx) * term_xy + D(cd_vxy) * term_yy + D(cd_vxz) * term_yz) * ddenom;
        double gradvxdz = (D(cd_vxx) * term_xz + D(cd_vxy) * term_yz + D(cd_vxz) * term_zz) * ddenom;
        double gradvydx = (D(cd_vyx) * term_xx + D(cd_vyy) * term_xy + D(cd_vyz) * term_xz) * ddenom;
        double gradvydy = (D(cd_vyx) * term_xy + D(cd_vyy) * term_yy + D(cd_vyz) * term_yz) * ddenom;
        double gradvydz = (D(cd_vyx) * term_xz + D(cd_vyy) * term_yz + D(cd_vyz) * term_zz) * ddenom;
        double gradvzdx = (D(cd_vzx) * term_xx + D(cd_vzy) * term_xy + D(cd_vzz) * term_xz) * ddenom;
        double gradvzdy = (D(cd_vzx) * term_xy + D(cd_vzy) * term_yy + D(cd_vzz) * term_yz) * ddenom;
        double gradvzdz = (D(cd_vzx) * term_xz + D(cd_vzy) * term_yz + D(cd_vzz) * term_zz) * ddenom;

        dvxdx = -gradvxdx; dvxdy = -gradvxdy; dvxdz = -gradvxdz;
        dvydx = -gradvydx; dvydy = -gradvydy; dvydz = -gradvydz;
        dvzdx = -gradvzdx; dvzdy = -gradvzdy; dvzdz = -gradvzdz;
    } else {
        div_a = -termnorm * (D(cd_axx) + D(cd_ayy) + D(cd_azz));
        dvxdx = -termnorm * D(cd_vxx);
        dvxdy = -termnorm * D(cd_vxy);
        dvxdz = -termnorm * D(cd_vxz);
        dvydx = -termnorm * D(cd_vyx);
        dvydy = -termnorm * D(cd_vyy);
        dvydz = -termnorm * D(cd_vyz);
        dvzdx = -termnorm * D(cd_vzx);
        dvzdy = -termnorm * D(cd_vzy);
        dvzdz = -termnorm * D(cd_vzz);
    }

    double divvdt = div_a - (dvxdx * dvxdx + dvydy * dvydy + dvzdz * dvzdz + 2.0 * (dvxdy * dvydx + dvxdz * dvzdx + dvydz * dvzdy));
    double divv = dvxdx + dvydy + dvzdz;
    double curlvx = dvzdy - dvydz;
    double curlvy = dvxdz - dvzdx;
    double curlvz = dvydx - dvxdy;
    double compression = dmax(-divv, 0.0);
    double fac = compression * compression;
    double trace_s = curlvx * curlvx + curlvy * curlvy + curlvz * curlvz;
    double denominator = fac + trace_s;
    double xilimiter = denominator > 0.0 ? fac / denominator : 1.0;

    D(divvdt) = divvdt;
    D(xilimiter) = xilimiter;
    double source = 10.0 * D(h) * D(h) * D(xilimiter) * dmax(-D(divvdt), 0.0);
    double temp = D(cs) * D(cs);
    double alphaloc = dmax(dmin(source / temp, 1.0), 0.0);
    D(alphaloc) = dmax(P->alpha_min, P->alpha_max * alphaloc);
}

/* StrainDiagnostics.initialize, equations.py:434-471 */
static void strain_diagnostics(orc_t* o, int64_t d) {
    double rxx = D(cd_rxx), rxy = D(cd_rxy), rxz = D(cd_rxz);
    double ryy = D(cd_ryy), ryz = D(cd_ryz), rzz = D(cd_rzz);
    double denom = rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
    if (fabs(denom) > FLOAT_TINY) {
        double ddenom = 1.0 / denom;
        double term_xx = ryy * rzz - ryz * ryz;
        double term_xy = rxz * ryz - rzz * rxy;
        double term_xz = rxy * ryz - rxz * ryy;
        double term_yy = rzz * rxx - rxz * rxz;
        double term_yz = rxy * rxz - rxx * ryz;
        double term_zz = rxx * ryy - rxy * rxy;
        D(dvdx_xx) = -(D(cd_vxx) * term_xx + D(cd_vxy) * term_xy + D(cd_vxz) * term_xz) * ddenom;
        D(dvdx_xy) = -(D(cd_vxx) * term_xy + D(cd_vxy) * term_yy + D(cd_vxz) * term_yz) * ddenom;
        D(dvdx_xz) = -(D(cd_vxx) * term_xz + D(cd_vxy) * term_yz + D(cd_vxz) * term_zz) * ddenom;
        D(dvdx_yx) = -(D(cd_vyx) * term_xx + D(cd_vyy) * term_xy + D(cd_vyz) * term_xz) * ddenom;
        D(dvdx_yy) = -(D(cd_vyx) * term_xy + D(cd_vyy) * term_yy + D(cd_vyz) * term_yz) * ddenom;
        D(dvdx_yz) = -(D(cd_vyx) * term_xz + D(cd_vyy) * term_yz + D(cd_vyz) * term_zz) * ddenom;
        D(dvdx_zx) = -(D(cd_vzx) * term_xx + D(cd_vzy) * term_xy + D(cd_vzz) * term_xz) * ddenom;
        D(dvdx_zy) = -(D(cd_vzx) * term_xy + D(cd_vzy) * term_yy + D(cd_vzz) * term_yz) * ddenom;
        D(dvdx_zz) = -(D(cd_vzx) * term_xz + D(cd_vzy) * term_yz + D(cd_vzz) * term_zz) * ddenom;
    } else {
        double termnorm = D(omega) / D(rho);
        D(dvdx_xx) = -termnorm * D(cd_vxx);
        D(dvdx_xy) = -termnorm * D(cd_vxy);
        D(dvdx_xz) = -termnorm * D(cd_vxz);
        D(dvdx_yx) = -termnorm * D(cd_vyx);
        D(dvdx_yy) = -termnorm * D(cd_vyy);
        D(dvdx_yz) = -termnorm * D(cd_vyz);
        D(dvdx_zx) = -termnorm * D(cd_vzx);
        D(dvdx_zy) = -termnorm * D(cd_vzy);
        D(dvdx_zz) = -termnorm * D(cd_vzz);
    }
}

/* MagneticStateReconstruction.initialize, equations.py:476-485 */
static void b_reconstruction(orc_t* o, int64_t d) {
    double bxpred = D(Bevolxpred) * D(rho);
    double bypred = D(Bevolypred) * D(rho);
    double bzpred = D(Bevolzpred) * D(rho);
    D(Bx) = bxpred;
    D(By) = bypred;
    D(Bz) = bzpred;
    D(Bxpred) = bxpred;
    D(Bypred) = bypred;
    D(Bzpred) = bzpred;
}

/* MagneticStateRates.initialize, equations.py:498-504 */
static void magrates_initialize(orc_t* o, int64_t d) {
    D(aBevolx) = 0.0;
    D(aBevoly) = 0.0;
    D(aBevolz) = 0.0;
    D(apsi) = 0.0;
    D(divBsymm) = 0.0;
    D(divBdiff) = 0.0;
}

/* MagneticStateRates.loop, equations.py:507-559 */
static void magrates_loop(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const sym_t* y) {
    const double mu0 = P->mu0;
    double runix, runiy, runiz;
    if (y->rij > PAIR_DISTANCE_EPSILON) {
        runix = y->xij[0] / y->rij;
        runiy = y->xij[1] / y->rij;
        runiz = y->xij[2] / y->rij;
    } else {
        runix = 0.0; runiy = 0.0; runiz = 0.0;
    }
    double grad_i = D(omega) * (runix * y->dwi[0] + runiy * y->dwi[1] + runiz * y->dwi[2]);
    double grad_j = S(omega) * (runix * y->dwj[0] + runiy * y->dwj[1] + runiz * y->dwj[2]);
    double rho_i = D(rho);
    double rho_j = S(rho);
    double bx_i = D(Bxpred), by_i = D(Bypred), bz_i = D(Bzpred);
    double bx_j = S(Bxpred), by_j = S(Bypred), bz_j = S(Bzpred);
    double proj_bi = bx_i * runix + by_i * runiy + bz_i * runiz;
    double proj_bj = bx_j * runix + by_j * runiy + bz_j * runiz;
    double dbx = bx_i - bx_j;
    double dby = by_i - by_j;
    double dbz = bz_i - bz_j;
    double proj_db = dbx * runix + dby * runiy + dbz * runiz;
    double delta_b2 = dbx * dbx + dby * dby + dbz * dbz;
    double rho_i2_inv = 1.0 / (rho_i * rho_i);
    double rho_j2_inv = 1.0 / (rho_j * rho_j);
    double pmj_rho21_grad_i = S(m) * rho_i2_inv * grad_i;
    double pmj_rho21_grad_j = S(m) * rho_j2_inv * grad_j;
    double d_brho_term = -(pmj_rho21_grad_i * proj_bi);
    double dvx = D(upred) - S(upred);
    double dvy = D(vpred) - S(vpred);
    double dvz = D(wpred) - S(wpred);
    double projv = dvx * runix + dvy * runiy + dvz * runiz;
    double dvxt = dvx - projv * runix;
    double dvyt = dvy - projv * runiy;
    double dvzt = dvz - projv * runiz;
    double vsig_b = sqrt(dvxt * dvxt + dvyt * dvyt + dvzt * dvzt);
    double d_b_diss_term = 0.5 * (S(m) * rho_i2_inv * grad_i + S(m) * rho_j2_inv * grad_j) * vsig_b * P->artificial_magnetic_dissipation;
    double vwave_i = sqrt(D(cs) * D(cs) + (bx_i * bx_i + by_i * by_i + bz_i * bz_i) / (mu0 * rho_i));
    double vwave_j = sqrt(S(cs) * S(cs) + (bx_j * bx_j + by_j * by_j + bz_j * bz_j) / (mu0 * rho_j));
    double ch_i = P->cleaning_speed_factor * vwave_i;
    double ch_j = P->cleaning_speed_factor * vwave_j;
    double dpsi_term = pmj_rho21_grad_i * D(psipred) * ch_i + pmj_rho21_grad_j * S(psipred) * ch_j;

    D(aBevolx) += d_brho_term * dvx + d_b_diss_term * dbx - dpsi_term * runix;
    D(aBevoly) += d_brho_term * dvy + d_b_diss_term * dby - dpsi_term * runiy;
    D(aBevolz) += d_brho_term * dvz + d_b_diss_term * dbz - dpsi_term * runiz;
    D(ae) += -0.5 * delta_b2 * d_b_diss_term / mu0;
    D(divBsymm) += pmj_rho21_grad_i * proj_bi + pmj_rho21_grad_j * proj_bj;
    D(divBdiff) += -S(m) * proj_db * grad_i;
}

/* MagneticStateRates.post_loop, equations.py:562-567 */
static void magrates_post_loop(orc_t* o, const sphmhd_params* P, int64_t d) {
    const double mu0 = P->mu0;
    double cleaning_active = P->cleaning_speed_factor > 0.0 ? 1.0 : 0.0;
    double vwave = sqrt(D(cs) * D(cs) + (D(Bxpred) * D(Bxpred) + D(Bypred) * D(Bypred) + D(Bzpred) * D(Bzpred)) / (mu0 * D(rho)));
    double ch = P->cleaning_speed_factor * vwave;
    D(apsi) = -ch * D(divBdiff) / D(rho) - P->cleaning_damping_factor * D(psipred) * ch / D(h) - 0.5 * cleaning_active * D(psipred) * D(div_for_psi);
    D(dt_cfl) = dmax(D(dt_cfl), vwave);
    D(dt_cfl) = dmax(D(dt_cfl), ch);
}

/* AdaptiveTimestep.initialize, equations.py:578-585 */
static void adaptive_timestep(orc_t* o, const sphmhd_params* P, int64_t d) {
    double c_cour = P->timestep_factor * COURANT_FACTOR;
    double c_force = P->timestep_factor * FORCE_FACTOR;
    double dt_adapt = 1.0e29;
    if (D(dt_cfl) > 0.0) dt_adapt = c_cour * D(h) / D(dt_cfl);
    double f2 = D(au) * D(au) + D(av) * D(av) + D(aw) * D(aw);
    if (f2 > 0.0) dt_adapt = dmin(dt_adapt, c_force * sqrt(D(h) / sqrt(f2)));
    D(dt_adapt) = dt_adapt;
}

/* EnergyLimiter.initialize, equations.py:594-598 */
static void energy_limiter(orc_t* o, const sphmhd_params* P, int64_t d) {
    double c_cour = P->timestep_factor * COURANT_FACTOR;
    if (D(dt_cfl) > 0.0) {
        double dtc = c_cour * D(h) / D(dt_cfl);
        if (D(epred) + dtc * D(ae) < 0.0 && D(epred) > 0.0) D(ae) /= 1.0 - dtc * D(ae) / D(epred);
    }
}

/* DivBCorrection.initialize, equations.py:607-624 */
static void divb_correction(orc_t* o, const sphmhd_params* P, int64_t d) {
    const double mu0 = P->mu0;
    double bx = D(Bxpred), by = D(Bypred), bz = D(Bzpred);
    double b2 = bx * bx + by * by + bz * bz;
    double divb = D(divBsymm);
    double frac_divb = 0.0;
    if (b2 > 0.0) {
        double beta = 2.0 * mu0 * D(p) / b2;
        frac_divb = dmax(0.0, dmin(1.0, (10.0 - beta) * 0.125));
    }
    double ax = -bx * divb * frac_divb / mu0;
    double ay = -by * divb * frac_divb / mu0;
    double az = -bz * divb * frac_divb / mu0;
    D(au) += ax;
    D(av) += ay;
    D(aw) += az;
    D(divBsymm) = D(rho) * divb;
}

/* ------------------------------------------------------------------ */
/* Integrator (integrator.py:12-82), literal                           */
/* ------------------------------------------------------------------ */

static void step_initialize(orc_t* o, int64_t d) {
    D(x0) = D(x);
    D(y0) = D(y);
    D(z0) = D(z);
    D(u0) = D(u);
    D(v0) = D(v);
    D(w0) = D(w);
    D(upred) = D(u);
    D(vpred) = D(v);
    D(wpred) = D(w);
    D(e0) = D(e);
    D(epred) = D(e);
    D(h0) = D(h);
    D(converged) = 0;
    D(omega) = 1.0;
    D(alpha10) = D(alpha1);
    D(Bevolx0) = D(Bevolx);
    D(Bevoly0) = D(Bevoly);
    D(Bevolz0) = D(Bevolz);
    D(psi0) = D(psi);
    D(Bevolxpred) = D(Bevolx);
    D(Bevolypred) = D(Bevoly);
    D(Bevolzpred) = D(Bevolz);
    D(psipred) = D(psi);
}

/* dim == 2: positions are 2-D (z is neither integrated nor wrapped) while
 * v and B keep three components; the reference itself is 3-D only. */
static void step_stage1(orc_t* o, int dim, int64_t d, double dt) {
    if (o->itype[d] == 1) {
        double dtb2 = 0.5 * dt;
        D(u) = D(u0) + dtb2 * D(au);
        D(v) = D(v0) + dtb2 * D(av);
        D(w) = D(w0) + dtb2 * D(aw);
        D(upred) = D(u0) + dt * D(au);
        D(vpred) = D(v0) + dt * D(av);
        D(wpred) = D(w0) + dt * D(aw);
        D(x) = D(x0) + dt * D(u);
        D(y) = D(y0) + dt * D(v);
        if (dim == 3) D(z) = D(z0) + dt * D(w);
        D(e) = D(e0) + dtb2 * D(ae);
        D(epred) = D(e0) + dt * D(ae);
        D(h) = D(h0) + dt * D(ah);
        double decay_rate = 0.1 * D(cs) / D(h);
        if (D(alpha10) < D(alphaloc)) D(alpha1) = D(alphaloc);
        else D(alpha1) = (D(alpha10) + dt * D(alphaloc) * decay_rate) / (1.0 + dt * decay_rate);
        D(Bevolx) = D(Bevolx0) + dtb2 * D(aBevolx);
        D(Bevoly) = D(Bevoly0) + dtb2 * D(aBevoly);
        D(Bevolz) = D(Bevolz0) + dtb2 * D(aBevolz);
        D(psi) = D(psi0) + dtb2 * D(apsi);
        D(Bevolxpred) = D(Bevolx0) + dt * D(aBevolx);
        D(Bevolypred) = D(Bevoly0) + dt * D(aBevoly);
        D(Bevolzpred) = D(Bevolz0) + dt * D(aBevolz);
        D(psipred) = D(psi0) + dt * D(apsi);
    }
}

static void step_stage2(orc_t* o, const sphmhd_params* P, int64_t d, double dt) {
    if (o->itype[d] == 1) {
        double gamma1 = P->gamma - 1.0;
        double dtb2 = 0.5 * dt;
        D(u) += dtb2 * D(au);
        D(v) += dtb2 * D(av);
        D(w) += dtb2 * D(aw);
        D(e) += dtb2 * D(ae);
        D(p) = gamma1 * D(rho) * D(e);
        D(Bevolx) += dtb2 * D(aBevolx);
        D(Bevoly) += dtb2 * D(aBevoly);
        D(Bevolz) += dtb2 * D(aBevolz);
        D(psi) += dtb2 * D(apsi);
        D(Bx) = D(Bevolx) * D(rho);
        D(By) = D(Bevoly) * D(rho);
        D(Bz) = D(Bevolz) * D(rho);
    }
}

/* ------------------------------------------------------------------ */
/* Neighbour search                                                    */
/* ------------------------------------------------------------------ */

static int32_t n_threads_req = 0;

int orc_set_threads(int n) {
    n_threads_req = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
    return 0;
}

static double inst_hmax(orc_t* o, int k) {
    double hm = 0.0;
    for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i)
        if (o->p[P_h][i] > hm) hm = o->p[P_h][i];
    return hm;
}

static double inst_hmin(orc_t* o, int k) {
    double hm = 0.0;
    for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i)
        if (i == o->off[k] || o->p[P_h][i] < hm) hm = o->p[P_h][i];
    return hm;
}

/* cells per h_max cell along an axis: k of the cell rule above */
static double cell_split(double hmax, double hmin) {
    if (hmin > 0.0 && hmax / hmin >= 2.0) return fmin(floor(hmax / hmin), 4.0);
    return 1.0;
}

/* margin of an open domain's grid box around the particles' extent
   [mn, mx]: 1% of the extent, at least 1e-9 of its largest coordinate
   (and of 1), so the box is never empty */
static double box_pad(double mn, double mx) {
    return fmax(0.01 * (mx - mn), 1.0e-9 * fmax(fmax(fabs(mn), fabs(mx)), 1.0));
}

/* Levels of h for the 3-D cell size of a single instance: the biased
   exponent and top four mantissa bits of a positive h (16 per octave);
   level_value(l) is the smallest h of level l. */
#define HIST_SLOTS 64
static int h_level(double h) {
    uint64_t b;
    memcpy(&b, &h, sizeof b);
    return (int)(b >> 48);
}
static double level_value(int l) { return ldexp(1.0 + (double)(l & 15) / 16.0, (l >> 4) - 1023); }

/* The cell size among c0 (the h_max rule) and half of R times the upper
   edge (capped at h_max) of each occupied level that minimises the modelled
   scan cost: per particle of level edge u, s = max(1, ceil(R u / c)) cells
   per axis, (2s+1)^3 cells of (c/u)^3 h-volumes holding hfact^3 particles
   each, plus 6 candidates' worth per stencil row ((2s+1)^2), times 1.15
   where the h_max stencil spans more than two cells (the lists then prune
   and bound each particle's reach); ties keep the larger cell.  cnt[j]:
   particles of level l0 + j, j < nl. */
static double cell_model_3d(double R, double hfact, double hm, int l0, int nl, const int64_t* cnt, double c0) {
    const double k3 = hfact * hfact * hfact;
    double best = c0, bcost = 0.0;
    for (int cand = -1; cand < nl; ++cand) {
        if (cand >= 0 && !cnt[cand]) continue;
        double c = c0;
        if (cand >= 0) {
            double u = level_value(l0 + cand + 1);
            c = 0.5 * (R * (u < hm ? u : hm));
        }
        double cost = 0.0;
        for (int j = 0; j < nl; ++j) {
            if (!cnt[j]) continue;
            double v = level_value(l0 + j + 1);
            double u = v < hm ? v : hm;
            double sj = ceil(R * u / c);
            if (sj < 1.0) sj = 1.0;
            double w = 2.0 * sj + 1.0;
            double x = w * c / u;
            cost += (double)cnt[j] * (k3 * (x * x * x) + 6.0 * (w * w));
        }
        if (ceil(R * hm / c) > 2.0) cost = 1.15 * cost;
        if (cand < 0 || cost < bcost || (cost == bcost && c > best)) {
            best = c;
            bcost = cost;
        }
    }
    return best;
}

static void grid_setup(orc_t* o, int k, double hmax, double hmin) {
    const sphmhd_params* P = &o->prm[k];
    grid_t* g = &o->grid[k];
    int64_t nk = o->off[k + 1] - o->off[k];
    int64_t cap = kCellsPerParticle * nk > 1 ? kCellsPerParticle * nk : 1;
    double lo[3] = {0.0, 0.0, 0.0}, L[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < P->dim; ++a) {
        lo[a] = P->lo[a];
        L[a] = P->hi[a] - P->lo[a];
        if (!P->periodic && nk > 0) {
            /* open domain: the particles' box at this evaluation */
            double mn = o->p[P_x + a][o->off[k]], mx = mn;
            for (int64_t i = o->off[k] + 1; i < o->off[k + 1]; ++i) {
                double v = o->p[P_x + a][i];
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
            double pad = box_pad(mn, mx);
            double hi = mx + pad;
            lo[a] = mn - pad;
            L[a] = hi - lo[a];
        }
    }
    double R = orc_kernel_radius(P->kernel);
    /* in 3-D at least 2 cells per R h_max (twice the h split, at most 4) */
    double split = P->dim == 3 ? fmin(2.0 * cell_split(hmax, hmin), 4.0) : cell_split(hmax, hmin);
    double c = R * hmax / split;
    /* a single 3-D instance while h spans less than 64 levels: the model */
    int model = o->ninst == 1 && P->dim == 3 && hmin > 0.0 && hmax >= hmin &&
                h_level(hmax) - h_level(hmin) < HIST_SLOTS;
    if (model) {
        int64_t cnt[HIST_SLOTS];
        int l0 = h_level(hmin), nl = h_level(hmax) - l0 + 1;
        for (int j = 0; j < HIST_SLOTS; ++j) cnt[j] = 0;
        for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) {
            double h = o->p[P_h][i];
            if (h > 0.0) cnt[h_level(h) - l0]++;
        }
        c = cell_model_3d(R, P->hfact, hmax, l0, nl, cnt, c);
        c = fmax(c, R * hmax / 8.0);
        split = R * hmax / c;
    }
    double lmax = 0.0;
    for (int a = 0; a < P->dim; ++a) lmax = dmax(lmax, L[a]);
    if (!(c > 0.0) || !(c <= lmax)) { /* h_max of 0, NaN or above the box: one cell */
        c = lmax;
        split = 1.0;
    }
    for (int guard = 0; guard < 80; ++guard) {
        int64_t prod = 1;
        for (int a = 0; a < 3; ++a) {
            if (a < P->dim) {
                double nf = floor(L[a] / c);
                int na = nf < 1.0 ? 1 : (nf > 1.0e9 ? 1000000000 : (int)nf);
                g->n[a] = na;
                prod *= na;
                if (prod > cap) break;
            } else {
                g->n[a] = 1;
            }
        }
        if (prod <= cap) break;
        /* too many cells: first undo the split, then coarsen */
        if (model) {
            c = 2.0 * c;
        } else if (split > 1.0) {
            split -= 1.0;
            c = R * hmax / split;
        } else {
            c = 2.0 * c;
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (a < P->dim) {
            g->lo[a] = lo[a];
            g->L[a] = L[a];
            g->ceff[a] = g->L[a] / g->n[a];
            g->halfL[a] = 0.5 * g->L[a];
        } else {
            g->lo[a] = 0.0;
            g->L[a] = 0.0;
            g->ceff[a] = 1.0;
            g->halfL[a] = 0.0;
        }
    }
    g->ncell = (int64_t)g->n[0] * g->n[1] * g->n[2];
}

static int cell_coord(const grid_t* g, const sphmhd_params* P, int a, double x) {
    if (a >= P->dim) return 0;
    double f = floor((x - g->lo[a]) / g->ceff[a]);
    if (f < 0.0) return 0;
    if (f > (double)(g->n[a] - 1)) return g->n[a] - 1;
    return (int)f;
}

/* Bin all instances (sorted by global key, stable in storage order). */
static void nnps_build(orc_t* o) {
    int64_t base = 0;
    for (int k = 0; k < o->ninst; ++k) {
        grid_setup(o, k, inst_hmax(o, k), inst_hmin(o, k));
        o->grid[k].base = base;
        base += o->grid[k].ncell;
    }
    if (base > o->cell_cap) {
        free(o->cstart);
        free(o->cend);
        o->cell_cap = base;
        o->cstart = (int64_t*)malloc(sizeof(int64_t) * base);
        o->cend = (int64_t*)malloc(sizeof(int64_t) * base);
    }
    uint32_t* gkey = (uint32_t*)malloc(sizeof(uint32_t) * (o->n > 0 ? o->n : 1));
    for (int k = 0; k < o->ninst; ++k) {
        const sphmhd_params* P = &o->prm[k];
        const grid_t* g = &o->grid[k];
        for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) {
            int ix = cell_coord(g, P, 0, o->p[P_x][i]);
            int iy = cell_coord(g, P, 1, o->p[P_y][i]);
            int iz = cell_coord(g, P, 2, o->p[P_z][i]);
            int64_t local = ix + (int64_t)g->n[0] * (iy + (int64_t)g->n[1] * iz);
            o->cellof[i] = (int32_t)local;
            gkey[i] = (uint32_t)(g->base + local);
        }
    }
    /* stable counting sort by global key */
    for (int64_t c = 0; c < base; ++c) { o->cstart[c] = 0; o->cend[c] = 0; }
    for (int64_t i = 0; i < o->n; ++i) o->cend[gkey[i]]++;
    int64_t run = 0;
    for (int64_t c = 0; c < base; ++c) {
        int64_t cnt = o->cend[c];
        o->cstart[c] = run;
        o->cend[c] = run;
        run += cnt;
    }
    for (int64_t i = 0; i < o->n; ++i) {
        int64_t pos = o->cend[gkey[i]]++;
        o->perm[pos] = (int32_t)i;
        o->key[pos] = gkey[i];
    }
    free(gkey);
}

static int axis_cells(const grid_t* g, int a, int c, int s, int periodic, int* out) {
    int n = g->n[a];
    int cnt = 0;
    if (periodic) {
        if (2 * s + 1 >= n) {
            for (int q = 0; q < n; ++q) out[cnt++] = q;
        } else {
            for (int q = -s; q <= s; ++q) out[cnt++] = ((c + q) % n + n) % n;
        }
    } else {
        int lo = c - s < 0 ? 0 : c - s;
        int hi = c + s > n - 1 ? n - 1 : c + s;
        for (int q = lo; q <= hi; ++q) out[cnt++] = q;
    }
    return cnt;
}

static int stencil_radius(const grid_t* g, const sphmhd_params* P, int a, double hmax_now) {
    if (a >= P->dim) return 0;
    double R = orc_kernel_radius(P->kernel);
    double s = ceil(R * hmax_now / g->ceff[a]);
    if (!(s < 1.0e6)) return 1000000;
    if (s < 1.0) return 1;
    return (int)s;
}

static double min_image(double d, const grid_t* g, int a, int periodic) {
    if (periodic && g->L[a] > 0.0) {
        if (d > g->halfL[a]) d -= g->L[a];
        else if (d < -g->halfL[a]) d += g->L[a];
    }
    return d;
}

/* Visit the neighbours of storage index i (instance k) in traversal order. */
typedef void (*visit_fn)(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const double* xij, double r2, void* ctx);

static void for_neighbours(orc_t* o, int k, int64_t i, const int* srad, visit_fn fn, void* ctx) {
    const sphmhd_params* P = &o->prm[k];
    const grid_t* g = &o->grid[k];
    const double R = orc_kernel_radius(P->kernel);
    int cx = cell_coord(g, P, 0, o->p[P_x][i]);
    int cy = cell_coord(g, P, 1, o->p[P_y][i]);
    int cz = cell_coord(g, P, 2, o->p[P_z][i]);
    int *xs = (int*)malloc(sizeof(int) * (g->n[0] + 2 * srad[0] + 2));
    int *ys = (int*)malloc(sizeof(int) * (g->n[1] + 2 * srad[1] + 2));
    int *zs = (int*)malloc(sizeof(int) * (g->n[2] + 2 * srad[2] + 2));
    int nxs = axis_cells(g, 0, cx, srad[0], P->periodic, xs);
    int nys = axis_cells(g, 1, cy, srad[1], P->periodic, ys);
    int nzs = axis_cells(g, 2, cz, srad[2], P->periodic, zs);
    double xi = o->p[P_x][i], yi = o->p[P_y][i], zi = o->p[P_z][i], hi = o->p[P_h][i];
    for (int a = 0; a < nzs; ++a)
        for (int b = 0; b < nys; ++b)
            for (int c = 0; c < nxs; ++c) {
                int64_t cell = g->base + xs[c] + (int64_t)g->n[0] * (ys[b] + (int64_t)g->n[1] * zs[a]);
                for (int64_t t = o->cstart[cell]; t < o->cend[cell]; ++t) {
                    int64_t j = o->perm[t];
                    double xij[3];
                    xij[0] = min_image(xi - o->p[P_x][j], g, 0, P->periodic);
                    xij[1] = min_image(yi - o->p[P_y][j], g, 1, P->periodic);
                    xij[2] = P->dim == 3 ? min_image(zi - o->p[P_z][j], g, 2, P->periodic) : zi - o->p[P_z][j];
                    double r2 = xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2];
                    double hm = dmax(hi, o->p[P_h][j]);
                    double rad = R * hm;
                    if (r2 < rad * rad) fn(o, P, i, j, xij, r2, ctx);
                }
            }
    free(xs);
    free(ys);
    free(zs);
}

/* ------------------------------------------------------------------ */
/* Group driver (scheme.py:70-81)                                      */
/* ------------------------------------------------------------------ */

enum { PASS_DENSITY, PASS_CD, PASS_MHDACC, PASS_MAGSTRESS, PASS_MAGRATES };

/* Hooks of the operation-counting build: each pair symbol counts apart. */
#ifndef ORC_SYM_PART
#define ORC_SYM_PART(k)
#endif

static void make_sym(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const double* xij, double r2, sym_t* y) {
    y->xij[0] = xij[0];
    y->xij[1] = xij[1];
    y->xij[2] = xij[2];
    y->r2ij = r2;
    ORC_SYM_PART(0);
    y->rij = sqrt(r2);
    ORC_SYM_PART(1);
    y->rhoij = 0.5 * (D(rho) + S(rho));
    ORC_SYM_PART(2);
    y->wi = orc_kernel_w(P->kernel, P->dim, y->rij, D(h));
    ORC_SYM_PART(3);
    orc_kernel_grad(P->kernel, P->dim, y->xij, y->rij, D(h), y->dwi);
    ORC_SYM_PART(4);
    orc_kernel_grad(P->kernel, P->dim, y->xij, y->rij, S(h), y->dwj);
    ORC_SYM_PART(5);
    y->ghi = orc_kernel_gradh(P->kernel, P->dim, y->rij, D(h));
}

/* Hooks of the operation-counting build (flopcount.cpp): a pair's symbols
   (kernel values and gradients) and its pass's body are counted apart;
   empty in the plain build. */
#ifndef ORC_PAIR_BEGIN
#define ORC_PAIR_BEGIN(pass)
#define ORC_PAIR_BODY(pass)
#define ORC_PAIR_END()
#endif

static void visit_pass(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const double* xij, double r2, void* ctx) {
    int pass = *(int*)ctx;
    sym_t y;
    ORC_PAIR_BEGIN(pass);
    make_sym(o, P, d, s, xij, r2, &y);
    ORC_PAIR_BODY(pass);
    switch (pass) {
    case PASS_DENSITY: density_loop(o, d, s, &y); break;
    case PASS_CD: cd_loop(o, d, s, &y); break;
    case PASS_MHDACC: mhdacc_loop(o, P, d, s, &y); break;
    case PASS_MAGSTRESS: magstress_loop(o, P, d, s, &y); break;
    case PASS_MAGRATES: magrates_loop(o, P, d, s, &y); break;
    }
    ORC_PAIR_END();
}

static void pair_pass(orc_t* o, int k, int pass) {
    const sphmhd_params* P = &o->prm[k];
    const grid_t* g = &o->grid[k];
    double hm = inst_hmax(o, k);
    int srad[3];
    for (int a = 0; a < 3; ++a) srad[a] = stencil_radius(g, P, a, hm);
    int64_t a0 = o->off[k], a1 = o->off[k + 1];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = a0; i < a1; ++i) {
        int ps = pass;
        for_neighbours(o, k, i, srad, visit_pass, &ps);
    }
}

static int eval_instance(orc_t* o, int k) {
    const sphmhd_params* P = &o->prm[k];
    int64_t a0 = o->off[k], a1 = o->off[k + 1];
    int status = 0;

    /* G0: DensityIteration, iterated (scheme.py:71) */
    int iters = 0;
    for (;;) {
        for (int64_t i = a0; i < a1; ++i) density_initialize(o, i);
        pair_pass(o, k, PASS_DENSITY);
        int conv = 1;
#pragma omp parallel for reduction(min : conv)
        for (int64_t i = a0; i < a1; ++i) {
            int f = density_post_loop(o, P, i);
            if (f < conv) conv = f;
        }
        iters++;
        if (!P->density_iterate) break;
        if (iters >= P->min_iterations && conv > 0) break;
        if (iters >= P->max_iterations) { status = SPHMHD_WNOCONV; break; }
    }
    o->last_iters[k] = iters;

    for (int64_t i = a0; i < a1; ++i) {
        ah_from_rho(o, P, i);   /* G1 */
        rho_from_h(o, P, i);
        ideal_gas_eos(o, P, i); /* G2 */
    }
    /* G3: Cullen-Dehnen switch (reads the previous evaluation's au, av, aw) */
    for (int64_t i = a0; i < a1; ++i) cd_initialize(o, i);
    pair_pass(o, k, PASS_CD);
    for (int64_t i = a0; i < a1; ++i) cd_post_loop(o, P, i);
    /* G4 */
    for (int64_t i = a0; i < a1; ++i) {
        strain_diagnostics(o, i);
        b_reconstruction(o, i);
    }
    /* G5 */
    for (int64_t i = a0; i < a1; ++i) mhdacc_initialize(o, i);
    pair_pass(o, k, PASS_MHDACC);
    for (int64_t i = a0; i < a1; ++i) mhdacc_post_loop(o, i);
    /* G6 */
    for (int64_t i = a0; i < a1; ++i) ah_from_div(o, P, i);
    /* G7 */
    pair_pass(o, k, PASS_MAGSTRESS);
    /* G8 */
    for (int64_t i = a0; i < a1; ++i) magrates_initialize(o, i);
    pair_pass(o, k, PASS_MAGRATES);
    for (int64_t i = a0; i < a1; ++i) magrates_post_loop(o, P, i);
    /* G9 */
    double dtmin = 1.0e300;
    for (int64_t i = a0; i < a1; ++i) {
        energy_limiter(o, P, i);
        divb_correction(o, P, i);
        adaptive_timestep(o, P, i);
        if (o->p[P_dt_adapt][i] < dtmin) dtmin = o->p[P_dt_adapt][i];
        double a2 = o->p[P_au][i] * o->p[P_au][i] + o->p[P_av][i] * o->p[P_av][i] + o->p[P_aw][i] * o->p[P_aw][i];
        if (!isfinite(a2) && status == 0) status = SPHMHD_WNAN;
    }
    o->dt_pending[k] = dtmin;
    return status;
}

int orc_eval(orc_t* o) {
    nnps_build(o);
    int status = 0;
    for (int k = 0; k < o->ninst; ++k) {
        int s = eval_instance(o, k);
        if (s > status) status = s;
    }
    return status;
}

/* ------------------------------------------------------------------ */
/* PEC stages, clock, warm start                                       */
/* ------------------------------------------------------------------ */

static void wrap_positions(orc_t* o, int k) {
    const sphmhd_params* P = &o->prm[k];
    if (!P->periodic) return;
    for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) {
        for (int a = 0; a < P->dim; ++a) {
            int prop = a == 0 ? P_x : (a == 1 ? P_y : P_z);
            double L = P->hi[a] - P->lo[a];
            double x = o->p[prop][i];
            if (x < P->lo[a]) x += L;
            else if (x >= P->hi[a]) x -= L;
            o->p[prop][i] = x;
        }
    }
}

int orc_initialize(orc_t* o) {
    for (int64_t i = 0; i < o->n; ++i) step_initialize(o, i);
    return 0;
}

int orc_stage1(orc_t* o) {
    for (int k = 0; k < o->ninst; ++k) {
        for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) step_stage1(o, o->prm[k].dim, i, o->dt[k]);
        wrap_positions(o, k);
    }
    return 0;
}

int orc_stage2(orc_t* o) {
    for (int k = 0; k < o->ninst; ++k) {
        for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) step_stage2(o, &o->prm[k], i, o->dt[k]);
        wrap_positions(o, k);
    }
    return 0;
}

/* solver.py:57-61 with the adaptive rule of Appendix A7 */
static double next_dt(const sphmhd_params* P, double t, double dt_now, double pending) {
    double dt = P->adaptive_timestep ? pending : dt_now;
    if (t + dt > P->tf) dt = P->tf - t;
    return dt;
}

int orc_advance_clock(orc_t* o) {
    for (int k = 0; k < o->ninst; ++k) {
        o->t[k] += o->dt[k];
        o->dt[k] = next_dt(&o->prm[k], o->t[k], o->dt[k], o->dt_pending[k]);
    }
    return 0;
}

/* Up to nsteps PEC steps; the run ends after the step that brings every
 * instance's clock to its tf (t + dt >= tf), so a tf-bounded run takes no
 * zero-length steps (same contract as sphmhd_step). */
int orc_step(orc_t* o, int64_t nsteps) {
    int status = 0;
    for (int64_t s = 0; s < nsteps; ++s) {
        orc_initialize(o);
        orc_stage1(o);
        int st = orc_eval(o);
        if (st > status) status = st;
        int at_tf = 1;
        for (int k = 0; k < o->ninst; ++k)
            if (!(o->t[k] + o->dt[k] >= o->prm[k].tf)) at_tf = 0;
        orc_stage2(o);
        orc_advance_clock(o);
        o->steps += 1;
        if (at_tf) break;
    }
    return status;
}

int64_t orc_steps_taken(orc_t* o) { return o->steps; }

/* application.py:100-108 */
int orc_sync_initial_state(orc_t* o) {
    for (int64_t d = 0; d < o->n; ++d) {
        D(Bevolx) = D(Bx) / D(rho);
        D(Bevoly) = D(By) / D(rho);
        D(Bevolz) = D(Bz) / D(rho);
        D(psi) = 0.0;
        D(Bevolxpred) = D(Bevolx);
        D(Bevolypred) = D(Bevoly);
        D(Bevolzpred) = D(Bevolz);
        D(psipred) = D(psi);
        D(upred) = D(u);
        D(vpred) = D(v);
        D(wpred) = D(w);
        D(epred) = D(e);
    }
    return 0;
}

static int all_converged(orc_t* o) {
    for (int64_t i = 0; i < o->n; ++i)
        if (o->p[P_converged][i] != 1.0) return 0;
    return 1;
}

/* application.py:110-147 (no per-case thermodynamics refresh) */
int orc_warm_start(orc_t* o) {
    int64_t n = o->n;
    double* b0 = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    memcpy(b0, o->p[P_Bx], sizeof(double) * n);
    memcpy(b0 + n, o->p[P_By], sizeof(double) * n);
    memcpy(b0 + 2 * n, o->p[P_Bz], sizeof(double) * n);
    memcpy(o->p[P_h0], o->p[P_h], sizeof(double) * n);
    orc_sync_initial_state(o);
    orc_eval(o);
    if (!all_converged(o)) { free(b0); return SPHMHD_ENOTCONV; }
    memcpy(o->p[P_Bx], b0, sizeof(double) * n);
    memcpy(o->p[P_By], b0 + n, sizeof(double) * n);
    memcpy(o->p[P_Bz], b0 + 2 * n, sizeof(double) * n);
    free(b0);
    memcpy(o->p[P_h0], o->p[P_h], sizeof(double) * n);
    orc_sync_initial_state(o);
    for (int r = 0; r < 2; ++r) orc_eval(o);
    if (!all_converged(o)) return SPHMHD_ENOTCONV;
    /* _compute_initial_dt, application.py:134-147 */
    for (int k = 0; k < o->ninst; ++k) {
        const sphmhd_params* P = &o->prm[k];
        double dtc = 1.0e300, dtf = 1.0e300;
        int any_force = 0;
        for (int64_t d = o->off[k]; d < o->off[k + 1]; ++d) {
            if (!(D(dt_cfl) > 0.0)) return SPHMHD_EINVAL;
            double c = P->timestep_factor * COURANT_FACTOR * D(h) / D(dt_cfl);
            if (c < dtc) dtc = c;
            double fn = sqrt(D(au) * D(au) + D(av) * D(av) + D(aw) * D(aw));
            if (fn > 0.0) {
                any_force = 1;
                double f = P->timestep_factor * FORCE_FACTOR * sqrt(D(h) / fn);
                if (f < dtf) dtf = f;
            }
        }
        double dt = any_force ? dmin(dtc, dtf) : dtc;
        if (!(dt > 0.0)) return SPHMHD_EINVAL;
        /* MHDSolver.solve then re-reads the step (solver.py:40) */
        o->dt[k] = next_dt(P, o->t[k], dt, o->dt_pending[k]);
    }
    return 0;
}

int orc_get_clock(orc_t* o, double* t, double* dt) {
    for (int k = 0; k < o->ninst; ++k) { t[k] = o->t[k]; dt[k] = o->dt[k]; }
    return 0;
}

int orc_set_clock(orc_t* o, const double* t, const double* dt) {
    for (int k = 0; k < o->ninst; ++k) { o->t[k] = t[k]; o->dt[k] = dt[k]; }
    return 0;
}

int orc_get_pending_dt(orc_t* o, double* dt) {
    for (int k = 0; k < o->ninst; ++k) dt[k] = o->dt_pending[k];
    return 0;
}

int orc_last_iterations(orc_t* o, int32_t* iters) {
    for (int k = 0; k < o->ninst; ++k) iters[k] = o->last_iters[k];
    return 0;
}

int orc_totals(orc_t* o, double* out) {
    for (int k = 0; k < o->ninst; ++k) {
        const sphmhd_params* P = &o->prm[k];
        double* T = out + (int64_t)k * SPHMHD_NTOTALS;
        for (int q = 0; q < SPHMHD_NTOTALS; ++q) T[q] = 0.0;
        for (int64_t d = o->off[k]; d < o->off[k + 1]; ++d) {
            double m = D(m), rho = D(rho);
            double v2 = D(u) * D(u) + D(v) * D(v) + D(w) * D(w);
            double b2 = D(Bx) * D(Bx) + D(By) * D(By) + D(Bz) * D(Bz);
            T[SPHMHD_T_MASS] += m;
            T[SPHMHD_T_MOMX] += m * D(u);
            T[SPHMHD_T_MOMY] += m * D(v);
            T[SPHMHD_T_MOMZ] += m * D(w);
            T[SPHMHD_T_EKIN] += 0.5 * m * v2;
            T[SPHMHD_T_ETHERM] += m * D(e);
            T[SPHMHD_T_EMAG] += 0.5 * m * b2 / (P->mu0 * rho);
            if (b2 > 0.0) T[SPHMHD_T_DIVB] += (m / rho) * D(h) * fabs(D(divBsymm)) / sqrt(b2);
            T[SPHMHD_T_VOLUME] += m / rho;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Exports                                                             */
/* ------------------------------------------------------------------ */

int orc_cells(orc_t* o, uint32_t* keys, int32_t* perm) {
    nnps_build(o);
    memcpy(keys, o->key, sizeof(uint32_t) * o->n);
    memcpy(perm, o->perm, sizeof(int32_t) * o->n);
    return 0;
}

typedef struct { int32_t* buf; int64_t cnt; int64_t cap; } nbuf_t;

static void visit_collect(orc_t* o, const sphmhd_params* P, int64_t d, int64_t s, const double* xij, double r2, void* ctx) {
    (void)o; (void)P; (void)d; (void)xij; (void)r2;
    nbuf_t* b = (nbuf_t*)ctx;
    if (b->cnt == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 64;
        b->buf = (int32_t*)realloc(b->buf, sizeof(int32_t) * b->cap);
    }
    b->buf[b->cnt++] = (int32_t)s;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

int orc_neighbours(orc_t* o, int64_t* offsets, int32_t* idx, int64_t cap, int64_t* total) {
    nnps_build(o);
    int64_t pos = 0;
    nbuf_t b = {NULL, 0, 0};
    for (int k = 0; k < o->ninst; ++k) {
        const sphmhd_params* P = &o->prm[k];
        const grid_t* g = &o->grid[k];
        double hm = inst_hmax(o, k);
        int srad[3];
        for (int a = 0; a < 3; ++a) srad[a] = stencil_radius(g, P, a, hm);
        for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) {
            b.cnt = 0;
            for_neighbours(o, k, i, srad, visit_collect, &b);
            qsort(b.buf, b.cnt, sizeof(int32_t), cmp_i32);
            offsets[i] = pos;
            for (int64_t q = 0; q < b.cnt; ++q) {
                if (pos < cap) idx[pos] = b.buf[q];
                pos++;
            }
        }
    }
    offsets[o->n] = pos;
    free(b.buf);
    *total = pos;
    return pos > cap ? SPHMHD_ECAP : 0;
}

int orc_neighbours_brute(orc_t* o, int64_t* offsets, int32_t* idx, int64_t cap, int64_t* total) {
    int64_t pos = 0;
    for (int k = 0; k < o->ninst; ++k) {
        const sphmhd_params* P = &o->prm[k];
        double R = orc_kernel_radius(P->kernel);
        for (int64_t i = o->off[k]; i < o->off[k + 1]; ++i) {
            offsets[i] = pos;
            for (int64_t j = o->off[k]; j < o->off[k + 1]; ++j) {
                double r2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    int prop = a == 0 ? P_x : (a == 1 ? P_y : P_z);
                    double dd = o->p[prop][i] - o->p[prop][j];
                    if (a < P->dim && P->periodic) {
                        double L = P->hi[a] - P->lo[a];
                        if (dd > 0.5 * L) dd -= L;
                        else if (dd < -(0.5 * L)) dd += L;
                    }
                    r2 = r2 + dd * dd;
                }
                double rad = R * dmax(o->p[P_h][i], o->p[P_h][j]);
                if (r2 < rad * rad) {
                    if (pos < cap) idx[pos] = (int32_t)j;
                    pos++;
                }
            }
        }
    }
    offsets[o->n] = pos;
    *total = pos;
    return pos > cap ? SPHMHD_ECAP : 0;
}

/* ------------------------------------------------------------------ */
/* Single-method calls (for the golden vectors of the reference bodies) */
/* ------------------------------------------------------------------ */

int orc_call(orc_t* o, int body, int64_t d, int64_t s, const double* v) {
    const sphmhd_params* P = &o->prm[0];
    sym_t y;
    y.xij[0] = v[0]; y.xij[1] = v[1]; y.xij[2] = v[2];
    y.rij = v[3]; y.r2ij = v[4]; y.rhoij = v[5]; y.wi = v[6];
    y.dwi[0] = v[7]; y.dwi[1] = v[8]; y.dwi[2] = v[9];
    y.dwj[0] = v[10]; y.dwj[1] = v[11]; y.dwj[2] = v[12];
    y.ghi = v[13];
    double dt = v[14];
    switch (body) {
    case ORC_DENSITY_INIT: density_initialize(o, d); return 0;
    case ORC_DENSITY_LOOP: density_loop(o, d, s, &y); return 0;
    case ORC_DENSITY_POST: return density_post_loop(o, P, d);
    case ORC_EOS: ideal_gas_eos(o, P, d); return 0;
    case ORC_RHO_FROM_H: rho_from_h(o, P, d); return 0;
    case ORC_AH_FROM_RHO: ah_from_rho(o, P, d); return 0;
    case ORC_AH_FROM_DIV: ah_from_div(o, P, d); return 0;
    case ORC_MAGSTRESS_LOOP: magstress_loop(o, P, d, s, &y); return 0;
    case ORC_MHDACC_INIT: mhdacc_initialize(o, d); return 0;
    case ORC_MHDACC_LOOP: mhdacc_loop(o, P, d, s, &y); return 0;
    case ORC_MHDACC_POST: mhdacc_post_loop(o, d); return 0;
    case ORC_CD_INIT: cd_initialize(o, d); return 0;
    case ORC_CD_LOOP: cd_loop(o, d, s, &y); return 0;
    case ORC_CD_POST: cd_post_loop(o, P, d); return 0;
    case ORC_STRAIN: strain_diagnostics(o, d); return 0;
    case ORC_BRECON: b_reconstruction(o, d); return 0;
    case ORC_MAGRATES_INIT: magrates_initialize(o, d); return 0;
    case ORC_MAGRATES_LOOP: magrates_loop(o, P, d, s, &y); return 0;
    case ORC_MAGRATES_POST: magrates_post_loop(o, P, d); return 0;
    case ORC_ADAPTIVE_DT: adaptive_timestep(o, P, d); return 0;
    case ORC_ENERGY_LIMITER: energy_limiter(o, P, d); return 0;
    case ORC_DIVB_CORRECTION: divb_correction(o, P, d); return 0;
    case ORC_STEP_INITIALIZE: step_initialize(o, d); return 0;
    case ORC_STEP_STAGE1: step_stage1(o, P->dim, d, dt); return 0;
    case ORC_STEP_STAGE2: step_stage2(o, P, d, dt); return 0;
    }
    return SPHMHD_EINVAL;
}

/* ------------------------------------------------------------------ */
/* Lifetime                                                            */
/* ------------------------------------------------------------------ */

int orc_abi_version(void) { return SPHMHD_ABI_VERSION; }
int orc_nprops(void) { return NPROP; }
const char* orc_prop_name(int k) { return (k >= 0 && k < NPROP) ? PROP_NAMES[k] : NULL; }

orc_t* orc_create(const sphmhd_params* params, int32_t n_instances, const int64_t* offsets) {
    if (n_instances < 1 || offsets[0] != 0) return NULL;
    for (int k = 0; k < n_instances; ++k) {
        if (offsets[k + 1] < offsets[k]) return NULL;
        if (params[k].dim != 2 && params[k].dim != 3) return NULL;
    }
    orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
    o->ninst = n_instances;
    o->n = offsets[n_instances];
    int64_t n = o->n > 0 ? o->n : 1;
    o->off = (int64_t*)malloc(sizeof(int64_t) * (n_instances + 1));
    memcpy(o->off, offsets, sizeof(int64_t) * (n_instances + 1));
    o->prm = (sphmhd_params*)malloc(sizeof(sphmhd_params) * n_instances);
    memcpy(o->prm, params, sizeof(sphmhd_params) * n_instances);
    for (int q = 0; q < NPROP; ++q) o->p[q] = (double*)calloc(n, sizeof(double));
    o->itype = (int32_t*)malloc(sizeof(int32_t) * n);
    for (int64_t i = 0; i < n; ++i) o->itype[i] = 1; /* particles.py:7 default */
    o->t = (double*)calloc(n_instances, sizeof(double));
    o->dt = (double*)calloc(n_instances, sizeof(double));
    o->dt_pending = (double*)calloc(n_instances, sizeof(double));
    o->last_iters = (int32_t*)calloc(n_instances, sizeof(int32_t));
    o->grid = (grid_t*)calloc(n_instances, sizeof(grid_t));
    o->key = (uint32_t*)malloc(sizeof(uint32_t) * n);
    o->perm = (int32_t*)malloc(sizeof(int32_t) * n);
    o->cellof = (int32_t*)malloc(sizeof(int32_t) * n);
    return o;
}

void orc_destroy(orc_t* o) {
    if (!o) return;
    for (int q = 0; q < NPROP; ++q) free(o->p[q]);
    free(o->itype);
    free(o->off);
    free(o->prm);
    free(o->t);
    free(o->dt);
    free(o->dt_pending);
    free(o->last_iters);
    free(o->grid);
    free(o->key);
    free(o->perm);
    free(o->cellof);
    free(o->cstart);
    free(o->cend);
    free(o);
}

double* orc_prop(orc_t* o, const char* name) {
    for (int q = 0; q < NPROP; ++q)
        if (strcmp(PROP_NAMES[q], name) == 0) return o->p[q];
    return NULL;
}

int32_t* orc_itype(orc_t* o) { return o->itype; }
int64_t orc_size(orc_t* o) { return o->n; }
