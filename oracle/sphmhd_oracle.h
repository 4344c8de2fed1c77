/*
 * sphmhd_oracle.h -- FP64 CPU restatement of the reference SPH-MHD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product never does.
 *
 * It restates, one function per reference method:
 *   src/equations.py:14-624   (the 14 Equation classes)
 *   src/integrator.py:6-82    (MHDStep)
 *   src/scheme.py:67-81       (the 10-group plan and its iteration limits)
 *   src/solver.py:42-61       (step loop and time-step update)
 *   src/application.py:99-147 (warm start and initial dt)
 * plus the PySPH semantics the reference relies on (kernels, neighbour
 * search, group iteration), defined as documented in sphmhd_oracle.c.
 */
#ifndef SPHMHD_ORACLE_H
#define SPHMHD_ORACLE_H

#include <stdint.h>
#include "../include/sphmhd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* Body identifiers for orc_call (one per reference method). */
enum {
    ORC_DENSITY_INIT = 0, ORC_DENSITY_LOOP, ORC_DENSITY_POST,
    ORC_EOS,
    ORC_RHO_FROM_H,
    ORC_AH_FROM_RHO,
    ORC_AH_FROM_DIV,
    ORC_MAGSTRESS_LOOP,
    ORC_MHDACC_INIT, ORC_MHDACC_LOOP, ORC_MHDACC_POST,
    ORC_CD_INIT, ORC_CD_LOOP, ORC_CD_POST,
    ORC_STRAIN,
    ORC_BRECON,
    ORC_MAGRATES_INIT, ORC_MAGRATES_LOOP, ORC_MAGRATES_POST,
    ORC_ADAPTIVE_DT,
    ORC_ENERGY_LIMITER,
    ORC_DIVB_CORRECTION,
    ORC_STEP_INITIALIZE, ORC_STEP_STAGE1, ORC_STEP_STAGE2,
    ORC_NBODIES
};

int orc_abi_version(void);
int orc_nprops(void);
const char* orc_prop_name(int k);

orc_t* orc_create(const sphmhd_params* params, int32_t n_instances, const int64_t* offsets);
void orc_destroy(orc_t* o);
int orc_set_threads(int n);

/* Property storage (all reference properties, particles.py:1-162). */
double* orc_prop(orc_t* o, const char* name);
int32_t* orc_itype(orc_t* o);
int64_t orc_size(orc_t* o);

/* Group plan G0..G9 (scheme.py:70-81) for every instance. */
int orc_eval(orc_t* o);
int orc_initialize(orc_t* o);
int orc_stage1(orc_t* o);
int orc_stage2(orc_t* o);
int orc_advance_clock(orc_t* o);
int orc_step(orc_t* o, int64_t nsteps);
int64_t orc_steps_taken(orc_t* o);
int orc_sync_initial_state(orc_t* o);
int orc_warm_start(orc_t* o);
int orc_get_clock(orc_t* o, double* t, double* dt);
int orc_set_clock(orc_t* o, const double* t, const double* dt);
int orc_get_pending_dt(orc_t* o, double* dt);
int orc_totals(orc_t* o, double* out);
int orc_last_iterations(orc_t* o, int32_t* iters);

/* Neighbour search exports (same contract as sphmhd_cells / sphmhd_neighbours). */
int orc_cells(orc_t* o, uint32_t* keys, int32_t* perm);
int orc_neighbours(orc_t* o, int64_t* offsets, int32_t* idx, int64_t cap, int64_t* total);
/* O(N^2) enumeration of the same radius test, for pinning orc_neighbours. */
int orc_neighbours_brute(orc_t* o, int64_t* offsets, int32_t* idx, int64_t cap, int64_t* total);

/* Kernel functions of the PySPH kernels (W, dW/dr . x/r, dW/dh). */
double orc_kernel_w(int kernel, int dim, double r, double h);
double orc_kernel_dwdq(int kernel, int dim, double r, double h);
void orc_kernel_grad(int kernel, int dim, const double* xij, double r, double h, double* grad);
double orc_kernel_gradh(int kernel, int dim, double r, double h);
double orc_kernel_radius(int kernel);

/* Call one reference method on particles d (and s) of instance 0.
 * sym = {XIJ[3], RIJ, R2IJ, RHOIJ, WI, DWI[3], DWJ[3], GHI, dt}.
 * Returns equation_has_converged after DensityIteration.post_loop, else 0. */
int orc_call(orc_t* o, int body, int64_t d, int64_t s, const double* sym);

#ifdef __cplusplus
}
#endif

#endif
