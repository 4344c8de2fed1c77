/*
 * sphmhd.h -- C ABI of the MI355X (gfx950) SPH-MHD step library (libsphmhd.so).
 *
 * The library replaces the PySPH evaluation of the reference's equation-group
 * plan (reference: src/scheme.py:67-81) and its PEC integrator stages
 * (src/integrator.py:6-82) with hand-written HIP kernels.  Every entry point
 * below names the reference interface it stands in for.
 *
 * Conventions
 *   - Every function returns 0 on success, a negative code on error and a
 *     positive code for a warning (see sphmhd_status).  No C++ exception ever
 *     crosses this boundary; sphmhd_last_error() gives the message of the
 *     most recent non-zero return on the calling thread.
 *   - Field pointers are DEVICE pointers to FP64 arrays in particle storage
 *     order.  They are borrowed: the caller keeps them alive while bound.
 *   - One host thread per context.  All device work is ordered on the stream
 *     given to sphmhd_create (NULL = the default stream).
 *   - A context holds one or more independent instances (an ensemble).  The
 *     particles of instance k occupy storage [offsets[k], offsets[k+1]).  A
 *     plain run is an ensemble of one.
 */
#ifndef SPHMHD_H
#define SPHMHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHMHD_ABI_VERSION 8

typedef enum sphmhd_status {
    SPHMHD_OK = 0,
    SPHMHD_EINVAL = -1,     /* bad argument */
    SPHMHD_EHIP = -2,       /* HIP runtime error */
    SPHMHD_ECAP = -3,       /* a capacity was exceeded */
    SPHMHD_ESTATE = -4,     /* call out of order (e.g. no fields bound) */
    SPHMHD_ENOTCONV = -5,   /* warm start: a particle did not converge
                               (the reference asserts, application.py:120,131) */
    SPHMHD_WNOCONV = 1,     /* density iteration hit max_iterations */
    SPHMHD_WNAN = 2         /* a non-finite acceleration was produced */
} sphmhd_status;

/* SPH kernels, in the order of the reference's name map (scheme.py:48-56). */
typedef enum sphmhd_kernel_id {
    SPHMHD_CUBIC = 0,
    SPHMHD_QUINTIC = 1,
    SPHMHD_WENDLAND_C2 = 2,
    SPHMHD_WENDLAND_C4 = 3,
    SPHMHD_WENDLAND_C6 = 4,
    SPHMHD_GAUSSIAN = 5,
    SPHMHD_SUPER_GAUSSIAN = 6
} sphmhd_kernel_id;

/* Per-instance physics and numerics (config.py:8-54 plus the scheme's
 * constants).  dim, kernel and periodic must agree across the instances of
 * one context; every other member may differ per instance. */
typedef struct sphmhd_params {
    int32_t dim;                /* 2 or 3 (the reference hard-wires 3, scheme.py:57) */
    int32_t kernel;             /* sphmhd_kernel_id */
    int32_t density_iterate;    /* 1: numerics.density == "iterate", 0: "single" */
    int32_t min_iterations;     /* scheme.py:71: 2 when iterating, else 0 */
    int32_t max_iterations;     /* scheme.py:28,71: 250 when iterating, else 1 */
    int32_t periodic;           /* 1: minimum-image periodic box [lo, hi) on every axis */
    int32_t adaptive_timestep;  /* solver.adaptive_timestep */
    int32_t reserved;
    double lo[3];
    double hi[3];
    double gamma;
    double mu0;
    double hfact;                               /* DensityIteration k */
    double artificial_viscosity;
    double artificial_thermal_conductivity;
    double artificial_magnetic_dissipation;
    double cleaning_speed_factor;
    double cleaning_damping_factor;
    double timestep_factor;
    double alpha_min;           /* floor of alpha_loc; 0 reproduces equations.py:428 */
    double alpha_max;           /* multiplier of alpha_loc; 1 reproduces equations.py:428 */
    double tf;                  /* final time: dt is clipped to tf - t */
} sphmhd_params;

/* Bound particle properties.  The names are the reference's property names
 * (particles.py:1-162).  Properties the reference writes but never reads
 * again (grho*, arho, rho_sum, dwdh, dt_force, cd_*) are not part of
 * the device path; dvdx_* come on request (sphmhd_bind_strain). */
typedef enum sphmhd_field {
    SPHMHD_F_X = 0, SPHMHD_F_Y, SPHMHD_F_Z,
    SPHMHD_F_U, SPHMHD_F_V, SPHMHD_F_W,
    SPHMHD_F_RHO, SPHMHD_F_H, SPHMHD_F_M, SPHMHD_F_CS, SPHMHD_F_P, SPHMHD_F_E,
    SPHMHD_F_AU, SPHMHD_F_AV, SPHMHD_F_AW, SPHMHD_F_AE, SPHMHD_F_AH,
    SPHMHD_F_X0, SPHMHD_F_Y0, SPHMHD_F_Z0,
    SPHMHD_F_U0, SPHMHD_F_V0, SPHMHD_F_W0,
    SPHMHD_F_UPRED, SPHMHD_F_VPRED, SPHMHD_F_WPRED,
    SPHMHD_F_EPRED, SPHMHD_F_E0, SPHMHD_F_H0,
    SPHMHD_F_DIV, SPHMHD_F_DT_CFL, SPHMHD_F_DT_ADAPT,
    SPHMHD_F_OMEGA, SPHMHD_F_CONVERGED,
    SPHMHD_F_ALPHA1, SPHMHD_F_ALPHA10,
    SPHMHD_F_BX, SPHMHD_F_BY, SPHMHD_F_BZ,
    SPHMHD_F_BEVOLX, SPHMHD_F_BEVOLY, SPHMHD_F_BEVOLZ, SPHMHD_F_PSI,
    SPHMHD_F_BEVOLXPRED, SPHMHD_F_BEVOLYPRED, SPHMHD_F_BEVOLZPRED, SPHMHD_F_PSIPRED,
    SPHMHD_F_BXPRED, SPHMHD_F_BYPRED, SPHMHD_F_BZPRED,
    SPHMHD_F_ABEVOLX, SPHMHD_F_ABEVOLY, SPHMHD_F_ABEVOLZ, SPHMHD_F_APSI,
    SPHMHD_F_BEVOLX0, SPHMHD_F_BEVOLY0, SPHMHD_F_BEVOLZ0, SPHMHD_F_PSI0,
    SPHMHD_F_DIVBSYMM, SPHMHD_F_DIVBDIFF, SPHMHD_F_DIV_FOR_PSI,
    SPHMHD_F_ALPHALOC, SPHMHD_F_XILIMITER, SPHMHD_F_DIVVDT,
    SPHMHD_NFIELDS
} sphmhd_field;

/* Conserved totals per instance, in this order (sphmhd_totals). */
enum {
    SPHMHD_T_MASS = 0,
    SPHMHD_T_MOMX, SPHMHD_T_MOMY, SPHMHD_T_MOMZ,
    SPHMHD_T_EKIN,     /* sum m |v|^2 / 2 */
    SPHMHD_T_ETHERM,   /* sum m e */
    SPHMHD_T_EMAG,     /* sum m |B|^2 / (2 mu0 rho) */
    SPHMHD_T_DIVB,     /* sum (m/rho) h |divB| / |B|  over particles with |B| > 0 */
    SPHMHD_T_VOLUME,   /* sum m/rho */
    SPHMHD_NTOTALS
};

typedef struct sphmhd_stats {
    int64_t evals;               /* force evaluations since create */
    int64_t density_iterations;  /* density passes summed over instances */
    int64_t pairs_force;         /* neighbour entries of the last force pass */
    int64_t cells;               /* cells of the last neighbour-search build */
    int32_t last_iterations_max; /* most density passes any instance needed in the last eval */
    int32_t max_neighbours;      /* largest neighbour count in the last force pass */
    int64_t force_redos;         /* force passes redone (iteration needed more passes than
                                    were launched ahead, or a neighbour list grew) */
    int64_t graph_replays;       /* steps whose launches were replayed from a captured hipGraph */
    int32_t block_density;       /* workgroup size of the density passes */
    int32_t block_lists;         /* workgroup size of the prep + neighbour-list launches */
    int32_t block_force;         /* workgroup size of the force pass */
    int32_t launch_tuned;        /* 1 once the launch shapes were chosen by timing */
    int64_t steps;               /* PEC steps completed by sphmhd_step / sphmhd_tangent_step */
    int32_t persist_density;     /* 1: the density passes run persistent (work queues), 0: plain */
    int32_t persist_lists;       /* the same for the neighbour-list build */
    int32_t persist_force;       /* the same for the force pass */
    int32_t reserved;
} sphmhd_stats;

/* Device time of the hot kernels, measured with HIP events recorded on the
 * context's stream around each launch (enabled by sphmhd_set_timing). */
typedef struct sphmhd_timing {
    double force_ms;          /* fused force pass (G3+G5..G9) */
    double density_ms;        /* density + Newton passes (G0) */
    double nnps_ms;           /* cell keys + cell sort + cell table */
    double lists_ms;          /* G1/G2/G4 finalize + force pack + neighbour lists */
    double stage_ms;          /* integrator stage kernels */
    int64_t force_launches;
    int64_t density_launches;
    int64_t nnps_builds;
    int64_t list_builds;
    int64_t stage_launches;
} sphmhd_timing;

typedef struct sphmhd_ctx sphmhd_ctx;

int sphmhd_abi_version(void);
const char* sphmhd_last_error(void);

/* Create a context for n_instances instances; params[k] belongs to instance
 * k and offsets has n_instances + 1 ascending entries (offsets[0] == 0).
 * Replaces the PySPH solver set-up that compiles get_equations()
 * (scheme.py:47-81) and allocates NNPS state. */
int sphmhd_create(const sphmhd_params* params, int32_t n_instances,
                  const int64_t* offsets, void* hip_stream, sphmhd_ctx** out);
int sphmhd_destroy(sphmhd_ctx* ctx);

/* Bind device arrays: fields[f] for every sphmhd_field f (z, z0 may be NULL
 * when dim == 2) and itype (int32, particles.py:7).  No copies are made.
 * Replaces ParticleArray / pa.gpu property binding. */
int sphmhd_bind(sphmhd_ctx* ctx, double* const* fields, int32_t* itype);

/* StrainDiagnostics (equations.py:432-471, group G4 of scheme.py:75): the
 * velocity gradient dvdx_ab = d v_a / d x_b of the Cullen-Dehnen sums.  No
 * later group reads it, so it is produced only on request: bind 9 device
 * FP64 rows (dvdx_xx, dvdx_xy, dvdx_xz, dvdx_yx, ..., dvdx_zz, storage order,
 * borrowed) and every evaluation that writes the diagnostic rows (all but the
 * non-final steps of a multi-step sphmhd_step call) fills them; NULL unbinds. */
int sphmhd_bind_strain(sphmhd_ctx* ctx, double* const* dvdx);

/* One force evaluation: neighbour-search update, then groups G0..G9 of
 * scheme.py:70-81 (density iteration to convergence, EOS, Cullen-Dehnen
 * switch, MHD and Maxwell-stress forces, induction and cleaning, energy
 * limiter, divB correction, adaptive time step).  Leaves min dt_adapt per
 * instance as the pending next time step.  Replaces
 * AccelerationEval.compute / Integrator.initial_acceleration. */
int sphmhd_eval(sphmhd_ctx* ctx);

/* MHDStep.initialize / stage1 / stage2 (integrator.py:12-82), with the
 * periodic wrap of PySPH's update_domain after each stage.  stage1 and
 * stage2 use each instance's current dt (sphmhd_get_clock). */
int sphmhd_pec_initialize(sphmhd_ctx* ctx);
int sphmhd_pec_stage1(sphmhd_ctx* ctx);
int sphmhd_pec_stage2(sphmhd_ctx* ctx);

/* t += dt; dt := the pending step from the last evaluation (min dt_adapt,
 * clipped to tf - t).  Replaces solver.py:57-61. */
int sphmhd_advance_clock(sphmhd_ctx* ctx);

/* Up to nsteps full PEC steps with the clock advanced after each: the body of
 * MHDSolver.solve (solver.py:42-73) without dumps.  The call stops early,
 * after the step that brings every instance's clock to its tf (t + dt >= tf;
 * dt is clipped to tf - t), so a run bounded by tf never takes the zero-length
 * steps that would follow; stats.steps counts the steps taken.  Kernels are fused
 * (stage2 of step n with initialize+stage1 of step n+1); the fields at
 * return are exactly those after the last stage2.  Steps after the first
 * replay a captured hipGraph and are enqueued one ahead of the host's wait
 * (a step whose evaluation needs more density passes or longer lists makes
 * the step after it exit at entry, and both are finished before going on).
 * Environment: SPHMHD_GRAPHS=0 launches every kernel plainly (no graphs, no
 * enqueueing ahead); SPHMHD_SPEC=k launches k density passes ahead of the
 * force pass instead of the adaptive count.  Neither changes the results. */
int sphmhd_step(sphmhd_ctx* ctx, int64_t nsteps);

/* MHDApplication._sync_initial_mhd_state (application.py:100-108):
 * Bevol = B / rho, psi = 0, predicted = current. */
int sphmhd_sync_initial_state(sphmhd_ctx* ctx);

/* MHDApplication._warm_initial_derivatives + _compute_initial_dt
 * (application.py:110-147) without the per-case thermodynamics hook. */
int sphmhd_warm_start(sphmhd_ctx* ctx);

/* Per-instance clock (t, dt) as host arrays of n_instances doubles. */
int sphmhd_get_clock(sphmhd_ctx* ctx, double* t, double* dt);
int sphmhd_set_clock(sphmhd_ctx* ctx, const double* t, const double* dt);
/* Pending dt (min dt_adapt over each instance from the last evaluation). */
int sphmhd_get_pending_dt(sphmhd_ctx* ctx, double* dt);

/* Conserved totals, host array of n_instances * SPHMHD_NTOTALS doubles. */
int sphmhd_totals(sphmhd_ctx* ctx, double* out);

/* sphmhd_totals without any allocation.  No counterpart in the reference; it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Enqueues the reduction
 * on the context's stream and writes n_instances * SPHMHD_NTOTALS doubles to
 * the DEVICE array out_dev (no host copy, no synchronisation).  The partial
 * sums use a static device array of the library (2 MiB per GPU, held until
 * the process ends), so at most 16384 instances (SPHMHD_ECAP above that),
 * and two calls that overlap on one GPU race.  Values equal sphmhd_totals'
 * up to the order of the additions. */
int sphmhd_totals_noalloc(sphmhd_ctx* ctx, double* out_dev);

/* Neighbour-search exports for parity checks, into HOST buffers of n
 * entries: sorted cell key (instance-local key plus the instance's cell
 * base) and the storage index at each sorted position. */
int sphmhd_cells(sphmhd_ctx* ctx, uint32_t* keys, int32_t* perm);
/* Neighbour lists of the current h, symmetric radius test
 * r^2 < (R max(h_i, h_j))^2 with minimum image: offsets[n + 1] and
 * idx[cap] (storage indices, ascending per particle), *total = entries. */
int sphmhd_neighbours(sphmhd_ctx* ctx, int64_t* offsets, int32_t* idx,
                      int64_t cap, int64_t* total);

int sphmhd_get_stats(sphmhd_ctx* ctx, sphmhd_stats* out);

/* Enable (1) or disable (0) event timing; enabling resets the sums. */
int sphmhd_set_timing(sphmhd_ctx* ctx, int32_t enable);
int sphmhd_get_timing(sphmhd_ctx* ctx, sphmhd_timing* out);

/* Block until the context's stream is idle. */
int sphmhd_synchronize(sphmhd_ctx* ctx);

/* Close-packed (HCP) lattice positions generated on the GPU: the
 * reference's close_packed_lattice (src/cases/lattice.py:4-34), bit for bit,
 * for the initial conditions of its cases.  bounds = {xmin, xmax, ymin,
 * ymax, zmin, zmax}; x, y, z: device arrays of capacity cap; *count receives
 * the number of sites (SPHMHD_ECAP if it exceeds cap, nothing written).
 * Enqueued on hip_stream (0: the null stream); no context needed. */
int sphmhd_hcp_lattice(const double* bounds, double spacing, double* x, double* y, double* z,
                       int64_t cap, int64_t* count, void* hip_stream);

/* Initial fields of a case, computed on the GPU, one thread per particle.
 * Replaces the fields half of the cases' create_mhd_particles
 * (src/cases/orstang.py:18-44, mhdrotor.py:18-53, mhdblast.py:17-43,
 * alfven.py:21-49, jadvect.py:18-49, mhdshock.py:21-62, mhdsine.py:18-42,
 * mhdvortex.py:28-54, mhdwave.py:18-43, jet.py:68-193).  The benchmark
 * lattices (the first four ids) and the jets also write x, y, z; the other
 * reference cases read the positions already in fields[X..Z] (the
 * close-packed lattice of sphmhd_hcp_lattice, masked and joined as the case
 * does).  params: the case's particle-independent constants (masses,
 * smoothing lengths, the lattice step ...), at most 48, in the order the
 * host wrapper documents; fields: the SPHMHD_NFIELDS row pointers of a bound
 * layout (x..Bz, alpha1 required, the others untouched); itype: written by
 * SPHMHD_CASE_MHDSHOCK (walls 3, fluid 1) when non-null.  For the jets,
 * aux_dev holds the single jet's (axial, y, z) rows of n_aux particles each
 * and n is n_aux times the jet count.  Arithmetic in the reference's order
 * with contraction off: equal to the host builders' bit for bit except
 * through sin, cos, exp, tanh, cosh and pow (a few ulp).  Enqueued on
 * hip_stream; no context needed. */
enum {
    SPHMHD_CASE_ORSZAG_TANG_2D = 1, /* square lattice, cell centres */
    SPHMHD_CASE_ROTOR_2D = 2,
    SPHMHD_CASE_BLAST_3D = 3,       /* cubic lattice, x fastest */
    SPHMHD_CASE_ALFVEN_2D = 4,
    SPHMHD_CASE_ORSTANG = 5,        /* reference cases, positions given */
    SPHMHD_CASE_MHDROTOR = 6,
    SPHMHD_CASE_MHDBLAST = 7,
    SPHMHD_CASE_ALFVEN = 8,
    SPHMHD_CASE_JADVECT = 9,
    SPHMHD_CASE_MHDSHOCK = 10,
    SPHMHD_CASE_MHDSINE = 11,
    SPHMHD_CASE_MHDVORTEX = 12,
    SPHMHD_CASE_MHDWAVE = 13,
    SPHMHD_CASE_JETS = 14           /* 1jet, 3jet, 12jet */
};
int sphmhd_case_fields(int32_t case_id, const double* params, int32_t nparams, double* const* fields,
                       int32_t* itype, int64_t n, int64_t n_aux, const double* aux_dev, void* hip_stream);

/* Launch shapes.  No counterpart in the reference (PySPH picks its own launch
 * sizes); the results of a step do not depend on them, only its speed does.
 * blocks[3] = workgroup sizes of the density passes, the prep + neighbour-list
 * launches and the force pass, each 64, 128 or 256.  By default the first
 * steps of a context's first sphmhd_step calls run every shape (with HIP
 * events; the results are those of any shape) and keep the fastest per
 * launch; sphmhd_set_launch fixes them instead, and blocks == NULL tunes
 * again.  The environment variable SPHMHD_TUNE=0 disables tuning (all 256). */
int sphmhd_set_launch(sphmhd_ctx* ctx, const int32_t* blocks);
int sphmhd_get_launch(sphmhd_ctx* ctx, int32_t* blocks, int32_t* tuned);

/* Launch modes of the same three launch classes.  No counterpart in the
 * reference.  Plain: one workgroup per tile of sorted particles.  Persistent:
 * as many workgroups as the GPU holds at once, each pulling tiles (of any
 * instance of the ensemble) from per-XCD work queues until all are empty.
 * modes[3] = 0 or 1 fixes them; NULL lets the tuner choose them with the
 * workgroup sizes (the default; SPHMHD_PERSIST=0/1 fixes every class).  The
 * results never depend on the mode; sphmhd_stats reports the modes in use. */
int sphmhd_set_persistent(sphmhd_ctx* ctx, const int32_t* modes);

/* ---- Tangent-linear step -------------------------------------------------
 * No counterpart in the reference, which has no derivatives; it serves the
 * gradients of final-time energies with respect to the initial state and the
 * parameters below.  Every field carries SPHMHD_NTANGENTS directional
 * derivatives beside its value; one tangent step advances values and
 * derivatives together (forward mode), taking exactly the primal step's
 * branches: the density Newton iteration as run, neighbour sets and cells
 * held fixed, the adaptive dt through the particle attaining the minimum.
 * Masses carry no tangent. */
#define SPHMHD_NTANGENTS 3

enum sphmhd_tparam {
    SPHMHD_TP_ARTIFICIAL_VISCOSITY = 0,
    SPHMHD_TP_ARTIFICIAL_THERMAL_CONDUCTIVITY,
    SPHMHD_TP_ARTIFICIAL_MAGNETIC_DISSIPATION,
    SPHMHD_TP_CLEANING_SPEED_FACTOR,
    SPHMHD_TP_CLEANING_DAMPING_FACTOR,
    SPHMHD_TP_ALPHA_MAX,
    SPHMHD_NTPARAMS
};

/* Tangent rows: SPHMHD_NTANGENTS * SPHMHD_NFIELDS device pointers,
 * direction-major (tangent d of field f at [d * SPHMHD_NFIELDS + f]), each n
 * doubles in storage order, borrowed like the bound fields; null where the
 * bound field is null.  The caller sets the initial tangents. */
int sphmhd_tangent_bind(sphmhd_ctx* ctx, double* const* tangent_fields);
/* Parameter tangents, host array of n_instances * SPHMHD_NTPARAMS *
 * SPHMHD_NTANGENTS doubles (d param / d direction); also zeroes the tangents
 * of the current and pending dt. */
int sphmhd_tangent_seed(sphmhd_ctx* ctx, const double* param_tangents);
/* sphmhd_warm_start / sphmhd_step over values and tangents. */
int sphmhd_tangent_warm_start(sphmhd_ctx* ctx);
int sphmhd_tangent_step(sphmhd_ctx* ctx, int64_t nsteps);
/* Totals with tangents, host array n_instances * SPHMHD_NTOTALS *
 * (1 + SPHMHD_NTANGENTS): the value, then its tangents. */
int sphmhd_tangent_totals(sphmhd_ctx* ctx, double* out);
/* Tangents of the current dt, host array n_instances * SPHMHD_NTANGENTS. */
int sphmhd_tangent_get_dt(sphmhd_ctx* ctx, double* dt_tangent);

/* ---- Adjoint (reverse-mode) step with checkpointed recomputation --------
 * No counterpart in the reference, which has no derivatives.  It gives the
 * gradient of a scalar function of the final state of a rollout (the warm
 * start of application.py:110-147 and PEC steps of integrator.py:12-82 with
 * the solver's dt of solver.py:57-61) with respect to every property of the
 * initial state and to the parameters of sphmhd_tparam, in one reverse sweep
 * whatever their number.  The derivative is that of the discrete map the
 * forward step computes, as for the tangent-linear step: neighbour sets and
 * cells fixed, the density Newton iteration as run, branches and min/max as
 * taken, the adaptive dt through the particle attaining the minimum, a step
 * clipped to tf - t without derivative.  The rollout's states are kept every
 * `every` steps; the reverse sweep recomputes each segment from its
 * checkpoint, then each step's evaluation before reversing it, so memory
 * grows with steps / every + every states, not with the rollout.
 *
 * Bar rows: SPHMHD_NFIELDS device pointers (null where the bound field is
 * null), n doubles each in storage order, borrowed like the fields; bar row f
 * holds d(objective)/d(property f) of the current state. */
int sphmhd_adjoint_bind(sphmhd_ctx* ctx, double* const* bar_fields);
/* From the current state (the rollout's initial state): the warm start when
 * warm != 0, then nsteps PEC steps (fewer when the clocks reach tf), keeping
 * the state every `every` steps.  The fields then hold the final state.
 * Returns the largest warning of the forward calls. */
int sphmhd_adjoint_record(sphmhd_ctx* ctx, int32_t warm, int64_t nsteps, int32_t every);
/* Steps the last record took (-1: none recorded). */
int sphmhd_adjoint_steps(sphmhd_ctx* ctx, int64_t* steps);
/* Adds to the bar rows the gradient of sum_k sum_q weights[q] total_q(k)
 * (the totals of sphmhd_totals, SPHMHD_NTOTALS weights, the divB total's
 * weight 0) at the current state: seeds the reverse sweep after a record. */
int sphmhd_adjoint_seed_totals(sphmhd_ctx* ctx, const double* weights);
/* The reverse sweep of the recorded rollout: the bar rows, holding the
 * objective's gradient with respect to the final state, end holding it with
 * respect to the initial state, to which the fields are restored;
 * param_bars (host, n_instances * SPHMHD_NTPARAMS, or null) receives the
 * gradient with respect to the parameters of sphmhd_tparam.  SPHMHD_ECAP: a
 * density iteration ran more than 32 passes; SPHMHD_ESTATE: a recomputed
 * density iteration ended elsewhere than the recorded one. */
int sphmhd_adjoint_reverse(sphmhd_ctx* ctx, double* param_bars);

#ifdef __cplusplus
}
#endif

#endif /* SPHMHD_H */
