/*
 * lfdem_amd.h -- C ABI of the MI355X (gfx950) Stokes-resistance library.
 *
 * Two entry families:
 *
 *  1. lfd_stokes_*  -- drop-in for LF_DEM's StokesSolver (StokesSolver.h:31-286):
 *     the same fill -> factorize -> RHS -> solve protocol and packed 6x6 block
 *     formats (MatrixBlocks.h:5-84), solved on the GPU.
 *  2. lfd_system_*  -- the device-resident step loop of LF_DEM's System
 *     (System.cpp:888-983): cell list, pair lifecycle, assembly, solve and the
 *     predictor-corrector move, all kept in HBM between steps.
 *
 * Conventions: plain pointers and sizes, no exceptions; every call returns an
 * lfd_status (0 = success); lfd_last_error() describes the last failure of the
 * calling thread.  Handles are not thread safe; use one per host thread.
 * RHS / solution vectors are interleaved per particle as [F|U xyz, T|W xyz]
 * (StokesSolver.h:38-40), FP64 at the boundary.
 */
#ifndef LFDEM_AMD_H
#define LFDEM_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

typedef enum lfd_status {
	LFD_OK = 0,
	LFD_ERR_ARG = -1,            /* bad argument or call order */
	LFD_ERR_HIP = -2,            /* HIP runtime failure */
	LFD_ERR_NOT_SPD = -3,        /* a diagonal block is not positive definite */
	LFD_ERR_NOT_CONVERGED = -4,  /* Krylov / refinement limit reached */
	LFD_ERR_CAPACITY = -5,       /* pair or per-row capacity exceeded */
	LFD_ERR_BOX = -6,            /* position outside the cell grid (BoxSet.cpp:775-778) */
	LFD_ERR_UNSUPPORTED = -7     /* feature outside this library's scope */
} lfd_status;

/* Physical parameters, hydrodynamic units (gamma-dot = 1); defaults follow
 * SimulationInit.cpp:941-1014 with the configuration of the benchmark
 * workloads (kn=1e4, kt=5e3, mu=1, tau_n=1e-3, delta=1e-3, gap 0.5). */
typedef struct lfd_params {
	double shear_rate, theta_shear, lub_max_gap, lub_reduce, kn, kt;
	double mu_static, mu_dynamic, tau_n, tau_t, sd_coeff, repulsion;
	double repulsive_length, disp_max, dt_max, dt_min;
	int friction_model;     /* 0 frictionless, 1 Coulomb (Contact.cpp:255-291) */
	int lubrication_model;  /* 1 normal, 2 tangential */
	int fixed_dt;           /* 1: every step is dt_max, no adaptTimeStep (ParameterSet::fixed_dt,
	                           System.cpp:607-610, 1094, 1130, 1238-1243) */
} lfd_params;

typedef enum lfd_solver_kind {
	LFD_SOLVER_MIXED = 0,   /* FP64 refinement around persistent FP32 block-Jacobi PCG, blocks in LDS */
	LFD_SOLVER_PLAIN = 1,   /* FP64 block-Jacobi PCG, one kernel per operation */
	LFD_SOLVER_MIXED_GLOBAL = 2, /* as MIXED, blocks read from HBM each iteration */
	LFD_SOLVER_PLAIN_DEVICE = 3  /* as PLAIN with the scalars on the device: no host round trip per
	                                iteration (the host checks every 16 iterations); same iterates */
} lfd_solver_kind;

/* Preconditioner of the MIXED solver on the device-resident paths. */
typedef enum lfd_precond {
	LFD_PRECOND_BLOCK_JACOBI = 0,  /* inverse of each 6x6 diagonal block */
	LFD_PRECOND_AGGREGATE = 1      /* inverse of the 48x48 diagonal block of groups of up to 8 strongly
	                                  coupled particles.  Up to 8192 particles an instance is grouped
	                                  whole; a larger one of a system or batch (not the StokesSolver
	                                  drop-in) is grouped in cell-ordered chunks and solved by a
	                                  solver launch of its own, with 16 waves per workgroup (up to
	                                  16 x 8 x the device's compute units particles, 32768 on MI355X);
	                                  a larger instance still runs 6x6 block-Jacobi */
} lfd_precond;

/* How the persistent solver kernels are launched.  They synchronise their
 * workgroups with bounded spin waits, so every workgroup of a launch must be
 * resident at once.  A cooperative launch guarantees it and costs about 20 us
 * of idle device time per solve, and it cannot be captured in a step graph; a
 * plain launch relies on the device having room, which holds when no other
 * solve runs on the device at the same time (otherwise the solve ends with
 * LFD_ERR_SYNC_TIMEOUT, never hangs).  A host that runs solves on one device
 * from several streams at once should ask for cooperative launches.  The
 * solver that streams the matrix from HBM (LFD_SOLVER_MIXED_GLOBAL) is always
 * launched cooperatively. */
typedef enum lfd_launch {
	LFD_LAUNCH_AUTO = 0,           /* plain when the launch fits the device at once (every launch the
	                                  library makes does), else cooperative */
	LFD_LAUNCH_COOPERATIVE = 1,
	LFD_LAUNCH_PLAIN = 2
} lfd_launch;

typedef struct lfd_solve_opts {
	double tol;             /* target ||b - R x|| / ||b|| in FP64 */
	double inner_tol;       /* residual drop per FP32 inner solve */
	int max_outer, max_inner;
	int lanes_per_row;      /* SpMV lanes per 6x6 block row: 1, 2, 4, 8 */
	int groups;             /* workgroups of the persistent solver, 0 = auto */
	int kind;               /* lfd_solver_kind */
	int precond;            /* lfd_precond */
	int regroup_interval;   /* LFD_PRECOND_AGGREGATE: steps between regroupings (the block inverses
	                           are recomputed at every assembly) */
	int launch;             /* lfd_launch */
	int max_rhs;            /* right-hand sides one launch of the LDS-staged solver (block-Jacobi,
	                           or the aggregate solver of a large instance) takes together when a step solves several with one matrix (the rate
	                           independent and rate proportional parts of stress control, the four
	                           parts of a stress evaluation): 0 = two, 1 = one */
	int agg_waves;          /* LFD_PRECOND_AGGREGATE: waves of 8 particles per workgroup of the solver,
	                           8, 12 or 16 (0 = 8).  With 8 (two wavefronts per SIMD) every lane keeps
	                           its row of the wave's 48x48 inverse in registers; with 12 the
	                           inverses are staged in LDS and one launch takes 1.5x the waves
	                           (fewer cooperative launches for a large batch).  16: every instance
	                           takes the large-instance form (inverses in LDS, only the couplings
	                           between waves staged, two right-hand sides per launch), which
	                           instances above 8192 particles take whatever is asked. */
} lfd_solve_opts;

typedef struct lfd_solve_stats {
	int iters_inner, iters_outer;
	double res_norm, rhs_norm;
} lfd_solve_stats;

const char *lfd_last_error(void);
void lfd_default_params(lfd_params *p);
void lfd_default_solve_opts(lfd_solve_opts *o);
int lfd_device_count(void);

/* ------------------------------------------------------------------ */
/* StokesSolver drop-in                                                 */
/* ------------------------------------------------------------------ */
typedef struct lfd_stokes lfd_stokes;

/* StokesSolver::init (StokesSolver.cpp:77-106).  Only np_mobile == np_total
 * is supported (the mixed mobile/fixed problem is out of scope). */
int lfd_stokes_create(int np_total, int np_mobile, lfd_stokes **out);
void lfd_stokes_destroy(lfd_stokes *s);
int lfd_stokes_set_opts(lfd_stokes *s, const lfd_solve_opts *o);
/* StokesSolver::resetResistanceMatrix (StokesSolver.cpp:474-511); drag holds
 * np_total DBlocks (15 doubles each); nb_blocks_hint may overcount. */
int lfd_stokes_reset(lfd_stokes *s, int nb_blocks_hint, const double *drag_dblocks, int pattern_changed);
/* StokesSolver::startNewColumn (StokesSolver.cpp:921-935) */
int lfd_stokes_start_new_column(lfd_stokes *s);
/* StokesSolver::addResistanceBlocks (StokesSolver.cpp:121-129): i is the
 * current column, j > i in any order within the column (StokesSolver.h:104-105;
 * a row twice in one column fails at lfd_stokes_fill_done); dbi/dbj DBlocks
 * (15) added to the diagonal blocks of i and j, odb the ODBlock (18) at
 * (row j, column i). */
int lfd_stokes_add_blocks(lfd_stokes *s, int i, int j, const double *dbi, const double *dbj, const double *odb);
/* StokesSolver::matrixFillingDone (StokesSolver.cpp:937-948): uploads the
 * matrix and builds the preconditioner (the "factorization"). */
int lfd_stokes_fill_done(lfd_stokes *s);
/* RHS (StokesSolver.cpp:513-624) */
int lfd_stokes_reset_rhs(lfd_stokes *s);
int lfd_stokes_reset_rhs_torque(lfd_stokes *s);
int lfd_stokes_add_rhs_force(lfd_stokes *s, int i, const double f[3]);
int lfd_stokes_add_rhs_torque(lfd_stokes *s, int i, const double t[3]);
int lfd_stokes_set_rhs_force(lfd_stokes *s, int i, const double f[3]);
int lfd_stokes_set_rhs_torque(lfd_stokes *s, int i, const double t[3]);
int lfd_stokes_set_rhs(lfd_stokes *s, const double *ft);
int lfd_stokes_add_rhs(lfd_stokes *s, const double *ft);
/* StokesSolver::solve (StokesSolver.cpp:671-703): uw has room for
 * uw_len doubles; LFD_ERR_ARG when uw_len < 6*np_mobile (the reference throws). */
int lfd_stokes_solve(lfd_stokes *s, double *uw, int uw_len, lfd_solve_stats *stats);
/* StokesSolver::solvingIsDone (StokesSolver.cpp:847-851) */
int lfd_stokes_solving_done(lfd_stokes *s);
/* y = R x with the current matrix (StokesSolver::multiplySolutionByResMat's role) */
int lfd_stokes_multiply(lfd_stokes *s, const double *x, double *y);

/* Allocation-free solve.  No counterpart in LF_DEM: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It does what lfd_stokes_fill_done
 * followed by lfd_stokes_solve does, with the mixed FP32/FP64 6x6
 * block-Jacobi solver, on a matrix, right-hand side and solution held by the
 * caller in device memory:
 *   dblocks   np summed diagonal DBlocks (15 doubles each: drag plus the pair
 *             contributions, as after the column protocol),
 *   p0, p1    nblocks pairs p0 < p1 < np, sorted by (p0, p1), no repeats,
 *   odblocks  their ODBlocks (18 doubles) at (row p1, column p0),
 *   rhs       6 np doubles [F xyz, T xyz] per particle,
 *   x         6 np doubles, receives [U xyz, W xyz] per particle.
 * opts may be NULL (defaults; kind, precond, launch and max_rhs are ignored);
 * stream is the caller's hipStream_t (NULL: the null stream); the call ends
 * with a synchronisation of that stream to return its status.  All work
 * arrays sit in a static device buffer of LFD_NOALLOC_BYTES per GPU, held
 * until the process ends: at most LFD_NOALLOC_MAX_NP particles and
 * LFD_NOALLOC_MAX_BLOCKS blocks, larger sizes return LFD_ERR_CAPACITY.  Calls
 * on one GPU share that buffer: calls that overlap on one GPU race. */
#define LFD_NOALLOC_MAX_NP 16384
#define LFD_NOALLOC_MAX_BLOCKS 131072
#define LFD_NOALLOC_BYTES 25165824
int lfd_stokes_solve_noalloc(int np, int nblocks, const double *dblocks, const int *p0, const int *p1,
                             const double *odblocks, const double *rhs, double *x, const lfd_solve_opts *opts,
                             void *stream, lfd_solve_stats *stats);

/* ------------------------------------------------------------------ */
/* Device-resident System step loop                                     */
/* ------------------------------------------------------------------ */
typedef struct lfd_system lfd_system;

/* System::setupConfiguration (System.cpp:468-521, 546-611): cell list,
 * initial pairs, drag seed.  pos: np x 3 in [0, L); radius: np. */
int lfd_system_create(int np, const double *pos, const double *radius, const double L[3],
                      const double shear_disp[3], const lfd_params *p, const lfd_solve_opts *o,
                      lfd_system **out);
void lfd_system_destroy(lfd_system *s);
/* System::timeEvolution first-call block (System.cpp:1219-1228) */
int lfd_system_first_update(lfd_system *s);
/* n predictor-corrector steps (System.cpp:1241-1256, 888-983); strain_end < 0
 * disables the end-of-interval dt clamp.  Enqueued; errors are checked at the end. */
int lfd_system_step(lfd_system *s, int nsteps, double strain_end);
/* As lfd_system_step, each step replayed from a captured HIP graph (captured
 * on first use, once per parity of the pair buffers, and again when the
 * solver jobs, options or control change).  Steps run eagerly instead when
 * the solver form needs host round trips (LFD_SOLVER_PLAIN), a cooperative
 * launch, or while profiling.  LF_DEM has no counterpart: this is the
 * capture of the step loop of System.cpp:1241-1256. */
int lfd_system_step_graph(lfd_system *s, int nsteps, double strain_end);
/* Number of step graphs held (0 to 2, one per pair-buffer parity). */
int lfd_system_graphs(const lfd_system *s);
int lfd_system_sync(lfd_system *s);
/* individual phases (tests) */
int lfd_system_check_new_interactions(lfd_system *s);
int lfd_system_update_interactions(lfd_system *s, int in_predictor, double dt);
int lfd_system_assemble(lfd_system *s);
int lfd_system_compute_velocities(lfd_system *s, int in_predictor);
/* state access */
int lfd_system_num_pairs(lfd_system *s);
int lfd_system_get_pairs(lfd_system *s, int cap, int *p0, int *p1, int *state, double *geom5, double *disp_tan);
/* Per interaction, in the order of lfd_system_get_pairs, the force columns of
 * Simulation::outputIntFileTxt (Simulation.cpp:941-1006), LFD_PAIR_OUT values:
 * normal part of the lubrication force on p0 (positive for compression,
 * Lubrication.cpp:659-694, zero unless the gap is positive), its tangential
 * part (3); the contact force's normal part F.n (signed; the file prints its
 * norm) and tangential part (3), spring + dashpot (Contact.cpp:235-253,
 * 487-501); the repulsive force norm (RepulsiveForce.h:63-66); the contact
 * xF's viscosity contribution (rvec (x) F_c):E_inf / rate^2.  Computed on the
 * device from the current pairs and velocities; returns the pair count. */
#define LFD_PAIR_OUT 10
int lfd_system_get_pair_output(lfd_system *s, int cap, double *out);
/* System::getNonAffineDisp (System.h:529-532): the corrector's non-affine
 * velocity times dt summed over the steps (System.cpp:980-982), n x 3. */
int lfd_system_get_na_disp(lfd_system *s, double *out3);
/* System::getContacts (System.cpp:306-318): p0, p1, disp_tan[3], disp_rolling[3]
 * of every active contact (disp_rolling is zero: rolling friction is not
 * modelled); returns the count (only the count when p0 is NULL). */
int lfd_system_get_contacts(lfd_system *s, int cap, int *p0, int *p1, double *disp_tan, double *disp_rolling);
/* System::setContacts (System.cpp:285-304), after lfd_system_first_update as
 * setupConfiguration does after checkNewInteraction: the listed contacts'
 * tangential displacements are restored on their interactions (each must
 * exist, else LFD_ERR_ARG; nonzero disp_rolling gives LFD_ERR_UNSUPPORTED),
 * then the pairs are updated with dt = 0. */
int lfd_system_set_contacts(lfd_system *s, int n, const int *p0, const int *p1, const double *disp_tan,
                            const double *disp_rolling);
/* The clock of a checkpoint (States.h:13-17): time and cumulated strain. */
int lfd_system_set_clock(lfd_system *s, double time, double strain);
int lfd_system_get_state(lfd_system *s, double *pos, double *vel6, double *na6, double *scalars8);
int lfd_system_get_cells(lfd_system *s, int *dims4, int *labels);
int lfd_system_get_blocks(lfd_system *s, int cap, double *dblocks, double *odblocks, int *p0, int *p1);
int lfd_system_get_rhs(lfd_system *s, double *rhs);
int lfd_system_get_solve_stats(lfd_system *s, lfd_solve_stats *st, long long *total_inner);
/* Waves (groups of up to 8 particles whose 48x48 diagonal block is inverted)
 * of the aggregate preconditioner; 0 under 6x6 block-Jacobi.  No counterpart
 * in the reference (its factorisation is CHOLMOD's). */
int lfd_system_precond_waves(lfd_system *s, int *waves);
/* Solver slot of each of the n particles under the aggregate preconditioner,
 * counted from the system's first slot: the particles of wave w hold slots
 * 8w .. 8w+7 and share one inverted diagonal block; -1 for every particle
 * under 6x6 block-Jacobi.  No counterpart in the reference. */
int lfd_system_solver_slots(lfd_system *s, int *slots);
int lfd_system_set_opts(lfd_system *s, const lfd_solve_opts *o);
/* Solver-launch timing with HIP events on the library's stream: returns the
 * summed duration (ms) and count of the launches recorded since the previous
 * call, then enables (enable != 0) or disables recording.  No LF_DEM
 * counterpart (LF_DEM's own timing is wall clock, Simulation.cpp:478-489). */
int lfd_system_profile(lfd_system *s, int enable, double *solve_ms_total, int *solve_launches);

/* ---- stress and stress control ---- */
/* Imposed-flow control: the reference's rate (-r) and stress (-s) modes
 * (Simulation.cpp:241-264, System.cpp:2273-2299).  Under stress control every
 * velocity computation solves the rate independent and rate proportional
 * parts and sets the shear rate from target_stress (System::computeShearRate,
 * System.cpp:2017-2055); target_stress is in the internal stress unit. */
#define LFD_RATE_CONTROL 0
#define LFD_STRESS_CONTROL 1
int lfd_system_set_control(lfd_system *s, int control, double target_stress);
/* init_strain_shear_rate_limit / init_shear_rate_limit (System.cpp:90-91, 2050-2054) */
int lfd_system_set_shear_rate_limit(lfd_system *s, double init_strain_limit, double init_rate_limit);
/* One predictor-corrector step with the stress computed after the corrector
 * and no new adaptTimeStep before the predictor: the last step of
 * System::timeEvolution (System.cpp:1270-1273). */
int lfd_system_step_stress(lfd_system *s, double strain_end);
/* System::timeEvolution to strain_end >= 0 (System.cpp:1203-1279): steps while
 * the strain is below strain_end - 1e-8 (keepRunning, System.cpp:1189-1201),
 * then the stress step.  *nsteps receives the number of steps taken. */
int lfd_system_time_evolution(lfd_system *s, double strain_end, int *nsteps);
/* Velocities by components (computeVelocityByComponents, System.cpp:1947-1965)
 * and the stress at the current configuration (calcStressPerParticle +
 * calcStress, Stress.cpp:194-287, 342-384). */
int lfd_system_compute_stress(lfd_system *s, int in_predictor);
/* Last computed stress: groups24 = contact, dashpot, hydro, repulsion (6 each,
 * (xx, xy, xz, yz, yy, zz), already divided by the volume, hydro including the
 * solvent 2E/(6 pi)); total6; the shear rate; the viscosity
 * 0.5 S:E/(E:E) (Simulation.cpp:662-666); particle6 (np x 6, optional) the
 * per-particle total stress (calcTotalStressPerParticle, Stress.cpp:289-300).
 * Any output pointer may be NULL. */
int lfd_system_get_stress(lfd_system *s, double *groups24, double *total6, double *shear_rate,
                          double *viscosity, double *particle6);
/* Velocity part (0 contact, 1 dashpot, 2 hydro, 3 repulsion) of the last
 * stress computation, np x 6 (U, W) non-affine; rate proportional parts
 * already scaled by the shear rate under stress control. */
int lfd_system_get_parts(lfd_system *s, int part, double *vel6);

/* ------------------------------------------------------------------ */
/* Batches of independent instances (sweeps and served simulations)    */
/* ------------------------------------------------------------------ */
/* No LF_DEM counterpart: LF_DEM runs one System per process.  A batch steps
 * many independent Systems (each with its own N, box, parameters, shear
 * displacement and pair list) with one set of kernels per phase and one
 * cooperative solver launch per 256 workgroups; lfd_system_* is a batch of
 * one.  Every instance's trajectory is the one it has alone: instances share
 * no arithmetic.  Capacities are fixed at creation: particle slots (an
 * instance takes N rounded up to 128), cells and instance handles. */
typedef struct lfd_batch lfd_batch;
int lfd_batch_create(int slot_cap, int cell_cap, int inst_cap, int control, const lfd_solve_opts *o,
                     lfd_batch **out);
void lfd_batch_destroy(lfd_batch *b);
/* Slots, cells and device bytes an instance needs (admission by HBM
 * footprint); host-only, no device call. */
int lfd_batch_footprint(int np, const double *radius, const double L[3], const lfd_params *p,
                        int *slots, int *cells, long long *bytes);
/* 1 when an instance of that many slots and cells can be admitted now, else 0 */
int lfd_batch_room(lfd_batch *b, int slots, int cells);
/* Admission: System::setupSystem + the first-call block of timeEvolution
 * (System.cpp:468-611, 1219-1228) for the new instance only, between steps.
 * strain_end < 0: no end-of-interval clamp; target_stress is used under
 * stress control.  LFD_ERR_CAPACITY (nothing changed) when it does not fit. */
int lfd_batch_add(lfd_batch *b, int np, const double *pos, const double *radius, const double L[3],
                  const double shear_disp[3], const lfd_params *p, double strain_end, double target_stress,
                  int *id);
/* Retirement: frees the instance's slots, cells and handle. */
int lfd_batch_remove(lfd_batch *b, int id);
/* nsteps predictor-corrector steps of every live instance (enqueued). */
int lfd_batch_step(lfd_batch *b, int nsteps);
/* As lfd_batch_step through captured step graphs (see lfd_system_step_graph). */
int lfd_batch_step_graph(lfd_batch *b, int nsteps);
int lfd_batch_graphs(const lfd_batch *b);
int lfd_batch_sync(lfd_batch *b);
/* scalars8 as lfd_system_get_state; status: the instance's own lfd_status */
int lfd_batch_query(lfd_batch *b, int id, double *scalars8, lfd_solve_stats *st, int *status);
/* Every live instance in one device read: ids, strain and own lfd_status;
 * returns the count (at most cap). */
int lfd_batch_status(lfd_batch *b, int cap, int *ids, double *strain, int *status);
int lfd_batch_get_state(lfd_batch *b, int id, double *pos, double *vel6, double *na6, double *scalars8);
/* lfd_system_get_pair_output / lfd_system_get_na_disp of instance id */
int lfd_batch_get_pair_output(lfd_batch *b, int id, int cap, double *out);
int lfd_batch_get_na_disp(lfd_batch *b, int id, double *out3);
/* As lfd_system_precond_waves, for instance id of a batch. */
int lfd_batch_precond_waves(lfd_batch *b, int id, int *waves);
/* As lfd_system_solver_slots, for instance id of a batch. */
int lfd_batch_solver_slots(lfd_batch *b, int id, int *slots);
int lfd_batch_get_pairs(lfd_batch *b, int id, int cap, int *p0, int *p1, int *state, double *geom5,
                        double *disp_tan);
int lfd_batch_num_pairs(lfd_batch *b, int id);
int lfd_batch_info(lfd_batch *b, int *live, int *slot_cap, int *cell_cap, int *inst_cap, int *solver_launches,
                   long long *bytes);
int lfd_batch_profile(lfd_batch *b, int enable, double *solve_ms_total, int *solve_launches);

#ifdef __cplusplus
}
#endif
#endif
