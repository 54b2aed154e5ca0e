/*
 * lfdem_oracle.h -- CPU FP64 restatement of LF_DEM's per-step Stokes-resistance
 * path (cell list, pair lifecycle, lubrication / contact / dashpot / repulsion,
 * resistance-matrix assembly, sparse Cholesky solve, predictor-corrector step).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function cites the LF_DEM source (file:line, relative to the LF_DEM
 * repository root) whose semantics it restates.  Arithmetic follows the
 * reference's operation order; the library is built with -ffp-contract=off.
 *
 * Packed block formats (identical to LF_DEM's MatrixBlocks.h:5-46,58-84):
 *   DBlock  (15 doubles): col0[5] col1[3] col2[1] col3[3] col4[2] col5[1]
 *   ODBlock (18 doubles): col0[5] col1[3] col2[1] col3[5] col4[3] col5[1]
 */
#ifndef LFDEM_ORACLE_H
#define LFDEM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DBLOCK 15
#define ORC_ODBLOCK 18

typedef struct orc_params {
	double shear_rate;        /* gamma-dot (1 in hydrodynamic units) */
	double theta_shear;       /* shear direction (System.cpp:2006-2015) */
	double lub_max_gap;       /* lubrication cutoff gap (0.5) */
	double lub_reduce;        /* regularisation delta (1e-3) */
	double kn, kt;            /* spring constants */
	double mu_static, mu_dynamic;
	double tau_n, tau_t;      /* contact_relaxation_time, _tan */
	double sd_coeff;          /* Stokes-drag floor coefficient */
	double repulsion;         /* repulsive force scale F_R^0 */
	double repulsive_length;  /* screening length lambda_D; <=0 disables */
	double disp_max, dt_max, dt_min;
	int friction_model;       /* 0 frictionless, 1 Coulomb */
	int lubrication_model;    /* 1 normal, 2 tangential */
	int fixed_dt;             /* 1: constant step dt_max, no adaptTimeStep and no maximum-velocity
	                             update (System.cpp:607-610, 1094, 1130, 1238-1243, 2309) */
} orc_params;

void orc_default_params(orc_params *p);

/* ---- geometry (System.cpp:2478-2548, 2665-2674) ---- */
int orc_periodize(const double L[3], const double shear_disp[3], double pos[3]);
int orc_periodize_diff(const double L[3], const double shear_disp[3], double d[3]);
double orc_lub_range(double ri, double rj, double lub_max_gap);

/* ---- pair physics on a single pair (unit tests) ---- */
void orc_lub_constants(double a0, double a1, double out[36]);
void orc_dashpot_coeffs(double a0, double a1, const orc_params *p, double out[16]);
void orc_lub_blocks(double a0, double a1, double reduced_gap, const double nvec[3],
                    const orc_params *p, double db0[15], double db1[15], double od[18]);
void orc_dashpot_blocks(double a0, double a1, const double nvec[3], const orc_params *p,
                        double db0[15], double db1[15], double od[18]);
void orc_dblock_dense(const double b[15], double M[36]);
void orc_odblock_dense(const double b[18], double M[36]);

/* ---- CSC export, lower triangle (StokesSolver.cpp:334-379, 889-919) ---- */
long orc_csc_nnz(int np, int nod);
int orc_export_csc(int np, const double *dblocks, int nod, const int *od_col_ptr,
                   const int *od_rows, const double *odblocks,
                   int *colp, int *rowi, double *val);

/* ---- brute-force pair set (all i<j with sq_dist < range^2) ---- */
int orc_bruteforce_pairs(int np, const double *pos, const double *radius, const double L[3],
                         const double shear_disp[3], double lub_max_gap,
                         int *p0, int *p1, int cap);

/* ---- system ---- */
typedef struct orc_system orc_system;
orc_system *orc_create(int np, const double *pos, const double *radius, const double L[3],
                       const double shear_disp[3], const orc_params *p);
void orc_destroy(orc_system *s);
int orc_setup(orc_system *s);          /* initializeBoxing + checkNewInteraction + drag seed */
int orc_first_update(orc_system *s);   /* timeEvolution first-call block */
int orc_step(orc_system *s, double strain_end);   /* one predictor-corrector step */
const char *orc_last_error(orc_system *s);

/* phases, callable one at a time */
int orc_boxset_update(orc_system *s);
int orc_check_new_interaction(orc_system *s);
int orc_update_interactions(orc_system *s);
int orc_build_blocks(orc_system *s, double *dblocks, int *od_col_ptr, int *od_rows,
                     double *odblocks, int cap);
int orc_build_rhs(orc_system *s, double *rhs);
int orc_solve_current(orc_system *s, const double *rhs, double *x);
int orc_compute_velocities(orc_system *s);

/* ---- force / velocity components, stress, stress control ---- */
/* force components in LF_DEM's std::map order (System.cpp:169-215) */
enum { ORC_CONTACT = 0, ORC_DASHPOT = 1, ORC_HYDRO = 2, ORC_REPULSION = 3, ORC_NCOMP = 4 };
/* stress groups (Stress.cpp:16-73), map order */
enum { ORC_G_CONTACT = 0, ORC_G_DASHPOT = 1, ORC_G_HYDRO = 2, ORC_G_REPULSION = 3, ORC_NGROUP = 4 };
int orc_build_rhs_component(orc_system *s, int comp, double *rhs);
/* control: 0 rate, 1 stress (target_stress in internal force units per area) */
int orc_set_control(orc_system *s, int control, double target_stress);
int orc_set_shear_rate_limit(orc_system *s, double init_strain_limit, double init_rate_limit);
/* computeVelocities(divided_velocities=true) incl. the stress-control rate */
int orc_compute_velocities_components(orc_system *s);
/* calcStressPerParticle + calcStress at the current velocities */
int orc_calc_stress(orc_system *s);
/* group stresses (4 x 6, already / volume, hydro with 2E/(6 pi)), total (6),
 * per-particle total stress (np x 6, sum of all components) */
int orc_get_stress(const orc_system *s, double *groups, double *total, double *particle);
int orc_get_components(const orc_system *s, int comp, double *vel, double *ang);
/* GU stresslet per particle (np x 6) for given velocities (np x 3 each) */
int orc_gu_stress(orc_system *s, const double *vel, const double *ang, double *S);
double orc_shear_rate(const orc_system *s);
double orc_viscosity(const orc_system *s);
/* one predictor-corrector step with calc_stress (stress computed after the
 * corrector); adapt_first = 0 for the last step of timeEvolution */
int orc_step_ex(orc_system *s, double strain_end, int calc_stress, int adapt_first);
/* System::timeEvolution to strain_end: steps while keepRunning, then the
 * stress step; returns the number of steps or -1 */
int orc_time_evolution(orc_system *s, double strain_end);
int orc_set_in_predictor(orc_system *s, int flag);
int orc_set_dt(orc_system *s, double dt);

/* state access */
int orc_num_pairs(const orc_system *s);
/* System::getNonAffineDisp (System.h:529-532): N x 3 */
void orc_get_na_disp(const orc_system *s, double *out);
/* Per interaction (pair order) the force columns of Simulation::outputIntFileTxt:
 * lubrication normal part, lubrication tangential (3), contact normal norm,
 * contact tangential (3), repulsive norm, contact xF viscosity contribution */
#define ORC_PAIR_OUT 10
int orc_pair_output(const orc_system *s, double *out);
int orc_get_pairs(const orc_system *s, int *p0, int *p1, int *contact_state,
                  double *geom /* P x 5: nx ny nz r gap */, double *disp_tan /* P x 3 */);
int orc_set_pair_history(orc_system *s, int n, const int *p0, const int *p1,
                         const int *state, const double *disp_tan, const double *prev_disp_tan);
int orc_get_state(const orc_system *s, double *pos, double *vel, double *ang_vel,
                  double *na_vel, double *na_ang, double *scalars /* 8 */);
int orc_set_positions(orc_system *s, const double *pos, const double shear_disp[3]);
/* total velocities and angular velocities (np x 3 each) as the next pair update reads them */
int orc_set_velocities(orc_system *s, const double *vel, const double *ang);
int orc_get_box_info(const orc_system *s, int *dims /* 4: nx ny nz boxed */, int *labels);
int orc_get_neighbor_cells(const orc_system *s, int cell, int *out, int cap);

/* ---- wall time per phase (the CPU baseline's split), seconds ---- */
enum { ORC_PH_MOVE = 0,      /* displacements, Lees-Edwards boxing, cell membership */
       ORC_PH_NEIGHBOR = 1,  /* checkNewInteraction */
       ORC_PH_PAIRS = 2,     /* updateInteractions */
       ORC_PH_ASSEMBLE = 3,  /* resistance blocks and right-hand sides */
       ORC_PH_ORDER = 4,     /* minimum-degree ordering and symbolic pattern */
       ORC_PH_FACTOR = 5,    /* numeric factorisation */
       ORC_PH_SOLVE = 6,     /* triangular solves */
       ORC_NPHASE = 7 };
void orc_phase_times(double out[ORC_NPHASE]);
void orc_phase_reset(void);

/* ---- sparse Cholesky on 6x6-block SPD matrices ---- */
typedef struct orc_chol orc_chol;
orc_chol *orc_chol_factor(int np, const double *dblocks, const int *od_col_ptr,
                          const int *od_rows, const double *odblocks, int *status);
void orc_chol_solve(const orc_chol *c, const double *b, double *x);
long orc_chol_nnz_blocks(const orc_chol *c);
double orc_chol_flops(const orc_chol *c);
void orc_chol_free(orc_chol *c);

#ifdef __cplusplus
}
#endif
#endif
