/*
 * lfdem_oracle.c -- CPU FP64 restatement of LF_DEM's Stokes-resistance step
 * path.  TEST INFRASTRUCTURE ONLY (the checker for the HIP product path; see
 * lfdem_oracle.h).  Built with -O2 -ffp-contract=off so that every product and
 * sum is rounded exactly as the reference's default (no -march) build rounds it.
 *
 * Citations are LF_DEM file:line.  Deliberate departures from the reference:
 *  - pairs are kept sorted by (p0,p1) instead of in creation order with
 *    swap-with-last removal (System.cpp:1395-1400); per-particle sums then
 *    run in ascending partner order, which is the reference's own order for
 *    the diagonal blocks (System.cpp:1459-1471).
 *  - the Interaction copy quirk (Interaction.cpp:38-66: copies re-initialise
 *    contact history on std::vector growth, and the element swapped into a
 *    removed slot skips that pass's update) is not reproduced: every pair is
 *    updated exactly once per pass and keeps its history.
 *  - the solve is a project-written sparse Cholesky (minimum-degree order,
 *    left-looking 6x6-block factorisation) standing in for CHOLMOD, which is
 *    not part of the reference tree.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c99 */
#include "lfdem_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <time.h>

/* Wall time per phase of the step path, accumulated over all systems of the
 * process (the CPU baseline's per-phase split; single-threaded use). */
static double g_phase[ORC_NPHASE];
static double ph_now(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec+1e-9*(double)t.tv_nsec;
}
void orc_phase_times(double out[ORC_NPHASE]) { memcpy(out, g_phase, sizeof(g_phase)); }
void orc_phase_reset(void) { memset(g_phase, 0, sizeof(g_phase)); }
#define PH(k, stmt) do { double t0_ = ph_now(); stmt; g_phase[k] += ph_now()-t0_; } while (0)

/* ======================================================================= */
/* small vector helpers, same operation order as vec3d.h                   */
/* ======================================================================= */
static inline double dot3(const double a[3], const double b[3])
{
	return a[0]*b[0]+a[1]*b[1]+a[2]*b[2];          /* vec3d.h:251-255 */
}
static inline void cross3(const double a[3], const double b[3], double c[3])
{
	double x = a[1]*b[2]-a[2]*b[1];                 /* vec3d.h:277-283 */
	double y = a[2]*b[0]-a[0]*b[2];
	double z = a[0]*b[1]-a[1]*b[0];
	c[0] = x; c[1] = y; c[2] = z;
}
/* dot(Sym2Tensor, vec3d), elm = (xx, xy, xz, yz, yy, zz) (Sym2Tensor.h:231-238) */
static inline void symdot(const double E[6], const double v[3], double o[3])
{
	double x = E[0]*v[0]+E[1]*v[1]+E[2]*v[2];
	double y = E[1]*v[0]+E[4]*v[1]+E[3]*v[2];
	double z = E[2]*v[0]+E[3]*v[1]+E[5]*v[2];
	o[0] = x; o[1] = y; o[2] = z;
}

void orc_default_params(orc_params *p)
{
	/* SimulationInit.cpp:941-1014 defaults, hydrodynamic units (gamma-dot=1) */
	memset(p, 0, sizeof(*p));
	p->shear_rate = 1;
	p->theta_shear = 0;
	p->lub_max_gap = 0.5;
	p->lub_reduce = 1e-3;
	p->kn = 1e4;
	p->kt = 5e3;
	p->mu_static = 1;
	p->mu_dynamic = 1;
	p->tau_n = 1e-3;
	p->tau_t = 0;
	p->sd_coeff = 1;
	p->repulsion = 5;
	p->repulsive_length = 0.02;
	p->disp_max = 1e-3;
	p->dt_max = -1;
	p->dt_min = -1;
	p->friction_model = 1;
	p->lubrication_model = 2;
}

/* ======================================================================= */
/* Lees-Edwards geometry                                                   */
/* ======================================================================= */
int orc_periodize(const double L[3], const double sd[3], double pos[3])
{
	/* System.cpp:2478-2507 */
	int z_shift = 0;
	if (pos[2] >= L[2]) {
		pos[2] -= L[2];
		pos[0] -= sd[0]; pos[1] -= sd[1]; pos[2] -= sd[2];
		z_shift = -1;
	} else if (pos[2] < 0) {
		pos[2] += L[2];
		pos[0] += sd[0]; pos[1] += sd[1]; pos[2] += sd[2];
		z_shift = 1;
	}
	while (pos[0] >= L[0]) pos[0] -= L[0];
	while (pos[0] < 0) pos[0] += L[0];
	if (pos[1] >= L[1]) pos[1] -= L[1];
	else if (pos[1] < 0) pos[1] += L[1];
	return z_shift;
}

int orc_periodize_diff(const double L[3], const double sd[3], double d[3])
{
	/* System.cpp:2517-2548; lx_half etc. are L/2 (System.h setBoxSize) */
	double hx = L[0]/2, hy = L[1]/2, hz = L[2]/2;
	int zshift = 0;
	if (d[2] > hz) {
		d[2] -= L[2];
		d[0] -= sd[0]; d[1] -= sd[1]; d[2] -= sd[2];
		zshift = -1;
	} else if (d[2] < -hz) {
		d[2] += L[2];
		d[0] += sd[0]; d[1] += sd[1]; d[2] += sd[2];
		zshift = 1;
	}
	while (d[0] > hx) d[0] -= L[0];
	while (d[0] < -hx) d[0] += L[0];
	if (d[1] > hy) d[1] -= L[1];
	else if (d[1] < -hy) d[1] += L[1];
	return zshift;
}

double orc_lub_range(double ri, double rj, double g)
{
	/* System.cpp:2665-2674 */
	double ratio = ri/rj;
	if (ratio < 2 && ratio > 0.5) {
		return (2+g)*0.5*(ri+rj);
	} else {
		double mr = (ri < rj ? ri : rj);
		return ri+rj+g*mr;
	}
}

/* ======================================================================= */
/* resistance functions (LubricationFunctions.h:12-90)                     */
/* ======================================================================= */
static double g1_XA(double l) { double p = l+1; return 2*l*l/(p*p*p); }
static double g2_YA(double l) { double p = l+1; return (4./15)*l*(2+l+2*l*l)/(p*p*p); }
static double g2_YB(double l) { double p = l+1; return -(1./5)*l*(4+l)/(p*p); }
static double g2_YC(double l) { return (2./5)*l/(l+1); }
static double g4_YC(double l) { double p = l+1; return (4./5)*l*l/(p*p*p*p); }
static double g1_XG(double l) { double p = l+1; return 3*l*l/(p*p*p); }
static double g2_YG(double l) { double p = l+1; return (1./10)*l*(4-l+7*l*l)/(p*p*p); }
static double g2_YH(double l) { double p = l+1; return (1./10)*l*(2-l)/(p*p); }
static double g5_YH(double l) { double p = l+1; return (1./20)*l*l*(1+7*l)/(p*p); }
static double g1_XM(double l) { double p = l+1; return (6./5)*l*l/(p*p*p); }
static double g4_XM(double l) { double p = l+1; return (6./5)*l*l*l/(p*p*p); }
static double g2_YM(double l) { double p = l+1; return (6./25)*l*(1-l+4*l*l)/(p*p*p); }
static double g5_YM(double l) { double p = l+1; return (3./50)*l*l*(7-10*l+7*l*l)/(p*p*p); }

enum { cXA = 0, cYA, cYB, cYC, cXG, cYG, cYH, cXM, cYM };

void orc_lub_constants(double a0, double a1, double c[36])
{
	/* Lubrication::calcLubConstants, Lubrication.cpp:165-294.
	 * out[4*k + m]: k in {XA, YA, YB, YC, XG, YG, YH, XM, YM}, m in 0..3 */
	double ro = a0+a1;
	double ro_12 = ro/2;
	double lambda = a1/a0;
	double invlambda = 1/lambda;
	double lambda_square = lambda*lambda;
	double lambda_cubic = lambda*lambda*lambda;
	double lambda_p_1 = 1+lambda;
	double lambda_p_1_square = lambda_p_1*lambda_p_1;
	double lambda_p_1_cubic = lambda_p_1_square*lambda_p_1;
	double a0a0_23 = a0*a0*(2.0/3);
	double a1a1_23 = a1*a1*(2.0/3);
	double roro_16 = ro*ro*(1.0/6);
	double a0a0a0 = a0*a0*a0;
	double a1a1a1 = a1*a1*a1;
	double rororo = ro*ro*ro;
	double a0a0a0_43 = a0a0a0*(4.0/3);
	double a1a1a1_43 = a1a1a1*(4.0/3);
	double rororo_16 = rororo*(1.0/6);
	double a0a0a0_109 = a0a0a0*(10.0/9);
	double a1a1a1_109 = a1a1a1*(10.0/9);
	double rororo_536 = rororo*(5.0/36);
	double g, gi, g4;
	double *o;

	g = g1_XA(lambda); gi = g1_XA(invlambda);
	o = c+4*cXA; o[0] = a0*g; o[1] = ro_12*(-2/lambda_p_1)*g; o[2] = o[1]; o[3] = a1*gi;
	g = g2_YA(lambda); gi = g2_YA(invlambda);
	o = c+4*cYA; o[0] = a0*g; o[1] = ro_12*(-2/lambda_p_1)*g; o[2] = o[1]; o[3] = a1*gi;
	g = g2_YB(lambda); gi = g2_YB(invlambda);
	o = c+4*cYB; o[0] = a0a0_23*g; o[1] = roro_16*(-4/lambda_p_1_square)*g;
	o[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; o[3] = -a1a1_23*gi;
	g = g2_YC(lambda); gi = g2_YC(invlambda); g4 = g4_YC(lambda);
	o = c+4*cYC; o[0] = a0a0a0_43*g; o[1] = rororo_16*g4; o[2] = o[1]; o[3] = a1a1a1_43*gi;
	g = g1_XG(lambda); gi = g1_XG(invlambda);
	o = c+4*cXG; o[0] = a0a0_23*g; o[1] = roro_16*(-4/lambda_p_1_square)*g;
	o[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; o[3] = -a1a1_23*gi;
	g = g2_YG(lambda); gi = g2_YG(invlambda);
	o = c+4*cYG; o[0] = a0a0_23*g; o[1] = roro_16*(-4/lambda_p_1_square)*g;
	o[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; o[3] = -a1a1_23*gi;
	{
		double g5 = g5_YH(lambda), g5i = g5_YH(invlambda);
		g = g2_YH(lambda); gi = g2_YH(invlambda);
		o = c+4*cYH; o[0] = a0a0a0_43*g; o[1] = rororo_16*(8/lambda_p_1_cubic)*g5;
		o[2] = rororo_16*(8*lambda_cubic/lambda_p_1_cubic)*g5i; o[3] = a1a1a1_43*gi;
	}
	g = g1_XM(lambda); gi = g1_XM(invlambda); g4 = g4_XM(lambda);
	o = c+4*cXM; o[0] = a0a0a0_109*g; o[1] = rororo_536*(8/lambda_p_1_cubic)*g4; o[2] = o[1];
	o[3] = a1a1a1_109*gi;
	{
		double g5 = g5_YM(lambda);
		g = g2_YM(lambda); gi = g2_YM(invlambda);
		o = c+4*cYM; o[0] = a0a0a0_109*g; o[1] = rororo_536*(8/lambda_p_1_cubic)*g5; o[2] = o[1];
		o[3] = a1a1a1_109*gi;
	}
}

/* ContactDashpot::setDashpotResistanceCoeffs + calcDashpotResistances
 * (ContactDashpot.cpp:65-183).  out = XA[4] YA[4] YB[4] YC[4]. */
static double dashpot_tangential_coeff(const orc_params *p)
{
	int friction = p->friction_model != 0;
	if (!friction && p->lubrication_model != 2) return 0;
	if (p->tau_t > 0) return 6*p->kt*p->tau_t;   /* only valid without lubrication */
	if (p->lub_reduce < 1) return log(1/p->lub_reduce);
	return 0;
}
static double dashpot_normal_coeff(const orc_params *p)
{
	if (p->tau_n > 0) return 4*p->kn*p->tau_n;
	return 1/p->lub_reduce;
}

void orc_dashpot_coeffs(double a0, double a1, const orc_params *p, double out[16])
{
	double normal_coeff = dashpot_normal_coeff(p);
	double tangential_coeff = dashpot_tangential_coeff(p);
	double ro = a0+a1, ro_12 = ro/2;
	double lambda = a1/a0;
	double invlambda = 1/lambda;
	double lambda_square = lambda*lambda;
	double lambda_p_1 = 1+lambda;
	double lambda_p_1_square = lambda_p_1*lambda_p_1;
	double a0a0_23 = a0*a0*(2.0/3);
	double a1a1_23 = a1*a1*(2.0/3);
	double roro_16 = ro*ro*(1.0/6);
	double a0a0a0_43 = a0*a0*a0*(4.0/3);
	double a1a1a1_43 = a1*a1*a1*(4.0/3);
	double rororo_16 = ro*ro*ro*(1.0/6);
	double *XA = out, *YA = out+4, *YB = out+8, *YC = out+12;
	int j;
	memset(out, 0, 16*sizeof(double));
	{
		double g = g1_XA(lambda), gi = g1_XA(invlambda);
		XA[0] = a0*g;
		XA[1] = ro_12*(-2/lambda_p_1)*g;
		XA[2] = ro_12*XA[1];
		XA[3] = a1*gi;
		for (j = 0; j < 4; j++) XA[j] *= normal_coeff;
	}
	if (tangential_coeff == 0) return;
	{
		double g = g2_YA(lambda), gi = g2_YA(invlambda);
		YA[0] = a0*g;
		YA[1] = ro_12*(-2/lambda_p_1)*g;
		YA[2] = ro_12*YA[1];
		YA[3] = a1*gi;
		for (j = 0; j < 4; j++) YA[j] *= tangential_coeff;
	}
	{
		double g = g2_YB(lambda), gi = g2_YB(invlambda);
		YB[0] = a0a0_23*g;
		YB[1] = roro_16*(-4/lambda_p_1_square)*g;
		YB[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi;
		YB[3] = -a1a1_23*gi;
		for (j = 0; j < 4; j++) YB[j] *= tangential_coeff;
	}
	{
		double g = g2_YC(lambda), gi = g2_YC(invlambda), g4 = g4_YC(lambda);
		YC[0] = a0a0a0_43*g;
		YC[1] = rororo_16*g4;
		YC[2] = rororo_16*YC[1];
		YC[3] = a1a1a1_43*gi;
		for (j = 0; j < 4; j++) YC[j] *= tangential_coeff;
	}
}

/* ---- block builders -------------------------------------------------- */
/* Lubrication::RFU_ODBlock_squeeze_tangential (Lubrication.cpp:437-467) and
 * ContactDashpot::RFU_ODBlock tangential branch (ContactDashpot.cpp:189-214):
 * the same expression with X/Y coefficient slots 1 and YB slot 2. */
static void od_tangential(double XA1, double YA1, double YB1, double YB2, double YC1,
                          const double n[3], double b[18])
{
	double XA1_YA1 = XA1-YA1;
	b[0] = YA1+XA1_YA1*n[0]*n[0];
	b[1] = XA1_YA1*n[0]*n[1];
	b[2] = XA1_YA1*n[0]*n[2];
	b[3] = -YB2*n[2];
	b[4] = YB2*n[1];
	b[5] = YA1+XA1_YA1*n[1]*n[1];
	b[6] = XA1_YA1*n[1]*n[2];
	b[7] = -YB2*n[0];
	b[8] = YA1+XA1_YA1*n[2]*n[2];
	b[9] = YB1*n[2];
	b[10] = -YB1*n[1];
	b[11] = YC1*(1-n[0]*n[0]);
	b[12] = -YC1*n[0]*n[1];
	b[13] = -YC1*n[0]*n[2];
	b[14] = YB1*n[0];
	b[15] = YC1*(1-n[1]*n[1]);
	b[16] = -YC1*n[1]*n[2];
	b[17] = YC1*(1-n[2]*n[2]);
}
/* normal-only off-diagonal block (Lubrication.cpp:406-434, ContactDashpot.cpp:216-239) */
static void od_normal(double XA1, const double n[3], double b[18])
{
	memset(b, 0, 18*sizeof(double));
	b[0] = XA1*n[0]*n[0];
	b[1] = XA1*n[0]*n[1];
	b[2] = XA1*n[0]*n[2];
	b[5] = XA1*n[1]*n[1];
	b[6] = XA1*n[1]*n[2];
	b[8] = XA1*n[2]*n[2];
}
/* normal-only diagonal block (Lubrication.cpp:470-517, ContactDashpot.cpp:296-337) */
static void d_normal(double XA, const double n[3], double b[15])
{
	memset(b, 0, 15*sizeof(double));
	b[0] = XA*n[0]*n[0];
	b[1] = XA*n[0]*n[1];
	b[2] = XA*n[0]*n[2];
	b[5] = XA*n[1]*n[1];
	b[6] = XA*n[1]*n[2];
	b[8] = XA*n[2]*n[2];
}
/* Lubrication::RFU_DBlocks_squeeze_tangential (Lubrication.cpp:520-580) */
static void d_lub_tangential(double XA, double YA, double YB, double YC, const double n[3], double b[15])
{
	double n0n0 = n[0]*n[0], n0n1 = n[0]*n[1], n0n2 = n[0]*n[2];
	double n1n1 = n[1]*n[1], n1n2 = n[1]*n[2], n2n2 = n[2]*n[2];
	double one_n0n0 = 1-n0n0, one_n1n1 = 1-n1n1, one_n2n2 = 1-n2n2;
	double XAmYA = XA-YA;
	b[0] = XA*n0n0+YA*one_n0n0;
	b[1] = XAmYA*n0n1;
	b[2] = XAmYA*n0n2;
	b[3] = -YB*n[2];
	b[4] = YB*n[1];
	b[5] = XA*n1n1+YA*one_n1n1;
	b[6] = XAmYA*n1n2;
	b[7] = -YB*n[0];
	b[8] = XA*n2n2+YA*one_n2n2;
	b[9] = YC*one_n0n0;
	b[10] = -YC*n0n1;
	b[11] = -YC*n0n2;
	b[12] = YC*one_n1n1;
	b[13] = -YC*n1n2;
	b[14] = YC*one_n2n2;
}
/* ContactDashpot::RFU_DBlocks tangential branch (ContactDashpot.cpp:249-294) */
static void d_dash_tangential(double XA, double YA, double YB, double YC, const double n[3], double b[15])
{
	double XA_YA = XA-YA;
	b[0] = YA+XA_YA*n[0]*n[0];
	b[1] = XA_YA*n[0]*n[1];
	b[2] = XA_YA*n[0]*n[2];
	b[3] = -YB*n[2];
	b[4] = YB*n[1];
	b[5] = YA+XA_YA*n[1]*n[1];
	b[6] = XA_YA*n[1]*n[2];
	b[7] = -YB*n[0];
	b[8] = YA+XA_YA*n[2]*n[2];
	b[9] = YC*(1-n[0]*n[0]);
	b[10] = -YC*n[0]*n[1];
	b[11] = -YC*n[0]*n[2];
	b[12] = YC*(1-n[1]*n[1]);
	b[13] = -YC*n[1]*n[2];
	b[14] = YC*(1-n[2]*n[2]);
}

void orc_lub_blocks(double a0, double a1, double gap, const double n[3], const orc_params *p,
                    double db0[15], double db1[15], double od[18])
{
	double c[36], lc, llc, XA[4], YA[4], YB[4], YC[4];
	int j;
	orc_lub_constants(a0, a1, c);
	/* Lubrication::updateResistanceCoeff (Lubrication.cpp:696-706) */
	lc = 1/(gap+p->lub_reduce);
	llc = lc > 1. ? log(lc) : 0.;
	for (j = 0; j < 4; j++) {
		XA[j] = c[4*cXA+j]*lc;
		YA[j] = c[4*cYA+j]*llc;
		YB[j] = c[4*cYB+j]*llc;
		YC[j] = c[4*cYC+j]*llc;
	}
	if (p->lubrication_model == 2) {
		d_lub_tangential(XA[0], YA[0], YB[0], YC[0], n, db0);
		d_lub_tangential(XA[3], YA[3], YB[3], YC[3], n, db1);
		od_tangential(XA[1], YA[1], YB[1], YB[2], YC[1], n, od);
	} else {
		d_normal(XA[0], n, db0);
		d_normal(XA[3], n, db1);
		od_normal(XA[1], n, od);
	}
}

void orc_dashpot_blocks(double a0, double a1, const double n[3], const orc_params *p,
                        double db0[15], double db1[15], double od[18])
{
	double c[16];
	orc_dashpot_coeffs(a0, a1, p, c);
	if (dashpot_tangential_coeff(p) > 0) {
		d_dash_tangential(c[0], c[4], c[8], c[12], n, db0);
		d_dash_tangential(c[3], c[7], c[11], c[15], n, db1);
		od_tangential(c[1], c[5], c[9], c[10], c[13], n, od);
	} else {
		d_normal(c[0], n, db0);
		d_normal(c[3], n, db1);
		od_normal(c[1], n, od);
	}
}

/* Dense 6x6 (row-major) of a packed diagonal block, from the lower-triangle
 * CSC values written by StokesSolver::insertDBlockValues (StokesSolver.cpp:154-184)
 * with row layout db_layout (StokesSolver.cpp:94-100); symmetric. */
void orc_dblock_dense(const double b[15], double M[36])
{
	double L[36];
	int r, c;
	memset(L, 0, sizeof(L));
#define LO(r, c) L[6*(r)+(c)]
	LO(0,0) = b[0]; LO(1,0) = b[1]; LO(2,0) = b[2]; LO(4,0) = b[3]; LO(5,0) = b[4];
	LO(1,1) = b[5]; LO(2,1) = b[6]; LO(3,1) = -b[3]; LO(5,1) = b[7];
	LO(2,2) = b[8]; LO(3,2) = -b[4]; LO(4,2) = -b[7];
	LO(3,3) = b[9]; LO(4,3) = b[10]; LO(5,3) = b[11];
	LO(4,4) = b[12]; LO(5,4) = b[13];
	LO(5,5) = b[14];
#undef LO
	for (r = 0; r < 6; r++)
		for (c = 0; c < 6; c++)
			M[6*r+c] = r >= c ? L[6*r+c] : L[6*c+r];
}

/* Dense 6x6 (row-major) of a packed off-diagonal block sitting at (row block
 * j, column block i), from StokesSolver::insertODBlockValues
 * (StokesSolver.cpp:186-231) with row layout odb_layout (StokesSolver.cpp:87-93). */
void orc_odblock_dense(const double b[18], double M[36])
{
	memset(M, 0, 36*sizeof(double));
#define O(r, c) M[6*(r)+(c)]
	O(0,0) = b[0]; O(1,0) = b[1]; O(2,0) = b[2]; O(4,0) = b[3]; O(5,0) = b[4];
	O(0,1) = b[1]; O(1,1) = b[5]; O(2,1) = b[6]; O(3,1) = -b[3]; O(5,1) = b[7];
	O(0,2) = b[2]; O(1,2) = b[6]; O(2,2) = b[8]; O(3,2) = -b[4]; O(4,2) = -b[7];
	O(1,3) = b[9]; O(2,3) = b[10]; O(3,3) = b[11]; O(4,3) = b[12]; O(5,3) = b[13];
	O(0,4) = -b[9]; O(2,4) = b[14]; O(3,4) = b[12]; O(4,4) = b[15]; O(5,4) = b[16];
	O(0,5) = -b[10]; O(1,5) = -b[14]; O(3,5) = b[13]; O(4,5) = b[16]; O(5,5) = b[17];
#undef O
}

/* ======================================================================= */
/* CSC export (StokesSolver.cpp:334-379, 154-299, 889-919)                 */
/* ======================================================================= */
static const int db_layout[6][5] = {{0,1,2,4,5},{1,2,3,5},{2,3,4},{3,4,5},{4,5},{5}};
static const int db_cnt[6] = {5,4,3,3,2,1};
static const int odb_layout[6][5] = {{0,1,2,4,5},{0,1,2,3,5},{0,1,2,3,4},{1,2,3,4,5},{0,2,3,4,5},{0,1,3,4,5}};

long orc_csc_nnz(int np, int nod) { return 18L*np+30L*nod; }

int orc_export_csc(int np, const double *dblocks, int nod, const int *cp, const int *rows,
                   const double *odb, int *colp, int *rowi, double *val)
{
	int j, col, k;
	(void)nod;
	for (j = 0; j < np; j++) {
		int ix[6];
		int j6 = 6*j;
		int od_nz = 5*(cp[j+1]-cp[j]);
		const double *b = dblocks+15*j;
		double dv[6][5];
		ix[0] = 18*j+30*cp[j];
		for (col = 1; col < 6; col++) ix[col] = ix[col-1]+db_cnt[col-1]+od_nz;
		for (col = 0; col < 6; col++) colp[j6+col] = ix[col];
		/* insertDBlockValues order */
		dv[0][0] = b[0]; dv[0][1] = b[1]; dv[0][2] = b[2]; dv[0][3] = b[3]; dv[0][4] = b[4];
		dv[1][0] = b[5]; dv[1][1] = b[6]; dv[1][2] = -b[3]; dv[1][3] = b[7];
		dv[2][0] = b[8]; dv[2][1] = -b[4]; dv[2][2] = -b[7];
		dv[3][0] = b[9]; dv[3][1] = b[10]; dv[3][2] = b[11];
		dv[4][0] = b[12]; dv[4][1] = b[13];
		dv[5][0] = b[14];
		for (col = 0; col < 6; col++) {
			int s;
			for (s = 0; s < db_cnt[col]; s++) {
				rowi[ix[col]+s] = j6+db_layout[col][s];
				val[ix[col]+s] = dv[col][s];
			}
			ix[col] += db_cnt[col];
		}
		for (k = cp[j]; k < cp[j+1]; k++) {
			const double *o = odb+18*(long)k;
			int top = 6*rows[k];
			double ov[6][5] = {
				{o[0], o[1], o[2], o[3], o[4]},
				{o[1], o[5], o[6], -o[3], o[7]},
				{o[2], o[6], o[8], -o[4], -o[7]},
				{o[9], o[10], o[11], o[12], o[13]},
				{-o[9], o[14], o[12], o[15], o[16]},
				{-o[10], -o[14], o[13], o[16], o[17]}};
			for (col = 0; col < 6; col++) {
				int s;
				for (s = 0; s < 5; s++) {
					rowi[ix[col]+s] = top+odb_layout[col][s];
					val[ix[col]+s] = ov[col][s];
				}
				ix[col] += 5;
			}
		}
	}
	colp[6*np] = colp[6*np-1]+1;
	return 0;
}

/* ======================================================================= */
/* brute force pair set                                                    */
/* ======================================================================= */
int orc_bruteforce_pairs(int np, const double *pos, const double *rad, const double L[3],
                         const double sd[3], double g, int *p0, int *p1, int cap)
{
	int n = 0, i, j;
	for (i = 0; i < np-1; i++) {
		for (j = i+1; j < np; j++) {
			double d[3] = {pos[3*j]-pos[3*i], pos[3*j+1]-pos[3*i+1], pos[3*j+2]-pos[3*i+2]};
			double sq, range;
			orc_periodize_diff(L, sd, d);
			sq = d[0]*d[0]+d[1]*d[1]+d[2]*d[2];
			range = orc_lub_range(rad[i], rad[j], g);
			if (sq < range*range) {
				if (n < cap) { p0[n] = i; p1[n] = j; }
				n++;
			}
		}
	}
	return n;
}

/* ======================================================================= */
/* system state                                                            */
/* ======================================================================= */
typedef struct {
	int p0, p1;
	double range, ro, a0, a1;
	double rvec[3], nvec[3], r, gap;
	int z_offset;
	/* Contact (Contact.h:36-175) */
	int contact_active, state;
	double kn_s, kt_s;
	double disp_tan[3], prev_disp_tan[3];
	double f_normal_norm, normal_load;
	double f_normal[3], f_tan[3], f_total[3];
	double rsv_sq;
	/* ContactDashpot */
	double dc[16];
	/* Lubrication */
	int lub_active;
	double lc[36];
	double lub_coeff, log_lub_coeff;
	double XA[4], YA[4], YB[4], YC[4], XG[4], YG[4], YH[4];
	/* RepulsiveForce */
	double geo_factor, rep_norm, rep_force[3];
} orc_pair;

/* stress components in LF_DEM's std::map order (Stress.cpp:16-73): the
 * M:E term, GU of the four velocity components, the contact xF (rate
 * control: one slot; stress control: rate independent, rate proportional)
 * and the repulsive xF */
enum { SC_ME = 0, SC_GU = 1, SC_XF_CONTACT = 5, SC_XF_CONTACT_PROP = 6, SC_XF_REP = 7, ORC_NSC = 8 };

struct orc_system {
	orc_params p;
	int np;
	double L[3];
	double *pos, *radius;
	double *vel, *ang, *na_vel, *na_ang, *vel_pred, *ang_pred, *u_inf;
	double *na_disp;                /* System::na_disp (System.h:89) */
	double shear_disp[3], vel_difference[3];
	double E[6], omega[3];
	double dt, time, strain;
	double max_velocity, max_sliding_velocity;
	int in_predictor;
	int friction, repulsive;
	double tangential_coeff;
	/* BoxSet */
	int boxed;
	unsigned nx, ny, nz, yz, nbox;
	double bx, by, bz;
	double top_probe[16][3], bot_probe[16][3];
	int *ctype;                    /* 0 bulk, 1 bottom, 2 top, 3 top-bottom */
	int *static_nb, *n_static;     /* 26 per box */
	int *moving_nb, *n_moving;     /* 32 per box */
	int *box_of;
	int *snap_start, *snap_mem;    /* neighborhood snapshot (members per box) */
	/* pairs, sorted by (p0, p1) */
	orc_pair *pairs;
	int npairs, cap;
	/* scratch */
	double *dblocks, *odblocks;
	int *od_cp, *od_rows;
	int od_cap;
	/* velocity components and stress (System.h na_velo_components,
	 * stress_components, total_stress_groups) */
	int control;                   /* 0 rate, 1 stress */
	double target_stress, init_strain_limit, init_rate_limit;
	double *comp_vel, *comp_ang;   /* ORC_NCOMP x 3 np */
	int comp_present[ORC_NCOMP];
	double *pstress;               /* ORC_NSC x 6 np, stress components per particle */
	double groups[ORC_NGROUP][6], total_stress[6];
	char err[256];
};

const char *orc_last_error(orc_system *s) { return s->err; }

static int set_err(orc_system *s, const char *m) { snprintf(s->err, sizeof(s->err), "%s", m); return -1; }

/* ---- BoxSet (BoxSet.cpp:7-70, 170-306, 408-442, 498-506, 764-823) ---- */
static int which_box(const orc_system *s, const double pos[3], int *label)
{
	unsigned ix = (unsigned)(pos[0]/s->bx);
	unsigned iy = (unsigned)(pos[1]/s->by);
	unsigned iz = (unsigned)(pos[2]/s->bz);
	unsigned l = ix*s->yz+iy*s->nz+iz;
	if (l >= s->nbox) return -1;
	*label = (int)l;
	return 0;
}

static int which_box_periodized(const orc_system *s, const double pos[3], int *label)
{
	double q[3] = {pos[0], pos[1], pos[2]};
	orc_periodize(s->L, s->shear_disp, q);
	return which_box(s, q, label);
}

static void box_center(const orc_system *s, int label, double c[3])
{
	unsigned ix = label/s->yz, iy = (label%s->yz)/s->nz, iz = label%s->nz;
	c[0] = s->bx*(ix+0.5); c[1] = s->by*(iy+0.5); c[2] = s->bz*(iz+0.5);
}

static int boxset_init(orc_system *s, double max_range)
{
	unsigned b;
	s->nx = (unsigned)(s->L[0]/max_range);
	s->ny = (unsigned)(s->L[1]/max_range);
	s->nz = (unsigned)(s->L[2]/max_range);
	if (s->nx == 0) s->nx = 1;
	if (s->ny == 0) s->ny = 1;
	if (s->nz == 0) s->nz = 1;
	if (s->nx < 4 && s->ny < 4 && s->nz < 4) {
		s->boxed = 0;
		s->nx = s->ny = s->nz = 1;
		s->bx = s->L[0]; s->by = s->L[1]; s->bz = s->L[2];
	} else {
		s->boxed = 1;
		s->bx = s->L[0]/s->nx; s->by = s->L[1]/s->ny; s->bz = s->L[2]/s->nz;
	}
	s->yz = s->ny*s->nz;
	s->nbox = s->nx*s->yz;
	s->ctype = calloc(s->nbox, sizeof(int));
	s->static_nb = calloc(26*s->nbox, sizeof(int));
	s->n_static = calloc(s->nbox, sizeof(int));
	s->moving_nb = calloc(32*s->nbox, sizeof(int));
	s->n_moving = calloc(s->nbox, sizeof(int));
	s->snap_start = calloc(s->nbox+1, sizeof(int));
	if (!s->boxed) return 0;
	{
		/* probing positions (BoxSet.cpp:47-61) */
		int m1p1[2] = {-1, 1}, ai, bi, k = 0;
		for (ai = 0; ai < 2; ai++) {
			for (bi = 0; bi < 2; bi++) {
				double a = m1p1[ai], bb = m1p1[bi];
				double fc[3] = {1.4999999*(a*s->bx), 1.4999999*(bb*s->by), 1.4999999*s->bz};
				double *t = s->top_probe[k];
				t[0] = fc[0]; t[1] = fc[1]; t[2] = fc[2];
				t = s->top_probe[k+1]; t[0] = fc[0]-a*s->bx; t[1] = fc[1]-0; t[2] = fc[2]-0;
				t = s->top_probe[k+2]; t[0] = fc[0]-a*s->bx; t[1] = fc[1]-bb*s->by; t[2] = fc[2]-0;
				t = s->top_probe[k+3]; t[0] = fc[0]-0; t[1] = fc[1]-bb*s->by; t[2] = fc[2]-0;
				fc[0] = 1.4999999*(a*s->bx); fc[1] = 1.4999999*(bb*s->by); fc[2] = 1.4999999*(-s->bz);
				t = s->bot_probe[k]; t[0] = fc[0]; t[1] = fc[1]; t[2] = fc[2];
				t = s->bot_probe[k+1]; t[0] = fc[0]-a*s->bx; t[1] = fc[1]-0; t[2] = fc[2]-0;
				t = s->bot_probe[k+2]; t[0] = fc[0]-a*s->bx; t[1] = fc[1]-bb*s->by; t[2] = fc[2]-0;
				t = s->bot_probe[k+3]; t[0] = fc[0]-0; t[1] = fc[1]-bb*s->by; t[2] = fc[2]-0;
				k += 4;
			}
		}
	}
	for (b = 0; b < s->nbox; b++) {
		unsigned iz = b%s->nz;
		int t;
		double c[3];
		int a, bb, cc;
		if (iz == 0 && iz < s->nz-1) t = 1;
		else if (iz == s->nz-1 && iz > 0) t = 2;
		else if (iz == 0 && iz == s->nz-1) t = 3;
		else t = 0;
		s->ctype[b] = t;
		box_center(s, b, c);
		for (a = -1; a <= 1; a++) {
			for (bb = -1; bb <= 1; bb++) {
				int clo = (t == 0) ? -1 : (t == 2 ? -1 : 0);
				int chi = (t == 0) ? 1 : (t == 1 ? 1 : 0);
				for (cc = clo; cc <= chi; cc++) {
					double q[3] = {c[0]+a*s->bx, c[1]+bb*s->by, c[2]+cc*s->bz};
					int l;
					if (which_box_periodized(s, q, &l)) return -1;
					if (l != (int)b) s->static_nb[26*b+s->n_static[b]++] = l;
				}
			}
		}
	}
	return 0;
}

/* BoxSet::updateNeighbors (BoxSet.cpp:408-442) */
static int boxset_update_moving(orc_system *s)
{
	unsigned b;
	int k;
	if (!s->boxed) return 0;
	for (b = 0; b < s->nbox; b++) {
		int t = s->ctype[b];
		double c[3];
		s->n_moving[b] = 0;
		if (t == 0) continue;
		box_center(s, b, c);
		if (t == 2 || t == 3) {
			for (k = 0; k < 16; k++) {
				double q[3] = {c[0]+s->top_probe[k][0], c[1]+s->top_probe[k][1], c[2]+s->top_probe[k][2]};
				int l;
				if (which_box_periodized(s, q, &l)) return -1;
				if (l != (int)b) s->moving_nb[32*b+s->n_moving[b]++] = l;
			}
		}
		if (t == 1 || t == 3) {
			for (k = 0; k < 16; k++) {
				double q[3] = {c[0]+s->bot_probe[k][0], c[1]+s->bot_probe[k][1], c[2]+s->bot_probe[k][2]};
				int l;
				if (which_box_periodized(s, q, &l)) return -1;
				if (l != (int)b) s->moving_nb[32*b+s->n_moving[b]++] = l;
			}
		}
	}
	return 0;
}

/* BoxSet::update (BoxSet.cpp:498-506): moving neighbours, then the per-box
 * neighbourhood snapshot (Box.cpp:45-87).  The snapshot is kept as the box
 * membership at this instant; neighbourhood(i) is assembled from it. */
int orc_boxset_update(orc_system *s)
{
	unsigned b;
	int i;
	if (boxset_update_moving(s)) return set_err(s, "BoxSet: probe out of boundaries");
	memset(s->snap_start, 0, (s->nbox+1)*sizeof(int));
	for (i = 0; i < s->np; i++) s->snap_start[s->box_of[i]+1]++;
	for (b = 0; b < s->nbox; b++) s->snap_start[b+1] += s->snap_start[b];
	{
		int *fill = malloc(s->nbox*sizeof(int));
		for (b = 0; b < s->nbox; b++) fill[b] = s->snap_start[b];
		for (i = 0; i < s->np; i++) s->snap_mem[fill[s->box_of[i]]++] = i;   /* ascending, like std::set<int> */
		free(fill);
	}
	return 0;
}

static int box_particle(orc_system *s, int i)
{
	/* BoxSet::box (BoxSet.cpp:812-823) */
	int l;
	if (which_box(s, s->pos+3*i, &l)) return set_err(s, "BoxSet: trying to box position out of boundaries");
	s->box_of[i] = l;
	return 0;
}

int orc_get_box_info(const orc_system *s, int *dims, int *labels)
{
	dims[0] = (int)s->nx; dims[1] = (int)s->ny; dims[2] = (int)s->nz; dims[3] = s->boxed;
	if (labels) memcpy(labels, s->box_of, s->np*sizeof(int));
	return (int)s->nbox;
}

int orc_get_neighbor_cells(const orc_system *s, int cell, int *out, int cap)
{
	int n = 0, k;
	if (n < cap) out[n] = cell;
	n++;
	for (k = 0; k < s->n_static[cell]; k++) { if (n < cap) out[n] = s->static_nb[26*cell+k]; n++; }
	for (k = 0; k < s->n_moving[cell]; k++) { if (n < cap) out[n] = s->moving_nb[32*cell+k]; n++; }
	return n;
}

/* ---- pair lifecycle --------------------------------------------------- */
static void pair_geometry(orc_system *s, orc_pair *q)
{
	/* Interaction::calcNormalVectorDistanceGap (Interaction.cpp:255-266) */
	const double *x0 = s->pos+3*q->p0, *x1 = s->pos+3*q->p1;
	q->rvec[0] = x1[0]-x0[0]; q->rvec[1] = x1[1]-x0[1]; q->rvec[2] = x1[2]-x0[2];
	q->z_offset = orc_periodize_diff(s->L, s->shear_disp, q->rvec);
	q->r = sqrt(q->rvec[0]*q->rvec[0]+q->rvec[1]*q->rvec[1]+q->rvec[2]*q->rvec[2]);
	q->nvec[0] = q->rvec[0]/q->r; q->nvec[1] = q->rvec[1]/q->r; q->nvec[2] = q->rvec[2]/q->r;
	q->gap = 2*q->r/q->ro-2;
}

static void contact_activate(orc_system *s, orc_pair *q)
{
	/* Contact::activate (Contact.cpp:73-101) */
	q->contact_active = 1;
	q->f_normal_norm = 0;
	q->f_normal[0] = q->f_normal[1] = q->f_normal[2] = 0;
	q->normal_load = 0;
	q->f_total[0] = q->f_total[1] = q->f_total[2] = 0;
	if (s->friction) {
		q->state = 2;
		memset(q->disp_tan, 0, sizeof(q->disp_tan));
		memset(q->prev_disp_tan, 0, sizeof(q->prev_disp_tan));
	} else {
		q->state = 1;
	}
}

static void contact_deactivate(orc_system *s, orc_pair *q)
{
	/* Contact::deactivate (Contact.cpp:103-117) */
	q->state = 0;
	q->contact_active = 0;
	q->f_normal_norm = 0;
	memset(q->f_normal, 0, sizeof(q->f_normal));
	memset(q->f_total, 0, sizeof(q->f_total));
	if (s->friction) {
		memset(q->disp_tan, 0, sizeof(q->disp_tan));
		memset(q->f_tan, 0, sizeof(q->f_tan));
	}
}

static void lub_update_coeff(orc_system *s, orc_pair *q)
{
	/* Lubrication::updateResistanceCoeff + calcXYFunctions (Lubrication.cpp:696-706, 297-322) */
	int j;
	if (q->gap > 0) {
		double coeff = 1/(q->gap+s->p.lub_reduce);
		q->lub_coeff = coeff;
		q->log_lub_coeff = coeff > 1. ? log(coeff) : 0.;
		for (j = 0; j < 4; j++) {
			q->XA[j] = q->lc[4*cXA+j]*q->lub_coeff;
			q->XG[j] = q->lc[4*cXG+j]*q->lub_coeff;
		}
		if (s->p.lubrication_model == 2) {
			for (j = 0; j < 4; j++) {
				q->YA[j] = q->lc[4*cYA+j]*q->log_lub_coeff;
				q->YB[j] = q->lc[4*cYB+j]*q->log_lub_coeff;
				q->YC[j] = q->lc[4*cYC+j]*q->log_lub_coeff;
				q->YG[j] = q->lc[4*cYG+j]*q->log_lub_coeff;
				q->YH[j] = q->lc[4*cYH+j]*q->log_lub_coeff;
			}
		}
	}
}

static void lub_update_activation(orc_pair *q)
{
	/* Lubrication::updateActivationState (Lubrication.cpp:137-146) */
	q->lub_active = (q->r < q->range && q->r > q->ro);
}

static void pair_create(orc_system *s, orc_pair *q, int i, int j, double range)
{
	/* Interaction ctor + activateForceMembers (Interaction.cpp:13-36, 137-156) */
	memset(q, 0, sizeof(*q));
	q->p0 = i < j ? i : j;
	q->p1 = i < j ? j : i;
	q->range = range;
	q->a0 = s->radius[q->p0];
	q->a1 = s->radius[q->p1];
	q->ro = q->a0+q->a1;
	/* RepulsiveForce::activate (RepulsiveForce.cpp:23-40) */
	q->geo_factor = q->a0*q->a1/(q->a0+q->a1);
	pair_geometry(s, q);
	/* Contact::setInteractionData / setSpringConstants (Contact.cpp:35-71) */
	{
		double ro_12 = (q->a0+q->a1)/2;
		q->kn_s = ro_12*ro_12*s->p.kn;
		q->kt_s = s->friction ? ro_12*s->p.kt : 0;
	}
	orc_dashpot_coeffs(q->a0, q->a1, &s->p, q->dc);
	if (q->gap <= 0) contact_activate(s, q);
	/* Lubrication::setParticleData + activation (Lubrication.cpp:112-123) */
	orc_lub_constants(q->a0, q->a1, q->lc);
	lub_update_activation(q);
	if (q->lub_active) lub_update_coeff(s, q);
}

static void contact_increment(orc_system *s, orc_pair *q)
{
	/* Contact::incrementTangentialDisplacement (Contact.cpp:125-160) */
	const double *vd = s->vel_difference;
	double zo = (double)q->z_offset;
	double off[3] = {zo*vd[0], zo*vd[1], zo*vd[2]};
	const double *v0 = s->vel+3*q->p0, *v1 = s->vel+3*q->p1;
	const double *w0 = s->ang+3*q->p0, *w1 = s->ang+3*q->p1;
	double tr[3], aw[3], rot[3], rsv[3], d;
	tr[0] = v1[0]-v0[0]+off[0]; tr[1] = v1[1]-v0[1]+off[1]; tr[2] = v1[2]-v0[2]+off[2];
	aw[0] = q->a0*w0[0]+q->a1*w1[0]; aw[1] = q->a0*w0[1]+q->a1*w1[1]; aw[2] = q->a0*w0[2]+q->a1*w1[2];
	cross3(aw, q->nvec, rot);
	rot[0] = -rot[0]; rot[1] = -rot[1]; rot[2] = -rot[2];
	rsv[0] = tr[0]+rot[0]; rsv[1] = tr[1]+rot[1]; rsv[2] = tr[2]+rot[2];
	d = dot3(rsv, q->nvec);
	rsv[0] -= d*q->nvec[0]; rsv[1] -= d*q->nvec[1]; rsv[2] -= d*q->nvec[2];
	q->rsv_sq = rsv[0]*rsv[0]+rsv[1]*rsv[1]+rsv[2]*rsv[2];
	if (s->in_predictor) memcpy(q->prev_disp_tan, q->disp_tan, sizeof(q->disp_tan));
	q->disp_tan[0] = q->prev_disp_tan[0]+rsv[0]*s->dt;
	q->disp_tan[1] = q->prev_disp_tan[1]+rsv[1]*s->dt;
	q->disp_tan[2] = q->prev_disp_tan[2]+rsv[2]*s->dt;
}

static void contact_spring(orc_system *s, orc_pair *q)
{
	/* Contact::calcContactSpringForce + frictionlaw_standard (Contact.cpp:200-233, 255-298) */
	const double *n = q->nvec;
	if (!q->contact_active) return;
	q->f_normal_norm = -q->kn_s*q->gap;
	q->f_normal[0] = -q->f_normal_norm*n[0];
	q->f_normal[1] = -q->f_normal_norm*n[1];
	q->f_normal[2] = -q->f_normal_norm*n[2];
	if (s->friction) {
		double d = dot3(q->disp_tan, n), sq, sup;
		q->disp_tan[0] -= d*n[0]; q->disp_tan[1] -= d*n[1]; q->disp_tan[2] -= d*n[2];
		q->f_tan[0] = q->kt_s*q->disp_tan[0];
		q->f_tan[1] = q->kt_s*q->disp_tan[1];
		q->f_tan[2] = q->kt_s*q->disp_tan[2];
		sq = q->f_tan[0]*q->f_tan[0]+q->f_tan[1]*q->f_tan[1]+q->f_tan[2]*q->f_tan[2];
		q->normal_load = q->f_normal_norm;
		sup = (q->state == 2) ? s->p.mu_static*q->normal_load : s->p.mu_dynamic*q->normal_load;
		if (sq > sup*sup) {
			q->state = 3;
			sup = s->p.mu_dynamic*q->normal_load;
		} else {
			q->state = 2;
		}
		if (q->state == 3) {
			double f = sup/sqrt(sq);
			q->disp_tan[0] *= f; q->disp_tan[1] *= f; q->disp_tan[2] *= f;
			q->f_tan[0] = q->kt_s*q->disp_tan[0];
			q->f_tan[1] = q->kt_s*q->disp_tan[1];
			q->f_tan[2] = q->kt_s*q->disp_tan[2];
		}
	}
	if (q->state <= 1) {
		memcpy(q->f_total, q->f_normal, sizeof(q->f_total));
	} else {
		q->f_total[0] = q->f_normal[0]+q->f_tan[0];
		q->f_total[1] = q->f_normal[1]+q->f_tan[1];
		q->f_total[2] = q->f_normal[2]+q->f_tan[2];
	}
}

static void repulsion_force(orc_system *s, orc_pair *q)
{
	/* RepulsiveForce::calcReducedForceNorm + calcScaledForce (RepulsiveForce.cpp:42-85),
	 * no vdW, no cutoff (repulsive_max_length = -1) */
	double gap = q->r-q->ro, red;
	if (gap > 0) {
		red = exp(-gap/s->p.repulsive_length);
		red *= q->geo_factor;
	} else {
		red = q->geo_factor;
	}
	q->rep_norm = s->p.repulsion*red;
	q->rep_force[0] = -q->rep_norm*q->nvec[0];
	q->rep_force[1] = -q->rep_norm*q->nvec[1];
	q->rep_force[2] = -q->rep_norm*q->nvec[2];
}

/* Interaction::updateState + updateContactState (Interaction.cpp:171-236);
 * returns 1 if the pair leaves (r > range) */
static int pair_update(orc_system *s, orc_pair *q)
{
	if (q->contact_active && s->friction) contact_increment(s, q);
	pair_geometry(s, q);
	if (q->r > q->range) {
		if (q->contact_active) contact_deactivate(s, q);
		q->lub_active = 0;
		return 1;
	}
	if (q->contact_active) {
		if (q->gap > 0) contact_deactivate(s, q);
	} else {
		if (q->gap <= 0) contact_activate(s, q);
	}
	if (q->contact_active) contact_spring(s, q);
	lub_update_activation(q);
	if (q->lub_active) lub_update_coeff(s, q);
	if (s->repulsive) repulsion_force(s, q);
	return 0;
}

int orc_update_interactions(orc_system *s)
{
	/* System::updateInteractions (System.cpp:1375-1403) */
	double sq_max = 0;
	int k, w = 0;
	for (k = 0; k < s->npairs; k++) {
		orc_pair *q = &s->pairs[k];
		int drop = pair_update(s, q);
		if (s->friction && q->contact_active) {
			if (q->rsv_sq > sq_max) sq_max = q->rsv_sq;
		}
		if (!drop) {
			if (w != k) s->pairs[w] = *q;
			w++;
		}
	}
	s->npairs = w;
	s->max_sliding_velocity = sqrt(sq_max);
	return 0;
}

static int pair_find(const orc_system *s, int i, int j)
{
	int lo = 0, hi = s->npairs-1;
	while (lo <= hi) {
		int mid = (lo+hi)/2;
		const orc_pair *q = &s->pairs[mid];
		if (q->p0 == i && q->p1 == j) return mid;
		if (q->p0 < i || (q->p0 == i && q->p1 < j)) lo = mid+1;
		else hi = mid-1;
	}
	return -1;
}

static int cmp_pair(const void *a, const void *b)
{
	const orc_pair *x = a, *y = b;
	if (x->p0 != y->p0) return x->p0 < y->p0 ? -1 : 1;
	if (x->p1 != y->p1) return x->p1 < y->p1 ? -1 : 1;
	return 0;
}

int orc_check_new_interaction(orc_system *s)
{
	/* System::checkNewInteraction (System.cpp:1328-1353) over
	 * BoxSet::neighborhood(i) = the snapshot of boxMap[i]'s neighbourhood
	 * (Box.cpp:45-87): own box, static neighbours, moving neighbours.
	 * New pairs are collected aside and merged once, so the binary search
	 * over the existing (sorted) pairs stays valid during the scan. */
	int i, added = 0, ncap = 256;
	orc_pair *fresh = malloc(ncap*sizeof(orc_pair));
	for (i = 0; i < s->np-1; i++) {
		int cells[1+26+32], nc, c, row0 = added;
		nc = orc_get_neighbor_cells(s, s->box_of[i], cells, 59);
		for (c = 0; c < nc; c++) {
			int b = cells[c], m;
			for (m = s->snap_start[b]; m < s->snap_start[b+1]; m++) {
				int j = s->snap_mem[m], k, dup = 0;
				double d[3], sq, range;
				if (j <= i) continue;
				if (pair_find(s, i, j) >= 0) continue;
				for (k = row0; k < added; k++) if (fresh[k].p1 == j) { dup = 1; break; }
				if (dup) continue;
				d[0] = s->pos[3*j]-s->pos[3*i];
				d[1] = s->pos[3*j+1]-s->pos[3*i+1];
				d[2] = s->pos[3*j+2]-s->pos[3*i+2];
				orc_periodize_diff(s->L, s->shear_disp, d);
				sq = d[0]*d[0]+d[1]*d[1]+d[2]*d[2];
				range = orc_lub_range(s->radius[i], s->radius[j], s->p.lub_max_gap);
				if (sq < range*range) {
					if (added == ncap) { ncap *= 2; fresh = realloc(fresh, ncap*sizeof(orc_pair)); }
					pair_create(s, &fresh[added], i, j, range);
					added++;
				}
			}
		}
	}
	if (added) {
		if (s->npairs+added > s->cap) {
			s->cap = s->npairs+added+1024;
			s->pairs = realloc(s->pairs, s->cap*sizeof(orc_pair));
		}
		memcpy(s->pairs+s->npairs, fresh, added*sizeof(orc_pair));
		s->npairs += added;
		qsort(s->pairs, s->npairs, sizeof(orc_pair), cmp_pair);
	}
	free(fresh);
	return added;
}

/* ---- matrix build ----------------------------------------------------- */
static void dblock_add(double *d, const double *b) { int k; for (k = 0; k < 15; k++) d[k] += b[k]; }

int orc_build_blocks(orc_system *s, double *dblocks, int *cp, int *rows, double *odb, int cap)
{
	/* System::buildResistanceMatrix (System.cpp:1429-1473) with the drag seed
	 * (System.cpp:558-575) and StokesSolver::addResistanceBlocks column
	 * protocol (StokesSolver.cpp:121-152, 921-948).  Pairs are visited by
	 * column i ascending, partner j ascending. */
	int i, k = 0, n = 0;
	for (i = 0; i < s->np; i++) {
		double *d = dblocks+15*i;
		double a = s->radius[i];
		double FU = s->p.sd_coeff*a;
		double TW = s->p.sd_coeff*(a*a*a)*4.0/3;
		memset(d, 0, 15*sizeof(double));
		d[0] = FU; d[5] = FU; d[8] = FU; d[9] = TW; d[12] = TW; d[14] = TW;
	}
	for (i = 0; i < s->np; i++) {
		cp[i] = n;
		while (k < s->npairs && s->pairs[k].p0 == i) {
			orc_pair *q = &s->pairs[k];
			double db0[15], db1[15], od[18];
			int has = q->contact_active || q->lub_active;
			if (has) {
				if (q->contact_active) {
					/* ContactDashpot blocks (ContactDashpot.cpp:186-340) */
					const double *c = q->dc;
					if (s->tangential_coeff > 0) {
						d_dash_tangential(c[0], c[4], c[8], c[12], q->nvec, db0);
						d_dash_tangential(c[3], c[7], c[11], c[15], q->nvec, db1);
						od_tangential(c[1], c[5], c[9], c[10], c[13], q->nvec, od);
					} else {
						d_normal(c[0], q->nvec, db0);
						d_normal(c[3], q->nvec, db1);
						od_normal(c[1], q->nvec, od);
					}
				} else {
					if (s->p.lubrication_model == 2) {
						d_lub_tangential(q->XA[0], q->YA[0], q->YB[0], q->YC[0], q->nvec, db0);
						d_lub_tangential(q->XA[3], q->YA[3], q->YB[3], q->YC[3], q->nvec, db1);
						od_tangential(q->XA[1], q->YA[1], q->YB[1], q->YB[2], q->YC[1], q->nvec, od);
					} else {
						d_normal(q->XA[0], q->nvec, db0);
						d_normal(q->XA[3], q->nvec, db1);
						od_normal(q->XA[1], q->nvec, od);
					}
				}
				dblock_add(dblocks+15*q->p0, db0);
				dblock_add(dblocks+15*q->p1, db1);
				if (n < cap) {
					rows[n] = q->p1;
					memcpy(odb+18*(long)n, od, sizeof(od));
				}
				n++;
			}
			k++;
		}
	}
	cp[s->np] = n;
	return n;
}

/* ---- forces / RHS ----------------------------------------------------- */
static void compute_uinf(orc_system *s)
{
	/* System::computeUInf (System.cpp:2357-2367) */
	int i;
	for (i = 0; i < s->np; i++) {
		double e[3], w[3];
		symdot(s->E, s->pos+3*i, e);
		cross3(s->omega, s->pos+3*i, w);
		s->u_inf[3*i] = e[0]+w[0];
		s->u_inf[3*i+1] = e[1]+w[1];
		s->u_inf[3*i+2] = e[2]+w[2];
	}
}

static void add3(double *a, const double *b) { a[0] += b[0]; a[1] += b[1]; a[2] += b[2]; }
static void sub3(double *a, const double *b) { a[0] -= b[0]; a[1] -= b[1]; a[2] -= b[2]; }

/* Force on p0 from the dashpot at the given (total or non-affine) velocities
 * (ContactDashpot.cpp:342-385; getForceOnP0_nonaffine :387-424 drops the LE
 * velocity jump). */
static void dashpot_force(const orc_system *s, const orc_pair *q, const double *v0, const double *v1,
                          const double *w0, const double *w1, int nonaffine, double f[3])
{
	const double *n = q->nvec, *c = q->dc;
	double vi[3], vj[3], t[3], d;
	memcpy(vi, v0, sizeof(vi));
	memcpy(vj, v1, sizeof(vj));
	if (!nonaffine) {
		double zo = (double)q->z_offset;
		vj[0] += zo*s->vel_difference[0];
		vj[1] += zo*s->vel_difference[1];
		vj[2] += zo*s->vel_difference[2];
	}
	t[0] = c[0]*vi[0]+c[1]*vj[0];
	t[1] = c[0]*vi[1]+c[1]*vj[1];
	t[2] = c[0]*vi[2]+c[1]*vj[2];
	d = -dot3(t, n);
	f[0] = d*n[0]; f[1] = d*n[1]; f[2] = d*n[2];
	if (s->tangential_coeff > 0) {
		double di = dot3(n, vi), dj = dot3(n, vj), a[3], b[3], ci[3], cj[3];
		a[0] = vi[0]-n[0]*di; a[1] = vi[1]-n[1]*di; a[2] = vi[2]-n[2]*di;
		b[0] = vj[0]-n[0]*dj; b[1] = vj[1]-n[1]*dj; b[2] = vj[2]-n[2]*dj;
		f[0] += -c[4]*a[0]-c[5]*b[0];
		f[1] += -c[4]*a[1]-c[5]*b[1];
		f[2] += -c[4]*a[2]-c[5]*b[2];
		cross3(n, w0, ci);
		cross3(n, w1, cj);
		f[0] += -c[8]*ci[0]-c[10]*cj[0];
		f[1] += -c[8]*ci[1]-c[10]*cj[1];
		f[2] += -c[8]*ci[2]-c[10]*cj[2];
	}
}

/* Per-particle force F and torque T (3 np each, zeroed here) of one force
 * component (System.cpp:169-215): ORC_CONTACT, ORC_DASHPOT, ORC_HYDRO,
 * ORC_REPULSION.  Returns 1 when the component carries a torque, 0 when not,
 * -1 when the component is not declared for this system. */
static int component_forces(orc_system *s, int comp, double *F, double *T)
{
	int np = s->np, k;
	memset(F, 0, 3*(size_t)np*sizeof(double));
	memset(T, 0, 3*(size_t)np*sizeof(double));
	if (comp == ORC_CONTACT) {
		/* System::setContactForceToParticle (System.cpp:1800-1814; Contact.cpp:411-437) */
		for (k = 0; k < s->npairs; k++) {
			orc_pair *q = &s->pairs[k];
			if (!q->contact_active) continue;
			add3(F+3*q->p0, q->f_total);
			sub3(F+3*q->p1, q->f_total);
			if (q->state >= 2) {
				double t[3], t0[3], t1[3];
				cross3(q->nvec, q->f_tan, t);
				t0[0] = q->a0*t[0]; t0[1] = q->a0*t[1]; t0[2] = q->a0*t[2];
				t1[0] = q->a1*t[0]; t1[1] = q->a1*t[1]; t1[2] = q->a1*t[2];
				add3(T+3*q->p0, t0);
				add3(T+3*q->p1, t1);
			}
		}
		return s->friction ? 1 : 0;
	}
	if (comp == ORC_DASHPOT) {
		/* R_FU^dash U_inf (System.cpp:1727-1750; ContactDashpot.cpp:426-441) */
		if (s->p.shear_rate == 0) return -1;
		for (k = 0; k < s->npairs; k++) {
			orc_pair *q = &s->pairs[k];
			double f[3];
			if (!q->contact_active) continue;
			dashpot_force(s, q, s->u_inf+3*q->p0, s->u_inf+3*q->p1, s->omega, s->omega, 0, f);
			add3(F+3*q->p0, f);
			sub3(F+3*q->p1, f);
			if (s->tangential_coeff > 0) {
				double tq[3], t0[3], t1[3], ratio = q->a1/q->a0;
				cross3(q->nvec, f, tq);
				t0[0] = q->a0*tq[0]; t0[1] = q->a0*tq[1]; t0[2] = q->a0*tq[2];
				t1[0] = ratio*t0[0]; t1[1] = ratio*t0[1]; t1[2] = ratio*t0[2];
				add3(T+3*q->p0, t0);
				add3(T+3*q->p1, t1);
			}
		}
		return 1;
	}
	if (comp == ORC_HYDRO) {
		/* R_FE:E_inf (Lubrication.cpp:378-403 / 332-355, System.cpp:1752-1798) */
		if (s->p.shear_rate == 0) return -1;
		for (k = 0; k < s->npairs; k++) {
			orc_pair *q = &s->pairs[k];
			const double *n = q->nvec;
			double En[3], nnE, GEi[3], GEj[3];
			if (!q->lub_active) continue;
			symdot(s->E, n, En);
			nnE = dot3(n, En);
			if (s->p.lubrication_model == 2) {
				double YG0_YG2 = q->YG[0]+q->YG[2];
				double YG1_YG3 = q->YG[1]+q->YG[3];
				double cGE_i = (q->XG[0]+q->XG[2]-2*YG0_YG2)*nnE;
				double cGE_j = (q->XG[1]+q->XG[3]-2*YG1_YG3)*nnE;
				double cHE_i = q->YH[0]+q->YH[2];
				double cHE_j = q->YH[1]+q->YH[3];
				double x[3], HEi[3], HEj[3], hi, hj, gi = YG0_YG2*2, gj = YG1_YG3*2;
				int m;
				for (m = 0; m < 3; m++) {
					GEi[m] = cGE_i*n[m]+gi*En[m];
					GEj[m] = cGE_j*n[m]+gj*En[m];
				}
				cross3(En, n, x);
				hi = -2*cHE_i; hj = -2*cHE_j;
				for (m = 0; m < 3; m++) { HEi[m] = hi*x[m]; HEj[m] = hj*x[m]; }
				add3(F+3*q->p0, GEi); add3(F+3*q->p1, GEj);
				add3(T+3*q->p0, HEi); add3(T+3*q->p1, HEj);
			} else {
				double ci = (q->XG[0]+q->XG[2])*nnE, cj = (q->XG[1]+q->XG[3])*nnE;
				GEi[0] = ci*n[0]; GEi[1] = ci*n[1]; GEi[2] = ci*n[2];
				GEj[0] = cj*n[0]; GEj[1] = cj*n[1]; GEj[2] = cj*n[2];
				add3(F+3*q->p0, GEi); add3(F+3*q->p1, GEj);
			}
		}
		return s->p.lubrication_model == 2 ? 1 : 0;
	}
	if (comp == ORC_REPULSION) {
		/* System::setRepulsiveForceToParticle (System.cpp:1816-1828; RepulsiveForce.cpp:101-105) */
		if (!s->repulsive) return -1;
		for (k = 0; k < s->npairs; k++) {
			orc_pair *q = &s->pairs[k];
			add3(F+3*q->p0, q->rep_force);
			sub3(F+3*q->p1, q->rep_force);
		}
		return 0;
	}
	return -1;
}

int orc_build_rhs(orc_system *s, double *rhs)
{
	/* System::computeVelocityWithoutComponents: the force components in
	 * std::map order contact, dashpot, hydro, repulsion added into the RHS
	 * (System.cpp:1932-1945, 1873-1885; StokesSolver.cpp:513-549) */
	int np = s->np, i, comp;
	double *F = malloc(6*(size_t)np*sizeof(double)), *T = F+3*(size_t)np;
	compute_uinf(s);
	memset(rhs, 0, 6*(size_t)np*sizeof(double));
	for (comp = 0; comp < ORC_NCOMP; comp++) {
		int tq = component_forces(s, comp, F, T);
		if (tq < 0) continue;
		for (i = 0; i < np; i++) {
			add3(rhs+6*i, F+3*i);
			if (tq) add3(rhs+6*i+3, T+3*i);
		}
	}
	free(F);
	return 0;
}

int orc_build_rhs_component(orc_system *s, int comp, double *rhs)
{
	/* System::setSolverRHS (System.cpp:1858-1871): force and torque set,
	 * torque zero for a component without torque; U_inf of the current
	 * positions first, as computeVelocities does (System.cpp:2290-2292) */
	int np = s->np, i, tq;
	double *F = malloc(6*(size_t)np*sizeof(double)), *T = F+3*(size_t)np;
	compute_uinf(s);
	tq = component_forces(s, comp, F, T);
	memset(rhs, 0, 6*(size_t)np*sizeof(double));
	if (tq >= 0)
		for (i = 0; i < np; i++) {
			memcpy(rhs+6*i, F+3*i, 3*sizeof(double));
			if (tq) memcpy(rhs+6*i+3, T+3*i, 3*sizeof(double));
		}
	free(F);
	return tq < 0 ? -1 : 0;
}

/* ======================================================================= */
/* sparse Cholesky: minimum degree ordering + left-looking 6x6 blocks      */
/* ======================================================================= */
struct orc_chol {
	int n;
	int *perm, *iperm;
	int *cp, *ri;          /* pattern of L below the diagonal, new ordering, sorted */
	double *Ld, *Lo;       /* diagonal (36 per column) and off-diagonal (36 per entry) blocks, row-major */
	double flops;
};

static int chol6(double *A)
{
	/* in-place dense Cholesky of a row-major 6x6 SPD block, lower triangle */
	int i, j, k;
	for (j = 0; j < 6; j++) {
		double d = A[6*j+j];
		for (k = 0; k < j; k++) d -= A[6*j+k]*A[6*j+k];
		if (!(d > 0)) return -1;
		d = sqrt(d);
		A[6*j+j] = d;
		for (i = j+1; i < 6; i++) {
			double v = A[6*i+j];
			for (k = 0; k < j; k++) v -= A[6*i+k]*A[6*j+k];
			A[6*i+j] = v/d;
		}
		for (i = 0; i < j; i++) A[6*i+j] = 0;
	}
	return 0;
}

orc_chol *orc_chol_factor(int np, const double *dblocks, const int *cp, const int *rows,
                          const double *odb, int *status)
{
	const double t_sym = ph_now();
	orc_chol *c = calloc(1, sizeof(orc_chol));
	int n = np, words = (np+63)/64;
	uint64_t *adj = calloc((size_t)n*words, sizeof(uint64_t));
	int *deg = calloc(n, sizeof(int)), *elim = calloc(n, sizeof(int));
	int *nb = malloc(n*sizeof(int));
	int *pat_start, *pat;            /* pattern per elimination step (original indices) */
	size_t pat_cap = 16*(size_t)n, pat_n = 0;
	int i, k, step;
	*status = 0;
	c->n = n;
	c->perm = malloc(n*sizeof(int));
	c->iperm = malloc(n*sizeof(int));
	pat_start = malloc((n+1)*sizeof(int));
	pat = malloc(pat_cap*sizeof(int));
	for (i = 0; i < n; i++) {
		for (k = cp[i]; k < cp[i+1]; k++) {
			int j = rows[k];
			adj[(size_t)i*words+j/64] |= 1ULL << (j%64);
			adj[(size_t)j*words+i/64] |= 1ULL << (i%64);
		}
	}
	for (i = 0; i < n; i++) {
		int w, d = 0;
		for (w = 0; w < words; w++) d += __builtin_popcountll(adj[(size_t)i*words+w]);
		deg[i] = d;
	}
	/* minimum degree elimination on the explicit elimination graph */
	for (step = 0; step < n; step++) {
		int v = -1, dv = 0x7fffffff, cnt = 0, w, a, b;
		for (i = 0; i < n; i++) if (!elim[i] && deg[i] < dv) { dv = deg[i]; v = i; }
		c->perm[step] = v;
		elim[v] = 1;
		for (w = 0; w < words; w++) {
			uint64_t m = adj[(size_t)v*words+w];
			while (m) {
				int bit = __builtin_ctzll(m);
				nb[cnt++] = 64*w+bit;
				m &= m-1;
			}
		}
		pat_start[step] = (int)pat_n;
		if (pat_n+cnt > pat_cap) { while (pat_n+cnt > pat_cap) pat_cap *= 2; pat = realloc(pat, pat_cap*sizeof(int)); }
		memcpy(pat+pat_n, nb, cnt*sizeof(int));
		pat_n += cnt;
		for (a = 0; a < cnt; a++) {
			uint64_t *ra = adj+(size_t)nb[a]*words;
			ra[v/64] &= ~(1ULL << (v%64));
			for (b = 0; b < cnt; b++) {
				int u = nb[b];
				if (u != nb[a]) ra[u/64] |= 1ULL << (u%64);
			}
		}
		for (a = 0; a < cnt; a++) {
			int d = 0;
			for (w = 0; w < words; w++) d += __builtin_popcountll(adj[(size_t)nb[a]*words+w]);
			deg[nb[a]] = d;
		}
	}
	pat_start[n] = (int)pat_n;
	free(adj); free(deg); free(elim);
	for (i = 0; i < n; i++) c->iperm[c->perm[i]] = i;
	/* pattern in the new ordering, sorted ascending */
	c->cp = malloc((n+1)*sizeof(int));
	c->ri = malloc((pat_n ? pat_n : 1)*sizeof(int));
	for (k = 0; k <= n; k++) c->cp[k] = pat_start[k];
	for (k = 0; k < n; k++) {
		int s0 = pat_start[k], s1 = pat_start[k+1], a, b;
		for (a = s0; a < s1; a++) c->ri[a] = c->iperm[pat[a]];
		for (a = s0+1; a < s1; a++) {       /* insertion sort */
			int x = c->ri[a];
			b = a-1;
			while (b >= s0 && c->ri[b] > x) { c->ri[b+1] = c->ri[b]; b--; }
			c->ri[b+1] = x;
		}
	}
	free(pat); free(pat_start);
	g_phase[ORC_PH_ORDER] += ph_now()-t_sym;
	/* numeric: left-looking over block columns */
	{
		const double t_num = ph_now();
		size_t nnz = (size_t)c->cp[n];
		double *W;
		int *map = malloc(n*sizeof(int));
		int *head = malloc(n*sizeof(int)), *nxt = malloc(n*sizeof(int)), *ptr = malloc(n*sizeof(int));
		/* original matrix access: column lists per node in both directions */
		int *acnt = calloc(n+1, sizeof(int)), *aidx, *ablk, *afl;
		int nod = cp[n];
		c->Ld = calloc(36*(size_t)n, sizeof(double));
		c->Lo = calloc(36*(nnz ? nnz : 1), sizeof(double));
		W = malloc(36*(size_t)(n+1)*sizeof(double));
		for (k = 0; k < nod; k++) { acnt[rows[k]+1]++; }
		for (i = 0; i < n; i++) acnt[i+1] += cp[i+1]-cp[i];
		for (i = 0; i < n; i++) acnt[i+1] += acnt[i];
		aidx = malloc((2*(size_t)nod+1)*sizeof(int));
		ablk = malloc((2*(size_t)nod+1)*sizeof(int));
		afl = malloc((2*(size_t)nod+1)*sizeof(int));
		{
			int *f = malloc(n*sizeof(int));
			for (i = 0; i < n; i++) f[i] = acnt[i];
			for (i = 0; i < n; i++) {
				for (k = cp[i]; k < cp[i+1]; k++) {
					int j = rows[k];
					/* block at (row j, col i): stored for node i (transposed use) and node j */
					aidx[f[i]] = j; ablk[f[i]] = k; afl[f[i]] = 1; f[i]++;
					aidx[f[j]] = i; ablk[f[j]] = k; afl[f[j]] = 0; f[j]++;
				}
			}
			free(f);
		}
		for (i = 0; i < n; i++) { map[i] = -1; head[i] = -1; }
		c->flops = 0;
		for (k = 0; k < n; k++) {
			int v = c->perm[k], s0 = c->cp[k], s1 = c->cp[k+1], m, jcol;
			double *Wd = W+36*(size_t)n;     /* diagonal work block */
			orc_dblock_dense(dblocks+15*(size_t)v, Wd);
			for (m = s0; m < s1; m++) {
				map[c->ri[m]] = m-s0;
				memset(W+36*(size_t)(m-s0), 0, 36*sizeof(double));
			}
			/* scatter column k of the permuted matrix: rows u with iperm[u] > k */
			for (m = acnt[v]; m < acnt[v+1]; m++) {
				int u = aidx[m], ku = c->iperm[u], slot, r, cc;
				double B[36];
				if (ku <= k) continue;
				slot = map[ku];
				orc_odblock_dense(odb+18*(size_t)ablk[m], B);
				if (afl[m]) {
					/* block (row u, col v) = B */
					for (r = 0; r < 36; r++) W[36*(size_t)slot+r] = B[r];
				} else {
					/* block (row u, col v) = B^T, since B sits at (row v, col u) */
					for (r = 0; r < 6; r++) for (cc = 0; cc < 6; cc++) W[36*(size_t)slot+6*r+cc] = B[6*cc+r];
				}
			}
			/* updates from columns j whose next pending row is k */
			jcol = head[k];
			while (jcol >= 0) {
				int nj = nxt[jcol];
				int p = ptr[jcol], e = c->cp[jcol+1], q;
				const double *Lkj = c->Lo+36*(size_t)p;
				for (q = p; q < e; q++) {
					int row = c->ri[q], r, cc, t;
					const double *Lij = c->Lo+36*(size_t)q;
					double *dst = (row == k) ? Wd : W+36*(size_t)map[row];
					/* dst -= Lij * Lkj^T */
					for (r = 0; r < 6; r++)
						for (cc = 0; cc < 6; cc++) {
							double acc = 0;
							for (t = 0; t < 6; t++) acc += Lij[6*r+t]*Lkj[6*cc+t];
							dst[6*r+cc] -= acc;
						}
					c->flops += 432;
				}
				if (p+1 < e) {
					int nr = c->ri[p+1];
					ptr[jcol] = p+1;
					nxt[jcol] = head[nr];
					head[nr] = jcol;
				}
				jcol = nj;
			}
			if (chol6(Wd)) { *status = -1; }
			memcpy(c->Ld+36*(size_t)k, Wd, 36*sizeof(double));
			/* L_ik = W_i * L_kk^{-T}: solve X * Lkk^T = W_i row by row */
			for (m = s0; m < s1; m++) {
				double *Wi = W+36*(size_t)(m-s0), *X = c->Lo+36*(size_t)m;
				int r, cc, t;
				for (r = 0; r < 6; r++) {
					for (cc = 0; cc < 6; cc++) {
						double v2 = Wi[6*r+cc];
						for (t = 0; t < cc; t++) v2 -= X[6*r+t]*Wd[6*cc+t];
						X[6*r+cc] = v2/Wd[6*cc+cc];
					}
				}
				c->flops += 216;
				map[c->ri[m]] = -1;
			}
			if (s1 > s0) {
				int nr = c->ri[s0];
				ptr[k] = s0;
				nxt[k] = head[nr];
				head[nr] = k;
			}
		}
		free(W); free(map); free(head); free(nxt); free(ptr);
		free(acnt); free(aidx); free(ablk); free(afl);
		g_phase[ORC_PH_FACTOR] += ph_now()-t_num;
	}
	free(nb);
	return c;
}

void orc_chol_solve(const orc_chol *c, const double *b, double *x)
{
	int n = c->n, k, m, r, t;
	double *y = malloc(6*(size_t)n*sizeof(double));
	for (k = 0; k < n; k++) memcpy(y+6*k, b+6*(size_t)c->perm[k], 6*sizeof(double));
	/* forward: L y = Pb (column oriented) */
	for (k = 0; k < n; k++) {
		const double *D = c->Ld+36*(size_t)k;
		double *yk = y+6*k;
		for (r = 0; r < 6; r++) {
			double v = yk[r];
			for (t = 0; t < r; t++) v -= D[6*r+t]*yk[t];
			yk[r] = v/D[6*r+r];
		}
		for (m = c->cp[k]; m < c->cp[k+1]; m++) {
			const double *Lb = c->Lo+36*(size_t)m;
			double *yi = y+6*(size_t)c->ri[m];
			for (r = 0; r < 6; r++) {
				double v = 0;
				for (t = 0; t < 6; t++) v += Lb[6*r+t]*yk[t];
				yi[r] -= v;
			}
		}
	}
	/* backward: L^T x = y */
	for (k = n-1; k >= 0; k--) {
		const double *D = c->Ld+36*(size_t)k;
		double *yk = y+6*k;
		for (m = c->cp[k]; m < c->cp[k+1]; m++) {
			const double *Lb = c->Lo+36*(size_t)m;
			const double *yi = y+6*(size_t)c->ri[m];
			for (t = 0; t < 6; t++) {
				double v = 0;
				for (r = 0; r < 6; r++) v += Lb[6*r+t]*yi[r];
				yk[t] -= v;
			}
		}
		for (r = 5; r >= 0; r--) {
			double v = yk[r];
			for (t = r+1; t < 6; t++) v -= D[6*t+r]*yk[t];
			yk[r] = v/D[6*r+r];
		}
	}
	for (k = 0; k < n; k++) memcpy(x+6*(size_t)c->perm[k], y+6*k, 6*sizeof(double));
	free(y);
}

long orc_chol_nnz_blocks(const orc_chol *c) { return c->cp[c->n]+c->n; }
double orc_chol_flops(const orc_chol *c) { return c->flops; }

void orc_chol_free(orc_chol *c)
{
	if (!c) return;
	free(c->perm); free(c->iperm); free(c->cp); free(c->ri); free(c->Ld); free(c->Lo);
	free(c);
}

/* ======================================================================= */
/* system lifecycle and step                                               */
/* ======================================================================= */
static void set_shear_rate(orc_system *s, double rate)
{
	/* System::set_shear_rate + setShearDirection (System.cpp:1972-1982, 2006-2015) */
	double ct = cos(s->p.theta_shear), st = sin(s->p.theta_shear);
	double Eh[6] = {0, 0, ct/2, st/2, 0, 0};
	double Oh[3] = {-0.5*st, 0.5*ct, 0};
	int k;
	s->p.shear_rate = rate;
	for (k = 0; k < 6; k++) s->E[k] = Eh[k]*rate;
	for (k = 0; k < 3; k++) s->omega[k] = Oh[k]*rate;
	/* setVelocityDifference (System.cpp:1967-1970) */
	{
		double z[3] = {0, 0, s->L[2]}, e[3];
		symdot(s->E, z, e);
		s->vel_difference[0] = 2*e[0]; s->vel_difference[1] = 2*e[1]; s->vel_difference[2] = 2*e[2];
	}
}

orc_system *orc_create(int np, const double *pos, const double *radius, const double L[3],
                       const double shear_disp[3], const orc_params *p)
{
	orc_system *s = calloc(1, sizeof(orc_system));
	s->p = *p;
	s->np = np;
	memcpy(s->L, L, 3*sizeof(double));
	memcpy(s->shear_disp, shear_disp, 3*sizeof(double));
	s->pos = malloc(3*(size_t)np*sizeof(double));
	s->radius = malloc((size_t)np*sizeof(double));
	memcpy(s->pos, pos, 3*(size_t)np*sizeof(double));
	memcpy(s->radius, radius, (size_t)np*sizeof(double));
	s->vel = calloc(3*(size_t)np, sizeof(double));
	s->ang = calloc(3*(size_t)np, sizeof(double));
	s->na_vel = calloc(3*(size_t)np, sizeof(double));
	s->na_ang = calloc(3*(size_t)np, sizeof(double));
	s->vel_pred = calloc(3*(size_t)np, sizeof(double));
	s->ang_pred = calloc(3*(size_t)np, sizeof(double));
	s->u_inf = calloc(3*(size_t)np, sizeof(double));
	s->na_disp = calloc(3*(size_t)np, sizeof(double));
	s->box_of = calloc(np, sizeof(int));
	s->snap_mem = calloc(np, sizeof(int));
	s->friction = p->friction_model != 0;
	s->repulsive = p->repulsive_length > 0;
	s->tangential_coeff = dashpot_tangential_coeff(p);
	s->dt = p->fixed_dt ? p->dt_max : 1e-4;   /* dt = p.dt (System.cpp:607) */
	s->init_strain_limit = 0;       /* System.cpp:90-91 */
	s->init_rate_limit = 999;
	s->comp_vel = calloc(ORC_NCOMP*3*(size_t)np, sizeof(double));
	s->comp_ang = calloc(ORC_NCOMP*3*(size_t)np, sizeof(double));
	s->pstress = calloc(ORC_NSC*6*(size_t)np, sizeof(double));
	set_shear_rate(s, p->shear_rate);
	return s;
}

void orc_destroy(orc_system *s)
{
	if (!s) return;
	free(s->pos); free(s->radius); free(s->vel); free(s->ang); free(s->na_vel); free(s->na_ang);
	free(s->vel_pred); free(s->ang_pred); free(s->u_inf); free(s->na_disp); free(s->box_of); free(s->snap_mem);
	free(s->ctype); free(s->static_nb); free(s->n_static); free(s->moving_nb); free(s->n_moving);
	free(s->snap_start); free(s->pairs); free(s->dblocks); free(s->odblocks); free(s->od_cp);
	free(s->od_rows);
	free(s->comp_vel); free(s->comp_ang); free(s->pstress);
	free(s);
}

int orc_setup(orc_system *s)
{
	/* System::initializeBoxing (System.cpp:613-652) + checkNewInteraction */
	double max_range = 0;
	int i, j;
	for (i = 0; i < s->np-1; i++)
		for (j = i+1; j < s->np; j++) {
			double r = orc_lub_range(s->radius[i], s->radius[j], s->p.lub_max_gap);
			if (r > max_range) max_range = r;
		}
	if (boxset_init(s, max_range)) return set_err(s, "BoxSet: init out of boundaries");
	for (i = 0; i < s->np; i++) if (box_particle(s, i)) return -1;
	if (orc_boxset_update(s)) return -1;
	orc_check_new_interaction(s);
	return 0;
}

int orc_first_update(orc_system *s)
{
	/* System::timeEvolution first-call block (System.cpp:1219-1228) */
	double dt_bak = s->dt;
	s->dt = 0;
	orc_check_new_interaction(s);
	s->in_predictor = 1;
	orc_update_interactions(s);
	s->in_predictor = 0;
	s->dt = dt_bak;
	return 0;
}

static int ensure_block_scratch(orc_system *s)
{
	int need = s->npairs > 0 ? s->npairs : 1;
	if (!s->dblocks) {
		s->dblocks = malloc(15*(size_t)s->np*sizeof(double));
		s->od_cp = malloc(((size_t)s->np+1)*sizeof(int));
	}
	if (need > s->od_cap) {
		s->od_cap = need;
		s->odblocks = realloc(s->odblocks, 18*(size_t)need*sizeof(double));
		s->od_rows = realloc(s->od_rows, (size_t)need*sizeof(int));
	}
	return 0;
}

int orc_solve_current(orc_system *s, const double *rhs, double *x)
{
	orc_chol *c;
	int st;
	ensure_block_scratch(s);
	PH(ORC_PH_ASSEMBLE, orc_build_blocks(s, s->dblocks, s->od_cp, s->od_rows, s->odblocks, s->od_cap));
	c = orc_chol_factor(s->np, s->dblocks, s->od_cp, s->od_rows, s->odblocks, &st);
	PH(ORC_PH_SOLVE, orc_chol_solve(c, rhs, x));
	orc_chol_free(c);
	return st;
}

int orc_compute_velocities(orc_system *s)
{
	/* System::computeVelocities, rate control, no components (System.cpp:2273-2319) */
	int i, st;
	double *rhs = malloc(12*(size_t)s->np*sizeof(double)), *x = rhs+6*(size_t)s->np;
	PH(ORC_PH_ASSEMBLE, orc_build_rhs(s, rhs));
	st = orc_solve_current(s, rhs, x);
	for (i = 0; i < s->np; i++) {
		memcpy(s->na_vel+3*i, x+6*i, 3*sizeof(double));
		memcpy(s->na_ang+3*i, x+6*i+3, 3*sizeof(double));
	}
	free(rhs);
	if (!s->p.fixed_dt && s->in_predictor) {
		/* computeMaxNAVelocity (System.cpp:1887-1902, called as at 2309) */
		double m = 0;
		for (i = 0; i < s->np; i++) {
			const double *v = s->na_vel+3*i;
			double q = v[0]*v[0]+v[1]*v[1]+v[2]*v[2];
			if (q > m) m = q;
		}
		s->max_velocity = sqrt(m);
	}
	/* adjustVelocityPeriodicBoundary (System.cpp:2369-2392) */
	for (i = 0; i < s->np; i++) {
		int k;
		for (k = 0; k < 3; k++) {
			s->vel[3*i+k] = s->na_vel[3*i+k]+s->u_inf[3*i+k];
			s->ang[3*i+k] = s->na_ang[3*i+k]+s->omega[k];
		}
	}
	return st;
}

static void adapt_time_step(orc_system *s, double strain_end)
{
	/* System::adaptTimeStep (System.cpp:1013-1066), time_end disabled; never
	 * called with fixed_dt (System.cpp:1094, 1130, 1243) */
	double rate = s->p.shear_rate;
	if (s->p.fixed_dt) return;
	if (s->max_velocity > 0 || s->max_sliding_velocity > 0) {
		if (s->max_velocity > s->max_sliding_velocity) s->dt = s->p.disp_max/s->max_velocity;
		else s->dt = s->p.disp_max/s->max_sliding_velocity;
	} else {
		s->dt = s->p.disp_max/rate;
	}
	if (s->dt*rate > s->p.disp_max) s->dt = s->p.disp_max/rate;
	if (s->p.dt_max > 0 && s->dt > s->p.dt_max) s->dt = s->p.dt_max;
	if (s->p.dt_min > 0 && s->dt < s->p.dt_min) s->dt = s->p.dt_min;
	if (strain_end >= 0) {
		if (fabs(s->dt*rate) > strain_end-s->strain) s->dt = fabs((strain_end-s->strain)/rate);
	}
}

static int time_step_boxing(orc_system *s)
{
	/* System::timeStepBoxing (System.cpp:672-717), simple shear */
	double z[3] = {0, 0, 1}, e[3], inc[3];
	int m;
	s->strain += s->p.shear_rate*s->dt;
	symdot(s->E, z, e);
	inc[0] = 2*e[0]*s->dt; inc[1] = 2*e[1]*s->dt; inc[2] = 2*e[2]*s->dt;
	s->shear_disp[0] += inc[0]*s->L[2];
	s->shear_disp[1] += inc[1]*s->L[2];
	s->shear_disp[2] += inc[2]*s->L[2];
	m = (int)(s->shear_disp[0]/s->L[0]);
	if (s->shear_disp[0] < 0) m--;
	s->shear_disp[0] = s->shear_disp[0]-m*s->L[0];
	m = (int)(s->shear_disp[1]/s->L[1]);
	if (s->shear_disp[1] < 0) m--;
	s->shear_disp[1] = s->shear_disp[1]-m*s->L[1];
	return orc_boxset_update(s);
}

static int displacement(orc_system *s, int i, const double dr[3])
{
	/* System::displacement (System.cpp:2418-2443) */
	double *x = s->pos+3*i;
	int zs;
	x[0] += dr[0]; x[1] += dr[1]; x[2] += dr[2];
	zs = orc_periodize(s->L, s->shear_disp, x);
	if (zs != 0) {
		double f = (double)zs;
		s->vel[3*i] += f*s->vel_difference[0];
		s->vel[3*i+1] += f*s->vel_difference[1];
		s->vel[3*i+2] += f*s->vel_difference[2];
	}
	return box_particle(s, i);
}

/* ---- velocity components, stress, stress control ---------------------- */
static void outer_sym3(const double a[3], const double b[3], double o[6])
{
	/* Sym2Tensor.h:256-264, (xx, xy, xz, yz, yy, zz) */
	o[0] = a[0]*b[0];
	o[1] = 0.5*(a[0]*b[1]+a[1]*b[0]);
	o[2] = 0.5*(a[0]*b[2]+a[2]*b[0]);
	o[3] = 0.5*(a[1]*b[2]+a[2]*b[1]);
	o[4] = a[1]*b[1];
	o[5] = a[2]*b[2];
}
static void outer3(const double a[3], double o[6])
{
	/* Sym2Tensor.h:266-274 */
	o[0] = a[0]*a[0]; o[1] = a[0]*a[1]; o[2] = a[0]*a[2];
	o[3] = a[1]*a[2]; o[4] = a[1]*a[1]; o[5] = a[2]*a[2];
}
static double doubledot6(const double a[6], const double b[6])
{
	/* Sym2Tensor.h:246-254 */
	return a[0]*b[0]+2*a[1]*b[1]+2*a[2]*b[2]+2*a[3]*b[3]+a[4]*b[4]+a[5]*b[5];
}
static double volume(const orc_system *s) { return s->L[0]*s->L[1]*s->L[2]; }
static const double ORC_PI = 3.14159265358979323846;   /* M_PI */

int orc_set_control(orc_system *s, int control, double target_stress)
{
	s->control = control;
	s->target_stress = target_stress;
	return 0;
}

int orc_set_shear_rate_limit(orc_system *s, double init_strain_limit, double init_rate_limit)
{
	s->init_strain_limit = init_strain_limit;
	s->init_rate_limit = init_rate_limit;
	return 0;
}

double orc_shear_rate(const orc_system *s) { return s->p.shear_rate; }

/* XM, YM of a lubricated pair at the current gap (calcXYFunctionsStress /
 * calcXFunctionsStress, Lubrication.cpp:305-330) */
static void lub_stress_functions(const orc_system *s, const orc_pair *q, double XM[4], double YM[4])
{
	int j;
	for (j = 0; j < 4; j++) {
		XM[j] = q->lc[4*cXM+j]*q->lub_coeff;
		YM[j] = s->p.lubrication_model == 2 ? q->lc[4*cYM+j]*q->log_lub_coeff : 0;
	}
}

/* Lubrication::addGUStresslet (Lubrication.cpp:587-615) */
static void add_gu(const orc_system *s, const orc_pair *q, const double YM[4], const double *vi, const double *vj,
                   const double *oi, const double *oj, double *Si, double *Sj)
{
	const double *n = q->nvec;
	double nvi = dot3(n, vi), nvj = dot3(n, vj), nn[6], ci, cj;
	int k;
	outer3(n, nn);
	ci = -(q->XG[0]*nvi+q->XG[1]*nvj);
	cj = -(q->XG[2]*nvi+q->XG[3]*nvj);
	for (k = 0; k < 6; k++) { Si[k] += ci*nn[k]; Sj[k] += cj*nn[k]; }
	if (s->p.lubrication_model != 2) return;
	{
		double ti[6], tj[6], t2i[6], t2j[6], wi[3], wj[3];
		outer_sym3(n, vi, ti);
		outer_sym3(n, vj, tj);
		for (k = 0; k < 6; k++) { ti[k] = ti[k]-nvi*nn[k]; tj[k] = tj[k]-nvj*nn[k]; }
		for (k = 0; k < 6; k++) {
			Si[k] += -2*(q->YG[0]*ti[k]+q->YG[1]*tj[k]);
			Sj[k] += -2*(q->YG[2]*ti[k]+q->YG[3]*tj[k]);
		}
		cross3(oi, n, wi);
		cross3(oj, n, wj);
		outer_sym3(n, wi, t2i);
		outer_sym3(n, wj, t2j);
		for (k = 0; k < 6; k++) {
			Si[k] += -2*(YM[0]*t2i[k]+YM[1]*t2j[k]);
			Sj[k] += -2*(YM[2]*t2i[k]+YM[3]*t2j[k]);
		}
	}
}

/* Lubrication::addMEStresslet (Lubrication.cpp:617-650) */
static void add_me(const orc_system *s, const orc_pair *q, const double XM[4], const double YM[4],
                   double *Si, double *Sj)
{
	const double *n = q->nvec;
	double En[3], nnE, nn[6], coeff, ci, cj;
	int k;
	symdot(s->E, n, En);
	nnE = dot3(n, En);
	outer3(n, nn);
	coeff = 1.5*nnE;
	ci = coeff*(XM[0]+XM[1]);
	cj = coeff*(XM[2]+XM[3]);
	for (k = 0; k < 6; k++) { Si[k] += ci*nn[k]; Sj[k] += cj*nn[k]; }
	if (s->p.lubrication_model == 2) {
		double nEn[6], t[6], yi = YM[0]+YM[1], yj = YM[2]+YM[3];
		outer_sym3(n, En, nEn);
		for (k = 0; k < 6; k++) t[k] = 2*(nEn[k]-nnE*nn[k]);
		for (k = 0; k < 6; k++) { Si[k] += yi*t[k]; Sj[k] += yj*t[k]; }
	}
}

static void add_scaled(double *S, double a, const double t[6])
{
	int k;
	for (k = 0; k < 6; k++) S[k] += a*t[k];
}

/* System::calcStressPerParticle (Stress.cpp:194-287) with
 * calcContactXFPerParticleStressControlled (Stress.cpp:145-192) */
static void calc_stress_per_particle(orc_system *s)
{
	size_t np = (size_t)s->np;
	double *ps = s->pstress;
	int k, c, i;
	memset(ps, 0, ORC_NSC*6*np*sizeof(double));
	/* GU per velocity component, M:E (lubricated pairs) */
	for (k = 0; k < s->npairs; k++) {
		const orc_pair *q = &s->pairs[k];
		double XM[4], YM[4];
		if (!q->lub_active) continue;
		lub_stress_functions(s, q, XM, YM);
		add_me(s, q, XM, YM, ps+6*(size_t)q->p0, ps+6*(size_t)q->p1);
		for (c = 0; c < ORC_NCOMP; c++) {
			const double *v = s->comp_vel+3*np*c, *w = s->comp_ang+3*np*c;
			double *S = ps+6*np*(SC_GU+c);
			if (!s->comp_present[c]) continue;
			add_gu(s, q, YM, v+3*q->p0, v+3*q->p1, w+3*q->p0, w+3*q->p1, S+6*(size_t)q->p0, S+6*(size_t)q->p1);
		}
	}
	if (s->control == 0) {
		/* Contact::addUpStress (Contact.cpp:439-458): xF of spring + dashpot at the total velocities */
		double *S = ps+6*np*SC_XF_CONTACT;
		for (k = 0; k < s->npairs; k++) {
			const orc_pair *q = &s->pairs[k];
			double f[3], F[3], st[6], rij;
			if (!q->contact_active) continue;
			dashpot_force(s, q, s->vel+3*q->p0, s->vel+3*q->p1, s->ang+3*q->p0, s->ang+3*q->p1, 0, f);
			F[0] = q->f_total[0]+f[0]; F[1] = q->f_total[1]+f[1]; F[2] = q->f_total[2]+f[2];
			outer_sym3(q->rvec, F, st);
			rij = q->a0+q->a1;
			add_scaled(S+6*(size_t)q->p0, q->a0/rij, st);
			add_scaled(S+6*(size_t)q->p1, q->a1/rij, st);
		}
	} else {
		/* gatherVelocitiesByRateDependencies (Stress.cpp:116-143) */
		double *pv = calloc(12*np, sizeof(double));
		double *pw = pv+3*np, *iv = pv+6*np, *iw = pv+9*np;
		double *Sind = ps+6*np*SC_XF_CONTACT, *Spro = ps+6*np*SC_XF_CONTACT_PROP;
		static const int prop[ORC_NCOMP] = {0, 1, 1, 0};
		for (c = 0; c < ORC_NCOMP; c++) {
			const double *v = s->comp_vel+3*np*c, *w = s->comp_ang+3*np*c;
			if (!s->comp_present[c]) continue;
			for (i = 0; i < (int)(3*np); i++) {
				if (prop[c]) { pv[i] += v[i]; pw[i] += w[i]; }
				else { iv[i] += v[i]; iw[i] += w[i]; }
			}
		}
		for (i = 0; i < (int)np; i++) {
			add3(pv+3*i, s->u_inf+3*i);
			add3(pw+3*i, s->omega);
		}
		for (k = 0; k < s->npairs; k++) {
			const orc_pair *q = &s->pairs[k];
			double st[6], f[3], rij = s->radius[q->p0]+s->radius[q->p1];
			if (!q->contact_active) continue;
			/* Contact::addUpStressSpring (Contact.cpp:460-467) */
			outer_sym3(q->rvec, q->f_total, st);
			add_scaled(Sind+6*(size_t)q->p0, q->a0/(q->a0+q->a1), st);
			add_scaled(Sind+6*(size_t)q->p1, q->a1/(q->a0+q->a1), st);
			dashpot_force(s, q, pv+3*q->p0, pv+3*q->p1, pw+3*q->p0, pw+3*q->p1, 0, f);
			outer_sym3(q->rvec, f, st);
			add_scaled(Spro+6*(size_t)q->p0, s->radius[q->p0]/rij, st);
			add_scaled(Spro+6*(size_t)q->p1, s->radius[q->p1]/rij, st);
			dashpot_force(s, q, iv+3*q->p0, iv+3*q->p1, iw+3*q->p0, iw+3*q->p1, 1, f);
			outer_sym3(q->rvec, f, st);
			add_scaled(Sind+6*(size_t)q->p0, s->radius[q->p0]/rij, st);
			add_scaled(Sind+6*(size_t)q->p1, s->radius[q->p1]/rij, st);
		}
		free(pv);
	}
	if (s->repulsive) {
		/* RepulsiveForce::addUpStressXF (RepulsiveForce.cpp:111-131): half to each particle */
		double *S = ps+6*np*SC_XF_REP;
		for (k = 0; k < s->npairs; k++) {
			const orc_pair *q = &s->pairs[k];
			double st[6], sc[6];
			int m;
			outer_sym3(q->rvec, q->rep_force, st);
			for (m = 0; m < 6; m++) sc[m] = 0.5*st[m];
			add3(S+6*(size_t)q->p0, sc); add3(S+6*(size_t)q->p0+3, sc+3);
			add3(S+6*(size_t)q->p1, sc); add3(S+6*(size_t)q->p1+3, sc+3);
		}
	}
}

/* stress component slot -> group, rate dependence (1 proportional, 0 independent) */
static const int sc_group[ORC_NSC] = {ORC_G_HYDRO, ORC_G_CONTACT, ORC_G_DASHPOT, ORC_G_HYDRO, ORC_G_REPULSION,
                                      ORC_G_CONTACT, ORC_G_CONTACT, ORC_G_REPULSION};
static const int sc_prop[ORC_NSC] = {1, 0, 1, 1, 0, 0, 1, 0};

static int sc_present(const orc_system *s, int k)
{
	if (k >= SC_GU && k < SC_GU+ORC_NCOMP) return s->comp_present[k-SC_GU];
	if (k == SC_XF_CONTACT_PROP) return s->control == 1;
	if (k == SC_XF_REP) return s->repulsive;
	return 1;
}

/* StressComponent::getTotalStress (StressComponent.h:26-33): sum over particles */
static void sc_total(const orc_system *s, int k, double t[6])
{
	const double *S = s->pstress+6*(size_t)s->np*k;
	int i, m;
	for (m = 0; m < 6; m++) t[m] = 0;
	for (i = 0; i < s->np; i++) for (m = 0; m < 6; m++) t[m] += S[6*i+m];
}

/* System::computeShearRate + gatherStressesByRateDependencies
 * (System.cpp:2017-2055; Stress.cpp:320-340) */
static void compute_shear_rate(orc_system *s)
{
	double prop[6] = {0, 0, 0, 0, 0, 0}, indep[6] = {0, 0, 0, 0, 0, 0}, t[6], vinv, rp, ri;
	int k, m;
	calc_stress_per_particle(s);
	for (k = 0; k < ORC_NSC; k++) {
		if (!sc_present(s, k)) continue;
		sc_total(s, k, t);
		for (m = 0; m < 6; m++) {
			if (sc_prop[k]) prop[m] += t[m];
			else indep[m] += t[m];
		}
	}
	vinv = 1.0/volume(s);
	for (m = 0; m < 6; m++) { prop[m] *= vinv; indep[m] *= vinv; }
	for (m = 0; m < 6; m++) prop[m] += (2*s->E[m])/(6*ORC_PI);
	rp = doubledot6(prop, s->E);
	ri = doubledot6(indep, s->E);
	if (rp != 0) set_shear_rate(s, (s->target_stress-ri)/rp);
	if (s->strain < s->init_strain_limit && s->p.shear_rate > s->init_rate_limit)
		set_shear_rate(s, s->init_rate_limit);
}

int orc_compute_velocities_components(orc_system *s)
{
	/* System::computeVelocities(divided_velocities) (System.cpp:2273-2319)
	 * with computeVelocityByComponents (System.cpp:1947-1965) */
	size_t np = (size_t)s->np;
	int i, c, st = 0;
	double *rhs = malloc(12*np*sizeof(double)), *x = rhs+6*np;
	orc_chol *ch;
	if (s->control == 1) set_shear_rate(s, 1);
	compute_uinf(s);
	ensure_block_scratch(s);
	PH(ORC_PH_ASSEMBLE, orc_build_blocks(s, s->dblocks, s->od_cp, s->od_rows, s->odblocks, s->od_cap));
	ch = orc_chol_factor(s->np, s->dblocks, s->od_cp, s->od_rows, s->odblocks, &st);
	for (c = 0; c < ORC_NCOMP; c++) {
		double *v = s->comp_vel+3*np*c, *w = s->comp_ang+3*np*c;
		PH(ORC_PH_ASSEMBLE, s->comp_present[c] = orc_build_rhs_component(s, c, rhs) == 0);
		if (!s->comp_present[c]) continue;
		PH(ORC_PH_SOLVE, orc_chol_solve(ch, rhs, x));
		for (i = 0; i < (int)np; i++) {
			memcpy(v+3*i, x+6*i, 3*sizeof(double));
			memcpy(w+3*i, x+6*i+3, 3*sizeof(double));
		}
	}
	orc_chol_free(ch);
	free(rhs);
	if (s->control == 1) {
		/* computeShearRate, then rescaleVelHydroStressControlled (System.cpp:2264-2271) */
		compute_shear_rate(s);
		for (c = 0; c < ORC_NCOMP; c++) {
			double *v = s->comp_vel+3*np*c, *w = s->comp_ang+3*np*c;
			if (!s->comp_present[c] || !(c == ORC_DASHPOT || c == ORC_HYDRO)) continue;
			for (i = 0; i < (int)(3*np); i++) { v[i] *= s->p.shear_rate; w[i] *= s->p.shear_rate; }
		}
	}
	/* sumUpVelocityComponents (System.cpp:2236-2250) */
	memset(s->na_vel, 0, 3*np*sizeof(double));
	memset(s->na_ang, 0, 3*np*sizeof(double));
	for (c = 0; c < ORC_NCOMP; c++) {
		const double *v = s->comp_vel+3*np*c, *w = s->comp_ang+3*np*c;
		if (!s->comp_present[c]) continue;
		for (i = 0; i < (int)np; i++) { add3(s->na_vel+3*i, v+3*i); add3(s->na_ang+3*i, w+3*i); }
	}
	if (!s->p.fixed_dt && s->in_predictor) {   /* System.cpp:2309 */
		double m = 0;
		for (i = 0; i < (int)np; i++) {
			const double *v = s->na_vel+3*i;
			double q = v[0]*v[0]+v[1]*v[1]+v[2]*v[2];
			if (q > m) m = q;
		}
		s->max_velocity = sqrt(m);
	}
	/* adjustVelocityPeriodicBoundary (System.cpp:2369-2392) */
	if (s->control == 1) compute_uinf(s);
	for (i = 0; i < (int)np; i++) {
		int k;
		for (k = 0; k < 3; k++) {
			s->vel[3*i+k] = s->na_vel[3*i+k]+s->u_inf[3*i+k];
			s->ang[3*i+k] = s->na_ang[3*i+k]+s->omega[k];
		}
	}
	return st;
}

int orc_calc_stress(orc_system *s)
{
	/* calcStressPerParticle + System::calcStress (Stress.cpp:342-384) */
	double t[6], vinv = 1.0/volume(s);
	int k, m, g;
	calc_stress_per_particle(s);
	memset(s->groups, 0, sizeof(s->groups));
	for (k = 0; k < ORC_NSC; k++) {
		if (!sc_present(s, k)) continue;
		sc_total(s, k, t);
		for (m = 0; m < 6; m++) s->groups[sc_group[k]][m] += t[m];
	}
	for (g = 0; g < ORC_NGROUP; g++) for (m = 0; m < 6; m++) s->groups[g][m] *= vinv;
	for (m = 0; m < 6; m++) s->groups[ORC_G_HYDRO][m] += (2*s->E[m])/(6*ORC_PI);
	for (m = 0; m < 6; m++) s->total_stress[m] = 0;
	for (g = 0; g < ORC_NGROUP; g++) for (m = 0; m < 6; m++) s->total_stress[m] += s->groups[g][m];
	return 0;
}

int orc_get_stress(const orc_system *s, double *groups, double *total, double *particle)
{
	int i, k, m;
	if (groups) memcpy(groups, s->groups, sizeof(s->groups));
	if (total) memcpy(total, s->total_stress, sizeof(s->total_stress));
	if (particle) {
		/* calcTotalStressPerParticle (Stress.cpp:289-300) */
		memset(particle, 0, 6*(size_t)s->np*sizeof(double));
		for (k = 0; k < ORC_NSC; k++) {
			const double *S = s->pstress+6*(size_t)s->np*k;
			if (!sc_present(s, k)) continue;
			for (i = 0; i < s->np; i++) for (m = 0; m < 6; m++) particle[6*i+m] += S[6*i+m];
		}
	}
	return 0;
}

int orc_gu_stress(orc_system *s, const double *vel, const double *ang, double *S)
{
	/* addUpInteractionStressGU (Stress.cpp:77-97) for given velocities */
	int k;
	memset(S, 0, 6*(size_t)s->np*sizeof(double));
	for (k = 0; k < s->npairs; k++) {
		const orc_pair *q = &s->pairs[k];
		double XM[4], YM[4];
		if (!q->lub_active) continue;
		lub_stress_functions(s, q, XM, YM);
		add_gu(s, q, YM, vel+3*q->p0, vel+3*q->p1, ang+3*q->p0, ang+3*q->p1, S+6*(size_t)q->p0, S+6*(size_t)q->p1);
	}
	return 0;
}

int orc_get_components(const orc_system *s, int comp, double *vel, double *ang)
{
	size_t np = (size_t)s->np;
	if (comp < 0 || comp >= ORC_NCOMP || !s->comp_present[comp]) return -1;
	if (vel) memcpy(vel, s->comp_vel+3*np*comp, 3*np*sizeof(double));
	if (ang) memcpy(ang, s->comp_ang+3*np*comp, 3*np*sizeof(double));
	return 0;
}

double orc_viscosity(const orc_system *s)
{
	/* Simulation::outputData (Simulation.cpp:662-666) */
	double Eh[6], ct = cos(s->p.theta_shear), st = sin(s->p.theta_shear);
	const double *E = s->E;
	if (s->p.shear_rate == 0) {
		Eh[0] = 0; Eh[1] = 0; Eh[2] = ct/2; Eh[3] = st/2; Eh[4] = 0; Eh[5] = 0;
		E = Eh;
	}
	return 0.5*doubledot6(s->total_stress, E)/doubledot6(E, E);
}

static int velocities(orc_system *s, int calc_stress)
{
	/* System::computeVelocities (System.cpp:2273-2319): by components when
	 * the stress is wanted or under stress control */
	if (calc_stress || s->control == 1) return orc_compute_velocities_components(s);
	return orc_compute_velocities(s);
}

int orc_step_ex(orc_system *s, double strain_end, int calc_stress, int adapt_first)
{
	/* System::timeEvolution loop body + timeEvolutionPredictorCorrectorMethod
	 * (System.cpp:1241-1256, 888-983, 1124-1187) */
	int i, st = 0;
	double t0;
	if (adapt_first) adapt_time_step(s, strain_end);
	/* predictor */
	s->in_predictor = 1;
	st |= velocities(s, calc_stress);
	if (calc_stress) orc_calc_stress(s);
	adapt_time_step(s, strain_end);
	s->time += s->dt;
	t0 = ph_now();
	if (time_step_boxing(s)) return -1;
	for (i = 0; i < s->np; i++) {
		double dr[3] = {s->vel[3*i]*s->dt, s->vel[3*i+1]*s->dt, s->vel[3*i+2]*s->dt};
		if (displacement(s, i, dr)) return -1;
	}
	g_phase[ORC_PH_MOVE] += ph_now()-t0;
	PH(ORC_PH_PAIRS, orc_update_interactions(s));
	memcpy(s->vel_pred, s->vel, 3*(size_t)s->np*sizeof(double));
	memcpy(s->ang_pred, s->ang, 3*(size_t)s->np*sizeof(double));
	/* corrector */
	s->in_predictor = 0;
	st |= velocities(s, calc_stress);
	if (calc_stress) orc_calc_stress(s);
	for (i = 0; i < 3*s->np; i++) {
		s->vel[i] = 0.5*(s->vel[i]+s->vel_pred[i]);
		s->ang[i] = 0.5*(s->ang[i]+s->ang_pred[i]);
	}
	t0 = ph_now();
	for (i = 0; i < s->np; i++) {
		double dr[3] = {(s->vel[3*i]-s->vel_pred[3*i])*s->dt,
		                (s->vel[3*i+1]-s->vel_pred[3*i+1])*s->dt,
		                (s->vel[3*i+2]-s->vel_pred[3*i+2])*s->dt};
		if (displacement(s, i, dr)) return -1;
	}
	g_phase[ORC_PH_MOVE] += ph_now()-t0;
	PH(ORC_PH_NEIGHBOR, orc_check_new_interaction(s));
	PH(ORC_PH_PAIRS, orc_update_interactions(s));
	/* the corrector's non-affine velocity times dt (System.cpp:980-982) */
	for (i = 0; i < 3*s->np; i++) s->na_disp[i] += s->na_vel[i]*s->dt;
	return st;
}

int orc_step(orc_system *s, double strain_end)
{
	return orc_step_ex(s, strain_end, 0, 1);
}

int orc_time_evolution(orc_system *s, double strain_end)
{
	/* System::timeEvolution (System.cpp:1203-1279), time_end disabled:
	 * steps while keepRunning (System.cpp:1189-1201), then the last step
	 * with the stress, without a new adaptTimeStep.  With fixed_dt the loop
	 * stops one step earlier (loop_time_adjust, System.cpp:1237-1242). */
	int n = 0;
	const double se = strain_end-(s->p.fixed_dt ? s->dt : 0);
	while (!(s->strain > se-1e-8 && se >= 0)) {
		if (orc_step_ex(s, strain_end, 0, 1) < 0) return -1;
		n++;
	}
	if (orc_step_ex(s, strain_end, 1, 0) < 0) return -1;
	return n+1;
}

int orc_set_in_predictor(orc_system *s, int flag) { s->in_predictor = flag; return 0; }
int orc_set_dt(orc_system *s, double dt) { s->dt = dt; return 0; }

int orc_num_pairs(const orc_system *s) { return s->npairs; }

void orc_get_na_disp(const orc_system *s, double *out)
{
	memcpy(out, s->na_disp, 3*(size_t)s->np*sizeof(double));
}

/* Per interaction, in pair order, the force columns of
 * Simulation::outputIntFileTxt (Simulation.cpp:941-1006), ORC_PAIR_OUT values:
 *  [0]   normal part of the lubrication force on p0, positive for compression:
 *        -F.n with F = Lubrication::getTotalForce (Lubrication.cpp:659-694) at
 *        the non-affine velocities, zero unless the gap is positive;
 *  [1-3] its tangential part F - (F.n) n (Lubrication.h:120-124);
 *  [4]   normal part of the contact force F_c.n (signed; the file prints its
 *        norm) and [5-7] its tangential part, F_c = spring + dashpot at the
 *        total velocities (Contact.cpp:235-253, 487-501), zero without contact;
 *  [8]   norm of the repulsive force (RepulsiveForce.h:63-66);
 *  [9]   viscosity contribution of the contact xF,
 *        (rvec (x) F_c) : E_inf / rate^2 (Contact.cpp:439-450 at this state). */
int orc_pair_output(const orc_system *s, double *out)
{
	int k, m;
	const double rate = s->p.shear_rate;
	for (k = 0; k < s->npairs; k++) {
		const orc_pair *q = &s->pairs[k];
		const double *n = q->nvec;
		double *o = out+ORC_PAIR_OUT*(size_t)k;
		for (m = 0; m < ORC_PAIR_OUT; m++) o[m] = 0;
		if (q->lub_active && q->gap > 0) {
			const double *vi = s->na_vel+3*q->p0, *vj = s->na_vel+3*q->p1;
			double f[3], t[3], d;
			for (m = 0; m < 3; m++) t[m] = q->XA[0]*vi[m]+q->XA[1]*vj[m];
			d = -dot3(t, n);
			for (m = 0; m < 3; m++) f[m] = d*n[m];
			if (s->p.lubrication_model == 2) {
				const double *oi = s->na_ang+3*q->p0, *oj = s->na_ang+3*q->p1;
				double di = dot3(n, vi), dj = dot3(n, vj), ci[3], cj[3];
				cross3(n, oi, ci);
				cross3(n, oj, cj);
				for (m = 0; m < 3; m++) {
					f[m] += -q->YA[0]*(vi[m]-n[m]*di)-q->YA[1]*(vj[m]-n[m]*dj);
					f[m] += -q->YB[0]*ci[m]-q->YB[2]*cj[m];
				}
			}
			if (rate != 0) {
				double En[3], nnE;
				symdot(s->E, n, En);
				nnE = dot3(n, En);
				if (s->p.lubrication_model == 2) {
					double yg = q->YG[0]+q->YG[2];
					double c = (q->XG[0]+q->XG[2]-2*yg)*nnE;
					for (m = 0; m < 3; m++) f[m] += c*n[m]+2*yg*En[m];
				} else {
					double c = (q->XG[0]+q->XG[2])*nnE;
					for (m = 0; m < 3; m++) f[m] += c*n[m];
				}
			}
			d = dot3(f, n);
			o[0] = -d;
			for (m = 0; m < 3; m++) o[1+m] = f[m]-d*n[m];
		}
		if (q->contact_active) {
			double fd[3], fc[3], d, st[6];
			dashpot_force(s, q, s->vel+3*q->p0, s->vel+3*q->p1, s->ang+3*q->p0, s->ang+3*q->p1, 0, fd);
			for (m = 0; m < 3; m++) fc[m] = q->f_total[m]+fd[m];
			d = dot3(fc, n);
			o[4] = d;
			for (m = 0; m < 3; m++) o[5+m] = fc[m]-d*n[m];
			/* Sym2Tensor outer_sym (xx, xy, xz, yz, yy, zz) and its double contraction */
			st[0] = q->rvec[0]*fc[0];
			st[1] = 0.5*(q->rvec[0]*fc[1]+q->rvec[1]*fc[0]);
			st[2] = 0.5*(q->rvec[0]*fc[2]+q->rvec[2]*fc[0]);
			st[3] = 0.5*(q->rvec[1]*fc[2]+q->rvec[2]*fc[1]);
			st[4] = q->rvec[1]*fc[1];
			st[5] = q->rvec[2]*fc[2];
			if (rate != 0)
				o[9] = (st[0]*s->E[0]+2*st[1]*s->E[1]+2*st[2]*s->E[2]+2*st[3]*s->E[3]+st[4]*s->E[4]
				        +st[5]*s->E[5])/(rate*rate);
		}
		o[8] = q->rep_norm;
	}
	return s->npairs;
}

int orc_get_pairs(const orc_system *s, int *p0, int *p1, int *cs, double *geom, double *dtan)
{
	int k;
	for (k = 0; k < s->npairs; k++) {
		const orc_pair *q = &s->pairs[k];
		if (p0) p0[k] = q->p0;
		if (p1) p1[k] = q->p1;
		if (cs) cs[k] = q->state;
		if (geom) {
			geom[5*k] = q->nvec[0]; geom[5*k+1] = q->nvec[1]; geom[5*k+2] = q->nvec[2];
			geom[5*k+3] = q->r; geom[5*k+4] = q->gap;
		}
		if (dtan) memcpy(dtan+3*k, q->disp_tan, 3*sizeof(double));
	}
	return s->npairs;
}

int orc_set_pair_history(orc_system *s, int n, const int *p0, const int *p1, const int *state,
                         const double *disp_tan, const double *prev_disp_tan)
{
	int m, miss = 0;
	for (m = 0; m < n; m++) {
		int k = pair_find(s, p0[m], p1[m]);
		if (k < 0) { miss++; continue; }
		if (state) s->pairs[k].state = state[m];
		if (disp_tan) memcpy(s->pairs[k].disp_tan, disp_tan+3*m, 3*sizeof(double));
		if (prev_disp_tan) memcpy(s->pairs[k].prev_disp_tan, prev_disp_tan+3*m, 3*sizeof(double));
	}
	return miss;
}

int orc_get_state(const orc_system *s, double *pos, double *vel, double *ang, double *na_vel,
                  double *na_ang, double *sc)
{
	size_t b = 3*(size_t)s->np*sizeof(double);
	if (pos) memcpy(pos, s->pos, b);
	if (vel) memcpy(vel, s->vel, b);
	if (ang) memcpy(ang, s->ang, b);
	if (na_vel) memcpy(na_vel, s->na_vel, b);
	if (na_ang) memcpy(na_ang, s->na_ang, b);
	if (sc) {
		sc[0] = s->shear_disp[0]; sc[1] = s->shear_disp[1]; sc[2] = s->shear_disp[2];
		sc[3] = s->dt; sc[4] = s->time; sc[5] = s->strain;
		sc[6] = s->max_velocity; sc[7] = s->max_sliding_velocity;
	}
	return 0;
}

int orc_set_positions(orc_system *s, const double *pos, const double sd[3])
{
	int i;
	memcpy(s->pos, pos, 3*(size_t)s->np*sizeof(double));
	if (sd) memcpy(s->shear_disp, sd, 3*sizeof(double));
	if (s->nbox) for (i = 0; i < s->np; i++) if (box_particle(s, i)) return -1;
	return 0;
}

int orc_set_velocities(orc_system *s, const double *vel, const double *ang)
{
	memcpy(s->vel, vel, 3*(size_t)s->np*sizeof(double));
	memcpy(s->ang, ang, 3*(size_t)s->np*sizeof(double));
	return 0;
}
