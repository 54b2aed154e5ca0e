// lfd_physics.hpp -- per-pair physics of LF_DEM's Stokes-resistance step, as
// host+device inline functions shared by every kernel.
//
// Arithmetic follows the reference's operation order (the library is built
// with -ffp-contract=off), so FP64 results are bit-identical to a CPU build of
// the same expressions.  LF_DEM citations are file:line in its source tree.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define LFD_HD __host__ __device__ inline

namespace lfd {

struct Params {
    double shear_rate, theta_shear, lub_max_gap, lub_reduce, kn, kt;
    double mu_static, mu_dynamic, tau_n, tau_t, sd_coeff, repulsion;
    double repulsive_length, disp_max, dt_max, dt_min;
    int friction_model, lubrication_model;
    int fixed_dt;                            // constant step dt_max (System.cpp:607-610, 1238-1243)
};

// Derived constants shared by all kernels of one instance.
struct Consts {
    double L[3], half[3];
    double Eh[6], Oh[3];                     // strain rate and vorticity at unit shear rate (System.cpp:2006-2015)
    double normal_coeff, tangential_coeff;   // dashpot (ContactDashpot.cpp:65-105)
    double lub_max_gap, lub_reduce, kn, kt, mu_s, mu_d, sd_coeff;
    double repulsion, rep_len;
    int friction, repulsive, tangential;    // tangential lubrication model
};

// pair flags
enum : int { F_CONTACT = 1, F_LUB = 2, F_STATE_SHIFT = 2 };  // state in bits 2..3
LFD_HD int pair_state(int f) { return (f >> F_STATE_SHIFT) & 3; }
LFD_HD int with_state(int f, int s) { return (f & ~(3 << F_STATE_SHIFT)) | (s << F_STATE_SHIFT); }

// One interacting pair (p0 < p1); mirrors the state an LF_DEM Interaction
// carries between steps (Interaction.h:30-115, Contact.h:36-175).
struct alignas(16) PairRec {
    double n[3], r, gap;
    double disp_tan[3], prev_disp_tan[3], rsv_sq;
    double f_total[3], f_tan[3];
    double rep_norm;
    int p0, p1, zoff, flags;
    int pad[2];
};

LFD_HD double dot3(const double* a, const double* b) { return a[0]*b[0]+a[1]*b[1]+a[2]*b[2]; }
LFD_HD void cross3(const double* a, const double* b, double* c)
{
    double x = a[1]*b[2]-a[2]*b[1];
    double y = a[2]*b[0]-a[0]*b[2];
    double z = a[0]*b[1]-a[1]*b[0];
    c[0] = x; c[1] = y; c[2] = z;
}
// dot(Sym2Tensor, vec3d) with elm = (xx, xy, xz, yz, yy, zz) (Sym2Tensor.h:231-238)
LFD_HD void symdot(const double* E, const double* v, double* o)
{
    double x = E[0]*v[0]+E[1]*v[1]+E[2]*v[2];
    double y = E[1]*v[0]+E[4]*v[1]+E[3]*v[2];
    double z = E[2]*v[0]+E[3]*v[1]+E[5]*v[2];
    o[0] = x; o[1] = y; o[2] = z;
}

// Imposed flow at the current shear rate: System::set_shear_rate with
// setVelocityDifference (System.cpp:1967-1982).  Kept in device memory so that
// stress control can change the rate between kernels.
struct Flow { double rate, E[6], omega[3], vel_diff[3]; };
LFD_HD void make_flow(const Consts& C, double rate, Flow& F)
{
    F.rate = rate;
    for (int k = 0; k < 6; k++) F.E[k] = C.Eh[k]*rate;
    for (int k = 0; k < 3; k++) F.omega[k] = C.Oh[k]*rate;
    double z[3] = {0, 0, C.L[2]}, e[3];
    symdot(F.E, z, e);
    for (int k = 0; k < 3; k++) F.vel_diff[k] = 2*e[k];
}

// System::periodize (System.cpp:2478-2507)
LFD_HD int periodize(const double* L, const double* sd, double* p)
{
    int z = 0;
    if (p[2] >= L[2]) { p[2] -= L[2]; p[0] -= sd[0]; p[1] -= sd[1]; p[2] -= sd[2]; z = -1; }
    else if (p[2] < 0) { p[2] += L[2]; p[0] += sd[0]; p[1] += sd[1]; p[2] += sd[2]; z = 1; }
    while (p[0] >= L[0]) p[0] -= L[0];
    while (p[0] < 0) p[0] += L[0];
    if (p[1] >= L[1]) p[1] -= L[1];
    else if (p[1] < 0) p[1] += L[1];
    return z;
}

// System::periodizeDiff (System.cpp:2517-2548)
LFD_HD int periodize_diff(const double* L, const double* half, const double* sd, double* d)
{
    int z = 0;
    if (d[2] > half[2]) { d[2] -= L[2]; d[0] -= sd[0]; d[1] -= sd[1]; d[2] -= sd[2]; z = -1; }
    else if (d[2] < -half[2]) { d[2] += L[2]; d[0] += sd[0]; d[1] += sd[1]; d[2] += sd[2]; z = 1; }
    while (d[0] > half[0]) d[0] -= L[0];
    while (d[0] < -half[0]) d[0] += L[0];
    if (d[1] > half[1]) d[1] -= L[1];
    else if (d[1] < -half[1]) d[1] += L[1];
    return z;
}

// System::calcLubricationRange (System.cpp:2665-2674)
LFD_HD double lub_range(double ri, double rj, double g)
{
    double ratio = ri/rj;
    if (ratio < 2 && ratio > 0.5) return (2+g)*0.5*(ri+rj);
    double mr = ri < rj ? ri : rj;
    return ri+rj+g*mr;
}

// BoxSet::whichBox (BoxSet.cpp:764-781): -1 when out of boundaries
LFD_HD int which_box(const double* p, double bx, double by, double bz, unsigned yz, unsigned nz, unsigned nbox)
{
    unsigned ix = (unsigned)(p[0]/bx);
    unsigned iy = (unsigned)(p[1]/by);
    unsigned iz = (unsigned)(p[2]/bz);
    unsigned l = ix*yz+iy*nz+iz;
    return l >= nbox ? -1 : (int)l;
}

// ---- resistance functions (LubricationFunctions.h:12-90) ----
LFD_HD double g1_XA(double l) { double p = l+1; return 2*l*l/(p*p*p); }
LFD_HD double g2_YA(double l) { double p = l+1; return (4./15)*l*(2+l+2*l*l)/(p*p*p); }
LFD_HD double g2_YB(double l) { double p = l+1; return -(1./5)*l*(4+l)/(p*p); }
LFD_HD double g2_YC(double l) { return (2./5)*l/(l+1); }
LFD_HD double g4_YC(double l) { double p = l+1; return (4./5)*l*l/(p*p*p*p); }
LFD_HD double g1_XG(double l) { double p = l+1; return 3*l*l/(p*p*p); }
LFD_HD double g2_YG(double l) { double p = l+1; return (1./10)*l*(4-l+7*l*l)/(p*p*p); }
LFD_HD double g2_YH(double l) { double p = l+1; return (1./10)*l*(2-l)/(p*p); }
LFD_HD double g5_YH(double l) { double p = l+1; return (1./20)*l*l*(1+7*l)/(p*p); }

// Constants of the singular resistance functions for radii (a0, a1)
// (Lubrication::calcLubConstants, Lubrication.cpp:165-269).
struct LubConst {
    double XA[4], YA[4], YB[4], YC[4], XG[4], YG[4], YH[4];
};

LFD_HD void lub_constants(double a0, double a1, LubConst& c)
{
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
    double g, gi;
    g = g1_XA(lambda); gi = g1_XA(invlambda);
    c.XA[0] = a0*g; c.XA[1] = ro_12*(-2/lambda_p_1)*g; c.XA[2] = c.XA[1]; c.XA[3] = a1*gi;
    g = g2_YA(lambda); gi = g2_YA(invlambda);
    c.YA[0] = a0*g; c.YA[1] = ro_12*(-2/lambda_p_1)*g; c.YA[2] = c.YA[1]; c.YA[3] = a1*gi;
    g = g2_YB(lambda); gi = g2_YB(invlambda);
    c.YB[0] = a0a0_23*g; c.YB[1] = roro_16*(-4/lambda_p_1_square)*g;
    c.YB[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; c.YB[3] = -a1a1_23*gi;
    {
        double g4 = g4_YC(lambda);
        g = g2_YC(lambda); gi = g2_YC(invlambda);
        c.YC[0] = a0a0a0_43*g; c.YC[1] = rororo_16*g4; c.YC[2] = c.YC[1]; c.YC[3] = a1a1a1_43*gi;
    }
    g = g1_XG(lambda); gi = g1_XG(invlambda);
    c.XG[0] = a0a0_23*g; c.XG[1] = roro_16*(-4/lambda_p_1_square)*g;
    c.XG[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; c.XG[3] = -a1a1_23*gi;
    g = g2_YG(lambda); gi = g2_YG(invlambda);
    c.YG[0] = a0a0_23*g; c.YG[1] = roro_16*(-4/lambda_p_1_square)*g;
    c.YG[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; c.YG[3] = -a1a1_23*gi;
    {
        double g5 = g5_YH(lambda), g5i = g5_YH(invlambda);
        g = g2_YH(lambda); gi = g2_YH(invlambda);
        c.YH[0] = a0a0a0_43*g; c.YH[1] = rororo_16*(8/lambda_p_1_cubic)*g5;
        c.YH[2] = rororo_16*(8*lambda_cubic/lambda_p_1_cubic)*g5i; c.YH[3] = a1a1a1_43*gi;
    }
}

// Stresslet resistance constants XM, YM of a pair (Lubrication.cpp:278-293,
// LubricationFunctions.h); the XG, YG of LubConst enter the GU stresslet too.
LFD_HD double g1_XM(double l) { double p = l+1; return (6./5)*l*l/(p*p*p); }
LFD_HD double g4_XM(double l) { double p = l+1; return (6./5)*l*l*l/(p*p*p); }
LFD_HD double g2_YM(double l) { double p = l+1; return (6./25)*l*(1-l+4*l*l)/(p*p*p); }
LFD_HD double g5_YM(double l) { double p = l+1; return (3./50)*l*l*(7-10*l+7*l*l)/(p*p*p); }
LFD_HD void lub_stress_constants(double a0, double a1, double* cXM, double* cYM)
{
    double ro = a0+a1;
    double lambda = a1/a0;
    double invlambda = 1/lambda;
    double lambda_p_1 = 1+lambda;
    double lambda_p_1_square = lambda_p_1*lambda_p_1;
    double lambda_p_1_cubic = lambda_p_1_square*lambda_p_1;
    double a0a0a0 = a0*a0*a0;
    double a1a1a1 = a1*a1*a1;
    double rororo = ro*ro*ro;
    double a0a0a0_109 = a0a0a0*(10.0/9);
    double a1a1a1_109 = a1a1a1*(10.0/9);
    double rororo_536 = rororo*(5.0/36);
    double g = g1_XM(lambda), gi = g1_XM(invlambda), g4 = g4_XM(lambda);
    cXM[0] = a0a0a0_109*g; cXM[1] = rororo_536*(8/lambda_p_1_cubic)*g4; cXM[2] = cXM[1]; cXM[3] = a1a1a1_109*gi;
    double g5 = g5_YM(lambda);
    g = g2_YM(lambda); gi = g2_YM(invlambda);
    cYM[0] = a0a0a0_109*g; cYM[1] = rororo_536*(8/lambda_p_1_cubic)*g5; cYM[2] = cYM[1]; cYM[3] = a1a1a1_109*gi;
}

// Dashpot resistances (ContactDashpot::calcDashpotResistances, ContactDashpot.cpp:107-183)
struct DashCoef { double XA[4], YA[4], YB[4], YC[4]; };

LFD_HD void dashpot_coeffs(double a0, double a1, double normal_coeff, double tangential_coeff, DashCoef& c)
{
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
    for (int j = 0; j < 4; j++) { c.XA[j] = 0; c.YA[j] = 0; c.YB[j] = 0; c.YC[j] = 0; }
    double g = g1_XA(lambda), gi = g1_XA(invlambda);
    c.XA[0] = a0*g; c.XA[1] = ro_12*(-2/lambda_p_1)*g; c.XA[2] = ro_12*c.XA[1]; c.XA[3] = a1*gi;
    for (int j = 0; j < 4; j++) c.XA[j] *= normal_coeff;
    if (tangential_coeff == 0) return;
    g = g2_YA(lambda); gi = g2_YA(invlambda);
    c.YA[0] = a0*g; c.YA[1] = ro_12*(-2/lambda_p_1)*g; c.YA[2] = ro_12*c.YA[1]; c.YA[3] = a1*gi;
    for (int j = 0; j < 4; j++) c.YA[j] *= tangential_coeff;
    g = g2_YB(lambda); gi = g2_YB(invlambda);
    c.YB[0] = a0a0_23*g; c.YB[1] = roro_16*(-4/lambda_p_1_square)*g;
    c.YB[2] = roro_16*(4*lambda_square/lambda_p_1_square)*gi; c.YB[3] = -a1a1_23*gi;
    for (int j = 0; j < 4; j++) c.YB[j] *= tangential_coeff;
    double g4 = g4_YC(lambda);
    g = g2_YC(lambda); gi = g2_YC(invlambda);
    c.YC[0] = a0a0a0_43*g; c.YC[1] = rororo_16*g4; c.YC[2] = rororo_16*c.YC[1]; c.YC[3] = a1a1a1_43*gi;
    for (int j = 0; j < 4; j++) c.YC[j] *= tangential_coeff;
}

// ---- 6x6 blocks in LF_DEM's packed layout (MatrixBlocks.h:5-84) ----
// ODBlock tangential form (Lubrication.cpp:437-467, ContactDashpot.cpp:189-214)
LFD_HD void od_tangential(double XA1, double YA1, double YB1, double YB2, double YC1, const double* n, double* b)
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
// normal-only forms (Lubrication.cpp:406-434, 470-517; ContactDashpot.cpp:216-239, 296-337)
LFD_HD void od_normal(double XA1, const double* n, double* b)
{
    for (int k = 0; k < 18; k++) b[k] = 0;
    b[0] = XA1*n[0]*n[0]; b[1] = XA1*n[0]*n[1]; b[2] = XA1*n[0]*n[2];
    b[5] = XA1*n[1]*n[1]; b[6] = XA1*n[1]*n[2]; b[8] = XA1*n[2]*n[2];
}
LFD_HD void d_normal(double XA, const double* n, double* b)
{
    for (int k = 0; k < 15; k++) b[k] = 0;
    b[0] = XA*n[0]*n[0]; b[1] = XA*n[0]*n[1]; b[2] = XA*n[0]*n[2];
    b[5] = XA*n[1]*n[1]; b[6] = XA*n[1]*n[2]; b[8] = XA*n[2]*n[2];
}
// DBlock of lubrication, tangential form (Lubrication.cpp:520-580)
LFD_HD void d_lub_tangential(double XA, double YA, double YB, double YC, const double* n, double* b)
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
// DBlock of the contact dashpot, tangential form (ContactDashpot.cpp:249-294)
LFD_HD void d_dash_tangential(double XA, double YA, double YB, double YC, const double* n, double* b)
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

// Lubrication coefficient from the reduced gap (Lubrication.cpp:696-706)
LFD_HD void lub_coeffs(double gap, double delta, double& lc, double& llc)
{
    lc = 1/(gap+delta);
    llc = lc > 1. ? log(lc) : 0.;
}

// Resistance blocks of one pair: which = 0 none, 1 dashpot, 2 lubrication.
// Interaction::RFU_DBlocks / RFU_ODBlock (Interaction.cpp:247-311).
LFD_HD int pair_blocks(const PairRec& q, double a0, double a1, const Consts& C,
                       double* db0, double* db1, double* od)
{
    if (q.flags & F_CONTACT) {
        DashCoef c;
        dashpot_coeffs(a0, a1, C.normal_coeff, C.tangential_coeff, c);
        if (C.tangential_coeff > 0) {
            d_dash_tangential(c.XA[0], c.YA[0], c.YB[0], c.YC[0], q.n, db0);
            d_dash_tangential(c.XA[3], c.YA[3], c.YB[3], c.YC[3], q.n, db1);
            od_tangential(c.XA[1], c.YA[1], c.YB[1], c.YB[2], c.YC[1], q.n, od);
        } else {
            d_normal(c.XA[0], q.n, db0);
            d_normal(c.XA[3], q.n, db1);
            od_normal(c.XA[1], q.n, od);
        }
        return 1;
    }
    if (q.flags & F_LUB) {
        LubConst c;
        lub_constants(a0, a1, c);
        double lc, llc;
        lub_coeffs(q.gap, C.lub_reduce, lc, llc);
        double XA[4], YA[4], YB[4], YC[4];
        for (int j = 0; j < 4; j++) {
            XA[j] = c.XA[j]*lc; YA[j] = c.YA[j]*llc; YB[j] = c.YB[j]*llc; YC[j] = c.YC[j]*llc;
        }
        if (C.tangential) {
            d_lub_tangential(XA[0], YA[0], YB[0], YC[0], q.n, db0);
            d_lub_tangential(XA[3], YA[3], YB[3], YC[3], q.n, db1);
            od_tangential(XA[1], YA[1], YB[1], YB[2], YC[1], q.n, od);
        } else {
            d_normal(XA[0], q.n, db0);
            d_normal(XA[3], q.n, db1);
            od_normal(XA[1], q.n, od);
        }
        return 2;
    }
    for (int k = 0; k < 15; k++) { db0[k] = 0; db1[k] = 0; }
    for (int k = 0; k < 18; k++) od[k] = 0;
    return 0;
}

// ---- pair state updates ----
// Interaction::calcNormalVectorDistanceGap (Interaction.cpp:255-266)
LFD_HD void pair_geometry(PairRec& q, const double* x0, const double* x1, double ro, const double* L,
                          const double* half, const double* sd)
{
    double rv[3] = {x1[0]-x0[0], x1[1]-x0[1], x1[2]-x0[2]};
    q.zoff = periodize_diff(L, half, sd, rv);
    q.r = sqrt(rv[0]*rv[0]+rv[1]*rv[1]+rv[2]*rv[2]);
    q.n[0] = rv[0]/q.r; q.n[1] = rv[1]/q.r; q.n[2] = rv[2]/q.r;
    q.gap = 2*q.r/ro-2;
}

// Contact::activate / deactivate (Contact.cpp:73-117)
LFD_HD void contact_activate(PairRec& q, int friction)
{
    q.flags |= F_CONTACT;
    for (int k = 0; k < 3; k++) q.f_total[k] = 0;
    if (friction) {
        q.flags = with_state(q.flags, 2);
        for (int k = 0; k < 3; k++) { q.disp_tan[k] = 0; q.prev_disp_tan[k] = 0; }
    } else {
        q.flags = with_state(q.flags, 1);
    }
}
LFD_HD void contact_deactivate(PairRec& q, int friction)
{
    q.flags &= ~F_CONTACT;
    q.flags = with_state(q.flags, 0);
    for (int k = 0; k < 3; k++) q.f_total[k] = 0;
    if (friction) for (int k = 0; k < 3; k++) { q.disp_tan[k] = 0; q.f_tan[k] = 0; }
}

// Contact::calcContactSpringForce + frictionlaw_standard (Contact.cpp:200-298)
LFD_HD void contact_spring(PairRec& q, double a0, double a1, const Consts& C)
{
    double ro_12 = (a0+a1)/2;
    double kn_s = ro_12*ro_12*C.kn;
    const double* n = q.n;
    double fnn = -kn_s*q.gap;
    double fn[3] = {-fnn*n[0], -fnn*n[1], -fnn*n[2]};
    int state = pair_state(q.flags);
    if (C.friction) {
        double kt_s = ro_12*C.kt;
        double d = dot3(q.disp_tan, n);
        q.disp_tan[0] -= d*n[0]; q.disp_tan[1] -= d*n[1]; q.disp_tan[2] -= d*n[2];
        q.f_tan[0] = kt_s*q.disp_tan[0]; q.f_tan[1] = kt_s*q.disp_tan[1]; q.f_tan[2] = kt_s*q.disp_tan[2];
        double sq = q.f_tan[0]*q.f_tan[0]+q.f_tan[1]*q.f_tan[1]+q.f_tan[2]*q.f_tan[2];
        double load = fnn;
        double sup = (state == 2) ? C.mu_s*load : C.mu_d*load;
        if (sq > sup*sup) { state = 3; sup = C.mu_d*load; }
        else state = 2;
        if (state == 3) {
            double f = sup/sqrt(sq);
            q.disp_tan[0] *= f; q.disp_tan[1] *= f; q.disp_tan[2] *= f;
            q.f_tan[0] = kt_s*q.disp_tan[0]; q.f_tan[1] = kt_s*q.disp_tan[1]; q.f_tan[2] = kt_s*q.disp_tan[2];
        }
        q.flags = with_state(q.flags, state);
    }
    if (state <= 1) { q.f_total[0] = fn[0]; q.f_total[1] = fn[1]; q.f_total[2] = fn[2]; }
    else { q.f_total[0] = fn[0]+q.f_tan[0]; q.f_total[1] = fn[1]+q.f_tan[1]; q.f_total[2] = fn[2]+q.f_tan[2]; }
}

// RepulsiveForce::calcReducedForceNorm + calcScaledForce (RepulsiveForce.cpp:42-85), no vdW / cutoff
LFD_HD double repulsion_norm(double r, double ro, double a0, double a1, const Consts& C)
{
    double geo = a0*a1/(a0+a1);
    double gap = r-ro, red;
    if (gap > 0) { red = exp(-gap/C.rep_len); red *= geo; }
    else red = geo;
    return C.repulsion*red;
}

// Contact::incrementTangentialDisplacement (Contact.cpp:125-160); v6 = [U, W] per particle
LFD_HD void contact_increment(PairRec& q, const double* v0, const double* v1, double a0, double a1,
                              const double* vel_diff, double dt, int in_predictor)
{
    double zo = (double)q.zoff;
    double off[3] = {zo*vel_diff[0], zo*vel_diff[1], zo*vel_diff[2]};
    double tr[3], aw[3], rot[3], rsv[3];
    for (int k = 0; k < 3; k++) tr[k] = v1[k]-v0[k]+off[k];
    for (int k = 0; k < 3; k++) aw[k] = a0*v0[3+k]+a1*v1[3+k];
    cross3(aw, q.n, rot);
    for (int k = 0; k < 3; k++) rsv[k] = tr[k]+(-rot[k]);
    double d = dot3(rsv, q.n);
    for (int k = 0; k < 3; k++) rsv[k] -= d*q.n[k];
    q.rsv_sq = rsv[0]*rsv[0]+rsv[1]*rsv[1]+rsv[2]*rsv[2];
    if (in_predictor) for (int k = 0; k < 3; k++) q.prev_disp_tan[k] = q.disp_tan[k];
    for (int k = 0; k < 3; k++) q.disp_tan[k] = q.prev_disp_tan[k]+rsv[k]*dt;
}

// New pair: Interaction ctor + activateForceMembers (Interaction.cpp:13-36, 137-156)
LFD_HD void pair_create(PairRec& q, int i, int j, const double* xi, const double* xj, double ri, double rj,
                        const Consts& C, const double* sd)
{
    q.p0 = i; q.p1 = j;
    q.flags = 0;
    for (int k = 0; k < 3; k++) {
        q.disp_tan[k] = 0; q.prev_disp_tan[k] = 0; q.f_total[k] = 0; q.f_tan[k] = 0;
    }
    q.rsv_sq = 0;
    q.rep_norm = 0;
    q.pad[0] = q.pad[1] = 0;
    double ro = ri+rj;
    pair_geometry(q, xi, xj, ro, C.L, C.half, sd);
    if (q.gap <= 0) contact_activate(q, C.friction);
    double range = lub_range(ri, rj, C.lub_max_gap);
    if (q.r < range && q.r > ro) q.flags |= F_LUB;
}

// Interaction::updateState (Interaction.cpp:171-236); returns 1 if the pair leaves (r > range)
LFD_HD int pair_update(PairRec& q, const double* x0, const double* x1, double a0, double a1,
                       const double* v0, const double* v1, const Consts& C, const double* vel_diff,
                       const double* sd, double dt, int in_predictor)
{
    if ((q.flags & F_CONTACT) && C.friction) contact_increment(q, v0, v1, a0, a1, vel_diff, dt, in_predictor);
    double ro = a0+a1;
    pair_geometry(q, x0, x1, ro, C.L, C.half, sd);
    double range = lub_range(a0, a1, C.lub_max_gap);
    if (q.r > range) {
        if (q.flags & F_CONTACT) contact_deactivate(q, C.friction);
        q.flags &= ~F_LUB;
        return 1;
    }
    if (q.flags & F_CONTACT) {
        if (q.gap > 0) contact_deactivate(q, C.friction);
    } else if (q.gap <= 0) {
        contact_activate(q, C.friction);
    }
    if (q.flags & F_CONTACT) contact_spring(q, a0, a1, C);
    if (q.r < range && q.r > ro) q.flags |= F_LUB; else q.flags &= ~F_LUB;
    if (C.repulsive) q.rep_norm = repulsion_norm(q.r, ro, a0, a1, C);
    return 0;
}

}  // namespace lfd
