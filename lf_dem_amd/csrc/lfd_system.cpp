// lfd_system.cpp -- host side of the MI355X Stokes-resistance library:
// the device-resident System step loop, the StokesSolver staging object and
// the C ABI declared in include/lfdem_amd.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "lfd_internal.hpp"
#include "../../include/lfdem_amd.h"

using namespace lfd;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(LFD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, (n ? n : 1)*sizeof(T));
}

extern "C" const char* lfd_last_error(void) { return g_err.c_str(); }

extern "C" void lfd_default_params(lfd_params* p)
{
    std::memset(p, 0, sizeof(*p));
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

extern "C" void lfd_default_solve_opts(lfd_solve_opts* o)
{
    o->tol = 1e-12;
    o->inner_tol = 1e-2;
    o->max_outer = 40;
    o->max_inner = 2000;
    o->lanes_per_row = 4;
    o->groups = 0;
    o->kind = LFD_SOLVER_MIXED;
}

extern "C" int lfd_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static Params to_params(const lfd_params* p)
{
    Params P;
    P.shear_rate = p->shear_rate; P.theta_shear = p->theta_shear; P.lub_max_gap = p->lub_max_gap;
    P.lub_reduce = p->lub_reduce; P.kn = p->kn; P.kt = p->kt; P.mu_static = p->mu_static;
    P.mu_dynamic = p->mu_dynamic; P.tau_n = p->tau_n; P.tau_t = p->tau_t; P.sd_coeff = p->sd_coeff;
    P.repulsion = p->repulsion; P.repulsive_length = p->repulsive_length; P.disp_max = p->disp_max;
    P.dt_max = p->dt_max; P.dt_min = p->dt_min; P.friction_model = p->friction_model;
    P.lubrication_model = p->lubrication_model;
    return P;
}

static SolveOpts to_opts(const lfd_solve_opts* o)
{
    SolveOpts so;
    so.tol = o->tol; so.inner_tol = o->inner_tol; so.max_outer = o->max_outer; so.max_inner = o->max_inner;
    so.lanes_per_row = o->lanes_per_row; so.block = 256; so.groups = o->groups;
    return so;
}

// Constants derived from the parameters (System.cpp:1972-1982, 2006-2015;
// ContactDashpot.cpp:65-105)
static Consts make_consts(const Params& P, const double L[3])
{
    Consts C;
    std::memset(&C, 0, sizeof(C));
    for (int k = 0; k < 3; k++) { C.L[k] = L[k]; C.half[k] = L[k]/2; }
    double ct = std::cos(P.theta_shear), st = std::sin(P.theta_shear);
    double Eh[6] = {0, 0, ct/2, st/2, 0, 0};
    double Oh[3] = {-0.5*st, 0.5*ct, 0};
    for (int k = 0; k < 6; k++) C.Eh[k] = Eh[k];
    for (int k = 0; k < 3; k++) C.Oh[k] = Oh[k];
    C.normal_coeff = P.tau_n > 0 ? 4*P.kn*P.tau_n : 1/P.lub_reduce;
    int friction = P.friction_model != 0;
    if (!friction && P.lubrication_model != 2) C.tangential_coeff = 0;
    else if (P.tau_t > 0) C.tangential_coeff = 6*P.kt*P.tau_t;
    else C.tangential_coeff = P.lub_reduce < 1 ? std::log(1/P.lub_reduce) : 0;
    C.lub_max_gap = P.lub_max_gap;
    C.lub_reduce = P.lub_reduce;
    C.kn = P.kn; C.kt = P.kt; C.mu_s = P.mu_static; C.mu_d = P.mu_dynamic; C.sd_coeff = P.sd_coeff;
    C.repulsion = P.repulsion; C.rep_len = P.repulsive_length;
    C.friction = friction;
    C.repulsive = P.repulsive_length > 0;
    C.tangential = P.lubrication_model == 2;
    return C;
}

// ===================================================================== solver storage
struct Solver {
    int n = 0, pcap = 0;
    SolveOpts opts{};
    int kind = LFD_SOLVER_MIXED;
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof;   // per solve launch
    double *diag64 = nullptr, *od64 = nullptr, *minv64 = nullptr;
    float *diag32 = nullptr, *od32 = nullptr, *minv32 = nullptr;
    int *p0 = nullptr, *p1 = nullptr, *row_ptr = nullptr, *tptr = nullptr, *tidx = nullptr;
    SolverWork w{};
    hipError_t alloc(int n_, int pcap_)
    {
        n = n_; pcap = pcap_;
        size_t m = 6*(size_t)n;
        hipError_t e = hipSuccess;
#define A(x) if (e == hipSuccess) e = (x)
        A(dalloc(&diag64, 15*(size_t)n)); A(dalloc(&diag32, 15*(size_t)n));
        A(dalloc(&minv64, 36*(size_t)n)); A(dalloc(&minv32, 36*(size_t)n));
        A(dalloc(&od64, 18*(size_t)pcap)); A(dalloc(&od32, 18*(size_t)pcap));
        A(dalloc(&p0, pcap)); A(dalloc(&p1, pcap)); A(dalloc(&tidx, pcap));
        A(dalloc(&row_ptr, n+1)); A(dalloc(&tptr, n+1));
        A(dalloc(&w.r64, m)); A(dalloc(&w.x64, m));
        A(dalloc(&w.r, m)); A(dalloc(&w.z, m)); A(dalloc(&w.p, m)); A(dalloc(&w.q, m)); A(dalloc(&w.w, m));
        A(dalloc(&w.d, m));
        A(dalloc(&w.pr, m)); A(dalloc(&w.pz, m)); A(dalloc(&w.pp, m)); A(dalloc(&w.pq, m));
        w.max_groups = 2048;
        A(dalloc(&w.partials, 8*(size_t)w.max_groups));
        A(dalloc(&w.barrier, 16384));
        A(hipMemset(w.barrier, 0, 16384*sizeof(unsigned)));
#undef A
        return e;
    }
    void release()
    {
        void* ptrs[] = {diag64, diag32, minv64, minv32, od64, od32, p0, p1, tidx, row_ptr, tptr,
                        w.r64, w.x64, w.r, w.z, w.p, w.q, w.w, w.d, w.pr, w.pz, w.pp, w.pq, w.partials,
                        w.barrier};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        for (auto& e : prof) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        prof.clear();
    }
    BsrView view(const int* npairs) const
    {
        BsrView v;
        v.n = n; v.npairs = npairs; v.p0 = p0; v.p1 = p1; v.row_ptr = row_ptr; v.tptr = tptr; v.tidx = tidx;
        v.diag64 = diag64; v.diag32 = diag32; v.od64 = od64; v.od32 = od32; v.minv32 = minv32; v.minv64 = minv64;
        return v;
    }
};

static int check_scalars(DevScalars* sc_dev, hipStream_t s, DevScalars* out = nullptr)
{
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, sc_dev, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (out) *out = h;
    if (h.error & E_BOX_OOB) return fail(LFD_ERR_BOX, "BoxSet: trying to box position out of boundaries");
    if (h.error & (E_FOUND_OVERFLOW | E_PAIR_OVERFLOW)) return fail(LFD_ERR_CAPACITY, "pair capacity exceeded");
    if (h.error & E_NOT_SPD) return fail(LFD_ERR_NOT_SPD, "resistance diagonal block not positive definite");
    if (h.error & E_SYNC_TIMEOUT) return fail(LFD_ERR_HIP, "solver workgroups lost synchronisation (spin limit)");
    if (h.error & E_NOT_CONVERGED) return fail(LFD_ERR_NOT_CONVERGED, "solver did not reach the tolerance");
    return LFD_OK;
}

static int run_solve(Solver& S, const int* npairs_dev, const double* b, double* x, DevScalars* sc,
                     hipStream_t s)
{
    BsrView A = S.view(npairs_dev);
    if (S.kind == LFD_SOLVER_PLAIN) {
        int it = 0;
        double res = 0;
        HIPCHK(solve_plain(A, b, x, S.w, sc, S.opts, s, &it, &res));
        DevScalars h;
        HIPCHK(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        h.iters_inner = it; h.iters_outer = 1; h.res_norm = res; h.rhs_norm = 1; h.total_inner += it;
        if (!(res <= S.opts.tol)) h.error |= E_NOT_CONVERGED;
        HIPCHK(hipMemcpyAsync(sc, &h, sizeof(h), hipMemcpyHostToDevice, s));
        return LFD_OK;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (S.profiling) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, s));
    }
    hipError_t e = hipErrorCooperativeLaunchTooLarge;
    if (S.kind == LFD_SOLVER_MIXED) e = solve_lds(A, b, x, S.w, sc, S.opts, s);
    if (e == hipErrorCooperativeLaunchTooLarge) {
        (void)hipGetLastError();
        e = solve_mixed(A, b, x, S.w, sc, S.opts, s);
    }
    HIPCHK(e);
    if (S.profiling) {
        HIPCHK(hipEventRecord(e1, s));
        S.prof.emplace_back(e0, e1);
    }
    return LFD_OK;
}

// Sum of the durations of the profiled solver launches (HIP events on the
// library's stream), then the record is cleared.
static int collect_profile(Solver& S, hipStream_t s, double* ms_total, int* launches)
{
    HIPCHK(hipStreamSynchronize(s));
    double tot = 0;
    for (auto& e : S.prof) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (ms_total) *ms_total = tot;
    if (launches) *launches = (int)S.prof.size();
    S.prof.clear();
    return LFD_OK;
}

// ===================================================================== System
struct lfd_system {
    int n = 0, pcap = 0, K = 64;
    Params P{};
    Consts C{};
    CellGrid g{};
    hipStream_t s = nullptr;
    double4* pos = nullptr;
    double *vel6 = nullptr, *na6 = nullptr, *pred6 = nullptr, *rhs = nullptr;
    int *box_of = nullptr, *snap_start = nullptr, *snap_mem = nullptr, *home_start = nullptr, *home_mem = nullptr;
    int *bin_count = nullptr, *bin_cursor = nullptr, *found = nullptr, *found_cnt = nullptr;
    PairRec* pairs[2] = {nullptr, nullptr};
    int cur = 0;
    int *keep = nullptr, *keep_scan = nullptr, *tcnt = nullptr, *merge_cnt = nullptr, *scan_tmp = nullptr;
    int* probe = nullptr;
    double *pdb = nullptr, *prhs = nullptr;
    // velocity parts (contact, dashpot, hydro, repulsion / rate independent,
    // rate proportional) with their right-hand sides, and the stress buffers
    double *rhs_parts = nullptr, *vel_parts = nullptr, *pst = nullptr, *ps = nullptr, *st_partial = nullptr;
    int control = LFD_RATE_CONTROL;
    int have_stress = 0;
    DevScalars* sc = nullptr;
    Solver S;
    ~lfd_system();
};

lfd_system::~lfd_system()
{
    if (s) (void)hipStreamSynchronize(s);
    void* ptrs[] = {pos, vel6, na6, pred6, rhs, box_of, snap_start, snap_mem, home_start, home_mem, bin_count,
                    bin_cursor, found, found_cnt, pairs[0], pairs[1], keep, keep_scan, tcnt, merge_cnt,
                    scan_tmp, sc, g.nb, g.nnb, g.nstatic, g.ctype, probe, pdb, prhs, rhs_parts, vel_parts,
                    pst, ps, st_partial};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    S.release();
    if (s) (void)hipStreamDestroy(s);
}

// BoxSet::init (BoxSet.cpp:7-70, 170-306): grid, probes and static neighbours
static void build_grid_host(lfd_system* y, double max_range, const double sd[3], std::vector<int>& nb,
                            std::vector<int>& nnb, std::vector<int>& ctype)
{
    CellGrid& g = y->g;
    const double* L = y->C.L;
    unsigned nx = (unsigned)(L[0]/max_range), ny = (unsigned)(L[1]/max_range), nz = (unsigned)(L[2]/max_range);
    if (nx == 0) nx = 1;
    if (ny == 0) ny = 1;
    if (nz == 0) nz = 1;
    if (nx < 4 && ny < 4 && nz < 4) {
        g.boxed = 0;
        nx = ny = nz = 1;
        g.bx = L[0]; g.by = L[1]; g.bz = L[2];
    } else {
        g.boxed = 1;
        g.bx = L[0]/nx; g.by = L[1]/ny; g.bz = L[2]/nz;
    }
    g.nx = nx; g.ny = ny; g.nz = nz; g.yz = ny*nz; g.nbox = nx*ny*nz;
    int k = 0;
    for (int a : {-1, 1})
        for (int b : {-1, 1}) {
            double fc[3] = {1.4999999*(a*g.bx), 1.4999999*(b*g.by), 1.4999999*g.bz};
            double t[4][3] = {{fc[0], fc[1], fc[2]}, {fc[0]-a*g.bx, fc[1]-0, fc[2]-0},
                              {fc[0]-a*g.bx, fc[1]-b*g.by, fc[2]-0}, {fc[0]-0, fc[1]-b*g.by, fc[2]-0}};
            double fb[3] = {1.4999999*(a*g.bx), 1.4999999*(b*g.by), 1.4999999*(-g.bz)};
            double u[4][3] = {{fb[0], fb[1], fb[2]}, {fb[0]-a*g.bx, fb[1]-0, fb[2]-0},
                              {fb[0]-a*g.bx, fb[1]-b*g.by, fb[2]-0}, {fb[0]-0, fb[1]-b*g.by, fb[2]-0}};
            for (int m = 0; m < 4; m++)
                for (int c = 0; c < 3; c++) { g.probe_top[k+m][c] = t[m][c]; g.probe_bot[k+m][c] = u[m][c]; }
            k += 4;
        }
    nb.assign(64*(size_t)g.nbox, 0);
    nnb.assign(g.nbox, 0);
    ctype.assign(g.nbox, 0);
    for (int bidx = 0; bidx < g.nbox; bidx++) {
        unsigned iz = bidx%g.nz, iy = (bidx%g.yz)/g.nz, ix = bidx/g.yz;
        int t;
        if (iz == 0 && iz < (unsigned)g.nz-1) t = 1;
        else if (iz == (unsigned)g.nz-1 && iz > 0) t = 2;
        else if (iz == 0 && iz == (unsigned)g.nz-1) t = 3;
        else t = 0;
        ctype[bidx] = t;
        int cnt = 0;
        nb[64*(size_t)bidx+cnt++] = bidx;
        if (g.boxed) {
            double c[3] = {g.bx*(ix+0.5), g.by*(iy+0.5), g.bz*(iz+0.5)};
            int clo = (t == 0 || t == 2) ? -1 : 0, chi = (t == 0 || t == 1) ? 1 : 0;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++)
                    for (int cc = clo; cc <= chi; cc++) {
                        double q[3] = {c[0]+a*g.bx, c[1]+b*g.by, c[2]+cc*g.bz};
                        periodize(L, sd, q);
                        int l = which_box(q, g.bx, g.by, g.bz, (unsigned)g.yz, (unsigned)g.nz, (unsigned)g.nbox);
                        if (l < 0) continue;
                        bool dup = false;
                        for (int m = 0; m < cnt; m++) if (nb[64*(size_t)bidx+m] == l) dup = true;
                        if (!dup) nb[64*(size_t)bidx+cnt++] = l;
                    }
        }
        nnb[bidx] = cnt;
    }
}

static int sys_bin_snapshot(lfd_system* y)
{
    // BoxSet::update (BoxSet.cpp:498-506): moving neighbours, then the
    // neighbourhood snapshot (cell membership at this instant)
    launch_moving_neighbors(y->g, y->sc, y->C.L, y->probe, y->sc, y->s);
    launch_bin(y->box_of, y->n, y->g.nbox, y->bin_count, y->snap_start, y->snap_mem, y->bin_cursor, y->scan_tmp, y->s);
    return LFD_OK;
}

static int sys_check_new(lfd_system* y)
{
    // System::checkNewInteraction (System.cpp:1328-1353)
    launch_bin(y->box_of, y->n, y->g.nbox, y->bin_count, y->home_start, y->home_mem, y->bin_cursor, y->scan_tmp, y->s);
    launch_find_pairs(y->pos, y->n, y->g, y->home_start, y->home_mem, y->snap_start, y->snap_mem, y->sc, y->C,
                      y->found, y->found_cnt, y->K, y->sc, y->s);
    PairRec* oldp = y->pairs[y->cur];
    PairRec* newp = y->pairs[y->cur^1];
    launch_merge_pairs(y->n, oldp, y->S.p1, y->S.row_ptr, y->found, y->found_cnt, y->K, y->merge_cnt, y->keep_scan,
                       y->keep, newp, y->pos, y->C, y->sc, y->pcap, y->scan_tmp, y->s);
    y->cur ^= 1;
    launch_pair_index(newp, y->n, y->S.row_ptr, y->S.p0, y->S.p1, y->S.tptr, y->S.tidx, y->tcnt, y->keep, y->sc,
                      y->pcap, y->scan_tmp, y->s);
    return LFD_OK;
}

static int sys_update(lfd_system* y, int in_predictor, double dt_override)
{
    // System::updateInteractions (System.cpp:1375-1403)
    PairRec* cp = y->pairs[y->cur];
    PairRec* np_ = y->pairs[y->cur^1];
    launch_update_pairs(cp, y->pos, y->vel6, y->C, y->sc, y->keep, y->pcap, in_predictor, dt_override, y->s);
    launch_compact_pairs(cp, np_, y->keep, y->keep_scan, y->n, y->S.row_ptr, y->S.p0, y->S.p1, y->S.tptr,
                         y->S.tidx, y->tcnt, y->sc, y->pcap, y->scan_tmp, y->s);
    y->cur ^= 1;
    return LFD_OK;
}

static int sys_assemble(lfd_system* y, int nparts = 0)
{
    launch_assemble(y->pairs[y->cur], y->pos, y->n, y->C, y->sc, y->S.row_ptr, y->S.tptr, y->S.tidx, y->S.diag64,
                    y->S.diag32, y->S.od64, y->S.od32, y->S.minv64, y->S.minv32, y->rhs, y->pdb, y->prhs, y->pcap,
                    y->rhs_parts, nparts, y->sc, y->s);
    return LFD_OK;
}

// calcStressPerParticle + calcStress (mode 1) or computeShearRate (mode 0)
static void sys_stress(lfd_system* y, int nparts, int mode)
{
    launch_stress(y->pairs[y->cur], y->pos, y->C, y->sc, y->vel_parts, nparts, y->vel6, y->n, y->control, mode,
                  y->S.row_ptr, y->S.tptr, y->S.tidx, y->pst, y->ps, y->st_partial, y->pcap, y->s);
}

static int sys_compute_velocities(lfd_system* y, int in_predictor, int calc_stress = 0)
{
    // System::computeVelocities (System.cpp:2273-2319)
    if (y->control == LFD_RATE_CONTROL && !calc_stress) {
        sys_assemble(y);
        int rc = run_solve(y->S, &y->sc->npairs, y->rhs, y->na6, y->sc, y->s);
        if (rc) return rc;
        launch_finish_velocity(y->n, y->pos, y->na6, y->vel6, y->C, y->sc, in_predictor, y->s);
        return LFD_OK;
    }
    // velocities by parts (computeVelocityByComponents, System.cpp:1947-1965):
    // four components when the stress is wanted, else (stress control) the
    // rate independent and rate proportional sums
    const int nparts = calc_stress ? 4 : 2;
    const size_t m6 = 6*(size_t)y->n;
    if (y->control == LFD_STRESS_CONTROL) launch_set_rate(y->sc, y->C, 1.0, y->s);
    sys_assemble(y, nparts);
    for (int p = 0; p < nparts; p++) {
        int rc = run_solve(y->S, &y->sc->npairs, y->rhs_parts+p*m6, y->vel_parts+p*m6, y->sc, y->s);
        if (rc) return rc;
    }
    if (y->control == LFD_STRESS_CONTROL) sys_stress(y, nparts, 0);     // new shear rate
    launch_sum_parts(y->vel_parts, nparts, y->na6, y->n, y->sc, y->control == LFD_STRESS_CONTROL, y->s);
    launch_finish_velocity(y->n, y->pos, y->na6, y->vel6, y->C, y->sc, in_predictor, y->s);
    if (calc_stress) {
        sys_stress(y, nparts, 1);
        y->have_stress = 1;
    }
    return LFD_OK;
}

static int sys_step(lfd_system* y, double strain_end, int calc_stress = 0, int adapt_first = 1)
{
    // System::timeEvolution loop body + timeEvolutionPredictorCorrectorMethod
    // (System.cpp:1241-1256, 888-983, 1124-1187); the stress, when wanted, is
    // the corrector's
    int rc;
    if (adapt_first) launch_adapt_dt(y->sc, y->P, strain_end, 0, y->s);
    if ((rc = sys_compute_velocities(y, 1))) return rc;
    launch_adapt_dt(y->sc, y->P, strain_end, 1, y->s);
    launch_boxing(y->sc, y->C, y->s);
    sys_bin_snapshot(y);
    launch_displace(y->pos, y->vel6, y->pred6, y->n, y->C, y->sc, y->g, y->box_of, 0, y->s);
    sys_update(y, 1, -1.0);
    launch_copy_double(y->pred6, y->vel6, 6*(size_t)y->n, y->s);
    if ((rc = sys_compute_velocities(y, 0, calc_stress))) return rc;
    launch_displace(y->pos, y->vel6, y->pred6, y->n, y->C, y->sc, y->g, y->box_of, 1, y->s);
    sys_check_new(y);
    sys_update(y, 0, -1.0);
    return LFD_OK;
}

extern "C" int lfd_system_create(int np, const double* pos, const double* radius, const double L[3],
                                 const double shear_disp[3], const lfd_params* p, const lfd_solve_opts* o,
                                 lfd_system** out)
{
    if (!out || np < 2 || !pos || !radius || !L || !p) return fail(LFD_ERR_ARG, "lfd_system_create: bad argument");
    *out = nullptr;
    lfd_system* y = new lfd_system();
    y->n = np;
    y->P = to_params(p);
    y->C = make_consts(y->P, L);
    lfd_solve_opts od;
    lfd_default_solve_opts(&od);
    if (!o) o = &od;
    y->S.opts = to_opts(o);
    y->S.kind = o->kind;
    y->pcap = 24*np+1024;
    double sd[3] = {0, 0, 0};
    if (shear_disp) for (int k = 0; k < 3; k++) sd[k] = shear_disp[k];
    // max interaction range over the (radius_i, radius_j), i < j, that occur
    // (System::initializeBoxing, System.cpp:613-652)
    std::vector<double> uniq(radius, radius+np);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::vector<int> first(uniq.size(), np), last(uniq.size(), -1);
    for (int i = 0; i < np; i++) {
        size_t u = std::lower_bound(uniq.begin(), uniq.end(), radius[i])-uniq.begin();
        first[u] = std::min(first[u], i);
        last[u] = std::max(last[u], i);
    }
    double max_range = 0;
    for (size_t a = 0; a < uniq.size(); a++)
        for (size_t b = 0; b < uniq.size(); b++) {
            if (!(first[a] < last[b])) continue;     // some i (radius a) < j (radius b)
            double r = lub_range(uniq[a], uniq[b], y->P.lub_max_gap);
            if (r > max_range) max_range = r;
        }
    std::vector<int> nb, nnb, ctype;
    build_grid_host(y, max_range, sd, nb, nnb, ctype);
    hipError_t e = hipStreamCreateWithFlags(&y->s, hipStreamNonBlocking);
    int n = np, nbox = y->g.nbox, pcap = y->pcap;
    size_t m6 = 6*(size_t)n;
#define A(x) if (e == hipSuccess) e = (x)
    A(dalloc(&y->pos, n)); A(dalloc(&y->vel6, m6)); A(dalloc(&y->na6, m6)); A(dalloc(&y->pred6, m6));
    A(dalloc(&y->rhs, m6)); A(dalloc(&y->box_of, n)); A(dalloc(&y->snap_start, nbox+1));
    A(dalloc(&y->snap_mem, n)); A(dalloc(&y->home_start, nbox+1)); A(dalloc(&y->home_mem, n));
    A(dalloc(&y->bin_count, nbox)); A(dalloc(&y->bin_cursor, nbox)); A(dalloc(&y->found, (size_t)n*y->K));
    A(dalloc(&y->found_cnt, n)); A(dalloc(&y->pairs[0], pcap)); A(dalloc(&y->pairs[1], pcap));
    A(dalloc(&y->keep, pcap)); A(dalloc(&y->keep_scan, (size_t)std::max(pcap, n)+1)); A(dalloc(&y->tcnt, n));
    A(dalloc(&y->merge_cnt, n)); A(dalloc(&y->scan_tmp, 4096)); A(dalloc(&y->sc, 1));
    A(dalloc(&y->g.nb, 64*(size_t)nbox)); A(dalloc(&y->g.nnb, nbox)); A(dalloc(&y->g.nstatic, nbox));
    A(dalloc(&y->g.ctype, nbox)); A(dalloc(&y->probe, 32*(size_t)nbox));
    A(dalloc(&y->pdb, 30*(size_t)pcap)); A(dalloc(&y->prhs, 42*(size_t)pcap));
    A(dalloc(&y->rhs_parts, 4*m6)); A(dalloc(&y->vel_parts, 4*m6));
    A(dalloc(&y->pst, 2*(size_t)STRESS_VALS*pcap)); A(dalloc(&y->ps, (size_t)STRESS_VALS*n));
    A(dalloc(&y->st_partial, (size_t)STRESS_VALS*((n+STRESS_CHUNK-1)/STRESS_CHUNK)));
    A(y->S.alloc(n, pcap));
#undef A
    if (e != hipSuccess) {
        delete y;
        return fail(LFD_ERR_HIP, std::string("lfd_system_create: ") + hipGetErrorString(e));
    }
    std::vector<double4> h4(n);
    for (int i = 0; i < n; i++) h4[i] = make_double4(pos[3*i], pos[3*i+1], pos[3*i+2], radius[i]);
    DevScalars hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.dt = 1e-4;
    for (int k = 0; k < 3; k++) hs.shear_disp[k] = sd[k];
    make_flow(y->C, y->P.shear_rate, hs.flow);
    hs.init_strain_limit = 0;       // System.cpp:90-91
    hs.init_rate_limit = 999;
    int rc = LFD_OK;
    auto up = [&](void* d, const void* h, size_t b) {
        if (e == hipSuccess) e = hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, y->s);
    };
    up(y->pos, h4.data(), n*sizeof(double4));
    up(y->sc, &hs, sizeof(hs));
    up(y->g.nb, nb.data(), nb.size()*sizeof(int));
    up(y->g.nnb, nnb.data(), nnb.size()*sizeof(int));
    up(y->g.nstatic, nnb.data(), nnb.size()*sizeof(int));
    up(y->g.ctype, ctype.data(), ctype.size()*sizeof(int));
    if (e != hipSuccess) {
        delete y;
        return fail(LFD_ERR_HIP, std::string("lfd_system_create: ") + hipGetErrorString(e));
    }
    launch_fill_double(y->vel6, 0.0, m6, y->s);
    launch_fill_double(y->na6, 0.0, m6, y->s);
    launch_fill_double(y->pred6, 0.0, m6, y->s);
    launch_fill_double(y->vel_parts, 0.0, 4*m6, y->s);
    launch_fill_int(y->S.row_ptr, 0, n+1, y->s);
    launch_box_labels(y->pos, n, y->g, y->box_of, y->sc, y->s);
    sys_bin_snapshot(y);
    sys_check_new(y);
    rc = check_scalars(y->sc, y->s);
    if (rc) { delete y; return rc; }
    *out = y;
    return LFD_OK;
}

extern "C" void lfd_system_destroy(lfd_system* y) { delete y; }

extern "C" int lfd_system_profile(lfd_system* y, int enable, double* solve_ms_total, int* solve_launches)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = collect_profile(y->S, y->s, solve_ms_total, solve_launches);
    y->S.profiling = enable != 0;
    return rc;
}

extern "C" int lfd_system_set_opts(lfd_system* y, const lfd_solve_opts* o)
{
    if (!y || !o) return fail(LFD_ERR_ARG, "lfd_system_set_opts: bad argument");
    y->S.opts = to_opts(o);
    y->S.kind = o->kind;
    return LFD_OK;
}

extern "C" int lfd_system_first_update(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    sys_check_new(y);
    sys_update(y, 1, 0.0);
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_step(lfd_system* y, int nsteps, double strain_end)
{
    if (!y || nsteps < 0) return fail(LFD_ERR_ARG, "lfd_system_step: bad argument");
    for (int k = 0; k < nsteps; k++) {
        int rc = sys_step(y, strain_end);
        if (rc) return rc;
    }
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_set_control(lfd_system* y, int control, double target_stress)
{
    if (!y || (control != LFD_RATE_CONTROL && control != LFD_STRESS_CONTROL))
        return fail(LFD_ERR_ARG, "lfd_system_set_control: bad argument");
    y->control = control;
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, y->sc, sizeof(h), hipMemcpyDeviceToHost, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    h.target_stress = target_stress;
    HIPCHK(hipMemcpyAsync(y->sc, &h, sizeof(h), hipMemcpyHostToDevice, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    return LFD_OK;
}

extern "C" int lfd_system_set_shear_rate_limit(lfd_system* y, double init_strain_limit, double init_rate_limit)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, y->sc, sizeof(h), hipMemcpyDeviceToHost, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    h.init_strain_limit = init_strain_limit;
    h.init_rate_limit = init_rate_limit;
    HIPCHK(hipMemcpyAsync(y->sc, &h, sizeof(h), hipMemcpyHostToDevice, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    return LFD_OK;
}

extern "C" int lfd_system_step_stress(lfd_system* y, double strain_end)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = sys_step(y, strain_end, 1, 0);
    if (rc) return rc;
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_time_evolution(lfd_system* y, double strain_end, int* nsteps)
{
    if (!y || !(strain_end >= 0)) return fail(LFD_ERR_ARG, "lfd_system_time_evolution: bad argument");
    int n = 0;
    for (;;) {
        DevScalars h;
        int rc = check_scalars(y->sc, y->s, &h);
        if (rc) return rc;
        if (h.strain > strain_end-1e-8) break;
        if ((rc = sys_step(y, strain_end))) return rc;
        n++;
    }
    int rc = sys_step(y, strain_end, 1, 0);
    if (rc) return rc;
    if (nsteps) *nsteps = n+1;
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_compute_stress(lfd_system* y, int in_predictor)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = sys_compute_velocities(y, in_predictor, 1);
    if (rc) return rc;
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_get_stress(lfd_system* y, double* groups24, double* total6, double* shear_rate,
                                     double* viscosity, double* particle6)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    if (!y->have_stress) return fail(LFD_ERR_ARG, "lfd_system_get_stress: no stress computed yet");
    DevScalars h;
    int rc = check_scalars(y->sc, y->s, &h);
    if (rc) return rc;
    if (groups24) std::memcpy(groups24, h.stress_groups, 24*sizeof(double));
    if (total6) std::memcpy(total6, h.stress_total, 6*sizeof(double));
    if (shear_rate) *shear_rate = h.flow.rate;
    if (viscosity) {
        // Simulation::outputData (Simulation.cpp:662-666)
        const double* E = h.flow.E;
        if (h.flow.rate == 0) E = y->C.Eh;
        auto dd = [](const double* a, const double* b) {
            return a[0]*b[0]+2*a[1]*b[1]+2*a[2]*b[2]+2*a[3]*b[3]+a[4]*b[4]+a[5]*b[5];
        };
        *viscosity = 0.5*dd(h.stress_total, E)/dd(E, E);
    }
    if (particle6) {
        std::vector<double> ps((size_t)STRESS_VALS*y->n);
        HIPCHK(hipMemcpyAsync(ps.data(), y->ps, ps.size()*sizeof(double), hipMemcpyDeviceToHost, y->s));
        HIPCHK(hipStreamSynchronize(y->s));
        for (int i = 0; i < y->n; i++)
            for (int m = 0; m < 6; m++) {
                double t = 0;
                for (int sl = 0; sl < STRESS_VALS/6; sl++) t += ps[(size_t)STRESS_VALS*i+6*sl+m];
                particle6[6*(size_t)i+m] = t;
            }
    }
    return LFD_OK;
}

extern "C" int lfd_system_get_parts(lfd_system* y, int part, double* vel6)
{
    if (!y || part < 0 || part > 3 || !vel6) return fail(LFD_ERR_ARG, "lfd_system_get_parts: bad argument");
    const size_t m6 = 6*(size_t)y->n;
    HIPCHK(hipMemcpyAsync(vel6, y->vel_parts+part*m6, m6*sizeof(double), hipMemcpyDeviceToHost, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    return LFD_OK;
}

extern "C" int lfd_system_sync(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_check_new_interactions(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    sys_check_new(y);
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_update_interactions(lfd_system* y, int in_predictor, double dt)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    sys_update(y, in_predictor, dt);
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_assemble(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    sys_assemble(y);
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_compute_velocities(lfd_system* y, int in_predictor)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = sys_compute_velocities(y, in_predictor);
    if (rc) return rc;
    return check_scalars(y->sc, y->s);
}

extern "C" int lfd_system_num_pairs(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    DevScalars h;
    int rc = check_scalars(y->sc, y->s, &h);
    if (rc) return rc;
    return h.npairs;
}

extern "C" int lfd_system_get_pairs(lfd_system* y, int cap, int* p0, int* p1, int* state, double* geom5,
                                    double* disp_tan)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    DevScalars h;
    int rc = check_scalars(y->sc, y->s, &h);
    if (rc) return rc;
    int np_ = h.npairs;
    if (np_ > cap) return fail(LFD_ERR_ARG, "lfd_system_get_pairs: cap too small");
    std::vector<PairRec> v(np_);
    if (np_) HIPCHK(hipMemcpy(v.data(), y->pairs[y->cur], np_*sizeof(PairRec), hipMemcpyDeviceToHost));
    for (int k = 0; k < np_; k++) {
        const PairRec& q = v[k];
        if (p0) p0[k] = q.p0;
        if (p1) p1[k] = q.p1;
        if (state) state[k] = pair_state(q.flags);
        if (geom5) {
            geom5[5*k] = q.n[0]; geom5[5*k+1] = q.n[1]; geom5[5*k+2] = q.n[2];
            geom5[5*k+3] = q.r; geom5[5*k+4] = q.gap;
        }
        if (disp_tan) for (int c = 0; c < 3; c++) disp_tan[3*k+c] = q.disp_tan[c];
    }
    return np_;
}

extern "C" int lfd_system_get_state(lfd_system* y, double* pos, double* vel6, double* na6, double* sc8)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    DevScalars h;
    int rc = check_scalars(y->sc, y->s, &h);
    if (rc) return rc;
    if (pos) {
        std::vector<double4> v(y->n);
        HIPCHK(hipMemcpy(v.data(), y->pos, y->n*sizeof(double4), hipMemcpyDeviceToHost));
        for (int i = 0; i < y->n; i++) { pos[3*i] = v[i].x; pos[3*i+1] = v[i].y; pos[3*i+2] = v[i].z; }
    }
    if (vel6) HIPCHK(hipMemcpy(vel6, y->vel6, 6*(size_t)y->n*sizeof(double), hipMemcpyDeviceToHost));
    if (na6) HIPCHK(hipMemcpy(na6, y->na6, 6*(size_t)y->n*sizeof(double), hipMemcpyDeviceToHost));
    if (sc8) {
        sc8[0] = h.shear_disp[0]; sc8[1] = h.shear_disp[1]; sc8[2] = h.shear_disp[2];
        sc8[3] = h.dt; sc8[4] = h.time; sc8[5] = h.strain;
        sc8[6] = std::sqrt(__builtin_bit_cast(double, h.max_vel_bits));
        sc8[7] = std::sqrt(__builtin_bit_cast(double, h.max_slide_bits));
    }
    return LFD_OK;
}

extern "C" int lfd_system_get_cells(lfd_system* y, int* dims4, int* labels)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = check_scalars(y->sc, y->s);
    if (rc) return rc;
    if (dims4) { dims4[0] = y->g.nx; dims4[1] = y->g.ny; dims4[2] = y->g.nz; dims4[3] = y->g.boxed; }
    if (labels) HIPCHK(hipMemcpy(labels, y->box_of, y->n*sizeof(int), hipMemcpyDeviceToHost));
    return y->g.nbox;
}

extern "C" int lfd_system_get_blocks(lfd_system* y, int cap, double* dblocks, double* odblocks, int* p0, int* p1)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    DevScalars h;
    int rc = check_scalars(y->sc, y->s, &h);
    if (rc) return rc;
    int np_ = h.npairs;
    if (np_ > cap) return fail(LFD_ERR_ARG, "lfd_system_get_blocks: cap too small");
    if (dblocks) HIPCHK(hipMemcpy(dblocks, y->S.diag64, 15*(size_t)y->n*sizeof(double), hipMemcpyDeviceToHost));
    if (odblocks && np_) HIPCHK(hipMemcpy(odblocks, y->S.od64, 18*(size_t)np_*sizeof(double), hipMemcpyDeviceToHost));
    if (p0 && np_) HIPCHK(hipMemcpy(p0, y->S.p0, np_*sizeof(int), hipMemcpyDeviceToHost));
    if (p1 && np_) HIPCHK(hipMemcpy(p1, y->S.p1, np_*sizeof(int), hipMemcpyDeviceToHost));
    return np_;
}

extern "C" int lfd_system_get_rhs(lfd_system* y, double* rhs)
{
    if (!y || !rhs) return fail(LFD_ERR_ARG, "bad argument");
    int rc = check_scalars(y->sc, y->s);
    if (rc) return rc;
    HIPCHK(hipMemcpy(rhs, y->rhs, 6*(size_t)y->n*sizeof(double), hipMemcpyDeviceToHost));
    return LFD_OK;
}

extern "C" int lfd_system_get_solve_stats(lfd_system* y, lfd_solve_stats* st, long long* total_inner)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, y->sc, sizeof(h), hipMemcpyDeviceToHost, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    if (st) { st->iters_inner = h.iters_inner; st->iters_outer = h.iters_outer; st->res_norm = h.res_norm; st->rhs_norm = h.rhs_norm; }
    if (total_inner) *total_inner = h.total_inner;
    return LFD_OK;
}

// ===================================================================== StokesSolver staging
struct lfd_stokes {
    int np = 0, np_mobile = 0;
    int cap = 0;
    int current_column = 0;
    bool filled = false;
    std::vector<double> dblocks;          // np * 15
    std::vector<int> op0, op1;            // column i, row j
    std::vector<double> odb;              // 18 per block
    std::vector<double> rhs;              // 6 * np
    hipStream_t s = nullptr;
    double *d_rhs = nullptr, *d_x = nullptr, *d_y = nullptr;
    int* d_npairs = nullptr;
    DevScalars* sc = nullptr;
    Solver S;
    ~lfd_stokes()
    {
        if (s) (void)hipStreamSynchronize(s);
        void* ptrs[] = {d_rhs, d_x, d_y, d_npairs, sc};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        S.release();
        if (s) (void)hipStreamDestroy(s);
    }
};

extern "C" int lfd_stokes_create(int np_total, int np_mobile, lfd_stokes** out)
{
    if (!out || np_total < 1) return fail(LFD_ERR_ARG, "lfd_stokes_create: bad argument");
    if (np_mobile != np_total) return fail(LFD_ERR_UNSUPPORTED, "lfd_stokes_create: fixed particles are not supported");
    lfd_stokes* st = new lfd_stokes();
    st->np = np_total;
    st->np_mobile = np_mobile;
    st->dblocks.assign(15*(size_t)np_total, 0.0);
    st->rhs.assign(6*(size_t)np_total, 0.0);
    lfd_solve_opts o;
    lfd_default_solve_opts(&o);
    st->S.opts = to_opts(&o);
    st->S.kind = o.kind;
    hipError_t e = hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking);
    size_t m = 6*(size_t)np_total;
    if (e == hipSuccess) e = dalloc(&st->d_rhs, m);
    if (e == hipSuccess) e = dalloc(&st->d_x, m);
    if (e == hipSuccess) e = dalloc(&st->d_y, m);
    if (e == hipSuccess) e = dalloc(&st->d_npairs, 1);
    if (e == hipSuccess) e = dalloc(&st->sc, 1);
    if (e == hipSuccess) e = st->S.alloc(np_total, 0);
    if (e != hipSuccess) {
        delete st;
        return fail(LFD_ERR_HIP, std::string("lfd_stokes_create: ") + hipGetErrorString(e));
    }
    *out = st;
    return LFD_OK;
}

extern "C" void lfd_stokes_destroy(lfd_stokes* st) { delete st; }

extern "C" int lfd_stokes_set_opts(lfd_stokes* st, const lfd_solve_opts* o)
{
    if (!st || !o) return fail(LFD_ERR_ARG, "bad argument");
    st->S.opts = to_opts(o);
    st->S.kind = o->kind;
    return LFD_OK;
}

extern "C" int lfd_stokes_reset(lfd_stokes* st, int hint, const double* drag, int pattern_changed)
{
    if (!st || !drag) return fail(LFD_ERR_ARG, "lfd_stokes_reset: bad argument");
    (void)pattern_changed;
    std::memcpy(st->dblocks.data(), drag, st->dblocks.size()*sizeof(double));
    st->op0.clear(); st->op1.clear(); st->odb.clear();
    if (hint > 0) { st->op0.reserve(hint); st->op1.reserve(hint); st->odb.reserve(18*(size_t)hint); }
    st->current_column = 0;
    st->filled = false;
    return LFD_OK;
}

extern "C" int lfd_stokes_start_new_column(lfd_stokes* st)
{
    if (!st) return fail(LFD_ERR_ARG, "null handle");
    if (st->current_column >= st->np) return fail(LFD_ERR_ARG, "startNewColumn: too many columns");
    st->current_column++;
    return LFD_OK;
}

extern "C" int lfd_stokes_add_blocks(lfd_stokes* st, int i, int j, const double* dbi, const double* dbj,
                                     const double* odb)
{
    if (!st || !dbi || !dbj || !odb) return fail(LFD_ERR_ARG, "addResistanceBlocks: bad argument");
    if (i != st->current_column-1) return fail(LFD_ERR_ARG, "addResistanceBlocks: i is not the current column");
    if (j <= i || j >= st->np) return fail(LFD_ERR_ARG, "addResistanceBlocks: need i < j < np");
    if (!st->op0.empty() && st->op0.back() == i && st->op1.back() >= j)
        return fail(LFD_ERR_ARG, "addResistanceBlocks: rows within a column must be ascending");
    for (int m = 0; m < 15; m++) { st->dblocks[15*(size_t)i+m] += dbi[m]; st->dblocks[15*(size_t)j+m] += dbj[m]; }
    st->op0.push_back(i);
    st->op1.push_back(j);
    st->odb.insert(st->odb.end(), odb, odb+18);
    return LFD_OK;
}

extern "C" int lfd_stokes_fill_done(lfd_stokes* st)
{
    if (!st) return fail(LFD_ERR_ARG, "null handle");
    int n = st->np, P = (int)st->op0.size();
    if (P > st->cap) {
        int ncap = std::max(P, 2*st->cap);
        Solver fresh;
        fresh.opts = st->S.opts;
        fresh.kind = st->S.kind;
        st->S.release();
        st->S = fresh;
        HIPCHK(st->S.alloc(n, ncap));
        st->cap = ncap;
    }
    std::vector<int> row_ptr(n+1, 0), tcnt(n+1, 0), tidx(std::max(P, 1));
    for (int k = 0; k < P; k++) { row_ptr[st->op0[k]+1]++; tcnt[st->op1[k]+1]++; }
    for (int i = 0; i < n; i++) { row_ptr[i+1] += row_ptr[i]; tcnt[i+1] += tcnt[i]; }
    {
        std::vector<int> cur(tcnt.begin(), tcnt.end()-1);
        for (int k = 0; k < P; k++) tidx[cur[st->op1[k]]++] = k;    // ascending k == ascending p0
    }
    hipStream_t s = st->s;
    HIPCHK(hipMemcpyAsync(st->S.diag64, st->dblocks.data(), 15*(size_t)n*sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(st->S.row_ptr, row_ptr.data(), (n+1)*sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(st->S.tptr, tcnt.data(), (n+1)*sizeof(int), hipMemcpyHostToDevice, s));
    if (P) {
        HIPCHK(hipMemcpyAsync(st->S.od64, st->odb.data(), 18*(size_t)P*sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(st->S.p0, st->op0.data(), P*sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(st->S.p1, st->op1.data(), P*sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(st->S.tidx, tidx.data(), P*sizeof(int), hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(st->d_npairs, &P, sizeof(int), hipMemcpyHostToDevice, s));
    DevScalars h;
    std::memset(&h, 0, sizeof(h));
    h.npairs = P;
    HIPCHK(hipMemcpyAsync(st->sc, &h, sizeof(h), hipMemcpyHostToDevice, s));
    launch_prepare_matrix(n, st->d_npairs, st->S.diag64, st->S.diag32, st->S.od64, st->S.od32, st->S.minv64,
                          st->S.minv32, st->cap, st->sc, s);
    int rc = check_scalars(st->sc, s);
    if (rc) return rc;
    st->filled = true;
    return LFD_OK;
}

static bool rhs_index_ok(lfd_stokes* st, int i) { return st && i >= 0 && i < st->np; }

extern "C" int lfd_stokes_reset_rhs(lfd_stokes* st)
{
    if (!st) return fail(LFD_ERR_ARG, "null handle");
    std::fill(st->rhs.begin(), st->rhs.end(), 0.0);
    return LFD_OK;
}
extern "C" int lfd_stokes_reset_rhs_torque(lfd_stokes* st)
{
    if (!st) return fail(LFD_ERR_ARG, "null handle");
    for (size_t i = 3; i < st->rhs.size(); i += 6) st->rhs[i] = st->rhs[i+1] = st->rhs[i+2] = 0;
    return LFD_OK;
}
extern "C" int lfd_stokes_add_rhs_force(lfd_stokes* st, int i, const double f[3])
{
    if (!rhs_index_ok(st, i) || !f) return fail(LFD_ERR_ARG, "addToRHSForce: bad argument");
    for (int k = 0; k < 3; k++) st->rhs[6*(size_t)i+k] += f[k];
    return LFD_OK;
}
extern "C" int lfd_stokes_add_rhs_torque(lfd_stokes* st, int i, const double t[3])
{
    if (!rhs_index_ok(st, i) || !t) return fail(LFD_ERR_ARG, "addToRHSTorque: bad argument");
    for (int k = 0; k < 3; k++) st->rhs[6*(size_t)i+3+k] += t[k];
    return LFD_OK;
}
extern "C" int lfd_stokes_set_rhs_force(lfd_stokes* st, int i, const double f[3])
{
    if (!rhs_index_ok(st, i) || !f) return fail(LFD_ERR_ARG, "setRHSForce: bad argument");
    for (int k = 0; k < 3; k++) st->rhs[6*(size_t)i+k] = f[k];
    return LFD_OK;
}
extern "C" int lfd_stokes_set_rhs_torque(lfd_stokes* st, int i, const double t[3])
{
    if (!rhs_index_ok(st, i) || !t) return fail(LFD_ERR_ARG, "setRHSTorque: bad argument");
    for (int k = 0; k < 3; k++) st->rhs[6*(size_t)i+3+k] = t[k];
    return LFD_OK;
}
extern "C" int lfd_stokes_set_rhs(lfd_stokes* st, const double* ft)
{
    if (!st || !ft) return fail(LFD_ERR_ARG, "setRHS: bad argument");
    std::memcpy(st->rhs.data(), ft, st->rhs.size()*sizeof(double));
    return LFD_OK;
}
extern "C" int lfd_stokes_add_rhs(lfd_stokes* st, const double* ft)
{
    if (!st || !ft) return fail(LFD_ERR_ARG, "addToRHS: bad argument");
    for (size_t k = 0; k < st->rhs.size(); k++) st->rhs[k] += ft[k];
    return LFD_OK;
}

extern "C" int lfd_stokes_solve(lfd_stokes* st, double* uw, int uw_len, lfd_solve_stats* stats)
{
    if (!st || !uw) return fail(LFD_ERR_ARG, "solve: bad argument");
    if (!st->filled) return fail(LFD_ERR_ARG, "solve: matrixFillingDone has not been called");
    if (uw_len < 6*st->np_mobile) return fail(LFD_ERR_ARG, "StokesSolver::solve: allocated solution vector too small");
    hipStream_t s = st->s;
    size_t m = 6*(size_t)st->np;
    HIPCHK(hipMemcpyAsync(st->d_rhs, st->rhs.data(), m*sizeof(double), hipMemcpyHostToDevice, s));
    launch_fill_double(st->d_x, 0.0, m, s);
    int rc = run_solve(st->S, st->d_npairs, st->d_rhs, st->d_x, st->sc, s);
    if (rc) return rc;
    DevScalars h;
    rc = check_scalars(st->sc, s, &h);
    HIPCHK(hipMemcpy(uw, st->d_x, m*sizeof(double), hipMemcpyDeviceToHost));
    if (stats) { stats->iters_inner = h.iters_inner; stats->iters_outer = h.iters_outer; stats->res_norm = h.res_norm; stats->rhs_norm = h.rhs_norm; }
    if (rc) {
        // clear the sticky bit so the next solve reports its own outcome
        h.error = 0;
        (void)hipMemcpy(st->sc, &h, sizeof(h), hipMemcpyHostToDevice);
    }
    return rc;
}

extern "C" int lfd_stokes_solving_done(lfd_stokes* st)
{
    if (!st) return fail(LFD_ERR_ARG, "null handle");
    return LFD_OK;
}

extern "C" int lfd_stokes_multiply(lfd_stokes* st, const double* x, double* y)
{
    if (!st || !x || !y) return fail(LFD_ERR_ARG, "multiply: bad argument");
    if (!st->filled) return fail(LFD_ERR_ARG, "multiply: matrixFillingDone has not been called");
    size_t m = 6*(size_t)st->np;
    HIPCHK(hipMemcpyAsync(st->d_x, x, m*sizeof(double), hipMemcpyHostToDevice, st->s));
    launch_spmv64(st->S.view(st->d_npairs), st->d_x, st->d_y, st->s);
    HIPCHK(hipMemcpyAsync(y, st->d_y, m*sizeof(double), hipMemcpyDeviceToHost, st->s));
    HIPCHK(hipStreamSynchronize(st->s));
    return LFD_OK;
}
