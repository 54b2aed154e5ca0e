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

int lfd::set_last_error(int code, const std::string& msg) { return fail(code, msg); }

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
    o->precond = LFD_PRECOND_AGGREGATE;
    o->regroup_interval = 4;
    o->launch = LFD_LAUNCH_AUTO;
    o->max_rhs = 0;
    o->agg_waves = 0;
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
    P.lubrication_model = p->lubrication_model; P.fixed_dt = p->fixed_dt;
    return P;
}

static SolveOpts to_opts(const lfd_solve_opts* o)
{
    SolveOpts so;
    so.tol = o->tol; so.inner_tol = o->inner_tol; so.max_outer = o->max_outer; so.max_inner = o->max_inner;
    so.lanes_per_row = o->lanes_per_row; so.block = 256; so.groups = o->groups; so.launch = o->launch;
    so.max_rhs = o->max_rhs == 1 ? 1 : 2;
    so.agg_waves = o->agg_waves == AGG_WAVES_WIDE ? AGG_WAVES_WIDE
                 : o->agg_waves == AGG_LARGE_WAVES ? AGG_LARGE_WAVES : AGG_WAVES_REG;
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
    size_t nbar = 0;          // tree-barrier words in w.barrier
    // batched LDS solver: job table and its cooperative launches (first job, count);
    // no launch means the single instance `single` is solved by the global-memory form
    SolveJob* jobs = nullptr;
    int jobs_cap = 0;
    std::vector<std::pair<int, int>> launches;
    int single = -1;          // the only instance, or -1 when there are several
    // aggregate preconditioner: solver slots, groups and inverses, the
    // grouping jobs (one per instance that uses it) and the aggregate
    // solver's job table and launches
    int precond = LFD_PRECOND_AGGREGATE;
    int scap = 0;             // solver slots (AGG_SLOTS_PER_BLOCK per 128 particle slots)
    AggView agg{};
    int *wave_inst = nullptr, *wave_local = nullptr;
    std::vector<int> h_wave_inst, h_wave_local;
    AggJob* agg_jobs = nullptr;
    int n_agg_jobs = 0;
    SolveJob* ajobs = nullptr;
    int ajobs_cap = 0;
    std::vector<std::pair<int, int>> agg_launches;
    // large instances (more waves than the 8-wavefront aggregate solver holds
    // at once): grouped in spatial chunks, one launch of the 16-wave solver each
    std::vector<std::pair<int, int>> large_launches;
    std::vector<int> h_chunk;
    bool allow_large = false; // the chunked grouping takes its cell order from a batch's cell bins
    int regroup_interval = 4; // steps between an instance's groupings (the inverses follow every assembly)
    long long gen = 0;        // bumped whenever the solver's launch sequence changes
    hipError_t alloc(int n_, int pcap_, int icap = 0)
    {
        n = n_; pcap = pcap_;
        size_t m = 6*(size_t)n;
        const int rb = (n+LDS_ROWS-1)/LDS_ROWS;
        hipError_t e = hipSuccess;
#define A(x) if (e == hipSuccess) e = (x)
        A(dalloc(&diag64, 15*(size_t)n)); A(dalloc(&diag32, 15*(size_t)n));
        A(dalloc(&minv64, 36*(size_t)n)); A(dalloc(&minv32, 36*(size_t)n));
        A(dalloc(&od64, 18*(size_t)pcap)); A(dalloc(&od32, 18*(size_t)pcap));
        A(dalloc(&p0, pcap)); A(dalloc(&p1, pcap)); A(dalloc(&tidx, pcap));
        A(dalloc(&row_ptr, n+1)); A(dalloc(&tptr, n+1));
        A(dalloc(&w.r64, m)); A(dalloc(&w.x64, m));
        // w.z and w.q hold two vectors for the LDS solver's paired right-hand sides
        A(dalloc(&w.r, m)); A(dalloc(&w.z, 2*m)); A(dalloc(&w.p, m)); A(dalloc(&w.q, 2*m)); A(dalloc(&w.w, m));
        A(dalloc(&w.d, m));
        A(dalloc(&w.pr, m)); A(dalloc(&w.pz, m)); A(dalloc(&w.pp, m)); A(dalloc(&w.pq, m));
        w.max_groups = 2048;
        A(dalloc(&w.partials, std::max<size_t>(8*(size_t)w.max_groups, LDS_PART_PER_BLOCK*(size_t)rb+64)));
        // tree-barrier words: 128 per row block for the batched solver, 16384 for the global form
        nbar = std::max<size_t>(16384, 128*(size_t)rb+256);
        A(dalloc(&w.barrier, nbar));
        A(hipMemset(w.barrier, 0, nbar*sizeof(unsigned)));
        jobs_cap = rb+1;
        A(dalloc(&jobs, jobs_cap));
        if (pcap > 0) {
            scap = rb*AGG_SLOTS_PER_BLOCK;
            const int nwv = scap/AGG_CAP;
            A(dalloc(&agg.perm, scap)); A(dalloc(&agg.iperm, n)); A(dalloc(&agg.waves, std::max(icap, rb)+1));
            A(dalloc(&wave_inst, nwv)); A(dalloc(&wave_local, nwv)); A(dalloc(&agg.adj, AGG_ADJ_K*(size_t)n));
            A(dalloc(&agg.inv, (size_t)nwv*AGG_PACK));
            // zb, qb: two right-hand sides of the large-instance solver
            A(dalloc(&agg.zb, 12*(size_t)scap)); A(dalloc(&agg.qb, 12*(size_t)scap));
            agg.scap = scap;
            A(dalloc(&agg.wexact, nwv));
            A(dalloc(&agg.chunk, std::max(icap, 1))); A(dalloc(&agg.order, n)); A(dalloc(&agg.rank, n));
            h_chunk.assign(std::max(icap, 1), 0);
            A(dalloc(&agg.db, 6*(size_t)scap));
            A(dalloc(&agg.dsum[0], 6*(size_t)scap)); A(dalloc(&agg.dsum[1], 6*(size_t)scap));
            A(dalloc(&agg.r0buf, 6*(size_t)scap));
            agg.part_len = (size_t)AGG_PART_PER_BLOCK*rb+64;
            A(dalloc(&agg.part, agg.part_len));
            A(dalloc(&agg_jobs, rb+1));
            ajobs_cap = 3*rb+std::max(icap, 1)+64;    // up to ceil(2.5 m) workgroups per instance of m blocks
            A(dalloc(&ajobs, ajobs_cap));
            agg.wave_inst = wave_inst;
            agg.wave_local = wave_local;
            h_wave_inst.assign(nwv, -1);
            h_wave_local.assign(nwv, 0);
            if (e == hipSuccess) {
                launch_fill_int(agg.waves, 0, std::max(icap, rb)+1, nullptr);
                launch_fill_int(wave_inst, -1, nwv, nullptr);
                launch_fill_int(agg.iperm, -1, n, nullptr);
                launch_fill_int(agg.perm, -1, scap, nullptr);
                launch_fill_int(agg.chunk, 0, std::max(icap, 1), nullptr);
                launch_fill_int(agg.wexact, 1, nwv, nullptr);
                e = hipDeviceSynchronize();
            }
        }
#undef A
        return e;
    }
    void release()
    {
        void* ptrs[] = {diag64, diag32, minv64, minv32, od64, od32, p0, p1, tidx, row_ptr, tptr,
                        w.r64, w.x64, w.r, w.z, w.p, w.q, w.w, w.d, w.pr, w.pz, w.pp, w.pq, w.partials,
                        w.barrier, jobs, agg.perm, agg.iperm, agg.waves, wave_inst, wave_local, agg.adj, agg.inv, agg.zb,
                        agg.qb, agg.db, agg.dsum[0], agg.dsum[1], agg.r0buf, agg.part, agg_jobs, ajobs,
                        agg.wexact, agg.chunk, agg.order, agg.rank};
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

// Any persistent solver launch in the job table.
static bool has_persistent(const Solver& S)
{
    return !(S.launches.empty() && S.agg_launches.empty() && S.large_launches.empty());
}

struct InstSpan { int inst, row0, n; };

// Job table of the batched solvers.  An instance that can take the aggregate
// preconditioner gets its budget of aggregate-solver workgroups (those beyond
// the waves in use exit at once); the others get ceil(n/128) workgroups of the
// 6x6 block-Jacobi solver.  Each kind is packed in slot order into cooperative
// launches that fit the device.  An instance too large for one launch can
// only run alone, on the global-memory form.
static int set_jobs(Solver& S, const std::vector<InstSpan>& spans, hipStream_t s)
{
    S.gen++;
    S.launches.clear();
    S.agg_launches.clear();
    S.large_launches.clear();
    std::fill(S.h_chunk.begin(), S.h_chunk.end(), 0);
    S.n_agg_jobs = 0;
    S.single = spans.size() == 1 ? spans[0].inst : -1;
    if (spans.empty()) return LFD_OK;
    const int maxg = std::min(lds_solver_max_groups(), S.w.max_groups);
    // waves per workgroup of the 8- or 12-wave aggregate solver (16 asks for
    // the large-instance form, which solve_agg runs as 8 where it is not
    // available)
    const int aw = S.opts.agg_waves == AGG_WAVES_WIDE ? AGG_WAVES_WIDE : AGG_WAVES_REG;
    // agg_waves 16: every aggregate instance takes the large-instance form
    // (where it is available; otherwise the 8-wave form)
    const bool all_large = S.opts.agg_waves == AGG_LARGE_WAVES && S.allow_large;
    const int amaxg = S.scap > 0 && !all_large ? agg_solver_max_groups(aw) : 0;
    const int lmaxg = S.scap > 0 && S.allow_large ? agg_large_max_groups() : 0;
    std::vector<SolveJob> jobs, ajobs, ljobs;
    std::vector<AggJob> gjobs;
    std::fill(S.h_wave_inst.begin(), S.h_wave_inst.end(), -1);
    int first = 0, afirst = 0, lfirst = 0;
    for (const InstSpan& sp : spans) {
        const int m = (sp.n+LDS_ROWS-1)/LDS_ROWS;
        const int bw = m*AGG_SLOTS_PER_BLOCK/AGG_CAP;
        const int bg = (bw+aw-1)/aw;
        if (S.scap > 0 && S.precond == LFD_PRECOND_AGGREGATE && S.kind == LFD_SOLVER_MIXED && sp.n <= AGG_NMAX &&
            !all_large && bg <= amaxg) {
            const int soff = sp.row0/LDS_ROWS*AGG_SLOTS_PER_BLOCK;
            if ((int)ajobs.size()-afirst+bg > amaxg) {
                S.agg_launches.emplace_back(afirst, (int)ajobs.size()-afirst);
                afirst = (int)ajobs.size();
            }
            for (int g = 0; g < bg; g++) {
                SolveJob j{};
                j.inst = sp.inst; j.lwg = g; j.G = bg; j.row0 = soff; j.row_end = soff+AGG_CAP*bw;
                j.key = sp.row0/LDS_ROWS;
                ajobs.push_back(j);
            }
            AggJob a{};
            a.inst = sp.inst; a.poff = sp.row0; a.n = sp.n; a.soff = soff; a.budget = bw;
            gjobs.push_back(a);
            for (int v = 0; v < bw; v++) {
                S.h_wave_inst[soff/AGG_CAP+v] = sp.inst;
                S.h_wave_local[soff/AGG_CAP+v] = v;
            }
            continue;
        }
        if (S.scap > 0 && S.precond == LFD_PRECOND_AGGREGATE && S.kind == LFD_SOLVER_MIXED && lmaxg > 0 &&
            sp.inst < (int)S.h_chunk.size()) {
            // a large instance (or any, with agg_waves 16): grouped in spatial
            // chunks of at most AGG_NMAX particles (cell order), each packed
            // into its own budget of waves; the instance's waves are about 4%
            // more than its particles / 8 (packing with whole groups), at most
            // what one launch holds.  Instances share a launch while it fits.
            const int nch = (sp.n+AGG_NMAX-1)/AGG_NMAX, cs = (sp.n+nch-1)/nch;
            std::vector<int> bud(nch);
            int need = 0;
            for (int c = 0; c < nch; c++) {
                bud[c] = (std::min(cs, sp.n-c*cs)+AGG_CAP-1)/AGG_CAP;
                need += bud[c];
            }
            const int want = (int)std::ceil(sp.n/(AGG_CAP*0.96));
            const int W = std::min({AGG_LARGE_WAVES*lmaxg, std::max(want, need), bw});
            if (W >= need) {
                for (int c = 0, spare = W-need; spare > 0; c = (c+1)%nch, spare--) bud[c]++;
                const int soff = sp.row0/LDS_ROWS*AGG_SLOTS_PER_BLOCK;
                const int Gl = (W+AGG_LARGE_WAVES-1)/AGG_LARGE_WAVES;
                if ((int)ljobs.size()-lfirst+Gl > lmaxg) {
                    S.large_launches.emplace_back(lfirst, (int)ljobs.size()-lfirst);
                    lfirst = (int)ljobs.size();
                }
                for (int g = 0; g < Gl; g++) {
                    SolveJob j{};
                    j.inst = sp.inst; j.lwg = g; j.G = Gl; j.row0 = soff; j.row_end = soff+AGG_CAP*W;
                    j.key = sp.row0/LDS_ROWS;
                    ljobs.push_back(j);
                }
                for (int c = 0, so = soff; c < nch; so += AGG_CAP*bud[c], c++) {
                    AggJob a{};
                    a.inst = sp.inst; a.poff = sp.row0+c*cs; a.n = std::min(cs, sp.n-c*cs); a.soff = so;
                    a.budget = bud[c]; a.chunked = 1; a.total = W;
                    gjobs.push_back(a);
                }
                for (int v = 0; v < W; v++) {
                    S.h_wave_inst[soff/AGG_CAP+v] = sp.inst;
                    S.h_wave_local[soff/AGG_CAP+v] = v;
                }
                S.h_chunk[sp.inst] = cs;
                continue;
            }
        }
        const int G = (sp.n+LDS_ROWS-1)/LDS_ROWS;
        if (G > maxg) {
            if (spans.size() == 1) return LFD_OK;
            return fail(LFD_ERR_CAPACITY, "an instance of this size only runs alone in its batch");
        }
        if ((int)jobs.size()-first+G > maxg) {
            S.launches.emplace_back(first, (int)jobs.size()-first);
            first = (int)jobs.size();
        }
        for (int g = 0; g < G; g++) {
            SolveJob j{};
            j.inst = sp.inst; j.lwg = g; j.G = G; j.row0 = sp.row0; j.row_end = sp.row0+sp.n;
            j.key = sp.row0/LDS_ROWS;
            jobs.push_back(j);
        }
    }
    if ((int)jobs.size() > first) S.launches.emplace_back(first, (int)jobs.size()-first);
    if ((int)ajobs.size() > afirst) S.agg_launches.emplace_back(afirst, (int)ajobs.size()-afirst);
    if ((int)ljobs.size() > lfirst) S.large_launches.emplace_back(lfirst, (int)ljobs.size()-lfirst);
    // the large instances' jobs follow the aggregate solver's
    for (auto& L : S.large_launches) L.first += (int)ajobs.size();
    ajobs.insert(ajobs.end(), ljobs.begin(), ljobs.end());
    if ((int)jobs.size() > S.jobs_cap || (int)ajobs.size() > S.ajobs_cap)
        return fail(LFD_ERR_CAPACITY, "solver job table capacity exceeded");
    // An instance's barrier words start at its first row block and may run
    // into the next blocks (one leaf counter per 8 workgroups): a layout of
    // the instances other than the last one's leaves other instances' words,
    // generation counters among them, where the new one keeps its counters.
    // All start again from zero.
    launch_fill_int(reinterpret_cast<int*>(S.w.barrier), 0, S.nbar, s);
    if (!jobs.empty())
        HIPCHK(hipMemcpyAsync(S.jobs, jobs.data(), jobs.size()*sizeof(SolveJob), hipMemcpyHostToDevice, s));
    if (!ajobs.empty()) {
        HIPCHK(hipMemcpyAsync(S.ajobs, ajobs.data(), ajobs.size()*sizeof(SolveJob), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(S.agg_jobs, gjobs.data(), gjobs.size()*sizeof(AggJob), hipMemcpyHostToDevice, s));
    }
    if (S.scap > 0) {
        HIPCHK(hipMemcpyAsync(S.agg.chunk, S.h_chunk.data(), S.h_chunk.size()*sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(S.wave_inst, S.h_wave_inst.data(), S.h_wave_inst.size()*sizeof(int),
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(S.wave_local, S.h_wave_local.data(), S.h_wave_local.size()*sizeof(int),
                              hipMemcpyHostToDevice, s));
    }
    S.n_agg_jobs = (int)gjobs.size();
    HIPCHK(hipStreamSynchronize(s));
    return LFD_OK;
}

static int status_of(int err)
{
    if (err & E_BOX_OOB) return fail(LFD_ERR_BOX, "BoxSet: trying to box position out of boundaries");
    if (err & (E_FOUND_OVERFLOW | E_PAIR_OVERFLOW)) return fail(LFD_ERR_CAPACITY, "pair capacity exceeded");
    if (err & E_NOT_SPD) return fail(LFD_ERR_NOT_SPD, "resistance diagonal block not positive definite");
    if (err & E_SYNC_TIMEOUT) return fail(LFD_ERR_HIP, "solver workgroups lost synchronisation (spin limit)");
    if (err & E_NOT_CONVERGED) return fail(LFD_ERR_NOT_CONVERGED, "solver did not reach the tolerance");
    return LFD_OK;
}

static int check_scalars(DevScalars* sc_dev, hipStream_t s, DevScalars* out = nullptr)
{
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, sc_dev, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (out) *out = h;
    return status_of(h.error);
}

// One solve of every instance: the batched LDS launches, or for a lone
// instance the plain form or the global-memory form.
// nrhs right-hand sides at b + k*rs, solutions at x + k*rs; the LDS solver
// takes them two per launch (unless opts.max_rhs is 1), the other forms one
// at a time
static int run_solve(Solver& S, const int* npairs_dev, const double* b, double* x, DevScalars* sc,
                     hipStream_t s, int nrhs = 1, size_t rs = 0)
{
    BsrView A = S.view(npairs_dev);
    // two right-hand sides per launch: the block-Jacobi and large-instance solvers
    const bool pairs = S.kind == LFD_SOLVER_MIXED && S.agg_launches.empty() && has_persistent(S);
    if (nrhs > 1 && !(pairs && S.opts.max_rhs > 1)) {
        for (int k = 0; k < nrhs; k++) {
            int rc = run_solve(S, npairs_dev, b+k*rs, x+k*rs, sc, s);
            if (rc) return rc;
        }
        return LFD_OK;
    }
    const bool plain = S.kind == LFD_SOLVER_PLAIN || S.kind == LFD_SOLVER_PLAIN_DEVICE;
    if (plain || S.kind == LFD_SOLVER_MIXED_GLOBAL || !has_persistent(S)) {
        if (S.single < 0) return fail(LFD_ERR_UNSUPPORTED, "this solver form takes a single instance");
        sc += S.single;
    }
    if (plain) {
        int it = 0;
        double res = 0;
        if (S.kind == LFD_SOLVER_PLAIN) HIPCHK(solve_plain(A, b, x, S.w, sc, S.opts, s, &it, &res));
        else HIPCHK(solve_plain_dev(A, b, x, S.w, sc, S.opts, s, &it, &res));
        DevScalars h;
        HIPCHK(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        h.iters_inner = it; h.iters_outer = 1; h.res_norm = res; h.rhs_norm = 1; h.total_inner += it;
        if (!(res <= S.opts.tol)) h.error |= E_NOT_CONVERGED;
        HIPCHK(hipMemcpyAsync(sc, &h, sizeof(h), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return LFD_OK;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (S.profiling) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, s));
    }
    if (S.kind == LFD_SOLVER_MIXED && has_persistent(S)) {
        for (auto& L : S.agg_launches)
            HIPCHK(solve_agg(A, S.agg, b, x, S.w, sc, S.ajobs+L.first, L.second, S.opts, s));
        for (auto& L : S.large_launches)
            for (int k = 0; k < nrhs; k += 2) {
                const int nk = std::min(2, nrhs-k);
                HIPCHK(solve_agg_large(A, S.agg, b+k*rs, x+k*rs, nk, rs, S.w, sc, S.ajobs+L.first, L.second,
                                       S.opts, s));
            }
        for (auto& L : S.launches)
            for (int k = 0; k < nrhs; k += 2) {
                const int nk = std::min(2, nrhs-k);
                HIPCHK(solve_lds(A, b+k*rs, x+k*rs, nk, rs, S.w, sc, S.jobs+L.first, L.second, S.opts, s));
            }
    } else {
        HIPCHK(solve_mixed(A, b, x, S.w, sc, S.opts, s));
    }
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

// ===================================================================== batch of instances
// Every array is indexed by a global particle slot, cell or pair; instance k
// owns the slots [poff, poff + round_up(n, 128)) and the cells
// [coff, coff + nbox).  Pairs never cross instances, so the one (p0, p1)-sorted
// pair list is the concatenation of the instances' own sorted lists and every
// per-instance result is independent of the other instances.
struct InstHost {
    bool live = false;
    int slots = 0;
    int have_stress = 0;
    InstConst ic{};
};

struct lfd_batch {
    int M = 0, ccap = 0, icap = 0, pcap = 0, K = 64;
    int ninst = 0;                // instance handles in use (high-water mark)
    int control = LFD_RATE_CONTROL;
    hipStream_t s = nullptr;
    double4* pos = nullptr;
    double *vel6 = nullptr, *na6 = nullptr, *pred6 = nullptr, *rhs = nullptr;
    double* na_disp = nullptr;    // [3 per slot] accumulated non-affine displacement
    int *inst_of = nullptr, *box_of = nullptr;
    int *snap_start = nullptr, *snap_mem = nullptr, *home_start = nullptr, *home_mem = nullptr;
    int *bin_count = nullptr, *bin_cursor = nullptr, *found = nullptr, *found_cnt = nullptr;
    PairRec* pairs[2] = {nullptr, nullptr};
    int cur = 0;
    int *keep = nullptr, *keep_scan = nullptr, *tcnt = nullptr, *merge_cnt = nullptr, *scan_tmp = nullptr;
    int* probe = nullptr;
    double *pdb = nullptr, *prhs = nullptr;
    // velocity parts (contact, dashpot, hydro, repulsion / rate independent,
    // rate proportional) with their right-hand sides, and the stress buffers
    double *rhs_parts = nullptr, *vel_parts = nullptr, *pst = nullptr, *ps = nullptr, *st_partial = nullptr;
    CellArrays ca{};
    InstConst* d_ic = nullptr;
    DevScalars* sc = nullptr;
    BatchScalars* bs = nullptr;
    Solver S;
    std::vector<InstHost> inst;
    // captured single steps, one per parity of the pair buffers at the start
    // of the step, with the key they were captured under and the parity flip
    // a step makes
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    long long gkey[2] = {-1, -1};
    int gflip[2] = {0, 0};
    ~lfd_batch();
};

// The single-instance System is a batch holding instance 0.
struct lfd_system : lfd_batch {};

lfd_batch::~lfd_batch()
{
    if (s) (void)hipStreamSynchronize(s);
    void* ptrs[] = {pos, vel6, na6, pred6, rhs, na_disp, inst_of, box_of, snap_start, snap_mem, home_start, home_mem,
                    bin_count, bin_cursor, found, found_cnt, pairs[0], pairs[1], keep, keep_scan, tcnt, merge_cnt,
                    scan_tmp, probe, pdb, prhs, rhs_parts, vel_parts, pst, ps, st_partial, ca.nb, ca.nnb,
                    ca.nstatic, ca.ctype, ca.cell_inst, d_ic, sc, bs};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    for (auto& g : gexec) if (g) (void)hipGraphExecDestroy(g);
    S.release();
    if (s) (void)hipStreamDestroy(s);
}

static int round_up(int a, int m) { return (a+m-1)/m*m; }

// Device bytes per particle slot and per cell of a batch (pair storage is
// sized per slot), used for admission by HBM footprint.
static void batch_unit_bytes(size_t* per_slot, size_t* per_cell)
{
    const size_t pp = 16;     // pair capacity per slot
    size_t slot = sizeof(double4)+4*6*8+3*8+2*4+2*4+(size_t)64*4+4*4  // pos, vel6/na6/pred6/rhs, na_disp, labels, found
                  +8*6*8+STRESS_VALS*8                               // velocity parts, rhs parts, particle stress
                  +(15+36)*(8+4)+6*8*2+6*4*6+6*8*4+3*4;              // solver rows and vectors
    slot += pp*(2*sizeof(PairRec)+2*4+(30+42)*8+2*STRESS_VALS*8+18*(8+4)+3*4);
    if (per_slot) *per_slot = slot;
    if (per_cell) *per_cell = (size_t)(64+32+8)*4;
}

static int batch_init(lfd_batch* B, int slot_cap, int cell_cap, int inst_cap, const lfd_solve_opts* o, int control)
{
    lfd_solve_opts od;
    lfd_default_solve_opts(&od);
    if (!o) o = &od;
    B->S.opts = to_opts(o);
    B->S.kind = o->kind;
    B->S.precond = o->precond;
    B->S.regroup_interval = std::max(1, o->regroup_interval);
    B->control = control;
    B->M = round_up(std::max(slot_cap, 1), LDS_ROWS);
    B->ccap = std::max(cell_cap, 1);
    B->icap = std::max(inst_cap, 1);
    B->pcap = 16*B->M+1024;
    B->inst.assign(B->icap, InstHost());
    const int M = B->M, ccap = B->ccap, pcap = B->pcap;
    const size_t m6 = 6*(size_t)M;
    hipError_t e = hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking);
#define A(x) if (e == hipSuccess) e = (x)
    A(dalloc(&B->pos, M)); A(dalloc(&B->vel6, m6)); A(dalloc(&B->na6, m6)); A(dalloc(&B->pred6, m6));
    A(dalloc(&B->rhs, m6)); A(dalloc(&B->na_disp, 3*(size_t)M)); A(dalloc(&B->inst_of, M)); A(dalloc(&B->box_of, M));
    A(dalloc(&B->snap_start, ccap+1)); A(dalloc(&B->snap_mem, M)); A(dalloc(&B->home_start, ccap+1));
    A(dalloc(&B->home_mem, M)); A(dalloc(&B->bin_count, ccap)); A(dalloc(&B->bin_cursor, ccap));
    A(dalloc(&B->found, (size_t)M*B->K)); A(dalloc(&B->found_cnt, M));
    A(dalloc(&B->pairs[0], pcap)); A(dalloc(&B->pairs[1], pcap));
    A(dalloc(&B->keep, pcap)); A(dalloc(&B->keep_scan, (size_t)std::max(pcap, M)+1)); A(dalloc(&B->tcnt, M));
    A(dalloc(&B->merge_cnt, M)); A(dalloc(&B->scan_tmp, pcap/2048+64)); A(dalloc(&B->probe, 32*(size_t)ccap));
    A(dalloc(&B->pdb, 30*(size_t)pcap)); A(dalloc(&B->prhs, 42*(size_t)pcap));
    A(dalloc(&B->rhs_parts, 4*m6)); A(dalloc(&B->vel_parts, 4*m6));
    A(dalloc(&B->pst, 2*(size_t)STRESS_VALS*pcap)); A(dalloc(&B->ps, (size_t)STRESS_VALS*M));
    A(dalloc(&B->st_partial, (size_t)STRESS_VALS*(M/STRESS_CHUNK+1)));
    A(dalloc(&B->ca.nb, 64*(size_t)ccap)); A(dalloc(&B->ca.nnb, ccap)); A(dalloc(&B->ca.nstatic, ccap));
    A(dalloc(&B->ca.ctype, ccap)); A(dalloc(&B->ca.cell_inst, ccap));
    A(dalloc(&B->d_ic, B->icap)); A(dalloc(&B->sc, B->icap)); A(dalloc(&B->bs, 1));
    B->S.allow_large = true;
    A(B->S.alloc(M, pcap, B->icap));
#undef A
    if (e != hipSuccess) return fail(LFD_ERR_HIP, std::string("batch allocation: ")+hipGetErrorString(e));
    B->ca.ncell = ccap;
    hipStream_t s = B->s;
    launch_fill_int(B->inst_of, -1, M, s);
    launch_fill_int(B->box_of, -1, M, s);
    launch_fill_int(B->ca.cell_inst, -1, ccap, s);
    launch_fill_int(B->ca.nnb, 0, ccap, s);
    launch_fill_int(B->ca.nstatic, 0, ccap, s);
    launch_fill_int(B->ca.ctype, 0, ccap, s);
    launch_fill_int(B->S.row_ptr, 0, M+1, s);
    launch_fill_int(B->S.tptr, 0, M+1, s);
    launch_fill_int((int*)B->bs, 0, sizeof(BatchScalars)/sizeof(int), s);
    launch_fill_int((int*)B->d_ic, 0, B->icap*sizeof(InstConst)/sizeof(int), s);
    launch_fill_int((int*)B->sc, 0, B->icap*sizeof(DevScalars)/sizeof(int), s);
    launch_fill_double(B->vel6, 0.0, m6, s);
    launch_fill_double(B->na6, 0.0, m6, s);
    launch_fill_double(B->na_disp, 0.0, 3*(size_t)M, s);
    launch_fill_double(B->pred6, 0.0, m6, s);
    launch_fill_double(B->vel_parts, 0.0, 4*m6, s);
    HIPCHK(hipStreamSynchronize(s));
    return LFD_OK;
}

// BoxSet::init (BoxSet.cpp:7-70, 170-306): grid, probes and static neighbours
// of one instance (local cell labels)
static void build_grid_host(InstConst& g, double max_range, const double sd[3], std::vector<int>& nb,
                            std::vector<int>& nnb, std::vector<int>& ctype)
{
    const double* L = g.C.L;
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

// Constants, cell grid and static neighbour lists (local cell labels, at the
// initial shear displacement sd) of a prospective instance.
static int describe_instance(int np, const double* radius, const double L[3], const lfd_params* p,
                             const double sd[3], InstConst& ic, std::vector<int>& nb, std::vector<int>& nnb,
                             std::vector<int>& ctype)
{
    if (np < 2 || !radius || !L || !p) return fail(LFD_ERR_ARG, "instance: bad argument");
    for (int k = 0; k < 3; k++)
        if (!(L[k] > 0)) return fail(LFD_ERR_ARG, "instance: box lengths must be positive");
    std::memset(&ic, 0, sizeof(ic));
    ic.P = to_params(p);
    ic.C = make_consts(ic.P, L);
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
            double r = lub_range(uniq[a], uniq[b], ic.P.lub_max_gap);
            if (r > max_range) max_range = r;
        }
    build_grid_host(ic, max_range, sd, nb, nnb, ctype);
    ic.n = np;
    return LFD_OK;
}

// First fit of `len` units among the live ranges (start, length) in [0, cap).
static int first_fit(std::vector<std::pair<int, int>> used, int cap, int len)
{
    std::sort(used.begin(), used.end());
    int at = 0;
    for (auto& u : used) {
        if (u.first-at >= len) return at;
        at = std::max(at, u.first+u.second);
    }
    return cap-at >= len ? at : -1;
}

static std::vector<InstSpan> live_spans(const lfd_batch* B)
{
    std::vector<InstSpan> v;
    for (int k = 0; k < B->ninst; k++)
        if (B->inst[k].live) v.push_back({k, B->inst[k].ic.poff, B->inst[k].ic.n});
    std::sort(v.begin(), v.end(), [](const InstSpan& a, const InstSpan& b) { return a.row0 < b.row0; });
    return v;
}

static int batch_check(lfd_batch* B, int id, DevScalars* out = nullptr)
{
    BatchScalars hb;
    HIPCHK(hipMemcpyAsync(&hb, B->bs, sizeof(hb), hipMemcpyDeviceToHost, B->s));
    std::vector<DevScalars> h(std::max(B->ninst, 1));
    if (B->ninst) HIPCHK(hipMemcpyAsync(h.data(), B->sc, B->ninst*sizeof(DevScalars), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    if (out && id >= 0 && id < B->ninst) *out = h[id];
    int rc = status_of(hb.error);
    if (rc) return rc;
    for (int k = 0; k < B->ninst; k++) {
        if (!B->inst[k].live || (id >= 0 && k != id)) continue;
        if ((rc = status_of(h[k].error))) return rc;
    }
    return LFD_OK;
}

static void bat_bin_snapshot(lfd_batch* B)
{
    // BoxSet::update (BoxSet.cpp:498-506): moving neighbours, then the
    // neighbourhood snapshot (cell membership at this instant)
    launch_moving_neighbors(B->ca, B->d_ic, B->sc, B->probe, B->s);
    launch_bin(B->box_of, B->M, B->ccap, B->bin_count, B->snap_start, B->snap_mem, B->bin_cursor, B->scan_tmp, B->s);
}

static void bat_check_new(lfd_batch* B, int only = -1)
{
    // System::checkNewInteraction (System.cpp:1328-1353)
    launch_bin(B->box_of, B->M, B->ccap, B->bin_count, B->home_start, B->home_mem, B->bin_cursor, B->scan_tmp, B->s);
    launch_find_pairs(B->pos, B->M, B->ca, B->d_ic, B->home_start, B->home_mem, B->snap_start, B->snap_mem, B->sc,
                      B->found, B->found_cnt, B->K, only, B->s);
    PairRec* oldp = B->pairs[B->cur];
    PairRec* newp = B->pairs[B->cur^1];
    launch_merge_pairs(B->M, oldp, B->S.p1, B->S.row_ptr, B->found, B->found_cnt, B->K, B->merge_cnt, B->keep_scan,
                       B->keep, newp, B->pos, B->inst_of, B->d_ic, B->sc, B->bs, B->pcap, B->scan_tmp, B->s);
    B->cur ^= 1;
    launch_pair_index(newp, B->M, B->S.row_ptr, B->S.p0, B->S.p1, B->S.tptr, B->S.tidx, B->tcnt, B->keep, B->bs,
                      B->pcap, B->scan_tmp, B->s);
}

static void bat_update(lfd_batch* B, int in_predictor, double dt_override, int only = -1)
{
    // System::updateInteractions (System.cpp:1375-1403)
    PairRec* cp = B->pairs[B->cur];
    PairRec* np_ = B->pairs[B->cur^1];
    launch_update_pairs(cp, B->pos, B->vel6, B->inst_of, B->d_ic, B->sc, B->ninst, B->bs, B->keep, B->pcap,
                        in_predictor, dt_override, only, B->s);
    launch_compact_pairs(cp, np_, B->keep, B->keep_scan, B->M, B->S.row_ptr, B->S.p0, B->S.p1, B->S.tptr,
                         B->S.tidx, B->tcnt, B->bs, B->pcap, B->scan_tmp, B->s);
    B->cur ^= 1;
}

static void bat_assemble(lfd_batch* B, int nparts = 0)
{
    launch_assemble(B->pairs[B->cur], B->pos, B->M, B->inst_of, B->d_ic, B->sc, B->bs, B->S.row_ptr, B->S.tptr,
                    B->S.tidx, B->S.diag64, B->S.diag32, B->S.od64, B->S.od32, B->S.minv64, B->S.minv32, B->rhs,
                    B->pdb, B->prhs, B->pcap, B->rhs_parts, nparts, B->s);
    // aggregate preconditioner: the groups of the instances that are due, then
    // the wave inverses of this matrix
    if (B->S.n_agg_jobs > 0) {
        const BsrView A = B->S.view(&B->bs->npairs);
        const bool chunked = std::any_of(B->S.h_chunk.begin(), B->S.h_chunk.end(), [](int c) { return c > 0; });
        launch_agg_group(A, B->S.agg, B->S.agg_jobs, B->S.n_agg_jobs, B->inst_of, B->d_ic, B->sc,
                         B->S.regroup_interval, B->M, B->ninst, chunked ? B->home_start : nullptr,
                         chunked ? B->home_mem : nullptr, B->ccap, B->s);
        launch_agg_inverse(A, B->S.agg, B->S.scap/AGG_CAP, B->s);
    }
}

// calcStressPerParticle + calcStress (mode 1) or computeShearRate (mode 0)
static void bat_stress(lfd_batch* B, int nparts, int mode)
{
    launch_stress(B->pairs[B->cur], B->pos, B->inst_of, B->d_ic, B->ninst, B->sc, B->bs, B->vel_parts, nparts,
                  B->vel6, B->M, B->control, mode, B->S.row_ptr, B->S.tptr, B->S.tidx, B->pst, B->ps,
                  B->st_partial, B->pcap, B->s);
}

static int bat_velocities(lfd_batch* B, int in_predictor, int calc_stress = 0)
{
    // System::computeVelocities (System.cpp:2273-2319)
    if (B->control == LFD_RATE_CONTROL && !calc_stress) {
        bat_assemble(B);
        int rc = run_solve(B->S, &B->bs->npairs, B->rhs, B->na6, B->sc, B->s);
        if (rc) return rc;
        launch_finish_velocity(B->M, B->pos, B->na6, B->vel6, B->inst_of, B->sc, B->ninst, in_predictor, B->s);
        return LFD_OK;
    }
    // velocities by parts (computeVelocityByComponents, System.cpp:1947-1965):
    // four components when the stress is wanted, else (stress control) the
    // rate independent and rate proportional sums
    const int nparts = calc_stress ? 4 : 2;
    const size_t m6 = 6*(size_t)B->M;
    if (B->control == LFD_STRESS_CONTROL) launch_set_rate(B->sc, B->d_ic, B->ninst, 1.0, B->s);
    bat_assemble(B, nparts);
    int rc = run_solve(B->S, &B->bs->npairs, B->rhs_parts, B->vel_parts, B->sc, B->s, nparts, m6);
    if (rc) return rc;
    if (B->control == LFD_STRESS_CONTROL) bat_stress(B, nparts, 0);     // new shear rate
    launch_sum_parts(B->vel_parts, nparts, B->na6, B->M, B->inst_of, B->sc, B->control == LFD_STRESS_CONTROL, B->s);
    launch_finish_velocity(B->M, B->pos, B->na6, B->vel6, B->inst_of, B->sc, B->ninst, in_predictor, B->s);
    if (calc_stress) {
        bat_stress(B, nparts, 1);
        for (auto& h : B->inst) if (h.live) h.have_stress = 1;
    }
    return LFD_OK;
}

static int bat_step(lfd_batch* B, int calc_stress = 0, int adapt_first = 1)
{
    // System::timeEvolution loop body + timeEvolutionPredictorCorrectorMethod
    // (System.cpp:1241-1256, 888-983, 1124-1187); the stress, when wanted, is
    // the corrector's
    int rc;
    if (adapt_first) launch_adapt_dt(B->sc, B->d_ic, B->ninst, 0, B->s);
    launch_agg_age(B->sc, B->ninst, B->s);
    if ((rc = bat_velocities(B, 1))) return rc;
    launch_adapt_dt(B->sc, B->d_ic, B->ninst, 1, B->s);
    launch_boxing(B->sc, B->d_ic, B->ninst, B->s);
    bat_bin_snapshot(B);
    launch_displace(B->pos, B->vel6, B->pred6, B->M, B->inst_of, B->d_ic, B->sc, B->box_of, 0, B->na6, B->na_disp,
                    B->s);
    bat_update(B, 1, -1.0);
    launch_copy_double(B->pred6, B->vel6, 6*(size_t)B->M, B->s);
    if ((rc = bat_velocities(B, 0, calc_stress))) return rc;
    launch_displace(B->pos, B->vel6, B->pred6, B->M, B->inst_of, B->d_ic, B->sc, B->box_of, 1, B->na6, B->na_disp,
                    B->s);
    bat_check_new(B);
    bat_update(B, 0, -1.0);
    return LFD_OK;
}

// ------------------------------------------------------------------ captured steps
// A step is a fixed sequence of kernel launches on the batch's stream with no
// host round trip, so it can be captured once into a HIP graph and replayed.
// The capture depends on the pair-buffer parity at the start of the step
// (host-side ping-pong), the solver's job table and options (Solver::gen) and
// the control; it is redone when any of them changes.  Not captured: the plain
// FP64 solver (host-side scalars), profiling (events), and solver launches
// that must be cooperative.
static long long graph_key(const lfd_batch* B)
{
    return 2*B->S.gen+(B->control == LFD_STRESS_CONTROL ? 1 : 0);
}

static bool capturable(const lfd_batch* B)
{
    const Solver& S = B->S;
    if (S.kind != LFD_SOLVER_MIXED || S.profiling || !has_persistent(S)) return false;
    auto plain = [&](int nwg, int cap) { return S.opts.launch == LFD_LAUNCH_PLAIN ||
                                                (S.opts.launch == LFD_LAUNCH_AUTO && nwg <= cap); };
    for (auto& L : S.agg_launches) if (!plain(L.second, agg_solver_max_groups(S.opts.agg_waves))) return false;
    for (auto& L : S.launches) if (!plain(L.second, lds_solver_max_groups())) return false;
    for (auto& L : S.large_launches) if (!plain(L.second, agg_large_max_groups())) return false;
    return true;
}

static int bat_capture(lfd_batch* B, int parity)
{
    const int cur0 = B->cur;
    B->cur = parity;
    HIPCHK(hipStreamBeginCapture(B->s, hipStreamCaptureModeThreadLocal));
    const int rc = bat_step(B);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->s, &g);
    const int flip = B->cur^parity;
    B->cur = cur0;
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : fail(LFD_ERR_HIP, "step capture failed");
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(LFD_ERR_HIP, "step graph instantiation failed");
    if (B->gexec[parity]) (void)hipGraphExecDestroy(B->gexec[parity]);
    B->gexec[parity] = x;
    B->gkey[parity] = graph_key(B);
    B->gflip[parity] = flip;
    return LFD_OK;
}

static int bat_step_graph(lfd_batch* B, int nsteps)
{
    if (!capturable(B)) {
        for (int k = 0; k < nsteps; k++) {
            const int rc = bat_step(B);
            if (rc) return rc;
        }
        return LFD_OK;
    }
    for (int k = 0; k < nsteps; k++) {
        const int p = B->cur;
        if (!B->gexec[p] || B->gkey[p] != graph_key(B)) {
            const int rc = bat_capture(B, p);
            if (rc) return rc;
        }
        HIPCHK(hipGraphLaunch(B->gexec[p], B->s));
        B->cur ^= B->gflip[p];
    }
    return LFD_OK;
}

// Admission of one instance: slots, cells and a handle are taken first fit;
// the particles are binned and the instance's pairs found (System::setupSystem,
// System.cpp:468-611), touching no other instance's pairs.  first_update also
// runs the first-call block of System::timeEvolution (System.cpp:1219-1228).
static int bat_add(lfd_batch* B, int np, const double* pos, const double* radius, const double L[3],
                   const double shear_disp[3], const lfd_params* p, double strain_end, int first_update, int* id_out)
{
    if (!pos) return fail(LFD_ERR_ARG, "instance: bad argument");
    double sd[3] = {0, 0, 0};
    if (shear_disp) for (int k = 0; k < 3; k++) sd[k] = shear_disp[k];
    InstConst ic;
    std::vector<int> nb, nnb, ctype;
    int rc = describe_instance(np, radius, L, p, sd, ic, nb, nnb, ctype);
    if (rc) return rc;
    const int slots = round_up(np, LDS_ROWS);
    std::vector<std::pair<int, int>> us, uc;
    int id = -1;
    for (int k = 0; k < B->icap; k++) {
        if (B->inst[k].live) {
            us.emplace_back(B->inst[k].ic.poff, B->inst[k].slots);
            uc.emplace_back(B->inst[k].ic.coff, B->inst[k].ic.nbox);
        } else if (id < 0) id = k;
    }
    const int poff = first_fit(us, B->M, slots);
    const int coff = first_fit(uc, B->ccap, ic.nbox);
    if (id < 0 || poff < 0 || coff < 0) return fail(LFD_ERR_CAPACITY, "batch: no room for the instance");
    ic.poff = poff;
    ic.coff = coff;
    // the solver must be able to take the new set of instances
    std::vector<InstSpan> spans = live_spans(B);
    spans.push_back({id, poff, np});
    std::sort(spans.begin(), spans.end(), [](const InstSpan& a, const InstSpan& b) { return a.row0 < b.row0; });
    if ((rc = set_jobs(B->S, spans, B->s))) {
        set_jobs(B->S, live_spans(B), B->s);
        return rc;
    }
    hipStream_t s = B->s;
    std::vector<double4> h4(np);
    for (int i = 0; i < np; i++) h4[i] = make_double4(pos[3*i], pos[3*i+1], pos[3*i+2], radius[i]);
    for (auto& v : nb) v += coff;
    DevScalars hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.dt = p->fixed_dt ? p->dt_max : 1e-4;     // dt = p.dt (System.cpp:607)
    for (int k = 0; k < 3; k++) hs.shear_disp[k] = sd[k];
    make_flow(ic.C, ic.P.shear_rate, hs.flow);
    hs.group_age = -1;
    hs.init_strain_limit = 0;       // System.cpp:90-91
    hs.init_rate_limit = 999;
    hs.strain_end = strain_end;
    const size_t o6 = 6*(size_t)poff, s6 = 6*(size_t)slots;
    HIPCHK(hipMemcpyAsync(B->pos+poff, h4.data(), np*sizeof(double4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B->sc+id, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B->d_ic+id, &ic, sizeof(ic), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B->ca.nb+64*(size_t)coff, nb.data(), nb.size()*sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B->ca.nnb+coff, nnb.data(), nnb.size()*sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B->ca.nstatic+coff, nnb.data(), nnb.size()*sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B->ca.ctype+coff, ctype.data(), ctype.size()*sizeof(int), hipMemcpyHostToDevice, s));
    launch_fill_int(B->ca.cell_inst+coff, id, ic.nbox, s);
    launch_fill_int(B->inst_of+poff, id, np, s);
    launch_fill_double(B->vel6+o6, 0.0, s6, s);
    launch_fill_double(B->na6+o6, 0.0, s6, s);
    launch_fill_double(B->na_disp+3*(size_t)poff, 0.0, 3*(size_t)np, s);
    launch_fill_double(B->pred6+o6, 0.0, s6, s);
    for (int q = 0; q < 4; q++) launch_fill_double(B->vel_parts+q*6*(size_t)B->M+o6, 0.0, s6, s);
    InstHost& h = B->inst[id];
    h.live = true;
    h.slots = slots;
    h.have_stress = 0;
    h.ic = ic;
    B->ninst = std::max(B->ninst, id+1);
    launch_box_labels(B->pos+poff, B->inst_of+poff, np, B->d_ic, B->box_of+poff, B->sc, s);
    bat_bin_snapshot(B);
    bat_check_new(B, id);
    if (first_update) {
        bat_check_new(B, id);
        bat_update(B, 1, 0.0, id);
    }
    if ((rc = batch_check(B, id))) return rc;
    if (id_out) *id_out = id;
    return LFD_OK;
}

// Retirement: the instance's slots and cells are freed and its pairs dropped;
// the other instances' pairs keep their order.
static int bat_remove(lfd_batch* B, int id)
{
    if (id < 0 || id >= B->ninst || !B->inst[id].live) return fail(LFD_ERR_ARG, "batch: no such instance");
    InstHost& h = B->inst[id];
    hipStream_t s = B->s;
    launch_fill_int(B->inst_of+h.ic.poff, -1, h.slots, s);
    launch_fill_int(B->box_of+h.ic.poff, -1, h.slots, s);
    launch_fill_int(B->ca.cell_inst+h.ic.coff, -1, h.ic.nbox, s);
    launch_fill_int(B->ca.nnb+h.ic.coff, 0, h.ic.nbox, s);
    launch_fill_int((int*)(B->d_ic+id), 0, sizeof(InstConst)/sizeof(int), s);
    bat_update(B, 0, 0.0, -2);
    h = InstHost();
    while (B->ninst > 0 && !B->inst[B->ninst-1].live) B->ninst--;
    int rc = set_jobs(B->S, live_spans(B), s);
    if (rc) return rc;
    BatchScalars hb;
    HIPCHK(hipMemcpyAsync(&hb, B->bs, sizeof(hb), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return status_of(hb.error);
}

// ---------------------------------------------------------------- per-instance views
static const InstHost* live_inst(lfd_batch* B, int id)
{
    if (!B || id < 0 || id >= B->ninst || !B->inst[id].live) {
        fail(LFD_ERR_ARG, "no such instance");
        return nullptr;
    }
    return &B->inst[id];
}

// pair range [k0, k1) of an instance in the global sorted list
static int pair_range(lfd_batch* B, const InstHost& h, int& k0, int& k1)
{
    HIPCHK(hipMemcpyAsync(&k0, B->S.row_ptr+h.ic.poff, sizeof(int), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipMemcpyAsync(&k1, B->S.row_ptr+h.ic.poff+h.ic.n, sizeof(int), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    return LFD_OK;
}

static int inst_num_pairs(lfd_batch* B, int id)
{
    const InstHost* h = live_inst(B, id);
    if (!h) return LFD_ERR_ARG;
    int rc = batch_check(B, id);
    if (rc) return rc;
    int k0 = 0, k1 = 0;
    if ((rc = pair_range(B, *h, k0, k1))) return rc;
    return k1-k0;
}

static int inst_get_pairs(lfd_batch* B, int id, int cap, int* p0, int* p1, int* state, double* geom5,
                          double* disp_tan)
{
    const InstHost* h = live_inst(B, id);
    if (!h) return LFD_ERR_ARG;
    int rc = batch_check(B, id);
    if (rc) return rc;
    int k0 = 0, k1 = 0;
    if ((rc = pair_range(B, *h, k0, k1))) return rc;
    const int np_ = k1-k0, off = h->ic.poff;
    if (np_ > cap) return fail(LFD_ERR_ARG, "get_pairs: cap too small");
    std::vector<PairRec> v(np_);
    if (np_) HIPCHK(hipMemcpy(v.data(), B->pairs[B->cur]+k0, np_*sizeof(PairRec), hipMemcpyDeviceToHost));
    for (int k = 0; k < np_; k++) {
        const PairRec& q = v[k];
        if (p0) p0[k] = q.p0-off;
        if (p1) p1[k] = q.p1-off;
        if (state) state[k] = pair_state(q.flags);
        if (geom5) {
            geom5[5*k] = q.n[0]; geom5[5*k+1] = q.n[1]; geom5[5*k+2] = q.n[2];
            geom5[5*k+3] = q.r; geom5[5*k+4] = q.gap;
        }
        if (disp_tan) for (int c = 0; c < 3; c++) disp_tan[3*k+c] = q.disp_tan[c];
    }
    return np_;
}

// System::getContacts (System.cpp:306-318): the state of every active
// contact, in pair order.  With p0 null only the count is returned.
static int inst_get_contacts(lfd_batch* B, int id, int cap, int* p0, int* p1, double* disp_tan,
                             double* disp_rolling)
{
    const InstHost* h = live_inst(B, id);
    if (!h) return LFD_ERR_ARG;
    int rc = batch_check(B, id);
    if (rc) return rc;
    int k0 = 0, k1 = 0;
    if ((rc = pair_range(B, *h, k0, k1))) return rc;
    const int off = h->ic.poff;
    std::vector<PairRec> v(k1-k0);
    if (k1 > k0) HIPCHK(hipMemcpy(v.data(), B->pairs[B->cur]+k0, (k1-k0)*sizeof(PairRec), hipMemcpyDeviceToHost));
    int m = 0;
    for (const PairRec& q : v) {
        if (!(q.flags & F_CONTACT)) continue;
        if (p0) {
            if (m >= cap) return fail(LFD_ERR_ARG, "get_contacts: cap too small");
            p0[m] = q.p0-off;
            p1[m] = q.p1-off;
            for (int c = 0; c < 3; c++) {
                if (disp_tan) disp_tan[3*m+c] = q.disp_tan[c];
                if (disp_rolling) disp_rolling[3*m+c] = 0.0;     // rolling friction is not modelled
            }
        }
        m++;
    }
    return m;
}

// System::setContacts (System.cpp:285-304): the tangential displacement of
// each listed contact is set on its interaction, which must exist (pairs are
// found from the positions first, as setupConfiguration does before it).  The
// pairs are then updated with dt = 0, so that the contact forces follow the
// restored displacements before the first velocity computation.
static int inst_set_contacts(lfd_batch* B, int id, int n, const int* p0, const int* p1, const double* disp_tan,
                             const double* disp_rolling)
{
    const InstHost* h = live_inst(B, id);
    if (!h || n < 0 || (n > 0 && (!p0 || !p1 || !disp_tan))) return fail(LFD_ERR_ARG, "set_contacts: bad argument");
    int rc = batch_check(B, id);
    if (rc) return rc;
    int k0 = 0, k1 = 0;
    if ((rc = pair_range(B, *h, k0, k1))) return rc;
    const int off = h->ic.poff;
    std::vector<PairRec> v(k1-k0);
    if (k1 > k0) HIPCHK(hipMemcpy(v.data(), B->pairs[B->cur]+k0, (k1-k0)*sizeof(PairRec), hipMemcpyDeviceToHost));
    for (int c = 0; c < n; c++) {
        if (disp_rolling)
            for (int d = 0; d < 3; d++)
                if (disp_rolling[3*c+d] != 0.0)
                    return fail(LFD_ERR_UNSUPPORTED, "set_contacts: rolling displacements (rolling friction) are not modelled");
        const int a = std::min(p0[c], p1[c])+off, b = std::max(p0[c], p1[c])+off;
        auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(a, b),
                                   [](const PairRec& q, const std::pair<int, int>& k) {
                                       return q.p0 < k.first || (q.p0 == k.first && q.p1 < k.second); });
        if (it == v.end() || it->p0 != a || it->p1 != b)
            return fail(LFD_ERR_ARG, "set_contacts: no interaction for a listed contact");
        for (int d = 0; d < 3; d++) it->disp_tan[d] = disp_tan[3*c+d];
    }
    if (k1 > k0) HIPCHK(hipMemcpy(B->pairs[B->cur]+k0, v.data(), (k1-k0)*sizeof(PairRec), hipMemcpyHostToDevice));
    bat_update(B, 1, 0.0, id);
    return batch_check(B, id);
}

static void fill_scalars8(const DevScalars& h, double* sc8)
{
    sc8[0] = h.shear_disp[0]; sc8[1] = h.shear_disp[1]; sc8[2] = h.shear_disp[2];
    sc8[3] = h.dt; sc8[4] = h.time; sc8[5] = h.strain;
    sc8[6] = std::sqrt(__builtin_bit_cast(double, h.max_vel_bits));
    sc8[7] = std::sqrt(__builtin_bit_cast(double, h.max_slide_bits));
}

// System::getNonAffineDisp (System.h:529-532): n x 3
static int inst_get_na_disp(lfd_batch* B, int id, double* out)
{
    const InstHost* h = live_inst(B, id);
    if (!h || !out) return fail(LFD_ERR_ARG, "get_na_disp: bad argument");
    int rc = batch_check(B, id);
    if (rc) return rc;
    HIPCHK(hipMemcpy(out, B->na_disp+3*(size_t)h->ic.poff, 3*(size_t)h->ic.n*sizeof(double),
                     hipMemcpyDeviceToHost));
    return LFD_OK;
}

// The force columns of Simulation::outputIntFileTxt per interaction, in the
// order of get_pairs, computed on the device from the current pairs and
// velocities (k_pair_output); the per-pair stress buffer is the scratch.
static int inst_get_pair_output(lfd_batch* B, int id, int cap, double* out)
{
    const InstHost* h = live_inst(B, id);
    if (!h) return fail(LFD_ERR_ARG, "get_pair_output: bad instance");
    int rc = batch_check(B, id);
    if (rc) return rc;
    int k0 = 0, k1 = 0;
    if ((rc = pair_range(B, *h, k0, k1))) return rc;
    const int np_ = k1-k0;
    if (np_ > cap || (np_ > 0 && !out)) return fail(LFD_ERR_ARG, "get_pair_output: cap too small");
    if ((size_t)np_*PAIR_OUT > 2*(size_t)STRESS_VALS*B->pcap) return fail(LFD_ERR_CAPACITY, "get_pair_output");
    launch_pair_output(B->pairs[B->cur], k0, k1, B->pos, B->inst_of, B->d_ic, B->sc, B->vel6, B->na6, B->pst, B->s);
    HIPCHK(hipGetLastError());
    if (np_) HIPCHK(hipMemcpyAsync(out, B->pst, (size_t)np_*PAIR_OUT*sizeof(double), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    return np_;
}

static int inst_get_state(lfd_batch* B, int id, double* pos, double* vel6, double* na6, double* sc8)
{
    const InstHost* h = live_inst(B, id);
    if (!h) return LFD_ERR_ARG;
    DevScalars hs;
    int rc = batch_check(B, id, &hs);
    if (rc) return rc;
    const int n = h->ic.n, off = h->ic.poff;
    if (pos) {
        std::vector<double4> v(n);
        HIPCHK(hipMemcpy(v.data(), B->pos+off, n*sizeof(double4), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) { pos[3*i] = v[i].x; pos[3*i+1] = v[i].y; pos[3*i+2] = v[i].z; }
    }
    if (vel6) HIPCHK(hipMemcpy(vel6, B->vel6+6*(size_t)off, 6*(size_t)n*sizeof(double), hipMemcpyDeviceToHost));
    if (na6) HIPCHK(hipMemcpy(na6, B->na6+6*(size_t)off, 6*(size_t)n*sizeof(double), hipMemcpyDeviceToHost));
    if (sc8) fill_scalars8(hs, sc8);
    return LFD_OK;
}

static int inst_get_stress(lfd_batch* B, int id, double* groups24, double* total6, double* shear_rate,
                           double* viscosity, double* particle6)
{
    const InstHost* h = live_inst(B, id);
    if (!h) return LFD_ERR_ARG;
    if (!h->have_stress) return fail(LFD_ERR_ARG, "get_stress: no stress computed yet");
    DevScalars hs;
    int rc = batch_check(B, id, &hs);
    if (rc) return rc;
    if (groups24) std::memcpy(groups24, hs.stress_groups, 24*sizeof(double));
    if (total6) std::memcpy(total6, hs.stress_total, 6*sizeof(double));
    if (shear_rate) *shear_rate = hs.flow.rate;
    if (viscosity) {
        // Simulation::outputData (Simulation.cpp:662-666)
        const double* E = hs.flow.E;
        if (hs.flow.rate == 0) E = h->ic.C.Eh;
        auto dd = [](const double* a, const double* b) {
            return a[0]*b[0]+2*a[1]*b[1]+2*a[2]*b[2]+2*a[3]*b[3]+a[4]*b[4]+a[5]*b[5];
        };
        *viscosity = 0.5*dd(hs.stress_total, E)/dd(E, E);
    }
    if (particle6) {
        const int n = h->ic.n;
        std::vector<double> ps((size_t)STRESS_VALS*n);
        HIPCHK(hipMemcpyAsync(ps.data(), B->ps+(size_t)STRESS_VALS*h->ic.poff, ps.size()*sizeof(double),
                              hipMemcpyDeviceToHost, B->s));
        HIPCHK(hipStreamSynchronize(B->s));
        for (int i = 0; i < n; i++)
            for (int m = 0; m < 6; m++) {
                double t = 0;
                for (int sl = 0; sl < STRESS_VALS/6; sl++) t += ps[(size_t)STRESS_VALS*i+6*sl+m];
                particle6[6*(size_t)i+m] = t;
            }
    }
    return LFD_OK;
}

static int inst_set_scalars(lfd_batch* B, int id, const double* target_stress, const double* strain_limit,
                            const double* rate_limit, const double* strain_end, const double* time = nullptr,
                            const double* strain = nullptr)
{
    if (!live_inst(B, id)) return LFD_ERR_ARG;
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, B->sc+id, sizeof(h), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    if (target_stress) h.target_stress = *target_stress;
    if (strain_limit) h.init_strain_limit = *strain_limit;
    if (rate_limit) h.init_rate_limit = *rate_limit;
    if (strain_end) h.strain_end = *strain_end;
    if (time) h.time = *time;
    if (strain) h.strain = *strain;
    HIPCHK(hipMemcpyAsync(B->sc+id, &h, sizeof(h), hipMemcpyHostToDevice, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    return LFD_OK;
}

// ===================================================================== System (a batch of one)
extern "C" int lfd_system_create(int np, const double* pos, const double* radius, const double L[3],
                                 const double shear_disp[3], const lfd_params* p, const lfd_solve_opts* o,
                                 lfd_system** out)
{
    if (!out || np < 2 || !pos || !radius || !L || !p) return fail(LFD_ERR_ARG, "lfd_system_create: bad argument");
    *out = nullptr;
    InstConst ic;
    std::vector<int> nb, nnb, ctype;
    const double sd0[3] = {0, 0, 0};
    int rc = describe_instance(np, radius, L, p, shear_disp ? shear_disp : sd0, ic, nb, nnb, ctype);
    if (rc) return rc;
    lfd_system* y = new lfd_system();
    rc = batch_init(y, np, ic.nbox, 1, o, LFD_RATE_CONTROL);
    int id = -1;
    if (!rc) rc = bat_add(y, np, pos, radius, L, shear_disp, p, -1.0, 0, &id);
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
    y->S.precond = o->precond;
    y->S.regroup_interval = std::max(1, o->regroup_interval);
    return set_jobs(y->S, live_spans(y), y->s);
}

extern "C" int lfd_system_first_update(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    bat_check_new(y);
    bat_update(y, 1, 0.0);
    return batch_check(y, 0);
}

extern "C" int lfd_system_step(lfd_system* y, int nsteps, double strain_end)
{
    if (!y || nsteps < 0) return fail(LFD_ERR_ARG, "lfd_system_step: bad argument");
    launch_set_strain_end(y->sc, y->ninst, strain_end, y->s);
    for (int k = 0; k < nsteps; k++) {
        int rc = bat_step(y);
        if (rc) return rc;
    }
    return batch_check(y, 0);
}

extern "C" int lfd_system_step_graph(lfd_system* y, int nsteps, double strain_end)
{
    if (!y || nsteps < 0) return fail(LFD_ERR_ARG, "lfd_system_step_graph: bad argument");
    launch_set_strain_end(y->sc, y->ninst, strain_end, y->s);
    const int rc = bat_step_graph(y, nsteps);
    if (rc) return rc;
    return batch_check(y, 0);
}

extern "C" int lfd_system_graphs(const lfd_system* y)
{
    return y ? (y->gexec[0] != nullptr)+(y->gexec[1] != nullptr) : 0;
}

extern "C" int lfd_system_set_control(lfd_system* y, int control, double target_stress)
{
    if (!y || (control != LFD_RATE_CONTROL && control != LFD_STRESS_CONTROL))
        return fail(LFD_ERR_ARG, "lfd_system_set_control: bad argument");
    y->control = control;
    return inst_set_scalars(y, 0, &target_stress, nullptr, nullptr, nullptr);
}

extern "C" int lfd_system_set_shear_rate_limit(lfd_system* y, double init_strain_limit, double init_rate_limit)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_set_scalars(y, 0, nullptr, &init_strain_limit, &init_rate_limit, nullptr);
}

extern "C" int lfd_system_step_stress(lfd_system* y, double strain_end)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    launch_set_strain_end(y->sc, y->ninst, strain_end, y->s);
    int rc = bat_step(y, 1, 0);
    if (rc) return rc;
    return batch_check(y, 0);
}

extern "C" int lfd_system_time_evolution(lfd_system* y, double strain_end, int* nsteps)
{
    if (!y || !(strain_end >= 0)) return fail(LFD_ERR_ARG, "lfd_system_time_evolution: bad argument");
    launch_set_strain_end(y->sc, y->ninst, strain_end, y->s);
    int n = 0;
    const bool fixed = y->inst[0].ic.P.fixed_dt != 0;
    for (;;) {
        DevScalars h;
        int rc = batch_check(y, 0, &h);
        if (rc) return rc;
        // keepRunning(strain_end - loop_time_adjust) (System.cpp:1189-1201, 1237-1242)
        const double se = strain_end-(fixed ? h.dt : 0.0);
        if (h.strain > se-1e-8 && se >= 0) break;
        if ((rc = bat_step(y))) return rc;
        n++;
    }
    int rc = bat_step(y, 1, 0);
    if (rc) return rc;
    if (nsteps) *nsteps = n+1;
    return batch_check(y, 0);
}

extern "C" int lfd_system_compute_stress(lfd_system* y, int in_predictor)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = bat_velocities(y, in_predictor, 1);
    if (rc) return rc;
    return batch_check(y, 0);
}

extern "C" int lfd_system_get_stress(lfd_system* y, double* groups24, double* total6, double* shear_rate,
                                     double* viscosity, double* particle6)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_stress(y, 0, groups24, total6, shear_rate, viscosity, particle6);
}

extern "C" int lfd_system_get_parts(lfd_system* y, int part, double* vel6)
{
    if (!y || part < 0 || part > 3 || !vel6) return fail(LFD_ERR_ARG, "lfd_system_get_parts: bad argument");
    const InstHost& h = y->inst[0];
    const size_t m6 = 6*(size_t)y->M;
    HIPCHK(hipMemcpyAsync(vel6, y->vel_parts+part*m6+6*(size_t)h.ic.poff, 6*(size_t)h.ic.n*sizeof(double),
                          hipMemcpyDeviceToHost, y->s));
    HIPCHK(hipStreamSynchronize(y->s));
    return LFD_OK;
}

extern "C" int lfd_system_sync(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return batch_check(y, 0);
}

extern "C" int lfd_system_check_new_interactions(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    bat_check_new(y);
    return batch_check(y, 0);
}

extern "C" int lfd_system_update_interactions(lfd_system* y, int in_predictor, double dt)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    bat_update(y, in_predictor, dt);
    return batch_check(y, 0);
}

extern "C" int lfd_system_assemble(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    bat_assemble(y);
    return batch_check(y, 0);
}

extern "C" int lfd_system_compute_velocities(lfd_system* y, int in_predictor)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = bat_velocities(y, in_predictor);
    if (rc) return rc;
    return batch_check(y, 0);
}

extern "C" int lfd_system_num_pairs(lfd_system* y)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_num_pairs(y, 0);
}

extern "C" int lfd_system_get_pairs(lfd_system* y, int cap, int* p0, int* p1, int* state, double* geom5,
                                    double* disp_tan)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_pairs(y, 0, cap, p0, p1, state, geom5, disp_tan);
}

extern "C" int lfd_system_get_contacts(lfd_system* y, int cap, int* p0, int* p1, double* disp_tan,
                                       double* disp_rolling)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_contacts(y, 0, cap, p0, p1, disp_tan, disp_rolling);
}

extern "C" int lfd_system_set_contacts(lfd_system* y, int n, const int* p0, const int* p1, const double* disp_tan,
                                       const double* disp_rolling)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_set_contacts(y, 0, n, p0, p1, disp_tan, disp_rolling);
}

extern "C" int lfd_system_set_clock(lfd_system* y, double time, double strain)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_set_scalars(y, 0, nullptr, nullptr, nullptr, nullptr, &time, &strain);
}

extern "C" int lfd_system_get_na_disp(lfd_system* y, double* out3)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_na_disp(y, 0, out3);
}

extern "C" int lfd_system_get_pair_output(lfd_system* y, int cap, double* out10)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_pair_output(y, 0, cap, out10);
}

extern "C" int lfd_system_get_state(lfd_system* y, double* pos, double* vel6, double* na6, double* sc8)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_state(y, 0, pos, vel6, na6, sc8);
}

extern "C" int lfd_system_get_cells(lfd_system* y, int* dims4, int* labels)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = batch_check(y, 0);
    if (rc) return rc;
    const InstConst& g = y->inst[0].ic;
    if (dims4) { dims4[0] = g.nx; dims4[1] = g.ny; dims4[2] = g.nz; dims4[3] = g.boxed; }
    if (labels) {
        HIPCHK(hipMemcpy(labels, y->box_of+g.poff, g.n*sizeof(int), hipMemcpyDeviceToHost));
        for (int i = 0; i < g.n; i++) labels[i] -= g.coff;
    }
    return g.nbox;
}

extern "C" int lfd_system_get_blocks(lfd_system* y, int cap, double* dblocks, double* odblocks, int* p0, int* p1)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    int rc = batch_check(y, 0);
    if (rc) return rc;
    const InstHost& h = y->inst[0];
    int k0 = 0, k1 = 0;
    if ((rc = pair_range(y, h, k0, k1))) return rc;
    const int np_ = k1-k0, off = h.ic.poff;
    if (np_ > cap) return fail(LFD_ERR_ARG, "lfd_system_get_blocks: cap too small");
    if (dblocks) HIPCHK(hipMemcpy(dblocks, y->S.diag64+15*(size_t)off, 15*(size_t)h.ic.n*sizeof(double),
                                  hipMemcpyDeviceToHost));
    if (odblocks && np_) HIPCHK(hipMemcpy(odblocks, y->S.od64+18*(size_t)k0, 18*(size_t)np_*sizeof(double),
                                          hipMemcpyDeviceToHost));
    if (p0 && np_) {
        HIPCHK(hipMemcpy(p0, y->S.p0+k0, np_*sizeof(int), hipMemcpyDeviceToHost));
        for (int k = 0; k < np_; k++) p0[k] -= off;
    }
    if (p1 && np_) {
        HIPCHK(hipMemcpy(p1, y->S.p1+k0, np_*sizeof(int), hipMemcpyDeviceToHost));
        for (int k = 0; k < np_; k++) p1[k] -= off;
    }
    return np_;
}

extern "C" int lfd_system_get_rhs(lfd_system* y, double* rhs)
{
    if (!y || !rhs) return fail(LFD_ERR_ARG, "bad argument");
    int rc = batch_check(y, 0);
    if (rc) return rc;
    const InstConst& g = y->inst[0].ic;
    HIPCHK(hipMemcpy(rhs, y->rhs+6*(size_t)g.poff, 6*(size_t)g.n*sizeof(double), hipMemcpyDeviceToHost));
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

// Waves (8-particle groups) of an instance's aggregate preconditioner; 0 when
// the instance is solved with the 6x6 block-Jacobi preconditioner.
static int inst_precond_waves(lfd_batch* B, int id, int* waves)
{
    const InstHost* h = live_inst(B, id);
    if (!h || !waves) return fail(LFD_ERR_ARG, "precond_waves: bad argument");
    *waves = 0;
    for (int k = 0; k < B->S.n_agg_jobs; k++) {
        AggJob j;
        HIPCHK(hipMemcpyAsync(&j, B->S.agg_jobs+k, sizeof(j), hipMemcpyDeviceToHost, B->s));
        HIPCHK(hipStreamSynchronize(B->s));
        if (j.inst != id) continue;
        HIPCHK(hipMemcpyAsync(waves, B->S.agg.waves+id, sizeof(int), hipMemcpyDeviceToHost, B->s));
        HIPCHK(hipStreamSynchronize(B->s));
    }
    return LFD_OK;
}

extern "C" int lfd_system_precond_waves(lfd_system* y, int* waves)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_precond_waves(y, 0, waves);
}

// Solver slot of each particle of an instance under the aggregate
// preconditioner, counted from the instance's first slot (wave = slot / 8);
// -1 for every particle under 6x6 block-Jacobi.
static int inst_solver_slots(lfd_batch* B, int id, int* slots)
{
    const InstHost* h = live_inst(B, id);
    if (!h || !slots) return fail(LFD_ERR_ARG, "solver_slots: bad argument");
    int waves = 0, rc = inst_precond_waves(B, id, &waves);
    if (rc) return rc;
    const int n = h->ic.n;
    if (waves == 0) {
        std::fill(slots, slots+n, -1);
        return LFD_OK;
    }
    HIPCHK(hipMemcpy(slots, B->S.agg.iperm+h->ic.poff, n*sizeof(int), hipMemcpyDeviceToHost));
    const int soff = h->ic.poff/LDS_ROWS*AGG_SLOTS_PER_BLOCK;
    for (int k = 0; k < n; k++) slots[k] -= soff;
    return LFD_OK;
}

extern "C" int lfd_system_solver_slots(lfd_system* y, int* slots)
{
    if (!y) return fail(LFD_ERR_ARG, "null handle");
    return inst_solver_slots(y, 0, slots);
}

// ===================================================================== batch API
extern "C" int lfd_batch_create(int slot_cap, int cell_cap, int inst_cap, int control, const lfd_solve_opts* o,
                                lfd_batch** out)
{
    if (!out || slot_cap < 2 || cell_cap < 1 || inst_cap < 1 ||
        (control != LFD_RATE_CONTROL && control != LFD_STRESS_CONTROL))
        return fail(LFD_ERR_ARG, "lfd_batch_create: bad argument");
    *out = nullptr;
    if (o && o->kind != LFD_SOLVER_MIXED) return fail(LFD_ERR_UNSUPPORTED, "lfd_batch_create: batches use the LDS solver");
    lfd_batch* B = new lfd_batch();
    int rc = batch_init(B, slot_cap, cell_cap, inst_cap, o, control);
    if (rc) { delete B; return rc; }
    *out = B;
    return LFD_OK;
}

extern "C" void lfd_batch_destroy(lfd_batch* B) { delete B; }

extern "C" int lfd_batch_footprint(int np, const double* radius, const double L[3], const lfd_params* p,
                                   int* slots, int* cells, long long* bytes)
{
    InstConst ic;
    std::vector<int> nb, nnb, ctype;
    const double sd0[3] = {0, 0, 0};
    int rc = describe_instance(np, radius, L, p, sd0, ic, nb, nnb, ctype);
    if (rc) return rc;
    size_t ps = 0, pc = 0;
    batch_unit_bytes(&ps, &pc);
    const int sl = round_up(np, LDS_ROWS);
    if (slots) *slots = sl;
    if (cells) *cells = ic.nbox;
    if (bytes) *bytes = (long long)(ps*sl+pc*ic.nbox);
    return LFD_OK;
}

extern "C" int lfd_batch_room(lfd_batch* B, int slots, int cells)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    std::vector<std::pair<int, int>> us, uc;
    bool handle = false;
    for (int k = 0; k < B->icap; k++) {
        if (B->inst[k].live) {
            us.emplace_back(B->inst[k].ic.poff, B->inst[k].slots);
            uc.emplace_back(B->inst[k].ic.coff, B->inst[k].ic.nbox);
        } else handle = true;
    }
    return handle && first_fit(us, B->M, round_up(slots, LDS_ROWS)) >= 0 && first_fit(uc, B->ccap, cells) >= 0;
}

extern "C" int lfd_batch_add(lfd_batch* B, int np, const double* pos, const double* radius, const double L[3],
                             const double shear_disp[3], const lfd_params* p, double strain_end,
                             double target_stress, int* id)
{
    if (!B || !id) return fail(LFD_ERR_ARG, "lfd_batch_add: bad argument");
    int rc = bat_add(B, np, pos, radius, L, shear_disp, p, strain_end, 1, id);
    if (rc) return rc;
    if (B->control == LFD_STRESS_CONTROL) rc = inst_set_scalars(B, *id, &target_stress, nullptr, nullptr, nullptr);
    return rc;
}

extern "C" int lfd_batch_remove(lfd_batch* B, int id)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    return bat_remove(B, id);
}

extern "C" int lfd_batch_step(lfd_batch* B, int nsteps)
{
    if (!B || nsteps < 0) return fail(LFD_ERR_ARG, "lfd_batch_step: bad argument");
    bool any = false;
    for (int k = 0; k < B->ninst; k++) any = any || B->inst[k].live;
    if (!any) return nsteps ? fail(LFD_ERR_ARG, "lfd_batch_step: empty batch") : LFD_OK;
    for (int k = 0; k < nsteps; k++) {
        int rc = bat_step(B);
        if (rc) return rc;
    }
    return LFD_OK;
}

extern "C" int lfd_batch_step_graph(lfd_batch* B, int nsteps)
{
    if (!B || nsteps < 0) return fail(LFD_ERR_ARG, "lfd_batch_step_graph: bad argument");
    bool any = false;
    for (int k = 0; k < B->ninst; k++) any = any || B->inst[k].live;
    if (!any) return nsteps ? fail(LFD_ERR_ARG, "lfd_batch_step_graph: empty batch") : LFD_OK;
    return bat_step_graph(B, nsteps);
}

extern "C" int lfd_batch_graphs(const lfd_batch* B)
{
    return B ? (B->gexec[0] != nullptr)+(B->gexec[1] != nullptr) : 0;
}

extern "C" int lfd_batch_sync(lfd_batch* B)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    return batch_check(B, -1);
}

extern "C" int lfd_batch_query(lfd_batch* B, int id, double* scalars8, lfd_solve_stats* st, int* status)
{
    if (!live_inst(B, id)) return LFD_ERR_ARG;
    DevScalars h;
    HIPCHK(hipMemcpyAsync(&h, B->sc+id, sizeof(h), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    if (scalars8) fill_scalars8(h, scalars8);
    if (st) { st->iters_inner = h.iters_inner; st->iters_outer = h.iters_outer; st->res_norm = h.res_norm; st->rhs_norm = h.rhs_norm; }
    if (status) {
        std::string keep = g_err;
        *status = status_of(h.error);
        g_err = keep;
    }
    return LFD_OK;
}

extern "C" int lfd_batch_status(lfd_batch* B, int cap, int* ids, double* strain, int* status)
{
    if (!B || cap < 0) return fail(LFD_ERR_ARG, "lfd_batch_status: bad argument");
    std::vector<DevScalars> h(std::max(B->ninst, 1));
    if (B->ninst) HIPCHK(hipMemcpyAsync(h.data(), B->sc, B->ninst*sizeof(DevScalars), hipMemcpyDeviceToHost, B->s));
    BatchScalars hb;
    HIPCHK(hipMemcpyAsync(&hb, B->bs, sizeof(hb), hipMemcpyDeviceToHost, B->s));
    HIPCHK(hipStreamSynchronize(B->s));
    if (hb.error) return status_of(hb.error);
    std::string keep = g_err;
    int m = 0;
    for (int k = 0; k < B->ninst; k++) {
        if (!B->inst[k].live) continue;
        if (m >= cap) return fail(LFD_ERR_ARG, "lfd_batch_status: cap too small");
        if (ids) ids[m] = k;
        if (strain) strain[m] = h[k].strain;
        if (status) status[m] = status_of(h[k].error);
        m++;
    }
    g_err = keep;
    return m;
}

extern "C" int lfd_batch_precond_waves(lfd_batch* B, int id, int* waves)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    return inst_precond_waves(B, id, waves);
}

extern "C" int lfd_batch_solver_slots(lfd_batch* B, int id, int* slots)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    return inst_solver_slots(B, id, slots);
}

extern "C" int lfd_batch_get_state(lfd_batch* B, int id, double* pos, double* vel6, double* na6, double* scalars8)
{
    return inst_get_state(B, id, pos, vel6, na6, scalars8);
}

extern "C" int lfd_batch_get_na_disp(lfd_batch* B, int id, double* out3)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_na_disp(B, id, out3);
}

extern "C" int lfd_batch_get_pair_output(lfd_batch* B, int id, int cap, double* out10)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    return inst_get_pair_output(B, id, cap, out10);
}

extern "C" int lfd_batch_get_pairs(lfd_batch* B, int id, int cap, int* p0, int* p1, int* state, double* geom5,
                                   double* disp_tan)
{
    return inst_get_pairs(B, id, cap, p0, p1, state, geom5, disp_tan);
}

extern "C" int lfd_batch_num_pairs(lfd_batch* B, int id) { return inst_num_pairs(B, id); }

extern "C" int lfd_batch_info(lfd_batch* B, int* live, int* slot_cap, int* cell_cap, int* inst_cap,
                              int* solver_launches, long long* bytes)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    int nl = 0;
    for (auto& h : B->inst) nl += h.live;
    if (live) *live = nl;
    if (slot_cap) *slot_cap = B->M;
    if (cell_cap) *cell_cap = B->ccap;
    if (inst_cap) *inst_cap = B->icap;
    if (solver_launches)
        *solver_launches = (int)(B->S.launches.size()+B->S.agg_launches.size()+B->S.large_launches.size());
    if (bytes) {
        size_t ps = 0, pc = 0;
        batch_unit_bytes(&ps, &pc);
        *bytes = (long long)(ps*B->M+pc*B->ccap);
    }
    return LFD_OK;
}

extern "C" int lfd_batch_profile(lfd_batch* B, int enable, double* solve_ms_total, int* solve_launches)
{
    if (!B) return fail(LFD_ERR_ARG, "null handle");
    int rc = collect_profile(B->S, B->s, solve_ms_total, solve_launches);
    B->S.profiling = enable != 0;
    return rc;
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
    st->S.precond = o->precond;
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
    for (int m = 0; m < 15; m++) { st->dblocks[15*(size_t)i+m] += dbi[m]; st->dblocks[15*(size_t)j+m] += dbj[m]; }
    st->op0.push_back(i);
    st->op1.push_back(j);
    st->odb.insert(st->odb.end(), odb, odb+18);
    return LFD_OK;
}

// Rows of a column may arrive in any order (StokesSolver.h:104-105): the
// column's entries are put in ascending row order here, the order the
// device's pair list and its transpose index assume; a row given twice in
// one column is an error.
static int sort_columns(lfd_stokes* st)
{
    const int P = (int)st->op0.size();
    std::vector<int> perm(P);
    for (int k = 0; k < P; k++) perm[k] = k;
    bool moved = false;
    for (int a = 0; a < P;) {
        int b = a;
        while (b < P && st->op0[b] == st->op0[a]) b++;
        std::stable_sort(perm.begin()+a, perm.begin()+b, [&](int x, int y) { return st->op1[x] < st->op1[y]; });
        for (int k = a+1; k < b; k++)
            if (st->op1[perm[k]] == st->op1[perm[k-1]])
                return fail(LFD_ERR_ARG, "matrixFillingDone: a row appears twice in column " + std::to_string(st->op0[a]));
        for (int k = a; k < b; k++) moved |= perm[k] != k;
        a = b;
    }
    if (!moved) return LFD_OK;
    std::vector<int> r1(P);
    std::vector<double> od(18*(size_t)P);
    for (int k = 0; k < P; k++) {
        r1[k] = st->op1[perm[k]];
        std::memcpy(&od[18*(size_t)k], &st->odb[18*(size_t)perm[k]], 18*sizeof(double));
    }
    st->op1.swap(r1);
    st->odb.swap(od);
    return LFD_OK;
}

extern "C" int lfd_stokes_fill_done(lfd_stokes* st)
{
    if (!st) return fail(LFD_ERR_ARG, "null handle");
    if (int rc = sort_columns(st)) return rc;
    int n = st->np, P = (int)st->op0.size();
    if (P > st->cap) {
        int ncap = std::max(P, 2*st->cap);
        Solver fresh;
        fresh.opts = st->S.opts;
        fresh.kind = st->S.kind;
        fresh.precond = st->S.precond;
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
    h.group_age = -1;
    HIPCHK(hipMemcpyAsync(st->sc, &h, sizeof(h), hipMemcpyHostToDevice, s));
    launch_prepare_matrix(n, st->d_npairs, st->S.diag64, st->S.diag32, st->S.od64, st->S.od32, st->S.minv64,
                          st->S.minv32, st->cap, st->sc, s);
    int rc = check_scalars(st->sc, s);
    if (rc) return rc;
    if ((rc = set_jobs(st->S, std::vector<InstSpan>{{0, 0, n}}, s))) return rc;
    if (st->S.n_agg_jobs > 0) {
        const BsrView A = st->S.view(st->d_npairs);
        launch_agg_group(A, st->S.agg, st->S.agg_jobs, st->S.n_agg_jobs, nullptr, nullptr, st->sc, 1, n, 1,
                         nullptr, nullptr, 0, s);
        launch_agg_inverse(A, st->S.agg, st->S.scap/AGG_CAP, s);
    }
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
