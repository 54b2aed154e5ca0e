// lfd_internal.hpp -- device data layout and kernel launchers of the
// MI355X Stokes-resistance library (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include "lfd_physics.hpp"

namespace lfd {

// Error bits raised by kernels into DevScalars::error
enum : int {
    E_BOX_OOB = 1,        // a position fell outside the cell grid (BoxSet.cpp:775-778)
    E_FOUND_OVERFLOW = 2, // more new pairs for one particle than the per-row slot count
    E_PAIR_OVERFLOW = 4,  // pair capacity exceeded
    E_NOT_SPD = 8,        // a 6x6 diagonal block is not positive definite
    E_NOT_CONVERGED = 16, // Krylov iteration limit reached before the tolerance
    E_SYNC_TIMEOUT = 32,  // a solver workgroup gave up waiting for the others (bounded spin)
};

// Per-instance scalars kept on the device so a step needs no host round trip.
struct DevScalars {
    double dt, time, strain;
    double shear_disp[3];
    unsigned long long max_vel_bits;    // max |U_na| (bit pattern of a non-negative double)
    unsigned long long max_slide_bits;  // max sliding velocity squared
    double res_norm, rhs_norm;          // last solve: ||b - A x||, ||b||
    double strain_end;                  // the time step is clipped to reach it (< 0: no limit)
    int npairs;                         // StokesSolver staging: block count
    int error;
    int in_predictor;
    int iters_inner, iters_outer;       // last solve
    int total_inner;                    // accumulated over the instance's life
    int group_age;                      // steps since the preconditioner grouping (-1: none yet)
    Flow flow;                          // imposed flow at the current shear rate
    // stress control (System::computeShearRate, System.cpp:2017-2055)
    double target_stress, init_strain_limit, init_rate_limit;
    // last stress computation: groups contact, dashpot, hydro, repulsion
    // (already / volume, hydro with the solvent 2E/(6 pi)) and their total
    double stress_groups[4][6], stress_total[6];
};

// Counters of the whole batch.
struct BatchScalars {
    int npairs;       // pairs of all instances (one list sorted by (p0, p1))
    int error;        // errors not owned by one instance (pair capacity)
    int pad[2];
};

// Everything constant about one instance between admissions: its physics,
// cell grid and where its particles and cells sit in the batch arrays.
// Particles of instance k occupy slots [poff, poff + n) (poff a multiple of
// 128), its cells [coff, coff + nbox).
struct InstConst {
    Consts C;
    Params P;
    double bx, by, bz;
    int nx, ny, nz, yz, nbox, boxed;
    int coff, poff, n, pad;
    double probe_top[16][3], probe_bot[16][3];
};

// Cell arrays of the batch (global cell index; -1 in cell_inst for a free cell).
struct CellArrays {
    int ncell;     // cells in use (grid bound)
    int* nb;       // [ncell*64] neighbour cells incl. self, deduplicated, global index
    int* nnb;      // [ncell]
    int* nstatic;  // [ncell] count of self+static entries (moving ones follow)
    int* ctype;    // [ncell] 0 bulk, 1 bottom, 2 top, 3 top-bottom
    int* cell_inst;
};

// Work item of one workgroup of the batched LDS solver.
struct SolveJob {
    int inst;      // instance (index into the DevScalars array)
    int lwg;       // workgroup index within the instance
    int G;         // workgroups of the instance
    int row0;      // first slot of the instance (solver slot for the aggregate solver)
    int row_end;   // one past its last particle (budget end for the aggregate solver)
    int key;       // first 128-slot particle block of the instance: its partial records and barrier words
    int pad[2];
};

// Aggregate block-Jacobi preconditioner.  Strongly coupled particles are
// grouped (at most 8 per group) and groups are packed into "waves" of at most
// 8 particles; the solver's rows are the waves' slots (8 per wave, -1 for
// padding), so one 64-lane wavefront holds one wave and applies the inverse
// of the wave's 48 x 48 diagonal block.  Each 128-particle slot block of an
// instance owns AGG_SLOTS_PER_BLOCK solver slots (AGG_SLOTS_PER_BLOCK / 8
// waves), enough for any packing of at least 6.4 particles per wave; a
// grouping that needs more falls back to consecutive particles, 8 per wave.
static constexpr int AGG_CAP = 8;                  // particles per wave
static constexpr int AGG_DIM = 6*AGG_CAP;          // rows of a wave block
static constexpr int AGG_PACK = AGG_DIM*(AGG_DIM+1)/2;
static constexpr int AGG_SLOTS_PER_BLOCK = 160;
static constexpr int AGG_WAVES_REG = 8;            // wavefronts (waves) per solver workgroup (default):
                                                   // inverse rows in registers
static constexpr int AGG_WAVES_WIDE = 12;          // the other choice of lfd_solve_opts.agg_waves:
                                                   // inverses in LDS, more waves per launch
static constexpr int AGG_ADJ_K = 8;               // strongest couplings kept per particle for the grouping
static constexpr int AGG_NMAX = 8192;              // largest instance the grouping kernel takes
// Partial-sum records of the persistent solvers: every workgroup of an
// instance polls every record of it.  Records packed into a few cache lines
// put all of those loads on a few memory channels; one 256-byte block per
// record spreads them, which pays up to about a hundred workgroups (measured
// per solve of k_solve_agg: -5.8% at 33 workgroups, -5.6% at 66, -3.2% at 97,
// +1.0% at 132; +6.7% for k_solve_agg_large at 256 workgroups of six-value
// records, where the spread records take 2.7x the cache lines per poll).
static constexpr int REC_STRIDE = 32;              // doubles per record slot when spread
static constexpr int REC_SPREAD_MAX_G = 112;       // spread up to this many workgroups, packed above
static constexpr int AGG_PART_PER_BLOCK = 512;     // partial-record doubles per 128-slot block (four
                                                   // buffers of REC_STRIDE per workgroup, up to 2.5
                                                   // workgroups per block plus one)

struct AggView {
    int* perm;             // [solver slots] particle slot, -1 for padding
    int* iperm;            // [particle slots] solver slot, -1 outside aggregate instances
    int* waves;            // [instances] waves in use
    const int* wave_inst;  // [solver waves] owning instance, -1 when free
    const int* wave_local; // [solver waves] wave index within the owning instance
    unsigned* adj;         // [AGG_ADJ_K * slots] per particle: strongest couplings as (weight << 14 | partner)
    float* inv;            // [solver waves * AGG_PACK] wave block inverse, packed lower triangle (FP32)
    float* zb;             // [6 * solver slots] exchanged Krylov vectors
    float* qb;
    float* db;             // [6 * solver slots] inner solution of the current refinement pass
    double* dsum[2];       // [6 * solver slots] accumulated correction, alternating by pass
    double* r0buf;         // [6 * solver slots] b - A x0
    double* part;          // partial records, AGG_PART_PER_BLOCK per 128-slot block
    size_t part_len;
    int scap;              // solver slots (stride between the right-hand sides of zb and qb)
    int* wexact;           // [solver waves] 1 when inv is the inverse of the wave's whole block
    // instances grouped in spatial chunks (more particles than one grouping
    // workgroup takes): positions are the particles' places in cell order
    int* chunk;            // [instances] particles per chunk, 0 when grouped whole
    int* order;            // [particle slots] particle at each position
    int* rank;             // [particle slots] position of each particle
};

// One instance of the grouping kernel.
struct AggJob {
    int inst, poff, n, soff, budget;   // budget: waves available to the instance
    int chunked;                       // 1: a spatial chunk of a large instance, poff is a position
    int total;                         // chunked: waves of the whole instance
    int pad;
};

// 6x6-block sparse symmetric matrix.  Off-diagonal blocks are stored once per
// pair (p0 < p1) in LF_DEM's packed ODBlock layout, located at (row p1, col p0);
// row p0 uses the transpose.  Upper incidence of row k = pairs [row_ptr[k], row_ptr[k+1])
// (pairs sorted by (p0, p1)); lower incidence = tidx[tptr[k] .. tptr[k+1]).
struct BsrView {
    int n;                 // block rows (particles)
    const int* npairs;     // device count
    const int* p0;         // [P] (strided view into PairRec is avoided: compact arrays)
    const int* p1;
    const int* row_ptr;    // [n+1]
    const int* tptr;       // [n+1]
    const int* tidx;       // [P]
    const double* diag64;  // [n*15]
    const float* diag32;   // [n*15]
    const double* od64;    // [P*18]
    const float* od32;     // [P*18]
    const float* minv32;   // [n*36] block-Jacobi inverse
    const double* minv64;  // [n*36]
};

struct SolverWork {
    // FP64
    double* r64; double* x64;
    // FP32 Krylov vectors
    float *r, *z, *p, *q, *w, *d;
    // FP64 vectors for the plain form
    double *pr, *pz, *pp, *pq;
    double* partials;      // [max_groups * 8]
    unsigned* barrier;     // [16384] tree-barrier words, zeroed once at allocation
    int max_groups;
};

struct SolveOpts {
    double tol;            // relative residual target ||b - A x|| / ||b|| (FP64)
    double inner_tol;      // relative residual drop per FP32 inner solve
    int max_outer, max_inner;
    int lanes_per_row;     // 1, 2, 4 or 8 (SpMV lanes per block row)
    int block;             // workgroup size
    int groups;            // workgroups of the persistent kernel (0 = auto)
    int launch;            // lfd_launch
    int max_rhs;           // right-hand sides one LDS solver launch takes together (1 or 2)
    int agg_waves;         // waves per workgroup of the aggregate solver (8 or 12)
};

// Launch of a persistent solver kernel of G workgroups when at most `cap`
// fit on the device at once: plain when that is asked for, or by default
// when the launch fits the device (the library's kernels run one at a time
// on its stream, so an otherwise idle device holds them all), cooperative
// when that is asked for.
inline hipError_t launch_persistent(const void* fn, int G, int block, void** args, const SolveOpts& o, int cap,
                                    hipStream_t s)
{
    const bool plain = o.launch == 2 || (o.launch == 0 && G <= cap);
    if (plain) return hipLaunchKernel(fn, dim3(G), dim3(block), args, 0, s);
    return hipLaunchCooperativeKernel(fn, dim3(G), dim3(block), args, 0, s);
}

// Records msg as the calling thread's lfd_last_error() and returns code.
int set_last_error(int code, const std::string& msg);

// ---------------- launchers (lfd_kernels.hip) ----------------
void launch_fill_int(int* p, int v, size_t n, hipStream_t s);
void launch_fill_double(double* p, double v, size_t n, hipStream_t s);
void exclusive_scan_int(const int* in, int* out, const int* count_dev, int cap, int* tmp, hipStream_t s,
                        int* out2 = nullptr);

// cell list and pair discovery; `only` >= 0 restricts the search to one instance
void launch_box_labels(const double4* pos, const int* inst_of, int n, const InstConst* IC, int* box_of,
                       DevScalars* sc, hipStream_t s);
void launch_bin(const int* box_of, int n, int nbox, int* count, int* start, int* mem, int* cursor,
                int* scan_tmp, hipStream_t s);
void launch_moving_neighbors(CellArrays ca, const InstConst* IC, const DevScalars* sc, int* probe, hipStream_t s);
void launch_find_pairs(const double4* pos, int n, CellArrays ca, const InstConst* IC, const int* home_start,
                       const int* home_mem, const int* snap_start, const int* snap_mem, DevScalars* sc,
                       int* found, int* found_cnt, int K, int only, hipStream_t s);
void launch_merge_pairs(int n, const PairRec* old_pairs, const int* old_p1, const int* old_row_ptr, const int* found,
                        const int* found_cnt, int K, int* cnt, int* new_row_ptr, int* src, PairRec* new_pairs,
                        const double4* pos, const int* inst_of, const InstConst* IC, const DevScalars* sc,
                        BatchScalars* bs, int pcap, int* scan_tmp, hipStream_t s);
// only >= 0: other instances' pairs are kept unchanged; only == -2: drop the
// pairs of free slots and change nothing else
void launch_update_pairs(PairRec* pairs, const double4* pos, const double* vel6, const int* inst_of,
                         const InstConst* IC, DevScalars* sc, int ninst, const BatchScalars* bs, int* keep,
                         int pcap, int in_predictor, double dt_override, int only, hipStream_t s);
void launch_compact_pairs(const PairRec* in, PairRec* out, const int* keep, int* keep_scan, int n,
                          int* row_ptr, int* p0, int* p1, int* tptr, int* tidx, int* tcnt, BatchScalars* bs,
                          int pcap, int* scan_tmp, hipStream_t s);
void launch_pair_index(const PairRec* pairs, int n, int* row_ptr, int* p0, int* p1, int* tptr, int* tidx,
                       int* tcnt, int* tmp, const BatchScalars* bs, int pcap, int* scan_tmp, hipStream_t s);
void launch_assemble(const PairRec* pairs, const double4* pos, int n, const int* inst_of, const InstConst* IC,
                     DevScalars* sc, const BatchScalars* bs, const int* row_ptr, const int* tptr, const int* tidx,
                     double* diag64, float* diag32, double* od64, float* od32, double* minv64, float* minv32,
                     double* rhs, double* pdb, double* prhs, int pcap, double* rhs_parts, int nparts, hipStream_t s);
// stress (mode 0: stress-control shear rate, mode 1: output groups) and velocity parts
static constexpr int STRESS_VALS = 48;      // 8 slots x 6 per particle / pair side
static constexpr int STRESS_CHUNK = 128;
void launch_stress(const PairRec* pairs, const double4* pos, const int* inst_of, const InstConst* IC, int ninst,
                   DevScalars* sc, const BatchScalars* bs, const double* parts, int nparts, const double* vel6,
                   int n, int control, int mode, const int* row_ptr, const int* tptr, const int* tidx, double* pst,
                   double* ps, double* partial, int pcap, hipStream_t s);
void launch_sum_parts(double* parts, int nparts, double* na6, int n, const int* inst_of, const DevScalars* sc,
                      int rescale, hipStream_t s);
void launch_set_rate(DevScalars* sc, const InstConst* IC, int ninst, double rate, hipStream_t s);
void launch_prepare_matrix(int n, const int* npairs, const double* diag64, float* diag32, const double* od64,
                           float* od32, double* minv64, float* minv32, int pcap, DevScalars* err, hipStream_t s);
void launch_finish_velocity(int n, const double4* pos, const double* na6, double* vel6, const int* inst_of,
                            DevScalars* sc, int ninst, int in_predictor, hipStream_t s);
void launch_adapt_dt(DevScalars* sc, const InstConst* IC, int ninst, int advance_time, hipStream_t s);
void launch_boxing(DevScalars* sc, const InstConst* IC, int ninst, hipStream_t s);
void launch_set_strain_end(DevScalars* sc, int ninst, double strain_end, hipStream_t s);
void launch_displace(double4* pos, double* vel6, const double* pred6, int n, const int* inst_of,
                     const InstConst* IC, DevScalars* sc, int* box_of, int mode, const double* na6, double* na_disp,
                     hipStream_t s);
// per-interaction force columns of the interaction snapshot (PAIR_OUT per pair)
static constexpr int PAIR_OUT = 10;
void launch_pair_output(const PairRec* pairs, int k0, int k1, const double4* pos, const int* inst_of,
                        const InstConst* IC, const DevScalars* sc, const double* vel6, const double* na6, double* out,
                        hipStream_t s);
void launch_copy_double(double* dst, const double* src, size_t n, hipStream_t s);

// solver
int persistent_groups(const SolveOpts& o);
hipError_t solve_mixed(const BsrView& A, const double* b, double* x, SolverWork& w, DevScalars* sc,
                       const SolveOpts& o, hipStream_t s);
// Batched LDS-resident solver: one persistent launch of nwg workgroups
// described by jobs[0..nwg) (all instances of the launch) for nrhs (1 or 2)
// right-hand sides at b + k*rs (solutions at x + k*rs).  Partial records and
// barrier words of an instance sit at LDS_PART_PER_BLOCK and 128 per row block
// from its first row block (row0 / 128); the FP32 vectors w.z, w.q hold nrhs
// vectors.
static constexpr int LDS_ROWS = 128;
static constexpr int LDS_PART_PER_BLOCK = 2*REC_STRIDE;   // two buffers, one record per workgroup
hipError_t solve_lds(const BsrView& A, const double* b, double* x, int nrhs, size_t rs, SolverWork& w,
                     DevScalars* sc, const SolveJob* jobs, int nwg, const SolveOpts& o, hipStream_t s);
int lds_solver_max_groups();
// aggregate preconditioner (lfd_agg.hip)
// Grouping of the instances in jobs that are due (never grouped, or grouped
// `interval` or more steps ago); inst_of / IC null: one instance at slot 0.
// Chunked instances take their cell order from the cell bins (home_start
// over ncell cells, home_mem): pass them when some instance of the batch is
// chunked, null otherwise.
void launch_agg_group(const BsrView& A, const AggView& G, const AggJob* jobs, int njobs, const int* inst_of,
                      const InstConst* IC, DevScalars* sc, int interval, int nslots, int ninst,
                      const int* home_start, const int* home_mem, int ncell, hipStream_t s);
// one step more since each instance's grouping
void launch_agg_age(DevScalars* sc, int ninst, hipStream_t s);
void launch_agg_inverse(const BsrView& A, const AggView& G, int nwaves, hipStream_t s);
// o.agg_waves waves (8 or 12) per workgroup
hipError_t solve_agg(const BsrView& A, const AggView& G, const double* b, double* x, SolverWork& w,
                     DevScalars* sc, const SolveJob* jobs, int nwg, const SolveOpts& o, hipStream_t s);
int agg_solver_max_groups(int waves);
// Large instances (lfd_agg_large.hip): AGG_LARGE_WAVES waves per workgroup,
// one or two right-hand sides per launch, the waves of the instance dealt
// to its workgroups in turn.
static constexpr int AGG_LARGE_WAVES = 16;
hipError_t solve_agg_large(const BsrView& A, const AggView& G, const double* b, double* x, int nrhs, size_t rs,
                           SolverWork& w, DevScalars* sc, const SolveJob* jobs, int nwg, const SolveOpts& o,
                           hipStream_t s);
int agg_large_max_groups();
hipError_t solve_plain(const BsrView& A, const double* b, double* x, SolverWork& w, DevScalars* sc,
                       const SolveOpts& o, hipStream_t s, int* iters_out, double* res_out);
// the same iteration with its scalars on the device (no host round trip per iteration)
hipError_t solve_plain_dev(const BsrView& A, const double* b, double* x, SolverWork& w, DevScalars* sc,
                       const SolveOpts& o, hipStream_t s, int* iters_out, double* res_out);
void launch_spmv64(const BsrView& A, const double* x, double* y, hipStream_t s);
void launch_spmv32(const BsrView& A, const float* x, float* y, int lanes_per_row, hipStream_t s);

}  // namespace lfd
