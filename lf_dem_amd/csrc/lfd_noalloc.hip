// lfd_noalloc.hip -- lfd_stokes_solve_noalloc: the drop-in's matrixFillingDone
// + solve (lfd_stokes_fill_done + lfd_stokes_solve) on a matrix, right-hand
// side and solution the caller keeps in device memory, with every work array
// in one file-scope device buffer.  No device memory is allocated or freed:
// for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The work arrays (FP32 copies of the blocks, 6x6 block-Jacobi inverses, the
// lower incidence index, Krylov vectors, partial records, barrier words, the
// job table and the instance scalars) are carved from `scratch`, sized for at
// most NA_MAX_NP particles and NA_MAX_BLOCKS pair blocks.  Each GPU holds its
// own copy of the array; its address on the current device is looked up once
// per device with hipGetSymbolAddress and used for every kernel argument.
// The array keeps the last call's contents: everything a call reads before
// writing (index counters, barrier words, scalars, the solution) is cleared
// by fill kernels at the start of the call.  Calls that overlap on one GPU
// share the array and race.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include "lfd_internal.hpp"
#include "../../include/lfdem_amd.h"

using namespace lfd;

namespace {

constexpr int NA_MAX_NP = LFD_NOALLOC_MAX_NP;
constexpr int NA_MAX_BLOCKS = LFD_NOALLOC_MAX_BLOCKS;
constexpr int NA_RB = (NA_MAX_NP+LDS_ROWS-1)/LDS_ROWS;        // row blocks (solver workgroups)
constexpr int NA_MAX_GROUPS = 2048;                           // SolverWork::max_groups of the library
constexpr size_t NA_PARTIALS = 8*(size_t)NA_MAX_GROUPS > LDS_PART_PER_BLOCK*(size_t)NA_RB+64
                                    ? 8*(size_t)NA_MAX_GROUPS : LDS_PART_PER_BLOCK*(size_t)NA_RB+64;
constexpr size_t NA_BARRIER = 16384 > 128*NA_RB+256 ? 16384 : 128*NA_RB+256;
constexpr size_t NA_ERR_INPUT = 64;                           // DevScalars::error bit: bad block list

constexpr size_t align256(size_t b) { return (b+255)/256*256; }

// byte layout of the buffer
struct Layout {
    size_t diag32, minv64, minv32, od32, row_ptr, tptr, tcnt, tidx, r64, x64, r, z, p, q, w, d, partials, barrier,
        jobs, sc, npairs, total;
};
constexpr Layout make_layout()
{
    Layout L{};
    const size_t n = NA_MAX_NP, P = NA_MAX_BLOCKS, m = 6*n;
    size_t o = 0;
    L.diag32 = o; o += align256(15*n*sizeof(float));
    L.minv64 = o; o += align256(36*n*sizeof(double));
    L.minv32 = o; o += align256(36*n*sizeof(float));
    L.od32 = o; o += align256(18*P*sizeof(float));
    L.row_ptr = o; o += align256((n+1)*sizeof(int));
    L.tptr = o; o += align256((n+1)*sizeof(int));
    L.tcnt = o; o += align256((n+1)*sizeof(int));
    L.tidx = o; o += align256(P*sizeof(int));
    L.r64 = o; o += align256(m*sizeof(double));
    L.x64 = o; o += align256(m*sizeof(double));
    L.r = o; o += align256(m*sizeof(float));
    L.z = o; o += align256(2*m*sizeof(float));
    L.p = o; o += align256(m*sizeof(float));
    L.q = o; o += align256(2*m*sizeof(float));
    L.w = o; o += align256(m*sizeof(float));
    L.d = o; o += align256(m*sizeof(float));
    L.partials = o; o += align256(NA_PARTIALS*sizeof(double));
    L.barrier = o; o += align256(NA_BARRIER*sizeof(unsigned));
    L.jobs = o; o += align256((NA_RB+1)*sizeof(SolveJob));
    L.sc = o; o += align256(sizeof(DevScalars));
    L.npairs = o; o += align256(sizeof(int));
    L.total = o;
    return L;
}
constexpr Layout LAYOUT = make_layout();
static_assert(LAYOUT.total <= LFD_NOALLOC_BYTES, "LFD_NOALLOC_BYTES in lfdem_amd.h must hold the layout");

}  // namespace

// The work buffer (per GPU, for the life of the process).
static __device__ __align__(256) unsigned char scratch[LFD_NOALLOC_BYTES];

namespace {

// ---- device-side set-up of the index, the job table and the scalars
__global__ void k_na_setup(int n, int P, const int* __restrict__ p0, int* row_ptr, int* tcnt, SolveJob* jobs,
                           int G, DevScalars* sc, int* npairs)
{
    const int t = blockIdx.x*blockDim.x+threadIdx.x;
    if (t <= n) {
        // upper incidence: first block of column t (blocks sorted by p0)
        int lo = 0, hi = P;
        while (lo < hi) {
            const int mid = (lo+hi)/2;
            if (p0[mid] < t) lo = mid+1; else hi = mid;
        }
        row_ptr[t] = lo;
        tcnt[t] = 0;
    }
    if (t < G) {
        SolveJob j{};
        j.inst = 0; j.lwg = t; j.G = G; j.row0 = 0; j.row_end = n; j.key = 0;
        jobs[t] = j;
    }
    if (t == 0) {
        DevScalars h{};
        h.npairs = P;
        h.group_age = -1;
        h.strain_end = -1;
        *sc = h;
        *npairs = P;
    }
}

// the check of the block list: p0 < p1 < n, sorted by (p0, p1) without repeats
__device__ inline bool na_valid(int k, int n, const int* __restrict__ p0, const int* __restrict__ p1)
{
    const int i = p0[k], j = p1[k];
    bool ok = i >= 0 && i < j && j < n;
    if (ok && k > 0) {
        const int pi = p0[k-1], pj = p1[k-1];
        ok = pi < i || (pi == i && pj < j);
    }
    return ok;
}

// lower incidence counts of the valid blocks; an invalid one raises the
// input error (the host then stops before the solver)
__global__ void k_na_count(int n, int P, const int* __restrict__ p0, const int* __restrict__ p1, int* tcnt,
                           DevScalars* sc)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= P) return;
    const int j = p1[k];
    if (!na_valid(k, n, p0, p1)) {
        atomicOr(&sc->error, (int)NA_ERR_INPUT);
        return;
    }
    atomicAdd(&tcnt[j], 1);
}

// exclusive scan of n+1 counts (n <= NA_MAX_NP) in one workgroup
constexpr int NA_SCAN_T = 1024;
constexpr int NA_SCAN_PER = (NA_MAX_NP+1+NA_SCAN_T-1)/NA_SCAN_T;
__global__ __launch_bounds__(NA_SCAN_T) void k_na_scan(int n, const int* __restrict__ cnt, int* out)
{
    __shared__ int part[NA_SCAN_T];
    const int t = threadIdx.x, base = t*NA_SCAN_PER;
    int v[NA_SCAN_PER], s = 0;
    for (int e = 0; e < NA_SCAN_PER; e++) {
        v[e] = base+e < n ? cnt[base+e] : 0;
        s += v[e];
    }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < NA_SCAN_T; off <<= 1) {
        const int a = t >= off ? part[t-off] : 0;
        __syncthreads();
        part[t] += a;
        __syncthreads();
    }
    int run = part[t]-s;
    for (int e = 0; e < NA_SCAN_PER; e++) {
        if (base+e <= n) out[base+e] = run;
        run += v[e];
    }
}

// place each valid block in the lower incidence list of its row p1 (the
// same blocks k_na_count counted, so every list stays inside its range)
__global__ void k_na_place(int n, int P, const int* __restrict__ p0, const int* __restrict__ p1,
                           const int* __restrict__ tptr, int* cursor, int* tidx)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= P || !na_valid(k, n, p0, p1)) return;
    const int j = p1[k];
    tidx[tptr[j]+atomicAdd(&cursor[j], 1)] = k;
}

// ascending block index within each list (the order of the library's own index)
__global__ void k_na_sort(int n, const int* __restrict__ tptr, int* tidx)
{
    const int j = blockIdx.x*blockDim.x+threadIdx.x;
    if (j >= n) return;
    const int a = tptr[j], b = tptr[j+1];
    for (int u = a+1; u < b; u++) {
        const int v = tidx[u];
        int w = u-1;
        while (w >= a && tidx[w] > v) { tidx[w+1] = tidx[w]; w--; }
        tidx[w+1] = v;
    }
}

std::mutex g_mu;
unsigned char* g_ptr[64] = {};      // the buffer's address on each device

int na_fail(int code, const std::string& msg) { return set_last_error(code, "lfd_stokes_solve_noalloc: "+msg); }

}  // namespace

extern "C" int lfd_stokes_solve_noalloc(int np, int nblocks, const double* dblocks, const int* p0, const int* p1,
                                        const double* odblocks, const double* rhs, double* x,
                                        const lfd_solve_opts* opts, void* stream, lfd_solve_stats* stats)
{
    if (np < 1 || nblocks < 0 || !dblocks || !rhs || !x || (nblocks > 0 && (!p0 || !p1 || !odblocks)))
        return na_fail(LFD_ERR_ARG, "bad argument");
    if (np > NA_MAX_NP || nblocks > NA_MAX_BLOCKS)
        return na_fail(LFD_ERR_CAPACITY, "at most " + std::to_string(NA_MAX_NP) + " particles and " +
                                             std::to_string(NA_MAX_BLOCKS) + " blocks (got " + std::to_string(np) +
                                             ", " + std::to_string(nblocks) + ")");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) return na_fail(LFD_ERR_HIP, "no current device");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_ptr[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return na_fail(LFD_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
            g_ptr[dev] = static_cast<unsigned char*>(p);
        }
        base = g_ptr[dev];
    }
    lfd_solve_opts uo;
    if (opts) uo = *opts; else lfd_default_solve_opts(&uo);
    SolveOpts o;
    o.tol = uo.tol; o.inner_tol = uo.inner_tol; o.max_outer = uo.max_outer; o.max_inner = uo.max_inner;
    o.lanes_per_row = uo.lanes_per_row; o.block = 256; o.groups = uo.groups; o.max_rhs = 1;
    o.agg_waves = AGG_WAVES_REG;
    // the persistent solver's workgroups wait for each other: a plain launch
    // needs them all resident, which the library takes as given when the grid
    // needs at most half of the device (a cooperative launch is not used here)
    const int G = (np+LDS_ROWS-1)/LDS_ROWS;
    const int cap = lds_solver_max_groups();
    if (2*G > cap) return na_fail(LFD_ERR_CAPACITY, "the solver grid needs more than half of this device");
    o.launch = LFD_LAUNCH_PLAIN;
    hipStream_t s = static_cast<hipStream_t>(stream);

    const Layout& L = LAYOUT;
    auto at = [&](size_t off) { return base+off; };
    BsrView A;
    A.n = np;
    A.npairs = reinterpret_cast<int*>(at(L.npairs));
    A.p0 = p0; A.p1 = p1;
    A.row_ptr = reinterpret_cast<int*>(at(L.row_ptr));
    A.tptr = reinterpret_cast<int*>(at(L.tptr));
    A.tidx = reinterpret_cast<int*>(at(L.tidx));
    A.diag64 = dblocks;
    A.diag32 = reinterpret_cast<float*>(at(L.diag32));
    A.od64 = odblocks;
    A.od32 = reinterpret_cast<float*>(at(L.od32));
    A.minv32 = reinterpret_cast<float*>(at(L.minv32));
    A.minv64 = reinterpret_cast<double*>(at(L.minv64));
    SolverWork w{};
    w.r64 = reinterpret_cast<double*>(at(L.r64)); w.x64 = reinterpret_cast<double*>(at(L.x64));
    w.r = reinterpret_cast<float*>(at(L.r)); w.z = reinterpret_cast<float*>(at(L.z));
    w.p = reinterpret_cast<float*>(at(L.p)); w.q = reinterpret_cast<float*>(at(L.q));
    w.w = reinterpret_cast<float*>(at(L.w)); w.d = reinterpret_cast<float*>(at(L.d));
    w.partials = reinterpret_cast<double*>(at(L.partials));
    w.barrier = reinterpret_cast<unsigned*>(at(L.barrier));
    w.max_groups = NA_MAX_GROUPS;
    int* row_ptr = reinterpret_cast<int*>(at(L.row_ptr));
    int* tptr = reinterpret_cast<int*>(at(L.tptr));
    int* tcnt = reinterpret_cast<int*>(at(L.tcnt));
    int* tidx = reinterpret_cast<int*>(at(L.tidx));
    SolveJob* jobs = reinterpret_cast<SolveJob*>(at(L.jobs));
    DevScalars* sc = reinterpret_cast<DevScalars*>(at(L.sc));
    int* npairs = reinterpret_cast<int*>(at(L.npairs));

    // fresh state: barrier words, counters, scalars (fill kernels, no memset)
    launch_fill_int(reinterpret_cast<int*>(w.barrier), 0, NA_BARRIER, s);
    const int T = 256;
    const int span = std::max(np+1, G);
    k_na_setup<<<(span+T-1)/T, T, 0, s>>>(np, nblocks, p0, row_ptr, tcnt, jobs, G, sc, npairs);
    if (nblocks > 0) k_na_count<<<(nblocks+T-1)/T, T, 0, s>>>(np, nblocks, p0, p1, tcnt, sc);
    k_na_scan<<<1, NA_SCAN_T, 0, s>>>(np, tcnt, tptr);
    launch_fill_int(tcnt, 0, (size_t)np+1, s);
    if (nblocks > 0) {
        k_na_place<<<(nblocks+T-1)/T, T, 0, s>>>(np, nblocks, p0, p1, tptr, tcnt, tidx);
        k_na_sort<<<(np+T-1)/T, T, 0, s>>>(np, tptr, tidx);
    }
    launch_prepare_matrix(np, npairs, dblocks, reinterpret_cast<float*>(at(L.diag32)), odblocks,
                          reinterpret_cast<float*>(at(L.od32)), reinterpret_cast<double*>(at(L.minv64)),
                          reinterpret_cast<float*>(at(L.minv32)), nblocks, sc, s);
    launch_fill_double(x, 0.0, 6*(size_t)np, s);
    e = hipGetLastError();
    if (e != hipSuccess) return na_fail(LFD_ERR_HIP, hipGetErrorString(e));
    // a bad block list is reported before the solver runs on it
    DevScalars h;
    e = hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return na_fail(LFD_ERR_HIP, hipGetErrorString(e));
    if (h.error & NA_ERR_INPUT)
        return na_fail(LFD_ERR_ARG, "blocks must satisfy p0 < p1 < np, sorted by (p0, p1) without repeats");
    if (h.error & E_NOT_SPD) return na_fail(LFD_ERR_NOT_SPD, "a diagonal block is not positive definite");
    e = solve_lds(A, rhs, x, 1, 0, w, sc, jobs, G, o, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return na_fail(LFD_ERR_HIP, hipGetErrorString(e));
    if (stats) {
        stats->iters_inner = h.iters_inner; stats->iters_outer = h.iters_outer;
        stats->res_norm = h.res_norm; stats->rhs_norm = h.rhs_norm;
    }
    if (h.error & E_SYNC_TIMEOUT) return na_fail(LFD_ERR_HIP, "solver workgroups lost synchronisation (spin limit)");
    if (h.error & E_NOT_CONVERGED) return na_fail(LFD_ERR_NOT_CONVERGED, "the solver did not reach the tolerance");
    return LFD_OK;
}
