// lfd_agg.hip -- aggregate block-Jacobi preconditioner for the resistance
// matrix and the persistent mixed-precision solver that applies it (gfx950).
//
// Why: lubrication couples near-contact pairs with resistances ~ 1/(h + delta)
// that are up to a thousand times the Stokes drag.  With 6x6 block-Jacobi the
// slowest modes are rigid motions of such strongly coupled clusters (scaled
// eigenvalues ~ drag / coupling), so PCG needs ~250 iterations at N = 4000,
// phi = 0.56.  Inverting the diagonal block of each group of up to 8 strongly
// coupled particles instead removes most of them (about 0.65x the iterations).
//
// Pipeline, per regrouping (every few steps of an instance) and per assembly:
//   k_agg_adjacency  each particle's 8 strongest couplings (weights from the
//                    pairs' FP64 blocks), sorted;
//   k_agg_group      one workgroup per instance, everything in LDS: three
//                    levels of parallel matching (each group merges with its
//                    mutually strongest neighbour group while the size stays
//                    <= 2, 4, 8), then the groups are packed into waves of
//                    <= 8 particles by a best-fit-decreasing rule evaluated on
//                    the size counts, and every particle gets its solver slot
//                    (slot -> particle and particle -> slot maps);
//   k_agg_inverse    one wavefront per wave: FP64 Gauss-Jordan inverse of the
//                    wave's 48 x 48 diagonal block (all couplings between its
//                    particles), packed lower triangle in FP32;
//   k_solve_agg      FP64 iterative refinement around pipelined FP32 PCG with
//                    the wave inverses as preconditioner; one wavefront per
//                    wave, 12 waves per workgroup, blocks and inverses staged
//                    in LDS once per solve.
#include <hip/hip_runtime.h>
#include "lfd_internal.hpp"
#include "lfd_device.hpp"

namespace lfd {

static constexpr int AG_LEVELS = 3, AG_ROUNDS = 3;
static constexpr int AGB_THREADS = 1024;
static constexpr int AG_RMAX = 64;                  // run-table entries of the packing
static constexpr int AG_HMASK = 16383;              // 14-bit local partner index in a matching key
static constexpr int AG_K = AGG_ADJ_K;
static constexpr int AG_CH = AGG_NMAX/AGB_THREADS;  // particles of one grouping thread at most
static_assert(AG_K % 4 == 0, "adjacency rows are stored as uint4");

// ------------------------------------------------------------------ coupling adjacency
// Per particle, its AG_K strongest couplings as (quantised weight << 14 |
// partner's local index), in descending order of that key, zero-padded, at
// adj[AG_K * slot].  The weight |B|_F / sqrt(|D_p0|_F |D_p1|_F) over the
// packed FP64 values is the size of the coupling relative to the two
// particles' own resistances; the 18-bit quantisation only orders the
// candidates.  A group of at most 8 particles merges along its members'
// strongest links, so the weaker ones are never candidates that matter.
__device__ inline unsigned agg_weight_q(const BsrView& A, int k)
{
    const double* o = A.od64+18*(size_t)k;
    const double* d0 = A.diag64+15*(size_t)A.p0[k];
    const double* d1 = A.diag64+15*(size_t)A.p1[k];
    double so = 0, s0 = 0, s1 = 0;
    for (int m = 0; m < 18; m++) so += o[m]*o[m];
    for (int m = 0; m < 15; m++) { s0 += d0[m]*d0[m]; s1 += d1[m]*d1[m]; }
    const double den = sqrt(sqrt(s0*s1));
    const double w = den > 0 ? sqrt(so)/den : 0.0;
    return (unsigned)fmin(fmax(w, 0.0)*262144.0, 262143.0);
}

// An instance is regrouped at its first assembly and then at the predictor
// assembly of every `interval`-th step of its own, so its preconditioner
// never depends on the other instances of the batch.
__device__ inline bool agg_due(const DevScalars* sc, int interval)
{
    return sc->group_age < 0 || sc->group_age >= interval;
}

__global__ void k_agg_age(DevScalars* sc, int ninst)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k < ninst && sc[k].group_age >= 0) sc[k].group_age++;
}

// Positions of the particles of chunked instances that are due: their places
// in the cell bins (cells in label order, members in ascending particle
// order), counted from the instance's particle offset.
__global__ void k_agg_order(const int* __restrict__ home_start, const int* __restrict__ home_mem, int ncell,
                            const int* __restrict__ inst_of, const InstConst* __restrict__ IC,
                            const DevScalars* __restrict__ scs, int interval, AggView G)
{
    const int q = blockIdx.x*blockDim.x+threadIdx.x;
    if (q >= home_start[ncell]) return;
    const int i = home_mem[q];
    const int inst = inst_of[i];
    if (inst < 0 || G.chunk[inst] <= 0 || !agg_due(scs+inst, interval)) return;
    const int p = IC[inst].poff+(q-home_start[IC[inst].coff]);
    G.order[p] = i;
    G.rank[i] = p;
}

// Adjacency rows are indexed by particle, or by position for a chunked
// instance, whose couplings to other chunks are left out.
__global__ void k_agg_adjacency(BsrView A, AggView G, const int* __restrict__ inst_of,
                                const InstConst* __restrict__ IC, const DevScalars* __restrict__ scs, int interval,
                                int nslots)
{
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= nslots) return;
    int poff = 0, inst = 0;
    if (inst_of) {
        inst = inst_of[i];
        if (inst < 0) return;
        poff = IC[inst].poff;
    }
    if (!agg_due(scs+inst, interval)) return;
    const int cs = inst_of ? G.chunk[inst] : 0;
    int row = i, cbase = poff, cidx = 0;
    if (cs > 0) {
        row = G.rank[i];
        cidx = (row-poff)/cs;
        cbase = poff+cidx*cs;
    }
    unsigned top[AG_K];
#pragma unroll
    for (int j = 0; j < AG_K; j++) top[j] = 0;
    auto offer = [&](unsigned w, int partner) {          // insertion into the descending list
        if (!w) return;
        int pl = partner;
        if (cs > 0) {
            pl = G.rank[partner];
            if ((pl-poff)/cs != cidx) return;
        }
        unsigned a = w << 14 | (unsigned)(pl-cbase);
#pragma unroll
        for (int j = 0; j < AG_K; j++) {
            const unsigned t = top[j];
            top[j] = max(a, t);
            a = min(a, t);
        }
    };
    for (int e = A.tptr[i]; e < A.tptr[i+1]; e++) {
        const int k = A.tidx[e];
        offer(agg_weight_q(A, k), A.p0[k]);
    }
    for (int k = A.row_ptr[i]; k < A.row_ptr[i+1]; k++) offer(agg_weight_q(A, k), A.p1[k]);
    uint4* out = reinterpret_cast<uint4*>(G.adj+(size_t)AG_K*row);
#pragma unroll
    for (int j = 0; j < AG_K/4; j++) out[j] = make_uint4(top[4*j], top[4*j+1], top[4*j+2], top[4*j+3]);
}

// The chunked instances that were due are grouped now.
__global__ void k_agg_done(DevScalars* sc, const int* __restrict__ chunk, int ninst, int interval)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k < ninst && chunk[k] > 0 && agg_due(sc+k, interval)) sc[k].group_age = 0;
}

// Inclusive scan over the workgroup of two 64-bit words (four 16-bit counters each).
__device__ inline void block_scan2(unsigned long long& a, unsigned long long& b, unsigned long long* lds)
{
    const int lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE, nw = blockDim.x/WAVE;
    for (int o = 1; o < WAVE; o <<= 1) {
        unsigned long long ta = __shfl_up(a, o), tb = __shfl_up(b, o);
        if (lane >= o) { a += ta; b += tb; }
    }
    if (lane == WAVE-1) { lds[2*wid] = a; lds[2*wid+1] = b; }
    __syncthreads();
    unsigned long long pa = 0, pb = 0;
    for (int q = 0; q < wid; q++) { pa += lds[2*q]; pb += lds[2*q+1]; }
    a += pa; b += pb;
    (void)nw;
    __syncthreads();
}

__device__ inline int cls_count(unsigned long long lo, unsigned long long hi, int s)
{
    return s <= 4 ? (int)((lo >> (16*(s-1))) & 0xFFFF) : (int)((hi >> (16*(s-5))) & 0xFFFF);
}

// Best-fit decreasing on the counts cnt[1..8] of groups of each size
// (consumed): open a wave with the largest remaining group and fill it with
// the largest groups that fit; runs of identical waves are recorded once in
// the run table.  Returns the number of waves, or -1 when the run table is
// full.
__device__ inline int bfd_counts(int* cnt, unsigned* run_comp, int* run_rep, int* run_bin0, int& nr_out)
{
    int nr = 0, bins = 0;
    for (int s = 8; s >= 1; s--) {
        while (cnt[s] > 0) {
            int take[9], before[9];
            for (int t = 0; t <= 8; t++) { take[t] = 0; before[t] = cnt[t]; }
            take[s] = 1;
            cnt[s]--;
            int rem = 8-s, limited = 0;
            for (int t = min(rem, s); t >= 1; t--) {
                if (rem < t) continue;
                const int want = rem/t, k = min(cnt[t], want);
                if (k < want && k > 0) limited = 1;
                take[t] += k;
                cnt[t] -= k;
                rem -= k*t;
            }
            int rep = limited ? 1 : 1 << 30;
            if (!limited)
                for (int t = 1; t <= 8; t++) if (take[t]) rep = min(rep, before[t]/take[t]);
            unsigned comp = 0;
            for (int t = 1; t <= 8; t++) {
                cnt[t] = before[t]-rep*take[t];
                comp |= (unsigned)take[t] << (4*(t-1));
            }
            if (nr == AG_RMAX) { nr_out = nr; return -1; }
            run_comp[nr] = comp;
            run_rep[nr] = rep;
            run_bin0[nr] = bins;
            nr++;
            bins += rep;
        }
    }
    nr_out = nr;
    return bins;
}

// ------------------------------------------------------------------ grouping and packing
// One workgroup per instance, or per spatial chunk of a large instance.
// Local indices l = particle - poff, or position - poff for a chunk.
__global__ __launch_bounds__(AGB_THREADS) void k_agg_group(BsrView A, AggView G, const AggJob* __restrict__ jobs,
                                                          DevScalars* scs, int interval)
{
    __shared__ short grp[AGG_NMAX];            // group leader (local index)
    __shared__ unsigned char gsz[AGG_NMAX];    // group size, at the leader
    __shared__ unsigned char done[AGG_NMAX];   // merged in this level
    __shared__ unsigned best[AGG_NMAX];        // strongest eligible neighbour group (matching key); later list heads
    __shared__ short relab[AGG_NMAX];          // new leader of a merged group
    __shared__ short nxt[AGG_NMAX];            // member lists
    __shared__ short gbin[AGG_NMAX];           // wave of a group (at the leader)
    __shared__ unsigned char goff[AGG_NMAX];   // first position of the group in its wave
    __shared__ unsigned long long sc_lds[2*AGB_THREADS/WAVE];
    __shared__ unsigned run_comp[AG_RMAX];     // per run: count of each size class in one wave (nibbles)
    __shared__ int run_rep[AG_RMAX], run_bin0[AG_RMAX];
    __shared__ int s_nruns, s_bins, s_fallback;
    const AggJob J = jobs[blockIdx.x];
    if (!agg_due(scs+J.inst, interval)) return;
    const int n = J.n, poff = J.poff;
    const int tid = threadIdx.x, T = AGB_THREADS;
    const int ch = (n+T-1)/T, c0 = min(tid*ch, n), c1 = min(c0+ch, n);   // contiguous chunk for the ranks
    for (int l = tid; l < n; l += T) { grp[l] = (short)l; gsz[l] = 1; best[l] = 0; relab[l] = -1; }
    // the couplings of particles tid, tid + T, ... stay in registers
    unsigned ent[AG_CH][AG_K];
#pragma unroll
    for (int k = 0; k < AG_CH; k++) {
        const int l = tid+k*T;
        const uint4* row = reinterpret_cast<const uint4*>(G.adj+(size_t)AG_K*(poff+l));
#pragma unroll
        for (int j = 0; j < AG_K/4; j++) {
            const uint4 v = l < n ? row[j] : make_uint4(0, 0, 0, 0);
            ent[k][4*j] = v.x; ent[k][4*j+1] = v.y; ent[k][4*j+2] = v.z; ent[k][4*j+3] = v.w;
        }
    }
    __syncthreads();
    for (int lev = 0; lev < AG_LEVELS; lev++) {
        const int cap = 2 << lev;
        for (int l = tid; l < n; l += T) done[l] = 0;
        __syncthreads();
        for (int rnd = 0; rnd < AG_ROUNDS; rnd++) {
            // propose: every particle offers its group the strongest eligible
            // neighbour group among its couplings (the first eligible one in
            // key order); the group keeps the maximum of (weight, smaller
            // group index)
#pragma unroll
            for (int k = 0; k < AG_CH; k++)
#pragma unroll
                for (int j = 0; j < AG_K; j++) asm volatile("" : "+v"(ent[k][j]));   // keep derived values out of registers
#pragma unroll
            for (int k = 0; k < AG_CH; k++) {
                const int l = tid+k*T;
                if (l >= n) continue;
                const int g = grp[l];
                if (done[g]) continue;
                const int sg = gsz[g];
#pragma unroll
                for (int j = 0; j < AG_K; j++) {
                    const unsigned a = ent[k][j];
                    if (!a) break;
                    const int h = grp[a & AG_HMASK];
                    if (h == g || done[h] || sg+gsz[h] > cap) continue;
                    atomicMax(&best[g], (a >> 14) << 14 | (unsigned)(AG_HMASK-h));
                    break;
                }
            }
            __syncthreads();
            // accept mutual choices; the smaller leader absorbs the larger
            for (int l = tid; l < n; l += T) {
                if (grp[l] != l || !best[l]) continue;
                const int h = AG_HMASK-(int)(best[l] & AG_HMASK);
                if (l < h && best[h] && AG_HMASK-(int)(best[h] & AG_HMASK) == l) {
                    relab[h] = (short)l;
                    gsz[l] = (unsigned char)(gsz[l]+gsz[h]);
                    done[l] = 1;
                    done[h] = 1;
                }
            }
            __syncthreads();
            for (int l = tid; l < n; l += T) {
                const int r = relab[grp[l]];
                if (r >= 0) grp[l] = (short)r;
            }
            __syncthreads();
            for (int l = tid; l < n; l += T) { best[l] = 0; relab[l] = -1; }
            __syncthreads();
        }
    }
    // member lists (list heads reuse `best`)
    int* head = reinterpret_cast<int*>(best);
    for (int l = tid; l < n; l += T) head[l] = -1;
    __syncthreads();
    for (int l = tid; l < n; l += T) nxt[l] = (short)atomicExch(&head[grp[l]], l);
    __syncthreads();
    // ---- counts and ranks of the groups by size, in leader order
    unsigned long long lo = 0, hi = 0;
    for (int l = c0; l < c1; l++)
        if (grp[l] == l) {
            const int s = gsz[l];
            if (s <= 4) lo += 1ull << (16*(s-1));
            else hi += 1ull << (16*(s-5));
        }
    unsigned long long plo = lo, phi = hi;
    block_scan2(plo, phi, sc_lds);
    if (tid == T-1) { sc_lds[0] = plo; sc_lds[1] = phi; }
    plo -= lo; phi -= hi;                      // exclusive prefix of this chunk
    __syncthreads();
    if (tid == 0) {
        int cnt[9];
        for (int s = 1; s <= 8; s++) cnt[s] = cls_count(sc_lds[0], sc_lds[1], s);
        int nr = 0;
        const int bins = bfd_counts(cnt, run_comp, run_rep, run_bin0, nr);
        s_nruns = nr;
        s_bins = bins;
        // more waves than the budget: the last ones are dissolved below; only
        // a full run table, or a budget that cannot hold the particles at
        // all, falls back to consecutive particles
        s_fallback = bins < 0 || J.budget*AGG_CAP < n;
    }
    __syncthreads();
    const int kept = min(s_bins, J.budget);
    // particles in wave b (from the run table)
    auto wave_fill = [&](int b) {
        for (int q = 0; q < s_nruns; q++)
            if (b >= run_bin0[q] && b < run_bin0[q]+run_rep[q]) {
                int f = 0;
                for (int t = 1; t <= 8; t++) f += t*(int)((run_comp[q] >> (4*(t-1))) & 0xF);
                return f;
            }
        return 0;
    };
    // particles of the waves [b0, b1), summed run by run
    auto fill_between = [&](int b0, int b1) {
        int acc = 0;
        for (int q = 0; q < s_nruns; q++) {
            const int lo_b = max(b0, run_bin0[q]), hi_b = min(b1, run_bin0[q]+run_rep[q]);
            if (hi_b <= lo_b) continue;
            int f = 0;
            for (int t = 1; t <= 8; t++) f += t*(int)((run_comp[q] >> (4*(t-1))) & 0xF);
            acc += (hi_b-lo_b)*f;
        }
        return acc;
    };
    // the h-th free position of the kept waves, in wave order: 8 wave + pos
    auto hole_slot = [&](int h) {
        for (int q = 0; q < s_nruns; q++) {
            const int b0 = run_bin0[q], nb = min(run_rep[q], kept-b0);
            if (nb <= 0) continue;
            int f = 0;
            for (int t = 1; t <= 8; t++) f += t*(int)((run_comp[q] >> (4*(t-1))) & 0xF);
            const int per = AGG_CAP-f;
            if (h < per*nb) return AGG_CAP*(b0+h/per)+f+h%per;
            h -= per*nb;
        }
        return -1;
    };
    if (!s_fallback) {
        // wave and position of every group from its size class and rank
        int run_c[9];
        for (int s = 1; s <= 8; s++) run_c[s] = 0;
        for (int l = c0; l < c1; l++) {
            if (grp[l] != l) continue;
            gbin[l] = -1;
            const int s = gsz[l];
            const int r = cls_count(plo, phi, s)+run_c[s];
            run_c[s]++;
            int acc = 0;
            for (int q = 0; q < s_nruns; q++) {
                const unsigned comp = run_comp[q];
                const int ts = (int)((comp >> (4*(s-1))) & 0xF);
                if (!ts) continue;
                const int here = ts*run_rep[q];
                if (r < acc+here) {
                    const int u = r-acc, bin = run_bin0[q]+u/ts;
                    int off = 0;
                    for (int t = 8; t > s; t--) off += t*(int)((comp >> (4*(t-1))) & 0xF);
                    off += (u%ts)*s;
                    gbin[l] = (short)bin;
                    goff[l] = (unsigned char)off;
                    break;
                }
                acc += here;
            }
            if (gbin[l] < 0 || goff[l]+s > AGG_CAP) atomicOr(&s_fallback, 1);
        }
    }
    __syncthreads();
    const int fallback = s_fallback;
    // particle of local index l
    auto part = [&](int l) { return J.chunked ? G.order[poff+l] : poff+l; };
    // slots: every used slot and every padding slot of the waves in use is
    // written exactly once (perm) and every particle gets its slot (iperm).
    // The members of the waves beyond the budget (the last ones of the
    // packing, which hold the smallest groups) move one by one into the free
    // positions of the kept waves, in wave order.
    for (int l = tid; l < n; l += T) {
        if (fallback) {
            const int p = part(l);
            G.iperm[p] = J.soff+l;
            G.perm[J.soff+l] = p;
            continue;
        }
        if (grp[l] != l) continue;
        int mem[AGG_CAP], m = 0;
        for (int v = head[l]; v >= 0 && m < AGG_CAP; v = nxt[v]) mem[m++] = v;
        for (int a = 1; a < m; a++) {            // ascending local index
            const int v = mem[a];
            int b = a-1;
            while (b >= 0 && mem[b] > v) { mem[b+1] = mem[b]; b--; }
            mem[b+1] = v;
        }
        const int gb = gbin[l];
        const int moved0 = gb >= kept ? fill_between(kept, gb)+goff[l] : 0;
        for (int a = 0; a < m; a++) {
            const int p = part(mem[a]);
            const int sl = gb < kept ? AGG_CAP*gb+goff[l]+a : hole_slot(moved0+a);
            G.iperm[p] = J.soff+sl;
            G.perm[J.soff+sl] = p;
        }
    }
    const int used = fallback ? (n+AGG_CAP-1)/AGG_CAP : kept;
    if (fallback) {
        for (int q = n+tid; q < AGG_CAP*used; q += T) G.perm[J.soff+q] = -1;
    } else {
        // the free positions of the kept waves beyond the moved particles
        const int moved = fill_between(kept, s_bins);
        for (int b = tid; b < kept; b += T) {
            const int f = wave_fill(b);
            const int h0 = AGG_CAP*b-fill_between(0, b);          // free positions before wave b
            for (int q = f; q < AGG_CAP; q++)
                if (h0+(q-f) >= moved) G.perm[J.soff+AGG_CAP*b+q] = -1;
        }
    }
    // a chunk's waves beyond those it uses are empty (the solver of a
    // chunked instance runs every wave of the instance)
    if (J.chunked)
        for (int q = AGG_CAP*used+tid; q < AGG_CAP*J.budget; q += T) G.perm[J.soff+q] = -1;
    if (tid == 0) {
        if (J.chunked) {
            G.waves[J.inst] = J.total;
        } else {
            G.waves[J.inst] = used;
            scs[J.inst].group_age = 0;
        }
    }
}

void launch_agg_group(const BsrView& A, const AggView& G, const AggJob* jobs, int njobs, const int* inst_of,
                      const InstConst* IC, DevScalars* sc, int interval, int nslots, int ninst,
                      const int* home_start, const int* home_mem, int ncell, hipStream_t s)
{
    if (njobs <= 0) return;
    const bool chunked = home_start && inst_of;
    if (chunked)
        k_agg_order<<<(nslots+255)/256, 256, 0, s>>>(home_start, home_mem, ncell, inst_of, IC, sc, interval, G);
    k_agg_adjacency<<<(nslots+255)/256, 256, 0, s>>>(A, G, inst_of, IC, sc, interval, nslots);
    k_agg_group<<<njobs, AGB_THREADS, 0, s>>>(A, G, jobs, sc, interval);
    if (chunked) k_agg_done<<<(ninst+63)/64, 64, 0, s>>>(sc, G.chunk, ninst, interval);
}

void launch_agg_age(DevScalars* sc, int ninst, hipStream_t s)
{
    if (ninst > 0) k_agg_age<<<(ninst+63)/64, 64, 0, s>>>(sc, ninst);
}

// ------------------------------------------------------------------ wave inverses
static constexpr int AI_THREADS = WAVE;                       // one wavefront per wave block
static constexpr int AI_TILES = AGG_DIM/6;                    // 8 x 8 tiles of 6 x 6
static_assert(AI_TILES*AI_TILES == AI_THREADS, "one 6 x 6 tile per lane");

// One wavefront per solver wave.  The 48 x 48 diagonal block of the wave's
// particles (every coupling between them) is assembled in LDS from the FP64
// blocks: lanes 6p + b place column b of member p's own block, the 8 lanes of
// member p its pairs.  It is then inverted in FP64 by Gauss-Jordan
// elimination without pivoting (the block is SPD): lane t holds the 6 x 6
// tile (t / 8, t % 8) -- the coupling of members t / 8 and t % 8 -- in
// registers, and each step reads the pivot row and column from LDS.  Padding
// positions are the identity during the elimination and zero in the result,
// written as the packed lower triangle in FP32.  A block that is not positive
// definite in the elimination falls back to the particles' 6x6 block-Jacobi
// inverses.
__global__ __launch_bounds__(AI_THREADS) void k_agg_inverse(BsrView A, AggView G)
{
    __shared__ double M[AGG_DIM*AGG_DIM];
    __shared__ double rowb[AGG_DIM], colb[AGG_DIM], pivb;
    __shared__ int mem[AGG_CAP];
    const int w = blockIdx.x;
    const int inst = G.wave_inst[w];
    if (inst < 0 || G.wave_local[w] >= G.waves[inst]) return;
    const int t = threadIdx.x;
    if (t < AGG_CAP) mem[t] = G.perm[AGG_CAP*w+t];
    for (int a = t; a < AGG_DIM*AGG_DIM; a += AI_THREADS) M[a] = 0.0;
    __syncthreads();
    // members' own blocks (padding: identity), column b of member p by lane 6p + b
    if (t < AGG_DIM) {
        const int p = t/6, b = t%6, i = mem[p];
        if (i >= 0) {
            double eb[6] = {0, 0, 0, 0, 0, 0}, y[6] = {0, 0, 0, 0, 0, 0};
            eb[b] = 1;
            dmul<double, double>(A.diag64+15*(size_t)i, eb, y);
            for (int a = 0; a < 6; a++) M[(6*p+a)*AGG_DIM+6*p+b] = y[a];
        } else {
            M[t*AGG_DIM+t] = 1.0;
        }
    }
    // couplings inside the wave, as seen from member p: block (p, q)
    {
        const int p = t/AI_TILES, e0 = t%AI_TILES, i = mem[p];
        if (i >= 0) {
            const int l0 = A.tptr[i], nl = A.tptr[i+1]-l0, u0 = A.row_ptr[i], nt = nl+(A.row_ptr[i+1]-u0);
            for (int e = e0; e < nt; e += AI_TILES) {
                const bool lower = e < nl;              // pair (j, i), j < i: B at (row i, col j)
                const int k = lower ? A.tidx[l0+e] : u0+(e-nl);
                const int q = G.iperm[lower ? A.p0[k] : A.p1[k]]-AGG_CAP*w;
                if (q < 0 || q >= AGG_CAP) continue;
                const double* B = A.od64+18*(size_t)k;
                for (int b = 0; b < 6; b++) {
                    double eb[6] = {0, 0, 0, 0, 0, 0}, y[6] = {0, 0, 0, 0, 0, 0};
                    eb[b] = 1;
                    if (lower) omul<double, double>(B, eb, y);        // column b of B
                    else omul_t<double, double>(B, eb, y);            // column b of B^T
                    for (int a = 0; a < 6; a++) M[(6*p+a)*AGG_DIM+6*q+b] = y[a];
                }
            }
        }
    }
    __syncthreads();
    const int ti = t/AI_TILES, tj = t%AI_TILES, r0 = 6*ti, c0 = 6*tj;
    double R[6][6];
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) R[a][b] = M[(r0+a)*AGG_DIM+c0+b];
    // Step k, with p = a_kk:  a_kk <- 1/p,  a_kj <- a_kj/p,  a_ik <- -a_ik/p,
    // a_ij <- a_ij - a_ik a_kj / p.  The owners of row k and column k publish
    // them with a_kk replaced by 1 in the row and by -1 in the column and zero
    // their own entries, so that every entry is  a_ij - col_i (row_j / p).
    bool bad = false;
    for (int kt = 0; kt < AI_TILES; kt++) {
        const bool rown = ti == kt, cown = tj == kt;
#pragma unroll
        for (int kk = 0; kk < 6; kk++) {
            if (rown) {
#pragma unroll
                for (int b = 0; b < 6; b++) {
                    rowb[c0+b] = (cown && b == kk) ? 1.0 : R[kk][b];
                    R[kk][b] = (cown && b == kk) ? R[kk][b] : 0.0;
                }
            }
            if (cown) {
#pragma unroll
                for (int a = 0; a < 6; a++) colb[r0+a] = (rown && a == kk) ? -1.0 : R[a][kk];
                if (rown) pivb = R[kk][kk];
#pragma unroll
                for (int a = 0; a < 6; a++) R[a][kk] = 0.0;
            }
            __syncthreads();
            double piv = pivb;
            if (!(piv > 0) || !isfinite(piv)) { bad = true; piv = 1.0; }
            const double rp = 1.0/piv;
            double cl[6], rw[6];
#pragma unroll
            for (int a = 0; a < 6; a++) cl[a] = colb[r0+a];
#pragma unroll
            for (int b = 0; b < 6; b++) rw[b] = rowb[c0+b]*rp;
            __syncthreads();
#pragma unroll
            for (int a = 0; a < 6; a++)
#pragma unroll
                for (int b = 0; b < 6; b++) R[a][b] = fma(-cl[a], rw[b], R[a][b]);
        }
    }
    float* out = G.inv+(size_t)w*AGG_PACK;
    const bool live_r = mem[ti] >= 0, live_c = mem[tj] >= 0;
    if (t == 0) G.wexact[w] = bad ? 0 : 1;
    if (bad) {
        // 6x6 block-Jacobi inverses on the diagonal, zero elsewhere
        const int i = mem[ti];
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int b = 0; b < 6; b++) {
                const int r = r0+a, c = c0+b;
                if (c > r) continue;
                out[r*(r+1)/2+c] = (i >= 0 && ti == tj) ? A.minv32[36*(size_t)i+6*a+b] : 0.f;
            }
        return;
    }
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const int r = r0+a, c = c0+b;
            if (c > r) continue;
            out[r*(r+1)/2+c] = (live_r && live_c) ? (float)R[a][b] : 0.f;
        }
}

void launch_agg_inverse(const BsrView& A, const AggView& G, int nwaves, hipStream_t s)
{
    if (nwaves > 0) k_agg_inverse<<<nwaves, AI_THREADS, 0, s>>>(A, G);
}

// ------------------------------------------------------------------ persistent solver
static constexpr int AS_LPR = 8;
static constexpr int AS_ENTRY_CAP = 1024;           // staged blocks per workgroup (78 KB)
static constexpr int AS_POLL_ROUNDS = 3;            // records per lane per poll of the pipelined sums
static constexpr int AS_POLL_ROUNDS8 = 12;          // workgroups per lane per poll of the 8-value sums

struct AsRow { int e0, ne; };       // staged entries of a slot row (ne < 0: not staged)

// Same structure as the 6x6 block-Jacobi persistent solver (FP64 refinement
// with the translation coarse correction around pipelined FP32 PCG, classic
// PCG once a refinement pass fails to reduce the residual), with:
//  * rows = solver slots: lane c < 6 of the 8-lane group of slot s owns
//    component c of particle perm[s]; b and x stay in particle order, the
//    exchanged Krylov vectors are in slot order;
//  * the preconditioner z = W^-1 r for the wave W of the lane's wavefront:
//    every lane forms its row of the symmetric inverse with the 48 residual
//    components read from their lanes into scalar registers.
// AS_WAVES waves (wavefronts) per workgroup: 12 (three per SIMD) or 8 (two
// per SIMD, more workgroups for the same waves).  With 12 the wave inverses
// are staged in LDS (packed lower triangle); with 8 a wavefront may hold 256
// VGPRs, and each lane keeps its 48-entry row of the inverse in registers,
// so an application reads no LDS.
template <int AS_WAVES>
__global__ __launch_bounds__(AS_WAVES*WAVE) void k_solve_agg(BsrView A, AggView Gv, const double* __restrict__ b,
    double* __restrict__ x, SolverWork w, DevScalars* scs, SolveOpts o, const SolveJob* __restrict__ jobs)
{
    constexpr int AS_BLOCK = AS_WAVES*WAVE;
    constexpr int AS_ROWS = AS_BLOCK/AS_LPR;        // solver slots per workgroup
    constexpr bool INV_REG = AS_WAVES <= 8;         // inverse rows in registers (else LDS)
    __shared__ int ent_j[AS_ENTRY_CAP];             // partner slot; ~slot for the transposed block
    __shared__ float ent_b[AS_ENTRY_CAP*18];
    __shared__ float s_inv[INV_REG ? 1 : AS_WAVES*AGG_PACK];
    __shared__ AsRow rows[AS_ROWS];
    __shared__ double red[3*AS_WAVES+4];
    __shared__ double red8[8*AS_WAVES+8];
    __shared__ double s_tv[4];
    __shared__ double s_ei[9];
    __shared__ int s_fail;
    const SolveJob J = jobs[blockIdx.x];
    const int nw = Gv.waves[J.inst];
    const int G = (nw+AS_WAVES-1)/AS_WAVES, wg = J.lwg;
    if (wg >= G) return;                            // beyond the waves in use: not part of this solve
    const int sbase = J.row0+wg*AS_ROWS, send = J.row0+AGG_CAP*nw;
    DevScalars* sc = scs+J.inst;
    unsigned* bar = w.barrier+128*(size_t)J.key;
    unsigned seq = 0;
    const int lr = threadIdx.x/AS_LPR, c = threadIdx.x%AS_LPR;
    const unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE;
    const int slot = sbase+lr;
    const int i = slot < send ? Gv.perm[slot] : -1;
    const bool live = i >= 0;
    const bool own = live && c < 6;
    const size_t gi = 6*(size_t)(live ? i : 0)+c;   // particle order (b, x)
    const size_t gs = 6*(size_t)slot+c;             // slot order (exchanged vectors)
    const int kk = c < 6 ? 6*(int)(lane/AS_LPR)+c : 0;
    unsigned gen = ld_acquire(bar+32);
    if (threadIdx.x == 0) s_fail = 0;
    // this wavefront's wave (global index); exact: its inverse is the
    // inverse of the wave's whole block (not the 6x6 fallback)
    const bool wlive = (int)wid < nw-(sbase-J.row0)/AGG_CAP;
    const int wave = sbase/AGG_CAP+(int)wid;
    const bool exact = wlive && Gv.wexact[wave] != 0;
    // row kk of this wavefront's inverse (register form; zero for lanes 6, 7
    // of a row group and for waves beyond the instance's)
    float Mr[INV_REG ? AGG_DIM : 1];
    if constexpr (INV_REG) {
        const float* src = Gv.inv+(size_t)(wlive ? wave : sbase/AGG_CAP)*AGG_PACK;
        const int kb = kk*(kk+1)/2;
#pragma unroll
        for (int m = 0; m < AGG_DIM; m++) Mr[m] = (wlive && c < 6) ? src[m <= kk ? kb+m : m*(m+1)/2+kk] : 0.f;
    } else {
        const float* src = Gv.inv+(size_t)(sbase/AGG_CAP)*AGG_PACK;
        const int nload = min(AS_WAVES, nw-(sbase-J.row0)/AGG_CAP)*AGG_PACK;
        for (int a = threadIdx.x; a < AS_WAVES*AGG_PACK; a += AS_BLOCK) s_inv[a] = a < nload ? src[a] : 0.f;
    }
    // ---- stage the couplings of each row to other waves (all of them when
    // the wave's inverse is the 6x6 fallback): the product is Eisenstat's,
    // A u = r + E u for u = W^-1 r, W the waves' diagonal blocks
    int l0 = 0, nl = 0, u0 = 0, nt = 0;
    if (live) {
        l0 = A.tptr[i]; nl = A.tptr[i+1]-l0; u0 = A.row_ptr[i];
        nt = nl+(A.row_ptr[i+1]-u0);
    }
    // entry e of row i: the pair index and the partner's slot (~slot when
    // the row applies the transposed block)
    auto entry = [&](int e, int& k) -> int {
        if (e < nl) { k = A.tidx[l0+e]; return Gv.iperm[A.p0[k]]; }
        k = u0+(e-nl);
        return ~Gv.iperm[A.p1[k]];
    };
    auto kept = [&](int jj) { return !exact || (jj >= 0 ? jj : ~jj)/AGG_CAP != wave; };
    int mine = 0;
    for (int e = c; e < nt; e += AS_LPR) {
        int k;
        if (kept(entry(e, k))) mine++;
    }
    int pre = mine;                                  // inclusive prefix over the row's 8 lanes
    for (int d = 1; d < AS_LPR; d <<= 1) {
        const int t = __shfl_up(pre, d, AS_LPR);
        if (c >= d) pre += t;
    }
    const int rowtot = __shfl(pre, AS_LPR-1, AS_LPR);
    pre -= mine;
    if (c == 0) rows[lr].ne = live ? rowtot : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        int used = 0;
        for (int t = 0; t < AS_ROWS; t++) {
            const int deg = rows[t].ne;
            if (deg > 0 && used+deg <= AS_ENTRY_CAP) { rows[t].e0 = used; used += deg; }
            else { rows[t].e0 = 0; rows[t].ne = deg > 0 ? -1 : 0; }
        }
    }
    __syncthreads();
    const int ne = live ? rows[lr].ne : 0, e0 = live ? rows[lr].e0 : 0;
    if (ne > 0) {
        int sidx = e0+pre;
        for (int e = c; e < nt; e += AS_LPR) {
            int k;
            const int jj = entry(e, k);
            if (!kept(jj)) continue;
            ent_j[sidx] = jj;
            for (int m = 0; m < 18; m++) ent_b[18*sidx+m] = A.od32[18*(size_t)k+m];
            sidx++;
        }
    }
    __syncthreads();

    // component c of yin + E v for this row: neighbours' v from vbuf (slot
    // order); with v = W^-1 yin this is A v
    auto spmv_e = [&](const float* vbuf, float yin) -> float {
        float y[6] = {0, 0, 0, 0, 0, 0};
        if (live) {
            if (ne >= 0) {
                for (int e = c; e < ne; e += AS_LPR) {
                    const int jj = ent_j[e0+e];
                    const float* bl = ent_b+18*(e0+e);
                    float xj[6];
                    load6(vbuf+6*(size_t)(jj >= 0 ? jj : ~jj), xj);
                    if (jj >= 0) omul<float, float>(bl, xj, y);
                    else omul_t<float, float>(bl, xj, y);
                }
            } else {
                for (int e = c; e < nt; e += AS_LPR) {
                    int k;
                    const int jj = entry(e, k);
                    if (!kept(jj)) continue;
                    float xj[6];
                    load6(vbuf+6*(size_t)(jj >= 0 ? jj : ~jj), xj);
                    if (jj >= 0) omul<float, float>(A.od32+18*(size_t)k, xj, y);
                    else omul_t<float, float>(A.od32+18*(size_t)k, xj, y);
                }
            }
        }
        float me = 0;
        for (int m = 0; m < 6; m++) {
            const float t = sum8f(y[m]);
            if (m == c) me = t;
        }
        return own ? yin+me : 0.f;
    };
    // z = W^-1 r on the lane's wave: component m of r is read from its lane
    // (8 (m / 6) + m % 6) into a scalar register, so LDS serves only the
    // inverse (one read per product)
    auto precond = [&](float rv) -> float {
        const int x = __float_as_int(own ? rv : 0.f);
        if constexpr (INV_REG) {
            float z = 0.f;
#pragma unroll
            for (int m = 0; m < AGG_DIM; m++)
                z += Mr[m]*__int_as_float(__builtin_amdgcn_readlane(x, 8*(m/6)+m%6));
            return own ? z : 0.f;
        } else {
            const float* Mw = s_inv+wid*AGG_PACK;
            int kq = kk;
            asm volatile("" : "+v"(kq));   // the addresses are recomputed per call, not held in registers
            const int kb = kq*(kq+1)/2;
            float z = 0.f;
#pragma unroll
            for (int m = 0; m < AGG_DIM; m++) {
                const float sm = __int_as_float(__builtin_amdgcn_readlane(x, 8*(m/6)+m%6));
                const int a = m <= kq ? kb+m : m*(m+1)/2+kq;
                z += Mw[a]*sm;
            }
            return own ? z : 0.f;
        }
    };

    double* part = Gv.part+AGG_PART_PER_BLOCK*(size_t)J.key;
    const int rs3 = rec_stride(G, 3), rs8 = rec_stride(G, 8);   // record strides of the 3- and 8-value sums
    int buf = 0;
    auto publish3 = [&](double a, double bs, double cs) {
        a = wave_sum_d(a); bs = wave_sum_d(bs); cs = wave_sum_d(cs);
        if (lane == 0) { red[3*wid] = a; red[3*wid+1] = bs; red[3*wid+2] = cs; }
        drain_vmem();
        __syncthreads();
        double* pb = part+(size_t)buf*rs3*G;
        seq++;
        if (wid == 0) {
            double t0 = lane < AS_WAVES ? red[3*lane] : 0.0;
            double t1 = lane < AS_WAVES ? red[3*lane+1] : 0.0;
            double t2 = lane < AS_WAVES ? red[3*lane+2] : 0.0;
            const unsigned tg = seq & 0xFF;
            t0 = tag_partial(wave_sum_d(t0), tg);
            t1 = tag_partial(wave_sum_d(t1), tg);
            t2 = tag_partial(wave_sum_d(t2), tg);
            release_for_records();
            if (lane == 0) {
                st_dev(pb+rs3*wg, t0); st_dev(pb+rs3*wg+1, t1); st_dev(pb+rs3*wg+2, t2);
            }
            const unsigned long long* pq = reinterpret_cast<const unsigned long long*>(pb);
            bool ok = false;
            for (long spin = 0; s_fail == 0 && spin < SPIN_LIMIT; spin++) {
                double ac[3] = {0, 0, 0};
                const bool dn = poll_records<3, AS_POLL_ROUNDS>(pq, G, rs3, tg, lane, ac);
                if (__all(dn)) {
                    const double a0 = wave_sum_d(ac[0]), a1 = wave_sum_d(ac[1]), a2 = wave_sum_d(ac[2]);
                    if (lane == 0) { red[3*AS_WAVES] = a0; red[3*AS_WAVES+1] = a1; red[3*AS_WAVES+2] = a2; }
                    ok = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ok && lane == 0) s_fail = 1;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
    };
    // the instance totals, in place after publish3 (the slots are rewritten
    // only after the next call's first barrier)
    auto gather3_start = [&](double& t0, double& t1, double& t2) {
        t0 = red[3*AS_WAVES]; t1 = red[3*AS_WAVES+1]; t2 = red[3*AS_WAVES+2];
    };
    auto gather3_finish = [&](double&, double&, double&) { buf ^= 1; };
    auto gsum3 = [&](double& a, double& bs, double& cs) {
        publish3(a, bs, cs);
        gather3_start(a, bs, cs);
        gather3_finish(a, bs, cs);
    };

    // Sums of eight values over every workgroup of the instance, for the
    // refinement passes: own double-buffered records after the pipelined
    // ones, own round tags (a record is rewritten every second call).
    unsigned seq8 = 0;
    int buf8 = 0;
    auto gsum8 = [&](double* v) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const double t = wave_sum_d(v[m]);
            if (lane == 0) red8[8*wid+m] = t;
        }
        drain_vmem();
        __syncthreads();
        seq8++;
        if (wid == 0) {
            // lane L works on value L % 8 throughout: its workgroup total, then
            // that value of the workgroups L / 8, L / 8 + 8, ... of the instance
            // (a workgroup's record: 8 values at the start of its slot)
            double* pb = part+2*rs3*(size_t)G+(size_t)buf8*rs8*G;
            const unsigned tg = seq8 & 0xFF;
            const int m = lane & 7;
            double t = 0;
            for (int q = 0; q < AS_WAVES; q++) t += red8[8*q+m];
            t = tag_partial(t, tg);
            release_for_records();
            if (lane < 8) st_dev(pb+rs8*wg+m, t);
            const unsigned long long* pq = reinterpret_cast<const unsigned long long*>(pb)+m;
            bool ok = false;
            for (long spin = 0; s_fail == 0 && spin < SPIN_LIMIT; spin++) {
                double acc = 0;
                bool dn = true;
                for (int g0 = (int)(lane >> 3); g0 < G; g0 += 8*AS_POLL_ROUNDS8) {
                    unsigned long long v[AS_POLL_ROUNDS8];
#pragma unroll
                    for (int r = 0; r < AS_POLL_ROUNDS8; r++) {
                        const int g = g0+8*r;
                        v[r] = g < G ? ld_dev(pq+(size_t)rs8*g) : (unsigned long long)tg;
                    }
#pragma unroll
                    for (int r = 0; r < AS_POLL_ROUNDS8; r++) {
                        dn &= (v[r] & 0xFF) == tg;
                        acc += untag_partial(v[r]);
                    }
                }
                if (__all(dn)) {
                    acc += __shfl_xor(acc, 8);
                    acc += __shfl_xor(acc, 16);
                    acc += __shfl_xor(acc, 32);
                    if (lane < 8) red8[8*AS_WAVES+m] = acc;
                    ok = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ok && lane == 0) s_fail = 1;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = red8[8*AS_WAVES+m];
        buf8 ^= 1;
    };

    // Galerkin correction onto the three uniform translations, as in the
    // block-Jacobi solver: (A z_a) rows in registers, E^-1 in LDS
    double AZ[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        double y[6] = {0, 0, 0, 0, 0, 0};
        if (live) row_translation<AS_LPR>(A, i, c, a, y);
        for (int m = 0; m < 6; m++) if (m == c && own) AZ[a] = y[m];
    }
    bool use_coarse;
    {
        double ev[8] = {c == 0 ? AZ[0] : 0.0, c == 1 ? AZ[1] : 0.0, c == 2 ? AZ[2] : 0.0,
                        c == 0 ? AZ[1] : 0.0, c == 0 ? AZ[2] : 0.0, c == 1 ? AZ[2] : 0.0, 0.0, 0.0};
        gsum8(ev);
        const double e00 = ev[0], e11 = ev[1], e22 = ev[2], e01 = ev[3], e02 = ev[4], e12 = ev[5];
        const double c00 = e11*e22-e12*e12, c01 = e02*e12-e01*e22, c02 = e01*e12-e02*e11;
        const double det = e00*c00+e01*c01+e02*c02;
        use_coarse = det > 0 && isfinite(det) && e00 > 0;
        const double id = use_coarse ? 1.0/det : 0.0;
        if (threadIdx.x == 0) {
            s_ei[0] = c00*id; s_ei[1] = c01*id; s_ei[2] = c02*id;
            s_ei[4] = (e00*e22-e02*e02)*id; s_ei[5] = (e01*e02-e00*e12)*id; s_ei[8] = (e00*e11-e01*e01)*id;
            s_ei[3] = s_ei[1]; s_ei[6] = s_ei[2]; s_ei[7] = s_ei[5];
        }
        __syncthreads();
    }

    // Refinement passes.  x stays at its initial value x0 until the end; the
    // correction D = x - x0 is accumulated in FP64: D_0 = 0 and
    // D_k = (D_{k-1} + t_{k-1}) + s_{k-1} d_{k-1}, with t the translation
    // shift of the coarse correction and s d the scaled FP32 inner solution of
    // pass k-1.  The residual of pass k is b - A x0 - A D_k.  Every
    // workgroup forms D_k of its neighbours from the previous pass's D and the
    // exchanged d (both in slot order), so the end of a pass needs no barrier:
    // d is exchanged by the inner solver's own reductions.
    // A64 D_k for this row (component c)
    auto spmv64_d = [&](const double* Dp, bool first, const double* tv, double sv) __attribute__((always_inline)) -> double {
        double y[6] = {0, 0, 0, 0, 0, 0};
        auto dval = [&](int js, double* xj) __attribute__((always_inline)) {
            for (int m = 0; m < 6; m++) {
                double v = first ? 0.0 : Dp[6*(size_t)js+m];
                if (m < 3) v += tv[m];
                xj[m] = v+sv*(double)Gv.db[6*(size_t)js+m];
            }
        };
        if (live) {
            for (int e = c; e < nt; e += AS_LPR) {
                int k;
                const int jj = entry(e, k);
                double xj[6];
                dval(jj >= 0 ? jj : ~jj, xj);
                if (jj >= 0) omul<double, double>(A.od64+18*(size_t)k, xj, y);
                else omul_t<double, double>(A.od64+18*(size_t)k, xj, y);
            }
            if (c == 0) {
                double xi[6];
                dval(slot, xi);
                dmul<double, double>(A.diag64+15*(size_t)i, xi, y);
            }
        }
        for (int o2 = 1; o2 < AS_LPR; o2 <<= 1)
            for (int m = 0; m < 6; m++) y[m] += __shfl_xor(y[m], o2, AS_LPR);
        double mine = 0;
        for (int m = 0; m < 6; m++) if (m == c) mine = y[m];
        return mine;
    };

    double bb = 0, rr64 = 0, Dfin = 0;
    int outer, inner_total = 0;
    bool converged = false, use_classic = false, zero_rhs = false;
    // per lane in global memory (slot order): b - A x0 in r0buf, D in dsum,
    // the last inner solution in db; the last shift and scale in s_tv
    for (outer = 0; outer <= o.max_outer; outer++) {
        double r64 = 0, bi = 0, Dv = 0;
        if (outer == 0) {
            double yv[6] = {0, 0, 0, 0, 0, 0};
            if (live) spmv_row<double, double, AS_LPR>(A, A.od64, A.diag64+15*(size_t)i, x, i, c, yv);
            if (own) {
                for (int m = 0; m < 6; m++) if (m == c) r64 = yv[m];
                bi = b[gi];
                r64 = bi-r64;
                Gv.r0buf[gs] = r64;
            }
        } else {
            const double tv[3] = {s_tv[0], s_tv[1], s_tv[2]}, sv = s_tv[3];
            const double y = spmv64_d((outer & 1) ? Gv.dsum[0] : Gv.dsum[1], outer == 1, tv, sv);
            if (own) {
                const double Dold = outer == 1 ? 0.0 : ((outer & 1) ? Gv.dsum[0] : Gv.dsum[1])[gs];
                Dv = (Dold+(c == 0 ? tv[0] : c == 1 ? tv[1] : c == 2 ? tv[2] : 0.0))+sv*(double)Gv.db[gs];
                ((outer & 1) ? Gv.dsum[1] : Gv.dsum[0])[gs] = Dv;
                r64 = Gv.r0buf[gs]-y;
            }
        }
        double rv[8] = {r64*r64, bi*bi, c == 0 ? r64 : 0.0, c == 1 ? r64 : 0.0, c == 2 ? r64 : 0.0, 0.0, 0.0, 0.0};
        if (!own) rv[2] = rv[3] = rv[4] = 0.0;
        gsum8(rv);
        if (s_fail) break;
        const double s_rr = rv[0];
        if (outer == 0) bb = rv[1];
        if (bb == 0) {
            zero_rhs = true;
            rr64 = 0;
            converged = true;
            break;
        }
        if (outer > 0 && !(s_rr < 0.25*rr64)) use_classic = true;
        rr64 = s_rr;
        converged = sqrt(rr64) <= o.tol*sqrt(bb);
        if (converged || outer == o.max_outer) {
            Dfin = Dv;
            break;
        }
        double tv[3] = {0, 0, 0};
        if (use_coarse) {
            for (int a = 0; a < 3; a++) tv[a] = s_ei[3*a]*rv[2]+s_ei[3*a+1]*rv[3]+s_ei[3*a+2]*rv[4];
            if (own) r64 -= AZ[0]*tv[0]+AZ[1]*tv[1]+AZ[2]*tv[2];
        }
        if (threadIdx.x == 0) { s_tv[0] = tv[0]; s_tv[1] = tv[1]; s_tv[2] = tv[2]; s_tv[3] = sqrt(rr64); }
        float d = 0.f;
        const double scale = 1.0/sqrt(rr64);
        int it = 0;
        if (!use_classic) {
            float* mbuf[2] = {Gv.zb, Gv.qb};
            float r = own ? (float)(r64*scale) : 0.f;
            float u = precond(r);
            if (own) mbuf[1][gs] = u;
            double s_r2 = (double)r*r, z0 = 0, z1 = 0;
            gsum3(s_r2, z0, z1);
            const double stop2 = inner_stop2(o.inner_tol, s_r2, o.tol, bb, rr64);
            float wv = spmv_e(mbuf[1], r);          // A u = r + E u
            float zv = 0, qv = 0, sv2 = 0, pv = 0;
            double gamma_old = 1, alpha_old = 1;
            int par = 0;
            for (it = 0; it < o.max_inner; it++) {
                float mv = precond(wv);
                if (own) { mbuf[par][gs] = mv; Gv.db[gs] = d; }
                publish3((double)r*u, (double)wv*u, (double)r*r);
                double g, dl, rho;
                gather3_start(g, dl, rho);
                float nv = spmv_e(mbuf[par], wv);   // A m = w + E m
                gather3_finish(g, dl, rho);
                if (s_fail || rho <= stop2) break;
                double beta, alpha;
                if (it > 0) {
                    beta = g/gamma_old;
                    double den = dl-beta*g/alpha_old;
                    if (!(den > 0)) break;
                    alpha = g/den;
                } else {
                    if (!(dl > 0)) break;
                    beta = 0;
                    alpha = g/dl;
                }
                const float bf = (float)beta, af = (float)alpha;
                zv = nv+bf*zv;
                qv = mv+bf*qv;
                sv2 = wv+bf*sv2;
                pv = u+bf*pv;
                d += af*pv;
                r -= af*sv2;
                u -= af*qv;
                wv -= af*zv;
                gamma_old = g;
                alpha_old = alpha;
                par ^= 1;
            }
            if (it == o.max_inner) {                  // d changed after its last exchange
                if (own) Gv.db[gs] = d;
                tree_barrier(bar, G, wg, gen, &s_fail);
            }
        } else {
            float r = own ? (float)(r64*scale) : 0.f;
            float z = precond(r), p = z;
            if (own) Gv.zb[gs] = z;
            double s_rz = (double)r*z, s_r2 = (double)r*r, z0 = 0;
            gsum3(s_rz, s_r2, z0);
            double rz = s_rz;
            const double stop2 = inner_stop2(o.inner_tol, s_r2, o.tol, bb, rr64);
            float q = spmv_e(Gv.zb, r);             // A z = r + E z
            double s_pq = (double)p*q, z1 = 0, z2 = 0;
            gsum3(s_pq, z1, z2);
            double pq = s_pq;
            for (it = 0; it < o.max_inner; it++) {
                if (s_fail || !(pq > 0)) break;
                const float alpha = (float)(rz/pq);
                d += alpha*p;
                r -= alpha*q;
                z = precond(r);
                if (own) Gv.zb[gs] = z;
                double s_rz2 = (double)r*z, s_rr2 = (double)r*r, z3 = 0;
                gsum3(s_rz2, s_rr2, z3);
                if (s_rr2 <= stop2) { it++; break; }
                const float beta = (float)(s_rz2/rz);
                rz = s_rz2;
                float wz = spmv_e(Gv.zb, r);
                p = z+beta*p;
                q = wz+beta*q;
                double s_pq2 = (double)p*q, z4 = 0, z5 = 0;
                gsum3(s_pq2, z4, z5);
                pq = s_pq2;
            }
            if (own) Gv.db[gs] = d;
            tree_barrier(bar, G, wg, gen, &s_fail);
        }
        inner_total += it;
        if (s_fail) break;
    }
    if (own && !s_fail) x[gi] = zero_rhs ? 0.0 : x[gi]+Dfin;
    if (wg == 0 && threadIdx.x == 0) {
        sc->res_norm = sqrt(rr64);
        sc->rhs_norm = sqrt(bb);
        sc->iters_inner = inner_total;
        sc->iters_outer = outer;
        sc->total_inner += inner_total;
        if (!converged) atomicOr(&sc->error, E_NOT_CONVERGED);
    }
    if (threadIdx.x == 0 && s_fail) atomicOr(&sc->error, E_SYNC_TIMEOUT);
}

template <int W> static int agg_groups_of()
{
    static int cached = 0;
    if (!cached) {
        int dev = 0, ncu = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_solve_agg<W>, W*WAVE, 0) != hipSuccess || per < 1)
            per = 1;
        cached = ncu*per;
    }
    return cached;
}

int agg_solver_max_groups(int waves)
{
    return waves == AGG_WAVES_WIDE ? agg_groups_of<AGG_WAVES_WIDE>() : agg_groups_of<AGG_WAVES_REG>();
}

hipError_t solve_agg(const BsrView& A, const AggView& G, const double* b, double* x, SolverWork& w,
                     DevScalars* sc, const SolveJob* jobs, int nwg, const SolveOpts& o, hipStream_t s)
{
    const int W = o.agg_waves == AGG_WAVES_WIDE ? AGG_WAVES_WIDE : AGG_WAVES_REG;
    if (nwg > agg_solver_max_groups(W)) return hipErrorCooperativeLaunchTooLarge;
    BsrView Av = A;
    AggView Gv = G;
    SolverWork wv = w;
    SolveOpts ov = o;
    // no stale round tags in any instance's partial records
    launch_fill_double(G.part, 0.0, G.part_len, s);
    void* args[] = {&Av, &Gv, (void*)&b, (void*)&x, &wv, &sc, &ov, (void*)&jobs};
    const void* fn = W == AGG_WAVES_WIDE ? (const void*)k_solve_agg<AGG_WAVES_WIDE>
                                         : (const void*)k_solve_agg<AGG_WAVES_REG>;
    return launch_persistent(fn, nwg, W*WAVE, args, o, agg_solver_max_groups(W), s);
}

}  // namespace lfd
