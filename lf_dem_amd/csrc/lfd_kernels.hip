// lfd_kernels.hip -- HIP kernels for gfx950 (MI355X) of the Stokes-resistance
// step: cell binning, LDS-staged neighbour search with wavefront ballot
// compaction, pair lifecycle, deterministic 6x6-block assembly, block SpMV and
// the Krylov solvers.  Wavefronts are 64 lanes wide.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include "lfd_internal.hpp"
#include "lfd_device.hpp"

namespace cg = cooperative_groups;

namespace lfd {

__device__ inline void raise_error(DevScalars* sc, int bit) { atomicOr(&sc->error, bit); }

__device__ inline unsigned lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ inline unsigned popc_below(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ inline void atomic_max_nonneg(unsigned long long* addr, double v)
{
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// Max of non-negative doubles into sc[inst].*field over a whole wavefront (all
// 64 lanes must reach this call; lanes with inst < 0 or v < 0 contribute
// nothing).  When every contributing lane has the same instance the wave
// reduces first and issues one atomic, otherwise each lane issues its own.
template <unsigned long long DevScalars::*field>
__device__ inline void wave_max_into(DevScalars* sc, int inst, double v)
{
    const bool has = inst >= 0 && v >= 0;
    unsigned long long b = has ? (unsigned long long)__double_as_longlong(v) : 0ull;
    const unsigned long long m = __ballot(has);
    if (!m) return;
    const int lead = __builtin_ctzll(m);
    const int i0 = __shfl(inst, lead);
    if (__all(!has || inst == i0)) {
        for (int o = 1; o < WAVE; o <<= 1) {
            const unsigned long long t = __shfl_xor(b, o);
            b = t > b ? t : b;
        }
        if ((int)lane_id() == lead) atomicMax(&(sc[i0].*field), b);
    } else if (has) {
        atomicMax(&(sc[inst].*field), b);
    }
}

// ------------------------------------------------------------------ fills
__global__ void k_fill_int(int* p, int v, size_t n)
{
    size_t i = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_double(double* p, double v, size_t n)
{
    size_t i = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_copy_double(double* d, const double* s, size_t n)
{
    size_t i = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (i < n) d[i] = s[i];
}
void launch_fill_int(int* p, int v, size_t n, hipStream_t s)
{
    if (n) k_fill_int<<<(unsigned)((n+255)/256), 256, 0, s>>>(p, v, n);
}
void launch_fill_double(double* p, double v, size_t n, hipStream_t s)
{
    if (n) k_fill_double<<<(unsigned)((n+255)/256), 256, 0, s>>>(p, v, n);
}
void launch_copy_double(double* d, const double* src, size_t n, hipStream_t s)
{
    if (n) k_copy_double<<<(unsigned)((n+255)/256), 256, 0, s>>>(d, src, n);
}

// ------------------------------------------------------------------ scan
// Exclusive scan of in[0..count) into out[0..count], out[count] = total.
static constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T*SCAN_I;

__device__ inline int block_exclusive_scan(int v, int* lds, int& total)
{
    // wave-inclusive scan by shuffles, then scan of the wave totals in LDS
    unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE;
    int x = v;
    for (int o = 1; o < WAVE; o <<= 1) {
        int y = __shfl_up(x, o, WAVE);
        if ((int)lane >= o) x += y;
    }
    if (lane == WAVE-1) lds[wid] = x;
    __syncthreads();
    int nw = blockDim.x/WAVE;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < nw; k++) { int t = lds[k]; lds[k] = acc; acc += t; }
        lds[nw] = acc;
    }
    __syncthreads();
    int r = x-v+lds[wid];
    total = lds[nw];
    __syncthreads();
    return r;
}

__global__ void k_scan_reduce(const int* in, const int* count_dev, int cap, int* sums)
{
    int count = count_dev ? min(*count_dev, cap) : cap;
    int base = blockIdx.x*SCAN_TILE+threadIdx.x*SCAN_I;
    int acc = 0;
    for (int k = 0; k < SCAN_I; k++) { int i = base+k; if (i < count) acc += in[i]; }
    __shared__ int lds[SCAN_T/WAVE+1];
    int total;
    block_exclusive_scan(acc, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
__global__ void k_scan_top(int* sums, int nb)
{
    __shared__ int lds[SCAN_T/WAVE+1];
    int carry = 0;
    for (int base = 0; base < nb; base += SCAN_T) {
        int i = base+threadIdx.x;
        int v = i < nb ? sums[i] : 0;
        int total;
        int e = block_exclusive_scan(v, lds, total);
        if (i < nb) sums[i] = e+carry;
        carry += total;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}
__global__ void k_scan_down(const int* in, int* out, const int* count_dev, int cap, const int* sums)
{
    int count = count_dev ? min(*count_dev, cap) : cap;
    int base = blockIdx.x*SCAN_TILE+threadIdx.x*SCAN_I;
    int v[SCAN_I], acc = 0;
    for (int k = 0; k < SCAN_I; k++) { int i = base+k; v[k] = i < count ? in[i] : 0; acc += v[k]; }
    __shared__ int lds[SCAN_T/WAVE+1];
    int total;
    int e = block_exclusive_scan(acc, lds, total)+sums[blockIdx.x];
    for (int k = 0; k < SCAN_I; k++) {
        int i = base+k;
        if (i < count) out[i] = e;
        e += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[count] = sums[gridDim.x];
}
// Whole scan in one workgroup for arrays up to SCAN1_MAX entries: tiles of
// SCAN1_T*SCAN1_I entries with a running carry.  out2 (optional) receives a
// second copy of out[0..count) (a cursor array); it may alias in.
static constexpr int SCAN1_T = 1024, SCAN1_I = 32, SCAN1_MAX = 1 << 17;
__global__ __launch_bounds__(SCAN1_T) void k_scan_single(const int* in, int* out, int* out2, const int* count_dev,
                                                         int cap)
{
    __shared__ int lds[SCAN1_T/WAVE+1];
    const int count = count_dev ? min(*count_dev, cap) : cap;
    // 16-byte accesses when every array allows them (a thread's run of
    // SCAN1_I entries is one 128-byte line)
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out) |
                       reinterpret_cast<uintptr_t>(out2)) & 15) == 0;
    int carry = 0;
    for (int base = 0; base < count; base += SCAN1_T*SCAN1_I) {
        const int b0 = base+threadIdx.x*SCAN1_I;
        const bool full = vec && b0+SCAN1_I <= count;
        int v[SCAN1_I], acc = 0;
        if (full) {
#pragma unroll
            for (int k = 0; k < SCAN1_I/4; k++) {
                const int4 t = reinterpret_cast<const int4*>(in+b0)[k];
                v[4*k] = t.x; v[4*k+1] = t.y; v[4*k+2] = t.z; v[4*k+3] = t.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < SCAN1_I; k++) v[k] = b0+k < count ? in[b0+k] : 0;
        }
#pragma unroll
        for (int k = 0; k < SCAN1_I; k++) acc += v[k];
        int total;
        int e = block_exclusive_scan(acc, lds, total)+carry;
#pragma unroll
        for (int k = 0; k < SCAN1_I; k++) {             // v becomes the exclusive prefix
            const int t = v[k];
            v[k] = e;
            e += t;
        }
        if (full) {
#pragma unroll
            for (int k = 0; k < SCAN1_I/4; k++) {
                const int4 t = make_int4(v[4*k], v[4*k+1], v[4*k+2], v[4*k+3]);
                reinterpret_cast<int4*>(out+b0)[k] = t;
                if (out2) reinterpret_cast<int4*>(out2+b0)[k] = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < SCAN1_I; k++)
                if (b0+k < count) { out[b0+k] = v[k]; if (out2) out2[b0+k] = v[k]; }
        }
        carry += total;
    }
    if (threadIdx.x == 0) out[count] = carry;
}
__global__ void k_copy_int(int* d, const int* s, int n);
void exclusive_scan_int(const int* in, int* out, const int* count_dev, int cap, int* tmp, hipStream_t s,
                        int* out2)
{
    if (cap <= SCAN1_MAX) {
        k_scan_single<<<1, SCAN1_T, 0, s>>>(in, out, out2, count_dev, cap);
        return;
    }
    int nb = (cap+SCAN_TILE-1)/SCAN_TILE;
    if (nb < 1) nb = 1;
    k_scan_reduce<<<nb, SCAN_T, 0, s>>>(in, count_dev, cap, tmp);
    k_scan_top<<<1, SCAN_T, 0, s>>>(tmp, nb);
    k_scan_down<<<nb, SCAN_T, 0, s>>>(in, out, count_dev, cap, tmp);
    if (out2) k_copy_int<<<(cap+255)/256, 256, 0, s>>>(out2, out, cap);
}

// ------------------------------------------------------------------ cells
// Global cell label of every particle slot (-1 for a free slot).
__global__ void k_box_labels(const double4* pos, const int* inst_of, int n, const InstConst* IC, int* box_of,
                             DevScalars* sc)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n) return;
    const int k = inst_of[i];
    if (k < 0) { box_of[i] = -1; return; }
    const InstConst& ic = IC[k];
    double4 q = pos[i];
    double p[3] = {q.x, q.y, q.z};
    int l = which_box(p, ic.bx, ic.by, ic.bz, (unsigned)ic.yz, (unsigned)ic.nz, (unsigned)ic.nbox);
    if (l < 0) { raise_error(sc+k, E_BOX_OOB); l = 0; }
    box_of[i] = ic.coff+l;
}
void launch_box_labels(const double4* pos, const int* inst_of, int n, const InstConst* IC, int* box_of,
                       DevScalars* sc, hipStream_t s)
{
    k_box_labels<<<(n+255)/256, 256, 0, s>>>(pos, inst_of, n, IC, box_of, sc);
}

__global__ void k_bin_count(const int* box_of, int n, int* count)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n && box_of[i] >= 0) atomicAdd(&count[box_of[i]], 1);
}
__global__ void k_bin_scatter(const int* box_of, int n, int* cursor, int* mem)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n && box_of[i] >= 0) mem[atomicAdd(&cursor[box_of[i]], 1)] = i;
}
// members of a cell in ascending particle order (the order of Box's std::set<int>)
__global__ void k_bin_sort(const int* start, int nbox, int* mem)
{
    int b = blockIdx.x*blockDim.x+threadIdx.x;
    if (b >= nbox) return;
    int s0 = start[b], s1 = start[b+1];
    for (int a = s0+1; a < s1; a++) {
        int x = mem[a], c = a-1;
        while (c >= s0 && mem[c] > x) { mem[c+1] = mem[c]; c--; }
        mem[c+1] = x;
    }
}
__global__ void k_copy_int(int* d, const int* s, int n)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n) d[i] = s[i];
}
void launch_bin(const int* box_of, int n, int nbox, int* count, int* start, int* mem, int* cursor,
                int* scan_tmp, hipStream_t s)
{
    launch_fill_int(count, 0, nbox, s);
    k_bin_count<<<(n+255)/256, 256, 0, s>>>(box_of, n, count);
    exclusive_scan_int(count, start, nullptr, nbox, scan_tmp, s, cursor);
    k_bin_scatter<<<(n+255)/256, 256, 0, s>>>(box_of, n, cursor, mem);
    k_bin_sort<<<(nbox+255)/256, 256, 0, s>>>(start, nbox, mem);
}

// BoxSet::updateNeighbors (BoxSet.cpp:408-442): moving neighbours of top and
// bottom cells from the 16 probing positions, with the current LE shift.
// One thread per (cell, probe) computes the probe's cell label; a second pass
// appends the new labels to each cell's list (top probes first).
__global__ void k_probe_labels(CellArrays ca, const InstConst* IC, const DevScalars* sc, int* probe)
{
    int t = blockIdx.x*blockDim.x+threadIdx.x;
    int b = t >> 5, k = t & 31;
    if (b >= ca.ncell) return;
    const int inst = ca.cell_inst[b];
    int l = -1;
    if (inst >= 0) {
        const InstConst& ic = IC[inst];
        const int lb = b-ic.coff, ty = ca.ctype[b];
        const bool top = k < 16;
        const bool want = ic.boxed && ((top && (ty == 2 || ty == 3)) || (!top && (ty == 1 || ty == 3)));
        if (want) {
            unsigned ix = lb/ic.yz, iy = (lb%ic.yz)/ic.nz, iz = lb%ic.nz;
            double c[3] = {ic.bx*(ix+0.5), ic.by*(iy+0.5), ic.bz*(iz+0.5)};
            double sd[3] = {sc[inst].shear_disp[0], sc[inst].shear_disp[1], sc[inst].shear_disp[2]};
            const double* pr = top ? ic.probe_top[k] : ic.probe_bot[k-16];
            double q[3] = {c[0]+pr[0], c[1]+pr[1], c[2]+pr[2]};
            periodize(ic.C.L, sd, q);
            l = which_box(q, ic.bx, ic.by, ic.bz, (unsigned)ic.yz, (unsigned)ic.nz, (unsigned)ic.nbox);
            if (l < 0) raise_error(const_cast<DevScalars*>(sc)+inst, E_BOX_OOB);
            else if (l+ic.coff == b) l = -1;
            else l += ic.coff;
        }
    }
    probe[t] = l;
}
// One wave per cell: the static neighbours sit one per lane, the 32 probe
// labels are appended in probe order when no lane already holds them.
__global__ __launch_bounds__(256) void k_moving_merge(CellArrays ca, const InstConst* IC, const int* probe)
{
    const int b = blockIdx.x*(256/WAVE)+(int)(threadIdx.x/WAVE);
    const int lane = threadIdx.x & (WAVE-1);
    if (b >= ca.ncell) return;
    const int inst = ca.cell_inst[b];
    if (inst < 0) return;
    int cnt = ca.nstatic[b];
    int* nb = ca.nb+64*(size_t)b;
    if (ca.ctype[b] != 0 && IC[inst].boxed) {
        int mine = lane < cnt ? nb[lane] : -1;
        const int pl = lane < 32 ? probe[32*(size_t)b+lane] : -1;
        for (int k = 0; k < 32; k++) {
            const int l = __shfl(pl, k);
            if (l < 0) continue;
            const bool dup = __any(mine == l);
            if (!dup && cnt < 64) {
                if (lane == cnt) mine = l;
                cnt++;
            }
        }
        if (lane < cnt) nb[lane] = mine;
    }
    if (lane == 0) ca.nnb[b] = cnt;
}
void launch_moving_neighbors(CellArrays ca, const InstConst* IC, const DevScalars* sc, int* probe, hipStream_t s)
{
    if (ca.ncell == 0) return;
    k_probe_labels<<<(unsigned)((32*(size_t)ca.ncell+255)/256), 256, 0, s>>>(ca, IC, sc, probe);
    k_moving_merge<<<(ca.ncell+3)/4, 256, 0, s>>>(ca, IC, probe);
}

// ------------------------------------------------------------------ neighbour search
// One workgroup per home cell.  Candidates (snapshot members of the cell's
// neighbourhood, current positions) are staged in LDS in chunks; each wave
// takes one home particle, its 64 lanes test 64 candidates at a time, and hits
// are compacted with a ballot + masked popcount (System::checkNewInteraction,
// System.cpp:1328-1353, over BoxSet::neighborhood, Box.cpp:45-87).
static constexpr int FP_BLOCK = 256, FP_CHUNK = 512;

__global__ __launch_bounds__(FP_BLOCK) void k_find_pairs(const double4* __restrict__ pos, CellArrays ca,
    const InstConst* IC, const int* __restrict__ home_start, const int* __restrict__ home_mem,
    const int* __restrict__ snap_start, const int* __restrict__ snap_mem, DevScalars* sc,
    int* __restrict__ found, int* __restrict__ found_cnt, int K, int only)
{
    __shared__ int cells[64];
    __shared__ int coff[65];
    __shared__ int ncell;
    __shared__ int cidx[FP_CHUNK];
    __shared__ double4 cpos[FP_CHUNK];
    const int cell = blockIdx.x;
    const int h0 = home_start[cell], nh = home_start[cell+1]-h0;
    if (nh == 0) return;
    const int inst = ca.cell_inst[cell];
    if (inst < 0 || (only >= 0 && inst != only)) return;
    const Consts& C = IC[inst].C;
    const unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE, nwave = FP_BLOCK/WAVE;
    if (threadIdx.x == 0) {
        int nc = ca.nnb[cell], acc = 0;
        for (int c = 0; c < nc; c++) {
            int b = ca.nb[64*(size_t)cell+c];
            cells[c] = b;
            coff[c] = acc;
            acc += snap_start[b+1]-snap_start[b];
        }
        coff[nc] = acc;
        ncell = nc;
    }
    __syncthreads();
    const int nc = ncell;
    const int total = coff[nc];
    double sd[3] = {sc[inst].shear_disp[0], sc[inst].shear_disp[1], sc[inst].shear_disp[2]};
    const double L[3] = {C.L[0], C.L[1], C.L[2]}, half[3] = {C.half[0], C.half[1], C.half[2]};
    const double lmg = C.lub_max_gap;
    for (int base = 0; base < total; base += FP_CHUNK) {
        int nchunk = min(FP_CHUNK, total-base);
        for (int t = threadIdx.x; t < nchunk; t += FP_BLOCK) {
            int gi = base+t, c = 0;
            while (c+1 < nc && coff[c+1] <= gi) c++;
            int j = snap_mem[snap_start[cells[c]]+(gi-coff[c])];
            cidx[t] = j;
            cpos[t] = pos[j];
        }
        __syncthreads();
        for (int h = wid; h < nh; h += nwave) {
            const int i = home_mem[h0+h];
            const double4 xi = pos[i];
            int cnt = found_cnt[i];
            for (int t0 = 0; t0 < nchunk; t0 += WAVE) {
                int t = t0+(int)lane;
                bool hit = false;
                int j = -1;
                if (t < nchunk) {
                    j = cidx[t];
                    if (j > i) {
                        double4 xj = cpos[t];
                        double d[3] = {xj.x-xi.x, xj.y-xi.y, xj.z-xi.z};
                        periodize_diff(L, half, sd, d);
                        double sq = d[0]*d[0]+d[1]*d[1]+d[2]*d[2];
                        double range = lub_range(xi.w, xj.w, lmg);
                        hit = sq < range*range;
                    }
                }
                unsigned long long mask = __ballot(hit);
                if (hit) {
                    int slot = cnt+(int)popc_below(mask);
                    if (slot < K) found[(size_t)i*K+slot] = j;
                }
                cnt += __popcll(mask);
            }
            if (lane == 0) {
                found_cnt[i] = cnt;
                if (cnt > K) raise_error(sc+inst, E_FOUND_OVERFLOW);
            }
        }
        __syncthreads();
    }
}
void launch_find_pairs(const double4* pos, int n, CellArrays ca, const InstConst* IC, const int* home_start,
                       const int* home_mem, const int* snap_start, const int* snap_mem, DevScalars* sc,
                       int* found, int* found_cnt, int K, int only, hipStream_t s)
{
    launch_fill_int(found_cnt, 0, n, s);
    if (ca.ncell > 0)
        k_find_pairs<<<ca.ncell, FP_BLOCK, 0, s>>>(pos, ca, IC, home_start, home_mem, snap_start, snap_mem, sc,
                                                   found, found_cnt, K, only);
}

// ------------------------------------------------------------------ merge new pairs
// Row i: union of the existing pairs (i, j) and the pairs found this pass.
__global__ void k_merge_count(int n, const int* __restrict__ old_p1, const int* old_row_ptr, int* found,
                              const int* found_cnt, int K, int* cnt)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n) return;
    int nf = min(found_cnt[i], K);
    int* f = found+(size_t)i*K;
    for (int a = 1; a < nf; a++) {            // insertion sort of the found partners
        int x = f[a], c = a-1;
        while (c >= 0 && f[c] > x) { f[c+1] = f[c]; c--; }
        f[c+1] = x;
    }
    int o = old_row_ptr[i], oe = old_row_ptr[i+1], a = 0, m = 0;
    while (o < oe || a < nf) {
        if (a >= nf) { o++; }
        else if (o >= oe) { a++; }
        else {
            int jo = old_p1[o], jf = f[a];
            if (jo == jf) { o++; a++; }
            else if (jo < jf) o++;
            else a++;
        }
        m++;
    }
    cnt[i] = m;
}
// Source of every slot of the merged rows: an old pair index (>= 0) or
// ~(slot in the found list) for a new pair.
__global__ void k_merge_index(int n, const int* __restrict__ old_p1, const int* __restrict__ old_row_ptr,
                              const int* __restrict__ found, const int* __restrict__ found_cnt, int K,
                              const int* __restrict__ new_row_ptr, int* __restrict__ src, int pcap, BatchScalars* err)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n) return;
    int nf = min(found_cnt[i], K);
    const int* f = found+(size_t)i*K;
    int o = old_row_ptr[i], oe = old_row_ptr[i+1], a = 0;
    int w = new_row_ptr[i];
    while (o < oe || a < nf) {
        int v;
        if (a >= nf) v = o++;
        else if (o >= oe) { v = ~(i*K+a); a++; }
        else {
            int jo = old_p1[o], jf = f[a];
            if (jo == jf) { v = o++; a++; }
            else if (jo < jf) v = o++;
            else { v = ~(i*K+a); a++; }
        }
        if (w >= pcap) { atomicOr(&err->error, E_PAIR_OVERFLOW); return; }
        src[w++] = v;
    }
}
// One thread per merged slot: copy the old record or create the new pair.
__global__ __launch_bounds__(256) void k_merge_copy(const PairRec* __restrict__ old_pairs, const int* __restrict__ src,
    const int* __restrict__ found, int K, const int* __restrict__ total, PairRec* __restrict__ new_pairs,
    const double4* __restrict__ pos, const int* __restrict__ inst_of, const InstConst* IC, const DevScalars* sc,
    int pcap)
{
    int w = blockIdx.x*blockDim.x+threadIdx.x;
    if (w >= min(*total, pcap)) return;
    int v = src[w];
    if (v >= 0) { new_pairs[w] = old_pairs[v]; return; }
    v = ~v;
    int i = v/K, j = found[v];
    const int inst = inst_of[i];
    const Consts& C = IC[inst].C;
    double sd[3] = {sc[inst].shear_disp[0], sc[inst].shear_disp[1], sc[inst].shear_disp[2]};
    double4 xi = pos[i], xj = pos[j];
    double pi[3] = {xi.x, xi.y, xi.z}, pj[3] = {xj.x, xj.y, xj.z};
    PairRec q;
    pair_create(q, i, j, pi, pj, xi.w, xj.w, C, sd);
    new_pairs[w] = q;
}
__global__ void k_set_npairs_from(BatchScalars* bs, const int* total, int pcap)
{
    int t = *total;
    if (t > pcap) { atomicOr(&bs->error, E_PAIR_OVERFLOW); t = pcap; }
    bs->npairs = t;
}
void launch_merge_pairs(int n, const PairRec* old_pairs, const int* old_p1, const int* old_row_ptr, const int* found,
                        const int* found_cnt, int K, int* cnt, int* new_row_ptr, int* src, PairRec* new_pairs,
                        const double4* pos, const int* inst_of, const InstConst* IC, const DevScalars* sc,
                        BatchScalars* bs, int pcap, int* scan_tmp, hipStream_t s)
{
    k_merge_count<<<(n+255)/256, 256, 0, s>>>(n, old_p1, old_row_ptr, const_cast<int*>(found), found_cnt, K, cnt);
    exclusive_scan_int(cnt, new_row_ptr, nullptr, n, scan_tmp, s);
    k_merge_index<<<(n+255)/256, 256, 0, s>>>(n, old_p1, old_row_ptr, found, found_cnt, K, new_row_ptr, src, pcap, bs);
    k_merge_copy<<<(pcap+255)/256, 256, 0, s>>>(old_pairs, src, found, K, new_row_ptr+n, new_pairs, pos, inst_of,
                                                IC, sc, pcap);
    k_set_npairs_from<<<1, 1, 0, s>>>(bs, new_row_ptr+n, pcap);
}

// ------------------------------------------------------------------ pair update
// System::updateInteractions (System.cpp:1375-1403) without the swap-with-last
// removal: every pair is updated once and removals are compacted afterwards.
__global__ __launch_bounds__(256) void k_update_pairs(PairRec* pairs, const double4* __restrict__ pos,
    const double* __restrict__ vel6, const int* __restrict__ inst_of, const InstConst* IC, DevScalars* sc,
    const BatchScalars* bs, int* keep, int pcap, int in_predictor, double dt_override, int only)
{
    int k = blockIdx.x*blockDim.x+threadIdx.x;
    int np = bs->npairs;
    if (blockIdx.x*blockDim.x >= min(np, pcap)) return;      // whole workgroup beyond the pairs
    const bool valid = k < np && k < pcap;
    const int inst = valid ? inst_of[pairs[k].p0] : -1;
    const bool work = valid && inst >= 0 && only != -2 && (only < 0 || inst == only);
    double slide = -1.0;                       // squared sliding velocity of a frictional contact
    if (valid && !work) keep[k] = inst >= 0 ? 1 : 0;
    if (work) {
        PairRec q = pairs[k];
        const Consts& C = IC[inst].C;
        DevScalars* si = sc+inst;
        double4 x0 = pos[q.p0], x1 = pos[q.p1];
        double a[3] = {x0.x, x0.y, x0.z}, b[3] = {x1.x, x1.y, x1.z};
        double sd[3] = {si->shear_disp[0], si->shear_disp[1], si->shear_disp[2]};
        double dt = dt_override >= 0 ? dt_override : si->dt;
        const double vd[3] = {si->flow.vel_diff[0], si->flow.vel_diff[1], si->flow.vel_diff[2]};
        int drop = pair_update(q, a, b, x0.w, x1.w, vel6+6*(size_t)q.p0, vel6+6*(size_t)q.p1, C, vd, sd, dt,
                               in_predictor);
        pairs[k] = q;
        keep[k] = drop ? 0 : 1;
        if (C.friction && (q.flags & F_CONTACT)) slide = q.rsv_sq;
    }
    wave_max_into<&DevScalars::max_slide_bits>(sc, work ? inst : -1, slide);
}
// per-instance resets: bit 0 max sliding velocity, bit 1 max velocity
__global__ void k_reset_inst(DevScalars* sc, int ninst, int what, int only)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= ninst || (only >= 0 && k != only)) return;
    if (what & 1) sc[k].max_slide_bits = 0;
    if (what & 2) sc[k].max_vel_bits = 0;
}
void launch_update_pairs(PairRec* pairs, const double4* pos, const double* vel6, const int* inst_of,
                         const InstConst* IC, DevScalars* sc, int ninst, const BatchScalars* bs, int* keep,
                         int pcap, int in_predictor, double dt_override, int only, hipStream_t s)
{
    if (only != -2) k_reset_inst<<<(ninst+63)/64, 64, 0, s>>>(sc, ninst, 1, only);
    k_update_pairs<<<(pcap+255)/256, 256, 0, s>>>(pairs, pos, vel6, inst_of, IC, sc, bs, keep, pcap, in_predictor,
                                                  dt_override, only);
}

__global__ void k_compact(const PairRec* in, PairRec* out, const int* keep, const int* ks, const BatchScalars* bs)
{
    int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= bs->npairs) return;
    if (keep[k]) out[ks[k]] = in[k];
}
__global__ void k_set_npairs(BatchScalars* bs, const int* ks)
{
    bs->npairs = ks[bs->npairs];
}
void launch_compact_pairs(const PairRec* in, PairRec* out, const int* keep, int* keep_scan, int n,
                          int* row_ptr, int* p0, int* p1, int* tptr, int* tidx, int* tcnt, BatchScalars* bs,
                          int pcap, int* scan_tmp, hipStream_t s)
{
    exclusive_scan_int(keep, keep_scan, &bs->npairs, pcap, scan_tmp, s);
    k_compact<<<(pcap+255)/256, 256, 0, s>>>(in, out, keep, keep_scan, bs);
    k_set_npairs<<<1, 1, 0, s>>>(bs, keep_scan);
    launch_pair_index(out, n, row_ptr, p0, p1, tptr, tidx, tcnt, const_cast<int*>(keep), bs, pcap, scan_tmp, s);
}

// ------------------------------------------------------------------ pair index
__global__ void k_pair_arrays(const PairRec* pairs, int* p0, int* p1, int* tcnt, const BatchScalars* sc)
{
    int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= sc->npairs) return;
    int a = pairs[k].p0, b = pairs[k].p1;
    p0[k] = a; p1[k] = b;
    atomicAdd(&tcnt[b], 1);
}
__global__ void k_row_ptr(const int* p0, int n, int* row_ptr, const BatchScalars* sc)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i > n) return;
    int lo = 0, hi = sc->npairs;          // lower_bound of i in the sorted p0
    while (lo < hi) { int m = (lo+hi) >> 1; if (p0[m] < i) lo = m+1; else hi = m; }
    row_ptr[i] = lo;
}
__global__ void k_tfill(const int* p1, int* cursor, int* tmp, const BatchScalars* sc)
{
    int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= sc->npairs) return;
    tmp[atomicAdd(&cursor[p1[k]], 1)] = k;
}
// lower incidence lists in ascending pair index (= ascending p0): one thread
// per entry places it by its rank within its list
__global__ void k_tsort(const int* __restrict__ tptr, const int* __restrict__ p1, const int* __restrict__ tmp,
                        int* __restrict__ tidx, const BatchScalars* sc)
{
    const int a = blockIdx.x*blockDim.x+threadIdx.x;
    if (a >= sc->npairs) return;
    const int x = tmp[a], j = p1[x];
    const int s0 = tptr[j], s1 = tptr[j+1];
    int rank = 0;
    for (int q = s0; q < s1; q++) rank += tmp[q] < x;
    tidx[s0+rank] = x;
}
void launch_pair_index(const PairRec* pairs, int n, int* row_ptr, int* p0, int* p1, int* tptr, int* tidx,
                       int* tcnt, int* tmp, const BatchScalars* sc, int pcap, int* scan_tmp, hipStream_t s)
{
    launch_fill_int(tcnt, 0, n, s);
    k_pair_arrays<<<(pcap+255)/256, 256, 0, s>>>(pairs, p0, p1, tcnt, sc);
    k_row_ptr<<<(n+1+255)/256, 256, 0, s>>>(p0, n, row_ptr, sc);
    exclusive_scan_int(tcnt, tptr, nullptr, n, scan_tmp, s, tcnt);
    k_tfill<<<(pcap+255)/256, 256, 0, s>>>(p1, tcnt, tmp, sc);
    k_tsort<<<(pcap+255)/256, 256, 0, s>>>(tptr, p1, tmp, tidx, sc);
}

// ------------------------------------------------------------------ assembly
__device__ inline void uinf(const Flow& F, const double* x, double* u)
{
    // System::computeUInf (System.cpp:2357-2367)
    double e[3], w[3];
    symdot(F.E, x, e);
    cross3(F.omega, x, w);
    u[0] = e[0]+w[0]; u[1] = e[1]+w[1]; u[2] = e[2]+w[2];
}

// Right-hand-side pieces of one pair for one of its particles (side 0 = p0,
// side 1 = p1), accumulated in LF_DEM's component order: contact
// (Contact.cpp:411-437), dashpot R_FU^dash.U_inf (ContactDashpot.cpp:342-441),
// hydro R_FE:E_inf (Lubrication.cpp:332-403), repulsion (RepulsiveForce.cpp:101-105).
struct RhsAcc { double Fc[3], Tc[3], Fd[3], Td[3], Fh[3], Th[3], Fr[3]; };

__device__ inline void pair_rhs_side(const PairRec& q, int side, double a0, double a1, const double* x0,
                                     const double* x1, const Consts& C, const Flow& F, RhsAcc& R)
{
    const double* n = q.n;
    if (q.flags & F_CONTACT) {
        if (side == 0) for (int k = 0; k < 3; k++) R.Fc[k] += q.f_total[k];
        else for (int k = 0; k < 3; k++) R.Fc[k] -= q.f_total[k];
        if (pair_state(q.flags) >= 2) {
            double t[3];
            cross3(n, q.f_tan, t);
            double a = side == 0 ? a0 : a1;
            for (int k = 0; k < 3; k++) R.Tc[k] += a*t[k];
        }
        if (F.rate != 0) {
            DashCoef c;
            dashpot_coeffs(a0, a1, C.normal_coeff, C.tangential_coeff, c);
            double vi[3], vj[3], t[3], f[3];
            uinf(F, x0, vi);
            uinf(F, x1, vj);
            double zo = (double)q.zoff;
            for (int k = 0; k < 3; k++) vj[k] += zo*F.vel_diff[k];
            for (int k = 0; k < 3; k++) t[k] = c.XA[0]*vi[k]+c.XA[1]*vj[k];
            double d = -dot3(t, n);
            for (int k = 0; k < 3; k++) f[k] = d*n[k];
            if (C.tangential_coeff > 0) {
                double di = dot3(n, vi), dj = dot3(n, vj), ci[3];
                for (int k = 0; k < 3; k++) {
                    double av = vi[k]-n[k]*di, bv = vj[k]-n[k]*dj;
                    f[k] += -c.YA[0]*av-c.YA[1]*bv;
                }
                cross3(n, F.omega, ci);
                for (int k = 0; k < 3; k++) f[k] += -c.YB[0]*ci[k]-c.YB[2]*ci[k];
            }
            if (side == 0) for (int k = 0; k < 3; k++) R.Fd[k] += f[k];
            else for (int k = 0; k < 3; k++) R.Fd[k] -= f[k];
            if (C.tangential_coeff > 0) {
                double tq[3], t0[3];
                cross3(n, f, tq);
                for (int k = 0; k < 3; k++) t0[k] = a0*tq[k];
                if (side == 0) for (int k = 0; k < 3; k++) R.Td[k] += t0[k];
                else {
                    double ratio = a1/a0;
                    for (int k = 0; k < 3; k++) R.Td[k] += ratio*t0[k];
                }
            }
        }
    }
    if ((q.flags & F_LUB) && F.rate != 0) {
        LubConst c;
        lub_constants(a0, a1, c);
        double lc, llc;
        lub_coeffs(q.gap, C.lub_reduce, lc, llc);
        double XG[4], YG[4], YH[4];
        for (int j = 0; j < 4; j++) { XG[j] = c.XG[j]*lc; YG[j] = c.YG[j]*llc; YH[j] = c.YH[j]*llc; }
        double En[3];
        symdot(F.E, n, En);
        double nnE = dot3(n, En);
        if (C.tangential) {
            double YGa = side == 0 ? YG[0]+YG[2] : YG[1]+YG[3];
            double cGE = side == 0 ? (XG[0]+XG[2]-2*YGa)*nnE : (XG[1]+XG[3]-2*YGa)*nnE;
            double cHE = side == 0 ? YH[0]+YH[2] : YH[1]+YH[3];
            double gy = YGa*2, x[3];
            for (int k = 0; k < 3; k++) R.Fh[k] += cGE*n[k]+gy*En[k];
            cross3(En, n, x);
            double h = -2*cHE;
            for (int k = 0; k < 3; k++) R.Th[k] += h*x[k];
        } else {
            double cGE = side == 0 ? (XG[0]+XG[2])*nnE : (XG[1]+XG[3])*nnE;
            for (int k = 0; k < 3; k++) R.Fh[k] += cGE*n[k];
        }
    }
    if (C.repulsive) {
        for (int k = 0; k < 3; k++) {
            double rf = -q.rep_norm*n[k];
            if (side == 0) R.Fr[k] += rf; else R.Fr[k] -= rf;
        }
    }
}

// Lower triangle (row r >= col c) of the symmetric 6x6 block of a packed
// DBlock, as laid out by StokesSolver::insertDBlockValues (StokesSolver.cpp:154-184),
// packed row by row: index r*(r+1)/2 + c.
__device__ inline void dblock_lower_packed(const double* b, double* Lp)
{
    Lp[0] = b[0];
    Lp[1] = b[1]; Lp[2] = b[5];
    Lp[3] = b[2]; Lp[4] = b[6]; Lp[5] = b[8];
    Lp[6] = 0; Lp[7] = -b[3]; Lp[8] = -b[4]; Lp[9] = b[9];
    Lp[10] = b[3]; Lp[11] = 0; Lp[12] = -b[7]; Lp[13] = b[10]; Lp[14] = b[12];
    Lp[15] = b[4]; Lp[16] = b[7]; Lp[17] = 0; Lp[18] = b[11]; Lp[19] = b[13]; Lp[20] = b[14];
}
#define LP(r, c) Lp[(r)*((r)+1)/2+(c)]
// Block-Jacobi inverse of a packed DBlock through an in-place packed
// Cholesky factor; writes the dense row-major inverse in FP64 and FP32.
__device__ inline bool dblock_inverse(const double* b, double* inv64, float* inv32)
{
    double Lp[21];
    dblock_lower_packed(b, Lp);
    #pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = LP(j, j);
        #pragma unroll
        for (int k = 0; k < j; k++) d -= LP(j, k)*LP(j, k);
        if (!(d > 0)) return false;
        d = sqrt(d);
        LP(j, j) = d;
        #pragma unroll
        for (int i = j+1; i < 6; i++) {
            double v = LP(i, j);
            #pragma unroll
            for (int k = 0; k < j; k++) v -= LP(i, k)*LP(j, k);
            LP(i, j) = v/d;
        }
    }
    #pragma unroll
    for (int c = 0; c < 6; c++) {
        double y[6];
        #pragma unroll
        for (int r = 0; r < 6; r++) {
            double v = (r == c) ? 1.0 : 0.0;
            #pragma unroll
            for (int k = 0; k < r; k++) v -= LP(r, k)*y[k];
            y[r] = v/LP(r, r);
        }
        #pragma unroll
        for (int r = 5; r >= 0; r--) {
            double v = y[r];
            #pragma unroll
            for (int k = r+1; k < 6; k++) v -= LP(k, r)*y[k];
            y[r] = v/LP(r, r);
        }
        #pragma unroll
        for (int r = 0; r < 6; r++) { inv64[6*r+c] = y[r]; inv32[6*r+c] = (float)y[r]; }
    }
    return true;
}
#undef LP

// Per pair: off-diagonal block (FP64 and FP32 copies), the pair's two
// diagonal contributions and its right-hand-side terms for both particles.
// pdb[30*k]: DBlock of p0 then of p1; prhs[42*k]: RhsAcc of p0 then of p1.
static constexpr int RHS_TERMS = 21;
__global__ __launch_bounds__(128) void k_assemble_pairs(const PairRec* __restrict__ pairs,
    const double4* __restrict__ pos, const int* __restrict__ inst_of, const InstConst* IC, const DevScalars* sc,
    const BatchScalars* bs, double* __restrict__ od64, float* __restrict__ od32, double* __restrict__ pdb,
    double* __restrict__ prhs, int pcap)
{
    int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= bs->npairs || k >= pcap) return;
    const PairRec q = pairs[k];
    const int inst = inst_of[q.p0];
    const Consts& C = IC[inst].C;
    const double4 x04 = pos[q.p0], x14 = pos[q.p1];
    double db0[15], db1[15], od[18];
    pair_blocks(q, x04.w, x14.w, C, db0, db1, od);
    for (int m = 0; m < 18; m++) { od64[18*(size_t)k+m] = od[m]; od32[18*(size_t)k+m] = (float)od[m]; }
    for (int m = 0; m < 15; m++) { pdb[30*(size_t)k+m] = db0[m]; pdb[30*(size_t)k+15+m] = db1[m]; }
    const double x0[3] = {x04.x, x04.y, x04.z}, x1[3] = {x14.x, x14.y, x14.z};
    const Flow F = sc[inst].flow;
    for (int side = 0; side < 2; side++) {
        RhsAcc R;
        for (int m = 0; m < 3; m++) {
            R.Fc[m] = 0; R.Tc[m] = 0; R.Fd[m] = 0; R.Td[m] = 0; R.Fh[m] = 0; R.Th[m] = 0; R.Fr[m] = 0;
        }
        pair_rhs_side(q, side, x04.w, x14.w, x0, x1, C, F, R);
        double* o = prhs+42*(size_t)k+RHS_TERMS*side;
        for (int m = 0; m < 3; m++) {
            o[m] = R.Fc[m]; o[3+m] = R.Tc[m]; o[6+m] = R.Fd[m]; o[9+m] = R.Td[m];
            o[12+m] = R.Fh[m]; o[15+m] = R.Th[m]; o[18+m] = R.Fr[m];
        }
    }
}

// Per particle: diagonal block = drag seed (System.cpp:558-575) plus the
// pair contributions in ascending partner order (System.cpp:1459-1471),
// block-Jacobi inverse, and the RHS summed per force component in the same
// order, then across components in LF_DEM's std::map order
// (StokesSolver.cpp:513-549, System.cpp:1873-1885).
// Per-particle sums of the pair terms.  Eight lanes per particle, lane c
// summing components c, c+8, ... of the 15 diagonal-block and 21 RHS values,
// each in ascending partner order (drag seed, then pairs (j, i) with j < i,
// then pairs (i, j)); then one lane per particle inverts the block and
// combines the RHS from the workgroup's LDS copy.
static constexpr int GD_LANES = 8, GD_BLOCK = 256, GD_PER = GD_BLOCK/GD_LANES, GD_VALS = 15+RHS_TERMS;
static constexpr int GD_UNROLL = 4;

__global__ __launch_bounds__(GD_BLOCK) void k_gather_diag_rhs(const double* __restrict__ pdb,
    const double* __restrict__ prhs, const double4* __restrict__ pos, int n, const int* __restrict__ inst_of,
    const InstConst* IC, const int* __restrict__ row_ptr, const int* __restrict__ tptr,
    const int* __restrict__ tidx, double* __restrict__ diag64, float* __restrict__ diag32,
    double* __restrict__ minv64, float* __restrict__ minv32, double* __restrict__ rhs,
    double* __restrict__ rhs_parts, int nparts, DevScalars* err)
{
    __shared__ double sv[GD_PER][GD_VALS+1];
    const int lp = threadIdx.x/GD_LANES, c = threadIdx.x%GD_LANES;
    const int i = blockIdx.x*GD_PER+lp;
    const int inst = i < n ? inst_of[i] : -1;
    if (inst >= 0) {
        const Consts& C = IC[inst].C;
        const double a = pos[i].w;
        const double FU = C.sd_coeff*a;
        const double TW = C.sd_coeff*(a*a*a)*4.0/3;
        double acc[5];
        #pragma unroll
        for (int u = 0; u < 5; u++) {
            const int q = c+GD_LANES*u;
            double v = 0;
            if (q == 0 || q == 5 || q == 8) v = FU;
            if (q == 9 || q == 12 || q == 14) v = TW;
            acc[u] = v;
        }
        // pairs in batches of GD_UNROLL: the loads of a batch are issued
        // together, the sums keep the order of the pairs
        const int t0 = tptr[i], t1 = tptr[i+1];
        for (int t = t0; t < t1; t += GD_UNROLL) {           // pairs (j, i), j < i
            double v[GD_UNROLL][5];
            #pragma unroll
            for (int j = 0; j < GD_UNROLL; j++) {
                const bool in = t+j < t1;
                const size_t k = in ? tidx[t+j] : 0;
                #pragma unroll
                for (int u = 0; u < 5; u++) {
                    const int q = c+GD_LANES*u;
                    v[j][u] = !in ? 0.0 : q < 15 ? pdb[30*k+15+q] : q < GD_VALS ? prhs[42*k+RHS_TERMS+(q-15)] : 0.0;
                }
            }
            #pragma unroll
            for (int j = 0; j < GD_UNROLL; j++)
                #pragma unroll
                for (int u = 0; u < 5; u++) if (t+j < t1) acc[u] += v[j][u];
        }
        const int k0 = row_ptr[i], k1 = row_ptr[i+1];
        for (int kb = k0; kb < k1; kb += GD_UNROLL) {        // pairs (i, j), j > i
            double v[GD_UNROLL][5];
            #pragma unroll
            for (int j = 0; j < GD_UNROLL; j++) {
                const bool in = kb+j < k1;
                const size_t k = in ? (size_t)(kb+j) : 0;
                #pragma unroll
                for (int u = 0; u < 5; u++) {
                    const int q = c+GD_LANES*u;
                    v[j][u] = !in ? 0.0 : q < 15 ? pdb[30*k+q] : q < GD_VALS ? prhs[42*k+(q-15)] : 0.0;
                }
            }
            #pragma unroll
            for (int j = 0; j < GD_UNROLL; j++)
                #pragma unroll
                for (int u = 0; u < 5; u++) if (kb+j < k1) acc[u] += v[j][u];
        }
        #pragma unroll
        for (int u = 0; u < 5; u++) {
            const int q = c+GD_LANES*u;
            if (q < GD_VALS) sv[lp][q] = acc[u];
            if (q < 15) { diag64[15*(size_t)i+q] = acc[u]; diag32[15*(size_t)i+q] = (float)acc[u]; }
        }
    } else if (i < n) {
        // free slot: unit diagonal, zero right-hand sides
        #pragma unroll
        for (int u = 0; u < 5; u++) {
            const int q = c+GD_LANES*u;
            const double v = (q == 0 || q == 5 || q == 8 || q == 9 || q == 12 || q == 14) ? 1.0 : 0.0;
            if (q < GD_VALS) sv[lp][q] = q < 15 ? v : 0.0;
            if (q < 15) { diag64[15*(size_t)i+q] = v; diag32[15*(size_t)i+q] = (float)v; }
        }
    }
    __syncthreads();
    if (threadIdx.x >= GD_PER) return;
    const int ip = blockIdx.x*GD_PER+threadIdx.x;
    if (ip >= n) return;
    const int ik = inst_of[ip];
    const double* D = sv[threadIdx.x];
    const double* R = sv[threadIdx.x]+15;
    double Dl[15];
    for (int m = 0; m < 15; m++) Dl[m] = D[m];
    if (!dblock_inverse(Dl, minv64+36*(size_t)ip, minv32+36*(size_t)ip) && ik >= 0) raise_error(err+ik, E_NOT_SPD);
    const size_t m6 = 6*(size_t)n, o6 = 6*(size_t)ip;
    if (ik < 0) {
        for (int m = 0; m < 6; m++) {
            rhs[o6+m] = 0.0;
            for (int p = 0; p < nparts; p++) rhs_parts[p*m6+o6+m] = 0.0;
        }
        return;
    }
    const Consts& C = IC[ik].C;
    for (int m = 0; m < 3; m++) {
        double F = 0, T = 0;
        F += R[m];                           // contact
        if (C.friction) T += R[3+m];
        F += R[6+m];                         // dashpot
        T += R[9+m];
        F += R[12+m];                        // hydro
        if (C.tangential) T += R[15+m];
        if (C.repulsive) F += R[18+m];       // repulsion
        rhs[6*(size_t)ip+m] = F;
        rhs[6*(size_t)ip+3+m] = T;
    }
    // right-hand sides of the velocity parts (System::setSolverRHS, System.cpp:1858-1871)
    for (int m = 0; m < 3 && nparts == 4; m++) {
        rhs_parts[o6+m] = R[m];                                   // contact
        rhs_parts[o6+3+m] = C.friction ? R[3+m] : 0.0;
        rhs_parts[m6+o6+m] = R[6+m];                              // dashpot
        rhs_parts[m6+o6+3+m] = R[9+m];
        rhs_parts[2*m6+o6+m] = R[12+m];                           // hydro
        rhs_parts[2*m6+o6+3+m] = C.tangential ? R[15+m] : 0.0;
        rhs_parts[3*m6+o6+m] = C.repulsive ? R[18+m] : 0.0;       // repulsion
        rhs_parts[3*m6+o6+3+m] = 0.0;
    }
    for (int m = 0; m < 3 && nparts == 2; m++) {
        // rate independent (contact + repulsion), rate proportional (dashpot + hydro)
        double F = 0, T = 0;
        F += R[m];
        if (C.friction) T += R[3+m];
        if (C.repulsive) F += R[18+m];
        rhs_parts[o6+m] = F;
        rhs_parts[o6+3+m] = T;
        F = 0; T = 0;
        F += R[6+m];
        T += R[9+m];
        F += R[12+m];
        if (C.tangential) T += R[15+m];
        rhs_parts[m6+o6+m] = F;
        rhs_parts[m6+o6+3+m] = T;
    }
}
void launch_assemble(const PairRec* pairs, const double4* pos, int n, const int* inst_of, const InstConst* IC,
                     DevScalars* sc, const BatchScalars* bs, const int* row_ptr, const int* tptr, const int* tidx,
                     double* diag64, float* diag32, double* od64, float* od32, double* minv64, float* minv32,
                     double* rhs, double* pdb, double* prhs, int pcap, double* rhs_parts, int nparts, hipStream_t s)
{
    k_assemble_pairs<<<(pcap+127)/128, 128, 0, s>>>(pairs, pos, inst_of, IC, sc, bs, od64, od32, pdb, prhs, pcap);
    k_gather_diag_rhs<<<(n+GD_PER-1)/GD_PER, GD_BLOCK, 0, s>>>(pdb, prhs, pos, n, inst_of, IC, row_ptr, tptr, tidx,
                                                               diag64, diag32, minv64, minv32, rhs, rhs_parts,
                                                               nparts, sc);
}

// Matrix handed over through the StokesSolver staging protocol: FP32 copy and
// block-Jacobi inverse of host-provided blocks.
__global__ void k_prepare_od(const double* od64, float* od32, const int* npairs, int pcap)
{
    size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    size_t m = (size_t)min(*npairs, pcap)*18;
    if (k < m) od32[k] = (float)od64[k];
}
__global__ __launch_bounds__(128) void k_prepare_diag(int n, const double* diag64, float* diag32, double* minv64, float* minv32,
                               DevScalars* err)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n) return;
    double D[15];
    for (int m = 0; m < 15; m++) { D[m] = diag64[15*(size_t)i+m]; diag32[15*(size_t)i+m] = (float)D[m]; }
    if (!dblock_inverse(D, minv64+36*(size_t)i, minv32+36*(size_t)i)) raise_error(err, E_NOT_SPD);
}
void launch_prepare_matrix(int n, const int* npairs, const double* diag64, float* diag32, const double* od64,
                           float* od32, double* minv64, float* minv32, int pcap, DevScalars* err, hipStream_t s)
{
    size_t m = (size_t)pcap*18;
    if (m) k_prepare_od<<<(unsigned)((m+255)/256), 256, 0, s>>>(od64, od32, npairs, pcap);
    k_prepare_diag<<<(n+127)/128, 128, 0, s>>>(n, diag64, diag32, minv64, minv32, err);
}

// ------------------------------------------------------------------ velocity parts and stress
// Stress slots, in LF_DEM's std::map order of stress components
// (Stress.cpp:16-73): M:E, GU of up to four velocity parts, the contact xF
// (rate control: spring + dashpot; stress control: its rate independent
// part), the rate proportional contact xF (stress control), the repulsive xF.
// Velocity parts: four (contact, dashpot, hydro, repulsion) or two (rate
// independent, rate proportional).
enum : int { SL_ME = 0, SL_GU = 1, SL_XFC = 5, SL_XFC_PROP = 6, SL_XFR = 7, NSLOT = 8 };
static constexpr int ST_VALS = 6*NSLOT;

__device__ inline void outer_sym3(const double* a, const double* b, double* o)
{
    // Sym2Tensor.h:256-264, (xx, xy, xz, yz, yy, zz)
    o[0] = a[0]*b[0];
    o[1] = 0.5*(a[0]*b[1]+a[1]*b[0]);
    o[2] = 0.5*(a[0]*b[2]+a[2]*b[0]);
    o[3] = 0.5*(a[1]*b[2]+a[2]*b[1]);
    o[4] = a[1]*b[1];
    o[5] = a[2]*b[2];
}
__device__ inline void outer3(const double* a, double* o)
{
    // Sym2Tensor.h:266-274
    o[0] = a[0]*a[0]; o[1] = a[0]*a[1]; o[2] = a[0]*a[2];
    o[3] = a[1]*a[2]; o[4] = a[1]*a[1]; o[5] = a[2]*a[2];
}
__device__ inline double ddot6(const double* a, const double* b)
{
    // Sym2Tensor.h:246-254
    return a[0]*b[0]+2*a[1]*b[1]+2*a[2]*b[2]+2*a[3]*b[3]+a[4]*b[4]+a[5]*b[5];
}

// Force on p0 from the dashpot at the given velocities
// (ContactDashpot::getForceOnP0 / getForceOnP0_nonaffine, ContactDashpot.cpp:342-424)
__device__ inline void dashpot_force(const PairRec& q, const DashCoef& c, const Consts& C, const Flow& F,
                                     const double* v0, const double* v1, const double* w0, const double* w1,
                                     int nonaffine, double* f)
{
    const double* n = q.n;
    double vi[3] = {v0[0], v0[1], v0[2]}, vj[3] = {v1[0], v1[1], v1[2]}, t[3];
    if (!nonaffine) {
        double zo = (double)q.zoff;
        for (int k = 0; k < 3; k++) vj[k] += zo*F.vel_diff[k];
    }
    for (int k = 0; k < 3; k++) t[k] = c.XA[0]*vi[k]+c.XA[1]*vj[k];
    double d = -dot3(t, n);
    for (int k = 0; k < 3; k++) f[k] = d*n[k];
    if (C.tangential_coeff > 0) {
        double di = dot3(n, vi), dj = dot3(n, vj), ci[3], cj[3];
        for (int k = 0; k < 3; k++) {
            double av = vi[k]-n[k]*di, bv = vj[k]-n[k]*dj;
            f[k] += -c.YA[0]*av-c.YA[1]*bv;
        }
        cross3(n, w0, ci);
        cross3(n, w1, cj);
        for (int k = 0; k < 3; k++) f[k] += -c.YB[0]*ci[k]-c.YB[2]*cj[k];
    }
}

// Per pair and side, the contribution to every stress slot
// (calcStressPerParticle, Stress.cpp:194-287):
//   M:E and GU of each velocity part for lubricated pairs
//   (Lubrication::addMEStresslet / addGUStresslet, Lubrication.cpp:587-650);
//   contact xF: rate control Contact::addUpStress at the total velocities
//   (Contact.cpp:439-458); stress control the spring part and the non-affine
//   dashpot part (rate independent) and the dashpot part at the rate
//   proportional velocities (calcContactXFPerParticleStressControlled,
//   Stress.cpp:145-192);
//   repulsive xF, half to each particle (RepulsiveForce.cpp:111-131).
// pst[k][side][slot][6]; unused slots are written as zero.
__global__ __launch_bounds__(128) void k_stress_pairs(const PairRec* __restrict__ pairs, const double4* __restrict__ pos,
    const int* __restrict__ inst_of, const InstConst* IC, const DevScalars* sc, const BatchScalars* bs,
    const double* __restrict__ parts, int nparts, const double* __restrict__ vel6, int n, int control,
    double* __restrict__ pst, int pcap)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= bs->npairs || k >= pcap) return;
    const PairRec q = pairs[k];
    const int inst = inst_of[q.p0];
    const Consts& C = IC[inst].C;
    const Flow F = sc[inst].flow;
    const double4 x04 = pos[q.p0], x14 = pos[q.p1];
    const double a0 = x04.w, a1 = x14.w;
    const double* n3 = q.n;
    double* o0 = pst+2*(size_t)ST_VALS*k;
    double* o1 = o0+ST_VALS;
    for (int m = 0; m < ST_VALS; m++) { o0[m] = 0; o1[m] = 0; }
    const size_t m6 = 6*(size_t)n, i0 = 6*(size_t)q.p0, i1 = 6*(size_t)q.p1;
    // interparticle vector as Interaction::calcNormalVectorDistanceGap computes it
    double rv[3] = {x14.x-x04.x, x14.y-x04.y, x14.z-x04.z};
    {
        double sd[3] = {sc[inst].shear_disp[0], sc[inst].shear_disp[1], sc[inst].shear_disp[2]};
        periodize_diff(C.L, C.half, sd, rv);
    }
    if (q.flags & F_LUB) {
        LubConst lc;
        lub_constants(a0, a1, lc);
        double cXM[4], cYM[4];
        lub_stress_constants(a0, a1, cXM, cYM);
        double co, lco;
        lub_coeffs(q.gap, C.lub_reduce, co, lco);
        double XG[4], YG[4], XM[4], YM[4];
        for (int j = 0; j < 4; j++) {
            XG[j] = lc.XG[j]*co;
            YG[j] = C.tangential ? lc.YG[j]*lco : 0.0;
            XM[j] = cXM[j]*co;
            YM[j] = C.tangential ? cYM[j]*lco : 0.0;
        }
        double nn[6], En[3];
        outer3(n3, nn);
        symdot(F.E, n3, En);
        const double nnE = dot3(n3, En);
        {   // M:E
            const double coeff = 1.5*nnE;
            const double ci = coeff*(XM[0]+XM[1]), cj = coeff*(XM[2]+XM[3]);
            double t[6];
            for (int m = 0; m < 6; m++) t[m] = 0;
            if (C.tangential) {
                double nEn[6];
                outer_sym3(n3, En, nEn);
                for (int m = 0; m < 6; m++) t[m] = 2*(nEn[m]-nnE*nn[m]);
            }
            const double yi = YM[0]+YM[1], yj = YM[2]+YM[3];
            for (int m = 0; m < 6; m++) {
                o0[6*SL_ME+m] = ci*nn[m]+yi*t[m];
                o1[6*SL_ME+m] = cj*nn[m]+yj*t[m];
            }
        }
        for (int p = 0; p < nparts; p++) {   // GU
            const double* vi = parts+p*m6+i0;
            const double* vj = parts+p*m6+i1;
            const double nvi = dot3(n3, vi), nvj = dot3(n3, vj);
            const double ci = -(XG[0]*nvi+XG[1]*nvj), cj = -(XG[2]*nvi+XG[3]*nvj);
            double si[6], sj[6];
            for (int m = 0; m < 6; m++) { si[m] = ci*nn[m]; sj[m] = cj*nn[m]; }
            if (C.tangential) {
                double ti[6], tj[6], t2i[6], t2j[6], wi[3], wj[3];
                outer_sym3(n3, vi, ti);
                outer_sym3(n3, vj, tj);
                for (int m = 0; m < 6; m++) { ti[m] = ti[m]-nvi*nn[m]; tj[m] = tj[m]-nvj*nn[m]; }
                cross3(vi+3, n3, wi);
                cross3(vj+3, n3, wj);
                outer_sym3(n3, wi, t2i);
                outer_sym3(n3, wj, t2j);
                for (int m = 0; m < 6; m++) {
                    si[m] += -2*(YG[0]*ti[m]+YG[1]*tj[m]);
                    sj[m] += -2*(YG[2]*ti[m]+YG[3]*tj[m]);
                    si[m] += -2*(YM[0]*t2i[m]+YM[1]*t2j[m]);
                    sj[m] += -2*(YM[2]*t2i[m]+YM[3]*t2j[m]);
                }
            }
            for (int m = 0; m < 6; m++) { o0[6*(SL_GU+p)+m] = si[m]; o1[6*(SL_GU+p)+m] = sj[m]; }
        }
    }
    if (q.flags & F_CONTACT) {
        DashCoef c;
        dashpot_coeffs(a0, a1, C.normal_coeff, C.tangential_coeff, c);
        const double rij = a0+a1, w0 = a0/rij, w1 = a1/rij;
        double st[6], f[3];
        if (control == 0) {
            dashpot_force(q, c, C, F, vel6+i0, vel6+i1, vel6+i0+3, vel6+i1+3, 0, f);
            double Ft[3] = {q.f_total[0]+f[0], q.f_total[1]+f[1], q.f_total[2]+f[2]};
            outer_sym3(rv, Ft, st);
            for (int m = 0; m < 6; m++) { o0[6*SL_XFC+m] = w0*st[m]; o1[6*SL_XFC+m] = w1*st[m]; }
        } else {
            // rate proportional velocities (with U_inf) and rate independent ones
            // (gatherVelocitiesByRateDependencies, Stress.cpp:116-143)
            double vp[2][6], vi[2][6];
            for (int s = 0; s < 2; s++) {
                const size_t ii = s == 0 ? i0 : i1;
                for (int m = 0; m < 6; m++) {
                    double a = 0, b = 0;
                    if (nparts == 4) {
                        a += parts[m6+ii+m]; a += parts[2*m6+ii+m];
                        b += parts[ii+m]; b += parts[3*m6+ii+m];
                    } else {
                        a += parts[m6+ii+m];
                        b += parts[ii+m];
                    }
                    vp[s][m] = a; vi[s][m] = b;
                }
                const double4 xs = s == 0 ? x04 : x14;
                double x[3] = {xs.x, xs.y, xs.z}, u[3];
                uinf(F, x, u);
                for (int m = 0; m < 3; m++) { vp[s][m] += u[m]; vp[s][3+m] += F.omega[m]; }
            }
            double sp[6], si[6];
            outer_sym3(rv, q.f_total, st);
            for (int m = 0; m < 6; m++) si[m] = st[m];
            dashpot_force(q, c, C, F, vp[0], vp[1], vp[0]+3, vp[1]+3, 0, f);
            outer_sym3(rv, f, sp);
            dashpot_force(q, c, C, F, vi[0], vi[1], vi[0]+3, vi[1]+3, 1, f);
            outer_sym3(rv, f, st);
            for (int m = 0; m < 6; m++) {
                o0[6*SL_XFC+m] = w0*si[m]+w0*st[m];
                o1[6*SL_XFC+m] = w1*si[m]+w1*st[m];
                o0[6*SL_XFC_PROP+m] = w0*sp[m];
                o1[6*SL_XFC_PROP+m] = w1*sp[m];
            }
        }
    }
    if (C.repulsive) {
        double f[3], st[6];
        for (int m = 0; m < 3; m++) f[m] = -q.rep_norm*n3[m];
        outer_sym3(rv, f, st);
        for (int m = 0; m < 6; m++) { o0[6*SL_XFR+m] = 0.5*st[m]; o1[6*SL_XFR+m] = 0.5*st[m]; }
    }
}

// Per-interaction force columns of Simulation::outputIntFileTxt
// (Simulation.cpp:941-1006), PAIR_OUT values per pair of [k0, k1):
//  [0]   normal part of the lubrication force on p0, positive for compression:
//        -F.n, F = Lubrication::getTotalForce (Lubrication.cpp:659-694) at the
//        non-affine velocities plus the R_FE:E term of the pair's right-hand
//        side; zero unless the gap is positive;
//  [1-3] its tangential part (Lubrication.h:120-124);
//  [4]   normal part F_c.n of the contact force (signed) and [5-7] its
//        tangential part; F_c = spring + dashpot at the total velocities
//        (Contact.cpp:235-253, 487-501); zero without contact;
//  [8]   norm of the repulsive force (RepulsiveForce.h:63-66);
//  [9]   (rvec (x) F_c) : E_inf / rate^2, the viscosity contribution of the
//        contact xF (Contact.cpp:439-450) at this state.
__global__ void k_pair_output(const PairRec* __restrict__ pairs, int k0, int k1, const double4* __restrict__ pos,
    const int* __restrict__ inst_of, const InstConst* IC, const DevScalars* sc, const double* __restrict__ vel6,
    const double* __restrict__ na6, double* __restrict__ out)
{
    const int k = k0+blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= k1) return;
    const PairRec q = pairs[k];
    const int inst = inst_of[q.p0];
    const Consts& C = IC[inst].C;
    const Flow F = sc[inst].flow;
    const double4 x04 = pos[q.p0], x14 = pos[q.p1];
    const double a0 = x04.w, a1 = x14.w;
    const double* n = q.n;
    double* o = out+PAIR_OUT*(size_t)(k-k0);
    double v[PAIR_OUT];
    for (int m = 0; m < PAIR_OUT; m++) v[m] = 0;
    const size_t i0 = 6*(size_t)q.p0, i1 = 6*(size_t)q.p1;
    if ((q.flags & F_LUB) && q.gap > 0) {
        LubConst c;
        lub_constants(a0, a1, c);
        double lc, llc;
        lub_coeffs(q.gap, C.lub_reduce, lc, llc);
        const double* vi = na6+i0;
        const double* vj = na6+i1;
        double f[3], t[3];
        for (int m = 0; m < 3; m++) t[m] = c.XA[0]*lc*vi[m]+c.XA[1]*lc*vj[m];
        const double d = -dot3(t, n);
        for (int m = 0; m < 3; m++) f[m] = d*n[m];
        if (C.tangential) {
            const double di = dot3(n, vi), dj = dot3(n, vj);
            double ci[3], cj[3];
            cross3(n, vi+3, ci);
            cross3(n, vj+3, cj);
            for (int m = 0; m < 3; m++) {
                f[m] += -c.YA[0]*llc*(vi[m]-n[m]*di)-c.YA[1]*llc*(vj[m]-n[m]*dj);
                f[m] += -c.YB[0]*llc*ci[m]-c.YB[2]*llc*cj[m];
            }
        }
        RhsAcc R{};
        const double x0[3] = {x04.x, x04.y, x04.z}, x1[3] = {x14.x, x14.y, x14.z};
        PairRec ql = q;
        ql.flags &= ~F_CONTACT;           // the hydro term only
        Consts Cl = C;
        Cl.repulsive = 0;
        pair_rhs_side(ql, 0, a0, a1, x0, x1, Cl, F, R);
        for (int m = 0; m < 3; m++) f[m] += R.Fh[m];
        const double fn = dot3(f, n);
        v[0] = -fn;
        for (int m = 0; m < 3; m++) v[1+m] = f[m]-fn*n[m];
    }
    if (q.flags & F_CONTACT) {
        DashCoef c;
        dashpot_coeffs(a0, a1, C.normal_coeff, C.tangential_coeff, c);
        double fd[3], fc[3], rv[3] = {n[0]*q.r, n[1]*q.r, n[2]*q.r}, st[6];
        dashpot_force(q, c, C, F, vel6+i0, vel6+i1, vel6+i0+3, vel6+i1+3, 0, fd);
        for (int m = 0; m < 3; m++) fc[m] = q.f_total[m]+fd[m];
        const double d = dot3(fc, n);
        v[4] = d;
        for (int m = 0; m < 3; m++) v[5+m] = fc[m]-d*n[m];
        outer_sym3(rv, fc, st);
        if (F.rate != 0) v[9] = ddot6(st, F.E)/(F.rate*F.rate);
    }
    if (C.repulsive) v[8] = q.rep_norm;
    for (int m = 0; m < PAIR_OUT; m++) o[m] = v[m];
}
void launch_pair_output(const PairRec* pairs, int k0, int k1, const double4* pos, const int* inst_of,
                        const InstConst* IC, const DevScalars* sc, const double* vel6, const double* na6, double* out,
                        hipStream_t s)
{
    if (k1 > k0) k_pair_output<<<(k1-k0+127)/128, 128, 0, s>>>(pairs, k0, k1, pos, inst_of, IC, sc, vel6, na6, out);
}

// Per particle and slot, the sum of its pairs' contributions in ascending
// partner order; 64 lanes per particle, lane v < 48 owns value v.
__global__ __launch_bounds__(256) void k_stress_gather(const double* __restrict__ pst, int n,
    const int* __restrict__ row_ptr, const int* __restrict__ tptr, const int* __restrict__ tidx,
    double* __restrict__ ps)
{
    const int i = blockIdx.x*4+(int)(threadIdx.x/WAVE);
    const int v = threadIdx.x & (WAVE-1);
    if (i >= n || v >= ST_VALS) return;
    double acc = 0;
    for (int t = tptr[i]; t < tptr[i+1]; t++) acc += pst[2*(size_t)ST_VALS*tidx[t]+ST_VALS+v];
    for (int k = row_ptr[i]; k < row_ptr[i+1]; k++) acc += pst[2*(size_t)ST_VALS*k+v];
    ps[(size_t)ST_VALS*i+v] = acc;
}

// Sums over particles in fixed chunks (StressComponent::getTotalStress,
// StressComponent.h:26-33), one partial row of 48 values per chunk.
static constexpr int ST_CHUNK = 128;
// acc + p[0] + p[s] + ... + p[(m-1) s], added in that order; the loads are
// issued ST_BATCH at a time so that a sum waits for one memory latency per
// batch instead of one per term
static constexpr int ST_BATCH = 8;
__device__ inline double ordered_sum(double acc, const double* __restrict__ p, size_t s, int m)
{
    int k = 0;
    for (; k+ST_BATCH <= m; k += ST_BATCH) {
        double t[ST_BATCH];
#pragma unroll
        for (int q = 0; q < ST_BATCH; q++) t[q] = p[s*(size_t)(k+q)];
#pragma unroll
        for (int q = 0; q < ST_BATCH; q++) acc += t[q];
    }
    for (; k < m; k++) acc += p[s*(size_t)k];
    return acc;
}
__global__ void k_stress_chunks(const double* __restrict__ ps, int n, double* __restrict__ partial)
{
    const int v = threadIdx.x;
    if (v >= ST_VALS) return;
    const int i0 = blockIdx.x*ST_CHUNK, i1 = min(n, i0+ST_CHUNK);
    partial[(size_t)ST_VALS*blockIdx.x+v] = ordered_sum(0.0, ps+(size_t)ST_VALS*i0+v, ST_VALS, i1-i0);
}

// Totals per slot, then either the stress-control shear rate
// (System::computeShearRate + gatherStressesByRateDependencies,
// System.cpp:2017-2055, Stress.cpp:320-340) or the output groups
// (System::calcStress, Stress.cpp:342-384).
// One workgroup per instance, over the chunks of its slots.
__global__ void k_stress_finalize(const double* __restrict__ partial, const InstConst* IC, DevScalars* scs,
                                  int nparts, int control, int mode)
{
    __shared__ double tot[ST_VALS];
    const int v = threadIdx.x;
    const InstConst& ic = IC[blockIdx.x];
    DevScalars* sc = scs+blockIdx.x;
    if (ic.n <= 0) return;
    const Consts& C = ic.C;
    const int b0 = ic.poff/ST_CHUNK, b1 = (ic.poff+ic.n+ST_CHUNK-1)/ST_CHUNK;
    if (v < ST_VALS) tot[v] = ordered_sum(0.0, partial+(size_t)ST_VALS*b0+v, ST_VALS, b1-b0);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double vinv = 1.0/(C.L[0]*C.L[1]*C.L[2]);
    const double pi = 3.14159265358979323846;
    const Flow F = sc->flow;
    if (mode == 0) {
        double prop[6], indep[6];
        for (int m = 0; m < 6; m++) {
            double a = 0, b = 0;
            a += tot[6*SL_ME+m];
            if (nparts == 4) {
                b += tot[6*SL_GU+m];
                a += tot[6*(SL_GU+1)+m];
                a += tot[6*(SL_GU+2)+m];
                b += tot[6*(SL_GU+3)+m];
            } else {
                b += tot[6*SL_GU+m];
                a += tot[6*(SL_GU+1)+m];
            }
            b += tot[6*SL_XFC+m];
            a += tot[6*SL_XFC_PROP+m];
            b += tot[6*SL_XFR+m];
            prop[m] = a*vinv;
            indep[m] = b*vinv;
        }
        for (int m = 0; m < 6; m++) prop[m] += (2*F.E[m])/(6*pi);
        const double rp = ddot6(prop, F.E), ri = ddot6(indep, F.E);
        double rate = F.rate;
        if (rp != 0) rate = (sc->target_stress-ri)/rp;
        if (sc->strain < sc->init_strain_limit && rate > sc->init_rate_limit) rate = sc->init_rate_limit;
        Flow Fn;
        make_flow(C, rate, Fn);
        sc->flow = Fn;
        return;
    }
    // groups contact, dashpot, hydro, repulsion; components added in map order
    double g[4][6];
    for (int m = 0; m < 6; m++) {
        double c = 0, d = 0, h = 0, r = 0;
        h += tot[6*SL_ME+m];
        c += tot[6*SL_GU+m];
        d += tot[6*(SL_GU+1)+m];
        h += tot[6*(SL_GU+2)+m];
        r += tot[6*(SL_GU+3)+m];
        c += tot[6*SL_XFC+m];
        if (control == 1) c += tot[6*SL_XFC_PROP+m];
        r += tot[6*SL_XFR+m];
        g[0][m] = c*vinv; g[1][m] = d*vinv; g[2][m] = h*vinv; g[3][m] = r*vinv;
    }
    for (int m = 0; m < 6; m++) g[2][m] += (2*F.E[m])/(6*pi);
    for (int m = 0; m < 6; m++) {
        double t = 0;
        for (int q = 0; q < 4; q++) { sc->stress_groups[q][m] = g[q][m]; t += g[q][m]; }
        sc->stress_total[m] = t;
    }
}

// na = sum of the velocity parts in order; under stress control the rate
// proportional parts are first rescaled by the new shear rate, in place
// (rescaleVelHydroStressControlled + sumUpVelocityComponents, System.cpp:2236-2271).
__global__ void k_sum_parts(double* __restrict__ parts, int nparts, double* __restrict__ na6, int n,
                            const int* __restrict__ inst_of, const DevScalars* sc, int rescale)
{
    const size_t t = blockIdx.x*(size_t)blockDim.x+threadIdx.x, m6 = 6*(size_t)n;
    if (t >= m6) return;
    const int inst = inst_of[t/6];
    if (inst < 0) { na6[t] = 0.0; return; }
    const double rate = sc[inst].flow.rate;
    double acc = 0;
    for (int p = 0; p < nparts; p++) {
        double v = parts[p*m6+t];
        const bool prop = nparts == 4 ? (p == 1 || p == 2) : p == 1;
        if (rescale && prop) { v *= rate; parts[p*m6+t] = v; }
        acc += v;
    }
    na6[t] = acc;
}

__global__ void k_set_rate(DevScalars* sc, const InstConst* IC, int ninst, double rate)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= ninst || IC[k].n <= 0) return;
    Flow F;
    make_flow(IC[k].C, rate, F);
    sc[k].flow = F;
}

void launch_set_rate(DevScalars* sc, const InstConst* IC, int ninst, double rate, hipStream_t s)
{
    k_set_rate<<<(ninst+63)/64, 64, 0, s>>>(sc, IC, ninst, rate);
}

void launch_stress(const PairRec* pairs, const double4* pos, const int* inst_of, const InstConst* IC, int ninst,
                   DevScalars* sc, const BatchScalars* bs, const double* parts, int nparts, const double* vel6,
                   int n, int control, int mode, const int* row_ptr, const int* tptr, const int* tidx, double* pst,
                   double* ps, double* partial, int pcap, hipStream_t s)
{
    k_stress_pairs<<<(pcap+127)/128, 128, 0, s>>>(pairs, pos, inst_of, IC, sc, bs, parts, nparts, vel6, n, control,
                                                  pst, pcap);
    k_stress_gather<<<(n+3)/4, 256, 0, s>>>(pst, n, row_ptr, tptr, tidx, ps);
    const int nch = (n+ST_CHUNK-1)/ST_CHUNK;
    k_stress_chunks<<<nch, 64, 0, s>>>(ps, n, partial);
    k_stress_finalize<<<ninst, 64, 0, s>>>(partial, IC, sc, nparts, control, mode);
}

void launch_sum_parts(double* parts, int nparts, double* na6, int n, const int* inst_of, const DevScalars* sc,
                      int rescale, hipStream_t s)
{
    const size_t m6 = 6*(size_t)n;
    k_sum_parts<<<(unsigned)((m6+255)/256), 256, 0, s>>>(parts, nparts, na6, n, inst_of, sc, rescale);
}

// ------------------------------------------------------------------ velocities, time step, moves
__global__ void k_finish_velocity(int n, const double4* pos, const double* na6, double* vel6,
                                  const int* __restrict__ inst_of, DevScalars* scs, int in_predictor)
{
    // System::computeMaxNAVelocity + adjustVelocityPeriodicBoundary (System.cpp:1887-1902, 2369-2392)
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (blockIdx.x*blockDim.x >= n) return;
    const int inst = i < n ? inst_of[i] : -1;
    double vmax = -1.0;
    if (inst >= 0) {
        DevScalars* sc = scs+inst;
        double4 x4 = pos[i];
        double x[3] = {x4.x, x4.y, x4.z}, u[3];
        const Flow& F = sc->flow;
        uinf(F, x, u);
        const double* v = na6+6*(size_t)i;
        vmax = v[0]*v[0]+v[1]*v[1]+v[2]*v[2];
        for (int k = 0; k < 3; k++) {
            vel6[6*(size_t)i+k] = v[k]+u[k];
            vel6[6*(size_t)i+3+k] = v[3+k]+F.omega[k];
        }
    }
    if (in_predictor) wave_max_into<&DevScalars::max_vel_bits>(scs, inst, vmax);
}
void launch_finish_velocity(int n, const double4* pos, const double* na6, double* vel6, const int* inst_of,
                            DevScalars* sc, int ninst, int in_predictor, hipStream_t s)
{
    if (in_predictor) k_reset_inst<<<(ninst+63)/64, 64, 0, s>>>(sc, ninst, 2, -1);
    k_finish_velocity<<<(n+255)/256, 256, 0, s>>>(n, pos, na6, vel6, inst_of, sc, in_predictor);
}

// System::adaptTimeStep (System.cpp:1013-1066), time_end disabled; one
// thread per instance.  With fixed_dt the step stays dt_max (adaptTimeStep is
// not called, System.cpp:1094, 1130).
__global__ void k_adapt_dt(DevScalars* scs, const InstConst* IC, int ninst, int advance_time)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= ninst || IC[k].n <= 0) return;
    DevScalars* sc = scs+k;
    const Params& P = IC[k].P;
    if (P.fixed_dt) {
        sc->dt = P.dt_max;
        if (advance_time) sc->time += P.dt_max;
        return;
    }
    const double strain_end = sc->strain_end;
    double rate = sc->flow.rate;
    double mv = sqrt(bits_to_double(sc->max_vel_bits));
    double ms = sqrt(bits_to_double(sc->max_slide_bits));
    double dt;
    if (mv > 0 || ms > 0) dt = mv > ms ? P.disp_max/mv : P.disp_max/ms;
    else dt = P.disp_max/rate;
    if (dt*rate > P.disp_max) dt = P.disp_max/rate;
    if (P.dt_max > 0 && dt > P.dt_max) dt = P.dt_max;
    if (P.dt_min > 0 && dt < P.dt_min) dt = P.dt_min;
    if (strain_end >= 0) {
        if (fabs(dt*rate) > strain_end-sc->strain) dt = fabs((strain_end-sc->strain)/rate);
    }
    sc->dt = dt;
    if (advance_time) sc->time += dt;
}
void launch_adapt_dt(DevScalars* sc, const InstConst* IC, int ninst, int advance_time, hipStream_t s)
{
    k_adapt_dt<<<(ninst+63)/64, 64, 0, s>>>(sc, IC, ninst, advance_time);
}
__global__ void k_set_strain_end(DevScalars* sc, int ninst, double strain_end)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k < ninst) sc[k].strain_end = strain_end;
}
void launch_set_strain_end(DevScalars* sc, int ninst, double strain_end, hipStream_t s)
{
    k_set_strain_end<<<(ninst+63)/64, 64, 0, s>>>(sc, ninst, strain_end);
}

// System::timeStepBoxing (System.cpp:672-717), simple shear; one thread per instance
__global__ void k_boxing(DevScalars* scs, const InstConst* IC, int ninst)
{
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= ninst || IC[k].n <= 0) return;
    DevScalars* sc = scs+k;
    const Consts& C = IC[k].C;
    double dt = sc->dt;
    const Flow& F = sc->flow;
    sc->strain += F.rate*dt;
    double z[3] = {0, 0, 1}, e[3];
    symdot(F.E, z, e);
    double inc[3] = {2*e[0]*dt, 2*e[1]*dt, 2*e[2]*dt};
    for (int k = 0; k < 3; k++) sc->shear_disp[k] += inc[k]*C.L[2];
    int m = (int)(sc->shear_disp[0]/C.L[0]);
    if (sc->shear_disp[0] < 0) m--;
    sc->shear_disp[0] = sc->shear_disp[0]-m*C.L[0];
    m = (int)(sc->shear_disp[1]/C.L[1]);
    if (sc->shear_disp[1] < 0) m--;
    sc->shear_disp[1] = sc->shear_disp[1]-m*C.L[1];
}
void launch_boxing(DevScalars* sc, const InstConst* IC, int ninst, hipStream_t s)
{
    k_boxing<<<(ninst+63)/64, 64, 0, s>>>(sc, IC, ninst);
}

// System::displacement (System.cpp:2418-2443).  mode 0: predictor move by
// U dt; mode 1: corrector (velocity average, System.cpp:1164-1187).
// mode 0: predictor move; mode 1: corrector move, then the corrector's
// non-affine velocity times dt into na_disp (System.cpp:980-982)
__global__ void k_displace(double4* pos, double* vel6, const double* pred6, int n, const int* __restrict__ inst_of,
                           const InstConst* IC, DevScalars* scs, int* box_of, int mode, const double* na6,
                           double* na_disp)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n) return;
    const int inst = inst_of[i];
    if (inst < 0) return;
    DevScalars* sc = scs+inst;
    const InstConst& g = IC[inst];
    const Consts& C = g.C;
    double dt = sc->dt;
    double* v = vel6+6*(size_t)i;
    double dr[3];
    if (mode == 0) {
        for (int k = 0; k < 3; k++) dr[k] = v[k]*dt;
    } else {
        const double* vp = pred6+6*(size_t)i;
        for (int k = 0; k < 6; k++) v[k] = 0.5*(v[k]+vp[k]);
        for (int k = 0; k < 3; k++) dr[k] = (v[k]-vp[k])*dt;
        for (int k = 0; k < 3; k++) na_disp[3*(size_t)i+k] += na6[6*(size_t)i+k]*dt;
    }
    double4 x4 = pos[i];
    double x[3] = {x4.x+dr[0], x4.y+dr[1], x4.z+dr[2]};
    double sd[3] = {sc->shear_disp[0], sc->shear_disp[1], sc->shear_disp[2]};
    int zs = periodize(C.L, sd, x);
    if (zs != 0) {
        double f = (double)zs;
        for (int k = 0; k < 3; k++) v[k] += f*sc->flow.vel_diff[k];
    }
    pos[i] = make_double4(x[0], x[1], x[2], x4.w);
    int l = which_box(x, g.bx, g.by, g.bz, (unsigned)g.yz, (unsigned)g.nz, (unsigned)g.nbox);
    if (l < 0) { raise_error(sc, E_BOX_OOB); l = 0; }
    box_of[i] = g.coff+l;
}
void launch_displace(double4* pos, double* vel6, const double* pred6, int n, const int* inst_of,
                     const InstConst* IC, DevScalars* sc, int* box_of, int mode, const double* na6, double* na_disp,
                     hipStream_t s)
{
    k_displace<<<(n+255)/256, 256, 0, s>>>(pos, vel6, pred6, n, inst_of, IC, sc, box_of, mode, na6, na_disp);
}

template <int LPR>
__global__ void k_spmv64(BsrView A, const double* x, double* y)
{
    int g = blockIdx.x*blockDim.x+threadIdx.x;
    int i = g/LPR, sub = g%LPR;
    if (i >= A.n) return;
    double v[6];
    spmv_row<double, double, LPR>(A, A.od64, A.diag64+15*(size_t)i, x, i, sub, v);
    if (sub == 0) for (int c = 0; c < 6; c++) y[6*(size_t)i+c] = v[c];
}
void launch_spmv64(const BsrView& A, const double* x, double* y, hipStream_t s)
{
    int th = A.n*4;
    k_spmv64<4><<<(th+255)/256, 256, 0, s>>>(A, x, y);
}
// FP32 SpMV with the FP32 diagonal taken from the FP64 one on the fly
template <int LPR>
__global__ void k_spmv32(BsrView A, const float* x, float* y)
{
    int g = blockIdx.x*blockDim.x+threadIdx.x;
    int i = g/LPR, sub = g%LPR;
    if (i >= A.n) return;
    float v[6];
    spmv_row<float, float, LPR>(A, A.od32, (const float*)nullptr, x, i, sub, v);
    if (sub == 0) {
        float xi[6];
        for (int c = 0; c < 6; c++) xi[c] = x[6*(size_t)i+c];
        dmul<float, float>(A.diag32+15*(size_t)i, xi, v);
        for (int c = 0; c < 6; c++) y[6*(size_t)i+c] = v[c];
    }
}
void launch_spmv32(const BsrView& A, const float* x, float* y, int lpr, hipStream_t s)
{
    int th = A.n*lpr;
    switch (lpr) {
    case 1: k_spmv32<1><<<(th+255)/256, 256, 0, s>>>(A, x, y); break;
    case 2: k_spmv32<2><<<(th+255)/256, 256, 0, s>>>(A, x, y); break;
    case 8: k_spmv32<8><<<(th+255)/256, 256, 0, s>>>(A, x, y); break;
    default: k_spmv32<4><<<(th+255)/256, 256, 0, s>>>(A, x, y); break;
    }
}

// ------------------------------------------------------------------ persistent mixed-precision solver
// FP64 iterative refinement around an FP32 block-Jacobi PCG, in one
// cooperative launch.  Per inner iteration two grid barriers:
//   phase A: d += a p, r -= a q, z = M^-1 r, partial (r.z, r.r)
//   phase B: w = A z, p = z + b p, q = w + b q, partial p.q   (q = A p without a second SpMV)
// Partial sums are reduced by every workgroup in the same fixed order, so all
// workgroups take identical branch decisions.
static constexpr int PS_BLOCK = 256;

__device__ inline void block_sum2(double& a, double& b, double* lds)
{
    for (int o = WAVE/2; o > 0; o >>= 1) { a += __shfl_xor(a, o, WAVE); b += __shfl_xor(b, o, WAVE); }
    unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE;
    __syncthreads();
    if (lane == 0) { lds[2*wid] = a; lds[2*wid+1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0;
        for (unsigned k = 0; k < blockDim.x/WAVE; k++) { s0 += lds[2*k]; s1 += lds[2*k+1]; }
        lds[0] = s0; lds[1] = s1;
    }
    __syncthreads();
    a = lds[0]; b = lds[1];
    __syncthreads();
}
// every workgroup reduces the G partial pairs in the same order
__device__ inline void grid_reduce2(const double* part, int G, double& a, double& b, double* lds)
{
    double s0 = 0, s1 = 0;
    if (threadIdx.x < WAVE) {
        for (int g = threadIdx.x; g < G; g += WAVE) { s0 += part[2*g]; s1 += part[2*g+1]; }
        for (int o = WAVE/2; o > 0; o >>= 1) { s0 += __shfl_xor(s0, o, WAVE); s1 += __shfl_xor(s1, o, WAVE); }
        if (threadIdx.x == 0) { lds[16] = s0; lds[17] = s1; }
    }
    __syncthreads();
    a = lds[16]; b = lds[17];
    __syncthreads();
}

// block sum, publish into one of two partial buffers, grid barrier, then the
// same ordered reduction in every workgroup.  Alternating buffers means one
// barrier per reduction suffices: a buffer is rewritten only after the next
// barrier, by which time every workgroup has finished reading it.
__device__ inline void grid_sum2(cg::grid_group& grid, double& a, double& b, double* part, int& buf, int G,
                                 double* lds)
{
    block_sum2(a, b, lds);
    double* pb = part+(size_t)buf*2*G;
    if (threadIdx.x == 0) { pb[2*blockIdx.x] = a; pb[2*blockIdx.x+1] = b; }
    grid.sync();
    grid_reduce2(pb, G, a, b, lds);
    buf ^= 1;
}

template <int LPR>
__global__ __launch_bounds__(PS_BLOCK) void k_solve_mixed(BsrView A, const double* __restrict__ b,
    double* __restrict__ x, SolverWork w, DevScalars* sc, SolveOpts o)
{
    cg::grid_group grid = cg::this_grid();
    __shared__ double lds[32];
    const int G = gridDim.x;
    const int rows_per_block = PS_BLOCK/LPR;
    const int sub = threadIdx.x%LPR;
    const int stride = G*rows_per_block;
    const int row0 = blockIdx.x*rows_per_block+threadIdx.x/LPR;
    const int n = A.n;
    double* part = w.partials;
    int buf = 0;
    double bb = 0, rr64 = 0;
    int outer = 0, inner_total = 0;
    bool converged = false;
    for (outer = 0; outer <= o.max_outer; outer++) {
        // residual r64 = b - A64 x (and ||b||^2)
        double s_rr = 0, s_bb = 0;
        for (int i = row0; i < n; i += stride) {
            double y[6];
            spmv_row<double, double, LPR>(A, A.od64, A.diag64+15*(size_t)i, x, i, sub, y);
            if (sub == 0) {
                for (int c = 0; c < 6; c++) {
                    double bi = b[6*(size_t)i+c];
                    double r = bi-y[c];
                    w.r64[6*(size_t)i+c] = r;
                    s_rr += r*r;
                    s_bb += bi*bi;
                }
            }
        }
        grid_sum2(grid, s_rr, s_bb, part, buf, G, lds);
        if (outer == 0) bb = s_bb;
        if (bb == 0) {                     // zero right-hand side: the solution is zero
            for (int i = row0; i < n; i += stride)
                if (sub == 0) for (int c = 0; c < 6; c++) x[6*(size_t)i+c] = 0.0;
            rr64 = 0;
            converged = true;
            break;
        }
        rr64 = s_rr;
        converged = sqrt(rr64) <= o.tol*sqrt(bb);
        if (converged || outer == o.max_outer) break;
        // inner PCG on A32 d = r64 / ||r64|| ; d0 = 0, p = z = M^-1 r
        const double scale = 1.0/sqrt(rr64);
        double s_rz = 0, s_r2 = 0;
        for (int i = row0; i < n; i += stride) {
            if (sub != 0) continue;
            float r[6];
            for (int c = 0; c < 6; c++) r[c] = (float)(w.r64[6*(size_t)i+c]*scale);
            const float* Mi = A.minv32+36*(size_t)i;
            for (int a = 0; a < 6; a++) {
                float v = 0;
                for (int c = 0; c < 6; c++) v += Mi[6*a+c]*r[c];
                size_t k = 6*(size_t)i+a;
                w.r[k] = r[a]; w.z[k] = v; w.p[k] = v; w.d[k] = 0.f;
                s_rz += (double)r[a]*v;
                s_r2 += (double)r[a]*r[a];
            }
        }
        grid_sum2(grid, s_rz, s_r2, part, buf, G, lds);
        double rz = s_rz, r20 = s_r2;
        // q = A p
        double s_pq = 0, dummy = 0;
        for (int i = row0; i < n; i += stride) {
            float y[6];
            spmv_row<float, float, LPR>(A, A.od32, (const float*)nullptr, w.p, i, sub, y);
            if (sub == 0) {
                float xi[6];
                for (int c = 0; c < 6; c++) xi[c] = w.p[6*(size_t)i+c];
                dmul<float, float>(A.diag32+15*(size_t)i, xi, y);
                for (int c = 0; c < 6; c++) { w.q[6*(size_t)i+c] = y[c]; s_pq += (double)xi[c]*y[c]; }
            }
        }
        grid_sum2(grid, s_pq, dummy, part, buf, G, lds);
        double pq = s_pq;
        const double stop2 = inner_stop2(o.inner_tol, r20, o.tol, bb, rr64);
        int it;
        for (it = 0; it < o.max_inner; it++) {
            if (!(pq > 0)) break;
            const float alpha = (float)(rz/pq);
            double s_rz2 = 0, s_rr2 = 0;
            for (int i = row0; i < n; i += stride) {
                if (sub != 0) continue;
                float r[6];
                for (int c = 0; c < 6; c++) {
                    size_t k = 6*(size_t)i+c;
                    w.d[k] += alpha*w.p[k];
                    r[c] = w.r[k]-alpha*w.q[k];
                    w.r[k] = r[c];
                }
                const float* Mi = A.minv32+36*(size_t)i;
                for (int a = 0; a < 6; a++) {
                    float v = 0;
                    for (int c = 0; c < 6; c++) v += Mi[6*a+c]*r[c];
                    w.z[6*(size_t)i+a] = v;
                    s_rz2 += (double)r[a]*v;
                    s_rr2 += (double)r[a]*r[a];
                }
            }
            grid_sum2(grid, s_rz2, s_rr2, part, buf, G, lds);
            if (s_rr2 <= stop2) { it++; break; }
            const float beta = (float)(s_rz2/rz);
            rz = s_rz2;
            double s_pq2 = 0;
            for (int i = row0; i < n; i += stride) {
                float y[6];
                spmv_row<float, float, LPR>(A, A.od32, (const float*)nullptr, w.z, i, sub, y);
                if (sub == 0) {
                    float zi[6];
                    for (int c = 0; c < 6; c++) zi[c] = w.z[6*(size_t)i+c];
                    dmul<float, float>(A.diag32+15*(size_t)i, zi, y);
                    for (int c = 0; c < 6; c++) {
                        size_t k = 6*(size_t)i+c;
                        float pn = zi[c]+beta*w.p[k];
                        float qn = y[c]+beta*w.q[k];
                        w.p[k] = pn; w.q[k] = qn;
                        s_pq2 += (double)pn*qn;
                    }
                }
            }
            grid_sum2(grid, s_pq2, dummy, part, buf, G, lds);
            pq = s_pq2;
        }
        inner_total += it;
        // x += d * ||r64|| (FP64); the next residual reads neighbours' x
        const double inv = sqrt(rr64);
        for (int i = row0; i < n; i += stride) {
            if (sub != 0) continue;
            for (int c = 0; c < 6; c++) {
                size_t k = 6*(size_t)i+c;
                x[k] += (double)w.d[k]*inv;
            }
        }
        grid.sync();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->res_norm = sqrt(rr64);
        sc->rhs_norm = sqrt(bb);
        sc->iters_inner = inner_total;
        sc->iters_outer = outer;
        sc->total_inner += inner_total;
        if (!converged) atomicOr(&sc->error, E_NOT_CONVERGED);
    }
}

// ------------------------------------------------------------------ LDS-resident persistent solver
// FP64 iterative refinement around FP32 block-Jacobi PCG, restructured for
// latency on MI355X:
//  * 8 lanes per block row; lane c < 6 owns component c of every Krylov
//    vector of the row in registers, and row c of the diagonal block and of
//    its inverse;
//  * each workgroup stages the packed FP32 off-diagonal blocks of its rows in
//    LDS once per solve, so an SpMV touches HBM only for neighbours' vectors;
//  * the inner solver is pipelined PCG (Ghysels & Vanroose 2014): the
//    reduction of (r,u), (w,u), (r,r) and the exchange of m = M^-1 w share one
//    grid barrier per iteration; m is double-buffered so a workgroup may
//    publish m_{i+1} while others still read m_i; if a refinement pass fails
//    to halve the true residual the solve continues with classic PCG;
//  * cross-lane traffic uses DPP lane moves, block sums a DPP tree, and the
//    grid barrier is a two-level arrival tree with a generation word.
static constexpr int LS_BLOCK = 1024, LS_LPR = 8, LS_ROWS = LS_BLOCK/LS_LPR, LS_WAVES = LS_BLOCK/WAVE;
static constexpr int LS_ENTRY_CAP = 1536;         // staged blocks per workgroup (117 KB)

struct LdsRow { int e0, ne; };      // staged entries of a row (ne < 0: not staged)

// One right-hand side (the unused size_t is the right-hand-side stride of the
// paired form, so both take the same launch arguments).
__global__ __launch_bounds__(LS_BLOCK) void k_solve_lds_one(BsrView A, const double* __restrict__ b,
    double* __restrict__ x, size_t, SolverWork w, DevScalars* scs, SolveOpts o, const SolveJob* __restrict__ jobs)
{
    __shared__ int ent_j[LS_ENTRY_CAP];             // partner; ~partner for the transposed block
    __shared__ float ent_b[LS_ENTRY_CAP*18];
    __shared__ LdsRow rows[LS_ROWS];
    __shared__ double red[3*LS_WAVES+4];
    __shared__ int s_fail;
    // this workgroup's instance: rows [row0 + wg*LS_ROWS, ...) below row_end;
    // its partial records and barrier words follow from its first row block
    const SolveJob J = jobs[blockIdx.x];
    const int wg = J.lwg, G = J.G, n = J.row_end, rbase = J.row0+wg*LS_ROWS;
    DevScalars* sc = scs+J.inst;
    unsigned* bar = w.barrier+128*(size_t)J.key;
    unsigned seq = 0;
    double g0 = 0, g1 = 0, g2 = 0;        // wave 0: this lane's share of the polled partial sums
    const int lr = threadIdx.x/LS_LPR, c = threadIdx.x%LS_LPR;
    const unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE;
    const int i = rbase+lr;
    const bool live = i < n;
    const bool own = live && c < 6;
    const size_t gi = 6*(size_t)i+c;
    unsigned gen = ld_acquire(bar+32);
    if (threadIdx.x == 0) s_fail = 0;
    // ---- stage this workgroup's rows: degrees in parallel, serial prefix in LDS
    if (threadIdx.x < LS_ROWS) {
        int ii = rbase+threadIdx.x;
        rows[threadIdx.x].ne = ii < n ? (A.tptr[ii+1]-A.tptr[ii])+(A.row_ptr[ii+1]-A.row_ptr[ii]) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int used = 0;
        for (int t = 0; t < LS_ROWS; t++) {
            int ii = rbase+t, deg = rows[t].ne;
            if (ii < n && used+deg <= LS_ENTRY_CAP) { rows[t].e0 = used; used += deg; }
            else { rows[t].e0 = 0; rows[t].ne = ii < n ? -1 : 0; }
        }
    }
    __syncthreads();
    if (live && rows[lr].ne > 0) {
        const int l0 = A.tptr[i], nl = A.tptr[i+1]-l0, u0 = A.row_ptr[i];
        for (int e = c; e < rows[lr].ne; e += LS_LPR) {
            int k, j;
            if (e < nl) { k = A.tidx[l0+e]; j = A.p0[k]; }
            else { k = u0+(e-nl); j = ~A.p1[k]; }
            int s = rows[lr].e0+e;
            ent_j[s] = j;
            for (int m = 0; m < 18; m++) ent_b[18*s+m] = A.od32[18*(size_t)k+m];
        }
    }
    // row c of D and of D^-1 in registers (lanes 0..5)
    float Dr[6] = {0, 0, 0, 0, 0, 0}, Mr[6] = {0, 0, 0, 0, 0, 0};
    if (own) {
        for (int m = 0; m < 6; m++) {
            float e6[6] = {0, 0, 0, 0, 0, 0};
            e6[m] = 1.f;
            float y[6] = {0, 0, 0, 0, 0, 0};
            dmul<float, float>(A.diag32+15*(size_t)i, e6, y);   // column m of the symmetric D
            Dr[m] = y[c];
            Mr[m] = A.minv32[36*(size_t)i+6*c+m];
        }
    }
    __syncthreads();
    const int ne = live ? rows[lr].ne : 0, e0 = live ? rows[lr].e0 : 0;

    // component c of A32 v for this row: neighbours' v from vbuf, own v from the group's lanes
    auto spmv32 = [&](const float* vbuf, float vown) -> float {
        float y[6] = {0, 0, 0, 0, 0, 0};
        if (live) {
            if (ne >= 0) {
                for (int e = c; e < ne; e += LS_LPR) {
                    int jj = ent_j[e0+e];
                    const float* bl = ent_b+18*(e0+e);
                    float xj[6];
                    if (jj >= 0) {
                        load6(vbuf+6*(size_t)jj, xj);
                        omul<float, float>(bl, xj, y);
                    } else {
                        jj = ~jj;
                        load6(vbuf+6*(size_t)jj, xj);
                        omul_t<float, float>(bl, xj, y);
                    }
                }
            } else {
                const int l0 = A.tptr[i], nl = A.tptr[i+1]-l0, u0 = A.row_ptr[i];
                const int nt = nl+(A.row_ptr[i+1]-u0);
                for (int e = c; e < nt; e += LS_LPR) {
                    float xj[6];
                    if (e < nl) {
                        int k = A.tidx[l0+e], jj = A.p0[k];
                        load6(vbuf+6*(size_t)jj, xj);
                        omul<float, float>(A.od32+18*(size_t)k, xj, y);
                    } else {
                        int k = u0+(e-nl), jj = A.p1[k];
                        load6(vbuf+6*(size_t)jj, xj);
                        omul_t<float, float>(A.od32+18*(size_t)k, xj, y);
                    }
                }
            }
        }
        for (int m = 0; m < 6; m++) y[m] = sum8f(y[m]);
        float vv[6];
        bcast8_6(vown, c, vv);
        float yc = 0, mine = 0;
        for (int m = 0; m < 6; m++) { yc += Dr[m]*vv[m]; if (m == c) mine = y[m]; }
        return own ? mine+yc : 0.f;
    };
    auto precond = [&](float rv) -> float {
        float rr[6];
        bcast8_6(rv, c, rr);
        float z = 0;
        for (int m = 0; m < 6; m++) z += Mr[m]*rr[m];
        return own ? z : 0.f;
    };

    double* part = w.partials+LDS_PART_PER_BLOCK*(size_t)J.key;
    const int rs = rec_stride(G, 3);
    int buf = 0;
    // Grid sums of three values.  Block sums (DPP tree, fixed order) are
    // published as tagged partial records in the current buffer after every
    // wave's vector stores have completed and a release fence; wave 0 then
    // polls all records with device-scope loads until each carries this
    // round's tag, keeps its lanes' shares in g0..g2, and issues the acquire
    // fence that makes the other workgroups' vector stores visible.
    auto publish3 = [&](double a, double bs, double cs) {
        a = wave_sum_d(a); bs = wave_sum_d(bs); cs = wave_sum_d(cs);
        if (lane == 0) { red[3*wid] = a; red[3*wid+1] = bs; red[3*wid+2] = cs; }
        drain_vmem();
        __syncthreads();
        double* pb = part+(size_t)buf*rs*G;
        seq++;
        if (wid == 0) {
            double t0 = lane < LS_WAVES ? red[3*lane] : 0.0;
            double t1 = lane < LS_WAVES ? red[3*lane+1] : 0.0;
            double t2 = lane < LS_WAVES ? red[3*lane+2] : 0.0;
            const unsigned tg = seq & 0xFF;
            t0 = tag_partial(wave_sum_d(t0), tg);
            t1 = tag_partial(wave_sum_d(t1), tg);
            t2 = tag_partial(wave_sum_d(t2), tg);
            release_for_records();
            if (lane == 0) {
                st_dev(pb+rs*wg, t0); st_dev(pb+rs*wg+1, t1); st_dev(pb+rs*wg+2, t2);
            }
            const unsigned long long* pq = reinterpret_cast<const unsigned long long*>(pb);
            bool ok = false;
            for (long spin = 0; s_fail == 0 && spin < SPIN_LIMIT; spin++) {
                double a0 = 0, a1 = 0, a2 = 0;
                bool done = true;
                for (int q = lane; q < G; q += WAVE) {
                    const unsigned long long* r = pq+(size_t)rs*q;
                    const unsigned long long v0 = ld_dev(r), v1 = ld_dev(r+1), v2 = ld_dev(r+2);
                    done &= (v0 & 0xFF) == tg && (v1 & 0xFF) == tg && (v2 & 0xFF) == tg;
                    a0 += untag_partial(v0); a1 += untag_partial(v1); a2 += untag_partial(v2);
                }
                if (__all(done)) { g0 = a0; g1 = a1; g2 = a2; ok = true; break; }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ok && lane == 0) s_fail = 1;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
    };
    auto gather3_start = [&](double& t0, double& t1, double& t2) { t0 = g0; t1 = g1; t2 = g2; };
    auto gather3_finish = [&](double& t0, double& t1, double& t2) {
        if (wid == 0) {
            t0 = wave_sum_d(t0); t1 = wave_sum_d(t1); t2 = wave_sum_d(t2);
            if (lane == 0) { red[3*LS_WAVES] = t0; red[3*LS_WAVES+1] = t1; red[3*LS_WAVES+2] = t2; }
        }
        __syncthreads();
        t0 = red[3*LS_WAVES]; t1 = red[3*LS_WAVES+1]; t2 = red[3*LS_WAVES+2];
        __syncthreads();
        buf ^= 1;
    };
    auto gsum3 = [&](double& a, double& bs, double& cs) {
        publish3(a, bs, cs);
        gather3_start(a, bs, cs);
        gather3_finish(a, bs, cs);
    };

    // Coarse space of the three uniform translations z_a (U_i = e_a, W_i = 0).
    // Pair resistances vanish on a rigid translation, so only the Stokes drag
    // holds these modes; with a small drag coefficient they are near-null and
    // the FP32 inner solve cannot resolve them.  Each refinement pass first
    // applies the Galerkin correction x += Z E^-1 Z^T r, E = Z^T A Z.
    // s_az[a][thread]: (A z_a) at row i, component c; s_ei: E^-1
    __shared__ double s_az[3][LS_BLOCK];
    __shared__ double s_ei[9];
    double AZ[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        double y[6] = {0, 0, 0, 0, 0, 0};
        if (live) row_translation<LS_LPR>(A, i, c, a, y);
        for (int m = 0; m < 6; m++) if (m == c && own) AZ[a] = y[m];
        s_az[a][threadIdx.x] = AZ[a];
    }
    bool use_coarse;
    {
        double e00 = c == 0 ? AZ[0] : 0.0, e01 = c == 0 ? AZ[1] : 0.0, e02 = c == 0 ? AZ[2] : 0.0;
        double e11 = c == 1 ? AZ[1] : 0.0, e12 = c == 1 ? AZ[2] : 0.0, e22 = c == 2 ? AZ[2] : 0.0;
        gsum3(e00, e11, e22);
        gsum3(e01, e02, e12);
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

    double bb = 0, rr64 = 0;
    int outer, inner_total = 0;
    bool converged = false, use_classic = false;
    for (outer = 0; outer <= o.max_outer; outer++) {
        // FP64 residual from the FP64 blocks in HBM
        double yv[6] = {0, 0, 0, 0, 0, 0};
        if (live) spmv_row<double, double, LS_LPR>(A, A.od64, A.diag64+15*(size_t)i, x, i, c, yv);
        double r64 = 0, bi = 0;
        if (own) {
            for (int m = 0; m < 6; m++) if (m == c) r64 = yv[m];
            bi = b[gi];
            r64 = bi-r64;
        }
        double s_rr = r64*r64, s_bb = bi*bi, zr0 = c == 0 ? r64 : 0.0;
        gsum3(s_rr, s_bb, zr0);
        if (s_fail) break;
        if (outer == 0) bb = s_bb;
        if (bb == 0) {                     // zero right-hand side: the solution is zero
            if (own) x[gi] = 0.0;
            rr64 = 0;
            converged = true;
            break;
        }
        if (outer > 0 && !(s_rr < 0.25*rr64)) use_classic = true;
        rr64 = s_rr;
        converged = sqrt(rr64) <= o.tol*sqrt(bb);
        if (converged || outer == o.max_outer) break;
        if (use_coarse) {
            double zr1 = c == 1 ? r64 : 0.0, zr2 = c == 2 ? r64 : 0.0, z3 = 0;
            gsum3(zr1, zr2, z3);
            double cv[3];
            for (int a = 0; a < 3; a++) cv[a] = s_ei[3*a]*zr0+s_ei[3*a+1]*zr1+s_ei[3*a+2]*zr2;
            if (own) {
                if (c < 3) x[gi] += cv[c];
                r64 -= s_az[0][threadIdx.x]*cv[0]+s_az[1][threadIdx.x]*cv[1]+s_az[2][threadIdx.x]*cv[2];
            }
        }
        const double scale = 1.0/sqrt(rr64);
        float d = 0.f;
        int it = 0;
        if (!use_classic) {
            float* mbuf[2] = {w.z, w.q};
            float r = own ? (float)(r64*scale) : 0.f;
            float u = precond(r);
            if (own) mbuf[1][gi] = u;
            double s_r2 = (double)r*r, z0 = 0, z1 = 0;
            gsum3(s_r2, z0, z1);
            const double stop2 = inner_stop2(o.inner_tol, s_r2, o.tol, bb, rr64);
            float wv = spmv32(mbuf[1], u);
            float zv = 0, qv = 0, sv = 0, pv = 0;
            double gamma_old = 1, alpha_old = 1;
            int par = 0;
            for (it = 0; it < o.max_inner; it++) {
                float mv = precond(wv);
                if (own) mbuf[par][gi] = mv;
                publish3((double)r*u, (double)wv*u, (double)r*r);
                double g, dl, rho;
                gather3_start(g, dl, rho);
                float nv = spmv32(mbuf[par], mv);
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
                sv = wv+bf*sv;
                pv = u+bf*pv;
                d += af*pv;
                r -= af*sv;
                u -= af*qv;
                wv -= af*zv;
                gamma_old = g;
                alpha_old = alpha;
                par ^= 1;
            }
        } else {
            // classic PCG, two barriers per iteration
            float r = own ? (float)(r64*scale) : 0.f;
            float z = precond(r), p = z;
            if (own) w.z[gi] = z;
            double s_rz = (double)r*z, s_r2 = (double)r*r, z0 = 0;
            gsum3(s_rz, s_r2, z0);
            double rz = s_rz;
            const double stop2 = inner_stop2(o.inner_tol, s_r2, o.tol, bb, rr64);
            float q = spmv32(w.z, z);
            double s_pq = (double)p*q, z1 = 0, z2 = 0;
            gsum3(s_pq, z1, z2);
            double pq = s_pq;
            for (it = 0; it < o.max_inner; it++) {
                if (s_fail || !(pq > 0)) break;
                const float alpha = (float)(rz/pq);
                d += alpha*p;
                r -= alpha*q;
                z = precond(r);
                if (own) w.z[gi] = z;
                double s_rz2 = (double)r*z, s_rr2 = (double)r*r, z3 = 0;
                gsum3(s_rz2, s_rr2, z3);
                if (s_rr2 <= stop2) { it++; break; }
                const float beta = (float)(s_rz2/rz);
                rz = s_rz2;
                float wz = spmv32(w.z, z);
                p = z+beta*p;
                q = wz+beta*q;
                double s_pq2 = (double)p*q, z4 = 0, z5 = 0;
                gsum3(s_pq2, z4, z5);
                pq = s_pq2;
            }
        }
        inner_total += it;
        if (own) x[gi] += (double)d*sqrt(rr64);
        tree_barrier(bar, G, wg, gen, &s_fail);
        if (s_fail) break;
    }
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

// k_solve_lds for NR >= 2 right-hand sides in one launch (one right-hand
// side: k_solve_lds_one, whose register allocation measured faster for it):
// every grid sum carries 3 NR values and every SpMV reads each staged block
// once for the NR vectors, so the per-iteration synchronisation is paid
// once.  The right-hand side k lives at b + k*rs and its solution at
// x + k*rs; a right-hand side whose inner solve has stopped keeps its
// correction while the other one iterates, and one that has converged is
// left alone by the later refinement passes.
template <int NR>
__global__ __launch_bounds__(LS_BLOCK) void k_solve_lds(BsrView A, const double* __restrict__ b,
    double* __restrict__ x, size_t rs, SolverWork w, DevScalars* scs, SolveOpts o, const SolveJob* __restrict__ jobs)
{
    static_assert(NR >= 2, "one right-hand side: k_solve_lds_one");
    constexpr int NV = 3*NR;                         // values of one grid sum
    static_assert(NV <= REC_STRIDE, "a record fits its slot");
    __shared__ int ent_j[LS_ENTRY_CAP];             // partner; ~partner for the transposed block
    __shared__ float ent_b[LS_ENTRY_CAP*18];
    __shared__ LdsRow rows[LS_ROWS];
    __shared__ double red[NV*LS_WAVES+NV];
    __shared__ int s_fail;
    // this workgroup's instance: rows [row0 + wg*LS_ROWS, ...) below row_end;
    // its partial records and barrier words follow from its first row block
    const SolveJob J = jobs[blockIdx.x];
    const int wg = J.lwg, G = J.G, n = J.row_end, rbase = J.row0+wg*LS_ROWS;
    DevScalars* sc = scs+J.inst;
    unsigned* bar = w.barrier+128*(size_t)J.key;
    unsigned seq = 0;
    const int lr = threadIdx.x/LS_LPR, c = threadIdx.x%LS_LPR;
    const unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE;
    const int i = rbase+lr;
    const bool live = i < n;
    const bool own = live && c < 6;
    const size_t gi = 6*(size_t)i+c;
    const size_t vs = 6*(size_t)A.n;      // FP32 vector stride between right-hand sides
    unsigned gen = ld_acquire(bar+32);
    if (threadIdx.x == 0) s_fail = 0;
    // ---- stage this workgroup's rows: degrees in parallel, serial prefix in LDS
    if (threadIdx.x < LS_ROWS) {
        int ii = rbase+threadIdx.x;
        rows[threadIdx.x].ne = ii < n ? (A.tptr[ii+1]-A.tptr[ii])+(A.row_ptr[ii+1]-A.row_ptr[ii]) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int used = 0;
        for (int t = 0; t < LS_ROWS; t++) {
            int ii = rbase+t, deg = rows[t].ne;
            if (ii < n && used+deg <= LS_ENTRY_CAP) { rows[t].e0 = used; used += deg; }
            else { rows[t].e0 = 0; rows[t].ne = ii < n ? -1 : 0; }
        }
    }
    __syncthreads();
    if (live && rows[lr].ne > 0) {
        const int l0 = A.tptr[i], nl = A.tptr[i+1]-l0, u0 = A.row_ptr[i];
        for (int e = c; e < rows[lr].ne; e += LS_LPR) {
            int k, j;
            if (e < nl) { k = A.tidx[l0+e]; j = A.p0[k]; }
            else { k = u0+(e-nl); j = ~A.p1[k]; }
            int s = rows[lr].e0+e;
            ent_j[s] = j;
            for (int m = 0; m < 18; m++) ent_b[18*s+m] = A.od32[18*(size_t)k+m];
        }
    }
    // row c of D and of D^-1 in registers (lanes 0..5)
    float Dr[6] = {0, 0, 0, 0, 0, 0}, Mr[6] = {0, 0, 0, 0, 0, 0};
    if (own) {
        for (int m = 0; m < 6; m++) {
            float e6[6] = {0, 0, 0, 0, 0, 0};
            e6[m] = 1.f;
            float y[6] = {0, 0, 0, 0, 0, 0};
            dmul<float, float>(A.diag32+15*(size_t)i, e6, y);   // column m of the symmetric D
            Dr[m] = y[c];
            Mr[m] = A.minv32[36*(size_t)i+6*c+m];
        }
    }
    __syncthreads();
    const int ne = live ? rows[lr].ne : 0, e0 = live ? rows[lr].e0 : 0;

    // component c of A32 v_k for this row, k < NR: neighbours' v_k from
    // vbuf + k*vs, own v_k from the group's lanes
    auto spmv32 = [&](const float* vbuf, const float* vown, float* out) {
        float y[NR][6];
        for (int k = 0; k < NR; k++)
            for (int m = 0; m < 6; m++) y[k][m] = 0;
        if (live) {
            if (ne >= 0) {
                for (int e = c; e < ne; e += LS_LPR) {
                    int jj = ent_j[e0+e];
                    const float* bl = ent_b+18*(e0+e);
                    const bool tr = jj < 0;
                    if (tr) jj = ~jj;
                    for (int k = 0; k < NR; k++) {
                        float xj[6];
                        load6(vbuf+k*vs+6*(size_t)jj, xj);
                        if (!tr) omul<float, float>(bl, xj, y[k]);
                        else omul_t<float, float>(bl, xj, y[k]);
                    }
                }
            } else {
                const int l0 = A.tptr[i], nl = A.tptr[i+1]-l0, u0 = A.row_ptr[i];
                const int nt = nl+(A.row_ptr[i+1]-u0);
                for (int e = c; e < nt; e += LS_LPR) {
                    const bool tr = e >= nl;
                    const int k0 = tr ? u0+(e-nl) : A.tidx[l0+e];
                    const int jj = tr ? A.p1[k0] : A.p0[k0];
                    for (int k = 0; k < NR; k++) {
                        float xj[6];
                        load6(vbuf+k*vs+6*(size_t)jj, xj);
                        if (!tr) omul<float, float>(A.od32+18*(size_t)k0, xj, y[k]);
                        else omul_t<float, float>(A.od32+18*(size_t)k0, xj, y[k]);
                    }
                }
            }
        }
        for (int k = 0; k < NR; k++) {
            for (int m = 0; m < 6; m++) y[k][m] = sum8f(y[k][m]);
            float vv[6];
            bcast8_6(vown[k], c, vv);
            float yc = 0, mine = 0;
            for (int m = 0; m < 6; m++) { yc += Dr[m]*vv[m]; if (m == c) mine = y[k][m]; }
            out[k] = own ? mine+yc : 0.f;
        }
    };
    auto precond = [&](float rv) -> float {
        float rr[6];
        bcast8_6(rv, c, rr);
        float z = 0;
        for (int m = 0; m < 6; m++) z += Mr[m]*rr[m];
        return own ? z : 0.f;
    };

    double* part = w.partials+LDS_PART_PER_BLOCK*(size_t)J.key;
    const int rst = rec_stride(G, NV);   // record stride (rs: the right-hand sides' stride)
    int buf = 0;
    // Grid sums of NV values.  Block sums (DPP tree, fixed order) are
    // published as tagged partial records in the current buffer after every
    // wave's vector stores have completed and a release fence; wave 0 then
    // polls all records with device-scope loads until each carries this
    // round's tag, sums them (DPP tree) into red[NV*LS_WAVES..], and issues
    // the acquire fence that makes the other workgroups' vector stores
    // visible.  Every thread reads the totals after the publishing barrier;
    // wave 0 next writes them after the following publish's first barrier.
    auto publish = [&](double* v) {
        for (int k = 0; k < NV; k++) v[k] = wave_sum_d(v[k]);
        if (lane == 0)
            for (int k = 0; k < NV; k++) red[NV*wid+k] = v[k];
        drain_vmem();
        __syncthreads();
        double* pb = part+(size_t)buf*rst*G;
        seq++;
        if (wid == 0) {
            const unsigned tg = seq & 0xFF;
            double t[NV];
            for (int k = 0; k < NV; k++) {
                t[k] = lane < LS_WAVES ? red[NV*lane+k] : 0.0;
                t[k] = tag_partial(wave_sum_d(t[k]), tg);
            }
            release_for_records();
            if (lane == 0)
                for (int k = 0; k < NV; k++) st_dev(pb+rst*wg+k, t[k]);
            const unsigned long long* pq = reinterpret_cast<const unsigned long long*>(pb);
            bool ok = false;
            for (long spin = 0; s_fail == 0 && spin < SPIN_LIMIT; spin++) {
                double a[NV];
                for (int k = 0; k < NV; k++) a[k] = 0;
                bool done = true;
                for (int q = lane; q < G; q += WAVE) {
                    unsigned long long vq[NV];
                    for (int k = 0; k < NV; k++) vq[k] = ld_dev(pq+(size_t)rst*q+k);
                    for (int k = 0; k < NV; k++) done &= (vq[k] & 0xFF) == tg;
                    for (int k = 0; k < NV; k++) a[k] += untag_partial(vq[k]);
                }
                if (__all(done)) {
                    for (int k = 0; k < NV; k++) a[k] = wave_sum_d(a[k]);
                    if (lane == 0)
                        for (int k = 0; k < NV; k++) red[NV*LS_WAVES+k] = a[k];
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
    auto gather = [&](double* t) {
        for (int k = 0; k < NV; k++) t[k] = red[NV*LS_WAVES+k];
        buf ^= 1;
    };
    auto gsum = [&](double* v) {
        publish(v);
        gather(v);
    };

    // Coarse space of the three uniform translations z_a (U_i = e_a, W_i = 0).
    // Pair resistances vanish on a rigid translation, so only the Stokes drag
    // holds these modes; with a small drag coefficient they are near-null and
    // the FP32 inner solve cannot resolve them.  Each refinement pass first
    // applies the Galerkin correction x += Z E^-1 Z^T r, E = Z^T A Z.
    // s_az[a][thread]: (A z_a) at row i, component c; s_ei: E^-1
    __shared__ double s_az[3][LS_BLOCK];
    __shared__ double s_ei[9];
    double AZ[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        double y[6] = {0, 0, 0, 0, 0, 0};
        if (live) row_translation<LS_LPR>(A, i, c, a, y);
        for (int m = 0; m < 6; m++) if (m == c && own) AZ[a] = y[m];
        s_az[a][threadIdx.x] = AZ[a];
    }
    bool use_coarse;
    {
        // E = Z^T A Z: (e00, e11, e22, e01, e02, e12) in one grid sum
        double ev[NV];
        for (int k = 0; k < NV; k++) ev[k] = 0.0;
        ev[0] = c == 0 ? AZ[0] : 0.0; ev[1] = c == 1 ? AZ[1] : 0.0; ev[2] = c == 2 ? AZ[2] : 0.0;
        ev[3] = c == 0 ? AZ[1] : 0.0; ev[4] = c == 0 ? AZ[2] : 0.0; ev[5] = c == 1 ? AZ[2] : 0.0;
        gsum(ev);
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

    double bb[NR], rr64[NR];
    bool conv[NR];
    for (int k = 0; k < NR; k++) { bb[k] = 0; rr64[k] = 0; conv[k] = false; }
    int outer, inner_total = 0;
    bool use_classic = false;
    for (outer = 0; outer <= o.max_outer; outer++) {
        // FP64 residuals from the FP64 blocks in HBM
        double r64[NR], zr0[NR], v[NV];
        for (int k = 0; k < NR; k++) {
            double yv[6] = {0, 0, 0, 0, 0, 0};
            if (live) spmv_row<double, double, LS_LPR>(A, A.od64, A.diag64+15*(size_t)i, x+k*rs, i, c, yv);
            double rk = 0, bi = 0;
            if (own) {
                for (int m = 0; m < 6; m++) if (m == c) rk = yv[m];
                bi = b[k*rs+gi];
                rk = bi-rk;
            }
            r64[k] = rk;
            v[3*k] = rk*rk; v[3*k+1] = bi*bi; v[3*k+2] = c == 0 ? rk : 0.0;
        }
        gsum(v);
        if (s_fail) break;
        bool all_conv = true;
        for (int k = 0; k < NR; k++) {
            zr0[k] = v[3*k+2];
            if (outer == 0) bb[k] = v[3*k+1];
            if (bb[k] == 0) {              // zero right-hand side: the solution is zero
                if (own) x[k*rs+gi] = 0.0;
                rr64[k] = 0;
                conv[k] = true;
                continue;
            }
            if (outer > 0 && !conv[k] && !(v[3*k] < 0.25*rr64[k])) use_classic = true;
            rr64[k] = v[3*k];
            conv[k] = sqrt(rr64[k]) <= o.tol*sqrt(bb[k]);
            all_conv &= conv[k];
        }
        if (all_conv || outer == o.max_outer) break;
        if (use_coarse) {
            for (int k = 0; k < NR; k++) {
                v[3*k] = c == 1 ? r64[k] : 0.0;
                v[3*k+1] = c == 2 ? r64[k] : 0.0;
                v[3*k+2] = 0;
            }
            gsum(v);
            for (int k = 0; k < NR; k++) {
                if (conv[k]) continue;
                double cv[3];
                for (int a = 0; a < 3; a++) cv[a] = s_ei[3*a]*zr0[k]+s_ei[3*a+1]*v[3*k]+s_ei[3*a+2]*v[3*k+1];
                if (own) {
                    if (c < 3) x[k*rs+gi] += cv[c];
                    r64[k] -= s_az[0][threadIdx.x]*cv[0]+s_az[1][threadIdx.x]*cv[1]+s_az[2][threadIdx.x]*cv[2];
                }
            }
        }
        float d[NR], r[NR];
        bool act[NR];
        int itk[NR];
        for (int k = 0; k < NR; k++) {
            d[k] = 0.f;
            itk[k] = 0;
            act[k] = !conv[k];
            const double scale = act[k] ? 1.0/sqrt(rr64[k]) : 0.0;
            r[k] = own && act[k] ? (float)(r64[k]*scale) : 0.f;
        }
        if (!use_classic) {
            float* mbuf[2] = {w.z, w.q};
            float u[NR], wv[NR], zv[NR], qv[NR], sv[NR], pv[NR], mv[NR], nv[NR];
            double stop2[NR], gamma_old[NR], alpha_old[NR];
            for (int k = 0; k < NR; k++) {
                u[k] = precond(r[k]);
                if (own) mbuf[1][k*vs+gi] = u[k];
                v[3*k] = (double)r[k]*r[k]; v[3*k+1] = 0; v[3*k+2] = 0;
            }
            gsum(v);
            for (int k = 0; k < NR; k++) {
                stop2[k] = uni(inner_stop2(o.inner_tol, v[3*k], o.tol, bb[k], rr64[k]));
                zv[k] = qv[k] = sv[k] = pv[k] = 0;
                gamma_old[k] = alpha_old[k] = 1;
            }
            spmv32(mbuf[1], u, wv);
            int par = 0;
            for (int it = 0; it < o.max_inner; it++) {
                for (int k = 0; k < NR; k++) {
                    mv[k] = precond(wv[k]);
                    if (own) mbuf[par][k*vs+gi] = mv[k];
                    v[3*k] = (double)r[k]*u[k]; v[3*k+1] = (double)wv[k]*u[k]; v[3*k+2] = (double)r[k]*r[k];
                }
                publish(v);
                spmv32(mbuf[par], mv, nv);
                gather(v);
                if (s_fail) break;
                bool any = false;
                for (int k = 0; k < NR; k++) {
                    if (!act[k]) continue;
                    const double g = uni(v[3*k]), dl = uni(v[3*k+1]), rho = uni(v[3*k+2]);
                    if (rho <= stop2[k]) { act[k] = false; continue; }
                    double beta, alpha;
                    if (it > 0) {
                        beta = g/gamma_old[k];
                        double den = dl-beta*g/alpha_old[k];
                        if (!(den > 0)) { act[k] = false; continue; }
                        alpha = g/den;
                    } else {
                        if (!(dl > 0)) { act[k] = false; continue; }
                        beta = 0;
                        alpha = g/dl;
                    }
                    const float bf = (float)beta, af = (float)alpha;
                    zv[k] = nv[k]+bf*zv[k];
                    qv[k] = mv[k]+bf*qv[k];
                    sv[k] = wv[k]+bf*sv[k];
                    pv[k] = u[k]+bf*pv[k];
                    d[k] += af*pv[k];
                    r[k] -= af*sv[k];
                    u[k] -= af*qv[k];
                    wv[k] -= af*zv[k];
                    gamma_old[k] = g;
                    alpha_old[k] = uni(alpha);
                    itk[k]++;
                    any = true;
                }
                if (!any) break;
                par ^= 1;
            }
        } else {
            // classic PCG, two barriers per iteration
            float z[NR], p[NR], q[NR], wz[NR], beta[NR];
            double rz[NR], pq[NR], stop2[NR];
            for (int k = 0; k < NR; k++) {
                z[k] = precond(r[k]);
                p[k] = z[k];
                if (own) w.z[k*vs+gi] = z[k];
                v[3*k] = (double)r[k]*z[k]; v[3*k+1] = (double)r[k]*r[k]; v[3*k+2] = 0;
            }
            gsum(v);
            for (int k = 0; k < NR; k++) {
                rz[k] = v[3*k];
                stop2[k] = inner_stop2(o.inner_tol, v[3*k+1], o.tol, bb[k], rr64[k]);
            }
            spmv32(w.z, z, q);
            for (int k = 0; k < NR; k++) { v[3*k] = (double)p[k]*q[k]; v[3*k+1] = 0; v[3*k+2] = 0; }
            gsum(v);
            for (int k = 0; k < NR; k++) pq[k] = v[3*k];
            for (int it = 0; it < o.max_inner; it++) {
                if (s_fail) break;
                bool any = false;
                for (int k = 0; k < NR; k++) {
                    if (act[k] && !(pq[k] > 0)) act[k] = false;
                    if (!act[k]) continue;
                    const float alpha = (float)(rz[k]/pq[k]);
                    d[k] += alpha*p[k];
                    r[k] -= alpha*q[k];
                    z[k] = precond(r[k]);
                    if (own) w.z[k*vs+gi] = z[k];
                    itk[k]++;
                    any = true;
                }
                if (!any) break;
                for (int k = 0; k < NR; k++) { v[3*k] = (double)r[k]*z[k]; v[3*k+1] = (double)r[k]*r[k]; v[3*k+2] = 0; }
                gsum(v);
                any = false;
                for (int k = 0; k < NR; k++) {
                    if (!act[k]) continue;
                    if (v[3*k+1] <= stop2[k]) { act[k] = false; continue; }
                    beta[k] = (float)(v[3*k]/rz[k]);
                    rz[k] = v[3*k];
                    any = true;
                }
                if (!any) break;
                spmv32(w.z, z, wz);
                for (int k = 0; k < NR; k++) {
                    if (act[k]) {
                        p[k] = z[k]+beta[k]*p[k];
                        q[k] = wz[k]+beta[k]*q[k];
                    }
                    v[3*k] = (double)p[k]*q[k]; v[3*k+1] = 0; v[3*k+2] = 0;
                }
                gsum(v);
                for (int k = 0; k < NR; k++) pq[k] = v[3*k];
            }
        }
        int itmax = 0;
        for (int k = 0; k < NR; k++) {
            itmax = max(itmax, itk[k]);
            if (own && !conv[k]) x[k*rs+gi] += (double)d[k]*sqrt(rr64[k]);
        }
        inner_total += itmax;
        tree_barrier(bar, G, wg, gen, &s_fail);
        if (s_fail) break;
    }
    if (wg == 0 && threadIdx.x == 0) {
        // the right-hand side with the largest relative residual among the
        // nonzero ones (a zero one is solved exactly by x = 0)
        bool converged = true;
        double best = -1.0, rsel = rr64[0], bsel = bb[0];
#pragma unroll
        for (int k = 0; k < NR; k++) {
            converged &= conv[k];
            const double q = bb[k] > 0 ? rr64[k]/bb[k] : -1.0;
            if (q > best) { best = q; rsel = rr64[k]; bsel = bb[k]; }
        }
        sc->res_norm = sqrt(rsel);
        sc->rhs_norm = sqrt(bsel);
        sc->iters_inner = inner_total;
        sc->iters_outer = outer;
        sc->total_inner += inner_total;
        if (!converged) atomicOr(&sc->error, E_NOT_CONVERGED);
    }
    if (threadIdx.x == 0 && s_fail) atomicOr(&sc->error, E_SYNC_TIMEOUT);
}

int lds_solver_max_groups()
{
    static int cached = 0;
    if (!cached) {
        int dev = 0, ncu = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_solve_lds<2>, LS_BLOCK, 0) != hipSuccess || per < 1) per = 1;
        cached = ncu*per;
    }
    return cached;
}

// One workgroup of 1024 threads per 128 block rows; all workgroups of the
// launch (every instance in the job table) co-resident.  nrhs (1 or 2)
// right-hand sides at b + k*rs, solutions at x + k*rs.
hipError_t solve_lds(const BsrView& A, const double* b, double* x, int nrhs, size_t rs, SolverWork& w,
                     DevScalars* sc, const SolveJob* jobs, int nwg, const SolveOpts& o, hipStream_t s)
{
    if (nwg > lds_solver_max_groups()) return hipErrorCooperativeLaunchTooLarge;
    if (nrhs != 1 && nrhs != 2) return hipErrorInvalidValue;
    BsrView Av = A;
    SolverWork wv = w;
    SolveOpts ov = o;
    // no stale round tags in any instance's partial records
    launch_fill_double(w.partials, 0.0, LDS_PART_PER_BLOCK*(size_t)((A.n+LS_ROWS-1)/LS_ROWS), s);
    void* args[] = {&Av, (void*)&b, (void*)&x, &rs, &wv, &sc, &ov, (void*)&jobs};
    const void* fn = nrhs == 2 ? (const void*)k_solve_lds<2> : (const void*)k_solve_lds_one;
    return launch_persistent(fn, nwg, LS_BLOCK, args, o, lds_solver_max_groups(), s);
}

int persistent_groups(const SolveOpts& o)
{
    static int cached[9] = {0};
    int lpr = o.lanes_per_row;
    if (lpr != 1 && lpr != 2 && lpr != 4 && lpr != 8) lpr = 4;
    if (!cached[lpr]) {
        int dev = 0, ncu = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        hipError_t e = hipErrorInvalidValue;
        switch (lpr) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_solve_mixed<1>, PS_BLOCK, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_solve_mixed<2>, PS_BLOCK, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_solve_mixed<4>, PS_BLOCK, 0); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_solve_mixed<8>, PS_BLOCK, 0); break;
        }
        if (e != hipSuccess || per < 1) per = 1;
        cached[lpr] = ncu*per;
    }
    return cached[lpr];
}

hipError_t solve_mixed(const BsrView& A, const double* b, double* x, SolverWork& w, DevScalars* sc,
                       const SolveOpts& o, hipStream_t s)
{
    int lpr = o.lanes_per_row;
    if (lpr != 1 && lpr != 2 && lpr != 4 && lpr != 8) lpr = 4;
    int rows_per_block = PS_BLOCK/lpr;
    int need = (A.n+rows_per_block-1)/rows_per_block;
    int maxg = persistent_groups(o);
    int G = o.groups > 0 ? o.groups : need;
    if (G > maxg) G = maxg;
    if (G > w.max_groups) G = w.max_groups;
    if (G < 1) G = 1;
    BsrView Av = A;
    SolverWork wv = w;
    SolveOpts ov = o;
    void* args[] = {&Av, (void*)&b, (void*)&x, &wv, &sc, &ov};
    const void* fn = nullptr;
    switch (lpr) {
    case 1: fn = (const void*)k_solve_mixed<1>; break;
    case 2: fn = (const void*)k_solve_mixed<2>; break;
    case 8: fn = (const void*)k_solve_mixed<8>; break;
    default: fn = (const void*)k_solve_mixed<4>; break;
    }
    // cooperative always: the kernel's cg grid barrier needs the workspace
    // that only a cooperative launch provides
    return hipLaunchCooperativeKernel(fn, dim3(G), dim3(PS_BLOCK), args, 0, s);
}

// ------------------------------------------------------------------ plain FP64 PCG, one kernel per operation
__global__ void k_residual64(BsrView A, const double* b, const double* x, double* r)
{
    int g = blockIdx.x*blockDim.x+threadIdx.x;
    int i = g/4, sub = g%4;
    if (i >= A.n) return;
    double y[6];
    spmv_row<double, double, 4>(A, A.od64, A.diag64+15*(size_t)i, x, i, sub, y);
    if (sub == 0) for (int c = 0; c < 6; c++) r[6*(size_t)i+c] = b[6*(size_t)i+c]-y[c];
}
__global__ void k_precond64(BsrView A, const double* r, double* z)
{
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= A.n) return;
    const double* Mi = A.minv64+36*(size_t)i;
    for (int a = 0; a < 6; a++) {
        double v = 0;
        for (int c = 0; c < 6; c++) v += Mi[6*a+c]*r[6*(size_t)i+c];
        z[6*(size_t)i+a] = v;
    }
}
__global__ void k_dot64_partial(const double* a, const double* b, size_t m, double* part)
{
    __shared__ double lds[32];
    double s = 0, t = 0;
    for (size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x; k < m; k += (size_t)gridDim.x*blockDim.x)
        s += a[k]*b[k];
    block_sum2(s, t, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void k_dot64_final(const double* part, int G, double* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int g = 0; g < G; g++) s += part[g];
        *out = s;
    }
}
__global__ void k_axpy64(double* y, const double* x, double a, size_t m)
{
    size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (k < m) y[k] += a*x[k];
}
__global__ void k_xpay64(double* p, const double* z, double beta, size_t m)
{
    size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (k < m) p[k] = z[k]+beta*p[k];
}

static double host_dot(const double* a, const double* b, size_t m, double* part, double* out_dev, hipStream_t s)
{
    int G = 64;
    k_dot64_partial<<<G, 256, 0, s>>>(a, b, m, part);
    k_dot64_final<<<1, 64, 0, s>>>(part, G, out_dev);
    double v = 0;
    (void)hipMemcpyAsync(&v, out_dev, sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    return v;
}

hipError_t solve_plain(const BsrView& A, const double* b, double* x, SolverWork& w, DevScalars* sc,
                       const SolveOpts& o, hipStream_t s, int* iters_out, double* res_out)
{
    size_t m = 6*(size_t)A.n;
    unsigned gm = (unsigned)((m+255)/256);
    double* scal = w.partials+w.max_groups*8-1;
    k_residual64<<<(A.n*4+255)/256, 256, 0, s>>>(A, b, x, w.pr);
    k_precond64<<<(A.n+255)/256, 256, 0, s>>>(A, w.pr, w.pz);
    (void)hipMemcpyAsync(w.pp, w.pz, m*sizeof(double), hipMemcpyDeviceToDevice, s);
    double bb = host_dot(b, b, m, w.partials, scal, s);
    if (bb == 0) {                         // zero right-hand side: the solution is zero
        launch_fill_double(x, 0.0, m, s);
        if (iters_out) *iters_out = 0;
        if (res_out) *res_out = 0;
        return hipGetLastError();
    }
    double rz = host_dot(w.pr, w.pz, m, w.partials, scal, s);
    double rr = host_dot(w.pr, w.pr, m, w.partials, scal, s);
    double stop = o.tol*o.tol*bb;
    int it = 0;
    int maxit = o.max_inner*(o.max_outer > 0 ? o.max_outer : 1);
    while (rr > stop && it < maxit) {
        launch_spmv64(A, w.pp, w.pq, s);
        double pq = host_dot(w.pp, w.pq, m, w.partials, scal, s);
        if (!(pq > 0)) break;
        double alpha = rz/pq;
        k_axpy64<<<gm, 256, 0, s>>>(x, w.pp, alpha, m);
        k_axpy64<<<gm, 256, 0, s>>>(w.pr, w.pq, -alpha, m);
        rr = host_dot(w.pr, w.pr, m, w.partials, scal, s);
        it++;
        if (rr <= stop) break;
        k_precond64<<<(A.n+255)/256, 256, 0, s>>>(A, w.pr, w.pz);
        double rz_new = host_dot(w.pr, w.pz, m, w.partials, scal, s);
        double beta = rz_new/rz;
        rz = rz_new;
        k_xpay64<<<gm, 256, 0, s>>>(w.pp, w.pz, beta, m);
    }
    // true residual
    k_residual64<<<(A.n*4+255)/256, 256, 0, s>>>(A, b, x, w.pr);
    double rt = host_dot(w.pr, w.pr, m, w.partials, scal, s);
    if (iters_out) *iters_out = it;
    if (res_out) *res_out = bb > 0 ? sqrt(rt/bb) : 0;
    (void)sc;
    return hipGetLastError();
}

// ------------------------------------------------------------------ plain FP64 PCG, scalars on the device
// The iteration of solve_plain with its scalars kept on the device: each dot
// product ends in a one-thread kernel that also forms alpha or beta, and the
// vector kernels read them from device memory, so an iteration has no host
// round trip (8 kernels).  The host reads the state every PD_CHECK iterations;
// once the residual is below the stop (or pq is not positive, or the
// iteration limit is reached) a device flag turns the rest of the batch into
// no-ops, so the iteration count and the result are those of solve_plain.
// Scalars (doubles at ps): rz, pq, rr, alpha, beta, stop, done, it, maxit, bb.
enum { PD_RZ, PD_PQ, PD_RR, PD_ALPHA, PD_BETA, PD_STOP, PD_DONE, PD_IT, PD_MAXIT, PD_BB, PD_NSCAL };
static constexpr int PD_G = 64, PD_CHECK = 16;

__global__ void k_dot2_partial(const double* a, const double* b, const double* c, const double* d, size_t m,
                               double* part)
{
    __shared__ double lds[32];
    double s = 0, t = 0;
    for (size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x; k < m; k += (size_t)gridDim.x*blockDim.x) {
        s += a[k]*b[k];
        if (c) t += c[k]*d[k];
    }
    block_sum2(s, t, lds);
    if (threadIdx.x == 0) { part[2*blockIdx.x] = s; part[2*blockIdx.x+1] = t; }
}
__device__ inline void pd_sums(const double* part, double& s, double& t)
{
    s = 0; t = 0;
    for (int g = 0; g < PD_G; g++) { s += part[2*g]; t += part[2*g+1]; }
}
__global__ void k_pd_sum(const double* part, double* out)
{
    double s, t;
    pd_sums(part, s, t);
    *out = s;
}
// start: part holds (b.b, r.z), part2 (r.r, -)
__global__ void k_pd_start(const double* part, const double* part2, double* ps, double tol2, double maxit)
{
    double bb, rz, rr, z;
    pd_sums(part, bb, rz);
    pd_sums(part2, rr, z);
    ps[PD_BB] = bb; ps[PD_RZ] = rz; ps[PD_RR] = rr; ps[PD_STOP] = tol2*bb;
    ps[PD_IT] = 0; ps[PD_MAXIT] = maxit;
    ps[PD_DONE] = (bb == 0 || rr <= tol2*bb || maxit <= 0) ? 1.0 : 0.0;
}
__global__ void k_pd_alpha(const double* part, double* ps)
{
    if (ps[PD_DONE] != 0) return;
    double pq, z;
    pd_sums(part, pq, z);
    ps[PD_PQ] = pq;
    if (!(pq > 0)) { ps[PD_DONE] = 1; return; }
    ps[PD_ALPHA] = ps[PD_RZ]/pq;
}
__global__ void k_pd_update_xr(double* x, double* r, const double* p, const double* q, const double* ps, size_t m)
{
    if (ps[PD_DONE] != 0) return;
    const double a = ps[PD_ALPHA];
    size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (k < m) { x[k] += a*p[k]; r[k] += (-a)*q[k]; }
}
// part holds (r.r, r.z)
__global__ void k_pd_beta(const double* part, double* ps)
{
    if (ps[PD_DONE] != 0) return;
    double rr, rz_new;
    pd_sums(part, rr, rz_new);
    ps[PD_RR] = rr;
    ps[PD_IT] += 1;
    if (rr <= ps[PD_STOP] || ps[PD_IT] >= ps[PD_MAXIT]) { ps[PD_DONE] = 1; return; }
    ps[PD_BETA] = rz_new/ps[PD_RZ];
    ps[PD_RZ] = rz_new;
}
__global__ void k_pd_update_p(double* p, const double* z, const double* ps, size_t m)
{
    if (ps[PD_DONE] != 0) return;
    const double beta = ps[PD_BETA];
    size_t k = blockIdx.x*(size_t)blockDim.x+threadIdx.x;
    if (k < m) p[k] = z[k]+beta*p[k];
}

hipError_t solve_plain_dev(const BsrView& A, const double* b, double* x, SolverWork& w, DevScalars* sc,
                           const SolveOpts& o, hipStream_t s, int* iters_out, double* res_out)
{
    size_t m = 6*(size_t)A.n;
    unsigned gm = (unsigned)((m+255)/256);
    double* part = w.partials;                    // 2 PD_G partial sums
    double* part2 = w.partials+2*PD_G;
    double* ps = w.partials+4*PD_G;               // PD_NSCAL scalars
    const int maxit = o.max_inner*(o.max_outer > 0 ? o.max_outer : 1);
    k_residual64<<<(A.n*4+255)/256, 256, 0, s>>>(A, b, x, w.pr);
    k_precond64<<<(A.n+255)/256, 256, 0, s>>>(A, w.pr, w.pz);
    (void)hipMemcpyAsync(w.pp, w.pz, m*sizeof(double), hipMemcpyDeviceToDevice, s);
    k_dot2_partial<<<PD_G, 256, 0, s>>>(b, b, w.pr, w.pz, m, part);
    k_dot2_partial<<<PD_G, 256, 0, s>>>(w.pr, w.pr, nullptr, nullptr, m, part2);
    k_pd_start<<<1, 1, 0, s>>>(part, part2, ps, o.tol*o.tol, (double)maxit);
    double h[PD_NSCAL];
    (void)hipMemcpyAsync(h, ps, sizeof(h), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    if (h[PD_BB] == 0) {                          // zero right-hand side: the solution is zero
        launch_fill_double(x, 0.0, m, s);
        if (iters_out) *iters_out = 0;
        if (res_out) *res_out = 0;
        (void)sc;
        return hipGetLastError();
    }
    while (h[PD_DONE] == 0) {
        for (int k = 0; k < PD_CHECK; k++) {
            launch_spmv64(A, w.pp, w.pq, s);
            k_dot2_partial<<<PD_G, 256, 0, s>>>(w.pp, w.pq, nullptr, nullptr, m, part);
            k_pd_alpha<<<1, 1, 0, s>>>(part, ps);
            k_pd_update_xr<<<gm, 256, 0, s>>>(x, w.pr, w.pp, w.pq, ps, m);
            k_precond64<<<(A.n+255)/256, 256, 0, s>>>(A, w.pr, w.pz);
            k_dot2_partial<<<PD_G, 256, 0, s>>>(w.pr, w.pr, w.pr, w.pz, m, part);
            k_pd_beta<<<1, 1, 0, s>>>(part, ps);
            k_pd_update_p<<<gm, 256, 0, s>>>(w.pp, w.pz, ps, m);
        }
        (void)hipMemcpyAsync(h, ps, sizeof(h), hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
    }
    // true residual
    k_residual64<<<(A.n*4+255)/256, 256, 0, s>>>(A, b, x, w.pr);
    k_dot2_partial<<<PD_G, 256, 0, s>>>(w.pr, w.pr, nullptr, nullptr, m, part);
    k_pd_sum<<<1, 1, 0, s>>>(part, ps+PD_NSCAL);
    double rt = 0;
    (void)hipMemcpyAsync(&rt, ps+PD_NSCAL, sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    if (iters_out) *iters_out = (int)h[PD_IT];
    if (res_out) *res_out = sqrt(rt/h[PD_BB]);
    (void)sc;
    return hipGetLastError();
}

}  // namespace lfd
