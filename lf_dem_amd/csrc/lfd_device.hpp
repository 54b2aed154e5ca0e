// lfd_device.hpp -- device helpers shared by the HIP translation units of
// the Stokes-resistance library: packed 6x6 block products, block rows of
// the sparse matrix, DPP lane moves and wave sums, and the inter-workgroup
// synchronisation of the persistent solvers.  Wavefronts are 64 lanes wide.
#pragma once
#include <hip/hip_runtime.h>
#include "lfd_internal.hpp"

namespace lfd {

static constexpr int WAVE = 64;

// ------------------------------------------------------------------ block SpMV
// y = D x, y = M x, y = M^T x for LF_DEM's packed blocks (dense layout as in
// StokesSolver::insertDBlockValues / insertODBlockValues).
template <typename T, typename A>
__device__ inline void dmul(const T* d, const A* x, A* y)
{
    y[0] += (A)d[0]*x[0]+(A)d[1]*x[1]+(A)d[2]*x[2]+(A)d[3]*x[4]+(A)d[4]*x[5];
    y[1] += (A)d[1]*x[0]+(A)d[5]*x[1]+(A)d[6]*x[2]-(A)d[3]*x[3]+(A)d[7]*x[5];
    y[2] += (A)d[2]*x[0]+(A)d[6]*x[1]+(A)d[8]*x[2]-(A)d[4]*x[3]-(A)d[7]*x[4];
    y[3] += -(A)d[3]*x[1]-(A)d[4]*x[2]+(A)d[9]*x[3]+(A)d[10]*x[4]+(A)d[11]*x[5];
    y[4] += (A)d[3]*x[0]-(A)d[7]*x[2]+(A)d[10]*x[3]+(A)d[12]*x[4]+(A)d[13]*x[5];
    y[5] += (A)d[4]*x[0]+(A)d[7]*x[1]+(A)d[11]*x[3]+(A)d[13]*x[4]+(A)d[14]*x[5];
}
template <typename T, typename A>
__device__ inline void omul(const T* b, const A* x, A* y)
{
    y[0] += (A)b[0]*x[0]+(A)b[1]*x[1]+(A)b[2]*x[2]-(A)b[9]*x[4]-(A)b[10]*x[5];
    y[1] += (A)b[1]*x[0]+(A)b[5]*x[1]+(A)b[6]*x[2]+(A)b[9]*x[3]-(A)b[14]*x[5];
    y[2] += (A)b[2]*x[0]+(A)b[6]*x[1]+(A)b[8]*x[2]+(A)b[10]*x[3]+(A)b[14]*x[4];
    y[3] += -(A)b[3]*x[1]-(A)b[4]*x[2]+(A)b[11]*x[3]+(A)b[12]*x[4]+(A)b[13]*x[5];
    y[4] += (A)b[3]*x[0]-(A)b[7]*x[2]+(A)b[12]*x[3]+(A)b[15]*x[4]+(A)b[16]*x[5];
    y[5] += (A)b[4]*x[0]+(A)b[7]*x[1]+(A)b[13]*x[3]+(A)b[16]*x[4]+(A)b[17]*x[5];
}
template <typename T, typename A>
__device__ inline void omul_t(const T* b, const A* x, A* y)
{
    y[0] += (A)b[0]*x[0]+(A)b[1]*x[1]+(A)b[2]*x[2]+(A)b[3]*x[4]+(A)b[4]*x[5];
    y[1] += (A)b[1]*x[0]+(A)b[5]*x[1]+(A)b[6]*x[2]-(A)b[3]*x[3]+(A)b[7]*x[5];
    y[2] += (A)b[2]*x[0]+(A)b[6]*x[1]+(A)b[8]*x[2]-(A)b[4]*x[3]-(A)b[7]*x[4];
    y[3] += (A)b[9]*x[1]+(A)b[10]*x[2]+(A)b[11]*x[3]+(A)b[12]*x[4]+(A)b[13]*x[5];
    y[4] += -(A)b[9]*x[0]+(A)b[14]*x[2]+(A)b[12]*x[3]+(A)b[15]*x[4]+(A)b[16]*x[5];
    y[5] += -(A)b[10]*x[0]-(A)b[14]*x[1]+(A)b[13]*x[3]+(A)b[16]*x[4]+(A)b[17]*x[5];
}

// Row i of A x using LPR lanes: the lanes split row i's incident blocks and
// reduce with cross-lane shuffles.  Every lane of the group ends with y.
template <typename T, typename V, int LPR>
__device__ inline void spmv_row(const BsrView& A, const T* od, const T* diag, const V* x, int i, int sub, V* y)
{
    for (int c = 0; c < 6; c++) y[c] = 0;
    const int u0 = A.row_ptr[i], u1 = A.row_ptr[i+1];
    const int l0 = A.tptr[i], l1 = A.tptr[i+1];
    const int nl = l1-l0, nt = nl+(u1-u0);
    for (int e = sub; e < nt; e += LPR) {
        if (e < nl) {
            int k = A.tidx[l0+e];
            int j = A.p0[k];
            V xj[6];
            for (int c = 0; c < 6; c++) xj[c] = x[6*(size_t)j+c];
            omul<T, V>(od+18*(size_t)k, xj, y);       // block (row i = p1, col p0)
        } else {
            int k = u0+(e-nl);
            int j = A.p1[k];
            V xj[6];
            for (int c = 0; c < 6; c++) xj[c] = x[6*(size_t)j+c];
            omul_t<T, V>(od+18*(size_t)k, xj, y);     // transpose (row i = p0, col p1)
        }
    }
    if (sub == 0 && diag) {
        V xi[6];
        for (int c = 0; c < 6; c++) xi[c] = x[6*(size_t)i+c];
        dmul<T, V>(diag, xi, y);
    }
    for (int o = 1; o < LPR; o <<= 1)
        for (int c = 0; c < 6; c++) y[c] += __shfl_xor(y[c], o, LPR);
}

// Row i of A64 applied to the uniform translation e_a (a < 3): the sum of
// column a of the row's blocks (LPR lanes per row, result in every lane).
template <int LPR>
__device__ inline void row_translation(const BsrView& A, int i, int sub, int a, double* y)
{
    double ea[6] = {0, 0, 0, 0, 0, 0};
    ea[a] = 1;
    for (int c = 0; c < 6; c++) y[c] = 0;
    const int u0 = A.row_ptr[i], u1 = A.row_ptr[i+1];
    const int l0 = A.tptr[i], l1 = A.tptr[i+1];
    const int nl = l1-l0, nt = nl+(u1-u0);
    for (int e = sub; e < nt; e += LPR) {
        if (e < nl) omul<double, double>(A.od64+18*(size_t)A.tidx[l0+e], ea, y);
        else omul_t<double, double>(A.od64+18*(size_t)(u0+(e-nl)), ea, y);
    }
    if (sub == 0) dmul<double, double>(A.diag64+15*(size_t)i, ea, y);
    for (int o = 1; o < LPR; o <<= 1)
        for (int c = 0; c < 6; c++) y[c] += __shfl_xor(y[c], o, LPR);
}


// ---- DPP lane moves (a DPP row is 16 lanes)
template <int CTRL> __device__ inline int dpp_i(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }
template <int CTRL> __device__ inline float dpp_f(float v) { return __int_as_float(dpp_i<CTRL>(__float_as_int(v))); }
template <int CTRL> __device__ inline double dpp_d(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(dpp_i<CTRL>(hi), dpp_i<CTRL>(lo));
}
static constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;
// sum over each aligned group of 8 lanes, every lane of the group gets it
__device__ inline float sum8f(float v)
{
    v += dpp_f<DPP_XOR1>(v);
    v += dpp_f<DPP_XOR2>(v);
    v += dpp_f<DPP_HALF_MIRROR>(v);
    return v;
}
__device__ inline double sum16d(double v)
{
    v += dpp_d<DPP_XOR1>(v);
    v += dpp_d<DPP_XOR2>(v);
    v += dpp_d<DPP_HALF_MIRROR>(v);
    v += dpp_d<DPP_MIRROR>(v);
    return v;
}
// wave-uniform sum of a double over the 64 lanes, fixed order
__device__ inline double wave_sum_d(double v)
{
    v = sum16d(v);
    double r = 0;
    for (int k = 0; k < 4; k++) {
        int lo = __builtin_amdgcn_readlane(__double2loint(v), 16*k);
        int hi = __builtin_amdgcn_readlane(__double2hiint(v), 16*k);
        r += __hiloint2double(hi, lo);
    }
    return r;
}
// value of lane M of this lane's 8-lane group (sub = lane % 8)
template <int M> __device__ inline float bcast8(float v, int sub)
{
    constexpr int q = M & 3;
    constexpr int perm = q | (q << 2) | (q << 4) | (q << 6);
    float x = dpp_f<perm>(v);
    float y = dpp_f<DPP_HALF_MIRROR>(x);
    return ((sub < 4) == (M < 4)) ? x : y;
}
__device__ inline void bcast8_6(float v, int sub, float* o)
{
    o[0] = bcast8<0>(v, sub); o[1] = bcast8<1>(v, sub); o[2] = bcast8<2>(v, sub);
    o[3] = bcast8<3>(v, sub); o[4] = bcast8<4>(v, sub); o[5] = bcast8<5>(v, sub);
}

// Relaxed device-scope (sc1) accesses, performed at the device-coherent level.
template <class T> __device__ inline void st_dev(T* p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ inline T ld_dev(const T* p)
{
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until every memory operation this wave has issued is complete.
__device__ inline void drain_vmem()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Release at agent scope ahead of relaxed record stores or barrier arrivals.
// The fence writes back this L2's dirty lines (the vectors that other
// workgroups read once they see the record or the barrier: other XCDs' L2s
// do not snoop this one), and the following store must wait for that
// write-back to complete.  The compiler emits the write-back (buffer_wbl2) of
// a release fence or a release atomic but drops the wait for it when no other
// vector memory operation is outstanding, so a record or an arrival could
// reach another XCD before the data it publishes; the explicit wait restores
// it.  Every release of this library goes through here, followed by relaxed
// atomics.
__device__ inline void release_for_records()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    drain_vmem();
}
// barrier words: [0] root count, [32] generation, [64 + 32*leaf] leaf counts (128-byte spacing)
__device__ inline unsigned ld_acquire(unsigned* p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT); }

// Arrival and wait of one workgroup, run by a single thread after a
// __syncthreads: release before the arrival (release_for_records, then a
// relaxed increment), acquire after the last arrival of a leaf and of the
// root, acquire on the generation load.  The wait gives up after SPIN_LIMIT
// polls and reports false.
static constexpr long SPIN_LIMIT = 1L << 22;
__device__ inline bool tree_barrier_core(unsigned* bar, int G, int wg, unsigned& gen)
{
    const int leaf = wg >> 3, nleaf = (G+7) >> 3;
    const int lsize = min(8, G-(leaf << 3));
    unsigned* lc = bar+64+32*leaf;
    release_for_records();
    unsigned old = __hip_atomic_fetch_add(lc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == (unsigned)(lsize-1)) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(lc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        release_for_records();
        unsigned r = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (r == (unsigned)(nleaf-1)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            release_for_records();
            __hip_atomic_fetch_add(bar+32, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    for (long spin = 0; ld_acquire(bar+32) == gen; spin++) {
        if (spin >= SPIN_LIMIT) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    gen++;
    return true;
}

// A partial sum with the low 8 mantissa bits replaced by a round tag
// (relative perturbation below 3e-14).  Aligned 64-bit stores and loads are
// single-copy atomic, so a record read with the expected tag is that round's.
__device__ inline double tag_partial(double v, unsigned tag)
{
    return __longlong_as_double((__double_as_longlong(v) & ~0xFFll) | (long long)(tag & 0xFF));
}
__device__ inline double untag_partial(unsigned long long v)
{
    return __longlong_as_double((long long)(v & ~0xFFull));
}
// Distance in doubles between the partial records of K values of an
// instance of G workgroups: one 256-byte slot each, or packed.
__device__ inline int rec_stride(int G, int K)
{
    return G <= REC_SPREAD_MAX_G ? REC_STRIDE : K;
}
// One poll of tagged partial records of K values each (record q at
// pq[rs q .. rs q + K)): lane L reads records L, L + 64, L + 128, ... below
// nrec, R rounds at a time with every load of the R rounds issued before any
// is checked, so a poll of up to 64 R records waits for one memory latency
// instead of one per round.  The untagged values are added into acc[0..K) in
// record order; true when every record read carries tag tg.
template <int K, int R>
__device__ inline bool poll_records(const unsigned long long* pq, int nrec, int rs, unsigned tg, unsigned lane,
                                    double* acc)
{
    static_assert(K <= REC_STRIDE, "a record fits its slot");
    bool done = true;
    for (int base = 0; base < nrec; base += R*64) {
        unsigned long long v[R][K];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int q = base+64*r+(int)lane;
#pragma unroll
            for (int k = 0; k < K; k++) v[r][k] = q < nrec ? ld_dev(pq+(size_t)rs*q+k) : (unsigned long long)tg;
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int k = 0; k < K; k++) {
                done &= (v[r][k] & 0xFF) == tg;
                acc[k] += untag_partial(v[r][k]);
            }
    }
    return done;
}
// A workgroup-uniform double held in scalar registers (several right-hand
// sides need every vector register the loop can spare).
__device__ inline double uni(double v)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Squared stop of an inner FP32 solve, in units of the pass's scaled
// residual (||r|| = 1 before the coarse correction; s_r2 after it): a drop
// of inner_tol, or less when that is all the pass needs to bring the FP64
// residual to half the refinement target tol ||b|| (bb = ||b||^2, rr64 the
// pass's ||r||^2), so the last pass does not overshoot the target.
__device__ inline double inner_stop2(double inner_tol, double s_r2, double tol, double bb, double rr64)
{
    return fmax(inner_tol*inner_tol*s_r2, 0.25*tol*tol*bb/rr64);
}

// the 6 components of row j of an exchanged vector (8-byte aligned)
__device__ inline void load6(const float* p, float* o)
{
    const float2* q = reinterpret_cast<const float2*>(p);
    for (int h = 0; h < 3; h++) { float2 v = q[h]; o[2*h] = v.x; o[2*h+1] = v.y; }
}
// Full grid barrier (release / acquire) of the whole workgroup; skipped once
// the workgroup has failed, and *fail raised when the wait gives up.
__device__ inline void tree_barrier(unsigned* bar, int G, int wg, unsigned& gen, int* fail)
{
    __syncthreads();
    if (threadIdx.x == 0 && !*fail && !tree_barrier_core(bar, G, wg, gen)) *fail = 1;
    __syncthreads();
}

}  // namespace lfd
