// lfd_agg_large.hip -- the aggregate-preconditioned persistent solver for
// instances too large for the 8-wavefront form (gfx950), with one or two
// right-hand sides per launch.
//
// Why: the 8-wavefront form keeps each lane's row of its wave's inverse in
// registers, so one workgroup holds 8 waves (64 solver slots) and the device
// at most 2048 waves at once; its persistent launch needs every workgroup
// resident.  An instance of N = 32000 particles packs into about 4100 waves.
// Here a workgroup holds 16 waves (128 slots, one per row lane group, as
// the 6x6 block-Jacobi solver does) with the 16 wave inverses staged in LDS
// (75 KB), so 256 workgroups hold 4096 waves.
//
// Eisenstat's form of the product.  Every SpMV of preconditioned conjugate
// gradients is applied to a preconditioned vector, u = W^-1 r with W the
// block diagonal of the waves, and A = W + E with E the couplings between
// waves, so A u = r + E u.  Only the couplings between waves are staged and
// multiplied: the blocks inside a wave (every particle's own block and the
// couplings within its group) are held by the inverse.  With the FP32
// inverse the inner iteration solves with W~ + E, W~ the exact inverse of the
// stored FP32 inverse; it differs from A by about kappa(W) 6e-8 relative to
// W, far below the 1e-2 drop of an inner solve, and the FP64 refinement
// residual b - A x uses the FP64 blocks of A itself.  A wave whose inverse
// fell back to 6x6 block-Jacobi (wexact = 0) keeps all of its couplings in
// the product, where W is then the block diagonal of its particles.
//
// Refinement as in the 6x6 block-Jacobi solver: x is updated in place at
// the end of each pass (a grid barrier), the residual of the next pass is
// formed from the FP64 blocks in HBM in particle order, the Galerkin
// correction onto the three uniform translations precedes each inner solve;
// pipelined PCG in FP32 with FP64 sums, classic PCG once a pass fails to
// halve the residual.
#include <hip/hip_runtime.h>
#include "lfd_internal.hpp"
#include "lfd_device.hpp"

namespace lfd {

static constexpr int AL_WAVES = AGG_LARGE_WAVES;
static constexpr int AL_BLOCK = AL_WAVES*WAVE;
static constexpr int AL_LPR = 8;
static constexpr int AL_ROWS = AL_BLOCK/AL_LPR;      // solver slots per workgroup
static constexpr int AL_ENTRY_CAP = 1120;            // staged couplings per workgroup (83 KB)
static_assert(AL_ROWS == AL_WAVES*AGG_CAP, "one wave per wavefront");

struct AlRow { int e0, ne; };       // staged entries of a slot row (ne < 0: not staged)

// nr right-hand sides at b + k rs (solutions at x + k rs, particle order);
// the exchanged vectors of right-hand side k at zb / qb + 6 k scap (slot order).
template <int NR>
__global__ __launch_bounds__(AL_BLOCK) void k_solve_agg_large(BsrView A, AggView Gv, const double* __restrict__ b,
    double* __restrict__ x, size_t rs, SolverWork w, DevScalars* scs, SolveOpts o, const SolveJob* __restrict__ jobs)
{
    constexpr int NV = 3*NR;                         // values of one grid sum
    static_assert(NV <= REC_STRIDE, "a record fits its slot");
    __shared__ int ent_j[AL_ENTRY_CAP];             // partner slot; ~slot for the transposed block
    __shared__ float ent_b[AL_ENTRY_CAP*18];
    __shared__ float s_inv[AL_WAVES*AGG_PACK];
    __shared__ AlRow rows[AL_ROWS];
    __shared__ double red[NV*AL_WAVES+NV];
    __shared__ double s_ei[9];
    __shared__ int s_fail;
    const SolveJob J = jobs[blockIdx.x];
    const int wg = J.lwg, G = J.G;
    const int nw = Gv.waves[J.inst];
    DevScalars* sc = scs+J.inst;
    unsigned* bar = w.barrier+128*(size_t)J.key;
    unsigned seq = 0;
    const int lr = threadIdx.x/AL_LPR, c = threadIdx.x%AL_LPR;
    const unsigned lane = threadIdx.x & (WAVE-1), wid = threadIdx.x/WAVE;
    // the instance's waves are dealt to its workgroups in turn, so that each
    // holds an even share of every part of the packing (wave sizes and
    // coupling counts vary along it)
    const int wl = (int)wid*G+wg;
    const bool wlive = wl < nw;
    const int wave = J.row0/AGG_CAP+wl;
    const int slot = J.row0+AGG_CAP*wl+(int)(lane/AL_LPR);
    const int i = wlive ? Gv.perm[slot] : -1;
    const bool live = i >= 0;
    const bool own = live && c < 6;
    const size_t gi = 6*(size_t)(live ? i : 0)+c;        // particle order (b, x)
    const size_t gs = 6*(size_t)(wlive ? slot : 0)+c;    // slot order (exchanged vectors)
    const size_t vs = 6*(size_t)Gv.scap;                 // between right-hand sides
    const int kk = c < 6 ? 6*(int)(lane/AL_LPR)+c : 0;   // row of the wave block
    const bool exact = wlive && Gv.wexact[wave] != 0;
    unsigned gen = ld_acquire(bar+32);
    if (threadIdx.x == 0) s_fail = 0;
    // ---- the wave's inverse, then the couplings of each row to other waves
    {
        const float* src = Gv.inv+(size_t)(wlive ? wave : 0)*AGG_PACK;
        float* dst = s_inv+wid*AGG_PACK;
        for (int a = (int)lane; a < AGG_PACK; a += WAVE) dst[a] = wlive ? src[a] : 0.f;
    }
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
    for (int e = c; e < nt; e += AL_LPR) {
        int k;
        if (kept(entry(e, k))) mine++;
    }
    int pre = mine;                                  // inclusive prefix over the row's 8 lanes
    for (int d = 1; d < AL_LPR; d <<= 1) {
        const int t = __shfl_up(pre, d, AL_LPR);
        if (c >= d) pre += t;
    }
    const int rowtot = __shfl(pre, AL_LPR-1, AL_LPR);
    pre -= mine;
    if (c == 0) rows[lr].ne = live ? rowtot : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        int used = 0;
        for (int t = 0; t < AL_ROWS; t++) {
            const int deg = rows[t].ne;
            if (deg > 0 && used+deg <= AL_ENTRY_CAP) { rows[t].e0 = used; used += deg; }
            else { rows[t].e0 = 0; rows[t].ne = deg > 0 ? -1 : 0; }
        }
    }
    __syncthreads();
    const int ne = live ? rows[lr].ne : 0, e0 = live ? rows[lr].e0 : 0;
    if (ne > 0) {
        int s = e0+pre;
        for (int e = c; e < nt; e += AL_LPR) {
            int k;
            const int jj = entry(e, k);
            if (!kept(jj)) continue;
            ent_j[s] = jj;
            for (int m = 0; m < 18; m++) ent_b[18*s+m] = A.od32[18*(size_t)k+m];
            s++;
        }
    }
    __syncthreads();

    // out[k] = yin[k] + (E v_k) at this row, component c: neighbours' v_k
    // from vbuf + k vs (slot order)
    auto spmv_e = [&](const float* vbuf, const float* yin, float* out) {
        float y[NR][6];
        for (int k = 0; k < NR; k++)
            for (int m = 0; m < 6; m++) y[k][m] = 0;
        if (live) {
            if (ne >= 0) {
                for (int e = c; e < ne; e += AL_LPR) {
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
                for (int e = c; e < nt; e += AL_LPR) {
                    int k0;
                    int jj = entry(e, k0);
                    if (!kept(jj)) continue;
                    const bool tr = jj < 0;
                    if (tr) jj = ~jj;
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
            float me = 0;
            for (int m = 0; m < 6; m++) {
                const float t = sum8f(y[k][m]);
                if (m == c) me = t;
            }
            out[k] = own ? yin[k]+me : 0.f;
        }
    };
    // z_k = W^-1 r_k on the lane's wave: component m of r_k is read from its
    // lane (8 (m / 6) + m % 6) into a scalar register; one LDS read of the
    // inverse serves every right-hand side
    auto precond = [&](const float* rv, float* z) {
        const float* Mw = s_inv+wid*AGG_PACK;
        int kq = kk;
        asm volatile("" : "+v"(kq));   // the addresses are recomputed per call, not held in registers
        const int kb = kq*(kq+1)/2;
        int xr[NR];
        for (int k = 0; k < NR; k++) { xr[k] = __float_as_int(own ? rv[k] : 0.f); z[k] = 0.f; }
#pragma unroll
        for (int m = 0; m < AGG_DIM; m++) {
            const int a = m <= kq ? kb+m : m*(m+1)/2+kq;
            const float Mv = Mw[a];
            for (int k = 0; k < NR; k++)
                z[k] += Mv*__int_as_float(__builtin_amdgcn_readlane(xr[k], 8*(m/6)+m%6));
        }
        for (int k = 0; k < NR; k++) z[k] = own ? z[k] : 0.f;
    };

    double* part = Gv.part+AGG_PART_PER_BLOCK*(size_t)J.key;
    const int rst = rec_stride(G, NV);   // record stride (rs: the right-hand sides' stride)
    int buf = 0;
    // Grid sums of NV values: block sums (fixed order) published as tagged
    // records after every wave's vector stores have completed and a release
    // fence; wave 0 polls every record with device-scope loads until each
    // carries this round's tag, sums them, and issues the acquire fence that
    // makes the other workgroups' vector stores visible.
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
                t[k] = lane < AL_WAVES ? red[NV*lane+k] : 0.0;
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
                bool dn = true;
                for (int q = (int)lane; q < G; q += WAVE) {
                    unsigned long long vq[NV];
                    for (int k = 0; k < NV; k++) vq[k] = ld_dev(pq+(size_t)rst*q+k);
                    for (int k = 0; k < NV; k++) dn &= (vq[k] & 0xFF) == tg;
                    for (int k = 0; k < NV; k++) a[k] += untag_partial(vq[k]);
                }
                if (__all(dn)) {
                    for (int k = 0; k < NV; k++) a[k] = wave_sum_d(a[k]);
                    if (lane == 0)
                        for (int k = 0; k < NV; k++) red[NV*AL_WAVES+k] = a[k];
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
        for (int k = 0; k < NV; k++) t[k] = red[NV*AL_WAVES+k];
        buf ^= 1;
    };
    auto gsum = [&](double* v) {
        publish(v);
        gather(v);
    };

    // Coarse space of the three uniform translations: (A z_a) at this row in
    // registers, E = Z^T A Z by grid sums, E^-1 in LDS
    double AZ[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        double y[6] = {0, 0, 0, 0, 0, 0};
        if (live) row_translation<AL_LPR>(A, i, c, a, y);
        for (int m = 0; m < 6; m++) if (m == c && own) AZ[a] = y[m];
    }
    bool use_coarse;
    {
        double ev[6];
        const double e6[6] = {c == 0 ? AZ[0] : 0.0, c == 1 ? AZ[1] : 0.0, c == 2 ? AZ[2] : 0.0,
                              c == 0 ? AZ[1] : 0.0, c == 0 ? AZ[2] : 0.0, c == 1 ? AZ[2] : 0.0};
        for (int h = 0; h < 6; h += NV) {
            double v[NV];
            for (int k = 0; k < NV; k++) v[k] = h+k < 6 ? e6[h+k] : 0.0;
            gsum(v);
            for (int k = 0; k < NV && h+k < 6; k++) ev[h+k] = v[k];
        }
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
            if (live) spmv_row<double, double, AL_LPR>(A, A.od64, A.diag64+15*(size_t)i, x+k*rs, i, c, yv);
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
                    double sh = 0;
                    for (int a = 0; a < 3; a++) if (a == c) sh = cv[a];
                    if (c < 3) x[k*rs+gi] += sh;
                    r64[k] -= AZ[0]*cv[0]+AZ[1]*cv[1]+AZ[2]*cv[2];
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
            float* mbuf[2] = {Gv.zb, Gv.qb};
            float u[NR], wv[NR], zv[NR], qv[NR], sv[NR], pv[NR], mv[NR], nv[NR];
            double stop2[NR], gamma_old[NR], alpha_old[NR];
            precond(r, u);
            for (int k = 0; k < NR; k++) {
                if (own) mbuf[1][k*vs+gs] = u[k];
                v[3*k] = (double)r[k]*r[k]; v[3*k+1] = 0; v[3*k+2] = 0;
            }
            gsum(v);
            for (int k = 0; k < NR; k++) {
                stop2[k] = uni(inner_stop2(o.inner_tol, v[3*k], o.tol, bb[k], rr64[k]));
                zv[k] = qv[k] = sv[k] = pv[k] = 0;
                gamma_old[k] = alpha_old[k] = 1;
            }
            spmv_e(mbuf[1], r, wv);              // A u = r + E u
            int par = 0;
            for (int it = 0; it < o.max_inner; it++) {
                precond(wv, mv);
                for (int k = 0; k < NR; k++) {
                    if (own) mbuf[par][k*vs+gs] = mv[k];
                    v[3*k] = (double)r[k]*u[k]; v[3*k+1] = (double)wv[k]*u[k]; v[3*k+2] = (double)r[k]*r[k];
                }
                publish(v);
                spmv_e(mbuf[par], wv, nv);       // A m = w + E m
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
                        const double den = dl-beta*g/alpha_old[k];
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
            // classic PCG, two grid sums per iteration
            float z[NR], p[NR], q[NR], wz[NR], beta[NR];
            double rz[NR], pq[NR], stop2[NR];
            precond(r, z);
            for (int k = 0; k < NR; k++) {
                p[k] = z[k];
                if (own) Gv.zb[k*vs+gs] = z[k];
                v[3*k] = (double)r[k]*z[k]; v[3*k+1] = (double)r[k]*r[k]; v[3*k+2] = 0;
            }
            gsum(v);
            for (int k = 0; k < NR; k++) {
                rz[k] = v[3*k];
                stop2[k] = inner_stop2(o.inner_tol, v[3*k+1], o.tol, bb[k], rr64[k]);
            }
            spmv_e(Gv.zb, r, q);                 // A z = r + E z
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
                    itk[k]++;
                    any = true;
                }
                if (!any) break;
                precond(r, z);
                for (int k = 0; k < NR; k++) {
                    if (own && act[k]) Gv.zb[k*vs+gs] = z[k];
                    v[3*k] = (double)r[k]*z[k]; v[3*k+1] = (double)r[k]*r[k]; v[3*k+2] = 0;
                }
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
                spmv_e(Gv.zb, r, wz);            // A z = r + E z
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

int agg_large_max_groups()
{
    static int cached = 0;
    if (!cached) {
        int dev = 0, ncu = 0, p1 = 0, p2 = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&p1, k_solve_agg_large<1>, AL_BLOCK, 0) != hipSuccess) p1 = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&p2, k_solve_agg_large<2>, AL_BLOCK, 0) != hipSuccess) p2 = 1;
        cached = ncu*std::max(1, std::min(p1, p2));
    }
    return cached;
}

hipError_t solve_agg_large(const BsrView& A, const AggView& G, const double* b, double* x, int nrhs, size_t rs,
                           SolverWork& w, DevScalars* sc, const SolveJob* jobs, int nwg, const SolveOpts& o,
                           hipStream_t s)
{
    if (nwg > agg_large_max_groups()) return hipErrorCooperativeLaunchTooLarge;
    if (nrhs != 1 && nrhs != 2) return hipErrorInvalidValue;
    BsrView Av = A;
    AggView Gv = G;
    SolverWork wv = w;
    SolveOpts ov = o;
    // no stale round tags in any instance's partial records
    launch_fill_double(G.part, 0.0, G.part_len, s);
    void* args[] = {&Av, &Gv, (void*)&b, (void*)&x, &rs, &wv, &sc, &ov, (void*)&jobs};
    const void* fn = nrhs == 2 ? (const void*)k_solve_agg_large<2> : (const void*)k_solve_agg_large<1>;
    return launch_persistent(fn, nwg, AL_BLOCK, args, o, agg_large_max_groups(), s);
}

}  // namespace lfd
