// batched.hip -- mult of RB density columns at once (RB = 8 or 32) on one plan.
//
// The same pipeline as the single-column mult, with every buffer column-minor: ws[i RB + r],
// result[t RB + r], V[(cone n_hi + q) RB + r], F[(slot n_lo + c) RB + r].  Whatever depends
// on the geometry only is computed once per (target, source), (node, source), (node, child)
// or (target, source box) and applied to the RB columns: the P2P and factored kernel values,
// the upward table entries and cone lookups, the far-field cone location and centre factor.
// The Clenshaw sums read a coefficient row of RB columns per load (coalesced across the RHS
// lanes).  Compile-time interpolation orders only (orders 6 and 8 with tol <= 0, and the
// orders (6,8,7)/(3,5,4) of the tolerance estimate at order 6); other orders run the
// single-column mult per column.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "dev_common.hpp"
#include "device.hpp"

namespace ifgf
{
namespace
{
using namespace dev;

template <typename R>
__global__ void k_gather_b(const double* __restrict__ w_in, const int64_t* __restrict__ perm,
                           Cx<R>* __restrict__ w_out, int64_t n, int rb, int ncol,
                           double im_sign)
{
    // w_in: ncol columns of n values in caller order; columns ncol..rb-1 are zero
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * rb)
        return;
    const int64_t i = e / rb;
    const int r = int(e - i * rb);
    Cx<R> v = {0, 0};
    if (r < ncol)
    {
        const double* p = w_in + 2 * (int64_t(r) * n + perm[i]);
        v = {R(p[0]), R(im_sign * p[1])};
    }
    w_out[e] = v;
}

template <typename R>
__global__ void k_scatter_b(const Cx<R>* __restrict__ res, const int64_t* __restrict__ perm,
                            double* __restrict__ out, int64_t n, int rb, int ncol,
                            double re_sign, double im_sign)
{
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * rb)
        return;
    const int64_t i = e / rb;
    const int r = int(e - i * rb);
    if (r >= ncol)
        return;
    double* p = out + 2 * (int64_t(r) * n + perm[i]);
    p[0] = re_sign * double(res[e].x);
    p[1] = im_sign * double(res[e].y);
}

template <typename R> __global__ void k_zero_b(Cx<R>* p, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = {R(0), R(0)};
}

// value of column r of the fine interpolant in slot `slot` (coefficients RB apart).  Up to 50
// coefficients: loaded, then the Clenshaw sums; more (orders (3,5,4)): the sum of
// c_ijk T_i(x0) T_j(x1) T_k(x2) from the Chebyshev values, a row of coefficients at a time (the
// coefficients of a whole cone would not fit the registers)
template <int O0, int O1, int O2, typename R, int RB>
__device__ __forceinline__ Cx<R> fine_eval_s(const Cx<R>* __restrict__ F, int slot, int r, R x0,
                                             R x1, R x2)
{
    constexpr int N = O0 * O1 * O2;
    const Cx<R>* p = F + int64_t(slot) * N * RB + r;
    if constexpr (N <= 50)
    {
        Cx<R> c[N];
#pragma unroll
        for (int i = 0; i < N; i++)
            c[i] = p[i * RB];
        return clen3c<O0, O1, O2, R>(c, x0, x1, x2);
    }
    else
    {
        R t0[O0], t1[O1], t2[O2];
        t0[0] = t1[0] = t2[0] = R(1);
        if constexpr (O0 > 1)
            t0[1] = x0;
        if constexpr (O1 > 1)
            t1[1] = x1;
        if constexpr (O2 > 1)
            t2[1] = x2;
#pragma unroll
        for (int k = 2; k < O0; k++)
            t0[k] = R(2) * x0 * t0[k - 1] - t0[k - 2];
#pragma unroll
        for (int k = 2; k < O1; k++)
            t1[k] = R(2) * x1 * t1[k - 1] - t1[k - 2];
#pragma unroll
        for (int k = 2; k < O2; k++)
            t2[k] = R(2) * x2 * t2[k - 1] - t2[k - 2];
        Cx<R> out = {0, 0};
#pragma unroll
        for (int i2 = 0; i2 < O2; i2++)
        {
            Cx<R> pl = {0, 0};
#pragma unroll
            for (int i1 = 0; i1 < O1; i1++)
            {
                Cx<R> sm = {0, 0};
#pragma unroll
                for (int i0 = 0; i0 < O0; i0++)
                {
                    const Cx<R> c = p[(i0 + O0 * (i1 + O1 * i2)) * RB];
                    sm.x += c.x * t0[i0];
                    sm.y += c.y * t0[i0];
                }
                pl.x += sm.x * t1[i1];
                pl.y += sm.y * t1[i1];
            }
            out.x += pl.x * t2[i2];
            out.y += pl.y * t2[i2];
        }
        return out;
    }
}

// near field (ifgfoperator.hpp:900-960): as k_near (source leaves staged in LDS, lanes =
// (target, group)), the kernel value of each (source, target) pair applied to the RB columns
template <int KIND, typename R, int RB>
__global__ __launch_bounds__(64) void k_near_b(const WorkItem* __restrict__ items,
                                               int64_t nitems,
                                               const int32_t* __restrict__ near_off,
                                               const int32_t* __restrict__ near_src,
                                               const BoxGeo* __restrict__ geo,
                                               const int64_t* __restrict__ sbeg,
                                               const int64_t* __restrict__ send,
                                               const double* __restrict__ spts,
                                               const double* __restrict__ snrm,
                                               const Cx<R>* __restrict__ w,
                                               const double* __restrict__ tpts,
                                               Cx<R>* __restrict__ result, KParams kp)
{
    constexpr bool NRM = KIND == KD || KIND == KC;
    __shared__ double sp[3 * 64];
    __shared__ double sn[NRM ? 3 * 64 : 1];
    __shared__ Cx<R> sw[64 * RB];
    const int64_t it = xcd_block(kp.xlg);
    if (it >= nitems)
        return;
    const WorkItem wi = items[it];
    const int lane = threadIdx.x;
    const int Tp = pow2_at_least(wi.count);
    const int ti = lane & (Tp - 1);
    const int grp = lane / Tp, ngrp = 64 / Tp;
    const bool valid = ti < wi.count;
    const int64_t t = int64_t(wi.first) + (valid ? ti : 0);
    const double y0 = tpts[3 * t], y1 = tpts[3 * t + 1], y2 = tpts[3 * t + 2];
    Cx<R> acc[RB];
#pragma unroll
    for (int r = 0; r < RB; r++)
        acc[r] = {0, 0};
    const int32_t j0 = near_off[wi.box], j1 = near_off[wi.box + 1];
    for (int32_t j = j0; j < j1; j++)
    {
        const int32_t s = near_src[j];
        const KParams kb = with_k(kp, geo[s]);
        const int64_t b = sbeg[s], e = send[s];
        for (int64_t c0 = b; c0 < e; c0 += 64)
        {
            const int n = int(min<int64_t>(64, e - c0));
            __syncthreads();
            if (lane < n)
            {
                const int64_t i = c0 + lane;
                sp[3 * lane] = spts[3 * i];
                sp[3 * lane + 1] = spts[3 * i + 1];
                sp[3 * lane + 2] = spts[3 * i + 2];
                if constexpr (NRM)
                {
                    sn[3 * lane] = snrm[3 * i];
                    sn[3 * lane + 1] = snrm[3 * i + 1];
                    sn[3 * lane + 2] = snrm[3 * i + 2];
                }
            }
            for (int e2 = lane; e2 < n * RB; e2 += 64)
                sw[e2] = w[c0 * RB + e2];
            __syncthreads();
            for (int k = grp; k < n; k += ngrp)
            {
                const double d0 = sp[3 * k] - y0, d1 = sp[3 * k + 1] - y1, d2 = sp[3 * k + 2] - y2;
                double xn = 0;
                if constexpr (NRM)
                    xn = d0 * sn[3 * k] + d1 * sn[3 * k + 1] + d2 * sn[3 * k + 2];
                const Cx<R> g = p2p<KIND, R>(kb, R(d0), R(d1), R(d2), xn, Cx<R>{1, 0});
#pragma unroll
                for (int r = 0; r < RB; r++)
                    acc[r] = cadd(acc[r], cmul(g, sw[k * RB + r]));
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RB; r++)
        for (int m = Tp; m < 64; m <<= 1)
        {
            acc[r].x += __shfl_xor(acc[r].x, m, 64);
            acc[r].y += __shfl_xor(acc[r].y, m, 64);
        }
    if (grp == 0 && valid)
    {
        Cx<R>* o = result + t * RB;
#pragma unroll
        for (int r = 0; r < RB; r++)
            o[r] = cadd(o[r], acc[r]);
    }
}

// leaf sampling (ifgfoperator.hpp:1057-1073): thread = high-order node, the factored kernel
// of each (node, source) pair applied to the RB columns staged in LDS
template <int KIND, typename R, int RB>
__global__ __launch_bounds__(128) void k_leaf_b(const int32_t* __restrict__ work, int64_t nwork,
                                                const ConeDev* __restrict__ cones,
                                                const BoxGeo* __restrict__ geo,
                                                const int64_t* __restrict__ sbeg,
                                                const int64_t* __restrict__ send,
                                                const double* __restrict__ spts,
                                                const double* __restrict__ snrm,
                                                const Cx<R>* __restrict__ w, NodeTabs tabs,
                                                Cx<R>* __restrict__ V, KParams kp)
{
    constexpr int CH = 64; // sources per LDS chunk
    __shared__ R sx[CH], sy[CH], sz[CH];
    __shared__ R nx[CH], ny[CH], nz[CH];
    __shared__ Cx<R> sw[CH * RB];
    const int64_t wi = xcd_block(kp.xlg);
    if (wi >= nwork)
        return;
    const int32_t g = work[wi];
    const ConeDev cn = cones[g];
    const BoxGeo B = geo[cn.box];
    const KParams kb = with_k(kp, B);
    const int64_t b0 = sbeg[cn.box], b1 = send[cn.box];
    for (int q0 = 0; q0 < kp.n_hi; q0 += 128)
    {
        const int q = q0 + threadIdx.x;
        const bool active = q < kp.n_hi;
        double yd[3] = {0, 0, 0}, rr = 1;
        if (active)
            cone_node(B, cn.cell, q, kp.ho0, kp.ho1, kp.ho2, tabs, false, yd, rr);
        const R y0 = R(yd[0]), y1 = R(yd[1]), y2 = R(yd[2]), dc = R(rr);
        Cx<R> acc[RB];
#pragma unroll
        for (int r = 0; r < RB; r++)
            acc[r] = {0, 0};
        for (int64_t c0 = b0; c0 < b1; c0 += CH)
        {
            const int n = int(min<int64_t>(CH, b1 - c0));
            __syncthreads();
            if (threadIdx.x < n)
            {
                const int64_t i = c0 + threadIdx.x;
                sx[threadIdx.x] = R(spts[3 * i] - B.ctr[0]);
                sy[threadIdx.x] = R(spts[3 * i + 1] - B.ctr[1]);
                sz[threadIdx.x] = R(spts[3 * i + 2] - B.ctr[2]);
                if (snrm)
                {
                    nx[threadIdx.x] = R(snrm[3 * i]);
                    ny[threadIdx.x] = R(snrm[3 * i + 1]);
                    nz[threadIdx.x] = R(snrm[3 * i + 2]);
                }
            }
            for (int e = threadIdx.x; e < n * RB; e += 128)
                sw[e] = w[c0 * RB + e];
            __syncthreads();
            if (active)
                for (int k = 0; k < n; k++)
                {
                    R nn[3] = {0, 0, 0};
                    if (snrm)
                    {
                        nn[0] = nx[k];
                        nn[1] = ny[k];
                        nn[2] = nz[k];
                    }
                    const Cx<R> f = factored<KIND, R>(kb, y0, y1, y2, dc, sx[k], sy[k], sz[k],
                                                      nn, Cx<R>{1, 0});
#pragma unroll
                    for (int r = 0; r < RB; r++)
                        acc[r] = cadd(acc[r], cmul(f, sw[k * RB + r]));
                }
        }
        if (active)
        {
            Cx<R>* o = V + (int64_t(g) * kp.n_hi + q) * RB;
#pragma unroll
            for (int r = 0; r < RB; r++)
                o[r] = acc[r];
        }
    }
}

// upward pass with the per-level tables, parent cones grouped by cell (as k_up_cell):
// thread = (node, column); the table entry and the child cone lookup are shared by the
// column lanes of a node
template <typename R, int O0, int O1, int O2, int K, int RB>
__global__ __launch_bounds__(256) void k_up_cell_b(const WorkItem* __restrict__ items,
                                                   int64_t nitems,
                                                   const int32_t* __restrict__ glist,
                                                   const ConeDev* __restrict__ cones,
                                                   const BoxGeo* __restrict__ geo,
                                                   const int32_t* __restrict__ child,
                                                   const CellRef* __restrict__ ctab,
                                                   const UpEnt<R>* __restrict__ ents,
                                                   const int64_t* __restrict__ part_ent,
                                                   const Cx<R>* __restrict__ F,
                                                   Cx<R>* __restrict__ V, KParams kp)
{
    __shared__ int64_t toff[K][8];
    __shared__ int32_t gs[K];
    const int64_t it = xcd_block(kp.xlg);
    if (it >= nitems)
        return;
    const WorkItem wi = items[it];
    const int tid = threadIdx.x;
    if (tid < K * 8)
    {
        const int k = tid >> 3, c = tid & 7;
        int64_t off = -1;
        if (k < wi.count)
        {
            const int32_t g = glist[wi.first + k];
            if (c == 0)
                gs[k] = g;
            const int32_t ch = child[8 * cones[g].box + c];
            if (ch >= 0)
                off = geo[ch].table_off;
        }
        toff[k][c] = off;
    }
    __syncthreads();
    const int n = wi.count;
    const int r = tid % RB;
    for (int q = tid / RB; q < kp.n_hi; q += 256 / RB)
    {
        const UpEnt<R>* e = ents + part_ent[wi.pad] + (int64_t(wi.box) * kp.n_hi + q) * 8;
        Cx<R> acc[K];
#pragma unroll
        for (int k = 0; k < K; k++)
            acc[k] = {0, 0};
#pragma unroll 1
        for (int c = 0; c < 8; c++)
        {
            const UpEnt<R> E = e[c];
            if (E.cf < 0)
                continue;
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                if (k >= n)
                    continue;
                const int64_t to = toff[k][c];
                if (to < 0)
                    continue;
                const CellRef cr = ctab[to + E.cf];
                if (!((cr.mask >> E.sub) & 1u))
                    continue;
                const int slot = cr.fine_base + __popc(cr.mask & ((1u << E.sub) - 1u));
                const Cx<R> val =
                    fine_eval_s<O0, O1, O2, R, RB>(F, slot, r, E.xi0, E.xi1, E.xi2);
                acc[k] = cadd(acc[k], cmul(val, Cx<R>{E.tr, E.ti}));
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            if (k < n)
                V[(int64_t(gs[k]) * kp.n_hi + q) * RB + r] = acc[k];
    }
}

// LDS cache of the coefficient blocks (N x RB: the N coefficients of one fine cone for the RB
// columns) of up to CACHE distinct fine cones (slots), shared by a workgroup: clear(), then
// every thread insert()s its slot (an open-addressing table in LDS, HASH entries; -1 when
// the table is full), assign() numbers the distinct slots, load() copies their blocks, and
// block() gives a slot's coefficients: in LDS when cached, else in memory (the same values).
// Each step is separated from the next by a workgroup barrier.
template <typename R, int N, int RB, int CACHE, int HASH> struct SlotCache
{
    Cx<R>* coef;      // CACHE * N * RB
    int32_t* hash;    // HASH: slot or -1
    int32_t* cidx;    // HASH: cache index or -1
    int32_t* cached;  // CACHE: slots
    int32_t* ncache;
    __device__ void clear(int tid) const
    {
        if (tid < HASH)
            hash[tid] = -1;
        if (tid == 0)
            *ncache = 0;
    }
    __device__ int insert(int slot) const
    {
        if (slot < 0)
            return -1;
        int at = slot & (HASH - 1);
        for (int probe = 0; probe < HASH; probe++)
        {
            const int32_t prev = atomicCAS(&hash[at], -1, slot);
            if (prev == -1 || prev == slot)
                return at;
            at = (at + 1) & (HASH - 1);
        }
        return -1;
    }
    __device__ void assign(int tid) const
    {
        if (tid < HASH)
        {
            const int32_t sl = hash[tid];
            int ci = -1;
            if (sl >= 0)
            {
                ci = atomicAdd(ncache, 1);
                if (ci < CACHE)
                    cached[ci] = sl;
                else
                    ci = -1;
            }
            cidx[tid] = ci;
        }
    }
    __device__ void load(const Cx<R>* __restrict__ F, int tid, int nthreads) const
    {
        const int nc = min(*ncache, CACHE);
        for (int e = tid; e < nc * N * RB; e += nthreads)
        {
            const int b = e / (N * RB), o = e - b * (N * RB);
            coef[e] = F[int64_t(cached[b]) * N * RB + o];
        }
    }
    __device__ const Cx<R>* block(const Cx<R>* F, int slot, int h) const
    {
        const int ci = h >= 0 ? cidx[h] : -1;
        return ci >= 0 ? coef + ci * N * RB : F + int64_t(slot) * N * RB;
    }
};

// upward pass with tables, order 6 (the batched form of k_up_cell with the children's
// coefficients staged in LDS).  One workgroup per item (cell, up to K parent cones of that
// cell); for each cone and each child in order: the nodes' fine cones in the child (table
// entry + cell table) are collected, the distinct ones (up to UPS_CACHE) copied to LDS once as
// N x RB coefficient blocks, and thread = (node set, column) evaluates every node from there
// (a fine cone beyond the cache is read from memory: the same values).  Without the staging,
// each (node, child, column) read its 18 coefficients from L2, which the 4.6 KB blocks of 32
// columns thrash.
constexpr int UPS_CACHE = 8;  // coefficient blocks staged per (cone, child)
constexpr int UPS_HASH = 32; // open-addressing table of the distinct fine cones
template <typename R, int O0, int O1, int O2, int NH, int K, int RB>
__global__ __launch_bounds__(256) void k_up_stage_b(const WorkItem* __restrict__ items,
                                                    int64_t nitems,
                                                    const int32_t* __restrict__ glist,
                                                    const ConeDev* __restrict__ cones,
                                                    const int32_t* __restrict__ child,
                                                    const BoxGeo* __restrict__ geo,
                                                    const CellRef* __restrict__ ctab,
                                                    const UpEnt<R>* __restrict__ ents,
                                                    const int64_t* __restrict__ part_ent,
                                                    const Cx<R>* __restrict__ F,
                                                    Cx<R>* __restrict__ V, int xlg)
{
    constexpr int N = O0 * O1 * O2;
    constexpr int TQ = 256 / RB;             // node lanes
    constexpr int NQ = (NH + TQ - 1) / TQ;   // nodes per thread
    static_assert(NH <= 256, "one thread per node while collecting");
    __shared__ Cx<R> s_coef[UPS_CACHE * N * RB];
    __shared__ int32_t s_hash[UPS_HASH], s_cidx[UPS_HASH], s_cached[UPS_CACHE], s_ncache;
    const SlotCache<R, N, RB, UPS_CACHE, UPS_HASH> cache{s_coef, s_hash, s_cidx, s_cached,
                                                          &s_ncache};
    __shared__ int32_t s_qslot[NH];          // fine cone of each node in this child (or -1)
    __shared__ int32_t s_qh[NH];             // its hash position
    __shared__ R s_qx[3][NH];
    __shared__ Cx<R> s_qtf[NH];
    const int64_t it = xcd_block(xlg);
    if (it >= nitems)
        return;
    const WorkItem wi = items[it];
    const int tid = threadIdx.x;
    const int r = tid % RB, tq = tid / RB;
    const UpEnt<R>* ebase = ents + part_ent[wi.pad] + int64_t(wi.box) * NH * 8;
    for (int k = 0; k < wi.count; k++)
    {
        const int32_t g = glist[wi.first + k];
        const int32_t pbox = cones[g].box;
        Cx<R> acc[NQ];
#pragma unroll
        for (int j = 0; j < NQ; j++)
            acc[j] = {0, 0};
#pragma unroll 1
        for (int c = 0; c < 8; c++)
        {
            const int32_t ch = child[8 * pbox + c];
            if (ch < 0)
                continue; // uniform: the same child for every thread
            cache.clear(tid);
            __syncthreads();
            // collect: one thread per node
            if (tid < NH)
            {
                const UpEnt<R> E = ebase[int64_t(tid) * 8 + c];
                int slot = -1;
                if (E.cf >= 0)
                {
                    const CellRef cr = ctab[geo[ch].table_off + E.cf];
                    if ((cr.mask >> E.sub) & 1u)
                        slot = cr.fine_base + __popc(cr.mask & ((1u << E.sub) - 1u));
                }
                s_qslot[tid] = slot;
                s_qh[tid] = cache.insert(slot);
                s_qx[0][tid] = E.xi0;
                s_qx[1][tid] = E.xi1;
                s_qx[2][tid] = E.xi2;
                s_qtf[tid] = Cx<R>{E.tr, E.ti};
            }
            __syncthreads();
            cache.assign(tid);
            __syncthreads();
            cache.load(F, tid, 256);
            __syncthreads();
#pragma unroll
            for (int j = 0; j < NQ; j++)
            {
                const int q = tq + j * TQ;
                if (q >= NH)
                    break;
                const int slot = s_qslot[q];
                if (slot < 0)
                    continue;
                const Cx<R>* blk = cache.block(F, slot, s_qh[q]);
                Cx<R> cf[N];
#pragma unroll
                for (int i = 0; i < N; i++)
                    cf[i] = blk[i * RB + r];
                const Cx<R> val = clen3c<O0, O1, O2, R>(cf, s_qx[0][q], s_qx[1][q], s_qx[2][q]);
                acc[j] = cadd(acc[j], cmul(val, s_qtf[q]));
            }
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < NQ; j++)
        {
            const int q = tq + j * TQ;
            if (q < NH)
                V[(int64_t(g) * NH + q) * RB + r] = acc[j];
        }
    }
}

// upward pass without tables (k_up): thread = (node, column); each column lane locates the
// node in the children's grids itself
template <bool PH, typename R, int O0, int O1, int O2, int RB>
__global__ __launch_bounds__(256) void k_up_b(const int32_t* __restrict__ work, int64_t nwork,
                                              const ConeDev* __restrict__ cones,
                                              const BoxGeo* __restrict__ geo,
                                              const int32_t* __restrict__ child,
                                              const CellRef* __restrict__ ctab, NodeTabs tabs,
                                              const Cx<R>* __restrict__ F,
                                              Cx<R>* __restrict__ V, KParams kp)
{
    const int64_t wi = xcd_block(kp.xlg);
    if (wi >= nwork)
        return;
    const int32_t g = work[wi];
    const ConeDev cn = cones[g];
    const BoxGeo& P = geo[cn.box];
    const int r = threadIdx.x % RB;
    for (int q = threadIdx.x / RB; q < kp.n_hi; q += 256 / RB)
    {
        double y[3], rp;
        cone_node(P, cn.cell, q, kp.ho0, kp.ho1, kp.ho2, tabs, true, y, rp);
        Cx<R> acc = {0, 0};
        for (int c = 0; c < 8; c++)
        {
            const int32_t ch = child[8 * cn.box + c];
            if (ch < 0)
                continue;
            const BoxGeo& C = geo[ch];
            Loc L;
            if (!locate(C, ctab, y[0] - C.ctr[0], y[1] - C.ctr[1], y[2] - C.ctr[2], kp.f0,
                        kp.f1, kp.f2, L))
                continue;
            const Cx<R> val = fine_eval_s<O0, O1, O2, R, RB>(F, L.fine, r, R(L.xi[0]),
                                                             R(L.xi[1]), R(L.xi[2]));
            Cx<R> tf;
            if constexpr (!PH)
                tf = {R(rp / L.r), R(0)};
            else
                tf = cscale(expmk<R>(P.k_re, P.k_im, R(L.r - rp)), R(rp / L.r));
            acc = cadd(acc, cmul(val, tf));
        }
        V[(int64_t(g) * kp.n_hi + q) * RB + r] = acc;
    }
}

// the separable values -> fine-coefficients transform (k_transform_t) over instances
// (cone, column): NI instances per workgroup, all of one cone (NI divides RB)
template <typename R, int H0, int H1, int H2, int L0, int L1, int L2, int F0, int F1, int F2,
          int NI, int RB>
__global__ __launch_bounds__(256) void k_transform_b(int64_t c0, int64_t c1,
                                                     const ConeDev* __restrict__ cones,
                                                     const Cx<R>* __restrict__ V,
                                                     Cx<R>* __restrict__ Fout,
                                                     const R* __restrict__ W, int xlg)
{
    static_assert(RB % NI == 0, "instances of a workgroup share one cone");
    constexpr int M0 = F0 * L0, M1 = F1 * L1, M2 = F2 * L2;
    constexpr int NH = H0 * H1 * H2, NLO = L0 * L1 * L2, NSUB = F0 * F1 * F2;
    constexpr int XS = (NH > M0 * M1 * H2) ? NH : M0 * M1 * H2;
    constexpr int YS = (M0 * H1 * H2 > M0 * M1 * M2) ? M0 * H1 * H2 : M0 * M1 * M2;
    constexpr int NW0 = M0 * H0, NW1 = M1 * H1, NW2 = M2 * H2;
    __shared__ Cx<R> X[NI * XS];
    __shared__ Cx<R> Y[NI * YS];
    __shared__ unsigned char subl[32];
    __shared__ int nsub;
    const int tid = threadIdx.x;
    const int64_t inst0 = xcd_block(xlg) * NI; // first (cone, column) instance
    const int64_t cone = c0 + inst0 / RB;
    const int r0 = int(inst0 % RB);
    if (cone >= c1)
        return;
    for (int e = tid; e < NI * NH; e += 256)
    {
        const int k = e % NI, i = e / NI;
        X[k * XS + i] = V[(cone * NH + i) * RB + r0 + k];
    }
    if (tid == 0)
    {
        const uint32_t mask = cones[cone].mask;
        int n = 0;
        for (int s = 0; s < NSUB; s++)
            if ((mask >> s) & 1u)
                subl[n++] = (unsigned char)s;
        nsub = n;
    }
    __syncthreads();
    const R* W0 = W; // scalar loads at compile-time offsets (see k_transform_t)
    const R* W1 = W + NW0;
    const R* W2 = W + NW0 + NW1;
    for (int e = tid; e < NI * H1 * H2; e += 256)
    {
        const int k = e / (H1 * H2), p = e - k * (H1 * H2);
        Cx<R> in[H0];
#pragma unroll
        for (int t = 0; t < H0; t++)
            in[t] = X[k * XS + p * H0 + t];
#pragma unroll
        for (int j = 0; j < M0; j++)
        {
            Cx<R> a = {0, 0};
#pragma unroll
            for (int t = 0; t < H0; t++)
            {
                a.x += W0[j * H0 + t] * in[t].x;
                a.y += W0[j * H0 + t] * in[t].y;
            }
            Y[k * YS + p * M0 + j] = a;
        }
    }
    __syncthreads();
    for (int e = tid; e < NI * M0 * H2; e += 256)
    {
        const int k = e / (M0 * H2), p = e - k * (M0 * H2);
        const int j0 = p % M0, i2 = p / M0;
        Cx<R> in[H1];
#pragma unroll
        for (int t = 0; t < H1; t++)
            in[t] = Y[k * YS + j0 + M0 * (t + H1 * i2)];
#pragma unroll
        for (int j = 0; j < M1; j++)
        {
            Cx<R> a = {0, 0};
#pragma unroll
            for (int t = 0; t < H1; t++)
            {
                a.x += W1[j * H1 + t] * in[t].x;
                a.y += W1[j * H1 + t] * in[t].y;
            }
            X[k * XS + j0 + M0 * (j + M1 * i2)] = a;
        }
    }
    __syncthreads();
    for (int e = tid; e < NI * M0 * M1; e += 256)
    {
        const int k = e / (M0 * M1), p = e - k * (M0 * M1);
        Cx<R> in[H2];
#pragma unroll
        for (int t = 0; t < H2; t++)
            in[t] = X[k * XS + p + M0 * M1 * t];
#pragma unroll
        for (int j = 0; j < M2; j++)
        {
            Cx<R> a = {0, 0};
#pragma unroll
            for (int t = 0; t < H2; t++)
            {
                a.x += W2[j * H2 + t] * in[t].x;
                a.y += W2[j * H2 + t] * in[t].y;
            }
            Y[k * YS + p + M0 * M1 * j] = a;
        }
    }
    __syncthreads();
    const int n = nsub;
    const int base = cones[cone].fine_base;
    for (int e = tid; e < n * NLO * NI; e += 256)
    {
        const int k = e % NI, rest = e / NI;
        const int rk = rest / NLO, coef = rest - rk * NLO;
        const int sub = subl[rk];
        const int s0 = sub % F0, s1 = (sub / F0) % F1, s2 = sub / (F0 * F1);
        const int m0 = coef % L0, m1 = (coef / L0) % L1, m2 = coef / (L0 * L1);
        const int j0 = s0 * L0 + m0, j1 = s1 * L1 + m1, j2 = s2 * L2 + m2;
        Fout[(int64_t(base + rk) * NLO + coef) * RB + r0 + k] =
            Y[k * YS + j0 + M0 * (j1 + M1 * j2)];
    }
}

// far field (k_far) for RB columns: four waves; per step of four source boxes each wave
// locates the item's targets in one box (cone slot, local coordinates, centre factor into
// LDS), then every thread = (target, column) sums the interpolants of those boxes
template <bool PH, typename R, int O0, int O1, int O2, bool FAST, int RB>
__global__ __launch_bounds__(256) void k_far_b(const WorkItem* __restrict__ items,
                                               int64_t nitems,
                                               const int32_t* __restrict__ srange,
                                               const int32_t* __restrict__ far_src,
                                               const BoxGeo* __restrict__ geo,
                                               const CellRef* __restrict__ ctab,
                                               const Cx<R>* __restrict__ F,
                                               const double* __restrict__ tpts,
                                               const double* __restrict__ ang, int ang_base,
                                               int ang_n, Cx<R>* __restrict__ result,
                                               KParams kp)
{
    extern __shared__ double sang[];
    __shared__ int32_t s_slot[4][64];
    __shared__ R s_xi[4][3][64];
    __shared__ Cx<R> s_cf[4][64];
    constexpr int TL = 256 / RB;               // target lanes of the evaluation phase
    constexpr int TPT = (64 + TL - 1) / TL;    // targets per thread
    const int64_t it = xcd_block(kp.xlg);
    if (it >= nitems)
        return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if constexpr (FAST)
    {
        for (int i = tid; i < ang_n; i += 256)
            sang[i] = ang[ang_base + i];
        __syncthreads();
    }
    const WorkItem wi = items[it];
    const bool tvalid = lane < wi.count;
    const int64_t t = int64_t(wi.first) + (tvalid ? lane : 0);
    const double y0 = tpts[3 * t], y1 = tpts[3 * t + 1], y2 = tpts[3 * t + 2];
    const int r = tid % RB, tl = tid / RB;
    Cx<R> acc[TPT];
#pragma unroll
    for (int j = 0; j < TPT; j++)
        acc[j] = {0, 0};
    const int32_t j0 = srange[2 * it], j1 = srange[2 * it + 1];
    for (int32_t jb = j0; jb < j1; jb += 4)
    {
        {
            const int32_t j = jb + wave;
            int slot = -1;
            R x0 = 0, x1 = 0, x2 = 0;
            Cx<R> cf = {0, 0};
            if (j < j1)
            {
                const BoxGeo& B = geo[__builtin_amdgcn_readfirstlane(far_src[j])];
                bool ok = false;
                double rr = 1, ri = 1;
                if (tvalid)
                {
                    if constexpr (FAST)
                    {
                        LocF L;
                        const double* eth = sang + (2 * B.fth_off - ang_base);
                        const double* eph = sang + (2 * B.fph_off - ang_base);
                        ok = locate_fast<3>(B, eth, eph, ctab, y0 - B.ctr[0], y1 - B.ctr[1],
                                            y2 - B.ctr[2], kp.f0, kp.f1, kp.f2, L);
                        slot = L.fine;
                        rr = L.r;
                        ri = L.rinv;
                        x0 = L.xi[0];
                        x1 = L.xi[1];
                        x2 = L.xi[2];
                    }
                    else
                    {
                        Loc L;
                        ok = locate(B, ctab, y0 - B.ctr[0], y1 - B.ctr[1], y2 - B.ctr[2], kp.f0,
                                    kp.f1, kp.f2, L);
                        slot = L.fine;
                        rr = L.r;
                        ri = 1.0 / L.r;
                        x0 = R(L.xi[0]);
                        x1 = R(L.xi[1]);
                        x2 = R(L.xi[2]);
                    }
                }
                if (ok)
                {
                    if constexpr (!PH)
                        cf = {R(INV_4PI_D * ri), R(0)};
                    else
                        cf = cscale(expmk_far<R>(B.k_re, B.k_im, rr), R(INV_4PI_D * ri));
                }
                else
                    slot = -1;
            }
            s_slot[wave][lane] = slot;
            s_xi[wave][0][lane] = x0;
            s_xi[wave][1][lane] = x1;
            s_xi[wave][2][lane] = x2;
            s_cf[wave][lane] = cf;
        }
        __syncthreads();
        const int nb = min(4, j1 - jb);
        for (int b = 0; b < nb; b++)
        {
#pragma unroll
            for (int j = 0; j < TPT; j++)
            {
                const int tt = tl + j * TL;
                if (tt >= wi.count)
                    continue;
                const int slot = s_slot[b][tt];
                if (slot < 0)
                    continue;
                const Cx<R> val = fine_eval_s<O0, O1, O2, R, RB>(
                    F, slot, r, s_xi[b][0][tt], s_xi[b][1][tt], s_xi[b][2][tt]);
                acc[j] = cadd(acc[j], cmul(val, s_cf[b][tt]));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < TPT; j++)
    {
        const int tt = tl + j * TL;
        if (tt < wi.count)
        {
            Cx<R>* o = result + (int64_t(wi.first) + tt) * RB + r;
            *o = cadd(*o, acc[j]);
        }
    }
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
template <bool PH, typename R, int O0, int O1, int O2, int RB>
void launch_far_b(DevicePlan& D, hipStream_t st, const KParams& kp, int l, int64_t a, int64_t b,
                  const Cx<R>* F, Cx<R>* res)
{
    // l: the far group; a, b: its items
    const int abase = D.far_groups[l].ang_base, an = D.far_groups[l].ang_n;
    if constexpr (sizeof(R) == 4)
        if (an > 0)
        {
            k_far_b<PH, R, O0, O1, O2, true, RB>
                <<<dim3(unsigned(b - a)), dim3(256), size_t(an) * sizeof(double), st>>>(
                    D.far_items + a, b - a, D.far_srange + 2 * a, D.far_src, D.geo, D.ctab, F,
                    D.tgt_pts, D.fine_ang, abase, an, res, kp);
            return;
        }
    k_far_b<PH, R, O0, O1, O2, false, RB><<<dim3(unsigned(b - a)), dim3(256), 0, st>>>(
        D.far_items + a, b - a, D.far_srange + 2 * a, D.far_src, D.geo, D.ctab, F, D.tgt_pts,
        D.fine_ang, 0, 0, res, kp);
}

typedef float mf4 __attribute__((ext_vector_type(4)));

// one-wave workgroups: the wave's LDS writes complete before its later LDS reads (of other
// lanes' values) issue; the memory clobber keeps the compiler from moving LDS accesses across
__device__ __forceinline__ void lds_fence()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// values -> fine coefficients at orders (3,6,6)/(2,3,3) on FP32 MFMA (v_mfma_f32_16x16x4f32),
// one wave per (cone, 16 real columns): the three axis maps as 16 x 16 x 4 tiles, axis 1, then
// axis 2, then axis 0 (171 MFMAs when every sub-cone is active; the 6 x 3 and 9 x 6 maps fill
// 28% and 42% of a tile). Tile operands: lane l holds A[l % 16][l / 16] (map row, input index)
// and B[l / 16][l % 16] (input index, column); the result tile row 4 (l / 16) + i, column
// l % 16 is element i of lane l. The same map as k_transform_b up to rounding (the axes in
// another order), and slower than it on the 32-column spheroid (8.5 against 6.7 ms per mult):
// the pass moves 19 GB (values read, coefficients of the active sub-cones written) and the
// arithmetic is not what bounds it. Kept for that comparison: it runs only when the
// environment variable IFGF_TRANSFORM_MFMA is set when a batched pass is enqueued or captured.
template <int RB>
__global__ __launch_bounds__(64) void k_transform_mf(int64_t c0, int64_t c1,
                                                     const ConeDev* __restrict__ cones,
                                                     const float* __restrict__ V,
                                                     float* __restrict__ Fout,
                                                     const float* __restrict__ W, int xlg)
{
    constexpr int NC = 2 * RB, NT = NC / 16;
    constexpr int S1 = 100, S2 = 68; // padded row strides: lane groups 16 banks apart
    __shared__ float Y1[27 * S1];    // (t0, j1) x (t2, column)
    __shared__ float Y2[9 * S2];     // per j1: j2 x (t0, column)
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    const int64_t blk = xcd_block(xlg);
    const int64_t cone = c0 + blk / NT;
    if (cone >= c1)
        return;
    const int n = int(blk % NT) * 16 + c;
    const uint32_t mask = cones[cone].mask;
    const int64_t base = cones[cone].fine_base;
    const float* W0 = W;      // 6 x 3
    const float* W1 = W + 18; // 9 x 6
    const float* W2 = W + 72; // 9 x 6
    const float a1a = c < 9 ? W1[c * 6 + g] : 0.f;
    const float a1b = (c < 9 && g < 2) ? W1[c * 6 + g + 4] : 0.f;
    const float a2a = c < 9 ? W2[c * 6 + g] : 0.f;
    const float a2b = (c < 9 && g < 2) ? W2[c * 6 + g + 4] : 0.f;
    const float a0 = (c < 6 && g < 3) ? W0[c * 3 + g] : 0.f;
    const float* v = V + cone * (108 * NC) + n;
    float b[18][2];
#pragma unroll
    for (int q = 0; q < 18; q++) // q = t0 + 3 t2
    {
        const int t0 = q % 3, t2 = q / 3;
        b[q][0] = v[(t0 + 3 * (g + 6 * t2)) * NC];
        b[q][1] = g < 2 ? v[(t0 + 3 * (g + 4 + 6 * t2)) * NC] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < 18; q++)
    {
        const int t0 = q % 3, t2 = q / 3;
        mf4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1a, b[q][0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1b, b[q][1], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const int j1 = 4 * g + i;
            if (j1 < 9)
                Y1[(t0 * 9 + j1) * S1 + t2 * 16 + c] = acc[i];
        }
    }
    lds_fence();
    for (int j1 = 0; j1 < 9; j1++)
    {
        const int s1 = j1 / 3, m1 = j1 - 3 * s1;
        // sub-cones s0 + 3 s1 + 9 s2 for every s0, s2
        if (!((mask >> (3 * s1)) & 0x1C0E07u))
            continue;
#pragma unroll
        for (int t0 = 0; t0 < 3; t0++)
        {
            const float* y = Y1 + (t0 * 9 + j1) * S1 + c;
            mf4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a2a, y[g * 16], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a2b, g < 2 ? y[(g + 4) * 16] : 0.f, acc,
                                                       0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const int j2 = 4 * g + i;
                if (j2 < 9)
                    Y2[j2 * S2 + t0 * 16 + c] = acc[i];
            }
        }
        lds_fence();
        mf4 acc[9];
#pragma unroll
        for (int j2 = 0; j2 < 9; j2++)
        {
            acc[j2] = mf4{0.f, 0.f, 0.f, 0.f};
            acc[j2] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                a0, g < 3 ? Y2[j2 * S2 + g * 16 + c] : 0.f, acc[j2], 0, 0, 0);
        }
#pragma unroll
        for (int j2 = 0; j2 < 9; j2++)
        {
            const int s2 = j2 / 3, m2 = j2 - 3 * s2;
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const int j0 = 4 * g + i;
                const int sub = (j0 >> 1) + 3 * s1 + 9 * s2;
                if (j0 < 6 && ((mask >> sub) & 1u))
                {
                    const int rk = __popc(mask & ((1u << sub) - 1u));
                    const int coef = (j0 & 1) + 2 * (m1 + 3 * m2);
                    Fout[((base + rk) * 18 + coef) * NC + n] = acc[j2][i];
                }
            }
        }
        lds_fence();
    }
}

template <typename R, int H0, int H1, int H2, int L0, int L1, int L2, int NI, int RB>
void launch_transform_b(DevicePlan& D, hipStream_t st, int64_t c0, int64_t c1, const Cx<R>* V,
                        Cx<R>* F)
{
    const R* W = static_cast<const R*>(D.wmat);
    if constexpr (sizeof(R) == 4 && H0 == 3 && H1 == 6 && H2 == 6 && L0 == 2 && L1 == 3 &&
                  L2 == 3)
        if (std::getenv("IFGF_TRANSFORM_MFMA") != nullptr)
        {
            const int64_t blocks = (c1 - c0) * (2 * RB / 16);
            k_transform_mf<RB><<<dim3(unsigned(blocks)), dim3(64), 0, st>>>(
                c0, c1, D.cones, reinterpret_cast<const float*>(V), reinterpret_cast<float*>(F),
                W, D.shape.xcd_log2);
            return;
        }
    const int64_t blocks = (c1 - c0) * RB / NI;
    k_transform_b<R, H0, H1, H2, L0, L1, L2, 3, 3, 3, NI, RB>
        <<<dim3(unsigned(blocks)), dim3(256), 0, st>>>(c0, c1, D.cones, V, F, W,
                                                       D.shape.xcd_log2);
}

// everything of a batched mult between the gather of the densities and the scatter of the
// results: it reads and writes only the batch buffers, so one captured graph serves every call
// with these buffers
template <int KIND, typename R, int RB, int O0, int O1, int O2>
void run_batched_core(DevicePlan& D, hipStream_t st)
{
    constexpr bool PH = KIND != KL;
    constexpr bool P6 = O1 == 3;
    constexpr bool C3 = O0 == 3; // orders (6,8,7)/(3,5,4)
    constexpr int NI = C3 ? 1 : (sizeof(R) == 4 ? (P6 ? 4 : 2) : (P6 ? 2 : 1));
    const KParams kp = kparams(D);
    const NodeTabs tabs{D.ho_nodes, D.trig_th, D.trig_ph};
    Cx<R>* ws = static_cast<Cx<R>*>(D.b_ws);
    Cx<R>* res = static_cast<Cx<R>*>(D.b_res);
    Cx<R>* V = static_cast<Cx<R>*>(D.b_V);
    Cx<R>* F = static_cast<Cx<R>*>(D.b_F);
    const int64_t nt = D.ntgt * RB;
    k_zero_b<R><<<dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st>>>(res, nt);
    // near and far fields as in the single-column mult (SidePipe)
    const int overlap = D.n_near_items ? D.shape.overlap : 0;
    auto pipe = side_pipe(
        D, st, overlap,
        [&](hipStream_t s)
        {
            if (D.n_near_items)
                k_near_b<KIND, R, RB><<<dim3(unsigned(D.n_near_items)), dim3(64), 0, s>>>(
                    D.near_items, D.n_near_items, D.near_off, D.near_src, D.geo, D.src_begin,
                    D.src_end, D.src_pts, D.normals, ws, D.tgt_pts, res, kp);
        },
        [&](int g, hipStream_t s)
        {
            const DevicePlan::FarGroup& G = D.far_groups[g];
            launch_far_b<PH, R, O0, O1, O2, RB>(D, s, kp, g, G.item0, G.item1, F, res);
        });
    if (overlap != 1)
        pipe.fork();
    if (D.n_leaf_cones)
        k_leaf_b<KIND, R, RB><<<dim3(unsigned(D.n_leaf_cones)), dim3(128), 0, st>>>(
            D.leaf_cones, D.n_leaf_cones, D.cones, D.geo, D.src_begin, D.src_end, D.src_pts,
            D.normals, ws, tabs, V, kp);
    for (int l = D.levels - 1; l >= 0; l--)
    {
        const int64_t u0 = D.up_off[l], u1 = D.up_off[l + 1];
        const int64_t c0 = D.cone_lvl_off[l], c1 = D.cone_lvl_off[l + 1];
        const int64_t i0 = D.up_item_off[l], i1 = D.up_item_off[l + 1];
        if constexpr (P6 && sizeof(R) == 4)
        {
            if (i1 > i0)
            {
#define IFGF_UP_STAGE_B(K)                                                                      \
    k_up_stage_b<R, O0, O1, O2, 108, K, RB><<<dim3(unsigned(i1 - i0)), dim3(256), 0, st>>>(    \
        D.up_items + i0, i1 - i0, D.up_glist + D.up_gl_off[l], D.cones, D.child, D.geo, D.ctab, \
        static_cast<const UpEnt<R>*>(D.up_entries), D.up_part_ent, F, V, kp.xlg)
            switch (D.shape.up_k)
            {
            case 2: // items of at most 2 cones run as items of 4
            case 4:
                IFGF_UP_STAGE_B(4);
                break;
            case 16:
                IFGF_UP_STAGE_B(16);
                break;
            default:
                IFGF_UP_STAGE_B(8);
            }
#undef IFGF_UP_STAGE_B
            }
        }
        else if (i1 > i0)
        {
#define IFGF_UP_CELL_B(K)                                                                       \
    k_up_cell_b<R, O0, O1, O2, K, RB><<<dim3(unsigned(i1 - i0)), dim3(256), 0, st>>>(           \
        D.up_items + i0, i1 - i0, D.up_glist + D.up_gl_off[l], D.cones, D.geo, D.child, D.ctab, \
        static_cast<const UpEnt<R>*>(D.up_entries), D.up_part_ent, F, V, kp)
            switch (D.shape.up_k)
            {
            case 2: // items of at most 2 cones run as items of 4
            case 4:
                IFGF_UP_CELL_B(4);
                break;
            case 16:
                IFGF_UP_CELL_B(16);
                break;
            default:
                IFGF_UP_CELL_B(8);
            }
#undef IFGF_UP_CELL_B
        }
        if (u1 > u0)
            k_up_b<PH, R, O0, O1, O2, RB><<<dim3(unsigned(u1 - u0)), dim3(256), 0, st>>>(
                D.up_cones + u0, u1 - u0, D.cones, D.geo, D.child, D.ctab, tabs, F, V, kp);
        if (c1 > c0)
        {
            if constexpr (P6)
                launch_transform_b<R, 3, 6, 6, 2, 3, 3, NI, RB>(D, st, c0, c1, V, F);
            else if constexpr (C3)
                launch_transform_b<R, 6, 8, 7, 3, 5, 4, NI, RB>(D, st, c0, c1, V, F);
            else
                launch_transform_b<R, 5, 8, 8, 2, 5, 5, NI, RB>(D, st, c0, c1, V, F);
            pipe.fork();
        }
        pipe.ready(l);
    }
    pipe.finish();
}

// one batched mult: gather, the core (eagerly, or as the plan's captured batched graph), scatter
template <int KIND, typename R, int RB, int O0, int O1, int O2>
void run_batched(DevicePlan& D, const double* w_dev, double* out_dev, int ncol, hipStream_t st,
                 int io, bool graph)
{
    Cx<R>* ws = static_cast<Cx<R>*>(D.b_ws);
    Cx<R>* res = static_cast<Cx<R>*>(D.b_res);
    const double in_im = (io & IO_CONJ_IN) ? -1.0 : 1.0;
    const double out_re = (io & IO_NEGATE_OUT) ? -1.0 : 1.0;
    const double out_im = out_re * ((io & IO_CONJ_OUT) ? -1.0 : 1.0);
    const int64_t ns = D.nsrc * RB, nt = D.ntgt * RB;
    k_gather_b<R><<<dim3(unsigned((ns + 255) / 256)), dim3(256), 0, st>>>(
        w_dev, D.src_perm, ws, D.nsrc, RB, ncol, in_im);
    if (graph)
    {
        if (!D.b_graph_exec)
        {
            // captured on the plan's capture stream (the caller's may be the null stream,
            // which cannot be captured); kernels and event waits only
            if (!D.cap_stream)
            {
                hipStream_t cs;
                HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
                D.cap_stream = cs;
            }
            hipStream_t cs = static_cast<hipStream_t>(D.cap_stream);
            HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            run_batched_core<KIND, R, RB, O0, O1, O2>(D, cs);
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamEndCapture(cs, &g));
            hipGraphExec_t ex = nullptr;
            HIP_OK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
            D.b_graph = g;
            D.b_graph_exec = ex;
        }
        HIP_OK(hipGraphLaunch(static_cast<hipGraphExec_t>(D.b_graph_exec), st));
    }
    else
        run_batched_core<KIND, R, RB, O0, O1, O2>(D, st);
    k_scatter_b<R><<<dim3(unsigned((nt + 255) / 256)), dim3(256), 0, st>>>(
        res, D.tgt_perm, out_dev, D.ntgt, RB, ncol, out_re, out_im);
    HIP_OK(hipGetLastError());
}

template <typename R, int RB, int O0, int O1, int O2>
void run_kind(DevicePlan& D, const double* w, double* o, int ncol, hipStream_t st, int io,
              bool graph)
{
    switch (D.kind)
    {
    case KIND_HELMHOLTZ:
        return run_batched<KH, R, RB, O0, O1, O2>(D, w, o, ncol, st, io, graph);
    case KIND_MODHELMHOLTZ:
        return run_batched<KM, R, RB, O0, O1, O2>(D, w, o, ncol, st, io, graph);
    case KIND_LAPLACE:
        return run_batched<KL, R, RB, O0, O1, O2>(D, w, o, ncol, st, io, graph);
    case KIND_DOUBLELAYER:
        return run_batched<KD, R, RB, O0, O1, O2>(D, w, o, ncol, st, io, graph);
    case KIND_COMBINED:
        return run_batched<KC, R, RB, O0, O1, O2>(D, w, o, ncol, st, io, graph);
    case KIND_GRAD:
        return run_batched<KG, R, RB, O0, O1, O2>(D, w, o, ncol, st, io, graph);
    }
    throw std::runtime_error("unknown kernel kind");
}

template <typename R, int RB>
void run_order(DevicePlan& D, const double* w, double* o, int ncol, hipStream_t st, int io,
               bool graph)
{
    switch (order_class(D))
    {
    case 1:
        return run_kind<R, RB, 2, 3, 3>(D, w, o, ncol, st, io, graph);
    case 2:
        return run_kind<R, RB, 2, 5, 5>(D, w, o, ncol, st, io, graph);
    default:
        return run_kind<R, RB, 3, 5, 4>(D, w, o, ncol, st, io, graph);
    }
}
} // namespace

bool batched_supported(const DevicePlan& D)
{
    const int c = order_class(D);
    return c == 1 || c == 2 || c == 3;
}

void mult_batched(DevicePlan& D, const double* w_dev, double* out_dev, int rb, int ncol,
                  void* stream, int io, bool graph)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!batched_supported(D) || ncol < 1 || ncol > rb || rb != D.b_rb || (rb != 8 && rb != 32))
        throw std::runtime_error("mult_batched: unsupported orders, column count or buffers");
    if (D.prec == PREC_FP64)
    {
        if (rb == 8)
            run_order<double, 8>(D, w_dev, out_dev, ncol, st, io, graph);
        else
            run_order<double, 32>(D, w_dev, out_dev, ncol, st, io, graph);
    }
    else
    {
        if (rb == 8)
            run_order<float, 8>(D, w_dev, out_dev, ncol, st, io, graph);
        else
            run_order<float, 32>(D, w_dev, out_dev, ncol, st, io, graph);
    }
}
} // namespace ifgf
