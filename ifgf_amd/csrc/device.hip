// device.hip -- HIP kernels of IFGF operator application on gfx950 and their orchestration.
//
// One mult (ifgfoperator.hpp:246-867) runs as a level-synchronous pipeline on one stream:
//   gather weights -> near field (P2P) -> leaf sampling (all leaf cones) ->
//   for each source level, deepest first: upward transfer (non-leaf cones) and the fused
//   high-order-values -> fine-coefficients transform -> far-field evaluation per target
//   level -> scatter.
// Every kernel is templated on the arithmetic type R (double: the plain form; float: the
// mixed form, which keeps cone coordinates, phases and box geometry in FP64).  The kernels
// that evaluate or build fine interpolants are specialised on the interpolation orders of
// the common configurations (orders 6 and 8), with a runtime-order fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <type_traits>
#include <utility>

#include "dev_common.hpp"
#include "device.hpp"

namespace ifgf
{
namespace
{
using namespace dev;

// ------------------------------------------------------------------------------------
// simple kernels
// ------------------------------------------------------------------------------------
template <typename R>
__global__ void k_gather(const double* __restrict__ w_in, const int64_t* __restrict__ perm,
                         Cx<R>* __restrict__ w_out, int64_t n, double im_sign)
{
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t j = perm[i];
    w_out[i] = {R(w_in[2 * j]), R(im_sign * w_in[2 * j + 1])};
}

template <typename R>
__global__ void k_scatter(const Cx<R>* __restrict__ res, const int64_t* __restrict__ perm,
                          double* __restrict__ out, int64_t n, double re_sign, double im_sign)
{
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t j = perm[i];
    out[2 * j] = re_sign * double(res[i].x);
    out[2 * j + 1] = im_sign * double(res[i].y);
}

template <typename R> __global__ void k_zero(Cx<R>* p, int64_t n)
{
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = {R(0), R(0)};
}


// near field: for each target of the item, sum over the points of its near source leaves
// (ifgfoperator.hpp:900-960; octree.hpp:302-307).  The wave walks the source leaves in order
// and stages each (64 points at a time) in LDS with one coalesced load; lane = (target, group),
// the groups splitting the staged points; xor-shuffle reduction over the groups at the end.
template <int KIND, typename R>
__global__ __launch_bounds__(64) void k_near(const WorkItem* __restrict__ items, int64_t nitems,
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
    __shared__ Cx<R> sw[64];
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
    Cx<R> acc = {0, 0};
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
                sw[lane] = w[i];
            }
            __syncthreads();
            for (int k = grp; k < n; k += ngrp)
            {
                const double d0 = sp[3 * k] - y0, d1 = sp[3 * k + 1] - y1, d2 = sp[3 * k + 2] - y2;
                double xn = 0;
                if constexpr (NRM)
                    xn = d0 * sn[3 * k] + d1 * sn[3 * k + 1] + d2 * sn[3 * k + 2];
                acc = cadd(acc, p2p<KIND, R>(kb, R(d0), R(d1), R(d2), xn, sw[k]));
            }
        }
    }
    for (int m = Tp; m < 64; m <<= 1)
    {
        acc.x += __shfl_xor(acc.x, m, 64);
        acc.y += __shfl_xor(acc.y, m, 64);
    }
    if (grp == 0 && valid)
        result[t] = cadd(result[t], acc);
}

// the values -> fine coefficients transform of one cone by NT threads (the arithmetic of
// k_transform_t, pencil by pencil): X holds the cone's n_hi values on entry and is a work
// buffer of max(n_hi, M0 M1 H2) entries, Y follows it (M0 H1 H2 entries); the coefficients of
// the active sub-cones go to Fout from fine slot `base` on
template <int H0, int H1, int H2, int L0, int L1, int L2, int NT>
__device__ __forceinline__ void transform_one(Cx<float>* X, uint32_t mask, int base,
                                              const float* __restrict__ W,
                                              Cx<float>* __restrict__ Fout, int tid)
{
    constexpr int F0 = 3, F1 = 3, F2 = 3;
    constexpr int M0 = F0 * L0, M1 = F1 * L1, M2 = F2 * L2;
    constexpr int NH = H0 * H1 * H2, NLO = L0 * L1 * L2;
    constexpr int XS = (NH > M0 * M1 * H2) ? NH : M0 * M1 * H2;
    constexpr int NW0 = M0 * H0, NW1 = M1 * H1, NW2 = M2 * H2;
    constexpr bool RW = sizeof(float) * (NW0 + NW1 + NW2) > 512;
    Cx<float>* Y = X + XS;
    const float* W0 = W;
    const float* W1 = W + NW0;
    const float* W2 = W + NW0 + NW1;
    for (int p = tid; p < H1 * H2; p += NT)
    {
        Cx<float> in[H0];
#pragma unroll
        for (int t = 0; t < H0; t++)
            in[t] = X[p * H0 + t];
#pragma unroll
        for (int j = 0; j < M0; j++)
        {
            Cx<float> a = {0, 0};
#pragma unroll
            for (int t = 0; t < H0; t++)
            {
                a.x += W0[j * H0 + t] * in[t].x;
                a.y += W0[j * H0 + t] * in[t].y;
            }
            Y[p * M0 + j] = a;
            if constexpr (RW)
                asm volatile("" ::: "memory");
        }
    }
    __syncthreads();
    for (int p = tid; p < M0 * H2; p += NT)
    {
        const int j0 = p % M0, i2 = p / M0;
        Cx<float> in[H1];
#pragma unroll
        for (int t = 0; t < H1; t++)
            in[t] = Y[j0 + M0 * (t + H1 * i2)];
#pragma unroll
        for (int j = 0; j < M1; j++)
        {
            Cx<float> a = {0, 0};
#pragma unroll
            for (int t = 0; t < H1; t++)
            {
                a.x += W1[j * H1 + t] * in[t].x;
                a.y += W1[j * H1 + t] * in[t].y;
            }
            X[j0 + M0 * (j + M1 * i2)] = a;
            if constexpr (RW)
                asm volatile("" ::: "memory");
        }
    }
    __syncthreads();
    for (int p = tid; p < M0 * M1; p += NT)
    {
        const int j0 = p % M0, j1 = p / M0;
        const int s0 = j0 / L0, s1 = j1 / L1;
        const int m01 = (j0 - s0 * L0) + L0 * (j1 - s1 * L1);
        Cx<float> in[H2];
#pragma unroll
        for (int t = 0; t < H2; t++)
            in[t] = X[p + M0 * M1 * t];
#pragma unroll
        for (int j = 0; j < M2; j++)
        {
            Cx<float> a = {0, 0};
#pragma unroll
            for (int t = 0; t < H2; t++)
            {
                a.x += W2[j * H2 + t] * in[t].x;
                a.y += W2[j * H2 + t] * in[t].y;
            }
            const int s2 = j / L2, m2 = j - s2 * L2;
            const int sub = s0 + F0 * (s1 + F1 * s2);
            if ((mask >> sub) & 1u)
            {
                const int rk = __popc(mask & ((1u << sub) - 1u));
                Fout[int64_t(base + rk) * NLO + m01 + L0 * L1 * m2] = a;
            }
            if constexpr (RW)
                asm volatile("" ::: "memory");
        }
    }
}

// leaf sampling (ifgfoperator.hpp:1057-1073, :423-462): the factored kernel of a leaf's
// sources at the high-order nodes of its active coarse cones.  One 256-thread workgroup per
// leaf box, thread = (active cone, high-order node) pairs of the box, 256 at a time, the box's
// sources staged in LDS for each round: lanes stay busy however many cones the box has, and
// the sources are staged once per 256 pairs
// H0 > 0 (FP32, the compiled orders H / L): the box's cones are then transformed to their fine
// coefficients by the same workgroup (transform_one), their values read back from where this
// workgroup has just written them (L2), instead of by the separate transform kernel
template <int KIND, typename R, int H0 = 0, int H1 = 0, int H2 = 0, int L0 = 0, int L1 = 0,
          int L2 = 0>
__global__ __launch_bounds__(256) void k_leaf_box(const WorkItem* __restrict__ items,
                                                  int64_t nitems,
                                                  const ConeDev* __restrict__ cones,
                                                  const BoxGeo* __restrict__ geo,
                                                  const int64_t* __restrict__ sbeg,
                                                  const int64_t* __restrict__ send,
                                                  const double* __restrict__ spts,
                                                  const double* __restrict__ snrm,
                                                  const Cx<R>* __restrict__ w, NodeTabs tabs,
                                                  Cx<R>* __restrict__ V, KParams kp,
                                                  const float* __restrict__ W = nullptr,
                                                  Cx<float>* __restrict__ Fout = nullptr)
{
    constexpr int CH = 256;
    __shared__ R sx[CH], sy[CH], sz[CH];
    __shared__ R nx[CH], ny[CH], nz[CH];
    __shared__ Cx<R> sw[CH];
    const int64_t it = xcd_block(kp.xlg);
    if (it >= nitems)
        return;
    const WorkItem wi = items[it]; // box, its first cone, its number of cones
    const BoxGeo B = geo[wi.box];
    const KParams kb = with_k(kp, B);
    const int64_t b0 = sbeg[wi.box], b1 = send[wi.box];
    const int npair = wi.count * kp.n_hi;
    for (int p0 = 0; p0 < npair; p0 += CH)
    {
        const int e = p0 + int(threadIdx.x);
        const bool active = e < npair;
        const int k = active ? e / kp.n_hi : 0, q = active ? e - k * kp.n_hi : 0;
        const int32_t g = wi.first + k;
        double yd[3] = {0, 0, 0}, r = 1;
        if (active)
            cone_node(B, cones[g].cell, q, kp.ho0, kp.ho1, kp.ho2, tabs, false, yd, r);
        const R y0 = R(yd[0]), y1 = R(yd[1]), y2 = R(yd[2]), dc = R(r);
        Cx<R> acc = {0, 0};
        for (int64_t c0 = b0; c0 < b1; c0 += CH)
        {
            const int64_t i = c0 + threadIdx.x;
            __syncthreads();
            if (i < b1)
            {
                sx[threadIdx.x] = R(spts[3 * i] - B.ctr[0]);
                sy[threadIdx.x] = R(spts[3 * i + 1] - B.ctr[1]);
                sz[threadIdx.x] = R(spts[3 * i + 2] - B.ctr[2]);
                sw[threadIdx.x] = w[i];
                if (snrm)
                {
                    nx[threadIdx.x] = R(snrm[3 * i]);
                    ny[threadIdx.x] = R(snrm[3 * i + 1]);
                    nz[threadIdx.x] = R(snrm[3 * i + 2]);
                }
            }
            __syncthreads();
            const int n = int(min<int64_t>(CH, b1 - c0));
            if (active)
                for (int m = 0; m < n; m++)
                {
                    R nn[3] = {0, 0, 0};
                    if (snrm)
                    {
                        nn[0] = nx[m];
                        nn[1] = ny[m];
                        nn[2] = nz[m];
                    }
                    acc = cadd(acc, factored<KIND, R>(kb, y0, y1, y2, dc, sx[m], sy[m], sz[m], nn,
                                                      sw[m]));
                }
        }
        if (active)
            V[int64_t(g) * kp.n_hi + q] = acc;
    }
    if constexpr (H0 > 0)
    {
        static_assert(sizeof(R) == 4, "the fused transform is FP32");
        constexpr int NH = H0 * H1 * H2;
        constexpr int XS = (NH > 9 * L0 * L1 * H2) ? NH : 9 * L0 * L1 * H2;
        __shared__ Cx<float> T[XS + 3 * L0 * H1 * H2];
        __syncthreads(); // the box's values written by this workgroup are visible to it
        for (int k = 0; k < wi.count; k++)
        {
            const int32_t g = wi.first + k;
            for (int i = int(threadIdx.x); i < NH; i += 256)
                T[i] = V[int64_t(g) * NH + i];
            __syncthreads();
            transform_one<H0, H1, H2, L0, L1, L2, 256>(T, cones[g].mask, cones[g].fine_base, W,
                                                       Fout, int(threadIdx.x));
            __syncthreads();
        }
    }
}

// upward pass: the children's fine interpolants at the parent's high-order nodes times the
// transfer factor, summed over the children in order (ifgfoperator.hpp:631-688, 1639-1705)
template <bool PH, typename R, int O0, int O1, int O2>
__global__ __launch_bounds__(128) void k_up(const int32_t* __restrict__ work, int64_t nwork,
                                            const ConeDev* __restrict__ cones,
                                            const BoxGeo* __restrict__ geo,
                                            const int32_t* __restrict__ child,
                                            const CellRef* __restrict__ ctab, NodeTabs tabs,
                                            const Cx<R>* __restrict__ F, Cx<R>* __restrict__ V,
                                            KParams kp)
{
    const int64_t wi = xcd_block(kp.xlg);
    if (wi >= nwork)
        return;
    const int32_t g = work[wi];
    const ConeDev cn = cones[g];
    const BoxGeo& P = geo[cn.box];
    for (int q = threadIdx.x; q < kp.n_hi; q += blockDim.x)
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
            const Cx<R> val =
                fine_eval<O0, O1, O2, R>(F, L.fine, kp, R(L.xi[0]), R(L.xi[1]), R(L.xi[2]));
            // e^{-k (d - dp)} dp / d with d = |y - x_child|, dp = |y - x_parent|
            Cx<R> tf;
            if constexpr (!PH)
                tf = {R(rp / L.r), R(0)};
            else
                tf = cscale(expmk<R>(P.k_re, P.k_im, R(L.r - rp)), R(rp / L.r));
            acc = cadd(acc, cmul(val, tf));
        }
        V[int64_t(g) * kp.n_hi + q] = acc;
    }
}


// upward pass with the precomputed geometry (per parent level of a geometry, see UpLevel in
// plan.hpp): per (node, child) one table entry gives the child's fine cell, local coordinates
// and transfer factor.  A workgroup takes up to K parent cones of one cell (item.box = the
// cell's index in its table, item.pad = the table part, glist[item.first ..] the cones), so
// each entry is read once for K boxes; children are summed in order for every cone
// (ifgfoperator.hpp:631-688, 1639-1705).
template <typename R, int O0, int O1, int O2, int K>
__global__ __launch_bounds__(128) void k_up_cell(const WorkItem* __restrict__ items,
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
    __shared__ int64_t toff[K][8]; // table offset of child c of cone k; -1: no child
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
    for (int q = tid; q < kp.n_hi; q += blockDim.x)
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
                Cx<R> val;
                if constexpr (O0 > 0)
                    val = fine_eval_chunked<O0, O1, O2, R>(F, slot, E.xi0, E.xi1, E.xi2);
                else
                    val = fine_eval<O0, O1, O2, R>(F, slot, kp, E.xi0, E.xi1, E.xi2);
                acc[k] = cadd(acc[k], cmul(val, Cx<R>{E.tr, E.ti}));
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            if (k < n)
                V[int64_t(gs[k]) * kp.n_hi + q] = acc[k];
    }
}

// Chebyshev values T_1 .. T_{n-1} at x (T_0 = 1 is implied)
template <int N> __device__ __forceinline__ void cheb_t(float x, float (&t)[N])
{
    if constexpr (N > 1)
        t[1] = x;
#pragma unroll
    for (int k = 2; k < N; k++)
        t[k] = fmaf(2.0f * x, t[k - 1], -(k == 2 ? 1.0f : t[k - 2]));
}

// The interpolants of one fine cone (coefficients c, index i0 + O0 (i1 + O1 i2)) at G points
// from their Chebyshev values: sum_{i0,i1,i2} c T_i0(x0) T_i1(x1) T_i2(x2), summed over i0,
// then i1, then i2 (each coefficient read once for the G points)
template <int O0, int O1, int O2, int G>
__device__ __forceinline__ void eval_cone_g(const float2* __restrict__ c,
                                            const float (&t0)[G][O0], const float (&t1)[G][O1],
                                            const float (&t2)[G][O2], float2 (&out)[G])
{
#pragma unroll
    for (int i2 = 0; i2 < O2; i2++)
    {
        float2 pl[G];
#pragma unroll
        for (int i1 = 0; i1 < O1; i1++)
        {
            float2 row[O0];
#pragma unroll
            for (int i0 = 0; i0 < O0; i0++)
                row[i0] = c[i0 + O0 * (i1 + O1 * i2)];
#pragma unroll
            for (int g = 0; g < G; g++)
            {
                float2 sm = row[0];
#pragma unroll
                for (int i0 = 1; i0 < O0; i0++)
                {
                    sm.x = fmaf(row[i0].x, t0[g][i0], sm.x);
                    sm.y = fmaf(row[i0].y, t0[g][i0], sm.y);
                }
                if (i1 == 0)
                    pl[g] = sm;
                else
                {
                    pl[g].x = fmaf(sm.x, t1[g][i1], pl[g].x);
                    pl[g].y = fmaf(sm.y, t1[g][i1], pl[g].y);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++)
        {
            if (i2 == 0)
                out[g] = pl[g];
            else
            {
                out[g].x = fmaf(pl[g].x, t2[g][i2], out[g].x);
                out[g].y = fmaf(pl[g].y, t2[g][i2], out[g].y);
            }
        }
    }
}

// eval_cone_g with the coefficients in memory, read PB axis-2 planes at a time (the loads of a
// batch are not hoisted above the previous batch's arithmetic: fewer live registers)
template <int O0, int O1, int O2, int G, int PB = (O2 % 2 == 0) ? 2 : 1>
__device__ __forceinline__ void eval_cone_g_mem(const float2* c,
                                                const float (&t0)[G][O0], const float (&t1)[G][O1],
                                                const float (&t2)[G][O2], float2 (&out)[G])
{
    static_assert(O2 % PB == 0, "whole batches of planes");
    constexpr int NP = O0 * O1;
    float2 cp2[PB * NP];
#pragma unroll
    for (int i2 = 0; i2 < O2; i2++)
    {
        if (i2 % PB == 0)
        {
#pragma unroll
            for (int i = 0; i < PB * NP; i++)
                cp2[i] = c[NP * i2 + i];
        }
        const float2* cp = cp2 + NP * (i2 % PB);
        float2 pl[G];
#pragma unroll
        for (int i1 = 0; i1 < O1; i1++)
        {
#pragma unroll
            for (int g = 0; g < G; g++)
            {
                float2 sm = cp[O0 * i1];
#pragma unroll
                for (int i0 = 1; i0 < O0; i0++)
                {
                    sm.x = fmaf(cp[O0 * i1 + i0].x, t0[g][i0], sm.x);
                    sm.y = fmaf(cp[O0 * i1 + i0].y, t0[g][i0], sm.y);
                }
                if (i1 == 0)
                    pl[g] = sm;
                else
                {
                    pl[g].x = fmaf(sm.x, t1[g][i1], pl[g].x);
                    pl[g].y = fmaf(sm.y, t1[g][i1], pl[g].y);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++)
        {
            if (i2 == 0)
                out[g] = pl[g];
            else
            {
                out[g].x = fmaf(pl[g].x, t2[g][i2], out[g].x);
                out[g].y = fmaf(pl[g].y, t2[g][i2], out[g].y);
            }
        }
        // the next batch's address depends on this batch's sums: its loads follow them
        if (i2 % PB == PB - 1)
            asm volatile("" : "+v"(c) : "v"(out[0].x));
    }
}

// Upward pass with the grouped tables (FP32, compiled orders; ifgfoperator.hpp:631-688,
// 1639-1705).  A workgroup takes the parent cones of an item one at a time, each from its
// UpDesc record (read into LDS at once); all the cone's
// (child, group) tasks run at once: a group is up to UP_GRP nodes falling in one child cone,
// whose coefficients the thread reads once (from memory) for all of them, and writes each
// node's value times its transfer factor to the node's partial sum of that child in LDS.  The
// partial sums of every node are then added in child order, as the reference's.
template <int O0, int O1, int O2, int NHI, int NT, int H0, int H1, int H2, bool TF>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_up_grp(
    const WorkItem* __restrict__ items, int64_t nitems, const UpDesc* __restrict__ desc,
    const CellRef* __restrict__ ctab, const UpGrp* __restrict__ grp,
    const Cx<float>* __restrict__ F, Cx<float>* __restrict__ V, const float* __restrict__ W,
    Cx<float>* __restrict__ Fo, KParams kp)
{
    constexpr int NC = O0 * O1 * O2;
    static_assert(UP_GRP == 4, "records of four nodes");
    static_assert(NHI == H0 * H1 * H2, "the high orders");
    static_assert((H0 * H1 * H2 > 9 * O0 * O1 * H2 ? H0 * H1 * H2 : 9 * O0 * O1 * H2) +
                          3 * O0 * H1 * H2 <=
                      8 * NHI + 1,
                  "the transform's buffers fit in part");
    constexpr int G = UP_GRP;
    constexpr int NW = int(sizeof(UpDesc) / 16);
    // the cone's record: child tables, task ranges, the slots of the listed child cones
    __shared__ UpDesc sd;
    __shared__ float2 part[8 * NHI + 1]; // node q's contribution from child c (+ a dump slot)
    const int64_t it = xcd_block(kp.xlg);
    if (it >= nitems)
        return;
    const WorkItem wi = items[it];
    const int tid = threadIdx.x;
#pragma unroll 1
    for (int k = 0; k < wi.count; k++)
    {
        const int4* src = reinterpret_cast<const int4*>(desc + (wi.first + k));
        for (int i = tid; i < NW; i += NT)
            reinterpret_cast<int4*>(&sd)[i] = src[i];
        for (int i = tid; i < 8 * NHI; i += NT)
            part[i] = make_float2(0.0f, 0.0f);
        __syncthreads();
        const int tot = sd.tend[7];
        for (int t = tid; t < tot; t += NT)
        {
            int c = 0;
            while (t >= sd.tend[c])
                c++;
            // the record as six 16-byte loads: header, x0, x1, x2, tr, ti
            const float4* ep = reinterpret_cast<const float4*>(grp + (sd.tg0[c] + t));
            const float4 hd = ep[0], ex0 = ep[1], ex1 = ep[2], ex2 = ep[3];
            const int32_t cf = __float_as_int(hd.x), subd = __float_as_int(hd.y);
            const int d = subd >> 8;
            int slot;
            if (d < UP_DMAX)
                slot = sd.slot[c][d];
            else
            {
                const int sub = subd & 255;
                const CellRef cr = ctab[sd.toff[c] + cf];
                slot = ((cr.mask >> sub) & 1u) ? cr.fine_base + __popc(cr.mask & ((1u << sub) - 1u))
                                               : -1;
            }
            if (slot < 0)
                continue;
            const float xs0[G] = {ex0.x, ex0.y, ex0.z, ex0.w};
            const float xs1[G] = {ex1.x, ex1.y, ex1.z, ex1.w};
            const float xs2[G] = {ex2.x, ex2.y, ex2.z, ex2.w};
            float t0[G][O0], t1[G][O1], t2[G][O2];
#pragma unroll
            for (int gg = 0; gg < G; gg++)
            {
                cheb_t<O0>(xs0[gg], t0[gg]);
                cheb_t<O1>(xs1[gg], t1[gg]);
                cheb_t<O2>(xs2[gg], t2[gg]);
            }
            float2 val[G];
            eval_cone_g_mem<O0, O1, O2, G>(
                reinterpret_cast<const float2*>(F + int64_t(slot) * NC), t0, t1, t2, val);
            const float4 etr = ep[4], eti = ep[5];
            const float trs[G] = {etr.x, etr.y, etr.z, etr.w};
            const float tis[G] = {eti.x, eti.y, eti.z, eti.w};
            const uint32_t qq[2] = {uint32_t(__float_as_int(hd.z)), uint32_t(__float_as_int(hd.w))};
#pragma unroll
            for (int gg = 0; gg < G; gg++)
            {
                // an empty place writes to the dump slot (no branch: the evaluation stays one
                // straight sequence for the G points)
                const int q = int((qq[gg >> 1] >> (16 * (gg & 1))) & 0xffffu);
                const int at = q == 0xffff ? 8 * NHI : c * NHI + q;
                const float tr = trs[gg], ti = tis[gg];
                part[at] = make_float2(val[gg].x * tr - val[gg].y * ti,
                                       val[gg].x * ti + val[gg].y * tr);
            }
        }
        __syncthreads();
        if constexpr (!TF)
        {
            for (int q = tid; q < NHI; q += NT)
            {
                float2 a = make_float2(0.0f, 0.0f);
#pragma unroll
                for (int c = 0; c < 8; c++)
                {
                    const float2 p = part[c * NHI + q];
                    a = make_float2(a.x + p.x, a.y + p.y);
                }
                V[int64_t(sd.g) * NHI + q] = Cx<float>{a.x, a.y};
            }
            __syncthreads(); // before the next cone replaces sd and part
            continue;
        }
        // TF: the node values (children in order), then straight on to the cone's fine
        // coefficients: the values -> coefficients transform of k_transform_t for this one
        // cone, its buffers in part's place (the values never go to memory)
        constexpr int NQT = (NHI + NT - 1) / NT;
        float2 sum[NQT];
#pragma unroll
        for (int i = 0; i < NQT; i++)
        {
            const int q = tid + i * NT;
            float2 a = make_float2(0.0f, 0.0f);
            if (q < NHI)
#pragma unroll
                for (int c = 0; c < 8; c++)
                {
                    const float2 p = part[c * NHI + q];
                    a = make_float2(a.x + p.x, a.y + p.y);
                }
            sum[i] = a;
        }
        __syncthreads();
        Cx<float>* X = reinterpret_cast<Cx<float>*>(part);
#pragma unroll
        for (int i = 0; i < NQT; i++)
        {
            const int q = tid + i * NT;
            if (q < NHI)
                X[q] = Cx<float>{sum[i].x, sum[i].y};
        }
        __syncthreads();
        transform_one<H0, H1, H2, O0, O1, O2, NT>(X, sd.mask, sd.fine_base, W, Fo, tid);
        __syncthreads(); // before the next cone replaces sd and part
    }
}

// fused high-order values -> fine coefficients of the active sub-cones, separable per axis:
// D = (W2 x W1 x W0) V with W_d the (f o_lo) x o_hi axis maps (chebinterp.hpp:176-250,
// ifgfoperator.hpp:1539-1637 and the fine transform :548-570).  One thread per pencil of an
// axis; the active fine cones are written contiguously at the end.
template <typename R, int H0, int H1, int H2, int L0, int L1, int L2, int F0, int F1, int F2,
          int CPB>
__global__ __launch_bounds__(256) void k_transform_t(int64_t c0, int64_t c1,
                                                     const int32_t* __restrict__ list,
                                                     const ConeDev* __restrict__ cones,
                                                     const Cx<R>* __restrict__ V,
                                                     Cx<R>* __restrict__ Fout,
                                                     const R* __restrict__ W, int xlg)
{
    constexpr int M0 = F0 * L0, M1 = F1 * L1, M2 = F2 * L2;
    constexpr int NH = H0 * H1 * H2, NLO = L0 * L1 * L2, NSUB = F0 * F1 * F2;
    constexpr int XS = (NH > M0 * M1 * H2) ? NH : M0 * M1 * H2;
    // the axis-2 stage writes the active sub-cones' coefficients to memory directly: Y holds
    // the axis-0 stage only
    constexpr int YS = M0 * H1 * H2;
    constexpr int NW0 = M0 * H0, NW1 = M1 * H1, NW2 = M2 * H2;
    __shared__ Cx<R> X[CPB * XS];
    __shared__ Cx<R> Y[CPB * YS];
    __shared__ uint32_t cmask[CPB];
    __shared__ int cbase[CPB];
    __shared__ int64_t cid[CPB];
    const int tid = threadIdx.x;
    // cones [c0, c1), or with a list the cones list[c0 .. c1)
    const int64_t first = c0 + xcd_block(xlg) * CPB;
    const int ncone = int(min<int64_t>(CPB, c1 - first));
    if (tid < ncone)
    {
        const int64_t c = list ? int64_t(list[first + tid]) : first + tid;
        cid[tid] = c;
        cmask[tid] = cones[c].mask;
        cbase[tid] = cones[c].fine_base;
    }
    __syncthreads();
    for (int e = tid; e < ncone * NH; e += 256)
    {
        const int k = e / NH, i = e - k * NH;
        X[k * XS + i] = V[cid[k] * NH + i];
    }
    __syncthreads();
    // the axis maps at compile-time offsets from the kernel argument: scalar loads, the
    // weights stay in scalar registers as FMA operands.  Maps too large for the scalar
    // registers are read a row at a time (RW: a compiler barrier after each output keeps the
    // next row's loads from being hoisted, which would spill scalar registers to vector lanes)
    constexpr bool RW = sizeof(R) * (NW0 + NW1 + NW2) > 512;
    const R* W0 = W;
    const R* W1 = W + NW0;
    const R* W2 = W + NW0 + NW1;
    // axis 0: pencils (i1, i2): Y[j0 + M0 (i1 + H1 i2)] = sum_i0 W0[j0][i0] X[i0 + H0 (i1 + H1 i2)]
    for (int e = tid; e < ncone * H1 * H2; e += 256)
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
            if constexpr (RW)
                asm volatile("" ::: "memory");
        }
    }
    __syncthreads();
    // axis 1: pencils (j0, i2): X[j0 + M0 (j1 + M1 i2)] = sum_i1 W1[j1][i1] Y[j0 + M0 (i1 + H1 i2)]
    for (int e = tid; e < ncone * M0 * H2; e += 256)
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
            if constexpr (RW)
                asm volatile("" ::: "memory");
        }
    }
    __syncthreads();
    // axis 2: pencils (j0, j1): coefficient (j0, j1, j2) = sum_i2 W2[j2][i2] X[j0 + M0 (j1 + M1 i2)],
    // written to its sub-cone's block when the sub-cone is active (active blocks are stored
    // contiguously from the cone's fine_base, in sub-cone order)
    for (int e = tid; e < ncone * M0 * M1; e += 256)
    {
        const int k = e / (M0 * M1), p = e - k * (M0 * M1);
        const int j0 = p % M0, j1 = p / M0;
        const int s0 = j0 / L0, s1 = j1 / L1;
        const int m01 = (j0 - s0 * L0) + L0 * (j1 - s1 * L1);
        const uint32_t mask = cmask[k];
        const int base = cbase[k];
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
            const int s2 = j / L2, m2 = j - s2 * L2;
            const int sub = s0 + F0 * (s1 + F1 * s2);
            if ((mask >> sub) & 1u)
            {
                const int rk = __popc(mask & ((1u << sub) - 1u));
                Fout[int64_t(base + rk) * NLO + m01 + L0 * L1 * m2] = a;
            }
            if constexpr (RW)
                asm volatile("" ::: "memory");
        }
    }
}

// runtime-order version of the transform (any orders and factors)
template <typename R>
__global__ __launch_bounds__(256) void k_transform(int64_t c0, int64_t c1, int cpb, int xsz,
                                                   int ysz, const ConeDev* __restrict__ cones,
                                                   const Cx<R>* __restrict__ V,
                                                   Cx<R>* __restrict__ F, const R* __restrict__ W,
                                                   int w1off, int w2off, KParams kp)
{
    extern __shared__ unsigned char smem_raw[];
    R* Ws = reinterpret_cast<R*>(smem_raw);
    const int h0 = kp.ho0, h1 = kp.ho1, h2 = kp.ho2;
    const int m0 = kp.f0 * kp.lo0, m1 = kp.f1 * kp.lo1, m2 = kp.f2 * kp.lo2;
    const int nW = m0 * h0 + m1 * h1 + m2 * h2;
    const int wpad = (nW + 3) & ~3;
    Cx<R>* X = reinterpret_cast<Cx<R>*>(Ws + wpad);
    Cx<R>* Y = X + int64_t(cpb) * xsz;
    for (int i = threadIdx.x; i < nW; i += blockDim.x)
        Ws[i] = W[i];
    const int64_t first = c0 + xcd_block(kp.xlg) * cpb;
    const int ncone = int(min<int64_t>(cpb, c1 - first));
    const int nh = h0 * h1 * h2;
    for (int e = threadIdx.x; e < ncone * nh; e += blockDim.x)
    {
        const int k = e / nh, i = e - k * nh;
        X[k * xsz + i] = V[(first + k) * nh + i];
    }
    __syncthreads();
    const R* W0 = Ws;
    const R* W1 = Ws + w1off;
    const R* W2 = Ws + w2off;
    {
        const int n = m0 * h1 * h2;
        for (int e = threadIdx.x; e < ncone * n; e += blockDim.x)
        {
            const int k = e / n, i = e - k * n;
            const int j0 = i % m0, rest = i / m0;
            const Cx<R>* x = X + k * xsz + rest * h0;
            Cx<R> a = {0, 0};
            for (int t = 0; t < h0; t++)
            {
                const R wv = W0[j0 * h0 + t];
                a.x += wv * x[t].x;
                a.y += wv * x[t].y;
            }
            Y[k * ysz + i] = a;
        }
    }
    __syncthreads();
    {
        const int n = m0 * m1 * h2;
        for (int e = threadIdx.x; e < ncone * n; e += blockDim.x)
        {
            const int k = e / n, i = e - k * n;
            const int j0 = i % m0, j1 = (i / m0) % m1, i2 = i / (m0 * m1);
            const Cx<R>* y = Y + k * ysz + j0 + m0 * h1 * i2;
            Cx<R> a = {0, 0};
            for (int t = 0; t < h1; t++)
            {
                const R wv = W1[j1 * h1 + t];
                a.x += wv * y[m0 * t].x;
                a.y += wv * y[m0 * t].y;
            }
            X[k * xsz + i] = a;
        }
    }
    __syncthreads();
    {
        const int n = m0 * m1 * m2;
        const int lo0 = kp.lo0, lo1 = kp.lo1;
        for (int e = threadIdx.x; e < ncone * n; e += blockDim.x)
        {
            const int k = e / n, i = e - k * n;
            const int j01 = i % (m0 * m1), j2 = i / (m0 * m1);
            const Cx<R>* x = X + k * xsz + j01;
            Cx<R> a = {0, 0};
            for (int t = 0; t < h2; t++)
            {
                const R wv = W2[j2 * h2 + t];
                a.x += wv * x[m0 * m1 * t].x;
                a.y += wv * x[m0 * m1 * t].y;
            }
            const int j0 = j01 % m0, j1 = j01 / m0;
            const int s0 = j0 / lo0, s1 = j1 / lo1, s2 = j2 / kp.lo2;
            const int sub = s0 + kp.f0 * (s1 + kp.f1 * s2);
            const ConeDev cn = cones[first + k];
            if ((cn.mask >> sub) & 1u)
            {
                const int slot = cn.fine_base + __popc(cn.mask & ((1u << sub) - 1u));
                const int coef =
                    (j0 - s0 * lo0) + lo0 * ((j1 - s1 * lo1) + lo1 * (j2 - s2 * kp.lo2));
                F[int64_t(slot) * kp.n_lo + coef] = a;
            }
        }
    }
}

// far field: for each target of the item, the interpolants of its far source boxes times the
// centre factor (ifgfoperator.hpp:962-1012, 1708-1891).  FAST (mixed form): cones located by
// locate_fast with the level's fine-grid angle tables (doubles [ang_base, ang_base + ang_n) of
// ang, shared by every source box of the level) staged in LDS; otherwise the exact locate.
constexpr int FAR_STAGE = 4; // distinct cones per wave and source box staged in LDS

// one-wave exchanges through LDS: the wave's LDS writes have completed before its next LDS
// reads issue (other lanes read what a lane wrote); the memory clobber keeps the compiler from
// moving LDS accesses across
__device__ __forceinline__ void lds_wait()
{
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
}

template <bool PH, typename R, int O0, int O1, int O2, bool FAST, int WPB, bool DEC>
// the mixed order-6 kernel fits 64 VGPRs without spilling: ask for 8 waves per SIMD (the far
// field is latency-bound; 74 registers would allow 6); at larger orders 4 waves (128 VGPRs)
__global__ __launch_bounds__(64 * WPB)
__attribute__((amdgpu_waves_per_eu(FAST && sizeof(R) == 4 ? (O1 == 3 ? 8 : 4) : 1))) void k_far(const WorkItem* __restrict__ items,
                                                 int64_t nitems,
                                            const int32_t* __restrict__ srange,
                                            const int32_t* __restrict__ far_src,
                                            const BoxGeo* __restrict__ geo,
                                            const CellRef* __restrict__ ctab,
                                            const Cx<R>* __restrict__ F,
                                            const double* __restrict__ tpts,
                                            const double* __restrict__ ang, int ang_base,
                                            int ang_n, Cx<R>* __restrict__ result, KParams kp)
{
    extern __shared__ double sang[];
    __shared__ float4 fstage[WPB][FAST && O0 * O1 * O2 >= 50 ? FAR_STAGE * (O0 * O1 * O2 / 2) : 1];
    // WPB waves per workgroup, one item each, sharing the staged angle tables (one copy per
    // workgroup: the tables, not the registers, would otherwise bound the waves per CU)
    const int lane = WPB == 1 ? int(threadIdx.x) : int(threadIdx.x & 63);
    const int64_t it =
        WPB == 1 ? xcd_block(kp.xlg) : xcd_block(kp.xlg) * WPB + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (WPB == 1 && it >= nitems)
        return;
    if constexpr (FAST)
    {
        for (int i = threadIdx.x; i < ang_n; i += 64 * WPB)
            sang[i] = ang[ang_base + i];
        __syncthreads();
    }
    if (WPB > 1 && it >= nitems)
        return;
    const WorkItem wi = items[it];
    const int Tp = pow2_at_least(wi.count);
    const int ti = lane & (Tp - 1);
    const int grp = lane / Tp, ngrp = 64 / Tp;
    const bool valid = ti < wi.count;
    const int64_t t = int64_t(wi.first) + (valid ? ti : 0);
    const double y0 = tpts[3 * t], y1 = tpts[3 * t + 1], y2 = tpts[3 * t + 2];
    Cx<R> acc = {0, 0};
    const int32_t j0 = srange[2 * it], j1 = srange[2 * it + 1];
    auto body = [&](const BoxGeo& B)
    {
        int fine;
        double r, rinv;
        R x0, x1, x2;
        if constexpr (FAST)
        {
            LocF L;
            const double* eth = sang + (2 * B.fth_off - ang_base);
            const double* eph = sang + (2 * B.fph_off - ang_base);
            constexpr int FF = O0 > 0 ? 3 : 0; // the compiled order classes use factor 3
            if (!locate_fast<FF>(B, eth, eph, ctab, y0 - B.ctr[0], y1 - B.ctr[1], y2 - B.ctr[2],
                                 kp.f0, kp.f1, kp.f2, L))
                return;
            fine = L.fine;
            r = L.r;
            rinv = L.rinv;
            x0 = L.xi[0];
            x1 = L.xi[1];
            x2 = L.xi[2];
        }
        else
        {
            Loc L;
            if (!locate(B, ctab, y0 - B.ctr[0], y1 - B.ctr[1], y2 - B.ctr[2], kp.f0, kp.f1, kp.f2,
                        L))
                return;
            fine = L.fine;
            r = L.r;
            rinv = 1.0 / L.r;
            x0 = R(L.xi[0]);
            x1 = R(L.xi[1]);
            x2 = R(L.xi[2]);
        }
        Cx<R> val;
        if constexpr (FAST && O0 > 0)
        {
            // FP32, compiled orders: the interpolant from its Chebyshev values, the cone's
            // coefficients loaded first as whole float4
            float t0[1][O0], t1[1][O1], t2[1][O2];
            cheb_t<O0>(x0, t0[0]);
            cheb_t<O1>(x1, t1[0]);
            cheb_t<O2>(x2, t2[0]);
            float2 v[1];
            constexpr int NC = O0 * O1 * O2;
            constexpr int NQ = NC / 2; // float4 per cone
            float2 cc[NC];
            // at 50 and more coefficients the distinct cones of the wave's lanes (2.4 on
            // average at 1M points; up to FAR_STAGE) are read from memory once, by all lanes
            // together, into LDS, and every lane reads its cone there: the same values with
            // a tenth of the vector-memory traffic.  A wave with more distinct cones, and
            // every wave at fewer coefficients, reads memory directly.
            bool staged = false;
            if constexpr (NC >= 50)
            {
                const uint64_t act = __ballot(1);
                const int rank = __popcll(act & ((1ull << lane) - 1ull));
                const int nact = __popcll(act);
                uint64_t rem = act;
                int myslot = 0, nd = 0;
                int cs[FAR_STAGE] = {};
#pragma unroll 1
                while (rem != 0 && nd < FAR_STAGE)
                {
                    const int l0 = __ffsll((unsigned long long)rem) - 1;
                    const int f0 = __builtin_amdgcn_readlane(fine, l0);
                    const uint64_t m = __ballot(fine == f0);
                    if (fine == f0)
                        myslot = nd;
#pragma unroll
                    for (int e = 0; e < FAR_STAGE; e++)
                        cs[e] = nd == e ? f0 : cs[e];
                    rem &= ~m;
                    nd++;
                }
                if (rem == 0) // uniform
                {
                    float4* st4 = fstage[WPB == 1 ? 0 : (threadIdx.x >> 6)];
                    for (int i = rank; i < nd * NQ; i += nact)
                    {
                        const int d = i / NQ, k = i - d * NQ;
                        int cf = cs[0];
#pragma unroll
                        for (int e = 1; e < FAR_STAGE; e++)
                            cf = d == e ? cs[e] : cf;
                        st4[i] = reinterpret_cast<const float4*>(F + int64_t(cf) * NC)[k];
                    }
                    lds_wait(); // the wave's LDS writes are complete before it reads them
                    const float4* q4 = st4 + myslot * NQ;
#pragma unroll
                    for (int i = 0; i < NQ; i++)
                    {
                        const float4 q = q4[i];
                        cc[2 * i] = make_float2(q.x, q.y);
                        cc[2 * i + 1] = make_float2(q.z, q.w);
                    }
                    eval_cone_g<O0, O1, O2, 1>(cc, t0, t1, t2, v);
                    staged = true;
                }
            }
            if (!staged)
            {
                const float4* p4 = reinterpret_cast<const float4*>(F + int64_t(fine) * NC);
#pragma unroll
                for (int i = 0; i < NQ; i++)
                {
                    const float4 q = p4[i];
                    cc[2 * i] = make_float2(q.x, q.y);
                    cc[2 * i + 1] = make_float2(q.z, q.w);
                }
                eval_cone_g<O0, O1, O2, 1>(cc, t0, t1, t2, v);
            }
            // the next source box's staging writes stay behind this box's LDS reads
            asm volatile("" ::: "memory");
            val = {v[0].x, v[0].y};
        }
        else
            val = fine_eval<O0, O1, O2, R>(F, fine, kp, x0, x1, x2);
        Cx<R> cf;
        if constexpr (!PH)
            cf = {R(INV_4PI_D * rinv), R(0)};
        else
            cf = cscale(expmk_far<R, DEC>(B.k_re, B.k_im, r), R(INV_4PI_D * rinv));
        acc = cadd(acc, cmul(val, cf));
    };
    // the next source box's index is loaded one box ahead, so that its geometry loads do not
    // wait for it
    if (ngrp == 1)
    {
        // uniform source box: scalar loads of its geometry
        int32_t sn = j0 < j1 ? far_src[j0] : 0;
        for (int32_t j = j0; j < j1; j++)
        {
            const int32_t sb = __builtin_amdgcn_readfirstlane(sn);
            if (j + 1 < j1)
                sn = far_src[j + 1];
            body(geo[sb]);
        }
    }
    else
    {
        int32_t sn = j0 + grp < j1 ? far_src[j0 + grp] : 0;
        for (int32_t j = j0 + grp; j < j1; j += ngrp)
        {
            const int32_t sb = sn;
            if (j + ngrp < j1)
                sn = far_src[j + ngrp];
            body(geo[sb]);
        }
        for (int m = Tp; m < 64; m <<= 1)
        {
            acc.x += __shfl_xor(acc.x, m, 64);
            acc.y += __shfl_xor(acc.y, m, 64);
        }
    }
    if (grp == 0 && valid)
        result[t] = cadd(result[t], acc);
}

// ------------------------------------------------------------------------------------
// host helpers
// ------------------------------------------------------------------------------------
template <typename T> T* dalloc(size_t n, int64_t& total)
{
    T* p = nullptr;
    if (n == 0)
        n = 1;
    HIP_OK(hipMalloc(&p, n * sizeof(T)));
    total += int64_t(n * sizeof(T));
    return p;
}

template <typename T> T* dupload(const std::vector<T>& v, int64_t& total)
{
    T* p = dalloc<T>(v.size(), total);
    if (!v.empty())
        HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

static std::vector<WorkItem> make_items(const std::vector<int32_t>& boxes, const Tree& T)
{
    std::vector<WorkItem> out;
    for (int32_t b : boxes)
    {
        const int64_t s = T.begin[b], e = T.end[b];
        for (int64_t f = s; f < e; f += 64)
            out.push_back({b, int32_t(f), int32_t(std::min<int64_t>(64, e - f)), 0});
    }
    return out;
}




// up_level >= 0: the upward pass into parent level up_level; otherwise the far field of
// target level far_level
template <bool PH, typename R, int H0, int H1, int H2, int O0, int O1, int O2, bool DEC>
void launch_up_far_t(DevicePlan& D, hipStream_t st, const KParams& kp, const NodeTabs& tabs,
                     int up_level, int far_level, int64_t a, int64_t b, Cx<R>* F, Cx<R>* V,
                     Cx<R>* res, Cx<R>* Fo)
{
    const bool up = up_level >= 0;
    if (up)
    {
        const int64_t i0 = D.up_item_off[up_level], i1 = D.up_item_off[up_level + 1];
        if constexpr (sizeof(R) == 4 && O0 > 0)
        {
            // FP32, compiled orders: the grouped tables, the cones of an item one at a time,
            // the items in tiles of boxes x cells
            const int64_t t0 = D.up_tl_item_off[up_level], t1 = D.up_tl_item_off[up_level + 1];
            // (at the larger orders their transform included: the cones' fine coefficients
            // go to Fo, the values nowhere)
            constexpr bool TF = up_fuses_transform(O0 * O1 * O2);
            if (t1 > t0 && !D.up_tl_desc)
                throw std::runtime_error("upward pass: the per-cone records were not uploaded");
            if (t1 > t0)
                k_up_grp<O0, O1, O2, H0 * H1 * H2, 128, H0, H1, H2, TF>
                    <<<dim3(unsigned(t1 - t0)), dim3(128), 0, st>>>(
                        D.up_tl_items + t0, t1 - t0, D.up_tl_desc + D.up_gl_off[up_level],
                        D.ctab, static_cast<const UpGrp*>(D.up_grp), F, V,
                        static_cast<const float*>(D.wmat), Fo, kp);
        }
        else if (i1 > i0)
        {
#define IFGF_UP_CELL(K)                                                                         \
    k_up_cell<R, O0, O1, O2, K><<<dim3(unsigned(i1 - i0)), dim3(128), 0, st>>>(                \
        D.up_items + i0, i1 - i0, D.up_glist + D.up_gl_off[up_level], D.cones, D.geo, D.child,  \
        D.ctab, static_cast<const UpEnt<R>*>(D.up_entries), D.up_part_ent, F, V, kp)
            switch (D.shape.up_k)
            {
            case 2: // items of at most 2 cones run as items of 4
            case 4:
                IFGF_UP_CELL(4);
                break;
            case 16:
                IFGF_UP_CELL(16);
                break;
            default:
                IFGF_UP_CELL(8);
            }
#undef IFGF_UP_CELL
        }
        if (b > a)
            k_up<PH, R, O0, O1, O2><<<dim3(unsigned(b - a)), dim3(128), 0, st>>>(
                D.up_cones + a, b - a, D.cones, D.geo, D.child, D.ctab, tabs, F, V, kp);
    }
    else
    {
        // far_level: the far group; a, b: its items
        const int abase = D.far_groups[far_level].ang_base, an = D.far_groups[far_level].ang_n;
        const int64_t n = b - a;
#define IFGF_FAR(FAST, WPB, LDS, BASE, AN)                                                      \
    k_far<PH, R, O0, O1, O2, FAST, WPB, DEC>                                                    \
        <<<dim3(unsigned((n + WPB - 1) / WPB)), dim3(64 * WPB), LDS, st>>>(                     \
            D.far_items + a, n, D.far_srange + 2 * a, D.far_src, D.geo, D.ctab, F, D.tgt_pts,   \
            D.fine_ang, BASE, AN, res, kp)
        if constexpr (sizeof(R) == 4)
            if (an > 0)
            {
                const size_t lds = size_t(an) * sizeof(double);
                switch (D.shape.far_waves)
                {
                case 2:
                    IFGF_FAR(true, 2, lds, abase, an);
                    break;
                case 8:
                    IFGF_FAR(true, 8, lds, abase, an);
                    break;
                case 4:
                    IFGF_FAR(true, 4, lds, abase, an);
                    break;
                default:
                    IFGF_FAR(true, 1, lds, abase, an);
                }
                return;
            }
        IFGF_FAR(false, 1, 0, 0, 0);
#undef IFGF_FAR
    }
}

// DEC: the kernel family's wavenumber may have a real part (e^{-k r} decays); false for the
// real-kappa families, whose far field then forms the phase only
template <bool PH, typename R, bool DEC>
void launch_up_far(DevicePlan& D, hipStream_t st, const KParams& kp, const NodeTabs& tabs, int up,
                   int far, int64_t a, int64_t b, Cx<R>* F, Cx<R>* V, Cx<R>* res,
                   Cx<R>* Fo = nullptr)
{
    switch (order_class(D))
    {
    case 1:
        return launch_up_far_t<PH, R, 3, 6, 6, 2, 3, 3, DEC>(D, st, kp, tabs, up, far, a, b, F, V, res, Fo);
    case 2:
        return launch_up_far_t<PH, R, 5, 8, 8, 2, 5, 5, DEC>(D, st, kp, tabs, up, far, a, b, F, V, res, Fo);
    case 3:
        return launch_up_far_t<PH, R, 6, 8, 7, 3, 5, 4, DEC>(D, st, kp, tabs, up, far, a, b, F, V, res, Fo);
    default:
        return launch_up_far_t<PH, R, 0, 0, 0, 0, 0, 0, DEC>(D, st, kp, tabs, up, far, a, b, F, V, res, Fo);
    }
}

// transform cones per workgroup the compiled kernels offer per (order class, real size), the
// default first (LDS: 6.5 KB per cone at order 6 in FP32, 16.5 KB at order 8, 20 KB for
// (6,8,7)/(3,5,4); twice that in FP64)
static std::vector<int> transform_cpb_choices(int cls, size_t rsize)
{
    if (cls == 1)
        return rsize == 4 ? std::vector<int>{4, 2, 6, 8} : std::vector<int>{3, 1, 2, 4};
    if (cls == 2 || cls == 3)
        return rsize == 4 ? std::vector<int>{2, 1, 3} : std::vector<int>{1};
    return {0};
}

template <typename R>
void launch_transform(DevicePlan& D, hipStream_t st, const KParams& kp, int64_t c0, int64_t c1,
                      Cx<R>* V, Cx<R>* F, const int32_t* list = nullptr)
{
    const R* W = static_cast<const R*>(D.wmat);
    const int cls = order_class(D);
#define IFGF_TRANSFORM(H0, H1, H2, L0, L1, L2, CPB)                                            \
    k_transform_t<R, H0, H1, H2, L0, L1, L2, 3, 3, 3, CPB>                                      \
        <<<dim3(unsigned((c1 - c0 + CPB - 1) / CPB)), dim3(256), 0, st>>>(c0, c1, list, D.cones, V, \
                                                                          F, W,                 \
                                                                          kp.xlg)
    const int cpb_sel =
        D.shape.cpb > 0 ? D.shape.cpb : transform_cpb_choices(cls, sizeof(R))[0];
    if (cls == 1)
    {
        switch (cpb_sel)
        {
        case 1:
            IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 1);
            break;
        case 2:
            IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 2);
            break;
        case 3:
            IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 3);
            break;
        case 4:
            IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 4);
            break;
        case 8:
            if constexpr (sizeof(R) == 4)
            {
                IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 8);
                break;
            }
            [[fallthrough]];
        default:
            if constexpr (sizeof(R) == 4)
                IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 6);
            else
                IFGF_TRANSFORM(3, 6, 6, 2, 3, 3, 3);
        }
        return;
    }
    if (cls == 2)
    {
        switch (cpb_sel)
        {
        case 3:
            if constexpr (sizeof(R) == 4)
            {
                IFGF_TRANSFORM(5, 8, 8, 2, 5, 5, 3);
                break;
            }
            [[fallthrough]];
        case 2:
            if constexpr (sizeof(R) == 4)
            {
                IFGF_TRANSFORM(5, 8, 8, 2, 5, 5, 2);
                break;
            }
            [[fallthrough]];
        default:
            IFGF_TRANSFORM(5, 8, 8, 2, 5, 5, 1);
        }
        return;
    }
    if (cls == 3)
    {
        switch (cpb_sel)
        {
        case 3:
            if constexpr (sizeof(R) == 4)
            {
                IFGF_TRANSFORM(6, 8, 7, 3, 5, 4, 3);
                break;
            }
            [[fallthrough]];
        case 2:
            if constexpr (sizeof(R) == 4)
            {
                IFGF_TRANSFORM(6, 8, 7, 3, 5, 4, 2);
                break;
            }
            [[fallthrough]];
        default:
            IFGF_TRANSFORM(6, 8, 7, 3, 5, 4, 1);
        }
        return;
    }
#undef IFGF_TRANSFORM
    const int h0 = D.ho[0], h1 = D.ho[1], h2 = D.ho[2];
    const int m0 = D.f[0] * D.lo[0], m1 = D.f[1] * D.lo[1], m2 = D.f[2] * D.lo[2];
    // X holds the values and the axis-1 stage, Y the axis-0 stage; axis 2 goes to memory
    const int xsz = std::max(h0 * h1 * h2, m0 * m1 * h2);
    const int ysz = m0 * h1 * h2;
    const int nW = m0 * h0 + m1 * h1 + m2 * h2;
    const size_t per_cone = size_t(xsz + ysz) * sizeof(Cx<R>);
    const int cpb = int(std::max<size_t>(1, std::min<size_t>(8, (48 * 1024) / per_cone)));
    const size_t smem = ((nW + 3) & ~3) * sizeof(R) + cpb * per_cone;
    constexpr size_t LDS_MAX = 160 * 1024; // gfx950: 160 KB of LDS per CU
    if (smem > LDS_MAX)
        throw std::runtime_error("interpolation orders too large for the transform kernel's LDS");
    if (smem > 64 * 1024)
        HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_transform<R>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(smem)));
    k_transform<R><<<dim3(unsigned((c1 - c0 + cpb - 1) / cpb)), dim3(256), smem, st>>>(
        c0, c1, cpb, xsz, ysz, D.cones, V, F, W, D.woff[1], D.woff[2], kp);
}

// phase timing: events recorded around each segment on the mult's stream, read after the
// last segment (one synchronisation per mult); phases near (with gather and zeroing), leaf,
// upward, transform, far, and the whole mult
struct PhaseTimer
{
    bool on;
    hipStream_t st;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> segs;
    hipEvent_t open = nullptr;
    int open_phase = -1;
    void begin(int phase)
    {
        if (!on)
            return;
        HIP_OK(hipEventCreate(&open));
        HIP_OK(hipEventRecord(open, st));
        open_phase = phase;
    }
    void end()
    {
        if (!on)
            return;
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        HIP_OK(hipEventRecord(e, st));
        segs.push_back({open_phase, {open, e}});
    }
    void finish(float* out)
    {
        if (!on || segs.empty())
            return;
        HIP_OK(hipEventSynchronize(segs.back().second.second));
        for (int i = 0; i < 6; i++)
            out[i] = 0;
        for (auto& sg : segs)
        {
            float x = 0;
            HIP_OK(hipEventElapsedTime(&x, sg.second.first, sg.second.second));
            out[sg.first] += x;
        }
        float tot = 0;
        HIP_OK(hipEventElapsedTime(&tot, segs.front().second.first, segs.back().second.second));
        out[5] = tot;
        for (auto& sg : segs)
        {
            (void)hipEventDestroy(sg.second.first);
            (void)hipEventDestroy(sg.second.second);
        }
        segs.clear();
    }
};

// The Morton-order densities and results of one single-column mult: the plan's own vectors,
// or the allocation-free mult's fixed scratch (which also keeps the mult on one stream).
struct Stage
{
    void* ws;
    void* res;
    bool side; // the side stream of the overlapped near and far fields may be used
};
inline Stage plan_stage(const DevicePlan& D) { return {D.wsorted, D.result, true}; }

// everything of a mult between the gather of the densities and the scatter of the results:
// with the plan's stage it reads and writes only the plan's buffers, so one captured graph
// serves every call
// leaf sampling of nb leaf boxes; fused: the kernel also transforms the boxes' cones to their
// fine coefficients (in F)
template <int KIND, typename R>
void launch_leaf(DevicePlan& D, hipStream_t st, const WorkItem* boxes, int64_t nb, Cx<R>* ws,
                 const NodeTabs& tabs, Cx<R>* V, Cx<R>* F, const KParams& kp, bool fused)
{
    if constexpr (sizeof(R) == 4)
        if (fused)
        {
            const float* W = static_cast<const float*>(D.wmat);
#define IFGF_LEAF_TF(H0, H1, H2, L0, L1, L2)                                                    \
    k_leaf_box<KIND, R, H0, H1, H2, L0, L1, L2><<<dim3(unsigned(nb)), dim3(256), 0, st>>>(       \
        boxes, nb, D.cones, D.geo, D.src_begin, D.src_end, D.src_pts, D.normals, ws, tabs, V, kp, \
        W, F)
            switch (order_class(D))
            {
            case 2:
                IFGF_LEAF_TF(5, 8, 8, 2, 5, 5);
                return;
            case 3:
                IFGF_LEAF_TF(6, 8, 7, 3, 5, 4);
                return;
            default:
                throw std::runtime_error("fused leaf transform for an order class without one");
            }
#undef IFGF_LEAF_TF
        }
    k_leaf_box<KIND, R><<<dim3(unsigned(nb)), dim3(256), 0, st>>>(
        boxes, nb, D.cones, D.geo, D.src_begin, D.src_end, D.src_pts, D.normals, ws, tabs, V, kp);
}

template <int KIND, typename R>
void run_core(DevicePlan& D, hipStream_t st, PhaseTimer& pt, const Stage& sg)
{
    constexpr bool PH = KIND != KL;
    constexpr bool DEC = KIND != KH && KIND != KC; // real kappa: kr = 0
    // k_up_grp transforms the cones it computes (FP32, compiled orders, enough coefficients)
    const bool fused = sizeof(R) == 4 && up_fused_transform(D);
    const KParams kp = kparams(D);
    const NodeTabs tabs{D.ho_nodes, D.trig_th, D.trig_ph};
    Cx<R>* ws = static_cast<Cx<R>*>(sg.ws);
    Cx<R>* res = static_cast<Cx<R>*>(sg.res);
    Cx<R>* V = static_cast<Cx<R>*>(D.V);
    Cx<R>* F = static_cast<Cx<R>*>(D.F);
    // the values of level l and the fine coefficients of level l: the whole arrays, or with level
    // storage the level's buffer addressed as if it held every level (the kernels index by
    // global cone id and fine slot, and touch only the level's range)
    auto Vl = [&](int l) -> Cx<R>*
    {
        if (!D.ring)
            return V;
        return reinterpret_cast<Cx<R>*>(reinterpret_cast<uintptr_t>(D.V) -
                                        uintptr_t(D.cone_lvl_off[l]) * D.n_hi * sizeof(Cx<R>));
    };
    auto Fl = [&](int l) -> Cx<R>*
    {
        if (!D.ring)
            return F;
        return reinterpret_cast<Cx<R>*>(reinterpret_cast<uintptr_t>(D.F_bank[l & 1]) -
                                        uintptr_t(D.fine_lvl_off[l]) * D.n_lo * sizeof(Cx<R>));
    };
    // near and far fields: SidePipe (overlapped with the upward chain on the side stream when
    // the shape asks for it; phase timing runs everything in order on one stream)
    const int overlap = (pt.on || !D.n_near_items || !sg.side) ? 0 : D.shape.overlap;
    pt.begin(0);
    k_zero<R><<<dim3(unsigned((D.ntgt + 255) / 256)), dim3(256), 0, st>>>(res, D.ntgt);
    auto pipe = side_pipe(
        D, st, overlap,
        [&](hipStream_t s)
        {
            if (D.n_near_items)
                k_near<KIND, R><<<dim3(unsigned(D.n_near_items)), dim3(64), 0, s>>>(
                    D.near_items, D.n_near_items, D.near_off, D.near_src, D.geo, D.src_begin,
                    D.src_end, D.src_pts, D.normals, ws, D.tgt_pts, res, kp);
        },
        [&](int g, hipStream_t s)
        {
            const DevicePlan::FarGroup& G = D.far_groups[g];
            launch_up_far<PH, R, DEC>(D, s, kp, tabs, -1, g, G.item0, G.item1, Fl(G.src_level), V,
                                 res);
        });
    if (overlap != 1)
        pipe.fork(); // the near field first, on the main stream (inside the timed near phase)
    pt.end();
    if (!D.ring && D.n_leaf_boxes)
    {
        // every level's leaf cones at once (with level storage: a level at a time, below)
        pt.begin(1);
        launch_leaf<KIND, R>(D, st, D.leaf_boxes, D.n_leaf_boxes, ws, tabs, V, F, kp, fused);
        pt.end();
    }
    for (int l = D.levels - 1; l >= 0; l--)
    {
        const int64_t u0 = D.up_off[l], u1 = D.up_off[l + 1];
        const int64_t c0 = D.cone_lvl_off[l], c1 = D.cone_lvl_off[l + 1];
        if (c1 > c0)
            pipe.reuse(l + 2); // level storage: F_bank[l & 1] held level l + 2
        if (D.ring && D.leaf_lvl_off[l + 1] > D.leaf_lvl_off[l])
        {
            pt.begin(1);
            const int64_t a = D.leaf_lvl_off[l], nb = D.leaf_lvl_off[l + 1] - a;
            launch_leaf<KIND, R>(D, st, D.leaf_boxes + a, nb, ws, tabs, Vl(l), Fl(l), kp, fused);
            pt.end();
        }
        if (u1 > u0 || D.up_item_off[l + 1] > D.up_item_off[l])
        {
            pt.begin(2);
            launch_up_far<PH, R, DEC>(D, st, kp, tabs, l, -1, u0, u1, Fl(l + 1), Vl(l), res,
                                      Fl(l));
            pt.end();
        }
        if (c1 > c0)
        {
            pt.begin(3);
            if (fused) // leaf sampling and the grouped upward kernel transformed their cones
            {
                const int64_t t0 = D.tr_list_off[l], t1 = D.tr_list_off[l + 1];
                if (t1 > t0)
                    launch_transform<R>(D, st, kp, t0, t1, Vl(l), Fl(l), D.tr_list);
            }
            else
                launch_transform<R>(D, st, kp, c0, c1, Vl(l), Fl(l));
            pt.end();
            pipe.fork();
        }
        if (D.ring && pt.on)
        {
            pt.begin(4); // level storage: this level's far groups run here
            pipe.ready(l);
            pt.end();
        }
        else
            pipe.ready(l);
    }
    pt.begin(4);
    pipe.finish();
    pt.end();
}

// one mult: gather, the core (eagerly, or as the plan's captured graph), scatter
template <int KIND, typename R>
void run_mult(DevicePlan& D, const double* w_dev, double* out_dev, hipStream_t st, bool timed,
              int io, bool graph, const Stage& sg)
{
    Cx<R>* ws = static_cast<Cx<R>*>(sg.ws);
    Cx<R>* res = static_cast<Cx<R>*>(sg.res);
    PhaseTimer pt{timed, st, {}};
    const double in_im = (io & IO_CONJ_IN) ? -1.0 : 1.0;
    const double out_re = (io & IO_NEGATE_OUT) ? -1.0 : 1.0;
    const double out_im = out_re * ((io & IO_CONJ_OUT) ? -1.0 : 1.0);
    pt.begin(0);
    k_gather<R><<<dim3(unsigned((D.nsrc + 255) / 256)), dim3(256), 0, st>>>(w_dev, D.src_perm, ws,
                                                                           D.nsrc, in_im);
    pt.end();
    if (graph && !timed)
    {
        if (!D.graph_exec)
        {
            // capture on a private stream (the caller's may be the null stream, which cannot
            // be captured); kernels only, no memory set nodes
            hipStream_t cs;
            HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            D.cap_stream = cs;
            PhaseTimer none{false, cs, {}};
            HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            run_core<KIND, R>(D, cs, none, sg);
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamEndCapture(cs, &g));
            hipGraphExec_t ex = nullptr;
            HIP_OK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
            D.graph = g;
            D.graph_exec = ex;
        }
        HIP_OK(hipGraphLaunch(static_cast<hipGraphExec_t>(D.graph_exec), st));
    }
    else
        run_core<KIND, R>(D, st, pt, sg);
    pt.begin(0);
    k_scatter<R><<<dim3(unsigned((D.ntgt + 255) / 256)), dim3(256), 0, st>>>(
        res, D.tgt_perm, out_dev, D.ntgt, out_re, out_im);
    pt.end();
    HIP_OK(hipGetLastError());
    pt.finish(D.last_ms);
}

template <typename R>
void dispatch(DevicePlan& D, const double* w, double* o, hipStream_t st, bool timed, int io,
              bool graph, const Stage& sg)
{
    switch (D.kind)
    {
    case KIND_HELMHOLTZ:
        return run_mult<KH, R>(D, w, o, st, timed, io, graph, sg);
    case KIND_MODHELMHOLTZ:
        return run_mult<KM, R>(D, w, o, st, timed, io, graph, sg);
    case KIND_LAPLACE:
        return run_mult<KL, R>(D, w, o, st, timed, io, graph, sg);
    case KIND_DOUBLELAYER:
        return run_mult<KD, R>(D, w, o, st, timed, io, graph, sg);
    case KIND_COMBINED:
        return run_mult<KC, R>(D, w, o, st, timed, io, graph, sg);
    case KIND_GRAD:
        return run_mult<KG, R>(D, w, o, st, timed, io, graph, sg);
    }
    throw std::runtime_error("unknown kernel kind");
}

// The fixed device scratch of the allocation-free mult (device_mult_noalloc): the Morton-order
// densities and results of one column.  Every GPU the library runs on holds its own copy
// until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch), never by its name.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};

// the scratch's device address on the current device, looked up once per device
void* noalloc_scratch(std::string& err)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_SCRATCH_DEVICES)
    {
        err = "allocation-free mult: no current HIP device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev])
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            err = std::string("allocation-free mult: hipGetSymbolAddress(scratch) failed: ") +
                  hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

} // namespace

bool device_available()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return false;
    return n > 0;
}

int64_t device_transform_cones(const DevicePlan& D)
{
    return up_fused_transform(D) ? D.tr_list_off.back() : D.n_cones;
}

void device_free(DevicePlan& D)
{
    void* ptrs[] = {D.src_pts, D.tgt_pts, D.normals, D.src_perm, D.tgt_perm, D.src_begin,
                    D.src_end, D.child, D.geo, D.cones, D.ctab, D.ho_nodes, D.trig_th,
                    D.trig_ph, D.fine_ang, D.wmat, D.up_part_ent, D.up_entries, D.up_dkey,
                    D.up_doff, D.up_part_cell, D.up_goff, D.up_grp, D.up_items_k[0],
                    D.up_items_k[1], D.up_items_k[2], D.up_items_k[3], D.up_glist,
                    D.up_tl_items_k[0], D.up_tl_items_k[1], D.up_tl_items_k[2],
                    D.up_tl_items_k[3], D.up_tl_glist, D.up_tl_gcell, D.up_tl_desc, D.tr_list,
                    D.leaf_cones, D.leaf_boxes, D.up_cones, D.far_srange, D.far_src,
                    D.near_off, D.near_src, D.far_items, D.near_items, D.wsorted, D.result, D.V,
                    D.F, D.F_bank[0], D.F_bank[1], D.b_ws, D.b_res, D.b_V, D.b_F, D.b_in,
                    D.b_out, D.h_in, D.h_out};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    if (D.graph_exec)
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(D.graph_exec));
    if (D.graph)
        (void)hipGraphDestroy(static_cast<hipGraph_t>(D.graph));
    if (D.b_graph_exec)
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(D.b_graph_exec));
    if (D.b_graph)
        (void)hipGraphDestroy(static_cast<hipGraph_t>(D.b_graph));
    if (D.cap_stream)
        (void)hipStreamDestroy(static_cast<hipStream_t>(D.cap_stream));
    if (D.side_stream)
        (void)hipStreamDestroy(static_cast<hipStream_t>(D.side_stream));
    if (D.ev_fork)
        (void)hipEventDestroy(static_cast<hipEvent_t>(D.ev_fork));
    if (D.ev_near)
        (void)hipEventDestroy(static_cast<hipEvent_t>(D.ev_near));
    for (void* e : D.ev_level)
        if (e)
            (void)hipEventDestroy(static_cast<hipEvent_t>(e));
    for (void* e : D.ev_far_done)
        if (e)
            (void)hipEventDestroy(static_cast<hipEvent_t>(e));
    D = DevicePlan();
}

std::string device_upload(const Plan& P, int precision, DevicePlan& D, int ring)
{
    try
    {
        device_free(D);
        HIP_OK(hipGetDevice(&D.device));
        {
            // the allocation-free mult's scratch address on this device (once per device)
            std::string serr;
            if (!noalloc_scratch(serr))
                throw std::runtime_error(serr);
        }
        D.prec = precision;
        D.kind = P.K.kind;
        D.k_re = P.K.k.real();
        D.k_im = P.K.k.imag();
        D.kappa = P.K.kappa;
        D.dx = P.K.dx;
        for (int d = 0; d < 3; d++)
        {
            D.ho[d] = P.ho_order[d];
            D.lo[d] = P.lo_order[d];
            D.f[d] = P.factor[d];
        }
        D.n_hi = D.ho[0] * D.ho[1] * D.ho[2];
        D.n_lo = D.lo[0] * D.lo[1] * D.lo[2];
        D.nsub = D.f[0] * D.f[1] * D.f[2];
        if (D.nsub > 32)
            return "more than 32 fine sub-cones per coarse cone";
        D.nsrc = P.S.n;
        D.ntgt = P.T.n;
        D.levels = P.S.levels;
        D.n_cones = P.num_cones();
        D.n_fine = P.n_fine;
        if (D.n_cones > INT32_MAX || D.n_fine > INT32_MAX || D.nsrc > INT32_MAX ||
            D.ntgt > INT32_MAX)
            return "problem too large for 32-bit cone/point indices";
        int64_t& tot = D.bytes_total;
        D.src_pts = dupload(P.S.pts, tot);
        D.tgt_pts = dupload(P.T.pts, tot);
        if (!P.normals.empty())
            D.normals = dupload(P.normals, tot);
        D.src_perm = dupload(P.S.perm, tot);
        D.tgt_perm = dupload(P.T.perm, tot);
        D.src_begin = dupload(P.S.begin, tot);
        D.src_end = dupload(P.S.end, tot);
        D.child = dupload(P.S.child, tot);
        // node-angle tables per coarse-grid class (theta cells x nodes, phi cells x nodes), with
        // the plan's arguments (Grid::point) so node positions are bit-identical to the plan's
        const auto xth = chebnodes1d(D.ho[1]), xph = chebnodes1d(D.ho[2]);
        std::vector<double> th, ph;
        std::map<std::pair<int64_t, int64_t>, std::pair<int32_t, int32_t>> cls;
        const int64_t nb = P.S.num_boxes();
        std::vector<BoxGeo> geo(nb);
        for (int64_t g = 0; g < nb; g++)
        {
            BoxGeo& B = geo[g];
            for (int d = 0; d < 3; d++)
            {
                B.ctr[d] = P.S.ctr[3 * g + d];
                B.lo[d] = P.dom_lo[3 * g + d];
                B.hi[d] = P.dom_hi[3 * g + d];
                B.hf[d] = P.hf[3 * g + d];
                B.inv_hf[d] = 1.0 / P.hf[3 * g + d];
                B.hc[d] = P.hc[3 * g + d];
                B.nf[d] = int32_t(P.nf[3 * g + d]);
                B.nc[d] = int32_t(P.nc[3 * g + d]);
            }
            B.H = P.S.side[g];
            B.k_re = P.box_k.empty() ? P.K.k.real() : P.box_k[3 * g];
            B.k_im = P.box_k.empty() ? P.K.k.imag() : P.box_k[3 * g + 1];
            B.kappa = P.box_k.empty() ? P.K.kappa : P.box_k[3 * g + 2];
            B.table_off = P.table_off[g];
            const auto key = std::make_pair(P.nc[3 * g + 1], P.nc[3 * g + 2]);
            auto it = cls.find(key);
            if (it == cls.end())
            {
                const int32_t to = int32_t(th.size() / 2), po = int32_t(ph.size() / 2);
                for (int64_t c = 0; c < P.nc[3 * g + 1]; c++)
                    for (int j = 0; j < D.ho[1]; j++)
                    {
                        const double a = 0.5 * B.hc[1];
                        const double b = B.lo[1] + (double(c) + 0.5) * B.hc[1];
                        const double v = (xth[j] * a) + b;
                        th.push_back(std::cos(v));
                        th.push_back(std::sin(v));
                    }
                for (int64_t c = 0; c < P.nc[3 * g + 2]; c++)
                    for (int j = 0; j < D.ho[2]; j++)
                    {
                        const double a = 0.5 * B.hc[2];
                        const double b = B.lo[2] + (double(c) + 0.5) * B.hc[2];
                        const double v = (xph[j] * a) + b;
                        ph.push_back(std::cos(v));
                        ph.push_back(std::sin(v));
                    }
                it = cls.emplace(key, std::make_pair(to, po)).first;
            }
            B.th_off = it->second.first;
            B.ph_off = it->second.second;
        }
        {
            // fine-grid edge and centre angles per (nf1, nf2) class: (cos, sin) of
            // lo + k h (k = 0..n) then of lo + (k + 1/2) h (k < n), theta then phi
            std::vector<double> fa;
            std::map<std::pair<int64_t, int64_t>, std::pair<int32_t, int32_t>> fcls;
            for (int64_t g = 0; g < nb; g++)
            {
                BoxGeo& B = geo[g];
                const auto key = std::make_pair(P.nf[3 * g + 1], P.nf[3 * g + 2]);
                auto it = fcls.find(key);
                if (it == fcls.end())
                {
                    auto emit = [&](int axis) -> int32_t
                    {
                        const int32_t off = int32_t(fa.size() / 2);
                        const int64_t n = P.nf[3 * g + axis];
                        const double lo = B.lo[axis], h = B.hf[axis];
                        for (int64_t k = 0; k <= n; k++)
                        {
                            const double v = lo + double(k) * h;
                            fa.push_back(std::cos(v));
                            fa.push_back(std::sin(v));
                        }
                        for (int64_t k = 0; k < n; k++)
                        {
                            const double v = lo + (double(k) + 0.5) * h;
                            fa.push_back(std::cos(v));
                            fa.push_back(std::sin(v));
                        }
                        return off;
                    };
                    const int32_t a1 = emit(1);
                    const int32_t a2 = emit(2);
                    it = fcls.emplace(key, std::make_pair(a1, a2)).first;
                }
                B.fth_off = it->second.first;
                B.fph_off = it->second.second;
            }
            D.fine_ang = dupload(fa, tot);
        }
        D.geo = dupload(geo, tot);
        D.trig_th = dupload(th, tot);
        D.trig_ph = dupload(ph, tot);
        std::vector<ConeDev> cones(D.n_cones);
        for (int64_t c = 0; c < D.n_cones; c++)
            cones[c] = {P.cone_box[c], int32_t(P.cone_cell[c]), P.cone_mask[c],
                        int32_t(P.cone_fine[c])};
        D.cones = dupload(cones, tot);
        {
            std::vector<CellRef> ct(P.table.size(), CellRef{0, 0u});
            for (size_t i = 0; i < P.table.size(); i++)
                if (P.table[i] >= 0)
                    ct[i] = {cones[P.table[i]].fine_base, cones[P.table[i]].mask};
            D.ctab = dupload(ct, tot);
        }
        {
            auto r = chebnodes1d(D.ho[0]), t = chebnodes1d(D.ho[1]), p = chebnodes1d(D.ho[2]);
            std::vector<double> nodes;
            for (int i = 0; i < D.ho[2]; i++)
                for (int j = 0; j < D.ho[1]; j++)
                    for (int k = 0; k < D.ho[0]; k++)
                    {
                        nodes.push_back(r[k]);
                        nodes.push_back(t[j]);
                        nodes.push_back(p[i]);
                    }
            D.ho_nodes = dupload(nodes, tot);
        }
        {
            std::vector<double> W;
            for (int d = 0; d < 3; d++)
            {
                D.woff[d] = int(W.size());
                auto m = axis_map(D.ho[d], D.lo[d], D.f[d]);
                W.insert(W.end(), m.begin(), m.end());
            }
            D.woff[3] = int(W.size());
            if (precision == PREC_FP64)
                D.wmat = dupload(W, tot);
            else
            {
                std::vector<float> Wf(W.begin(), W.end());
                D.wmat = dupload(Wf, tot);
            }
        }
        // upward geometry tables: one part per level with a valid table (per geometry and level
        // in a plan of several geometries), entries concatenated; every parent box knows its
        // part
        std::vector<const UpLevel*> parts;
        std::vector<int32_t> box_part(nb, -1);
        if (P.up_parts.empty())
        {
            for (int l = 0; l < D.levels && l < int(P.up.size()); l++)
                if (P.up[l].valid)
                {
                    for (int64_t g = P.S.level_off[l]; g < P.S.level_off[l + 1]; g++)
                        box_part[g] = int32_t(parts.size());
                    parts.push_back(&P.up[l]);
                }
        }
        else
        {
            for (const auto& pp : P.up_parts)
                parts.push_back(&pp.U);
            box_part = P.box_up_part;
        }
        // the FP32 tables' row offsets kept on the host for k_up_grp's per-cone records (below)
        std::vector<int64_t> h_part_cell, h_goff;
        std::vector<int32_t> h_doff, h_dkey;
        {
            std::vector<UpEntF> ef;
            std::vector<UpEntD> ed;
            std::vector<int64_t> part_ent, part_cell;
            std::vector<int32_t> dkey, doff;
            std::vector<UpGrp> grp;
            std::vector<int64_t> goff;
            std::vector<int32_t> gcnt, gnode;
            for (const UpLevel* U : parts)
            {
                part_ent.push_back(int64_t(precision == PREC_FP64 ? ed.size() : ef.size()));
                part_cell.push_back(int64_t(doff.size() / 8));
                if (precision == PREC_FP64)
                {
                    for (const UpEntry& E : U->entries)
                        ed.push_back({E.cf, E.sub, E.xi[0], E.xi[1], E.xi[2], E.tf[0], E.tf[1]});
                    continue;
                }
                const size_t e0 = ef.size();
                for (const UpEntry& E : U->entries)
                    ef.push_back({E.cf, E.sub, float(E.xi[0]), float(E.xi[1]), float(E.xi[2]),
                                  float(E.tf[0]), float(E.tf[1]), -1});
                // the distinct child cones per (cell, octant); neighbouring nodes mostly share
                // the last one found
                const int64_t ncell = int64_t(U->entries.size()) / (int64_t(D.n_hi) * 8);
                for (int64_t cell = 0; cell < ncell; cell++)
                    for (int c = 0; c < 8; c++)
                    {
                        const int32_t k0 = int32_t(dkey.size() / 2);
                        doff.push_back(k0);
                        int last = -1;
                        for (int q = 0; q < D.n_hi; q++)
                        {
                            UpEntF& E = ef[e0 + size_t((cell * D.n_hi + q) * 8 + c)];
                            if (E.cf < 0)
                                continue;
                            const int nk = int(dkey.size() / 2) - k0;
                            auto is = [&](int i) {
                                return dkey[2 * (k0 + i)] == E.cf && dkey[2 * (k0 + i) + 1] == E.sub;
                            };
                            int d = (last >= 0 && is(last)) ? last : -1;
                            for (int i = 0; d < 0 && i < nk; i++)
                                if (is(i))
                                    d = i;
                            if (d < 0)
                            {
                                d = nk;
                                dkey.push_back(E.cf);
                                dkey.push_back(E.sub);
                            }
                            E.d = d;
                            last = d;
                        }
                        // the nodes regrouped by child cone (in list order, nodes ascending
                        // within a cone), runs of UP_GRP
                        const int nk = int(dkey.size() / 2) - k0;
                        goff.push_back(int64_t(grp.size()));
                        gcnt.assign(size_t(nk) + 1, 0);
                        for (int q = 0; q < D.n_hi; q++)
                        {
                            const UpEntF& E = ef[e0 + size_t((cell * D.n_hi + q) * 8 + c)];
                            if (E.cf >= 0)
                                gcnt[size_t(E.d) + 1]++;
                        }
                        for (int i = 0; i < nk; i++)
                            gcnt[i + 1] += gcnt[i];
                        gnode.assign(size_t(gcnt[nk]), 0);
                        for (int q = 0; q < D.n_hi; q++)
                        {
                            const UpEntF& E = ef[e0 + size_t((cell * D.n_hi + q) * 8 + c)];
                            if (E.cf >= 0)
                                gnode[size_t(gcnt[E.d]++)] = q;
                        }
                        for (size_t i0 = 0; i0 < gnode.size();)
                        {
                            const UpEntF& E0 = ef[e0 + size_t((cell * D.n_hi + gnode[i0]) * 8 + c)];
                            UpGrp G{};
                            G.cf = E0.cf;
                            G.subd = E0.sub | (E0.d << 8);
                            uint32_t qs[UP_GRP];
                            size_t i = i0;
                            for (int g = 0; g < UP_GRP; g++)
                            {
                                qs[g] = 0xffffu;
                                if (i < gnode.size() &&
                                    ef[e0 + size_t((cell * D.n_hi + gnode[i]) * 8 + c)].d == E0.d)
                                {
                                    const UpEntF& E = ef[e0 + size_t((cell * D.n_hi + gnode[i]) * 8 + c)];
                                    qs[g] = uint32_t(gnode[i]);
                                    G.x0[g] = E.xi0;
                                    G.x1[g] = E.xi1;
                                    G.x2[g] = E.xi2;
                                    G.tr[g] = E.tr;
                                    G.ti[g] = E.ti;
                                    i++;
                                }
                            }
                            G.q01 = qs[0] | (qs[1] << 16);
                            G.q23 = qs[2] | (qs[3] << 16);
                            grp.push_back(G);
                            i0 = i;
                        }
                    }
            }
            if (dkey.size() / 2 > size_t(INT32_MAX))
                throw std::runtime_error("upward tables: too many distinct child cones");
            if (grp.size() > size_t(INT32_MAX))
                throw std::runtime_error("upward tables: too many node groups");
            doff.push_back(int32_t(dkey.size() / 2));
            goff.push_back(int64_t(grp.size()));
            D.n_up_parts = int64_t(parts.size());
            D.up_part_ent = dupload(part_ent, tot);
            if (precision != PREC_FP64)
            {
                D.up_part_cell = dupload(part_cell, tot);
                D.up_doff = dupload(doff, tot);
                D.up_dkey = dupload(dkey, tot);
                std::vector<int32_t> goff32(goff.begin(), goff.end());
                D.up_goff = dupload(goff32, tot);
                D.up_grp = dupload(grp, tot);
                h_part_cell = std::move(part_cell);
                h_goff = std::move(goff);
                h_doff = std::move(doff);
                h_dkey = std::move(dkey);
            }
            if (precision == PREC_FP64)
                D.up_entries = dupload(ed, tot);
            else
                D.up_entries = dupload(ef, tot);
        }
        // work lists.  Per level: leaf cones (leaf sampling), parent cones of boxes with a table
        // grouped by (part, cell) into items of at most UP_K_CHOICES[i] cones of one cell (boxes in order
        // within a cell), and the other parent cones (per-box upward pass)
        std::vector<int32_t> leaf, up, glist, tl_glist, tl_gcell, tr_list;
        D.tr_list_off.assign(D.levels + 1, 0);
        std::vector<WorkItem> uitems[4], tl_items[4];
        D.up_off.assign(D.levels + 1, 0);
        D.up_gl_off.assign(D.levels + 1, 0);
        for (int kc = 0; kc < 4; kc++)
        {
            D.up_item_off_k[kc].assign(D.levels + 1, 0);
            D.up_tl_item_off_k[kc].assign(D.levels + 1, 0);
        }
        D.cone_lvl_off.assign(D.levels + 1, 0);
        for (int l = 0; l < D.levels; l++)
        {
            const int64_t b0 = P.S.level_off[l], b1 = P.S.level_off[l + 1];
            D.cone_lvl_off[l] = P.cone_off[b0];
            D.cone_lvl_off[l + 1] = P.cone_off[b1];
            D.up_off[l] = int64_t(up.size());
            D.up_gl_off[l] = int64_t(glist.size());
            for (int kc = 0; kc < 4; kc++)
            {
                D.up_item_off_k[kc][l] = int64_t(uitems[kc].size());
                D.up_tl_item_off_k[kc][l] = int64_t(tl_items[kc].size());
            }
            const int64_t g0 = int64_t(glist.size());
            D.tr_list_off[l] = int64_t(tr_list.size());
            for (int64_t g = b0; g < b1; g++)
                for (int64_t c = P.cone_off[g]; c < P.cone_off[g + 1]; c++)
                {
                    (P.S.leaf[g] ? leaf : (box_part[g] >= 0 ? glist : up)).push_back(int32_t(c));
                    if (!P.S.leaf[g] && box_part[g] < 0)
                        tr_list.push_back(int32_t(c));
                }
            auto part = [&](int32_t c) { return box_part[P.cone_box[c]]; };
            auto cu = [&](int32_t c) { return parts[part(c)]->cellmap[P.cone_cell[c]]; };
            // within a cell, cones of boxes with the same children next to one another: the
            // cones of an item then have their children in the same octants
            auto cmask = [&](int32_t c)
            {
                int m = 0;
                for (int o = 0; o < 8; o++)
                    m |= (P.S.child[8 * int64_t(P.cone_box[c]) + o] >= 0 ? 1 : 0) << o;
                return m;
            };
            std::stable_sort(glist.begin() + g0, glist.end(), [&](int32_t a, int32_t b)
                             {
                                 if (part(a) != part(b))
                                     return part(a) < part(b);
                                 if (cu(a) != cu(b))
                                     return cu(a) < cu(b);
                                 return cmask(a) < cmask(b);
                             });
            for (int kc = 0; kc < 4; kc++)
                for (int64_t i = g0; i < int64_t(glist.size());)
                {
                    const int32_t pa = part(glist[i]), u = cu(glist[i]);
                    if (u < 0)
                        throw std::runtime_error("upward table has no entry for an active cell");
                    int64_t j = i;
                    while (j < int64_t(glist.size()) && j - i < UP_K_CHOICES[kc] &&
                           part(glist[j]) == pa && cu(glist[j]) == u)
                        j++;
                    uitems[kc].push_back({u, int32_t(i - g0), int32_t(j - i), pa});
                    i = j;
                }
            // the grouped kernel's list: tiles of boxes x cells (measured at 1M points,
            // (6,8,7)/(3,5,4): 7.81 -> 7.4 ms for the upward phase, the same results)
            const size_t h0 = tl_glist.size();
            tl_glist.insert(tl_glist.end(), glist.begin() + g0, glist.end());
            auto tb = [&](int32_t c) { return int64_t(P.cone_box[c] - b0) / UP_TILE_BOXES; };
            auto tc = [&](int32_t c) { return cu(c) / UP_TILE_CELLS; };
            std::stable_sort(tl_glist.begin() + h0, tl_glist.end(), [&](int32_t a, int32_t b)
                             {
                                 if (part(a) != part(b))
                                     return part(a) < part(b);
                                 if (tb(a) != tb(b))
                                     return tb(a) < tb(b);
                                 if (tc(a) != tc(b))
                                     return tc(a) < tc(b);
                                 if (cu(a) != cu(b))
                                     return cu(a) < cu(b);
                                 return cmask(a) < cmask(b);
                             });
            for (size_t i = h0; i < tl_glist.size(); i++)
                tl_gcell.push_back(cu(tl_glist[i]));
            for (int kc = 0; kc < 4; kc++)
                for (size_t i = h0; i < tl_glist.size();)
                {
                    const int32_t a = tl_glist[i], pa = part(a);
                    size_t j = i;
                    while (j < tl_glist.size() && int(j - i) < UP_K_CHOICES[kc] &&
                           part(tl_glist[j]) == pa && tb(tl_glist[j]) == tb(a) &&
                           tc(tl_glist[j]) == tc(a))
                        j++;
                    tl_items[kc].push_back({cu(a), int32_t(i - h0), int32_t(j - i), pa});
                    i = j;
                }
        }
        D.up_off[D.levels] = int64_t(up.size());
        D.up_gl_off[D.levels] = int64_t(glist.size());
        for (int kc = 0; kc < 4; kc++)
        {
            D.up_item_off_k[kc][D.levels] = int64_t(uitems[kc].size());
            D.up_items_k[kc] = dupload(uitems[kc], tot);
            D.up_tl_item_off_k[kc][D.levels] = int64_t(tl_items[kc].size());
            D.up_tl_items_k[kc] = dupload(tl_items[kc], tot);
        }
        D.up_items = D.up_items_k[3]; // the default shape: 16 cones per item
        D.up_item_off = D.up_item_off_k[3];
        D.up_tl_items = D.up_tl_items_k[3];
        D.up_tl_item_off = D.up_tl_item_off_k[3];
        D.up_glist = dupload(glist, tot);
        D.up_tl_glist = dupload(tl_glist, tot);
        D.up_tl_gcell = dupload(tl_gcell, tot);
        if (precision != PREC_FP64 && order_class(D) != 0 && !tl_glist.empty())
        {
            // k_up_grp's per-cone records, from the plan's fixed tables: the children and their
            // cell-table offsets, the task ranges of each child's groups (the (cell, octant)
            // rows of up_goff), the fine slots of the listed child cones (up_doff / up_dkey
            // through the children's cell tables)
            std::vector<UpDesc> desc(tl_glist.size());
            auto fill = [&](size_t i0, size_t i1)
            {
                for (size_t i = i0; i < i1; i++)
                {
                    const int32_t g = tl_glist[i];
                    const int64_t box = P.cone_box[g];
                    const int64_t row = (h_part_cell[box_part[box]] + tl_gcell[i]) * 8;
                    UpDesc& u = desc[i];
                    u.g = g;
                    u.mask = cones[g].mask;
                    u.fine_base = cones[g].fine_base;
                    u.pad = 0;
                    int32_t tot_t = 0;
                    for (int c = 0; c < 8; c++)
                    {
                        const int32_t ch = P.S.child[8 * box + c];
                        const int64_t to = ch >= 0 ? geo[ch].table_off : -1;
                        u.toff[c] = to;
                        u.tg0[c] = int32_t(h_goff[row + c]) - tot_t;
                        if (to >= 0)
                            tot_t += int32_t(h_goff[row + c + 1] - h_goff[row + c]);
                        u.tend[c] = tot_t;
                        const int32_t nd = to >= 0 ? h_doff[row + c + 1] - h_doff[row + c] : 0;
                        for (int d = 0; d < UP_DMAX; d++)
                        {
                            int32_t sl = -1;
                            if (d < nd)
                            {
                                const int64_t kd = int64_t(h_doff[row + c]) + d;
                                const int32_t cf = h_dkey[2 * kd], sub = h_dkey[2 * kd + 1];
                                const int64_t tc = P.table[to + cf];
                                if (tc >= 0)
                                {
                                    const uint32_t m = cones[tc].mask;
                                    if ((m >> sub) & 1u)
                                        sl = cones[tc].fine_base +
                                             __builtin_popcount(m & ((1u << sub) - 1u));
                                }
                            }
                            u.slot[c][d] = sl;
                        }
                    }
                }
            };
            const size_t n = desc.size();
            const size_t nth = std::min<size_t>(16, std::max<size_t>(1, n / 4096));
            std::vector<std::thread> pool;
            for (size_t t = 1; t < nth; t++)
                pool.emplace_back(fill, n * t / nth, n * (t + 1) / nth);
            fill(0, n / nth);
            for (auto& th : pool)
                th.join();
            D.up_tl_desc = dupload(desc, tot);
        }
        D.tr_list_off[D.levels] = int64_t(tr_list.size());
        D.tr_list = dupload(tr_list, tot);
        D.leaf_cones = dupload(leaf, tot);
        D.n_leaf_cones = int64_t(leaf.size());
        {
            // leaf boxes with active cones: (box, first cone, number of cones)
            std::vector<WorkItem> lb;
            for (int64_t g = 0; g < P.S.num_boxes(); g++)
                if (P.S.leaf[g] && P.cone_off[g + 1] > P.cone_off[g])
                    lb.push_back({int32_t(g), int32_t(P.cone_off[g]),
                                  int32_t(P.cone_off[g + 1] - P.cone_off[g]), 0});
            D.leaf_boxes = dupload(lb, tot);
            D.n_leaf_boxes = int64_t(lb.size());
            // per level (boxes are level-major): the range of the leaf boxes, of the fine slots
            D.leaf_lvl_off.assign(D.levels + 1, 0);
            for (const WorkItem& w : lb)
                for (int l = 0; l < D.levels; l++)
                    if (w.box >= P.S.level_off[l] && w.box < P.S.level_off[l + 1])
                        D.leaf_lvl_off[l + 1]++;
            for (int l = 0; l < D.levels; l++)
                D.leaf_lvl_off[l + 1] += D.leaf_lvl_off[l];
            D.fine_lvl_off.assign(D.levels + 1, 0);
            for (int l = 0; l <= D.levels; l++)
                D.fine_lvl_off[l] = D.cone_lvl_off[l] < D.n_cones
                                        ? int64_t(P.cone_fine[D.cone_lvl_off[l]])
                                        : D.n_fine;
        }
        D.up_cones = dupload(up, tot);
        // far / near lists grouped by target box (pairs are sorted by target)
        const int64_t nt = P.T.num_boxes();
        auto csr = [&](const std::vector<Pair>& pairs, std::vector<int32_t>& off,
                       std::vector<int32_t>& src, std::vector<int32_t>& tboxes)
        {
            off.assign(nt + 1, 0);
            src.resize(pairs.size());
            for (size_t i = 0; i < pairs.size(); i++)
            {
                off[pairs[i].tgt + 1]++;
                src[i] = pairs[i].src;
            }
            for (int64_t t = 0; t < nt; t++)
            {
                if (off[t + 1] > 0)
                    tboxes.push_back(int32_t(t));
                off[t + 1] += off[t];
            }
        };
        std::vector<int32_t> foff, fsrc, ftb, noff, nsrcv, ntb;
        csr(P.far, foff, fsrc, ftb);
        csr(P.near, noff, nsrcv, ntb);
        D.far_src = dupload(fsrc, tot);
        D.near_off = dupload(noff, tot);
        D.near_src = dupload(nsrcv, tot);
        // far work: items of at most 64 targets of one target box against its far sources of
        // one source level (contiguous in its list: box ids are level-major), grouped by (source
        // level, target level), deepest source level first, then deepest target level
        std::vector<WorkItem> fitems;
        std::vector<int32_t> frange;
        D.far_groups.clear();
        std::vector<std::vector<int32_t>> tb_of_level(P.T.levels);
        for (int32_t t : ftb)
            tb_of_level[P.T.level_of(t)].push_back(t);
        for (int sl = P.S.levels - 1; sl >= 0; sl--)
            for (int tl = P.T.levels - 1; tl >= 0; tl--)
            {
                const int64_t i0 = int64_t(fitems.size());
                int64_t lo = INT64_MAX, hi = -1;
                for (int32_t t : tb_of_level[tl])
                {
                    // the sources ascend by box id, so by level: the level-sl run
                    int32_t j0 = foff[t];
                    const int32_t j1 = foff[t + 1];
                    while (j0 < j1 && P.S.level_of(fsrc[j0]) < sl)
                        j0++;
                    int32_t j = j0;
                    while (j < j1 && P.S.level_of(fsrc[j]) == sl)
                        j++;
                    if (j == j0)
                        continue;
                    for (int32_t m = j0; m < j; m++)
                    {
                        const BoxGeo& B = geo[fsrc[m]];
                        lo = std::min<int64_t>(lo, 2 * int64_t(B.fth_off));
                        hi = std::max<int64_t>(hi, 2 * int64_t(B.fph_off) + 2 * (2 * B.nf[2] + 1));
                    }
                    const int64_t s0 = P.T.begin[t], e0 = P.T.end[t];
                    for (int64_t f = s0; f < e0; f += 64)
                    {
                        fitems.push_back({t, int32_t(f), int32_t(std::min<int64_t>(64, e0 - f)), 0});
                        frange.push_back(j0);
                        frange.push_back(j);
                    }
                }
                if (int64_t(fitems.size()) == i0)
                    continue;
                // the angle tables of the group's grid classes, staged in LDS by the mixed far
                // kernel when they fit 48 KB
                const bool fits = hi > lo && (hi - lo) * int64_t(sizeof(double)) <= 48 * 1024;
                D.far_groups.push_back({i0, int64_t(fitems.size()), fits ? int(lo) : 0,
                                        fits ? int(hi - lo) : 0, sl, tl});
            }
        D.far_srange = dupload(frange, tot);
        D.far_items = dupload(fitems, tot);
        auto nitems = make_items(ntb, P.T);
        D.near_items = dupload(nitems, tot);
        D.n_near_items = int64_t(nitems.size());
        // work buffers
        const size_t cs = precision == PREC_FP64 ? 16 : 8;
        D.wsorted = dalloc<unsigned char>(cs * D.nsrc, tot);
        D.result = dalloc<unsigned char>(cs * D.ntgt, tot);
        const size_t full = cs * (size_t(std::max<int64_t>(1, D.n_cones)) * D.n_hi +
                                  size_t(std::max<int64_t>(1, D.n_fine)) * D.n_lo);
        size_t v_lvl = 1, f_lvl[2] = {1, 1};
        for (int l = 0; l < D.levels; l++)
        {
            v_lvl = std::max(v_lvl, size_t(D.cone_lvl_off[l + 1] - D.cone_lvl_off[l]) * D.n_hi);
            f_lvl[l & 1] =
                std::max(f_lvl[l & 1], size_t(D.fine_lvl_off[l + 1] - D.fine_lvl_off[l]) * D.n_lo);
        }
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        const size_t margin = size_t(4) << 30; // batch buffers, torch, the caller's vectors
        D.ring = ring == 1 || (full + margin > free_b && D.levels > 2);
        if (D.ring)
        {
            if (cs * (v_lvl + f_lvl[0] + f_lvl[1]) + margin > free_b)
                throw std::runtime_error(
                    "the values of one level and the fine coefficients of two levels (" +
                    std::to_string(cs * (v_lvl + f_lvl[0] + f_lvl[1]) >> 20) +
                    " MiB) do not fit the free device memory (" + std::to_string(free_b >> 20) +
                    " MiB)");
            D.V = dalloc<unsigned char>(cs * v_lvl, tot);
            D.F_bank[0] = dalloc<unsigned char>(cs * f_lvl[0], tot);
            D.F_bank[1] = dalloc<unsigned char>(cs * f_lvl[1], tot);
        }
        else
        {
            D.V = dalloc<unsigned char>(cs * size_t(std::max<int64_t>(1, D.n_cones)) * D.n_hi, tot);
            D.F = dalloc<unsigned char>(cs * size_t(std::max<int64_t>(1, D.n_fine)) * D.n_lo, tot);
        }
        HIP_OK(hipDeviceSynchronize());
    }
    catch (const std::exception& e)
    {
        device_free(D);
        return e.what();
    }
    return "";
}

// batch buffers for rb columns (all-or-nothing; false when the device memory is short)
static bool ensure_batch(DevicePlan& D, int rb, bool host_io)
{
    if (D.b_rb == rb && (!host_io || D.b_in))
        return true;
    // a captured batched graph holds the old buffers
    if (D.b_graph_exec)
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(D.b_graph_exec));
    if (D.b_graph)
        (void)hipGraphDestroy(static_cast<hipGraph_t>(D.b_graph));
    D.b_graph_exec = D.b_graph = nullptr;
    for (void* p : {D.b_ws, D.b_res, D.b_V, D.b_F, static_cast<void*>(D.b_in),
                    static_cast<void*>(D.b_out)})
        if (p)
            (void)hipFree(p);
    D.b_ws = D.b_res = D.b_V = D.b_F = nullptr;
    D.b_in = D.b_out = nullptr;
    D.b_rb = 0;
    D.b_bytes = 0;
    const size_t cs = D.prec == PREC_FP64 ? 16 : 8;
    const size_t n_ws = size_t(D.nsrc) * rb, n_res = size_t(D.ntgt) * rb;
    const size_t n_V = size_t(std::max<int64_t>(1, D.n_cones)) * D.n_hi * rb;
    const size_t n_F = size_t(std::max<int64_t>(1, D.n_fine)) * D.n_lo * rb;
    size_t need = cs * (n_ws + n_res + n_V + n_F);
    if (host_io)
        need += 16 * size_t(rb) * size_t(D.nsrc + D.ntgt);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need + (size_t(1) << 30) > free_b)
        return false;
    HIP_OK(hipMalloc(&D.b_ws, cs * n_ws));
    HIP_OK(hipMalloc(&D.b_res, cs * n_res));
    HIP_OK(hipMalloc(&D.b_V, cs * n_V));
    HIP_OK(hipMalloc(&D.b_F, cs * n_F));
    if (host_io)
    {
        HIP_OK(hipMalloc(&D.b_in, 16 * size_t(rb) * size_t(D.nsrc)));
        HIP_OK(hipMalloc(&D.b_out, 16 * size_t(rb) * size_t(D.ntgt)));
    }
    D.b_rb = rb;
    D.b_bytes = int64_t(need);
    return true;
}

std::string device_mult(DevicePlan& D, const void* w, void* out, int64_t nrhs, bool device_ptrs,
                        void* stream, int io, bool graph)
{
    try
    {
        hipStream_t st = static_cast<hipStream_t>(stream);
        const bool timed = D.force_timing || std::getenv("IFGF_PHASE_TIMING") != nullptr;
        if (D.shape.overlap && !D.side_stream)
        {
            // the side stream of the overlapped near and far fields and its events, created
            // outside any capture; default priority (as every stream here)
            hipStream_t s2;
            hipEvent_t e1, e2;
            HIP_OK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            HIP_OK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
            HIP_OK(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
            D.side_stream = s2;
            D.ev_fork = e1;
            D.ev_near = e2;
            D.ev_level.assign(std::max(1, D.levels), nullptr);
            for (void*& e : D.ev_level)
            {
                hipEvent_t ev;
                HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                e = ev;
            }
            D.ev_far_done.assign(std::max(1, D.levels), nullptr);
            for (void*& e : D.ev_far_done)
            {
                hipEvent_t ev;
                HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                e = ev;
            }
        }
        // several columns: the batched pipeline (8 or 32 columns per pass) when the orders
        // have compiled kernels and the batch buffers fit
        if (nrhs >= 2 && batched_supported(D) && !D.ring && std::getenv("IFGF_NO_BATCH") == nullptr)
        {
            int rb = D.shape.rb > 0 ? D.shape.rb : (nrhs <= 8 ? 8 : 32);
            bool ok = ensure_batch(D, rb, !device_ptrs);
            if (!ok && rb == 32)
                ok = ensure_batch(D, rb = 8, !device_ptrs);
            if (ok)
            {
                for (int64_t c = 0; c < nrhs; c += rb)
                {
                    const int ncol = int(std::min<int64_t>(rb, nrhs - c));
                    const double* wc = static_cast<const double*>(w) + 2 * c * D.nsrc;
                    double* oc = static_cast<double*>(out) + 2 * c * D.ntgt;
                    if (device_ptrs)
                    {
                        mult_batched(D, wc, oc, rb, ncol, st, io, graph && !timed);
                        continue;
                    }
                    HIP_OK(hipMemcpyAsync(D.b_in, wc, 16 * size_t(ncol) * size_t(D.nsrc),
                                          hipMemcpyHostToDevice, st));
                    mult_batched(D, D.b_in, D.b_out, rb, ncol, st, io, graph && !timed);
                    HIP_OK(hipMemcpyAsync(oc, D.b_out, 16 * size_t(ncol) * size_t(D.ntgt),
                                          hipMemcpyDeviceToHost, st));
                }
                if (!device_ptrs)
                    HIP_OK(hipStreamSynchronize(st));
                return "";
            }
        }
        // host densities and results go through two device vectors kept with the plan
        // (allocated on the first host-pointer mult, not per call)
        if (!device_ptrs && !D.h_in)
        {
            HIP_OK(hipMalloc(&D.h_in, 16 * size_t(D.nsrc)));
            HIP_OK(hipMalloc(&D.h_out, 16 * size_t(D.ntgt)));
        }
        double* wbuf = D.h_in;
        double* obuf = D.h_out;
        for (int64_t r = 0; r < nrhs; r++)
        {
            const double* wr = static_cast<const double*>(w) + 2 * r * D.nsrc;
            double* orr = static_cast<double*>(out) + 2 * r * D.ntgt;
            const double* wdev = wr;
            double* odev = orr;
            if (!device_ptrs)
            {
                HIP_OK(hipMemcpyAsync(wbuf, wr, 16 * size_t(D.nsrc), hipMemcpyHostToDevice, st));
                wdev = wbuf;
                odev = obuf;
            }
            if (D.prec == PREC_FP64)
                dispatch<double>(D, wdev, odev, st, timed, io, graph, plan_stage(D));
            else
                dispatch<float>(D, wdev, odev, st, timed, io, graph, plan_stage(D));
            if (!device_ptrs)
                HIP_OK(hipMemcpyAsync(orr, obuf, 16 * size_t(D.ntgt), hipMemcpyDeviceToHost, st));
        }
        if (!device_ptrs)
            HIP_OK(hipStreamSynchronize(st));
    }
    catch (const std::exception& e)
    {
        return e.what();
    }
    return "";
}

std::string device_mult_noalloc(DevicePlan& D, const double* w_dev, double* out_dev,
                                void* stream)
{
    try
    {
        int dev = -1;
        HIP_OK(hipGetDevice(&dev));
        if (dev != D.device)
            return "allocation-free mult: the current HIP device is not the operator's (init "
                   "ran on device " + std::to_string(D.device) + ")";
        const size_t cs = D.prec == PREC_FP64 ? 16 : 8;
        const size_t ws_bytes = (cs * size_t(D.nsrc) + 255) / 256 * 256;
        const size_t need = ws_bytes + cs * size_t(D.ntgt);
        if (need > NOALLOC_SCRATCH_BYTES)
            return "allocation-free mult: " + std::to_string(D.nsrc) + " sources and " +
                   std::to_string(D.ntgt) + " targets need " + std::to_string(need) +
                   " bytes of scratch, more than its " +
                   std::to_string(NOALLOC_SCRATCH_BYTES) + " (n_src + n_tgt <= " +
                   std::to_string(NOALLOC_SCRATCH_BYTES / cs - 32) + " in this precision)";
        std::string err;
        unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(err));
        if (!base)
            return err;
        // the results are zeroed by the mult's first kernel and the densities fully written by
        // the gather, so what an earlier call left in the scratch is never read
        const Stage sg{base, base + ws_bytes, false};
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (D.prec == PREC_FP64)
            dispatch<double>(D, w_dev, out_dev, st, false, 0, false, sg);
        else
            dispatch<float>(D, w_dev, out_dev, st, false, 0, false, sg);
    }
    catch (const std::exception& e)
    {
        return e.what();
    }
    return "";
}

std::string device_set_shape(DevicePlan& D, const LaunchShape& sh)
{
    if (sh.xcd_log2 < 0 || sh.xcd_log2 > 8)
        return "launch shape: XCD run length must be 2^0 .. 2^8";
    int kc = -1;
    for (int i = 0; i < 4; i++)
        if (UP_K_CHOICES[i] == sh.up_k)
            kc = i;
    if (kc < 0)
        return "launch shape: upward cones per item must be 2, 4, 8 or 16";
    if (sh.rb != 0 && sh.rb != 8 && sh.rb != 32)
        return "launch shape: batched columns per pass must be 8 or 32";
    if (sh.far_waves != 1 && sh.far_waves != 2 && sh.far_waves != 4 && sh.far_waves != 8)
        return "launch shape: far-field waves per workgroup must be 1, 2, 4 or 8";
    if (sh.overlap < 0 || sh.overlap > 2)
        return "launch shape: overlap is 0, 1 or 2";
    if (sh.cpb != 0)
    {
        const auto ch = transform_cpb_choices(order_class(D), D.prec == PREC_FP64 ? 8 : 4);
        if (std::find(ch.begin(), ch.end(), sh.cpb) == ch.end())
            return "launch shape: no transform kernel with that many cones per workgroup for "
                   "these orders and this precision";
    }
    const bool changed = sh.xcd_log2 != D.shape.xcd_log2 || sh.cpb != D.shape.cpb ||
                         sh.up_k != D.shape.up_k || sh.far_waves != D.shape.far_waves ||
                         sh.overlap != D.shape.overlap;
    D.shape = sh;
    D.up_items = D.up_items_k[kc];
    D.up_item_off = D.up_item_off_k[kc];
    D.up_tl_items = D.up_tl_items_k[kc];
    D.up_tl_item_off = D.up_tl_item_off_k[kc];
    if (changed)
        device_drop_graph(D); // the captured kernels have the old shapes
    return "";
}

void device_drop_graph(DevicePlan& D)
{
    // a captured graph holds the kernel arguments of its capture (launch shapes, the gradient
    // direction): it is captured again on the next graph mult
    if (D.graph_exec)
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(D.graph_exec));
    D.graph_exec = nullptr;
    if (D.graph)
        (void)hipGraphDestroy(static_cast<hipGraph_t>(D.graph));
    D.graph = nullptr;
    if (D.b_graph_exec)
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(D.b_graph_exec));
    D.b_graph_exec = nullptr;
    if (D.b_graph)
        (void)hipGraphDestroy(static_cast<hipGraph_t>(D.b_graph));
    D.b_graph = nullptr;
}

std::string device_autotune(DevicePlan& D, int64_t nrhs, int reps, LaunchShape& best)
{
    double* w = nullptr;
    double* o = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string err;
    try
    {
        nrhs = std::max<int64_t>(1, nrhs);
        reps = std::max(1, reps);
        // synthetic densities: a fixed pseudo-random pattern (the timings do not depend on it)
        std::vector<double> hw(2 * size_t(D.nsrc) * size_t(nrhs));
        uint64_t x = 0x9e3779b97f4a7c15ull;
        for (double& v : hw)
        {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            v = double(x >> 11) * 0x1.0p-53 - 0.5;
        }
        HIP_OK(hipMalloc(&w, hw.size() * sizeof(double)));
        HIP_OK(hipMalloc(&o, 2 * size_t(D.ntgt) * size_t(nrhs) * sizeof(double)));
        HIP_OK(hipMemcpy(w, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        const bool batched = nrhs >= 2 && batched_supported(D);
        // fastest of reps mults with the shape s (one untimed mult first): the phase's time
        // from the phase events, or the whole mult's (phase < 0)
        auto measure = [&](const LaunchShape& s, int phase) -> double
        {
            std::string e = device_set_shape(D, s);
            if (!e.empty())
                throw std::runtime_error(e);
            double t_best = 1e300;
            for (int r = -1; r < reps; r++)
            {
                D.force_timing = phase >= 0;
                HIP_OK(hipEventRecord(e0, st));
                e = device_mult(D, w, o, nrhs, true, st);
                HIP_OK(hipEventRecord(e1, st));
                HIP_OK(hipEventSynchronize(e1));
                D.force_timing = false;
                if (!e.empty())
                    throw std::runtime_error(e);
                float ms = 0;
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                const double t = phase >= 0 ? double(D.last_ms[phase]) : double(ms);
                if (r >= 0)
                    t_best = std::min(t_best, t);
            }
            return t_best;
        };
        LaunchShape cur = D.shape;
        auto pick = [&](int LaunchShape::*field, const std::vector<int>& choices, int phase)
        {
            int arg = cur.*field;
            double tb = 1e300;
            for (int c : choices)
            {
                LaunchShape s = cur;
                s.*field = c;
                const double t = measure(s, phase);
                if (t < tb)
                {
                    tb = t;
                    arg = c;
                }
            }
            cur.*field = arg;
        };
        pick(&LaunchShape::xcd_log2, {0, 2, 4, 6}, -1);
        if (!batched)
        {
            const auto ch = transform_cpb_choices(order_class(D), D.prec == PREC_FP64 ? 8 : 4);
            if (ch.size() > 1)
                pick(&LaunchShape::cpb, ch, 3);
        }
        pick(&LaunchShape::up_k, {2, 4, 8, 16}, batched ? -1 : 2);
        if (!batched && D.prec == PREC_MIXED)
            pick(&LaunchShape::far_waves, {1, 2, 4, 8}, 4);
        pick(&LaunchShape::overlap, {0, 1, 2}, -1);
        if (batched && nrhs > 8)
            pick(&LaunchShape::rb, {8, 32}, -1);
        err = device_set_shape(D, cur);
        best = cur;
    }
    catch (const std::exception& ex)
    {
        err = std::string("autotune: ") + ex.what();
        D.force_timing = false;
    }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (st)
        (void)hipStreamDestroy(st);
    if (w)
        (void)hipFree(w);
    if (o)
        (void)hipFree(o);
    return err;
}

void device_last_timings(const DevicePlan& D, double* out6)
{
    for (int i = 0; i < 6; i++)
        out6[i] = D.last_ms[i];
}

} // namespace ifgf
