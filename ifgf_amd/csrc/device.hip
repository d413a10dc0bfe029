// device.hip -- HIP kernels of IFGF operator application on gfx950 and their orchestration.
//
// One mult (ifgfoperator.hpp:246-867) runs as a level-synchronous pipeline on one stream:
//   gather weights -> near field (P2P) -> leaf sampling (all leaf cones) ->
//   for each source level, deepest first: upward transfer (non-leaf cones) and the fused
//   high-order-values -> fine-coefficients transform -> far-field evaluation per target
//   level -> scatter.
// Every kernel is templated on the arithmetic type R (double: the plain form; float: the
// mixed form, which keeps cone coordinates, phases and box geometry in FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "device.hpp"

namespace ifgf
{
namespace
{

#define HIP_OK(x)                                                                          \
    do                                                                                     \
    {                                                                                      \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

template <typename R> struct Cx
{
    R x, y;
};

template <typename R> __device__ __forceinline__ Cx<R> cadd(Cx<R> a, Cx<R> b)
{
    return {a.x + b.x, a.y + b.y};
}
template <typename R> __device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R> __device__ __forceinline__ Cx<R> cscale(Cx<R> a, R s)
{
    return {a.x * s, a.y * s};
}

struct KParams
{
    double k_re, k_im, kappa;
    int dx;
    int ho0, ho1, ho2;
    int lo0, lo1, lo2;
    int f0, f1, f2;
    int n_hi, n_lo;
};

constexpr double PI_D = 3.14159265358979323846;
constexpr double TWO_PI_D = 6.28318530717958647692;
constexpr double INV_4PI_D = 0.07957747154594766788;

enum
{
    KH = 0, // Helmholtz (real kappa, k = -i kappa)
    KM = 1, // modified Helmholtz (complex k)
    KL = 2, // Laplace
    KD = 3, // double layer
    KC = 4, // combined field
    KG = 5  // gradient
};

// e^{-k delta} for the arithmetic type R; delta in FP64 is reduced before the FP32 sincos
template <typename R> __device__ __forceinline__ Cx<R> expmk(double kr, double ki, R delta)
{
    R ph = R(-ki) * delta;
    R s, c;
    if constexpr (sizeof(R) == 4)
        sincosf(ph, &s, &c);
    else
        sincos(ph, &s, &c);
    if (kr != 0.0)
    {
        R m = exp(R(-kr) * delta);
        return {m * c, m * s};
    }
    return {c, s};
}

// e^{-k r} with a phase that can be large (hundreds of radians): reduce in FP64
template <typename R> __device__ __forceinline__ Cx<R> expmk_far(double kr, double ki, double r)
{
    double ph = -ki * r;
    R s, c;
    if constexpr (sizeof(R) == 4)
    {
        double n = rint(ph * (1.0 / TWO_PI_D));
        float red = float(fma(-n, TWO_PI_D, ph));
        sincosf(red, &s, &c);
    }
    else
        sincos(ph, &s, &c);
    if (kr != 0.0)
    {
        R m = R(exp(-kr * r));
        return {m * c, m * s};
    }
    return {c, s};
}

// ------------------------------------------------------------------------------------
// cone location: physical offset v (target - box centre, FP64) -> fine cone slot and local
// coordinates in [-1,1]^3 (util.hpp:156-171 cartToInterp2, cone_domain.hpp:173-184,232-256).
// The steps are IEEE FP64 without contraction so the cells agree with the host plan.
// ------------------------------------------------------------------------------------
struct Loc
{
    int fine;
    double r;
    double xi[3];
};

__device__ __forceinline__ bool locate(const BoxGeo& B, const int32_t* __restrict__ table,
                                       const ConeDev* __restrict__ cones, double v0, double v1,
                                       double v2, int f0, int f1, int f2, Loc& out)
{
#pragma clang fp contract(off)
    const double a = v0 * v0 + v1 * v1;
    double p[3];
    p[2] = atan2(v1, v0);
    p[1] = atan2(sqrt(a), v2);
    const double r = sqrt(a + v2 * v2);
    p[0] = B.H / r;
    out.r = r;
    double d2 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
        if (B.lo[k] > p[k])
        {
            double t = B.lo[k] - p[k];
            d2 += t * t;
        }
        else if (p[k] > B.hi[k])
        {
            double t = p[k] - B.hi[k];
            d2 += t * t;
        }
    }
    if (d2 > 0)
        return false;
    int q[3];
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
        int t = int(floor((p[j] - B.lo[j]) / B.hf[j]));
        q[j] = min(max(t, 0), B.nf[j] - 1);
    }
    const int c0 = q[0] / f0, c1 = q[1] / f1, c2 = q[2] / f2;
    const int cell = c0 + B.nc[0] * (c1 + B.nc[1] * c2);
    const int g = table[B.table_off + cell];
    if (g < 0)
        return false;
    const ConeDev cn = cones[g];
    const int sub = (q[0] - c0 * f0) + f0 * ((q[1] - c1 * f1) + f1 * (q[2] - c2 * f2));
    if (!((cn.mask >> sub) & 1u))
        return false;
    out.fine = cn.fine_base + __popc(cn.mask & ((1u << sub) - 1u));
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
        double aa = 0.5 * B.hf[j];
        double bb = B.lo[j] + (double(q[j]) + 0.5) * B.hf[j];
        out.xi[j] = (p[j] - bb) / aa;
    }
    return true;
}

// ------------------------------------------------------------------------------------
// tensor Clenshaw over a coefficient block c[i0 + o0 (i1 + o1 i2)] (chebinterp.hpp:252-346)
// ------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ Cx<R> clen_line(const Cx<R>* __restrict__ c, int n, R x)
{
    if (n == 1)
        return c[0];
    if (n == 2)
        return {c[0].x + x * c[1].x, c[0].y + x * c[1].y};
    const R x2 = x + x;
    Cx<R> b2 = c[n - 1];
    Cx<R> b1 = {x2 * b2.x + c[n - 2].x, x2 * b2.y + c[n - 2].y};
    for (int j = n - 3; j > 0; j--)
    {
        Cx<R> t = {c[j].x + x2 * b1.x - b2.x, c[j].y + x2 * b1.y - b2.y};
        b2 = b1;
        b1 = t;
    }
    return {x * b1.x - b2.x + c[0].x, x * b1.y - b2.y + c[0].y};
}

template <typename R>
__device__ __forceinline__ Cx<R> clen_plane(const Cx<R>* __restrict__ c, int o0, int o1, R x0,
                                            R x1)
{
    if (o1 == 1)
        return clen_line(c, o0, x0);
    if (o1 == 2)
    {
        Cx<R> a = clen_line(c, o0, x0), b = clen_line(c + o0, o0, x0);
        return {a.x + x1 * b.x, a.y + x1 * b.y};
    }
    const R x2 = x1 + x1;
    Cx<R> b2 = clen_line(c + o0 * (o1 - 1), o0, x0);
    Cx<R> t0 = clen_line(c + o0 * (o1 - 2), o0, x0);
    Cx<R> b1 = {x2 * b2.x + t0.x, x2 * b2.y + t0.y};
    for (int j = o1 - 3; j > 0; j--)
    {
        Cx<R> cj = clen_line(c + o0 * j, o0, x0);
        Cx<R> t = {cj.x + x2 * b1.x - b2.x, cj.y + x2 * b1.y - b2.y};
        b2 = b1;
        b1 = t;
    }
    Cx<R> c0 = clen_line(c, o0, x0);
    return {x1 * b1.x - b2.x + c0.x, x1 * b1.y - b2.y + c0.y};
}

template <typename R>
__device__ __forceinline__ Cx<R> clen3(const Cx<R>* __restrict__ c, int o0, int o1, int o2, R x0,
                                       R x1, R x2v)
{
    const int s = o0 * o1;
    if (o2 == 1)
        return clen_plane(c, o0, o1, x0, x1);
    if (o2 == 2)
    {
        Cx<R> a = clen_plane(c, o0, o1, x0, x1), b = clen_plane(c + s, o0, o1, x0, x1);
        return {a.x + x2v * b.x, a.y + x2v * b.y};
    }
    const R x2 = x2v + x2v;
    Cx<R> b2 = clen_plane(c + s * (o2 - 1), o0, o1, x0, x1);
    Cx<R> t0 = clen_plane(c + s * (o2 - 2), o0, o1, x0, x1);
    Cx<R> b1 = {x2 * b2.x + t0.x, x2 * b2.y + t0.y};
    for (int j = o2 - 3; j > 0; j--)
    {
        Cx<R> cj = clen_plane(c + s * j, o0, o1, x0, x1);
        Cx<R> t = {cj.x + x2 * b1.x - b2.x, cj.y + x2 * b1.y - b2.y};
        b2 = b1;
        b1 = t;
    }
    Cx<R> c0 = clen_plane(c, o0, o1, x0, x1);
    return {x2v * b1.x - b2.x + c0.x, x2v * b1.y - b2.y + c0.y};
}

// ------------------------------------------------------------------------------------
// kernel functions.  P2P: G(x_src - y) (helmholtz_ifgf.hpp:37-41 and siblings)
// ------------------------------------------------------------------------------------
template <int KIND, typename R>
__device__ __forceinline__ Cx<R> p2p(const KParams& kp, R x0, R x1, R x2, const double* nrm,
                                     Cx<R> w)
{
    // (x0,x1,x2) = source - target
    const R d2 = x0 * x0 + x1 * x1 + x2 * x2;
    const R d = sqrt(d2);
    const R inv4pi = R(INV_4PI_D);
    if constexpr (KIND == KH || KIND == KM || (KIND == KG))
    {
        if (KIND == KG && kp.dx >= 0)
        {
            if (d < R(1e-12))
                return {0, 0};
            const R xd = kp.dx == 0 ? x0 : (kp.dx == 1 ? x1 : x2);
            // -(1/4pi) (1/d^2) e^{-kd} (-k - 1/d) x_dx
            Cx<R> e = expmk<R>(kp.k_re, kp.k_im, d);
            Cx<R> mk = {R(-kp.k_re) - R(1) / d, R(-kp.k_im)};
            Cx<R> v = cmul(e, mk);
            R s = -inv4pi / d2 * xd;
            return cmul(w, cscale(v, s));
        }
        if (KIND == KG ? (d < R(1e-12)) : (d == R(0)))
            return {0, 0};
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, d);
        return cmul(w, cscale(e, inv4pi / d));
    }
    else if constexpr (KIND == KL)
    {
        if (d == R(0))
            return {0, 0};
        return cscale(w, inv4pi / d);
    }
    else if constexpr (KIND == KD)
    {
        if (d < R(1e-12))
            return {0, 0};
        const R xn = x0 * R(nrm[0]) + x1 * R(nrm[1]) + x2 * R(nrm[2]);
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, d);
        Cx<R> mk = {R(-kp.k_re) - R(1) / d, R(-kp.k_im)};
        Cx<R> v = cmul(e, mk);
        return cmul(w, cscale(v, -inv4pi / d2 * xn));
    }
    else // KC: combined_field_helmholtz_ifgf.hpp:43-68
    {
        const R invd = d2 < R(1e-14) ? R(0) : R(1) / d;
        const R dd = d2 * invd;
        const R nxy = -(R(nrm[0]) * x0 + R(nrm[1]) * x1 + R(nrm[2]) * x2);
        const R kap = R(kp.kappa);
        R s, c;
        if constexpr (sizeof(R) == 4)
            sincosf(kap * dd, &s, &c);
        else
            sincos(kap * dd, &s, &c);
        const R f = inv4pi * invd * invd * invd;
        const R re = f * (c * nxy + s * kap * (dd * nxy + d2));
        const R im = f * (-c * kap * (dd * nxy + d2) + s * nxy);
        return cmul(w, Cx<R>{re, im});
    }
}

// factored kernel at a node: y' = node - centre, x' = source - centre, dc = |y'|
// (helmholtz_ifgf.hpp:114-156 and siblings).  d - dc is formed as (|x'|^2 - 2 x'.y')/(d + dc)
// in the mixed form, which has no cancellation between the two large distances.
template <int KIND, typename R>
__device__ __forceinline__ Cx<R> factored(const KParams& kp, R y0, R y1, R y2, R dc, R x0, R x1,
                                          R x2, const R* nrm, Cx<R> w)
{
    const R z0 = x0 - y0, z1 = x1 - y1, z2 = x2 - y2; // source - node
    const R d2 = z0 * z0 + z1 * z1 + z2 * z2;
    const R d = sqrt(d2);
    R delta;
    if constexpr (sizeof(R) == 4)
        delta = (x0 * (x0 - R(2) * y0) + x1 * (x1 - R(2) * y1) + x2 * (x2 - R(2) * y2)) / (d + dc);
    else
        delta = d - dc;
    if constexpr (KIND == KH || KIND == KM)
    {
        if (d == R(0))
            return {0, 0};
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, delta);
        return cmul(w, cscale(e, dc / d));
    }
    else if constexpr (KIND == KL)
    {
        if (d == R(0))
            return {0, 0};
        return cscale(w, dc / d);
    }
    else if constexpr (KIND == KG)
    {
        const R id = d2 > R(1e-12) ? R(1) / d : R(0);
        if (kp.dx < 0)
        {
            Cx<R> e = expmk<R>(kp.k_re, kp.k_im, delta);
            return cmul(w, cscale(e, dc * id));
        }
        const R zd = kp.dx == 0 ? z0 : (kp.dx == 1 ? z1 : z2);
        const R ww = -(R(1) / d2) * dc * zd;
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, delta);
        Cx<R> mk = {R(-kp.k_re) - id, R(-kp.k_im)};
        return cmul(w, cscale(cmul(e, mk), ww));
    }
    else if constexpr (KIND == KD)
    {
        const R id = d2 > R(1e-12) ? R(1) / d : R(0);
        const R ww = -(R(1) / d2) * dc * (z0 * nrm[0] + z1 * nrm[1] + z2 * nrm[2]);
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, delta);
        Cx<R> mk = {R(-kp.k_re) - id, R(-kp.k_im)};
        return cmul(w, cscale(cmul(e, mk), ww));
    }
    else // KC: z = node - source
    {
        if (!(d2 > R(1e-14)))
            return {0, 0};
        const R nxy = -(z0 * nrm[0] + z1 * nrm[1] + z2 * nrm[2]);
        const R kap = R(kp.kappa);
        R s, c;
        if constexpr (sizeof(R) == 4)
            sincosf(kap * delta, &s, &c);
        else
            sincos(kap * delta, &s, &c);
        const R f = dc / (d2 * d);
        const R a = w.x * c - w.y * s, b = w.x * s + w.y * c;
        return {f * (a * nxy + b * kap * (d * nxy + d2)), f * (-a * kap * (d * nxy + d2) + b * nxy)};
    }
}

// centre factor CF(y - xc) at distance r (helmholtz_ifgf.hpp:43-65 single-point branch)
template <int KIND, typename R> __device__ __forceinline__ Cx<R> centre_factor(const KParams& kp, double r)
{
    const R a = R(INV_4PI_D / r);
    if constexpr (KIND == KL)
        return {a, 0};
    else
    {
        Cx<R> e = expmk_far<R>(kp.k_re, kp.k_im, r);
        return cscale(e, a);
    }
}

// ------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------
template <typename R>
__global__ void k_gather(const double* __restrict__ w_in, const int64_t* __restrict__ perm,
                         Cx<R>* __restrict__ w_out, int64_t n)
{
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t j = perm[i];
    w_out[i] = {R(w_in[2 * j]), R(w_in[2 * j + 1])};
}

template <typename R>
__global__ void k_scatter(const Cx<R>* __restrict__ res, const int64_t* __restrict__ perm,
                          double* __restrict__ out, int64_t n)
{
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t j = perm[i];
    out[2 * j] = double(res[i].x);
    out[2 * j + 1] = double(res[i].y);
}

template <typename R> __global__ void k_zero(Cx<R>* p, int64_t n)
{
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = {R(0), R(0)};
}

// lane layout of a 64-wide work item: Tp lanes per group (power of two >= targets), 64/Tp
// groups splitting the interaction list; reduction over groups by xor shuffles.
__device__ __forceinline__ int pow2_at_least(int n)
{
    int p = 1;
    while (p < n)
        p <<= 1;
    return p;
}

template <typename R> __device__ __forceinline__ R wave_xor(R v, int m)
{
    return __shfl_xor(v, m, 64);
}

// near field: for each target of the item, sum over the points of its near source leaves
// (ifgfoperator.hpp:900-960; octree.hpp:302-307)
template <int KIND, typename R>
__global__ __launch_bounds__(64) void k_near(const WorkItem* __restrict__ items, int64_t nitems,
                                             const int32_t* __restrict__ near_off,
                                             const int32_t* __restrict__ near_src,
                                             const int64_t* __restrict__ sbeg,
                                             const int64_t* __restrict__ send,
                                             const double* __restrict__ spts,
                                             const double* __restrict__ snrm,
                                             const Cx<R>* __restrict__ w,
                                             const double* __restrict__ tpts,
                                             Cx<R>* __restrict__ result, KParams kp)
{
    const int64_t it = blockIdx.x;
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
    for (int32_t j = j0 + grp; j < j1; j += ngrp)
    {
        const int32_t s = near_src[j];
        const int64_t b = sbeg[s], e = send[s];
        for (int64_t i = b; i < e; i++)
        {
            const R x0 = R(spts[3 * i] - y0), x1 = R(spts[3 * i + 1] - y1),
                    x2 = R(spts[3 * i + 2] - y2);
            acc = cadd(acc, p2p<KIND, R>(kp, x0, x1, x2, snrm ? snrm + 3 * i : nullptr, w[i]));
        }
    }
    for (int m = Tp; m < 64; m <<= 1)
    {
        acc.x += wave_xor(acc.x, m);
        acc.y += wave_xor(acc.y, m);
    }
    if (grp == 0 && valid)
        result[t] = cadd(result[t], acc);
}

// node of coarse cone `cell` of box B for reference node xi (cone_domain.hpp:145-158,
// util.hpp:96-116): returns y - centre and r = H / s
__device__ __forceinline__ void cone_node(const BoxGeo& B, int cell, const double* xi, double y[3],
                                          double& r)
{
#pragma clang fp contract(off)
    int idx[3];
    idx[0] = cell % B.nc[0];
    cell /= B.nc[0];
    idx[1] = cell % B.nc[1];
    idx[2] = cell / B.nc[1];
    double p[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        const double a = 0.5 * B.hc[i];
        const double b = B.lo[i] + (double(idx[i]) + 0.5) * B.hc[i];
        p[i] = (xi[i] * a) + b;
    }
    r = B.H / p[0];
    y[0] = r * cos(p[2]) * sin(p[1]);
    y[1] = r * sin(p[2]) * sin(p[1]);
    y[2] = r * cos(p[1]);
}

// leaf sampling: factored kernel of the leaf's sources at the high-order nodes of one
// active coarse cone (ifgfoperator.hpp:1057-1073, :423-462)
template <int KIND, typename R>
__global__ __launch_bounds__(128) void k_leaf(const int32_t* __restrict__ work, int64_t nwork,
                                              const ConeDev* __restrict__ cones,
                                              const BoxGeo* __restrict__ geo,
                                              const int64_t* __restrict__ sbeg,
                                              const int64_t* __restrict__ send,
                                              const double* __restrict__ spts,
                                              const double* __restrict__ snrm,
                                              const Cx<R>* __restrict__ w,
                                              const double* __restrict__ ho_nodes,
                                              Cx<R>* __restrict__ V, KParams kp)
{
    constexpr int CH = 128;
    __shared__ R sx[CH], sy[CH], sz[CH];
    __shared__ R nx[CH], ny[CH], nz[CH];
    __shared__ Cx<R> sw[CH];
    const int64_t wi = blockIdx.x;
    if (wi >= nwork)
        return;
    const int32_t g = work[wi];
    const ConeDev cn = cones[g];
    const BoxGeo B = geo[cn.box];
    const int64_t b0 = sbeg[cn.box], b1 = send[cn.box];
    for (int q0 = 0; q0 < kp.n_hi; q0 += CH)
    {
        const int q = q0 + threadIdx.x;
        const bool active = q < kp.n_hi;
        double yd[3] = {0, 0, 0}, r = 1;
        if (active)
            cone_node(B, cn.cell, ho_nodes + 3 * q, yd, r);
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
                for (int k = 0; k < n; k++)
                {
                    R nn[3] = {0, 0, 0};
                    if (snrm)
                    {
                        nn[0] = nx[k];
                        nn[1] = ny[k];
                        nn[2] = nz[k];
                    }
                    acc = cadd(acc, factored<KIND, R>(kp, y0, y1, y2, dc, sx[k], sy[k], sz[k], nn,
                                                      sw[k]));
                }
        }
        if (active)
            V[int64_t(g) * kp.n_hi + q] = acc;
    }
}

// upward pass: children's fine interpolants at the parent's high-order nodes times the
// transfer factor, summed over children in order (ifgfoperator.hpp:631-688, 1639-1705)
template <int KIND, typename R>
__global__ __launch_bounds__(128) void k_up(const int32_t* __restrict__ work, int64_t nwork,
                                            const ConeDev* __restrict__ cones,
                                            const BoxGeo* __restrict__ geo,
                                            const int32_t* __restrict__ child,
                                            const int32_t* __restrict__ table,
                                            const double* __restrict__ ho_nodes,
                                            const Cx<R>* __restrict__ F, Cx<R>* __restrict__ V,
                                            KParams kp)
{
    const int64_t wi = blockIdx.x;
    if (wi >= nwork)
        return;
    const int32_t g = work[wi];
    const ConeDev cn = cones[g];
    const BoxGeo P = geo[cn.box];
    for (int q = threadIdx.x; q < kp.n_hi; q += blockDim.x)
    {
        double y[3], rp;
        cone_node(P, cn.cell, ho_nodes + 3 * q, y, rp);
        Cx<R> acc = {0, 0};
        for (int c = 0; c < 8; c++)
        {
            const int32_t ch = child[8 * cn.box + c];
            if (ch < 0)
                continue;
            const BoxGeo& C = geo[ch];
            Loc L;
            const double v0 = y[0] - (C.ctr[0] - P.ctr[0]);
            const double v1 = y[1] - (C.ctr[1] - P.ctr[1]);
            const double v2 = y[2] - (C.ctr[2] - P.ctr[2]);
            if (!locate(C, table, cones, v0, v1, v2, kp.f0, kp.f1, kp.f2, L))
                continue;
            Cx<R> val = clen3<R>(F + int64_t(L.fine) * kp.n_lo, kp.lo0, kp.lo1, kp.lo2,
                                 R(L.xi[0]), R(L.xi[1]), R(L.xi[2]));
            // e^{-k (d - dp)} dp / d with d = |y - x_child|, dp = |y - x_parent|
            Cx<R> tf;
            if constexpr (KIND == KL)
                tf = {R(rp / L.r), R(0)};
            else
                tf = cscale(expmk<R>(kp.k_re, kp.k_im, R(L.r - rp)), R(rp / L.r));
            acc = cadd(acc, cmul(val, tf));
        }
        V[int64_t(g) * kp.n_hi + q] = acc;
    }
}

// fused high-order values -> fine coefficients of the active sub-cones, separable per axis:
// D = (W2 x W1 x W0) V with W_d the (f o_lo) x o_hi axis maps (chebinterp.hpp:176-250,
// ifgfoperator.hpp:1539-1637 and the fine transform :548-570)
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
    const int64_t first = c0 + int64_t(blockIdx.x) * cpb;
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
    // axis 0: Y[j0 + m0 (i1 + h1 i2)] = sum_i0 W0[j0][i0] X[i0 + h0 (i1 + h1 i2)]
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
    // axis 1: X[j0 + m0 (j1 + m1 i2)] = sum_i1 W1[j1][i1] Y[j0 + m0 (i1 + h1 i2)]
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
    // axis 2 and scatter to the active fine sub-cones
    {
        const int n = m0 * m1 * m2;
        const int lo0 = kp.lo0, lo1 = kp.lo1, lo2 = kp.lo2;
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
            const int s0 = j0 / lo0, s1 = j1 / lo1, s2 = j2 / lo2;
            const int sub = s0 + kp.f0 * (s1 + kp.f1 * s2);
            const ConeDev cn = cones[first + k];
            if ((cn.mask >> sub) & 1u)
            {
                const int slot = cn.fine_base + __popc(cn.mask & ((1u << sub) - 1u));
                const int coef = (j0 - s0 * lo0) + lo0 * ((j1 - s1 * lo1) + lo1 * (j2 - s2 * lo2));
                F[int64_t(slot) * kp.n_lo + coef] = a;
            }
        }
    }
}

// far field: for each target of the item, the interpolants of its far source boxes times the
// centre factor (ifgfoperator.hpp:962-1012, 1708-1891)
template <int KIND, typename R>
__global__ __launch_bounds__(64) void k_far(const WorkItem* __restrict__ items, int64_t nitems,
                                            const int32_t* __restrict__ far_off,
                                            const int32_t* __restrict__ far_src,
                                            const BoxGeo* __restrict__ geo,
                                            const int32_t* __restrict__ table,
                                            const ConeDev* __restrict__ cones,
                                            const Cx<R>* __restrict__ F,
                                            const double* __restrict__ tpts,
                                            Cx<R>* __restrict__ result, KParams kp)
{
    const int64_t it = blockIdx.x;
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
    const int32_t j0 = far_off[wi.box], j1 = far_off[wi.box + 1];
    for (int32_t j = j0 + grp; j < j1; j += ngrp)
    {
        const int32_t s = far_src[j];
        const BoxGeo& B = geo[s];
        Loc L;
        if (!locate(B, table, cones, y0 - B.ctr[0], y1 - B.ctr[1], y2 - B.ctr[2], kp.f0, kp.f1,
                    kp.f2, L))
            continue;
        Cx<R> val = clen3<R>(F + int64_t(L.fine) * kp.n_lo, kp.lo0, kp.lo1, kp.lo2, R(L.xi[0]),
                             R(L.xi[1]), R(L.xi[2]));
        acc = cadd(acc, cmul(val, centre_factor<KIND, R>(kp, L.r)));
    }
    for (int m = Tp; m < 64; m <<= 1)
    {
        acc.x += wave_xor(acc.x, m);
        acc.y += wave_xor(acc.y, m);
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

KParams kparams(const DevicePlan& D)
{
    KParams k;
    k.k_re = D.k_re;
    k.k_im = D.k_im;
    k.kappa = D.kappa;
    k.dx = D.dx;
    k.ho0 = D.ho[0];
    k.ho1 = D.ho[1];
    k.ho2 = D.ho[2];
    k.lo0 = D.lo[0];
    k.lo1 = D.lo[1];
    k.lo2 = D.lo[2];
    k.f0 = D.f[0];
    k.f1 = D.f[1];
    k.f2 = D.f[2];
    k.n_hi = D.n_hi;
    k.n_lo = D.n_lo;
    return k;
}

struct Timing
{
    float ms[6] = {0, 0, 0, 0, 0, 0};
};
Timing g_last;

template <int KIND, typename R>
void run_mult(DevicePlan& D, const double* w_dev, double* out_dev, hipStream_t st, bool timed)
{
    const KParams kp = kparams(D);
    Cx<R>* ws = static_cast<Cx<R>*>(D.wsorted);
    Cx<R>* res = static_cast<Cx<R>*>(D.result);
    Cx<R>* V = static_cast<Cx<R>*>(D.V);
    Cx<R>* F = static_cast<Cx<R>*>(D.F);
    hipEvent_t ev[7];
    if (timed)
        for (auto& e : ev)
            HIP_OK(hipEventCreate(&e));
    auto mark = [&](int i)
    {
        if (timed)
            HIP_OK(hipEventRecord(ev[i], st));
    };
    mark(0);
    k_gather<R><<<dim3(unsigned((D.nsrc + 255) / 256)), dim3(256), 0, st>>>(w_dev, D.src_perm, ws,
                                                                           D.nsrc);
    k_zero<R><<<dim3(unsigned((D.ntgt + 255) / 256)), dim3(256), 0, st>>>(res, D.ntgt);
    if (D.n_near_items)
        k_near<KIND, R><<<dim3(unsigned(D.n_near_items)), dim3(64), 0, st>>>(
            D.near_items, D.n_near_items, D.near_off, D.near_src, D.src_begin, D.src_end,
            D.src_pts, D.normals, ws, D.tgt_pts, res, kp);
    mark(1);
    if (D.n_leaf_cones)
        k_leaf<KIND, R><<<dim3(unsigned(D.n_leaf_cones)), dim3(128), 0, st>>>(
            D.leaf_cones, D.n_leaf_cones, D.cones, D.geo, D.src_begin, D.src_end, D.src_pts,
            D.normals, ws, D.ho_nodes, V, kp);
    mark(2);
    // separable transform launch shape
    const int h0 = D.ho[0], h1 = D.ho[1], h2 = D.ho[2];
    const int m0 = D.f[0] * D.lo[0], m1 = D.f[1] * D.lo[1], m2 = D.f[2] * D.lo[2];
    const int xsz = std::max(h0 * h1 * h2, m0 * m1 * h2);
    const int ysz = std::max(m0 * h1 * h2, m0 * m1 * m2);
    const int nW = m0 * h0 + m1 * h1 + m2 * h2;
    const size_t per_cone = size_t(xsz + ysz) * sizeof(Cx<R>);
    const int cpb = int(std::max<size_t>(1, std::min<size_t>(8, (48 * 1024) / per_cone)));
    const size_t smem = ((nW + 3) & ~3) * sizeof(R) + cpb * per_cone;
    float t_up = 0, t_tr = 0;
    for (int l = D.levels - 1; l >= 0; l--)
    {
        const int64_t u0 = D.up_off[l], u1 = D.up_off[l + 1];
        hipEvent_t a = nullptr, b = nullptr, c = nullptr;
        if (timed)
        {
            HIP_OK(hipEventCreate(&a));
            HIP_OK(hipEventCreate(&b));
            HIP_OK(hipEventCreate(&c));
            HIP_OK(hipEventRecord(a, st));
        }
        if (u1 > u0)
            k_up<KIND, R><<<dim3(unsigned(u1 - u0)), dim3(128), 0, st>>>(
                D.up_cones + u0, u1 - u0, D.cones, D.geo, D.child, D.table, D.ho_nodes, F, V, kp);
        if (timed)
            HIP_OK(hipEventRecord(b, st));
        const int64_t c0 = D.cone_lvl_off[l], c1 = D.cone_lvl_off[l + 1];
        if (c1 > c0)
        {
            const unsigned nb = unsigned((c1 - c0 + cpb - 1) / cpb);
            k_transform<R><<<dim3(nb), dim3(256), smem, st>>>(
                c0, c1, cpb, xsz, ysz, D.cones, V, F, static_cast<const R*>(D.wmat), D.woff[1],
                D.woff[2], kp);
        }
        if (timed)
        {
            HIP_OK(hipEventRecord(c, st));
            HIP_OK(hipEventSynchronize(c));
            float x = 0, y = 0;
            HIP_OK(hipEventElapsedTime(&x, a, b));
            HIP_OK(hipEventElapsedTime(&y, b, c));
            t_up += x;
            t_tr += y;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)hipEventDestroy(c);
        }
    }
    mark(3);
    for (size_t l = 0; l + 1 < D.far_item_off.size(); l++)
    {
        const int64_t a = D.far_item_off[l], b = D.far_item_off[l + 1];
        if (b > a)
            k_far<KIND, R><<<dim3(unsigned(b - a)), dim3(64), 0, st>>>(
                D.far_items + a, b - a, D.far_off, D.far_src, D.geo, D.table, D.cones, F,
                D.tgt_pts, res, kp);
    }
    mark(4);
    k_scatter<R><<<dim3(unsigned((D.ntgt + 255) / 256)), dim3(256), 0, st>>>(res, D.tgt_perm,
                                                                            out_dev, D.ntgt);
    mark(5);
    HIP_OK(hipGetLastError());
    if (timed)
    {
        HIP_OK(hipEventSynchronize(ev[5]));
        float x;
        HIP_OK(hipEventElapsedTime(&x, ev[0], ev[1]));
        g_last.ms[0] = x;
        HIP_OK(hipEventElapsedTime(&x, ev[1], ev[2]));
        g_last.ms[1] = x;
        g_last.ms[2] = t_up;
        g_last.ms[3] = t_tr;
        HIP_OK(hipEventElapsedTime(&x, ev[3], ev[4]));
        g_last.ms[4] = x;
        HIP_OK(hipEventElapsedTime(&x, ev[0], ev[5]));
        g_last.ms[5] = x;
        for (int i = 0; i < 6; i++)
            (void)hipEventDestroy(ev[i]);
        (void)hipEventDestroy(ev[6]);
    }
}

template <typename R> void dispatch(DevicePlan& D, const double* w, double* o, hipStream_t st, bool timed)
{
    switch (D.kind)
    {
    case KIND_HELMHOLTZ:
        return run_mult<KH, R>(D, w, o, st, timed);
    case KIND_MODHELMHOLTZ:
        return run_mult<KM, R>(D, w, o, st, timed);
    case KIND_LAPLACE:
        return run_mult<KL, R>(D, w, o, st, timed);
    case KIND_DOUBLELAYER:
        return run_mult<KD, R>(D, w, o, st, timed);
    case KIND_COMBINED:
        return run_mult<KC, R>(D, w, o, st, timed);
    case KIND_GRAD:
        return run_mult<KG, R>(D, w, o, st, timed);
    }
    throw std::runtime_error("unknown kernel kind");
}

} // namespace

bool device_available()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return false;
    return n > 0;
}

void device_free(DevicePlan& D)
{
    void* ptrs[] = {D.src_pts, D.tgt_pts, D.normals, D.src_perm, D.tgt_perm, D.src_begin,
                    D.src_end, D.child, D.geo, D.cones, D.table, D.ho_nodes, D.wmat,
                    D.leaf_cones, D.up_cones, D.far_off, D.far_src, D.near_off, D.near_src,
                    D.far_items, D.near_items, D.wsorted, D.result, D.V, D.F};
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    D = DevicePlan();
}

std::string device_upload(const Plan& P, int precision, DevicePlan& D)
{
    try
    {
        device_free(D);
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
                B.hc[d] = P.hc[3 * g + d];
                B.nf[d] = int32_t(P.nf[3 * g + d]);
                B.nc[d] = int32_t(P.nc[3 * g + d]);
            }
            B.H = P.S.side[g];
            B.table_off = P.table_off[g];
        }
        D.geo = dupload(geo, tot);
        std::vector<ConeDev> cones(D.n_cones);
        for (int64_t c = 0; c < D.n_cones; c++)
            cones[c] = {P.cone_box[c], int32_t(P.cone_cell[c]), P.cone_mask[c],
                        int32_t(P.cone_fine[c])};
        D.cones = dupload(cones, tot);
        D.table = dupload(P.table, tot);
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
        // work lists
        std::vector<int32_t> leaf, up;
        D.up_off.assign(D.levels + 1, 0);
        D.cone_lvl_off.assign(D.levels + 1, 0);
        for (int l = 0; l < D.levels; l++)
        {
            const int64_t b0 = P.S.level_off[l], b1 = P.S.level_off[l + 1];
            D.cone_lvl_off[l] = P.cone_off[b0];
            D.cone_lvl_off[l + 1] = P.cone_off[b1];
            D.up_off[l] = int64_t(up.size());
            for (int64_t g = b0; g < b1; g++)
                for (int64_t c = P.cone_off[g]; c < P.cone_off[g + 1]; c++)
                    (P.S.leaf[g] ? leaf : up).push_back(int32_t(c));
            D.up_off[l + 1] = int64_t(up.size());
        }
        D.leaf_cones = dupload(leaf, tot);
        D.n_leaf_cones = int64_t(leaf.size());
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
        D.far_off = dupload(foff, tot);
        D.far_src = dupload(fsrc, tot);
        D.near_off = dupload(noff, tot);
        D.near_src = dupload(nsrcv, tot);
        // far work items by target level (one launch per level: no target is written twice)
        std::vector<WorkItem> fitems;
        D.far_item_off.assign(1, 0);
        for (int l = 0; l < P.T.levels; l++)
        {
            std::vector<int32_t> bl;
            for (int32_t t : ftb)
                if (t >= P.T.level_off[l] && t < P.T.level_off[l + 1])
                    bl.push_back(t);
            auto it = make_items(bl, P.T);
            fitems.insert(fitems.end(), it.begin(), it.end());
            D.far_item_off.push_back(int64_t(fitems.size()));
        }
        D.far_items = dupload(fitems, tot);
        auto nitems = make_items(ntb, P.T);
        D.near_items = dupload(nitems, tot);
        D.n_near_items = int64_t(nitems.size());
        // work buffers
        const size_t cs = precision == PREC_FP64 ? 16 : 8;
        D.wsorted = dalloc<unsigned char>(cs * D.nsrc, tot);
        D.result = dalloc<unsigned char>(cs * D.ntgt, tot);
        D.V = dalloc<unsigned char>(cs * size_t(std::max<int64_t>(1, D.n_cones)) * D.n_hi, tot);
        D.F = dalloc<unsigned char>(cs * size_t(std::max<int64_t>(1, D.n_fine)) * D.n_lo, tot);
        HIP_OK(hipDeviceSynchronize());
    }
    catch (const std::exception& e)
    {
        device_free(D);
        return e.what();
    }
    return "";
}

std::string device_mult(DevicePlan& D, const void* w, void* out, int64_t nrhs, bool device_ptrs,
                        void* stream)
{
    try
    {
        hipStream_t st = static_cast<hipStream_t>(stream);
        const bool timed = std::getenv("IFGF_PHASE_TIMING") != nullptr;
        double* wbuf = nullptr;
        double* obuf = nullptr;
        if (!device_ptrs)
        {
            HIP_OK(hipMalloc(&wbuf, 16 * size_t(D.nsrc)));
            HIP_OK(hipMalloc(&obuf, 16 * size_t(D.ntgt)));
        }
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
                dispatch<double>(D, wdev, odev, st, timed);
            else
                dispatch<float>(D, wdev, odev, st, timed);
            if (!device_ptrs)
                HIP_OK(hipMemcpyAsync(orr, obuf, 16 * size_t(D.ntgt), hipMemcpyDeviceToHost, st));
        }
        if (!device_ptrs)
        {
            HIP_OK(hipStreamSynchronize(st));
            HIP_OK(hipFree(wbuf));
            HIP_OK(hipFree(obuf));
        }
    }
    catch (const std::exception& e)
    {
        return e.what();
    }
    return "";
}

void device_last_timings(const DevicePlan&, double* out6)
{
    for (int i = 0; i < 6; i++)
        out6[i] = g_last.ms[i];
}

} // namespace ifgf
