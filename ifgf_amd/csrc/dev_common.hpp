// dev_common.hpp -- device-side building blocks shared by the kernel translation units:
// complex arithmetic, kernel parameters, phases, cone location, the node map, Clenshaw
// evaluation, the kernel families (P2P and factored forms), work-item mapping and the
// upward-table entry layouts.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "device.hpp"
#include "hip_util.hpp"

namespace ifgf
{
namespace dev
{

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
    int xlg; // XCD run length log2 of the launch (xcd_block)
};

constexpr double PI_D = 3.14159265358979323846;
constexpr double TWO_PI_D = 6.28318530717958647692;
constexpr double INV_2PI_D = 0.15915494309189533577;
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

// sin/cos of a phase in radians.  FP32: reduce to [-1/2, 1/2] revolutions and use the
// hardware v_sin_f32 / v_cos_f32 (which take revolutions; max abs error 1.3e-7 measured on
// gfx950 over the reduced range).
template <typename R> __device__ __forceinline__ void sincos_r(R ph, R* s, R* c)
{
    if constexpr (sizeof(R) == 4)
    {
        float t = ph * 0.15915494309189535f;
        t = t - rintf(t);
        *s = __builtin_amdgcn_sinf(t);
        *c = __builtin_amdgcn_cosf(t);
    }
    else
        sincos(ph, s, c);
}

// e^{-k delta} for a moderate argument (|k delta| of at most a few hundred radians).
// DECAY = false: the caller's kernel family has a real wavenumber (k = -i kappa, kr = 0), and
// the factor e^{-kr delta} is not formed at all (the same values: it is 1 there)
template <typename R, bool DECAY = true>
__device__ __forceinline__ Cx<R> expmk(double kr, double ki, R delta)
{
    R s, c;
    sincos_r<R>(R(-ki) * delta, &s, &c);
    if (DECAY && kr != 0.0)
    {
        R m = exp(R(-kr) * delta);
        return {m * c, m * s};
    }
    return {c, s};
}

// e^{-k r} with r in FP64; in the mixed form the phase is reduced modulo 2 pi in FP64
// before the FP32 sincos (DECAY as for expmk)
template <typename R, bool DECAY = true>
__device__ __forceinline__ Cx<R> expmk_far(double kr, double ki, double r)
{
    const double ph = -ki * r;
    R s, c;
    if constexpr (sizeof(R) == 4)
    {
        double t = ph * INV_2PI_D; // revolutions, reduced exactly in FP64
        t = t - rint(t);
        s = __builtin_amdgcn_sinf(float(t));
        c = __builtin_amdgcn_cosf(float(t));
    }
    else
        sincos(ph, &s, &c);
    if (DECAY && kr != 0.0)
    {
        R m;
        if constexpr (sizeof(R) == 4)
            m = expf(float(-kr * r));
        else
            m = exp(-kr * r);
        return {m * c, m * s};
    }
    return {c, s};
}

// 1/sqrt(x) for x > 0 from v_rsq_f64 (about 2^-23 relative) and one Newton step (about
// 2e-14 relative)
__device__ __forceinline__ double rsq_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-0.5 * x * y, y, 0.5);
    return fma(y, e, y);
}

// ------------------------------------------------------------------------------------
// cone location: physical offset v (point - box centre, FP64) -> fine cone slot and local
// coordinates in [-1,1]^3 (util.hpp:156-171 cartToInterp2, cone_domain.hpp:173-184,232-256).
// The cell search multiplies by 1/h; within 1e-9 of a cell boundary it repeats the exact
// division the plan used, so the chosen cell is the plan's cell.
// ------------------------------------------------------------------------------------
struct Loc
{
    int fine;
    double r;
    double xi[3];
};

__device__ __forceinline__ bool locate(const BoxGeo& B, const CellRef* __restrict__ ctab,
                                       double v0, double v1, double v2, int f0, int f1, int f2,
                                       Loc& out)
{
#pragma clang fp contract(off)
    const double a = v0 * v0 + v1 * v1;
    double p[3];
    p[2] = atan2(v1, v0);
    p[1] = atan2(sqrt(a), v2);
    const double r = sqrt(a + v2 * v2);
    p[0] = B.H / r;
    out.r = r;
    // inside the closed domain box (theta and phi always are)
    if (p[0] < B.lo[0] || p[0] > B.hi[0])
        return false;
    int q[3];
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
        const double d = p[j] - B.lo[j];
        const double t = d * B.inv_hf[j];
        double fl = floor(t);
        const double fr = t - fl;
        if (fr < 1e-9 || fr > 1.0 - 1e-9)
            fl = floor(d / B.hf[j]);
        q[j] = min(max(int(fl), 0), B.nf[j] - 1);
        out.xi[j] = 2.0 * (t - double(q[j])) - 1.0;
    }
    const int c0 = q[0] / f0, c1 = q[1] / f1, c2 = q[2] / f2;
    const int cell = c0 + B.nc[0] * (c1 + B.nc[1] * c2);
    const CellRef cr = ctab[B.table_off + cell];
    const int sub = (q[0] - c0 * f0) + f0 * ((q[1] - c1 * f1) + f1 * (q[2] - c2 * f2));
    if (!((cr.mask >> sub) & 1u))
        return false;
    out.fine = cr.fine_base + __popc(cr.mask & ((1u << sub) - 1u));
    return true;
}

// Mixed form: the same cone location without FP64 atan2.  Candidate cells come from a cheap
// FP32 atan2 (the |err| <= 1e-5 polynomial of Abramowitz & Stegun 4.4.47); each candidate is
// verified against its cell edges b with an FP64 cross product (theta >= b  <=>
// rho cos b - v2 sin b >= 0; phi >= b  <=>  v1 cos b - v0 sin b >= 0) and moved by one cell if
// needed, so the cell is decided in FP64.  The local angle coordinates come from the offset
// rotated into the cell-centre frame: a small-angle atan in FP32, accurate to ~3e-7.
__device__ __forceinline__ float atan2_coarse(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float z = mn * __builtin_amdgcn_rcpf(mx);
    const float z2 = z * z;
    float p = z * (0.9998660f +
                   z2 * (-0.3302995f + z2 * (0.1801410f + z2 * (-0.0851330f + z2 * 0.0208351f))));
    if (ay > ax)
        p = 1.5707963267948966f - p;
    if (x < 0.f)
        p = 3.141592653589793f - p;
    return copysignf(p, y);
}

__device__ __forceinline__ float atan_small(float u)
{
    if (fabsf(u) <= 0.3f)
    {
        const float u2 = u * u; // Taylor series to u^13: error < 1e-9 on |u| <= 0.3
        return u * (1.f + u2 * (-1.f / 3 + u2 * (1.f / 5 + u2 * (-1.f / 7 + u2 * (1.f / 9 +
                                                      u2 * (-1.f / 11 + u2 * (1.f / 13)))))));
    }
    return atanf(u);
}

struct LocF
{
    int fine;
    double r;    // |v| (about 2e-14 relative)
    double rinv; // 1/|v|
    float xi[3];
};

// one angle axis of locate_fast: the cell from the FP32 candidate, the local coordinate from
// the offset rotated into the cell-centre frame (FP64 rotation, FP32 small-angle atan).  When
// |xi| comes out within 1e-5 of 1 the point is close to an edge (the candidate's error is about
// 1e-5 rad, far below half a cell): the cell is then decided by the FP64 cross products with the
// edges.  E: (cos, sin) of the n + 1 edges, then of the n centres.  (p, q): the offset in the
// axis' plane, angle = atan2(p, q).
__device__ __forceinline__ int angle_cell(const double* E, int n, double lo_shift, double inv_h,
                                          double p, double q, float& xi)
{
    const float cand = atan2_coarse(float(p), float(q)) + float(lo_shift);
    int c = min(max(int(cand * float(inv_h)), 0), n - 1);
    const float sc = float(2.0 * inv_h);
    const double* Cc = E + 2 * (n + 1) + 2 * c;
    double xr = q * Cc[0] + p * Cc[1], yr = p * Cc[0] - q * Cc[1];
    xi = atan_small(__fdividef(float(yr), float(xr))) * sc;
    if (fabsf(xi) >= 1.f - 1e-5f)
    {
        if (c > 0 && !(p * E[2 * c] - q * E[2 * c + 1] >= 0.0))
            c--;
        else if (c < n - 1 && p * E[2 * c + 2] - q * E[2 * c + 3] >= 0.0)
            c++;
        Cc = E + 2 * (n + 1) + 2 * c;
        xr = q * Cc[0] + p * Cc[1];
        yr = p * Cc[0] - q * Cc[1];
        xi = atan_small(__fdividef(float(yr), float(xr))) * sc;
    }
    return c;
}

// FF: the fine/coarse factor of every axis when known at compile time (else f0, f1, f2)
template <int FF = 0>
__device__ __forceinline__ bool locate_fast(const BoxGeo& B, const double* eth, const double* eph,
                                            const CellRef* __restrict__ ctab, double v0,
                                            double v1, double v2, int f0_, int f1_, int f2_,
                                            LocF& out)
{
    const int f0 = FF ? FF : f0_, f1 = FF ? FF : f1_, f2 = FF ? FF : f2_;
    double a, r2;
    {
#pragma clang fp contract(off)
        a = v0 * v0 + v1 * v1;
        r2 = a + v2 * v2;
    }
    const double rinv = rsq_nr(r2);
    double s = B.H * rinv;
    const double rho = a > 0.0 ? a * rsq_nr(a) : 0.0;
    out.r = r2 * rinv;
    out.rinv = rinv;
    int q[3];
    {
        // s decides the domain and the radial cell: near a bound or a cell edge it is
        // recomputed as the plan computes it (correctly rounded sqrt and division)
        double d = s - B.lo[0];
        double t = d * B.inv_hf[0];
        double fl = floor(t);
        double fr = t - fl;
        const double tol = 1e-12 * s;
        if (fabs(s - B.lo[0]) < tol || fabs(s - B.hi[0]) < tol || fr < 1e-9 ||
            fr > 1.0 - 1e-9)
        {
            s = B.H / sqrt(r2);
            if (s < B.lo[0] || s > B.hi[0])
                return false;
            d = s - B.lo[0];
            t = d * B.inv_hf[0];
            fl = floor(t);
            fr = t - fl;
            if (fr < 1e-9 || fr > 1.0 - 1e-9)
                fl = floor(d / B.hf[0]);
        }
        else if (s < B.lo[0] || s > B.hi[0])
            return false;
        q[0] = min(max(int(fl), 0), B.nf[0] - 1);
        out.xi[0] = float(2.0 * (t - double(q[0])) - 1.0);
    }
    const int n1 = B.nf[1], n2 = B.nf[2];
    // theta in [0, pi] (rho = 0 included: theta = 0 or pi)
    q[1] = angle_cell(eth, n1, 0.0, B.inv_hf[1], rho, v2, out.xi[1]);
    if (a != 0.0) // phi in [-pi, pi]
        q[2] = angle_cell(eph, n2, PI_D, B.inv_hf[2], v1, v0, out.xi[2]);
    else
    {
        // on the polar axis phi is atan2 of signed zeros (+-0 or +-pi), placed by the same
        // floor as the plan's
        const double p = signbit(v0) ? (signbit(v1) ? -PI_D : PI_D) : (signbit(v1) ? -0.0 : 0.0);
        const double d = p - B.lo[2];
        const double t = d * B.inv_hf[2];
        double fl = floor(t);
        const double fr = t - fl;
        if (fr < 1e-9 || fr > 1.0 - 1e-9)
            fl = floor(d / B.hf[2]);
        q[2] = min(max(int(fl), 0), n2 - 1);
        out.xi[2] = float(2.0 * (t - double(q[2])) - 1.0);
    }
    const int c0 = q[0] / f0, s0 = q[0] - c0 * f0;
    {
        const int c1 = q[1] / f1, c2 = q[2] / f2;
        const CellRef cr = ctab[B.table_off + c0 + B.nc[0] * (c1 + B.nc[1] * c2)];
        const int sub = s0 + f0 * ((q[1] - c1 * f1) + f1 * (q[2] - c2 * f2));
        if ((cr.mask >> sub) & 1u)
        {
            out.fine = cr.fine_base + __popc(cr.mask & ((1u << sub) - 1u));
            return true;
        }
    }
    // a point within rounding of a theta or phi edge which the plan's atan2 and floor put in
    // the neighbouring cell: the neighbours across the nearer edges (phi wraps around)
    const int d1 = out.xi[1] > 0.f ? 1 : -1, d2 = out.xi[2] > 0.f ? 1 : -1;
#pragma unroll 1
    for (int m = 1; m < 4; m++)
    {
        const int e1 = (m & 1) ? d1 : 0, e2 = (m & 2) ? d2 : 0;
        const int p1 = q[1] + e1;
        int p2 = q[2] + e2;
        p2 = p2 < 0 ? p2 + n2 : (p2 >= n2 ? p2 - n2 : p2);
        if (p1 < 0 || p1 >= n1)
            continue;
        const int c1 = p1 / f1, c2 = p2 / f2;
        const CellRef cr = ctab[B.table_off + c0 + B.nc[0] * (c1 + B.nc[1] * c2)];
        const int sub = s0 + f0 * ((p1 - c1 * f1) + f1 * (p2 - c2 * f2));
        if ((cr.mask >> sub) & 1u)
        {
            out.fine = cr.fine_base + __popc(cr.mask & ((1u << sub) - 1u));
            out.xi[1] -= float(2 * e1);
            out.xi[2] -= float(2 * e2);
            return true;
        }
    }
    return false;
}

// physical high-order node q of coarse cone `cell` of box B (cone_domain.hpp:145-158,
// util.hpp:96-116) with the plan's operation order; the cosines and sines of the node angles
// come from tables computed on the host from the same arguments.  absolute: y = centre + ...
// (as the plan forms it), otherwise y - centre.
struct NodeTabs
{
    const double* ho_xi; // n_hi x 3 reference nodes
    const double* th;    // (cos, sin) per (theta cell, theta node)
    const double* ph;    // (cos, sin) per (phi cell, phi node)
};

__device__ __forceinline__ void cone_node(const BoxGeo& B, int cell, int q, int h0, int h1, int h2,
                                          const NodeTabs& T, bool absolute, double y[3],
                                          double& r)
{
#pragma clang fp contract(off)
    const int c0 = cell % B.nc[0];
    const int c12 = cell / B.nc[0];
    const int c1 = c12 % B.nc[1], c2 = c12 / B.nc[1];
    const int j12 = q / h0;
    const int j1 = j12 % h1, j2 = j12 / h1;
    const double a = 0.5 * B.hc[0];
    const double b = B.lo[0] + (double(c0) + 0.5) * B.hc[0];
    const double s = (T.ho_xi[3 * q] * a) + b;
    r = B.H / s;
    const int it = 2 * (B.th_off + c1 * h1 + j1), ip = 2 * (B.ph_off + c2 * h2 + j2);
    const double ct = T.th[it], st = T.th[it + 1];
    const double cp = T.ph[ip], sp = T.ph[ip + 1];
    if (absolute)
    {
        y[0] = B.ctr[0] + r * cp * st;
        y[1] = B.ctr[1] + r * sp * st;
        y[2] = B.ctr[2] + r * ct;
    }
    else
    {
        y[0] = r * cp * st;
        y[1] = r * sp * st;
        y[2] = r * ct;
    }
}

// ------------------------------------------------------------------------------------
// tensor Clenshaw (chebinterp.hpp:252-346).  Compile-time orders: coefficients in registers;
// order 0: runtime fallback reading the block from memory.
// ------------------------------------------------------------------------------------
template <int N, typename R> __device__ __forceinline__ Cx<R> clen1c(const Cx<R>* c, R x)
{
    if constexpr (N == 1)
        return c[0];
    else if constexpr (N == 2)
        return {c[0].x + x * c[1].x, c[0].y + x * c[1].y};
    else
    {
        const R x2 = x + x;
        Cx<R> b2 = c[N - 1];
        Cx<R> b1 = {x2 * b2.x + c[N - 2].x, x2 * b2.y + c[N - 2].y};
#pragma unroll
        for (int j = N - 3; j > 0; j--)
        {
            Cx<R> t = {c[j].x + x2 * b1.x - b2.x, c[j].y + x2 * b1.y - b2.y};
            b2 = b1;
            b1 = t;
        }
        return {x * b1.x - b2.x + c[0].x, x * b1.y - b2.y + c[0].y};
    }
}

template <int O0, int O1, int O2, typename R>
__device__ __forceinline__ Cx<R> clen3c(const Cx<R>* c, R x0, R x1, R x2)
{
    Cx<R> plane[O2];
#pragma unroll
    for (int i2 = 0; i2 < O2; i2++)
    {
        Cx<R> line[O1];
#pragma unroll
        for (int i1 = 0; i1 < O1; i1++)
            line[i1] = clen1c<O0, R>(c + O0 * (i1 + O1 * i2), x0);
        plane[i2] = clen1c<O1, R>(line, x1);
    }
    return clen1c<O2, R>(plane, x2);
}

template <typename R> __device__ Cx<R> clen_line(const Cx<R>* c, int n, R x)
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

template <typename R> __device__ Cx<R> clen_plane(const Cx<R>* c, int o0, int o1, R x0, R x1)
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
__device__ Cx<R> clen3r(const Cx<R>* c, int o0, int o1, int o2, R x0, R x1, R x2v)
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

// the value of clen3c on the coefficients of fine slot `slot`, loaded a chunk of axis-2 planes
// at a time (chunks start on 16-byte boundaries), so that fewer coefficients are live at once;
// the same operations in the same order as clen3c
template <int O0, int O1, int O2, typename R>
__device__ __forceinline__ Cx<R> fine_eval_chunked(const Cx<R>* __restrict__ F, int slot, R x0,
                                                   R x1, R x2)
{
    constexpr int NP = O0 * O1;
    constexpr int PC = (sizeof(R) == 4 && (NP % 2) != 0) ? 2 : 1; // planes per chunk
    static_assert(O2 % PC == 0, "whole chunks of planes");
    const Cx<R>* p = F + int64_t(slot) * (NP * O2);
    Cx<R> plane[O2];
#pragma unroll
    for (int i2 = 0; i2 < O2; i2 += PC)
    {
        Cx<R> c[NP * PC];
        if constexpr (sizeof(R) == 4 && (NP * PC) % 2 == 0)
        {
            const float4* p4 = reinterpret_cast<const float4*>(p + NP * i2);
#pragma unroll
            for (int i = 0; i < NP * PC / 2; i++)
            {
                const float4 v = p4[i];
                c[2 * i] = {v.x, v.y};
                c[2 * i + 1] = {v.z, v.w};
            }
        }
        else
        {
#pragma unroll
            for (int i = 0; i < NP * PC; i++)
                c[i] = p[NP * i2 + i];
        }
#pragma unroll
        for (int k = 0; k < PC; k++)
        {
            Cx<R> line[O1];
#pragma unroll
            for (int i1 = 0; i1 < O1; i1++)
                line[i1] = clen1c<O0, R>(c + NP * k + O0 * i1, x0);
            plane[i2 + k] = clen1c<O1, R>(line, x1);
        }
    }
    return clen1c<O2, R>(plane, x2);
}

// value of the fine interpolant in slot `slot` at local coordinates xi
template <int O0, int O1, int O2, typename R>
__device__ __forceinline__ Cx<R> fine_eval(const Cx<R>* __restrict__ F, int slot, const KParams& kp,
                                           R x0, R x1, R x2)
{
    if constexpr (O0 > 0)
    {
        constexpr int N = O0 * O1 * O2;
        Cx<R> c[N];
        const Cx<R>* p = F + int64_t(slot) * N;
        if constexpr (sizeof(R) == 4 && (N % 2) == 0)
        {
            const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
            for (int i = 0; i < N / 2; i++)
            {
                const float4 v = p4[i];
                c[2 * i] = {v.x, v.y};
                c[2 * i + 1] = {v.z, v.w};
            }
        }
        else
        {
#pragma unroll
            for (int i = 0; i < N; i++)
                c[i] = p[i];
        }
        return clen3c<O0, O1, O2, R>(c, x0, x1, x2);
    }
    else
        return clen3r<R>(F + int64_t(slot) * kp.n_lo, kp.lo0, kp.lo1, kp.lo2, x0, x1, x2);
}

// ------------------------------------------------------------------------------------
// kernel functions.  P2P: G(x_src - y) (helmholtz_ifgf.hpp:37-41 and siblings)
// ------------------------------------------------------------------------------------
// distance and its inverse: in the mixed form from one v_rsq_f32 (1 ulp), in FP64 exactly
// (sqrt, then a division).  d2 > 0.
template <typename R> __device__ __forceinline__ void dist_inv(R d2, R& d, R& invd)
{
    if constexpr (sizeof(R) == 4)
    {
        invd = __builtin_amdgcn_rsqf(d2);
        d = d2 * invd;
    }
    else
    {
        d = sqrt(d2);
        invd = R(1) / d;
    }
}

// a / b: v_rcp_f32 (1 ulp) in the mixed form, a division in FP64
template <typename R> __device__ __forceinline__ R fdiv_r(R a, R b)
{
    if constexpr (sizeof(R) == 4)
        return a * __builtin_amdgcn_rcpf(b);
    else
        return a / b;
}

template <int KIND, typename R>
__device__ __forceinline__ Cx<R> p2p(const KParams& kp, R x0, R x1, R x2, double xn, Cx<R> w)
{
    // (x0,x1,x2) = source - target; xn = n . (source - target) with the source normal n,
    // formed by the caller in FP64 from the FP64 difference (double layer and combined field):
    // for close points on a smooth surface it is far smaller than |source - target| and loses
    // its leading digits if formed from the rounded difference
    const R d2 = x0 * x0 + x1 * x1 + x2 * x2;
    const R inv4pi = R(INV_4PI_D);
    if constexpr (KIND == KH || KIND == KM || (KIND == KG))
    {
        if (KIND == KG ? (d2 < R(1e-24)) : (d2 == R(0)))
            return {0, 0};
        R d, invd;
        dist_inv(d2, d, invd);
        if (KIND == KG && kp.dx >= 0)
        {
            const R xd = kp.dx == 0 ? x0 : (kp.dx == 1 ? x1 : x2);
            Cx<R> e = expmk<R>(kp.k_re, kp.k_im, d);
            Cx<R> mk = {R(-kp.k_re) - invd, R(-kp.k_im)};
            Cx<R> v = cmul(e, mk);
            R s = -inv4pi * (invd * invd) * xd;
            return cmul(w, cscale(v, s));
        }
        Cx<R> e = expmk<R, KIND != KH>(kp.k_re, kp.k_im, d);
        return cmul(w, cscale(e, inv4pi * invd));
    }
    else if constexpr (KIND == KL)
    {
        if (d2 == R(0))
            return {0, 0};
        R d, invd;
        dist_inv(d2, d, invd);
        return cscale(w, inv4pi * invd);
    }
    else if constexpr (KIND == KD)
    {
        if (d2 < R(1e-24))
            return {0, 0};
        R d, invd;
        dist_inv(d2, d, invd);
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, d);
        Cx<R> mk = {R(-kp.k_re) - invd, R(-kp.k_im)};
        Cx<R> v = cmul(e, mk);
        return cmul(w, cscale(v, -inv4pi * (invd * invd) * R(xn)));
    }
    else // KC: combined_field_helmholtz_ifgf.hpp:43-68
    {
        R d = 0, invd = 0;
        if (!(d2 < R(1e-14)))
            dist_inv(d2, d, invd);
        const R dd = d2 * invd;
        const R nxy = R(-xn);
        const R kap = R(kp.kappa);
        R s, c;
        sincos_r<R>(kap * dd, &s, &c);
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
    if (d2 == R(0)) // a source on a node: the kernels vanish there (G(0) = 0)
        return {0, 0};
    R d, invd;
    dist_inv(d2, d, invd);
    R delta;
    if constexpr (sizeof(R) == 4)
        delta = fdiv_r(x0 * (x0 - R(2) * y0) + x1 * (x1 - R(2) * y1) + x2 * (x2 - R(2) * y2),
                       d + dc);
    else
        delta = d - dc;
    if constexpr (KIND == KH || KIND == KM)
    {
        Cx<R> e = expmk<R, KIND != KH>(kp.k_re, kp.k_im, delta);
        return cmul(w, cscale(e, dc * invd));
    }
    else if constexpr (KIND == KL)
    {
        return cscale(w, dc * invd);
    }
    else if constexpr (KIND == KG)
    {
        const R id = d2 > R(1e-12) ? invd : R(0);
        if (kp.dx < 0)
        {
            Cx<R> e = expmk<R>(kp.k_re, kp.k_im, delta);
            return cmul(w, cscale(e, dc * id));
        }
        const R zd = kp.dx == 0 ? z0 : (kp.dx == 1 ? z1 : z2);
        const R ww = -(invd * invd) * dc * zd;
        Cx<R> e = expmk<R>(kp.k_re, kp.k_im, delta);
        Cx<R> mk = {R(-kp.k_re) - id, R(-kp.k_im)};
        return cmul(w, cscale(cmul(e, mk), ww));
    }
    else if constexpr (KIND == KD)
    {
        const R id = d2 > R(1e-12) ? invd : R(0);
        const R ww = -(invd * invd) * dc * (z0 * nrm[0] + z1 * nrm[1] + z2 * nrm[2]);
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
        sincos_r<R>(kap * delta, &s, &c);
        const R f = dc * (invd * invd * invd);
        const R a = w.x * c - w.y * s, b = w.x * s + w.y * c;
        return {f * (a * nxy + b * kap * (d * nxy + d2)),
                f * (-a * kap * (d * nxy + d2) + b * nxy)};
    }
}


// lane layout of a 64-wide work item: Tp lanes per group (power of two >= targets), 64/Tp
// groups splitting the interaction list; reduction over groups by xor shuffles.
// Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b mod 8) and every XCD has its
// own L2.  Neighbouring work items read much of the same data (the same source boxes, the same
// children, the same cells), so each XCD takes runs of 2^lg consecutive items, the runs dealt
// out round-robin so that costs drifting along a work list stay balanced: workgroup b handles
// item xcd_block(lg), a bijection of [0, gridDim.x) for every run length (the tail beyond the
// last whole round of runs keeps its order; run length 1 is the plain order).  The run length
// is a launch shape the autotuner picks per operator; it travels with each launch (KParams::xlg
// or a kernel argument), so launches of operators with different shapes never share state.
__device__ __forceinline__ int64_t xcd_block(int lg)
{
    const int64_t n = gridDim.x, b = blockIdx.x;
    const int64_t full = (n >> (3 + lg)) << (3 + lg);
    if (b >= full)
        return b;
    const int64_t x = b & 7, i = b >> 3;
    return ((((i >> lg) << 3) + x) << lg) + (i & ((int64_t(1) << lg) - 1));
}

// The near field and the far groups of a mult on the plan's side stream (host side).  The
// near field only writes the result (after the zeroing) and the far field adds to it (after
// the near field), one far group after another (a group: one source level and one target
// level, deepest source level first), each reading the coefficients of its source level only.
// mode 1: fork() starts the near field on the side stream (called after the deepest level's
// transform, so that it runs beside the memory-bound upward kernels rather than the VALU-bound
// leaf sampling); mode 2: fork() runs it on the main stream (called first).  In both, ready(l),
// called once the upward chain has written source level l, enqueues on the side stream the far
// groups of level l (and any deeper ones left), and finish() joins the side stream back.  Mode
// 0: the near field at fork() and the far groups at finish(), in the same order, on the main
// stream (the sums are the same in every mode).
template <class NearFn, class FarFn> struct SidePipe
{
    DevicePlan& D;
    hipStream_t st;
    bool on;     // far groups on the side stream (modes 1 and 2)
    bool near_side; // the near field on the side stream (mode 1)
    NearFn near; // near(stream)
    FarFn far;   // far(target level, stream)
    bool forked = false;
    int next;    // the next far group to enqueue (deepest source level first)
    // level storage (DevicePlan::ring) on one stream: the far groups of a source level run as
    // soon as its coefficients are written, before the level two above overwrites them
    bool inline_far;
    SidePipe(DevicePlan& d, hipStream_t s, int mode, NearFn n, FarFn f)
        : D(d), st(s), on(mode != 0 && d.side_stream), near_side(on && mode == 1), near(n),
          far(f), next(0), inline_far(!on && d.ring)
    {
    }
    hipStream_t side() const { return static_cast<hipStream_t>(D.side_stream); }
    void fork()
    {
        if (forked)
            return;
        forked = true;
        if (!near_side)
        {
            near(st);
            return;
        }
        HIP_OK(hipEventRecord(static_cast<hipEvent_t>(D.ev_fork), st));
        HIP_OK(hipStreamWaitEvent(side(), static_cast<hipEvent_t>(D.ev_fork), 0));
        near(side());
    }
    void ready(int done)
    {
        const int ng = int(D.far_groups.size());
        if (inline_far)
        {
            while (next < ng && D.far_groups[next].src_level >= done)
                far(next++, st);
            return;
        }
        if (!on || !forked)
            return;
        bool waited = false;
        while (next < ng && D.far_groups[next].src_level >= done)
        {
            if (!waited)
            {
                hipEvent_t e = static_cast<hipEvent_t>(D.ev_level[done]);
                HIP_OK(hipEventRecord(e, st));
                HIP_OK(hipStreamWaitEvent(side(), e, 0));
                waited = true;
            }
            far(next++, side());
        }
        if (D.ring)
            HIP_OK(hipEventRecord(static_cast<hipEvent_t>(D.ev_far_done[done]), side()));
    }
    // level storage on two streams: the main stream waits until the far groups of source level
    // `lvl` (and deeper) have read their coefficients, before their buffer is written again
    void reuse(int lvl)
    {
        if (!on || !D.ring || !forked || lvl >= D.levels)
            return;
        HIP_OK(hipStreamWaitEvent(st, static_cast<hipEvent_t>(D.ev_far_done[lvl]), 0));
    }
    void finish()
    {
        fork();
        if (on)
        {
            ready(0);
            HIP_OK(hipEventRecord(static_cast<hipEvent_t>(D.ev_near), side()));
            HIP_OK(hipStreamWaitEvent(st, static_cast<hipEvent_t>(D.ev_near), 0));
            return;
        }
        for (; next < int(D.far_groups.size()); next++)
            far(next, st);
    }
};
template <class NearFn, class FarFn>
SidePipe<NearFn, FarFn> side_pipe(DevicePlan& D, hipStream_t st, int mode, NearFn n, FarFn f)
{
    return SidePipe<NearFn, FarFn>(D, st, mode, n, f);
}

__device__ __forceinline__ int pow2_at_least(int n)
{
    int p = 1;
    while (p < n)
        p <<= 1;
    return p;
}

// upward-geometry table entries (host-built per parent level, see UpLevel in plan.hpp)
struct UpEntF
{
    int32_t cf, sub;
    float xi0, xi1, xi2, tr, ti;
    int32_t d; // index of (cf, sub) among the distinct child cones of its (cell, octant)
};
struct UpEntD
{
    int32_t cf, sub;
    double xi0, xi1, xi2, tr, ti;
};
template <typename R> using UpEnt = typename std::conditional<sizeof(R) == 4, UpEntF, UpEntD>::type;

// The upward table of one (cell, octant) regrouped for evaluation (FP32): the parent nodes
// whose positions fall in one child fine cone, in runs of at most UP_GRP nodes, so that one
// thread evaluates the interpolant of that cone at all of them from one read of its
// coefficients (96 B, six float4)
constexpr int UP_GRP = 4;
struct UpGrp
{
    int32_t cf;          // coarse cell of the child grid holding the fine cone
    int32_t subd;        // sub-cone index (bits 0-7) | index d of the cone in its list << 8
    uint32_t q01, q23;   // parent node indices, 16 bits each (0xffff: an empty place)
    float x0[UP_GRP], x1[UP_GRP], x2[UP_GRP]; // local coordinates in the fine cone
    float tr[UP_GRP], ti[UP_GRP];             // transfer factors
};
static_assert(sizeof(UpGrp) == 96, "six float4");

// the kernel parameters with the wavenumber of box B (plans of several geometries carry one
// per box)
__device__ __forceinline__ KParams with_k(KParams kp, const BoxGeo& B)
{
    kp.k_re = B.k_re;
    kp.k_im = B.k_im;
    kp.kappa = B.kappa;
    return kp;
}

inline KParams kparams(const DevicePlan& D)
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
    k.xlg = D.shape.xcd_log2;
    return k;
}

// which compile-time order set (high / low orders, fine factors 3) matches the plan:
// 1 = (3,6,6)/(2,3,3), order 6 without a tolerance; 2 = (5,8,8)/(2,5,5), order 8 without a
// tolerance; 3 = (6,8,7)/(3,5,4), what the tolerance estimate picks for order 6 at tol 1e-5 on
// the unit sphere at N = 1M, kappa = 100.  0: the runtime-order kernels.
inline int order_class(const DevicePlan& D)
{
    const bool f3 = D.f[0] == 3 && D.f[1] == 3 && D.f[2] == 3;
    auto is = [&](int h0, int h1, int h2, int l0, int l1, int l2) {
        return f3 && D.ho[0] == h0 && D.ho[1] == h1 && D.ho[2] == h2 && D.lo[0] == l0 &&
               D.lo[1] == l1 && D.lo[2] == l2;
    };
    if (is(3, 6, 6, 2, 3, 3))
        return 1;
    if (is(5, 8, 8, 2, 5, 5))
        return 2;
    if (is(6, 8, 7, 3, 5, 4))
        return 3;
    return 0;
}

// whether k_up_grp transforms its cones itself, by the number of fine coefficients: measured
// at 1M points, a gain at (6,8,7)/(3,5,4) (60; 0.44 ms per mult) and a loss at (3,6,6)/(2,3,3)
// (18; 0.1 ms)
constexpr bool up_fuses_transform(int n_lo) { return n_lo >= 50; }

// leaf sampling and k_up_grp transform their cones: FP32 at a compiled order class with enough
// fine coefficients (the standalone transform then takes the per-level lists of the other
// cones)
inline bool up_fused_transform(const DevicePlan& D)
{
    return D.prec != PREC_FP64 && order_class(D) != 0 && up_fuses_transform(D.n_lo);
}

} // namespace dev
} // namespace ifgf
