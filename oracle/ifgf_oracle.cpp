// ifgf_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU, FP64, IEEE-strict restatement of the IFGF operator-application path of
// TimHRO/ifgf (IfgfOperator::init / ::mult and the kernel hooks), written from the
// reference's source text.  It is the checker for the HIP library: tests,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it; the
// product (ifgf_amd/) never does.  Each function names the reference file:line it
// follows.  Deliberate, documented deviations from the reference:
//   * fine cones that are active but not covered by an active coarse cone are zero
//     (the reference leaves them uninitialised: ifgfoperator.hpp:892-893, :1624);
//   * upward contributions of sibling boxes at the first recursive level are summed
//     per parent in child order after the parallel phase (the reference adds them
//     concurrently, ifgfoperator.hpp:1181, a data race);
//   * estimateRefinement takes the maximum over all ten sampled boxes and samples
//     element 0 (ifgfoperator.hpp:122-136, :218; the reference's result depends on the
//     TBB partition and its element draw evaluates to 0);
//   * when no level has more than 32 boxes every level is processed iteratively
//     (the reference indexes past its level arrays in that case, ifgfoperator.hpp:292-308).
// Compiled without -ffast-math and with -ffp-contract=off.

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace orc
{
using cplx = std::complex<double>;
static const double PI = 3.14159265358979323846; // M_PI
static const size_t NONE = std::numeric_limits<size_t>::max();

// ----------------------------------------------------------------------------------------
// threading helper (stands in for the reference's TBB loops)
// ----------------------------------------------------------------------------------------
static void parallel_for(int64_t n, int nthreads, const std::function<void(int64_t, int)>& body)
{
    if (nthreads <= 1 || n < 2)
    {
        for (int64_t i = 0; i < n; i++)
            body(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    const int64_t chunk = std::max<int64_t>(1, n / (int64_t(nthreads) * 16));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back(
            [&, t]()
            {
                for (;;)
                {
                    int64_t b = next.fetch_add(chunk);
                    if (b >= n)
                        break;
                    int64_t e = std::min(n, b + chunk);
                    for (int64_t i = b; i < e; i++)
                        body(i, t);
                }
            });
    for (auto& th : pool)
        th.join();
}

// ----------------------------------------------------------------------------------------
// Axis-aligned box (boundingbox.hpp:8-32 on top of Eigen::AlignedBox)
// ----------------------------------------------------------------------------------------
struct Box3
{
    double lo[3], hi[3], c[3], side;
};

// boundingbox.hpp:13-26: centre = (max+min)/2, sideLength = max |diag_i|
static Box3 make_box(const double lo[3], const double hi[3])
{
    Box3 b;
    double m = 0;
    for (int i = 0; i < 3; i++)
    {
        b.lo[i] = lo[i];
        b.hi[i] = hi[i];
        b.c[i] = (hi[i] + lo[i]) / 2.0;
        m = std::max(m, std::abs(hi[i] - lo[i]));
    }
    b.side = m;
    return b;
}

// Eigen AlignedBox::squaredExteriorDistance: per axis, distance to the nearer face if
// outside, summed over axes 0,1,2.
static double sq_ext_dist(const double lo[3], const double hi[3], const double p[3])
{
    double d2 = 0;
    for (int k = 0; k < 3; k++)
    {
        if (lo[k] > p[k])
        {
            double a = lo[k] - p[k];
            d2 += a * a;
        }
        else if (p[k] > hi[k])
        {
            double a = p[k] - hi[k];
            d2 += a * a;
        }
    }
    return d2;
}

// Eigen AlignedBox::contains (closed box)
static bool contains(const Box3& b, const double p[3])
{
    return b.lo[0] <= p[0] && b.lo[1] <= p[1] && b.lo[2] <= p[2] && p[0] <= b.hi[0] &&
           p[1] <= b.hi[1] && p[2] <= b.hi[2];
}

static double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// ----------------------------------------------------------------------------------------
// Morton order (zorder_less.hpp:67-248): sign of z, y, x first; then the axis whose
// coordinates differ in the highest bit (ties to the higher axis); then signed compare.
// ----------------------------------------------------------------------------------------
static int float_exp(uint64_t bits) // zorder_less.hpp:89-105
{
    uint64_t u = bits & 0x7fffffffffffffffull;
    if (u == 0 || u >= 0x7ff0000000000000ull)
        return 0;
    u >>= 52;
    return (u == 0) ? -1022 : int(u) - 1023;
}

static int ilog2_u64(uint64_t x)
{
    int r = -1;
    while (x)
    {
        x >>= 1;
        r++;
    }
    return r;
}

static int float_xor_msb(double p, double q) // zorder_less.hpp:180-205
{
    if (p == q || p == -q)
        return std::numeric_limits<int>::min();
    uint64_t pu, qu;
    std::memcpy(&pu, &p, 8);
    std::memcpy(&qu, &q, 8);
    int pe = float_exp(pu), qe = float_exp(qu);
    if (pe == qe)
    {
        uint64_t x = (pu & 0x000fffffffffffffull) ^ (qu & 0x000fffffffffffffull);
        if (x > 0)
            return pe + ilog2_u64(x) - 52;
        return pe;
    }
    return std::max(pe, qe);
}

struct MortonLess // zorder_less.hpp:212-248
{
    Box3 bb;
    double diam[3];
    explicit MortonLess(const Box3& b) : bb(b)
    {
        for (int i = 0; i < 3; i++)
            diam[i] = b.hi[i] - b.lo[i];
    }
    bool operator()(const double* px, const double* qx) const
    {
        double p[3], q[3];
        for (int i = 0; i < 3; i++)
        {
            p[i] = (px[i] - bb.c[i]) / diam[i];
            q[i] = (qx[i] - bb.c[i]) / diam[i];
        }
        int x = std::numeric_limits<int>::min();
        int k = 0;
        for (int j = 2; j >= 0; j--)
        {
            if ((p[j] < 0.0) != (q[j] < 0.0))
                return p[j] < q[j];
            int y = float_xor_msb(p[j], q[j]);
            if (x < y)
            {
                x = y;
                k = j;
            }
        }
        return p[k] < q[k];
    }
};

// ----------------------------------------------------------------------------------------
// Cone grid on the (s, theta, phi) box (cone_domain.hpp:63-306)
// ----------------------------------------------------------------------------------------
struct ConeGrid
{
    size_t n[3] = {0, 0, 0};
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, h[3] = {1, 1, 1};
    std::vector<size_t> active;                 // insertion order (memId order)
    std::unordered_map<size_t, size_t> mem;     // cone id -> memId

    size_t total() const { return n[0] * n[1] * n[2]; }

    void setup(const size_t ne[3], const double blo[3], const double bhi[3]) // :73-84
    {
        for (int i = 0; i < 3; i++)
        {
            n[i] = ne[i];
            lo[i] = blo[i];
            hi[i] = bhi[i];
            h[i] = (hi[i] - lo[i]) / double(n[i]);
        }
    }

    void indices(size_t j, size_t idx[3]) const // :186-198
    {
        for (int i = 0; i < 3; i++)
        {
            idx[i] = j % n[i];
            j = j / n[i];
        }
    }

    size_t element_for_point(const double p[3]) const // :232-256
    {
        if (sq_ext_dist(lo, hi, p) > 0)
            return NONE;
        size_t idx = 0;
        size_t stride = 1;
        for (int j = 0; j < 3; j++)
        {
            int q = int(std::floor((p[j] - lo[j]) / h[j]));
            size_t ij = size_t(std::clamp(q, 0, int(n[j] - 1)));
            idx += ij * stride;
            stride *= n[j];
        }
        return idx;
    }

    void transform(size_t el, const double xi[3], double out[3]) const // :145-158
    {
        size_t idx[3];
        indices(el, idx);
        for (int i = 0; i < 3; i++)
        {
            double a = 0.5 * h[i];
            double b = lo[i] + (double(idx[i]) + 0.5) * h[i];
            out[i] = (xi[i] * a) + b;
        }
    }

    void transform_back(size_t el, const double p[3], double xi[3]) const // :173-184
    {
        size_t idx[3];
        indices(el, idx);
        for (int i = 0; i < 3; i++)
        {
            double a = 0.5 * h[i];
            double b = lo[i] + (double(idx[i]) + 0.5) * h[i];
            xi[i] = (p[i] - b) / a;
        }
    }

    bool is_active(size_t el) const { return mem.count(el) > 0; }
    size_t mem_id(size_t el) const { return mem.at(el); } // :125-132 (throws if inactive)
};

// util.hpp:96-116 (3-D branch)
static void interp_to_cart(const double p[3], const double xc[3], double H, double out[3])
{
    double r = H / p[0];
    out[0] = xc[0] + r * std::cos(p[2]) * std::sin(p[1]);
    out[1] = xc[1] + r * std::sin(p[2]) * std::sin(p[1]);
    out[2] = xc[2] + r * std::cos(p[1]);
}

// util.hpp:156-171 (cartToInterp2)
static void cart_to_interp(const double x[3], const double xc[3], double H, double out[3])
{
    double p0 = x[0] - xc[0], p1 = x[1] - xc[1], p2 = x[2] - xc[2];
    double a = p0 * p0 + p1 * p1;
    out[2] = std::atan2(p1, p0);
    out[1] = std::atan2(std::sqrt(a), p2);
    out[0] = H / std::sqrt(a + p2 * p2);
}

// ----------------------------------------------------------------------------------------
// Chebyshev machinery (chebinterp.hpp)
// ----------------------------------------------------------------------------------------
static std::vector<double> chebnodes1d(int n) // chebinterp.hpp:27-39
{
    std::vector<double> x(n);
    for (int i = 0; i < n; i++)
        x[i] = std::cos((2. * i + 1.0) / (2. * n) * PI);
    return x;
}

// chebinterp.hpp:88-114: 3-D nodes with axis 0 fastest; returns 3*n0*n1*n2 doubles
static std::vector<double> chebnodes3d(const int ns[3])
{
    auto r = chebnodes1d(ns[0]), t = chebnodes1d(ns[1]), p = chebnodes1d(ns[2]);
    std::vector<double> out(size_t(3) * ns[0] * ns[1] * ns[2]);
    for (int i = 0; i < ns[2]; i++)
        for (int j = 0; j < ns[1]; j++)
            for (int k = 0; k < ns[0]; k++)
            {
                size_t id = size_t(i) * ns[1] * ns[0] + size_t(j) * ns[0] + k;
                out[3 * id + 0] = r[k];
                out[3 * id + 1] = t[j];
                out[3 * id + 2] = p[i];
            }
    return out;
}

// chebinterp.hpp:116-149: T_idx(x_sigma) stored at [sigma*n + idx]
static std::vector<double> chebvals(int n)
{
    std::vector<double> a(size_t(n) * n);
    for (int idx = 0; idx < n; idx++)
        for (int s = 0; s < n; s++)
            a[size_t(s) * n + idx] = std::cos(idx * PI * (2. * s + 1.) / (2. * n));
    return a;
}

// chebinterp.hpp:176-250: separable values->coefficients, recursing over the last axis.
static void chebtransform_rec(const cplx* src, cplx* dst, const int* ns, int dim)
{
    const int nl = ns[dim - 1];
    int nsigma = 1;
    for (int i = 0; i < dim - 1; i++)
        nsigma *= ns[i];
    const int stride = nsigma;
    auto cv = chebvals(nl);
    for (int i = 0; i < nl * nsigma; i++)
        dst[i] = 0;
    if (dim == 1)
    {
        for (int idx = 0; idx < nl; idx++)
        {
            for (int s = 0; s < nl; s++)
            {
                double Td = cv[size_t(s) * nl + idx];
                dst[idx] += src[s] * Td;
            }
            dst[idx] *= (idx == 0 ? 1. : 2.) * 1. / nl;
        }
        return;
    }
    std::vector<cplx> M(size_t(nl) * nsigma);
    for (int idx = 0; idx < nl; idx++)
        chebtransform_rec(src + size_t(idx) * stride, M.data() + size_t(idx) * stride, ns,
                          dim - 1);
    for (int idx = 0; idx < nl; idx++)
    {
        for (int s = 0; s < nl; s++)
        {
            double Td = cv[size_t(s) * nl + idx];
            for (int q = 0; q < nsigma; q++)
                dst[size_t(idx) * stride + q] += M[size_t(s) * stride + q] * Td;
        }
        for (int q = 0; q < nsigma; q++)
            dst[size_t(idx) * stride + q] *= (idx == 0 ? 1 : 2) * 1. / nl;
    }
}

static void chebtransform3(const cplx* src, cplx* dst, const int ns[3])
{
    chebtransform_rec(src, dst, ns, 3);
}

// chebinterp.hpp:252-346: tensor Clenshaw, outermost recursion over the last axis (the
// coefficients of an axis are the Clenshaw sums over the axes before it), the recursion
// written out for one, two and three axes; no halving of c_0
static inline cplx clen_d1(double x, const cplx* c, int n)
{
    if (n == 1)
        return c[0];
    if (n == 2)
        return x * c[1] + c[0];
    cplx b2 = c[n - 1];
    cplx b1 = 2. * x * c[n - 1] + c[n - 2];
    for (int j = n - 3; j > 0; j--)
    {
        cplx tmp = c[j] + (2. * (b1 * x) - b2);
        b2 = b1;
        b1 = tmp;
    }
    return (b1 * x - b2) + c[0];
}

static inline cplx clen_d2(const double* x, const cplx* c, int n0, int n1)
{
    auto coef = [&](int j) { return clen_d1(x[0], c + size_t(j) * n0, n0); };
    const double xd = x[1];
    if (n1 == 1)
        return coef(0);
    if (n1 == 2)
        return xd * coef(1) + coef(0);
    cplx b2 = coef(n1 - 1);
    cplx b1 = 2. * b2 * xd + coef(n1 - 2);
    for (int j = n1 - 3; j > 0; j--)
    {
        cplx tmp = coef(j) + (2. * (b1 * xd) - b2);
        b2 = b1;
        b1 = tmp;
    }
    return (b1 * xd - b2) + coef(0);
}

static cplx clenshaw3(const double xi[3], const cplx* c, const int ns[3])
{
    const int n0 = ns[0], n1 = ns[1], n2 = ns[2];
    const size_t stride = size_t(n0) * n1;
    auto coef = [&](int j) { return clen_d2(xi, c + j * stride, n0, n1); };
    const double xd = xi[2];
    if (n2 == 1)
        return coef(0);
    if (n2 == 2)
        return xd * coef(1) + coef(0);
    cplx b2 = coef(n2 - 1);
    cplx b1 = 2. * b2 * xd + coef(n2 - 2);
    for (int j = n2 - 3; j > 0; j--)
    {
        cplx tmp = coef(j) + (2. * (b1 * xd) - b2);
        b2 = b1;
        b1 = tmp;
    }
    return (b1 * xd - b2) + coef(0);
}

// chebinterp.cpp:145-229 evaluated pointwise: the values of a tensor interpolant at the
// tensor grid pts0 x pts1 x pts2 (axis 0 fastest).
static void tp_evaluate(const std::vector<double>* pts, const cplx* c, const int ns[3], cplx* out)
{
    size_t n0 = pts[0].size(), n1 = pts[1].size(), n2 = pts[2].size();
    for (size_t i2 = 0; i2 < n2; i2++)
        for (size_t i1 = 0; i1 < n1; i1++)
            for (size_t i0 = 0; i0 < n0; i0++)
            {
                double x[3] = {pts[0][i0], pts[1][i1], pts[2][i2]};
                out[i0 + n0 * (i1 + n1 * i2)] = clenshaw3(x, c, ns);
            }
}

// util.hpp:214-234
static double compute_slice_norm(const cplx* data, size_t rows, const int ns[3], int axis,
                                 int layers)
{
    double v1 = 0, v2 = 0;
    for (size_t idx = 0; idx < rows; idx++)
    {
        size_t split[3];
        size_t j = idx;
        for (int i = 0; i < 3; i++)
        {
            split[i] = j % size_t(ns[i]);
            j /= size_t(ns[i]);
        }
        double n = std::norm(data[idx]);
        v1 += n;
        if (split[axis] == size_t(ns[axis] - layers))
            v2 += n;
    }
    return std::sqrt(v2) / std::max(1., std::sqrt(v1));
}

// ----------------------------------------------------------------------------------------
// Kernel families (helmholtz_ifgf.hpp, modified_helmholtz_ifgf.hpp, laplace_ifgf.hpp,
// double_layer_helmholtz_ifgf.hpp, combined_field_helmholtz_ifgf.hpp,
// grad_helmholtz_ifgf.hpp).  Every family writes its Green's function with a complex
// wavenumber k as e^{-k d}; Helmholtz with real kappa is k = -i kappa.
// ----------------------------------------------------------------------------------------
enum Kind
{
    K_HELMHOLTZ = 0,
    K_MODHELMHOLTZ = 1,
    K_LAPLACE = 2,
    K_DOUBLELAYER = 3,
    K_COMBINED = 4,
    K_GRAD = 5
};

struct Kernel
{
    int kind = K_HELMHOLTZ;
    cplx k{0, 0};      // complex wavenumber for the e^{-k d} families
    double kappa = 0;  // real wavenumber (Helmholtz, combined field)
    int dx = -1;       // Grad: -1 potential, 0/1/2 target derivative component
    double tol = -1;
    const double* normals = nullptr; // source normals in source Morton order

    // per-kernel orderForBox (helmholtz_ifgf.hpp:158-170 and siblings)
    void order_for_box(double, const int base[3], int step, int out[3]) const
    {
        for (int i = 0; i < 3; i++)
        {
            out[i] = base[i];
            if (step == 0)
            {
                out[i] = base[i] - 3;
                if (kind == K_HELMHOLTZ || kind == K_GRAD)
                    out[i] = std::max(out[i], 2);
            }
        }
    }

    // per-kernel elementsForBox (helmholtz_ifgf.hpp:172-195 and siblings)
    void elements_for_box(double H, const size_t base_in[3], int step, size_t out[3]) const
    {
        size_t base[3] = {base_in[0], base_in[1], base_in[2]};
        if (step == 0)
        {
            size_t f = (kind == K_MODHELMHOLTZ) ? 1 : 3;
            for (int i = 0; i < 3; i++)
                base[i] *= f;
        }
        if (kind == K_LAPLACE) // laplace_ifgf.hpp:147-157
        {
            for (int i = 0; i < 3; i++)
                out[i] = base[i];
            return;
        }
        for (int i = 0; i < 3; i++)
        {
            int idelta = 1;
            if (kind == K_HELMHOLTZ || kind == K_COMBINED)
            {
                double delta = std::max(kappa * H / 4., 1.);
                idelta = int(std::ceil(delta));
            }
            else if (kind == K_MODHELMHOLTZ) // modified_helmholtz_ifgf.hpp:172-180
            {
                int d = int(std::ceil(
                    std::max(std::abs(k.imag()) * H / (2 * (2 + k.real())), 1.0)));
                idelta = int(std::ceil(double(d)));
            }
            else // double layer / grad (double_layer_helmholtz_ifgf.hpp:205-209)
            {
                double delta = std::max(std::abs(k.imag()) * H / 4., 1.0) *
                               std::exp(-0.2 * (3.0 / std::sqrt(3.0)) * H * k.real());
                idelta = int(std::ceil(delta));
            }
            out[i] = std::max(base[i] * size_t(idelta), size_t(1));
        }
    }

    // ifgfoperator.hpp:1914; grad_helmholtz_ifgf.hpp:294-304
    double cutoff_limit(double H) const
    {
        if (kind == K_GRAD && tol > 0)
            return std::max(1e-4, std::abs(k.real()) * H / std::abs(std::log(tol)));
        return 1e-4;
    }

    // evaluateKernel single pair: x = source - target (the reference's argument order)
    cplx p2p(const double* xs, const double* y, int64_t src_idx) const
    {
        double x[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
        switch (kind)
        {
        case K_HELMHOLTZ: // helmholtz_ifgf.hpp:37-41
        {
            double d = norm3(x);
            return (d == 0) ? cplx(0) : (1 / (4 * PI)) * std::exp(cplx(0, kappa) * d) / d;
        }
        case K_MODHELMHOLTZ: // modified_helmholtz_ifgf.hpp:37-41
        {
            double d = norm3(x);
            return (d == 0) ? cplx(0) : (1 / (4 * PI)) * std::exp(-k * d) / d;
        }
        case K_LAPLACE: // laplace_ifgf.hpp:33-37
        {
            double d = norm3(x);
            return (d == 0) ? 0.0 : (1 / (4 * PI)) / d;
        }
        case K_DOUBLELAYER: // double_layer_helmholtz_ifgf.hpp:58-66
        {
            const double* n = normals + 3 * src_idx;
            double d = norm3(x);
            double xn = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
            return d < 1e-12 ? cplx(0.0)
                             : -(1.0 / (4.0 * PI)) * (1.0 / (d * d)) * std::exp(-k * d) *
                                   (-k - 1.0 / d) * xn;
        }
        case K_COMBINED: // combined_field_helmholtz_ifgf.hpp:43-68
        {
            const double* n = normals + 3 * src_idx;
            double d2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            double invd = d2 < 1e-14 ? 0.0 : (1.0 / std::sqrt(d2));
            double d = d2 * invd;
            double nxy = -(n[0] * x[0] + n[1] * x[1] + n[2] * x[2]);
            double s = std::sin(kappa * d), c = std::cos(kappa * d);
            double f = (1.0 / (4. * PI)) * invd * invd * invd;
            double re = f * (c * nxy + s * kappa * (d * nxy + d2));
            double im = f * (-c * kappa * (d * nxy + d2) + s * nxy);
            return cplx(re, im);
        }
        case K_GRAD: // grad_helmholtz_ifgf.hpp:44-58
        {
            double d = norm3(x);
            if (dx == -1)
                return (d < 1e-12) ? cplx(0) : (1 / (4 * PI)) * std::exp(-k * d) / d;
            return d < 1e-12 ? cplx(0.0)
                             : -(1.0 / (4.0 * PI)) * (1.0 / (d * d)) * std::exp(-k * d) *
                                   (-k - 1.0 / d) * x[dx];
        }
        }
        return 0;
    }

    // evaluateFactoredKernel: value at node y of the source sum, factored by the centre
    // factor; dc = |y - xc|.  Returns the contribution of one source with weight w.
    cplx factored(const double* xs, const double* y, cplx w, double dc, int64_t src_idx) const
    {
        switch (kind)
        {
        case K_HELMHOLTZ: // helmholtz_ifgf.hpp:114-156
        {
            double z[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
            double d = norm3(z);
            return w * std::exp(cplx(0, kappa) * (d - dc)) * dc / d;
        }
        case K_MODHELMHOLTZ: // modified_helmholtz_ifgf.hpp:114-136
        {
            double z[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
            double d = norm3(z);
            return (d == 0) ? cplx(0) : w * std::exp(-k * (d - dc)) * (dc) / d;
        }
        case K_LAPLACE: // laplace_ifgf.hpp:109-131
        {
            double z[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
            double d = norm3(z);
            return (d == 0) ? cplx(0) : w * (dc) / d;
        }
        case K_DOUBLELAYER: // double_layer_helmholtz_ifgf.hpp:131-159
        {
            const double* n = normals + 3 * src_idx;
            double z[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
            double d2 = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
            double id = (d2 > 1e-12) ? 1.0 / std::sqrt(d2) : 0;
            double d = d2 * id;
            double ww = (-(1.0 / (d2)) * dc * (z[0] * n[0] + z[1] * n[1] + z[2] * n[2]));
            return w * std::exp(-k * (d - dc)) * (-k - id) * ww;
        }
        case K_COMBINED: // combined_field_helmholtz_ifgf.hpp:131-182
        {
            const double* n = normals + 3 * src_idx;
            double z[3] = {y[0] - xs[0], y[1] - xs[1], y[2] - xs[2]};
            double d2 = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
            if (!(d2 > 1e-14))
                return 0;
            double nxy = z[0] * n[0] + z[1] * n[1] + z[2] * n[2];
            double id = 1.0 / std::sqrt(d2);
            double d = d2 * id;
            double s = std::sin(kappa * (d - dc)), c = std::cos(kappa * (d - dc));
            double f = dc / (d2 * d);
            double wr = w.real(), wi = w.imag();
            double re = f * ((wr * c - wi * s) * nxy + (wr * s + wi * c) * kappa * (d * nxy + d2));
            double im =
                f * (-(wr * c - wi * s) * kappa * (d * nxy + d2) + (wr * s + wi * c) * nxy);
            return cplx(re, im);
        }
        case K_GRAD: // grad_helmholtz_ifgf.hpp:170-214
        {
            double z[3] = {xs[0] - y[0], xs[1] - y[1], xs[2] - y[2]};
            double d2 = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
            double id = d2 > 1e-12 ? 1.0 / std::sqrt(d2) : 0.0;
            double d = d2 * id;
            if (dx == -1)
            {
                double ww = (dc)*id;
                return ww * w * std::exp(-k * (d - dc));
            }
            double ww = (-(1.0 / (d2)) * dc * (xs[dx] - y[dx]));
            return w * std::exp(-k * (d - dc)) * (-k - id) * ww;
        }
        }
        return 0;
    }

    // CF (single-point branch): the centre factor G-like function of y - xc
    cplx cf(const double* x) const
    {
        double d2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        double id = 1.0 / (std::sqrt(d2));
        double d = d2 * id;
        switch (kind)
        {
        case K_HELMHOLTZ: // helmholtz_ifgf.hpp:57-64
        case K_COMBINED:  // combined_field_helmholtz_ifgf.hpp:101-109
            return std::exp(cplx(0, kappa) * d) * id * (1 / (4.0 * PI));
        case K_LAPLACE: // laplace_ifgf.hpp:52-58
            return id * (1 / (4.0 * PI));
        default: // modified helmholtz, double layer, grad
            return std::exp(-k * d) * id * (1 / (4.0 * PI));
        }
    }

    // transfer_factor: factor that re-centres a child interpolant on its parent
    cplx transfer(const double* x, const double* xc, const double* pxc) const
    {
        double a[3] = {x[0] - xc[0], x[1] - xc[1], x[2] - xc[2]};
        double b[3] = {x[0] - pxc[0], x[1] - pxc[1], x[2] - pxc[2]};
        double d = norm3(a), dp = norm3(b);
        switch (kind)
        {
        case K_HELMHOLTZ: // helmholtz_ifgf.hpp:67-92
            return std::exp(cplx(0, kappa) * (d - dp)) * dp / d;
        case K_COMBINED: // combined_field_helmholtz_ifgf.hpp:112-127
            return std::exp(cplx(0, kappa) * (d - dp)) * (dp * (1.0 / d));
        case K_LAPLACE: // laplace_ifgf.hpp:61-85
            return dp / d;
        case K_GRAD: // grad_helmholtz_ifgf.hpp:95-123
        {
            double d2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            double invd = d2 > 1e-12 ? 1.0 / std::sqrt(d2) : 0.0;
            double dd = d2 * invd;
            return std::exp(-k * (dd - dp)) * (dp * invd);
        }
        case K_DOUBLELAYER: // double_layer_helmholtz_ifgf.hpp:98-113
            return std::exp(-k * (d - dp)) * (dp * (1.0 / d));
        default: // modified_helmholtz_ifgf.hpp:67-92
            return std::exp(-k * (d - dp)) * dp / d;
        }
    }
};

// ----------------------------------------------------------------------------------------
// Octree (octree.hpp)
// ----------------------------------------------------------------------------------------
struct Range
{
    int64_t b, e;
};

struct Node
{
    int64_t begin = 0, end = 0;
    int32_t parent = -1;
    int32_t child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    bool leaf = true;
    Box3 bb;
    std::vector<Range> far, near;    // target point ranges (octree.hpp:136-146)
    ConeGrid grid[2];                // step 0 (fine, low order), step 1 (coarse, high order)
    bool has_far_anc = false;        // hasFarTargetsIncludingAncestors (octree.hpp:947-955)
    bool parent_has_far = false;     // parentHasFarTargets (octree.hpp:957-978)
};

struct Tree
{
    size_t max_leaf = 0;
    int64_t n = 0;
    std::vector<std::vector<Node>> lv;
    std::vector<int64_t> perm; // sorted position -> original index
    std::vector<double> pts;   // sorted points, xyz interleaved
    double diameter = 0;
    Box3 root_box;

    int levels() const { return int(lv.size()); }

    // octree.hpp:1106-1184: greedy closed-box scan for each child in order j=0..7,
    // child j covering min + (diag/2)*(j&1, j>>1&1, j>>2&1).
    int32_t build_node(int32_t parent, int64_t b, int64_t e, const Box3& bb, int level)
    {
        if (b == e)
            return -1;
        if (level >= int(lv.size()))
            lv.emplace_back();
        int32_t id = int32_t(lv[level].size());
        lv[level].emplace_back();
        {
            Node& nd = lv[level][id];
            nd.begin = b;
            nd.end = e;
            nd.parent = parent;
            nd.bb = bb;
        }
        const int64_t N = e - b;
        if (size_t(N) <= max_leaf)
            return id;
        int64_t pi = b;
        for (int j = 0; j < 8; j++)
        {
            double mn[3], mx[3], size[3];
            for (int i = 0; i < 3; i++)
            {
                mn[i] = bb.lo[i];
                mx[i] = bb.hi[i];
                size[i] = 0.5 * (bb.hi[i] - bb.lo[i]);
            }
            double tup[3] = {double(j & 1), double((j >> 1) & 1), double((j >> 2) & 1)};
            for (int i = 0; i < 3; i++)
            {
                mn[i] += size[i] * tup[i];
                mx[i] = mn[i] + size[i];
            }
            Box3 cb = make_box(mn, mx);
            int64_t ep = pi;
            while (ep < e && contains(cb, &pts[3 * ep]))
                ++ep;
            int32_t c = build_node(id, pi, ep, cb, level + 1);
            if (c >= 0)
            {
                lv[level][id].child[j] = c;
                lv[level][id].leaf = false;
            }
            pi = ep;
        }
        return id;
    }

    // octree.hpp:202-238
    void build(const double* p, int64_t N, size_t maxLeaf)
    {
        max_leaf = maxLeaf;
        n = N;
        lv.clear();
        double mn[3] = {p[0], p[1], p[2]}, mx[3] = {p[0], p[1], p[2]};
        for (int64_t i = 0; i < N; i++)
            for (int d = 0; d < 3; d++)
            {
                mn[d] = std::min(mn[d], p[3 * i + d]);
                mx[d] = std::max(mx[d], p[3 * i + d]);
            }
        double nmn = norm3(mn), nmx = norm3(mx);
        for (int d = 0; d < 3; d++)
        {
            mn[d] -= 1e-8 * nmn;
            mx[d] += 1e-8 * nmx;
        }
        Box3 bb = make_box(mn, mx);
        root_box = bb;
        perm.resize(N);
        std::iota(perm.begin(), perm.end(), 0);
        MortonLess less(bb);
        std::sort(perm.begin(), perm.end(),
                  [&](int64_t a, int64_t b) { return less(p + 3 * a, p + 3 * b); });
        pts.resize(3 * size_t(N));
        for (int64_t i = 0; i < N; i++)
            for (int d = 0; d < 3; d++)
                pts[3 * i + d] = p[3 * perm[i] + d];
        double dg[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
        diameter = norm3(dg);
        build_node(-1, 0, N, bb, 0);
    }
};

// ----------------------------------------------------------------------------------------
// The operator (ifgfoperator.hpp)
// ----------------------------------------------------------------------------------------
struct InterpData
{
    int order[3] = {0, 0, 0};
    const ConeGrid* grid = nullptr;
    std::vector<cplx> values;
    size_t stride() const { return size_t(order[0]) * order[1] * order[2]; }
};

struct FieldInfo
{
    std::vector<int64_t> starts, idx;
};

struct Operator
{
    Kernel K;
    Tree S, T;
    size_t base_els[3];
    int base_order[3];
    int64_t nsrc = 0, ntgt = 0;
    int min_rec = 0;
    std::vector<FieldInfo> far_info, near_info;
    std::vector<double> normals_sorted;
    int nthreads = 1;
    std::function<bool(double)> cutoff = nullptr;
    struct Counter
    {
        int64_t near = 0, far = 0, tofar = 0;
    } counter;

    Operator(Kernel k, int64_t leaf, int order, int n_elem, double tol) : K(k)
    {
        // ifgfoperator.hpp:36-59
        base_els[0] = 1 * size_t(n_elem);
        base_els[1] = 2 * size_t(n_elem);
        base_els[2] = 4 * size_t(n_elem);
        S.max_leaf = size_t(leaf);
        T.max_leaf = size_t(leaf);
        for (int i = 0; i < 3; i++)
            base_order[i] = order;
        K.tol = tol;
    }

    bool has_far_anc(int level, int id) const { return S.lv[level][id].has_far_anc; }

    // octree.hpp:271-349: dual traversal
    void interact(int ls, int s, int lt, int t)
    {
        if (s < 0 || t < 0)
            return;
        Node& sn = S.lv[ls][s];
        const Node& tn = T.lv[lt][t];
        if (sn.begin == sn.end || tn.begin == tn.end)
            return;
        double ed = std::sqrt(sq_ext_dist(tn.bb.lo, tn.bb.hi, sn.bb.c));
        bool adm = ed >= std::sqrt(3.0) * sn.bb.side; // octree.hpp:192-197
        bool tofar = cutoff ? cutoff(ed) : false;
        if (adm)
        {
            if (tofar)
            {
                counter.tofar++;
                return;
            }
            sn.far.push_back({tn.begin, tn.end});
            counter.far++;
            return;
        }
        if (sn.leaf && tn.leaf)
        {
            sn.near.push_back({tn.begin, tn.end});
            counter.near++;
            return;
        }
        if (sn.leaf && !tn.leaf)
        {
            for (int j = 0; j < 8; j++)
                interact(ls, s, lt + 1, tn.child[j]);
            return;
        }
        if (tn.leaf && !sn.leaf)
        {
            for (int j = 0; j < 8; j++)
                interact(ls + 1, sn.child[j], lt, t);
            return;
        }
        if (sn.bb.side > tn.bb.side)
        {
            for (int j = 0; j < 8; j++)
                interact(ls + 1, sn.child[j], lt, t);
        }
        else
        {
            for (int j = 0; j < 8; j++)
                interact(ls, s, lt + 1, tn.child[j]);
        }
    }

    // ifgfoperator.hpp:108-244
    void estimate_refinement()
    {
        const int level = S.levels() - 1;
        const size_t nb = S.lv[level].size();
        const size_t sample = 10;
        const size_t stride = nb / sample;
        int best[3] = {0, 0, 0};
        for (size_t i = 0; i < sample; i++)
        {
            int r[3];
            estimate_on_box(level, int(i * stride), r);
            for (int d = 0; d < 3; d++)
                best[d] = std::max(best[d], r[d]);
        }
        for (int d = 0; d < 3; d++)
            base_order[d] = best[d];
    }

    // one sampling step of estimateRefinementOnBox (ifgfoperator.hpp:181-217): the factored
    // kernel of the box's sources with unit weights at the Chebyshev nodes of `order` in
    // element 0 of the box's coarse grid, and its Chebyshev coefficients.  ypts: the nodes in
    // Cartesian coordinates; info: H, centre (3), smin, the grid's elements (3)
    void estimate_sample(int level, int id, const int order[3], std::vector<double>& ypts,
                         std::vector<cplx>& data, std::vector<cplx>& trafo, double info[8])
    {
        const Node& nd = S.lv[level][id];
        const double H = nd.bb.side;
        const double* xc = nd.bb.c;
        const int64_t nS = nd.end - nd.begin;
        double smax = std::sqrt(3.0) / 3;
        double smin = K.cutoff_limit(H);
        double blo[3] = {smin, 0, -PI}, bhi[3] = {smax, PI, PI};
        size_t nels[3];
        K.elements_for_box(H, base_els, 1, nels);
        ConeGrid g;
        g.setup(nels, blo, bhi);
        auto nodes = chebnodes3d(order);
        const size_t nn = nodes.size() / 3;
        ypts.assign(3 * nn, 0.0);
        data.assign(nn, cplx(0, 0));
        trafo.assign(nn, cplx(0, 0));
        // the sampled element, (n_elements * Random()) / SIZE_MAX in the reference (:218),
        // evaluates to 0
        const size_t el = 0;
        for (size_t q = 0; q < nn; q++)
        {
            double sp[3], y[3];
            g.transform(el, &nodes[3 * q], sp);
            interp_to_cart(sp, xc, H, y);
            double dv[3] = {y[0] - xc[0], y[1] - xc[1], y[2] - xc[2]};
            double dc = norm3(dv);
            cplx acc = 0;
            for (int64_t i = 0; i < nS; i++)
                acc += K.factored(&S.pts[3 * (nd.begin + i)], y, cplx(1.0, 0.0), dc,
                                  nd.begin + i);
            data[q] = acc;
            for (int d = 0; d < 3; d++)
                ypts[3 * q + d] = y[d];
        }
        chebtransform3(data.data(), trafo.data(), order);
        info[0] = H;
        for (int d = 0; d < 3; d++)
        {
            info[1 + d] = xc[d];
            info[5 + d] = double(nels[d]);
        }
        info[4] = smin;
    }

    void estimate_on_box(int level, int id, int out[3]) // ifgfoperator.hpp:143-244
    {
        const int maxR = 10;
        const double H = S.lv[level][id].bb.side;
        int base[3] = {0, 0, 0};
        double err[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                         std::numeric_limits<double>::max()};
        int new_p[3] = {base_order[0], base_order[1], base_order[2]};
        auto maxc = [](const double* e) { return std::max(e[0], std::max(e[1], e[2])); };
        std::vector<double> ypts;
        std::vector<cplx> data, trafo;
        double info[8];
        while (maxc(err) > K.tol && std::max(base[0], std::max(base[1], base[2])) < maxR)
        {
            for (int d = 0; d < 3; d++)
                new_p[d] = base_order[d] + base[d];
            const int layers = 3;
            int plus[3] = {new_p[0] + layers + 1, new_p[1] + layers + 1, new_p[2] + layers + 1};
            int order[3];
            K.order_for_box(H, plus, 1, order);
            estimate_sample(level, id, order, ypts, data, trafo, info);
            for (int d = 0; d < 3; d++)
                err[d] = compute_slice_norm(trafo.data(), trafo.size(), order, d, layers);
            for (int d = 0; d < 3; d++)
                if (err[d] > K.tol)
                    base[d] += 1;
        }
        for (int d = 0; d < 3; d++)
            out[d] = new_p[d];
    }

    // octree.hpp:351-816 (TwoGrid branch, EXACT_INTERP_RANGE off)
    void interpolation_range()
    {
        for (int level = 0; level < S.levels(); level++)
        {
            auto& L = S.lv[level];
            parallel_for(int64_t(L.size()), nthreads,
                         [&](int64_t n, int)
                         {
                             Node& nd = L[n];
                             const double* xc = nd.bb.c;
                             const double H = nd.bb.side;
                             double smax = std::sqrt(3.0) / 3;
                             double smin = K.cutoff_limit(H);
                             if (smin >= smax)
                                 smin = smax - 0.05;
                             double blo[3] = {smin, 0, -PI}, bhi[3] = {smax, PI, PI};
                             size_t nf[3], nc[3];
                             K.elements_for_box(H, base_els, 0, nf);
                             K.elements_for_box(H, base_els, 1, nc);
                             ConeGrid& fine = nd.grid[0];
                             ConeGrid& coarse = nd.grid[1];
                             fine.setup(nf, blo, bhi);
                             coarse.setup(nc, blo, bhi);
                             std::vector<size_t> act[2];
                             std::unordered_set<size_t> seen[2];
                             auto mark = [&](const double s[3])
                             {
                                 size_t c0 = fine.element_for_point(s);
                                 size_t c1 = coarse.element_for_point(s);
                                 if (c1 != NONE && seen[1].insert(c1).second)
                                     act[1].push_back(c1);
                                 if (c0 != NONE && seen[0].insert(c0).second)
                                     act[0].push_back(c0);
                             };
                             for (const Range& r : nd.far)
                                 for (int64_t i = r.b; i < r.e; i++)
                                 {
                                     double s[3];
                                     cart_to_interp(&T.pts[3 * i], xc, H, s);
                                     mark(s);
                                 }
                             if (nd.parent_has_far)
                             {
                                 const Node& pn = S.lv[level - 1][nd.parent];
                                 const double* pxc = pn.bb.c;
                                 const double pH = pn.bb.side;
                                 int pord[3];
                                 K.order_for_box(pH, base_order, 1, pord);
                                 auto nodes = chebnodes3d(pord);
                                 const size_t nn = nodes.size() / 3;
                                 for (size_t el : pn.grid[1].active)
                                     for (size_t q = 0; q < nn; q++)
                                     {
                                         double sp[3], y[3], s[3];
                                         pn.grid[1].transform(el, &nodes[3 * q], sp);
                                         interp_to_cart(sp, pxc, pH, y);
                                         cart_to_interp(y, xc, H, s);
                                         mark(s);
                                     }
                             }
                             for (int st = 0; st < 2; st++)
                             {
                                 nd.grid[st].active = act[st];
                                 nd.grid[st].mem.clear();
                                 for (size_t m = 0; m < act[st].size(); m++)
                                     nd.grid[st].mem[act[st][m]] = m;
                             }
                         });
            if (level < min_rec)
            {
                far_info[level] = field_info(level, true);
                near_info[level] = field_info(level, false);
            }
        }
    }

    // octree.hpp:818-880 (sized by the number of targets)
    FieldInfo field_info(int level, bool far) const
    {
        FieldInfo f;
        f.starts.assign(size_t(ntgt) + 1, 0);
        std::vector<int64_t> cnt(size_t(ntgt), 0);
        const auto& L = S.lv[level];
        for (size_t n = 0; n < L.size(); n++)
            for (const Range& r : (far ? L[n].far : L[n].near))
                for (int64_t t = r.b; t < r.e; t++)
                    cnt[t]++;
        for (int64_t t = 0; t < ntgt; t++)
            f.starts[t + 1] = f.starts[t] + cnt[t];
        f.idx.resize(size_t(f.starts[ntgt]));
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t n = 0; n < L.size(); n++)
            for (const Range& r : (far ? L[n].far : L[n].near))
                for (int64_t t = r.b; t < r.e; t++)
                    f.idx[f.starts[t] + cnt[t]++] = int64_t(n);
        return f;
    }

    // ifgfoperator.hpp:65-106
    void init(const double* srcs, int64_t N, const double* tgts, int64_t M, const double* normals)
    {
        nsrc = N;
        ntgt = M;
        S.build(srcs, N, S.max_leaf);
        T.build(tgts, M, T.max_leaf);
        counter = Counter();
        interact(0, 0, 0, 0);
        if (normals) // double_layer_helmholtz_ifgf.hpp:29-35
        {
            normals_sorted.resize(3 * size_t(N));
            for (int64_t i = 0; i < N; i++)
                for (int d = 0; d < 3; d++)
                    normals_sorted[3 * i + d] = normals[3 * S.perm[i] + d];
            K.normals = normals_sorted.data();
        }
        // ancestor flags (octree.hpp:947-978)
        for (int l = 0; l < S.levels(); l++)
            for (Node& nd : S.lv[l])
            {
                nd.parent_has_far = false;
                if (l > 0)
                {
                    const Node& p = S.lv[l - 1][nd.parent];
                    nd.parent_has_far = !p.far.empty() || p.parent_has_far;
                }
                nd.has_far_anc = !nd.far.empty() || nd.parent_has_far;
            }
        if (K.tol > 0)
            estimate_refinement();
        else
            base_order[0] = std::max(base_order[0] - 3, 2);
        // the unclamped fine order (base - 3) of Laplace, modified Helmholtz, double layer and
        // combined field leaves interpolants without coefficients for small orders
        // (laplace_ifgf.hpp:133-145): the reference's behaviour is undefined there; refuse it
        for (int step = 0; step < 2; step++)
        {
            int o[3];
            K.order_for_box(1.0, base_order, step, o);
            for (int d = 0; d < 3; d++)
                if (o[d] < 1)
                    throw std::runtime_error("interpolation order below 1 (the fine order of "
                                             "this kernel is order-3)");
        }
        // RECURSIVE_MULT level choice (octree.hpp:372-380; ifgfoperator.hpp:292-300)
        min_rec = 0;
        for (; min_rec < S.levels(); min_rec++)
            if (S.lv[min_rec].size() > 32)
                break;
        far_info.assign(S.levels(), FieldInfo());
        near_info.assign(S.levels(), FieldInfo());
        interpolation_range();
    }

    // --- mult helpers ----------------------------------------------------------------

    void sample_leaf_cone(const Node& nd, const ConeGrid& g, size_t el, const int* ho,
                          const std::vector<double>& nodes, const cplx* w, cplx* out) const
    {
        const size_t nn = nodes.size() / 3;
        const double H = nd.bb.side;
        for (size_t q = 0; q < nn; q++)
        {
            double sp[3], y[3];
            g.transform(el, &nodes[3 * q], sp);
            interp_to_cart(sp, nd.bb.c, H, y);
            double dv[3] = {y[0] - nd.bb.c[0], y[1] - nd.bb.c[1], y[2] - nd.bb.c[2]};
            double dc = norm3(dv);
            cplx acc = 0;
            for (int64_t i = nd.begin; i < nd.end; i++)
                acc += K.factored(&S.pts[3 * i], y, w[i], dc, i);
            out[q] = acc;
        }
        (void)ho;
    }

    // ifgfoperator.hpp:1539-1637
    void coarse_to_fine(const Node& nd, const InterpData& hi, size_t memId, const int* lo_order,
                        InterpData& fine) const
    {
        const ConeGrid& hg = nd.grid[1];
        const ConeGrid& fg = nd.grid[0];
        size_t factor[3];
        for (int d = 0; d < 3; d++)
            factor[d] = fg.n[d] / hg.n[d];
        std::vector<double> pts[3];
        size_t Np = 1;
        for (int d = 0; d < 3; d++)
        {
            auto c1 = chebnodes1d(lo_order[d]);
            pts[d].resize(c1.size() * factor[d]);
            Np *= pts[d].size();
            for (size_t j = 0; j < factor[d]; j++)
            {
                const double h = 2;
                double mn = -1 + (double(j) * (h / double(factor[d])));
                double mx = (mn + (h / double(factor[d])));
                const double a = 0.5 * (mx - mn), b = 0.5 * (mx + mn);
                for (size_t q = 0; q < c1.size(); q++)
                    pts[d][j * c1.size() + q] = (c1[q] * a) + b;
            }
        }
        std::vector<cplx> tmp(Np);
        tp_evaluate(pts, &hi.values[memId * hi.stride()], hi.order, tmp.data());
        const size_t cid = hg.active[memId];
        size_t hid[3];
        hg.indices(cid, hid);
        const size_t fstride = fine.stride();
        size_t idx = 0;
        for (size_t l = 0; l < factor[2] * lo_order[2]; l++)
            for (size_t j = 0; j < factor[1] * lo_order[1]; j++)
                for (size_t k = 0; k < factor[0]; k++)
                {
                    size_t fel = (hid[2] * factor[2] + (l / lo_order[2])) * fg.n[1] * fg.n[0] +
                                 (hid[1] * factor[1] + (j / lo_order[1])) * fg.n[0] +
                                 (hid[0] * factor[0] + k);
                    if (fg.is_active(fel))
                    {
                        size_t fm = fg.mem_id(fel);
                        size_t fi = (l % lo_order[2]) * lo_order[0] * lo_order[1] +
                                    (j % lo_order[1]) * lo_order[0];
                        for (int q = 0; q < lo_order[0]; q++)
                            fine.values[fm * fstride + fi + q] = tmp[idx + q];
                    }
                    idx += lo_order[0];
                }
    }

    // interpolant evaluation at one physical point y for box nd with data d (fine grid);
    // returns 0 outside the cone domain (ifgfoperator.hpp:1708-1734 without the CF)
    cplx eval_interp(const Node& nd, const InterpData& d, const double* y) const
    {
        double s[3];
        cart_to_interp(y, nd.bb.c, nd.bb.side, s);
        size_t el = d.grid->element_for_point(s);
        if (el == NONE)
            return 0;
        size_t m = d.grid->mem_id(el);
        double xi[3];
        d.grid->transform_back(el, s, xi);
        return clenshaw3(xi, &d.values[m * d.stride()], d.order);
    }

    // ifgfoperator.hpp:1639-1705: child data at the parent's high-order nodes
    void transfer_interp(const Node& child, const InterpData& cd, const Node& parent,
                         const ConeGrid& pg, size_t pel, const std::vector<double>& pnodes,
                         cplx* acc) const
    {
        const size_t nn = pnodes.size() / 3;
        for (size_t q = 0; q < nn; q++)
        {
            double sp[3], y[3];
            pg.transform(pel, &pnodes[3 * q], sp);
            interp_to_cart(sp, parent.bb.c, parent.bb.side, y);
            cplx v = eval_interp(child, cd, y);
            acc[q] += v * K.transfer(y, child.bb.c, parent.bb.c);
        }
    }

    void hi_transform(InterpData& d) const
    {
        const size_t st = d.stride();
        std::vector<cplx> tmp(st);
        for (size_t m = 0; m < d.grid->active.size(); m++)
        {
            chebtransform3(&d.values[m * st], tmp.data(), d.order);
            std::copy(tmp.begin(), tmp.end(), d.values.begin() + m * st);
        }
    }

    // recursive_mult per source box (ifgfoperator.hpp:1014-1187); the upward contribution
    // to the parent's high-order values is returned in `up` (size of the parent data).
    void recursive_mult(int level, int id, const cplx* w, cplx* result,
                        const InterpData* parent, std::vector<cplx>* up) const
    {
        if (id < 0)
            return;
        const Node& nd = S.lv[level][id];
        if (nd.begin == nd.end)
            return;
        const double H = nd.bb.side;
        InterpData st;
        K.order_for_box(H, base_order, 1, st.order);
        st.grid = &nd.grid[1];
        auto ho_nodes = chebnodes3d(st.order);
        st.values.assign(st.grid->active.size() * st.stride(), cplx(0));
        // near field (ifgfoperator.hpp:900-960)
        for (const Range& r : nd.near)
            for (int64_t t = r.b; t < r.e; t++)
            {
                cplx acc = 0;
                for (int64_t i = nd.begin; i < nd.end; i++)
                    acc += w[i] * K.p2p(&S.pts[3 * i], &T.pts[3 * t], i);
                result[t] += acc;
            }
        if (nd.leaf)
        {
            for (size_t m = 0; m < st.grid->active.size(); m++)
                sample_leaf_cone(nd, *st.grid, st.grid->active[m], st.order, ho_nodes, w,
                                 &st.values[m * st.stride()]);
        }
        else
        {
            for (int c = 0; c < 8; c++)
            {
                std::vector<cplx> cup(st.values.size(), cplx(0));
                recursive_mult(level + 1, nd.child[c], w, result, &st, &cup);
                for (size_t q = 0; q < cup.size(); q++)
                    st.values[q] += cup[q];
            }
        }
        hi_transform(st);
        InterpData fine;
        K.order_for_box(H, base_order, 0, fine.order);
        fine.grid = &nd.grid[0];
        fine.values.assign(fine.grid->active.size() * fine.stride(), cplx(0));
        for (size_t m = 0; m < st.grid->active.size(); m++)
            coarse_to_fine(nd, st, m, fine.order, fine);
        hi_transform(fine);
        // far field (ifgfoperator.hpp:962-1012, 1736-1891)
        for (const Range& r : nd.far)
            for (int64_t t = r.b; t < r.e; t++)
            {
                const double* y = &T.pts[3 * t];
                cplx v = eval_interp(nd, fine, y);
                double dv[3] = {y[0] - nd.bb.c[0], y[1] - nd.bb.c[1], y[2] - nd.bb.c[2]};
                result[t] += v * K.cf(dv);
            }
        // upward (ifgfoperator.hpp:1137-1185)
        if (level == 0 || parent == nullptr)
            return;
        const Node& pn = S.lv[level - 1][nd.parent];
        if (pn.begin == pn.end || !pn.has_far_anc)
            return;
        const ConeGrid& pg = *parent->grid;
        auto pnodes = chebnodes3d(parent->order);
        const size_t pst = parent->stride();
        for (size_t m = 0; m < pg.active.size(); m++)
            transfer_interp(nd, fine, pn, pg, pg.active[m], pnodes, &(*up)[m * pst]);
    }

    // the recursive phase restricted to every stride-th box of the first recursive level;
    // returns the fraction of source points in the sampled subtrees (bounded CPU timing) and,
    // in thread_seconds, the sum of the subtrees' durations (each runs on one thread)
    double mult_sample(const cplx* weights, int stride, double* thread_seconds = nullptr) const
    {
        std::vector<cplx> w(nsrc);
        for (int64_t i = 0; i < nsrc; i++)
            w[i] = weights[S.perm[i]];
        if (min_rec >= S.levels() || min_rec == 0)
            return 0.0;
        const int pl = min_rec - 1;
        std::vector<InterpData> idata(S.lv[pl].size());
        for (size_t p = 0; p < S.lv[pl].size(); p++)
        {
            const Node& pn = S.lv[pl][p];
            K.order_for_box(pn.bb.side, base_order, 1, idata[p].order);
            idata[p].grid = &pn.grid[1];
            idata[p].values.assign(pn.grid[1].active.size() * idata[p].stride(), cplx(0));
        }
        const auto& Lr = S.lv[min_rec];
        std::vector<int64_t> ids;
        int64_t pts = 0;
        for (size_t i = 0; i < Lr.size(); i += size_t(stride))
        {
            ids.push_back(int64_t(i));
            pts += Lr[i].end - Lr[i].begin;
        }
        std::vector<std::vector<cplx>> thread_res(std::max(1, nthreads));
        for (auto& r : thread_res)
            r.assign(ntgt, cplx(0));
        std::vector<double> dur(ids.size(), 0.0);
        parallel_for(int64_t(ids.size()), nthreads,
                     [&](int64_t q, int th)
                     {
                         const auto t0 = std::chrono::steady_clock::now();
                         const Node& nd = Lr[ids[q]];
                         std::vector<cplx> up(idata[nd.parent].values.size(), cplx(0));
                         recursive_mult(min_rec, int(ids[q]), w.data(), thread_res[th].data(),
                                        &idata[nd.parent], &up);
                         dur[q] = std::chrono::duration<double>(
                                      std::chrono::steady_clock::now() - t0)
                                      .count();
                     });
        if (thread_seconds)
        {
            *thread_seconds = 0.0;
            for (double d : dur)
                *thread_seconds += d;
        }
        return double(pts) / double(nsrc);
    }

    // ifgfoperator.hpp:246-867
    std::vector<cplx> mult(const cplx* weights) const
    {
        std::vector<cplx> w(nsrc);
        for (int64_t i = 0; i < nsrc; i++)
            w[i] = weights[S.perm[i]];
        std::vector<cplx> result(ntgt, cplx(0));
        const int L = S.levels();
        int level = min_rec;
        // data of level `level-1` (high-order, coarse grid), indexed by box
        std::vector<InterpData> idata;
        if (min_rec < L)
        {
            const int pl = min_rec - 1;
            idata.resize(S.lv[pl].size());
            for (size_t p = 0; p < S.lv[pl].size(); p++)
            {
                const Node& pn = S.lv[pl][p];
                K.order_for_box(pn.bb.side, base_order, 1, idata[p].order);
                idata[p].grid = &pn.grid[1];
                idata[p].values.assign(pn.grid[1].active.size() * idata[p].stride(), cplx(0));
            }
            const auto& Lr = S.lv[min_rec];
            std::vector<std::vector<cplx>> thread_res(std::max(1, nthreads));
            for (auto& r : thread_res)
                r.assign(ntgt, cplx(0));
            std::vector<std::vector<cplx>> ups(Lr.size());
            parallel_for(int64_t(Lr.size()), nthreads,
                         [&](int64_t id, int th)
                         {
                             const Node& nd = Lr[id];
                             ups[id].assign(idata[nd.parent].values.size(), cplx(0));
                             recursive_mult(min_rec, int(id), w.data(), thread_res[th].data(),
                                            &idata[nd.parent], &ups[id]);
                         });
            for (auto& r : thread_res)
                for (int64_t t = 0; t < ntgt; t++)
                    result[t] += r[t];
            for (size_t p = 0; p < S.lv[pl].size(); p++)
                for (int c = 0; c < 8; c++)
                {
                    int cid = S.lv[pl][p].child[c];
                    if (cid < 0)
                        continue;
                    for (size_t q = 0; q < ups[cid].size(); q++)
                        idata[p].values[q] += ups[cid][q];
                }
            level = min_rec - 1;
        }
        else
        {
            level = L - 1;
        }

        for (; level >= 0; --level)
        {
            const auto& Lv = S.lv[level];
            // near field per target (ifgfoperator.hpp:360-391)
            const FieldInfo& nf = near_info[level];
            parallel_for(ntgt, nthreads,
                         [&](int64_t t, int)
                         {
                             for (int64_t q = nf.starts[t]; q < nf.starts[t + 1]; q++)
                             {
                                 const Node& nd = Lv[nf.idx[q]];
                                 cplx acc = 0;
                                 for (int64_t i = nd.begin; i < nd.end; i++)
                                     acc += w[i] * K.p2p(&S.pts[3 * i], &T.pts[3 * t], i);
                                 result[t] += acc;
                             }
                         });
            if (level == L - 1 || idata.empty())
            {
                idata.assign(Lv.size(), InterpData());
                for (size_t b = 0; b < Lv.size(); b++)
                {
                    K.order_for_box(Lv[b].bb.side, base_order, 1, idata[b].order);
                    idata[b].grid = &Lv[b].grid[1];
                    idata[b].values.assign(Lv[b].grid[1].active.size() * idata[b].stride(),
                                           cplx(0));
                }
            }
            // leaf sampling (ifgfoperator.hpp:423-462)
            parallel_for(int64_t(Lv.size()), nthreads,
                         [&](int64_t b, int)
                         {
                             const Node& nd = Lv[b];
                             if (!nd.leaf || !nd.has_far_anc)
                                 return;
                             auto nodes = chebnodes3d(idata[b].order);
                             for (size_t m = 0; m < nd.grid[1].active.size(); m++)
                                 sample_leaf_cone(nd, nd.grid[1], nd.grid[1].active[m],
                                                  idata[b].order, nodes, w.data(),
                                                  &idata[b].values[m * idata[b].stride()]);
                         });
            // high-order transform, coarse->fine, fine transform (:466-570)
            std::vector<InterpData> fdata(Lv.size());
            parallel_for(int64_t(Lv.size()), nthreads,
                         [&](int64_t b, int)
                         {
                             const Node& nd = Lv[b];
                             K.order_for_box(nd.bb.side, base_order, 0, fdata[b].order);
                             fdata[b].grid = &nd.grid[0];
                             fdata[b].values.assign(nd.grid[0].active.size() * fdata[b].stride(),
                                                    cplx(0));
                             if (!nd.has_far_anc)
                                 return;
                             hi_transform(idata[b]);
                             for (size_t m = 0; m < nd.grid[1].active.size(); m++)
                                 coarse_to_fine(nd, idata[b], m, fdata[b].order, fdata[b]);
                             hi_transform(fdata[b]);
                         });
            // far field per target (ifgfoperator.hpp:577-616)
            const FieldInfo& ff = far_info[level];
            parallel_for(ntgt, nthreads,
                         [&](int64_t t, int)
                         {
                             const double* y = &T.pts[3 * t];
                             for (int64_t q = ff.starts[t]; q < ff.starts[t + 1]; q++)
                             {
                                 const Node& nd = Lv[ff.idx[q]];
                                 cplx v = eval_interp(nd, fdata[ff.idx[q]], y);
                                 double dv[3] = {y[0] - nd.bb.c[0], y[1] - nd.bb.c[1],
                                                 y[2] - nd.bb.c[2]};
                                 result[t] += v * K.cf(dv);
                             }
                         });
            if (level < 1)
                break;
            // upward pass (ifgfoperator.hpp:631-688)
            const auto& Lp = S.lv[level - 1];
            std::vector<InterpData> pdata(Lp.size());
            parallel_for(int64_t(Lp.size()), nthreads,
                         [&](int64_t p, int)
                         {
                             const Node& pn = Lp[p];
                             K.order_for_box(pn.bb.side, base_order, 1, pdata[p].order);
                             pdata[p].grid = &pn.grid[1];
                             pdata[p].values.assign(pn.grid[1].active.size() * pdata[p].stride(),
                                                    cplx(0));
                             if (!pn.has_far_anc)
                                 return;
                             auto pnodes = chebnodes3d(pdata[p].order);
                             const size_t pst = pdata[p].stride();
                             for (size_t m = 0; m < pn.grid[1].active.size(); m++)
                                 for (int c = 0; c < 8; c++)
                                 {
                                     int cid = pn.child[c];
                                     if (cid < 0)
                                         continue;
                                     transfer_interp(Lv[cid], fdata[cid], pn, pn.grid[1],
                                                     pn.grid[1].active[m], pnodes,
                                                     &pdata[p].values[m * pst]);
                                 }
                         });
            idata.swap(pdata);
        }
        std::vector<cplx> out(ntgt);
        for (int64_t i = 0; i < ntgt; i++)
            out[T.perm[i]] = result[i];
        return out;
    }
};

} // namespace orc

// ----------------------------------------------------------------------------------------
// C interface for the tests (loaded with ctypes)
// ----------------------------------------------------------------------------------------
using namespace orc;

extern "C"
{
    struct OrcHandle
    {
        std::unique_ptr<Operator> op;
        std::string err;
        std::vector<std::vector<int64_t>> scratch;
    };

    void* orc_create(int kind, double k_re, double k_im, int64_t leaf, int order, int n_elem,
                     double tol, int dx)
    {
        Kernel K;
        K.kind = kind;
        K.k = cplx(k_re, k_im);
        K.kappa = k_re;
        K.dx = dx;
        auto* h = new OrcHandle();
        h->op.reset(new Operator(K, leaf, order, n_elem, tol));
        return h;
    }

    void orc_destroy(void* h) { delete static_cast<OrcHandle*>(h); }

    const char* orc_last_error(void* h) { return static_cast<OrcHandle*>(h)->err.c_str(); }

    int orc_init(void* hv, const double* srcs, int64_t N, const double* tgts, int64_t M,
                 const double* normals, int nthreads)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        try
        {
            h->op->nthreads = nthreads;
            h->op->init(srcs, N, tgts, M, normals);
        }
        catch (const std::exception& e)
        {
            h->err = e.what();
            return -1;
        }
        return 0;
    }

    int orc_mult(void* hv, const double* w, double* out, int nthreads)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        try
        {
            h->op->nthreads = nthreads;
            auto r = h->op->mult(reinterpret_cast<const cplx*>(w));
            std::memcpy(out, r.data(), r.size() * sizeof(cplx));
        }
        catch (const std::exception& e)
        {
            h->err = e.what();
            return -1;
        }
        return 0;
    }

    // cutOff callback of init (ifgfoperator.hpp:65-67) as a distance threshold: admissible
    // pairs farther than dmax are dropped (dmax <= 0: keep all)
    void orc_set_cutoff(void* hv, double dmax)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        if (dmax > 0)
            h->op->cutoff = [dmax](double d) { return d > dmax; };
        else
            h->op->cutoff = nullptr;
    }

    double orc_mult_sample(void* hv, const double* w, int stride, int nthreads,
                           double* thread_seconds)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        h->op->nthreads = nthreads;
        return h->op->mult_sample(reinterpret_cast<const cplx*>(w), stride, thread_seconds);
    }

    // util.hpp:214-234 on an interleaved complex array of rows = n0 n1 n2 values
    double orc_slice_norm(const double* data, int64_t rows, const int* ns, int axis, int layers)
    {
        return compute_slice_norm(reinterpret_cast<const cplx*>(data), size_t(rows), ns, axis,
                                  layers);
    }

    // one sampling step of the order estimate on source box `id` of `level` at `order`:
    // nodes (n x 3 Cartesian), samples and coefficients (n complex), info (8, as
    // Operator::estimate_sample); returns n = order[0] order[1] order[2]
    int64_t orc_estimate_sample(void* hv, int level, int64_t id, const int* order, double* pts,
                                double* data, double* trafo, double* info)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        std::vector<double> y;
        std::vector<cplx> d, t;
        h->op->estimate_sample(level, int(id), order, y, d, t, info);
        std::memcpy(pts, y.data(), y.size() * sizeof(double));
        std::memcpy(data, d.data(), d.size() * sizeof(cplx));
        std::memcpy(trafo, t.data(), t.size() * sizeof(cplx));
        return int64_t(d.size());
    }

    void orc_base_order(void* hv, int* out)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        for (int i = 0; i < 3; i++)
            out[i] = h->op->base_order[i];
    }

    int orc_levels(void* hv, int tree)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        return (tree == 0 ? h->op->S : h->op->T).levels();
    }

    int64_t orc_num_boxes(void* hv, int tree, int level)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        return int64_t((tree == 0 ? h->op->S : h->op->T).lv[level].size());
    }

    void orc_perm(void* hv, int tree, int64_t* out)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        const Tree& t = (tree == 0 ? h->op->S : h->op->T);
        std::memcpy(out, t.perm.data(), t.perm.size() * sizeof(int64_t));
    }

    void orc_root_box(void* hv, int tree, double* out)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        const Tree& t = (tree == 0 ? h->op->S : h->op->T);
        for (int i = 0; i < 3; i++)
        {
            out[i] = t.root_box.lo[i];
            out[3 + i] = t.root_box.hi[i];
        }
    }

    // per box: begin, end, parent, leaf, bbox lo/hi
    void orc_tree_level(void* hv, int tree, int level, int64_t* begin, int64_t* end,
                        int32_t* parent, int8_t* leaf, double* bbox)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        const Tree& t = (tree == 0 ? h->op->S : h->op->T);
        const auto& L = t.lv[level];
        for (size_t i = 0; i < L.size(); i++)
        {
            begin[i] = L[i].begin;
            end[i] = L[i].end;
            parent[i] = L[i].parent;
            leaf[i] = L[i].leaf ? 1 : 0;
            for (int d = 0; d < 3; d++)
            {
                bbox[6 * i + d] = L[i].bb.lo[d];
                bbox[6 * i + 3 + d] = L[i].bb.hi[d];
            }
        }
    }

    // interaction lists of source level `level`: counts then rows (box, tbegin, tend)
    int64_t orc_num_interactions(void* hv, int level, int far)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        int64_t n = 0;
        for (const Node& nd : h->op->S.lv[level])
            n += int64_t((far ? nd.far : nd.near).size());
        return n;
    }

    void orc_interactions(void* hv, int level, int far, int64_t* rows)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        int64_t q = 0;
        const auto& L = h->op->S.lv[level];
        for (size_t b = 0; b < L.size(); b++)
            for (const Range& r : (far ? L[b].far : L[b].near))
            {
                rows[3 * q + 0] = int64_t(b);
                rows[3 * q + 1] = r.b;
                rows[3 * q + 2] = r.e;
                q++;
            }
    }

    // cone grid of (level, step): per box n_elements[3]; active counts
    void orc_cone_counts(void* hv, int level, int step, int64_t* nel, int64_t* nactive)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        const auto& L = h->op->S.lv[level];
        for (size_t b = 0; b < L.size(); b++)
        {
            for (int d = 0; d < 3; d++)
                nel[3 * b + d] = int64_t(L[b].grid[step].n[d]);
            nactive[b] = int64_t(L[b].grid[step].active.size());
        }
    }

    // active cone ids of (level, step) concatenated over boxes, each box's list sorted
    void orc_cone_active(void* hv, int level, int step, int64_t* ids)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        const auto& L = h->op->S.lv[level];
        int64_t q = 0;
        for (size_t b = 0; b < L.size(); b++)
        {
            std::vector<size_t> a = L[b].grid[step].active;
            std::sort(a.begin(), a.end());
            for (size_t x : a)
                ids[q++] = int64_t(x);
        }
    }

    int orc_min_rec(void* hv) { return static_cast<OrcHandle*>(hv)->op->min_rec; }

    // --- primitives ------------------------------------------------------------------
    void orc_chebnodes1d(int n, double* out)
    {
        auto x = chebnodes1d(n);
        std::memcpy(out, x.data(), x.size() * sizeof(double));
    }

    void orc_chebtransform(const int* ns, const double* src, double* dst)
    {
        chebtransform3(reinterpret_cast<const cplx*>(src), reinterpret_cast<cplx*>(dst), ns);
    }

    void orc_clenshaw(const int* ns, const double* coef, const double* pts, int64_t n,
                      double* out)
    {
        auto* o = reinterpret_cast<cplx*>(out);
        for (int64_t i = 0; i < n; i++)
            o[i] = clenshaw3(pts + 3 * i, reinterpret_cast<const cplx*>(coef), ns);
    }

    void orc_morton_sort(const double* pts, int64_t n, int64_t* perm, double* box_out)
    {
        double mn[3] = {pts[0], pts[1], pts[2]}, mx[3] = {pts[0], pts[1], pts[2]};
        for (int64_t i = 0; i < n; i++)
            for (int d = 0; d < 3; d++)
            {
                mn[d] = std::min(mn[d], pts[3 * i + d]);
                mx[d] = std::max(mx[d], pts[3 * i + d]);
            }
        double nmn = norm3(mn), nmx = norm3(mx);
        for (int d = 0; d < 3; d++)
        {
            mn[d] -= 1e-8 * nmn;
            mx[d] += 1e-8 * nmx;
        }
        Box3 bb = make_box(mn, mx);
        for (int d = 0; d < 3; d++)
        {
            box_out[d] = mn[d];
            box_out[3 + d] = mx[d];
        }
        std::vector<int64_t> p(n);
        std::iota(p.begin(), p.end(), 0);
        MortonLess less(bb);
        std::sort(p.begin(), p.end(),
                  [&](int64_t a, int64_t b) { return less(pts + 3 * a, pts + 3 * b); });
        std::memcpy(perm, p.data(), n * sizeof(int64_t));
    }

    int orc_morton_less(const double* box, const double* p, const double* q)
    {
        Box3 bb = make_box(box, box + 3);
        return MortonLess(bb)(p, q) ? 1 : 0;
    }

    void orc_cart_to_interp(const double* x, const double* xc, double H, double* out)
    {
        cart_to_interp(x, xc, H, out);
    }

    void orc_interp_to_cart(const double* p, const double* xc, double H, double* out)
    {
        interp_to_cart(p, xc, H, out);
    }

    // kernel primitives for the tests
    void orc_kernel_eval(void* hv, const double* xs, const double* y, int64_t src_idx,
                         double* out)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        cplx v = h->op->K.p2p(xs, y, src_idx);
        out[0] = v.real();
        out[1] = v.imag();
    }

    // FP64 direct sum on selected targets (the reference tests' check, test_ifgf.cpp:143-165)
    void orc_direct(void* hv, const double* srcs, int64_t N, const double* w,
                    const double* normals, const double* tgts, const int64_t* idx, int64_t nidx,
                    double* out, int nthreads)
    {
        auto* h = static_cast<OrcHandle*>(hv);
        Kernel K = h->op->K;
        K.normals = normals;
        const cplx* ww = reinterpret_cast<const cplx*>(w);
        cplx* o = reinterpret_cast<cplx*>(out);
        parallel_for(nidx, nthreads,
                     [&](int64_t q, int)
                     {
                         const double* y = tgts + 3 * idx[q];
                         cplx acc = 0;
                         for (int64_t i = 0; i < N; i++)
                             acc += ww[i] * K.p2p(srcs + 3 * i, y, i);
                         o[q] = acc;
                     });
    }
}
